// gfx950 kernels and the device-resident solver for the 2D acoustic
// isotropic time step (potential formulation, one component per point).
//
// The reference's acoustic update (include/kokkos_kernels/domain_kernels.hpp:
// 72-141 with medium acoustic) is the elastic one with a scalar field:
//   source      a += -stf L / kappa      (medium/dim2/acoustic/isotropic/source.hpp:36-70)
//   stiffness   T = rho^-1 grad chi      (acoustic/isotropic/stress.hpp:15-56)
//               F = J (T . dxi/dx), r = -weighted divergence of F
//               (compute_stiffness_interaction.tpp:140-247)
//   Stacey      r += -(w |n|) (1 / (rho vp)) chi_dot   (stacey.tpp:293-321)
//   free surface r = 0 at tagged points of the element (dirichlet.hpp:20-45;
//               composite points get the Stacey term and then the zero,
//               composite_stacey_dirichlet.tpp:11-45)
//   mass        M = sum (1/kappa) w w J + (dt/2) w |n| / (rho vp)
//               (acoustic/isotropic/mass_matrix.tpp:1-14, stacey.tpp:817-853)
//   seismograms d/v/a: interpolated rho^-1 grad of chi / chi_dot / chi_ddot;
//               pressure: interpolated -chi_ddot (acoustic/isotropic/wavefield.hpp:12-97)
//
// Kernels mirror the elastic tile design (solver2d.hip): k_ac_tile runs the
// sources, the element stiffness and the colour-ordered LDS assembly of one
// tile, then the divide / corrector / next predictor of the tile's exclusive
// points; k_ac_shared finishes the points shared between tiles from the
// tiles' partials in fixed order; k_ac_seis_fused reduces the seismograms.
// One FP32 value per point per field: the tile LDS is about half the elastic
// footprint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "solver2d.hpp"

namespace smx {

using AcStreamSwap = StreamSwapT<hipStream_t>;

#define AC_HIP(call)                                                           \
  do {                                                                         \
    hipError_t err_ = (call);                                                  \
    if (err_ != hipSuccess)                                                    \
      throw SmxError(SMX_ERR_HIP, std::string(#call) + ": " +                  \
                                      hipGetErrorString(err_));                \
  } while (0)

namespace {

constexpr int AC_TILE_THREADS = NGLL * ELEM_TILE;
constexpr int AC_NODE_THREADS = 256;
constexpr int AC_SEIS_GROUPS = 8;

template <typename real> struct AcQuad {
  real hp[NGLL][NGLL]; // hp[i][j] = l'_j(xi_i)
  real w[NGLL];
};

struct AcTypes {
  int t[4];
};

enum AcMode : int { AC_STEP = 0, AC_FINAL = 1, AC_ASSEMBLE = 2 };

// s_u[nloc_max], s_a[nloc_max], then a region reused as the staged element
// potential, the xi integrand and the corrected velocity of the tile's points.
template <typename real> constexpr size_t ac_lds_bytes(int nloc_max) {
  return (2 * size_t(nloc_max) +
          std::max(size_t(NPT) * ELEM_TILE, size_t(nloc_max))) *
         sizeof(real);
}

// Seismogram term of one receiver entry: (wx f, wz f) for d / v / a,
// (h (-f), 0) for pressure (the reference negates the point value and then
// interpolates).
template <typename real>
__device__ __forceinline__ void ac_term(int type, real f, real wx, real wz,
                                        real h, real &tx, real &tz) {
  if (type == 3) {
    tx = h * (real(-1.0) * f);
    tz = 0;
  } else {
    tx = wx * f;
    tz = wz * f;
  }
}

template <typename real>
__device__ __forceinline__ real ac_shared_sum(const real *__restrict__ part,
                                              const int *__restrict__ shp_off,
                                              const int *__restrict__ shp_pos,
                                              int j) {
  real acc = 0;
  for (int k = shp_off[j]; k < shp_off[j + 1]; ++k)
    acc += part[shp_pos[k]];
  return acc;
}

// Exclusive points per thread whose mass inverse and velocity are requested
// before the colour-ordered assembly (an interior 8x8-element tile has 961
// exclusive points, about 3 per thread).
constexpr int kAcPrefetch = 4;

// The FP32 time step is held to 8 waves per SIMD (at most 64 VGPRs): its
// ~15 KB of LDS then lets 6 tiles share a CU (30 waves); measured 124 -> 100
// us per launch on the 800x600 workload together with the prefetch.
template <typename real, int MODE>
__global__ void __launch_bounds__(AC_TILE_THREADS) __attribute__((
    amdgpu_waves_per_eu(sizeof(real) == 4 && MODE == AC_STEP ? 8 : 1, 8)))
    k_ac_tile(const TileInfo *__restrict__ tiles, int nloc_max,
              const int *__restrict__ sh_list, const uint16_t *__restrict__ lidx,
              const uint8_t *__restrict__ ecolor, int ncolors,
              const real *__restrict__ geom, const real *__restrict__ mat,
              const real *__restrict__ geom_e, const real *__restrict__ mat_e,
              const int *__restrict__ sslot, const real *__restrict__ stacey,
              const uint32_t *__restrict__ dmask, real *__restrict__ U,
              real *__restrict__ V, real *__restrict__ A,
              const real *__restrict__ Minv, real *__restrict__ part,
              const int *__restrict__ src_li, const int *__restrict__ src_idx,
              const real *__restrict__ src_L, const real *__restrict__ stf,
              int nsrc_total, const int *__restrict__ trec_li,
              const int *__restrict__ trec_entry,
              const real *__restrict__ rent_w, const real *__restrict__ rent_h,
              real *__restrict__ seis_terms, int nentries, AcTypes types,
              int ntypes, int nsbs, const int *__restrict__ step_ptr,
              int istep, int ld, int sld, real dt, real dto2, real dt2o2,
              AcQuad<real> q, const int *__restrict__ cpl_li,
              const int *__restrict__ cpl_g, const real *__restrict__ cpl_w,
              const real *__restrict__ Ue, const int *__restrict__ cout_li,
              const int *__restrict__ cout_k, real *__restrict__ chi_dd,
              const int *__restrict__ sh_pad, int sh_stride) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  real *s_u = reinterpret_cast<real *>(lds_raw);
  real *s_a = s_u + nloc_max;
  real *xr = s_a + nloc_max;
  auto X = [&](int p, int l) -> real & { return xr[p * ELEM_TILE + l]; };
  real *s_v = xr;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int ix = __builtin_amdgcn_readfirstlane(tid >> 6);
  const TileInfo T = tiles[blockIdx.x];
  const int nloc = T.nex + T.nsh;
  const int slot = blockIdx.x * ELEM_TILE + lane;
  const bool valid = lane < T.ecount;
  const size_t pld = size_t(NPT) * ld;
  const int step = step_ptr ? *step_ptr : istep;

  int li[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz)
    li[iz] = valid ? int(lidx[size_t(iz * NGLL + ix) * ld + slot]) : 0;

  // ---- tile points: potential in, accumulator cleared (shared points'
  // global ids at a fixed stride per tile, padded with id 0, so that their
  // loads do not wait on the tile record)
  for (int s = tid; s < sh_stride; s += AC_TILE_THREADS) {
    const real u = U[sh_pad[size_t(blockIdx.x) * sh_stride + s]];
    if (s < T.nsh) {
      s_u[T.nex + s] = u;
      s_a[T.nex + s] = real(0);
    }
  }
  for (int i = tid; i < T.nex; i += AC_TILE_THREADS) {
    const int g = T.ex_off + i;
    s_u[i] = U[g];
    s_a[i] = MODE == AC_ASSEMBLE ? A[g] : real(0);
  }
  __syncthreads();
  // point sources first, in list order: -stf L / kappa
  if (tid == 0)
    for (int k = T.src_off; k < T.src_off + T.nsrc; ++k) {
      const int l = src_li[k], s = src_idx[k];
      const real f = stf[(size_t(step) * nsrc_total + s) * 2];
      s_a[l] += (real(-1.0) * f * src_L[2 * k]) / src_L[2 * k + 1];
    }
  // coupling with the solid of a coupled mesh: chi_dd += w (n . u), u the
  // elastic displacement after this step's predictor (coupling_terms/
  // acoustic_elastic.hpp:13-26), entries in tile order
  if (tid == 0)
    for (int k = T.cpl_off; k < T.cpl_off + T.ncpl; ++k) {
      const int g = cpl_g[k];
      s_a[cpl_li[k]] += cpl_w[3 * k] * (cpl_w[3 * k + 1] * Ue[2 * size_t(g)] +
                                        cpl_w[3 * k + 2] * Ue[2 * size_t(g) + 1]);
    }

  // ---- element potential (column ix, rows 0..4)
  real c[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    c[iz] = s_u[li[iz]];
    X(iz * NGLL + ix, lane) = c[iz];
  }
  __syncthreads();

  // ---- gradient in (xi, gamma) (include/algorithms/gradient.hpp:40-82)
  real dxi[NGLL], dga[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    real a = 0, b = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      a += q.hp[ix][l] * X(iz * NGLL + l, lane);
      b += q.hp[iz][l] * c[l];
    }
    dxi[iz] = a;
    dga[iz] = b;
  }
  __syncthreads(); // the staging area becomes the xi integrand

  // ---- T = rho^-1 grad chi, F = J T . dxi/dx (point/stress.hpp:166-183)
  const bool g_el = (T.flags & TILE_GEOM_ELEM) != 0;
  const bool m_el = (T.flags & TILE_MAT_ELEM) != 0;
  const size_t gst = g_el ? size_t(ld) : pld;
  real f1[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    const int p = iz * NGLL + ix;
    const size_t o = size_t(p) * ld + slot;
    const real *gp = g_el ? geom_e + slot : geom + o;
    const real xix = gp[0], xiz = gp[gst], gax = gp[2 * gst], gaz = gp[3 * gst],
               jac = gp[4 * gst];
    const real rinv = m_el ? mat_e[slot] : mat[o];
    const real dx = xix * dxi[iz] + gax * dga[iz];
    const real dz = xiz * dxi[iz] + gaz * dga[iz];
    const real tx = rinv * dx, tz = rinv * dz;
    X(p, lane) = jac * (tx * xix + tz * xiz);
    f1[iz] = jac * (tx * gax + tz * gaz);
  }
  __syncthreads();

  // ---- divergence (divergence.hpp:40-74), negation, Stacey, free surface
  const int bslot = valid ? sslot[slot] : -1;
  const uint32_t dm = valid ? dmask[slot] : 0u;
  real r[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    real t1 = 0, t2 = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      t1 += X(iz * NGLL + l, lane) * q.hp[l][ix] * q.w[l];
      t2 += f1[l] * q.hp[l][iz] * q.w[l];
    }
    real a = q.w[iz] * t1 + q.w[ix] * t2;
    a *= real(-1.0);
    const int p = iz * NGLL + ix;
    if (bslot >= 0) {
      const size_t so = size_t(p) * sld + bslot;
      const size_t sp = size_t(NPT) * sld;
      const real nrm = stacey[2 * sp + so], wgt = stacey[3 * sp + so],
                 rvpi = stacey[4 * sp + so];
      const int g = li[iz] < T.nex ? T.ex_off + li[iz]
                                   : sh_list[T.sh_off + li[iz] - T.nex];
      const real factor = wgt * nrm;
      a += real(-1.0) * factor * rvpi * V[g];
    }
    if ((dm >> p) & 1u)
      a = 0;
    r[iz] = a;
  }

  // mass inverse and velocity of this thread's first exclusive points,
  // requested before the colour-ordered assembly so their latency overlaps it
  real pre_m[kAcPrefetch], pre_v[kAcPrefetch];
#pragma unroll
  for (int k = 0; k < kAcPrefetch; ++k) {
    const int i = tid + k * AC_TILE_THREADS;
    if (i < T.nex) {
      pre_m[k] = Minv[T.ex_off + i];
      if (MODE != AC_ASSEMBLE)
        pre_v[k] = V[T.ex_off + i];
    }
  }

  // ---- assembly into the tile accumulator, colour by colour
  const int ecol = valid ? int(ecolor[slot]) : -1;
  for (int cc = 0; cc < ncolors; ++cc) {
    if (ecol == cc) {
#pragma unroll
      for (int iz = 0; iz < NGLL; ++iz)
        s_a[li[iz]] += r[iz];
    }
    __syncthreads();
  }

  // ---- exclusive points: finish the step; shared points: store partials
  const bool sample = seis_terms != nullptr && (step % nsbs) == 0;
  for (int i = tid, k = 0; i < T.nex; i += AC_TILE_THREADS, ++k) {
    const int g = T.ex_off + i;
    const real m = k < kAcPrefetch ? pre_m[k] : Minv[g];
    if (MODE == AC_ASSEMBLE) {
      A[g] = s_a[i] * m;
      continue;
    }
    real v = k < kAcPrefetch ? pre_v[k] : V[g];
    real a = s_a[i] * m;
    v += dto2 * a;
    if (sample || T.ncout > 0)
      s_a[i] = a;
    if (sample)
      s_v[i] = v;
    if (MODE == AC_STEP) {
      real u = s_u[i];
      u += dt * v + dt2o2 * a;
      v += dto2 * a;
      U[g] = u;
      V[g] = v;
    } else {
      V[g] = v;
      A[g] = a;
    }
  }
  for (int i = T.nex + tid; i < nloc; i += AC_TILE_THREADS)
    part[T.sh_off + i - T.nex] = s_a[i];
  // new accelerations of the tile's exclusive interface points, read by the
  // elastic update of this step
  if (MODE != AC_ASSEMBLE && T.ncout > 0) {
    __syncthreads();
    for (int k = T.cout_off + tid; k < T.cout_off + T.ncout; k += AC_TILE_THREADS)
      chi_dd[cout_k[k]] = s_a[cout_li[k]];
  }

  // ---- receiver terms on exclusive points (state of step n)
  if (MODE != AC_ASSEMBLE && sample && T.nrec > 0) {
    __syncthreads();
    for (int k = T.rec_off + tid; k < T.rec_off + T.nrec; k += AC_TILE_THREADS) {
      const int l = trec_li[k], id = trec_entry[k];
      const real wx = rent_w[2 * id], wz = rent_w[2 * id + 1], h = rent_h[id];
      for (int it = 0; it < ntypes; ++it) {
        const int type = types.t[it];
        const real f = type == 0 ? s_u[l] : (type == 1 ? s_v[l] : s_a[l]);
        real tx, tz;
        ac_term(type, f, wx, wz, h, tx, tz);
        real *o = seis_terms + (size_t(it) * nentries + id) * 2;
        o[0] = tx;
        o[1] = tz;
      }
    }
  }
}

// Operator path of a coupled mesh: chi_dd[k] = the new acceleration of
// interface point k (global point iface_g[k]).
template <typename real>
__global__ void k_ac_iface(const real *__restrict__ A,
                           const int *__restrict__ iface_g, int n,
                           real *__restrict__ chi_dd) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n)
    chi_dd[k] = A[iface_g[k]];
}

// Points shared between tiles: partials summed in tile order, then the same
// divide / corrector / predictor as the tile's exclusive points.
template <typename real, int MODE>
__global__ void __launch_bounds__(AC_NODE_THREADS)
    k_ac_shared(real *__restrict__ U, real *__restrict__ V, real *__restrict__ A,
                const real *__restrict__ Minv, const real *__restrict__ part,
                const int *__restrict__ shp_off, const int *__restrict__ shp_pos,
                int nshared, int nex_total, real dt, real dto2, real dt2o2,
                int *__restrict__ step_ptr, const int *__restrict__ sh_iface,
                real *__restrict__ chi_dd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0 && step_ptr && MODE != AC_ASSEMBLE)
    *step_ptr += 1;
  if (j >= nshared)
    return;
  const int g = nex_total + j;
  // the point's own 1/M, v (or a) and u first: they do not wait on the
  // partials
  const real m = Minv[g];
  const real a0 = MODE == AC_ASSEMBLE ? A[g] : real(0);
  real v = MODE == AC_ASSEMBLE ? real(0) : V[g];
  real u = MODE == AC_STEP ? U[g] : real(0);
  const real acc = ac_shared_sum(part, shp_off, shp_pos, j);
  if (MODE == AC_ASSEMBLE) {
    A[g] = (a0 + acc) * m;
    return;
  }
  real a = acc * m;
  v += dto2 * a;
  if (sh_iface) {
    const int k = sh_iface[j];
    if (k >= 0)
      chi_dd[k] = a;
  }
  if (MODE == AC_STEP) {
    u += dt * v + dt2o2 * a;
    v += dto2 * a;
    U[g] = u;
    V[g] = v;
  } else {
    V[g] = v;
    A[g] = a;
  }
}

// Fused-path seismograms between the tile launch and k_ac_shared: entries on
// shared points are finished here from the partials (same arithmetic as
// k_ac_shared); the 25 terms of a receiver are summed in point order.
template <typename real>
__global__ void k_ac_seis_fused(const real *__restrict__ U,
                                const real *__restrict__ V,
                                const real *__restrict__ Minv,
                                const real *__restrict__ part,
                                const int *__restrict__ shp_off,
                                const int *__restrict__ shp_pos,
                                const int *__restrict__ rent_g,
                                const int *__restrict__ rent_shared,
                                const real *__restrict__ rent_w,
                                const real *__restrict__ rent_h,
                                const real *__restrict__ seis_terms, int nrec,
                                AcTypes types, int ntypes, int nsbs,
                                const int *__restrict__ step_ptr, real dto2,
                                real *__restrict__ out) {
  __shared__ real s_t[AC_SEIS_GROUPS][NPT][2];
  const int grp = threadIdx.x >> 5, p = threadIdx.x & 31;
  const int pair = blockIdx.x * AC_SEIS_GROUPS + grp;
  const int r = pair % nrec, it = pair / nrec;
  const bool active = it < ntypes && p < NPT;
  // the entry's own data is requested before the step counter is known
  // (all in bounds for every active lane)
  const int id = r * NPT + p;
  int j = -1, g = 0;
  real t0 = 0, t1 = 0;
  if (active) {
    j = rent_shared[id];
    g = rent_g[id];
    const size_t o = (size_t(it) * nrec * NPT + id) * 2;
    t0 = seis_terms[o];
    t1 = seis_terms[o + 1];
  }
  const int step = *step_ptr;
  if (step % nsbs != 0)
    return;
  const int isig = step / nsbs;
  if (active) {
    const int type = types.t[it];
    real tx, tz;
    if (j < 0) {
      tx = t0;
      tz = t1;
    } else {
      real f;
      if (type == 0) {
        f = U[g];
      } else {
        // the point's own 1/M and v before its partial sums
        const real m = Minv[g];
        real v = V[g];
        real a = ac_shared_sum(part, shp_off, shp_pos, j) * m;
        v += dto2 * a;
        f = type == 1 ? v : a;
      }
      ac_term(type, f, rent_w[2 * id], rent_w[2 * id + 1], rent_h[id], tx, tz);
    }
    s_t[grp][p][0] = tx;
    s_t[grp][p][1] = tz;
  }
  __syncthreads();
  if (active && p == 0) {
    real sx = 0, sz = 0;
    for (int k = 0; k < NPT; ++k) {
      sx += s_t[grp][k][0];
      sz += s_t[grp][k][1];
    }
    const size_t o = ((size_t(isig) * ntypes + it) * nrec + r) * 2;
    out[o] = sx;
    out[o + 1] = sz;
  }
}

// Operator-path seismograms from the finished fields.
template <typename real>
__global__ void k_ac_seismo(const real *__restrict__ U,
                            const real *__restrict__ V,
                            const real *__restrict__ A,
                            const int *__restrict__ rent_g,
                            const real *__restrict__ rent_w,
                            const real *__restrict__ rent_h, int nrec,
                            AcTypes types, int ntypes, int isig,
                            real *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = t % nrec, it = t / nrec;
  if (it >= ntypes)
    return;
  const int type = types.t[it];
  const real *F = type == 0 ? U : (type == 1 ? V : A);
  real sx = 0, sz = 0;
  for (int p = 0; p < NPT; ++p) {
    const int id = r * NPT + p;
    real tx, tz;
    ac_term(type, F[rent_g[id]], rent_w[2 * id], rent_w[2 * id + 1], rent_h[id],
            tx, tz);
    sx += tx;
    sz += tz;
  }
  const size_t o = ((size_t(isig) * ntypes + it) * nrec + r) * 2;
  out[o] = sx;
  out[o + 1] = sz;
}

// Newmark pieces on [nglob_pad] arrays (operator path and first predictor).
template <typename real>
__global__ void __launch_bounds__(AC_NODE_THREADS)
    k_ac_predictor(real *__restrict__ U, real *__restrict__ V,
                   real *__restrict__ A, long n, real dt, real dto2,
                   real dt2o2) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  real u = U[i], v = V[i];
  const real a = A[i];
  u += dt * v + dt2o2 * a;
  v += dto2 * a;
  U[i] = u;
  V[i] = v;
  A[i] = 0;
}

template <typename real>
__global__ void __launch_bounds__(AC_NODE_THREADS)
    k_ac_corrector(real *__restrict__ V, const real *__restrict__ A, long n,
                   real dto2) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    V[i] += dto2 * A[i];
}

// Mass: M[ibool] += (1/kappa) (w_z w_x J) + the Stacey term with the
// velocity set to -dt/2; one launch per element colour.
template <typename real>
__global__ void k_ac_mass(real *__restrict__ M, const int *__restrict__ ibool,
                          const uint8_t *__restrict__ ecolor, int color,
                          const real *__restrict__ geom,
                          const real *__restrict__ kmr,
                          const int *__restrict__ sslot,
                          const real *__restrict__ stacey, int ld, int sld,
                          AcQuad<real> q, real dt) {
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(NPT) * ld)
    return;
  const int p = int(t / ld), e = int(t % ld);
  if (int(ecolor[e]) != color)
    return;
  const size_t pld = size_t(NPT) * ld;
  const int iz = p / NGLL, ix = p % NGLL;
  const real kinv = real(1.0) / kmr[t];
  const real f = q.w[iz] * q.w[ix] * geom[4 * pld + t];
  real m = kinv * f;
  const int bslot = sslot[e];
  if (bslot >= 0) {
    const size_t so = size_t(p) * sld + bslot, sp = size_t(NPT) * sld;
    const real nrm = stacey[2 * sp + so], wgt = stacey[3 * sp + so],
               rvpi = stacey[4 * sp + so];
    const real vf = real(-1.0) * dt * real(0.5);
    const real factor = wgt * nrm;
    m += real(-1.0) * factor * rvpi * vf;
  }
  M[ibool[t]] += m;
}

template <typename real>
__global__ void k_ac_invert(real *__restrict__ M, long n) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    M[i] = real(1.0) / M[i];
}

__global__ void k_ac_set_step(int *__restrict__ step_ptr, int value) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *step_ptr = value;
}

// ===================================================================
template <typename real> class SolverAc2D final : public Solver2DBase {
public:
  SolverAc2D(std::unique_ptr<HostAssembly2D> a, int graph_steps)
      : asm_(std::move(a)), gsteps_(graph_steps) {
    const HostAssembly2D &A = *asm_;
    AC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    AC_HIP(hipEventCreate(&ev0_));
    AC_HIP(hipEventCreate(&ev1_));
    n_ = (A.nglob + 255) / 256 * 256;
    tiles_ = upload_raw<TileInfo>(A.tiles);
    sh_list_ = upload_raw<int>(A.sh_list);
    {
      // shared-point ids at a fixed stride per tile for the tile gather
      for (const TileInfo &t : A.tiles)
        sh_stride_ = std::max(sh_stride_, t.nsh);
      std::vector<int> pad(size_t(std::max(1, A.ntiles)) * sh_stride_, 0);
      for (int t = 0; t < A.ntiles; ++t)
        for (int k = 0; k < A.tiles[t].nsh; ++k)
          pad[size_t(t) * sh_stride_ + k] = A.sh_list[A.tiles[t].sh_off + k];
      sh_pad_ = upload_raw<int>(pad);
    }
    shp_off_ = upload_raw<int>(A.shp_off);
    shp_pos_ = upload_raw<int>(A.shp_pos);
    lidx_ = upload_raw<uint16_t>(A.lidx);
    ecolor_ = upload_raw<uint8_t>(A.ecolor);
    ibool_ = upload_raw<int>(A.ibool);
    dmask_ = upload_raw<uint32_t>(A.dmask);
    geom_ = upload_real(A.geom);
    geom_e_ = upload_real(A.geom_e);
    kmr_ = upload_real(A.kappa_mu_rho);
    sslot_ = upload_raw<int>(A.stacey_slot);
    src_li_ = upload_raw<int>(A.src_li);
    src_idx_ = upload_raw<int>(A.src_idx);
    src_L_ = upload_real(A.src_L);
    stf_ = upload_real(A.stf);
    trec_li_ = upload_raw<int>(A.trec_li);
    trec_entry_ = upload_raw<int>(A.trec_entry);
    rent_g_ = upload_raw<int>(A.rent_g);
    rent_shared_ = upload_raw<int>(A.rent_shared);
    rent_h_ = upload_real(A.rent_h);
    rent_w_ = alloc<real>(A.rent_w.size());
    copy_to(rent_w_, sizeof(real) == 8
                         ? A.rent_w64
                         : std::vector<double>(A.rent_w.begin(), A.rent_w.end()));
    // 1/rho and 1/(rho vp): FP32 as the host formed them, or formed in double
    // for an FP64 context (an FP64 solve of the same FP32 model)
    std::vector<double> m(A.lam2mu_mu.begin(), A.lam2mu_mu.end());
    std::vector<double> st(A.stacey.begin(), A.stacey.end());
    std::vector<double> me(A.mat_e.begin(), A.mat_e.end());
    if (sizeof(real) == 8) {
      const size_t sp = A.imp64.size() / 2;
      std::copy(A.lam2mu64.begin(), A.lam2mu64.end(), m.begin());
      std::copy(A.imp64.begin(), A.imp64.begin() + sp, st.begin() + 4 * sp);
      std::copy(A.lam2mu64.begin(), A.lam2mu64.begin() + A.nspec_pad,
                me.begin());
    }
    mat_ = alloc<real>(m.size());
    copy_to(mat_, m);
    mat_e_ = alloc<real>(me.size());
    copy_to(mat_e_, me);
    stacey_ = alloc<real>(st.size());
    copy_to(stacey_, st);
    minv_ = alloc<real>(n_);
    part_ = alloc<real>(std::max<size_t>(1, A.sh_list.size()));
    cpl_li_ = upload_raw<int>(A.cpl_li);
    cpl_g_ = upload_raw<int>(A.cpl_idx);
    cpl_w_ = upload_real(A.cpl_w);
    cout_li_ = upload_raw<int>(A.cout_li);
    cout_k_ = upload_raw<int>(A.cout_k);
    sh_iface_ = upload_raw<int>(A.sh_iface);
    {
      // global point of every interface index k: exclusive entries through
      // their tile, shared ones through sh_iface
      std::vector<int32_t> ig(std::max(1, A.n_iface), -1);
      for (const TileInfo &t : A.tiles)
        for (int k = t.cout_off; k < t.cout_off + t.ncout; ++k)
          ig[A.cout_k[k]] = t.ex_off + A.cout_li[k];
      for (size_t j = 0; j < A.sh_iface.size(); ++j)
        if (A.sh_iface[j] >= 0)
          ig[A.sh_iface[j]] = A.nex_total + int32_t(j);
      for (int k = 0; k < A.n_iface; ++k)
        if (ig[k] < 0 || ig[k] >= A.nglob)
          throw SmxError(SMX_ERR_MESH, "interface point without a global point");
      iface_g_ = upload_raw<int>(ig);
    }
    lds_bytes_ = ac_lds_bytes<real>(std::max(1, A.nloc_max));
    for (const void *f : {reinterpret_cast<const void *>(&k_ac_tile<real, AC_STEP>),
                          reinterpret_cast<const void *>(&k_ac_tile<real, AC_FINAL>),
                          reinterpret_cast<const void *>(&k_ac_tile<real, AC_ASSEMBLE>)})
      AC_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 int(lds_bytes_)));
    step_ = alloc<int>(2);
    const int ntypes = int(A.seis_types.size());
    if (ntypes > 4)
      throw SmxError(SMX_ERR_ARG, "at most 4 seismogram types");
    for (int k = 0; k < 4; ++k)
      types_.t[k] = k < ntypes ? A.seis_types[k] : 0;
    seis_ = alloc<real>(std::max<size_t>(
        size_t(std::max(1, A.nsig)) * ntypes * std::max(1, A.nrec) * 2, 2));
    seis_terms_ = alloc<real>(
        std::max<size_t>(size_t(ntypes) * A.nrec * NPT * 2, 2));
    for (int i = 0; i < NGLL; ++i) {
      quad_.w[i] = A.quad.w[i];
      for (int j = 0; j < NGLL; ++j)
        quad_.hp[i][j] = A.quad.hprime[i][j];
    }
    dt_ = A.dt;
    dto2_ = A.dt_over2;
    dt2o2_ = A.dt2_over2;
    field(SMX_FIELD_FORWARD);
    AC_HIP(hipStreamSynchronize(stream_));
  }

  ~SolverAc2D() override {
    if (graph_exec_)
      (void)hipGraphExecDestroy(graph_exec_);
    if (graph_)
      (void)hipGraphDestroy(graph_);
    (void)hipStreamSynchronize(stream_);
    for (void *p : allocs_)
      (void)hipFree(p);
    (void)hipEventDestroy(ev0_);
    (void)hipEventDestroy(ev1_);
    (void)hipStreamDestroy(stream_);
  }

  int precision() const override { return sizeof(real); }
  const HostAssembly2D &assembly() const override { return *asm_; }
  int medium() const override { return SMX_MEDIUM_ACOUSTIC; }
  void on_stream(void *stream, const std::function<void()> &fn) override {
    AcStreamSwap sw(stream_, static_cast<hipStream_t>(stream));
    fn();
  }

  void initialize() override {
    const HostAssembly2D &A = *asm_;
    AC_HIP(hipMemsetAsync(minv_, 0, n_ * sizeof(real), stream_));
    const long n = long(NPT) * A.nspec_pad;
    for (int c = 0; c < A.nelem_colors; ++c) {
      k_ac_mass<real><<<dim3((n + 255) / 256), dim3(256), 0, stream_>>>(
          minv_, ibool_, ecolor_, c, geom_, kmr_, sslot_, stacey_, A.nspec_pad,
          A.nstacey_pad, quad_, real(A.dt));
      AC_HIP(hipGetLastError());
    }
    k_ac_invert<real><<<dim3((A.nglob + 255) / 256), dim3(256), 0, stream_>>>(
        minv_, A.nglob);
    AC_HIP(hipGetLastError());
    // Step graph captured, instantiated and uploaded once, outside every
    // run_forward call; a coupled driver captures both media itself.
    if (!chi_dd_) {
      build_graph(field(SMX_FIELD_FORWARD));
      AC_HIP(hipGraphUpload(graph_exec_, stream_));
    }
    AC_HIP(hipStreamSynchronize(stream_));
    initialized_ = true;
  }

  int predictor(int f) override {
    launch_predictor(field(f));
    return asm_->nglob;
  }

  int corrector(int f) override {
    Field &F = field(f);
    k_ac_corrector<real><<<node_grid(), dim3(AC_NODE_THREADS), 0, stream_>>>(
        F.v, F.a, n_, dto2_);
    AC_HIP(hipGetLastError());
    return asm_->nglob;
  }

  int update_wavefields(int f, int istep) override {
    require_init();
    const HostAssembly2D &A = *asm_;
    if (istep < 0 || istep >= A.nstep)
      throw SmxError(SMX_ERR_ARG, "istep outside [0, nstep)");
    Field &F = field(f);
    launch_tiles<AC_ASSEMBLE>(F, istep, nullptr, nullptr);
    launch_shared<AC_ASSEMBLE>(F, nullptr);
    if (chi_dd_ && A.n_iface > 0) {
      // the new accelerations of the interface points, read by the elastic
      // update of this step (the fused passes store them themselves)
      k_ac_iface<real><<<dim3((A.n_iface + 255) / 256), dim3(256), 0,
                          stream_>>>(F.a, iface_g_, A.n_iface, chi_dd_);
      AC_HIP(hipGetLastError());
    }
    return A.nspec;
  }

  void compute_seismograms(int f, int isig) override {
    const HostAssembly2D &A = *asm_;
    if (isig < 0 || isig >= A.nsig)
      throw SmxError(SMX_ERR_ARG, "seismogram step outside [0, nsig)");
    const int ntypes = int(A.seis_types.size());
    if (A.nrec == 0 || ntypes == 0)
      return;
    Field &F = field(f);
    const int n = A.nrec * ntypes;
    k_ac_seismo<real><<<dim3((n + 127) / 128), dim3(128), 0, stream_>>>(
        F.u, F.v, F.a, rent_g_, rent_w_, rent_h_, A.nrec, types_, ntypes, isig,
        seis_);
    AC_HIP(hipGetLastError());
  }

  void run_forward(int istep0, int nstep) override {
    require_init();
    const HostAssembly2D &A = *asm_;
    if (istep0 < 0 || nstep < 1 || istep0 + nstep > A.nstep)
      throw SmxError(SMX_ERR_ARG, "steps outside [0, nstep)");
    Field &F = field(SMX_FIELD_FORWARD);
    AC_HIP(hipEventRecord(ev0_, stream_));
    k_ac_set_step<<<dim3(1), dim3(64), 0, stream_>>>(step_, istep0);
    AC_HIP(hipGetLastError());
    launch_predictor(F);
    const int body = nstep - 1;
    const int reps = body / gsteps_;
    if (reps > 0) {
      build_graph(F);
      for (int r = 0; r < reps; ++r)
        AC_HIP(hipGraphLaunch(graph_exec_, stream_));
    }
    for (int s = reps * gsteps_; s < body; ++s)
      fused_step<AC_STEP>(F);
    fused_step<AC_FINAL>(F);
    AC_HIP(hipEventRecord(ev1_, stream_));
    last_launches_ = nstep;
    have_stats_ = true;
  }

  void synchronize() override { AC_HIP(hipStreamSynchronize(stream_)); }

  // ---- coupled meshes (the driver enqueues both media on its stream)
  void bind_coupling(void *elastic_u, void *chi_dd) override {
    if (sizeof(real) != 4)
      throw SmxError(SMX_ERR_ARG, "coupled meshes run in FP32");
    ue_ = static_cast<const real *>(elastic_u);
    chi_dd_ = static_cast<real *>(chi_dd);
  }
  void *coupling_field() override { return field(SMX_FIELD_FORWARD).u; }
  void enqueue_begin(void *stream, int istep0) override {
    require_init();
    AcStreamSwap sw(stream_, static_cast<hipStream_t>(stream));
    k_ac_set_step<<<dim3(1), dim3(64), 0, stream_>>>(step_, istep0);
    AC_HIP(hipGetLastError());
    launch_predictor(field(SMX_FIELD_FORWARD));
  }
  void enqueue_step(void *stream, bool final) override {
    AcStreamSwap sw(stream_, static_cast<hipStream_t>(stream));
    Field &F = field(SMX_FIELD_FORWARD);
    if (final)
      fused_step<AC_FINAL>(F);
    else
      fused_step<AC_STEP>(F);
  }

  void get_seismograms(void *host) override {
    const HostAssembly2D &A = *asm_;
    const size_t n = size_t(A.nsig) * A.seis_types.size() * A.nrec * 2;
    AC_HIP(hipMemcpyAsync(host, seis_, n * sizeof(real), hipMemcpyDeviceToHost,
                          stream_));
    AC_HIP(hipStreamSynchronize(stream_));
  }

  void get_field(int f, int which, void *host) override {
    AC_HIP(hipMemcpyAsync(host, pick(f, which), size_t(asm_->nglob) * sizeof(real),
                          hipMemcpyDeviceToHost, stream_));
    AC_HIP(hipStreamSynchronize(stream_));
  }

  void set_field(int f, int which, const void *host) override {
    AC_HIP(hipMemcpyAsync(pick(f, which), host, size_t(asm_->nglob) * sizeof(real),
                          hipMemcpyHostToDevice, stream_));
    AC_HIP(hipStreamSynchronize(stream_));
  }

  void *field_ptr(int f, int which) override { return pick(f, which); }

  void last_run_stats(double *ms, double *stiff_ms, int *launches) override {
    float t = 0;
    if (have_stats_) {
      AC_HIP(hipEventSynchronize(ev1_));
      AC_HIP(hipEventElapsedTime(&t, ev0_, ev1_));
    }
    if (ms)
      *ms = t;
    if (stiff_ms)
      *stiff_ms = last_stiff_ms_;
    if (launches)
      *launches = last_launches_;
  }

  double time_stiffness(int n) override {
    if (n < 1)
      throw SmxError(SMX_ERR_ARG, "n must be positive");
    require_init();
    const HostAssembly2D &A = *asm_;
    Field &F = field(SMX_FIELD_FORWARD);
    if (!scratch_.u) {
      scratch_.u = alloc<real>(n_);
      scratch_.v = alloc<real>(n_);
      scratch_.a = alloc<real>(n_);
      scratch_part_ = alloc<real>(std::max<size_t>(1, A.sh_list.size()));
    }
    AC_HIP(hipMemcpyAsync(scratch_.u, F.u, n_ * sizeof(real),
                          hipMemcpyDeviceToDevice, stream_));
    AC_HIP(hipMemcpyAsync(scratch_.v, F.v, n_ * sizeof(real),
                          hipMemcpyDeviceToDevice, stream_));
    launch_tiles<AC_STEP>(scratch_, 0, nullptr, nullptr, scratch_part_);
    AC_HIP(hipEventRecord(ev0_, stream_));
    for (int i = 0; i < n; ++i)
      launch_tiles<AC_STEP>(scratch_, 0, nullptr, nullptr, scratch_part_);
    AC_HIP(hipEventRecord(ev1_, stream_));
    AC_HIP(hipEventSynchronize(ev1_));
    float t = 0;
    AC_HIP(hipEventElapsedTime(&t, ev0_, ev1_));
    last_stiff_ms_ = double(t) / n;
    have_stats_ = false;
    return last_stiff_ms_;
  }

  // The adjoint and model updates are built for the elastic medium only.
  void set_model(const float *, const float *, const float *) override {
    throw SmxError(SMX_ERR_ARG, "set_model is built for elastic meshes only");
  }
  void run_forward_checkpointed(int) override {
    throw SmxError(SMX_ERR_ARG, "the adjoint is built for elastic meshes only");
  }
  void run_adjoint(const void *) override {
    throw SmxError(SMX_ERR_ARG, "the adjoint is built for elastic meshes only");
  }
  void get_gradient(double *) override {
    throw SmxError(SMX_ERR_ARG, "the adjoint is built for elastic meshes only");
  }
  void last_adjoint_stats(double *f, double *a) override {
    if (f)
      *f = 0;
    if (a)
      *a = 0;
  }
  void geometry_mode(int *mode, long long *mismatches) const override {
    if (mode)
      *mode = 0;
    if (mismatches)
      *mismatches = 0;
  }

private:
  struct Field {
    real *u = nullptr, *v = nullptr, *a = nullptr;
  };

  template <typename T> T *alloc(size_t n) {
    void *p = nullptr;
    AC_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    allocs_.push_back(p);
    AC_HIP(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), stream_));
    return static_cast<T *>(p);
  }
  template <typename T, typename H> T *upload_raw(const std::vector<H> &h) {
    static_assert(sizeof(T) == sizeof(H), "layout mismatch");
    T *d = alloc<T>(h.size());
    if (!h.empty()) {
      AC_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T),
                            hipMemcpyHostToDevice, stream_));
      AC_HIP(hipStreamSynchronize(stream_));
    }
    return d;
  }
  real *upload_real(const std::vector<float> &h) {
    real *d = alloc<real>(h.size());
    if (!h.empty())
      copy_to(d, std::vector<double>(h.begin(), h.end()));
    return d;
  }
  void copy_to(real *d, const std::vector<double> &h) {
    if (h.empty())
      return;
    std::vector<real> tmp(h.begin(), h.end());
    AC_HIP(hipMemcpyAsync(d, tmp.data(), tmp.size() * sizeof(real),
                          hipMemcpyHostToDevice, stream_));
    AC_HIP(hipStreamSynchronize(stream_));
  }

  Field &field(int f) {
    if (f < 0 || f > 2)
      throw SmxError(SMX_ERR_ARG, "unknown wavefield");
    Field &F = fields_[f];
    if (!F.u) {
      F.u = alloc<real>(n_);
      F.v = alloc<real>(n_);
      F.a = alloc<real>(n_);
    }
    return F;
  }
  real *pick(int f, int which) {
    Field &F = field(f);
    switch (which) {
    case SMX_U: return F.u;
    case SMX_V: return F.v;
    case SMX_A: return F.a;
    case SMX_MASS_INVERSE: return minv_;
    }
    throw SmxError(SMX_ERR_ARG, "unknown field component");
  }
  void require_init() {
    if (!initialized_)
      throw SmxError(SMX_ERR_STATE, "smx_initialize has not been called");
  }
  dim3 node_grid() const {
    return dim3(unsigned((n_ + AC_NODE_THREADS - 1) / AC_NODE_THREADS));
  }
  void launch_predictor(Field &F) {
    k_ac_predictor<real><<<node_grid(), dim3(AC_NODE_THREADS), 0, stream_>>>(
        F.u, F.v, F.a, n_, dt_, dto2_, dt2o2_);
    AC_HIP(hipGetLastError());
  }
  template <int MODE>
  void launch_tiles(Field &F, int istep, int *step_ptr, real *terms,
                    real *part = nullptr) {
    const HostAssembly2D &A = *asm_;
    k_ac_tile<real, MODE><<<dim3(A.ntiles), dim3(AC_TILE_THREADS), lds_bytes_,
                            stream_>>>(
        tiles_, A.nloc_max, sh_list_, lidx_, ecolor_, A.nelem_colors, geom_,
        mat_, geom_e_, mat_e_, sslot_, stacey_, dmask_, F.u, F.v, F.a, minv_,
        part ? part : part_, src_li_, src_idx_, src_L_, stf_,
        std::max(1, A.nsrc), trec_li_, trec_entry_, rent_w_, rent_h_, terms,
        A.nrec * NPT, types_, int(A.seis_types.size()), A.nsbs, step_ptr, istep,
        A.nspec_pad, A.nstacey_pad, dt_, dto2_, dt2o2_, quad_, cpl_li_, cpl_g_,
        cpl_w_, ue_, cout_li_, cout_k_, chi_dd_, sh_pad_, sh_stride_);
    AC_HIP(hipGetLastError());
  }
  template <int MODE> void launch_shared(Field &F, int *step_ptr) {
    const HostAssembly2D &A = *asm_;
    const int n = std::max(1, A.nshared);
    k_ac_shared<real, MODE><<<dim3((n + AC_NODE_THREADS - 1) / AC_NODE_THREADS),
                              dim3(AC_NODE_THREADS), 0, stream_>>>(
        F.u, F.v, F.a, minv_, part_, shp_off_, shp_pos_, A.nshared, A.nex_total,
        dt_, dto2_, dt2o2_, step_ptr, chi_dd_ ? sh_iface_ : nullptr, chi_dd_);
    AC_HIP(hipGetLastError());
  }
  template <int MODE> void fused_step(Field &F) {
    const HostAssembly2D &A = *asm_;
    const int ntypes = int(A.seis_types.size());
    const bool seis = A.nrec > 0 && ntypes > 0;
    launch_tiles<MODE>(F, 0, step_, seis ? seis_terms_ : nullptr);
    if (seis) {
      const int n = A.nrec * ntypes;
      k_ac_seis_fused<real><<<dim3((n + AC_SEIS_GROUPS - 1) / AC_SEIS_GROUPS),
                              dim3(32 * AC_SEIS_GROUPS), 0, stream_>>>(
          F.u, F.v, minv_, part_, shp_off_, shp_pos_, rent_g_, rent_shared_,
          rent_w_, rent_h_, seis_terms_, A.nrec, types_, ntypes, A.nsbs, step_,
          dto2_, seis_);
      AC_HIP(hipGetLastError());
    }
    launch_shared<MODE>(F, step_);
  }
  void build_graph(Field &F) {
    if (graph_exec_)
      return;
    AC_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      for (int s = 0; s < gsteps_; ++s)
        fused_step<AC_STEP>(F);
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(stream_, &tmp);
      if (tmp)
        (void)hipGraphDestroy(tmp);
      throw;
    }
    AC_HIP(hipStreamEndCapture(stream_, &graph_));
    AC_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  }

  std::unique_ptr<HostAssembly2D> asm_;
  int gsteps_ = 8;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<void *> allocs_;
  long n_ = 0;
  TileInfo *tiles_ = nullptr;
  int *sh_list_ = nullptr, *shp_off_ = nullptr, *shp_pos_ = nullptr,
      *ibool_ = nullptr, *sslot_ = nullptr, *src_li_ = nullptr,
      *src_idx_ = nullptr, *trec_li_ = nullptr, *trec_entry_ = nullptr,
      *rent_g_ = nullptr, *rent_shared_ = nullptr, *step_ = nullptr;
  uint16_t *lidx_ = nullptr;
  uint8_t *ecolor_ = nullptr;
  uint32_t *dmask_ = nullptr;
  // coupled meshes: coupling entries, the elastic displacement they read,
  // and the interface accelerations this solver writes for the elastic one
  int *cpl_li_ = nullptr, *cpl_g_ = nullptr, *cout_li_ = nullptr,
      *cout_k_ = nullptr, *sh_iface_ = nullptr, *iface_g_ = nullptr;
  real *cpl_w_ = nullptr, *chi_dd_ = nullptr;
  const real *ue_ = nullptr;
  real *geom_ = nullptr, *geom_e_ = nullptr, *mat_ = nullptr, *mat_e_ = nullptr,
       *kmr_ = nullptr, *stacey_ = nullptr, *src_L_ = nullptr, *stf_ = nullptr,
       *rent_h_ = nullptr, *rent_w_ = nullptr, *minv_ = nullptr,
       *part_ = nullptr, *seis_ = nullptr, *seis_terms_ = nullptr,
       *scratch_part_ = nullptr;
  size_t lds_bytes_ = 0;
  int *sh_pad_ = nullptr; // [ntiles][sh_stride_] global ids of shared points
  int sh_stride_ = 1;
  Field fields_[3], scratch_;
  AcQuad<real> quad_;
  AcTypes types_;
  real dt_ = 0, dto2_ = 0, dt2o2_ = 0;
  bool initialized_ = false, have_stats_ = false;
  double last_stiff_ms_ = 0;
  int last_launches_ = 0;
};

} // namespace

std::unique_ptr<Solver2DBase>
make_acoustic_solver2d(std::unique_ptr<HostAssembly2D> a, int precision,
                       int graph_steps) {
  if (precision == 4)
    return std::make_unique<SolverAc2D<float>>(std::move(a), graph_steps);
  if (precision == 8)
    return std::make_unique<SolverAc2D<double>>(std::move(a), graph_steps);
  throw SmxError(SMX_ERR_ARG, "precision must be 4 or 8");
}

} // namespace smx
