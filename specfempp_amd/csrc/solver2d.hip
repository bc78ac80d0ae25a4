// gfx950 kernels and the device-resident solver for the 2D elastic isotropic
// P-SV time step.
//
// Kernels (each cites the reference dispatch it replaces):
//   k_tile      compute_source_interaction + compute_stiffness_interaction
//               (include/kokkos_kernels/impl/compute_source_interaction.tpp:17-104,
//                compute_stiffness_interaction.tpp:140-247) for one tile of
//               elements: gather, GLL gradient, stress, J*sigma*dxi/dx,
//               weighted divergence, Stacey traction, assembly in LDS; then,
//               for the tile's exclusive points, divide_mass_matrix
//               (divide_mass_matrix.tpp:13-65), the Newmark corrector and the
//               next step's predictor (core/specfem/timescheme/newmark.tpp:13-128)
//               in the same pass.  For points shared with other tiles the
//               tile stores its partial sum in a private slot.
//   k_shared    sums the partials of each shared point in fixed order, then
//               the same divide / corrector / predictor.
//   k_seis_fused, k_seismo   compute_seismograms
//               (include/kokkos_kernels/impl/compute_seismogram.tpp:19-171).
//   k_predictor, k_corrector newmark predictor / corrector (operator path).
//   k_mass/k_invert  compute_mass_matrix + invert_mass_matrix
//               (compute_mass_matrix.tpp:23-122, invert_mass_matrix.tpp:13-55).
//
// k_tile launch geometry: one 320-thread workgroup (5 wavefronts) per tile of
// up to 64 elements; wavefront w handles GLL column ix = w of every element,
// lane l handles element l of the tile.  The tile's points live in LDS
// (displacement and accumulator); per-element arrays are [point][slot] so
// each per-point load of a wavefront is one contiguous segment.  Assembly in
// LDS runs colour by colour (elements of a colour share no point) and the
// shared points' partials are summed in tile order, so every sum has a fixed
// order, no atomics are used and runs are bitwise reproducible.  One launch
// covers all tiles: a tile writes only its exclusive points and its private
// partial slots, and reads shared points that only k_shared updates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <unordered_map>
#include <limits>
#include <mutex>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "solver2d.hpp"

namespace smx {

using StreamSwap = StreamSwapT<hipStream_t>;

#define SMX_HIP(call)                                                          \
  do {                                                                         \
    hipError_t err_ = (call);                                                  \
    if (err_ != hipSuccess)                                                    \
      throw SmxError(SMX_ERR_HIP, std::string(#call) + ": " +                  \
                                      hipGetErrorString(err_));                \
  } while (0)

template <typename real> struct QuadArgs {
  real hp[NGLL][NGLL]; // hp[i][j] = l'_j(xi_i)
  real w[NGLL];
  float xi[NGLL]; // GLL nodes (FP32, as the host assembly uses them)
};

// The host assembly's per-point geometry of a 4-node element at (xi, gamma),
// re-formed from its control nodes with the same FP32 operations in the same
// order and no contraction (shape_function_derivatives2d.cpp:7-20,
// core/specfem/jacobian/dim2/jacobian.cpp:25-51).  nd points at node 0 of
// the element's slot in [8][ld] (x0..x3, z0..z3).
__device__ __forceinline__ void geom_from_nodes2(const float *__restrict__ nd,
                                                 int ld, float xi, float ga,
                                                 float &xix, float &xiz,
                                                 float &gax, float &gaz,
                                                 float &jac) {
#pragma clang fp contract(off)
  const float sp = xi + 1, sm = xi - 1, tp = ga + 1, tm = ga - 1;
  const float dxi[4] = {0.25f * tm, -0.25f * tm, 0.25f * tp, -0.25f * tp};
  const float dga[4] = {0.25f * sm, -0.25f * sp, 0.25f * sp, -0.25f * sm};
  float xxi = 0.0f, zxi = 0.0f, xga = 0.0f, zga = 0.0f;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const float cx = nd[size_t(n) * ld], cz = nd[size_t(4 + n) * ld];
    xxi += dxi[n] * cx;
    zxi += dxi[n] * cz;
    xga += dga[n] * cx;
    zga += dga[n] * cz;
  }
  jac = xxi * zga - xga * zxi;
  xix = zga / jac;
  gax = -zxi / jac;
  xiz = -xga / jac;
  gaz = xxi / jac;
}

// Set-up check: 1 where the re-formed geometry of (point, slot) differs from
// the stored FP32 value in any bit (slots of padding elements count as equal).
__global__ void k_geom_check2(const float *__restrict__ nodes,
                              const float *__restrict__ geom,
                              const uint8_t *__restrict__ ecolor, int ld,
                              QuadArgs<float> q,
                              unsigned char *__restrict__ bad) {
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(NPT) * ld)
    return;
  const int p = int(t / ld), slot = int(t % ld);
  if (ecolor[slot] == 0xff) {
    bad[t] = 0;
    return;
  }
  float g[5];
  geom_from_nodes2(nodes + slot, ld, q.xi[p % NGLL], q.xi[p / NGLL], g[0],
                   g[1], g[2], g[3], g[4]);
  const size_t pld = size_t(NPT) * ld;
  bool diff = false;
  for (int c = 0; c < 5; ++c)
    diff = diff || (__float_as_uint(g[c]) != __float_as_uint(geom[c * pld + t]));
  bad[t] = diff ? 1 : 0;
}

template <typename real> struct Vec2;
template <> struct Vec2<float> {
  using T = float2;
};
template <> struct Vec2<double> {
  using T = double2;
};

template <typename real> struct Vec16;
template <> struct Vec16<float> {
  using T = float4;
  static constexpr int W = 4;
};
template <> struct Vec16<double> {
  using T = double2;
  static constexpr int W = 2;
};

constexpr int TILE_THREADS = NGLL * ELEM_TILE;
constexpr int NODE_THREADS = 256;
constexpr int SEIS_GROUPS = 8; // (type, receiver) pairs per k_seis_fused block
// Exclusive points per k_tile thread whose node-pass inputs are requested
// before the LDS assembly (an interior 8x8-element tile has 961 of them).
constexpr int kNodePrefetch = 4;

// What a tile launch does with the assembled points.
enum TileMode : int {
  MODE_STEP = 0,    // divide + corrector + next predictor; store u, v
  MODE_FINAL = 1,   // divide + corrector; store v, a (end of a run)
  MODE_ASSEMBLE = 2, // a := (a + sources + stiffness) / M (operator path)
  MODE_RECORD = 3    // MODE_STEP that also keeps (u~, v~, a') for the adjoint
};

// Forward-state record of step k, slot j = k mod every, in a buffer of
// every+1 slots of three [nglob_pad][2] arrays: the predicted displacement
// and velocity entering step k and the acceleration it produces.
template <typename V2>
__device__ __forceinline__ V2 *record_ptr(void *rec, long stride, int j,
                                          int c) {
  return reinterpret_cast<V2 *>(reinterpret_cast<char *>(rec) +
                                (size_t(j) * 3 + c) * size_t(stride));
}

struct SeisTypes {
  int t[4];
};

// ------------------------------------------------------------------------
// Newmark arithmetic shared by every kernel (same expressions everywhere, so
// the fused and operator paths round identically).
template <typename real>
__device__ __forceinline__ void predict(real &u, real &v, real &a, real dt,
                                        real dto2, real dt2o2) {
  u += dt * v + dt2o2 * a;
  v += dto2 * a;
  a = 0;
}

template <typename real>
__device__ __forceinline__ void divide_correct(real &a, real &v, real minv,
                                               real dto2) {
  a = a * minv;
  v += dto2 * a;
}

// Sum of the tile partials of shared point j, in tile order.
template <typename V2>
__device__ __forceinline__ V2 shared_sum(const V2 *__restrict__ part,
                                         const int *__restrict__ shp_off,
                                         const int *__restrict__ shp_pos,
                                         int j) {
  V2 acc;
  acc.x = 0;
  acc.y = 0;
  for (int k = shp_off[j]; k < shp_off[j + 1]; ++k) {
    const V2 p = part[shp_pos[k]];
    acc.x += p.x;
    acc.y += p.y;
  }
  return acc;
}

// ------------------------------------------------------------------------
// Dynamic LDS of k_tile: s_u[nloc_max], s_a[nloc_max] (point displacement and
// accumulator), then a region reused as the staged element displacement, the
// xi stress integrand and the corrected velocity of the tile's points.
template <typename real>
__host__ __device__ constexpr size_t tile_lds_bytes(int nloc_max) {
  return 2 * size_t(nloc_max) * 2 * sizeof(real) +
         (size_t(2 * NPT * ELEM_TILE) > size_t(nloc_max) * 2
              ? size_t(2 * NPT * ELEM_TILE)
              : size_t(nloc_max) * 2) *
             sizeof(real);
}

// The FP32 time-step variant is held to 7 waves per SIMD (at most 73
// VGPRs) so that register use does not cut below the 5 tiles per CU the LDS
// footprint allows.
// GN: the tiles not flagged element-constant re-form their geometry from
// the control nodes (instead of reading it per point).
// WPE: if nonzero, the minimum waves per SIMD the compiler must allow (an
// occupancy variant of the FP32 time-step kernel timed by the autotuner);
// the arithmetic is the same, only the register allocation changes.
template <typename real, int MODE, bool GN = false, int WPE = 0>
__global__ void __launch_bounds__(TILE_THREADS) __attribute__((amdgpu_waves_per_eu(
    WPE ? WPE : (sizeof(real) == 4 && MODE == MODE_STEP ? (GN ? 4 : 7) : 1),
    8))) k_tile(const TileInfo *__restrict__ tiles, int nloc_max,
           const int *__restrict__ sh_list, const uint16_t *__restrict__ lidx,
           const uint8_t *__restrict__ ecolor, int ncolors,
           const real *__restrict__ geom, const real *__restrict__ mat,
           const real *__restrict__ geom_e, const real *__restrict__ mat_e,
           const float *__restrict__ nodes,
           const int *__restrict__ sslot, const real *__restrict__ stacey,
           real *__restrict__ U, real *__restrict__ V, real *__restrict__ A,
           const real *__restrict__ Minv, real *__restrict__ part,
           const int *__restrict__ src_li, const int *__restrict__ src_idx,
           const real *__restrict__ src_L, const real *__restrict__ stf,
           int nsrc_total, const int *__restrict__ trec_li,
           const int *__restrict__ trec_entry, const real *__restrict__ rent_h,
           real *__restrict__ seis_terms, int nentries, SeisTypes types,
           int ntypes, int nsbs, int *__restrict__ step_ptr, int istep,
           int ld, int sld, real dt, real dto2, real dt2o2, QuadArgs<real> q,
           void *__restrict__ rec, long rec_stride, int every,
           const int *__restrict__ cpl_li, const int *__restrict__ cpl_k,
           const real *__restrict__ cpl_w, const real *__restrict__ chi_dd,
           const int *__restrict__ sh_pad, int sh_stride) {
  using V2 = typename Vec2<real>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V2 *s_u = reinterpret_cast<V2 *>(lds_raw);
  V2 *s_a = s_u + nloc_max;
  real *xr = reinterpret_cast<real *>(s_a + nloc_max);
  // xr as [2][NPT][ELEM_TILE]: staged displacement, then xi integrand
  auto X = [&](int c, int p, int l) -> real & {
    return xr[(c * NPT + p) * ELEM_TILE + l];
  };
  V2 *s_v = reinterpret_cast<V2 *>(xr);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int ix = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tile = blockIdx.x;
  const TileInfo T = tiles[tile];
  const int nloc = T.nex + T.nsh;
  const int slot = tile * ELEM_TILE + lane;
  const bool valid = lane < T.ecount;
  const size_t pld = size_t(NPT) * ld;
  V2 *__restrict__ U2 = reinterpret_cast<V2 *>(U);
  V2 *__restrict__ V2p = reinterpret_cast<V2 *>(V);
  V2 *__restrict__ A2 = reinterpret_cast<V2 *>(A);
  V2 *__restrict__ P2 = reinterpret_cast<V2 *>(part);
  const V2 *__restrict__ M2 = reinterpret_cast<const V2 *>(Minv);
  const int step = step_ptr ? *step_ptr : istep;
  V2 *rec_a = nullptr, *rec_u = nullptr, *rec_v = nullptr;
  if (MODE == MODE_RECORD) {
    const int j = step % every;
    rec_a = record_ptr<V2>(rec, rec_stride, j, 2);
    rec_u = record_ptr<V2>(rec, rec_stride, j + 1, 0);
    rec_v = record_ptr<V2>(rec, rec_stride, j + 1, 1);
    if (tile == 0 && tid == 0)
      step_ptr[1] = step; // k_shared reads the step from here
  }

  // tile-local ids of this thread's column (iz = 0..4, ix)
  int li[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz)
    li[iz] = valid ? int(lidx[size_t(iz * NGLL + ix) * ld + slot]) : 0;

  // ---- tile points: displacement in, accumulator cleared.  The global ids
  // of the tile's shared points sit at a fixed stride per tile (sh_pad,
  // padded with id 0), so their displacement loads wait on one load only,
  // not on the tile record as well.
  for (int s = tid; s < sh_stride; s += TILE_THREADS) {
    const V2 u = U2[sh_pad[size_t(tile) * sh_stride + s]];
    if (s < T.nsh) {
      V2 z;
      z.x = 0;
      z.y = 0;
      s_u[T.nex + s] = u;
      s_a[T.nex + s] = z;
    }
  }
  for (int i = tid; i < T.nex; i += TILE_THREADS) {
    const int g = T.ex_off + i;
    s_u[i] = U2[g];
    V2 z;
    z.x = 0;
    z.y = 0;
    if (MODE == MODE_ASSEMBLE)
      z = A2[g];
    s_a[i] = z;
  }
  __syncthreads();
  // point sources first, in list order (one thread: fixed order)
  if (tid == 0)
    for (int k = T.src_off; k < T.src_off + T.nsrc; ++k) {
      const int l = src_li[k], s = src_idx[k];
      V2 acc = s_a[l];
      acc.x += stf[(size_t(step) * nsrc_total + s) * 2 + 0] * src_L[2 * k];
      acc.y += stf[(size_t(step) * nsrc_total + s) * 2 + 1] * src_L[2 * k + 1];
      s_a[l] = acc;
    }
  // coupling with the fluid of a coupled mesh: a += (w n) chi_dd, the
  // acoustic acceleration of this step (include/medium/dim2/coupling_terms/
  // elastic_acoustic.hpp:13-26), entries in tile order
  if (tid == 0)
    for (int k = T.cpl_off; k < T.cpl_off + T.ncpl; ++k) {
      const int l = cpl_li[k];
      const real chi = chi_dd[cpl_k[k]];
      V2 acc = s_a[l];
      acc.x += cpl_w[3 * k] * chi;
      acc.y += cpl_w[3 * k + 1] * chi;
      s_a[l] = acc;
    }

  // ---- element displacement (this thread: column ix, rows 0..4)
  real ux[NGLL], uz[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    const int p = iz * NGLL + ix;
    const V2 uu = s_u[li[iz]];
    ux[iz] = uu.x;
    uz[iz] = uu.y;
    X(0, p, lane) = ux[iz];
    X(1, p, lane) = uz[iz];
  }
  __syncthreads();

  // ---- gradient (include/algorithms/gradient.hpp:40-82)
  real dxi_x[NGLL], dxi_z[NGLL], dga_x[NGLL], dga_z[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    real a0 = 0, a1 = 0, b0 = 0, b1 = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      a0 += q.hp[ix][l] * X(0, iz * NGLL + l, lane);
      b0 += q.hp[iz][l] * ux[l];
      a1 += q.hp[ix][l] * X(1, iz * NGLL + l, lane);
      b1 += q.hp[iz][l] * uz[l];
    }
    dxi_x[iz] = a0;
    dxi_z[iz] = a1;
    dga_x[iz] = b0;
    dga_z[iz] = b1;
  }
  __syncthreads(); // the staging area is reused for the xi integrand

  // ---- stress (stress.hpp:38-78) and F = J sigma dxi/dx (point/stress.hpp:166-183)
  real f1x[NGLL], f1z[NGLL];
  // element-constant geometry / moduli of flagged tiles come from the
  // per-element arrays (uniform choice of base and stride)
  const bool g_el = (T.flags & TILE_GEOM_ELEM) != 0;
  const bool m_el = (T.flags & TILE_MAT_ELEM) != 0;
  const size_t gst = g_el ? size_t(ld) : pld, mst = m_el ? size_t(ld) : pld;
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    const int p = iz * NGLL + ix;
    const size_t o = size_t(p) * ld + slot;
    const real *mp = m_el ? mat_e + slot : mat + o;
    real xix, xiz, gax, gaz, jac;
    if (GN && !g_el) {
      float a0, a1, a2, a3, a4;
      geom_from_nodes2(nodes + slot, ld, q.xi[ix], q.xi[iz], a0, a1, a2, a3,
                       a4);
      xix = a0, xiz = a1, gax = a2, gaz = a3, jac = a4;
    } else {
      const real *gp = g_el ? geom_e + slot : geom + o;
      xix = gp[0], xiz = gp[gst], gax = gp[2 * gst], gaz = gp[3 * gst];
      jac = gp[4 * gst];
    }
    const real l2m = mp[0], mu = mp[mst];
    const real lam = l2m - real(2.0) * mu;
    const real du_xx = xix * dxi_x[iz] + gax * dga_x[iz];
    const real du_xz = xiz * dxi_x[iz] + gaz * dga_x[iz];
    const real du_zx = xix * dxi_z[iz] + gax * dga_z[iz];
    const real du_zz = xiz * dxi_z[iz] + gaz * dga_z[iz];
    const real sxx = l2m * du_xx + lam * du_zz;
    const real szz = l2m * du_zz + lam * du_xx;
    const real sxz = mu * (du_xz + du_zx);
    X(0, p, lane) = jac * (sxx * xix + sxz * xiz);
    X(1, p, lane) = jac * (sxz * xix + szz * xiz);
    f1x[iz] = jac * (sxx * gax + sxz * gaz);
    f1z[iz] = jac * (sxz * gax + szz * gaz);
  }
  __syncthreads();

  // ---- divergence (divergence.hpp:40-74), negation, Stacey (stacey.tpp:46-80)
  const int bslot = valid ? sslot[slot] : -1;
  real rx[NGLL], rz[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    real t1x = 0, t1z = 0, t2x = 0, t2z = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      t1x += X(0, iz * NGLL + l, lane) * q.hp[l][ix] * q.w[l];
      t1z += X(1, iz * NGLL + l, lane) * q.hp[l][ix] * q.w[l];
      t2x += f1x[l] * q.hp[l][iz] * q.w[l];
      t2z += f1z[l] * q.hp[l][iz] * q.w[l];
    }
    real ax = q.w[iz] * t1x + q.w[ix] * t2x;
    real az = q.w[iz] * t1z + q.w[ix] * t2z;
    ax *= real(-1.0);
    az *= real(-1.0);
    if (bslot >= 0) {
      const int p = iz * NGLL + ix;
      const size_t so = size_t(p) * sld + bslot;
      const size_t sp = size_t(NPT) * sld;
      const real n0 = stacey[so], n1 = stacey[sp + so],
                 nrm = stacey[2 * sp + so], wgt = stacey[3 * sp + so],
                 rvp = stacey[4 * sp + so], rvs = stacey[5 * sp + so];
      const int g = li[iz] < T.nex ? T.ex_off + li[iz]
                                   : sh_list[T.sh_off + li[iz] - T.nex];
      const V2 vv = V2p[g];
      const real vx = vv.x, vz = vv.y;
      const real vn = vx * n0 + vz * n1;
      const real fx = ((vn * n0 / (nrm * nrm)) * (rvp - rvs)) + vx * rvs;
      const real fz = ((vn * n1 / (nrm * nrm)) * (rvp - rvs)) + vz * rvs;
      ax += real(-1.0) * fx * nrm * wgt;
      az += real(-1.0) * fz * nrm * wgt;
    }
    rx[iz] = ax;
    rz[iz] = az;
  }
  // mass inverse and velocity of this thread's first exclusive points,
  // requested now so their latency overlaps the colour-ordered assembly
  // (only this thread touches these points until it finishes them)
  V2 pre_m[kNodePrefetch], pre_v[kNodePrefetch];
#pragma unroll
  for (int k = 0; k < kNodePrefetch; ++k) {
    const int i = tid + k * TILE_THREADS;
    if (i < T.nex) {
      pre_m[k] = M2[T.ex_off + i];
      if (MODE != MODE_ASSEMBLE)
        pre_v[k] = V2p[T.ex_off + i];
    }
  }

  // ---- assembly into the tile accumulator, colour by colour
  const int ecol = valid ? int(ecolor[slot]) : -1;
  for (int c = 0; c < ncolors; ++c) {
    if (ecol == c) {
#pragma unroll
      for (int iz = 0; iz < NGLL; ++iz) {
        V2 acc = s_a[li[iz]];
        acc.x += rx[iz];
        acc.y += rz[iz];
        s_a[li[iz]] = acc;
      }
    }
    __syncthreads();
  }

  // ---- exclusive points: finish the step; shared points: store partials
  const bool sample = seis_terms != nullptr && (step % nsbs) == 0;
  for (int i = tid, k = 0; i < T.nex; i += TILE_THREADS, ++k) {
    const int g = T.ex_off + i;
    const V2 acc = s_a[i];
    V2 m, v;
    if (k < kNodePrefetch) {
      m = pre_m[k];
      v = pre_v[k];
    } else {
      m = M2[g];
      if (MODE != MODE_ASSEMBLE)
        v = V2p[g];
    }
    if (MODE == MODE_ASSEMBLE) {
      V2 a;
      a.x = acc.x * m.x;
      a.y = acc.y * m.y;
      A2[g] = a;
      continue;
    }
    V2 a = acc;
    divide_correct(a.x, v.x, m.x, dto2);
    divide_correct(a.y, v.y, m.y, dto2);
    if (sample) {
      s_a[i] = a;
      s_v[i] = v;
    }
    if (MODE == MODE_STEP || MODE == MODE_RECORD) {
      if (MODE == MODE_RECORD)
        rec_a[g] = a;
      V2 u = s_u[i];
      predict(u.x, v.x, a.x, dt, dto2, dt2o2);
      predict(u.y, v.y, a.y, dt, dto2, dt2o2);
      U2[g] = u;
      V2p[g] = v;
      if (MODE == MODE_RECORD) {
        rec_u[g] = u;
        rec_v[g] = v;
      }
    } else {
      V2p[g] = v;
      A2[g] = a;
    }
  }
  for (int i = T.nex + tid; i < nloc; i += TILE_THREADS)
    P2[T.sh_off + i - T.nex] = s_a[i];

  // ---- receiver terms on exclusive points: h * field at step n
  if (MODE != MODE_ASSEMBLE && sample && T.nrec > 0) {
    __syncthreads();
    for (int k = T.rec_off + tid; k < T.rec_off + T.nrec; k += TILE_THREADS) {
      const int l = trec_li[k], id = trec_entry[k];
      const real h = rent_h[id];
      for (int it = 0; it < ntypes; ++it) {
        if (types.t[it] == 3) // pressure: k_seis_pressure
          continue;
        const V2 f = types.t[it] == 0 ? s_u[l] : (types.t[it] == 1 ? s_v[l] : s_a[l]);
        real *o = seis_terms + (size_t(it) * nentries + id) * 2;
        o[0] = h * f.x;
        o[1] = h * f.y;
      }
    }
  }
}

// Seismogram sums folded into k_shared's first workgroup when every term
// was written by the tiles (displacement and pressure types only: the tiles
// then evaluate the entries on shared points too, from their gathered
// displacement).  on = 0: no such workgroup.
template <typename real> struct SeisSum {
  const real *terms; // [ntypes][nrec * 25][2]
  real *out;         // [nsig][ntypes][nrec][2]
  SeisTypes types;
  int nrec, ntypes, nsbs, on;
};

// s = sum of the 25 terms of each displacement (type, receiver) in point
// order from 0 (as k_seis_fused sums them) on sampled steps; then the step
// counter advances, after every thread has read it.
template <typename real>
__device__ void seis_sum_block(const SeisSum<real> &ss, int *__restrict__ step_ptr) {
  const int step = *step_ptr;
  if (step % ss.nsbs == 0) {
    const int isig = step / ss.nsbs, nent = ss.nrec * NPT;
    for (int q = threadIdx.x; q < ss.nrec * ss.ntypes; q += blockDim.x) {
      const int it = q / ss.nrec, r = q % ss.nrec;
      if (ss.types.t[it] != 0) // pressure: k_seis_pressure
        continue;
      const real *t = ss.terms + (size_t(it) * nent + size_t(r) * NPT) * 2;
      real sx = 0, sz = 0;
#pragma unroll
      for (int k = 0; k < NPT; ++k) {
        sx += t[2 * k];
        sz += t[2 * k + 1];
      }
      const size_t o = ((size_t(isig) * ss.ntypes + it) * ss.nrec + r) * 2;
      ss.out[o] = sx;
      ss.out[o + 1] = sz;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0)
    *step_ptr = step + 1;
}

// ------------------------------------------------------------------------
// Points shared between tiles: sum the tiles' partials in fixed order and
// finish the step as k_tile does for exclusive points.  Thread 0 advances
// the device step counter (fused modes), or the seismogram workgroup does
// when there is one (block 0; the points then start at block 1).
template <typename real, int MODE>
__global__ void __launch_bounds__(NODE_THREADS)
    k_shared(real *__restrict__ U, real *__restrict__ V, real *__restrict__ A,
             const real *__restrict__ Minv, const real *__restrict__ part,
             const int *__restrict__ shp_off, const int *__restrict__ shp_pos,
             int nshared, int nex_total, real dt, real dto2, real dt2o2,
             int *__restrict__ step_ptr, void *__restrict__ rec,
             long rec_stride, int every, SeisSum<real> ss) {
  using V2 = typename Vec2<real>::T;
  int blk = blockIdx.x;
  if (MODE == MODE_STEP && ss.on) {
    if (blk == 0) {
      seis_sum_block(ss, step_ptr);
      return;
    }
    --blk;
  }
  const int j = blk * blockDim.x + threadIdx.x;
  int step = 0;
  if (MODE == MODE_RECORD) {
    step = step_ptr[1];
    if (j == 0)
      step_ptr[0] = step + 1;
  } else if (j == 0 && step_ptr && MODE != MODE_ASSEMBLE &&
             !(MODE == MODE_STEP && ss.on)) {
    *step_ptr += 1;
  }
  if (j >= nshared)
    return;
  V2 *__restrict__ U2 = reinterpret_cast<V2 *>(U);
  V2 *__restrict__ V2p = reinterpret_cast<V2 *>(V);
  V2 *__restrict__ A2 = reinterpret_cast<V2 *>(A);
  const V2 *__restrict__ M2 = reinterpret_cast<const V2 *>(Minv);
  const int g = nex_total + j;
  // the point's own 1/M, v (or a) and u first: they do not wait on the
  // partials
  const V2 m = M2[g];
  V2 v, u, a0;
  v.x = v.y = u.x = u.y = a0.x = a0.y = 0;
  if (MODE == MODE_ASSEMBLE)
    a0 = A2[g];
  else
    v = V2p[g];
  if (MODE == MODE_STEP || MODE == MODE_RECORD)
    u = U2[g];
  V2 acc = shared_sum(reinterpret_cast<const V2 *>(part), shp_off, shp_pos, j);
  if (MODE == MODE_ASSEMBLE) {
    V2 a;
    a.x = (a0.x + acc.x) * m.x;
    a.y = (a0.y + acc.y) * m.y;
    A2[g] = a;
    return;
  }
  divide_correct(acc.x, v.x, m.x, dto2);
  divide_correct(acc.y, v.y, m.y, dto2);
  if (MODE == MODE_STEP || MODE == MODE_RECORD) {
    const int jr = MODE == MODE_RECORD ? step % every : 0;
    if (MODE == MODE_RECORD)
      record_ptr<V2>(rec, rec_stride, jr, 2)[g] = acc;
    predict(u.x, v.x, acc.x, dt, dto2, dt2o2);
    predict(u.y, v.y, acc.y, dt, dto2, dt2o2);
    U2[g] = u;
    V2p[g] = v;
    if (MODE == MODE_RECORD) {
      record_ptr<V2>(rec, rec_stride, jr + 1, 0)[g] = u;
      record_ptr<V2>(rec, rec_stride, jr + 1, 1)[g] = v;
    }
  } else {
    V2p[g] = v;
    A2[g] = acc;
  }
}

// ------------------------------------------------------------------------
// Fused-path seismograms, between the tile launch and k_shared: entries the
// tiles evaluated (those on exclusive points; all of them when only
// displacement and pressure are recorded, rent_shared = -1) are read from
// the terms; entries on shared points otherwise are evaluated here from the
// tile partials with the same divide / corrector arithmetic.  Sum over the
// 25 points in fixed order.  Runs on every step with velocity or
// acceleration types, else on the last step of a call only (the other steps
// sum in k_shared's first workgroup).
template <typename real>
__global__ void k_seis_fused(const real *__restrict__ U,
                             const real *__restrict__ V,
                             const real *__restrict__ Minv,
                             const real *__restrict__ part,
                             const int *__restrict__ shp_off,
                             const int *__restrict__ shp_pos,
                             const int *__restrict__ rent_g,
                             const int *__restrict__ rent_shared,
                             const real *__restrict__ rent_h,
                             const real *__restrict__ seis_terms, int nrec,
                             SeisTypes types, int ntypes, int nsbs,
                             const int *__restrict__ step_ptr, real dto2,
                             real *__restrict__ out) {
  using V2 = typename Vec2<real>::T;
  // 32 threads per (type, receiver): lane p < 25 evaluates point p, lane 0
  // then sums the 25 terms in point order.
  __shared__ real s_t[SEIS_GROUPS][NPT][2];
  const int grp = threadIdx.x >> 5, p = threadIdx.x & 31;
  const int pair = blockIdx.x * SEIS_GROUPS + grp;
  const int r = pair % nrec, it = pair / nrec;
  // pressure types are written by k_seis_pressure
  const bool active = it < ntypes && p < NPT && types.t[it] != 3;
  // the entry's own data is requested before the step counter is known
  // (all in bounds for every active lane)
  const int nentries = nrec * NPT;
  const int id = r * NPT + p;
  int j = -1, g = 0;
  real h = 0, t0 = 0, t1 = 0;
  if (active) {
    j = rent_shared[id];
    g = rent_g[id];
    h = rent_h[id];
    t0 = seis_terms[(size_t(it) * nentries + id) * 2];
    t1 = seis_terms[(size_t(it) * nentries + id) * 2 + 1];
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
      real fx, fz;
      if (type == 0) {
        fx = U[2 * size_t(g)];
        fz = U[2 * size_t(g) + 1];
      } else {
        real vx = V[2 * size_t(g)], vz = V[2 * size_t(g) + 1];
        const real m0 = Minv[2 * size_t(g)], m1 = Minv[2 * size_t(g) + 1];
        V2 acc = shared_sum(reinterpret_cast<const V2 *>(part), shp_off,
                            shp_pos, j);
        divide_correct(acc.x, vx, m0, dto2);
        divide_correct(acc.y, vz, m1, dto2);
        fx = type == 1 ? vx : acc.x;
        fz = type == 1 ? vz : acc.y;
      }
      tx = h * fx;
      tz = h * fz;
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

// Pressure seismograms of elastic P-SV receivers
// (medium/dim2/elastic/isotropic/wavefield.hpp:57-94): at every point of the
// receiver's element p = -(lambda + 2/3 mu)(du_x/dx + du_z/dz), formed in
// double from `real` lambda, mu and gradient and stored as `real`, then
// s = sum_p h_p p in point order (compute_seismogram.tpp:117-166).  The
// displacement read is the one at the end of the step: on the fused path
// this kernel runs before the step's tile pass, when U holds the predicted
// displacement that the corrector leaves unchanged; on the operator path it
// runs after the corrector.  One 32-lane group per receiver; lane p < 25
// evaluates point p.  Writes component 0 of every pressure type, 0 in
// component 1.
template <typename real>
__global__ void k_seis_pressure(const real *__restrict__ U,
                                const int *__restrict__ rent_g,
                                const real *__restrict__ rent_h,
                                const int *__restrict__ rec_slot,
                                const real *__restrict__ geom,
                                const real *__restrict__ mat, int ld,
                                QuadArgs<real> q, int nrec, SeisTypes types,
                                int ntypes, int nsbs,
                                const int *__restrict__ step_ptr, int isig,
                                real *__restrict__ out) {
  using V2 = typename Vec2<real>::T;
  __shared__ real s_t[SEIS_GROUPS][NPT];
  if (step_ptr) {
    const int step = *step_ptr;
    if (step % nsbs != 0)
      return;
    isig = step / nsbs;
  }
  const int grp = threadIdx.x >> 5, p = threadIdx.x & 31;
  const int r = blockIdx.x * SEIS_GROUPS + grp;
  const bool active = r < nrec && p < NPT;
  const V2 *__restrict__ U2 = reinterpret_cast<const V2 *>(U);
  if (active) {
    const int iz = p / NGLL, ix = p - iz * NGLL;
    const int base = r * NPT;
    real dxi_x = 0, dxi_z = 0, dga_x = 0, dga_z = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      const V2 a = U2[rent_g[base + iz * NGLL + l]]; // point (iz, l)
      const V2 b = U2[rent_g[base + l * NGLL + ix]]; // point (l, ix)
      dxi_x += q.hp[ix][l] * a.x;
      dxi_z += q.hp[ix][l] * a.y;
      dga_x += q.hp[iz][l] * b.x;
      dga_z += q.hp[iz][l] * b.y;
    }
    const size_t pld = size_t(NPT) * ld;
    const size_t o = size_t(p) * ld + rec_slot[r];
    const real xix = geom[o], xiz = geom[pld + o], gax = geom[2 * pld + o],
               gaz = geom[3 * pld + o];
    const real l2m = mat[o], mu = mat[pld + o];
    const real lam = l2m - real(2.0) * mu;
    const real du_xx = xix * dxi_x + gax * dga_x;
    const real du_zz = xiz * dxi_z + gaz * dga_z;
    const double coef = double(lam) + (2.0 / 3.0) * double(mu);
    const real pr = real(-1.0 * (coef * double(du_xx + du_zz)));
    s_t[grp][p] = rent_h[base + p] * pr;
  }
  __syncthreads();
  if (active && p == 0) {
    real s = 0;
    for (int k = 0; k < NPT; ++k)
      s += s_t[grp][k];
    for (int it = 0; it < ntypes; ++it)
      if (types.t[it] == 3) {
        const size_t o = ((size_t(isig) * ntypes + it) * nrec + r) * 2;
        out[o] = s;
        out[o + 1] = 0;
      }
  }
}

// Operator-path seismograms from the finished fields: s = sum_p h_p F[ibool].
template <typename real>
__global__ void k_seismo(const real *__restrict__ U, const real *__restrict__ V,
                         const real *__restrict__ A,
                         const int *__restrict__ rent_g,
                         const real *__restrict__ rent_h, int nrec,
                         SeisTypes types, int ntypes, int isig,
                         real *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = t % nrec, it = t / nrec;
  if (it >= ntypes || types.t[it] == 3) // pressure: k_seis_pressure
    return;
  const int type = types.t[it];
  const real *F = type == 0 ? U : (type == 1 ? V : A);
  real sx = 0, sz = 0;
  for (int p = 0; p < NPT; ++p) {
    const int id = r * NPT + p;
    const int g = rent_g[id];
    const real h = rent_h[id];
    sx += h * F[2 * size_t(g)];
    sz += h * F[2 * size_t(g) + 1];
  }
  const size_t o = ((size_t(isig) * ntypes + it) * nrec + r) * 2;
  out[o] = sx;
  out[o + 1] = sz;
}

// ------------------------------------------------------------------------
// Newmark pieces on [nglob][2] arrays, 16 B per thread per array.
template <typename real>
__global__ void __launch_bounds__(NODE_THREADS)
    k_predictor(real *__restrict__ U, real *__restrict__ V,
                real *__restrict__ A, long n2, real dt, real dto2, real dt2o2) {
  using VT = typename Vec16<real>::T;
  constexpr int W = Vec16<real>::W;
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i * W >= n2)
    return;
  real u[W], v[W], a[W];
  *reinterpret_cast<VT *>(u) = reinterpret_cast<const VT *>(U)[i];
  *reinterpret_cast<VT *>(v) = reinterpret_cast<const VT *>(V)[i];
  *reinterpret_cast<VT *>(a) = reinterpret_cast<const VT *>(A)[i];
#pragma unroll
  for (int k = 0; k < W; ++k)
    predict(u[k], v[k], a[k], dt, dto2, dt2o2);
  reinterpret_cast<VT *>(U)[i] = *reinterpret_cast<VT *>(u);
  reinterpret_cast<VT *>(V)[i] = *reinterpret_cast<VT *>(v);
  reinterpret_cast<VT *>(A)[i] = *reinterpret_cast<VT *>(a);
}

template <typename real>
__global__ void __launch_bounds__(NODE_THREADS)
    k_corrector(real *__restrict__ V, const real *__restrict__ A, long n2,
                real dto2) {
  using VT = typename Vec16<real>::T;
  constexpr int W = Vec16<real>::W;
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i * W >= n2)
    return;
  real v[W], a[W];
  *reinterpret_cast<VT *>(v) = reinterpret_cast<const VT *>(V)[i];
  *reinterpret_cast<VT *>(a) = reinterpret_cast<const VT *>(A)[i];
#pragma unroll
  for (int k = 0; k < W; ++k)
    v[k] += dto2 * a[k];
  reinterpret_cast<VT *>(V)[i] = *reinterpret_cast<VT *>(v);
}

__global__ void k_set_step(int *__restrict__ step_ptr, int value) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *step_ptr = value;
}

// ------------------------------------------------------------------------
// Mass matrix: M[ibool] += rho (w_z w_x J) + Stacey term with the velocity
// set to (-dt/2, -dt/2) (include/boundary_conditions/stacey/stacey.tpp:817-853).
// One launch per element colour (elements of a colour share no point, so
// the scatter is a plain read-modify-write in a fixed order).
template <typename real>
__global__ void k_mass(real *__restrict__ M, const int *__restrict__ ibool,
                       const uint8_t *__restrict__ ecolor, int color,
                       const real *__restrict__ geom,
                       const real *__restrict__ kmr,
                       const int *__restrict__ sslot,
                       const real *__restrict__ stacey, int ld, int sld,
                       QuadArgs<real> q, real dt) {
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(NPT) * ld)
    return;
  const int p = int(t / ld), e = int(t % ld);
  if (int(ecolor[e]) != color)
    return;
  const size_t pld = size_t(NPT) * ld;
  const int iz = p / NGLL, ix = p % NGLL;
  const real rho = kmr[2 * pld + t];
  const real jac = geom[4 * pld + t];
  const real f = q.w[iz] * q.w[ix] * jac;
  real m0 = rho * f, m1 = rho * f;
  const int bslot = sslot[e];
  if (bslot >= 0) {
    const size_t so = size_t(p) * sld + bslot, sp = size_t(NPT) * sld;
    const real n0 = stacey[so], n1 = stacey[sp + so], nrm = stacey[2 * sp + so],
               wgt = stacey[3 * sp + so], rvp = stacey[4 * sp + so],
               rvs = stacey[5 * sp + so];
    const real vf = real(-1.0) * dt * real(0.5);
    const real vn = vf * n0 + vf * n1;
    const real fx = ((vn * n0 / (nrm * nrm)) * (rvp - rvs)) + vf * rvs;
    const real fz = ((vn * n1 / (nrm * nrm)) * (rvp - rvs)) + vf * rvs;
    m0 += real(-1.0) * fx * nrm * wgt;
    m1 += real(-1.0) * fz * nrm * wgt;
  }
  const int g = ibool[t];
  M[2 * size_t(g)] += m0;
  M[2 * size_t(g) + 1] += m1;
}

template <typename real>
__global__ void k_invert(real *__restrict__ M, long n) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    M[i] = real(1.0) / M[i];
}

// ========================================================================
// Discrete adjoint of the fused time step.
//
// Forward step k maps (u, v, a)_k to (u, v, a)_{k+1} through
//   u~ = u + dt v + dt^2/2 a,  v~ = v + dt/2 a,
//   a' = W (f_k - K u~ - B v~),  W = M'^-1 (mass with the Stacey term),
//   u' = u~,  v' = v~ + dt/2 a'.
// Its transpose, with adjoint state (ub, vb, ab) and rb = W (ab + dt/2 vb),
// runs k = nstep-1 .. 0:
//   ub += H^T g_k                       (misfit source of seismogram k)
//   uu = ub - K rb,  vv = vb - B rb
//   ub = uu,  vb = vv + dt uu,  ab = dt/2 vv + dt^2/2 uu
// and the parameter gradient of step k is -rb^T (dK u~ + dB v~ + dM' a').
// The state kept between steps is (ub, vb, rb); ab only enters rb.  rb
// alternates between two buffers by step parity, so that a step reads rb_k
// everywhere while writing rb_{k-1}.
//
// Per-step gradient accumulators (final scaling by the quadrature weight
// and Jacobian, and the chain rule from the Stacey impedances to
// kappa/mu/rho, happen once on the host):
//   Gd[p][slot] += div(rb) div(u~),  Ge[p][slot] += eps(rb):eps(u~),
//   R[g] += rb(g) . a'(g),
//   SP, SS[p][bslot] += -rb^T dB/dP v~ - rb^T dM'/dP a' (and for S), with
//   P = rho vp and S = rho vs at absorbing points.

// Node update shared by k_adj_tile (exclusive points) and k_adj_shared.
template <typename real, typename V2>
__device__ __forceinline__ void adj_finish(const V2 acc, const V2 r, V2 &ub,
                                           V2 &vb, const V2 m, int b,
                                           const real *__restrict__ bsum,
                                           real dt, real dto2, real dt2o2,
                                           V2 &rbn) {
  // acc = -K rb (the element kernel's sign), so ub - K rb = ub + acc
  const real uux = ub.x + acc.x, uuz = ub.y + acc.y;
  real vvx = vb.x, vvz = vb.y;
  if (b >= 0) {
    const real bxx = bsum[3 * b], bxz = bsum[3 * b + 1], bzz = bsum[3 * b + 2];
    vvx = vvx - (bxx * r.x + bxz * r.y);
    vvz = vvz - (bxz * r.x + bzz * r.y);
  }
  ub.x = uux;
  ub.y = uuz;
  vb.x = vvx + dt * uux;
  vb.y = vvz + dt * uuz;
  const real abx = dto2 * vvx + dt2o2 * uux, abz = dto2 * vvz + dt2o2 * uuz;
  rbn.x = m.x * (abx + dto2 * vb.x);
  rbn.y = m.y * (abz + dto2 * vb.y);
}

// ub(node) += sum over its receiver entries of h * g_step (entry order).
template <typename real, typename V2>
__device__ __forceinline__ void adj_inject(V2 &ub, int j,
                                           const int *__restrict__ rin_off,
                                           const int *__restrict__ rin_ent,
                                           const real *__restrict__ rent_h,
                                           const real *__restrict__ gsrc,
                                           int nrec, int step) {
  for (int k = rin_off[j]; k < rin_off[j + 1]; ++k) {
    const int id = rin_ent[k];
    const int r = id / NPT;
    const real h = rent_h[id];
    const real *gs = gsrc + (size_t(step) * nrec + r) * 2;
    ub.x += h * gs[0];
    ub.y += h * gs[1];
  }
}

// Adjoint element pass of step k (= step_ptr[0]) over one tile: K rb
// assembled in LDS as in k_tile, gradient-density accumulation from rb and
// the recorded u~, the Stacey impedance terms, and the node update of the
// tile's exclusive points; shared points get their partial of -K rb.
template <typename real, bool GN = false>
__global__ void __launch_bounds__(TILE_THREADS)
    k_adj_tile(const TileInfo *__restrict__ tiles, int nloc_max,
               const int *__restrict__ sh_list,
               const uint16_t *__restrict__ lidx,
               const uint8_t *__restrict__ ecolor, int ncolors,
               const real *__restrict__ geom, const real *__restrict__ mat,
               const real *__restrict__ geom_e, const real *__restrict__ mat_e,
               const float *__restrict__ nodes,
               const int *__restrict__ sslot, const real *__restrict__ stacey,
               real *__restrict__ RB0, real *__restrict__ RB1,
               real *__restrict__ UB, real *__restrict__ VB,
               const real *__restrict__ Minv, real *__restrict__ part,
               void *__restrict__ rec, long rec_stride, int every,
               const int *__restrict__ bidx, const real *__restrict__ bsum,
               real *__restrict__ Gd, real *__restrict__ Ge,
               real *__restrict__ R, real *__restrict__ SP,
               real *__restrict__ SS, int *__restrict__ step_ptr, int ld,
               int sld, real dt, real dto2, real dt2o2, QuadArgs<real> q) {
  using V2 = typename Vec2<real>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V2 *s_r = reinterpret_cast<V2 *>(lds_raw);  // rb at the tile points
  V2 *s_a = s_r + nloc_max;                     // u~, then the accumulator
  real *xr = reinterpret_cast<real *>(s_a + nloc_max);
  auto X = [&](int c, int p, int l) -> real & {
    return xr[(c * NPT + p) * ELEM_TILE + l];
  };

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int ix = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tile = blockIdx.x;
  const TileInfo T = tiles[tile];
  const int nloc = T.nex + T.nsh;
  const int slot = tile * ELEM_TILE + lane;
  const bool valid = lane < T.ecount;
  const size_t pld = size_t(NPT) * ld;
  const int step = step_ptr[0];
  if (tile == 0 && tid == 0)
    step_ptr[1] = step; // k_adj_shared reads the step from here
  const V2 *__restrict__ RBc =
      reinterpret_cast<const V2 *>((step & 1) ? RB1 : RB0);
  V2 *__restrict__ RBn = reinterpret_cast<V2 *>((step & 1) ? RB0 : RB1);
  V2 *__restrict__ UB2 = reinterpret_cast<V2 *>(UB);
  V2 *__restrict__ VB2 = reinterpret_cast<V2 *>(VB);
  V2 *__restrict__ P2 = reinterpret_cast<V2 *>(part);
  const V2 *__restrict__ M2 = reinterpret_cast<const V2 *>(Minv);
  const int jr = step % every;
  const V2 *__restrict__ Ut = record_ptr<V2>(rec, rec_stride, jr, 0);
  const V2 *__restrict__ Vt = record_ptr<V2>(rec, rec_stride, jr, 1);
  const V2 *__restrict__ At = record_ptr<V2>(rec, rec_stride, jr, 2);

  int li[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz)
    li[iz] = valid ? int(lidx[size_t(iz * NGLL + ix) * ld + slot]) : 0;

  for (int i = tid; i < nloc; i += TILE_THREADS) {
    const int g = i < T.nex ? T.ex_off + i : sh_list[T.sh_off + i - T.nex];
    s_r[i] = RBc[g];
    s_a[i] = Ut[g];
  }
  __syncthreads();

  // ---- reference-coordinate derivatives of u~, then of rb
  real wx[NGLL], wz[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    const V2 w = s_a[li[iz]];
    wx[iz] = w.x;
    wz[iz] = w.y;
    X(0, iz * NGLL + ix, lane) = w.x;
    X(1, iz * NGLL + ix, lane) = w.y;
  }
  __syncthreads();
  real uxi_x[NGLL], uxi_z[NGLL], uga_x[NGLL], uga_z[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    real a0 = 0, a1 = 0, b0 = 0, b1 = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      a0 += q.hp[ix][l] * X(0, iz * NGLL + l, lane);
      b0 += q.hp[iz][l] * wx[l];
      a1 += q.hp[ix][l] * X(1, iz * NGLL + l, lane);
      b1 += q.hp[iz][l] * wz[l];
    }
    uxi_x[iz] = a0;
    uxi_z[iz] = a1;
    uga_x[iz] = b0;
    uga_z[iz] = b1;
  }
  __syncthreads(); // staging reused; s_a holds no u~ reads from here on
  for (int i = tid; i < nloc; i += TILE_THREADS) {
    V2 z;
    z.x = 0;
    z.y = 0;
    s_a[i] = z;
  }
  real ux[NGLL], uz[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    const V2 r = s_r[li[iz]];
    ux[iz] = r.x;
    uz[iz] = r.y;
    X(0, iz * NGLL + ix, lane) = r.x;
    X(1, iz * NGLL + ix, lane) = r.y;
  }
  __syncthreads();
  real dxi_x[NGLL], dxi_z[NGLL], dga_x[NGLL], dga_z[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    real a0 = 0, a1 = 0, b0 = 0, b1 = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      a0 += q.hp[ix][l] * X(0, iz * NGLL + l, lane);
      b0 += q.hp[iz][l] * ux[l];
      a1 += q.hp[ix][l] * X(1, iz * NGLL + l, lane);
      b1 += q.hp[iz][l] * uz[l];
    }
    dxi_x[iz] = a0;
    dxi_z[iz] = a1;
    dga_x[iz] = b0;
    dga_z[iz] = b1;
  }
  __syncthreads();

  // ---- gradient densities, stress of rb and the xi integrand
  real f1x[NGLL], f1z[NGLL];
  // element-constant geometry / moduli of flagged tiles come from the
  // per-element arrays (uniform choice of base and stride)
  const bool g_el = (T.flags & TILE_GEOM_ELEM) != 0;
  const bool m_el = (T.flags & TILE_MAT_ELEM) != 0;
  const size_t gst = g_el ? size_t(ld) : pld, mst = m_el ? size_t(ld) : pld;
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    const int p = iz * NGLL + ix;
    const size_t o = size_t(p) * ld + slot;
    const real *mp = m_el ? mat_e + slot : mat + o;
    real xix, xiz, gax, gaz, jac;
    if (GN && !g_el) {
      float a0, a1, a2, a3, a4;
      geom_from_nodes2(nodes + slot, ld, q.xi[ix], q.xi[iz], a0, a1, a2, a3,
                       a4);
      xix = a0, xiz = a1, gax = a2, gaz = a3, jac = a4;
    } else {
      const real *gp = g_el ? geom_e + slot : geom + o;
      xix = gp[0], xiz = gp[gst], gax = gp[2 * gst], gaz = gp[3 * gst];
      jac = gp[4 * gst];
    }
    const real l2m = mp[0], mu = mp[mst];
    const real lam = l2m - real(2.0) * mu;
    const real du_xx = xix * dxi_x[iz] + gax * dga_x[iz];
    const real du_xz = xiz * dxi_x[iz] + gaz * dga_x[iz];
    const real du_zx = xix * dxi_z[iz] + gax * dga_z[iz];
    const real du_zz = xiz * dxi_z[iz] + gaz * dga_z[iz];
    const real dw_xx = xix * uxi_x[iz] + gax * uga_x[iz];
    const real dw_xz = xiz * uxi_x[iz] + gaz * uga_x[iz];
    const real dw_zx = xix * uxi_z[iz] + gax * uga_z[iz];
    const real dw_zz = xiz * uxi_z[iz] + gaz * uga_z[iz];
    if (valid) {
      const real divr = du_xx + du_zz, divw = dw_xx + dw_zz;
      const real er = real(0.5) * (du_xz + du_zx);
      const real ew = real(0.5) * (dw_xz + dw_zx);
      Gd[o] += divr * divw;
      Ge[o] += du_xx * dw_xx + du_zz * dw_zz + real(2.0) * er * ew;
    }
    const real sxx = l2m * du_xx + lam * du_zz;
    const real szz = l2m * du_zz + lam * du_xx;
    const real sxz = mu * (du_xz + du_zx);
    X(0, p, lane) = jac * (sxx * xix + sxz * xiz);
    X(1, p, lane) = jac * (sxz * xix + szz * xiz);
    f1x[iz] = jac * (sxx * gax + sxz * gaz);
    f1z[iz] = jac * (sxz * gax + szz * gaz);
  }
  __syncthreads();

  // ---- divergence and sign (the element kernel's -K rb)
  real rx[NGLL], rz[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    real t1x = 0, t1z = 0, t2x = 0, t2z = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      t1x += X(0, iz * NGLL + l, lane) * q.hp[l][ix] * q.w[l];
      t1z += X(1, iz * NGLL + l, lane) * q.hp[l][ix] * q.w[l];
      t2x += f1x[l] * q.hp[l][iz] * q.w[l];
      t2z += f1z[l] * q.hp[l][iz] * q.w[l];
    }
    real ax = q.w[iz] * t1x + q.w[ix] * t2x;
    real az = q.w[iz] * t1z + q.w[ix] * t2z;
    rx[iz] = real(-1.0) * ax;
    rz[iz] = real(-1.0) * az;
  }

  // ---- Stacey impedance terms at this column's absorbing points
  const int bslot = valid ? sslot[slot] : -1;
  if (bslot >= 0) {
    const size_t sp = size_t(NPT) * sld;
#pragma unroll
    for (int iz = 0; iz < NGLL; ++iz) {
      const int p = iz * NGLL + ix;
      const size_t so = size_t(p) * sld + bslot;
      const real wgt = stacey[3 * sp + so];
      if (wgt != real(0)) {
        const real n0 = stacey[so], n1 = stacey[sp + so],
                   nrm = stacey[2 * sp + so];
        const int g = li[iz] < T.nex ? T.ex_off + li[iz]
                                     : sh_list[T.sh_off + li[iz] - T.nex];
        const V2 r = s_r[li[iz]], v = Vt[g], a = At[g];
        const real h0 = n0 / nrm, h1 = n1 / nrm, f = nrm * wgt;
        const real rn = r.x * h0 + r.y * h1, vn = v.x * h0 + v.y * h1;
        const real s1 = h0 + h1;
        const real dP = -(f * rn * vn + real(0.5) * dt * f *
                                            (r.x * a.x * h0 * s1 +
                                             r.y * a.y * h1 * s1));
        const real dS =
            -(f * (r.x * v.x + r.y * v.y - rn * vn) +
              real(0.5) * dt * f *
                  (r.x * a.x * (real(1.0) - h0 * s1) +
                   r.y * a.y * (real(1.0) - h1 * s1)));
        SP[so] += dP;
        SS[so] += dS;
      }
    }
  }

  // ---- assembly into the tile accumulator, colour by colour
  const int ecol = valid ? int(ecolor[slot]) : -1;
  for (int c = 0; c < ncolors; ++c) {
    if (ecol == c) {
#pragma unroll
      for (int iz = 0; iz < NGLL; ++iz) {
        V2 acc = s_a[li[iz]];
        acc.x += rx[iz];
        acc.y += rz[iz];
        s_a[li[iz]] = acc;
      }
    }
    __syncthreads();
  }

  // ---- exclusive points: adjoint node update (the loads of up to kAdjFin
  // points per thread requested before the first update); shared points:
  // partials
  constexpr int kAdjFin = 4;
  for (int i0 = tid; i0 < T.nex; i0 += kAdjFin * TILE_THREADS) {
    V2 a[kAdjFin], ub[kAdjFin], vb[kAdjFin], m[kAdjFin];
    real rr[kAdjFin];
    int bi[kAdjFin];
#pragma unroll
    for (int k = 0; k < kAdjFin; ++k) {
      const int i = i0 + k * TILE_THREADS;
      if (i < T.nex) {
        const int g = T.ex_off + i;
        a[k] = At[g];
        rr[k] = R[g];
        ub[k] = UB2[g];
        vb[k] = VB2[g];
        m[k] = M2[g];
        bi[k] = bidx[g];
      }
    }
#pragma unroll
    for (int k = 0; k < kAdjFin; ++k) {
      const int i = i0 + k * TILE_THREADS;
      if (i < T.nex) {
        const int g = T.ex_off + i;
        const V2 r = s_r[i];
        R[g] = rr[k] + (r.x * a[k].x + r.y * a[k].y);
        V2 rbn;
        adj_finish<real>(s_a[i], r, ub[k], vb[k], m[k], bi[k], bsum, dt, dto2,
                         dt2o2, rbn);
        UB2[g] = ub[k];
        VB2[g] = vb[k];
        RBn[g] = rbn;
      }
    }
  }
  for (int i = T.nex + tid; i < nloc; i += TILE_THREADS)
    P2[T.sh_off + i - T.nex] = s_a[i];
}

// Shared points of adjoint step k: partial sums in tile order and the node
// update; then the misfit source of step k-1 at every receiver node (shared
// ones by their updating thread, exclusive ones by the threads after
// nshared).  Thread 0 moves the step counter to k-1.
template <typename real>
__global__ void __launch_bounds__(NODE_THREADS)
    k_adj_shared(real *__restrict__ RB0, real *__restrict__ RB1,
                 real *__restrict__ UB, real *__restrict__ VB,
                 const real *__restrict__ Minv, const real *__restrict__ part,
                 const int *__restrict__ shp_off,
                 const int *__restrict__ shp_pos, int nshared, int nex_total,
                 void *__restrict__ rec, long rec_stride, int every,
                 const int *__restrict__ bidx, const real *__restrict__ bsum,
                 real *__restrict__ R, const int *__restrict__ sh_rin,
                 const int *__restrict__ rin_excl, int nrin_excl,
                 const int *__restrict__ rin_node,
                 const int *__restrict__ rin_off,
                 const int *__restrict__ rin_ent,
                 const real *__restrict__ rent_h,
                 const real *__restrict__ gsrc, int nrec,
                 int *__restrict__ step_ptr, real dt, real dto2, real dt2o2) {
  using V2 = typename Vec2<real>::T;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int step = step_ptr[1];
  if (t == 0)
    step_ptr[0] = step - 1;
  V2 *__restrict__ UB2 = reinterpret_cast<V2 *>(UB);
  if (t < nshared) {
    const int g = nex_total + t;
    const V2 *__restrict__ RBc =
        reinterpret_cast<const V2 *>((step & 1) ? RB1 : RB0);
    V2 *__restrict__ RBn = reinterpret_cast<V2 *>((step & 1) ? RB0 : RB1);
    V2 *__restrict__ VB2 = reinterpret_cast<V2 *>(VB);
    // the node's own state first: it does not wait on the partials
    const V2 r = RBc[g];
    const V2 a = record_ptr<V2>(rec, rec_stride, step % every, 2)[g];
    const real r0 = R[g];
    V2 ub = UB2[g], vb = VB2[g], rbn;
    const V2 m = reinterpret_cast<const V2 *>(Minv)[g];
    const int b = bidx[g];
    const V2 acc = shared_sum(reinterpret_cast<const V2 *>(part), shp_off,
                              shp_pos, t);
    R[g] = r0 + (r.x * a.x + r.y * a.y);
    adj_finish<real>(acc, r, ub, vb, m, b, bsum, dt, dto2, dt2o2, rbn);
    const int j = sh_rin[t];
    if (j >= 0 && step > 0)
      adj_inject<real>(ub, j, rin_off, rin_ent, rent_h, gsrc, nrec, step - 1);
    UB2[g] = ub;
    VB2[g] = vb;
    RBn[g] = rbn;
  } else if (t < nshared + nrin_excl && step > 0) {
    const int j = rin_excl[t - nshared];
    const int g = rin_node[j];
    V2 ub = UB2[g];
    adj_inject<real>(ub, j, rin_off, rin_ent, rent_h, gsrc, nrec, step - 1);
    UB2[g] = ub;
  }
}

// Misfit source of one step at every receiver node (the first adjoint step).
template <typename real>
__global__ void k_adj_inject(real *__restrict__ UB,
                             const int *__restrict__ rin_node,
                             const int *__restrict__ rin_off,
                             const int *__restrict__ rin_ent,
                             const real *__restrict__ rent_h,
                             const real *__restrict__ gsrc, int nrec,
                             int nnodes, int step) {
  using V2 = typename Vec2<real>::T;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnodes)
    return;
  V2 *__restrict__ UB2 = reinterpret_cast<V2 *>(UB);
  const int g = rin_node[j];
  V2 ub = UB2[g];
  adj_inject<real>(ub, j, rin_off, rin_ent, rent_h, gsrc, nrec, step);
  UB2[g] = ub;
}

// ------------------------------------------------------------------------
// A new model read from device memory (smx_set_model_device, FP32): the
// device form of set_model2d's arithmetic, in its FP32 / double order with
// no contraction, so that every derived array is bitwise the host's.
// Smallest element index holding a non-physical value (INT_MAX if none).
__global__ void k_model_check(const float *__restrict__ kappa,
                              const float *__restrict__ mu,
                              const float *__restrict__ rho, long n,
                              int *__restrict__ bad) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && !(rho[i] > 0 && mu[i] >= 0 && kappa[i] > 0))
    atomicMin(bad, int(i / NPT));
}

// The correctly rounded FP32 square root the host's std::sqrt gives: the
// device's v_sqrt_f32 is within 1 ulp only, while the double square root is
// correctly rounded, and rounding it to float is then exact (53 >= 2*24+2
// bits make the double rounding innocuous).
__device__ __forceinline__ float sqrt_rn(float x) {
  return static_cast<float>(sqrt(static_cast<double>(x)));
}

// (kappa, mu, rho) and (lambda+2mu, mu) at every GLL point, slot layout.
__global__ void k_model_in(const float *__restrict__ kappa,
                           const float *__restrict__ mu,
                           const float *__restrict__ rho,
                           const int *__restrict__ elem_slot, int nspec,
                           int ld, float *__restrict__ kmr,
                           float *__restrict__ mat) {
#pragma clang fp contract(off)
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(nspec) * NPT)
    return;
  const int e = int(t / NPT), p = int(t % NPT);
  const size_t pld = size_t(NPT) * ld, o = size_t(p) * ld + elem_slot[e];
  const float k = kappa[t], m = mu[t], r = rho[t];
  mat[o] = k + static_cast<float>(4.0 / 3.0) * m;
  mat[pld + o] = m;
  kmr[o] = k;
  kmr[pld + o] = m;
  kmr[2 * pld + o] = r;
}

// Stacey impedances rho vp, rho vs of the absorbing GLL points
// (elastic/isotropic/properties.hpp:77-87).
__global__ void k_model_stacey(const int *__restrict__ bpt_p,
                               const int *__restrict__ bpt_bslot,
                               const int *__restrict__ bpt_slot, int nbpt,
                               int ld, int sld, const float *__restrict__ kmr,
                               float *__restrict__ stacey) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbpt)
    return;
  const int p = bpt_p[k], bs = bpt_bslot[k];
  const size_t pld = size_t(NPT) * ld, sp = size_t(NPT) * sld;
  const size_t o = size_t(p) * ld + bpt_slot[k];
  const float kk = kmr[o], mm = kmr[pld + o], rr = kmr[2 * pld + o];
  const float l2m = kk + static_cast<float>(4.0 / 3.0) * mm;
  stacey[4 * sp + size_t(p) * sld + bs] = sqrt_rn(l2m / rr) * rr;
  stacey[5 * sp + size_t(p) * sld + bs] = sqrt_rn(rr * mm);
}

// Boundary matrices of the absorbing nodes, summed in double over each
// node's points in point-list order, stored in FP32.
__global__ void k_model_bsum(const int *__restrict__ boff,
                             const int *__restrict__ bk,
                             const int *__restrict__ bpt_p,
                             const int *__restrict__ bpt_bslot, int nbnode,
                             int sld, const float *__restrict__ stacey,
                             float *__restrict__ bsum) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nbnode)
    return;
  const size_t sp = size_t(NPT) * sld;
  double a0 = 0, a1 = 0, a2 = 0;
  for (int q = boff[j]; q < boff[j + 1]; ++q) {
    const int k = bk[q];
    const size_t so = size_t(bpt_p[k]) * sld + bpt_bslot[k];
    const double n0 = stacey[so], n1 = stacey[sp + so],
                 nrm = stacey[2 * sp + so], w = stacey[3 * sp + so];
    const double h0 = n0 / nrm, h1 = n1 / nrm, f = nrm * w;
    const double P = stacey[4 * sp + so], S = stacey[5 * sp + so];
    a0 += f * ((P - S) * h0 * h0 + S);
    a1 += f * ((P - S) * h0 * h1);
    a2 += f * ((P - S) * h1 * h1 + S);
  }
  bsum[3 * j + 0] = static_cast<float>(a0);
  bsum[3 * j + 1] = static_cast<float>(a1);
  bsum[3 * j + 2] = static_cast<float>(a2);
}

// Per-element copies of the moduli and the tiles' element-constant flag
// (refresh_compact): one 64-lane block per tile, lane j = element j.
__global__ void __launch_bounds__(ELEM_TILE)
    k_model_compact(TileInfo *__restrict__ tiles,
                    const float *__restrict__ mat,
                    float *__restrict__ mat_e, int ld) {
  const int t = blockIdx.x, j = threadIdx.x;
  const int ecount = tiles[t].ecount;
  const int sl = t * ELEM_TILE + j;
  const size_t pld = size_t(NPT) * ld;
  bool same = true;
  if (j < ecount) {
    for (int c = 0; c < 2; ++c)
      for (int p = 1; p < NPT; ++p)
        same = same && mat[c * pld + size_t(p) * ld + sl] == mat[c * pld + sl];
    for (int c = 0; c < 2; ++c)
      mat_e[c * size_t(ld) + sl] = mat[c * pld + sl];
  }
  const bool all = __all(same) != 0;
  if (j == 0)
    tiles[t].flags =
        (tiles[t].flags & ~TILE_MAT_ELEM) | (all ? TILE_MAT_ELEM : 0);
}

// The gradient in mesh element order, as get_gradient forms it on the host
// (w_iz w_ix J and the chain rule in double, plus the Stacey impedance
// terms at absorbing points).
template <typename real>
__global__ void k_gradient_out(const int *__restrict__ elem_slot,
                               const int *__restrict__ sslot,
                               const real *__restrict__ geom, QuadArgs<real> q,
                               const int *__restrict__ ibool,
                               const real *__restrict__ kmr,
                               const real *__restrict__ stacey,
                               const real *__restrict__ gd,
                               const real *__restrict__ ge,
                               const real *__restrict__ racc,
                               const real *__restrict__ sP,
                               const real *__restrict__ sS, int nspec, int ld,
                               int sld, double *__restrict__ out) {
#pragma clang fp contract(off)
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(nspec) * NPT)
    return;
  const int e = int(t / NPT), p = int(t % NPT);
  const int slot = elem_slot[e], bs = sslot[slot];
  const size_t pld = size_t(NPT) * ld, sp = size_t(NPT) * sld;
  const size_t o = size_t(p) * ld + slot;
  const int iz = p / NGLL, ix = p % NGLL;
  const double wj = double(q.w[iz]) * q.w[ix] * geom[4 * pld + o];
  const double Gd = gd[o], Ge = ge[o];
  double gk = -wj * Gd;
  double gm = -wj * (2.0 * Ge - 2.0 / 3.0 * Gd);
  double gr = -wj * double(racc[ibool[o]]);
  if (bs >= 0 && stacey[3 * sp + size_t(p) * sld + bs] != real(0)) {
    const double k = kmr[o], m = kmr[pld + o], rh = kmr[2 * pld + o];
    const double l2m = k + 4.0 / 3.0 * m;
    const double P = sqrt(rh * l2m), S = sqrt(rh * m);
    const double dP = sP[size_t(p) * sld + bs], dS = sS[size_t(p) * sld + bs];
    gk += dP * rh / (2.0 * P);
    gm += dP * (4.0 / 3.0) * rh / (2.0 * P);
    gr += dP * l2m / (2.0 * P);
    if (S > 0) {
      gm += dS * rh / (2.0 * S);
      gr += dS * m / (2.0 * S);
    }
  }
  const size_t nloc = size_t(nspec) * NPT;
  out[t] = gk;
  out[nloc + t] = gm;
  out[2 * nloc + t] = gr;
}

// ========================================================================
// Shared-point partials of smx_update_wavefields_noalloc: a fixed device
// array (one copy per GPU, held until the process ends), so that entry point
// calls no allocator.  Its address is looked up once per device, with that
// device current; host code uses only that address, never the array's name.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

void *noalloc_scratch(int device) {
  static std::mutex mu;
  static void *ptrs[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices)
    throw SmxError(SMX_ERR_ARG, "device id outside the supported range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[device]) {
    int cur = 0;
    SMX_HIP(hipGetDevice(&cur));
    if (cur != device)
      SMX_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device)
      SMX_HIP(hipSetDevice(cur));
    if (err != hipSuccess || !p)
      throw SmxError(SMX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(err));
    ptrs[device] = p;
  }
  return ptrs[device];
}

template <typename real> class Solver2D final : public Solver2DBase {
public:
  Solver2D(std::unique_ptr<HostAssembly2D> a, int graph_steps,
           bool nodes_geometry, int waves_per_eu)
      : asm_(std::move(a)) {
    gsteps_ = graph_steps;
    wpe_ = waves_per_eu;
    const HostAssembly2D &A = *asm_;
    SMX_HIP(hipGetDevice(&device_));
    SMX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    SMX_HIP(hipEventCreate(&ev0_));
    SMX_HIP(hipEventCreate(&ev1_));
    nglob_pad_ = (A.nglob + 255) / 256 * 256;
    n2_ = 2L * nglob_pad_;
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
    geom_ = upload_real(A.geom);
    mat_ = upload_real(A.lam2mu_mu);
    geom_e_ = upload_real(A.geom_e);
    mat_e_ = alloc<real>(A.mat_e.size());
    kmr_ = upload_real(A.kappa_mu_rho);
    sslot_ = upload_raw<int>(A.stacey_slot);
    stacey_ = upload_real(A.stacey);
    src_li_ = upload_raw<int>(A.src_li);
    src_idx_ = upload_raw<int>(A.src_idx);
    src_L_ = upload_real(A.src_L);
    stf_ = upload_real(A.stf);
    trec_li_ = upload_raw<int>(A.trec_li);
    trec_entry_ = upload_raw<int>(A.trec_entry);
    rent_g_ = upload_raw<int>(A.rent_g);
    rent_shared_ = upload_raw<int>(A.rent_shared);
    rent_h_ = upload_real(A.rent_h);
    rec_slot_ = upload_raw<int>(A.rec_ispec);
    minv_ = alloc<real>(n2_);
    part_ = alloc<real>(2 * std::max<size_t>(1, A.sh_list.size()));
    cpl_li_ = upload_raw<int>(A.cpl_li);
    cpl_k_ = upload_raw<int>(A.cpl_idx);
    cpl_w_ = upload_real(A.cpl_w);
    lds_bytes_ = tile_lds_bytes<real>(std::max(1, A.nloc_max));
    SMX_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_tile<real, MODE_STEP>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes_)));
    SMX_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_tile<real, MODE_FINAL>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes_)));
    SMX_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_tile<real, MODE_ASSEMBLE>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes_)));
    SMX_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_tile<real, MODE_RECORD>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes_)));
    SMX_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_adj_tile<real>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes_)));
    for (const void *f : {reinterpret_cast<const void *>(&k_tile<real, MODE_STEP, true>),
                          reinterpret_cast<const void *>(&k_tile<real, MODE_FINAL, true>),
                          reinterpret_cast<const void *>(&k_tile<real, MODE_ASSEMBLE, true>),
                          reinterpret_cast<const void *>(&k_tile<real, MODE_RECORD, true>),
                          reinterpret_cast<const void *>(&k_adj_tile<real, true>)})
      SMX_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(lds_bytes_)));
    if constexpr (std::is_same<real, float>::value)
      for (const void *f : {reinterpret_cast<const void *>(&k_tile<real, MODE_STEP, false, 6>),
                            reinterpret_cast<const void *>(&k_tile<real, MODE_STEP, false, 8>)})
        SMX_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds_bytes_)));
    // [0] step of the forward / recompute, [1] its copy for k_shared,
    // [2] step of the adjoint, [3] its copy for k_adj_shared
    step_ = alloc<int>(4);
    const int ntypes = int(A.seis_types.size());
    const size_t nseis =
        size_t(std::max(1, A.nsig)) * ntypes * std::max(1, A.nrec) * 2;
    seis_ = alloc<real>(std::max<size_t>(nseis, 2));
    seis_terms_ = alloc<real>(
        std::max<size_t>(size_t(ntypes) * A.nrec * NPT * 2, 2));
    for (int i = 0; i < NGLL; ++i) {
      quad_.w[i] = A.quad.w[i];
      quad_.xi[i] = A.quad.xi[i];
      for (int j = 0; j < NGLL; ++j)
        quad_.hp[i][j] = A.quad.hprime[i][j];
    }
    if (ntypes > 4)
      throw SmxError(SMX_ERR_ARG, "at most 4 seismogram types");
    for (int k = 0; k < 4; ++k)
      stypes_.t[k] = k < ntypes ? A.seis_types[k] : 0;
    has_pressure_ = A.nrec > 0 &&
                    std::find(A.seis_types.begin(), A.seis_types.end(),
                              int(SMX_SEIS_PRESSURE)) != A.seis_types.end();
    seis_in_tiles_ = A.nrec > 0;
    for (int t : A.seis_types)
      seis_in_tiles_ = seis_in_tiles_ && (t == 0 || t == int(SMX_SEIS_PRESSURE));
    for (int id = 0; id < A.nrec * NPT && seis_in_tiles_; ++id)
      seis_in_tiles_ = A.rent_shared[id] < 0;
    dt_ = A.dt;
    dto2_ = A.dt_over2;
    dt2o2_ = A.dt2_over2;
    field(SMX_FIELD_FORWARD);
    upload_material();
    if (nodes_geometry && !A.nodes.empty())
      try_geometry_from_nodes();
    SMX_HIP(hipStreamSynchronize(stream_));
  }

  ~Solver2D() override {
    drop_graph(graph_, graph_exec_);
    for (int b = 0; b < 2; ++b) {
      drop_graph(rec_graph_[b], rec_graph_exec_[b]);
      drop_graph(adj_graph_[b], adj_graph_exec_[b]);
    }
    (void)hipStreamSynchronize(stream_);
    if (adj_stream_)
      (void)hipStreamSynchronize(adj_stream_);
    for (void *p : allocs_)
      (void)hipFree(p);
    (void)hipEventDestroy(ev0_);
    (void)hipEventDestroy(ev1_);
    for (hipEvent_t e : adj_ev_)
      if (e)
        (void)hipEventDestroy(e);
    for (hipEvent_t e : {rec_done_[0], rec_done_[1], adj_done_[0],
                         adj_done_[1], sync_ev_[0], sync_ev_[1]})
      if (e)
        (void)hipEventDestroy(e);
    if (adj_stream_)
      (void)hipStreamDestroy(adj_stream_);
    (void)hipStreamDestroy(stream_);
  }

  int precision() const override { return sizeof(real); }
  const HostAssembly2D &assembly() const override { return *asm_; }
  int medium() const override { return SMX_MEDIUM_ELASTIC_PSV; }
  void on_stream(void *stream, const std::function<void()> &fn) override {
    StreamSwap sw(stream_, static_cast<hipStream_t>(stream));
    fn();
  }

  void initialize() override {
    const HostAssembly2D &A = *asm_;
    real *M = minv_;
    SMX_HIP(hipMemsetAsync(M, 0, n2_ * sizeof(real), stream_));
    const long n = long(NPT) * A.nspec_pad;
    for (int c = 0; c < A.nelem_colors; ++c) {
      k_mass<real><<<dim3((n + 255) / 256), dim3(256), 0, stream_>>>(
          M, ibool_, ecolor_, c, geom_, kmr_, sslot_, stacey_, A.nspec_pad,
          A.nstacey_pad, quad_, real(A.dt));
      SMX_HIP(hipGetLastError());
    }
    const long m = 2L * A.nglob;
    k_invert<real><<<dim3((m + 255) / 256), dim3(256), 0, stream_>>>(M, m);
    SMX_HIP(hipGetLastError());
    // The step graph is captured, instantiated and uploaded here, so no
    // run_forward call pays for it.  Under a coupled driver the driver
    // captures both media itself.
    if (!chi_dd_) {
      build_graph(field(SMX_FIELD_FORWARD));
      SMX_HIP(hipGraphUpload(graph_exec_, stream_));
    }
    SMX_HIP(hipStreamSynchronize(stream_));
    initialized_ = true;
  }

  // New graph length and occupancy variant: the step graphs are dropped and,
  // on an initialised context, the forward graph is captured and uploaded
  // again, so the next run_forward replays the new layout at once.
  void set_launch(int graph_steps, int waves_per_eu) override {
    check_launch(graph_steps, waves_per_eu, sizeof(real));
    if (chi_dd_)
      throw SmxError(SMX_ERR_STATE,
                     "the launch layout of a coupled medium is its driver's");
    SMX_HIP(hipStreamSynchronize(stream_));
    if (adj_stream_)
      SMX_HIP(hipStreamSynchronize(adj_stream_));
    drop_graph(graph_, graph_exec_);
    for (int b = 0; b < 2; ++b) {
      drop_graph(rec_graph_[b], rec_graph_exec_[b]);
      drop_graph(adj_graph_[b], adj_graph_exec_[b]);
    }
    gsteps_ = graph_steps;
    wpe_ = waves_per_eu;
    if (initialized_) {
      build_graph(field(SMX_FIELD_FORWARD));
      SMX_HIP(hipGraphUpload(graph_exec_, stream_));
      SMX_HIP(hipStreamSynchronize(stream_));
    }
  }

  int predictor(int f) override {
    Field &F = field(f);
    launch_predictor(F);
    return 2 * asm_->nglob;
  }

  int corrector(int f) override {
    Field &F = field(f);
    k_corrector<real><<<node_grid(), dim3(NODE_THREADS), 0, stream_>>>(
        F.v, F.a, n2_, dto2_);
    SMX_HIP(hipGetLastError());
    return 2 * asm_->nglob;
  }

  int update_wavefields(int f, int istep) override {
    require_init();
    const HostAssembly2D &A = *asm_;
    if (istep < 0 || istep >= A.nstep)
      throw SmxError(SMX_ERR_ARG, "istep outside [0, nstep)");
    Field &F = field(f);
    launch_tiles<MODE_ASSEMBLE>(F.u, F.v, F.a, nullptr, istep, nullptr);
    launch_shared<MODE_ASSEMBLE>(F.u, F.v, F.a, nullptr);
    return A.nspec;
  }

  // The same two launches on caller-held fields, partials in `scratch`.
  // Every partial slot is written by the tile pass before the shared-point
  // pass reads it, so nothing in the scratch array needs clearing.
  int update_wavefields_noalloc(const void *u, const void *v, void *a,
                                int istep) override {
    require_init();
    const HostAssembly2D &A = *asm_;
    if (istep < 0 || istep >= A.nstep)
      throw SmxError(SMX_ERR_ARG, "istep outside [0, nstep)");
    if (!u || !v || !a)
      throw SmxError(SMX_ERR_ARG, "null field pointer");
    const size_t slots = std::max<size_t>(1, A.sh_list.size());
    const size_t need = 2 * slots * sizeof(real);
    if (need > kNoAllocScratchBytes)
      throw SmxError(SMX_ERR_ARG,
                     "mesh needs " + std::to_string(need) +
                         " bytes of shared-point partials; the no-alloc "
                         "scratch holds " +
                         std::to_string(kNoAllocScratchBytes) + " (" +
                         std::to_string(kNoAllocScratchBytes / (2 * sizeof(real))) +
                         " partial slots at this precision)");
    real *part = static_cast<real *>(noalloc_scratch(device_));
    real *U = const_cast<real *>(static_cast<const real *>(u));
    real *V = const_cast<real *>(static_cast<const real *>(v));
    real *Aout = static_cast<real *>(a);
    launch_tiles<MODE_ASSEMBLE>(U, V, Aout, part, istep, nullptr);
    launch_shared<MODE_ASSEMBLE>(U, V, Aout, nullptr, nullptr, part);
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
    k_seismo<real><<<dim3((n + 127) / 128), dim3(128), 0, stream_>>>(
        F.u, F.v, F.a, rent_g_, rent_h_, A.nrec, stypes_, ntypes, isig, seis_);
    SMX_HIP(hipGetLastError());
    if (has_pressure_)
      launch_pressure(F.u, nullptr, isig);
  }

  void run_forward(int istep0, int nstep) override {
    require_init();
    const HostAssembly2D &A = *asm_;
    if (istep0 < 0 || nstep < 1 || istep0 + nstep > A.nstep)
      throw SmxError(SMX_ERR_ARG, "steps outside [0, nstep)");
    Field &F = field(SMX_FIELD_FORWARD);
    SMX_HIP(hipEventRecord(ev0_, stream_));
    k_set_step<<<dim3(1), dim3(64), 0, stream_>>>(step_, istep0);
    SMX_HIP(hipGetLastError());
    launch_predictor(F);
    const int body = nstep - 1; // steps followed by a fused predictor
    const int reps = body / gsteps_;
    if (reps > 0) {
      build_graph(F);
      for (int r = 0; r < reps; ++r)
        SMX_HIP(hipGraphLaunch(graph_exec_, stream_));
    }
    for (int s = reps * gsteps_; s < body; ++s)
      fused_step<MODE_STEP>(F);
    fused_step<MODE_FINAL>(F);
    SMX_HIP(hipEventRecord(ev1_, stream_));
    last_launches_ = nstep;
    have_stats_ = true;
  }

  void synchronize() override { SMX_HIP(hipStreamSynchronize(stream_)); }

  // ---- coupled meshes (the driver enqueues both media on its stream)
  void bind_coupling(void *, void *chi_dd) override {
    if (sizeof(real) != 4)
      throw SmxError(SMX_ERR_ARG, "coupled meshes run in FP32");
    chi_dd_ = static_cast<const real *>(chi_dd);
  }
  void *coupling_field() override { return field(SMX_FIELD_FORWARD).u; }
  void enqueue_begin(void *stream, int istep0) override {
    require_init();
    StreamSwap sw(stream_, static_cast<hipStream_t>(stream));
    k_set_step<<<dim3(1), dim3(64), 0, stream_>>>(step_, istep0);
    SMX_HIP(hipGetLastError());
    launch_predictor(field(SMX_FIELD_FORWARD));
  }
  void enqueue_step(void *stream, bool final) override {
    StreamSwap sw(stream_, static_cast<hipStream_t>(stream));
    Field &F = field(SMX_FIELD_FORWARD);
    if (final)
      fused_step<MODE_FINAL>(F);
    else
      fused_step<MODE_STEP>(F);
  }

  void get_seismograms(void *host) override {
    const HostAssembly2D &A = *asm_;
    const size_t n = size_t(A.nsig) * A.seis_types.size() * A.nrec * 2;
    SMX_HIP(hipMemcpyAsync(host, seis_, n * sizeof(real),
                           hipMemcpyDeviceToHost, stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
  }

  void get_field(int f, int which, void *host) override {
    const real *src = pick(f, which);
    SMX_HIP(hipMemcpyAsync(host, src, 2 * size_t(asm_->nglob) * sizeof(real),
                           hipMemcpyDeviceToHost, stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
  }

  void set_field(int f, int which, const void *host) override {
    real *dst = pick(f, which);
    SMX_HIP(hipMemcpyAsync(dst, host, 2 * size_t(asm_->nglob) * sizeof(real),
                           hipMemcpyHostToDevice, stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
  }

  void *field_ptr(int f, int which) override { return pick(f, which); }

  void last_run_stats(double *ms, double *stiff_ms, int *launches) override {
    float t = 0;
    if (have_stats_) {
      SMX_HIP(hipEventSynchronize(ev1_));
      SMX_HIP(hipEventElapsedTime(&t, ev0_, ev1_));
    }
    if (ms)
      *ms = t;
    if (stiff_ms)
      *stiff_ms = last_stiff_ms_;
    if (launches)
      *launches = last_launches_;
  }

  // Mean duration of one step's tile launches (every tile colour, MODE_STEP,
  // no seismograms) on scratch copies of u and v.
  double time_stiffness(int n) override {
    if (n < 1)
      throw SmxError(SMX_ERR_ARG, "n must be positive");
    require_init();
    Field &F = field(SMX_FIELD_FORWARD);
    if (!scratch_u_) {
      scratch_u_ = alloc<real>(n2_);
      scratch_v_ = alloc<real>(n2_);
      scratch_part_ = alloc<real>(2 * std::max<size_t>(1, asm_->sh_list.size()));
    }
    SMX_HIP(hipMemcpyAsync(scratch_u_, F.u, n2_ * sizeof(real),
                           hipMemcpyDeviceToDevice, stream_));
    SMX_HIP(hipMemcpyAsync(scratch_v_, F.v, n2_ * sizeof(real),
                           hipMemcpyDeviceToDevice, stream_));
    launch_tiles<MODE_STEP>(scratch_u_, scratch_v_, nullptr, scratch_part_, 0,
                            nullptr); // warm-up
    SMX_HIP(hipEventRecord(ev0_, stream_));
    for (int i = 0; i < n; ++i)
      launch_tiles<MODE_STEP>(scratch_u_, scratch_v_, nullptr, scratch_part_, 0,
                              nullptr);
    SMX_HIP(hipEventRecord(ev1_, stream_));
    SMX_HIP(hipEventSynchronize(ev1_));
    float t = 0;
    SMX_HIP(hipEventElapsedTime(&t, ev0_, ev1_));
    last_stiff_ms_ = double(t) / n;
    have_stats_ = false;
    return last_stiff_ms_;
  }

  // ---------------------------------------------------------------- adjoint
  void set_model(const float *kappa, const float *mu,
                 const float *rho) override {
    HostAssembly2D &A = *asm_;
    sync_host_model(); // the host arrays a device model left stale
    set_model2d(A, kappa, mu, rho);
    upload_material();
    ckpt_every_ = 0;
    initialize();
  }

  // The same update from device arrays: validated, then every derived
  // device array recomputed in place; the host copy is marked stale.
  void set_model_device(const float *kappa, const float *mu,
                        const float *rho) override {
    if (sizeof(real) != 4)
      return Solver2DBase::set_model_device(kappa, mu, rho);
    if (asm_->shot_elem_off.size() > 2)
      throw SmxError(SMX_ERR_STATE, "device-memory models are built for single-mesh contexts");
    const HostAssembly2D &A = *asm_;
    ensure_model_maps();
    ensure_adjoint(); // the boundary matrices are refreshed with the model
    const long n = long(A.nspec) * NPT;
    const int ok = std::numeric_limits<int>::max();
    SMX_HIP(hipMemcpyAsync(model_bad_, &ok, sizeof(int), hipMemcpyHostToDevice,
                           stream_));
    k_model_check<<<dim3((n + 255) / 256), dim3(256), 0, stream_>>>(
        kappa, mu, rho, n, model_bad_);
    SMX_HIP(hipGetLastError());
    int bad = ok;
    SMX_HIP(hipMemcpyAsync(&bad, model_bad_, sizeof(int), hipMemcpyDeviceToHost,
                           stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
    if (bad != ok)
      throw SmxError(SMX_ERR_ARG,
                     "non-physical material at element " + std::to_string(bad));
    float *kmr = reinterpret_cast<float *>(kmr_), *mat = reinterpret_cast<float *>(mat_);
    float *st = reinterpret_cast<float *>(stacey_);
    k_model_in<<<dim3((n + 255) / 256), dim3(256), 0, stream_>>>(
        kappa, mu, rho, elem_slot_, A.nspec, A.nspec_pad, kmr, mat);
    SMX_HIP(hipGetLastError());
    const int nbpt = int(A.bpt_g.size()), nbn = int(A.bnode_g.size());
    if (nbpt > 0) {
      k_model_stacey<<<dim3((nbpt + 255) / 256), dim3(256), 0, stream_>>>(
          bpt_p_, bpt_bslot_, bpt_slot_, nbpt, A.nspec_pad, A.nstacey_pad, kmr, st);
      SMX_HIP(hipGetLastError());
    }
    if (nbn > 0) {
      k_model_bsum<<<dim3((nbn + 255) / 256), dim3(256), 0, stream_>>>(
          bnode_off_, bnode_pts_, bpt_p_, bpt_bslot_, nbn, A.nstacey_pad, st,
          reinterpret_cast<float *>(bsum_));
      SMX_HIP(hipGetLastError());
    }
    k_model_compact<<<dim3(A.ntiles), dim3(ELEM_TILE), 0, stream_>>>(
        tiles_, mat, reinterpret_cast<float *>(mat_e_), A.nspec_pad);
    SMX_HIP(hipGetLastError());
    host_model_stale_ = true;
    ckpt_every_ = 0;
    initialize();
  }

  void get_gradient_device(double *out) override {
    if (sizeof(real) != 4)
      return Solver2DBase::get_gradient_device(out);
    if (!have_gradient_)
      throw SmxError(SMX_ERR_STATE, "no adjoint run");
    const HostAssembly2D &A = *asm_;
    ensure_model_maps();
    const long n = long(A.nspec) * NPT;
    k_gradient_out<real><<<dim3((n + 255) / 256), dim3(256), 0, stream_>>>(
        elem_slot_, sslot_, geom_, quad_, ibool_, kmr_, stacey_, gd_, ge_,
        racc_, sp_, ss_, A.nspec, A.nspec_pad, A.nstacey_pad, out);
    SMX_HIP(hipGetLastError());
    SMX_HIP(hipStreamSynchronize(stream_));
  }

  void sync_host_model() override {
    if (!host_model_stale_)
      return;
    HostAssembly2D &A = *asm_;
    download(A.kappa_mu_rho.data(), reinterpret_cast<const float *>(kmr_),
             A.kappa_mu_rho.size());
    download(A.lam2mu_mu.data(), reinterpret_cast<const float *>(mat_),
             A.lam2mu_mu.size());
    download(A.stacey.data(), reinterpret_cast<const float *>(stacey_),
             A.stacey.size());
    download(A.mat_e.data(), reinterpret_cast<const float *>(mat_e_),
             A.mat_e.size());
    download(A.tiles.data(), tiles_, A.tiles.size());
    if (bsum_ && !A.bsum.empty()) {
      std::vector<float> b(A.bsum.size());
      download(b.data(), reinterpret_cast<const float *>(bsum_), b.size());
      std::copy(b.begin(), b.end(), A.bsum.begin());
    }
    host_model_stale_ = false;
  }

  void run_forward_checkpointed(int every) override {
    require_init();
    const HostAssembly2D &A = *asm_;
    if (every < 1)
      throw SmxError(SMX_ERR_ARG, "checkpoint interval must be positive");
    every = std::min(every, A.nstep);
    const int nseg = (A.nstep + every - 1) / every;
    if (nseg > ckpt_cap_) {
      release(ckpt_);
      ckpt_ = alloc<real>(size_t(nseg) * 2 * n2_);
      ckpt_cap_ = nseg;
    }
    if (every != rec_every_) {
      for (int b = 0; b < 2; ++b) {
        release(rec_[b]);
        rec_[b] = alloc<real>(size_t(every + 1) * 3 * n2_);
        drop_graph(rec_graph_[b], rec_graph_exec_[b]);
        drop_graph(adj_graph_[b], adj_graph_exec_[b]);
      }
      rec_every_ = every;
    }
    ensure_adjoint_events();
    Field &F = field(SMX_FIELD_FORWARD);
    SMX_HIP(hipEventRecord(adj_ev_[0], stream_));
    for (real *p : {F.u, F.v, F.a}) // from rest
      SMX_HIP(hipMemsetAsync(p, 0, n2_ * sizeof(real), stream_));
    k_set_step<<<dim3(1), dim3(64), 0, stream_>>>(step_, 0);
    SMX_HIP(hipGetLastError());
    launch_predictor(F);
    for (int sgm = 0; sgm < nseg; ++sgm) {
      const int k0 = sgm * every, len = std::min(every, A.nstep - k0);
      real *ck = ckpt_ + size_t(sgm) * 2 * n2_;
      SMX_HIP(hipMemcpyAsync(ck, F.u, n2_ * sizeof(real),
                             hipMemcpyDeviceToDevice, stream_));
      SMX_HIP(hipMemcpyAsync(ck + n2_, F.v, n2_ * sizeof(real),
                             hipMemcpyDeviceToDevice, stream_));
      const bool last = sgm == nseg - 1;
      const int body = last ? len - 1 : len;
      const int reps = body / gsteps_;
      if (reps > 0) {
        build_graph(F);
        for (int r = 0; r < reps; ++r)
          SMX_HIP(hipGraphLaunch(graph_exec_, stream_));
      }
      for (int st = reps * gsteps_; st < body; ++st)
        fused_step<MODE_STEP>(F);
      if (last)
        fused_step<MODE_FINAL>(F);
    }
    SMX_HIP(hipEventRecord(adj_ev_[1], stream_));
    ckpt_every_ = every;
    have_adj_stats_ = false;
  }

  // Segments are processed newest first.  The recompute of segment s-1
  // (context stream, record buffer (s-1)%2) runs concurrently with the
  // adjoint steps of segment s (adjoint stream, record buffer s%2); events
  // order each recompute after the adjoint that last read its buffer and
  // each adjoint after the recompute of its segment.
  void run_adjoint(const void *gsrc_host) override {
    const HostAssembly2D &A = *asm_;
    if (!ckpt_every_)
      throw SmxError(SMX_ERR_STATE,
                     "no checkpointed forward run for the current model");
    ensure_adjoint();
    const size_t ng = size_t(A.nstep) * A.nrec * 2;
    if (ng)
      SMX_HIP(hipMemcpyAsync(gsrc_, gsrc_host, ng * sizeof(real),
                             hipMemcpyHostToDevice, stream_));
    Field &W = field(SMX_FIELD_BACKWARD);
    Field &B = field(SMX_FIELD_ADJOINT);
    const int every = ckpt_every_;
    const int nseg = (A.nstep + every - 1) / every;
    SMX_HIP(hipEventRecord(adj_ev_[2], stream_));
    for (real *p : {B.u, B.v, rb0_, rb1_, racc_})
      SMX_HIP(hipMemsetAsync(p, 0, n2_ * sizeof(real), stream_));
    for (real *p : {gd_, ge_})
      SMX_HIP(hipMemsetAsync(p, 0, size_t(NPT) * A.nspec_pad * sizeof(real),
                             stream_));
    for (real *p : {sp_, ss_})
      SMX_HIP(hipMemsetAsync(p, 0, size_t(NPT) * A.nstacey_pad * sizeof(real),
                             stream_));
    // fork: the adjoint stream starts after everything queued so far
    SMX_HIP(hipEventRecord(sync_ev_[0], stream_));
    SMX_HIP(hipStreamWaitEvent(adj_stream_, sync_ev_[0], 0));
    const int nrin = int(A.rin_node.size());
    if (nrin > 0) {
      k_adj_inject<real><<<dim3((nrin + 255) / 256), dim3(256), 0,
                           adj_stream_>>>(B.u, rin_node_, rin_off_, rin_ent_,
                                          rent_h_, gsrc_, A.nrec, nrin,
                                          A.nstep - 1);
      SMX_HIP(hipGetLastError());
    }
    for (int sgm = nseg - 1; sgm >= 0; --sgm) {
      const int b = sgm & 1;
      const int k0 = sgm * every, len = std::min(every, A.nstep - k0);
      const int reps = len / gsteps_;
      const real *ck = ckpt_ + size_t(sgm) * 2 * n2_;
      // recompute the segment from its checkpoint into record buffer b,
      // once the adjoint of segment sgm+2 (its last reader) is done
      if (sgm + 2 < nseg)
        SMX_HIP(hipStreamWaitEvent(stream_, adj_done_[b], 0));
      SMX_HIP(hipMemcpyAsync(W.u, ck, n2_ * sizeof(real),
                             hipMemcpyDeviceToDevice, stream_));
      SMX_HIP(hipMemcpyAsync(W.v, ck + n2_, n2_ * sizeof(real),
                             hipMemcpyDeviceToDevice, stream_));
      SMX_HIP(hipMemcpyAsync(rec_[b], ck, 2 * n2_ * sizeof(real),
                             hipMemcpyDeviceToDevice, stream_));
      k_set_step<<<dim3(1), dim3(64), 0, stream_>>>(step_, k0);
      SMX_HIP(hipGetLastError());
      if (reps > 0) {
        build_rec_graph(W, b);
        for (int r = 0; r < reps; ++r)
          SMX_HIP(hipGraphLaunch(rec_graph_exec_[b], stream_));
      }
      for (int st = reps * gsteps_; st < len; ++st)
        record_step(W, rec_[b]);
      SMX_HIP(hipEventRecord(rec_done_[b], stream_));
      // adjoint steps k0+len-1 .. k0 on the adjoint stream
      SMX_HIP(hipStreamWaitEvent(adj_stream_, rec_done_[b], 0));
      k_set_step<<<dim3(1), dim3(64), 0, adj_stream_>>>(step_ + 2,
                                                         k0 + len - 1);
      SMX_HIP(hipGetLastError());
      if (reps > 0) {
        build_adj_graph(B, b);
        for (int r = 0; r < reps; ++r)
          SMX_HIP(hipGraphLaunch(adj_graph_exec_[b], adj_stream_));
      }
      for (int st = reps * gsteps_; st < len; ++st)
        adjoint_step(B, rec_[b], adj_stream_);
      SMX_HIP(hipEventRecord(adj_done_[b], adj_stream_));
    }
    // join
    SMX_HIP(hipEventRecord(sync_ev_[1], adj_stream_));
    SMX_HIP(hipStreamWaitEvent(stream_, sync_ev_[1], 0));
    SMX_HIP(hipEventRecord(adj_ev_[3], stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
    have_adj_stats_ = true;
    have_gradient_ = true;
  }

  void get_gradient(double *out) override {
    if (!have_gradient_)
      throw SmxError(SMX_ERR_STATE, "no adjoint run");
    sync_host_model();
    const HostAssembly2D &A = *asm_;
    const int ld = A.nspec_pad, sld = A.nstacey_pad;
    const size_t pld = size_t(NPT) * ld, sp = size_t(NPT) * sld;
    std::vector<real> gd(pld), ge(pld), rr(n2_ / 2), sP(sp), sS(sp);
    download(gd.data(), gd_, pld);
    download(ge.data(), ge_, pld);
    download(rr.data(), racc_, n2_ / 2);
    download(sP.data(), sp_, sp);
    download(sS.data(), ss_, sp);
    const size_t nloc = size_t(A.nspec) * NPT;
    for (int e = 0; e < A.nspec; ++e) {
      const int slot = A.elem_slot[e], bs = A.stacey_slot[slot];
      for (int p = 0; p < NPT; ++p) {
        const size_t o = size_t(p) * ld + slot;
        const int iz = p / NGLL, ix = p % NGLL;
        const double wj =
            double(A.quad.w[iz]) * A.quad.w[ix] * A.geom[4 * pld + o];
        const double Gd = gd[o], Ge = ge[o];
        const int g = A.ibool[o];
        double gk = -wj * Gd;
        double gm = -wj * (2.0 * Ge - 2.0 / 3.0 * Gd);
        double gr = -wj * double(rr[g]);
        if (bs >= 0 && A.stacey[3 * sp + size_t(p) * sld + bs] != 0.0f) {
          const double k = A.kappa_mu_rho[o], m = A.kappa_mu_rho[pld + o],
                       rh = A.kappa_mu_rho[2 * pld + o];
          const double l2m = k + 4.0 / 3.0 * m;
          const double P = std::sqrt(rh * l2m), S = std::sqrt(rh * m);
          const double dP = sP[size_t(p) * sld + bs],
                       dS = sS[size_t(p) * sld + bs];
          gk += dP * rh / (2.0 * P);
          gm += dP * (4.0 / 3.0) * rh / (2.0 * P);
          gr += dP * l2m / (2.0 * P);
          if (S > 0) {
            gm += dS * rh / (2.0 * S);
            gr += dS * m / (2.0 * S);
          }
        }
        const size_t i = size_t(e) * NPT + p;
        out[i] = gk;
        out[nloc + i] = gm;
        out[2 * nloc + i] = gr;
      }
    }
  }

  void geometry_mode(int *mode, long long *mismatches) const override {
    if (mode)
      *mode = gmode_;
    if (mismatches)
      *mismatches = geom_mismatch_;
  }

  void last_adjoint_stats(double *fwd_ms, double *adj_ms) override {
    float tf = 0, ta = 0;
    if (have_adj_stats_) {
      SMX_HIP(hipEventSynchronize(adj_ev_[3]));
      SMX_HIP(hipEventElapsedTime(&tf, adj_ev_[0], adj_ev_[1]));
      SMX_HIP(hipEventElapsedTime(&ta, adj_ev_[2], adj_ev_[3]));
    }
    if (fwd_ms)
      *fwd_ms = tf;
    if (adj_ms)
      *adj_ms = ta;
  }

private:
  struct Field {
    real *u = nullptr, *v = nullptr, *a = nullptr;
  };
  int gsteps_ = 8; // time steps per replayed graph
  int wpe_ = 0;    // occupancy variant of k_tile<float, MODE_STEP> (0: default)

  template <typename T> T *alloc(size_t n) {
    void *p = nullptr;
    SMX_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    allocs_.push_back(p);
    SMX_HIP(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), stream_));
    return static_cast<T *>(p);
  }
  template <typename T, typename H> T *upload_raw(const std::vector<H> &h) {
    static_assert(sizeof(T) == sizeof(H), "layout mismatch");
    T *d = alloc<T>(h.size());
    if (!h.empty()) {
      SMX_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T),
                             hipMemcpyHostToDevice, stream_));
      SMX_HIP(hipStreamSynchronize(stream_));
    }
    return d;
  }
  real *upload_real(const std::vector<float> &h) {
    real *d = alloc<real>(h.size());
    if (!h.empty()) {
      std::vector<real> tmp(h.begin(), h.end());
      SMX_HIP(hipMemcpyAsync(d, tmp.data(), tmp.size() * sizeof(real),
                             hipMemcpyHostToDevice, stream_));
      SMX_HIP(hipStreamSynchronize(stream_));
    }
    return d;
  }

  Field &field(int f) {
    if (f < 0 || f > 2)
      throw SmxError(SMX_ERR_ARG, "unknown wavefield");
    Field &F = fields_[f];
    if (!F.u) {
      F.u = alloc<real>(n2_);
      F.v = alloc<real>(n2_);
      F.a = alloc<real>(n2_);
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
    constexpr int W = Vec16<real>::W;
    return dim3(unsigned((n2_ / W + NODE_THREADS - 1) / NODE_THREADS));
  }

  void launch_predictor(Field &F) {
    k_predictor<real><<<node_grid(), dim3(NODE_THREADS), 0, stream_>>>(
        F.u, F.v, F.a, n2_, dt_, dto2_, dt2o2_);
    SMX_HIP(hipGetLastError());
  }

  // One launch over every tile.  `part` defaults to the context's partials
  // buffer; seismogram terms are written only when `terms` is given.
  template <int MODE>
  void launch_tiles(real *U, real *V, real *Aout, real *part, int istep,
                    int *step_ptr, real *terms = nullptr,
                    real *rec = nullptr) {
    const HostAssembly2D &A = *asm_;
    const int ntypes = int(A.seis_types.size());
    auto kern = gmode_ == 2 ? &k_tile<real, MODE, true> : &k_tile<real, MODE, false>;
    if constexpr (MODE == MODE_STEP && std::is_same<real, float>::value) {
      if (gmode_ != 2 && wpe_ == 6)
        kern = &k_tile<real, MODE, false, 6>;
      else if (gmode_ != 2 && wpe_ == 8)
        kern = &k_tile<real, MODE, false, 8>;
    }
    kern<<<dim3(A.ntiles), dim3(TILE_THREADS), lds_bytes_, stream_>>>(
        tiles_, A.nloc_max, sh_list_, lidx_, ecolor_, A.nelem_colors, geom_,
        mat_, geom_e_, mat_e_, nodes_, sslot_, stacey_, U, V, Aout, minv_,
        part ? part : part_, src_li_,
        src_idx_, src_L_, stf_, std::max(1, A.nsrc), trec_li_, trec_entry_,
        rent_h_, terms, A.nrec * NPT, stypes_, ntypes, A.nsbs, step_ptr, istep,
        A.nspec_pad, A.nstacey_pad, dt_, dto2_, dt2o2_, quad_, rec,
        rec_stride(), std::max(1, rec_every_), cpl_li_, cpl_k_, cpl_w_,
        chi_dd_, sh_pad_, sh_stride_);
    SMX_HIP(hipGetLastError());
  }
  void launch_pressure(const real *U, const int *step_ptr, int isig) {
    const HostAssembly2D &A = *asm_;
    k_seis_pressure<real><<<dim3((A.nrec + SEIS_GROUPS - 1) / SEIS_GROUPS),
                            dim3(32 * SEIS_GROUPS), 0, stream_>>>(
        U, rent_g_, rent_h_, rec_slot_, geom_, mat_, A.nspec_pad, quad_,
        A.nrec, stypes_, int(A.seis_types.size()), A.nsbs, step_ptr, isig,
        seis_);
    SMX_HIP(hipGetLastError());
  }
  // seis: the step's seismogram sums in the first workgroup (MODE_STEP with
  // a step counter; every term written by the tiles)
  template <int MODE>
  void launch_shared(real *U, real *V, real *Aout, int *step_ptr,
                     real *rec = nullptr, real *part = nullptr,
                     bool seis = false) {
    const HostAssembly2D &A = *asm_;
    const int n = std::max(1, A.nshared);
    SeisSum<real> ss{seis_terms_, seis_, stypes_, A.nrec,
                     int(A.seis_types.size()), A.nsbs, 0};
    if (MODE == MODE_STEP && seis && step_ptr)
      ss.on = 1;
    k_shared<real, MODE><<<dim3((n + NODE_THREADS - 1) / NODE_THREADS + ss.on),
                           dim3(NODE_THREADS), 0, stream_>>>(
        U, V, Aout, minv_, part ? part : part_, shp_off_, shp_pos_, A.nshared,
        A.nex_total,
        dt_, dto2_, dt2o2_, step_ptr, rec, rec_stride(),
        std::max(1, rec_every_), ss);
    SMX_HIP(hipGetLastError());
  }
  long rec_stride() const { return long(n2_ * sizeof(real)); }

  // Tiles whose geometry is not element-constant re-form it from the control
  // nodes when the device reproduces every stored per-point value bit for bit
  // (checked once here); otherwise they keep reading the stored values.
  void try_geometry_from_nodes() {
    HostAssembly2D &A = *asm_;
    const int ld = A.nspec_pad;
    const size_t n = size_t(NPT) * ld;
    nodes_ = alloc<float>(A.nodes.size());
    SMX_HIP(hipMemcpyAsync(nodes_, A.nodes.data(), A.nodes.size() * sizeof(float),
                           hipMemcpyHostToDevice, stream_));
    float *gf = nullptr;
    unsigned char *bad = nullptr;
    SMX_HIP(hipMalloc(&gf, size_t(5) * n * sizeof(float)));
    SMX_HIP(hipMalloc(&bad, n));
    SMX_HIP(hipMemcpyAsync(gf, A.geom.data(), size_t(5) * n * sizeof(float),
                           hipMemcpyHostToDevice, stream_));
    QuadArgs<float> qf{};
    for (int i = 0; i < NGLL; ++i)
      qf.xi[i] = A.quad.xi[i];
    k_geom_check2<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_>>>(
        nodes_, gf, ecolor_, ld, qf, bad);
    SMX_HIP(hipGetLastError());
    std::vector<unsigned char> h(n);
    SMX_HIP(hipMemcpyAsync(h.data(), bad, n, hipMemcpyDeviceToHost, stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
    (void)hipFree(gf);
    (void)hipFree(bad);
    long long nbad = 0;
    for (unsigned char b : h)
      nbad += b;
    geom_mismatch_ = nbad;
    if (nbad != 0)
      return;
    gmode_ = 2;
    for (auto &t : A.tiles)
      if (!(t.flags & TILE_GEOM_ELEM))
        t.flags |= TILE_GEOM_NODES;
    SMX_HIP(hipMemcpyAsync(tiles_, A.tiles.data(),
                           A.tiles.size() * sizeof(TileInfo),
                           hipMemcpyHostToDevice, stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
  }

  // Material arrays on the device.  An FP32 context uses the FP32 values
  // the host formed (the reference's rounding); an FP64 context takes
  // lambda+2mu and the Stacey impedances formed in double from the same
  // FP32 model, so that it is an FP64 solve of that model.
  void upload_material() {
    const HostAssembly2D &A = *asm_;
    std::vector<double> m(A.lam2mu_mu.begin(), A.lam2mu_mu.end());
    std::vector<double> st(A.stacey.begin(), A.stacey.end());
    if (sizeof(real) == 8) {
      const size_t sp = A.imp64.size() / 2; // NPT * nstacey_pad
      std::copy(A.lam2mu64.begin(), A.lam2mu64.end(), m.begin());
      std::copy(A.imp64.begin(), A.imp64.end(), st.begin() + 4 * sp);
    }
    std::vector<double> me(A.mat_e.begin(), A.mat_e.end());
    if (sizeof(real) == 8)
      std::copy(A.lam2mu64.begin(), A.lam2mu64.begin() + A.nspec_pad,
                me.begin()); // point 0 of each slot
    copy_to(mat_, m);
    copy_to(mat_e_, me);
    SMX_HIP(hipMemcpyAsync(tiles_, A.tiles.data(),
                           A.tiles.size() * sizeof(TileInfo),
                           hipMemcpyHostToDevice, stream_));
    copy_to(stacey_, st);
    copy_to(kmr_, A.kappa_mu_rho);
    if (bsum_)
      copy_to(bsum_, sizeof(real) == 8 ? A.bsum64 : A.bsum);
  }

  // Recompute step: the fused step that also records (u~, v~, a') into
  // `rec` (context stream).
  void record_step(Field &W, real *rec) {
    launch_tiles<MODE_RECORD>(W.u, W.v, W.a, nullptr, 0, step_, nullptr, rec);
    launch_shared<MODE_RECORD>(W.u, W.v, W.a, step_, rec);
  }
  // One adjoint step reading the record `rec` (device step counter
  // step_[2]; own partials buffer, so that it may overlap a recompute).
  void adjoint_step(Field &B, real *rec, hipStream_t st) {
    const HostAssembly2D &A = *asm_;
    auto kern = gmode_ == 2 ? &k_adj_tile<real, true> : &k_adj_tile<real, false>;
    kern<<<dim3(A.ntiles), dim3(TILE_THREADS), lds_bytes_, st>>>(
        tiles_, A.nloc_max, sh_list_, lidx_, ecolor_, A.nelem_colors, geom_,
        mat_, geom_e_, mat_e_, nodes_, sslot_, stacey_, rb0_, rb1_, B.u, B.v,
        minv_,
        part_adj_, rec,
        rec_stride(), rec_every_, bidx_, bsum_, gd_, ge_, racc_, sp_, ss_,
        step_ + 2, A.nspec_pad, A.nstacey_pad, dt_, dto2_, dt2o2_, quad_);
    SMX_HIP(hipGetLastError());
    const int nex_rin = int(A.rin_excl.size());
    const int n = std::max(1, A.nshared + nex_rin);
    k_adj_shared<real><<<dim3((n + NODE_THREADS - 1) / NODE_THREADS),
                         dim3(NODE_THREADS), 0, st>>>(
        rb0_, rb1_, B.u, B.v, minv_, part_adj_, shp_off_, shp_pos_, A.nshared,
        A.nex_total, rec, rec_stride(), rec_every_, bidx_, bsum_, racc_,
        sh_rin_, rin_excl_, nex_rin, rin_node_, rin_off_, rin_ent_, rent_h_,
        gsrc_, A.nrec, step_ + 2, dt_, dto2_, dt2o2_);
    SMX_HIP(hipGetLastError());
  }
  template <typename F>
  void capture(hipGraph_t &g, hipGraphExec_t &ge, hipStream_t st, F body) {
    if (ge)
      return;
    SMX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      body();
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(st, &tmp);
      if (tmp)
        (void)hipGraphDestroy(tmp);
      throw;
    }
    SMX_HIP(hipStreamEndCapture(st, &g));
    SMX_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
  }
  void build_rec_graph(Field &W, int b) {
    capture(rec_graph_[b], rec_graph_exec_[b], stream_, [&] {
      for (int s = 0; s < gsteps_; ++s)
        record_step(W, rec_[b]);
    });
  }
  void build_adj_graph(Field &B, int b) {
    capture(adj_graph_[b], adj_graph_exec_[b], adj_stream_, [&] {
      for (int s = 0; s < gsteps_; ++s)
        adjoint_step(B, rec_[b], adj_stream_);
    });
  }
  static void drop_graph(hipGraph_t &g, hipGraphExec_t &ge) {
    if (ge)
      (void)hipGraphExecDestroy(ge);
    if (g)
      (void)hipGraphDestroy(g);
    g = nullptr;
    ge = nullptr;
  }
  template <typename T> void release(T *&p) {
    if (!p)
      return;
    SMX_HIP(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < allocs_.size(); ++i)
      if (allocs_[i] == static_cast<void *>(p)) {
        allocs_.erase(allocs_.begin() + i);
        break;
      }
    SMX_HIP(hipFree(p));
    p = nullptr;
  }
  template <typename H> void copy_to(real *d, const std::vector<H> &h) {
    std::vector<real> tmp(h.begin(), h.end());
    SMX_HIP(hipMemcpyAsync(d, tmp.data(), tmp.size() * sizeof(real),
                           hipMemcpyHostToDevice, stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
  }
  template <typename T> void download(T *h, const T *d, size_t n) {
    SMX_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost,
                           stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
  }
  void ensure_adjoint_events() {
    for (hipEvent_t &e : adj_ev_)
      if (!e)
        SMX_HIP(hipEventCreate(&e));
  }
  // Slot maps of the device-memory model and gradient, uploaded once: the
  // element slots, the absorbing points and each absorbing node's points.
  void ensure_model_maps() {
    if (elem_slot_)
      return;
    const HostAssembly2D &A = *asm_;
    elem_slot_ = upload_raw<int>(A.elem_slot);
    bpt_p_ = upload_raw<int>(A.bpt_p);
    bpt_bslot_ = upload_raw<int>(A.bpt_bslot);
    bpt_slot_ = upload_raw<int>(A.bpt_slot);
    std::unordered_map<int, int> index;
    for (size_t j = 0; j < A.bnode_g.size(); ++j)
      index[A.bnode_g[j]] = int(j);
    std::vector<int> count(A.bnode_g.size() + 1, 0), pts(A.bpt_g.size());
    for (size_t k = 0; k < A.bpt_g.size(); ++k)
      ++count[index.at(A.bpt_g[k]) + 1];
    for (size_t j = 0; j < A.bnode_g.size(); ++j)
      count[j + 1] += count[j];
    std::vector<int> fill(count.begin(), count.end() - 1);
    for (size_t k = 0; k < A.bpt_g.size(); ++k) // ascending k per node
      pts[fill[index.at(A.bpt_g[k])]++] = int(k);
    bnode_off_ = upload_raw<int>(count);
    bnode_pts_ = upload_raw<int>(pts);
    model_bad_ = alloc<int>(1);
  }
  // Device arrays of the adjoint, allocated on the first adjoint run.
  void ensure_adjoint() {
    if (rb0_)
      return;
    const HostAssembly2D &A = *asm_;
    SMX_HIP(hipStreamCreateWithFlags(&adj_stream_, hipStreamNonBlocking));
    for (hipEvent_t *e : {&rec_done_[0], &rec_done_[1], &adj_done_[0],
                          &adj_done_[1], &sync_ev_[0], &sync_ev_[1]})
      SMX_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    part_adj_ = alloc<real>(2 * std::max<size_t>(1, A.sh_list.size()));
    std::vector<int32_t> bidx(size_t(nglob_pad_), -1);
    for (size_t j = 0; j < A.bnode_g.size(); ++j)
      bidx[A.bnode_g[j]] = int32_t(j);
    bidx_ = upload_raw<int>(bidx);
    bsum_ = alloc<real>(A.bsum.size());
    copy_to(bsum_, sizeof(real) == 8 ? A.bsum64 : A.bsum);
    sh_rin_ = upload_raw<int>(A.sh_rin);
    rin_excl_ = upload_raw<int>(A.rin_excl);
    rin_node_ = upload_raw<int>(A.rin_node);
    rin_off_ = upload_raw<int>(A.rin_off);
    rin_ent_ = upload_raw<int>(A.rin_ent);
    gsrc_ = alloc<real>(std::max<size_t>(2, size_t(A.nstep) * A.nrec * 2));
    rb0_ = alloc<real>(n2_);
    rb1_ = alloc<real>(n2_);
    racc_ = alloc<real>(n2_);
    gd_ = alloc<real>(size_t(NPT) * A.nspec_pad);
    ge_ = alloc<real>(size_t(NPT) * A.nspec_pad);
    sp_ = alloc<real>(size_t(NPT) * A.nstacey_pad);
    ss_ = alloc<real>(size_t(NPT) * A.nstacey_pad);
  }
  // One fused step: tiles (sources + stiffness + assembly + finish of the
  // exclusive points), seismograms, shared points.
  template <int MODE> void fused_step(Field &F) {
    const HostAssembly2D &A = *asm_;
    const int ntypes = int(A.seis_types.size());
    const bool seis = A.nrec > 0 && ntypes > 0;
    if (has_pressure_) // reads the displacement before the tiles advance it
      launch_pressure(F.u, step_, 0);
    launch_tiles<MODE>(F.u, F.v, F.a, nullptr, 0, step_,
                       seis ? seis_terms_ : nullptr);
    // every term written by the tiles: the sums ride in k_shared's first
    // workgroup on the time steps of the graph
    const bool in_shared = seis && seis_in_tiles_ && MODE == MODE_STEP;
    if (seis && !in_shared) {
      const int n = A.nrec * ntypes;
      k_seis_fused<real><<<dim3((n + SEIS_GROUPS - 1) / SEIS_GROUPS),
                           dim3(32 * SEIS_GROUPS), 0, stream_>>>(
          F.u, F.v, minv_, part_, shp_off_, shp_pos_, rent_g_, rent_shared_,
          rent_h_, seis_terms_, A.nrec, stypes_, ntypes, A.nsbs, step_, dto2_,
          seis_);
      SMX_HIP(hipGetLastError());
    }
    launch_shared<MODE>(F.u, F.v, F.a, step_, nullptr, nullptr, in_shared);
  }
  void build_graph(Field &F) {
    capture(graph_, graph_exec_, stream_, [&] {
      for (int s = 0; s < gsteps_; ++s)
        fused_step<MODE_STEP>(F);
    });
  }

  std::unique_ptr<HostAssembly2D> asm_;
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipGraph_t graph_ = nullptr, rec_graph_[2] = {}, adj_graph_[2] = {};
  hipGraphExec_t graph_exec_ = nullptr, rec_graph_exec_[2] = {},
                 adj_graph_exec_[2] = {};
  std::vector<void *> allocs_;
  long nglob_pad_ = 0, n2_ = 0;
  TileInfo *tiles_ = nullptr;
  int *sh_list_ = nullptr, *ibool_ = nullptr, *sslot_ = nullptr,
      *src_li_ = nullptr, *src_idx_ = nullptr, *trec_li_ = nullptr,
      *trec_entry_ = nullptr, *rent_g_ = nullptr, *rent_shared_ = nullptr,
      *step_ = nullptr, *rec_slot_ = nullptr;
  bool has_pressure_ = false; // elastic pressure seismograms requested
  // every receiver term is written by the tiles (displacement and pressure
  // types only, no entry left to the shared-point pass)
  bool seis_in_tiles_ = false;
  // device-memory model: slot maps, and whether the host copy is stale
  int *elem_slot_ = nullptr, *bpt_p_ = nullptr, *bpt_bslot_ = nullptr,
      *bpt_slot_ = nullptr, *bnode_off_ = nullptr, *bnode_pts_ = nullptr,
      *model_bad_ = nullptr;
  bool host_model_stale_ = false;
  uint16_t *lidx_ = nullptr;
  uint8_t *ecolor_ = nullptr;
  real *geom_e_ = nullptr, *mat_e_ = nullptr;
  float *nodes_ = nullptr;
  int gmode_ = 0;
  long long geom_mismatch_ = 0;
  real *geom_ = nullptr, *mat_ = nullptr, *kmr_ = nullptr, *stacey_ = nullptr,
       *src_L_ = nullptr, *stf_ = nullptr, *rent_h_ = nullptr,
       *minv_ = nullptr, *part_ = nullptr, *seis_ = nullptr,
       *seis_terms_ = nullptr, *scratch_u_ = nullptr, *scratch_v_ = nullptr,
       *scratch_part_ = nullptr;
  int *shp_off_ = nullptr, *shp_pos_ = nullptr;
  // coupled meshes: entries of the fluid coupling and the fluid's
  // interface accelerations (owned by the acoustic solver)
  int *cpl_li_ = nullptr, *cpl_k_ = nullptr;
  real *cpl_w_ = nullptr;
  const real *chi_dd_ = nullptr;
  size_t lds_bytes_ = 0;
  int *sh_pad_ = nullptr; // [ntiles][sh_stride_] global ids of shared points
  int sh_stride_ = 1;
  Field fields_[3];
  QuadArgs<real> quad_;
  SeisTypes stypes_;
  real dt_ = 0, dto2_ = 0, dt2o2_ = 0;
  bool initialized_ = false, have_stats_ = false;
  double last_stiff_ms_ = 0;
  int last_launches_ = 0;
  // adjoint: checkpoints [nseg][u, v], the recompute record (every+1 slots
  // of u~, v~, a'), adjoint rb in two parity buffers, gradient accumulators
  real *ckpt_ = nullptr, *rec_[2] = {}, *part_adj_ = nullptr, *rb0_ = nullptr,
       *rb1_ = nullptr,
       *racc_ = nullptr, *gd_ = nullptr, *ge_ = nullptr, *sp_ = nullptr,
       *ss_ = nullptr, *bsum_ = nullptr, *gsrc_ = nullptr;
  int *bidx_ = nullptr, *sh_rin_ = nullptr, *rin_excl_ = nullptr,
      *rin_node_ = nullptr, *rin_off_ = nullptr, *rin_ent_ = nullptr;
  int ckpt_cap_ = 0, rec_every_ = 0, ckpt_every_ = 0;
  hipEvent_t adj_ev_[4] = {nullptr, nullptr, nullptr, nullptr};
  // adjoint stream (default priority) and its ordering events
  hipStream_t adj_stream_ = nullptr;
  hipEvent_t rec_done_[2] = {}, adj_done_[2] = {}, sync_ev_[2] = {};
  bool have_adj_stats_ = false, have_gradient_ = false;
};

void check_launch(int graph_steps, int waves_per_eu, int precision) {
  if (graph_steps < 1 || graph_steps > 256)
    throw SmxError(SMX_ERR_ARG, "graph_steps must lie in [1, 256]");
  if (waves_per_eu != 0 && waves_per_eu != 6 && waves_per_eu != 7 &&
      waves_per_eu != 8)
    throw SmxError(SMX_ERR_ARG, "waves_per_eu must be 0 (default), 6, 7 or 8");
  if (waves_per_eu != 0 && precision != 4)
    throw SmxError(SMX_ERR_ARG,
                   "waves_per_eu variants exist for FP32 contexts only");
}

std::unique_ptr<Solver2DBase> make_solver2d(std::unique_ptr<HostAssembly2D> a,
                                            int precision, int graph_steps,
                                            bool nodes_geometry,
                                            int waves_per_eu) {
  check_launch(graph_steps, 0, precision);
  if (a->medium == 1) {
    if (waves_per_eu != 0)
      throw SmxError(SMX_ERR_ARG,
                     "waves_per_eu variants exist for 2D elastic contexts only");
    return make_acoustic_solver2d(std::move(a), precision, graph_steps);
  }
  check_launch(graph_steps, waves_per_eu, precision);
  if (precision == 4)
    return std::make_unique<Solver2D<float>>(std::move(a), graph_steps,
                                             nodes_geometry, waves_per_eu);
  if (precision == 8)
    return std::make_unique<Solver2D<double>>(std::move(a), graph_steps,
                                              nodes_geometry, 0);
  throw SmxError(SMX_ERR_ARG, "precision must be 4 or 8");
}

} // namespace smx
