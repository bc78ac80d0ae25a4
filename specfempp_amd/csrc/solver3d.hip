// gfx950 kernels and the device-resident solver for the 3D elastic isotropic
// time step (hex8, NGLL = 5).
//
// Kernels (each cites the reference dispatch it replaces):
//   k3_tile     compute_stiffness_interaction for dim3
//               (include/kokkos_kernels/impl/compute_stiffness_interaction.tpp:
//               140-247 with the dim3 gradient, stress and divergence,
//               algorithms/gradient.hpp, medium/dim3/elastic/isotropic/
//               stress.hpp:46-99, point/stress.hpp:217-240,
//               algorithms/divergence.hpp:95-137), plus this build's 3D Stacey
//               faces, the source interaction (compute_source_interaction.tpp:
//               17-104), and for the tile's exclusive points divide_mass_matrix
//               + Newmark corrector + next predictor (divide_mass_matrix.tpp:
//               13-65, newmark.tpp:13-128).  One workgroup per tile of
//               elements: the tile's accumulator lives in LDS; its elements
//               are processed colour by colour in rounds of up to four
//               elements (128 threads each, one per GLL point), so that no two
//               elements of a round share a point and the assembly needs no
//               atomics.  Shared points leave one partial per tile.
//   k3_shared   the same finish for shared points: their partials summed in
//               tile order.
//   k3_seis     compute_seismograms (compute_seismogram.tpp:19-171).
// A step is k3_seis (displacement), k3_tile, k3_shared, k3_seis (velocity /
// acceleration) and the step counter; steps are captured into HIP graphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "solver3d.hpp"

namespace smx {

#define SMX3_HIP(call)                                                         \
  do {                                                                         \
    hipError_t err_ = (call);                                                  \
    if (err_ != hipSuccess)                                                    \
      throw SmxError(SMX_ERR_HIP, std::string(#call) + ": " +                  \
                                      hipGetErrorString(err_));                \
  } while (0)

namespace {

constexpr int ELEM_THREADS = 128; // >= NPT3 = 125
constexpr int NODE3_THREADS = 256;
constexpr int kGraphSteps3 = 8;

template <typename real> struct Quad3 {
  real hp[NGLL][NGLL]; // hp[i][j] = l'_j(xi_i)
  real w[NGLL];
};

struct Types3 {
  int t[4];
};

__device__ __forceinline__ int p3(int iz, int iy, int ix) {
  return (iz * NGLL + iy) * NGLL + ix;
}

// Where an element pass takes its geometry from.
enum GeomMode : int {
  GEOM_POINT = 0, // stored per GLL point
  GEOM_ELEM = 1,  // stored per element (affine elements)
  GEOM_NODES = 2  // re-formed from the 8 control nodes at every point
};

// The host assembly's per-point geometry, re-formed from the element's
// control nodes with the same FP32 operations in the same order and no
// contraction (shape-function derivatives in double rounded to float,
// core/specfem/shape_function/dim3/shape_function_derivatives3d.cpp:9-40;
// Jacobian and inverse in float, core/specfem/jacobian/dim3/jacobian.cpp:
// 27-74).  The solver checks at set-up that it reproduces the stored values
// bit for bit before using it.
__device__ __noinline__ void geom_from_nodes3(const float *__restrict__ nd,
                                              float xi, float et, float ga,
                                              float G[10]) {
#pragma clang fp contract(off)
  const double x = xi, e = et, g = ga;
  float xxi = 0, yxi = 0, zxi = 0, xet = 0, yet = 0, zet = 0, xga = 0,
        yga = 0, zga = 0;
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const double sx = (n == 1 || n == 2 || n == 5 || n == 6) ? 1.0 : -1.0;
    const double se = (n == 2 || n == 3 || n == 6 || n == 7) ? 1.0 : -1.0;
    const double sg = n >= 4 ? 1.0 : -1.0;
    const float d0 = static_cast<float>(0.125 * sx * (1.0 + se * e) * (1.0 + sg * g));
    const float d1 = static_cast<float>(0.125 * se * (1.0 + sx * x) * (1.0 + sg * g));
    const float d2 = static_cast<float>(0.125 * sg * (1.0 + sx * x) * (1.0 + se * e));
    const float cx = nd[n], cy = nd[8 + n], cz = nd[16 + n];
    xxi += d0 * cx;
    yxi += d0 * cy;
    zxi += d0 * cz;
    xet += d1 * cx;
    yet += d1 * cy;
    zet += d1 * cz;
    xga += d2 * cx;
    yga += d2 * cy;
    zga += d2 * cz;
  }
  const float J = xxi * (yet * zga - yga * zet) - xet * (yxi * zga - yga * zxi) +
                  xga * (yxi * zet - yet * zxi);
  G[0] = (yet * zga - yga * zet) / J;
  G[1] = (xga * zet - xet * zga) / J;
  G[2] = (xet * yga - xga * yet) / J;
  G[3] = (yga * zxi - yxi * zga) / J;
  G[4] = (xxi * zga - xga * zxi) / J;
  G[5] = (xga * yxi - xxi * yga) / J;
  G[6] = (yxi * zet - yet * zxi) / J;
  G[7] = (xet * zxi - xxi * zet) / J;
  G[8] = (xxi * yet - xet * yxi) / J;
  G[9] = J;
}

// Set-up check: 1 where the re-formed geometry of (element, point) differs
// from the stored FP32 value in any bit.
__global__ void k3_geom_check(const float *__restrict__ nodes,
                              const float *__restrict__ geom, int nspec,
                              const float *__restrict__ xi,
                              unsigned char *__restrict__ bad) {
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(nspec) * NPT3)
    return;
  const int slot = int(t / NPT3), p = int(t % NPT3);
  const int iz = p / (NGLL * NGLL), iy = (p / NGLL) % NGLL, ix = p % NGLL;
  float G[10];
  geom_from_nodes3(nodes + size_t(slot) * 24, xi[ix], xi[iy], xi[iz], G);
  bool diff = false;
  for (int k = 0; k < 10; ++k)
    diff = diff || (__float_as_uint(G[k]) !=
                    __float_as_uint(geom[(size_t(slot) * 10 + k) * NPT3 + p]));
  bad[t] = diff ? 1 : 0;
}

// Record buffers of the adjoint's recompute: per step j of a segment,
// a' [every][3][npad], the predicted u~ [every][3][npad] and v~ at the
// absorbing nodes [every][3][nb].
template <typename real> struct Rec3 {
  real *a, *u, *vb;
  const int *bidx; // node -> absorbing-node index or -1
  int every;
  long npad, nb;
};

// Modes of the tile and shared-point kernels.
enum Mode3 : int {
  M3_STEP = 0,     // finish the step and apply the next predictor
  M3_FINAL = 1,    // finish the step (corrected v and a stored)
  M3_RECORD = 2,   // as M3_STEP, and record a', u~, v~ for the adjoint
  M3_ASSEMBLE = 3  // a = (f - K u - B v) / M only (stiffness timing)
};

// Sampled-step outputs for velocity / acceleration seismograms (v, a null
// when no such seismogram is recorded).
template <typename real> struct Aux3 {
  real *v, *a;
  int nsbs;
};

// Finish one (point, component), i = c * npad + g, from its assembled force
// acc and its loaded 1/M, v and u: a = acc / M, v += dt/2 a, then the next
// predictor u += dt v + dt^2/2 a, v += dt/2 a (divide_mass_matrix.tpp:13-65,
// newmark.tpp:13-128).
template <typename real, int MODE>
__device__ __forceinline__ void finish_loaded(long i, long n3, real acc,
                                              real minv, real v, real u,
                                              real *__restrict__ U,
                                              real *__restrict__ V,
                                              real *__restrict__ A,
                                              real dt, real dto2, real dt2o2,
                                              const Rec3<real> &rec, int j,
                                              const Aux3<real> &aux,
                                              bool sample) {
  const real a = acc * minv;
  if (MODE == M3_ASSEMBLE) {
    A[i] = a;
    return;
  }
  v += dto2 * a;
  if (sample) {
    aux.v[i] = v;
    aux.a[i] = a;
  }
  if (MODE == M3_RECORD)
    rec.a[size_t(j) * n3 + i] = a;
  if (MODE == M3_FINAL) {
    V[i] = v;
    A[i] = a;
    return;
  }
  u += dt * v + dt2o2 * a;
  v += dto2 * a;
  U[i] = u;
  V[i] = v;
  if (MODE == M3_RECORD && j + 1 < rec.every) {
    rec.u[size_t(j + 1) * n3 + i] = u;
    const long g = i % rec.npad;
    const int c = int(i / rec.npad);
    const int b = rec.bidx[g];
    if (b >= 0)
      rec.vb[(size_t(j + 1) * 3 + c) * rec.nb + b] = v;
  }
}

// One tile of elements (see the file comment).  GM: where the geometry
// comes from (GeomMode); ME: moduli per element instead of per GLL point;
// RW: elements per round (block = RW x 128 threads).  LDS: the tile's
// accumulator [3][nloc_max], then per round slot the element displacement
// [3][125] and stress integrand [3][3][125], then the global ids of the
// tile's shared points [nsh_max] and its round slots [nround_max][RW]
// (int), so that no global load of a round waits on another.
template <typename real, int MODE, int GM, bool ME, int RW>
__global__ void __launch_bounds__(RW * ELEM_THREADS)
    k3_tile(const TileInfo3 *__restrict__ tiles,
            const int *__restrict__ round_slots,
            const uint16_t *__restrict__ lidx, const int *__restrict__ sh_list,
            const real *__restrict__ geom, const float *__restrict__ nodes,
            Quad3<float> xi3, const real *__restrict__ mat,
            const int *__restrict__ sslot, const real *__restrict__ stacey,
            real *__restrict__ U, real *__restrict__ V, real *__restrict__ A,
            const real *__restrict__ Minv, real *__restrict__ part, long npart,
            const int *__restrict__ tsrc_li, const int *__restrict__ tsrc_id,
            const real *__restrict__ tsrc_val, const real *__restrict__ stf,
            int nsrc, int nloc_max, int nsh_max, long npad, real dt, real dto2,
            real dt2o2, Quad3<real> q, const int *__restrict__ step_ptr,
            Rec3<real> rec, Aux3<real> aux) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  real *s_a = reinterpret_cast<real *>(lds_raw);
  const TileInfo3 T = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int sub = tid / ELEM_THREADS, p = tid % ELEM_THREADS;
  const bool pact = p < NPT3;
  const int pp = pact ? p : 0;
  const int iz = pp / (NGLL * NGLL), iy = (pp / NGLL) % NGLL, ix = pp % NGLL;
  real *se = s_a + 3 * size_t(nloc_max) + size_t(sub) * 12 * NPT3;
  real *sf = se + 3 * NPT3;
  int *s_sh = reinterpret_cast<int *>(s_a + 3 * size_t(nloc_max) +
                                      size_t(RW) * 12 * NPT3);
  int *s_slot = s_sh + nsh_max;
  const int nloc = T.nex + T.nsh;
  const int step = MODE == M3_ASSEMBLE ? 0 : *step_ptr;
  for (int i = tid; i < nloc; i += RW * ELEM_THREADS) {
    s_a[i] = 0;
    s_a[nloc_max + i] = 0;
    s_a[2 * nloc_max + i] = 0;
  }
  for (int i = tid; i < T.nsh; i += RW * ELEM_THREADS)
    s_sh[i] = sh_list[T.sh_off + i];
  for (int i = tid; i < T.nround * RW; i += RW * ELEM_THREADS)
    s_slot[i] = round_slots[size_t(T.round_off) * RW + i];
  __syncthreads();
  // point sources first, in list order (one thread: fixed order)
  if (MODE != M3_ASSEMBLE && tid == 0)
    for (int k = T.src_off; k < T.src_off + T.nsrc; ++k) {
      const int l = tsrc_li[k];
      const real f = stf[size_t(step) * nsrc + tsrc_id[k]];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        s_a[c * nloc_max + l] += tsrc_val[size_t(k) * 3 + c] * f;
    }
  __syncthreads();

  // Rounds are software-pipelined: the displacement, geometry and moduli of
  // round r + 1 are requested while round r computes, and the tile-local
  // point ids of round r + 2 (the global id then comes from LDS), so each
  // global load depends on nothing fetched in the same round (no point of
  // the tile is written before all rounds are done, and shared points only
  // by the next kernel, so the early reads see this step's values).
  auto slot_of = [&](int r) { return r < T.nround ? s_slot[r * RW + sub] : -1; };
  int l_pre = 0; // tile-local id of this thread's point in the next fetch
  auto fetch_l = [&](int r) {
    const int s = slot_of(r);
    if (pact && s >= 0)
      l_pre = lidx[size_t(s) * NPT3 + p];
  };
  int slot_n = -1, l_n = 0;
  long g_n = 0;
  real u_n[3] = {0, 0, 0}, G_n[10], L_n = 0, M_n = 0;
  auto fetch = [&](int r) {
    slot_n = slot_of(r);
    if (!(pact && slot_n >= 0))
      return;
    l_n = l_pre;
    g_n = l_n < T.nex ? long(T.ex_off) + l_n : long(s_sh[l_n - T.nex]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      u_n[c] = U[c * npad + g_n];
    if (GM != GEOM_NODES) {
#pragma unroll
      for (int k = 0; k < 10; ++k)
        G_n[k] = GM == GEOM_ELEM ? geom[size_t(slot_n) * 10 + k]
                                 : geom[(size_t(slot_n) * 10 + k) * NPT3 + p];
    }
    L_n = ME ? mat[size_t(slot_n) * 2] : mat[size_t(slot_n) * 2 * NPT3 + p];
    M_n = ME ? mat[size_t(slot_n) * 2 + 1] : mat[(size_t(slot_n) * 2 + 1) * NPT3 + p];
  };
  fetch_l(0);
  fetch(0);
  fetch_l(1);
  for (int r = 0; r < T.nround; ++r) {
    const int slot = slot_n;
    const bool act = pact && slot >= 0;
    const int l = l_n;
    const long g = g_n;
    real G[10];
#pragma unroll
    for (int k = 0; k < 10; ++k)
      G[k] = G_n[k];
    const real L = L_n, M = M_n;
    if (act)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        se[c * NPT3 + p] = u_n[c];
    __syncthreads();
    fetch(r + 1);
    fetch_l(r + 2);
    if (act) {
      real dxi[3], det[3], dga[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        real a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
        for (int k = 0; k < NGLL; ++k) {
          a0 += q.hp[ix][k] * se[c * NPT3 + p3(iz, iy, k)];
          a1 += q.hp[iy][k] * se[c * NPT3 + p3(iz, k, ix)];
          a2 += q.hp[iz][k] * se[c * NPT3 + p3(k, iy, ix)];
        }
        dxi[c] = a0;
        det[c] = a1;
        dga[c] = a2;
      }
      if (GM == GEOM_NODES) {
        float Gf[10];
        geom_from_nodes3(nodes + size_t(slot) * 24, xi3.w[ix], xi3.w[iy],
                         xi3.w[iz], Gf);
#pragma unroll
        for (int k = 0; k < 10; ++k)
          G[k] = Gf[k];
      }
      real du[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        du[c][0] = G[0] * dxi[c] + G[3] * det[c] + G[6] * dga[c];
        du[c][1] = G[1] * dxi[c] + G[4] * det[c] + G[7] * dga[c];
        du[c][2] = G[2] * dxi[c] + G[5] * det[c] + G[8] * dga[c];
      }
      const real lam = L - real(2.0) * M;
      real Ts[3][3];
      Ts[0][0] = L * du[0][0] + lam * (du[1][1] + du[2][2]);
      Ts[1][1] = L * du[1][1] + lam * (du[0][0] + du[2][2]);
      Ts[2][2] = L * du[2][2] + lam * (du[0][0] + du[1][1]);
      Ts[0][1] = Ts[1][0] = M * (du[0][1] + du[1][0]);
      Ts[0][2] = Ts[2][0] = M * (du[0][2] + du[2][0]);
      Ts[1][2] = Ts[2][1] = M * (du[1][2] + du[2][1]);
      const real J = G[9];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sf[(c * 3 + 0) * NPT3 + p] =
            J * (Ts[c][0] * G[0] + Ts[c][1] * G[1] + Ts[c][2] * G[2]);
        sf[(c * 3 + 1) * NPT3 + p] =
            J * (Ts[c][0] * G[3] + Ts[c][1] * G[4] + Ts[c][2] * G[5]);
        sf[(c * 3 + 2) * NPT3 + p] =
            J * (Ts[c][0] * G[6] + Ts[c][1] * G[7] + Ts[c][2] * G[8]);
      }
    }
    __syncthreads();
    if (act) {
      real rr[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        real t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
        for (int k = 0; k < NGLL; ++k) {
          t1 += sf[(c * 3 + 0) * NPT3 + p3(iz, iy, k)] * q.hp[k][ix] * q.w[k];
          t2 += sf[(c * 3 + 1) * NPT3 + p3(iz, k, ix)] * q.hp[k][iy] * q.w[k];
          t3 += sf[(c * 3 + 2) * NPT3 + p3(k, iy, ix)] * q.hp[k][iz] * q.w[k];
        }
        rr[c] = q.w[iz] * q.w[iy] * t1 + q.w[iz] * q.w[ix] * t2 +
                q.w[iy] * q.w[ix] * t3;
        rr[c] *= real(-1.0);
      }
      const int b = sslot[slot];
      if (b >= 0) {
        const real *st = stacey + size_t(b) * 7 * NPT3 + p;
        const real wgt = st[4 * NPT3];
        if (wgt != real(0)) {
          const real n[3] = {st[0], st[NPT3], st[2 * NPT3]};
          const real nrm = st[3 * NPT3], P = st[5 * NPT3], S = st[6 * NPT3];
          const real v[3] = {V[g], V[npad + g], V[2 * npad + g]};
          const real vn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const real fc = ((vn * n[c] / (nrm * nrm)) * (P - S)) + v[c] * S;
            rr[c] += real(-1.0) * fc * nrm * wgt;
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
        s_a[c * nloc_max + l] += rr[c];
    }
    __syncthreads();
  }

  // exclusive points: finish the step; shared points: store the partials
  const long n3 = 3 * npad;
  const bool sample = MODE != M3_ASSEMBLE && aux.v != nullptr && step % aux.nsbs == 0;
  const int j = MODE == M3_RECORD ? step % rec.every : 0;
  // up to kFin exclusive points per thread: all their loads are requested
  // before the first is used (the round registers are free by now)
  constexpr int kFin = 4;
  for (int i0 = tid; i0 < T.nex; i0 += kFin * RW * ELEM_THREADS) {
    real mv[kFin][3], vv[kFin][3], uv[kFin][3];
#pragma unroll
    for (int k = 0; k < kFin; ++k) {
      const int i = i0 + k * RW * ELEM_THREADS;
      if (i < T.nex) {
        const long g = long(T.ex_off) + i;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          mv[k][c] = Minv[c * npad + g];
          vv[k][c] = MODE == M3_ASSEMBLE ? real(0) : V[c * npad + g];
          uv[k][c] = (MODE == M3_STEP || MODE == M3_RECORD) ? U[c * npad + g]
                                                            : real(0);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kFin; ++k) {
      const int i = i0 + k * RW * ELEM_THREADS;
      if (i < T.nex) {
        const long g = long(T.ex_off) + i;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          finish_loaded<real, MODE>(c * npad + g, n3, s_a[c * nloc_max + i],
                                    mv[k][c], vv[k][c], uv[k][c], U, V, A, dt,
                                    dto2, dt2o2, rec, j, aux, sample);
      }
    }
  }
  for (int i = tid; i < T.nsh; i += RW * ELEM_THREADS)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      part[c * npart + T.sh_off + i] = s_a[c * nloc_max + T.nex + i];
}

// Shared points (global ids nex_total + s): the tiles' partials summed in
// tile order, then the same finish as the tiles' exclusive points.
template <typename real, int MODE>
__global__ void __launch_bounds__(NODE3_THREADS)
    k3_shared(int nsh, int nex_total, const int *__restrict__ sh_poff,
              const int *__restrict__ sh_pidx, const real *__restrict__ part,
              long npart, real *__restrict__ U, real *__restrict__ V,
              real *__restrict__ A, const real *__restrict__ Minv, long npad,
              real dt, real dto2, real dt2o2, int *__restrict__ step_ptr,
              Rec3<real> rec, Aux3<real> aux, int advance) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsh)
    return;
  // advance: this pass moves the step counter on (no thread of it reads the
  // counter then: only the recording mode and the velocity / acceleration
  // samples need the step)
  const bool need_step =
      MODE == M3_RECORD || (MODE != M3_ASSEMBLE && aux.v != nullptr);
  const int step = need_step ? *step_ptr : 0;
  if (advance && s == 0)
    *step_ptr += 1;
  const bool sample = MODE != M3_ASSEMBLE && aux.v != nullptr && step % aux.nsbs == 0;
  const int j = MODE == M3_RECORD ? step % rec.every : 0;
  const long g = long(nex_total) + s;
  const long n3 = 3 * npad;
  // the point's own 1/M, v and u first (they do not wait on the partials),
  // then the partials of the three components in tile order
  real mv[3], vv[3], uv[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    mv[c] = Minv[c * npad + g];
    vv[c] = MODE == M3_ASSEMBLE ? real(0) : V[c * npad + g];
    uv[c] = (MODE == M3_STEP || MODE == M3_RECORD) ? U[c * npad + g] : real(0);
  }
  const int k0 = sh_poff[s], k1 = sh_poff[s + 1];
  real acc[3] = {0, 0, 0};
  for (int k = k0; k < k1; ++k) {
    const int q = sh_pidx[k];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      acc[c] += part[c * npart + q];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c)
    finish_loaded<real, MODE>(c * npad + g, n3, acc[c], mv[c], vv[c], uv[c], U,
                              V, A, dt, dto2, dt2o2, rec, j, aux, sample);
}

// Seismograms of step `*step_ptr` (if sampled).  POST = false, before the
// step's tile pass: displacement types from U (= u~ of the step).  POST =
// true, after the shared-point pass: velocity / acceleration types from the
// step's corrected values (aux).  One wavefront per (type, receiver): lane l
// sums points l and l + 64 in order, then a fixed butterfly reduction (lane
// 0's order of additions is fixed, so the result is deterministic).  Grid:
// exactly nrec * ntypes blocks of 64 threads.
template <typename real, bool POST>
__global__ void __launch_bounds__(64)
    k3_seis(const real *__restrict__ U, const real *__restrict__ auxv,
            const real *__restrict__ auxa, const int *__restrict__ rec_g,
            const real *__restrict__ rec_h, int nrec, Types3 types, int ntypes,
            int nsbs, long npad, const int *__restrict__ step_ptr,
            real *__restrict__ out) {
  const int step = *step_ptr;
  if (step % nsbs != 0)
    return;
  const int isig = step / nsbs;
  const int r = blockIdx.x % nrec, it = blockIdx.x / nrec;
  const int type = types.t[it];
  if ((type != 0) != POST)
    return;
  const real *F = type == 0 ? U : (type == 1 ? auxv : auxa);
  const int lane = threadIdx.x;
  real s[3] = {0, 0, 0};
  for (int p = lane; p < NPT3; p += 64) {
    const int g = rec_g[size_t(r) * NPT3 + p];
    const real h = rec_h[size_t(r) * NPT3 + p];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      s[c] += h * F[c * npad + g];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      s[c] += __shfl_xor(s[c], off);
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      out[((size_t(isig) * ntypes + it) * nrec + r) * 3 + c] = s[c];
}

// v~ at the absorbing nodes into record slot 0 (segment start).
template <typename real>
__global__ void k3_gather_vb(const real *__restrict__ V,
                             const int *__restrict__ bnode, long nb, long npad,
                             real *__restrict__ vb0) {
  const long b = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nb)
    return;
  const int g = bnode[b];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    vb0[c * nb + b] = V[c * npad + g];
}

__global__ void k3_step_add(int *__restrict__ step_ptr, int d) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *step_ptr += d;
}

// Adjoint misfit source of step `*adj_step`: ub += H^T g_k at every node a
// receiver touches, each node summing its (receiver, weight) entries in
// order (no atomics; receivers may share nodes).
template <typename real>
__global__ void k3_adj_inject(const int *__restrict__ rin_node,
                              const int *__restrict__ rin_off,
                              const int *__restrict__ rin_r,
                              const real *__restrict__ rin_h, int nnode,
                              const real *__restrict__ gsrc, int nrec,
                              real *__restrict__ Ub, long npad,
                              const int *__restrict__ adj_step) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode)
    return;
  const real *gk = gsrc + size_t(*adj_step) * nrec * 3;
  real s[3] = {0, 0, 0};
  for (int e = rin_off[t]; e < rin_off[t + 1]; ++e) {
    const int r = rin_r[e];
    const real h = rin_h[e];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      s[c] += h * gk[size_t(r) * 3 + c];
  }
  const int g = rin_node[t];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    Ub[c * npad + g] += s[c];
}

// Adjoint colour pass of step k = *adj_step over the elements `slots` of
// one colour (one workgroup per element): with rb the adjoint acceleration
// field and u~_k the recorded forward displacement, assembles -K rb (as
// k3_tile assembles -K u, without the Stacey term) and
// accumulates the material densities div(rb) div(u~) (Gd) and
// eps(rb) : eps(u~) (Ge) per local point.
template <typename real, int GM, bool ME>
__global__ void __launch_bounds__(ELEM_THREADS)
    k3_adj_elem(const int *__restrict__ slots, const int *__restrict__ ibool,
                const real *__restrict__ geom, const float *__restrict__ nodes,
                Quad3<float> xi3, const real *__restrict__ mat,
                const real *__restrict__ Rb, const real *__restrict__ rec_u,
                int every, const int *__restrict__ adj_step,
                real *__restrict__ Kacc, real *__restrict__ Gd,
                real *__restrict__ Ge, long npad, Quad3<real> q) {
  __shared__ real s_r[3][NPT3];
  __shared__ real s_u[3][NPT3];
  __shared__ real s_f[3][3][NPT3];
  const int slot = slots[blockIdx.x];
  const int t = threadIdx.x;
  const bool act = t < NPT3;
  const int p = act ? t : 0;
  const int iz = p / (NGLL * NGLL), iy = (p / NGLL) % NGLL, ix = p % NGLL;
  const real *Ut = rec_u + size_t(*adj_step % every) * 3 * npad;
  int g = 0;
  if (act) {
    g = ibool[size_t(slot) * NPT3 + p];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s_r[c][p] = Rb[c * npad + g];
      s_u[c][p] = Ut[c * npad + g];
    }
  }
  __syncthreads();
  if (act) {
    real G[10];
    if (GM == GEOM_NODES) {
      float Gf[10];
      geom_from_nodes3(nodes + size_t(slot) * 24, xi3.w[ix], xi3.w[iy],
                       xi3.w[iz], Gf);
#pragma unroll
      for (int k = 0; k < 10; ++k)
        G[k] = Gf[k];
    } else {
#pragma unroll
      for (int k = 0; k < 10; ++k)
        G[k] = GM == GEOM_ELEM ? geom[size_t(slot) * 10 + k]
                               : geom[(size_t(slot) * 10 + k) * NPT3 + p];
    }
    real dr[3][3], du[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      real a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
      for (int l = 0; l < NGLL; ++l) {
        a0 += q.hp[ix][l] * s_r[c][p3(iz, iy, l)];
        a1 += q.hp[iy][l] * s_r[c][p3(iz, l, ix)];
        a2 += q.hp[iz][l] * s_r[c][p3(l, iy, ix)];
        b0 += q.hp[ix][l] * s_u[c][p3(iz, iy, l)];
        b1 += q.hp[iy][l] * s_u[c][p3(iz, l, ix)];
        b2 += q.hp[iz][l] * s_u[c][p3(l, iy, ix)];
      }
      dr[c][0] = G[0] * a0 + G[3] * a1 + G[6] * a2;
      dr[c][1] = G[1] * a0 + G[4] * a1 + G[7] * a2;
      dr[c][2] = G[2] * a0 + G[5] * a1 + G[8] * a2;
      du[c][0] = G[0] * b0 + G[3] * b1 + G[6] * b2;
      du[c][1] = G[1] * b0 + G[4] * b1 + G[7] * b2;
      du[c][2] = G[2] * b0 + G[5] * b1 + G[8] * b2;
    }
    const real divr = dr[0][0] + dr[1][1] + dr[2][2];
    const real divu = du[0][0] + du[1][1] + du[2][2];
    const real ee = dr[0][0] * du[0][0] + dr[1][1] * du[1][1] +
                    dr[2][2] * du[2][2] +
                    real(0.5) * ((dr[0][1] + dr[1][0]) * (du[0][1] + du[1][0]) +
                                 (dr[0][2] + dr[2][0]) * (du[0][2] + du[2][0]) +
                                 (dr[1][2] + dr[2][1]) * (du[1][2] + du[2][1]));
    const size_t o = size_t(slot) * NPT3 + p;
    Gd[o] += divr * divu;
    Ge[o] += ee;
    const real L = ME ? mat[size_t(slot) * 2] : mat[size_t(slot) * 2 * NPT3 + p];
    const real M = ME ? mat[size_t(slot) * 2 + 1]
                      : mat[(size_t(slot) * 2 + 1) * NPT3 + p];
    const real lam = L - real(2.0) * M;
    real T[3][3];
    T[0][0] = L * dr[0][0] + lam * (dr[1][1] + dr[2][2]);
    T[1][1] = L * dr[1][1] + lam * (dr[0][0] + dr[2][2]);
    T[2][2] = L * dr[2][2] + lam * (dr[0][0] + dr[1][1]);
    T[0][1] = T[1][0] = M * (dr[0][1] + dr[1][0]);
    T[0][2] = T[2][0] = M * (dr[0][2] + dr[2][0]);
    T[1][2] = T[2][1] = M * (dr[1][2] + dr[2][1]);
    const real J = G[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s_f[c][0][p] = J * (T[c][0] * G[0] + T[c][1] * G[1] + T[c][2] * G[2]);
      s_f[c][1][p] = J * (T[c][0] * G[3] + T[c][1] * G[4] + T[c][2] * G[5]);
      s_f[c][2][p] = J * (T[c][0] * G[6] + T[c][1] * G[7] + T[c][2] * G[8]);
    }
  }
  __syncthreads();
  if (!act)
    return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    real t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      t1 += s_f[c][0][p3(iz, iy, l)] * q.hp[l][ix] * q.w[l];
      t2 += s_f[c][1][p3(iz, l, ix)] * q.hp[l][iy] * q.w[l];
      t3 += s_f[c][2][p3(l, iy, ix)] * q.hp[l][iz] * q.w[l];
    }
    const real r = q.w[iz] * q.w[iy] * t1 + q.w[iz] * q.w[ix] * t2 +
                   q.w[iy] * q.w[ix] * t3;
    Kacc[c * npad + g] += real(-1.0) * r;
  }
}

// Adjoint state of the nodal update (device pointers and sizes).
template <typename real> struct AdjState3 {
  real *ub, *vb, *rb;          // [3][npad]
  const real *minv;            // [3][npad]
  const real *rec_a, *rec_vb;  // record buffers (a' per step, v~ at absorbing nodes)
  int every;
  real *R;                     // [npad] rb . a' per node
  const int *bidx;             // node -> absorbing-node index or -1
  const real *bmat;            // [6][nb] Stacey matrix per absorbing node
  real *T9, *D3;               // [9][nb], [3][nb]
  long nb, npad;
  real dt, dto2, dt2o2;
};

// Adjoint update of node g at step k (the transpose of the forward step),
// given its assembled -K rb (kr): with rb = W (ab + dt/2 vb),
//   uu = ub - K rb,  vv = vb - B rb,
//   ub = uu,  vb = vv + dt uu,  ab = dt/2 vv + dt^2/2 uu,  rb = W (ab + dt/2 vb)
// (rb of step k-1 is formed here; ab is not kept).  Accumulates rb . a'_k
// per node (density term) and, at the absorbing nodes, rb v~_k^T (9) and
// rb * a'_k per component (3) for the impedance terms; B at a node is the
// sum of its face points' symmetric 3x3 Stacey matrices.  adj_load reads
// what the update needs of node g (all three components), adj_apply
// applies it, adj_finish does both.
template <typename real> struct AdjNode {
  real rb[3], a[3], ub[3], vb[3], minv[3], R;
  int b;
};

template <typename real>
__device__ __forceinline__ void adj_load(long g, int j, const AdjState3<real> &S,
                                         AdjNode<real> &n) {
  const long npad = S.npad;
  const real *ap = S.rec_a + size_t(j) * 3 * npad;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long i = c * npad + g;
    n.rb[c] = S.rb[i];
    n.a[c] = ap[i];
    n.ub[c] = S.ub[i];
    n.vb[c] = S.vb[i];
    n.minv[c] = S.minv[i];
  }
  n.R = S.R[g];
  n.b = S.bidx[g];
}

template <typename real>
__device__ __forceinline__ void adj_apply(long g, const real kr[3], int j,
                                          const AdjState3<real> &S,
                                          const AdjNode<real> &n) {
  const long npad = S.npad, nb = S.nb;
  const real *rb = n.rb, *a = n.a;
  S.R[g] = n.R + (rb[0] * a[0] + rb[1] * a[1] + rb[2] * a[2]);
  real br[3] = {0, 0, 0};
  const int b = n.b;
  if (b >= 0) {
    const real *bm = S.bmat;
    const real mxx = bm[b], myy = bm[nb + b], mzz = bm[2 * nb + b],
               mxy = bm[3 * nb + b], mxz = bm[4 * nb + b], myz = bm[5 * nb + b];
    br[0] = mxx * rb[0] + mxy * rb[1] + mxz * rb[2];
    br[1] = mxy * rb[0] + myy * rb[1] + myz * rb[2];
    br[2] = mxz * rb[0] + myz * rb[1] + mzz * rb[2];
    const real *vt = S.rec_vb + size_t(j) * 3 * nb;
    const real v[3] = {vt[b], vt[nb + b], vt[2 * nb + b]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        S.T9[(c * 3 + d) * nb + b] += rb[c] * v[d];
      S.D3[c * nb + b] += rb[c] * a[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long i = c * npad + g;
    const real uu = n.ub[c] + kr[c];
    const real vv = n.vb[c] - br[c];
    const real vbn = vv + S.dt * uu;
    const real abn = S.dto2 * vv + S.dt2o2 * uu;
    S.ub[i] = uu;
    S.vb[i] = vbn;
    S.rb[i] = n.minv[c] * (abn + S.dto2 * vbn);
  }
}

template <typename real>
__device__ __forceinline__ void adj_finish(long g, const real kr[3], int j,
                                           const AdjState3<real> &S) {
  AdjNode<real> n;
  adj_load(g, j, S, n);
  adj_apply(g, kr, j, S, n);
}

// Adjoint node pass after the colour passes: every node with its assembled
// -K rb in Kacc (then cleared for the next step).
template <typename real>
__global__ void __launch_bounds__(NODE3_THREADS)
    k3_adj_node(real *__restrict__ Kacc, AdjState3<real> S, int nglob,
                const int *__restrict__ adj_step) {
  const long g = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nglob)
    return;
  const int j = *adj_step % S.every;
  real kr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    kr[c] = Kacc[c * S.npad + g];
  adj_finish(g, kr, j, S);
#pragma unroll
  for (int c = 0; c < 3; ++c)
    Kacc[c * S.npad + g] = 0;
}

// Adjoint step k = *adj_step as element tiles (the layout of k3_tile): per
// round, each element gathers rb and the recorded u~_k, forms both
// gradients, accumulates the densities div(rb) div(u~) (Gd) and
// eps(rb) : eps(u~) (Ge) per local point and assembles -K rb (without the
// Stacey term, which adj_finish applies as B rb) into the tile's LDS
// accumulator; the tile's exclusive nodes are then updated in place
// (adj_finish) and its shared nodes get partials.  LDS: accumulator
// [3][nloc_max], then per round slot rb [3][125], u~ [3][125] and the
// stress integrand [3][3][125], then the global ids of the tile's shared
// points [nsh_max] and its round slots [nround_max][RW] (int); each
// thread's point id of the next round is requested one round ahead, so the
// gather of a round waits on one global load.
template <typename real, int GM, bool ME, int RW>
__global__ void __launch_bounds__(RW * ELEM_THREADS)
    k3_adj_tile(const TileInfo3 *__restrict__ tiles,
                const int *__restrict__ round_slots,
                const uint16_t *__restrict__ lidx,
                const int *__restrict__ sh_list, const real *__restrict__ geom,
                const float *__restrict__ nodes, Quad3<float> xi3,
                const real *__restrict__ mat, const real *__restrict__ rec_u,
                real *__restrict__ part, long npart, real *__restrict__ Gd,
                real *__restrict__ Ge, int nloc_max, int nsh_max, Quad3<real> q,
                const int *__restrict__ adj_step, AdjState3<real> S) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  real *s_a = reinterpret_cast<real *>(lds_raw);
  const TileInfo3 T = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int sub = tid / ELEM_THREADS, p = tid % ELEM_THREADS;
  const bool pact = p < NPT3;
  const int pp = pact ? p : 0;
  const int iz = pp / (NGLL * NGLL), iy = (pp / NGLL) % NGLL, ix = pp % NGLL;
  real *sr = s_a + 3 * size_t(nloc_max) + size_t(sub) * 15 * NPT3;
  real *su = sr + 3 * NPT3;
  real *sf = su + 3 * NPT3;
  int *s_sh = reinterpret_cast<int *>(s_a + 3 * size_t(nloc_max) +
                                      size_t(RW) * 15 * NPT3);
  int *s_slot = s_sh + nsh_max;
  const long npad = S.npad;
  const int nloc = T.nex + T.nsh;
  const int j = *adj_step % S.every;
  const real *Ut = rec_u + size_t(j) * 3 * npad;
  for (int i = tid; i < nloc; i += RW * ELEM_THREADS) {
    s_a[i] = 0;
    s_a[nloc_max + i] = 0;
    s_a[2 * nloc_max + i] = 0;
  }
  for (int i = tid; i < T.nsh; i += RW * ELEM_THREADS)
    s_sh[i] = sh_list[T.sh_off + i];
  for (int i = tid; i < T.nround * RW; i += RW * ELEM_THREADS)
    s_slot[i] = round_slots[size_t(T.round_off) * RW + i];
  __syncthreads();
  auto slot_of = [&](int r) { return r < T.nround ? s_slot[r * RW + sub] : -1; };
  int l_next = 0;
  {
    const int s0 = slot_of(0);
    if (pact && s0 >= 0)
      l_next = lidx[size_t(s0) * NPT3 + p];
  }
  for (int r = 0; r < T.nround; ++r) {
    const int slot = slot_of(r);
    const bool act = pact && slot >= 0;
    int l = 0;
    if (act) {
      l = l_next;
      const long g = l < T.nex ? long(T.ex_off) + l : long(s_sh[l - T.nex]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sr[c * NPT3 + p] = S.rb[c * npad + g];
        su[c * NPT3 + p] = Ut[c * npad + g];
      }
    }
    {
      const int s1 = slot_of(r + 1);
      if (pact && s1 >= 0)
        l_next = lidx[size_t(s1) * NPT3 + p];
    }
    __syncthreads();
    if (act) {
      real G[10];
      if (GM == GEOM_NODES) {
        float Gf[10];
        geom_from_nodes3(nodes + size_t(slot) * 24, xi3.w[ix], xi3.w[iy],
                         xi3.w[iz], Gf);
#pragma unroll
        for (int k = 0; k < 10; ++k)
          G[k] = Gf[k];
      } else {
#pragma unroll
        for (int k = 0; k < 10; ++k)
          G[k] = GM == GEOM_ELEM ? geom[size_t(slot) * 10 + k]
                                 : geom[(size_t(slot) * 10 + k) * NPT3 + p];
      }
      real dr[3][3], du[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        real a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
        for (int k = 0; k < NGLL; ++k) {
          a0 += q.hp[ix][k] * sr[c * NPT3 + p3(iz, iy, k)];
          a1 += q.hp[iy][k] * sr[c * NPT3 + p3(iz, k, ix)];
          a2 += q.hp[iz][k] * sr[c * NPT3 + p3(k, iy, ix)];
          b0 += q.hp[ix][k] * su[c * NPT3 + p3(iz, iy, k)];
          b1 += q.hp[iy][k] * su[c * NPT3 + p3(iz, k, ix)];
          b2 += q.hp[iz][k] * su[c * NPT3 + p3(k, iy, ix)];
        }
        dr[c][0] = G[0] * a0 + G[3] * a1 + G[6] * a2;
        dr[c][1] = G[1] * a0 + G[4] * a1 + G[7] * a2;
        dr[c][2] = G[2] * a0 + G[5] * a1 + G[8] * a2;
        du[c][0] = G[0] * b0 + G[3] * b1 + G[6] * b2;
        du[c][1] = G[1] * b0 + G[4] * b1 + G[7] * b2;
        du[c][2] = G[2] * b0 + G[5] * b1 + G[8] * b2;
      }
      const real divr = dr[0][0] + dr[1][1] + dr[2][2];
      const real divu = du[0][0] + du[1][1] + du[2][2];
      const real ee = dr[0][0] * du[0][0] + dr[1][1] * du[1][1] +
                      dr[2][2] * du[2][2] +
                      real(0.5) * ((dr[0][1] + dr[1][0]) * (du[0][1] + du[1][0]) +
                                   (dr[0][2] + dr[2][0]) * (du[0][2] + du[2][0]) +
                                   (dr[1][2] + dr[2][1]) * (du[1][2] + du[2][1]));
      const size_t o = size_t(slot) * NPT3 + p;
      Gd[o] += divr * divu;
      Ge[o] += ee;
      const real L = ME ? mat[size_t(slot) * 2] : mat[size_t(slot) * 2 * NPT3 + p];
      const real M = ME ? mat[size_t(slot) * 2 + 1]
                        : mat[(size_t(slot) * 2 + 1) * NPT3 + p];
      const real lam = L - real(2.0) * M;
      real Ts[3][3];
      Ts[0][0] = L * dr[0][0] + lam * (dr[1][1] + dr[2][2]);
      Ts[1][1] = L * dr[1][1] + lam * (dr[0][0] + dr[2][2]);
      Ts[2][2] = L * dr[2][2] + lam * (dr[0][0] + dr[1][1]);
      Ts[0][1] = Ts[1][0] = M * (dr[0][1] + dr[1][0]);
      Ts[0][2] = Ts[2][0] = M * (dr[0][2] + dr[2][0]);
      Ts[1][2] = Ts[2][1] = M * (dr[1][2] + dr[2][1]);
      const real J = G[9];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sf[(c * 3 + 0) * NPT3 + p] =
            J * (Ts[c][0] * G[0] + Ts[c][1] * G[1] + Ts[c][2] * G[2]);
        sf[(c * 3 + 1) * NPT3 + p] =
            J * (Ts[c][0] * G[3] + Ts[c][1] * G[4] + Ts[c][2] * G[5]);
        sf[(c * 3 + 2) * NPT3 + p] =
            J * (Ts[c][0] * G[6] + Ts[c][1] * G[7] + Ts[c][2] * G[8]);
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        real t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
        for (int k = 0; k < NGLL; ++k) {
          t1 += sf[(c * 3 + 0) * NPT3 + p3(iz, iy, k)] * q.hp[k][ix] * q.w[k];
          t2 += sf[(c * 3 + 1) * NPT3 + p3(iz, k, ix)] * q.hp[k][iy] * q.w[k];
          t3 += sf[(c * 3 + 2) * NPT3 + p3(k, iy, ix)] * q.hp[k][iz] * q.w[k];
        }
        const real rr = q.w[iz] * q.w[iy] * t1 + q.w[iz] * q.w[ix] * t2 +
                        q.w[iy] * q.w[ix] * t3;
        s_a[c * nloc_max + l] += real(-1.0) * rr;
      }
    }
    __syncthreads();
  }
  // up to two exclusive nodes per thread: their loads are requested before
  // the first update (the round registers are free by now)
  constexpr int kAdjFin = 2;
  for (int i0 = tid; i0 < T.nex; i0 += kAdjFin * RW * ELEM_THREADS) {
    AdjNode<real> nd[kAdjFin];
#pragma unroll
    for (int k = 0; k < kAdjFin; ++k) {
      const int i = i0 + k * RW * ELEM_THREADS;
      if (i < T.nex)
        adj_load(long(T.ex_off) + i, j, S, nd[k]);
    }
#pragma unroll
    for (int k = 0; k < kAdjFin; ++k) {
      const int i = i0 + k * RW * ELEM_THREADS;
      if (i < T.nex) {
        const real kr[3] = {s_a[i], s_a[nloc_max + i], s_a[2 * nloc_max + i]};
        adj_apply(long(T.ex_off) + i, kr, j, S, nd[k]);
      }
    }
  }
  for (int i = tid; i < T.nsh; i += RW * ELEM_THREADS)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      part[c * npart + T.sh_off + i] = s_a[c * nloc_max + T.nex + i];
}

// Shared nodes of the adjoint step: partials summed in tile order, then
// adj_finish.
template <typename real>
__global__ void __launch_bounds__(NODE3_THREADS)
    k3_adj_shared(int nsh, int nex_total, const int *__restrict__ sh_poff,
                  const int *__restrict__ sh_pidx, const real *__restrict__ part,
                  long npart, AdjState3<real> S,
                  const int *__restrict__ adj_step) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsh)
    return;
  const int j = *adj_step % S.every;
  // the node's own state first (it does not wait on the partials), then the
  // partials of the three components in tile order
  AdjNode<real> nd;
  adj_load(long(nex_total) + s, j, S, nd);
  real kr[3] = {0, 0, 0};
  for (int k = sh_poff[s]; k < sh_poff[s + 1]; ++k) {
    const int q = sh_pidx[k];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      kr[c] += part[c * npart + q];
  }
  adj_apply(long(nex_total) + s, kr, j, S, nd);
}

template <typename real>
__global__ void __launch_bounds__(NODE3_THREADS)
    k3_pred(real *__restrict__ U, real *__restrict__ V, real *__restrict__ A,
            long n3, real dt, real dto2, real dt2o2) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n3)
    return;
  real u = U[i], v = V[i], a = A[i];
  u += dt * v + dt2o2 * a;
  v += dto2 * a;
  U[i] = u;
  V[i] = v;
  A[i] = 0;
}

__global__ void k3_set_step(int *__restrict__ step_ptr, int value) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *step_ptr = value;
}

} // namespace

template <typename real> class Solver3D final : public Solver3DBase {
public:
  Solver3D(std::unique_ptr<HostAssembly3D> a, int graph_steps,
           bool nodes_geometry)
      : asm_(std::move(a)), gsteps_(graph_steps) {
    const HostAssembly3D &A = *asm_;
    SMX3_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    SMX3_HIP(hipEventCreate(&ev0_));
    SMX3_HIP(hipEventCreate(&ev1_));
    npad_ = (long(A.nglob) + 255) / 256 * 256;
    ibool_ = upload<int>(A.ibool);
    geom_ = upload<real>(A.geom);
    gmode_ = A.geom_elem ? GEOM_ELEM : GEOM_POINT;
    for (int i = 0; i < NGLL; ++i)
      xi3_.w[i] = A.quad.xi[i];
    if (nodes_geometry && !A.geom_elem)
      try_geometry_from_nodes();
    sslot_ = upload<int>(A.stacey_slot);
    nst_ = std::max(1, A.nstacey);
    upload_material();
    stf_ = upload<real>(A.stf);
    // tile layout
    if (A.round_width != kRoundWidth)
      throw SmxError(SMX_ERR_ARG, "3D tile rounds must be " +
                                      std::to_string(kRoundWidth) +
                                      " elements wide for this precision");
    tiles_ = upload<TileInfo3>(A.tiles);
    round_slots_ = upload<int>(A.round_slots);
    lidx_ = upload<uint16_t>(A.lidx);
    sh_list_ = upload<int>(A.sh_list);
    sh_poff_ = upload<int>(A.sh_poff);
    sh_pidx_ = upload<int>(A.sh_pidx);
    color_slots_ = upload<int>(A.color_slots);
    tsrc_li_ = upload<int>(A.tsrc_li);
    tsrc_id_ = upload<int>(A.tsrc_id);
    tsrc_val_ = upload<real>(A.tsrc_val);
    npart_ = std::max(1L, long(A.npart));
    part_ = alloc<real>(3 * size_t(npart_));
    for (const TileInfo3 &t : A.tiles) {
      nsh_max_ = std::max(nsh_max_, t.nsh);
      nround_max_ = std::max(nround_max_, t.nround);
    }
    tile_lds_ = (size_t(3) * std::max(1, A.nloc_max) + size_t(kRoundWidth) * 12 * NPT3) *
                    sizeof(real) +
                (size_t(nsh_max_) + size_t(nround_max_) * kRoundWidth) * sizeof(int);
    {
      int dev = 0;
      SMX3_HIP(hipGetDevice(&dev));
      hipDeviceProp_t prop;
      SMX3_HIP(hipGetDeviceProperties(&prop, dev));
      if (tile_lds_ > prop.sharedMemPerBlock)
        throw SmxError(SMX_ERR_MESH,
                       "a 3D tile needs " + std::to_string(tile_lds_) +
                           " bytes of LDS (device limit " +
                           std::to_string(prop.sharedMemPerBlock) +
                           "); use fewer elements per tile");
      adj_lds_ = (size_t(3) * std::max(1, A.nloc_max) +
                  size_t(kRoundWidth) * 15 * NPT3) *
                     sizeof(real) +
                 (size_t(nsh_max_) + size_t(nround_max_) * kRoundWidth) * sizeof(int);
      adj_tiles_ = adj_lds_ <= prop.sharedMemPerBlock;
      if (adj_tiles_)
        for (const void *f :
             {reinterpret_cast<const void *>(&k3_adj_tile<real, GEOM_ELEM, true, kRoundWidth>),
              reinterpret_cast<const void *>(&k3_adj_tile<real, GEOM_ELEM, false, kRoundWidth>),
              reinterpret_cast<const void *>(&k3_adj_tile<real, GEOM_NODES, true, kRoundWidth>),
              reinterpret_cast<const void *>(&k3_adj_tile<real, GEOM_NODES, false, kRoundWidth>),
              reinterpret_cast<const void *>(&k3_adj_tile<real, GEOM_POINT, true, kRoundWidth>),
              reinterpret_cast<const void *>(&k3_adj_tile<real, GEOM_POINT, false, kRoundWidth>)})
          SMX3_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(adj_lds_)));
    }
    rec_g_ = upload<int>(A.rec_g);
    rec_h_ = upload<real>(A.rec_h);
    U_ = alloc<real>(3 * npad_);
    V_ = alloc<real>(3 * npad_);
    A_ = alloc<real>(3 * npad_);
    step_ = alloc<int>(1);
    const int ntypes = int(A.seis_types.size());
    seis_ = alloc<real>(std::max<size_t>(
        size_t(std::max(1, A.nsig)) * ntypes * std::max(1, A.nrec) * 3, 3));
    for (int k = 0; k < 4; ++k)
      types_.t[k] = k < ntypes ? A.seis_types[k] : 0;
    for (int t : A.seis_types) {
      need_disp_ = need_disp_ || t == 0;
      if (t != 0 && !aux_v_) {
        aux_v_ = alloc<real>(3 * npad_);
        aux_a_ = alloc<real>(3 * npad_);
      }
    }
    for (int i = 0; i < NGLL; ++i) {
      quad_.w[i] = A.quad.w[i];
      for (int j = 0; j < NGLL; ++j)
        quad_.hp[i][j] = A.quad.hprime[i][j];
    }
    dt_ = A.dt;
    dto2_ = A.dt_over2;
    dt2o2_ = A.dt2_over2;
    build_boundary_nodes();
    SMX3_HIP(hipStreamSynchronize(stream_));
  }

  ~Solver3D() override {
    if (graph_exec_)
      (void)hipGraphExecDestroy(graph_exec_);
    if (graph_)
      (void)hipGraphDestroy(graph_);
    (void)hipStreamSynchronize(stream_);
    for (void *p : allocs_)
      (void)hipFree(p);
    (void)hipEventDestroy(ev0_);
    (void)hipEventDestroy(ev1_);
    for (hipEvent_t e : adj_ev_)
      if (e)
        (void)hipEventDestroy(e);
    (void)hipStreamDestroy(stream_);
  }

  int precision() const override { return sizeof(real); }
  const HostAssembly3D &assembly() const override { return *asm_; }

  // The step graph is captured, instantiated and uploaded here (and again
  // after set_model), so no run_forward call pays for it.
  void initialize() override {
    prepare_graph();
    initialized_ = true;
  }
  void prepare_graph() {
    build_graph();
    SMX3_HIP(hipGraphUpload(graph_exec_, stream_));
    SMX3_HIP(hipStreamSynchronize(stream_));
  }

  void geometry_mode(int *mode, long long *mismatches) const override {
    if (mode)
      *mode = gmode_;
    if (mismatches)
      *mismatches = static_cast<long long>(geom_mismatch_);
  }

  void run_forward(int istep0, int nstep) override {
    if (!initialized_)
      throw SmxError(SMX_ERR_STATE, "smx_initialize has not been called");
    const HostAssembly3D &A = *asm_;
    if (istep0 < 0 || nstep < 1 || istep0 + nstep > A.nstep)
      throw SmxError(SMX_ERR_ARG, "steps outside [0, nstep)");
    SMX3_HIP(hipEventRecord(ev0_, stream_));
    k3_set_step<<<1, 64, 0, stream_>>>(step_, istep0);
    SMX3_HIP(hipGetLastError());
    k3_pred<real><<<node_grid(), NODE3_THREADS, 0, stream_>>>(
        U_, V_, A_, 3 * npad_, dt_, dto2_, dt2o2_);
    SMX3_HIP(hipGetLastError());
    const int body = nstep - 1;
    const int reps = body / gsteps_;
    if (reps > 0) {
      build_graph();
      for (int r = 0; r < reps; ++r)
        SMX3_HIP(hipGraphLaunch(graph_exec_, stream_));
    }
    for (int s = reps * gsteps_; s < body; ++s)
      step<false>();
    step<true>();
    SMX3_HIP(hipEventRecord(ev1_, stream_));
    have_stats_ = true;
    launches_ = nstep;
  }

  void synchronize() override { SMX3_HIP(hipStreamSynchronize(stream_)); }

  void get_seismograms(void *host) override {
    const HostAssembly3D &A = *asm_;
    const size_t n = size_t(A.nsig) * A.seis_types.size() * A.nrec * 3;
    SMX3_HIP(hipMemcpyAsync(host, seis_, n * sizeof(real),
                            hipMemcpyDeviceToHost, stream_));
    SMX3_HIP(hipStreamSynchronize(stream_));
  }

  void get_field(int which, void *host) override {
    const real *d = pick(which);
    const int n = asm_->nglob;
    std::vector<real> buf(size_t(3) * npad_);
    SMX3_HIP(hipMemcpyAsync(buf.data(), d, buf.size() * sizeof(real),
                            hipMemcpyDeviceToHost, stream_));
    SMX3_HIP(hipStreamSynchronize(stream_));
    real *h = static_cast<real *>(host);
    for (int g = 0; g < n; ++g)
      for (int c = 0; c < 3; ++c)
        h[size_t(g) * 3 + c] = buf[size_t(c) * npad_ + g];
  }

  void set_field(int which, const void *host) override {
    if (which == SMX_MASS_INVERSE)
      throw SmxError(SMX_ERR_ARG, "the mass matrix is read-only");
    real *d = pick(which);
    const int n = asm_->nglob;
    std::vector<real> buf(size_t(3) * npad_, real(0));
    const real *h = static_cast<const real *>(host);
    for (int g = 0; g < n; ++g)
      for (int c = 0; c < 3; ++c)
        buf[size_t(c) * npad_ + g] = h[size_t(g) * 3 + c];
    SMX3_HIP(hipMemcpyAsync(d, buf.data(), buf.size() * sizeof(real),
                            hipMemcpyHostToDevice, stream_));
    SMX3_HIP(hipStreamSynchronize(stream_));
  }

  void last_run_stats(double *ms, double *stiff_ms, int *launches) override {
    float t = 0;
    if (have_stats_) {
      SMX3_HIP(hipEventSynchronize(ev1_));
      SMX3_HIP(hipEventElapsedTime(&t, ev0_, ev1_));
    }
    if (ms)
      *ms = t;
    if (stiff_ms)
      *stiff_ms = stiff_ms_;
    if (launches)
      *launches = launches_;
  }

  double time_stiffness(int n) override {
    if (n < 1)
      throw SmxError(SMX_ERR_ARG, "n must be positive");
    if (!scratch_a_)
      scratch_a_ = alloc<real>(3 * npad_);
    launch_tile<M3_ASSEMBLE>(scratch_a_, Rec3<real>{}); // warm-up
    SMX3_HIP(hipEventRecord(ev0_, stream_));
    for (int i = 0; i < n; ++i)
      launch_tile<M3_ASSEMBLE>(scratch_a_, Rec3<real>{});
    SMX3_HIP(hipEventRecord(ev1_, stream_));
    SMX3_HIP(hipEventSynchronize(ev1_));
    float t = 0;
    SMX3_HIP(hipEventElapsedTime(&t, ev0_, ev1_));
    stiff_ms_ = double(t) / n;
    have_stats_ = false;
    return stiff_ms_;
  }

  // ---------------------------------------------------------------- adjoint
  void set_model(const float *kappa, const float *mu,
                 const float *rho) override {
    SMX3_HIP(hipStreamSynchronize(stream_));
    set_model3d(*asm_, kappa, mu, rho);
    upload_material();
    build_boundary_nodes();
    drop_graph(); // the captured launches hold the old material buffers
    ckpt_every_ = 0;
    have_gradient_ = false;
    SMX3_HIP(hipStreamSynchronize(stream_));
    if (initialized_)
      prepare_graph();
  }

  void run_forward_checkpointed(int every) override {
    if (!initialized_)
      throw SmxError(SMX_ERR_STATE, "smx_initialize has not been called");
    const HostAssembly3D &A = *asm_;
    if (every < 1)
      throw SmxError(SMX_ERR_ARG, "checkpoint interval must be positive");
    every = std::min(every, A.nstep);
    const int nseg = (A.nstep + every - 1) / every;
    const size_t f3 = size_t(3) * npad_;
    if (nseg > ckpt_cap_) {
      release(ckpt_);
      ckpt_ = alloc<real>(size_t(nseg) * 2 * f3);
      ckpt_cap_ = nseg;
    }
    ensure_events();
    SMX3_HIP(hipEventRecord(adj_ev_[0], stream_));
    for (real *p : {U_, V_, A_}) // from rest
      SMX3_HIP(hipMemsetAsync(p, 0, f3 * sizeof(real), stream_));
    k3_set_step<<<1, 64, 0, stream_>>>(step_, 0);
    SMX3_HIP(hipGetLastError());
    k3_pred<real><<<node_grid(), NODE3_THREADS, 0, stream_>>>(
        U_, V_, A_, 3 * npad_, dt_, dto2_, dt2o2_);
    SMX3_HIP(hipGetLastError());
    for (int sg = 0; sg < nseg; ++sg) {
      const int k0 = sg * every, len = std::min(every, A.nstep - k0);
      real *ck = ckpt_ + size_t(sg) * 2 * f3;
      SMX3_HIP(hipMemcpyAsync(ck, U_, f3 * sizeof(real), hipMemcpyDeviceToDevice,
                              stream_));
      SMX3_HIP(hipMemcpyAsync(ck + f3, V_, f3 * sizeof(real),
                              hipMemcpyDeviceToDevice, stream_));
      const bool last = sg == nseg - 1;
      const int body = last ? len - 1 : len;
      const int reps = body / gsteps_;
      if (reps > 0) {
        build_graph();
        for (int r = 0; r < reps; ++r)
          SMX3_HIP(hipGraphLaunch(graph_exec_, stream_));
      }
      for (int st = reps * gsteps_; st < body; ++st)
        step<false>();
      if (last)
        step<true>();
    }
    SMX3_HIP(hipEventRecord(adj_ev_[1], stream_));
    ckpt_every_ = every;
    have_gradient_ = false;
    have_adj_stats_ = false;
  }

  void run_adjoint(const void *gsrc_host) override {
    const HostAssembly3D &A = *asm_;
    if (!ckpt_every_)
      throw SmxError(SMX_ERR_STATE,
                     "no checkpointed forward run for the current model");
    const int every = ckpt_every_;
    const int nseg = (A.nstep + every - 1) / every;
    const size_t f3 = size_t(3) * npad_;
    ensure_adjoint(every);
    const size_t ng = size_t(A.nstep) * A.nrec * 3;
    if (ng)
      SMX3_HIP(hipMemcpyAsync(gsrc_, gsrc_host, ng * sizeof(real),
                              hipMemcpyHostToDevice, stream_));
    SMX3_HIP(hipEventRecord(adj_ev_[2], stream_));
    for (real *p : {ub_, vb_, rb_, kacc_})
      SMX3_HIP(hipMemsetAsync(p, 0, f3 * sizeof(real), stream_));
    for (real *p : {gd_, ge_})
      SMX3_HIP(hipMemsetAsync(p, 0, size_t(A.nspec) * NPT3 * sizeof(real), stream_));
    SMX3_HIP(hipMemsetAsync(racc_, 0, size_t(npad_) * sizeof(real), stream_));
    SMX3_HIP(hipMemsetAsync(t9_, 0, size_t(9) * nbn_ * sizeof(real), stream_));
    SMX3_HIP(hipMemsetAsync(d3_, 0, size_t(3) * nbn_ * sizeof(real), stream_));
    const Rec3<real> rec{rec_a_, rec_u_, rec_vb_, bidx_, every, npad_, nbn_alloc_};
    for (int sg = nseg - 1; sg >= 0; --sg) {
      const int k0 = sg * every, len = std::min(every, A.nstep - k0);
      const real *ck = ckpt_ + size_t(sg) * 2 * f3;
      // recompute the segment from its checkpoint, recording every step
      SMX3_HIP(hipMemcpyAsync(U_, ck, f3 * sizeof(real), hipMemcpyDeviceToDevice,
                              stream_));
      SMX3_HIP(hipMemcpyAsync(V_, ck + f3, f3 * sizeof(real),
                              hipMemcpyDeviceToDevice, stream_));
      SMX3_HIP(hipMemcpyAsync(rec_u_, ck, f3 * sizeof(real),
                              hipMemcpyDeviceToDevice, stream_));
      SMX3_HIP(hipMemsetAsync(A_, 0, f3 * sizeof(real), stream_));
      if (nbn_ > 0) {
        k3_gather_vb<real><<<dim3(unsigned((nbn_ + 255) / 256)), 256, 0, stream_>>>(
            V_, bnode_, nbn_alloc_, npad_, rec_vb_);
        SMX3_HIP(hipGetLastError());
      }
      k3_set_step<<<1, 64, 0, stream_>>>(step_, k0);
      SMX3_HIP(hipGetLastError());
      for (int st = 0; st < len; ++st)
        record_step(rec);
      // adjoint steps k0+len-1 .. k0
      k3_set_step<<<1, 64, 0, stream_>>>(adj_step_, k0 + len - 1);
      SMX3_HIP(hipGetLastError());
      for (int st = 0; st < len; ++st)
        adjoint_step(every);
    }
    SMX3_HIP(hipEventRecord(adj_ev_[3], stream_));
    have_gradient_ = true;
    have_adj_stats_ = true;
    grad_valid_ = false;
  }

  void get_gradient(double *out) override {
    if (!have_gradient_)
      throw SmxError(SMX_ERR_STATE, "no adjoint run");
    if (!grad_valid_)
      finalize_gradient();
    std::copy(grad_.begin(), grad_.end(), out);
  }

  void last_adjoint_stats(double *forward_ms, double *adjoint_ms) override {
    float f = 0, a = 0;
    if (have_adj_stats_) {
      SMX3_HIP(hipEventSynchronize(adj_ev_[3]));
      SMX3_HIP(hipEventElapsedTime(&f, adj_ev_[0], adj_ev_[1]));
      SMX3_HIP(hipEventElapsedTime(&a, adj_ev_[2], adj_ev_[3]));
    }
    if (forward_ms)
      *forward_ms = f;
    if (adjoint_ms)
      *adjoint_ms = a;
  }

private:
  template <typename T, typename H> T *upload(const std::vector<H> &h) {
    T *d = alloc<T>(h.size());
    if (!h.empty()) {
      std::vector<T> tmp(h.begin(), h.end());
      SMX3_HIP(hipMemcpyAsync(d, tmp.data(), tmp.size() * sizeof(T),
                              hipMemcpyHostToDevice, stream_));
      SMX3_HIP(hipStreamSynchronize(stream_));
    }
    return d;
  }
  template <typename T> T *alloc(size_t n) {
    void *p = nullptr;
    SMX3_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    allocs_.push_back(p);
    SMX3_HIP(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), stream_));
    return static_cast<T *>(p);
  }
  template <typename T> void release(T *&p) {
    if (!p)
      return;
    auto it = std::find(allocs_.begin(), allocs_.end(), static_cast<void *>(p));
    if (it != allocs_.end())
      allocs_.erase(it);
    SMX3_HIP(hipStreamSynchronize(stream_));
    (void)hipFree(p);
    p = nullptr;
  }
  void drop_graph() {
    if (graph_exec_)
      (void)hipGraphExecDestroy(graph_exec_);
    if (graph_)
      (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
  }
  void ensure_events() {
    for (hipEvent_t &e : adj_ev_)
      if (!e)
        SMX3_HIP(hipEventCreate(&e));
  }

  // Moduli, Stacey terms (impedances in double for FP64) and inverse mass.
  void upload_material() {
    const HostAssembly3D &A = *asm_;
    release(mat_);
    mat_ = sizeof(real) == 8 ? upload<real>(A.mat64) : upload<real>(A.mat);
    std::vector<double> st(A.stacey.begin(), A.stacey.end());
    if (sizeof(real) == 8)
      for (int b = 0; b < A.nstacey; ++b)
        for (int c = 0; c < 2; ++c)
          for (int p = 0; p < NPT3; ++p)
            st[(size_t(b) * 7 + 5 + c) * NPT3 + p] =
                A.stacey64[(size_t(b) * 2 + c) * NPT3 + p];
    release(stacey_);
    stacey_ = upload<real>(st);
    std::vector<double> mi(size_t(3) * npad_, 1.0);
    for (int c = 0; c < 3; ++c)
      for (int g = 0; g < A.nglob; ++g)
        mi[size_t(c) * npad_ + g] = sizeof(real) == 8
                                        ? A.minv64[size_t(c) * A.nglob + g]
                                        : A.minv[size_t(c) * A.nglob + g];
    release(minv_);
    minv_ = upload<real>(mi);
  }

  // Absorbing nodes (first touch in slot order), node -> index, and each
  // node's Stacey matrix B = sum over its face points of
  // |n| w [(rho vp - rho vs) n n^T / |n|^2 + rho vs I], formed in double.
  void build_boundary_nodes() {
    const HostAssembly3D &A = *asm_;
    std::vector<int> bidx(size_t(npad_), -1), bnode;
    std::vector<double> bm;
    for (int s = 0; s < A.nspec; ++s) {
      const int b = A.stacey_slot[s];
      if (b < 0)
        continue;
      auto at = [&](int c, int p) { return A.stacey[(size_t(b) * 7 + c) * NPT3 + p]; };
      for (int p = 0; p < NPT3; ++p) {
        if (at(4, p) == 0.0f)
          continue;
        const int g = A.ibool[size_t(s) * NPT3 + p];
        if (bidx[g] < 0) {
          bidx[g] = int(bnode.size());
          bnode.push_back(g);
          bm.resize(bm.size() + 6, 0.0);
        }
        const double n[3] = {at(0, p), at(1, p), at(2, p)};
        const double nrm = at(3, p), w = at(4, p);
        const double P = sizeof(real) == 8 ? A.stacey64[(size_t(b) * 2) * NPT3 + p]
                                           : double(at(5, p));
        const double S = sizeof(real) == 8 ? A.stacey64[(size_t(b) * 2 + 1) * NPT3 + p]
                                           : double(at(6, p));
        const double f = nrm * w, q = (P - S) / (nrm * nrm);
        double *m = &bm[size_t(bidx[g]) * 6];
        m[0] += f * (q * n[0] * n[0] + S);
        m[1] += f * (q * n[1] * n[1] + S);
        m[2] += f * (q * n[2] * n[2] + S);
        m[3] += f * q * n[0] * n[1];
        m[4] += f * q * n[0] * n[2];
        m[5] += f * q * n[1] * n[2];
      }
    }
    nbn_ = long(bnode.size());
    const long nb = std::max(1L, nbn_);
    std::vector<double> bmt(size_t(6) * nb, 0.0);
    for (long b = 0; b < nbn_; ++b)
      for (int k = 0; k < 6; ++k)
        bmt[size_t(k) * nb + b] = bm[size_t(b) * 6 + k];
    bidx_h_ = bidx;
    release(bidx_);
    release(bnode_);
    release(bmat_);
    bidx_ = upload<int>(bidx);
    bnode_ = upload<int>(bnode);
    bmat_ = upload<real>(bmt);
    nbn_alloc_ = nb;
    // adjoint buffers sized by nb are re-made on the next adjoint run
    release(t9_);
    release(d3_);
    rec_every_ = 0;
  }

  // Device buffers of the adjoint run (kept between runs).
  void ensure_adjoint(int every) {
    const HostAssembly3D &A = *asm_;
    const size_t f3 = size_t(3) * npad_;
    if (!ub_) {
      ub_ = alloc<real>(f3);
      vb_ = alloc<real>(f3);
      rb_ = alloc<real>(f3);
      kacc_ = alloc<real>(f3);
      racc_ = alloc<real>(size_t(npad_));
      gd_ = alloc<real>(size_t(A.nspec) * NPT3);
      ge_ = alloc<real>(size_t(A.nspec) * NPT3);
      adj_step_ = alloc<int>(1);
      gsrc_ = alloc<real>(std::max<size_t>(size_t(A.nstep) * A.nrec * 3, 1));
      // receiver injection: nodes touched by receivers, entries in
      // (receiver, point) order
      std::vector<std::vector<std::pair<int, float>>> lists;
      std::vector<int> idx(size_t(npad_), -1), nodes;
      for (int r = 0; r < A.nrec; ++r)
        for (int p = 0; p < NPT3; ++p) {
          const float h = A.rec_h[size_t(r) * NPT3 + p];
          if (h == 0.0f)
            continue;
          const int g = A.rec_g[size_t(r) * NPT3 + p];
          if (idx[g] < 0) {
            idx[g] = int(nodes.size());
            nodes.push_back(g);
            lists.emplace_back();
          }
          lists[idx[g]].push_back({r, h});
        }
      std::vector<int> off(1, 0), er;
      std::vector<float> eh;
      for (auto &l : lists) {
        for (auto &e : l) {
          er.push_back(e.first);
          eh.push_back(e.second);
        }
        off.push_back(int(er.size()));
      }
      nrin_ = int(nodes.size());
      rin_node_ = upload<int>(nodes);
      rin_off_ = upload<int>(off);
      rin_r_ = upload<int>(er);
      rin_h_ = upload<real>(eh);
    }
    if (!t9_) {
      t9_ = alloc<real>(size_t(9) * nbn_alloc_);
      d3_ = alloc<real>(size_t(3) * nbn_alloc_);
    }
    if (every != rec_every_) {
      release(rec_u_);
      release(rec_a_);
      release(rec_vb_);
      rec_u_ = alloc<real>(size_t(every) * f3);
      rec_a_ = alloc<real>(size_t(every) * f3);
      rec_vb_ = alloc<real>(size_t(every) * 3 * nbn_alloc_);
      rec_every_ = every;
    }
  }

  // One forward step of the recompute, recording a', u~ and v~.
  void record_step(const Rec3<real> &rec) {
    launch_tile<M3_RECORD>(A_, rec);
    launch_shared<M3_RECORD>(rec);
    k3_step_add<<<1, 64, 0, stream_>>>(step_, 1);
    SMX3_HIP(hipGetLastError());
  }

  AdjState3<real> adj_state(int every) const {
    return AdjState3<real>{ub_,    vb_,    rb_,   minv_, rec_a_, rec_vb_,
                           every,  racc_,  bidx_, bmat_, t9_,    d3_,
                           nbn_alloc_, npad_, dt_, dto2_, dt2o2_};
  }

  // One adjoint step (k = *adj_step, then k - 1): misfit source, then the
  // tile pass and the shared nodes, or (when a tile's adjoint staging does
  // not fit in LDS) the colour passes and the node pass.
  void adjoint_step(int every) {
    const HostAssembly3D &A = *asm_;
    if (nrin_ > 0) {
      k3_adj_inject<real><<<dim3(unsigned((nrin_ + 127) / 128)), 128, 0, stream_>>>(
          rin_node_, rin_off_, rin_r_, rin_h_, nrin_, gsrc_, A.nrec, ub_, npad_,
          adj_step_);
      SMX3_HIP(hipGetLastError());
    }
    if (adj_tiles_) {
      const AdjState3<real> S = adj_state(every);
      auto launch = [&](auto kern) {
        kern<<<dim3(unsigned(A.ntiles)), dim3(kRoundWidth * ELEM_THREADS),
               adj_lds_, stream_>>>(tiles_, round_slots_, lidx_, sh_list_, geom_,
                                    nodes_, xi3_, mat_, rec_u_, part_, npart_,
                                    gd_, ge_, std::max(1, A.nloc_max), nsh_max_,
                                    quad_, adj_step_, S);
      };
      const bool me = A.mat_elem;
      switch (gmode_) {
      case GEOM_ELEM:
        me ? launch(k3_adj_tile<real, GEOM_ELEM, true, kRoundWidth>)
           : launch(k3_adj_tile<real, GEOM_ELEM, false, kRoundWidth>);
        break;
      case GEOM_NODES:
        me ? launch(k3_adj_tile<real, GEOM_NODES, true, kRoundWidth>)
           : launch(k3_adj_tile<real, GEOM_NODES, false, kRoundWidth>);
        break;
      default:
        me ? launch(k3_adj_tile<real, GEOM_POINT, true, kRoundWidth>)
           : launch(k3_adj_tile<real, GEOM_POINT, false, kRoundWidth>);
      }
      SMX3_HIP(hipGetLastError());
      if (A.nsh_total > 0) {
        k3_adj_shared<real><<<dim3(unsigned((A.nsh_total + NODE3_THREADS - 1) /
                                            NODE3_THREADS)),
                              NODE3_THREADS, 0, stream_>>>(
            A.nsh_total, A.nex_total, sh_poff_, sh_pidx_, part_, npart_, S,
            adj_step_);
        SMX3_HIP(hipGetLastError());
      }
      k3_step_add<<<1, 64, 0, stream_>>>(adj_step_, -1);
      SMX3_HIP(hipGetLastError());
      return;
    }
    for (int c = 0; c < A.ncolors; ++c) {
      const int s0 = A.color_off[c], cnt = A.color_off[c + 1] - s0;
      if (cnt == 0)
        continue;
      auto launch = [&](auto kern) {
        kern<<<dim3(cnt), dim3(ELEM_THREADS), 0, stream_>>>(
            color_slots_ + s0, ibool_, geom_, nodes_, xi3_, mat_, rb_, rec_u_,
            every, adj_step_, kacc_, gd_, ge_, npad_, quad_);
      };
      const bool me = A.mat_elem;
      switch (gmode_) {
      case GEOM_ELEM:
        me ? launch(k3_adj_elem<real, GEOM_ELEM, true>)
           : launch(k3_adj_elem<real, GEOM_ELEM, false>);
        break;
      case GEOM_NODES:
        me ? launch(k3_adj_elem<real, GEOM_NODES, true>)
           : launch(k3_adj_elem<real, GEOM_NODES, false>);
        break;
      default:
        me ? launch(k3_adj_elem<real, GEOM_POINT, true>)
           : launch(k3_adj_elem<real, GEOM_POINT, false>);
      }
      SMX3_HIP(hipGetLastError());
    }
    k3_adj_node<real><<<dim3(unsigned((npad_ + NODE3_THREADS - 1) / NODE3_THREADS)),
                        NODE3_THREADS, 0, stream_>>>(kacc_, adj_state(every),
                                                     A.nglob, adj_step_);
    SMX3_HIP(hipGetLastError());
    k3_step_add<<<1, 64, 0, stream_>>>(adj_step_, -1);
    SMX3_HIP(hipGetLastError());
  }

  // Chain rule and quadrature on the host, in double:
  //   kappa: -wJ Gd,  mu: -wJ (2 Ge - 2/3 Gd),  rho: -wJ (rb . a')_node,
  // plus, at absorbing points, the rho vp / rho vs terms of B and of the
  // Stacey mass, differentiated through rho vp = sqrt(rho (kappa + 4/3 mu))
  // and rho vs = sqrt(rho mu).
  void finalize_gradient() {
    const HostAssembly3D &A = *asm_;
    SMX3_HIP(hipStreamSynchronize(stream_));
    const size_t nloc = size_t(A.nspec) * NPT3;
    std::vector<real> gd(nloc), ge(nloc), R(static_cast<size_t>(npad_)),
        t9(size_t(9) * nbn_alloc_), d3(size_t(3) * nbn_alloc_);
    SMX3_HIP(hipMemcpy(gd.data(), gd_, nloc * sizeof(real), hipMemcpyDeviceToHost));
    SMX3_HIP(hipMemcpy(ge.data(), ge_, nloc * sizeof(real), hipMemcpyDeviceToHost));
    SMX3_HIP(hipMemcpy(R.data(), racc_, R.size() * sizeof(real), hipMemcpyDeviceToHost));
    SMX3_HIP(hipMemcpy(t9.data(), t9_, t9.size() * sizeof(real), hipMemcpyDeviceToHost));
    SMX3_HIP(hipMemcpy(d3.data(), d3_, d3.size() * sizeof(real), hipMemcpyDeviceToHost));
    grad_.assign(3 * nloc, 0.0);
    std::vector<int> slot_elem(A.nspec);
    for (int e = 0; e < A.nspec; ++e)
      slot_elem[A.elem_slot[e]] = e;
    const long nb = nbn_alloc_;
    const double hdt = 0.5 * double(A.dt);
    for (int s = 0; s < A.nspec; ++s) {
      const size_t oe = size_t(slot_elem[s]) * NPT3;
      const int b = A.stacey_slot[s];
      for (int iz = 0; iz < NGLL; ++iz)
        for (int iy = 0; iy < NGLL; ++iy)
          for (int ix = 0; ix < NGLL; ++ix) {
            const int p = (iz * NGLL + iy) * NGLL + ix;
            const size_t i = size_t(s) * NPT3 + p;
            const int g = A.ibool[i];
            const double wj = double(A.quad.w[iz]) * A.quad.w[iy] * A.quad.w[ix] *
                              A.jac_pt[i];
            double gk = -wj * double(gd[i]);
            double gm = -wj * (2.0 * double(ge[i]) - 2.0 / 3.0 * double(gd[i]));
            double gr = -wj * double(R[g]);
            if (b >= 0 && A.stacey[(size_t(b) * 7 + 4) * NPT3 + p] != 0.0f) {
              auto at = [&](int c) { return double(A.stacey[(size_t(b) * 7 + c) * NPT3 + p]); };
              const double nv[3] = {at(0), at(1), at(2)};
              const double nrm = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
              const double nh[3] = {nv[0] / nrm, nv[1] / nrm, nv[2] / nrm};
              const double wt = at(4), s1 = nh[0] + nh[1] + nh[2];
              const long bi = bidx_h_[g];
              double rnvn = 0, tr = 0, tp = 0, ts = 0;
              for (int c = 0; c < 3; ++c) {
                tr += double(t9[size_t(c * 3 + c) * nb + bi]);
                for (int d = 0; d < 3; ++d)
                  rnvn += nh[c] * nh[d] * double(t9[size_t(c * 3 + d) * nb + bi]);
                const double D = d3[size_t(c) * nb + bi];
                tp += D * nh[c] * s1;
                ts += D * (1.0 - nh[c] * s1);
              }
              const double dP = -(nrm * wt * rnvn + hdt * nrm * wt * tp);
              const double dS = -(nrm * wt * (tr - rnvn) + hdt * nrm * wt * ts);
              const size_t o3 = size_t(s) * 3 * NPT3 + p;
              const double kk = A.kappa_mu_rho[o3], m = A.kappa_mu_rho[o3 + NPT3],
                           rh = A.kappa_mu_rho[o3 + 2 * NPT3];
              const double L = kk + 4.0 / 3.0 * m;
              const double P = std::sqrt(rh * L), S = std::sqrt(rh * m);
              gk += dP * rh / (2.0 * P);
              gm += dP * (4.0 / 3.0) * rh / (2.0 * P);
              gr += dP * L / (2.0 * P);
              if (S > 0) {
                gm += dS * rh / (2.0 * S);
                gr += dS * m / (2.0 * S);
              }
            }
            grad_[oe + p] = gk;
            grad_[nloc + oe + p] = gm;
            grad_[2 * nloc + oe + p] = gr;
          }
    }
    grad_valid_ = true;
  }

  real *pick(int which) {
    switch (which) {
    case SMX_U: return U_;
    case SMX_V: return V_;
    case SMX_A: return A_;
    case SMX_MASS_INVERSE: return minv_;
    }
    throw SmxError(SMX_ERR_ARG, "unknown field component");
  }
  dim3 node_grid() const {
    return dim3(unsigned((3 * npad_ + NODE3_THREADS - 1) / NODE3_THREADS));
  }

  // The tile pass of one step (A: where M3_FINAL / M3_ASSEMBLE store a).
  template <int MODE> void launch_tile(real *acc_out, const Rec3<real> &rec) {
    const HostAssembly3D &A = *asm_;
    const Aux3<real> aux{aux_v_, aux_a_, A.nsbs};
    auto launch = [&](auto kern) {
      kern<<<dim3(unsigned(A.ntiles)), dim3(kRoundWidth * ELEM_THREADS),
             tile_lds_, stream_>>>(
          tiles_, round_slots_, lidx_, sh_list_, geom_, nodes_, xi3_, mat_,
          sslot_, stacey_, U_, V_, acc_out, minv_, part_, npart_, tsrc_li_,
          tsrc_id_, tsrc_val_, stf_, std::max(1, A.nsrc), std::max(1, A.nloc_max),
          nsh_max_, npad_, dt_, dto2_, dt2o2_, quad_, step_, rec, aux);
    };
    const bool me = A.mat_elem;
    switch (gmode_) {
    case GEOM_ELEM:
      me ? launch(k3_tile<real, MODE, GEOM_ELEM, true, kRoundWidth>)
         : launch(k3_tile<real, MODE, GEOM_ELEM, false, kRoundWidth>);
      break;
    case GEOM_NODES:
      me ? launch(k3_tile<real, MODE, GEOM_NODES, true, kRoundWidth>)
         : launch(k3_tile<real, MODE, GEOM_NODES, false, kRoundWidth>);
      break;
    default:
      me ? launch(k3_tile<real, MODE, GEOM_POINT, true, kRoundWidth>)
         : launch(k3_tile<real, MODE, GEOM_POINT, false, kRoundWidth>);
    }
    SMX3_HIP(hipGetLastError());
  }

  // advance: the shared-point pass also moves the step counter on (the
  // caller checks that nothing in it reads the counter)
  template <int MODE> void launch_shared(const Rec3<real> &rec,
                                         bool advance = false) {
    const HostAssembly3D &A = *asm_;
    if (A.nsh_total == 0)
      return;
    const Aux3<real> aux{aux_v_, aux_a_, A.nsbs};
    k3_shared<real, MODE><<<dim3(unsigned((A.nsh_total + NODE3_THREADS - 1) /
                                          NODE3_THREADS)),
                            NODE3_THREADS, 0, stream_>>>(
        A.nsh_total, A.nex_total, sh_poff_, sh_pidx_, part_, npart_, U_, V_, A_,
        minv_, npad_, dt_, dto2_, dt2o2_, step_, rec, aux, int(advance));
    SMX3_HIP(hipGetLastError());
  }

  template <bool POST> void launch_seis() {
    const HostAssembly3D &A = *asm_;
    const int ntypes = int(A.seis_types.size());
    if (A.nrec == 0 || ntypes == 0 || (POST ? !aux_v_ : !need_disp_))
      return;
    k3_seis<real, POST><<<dim3(unsigned(A.nrec * ntypes)), 64, 0, stream_>>>(
        U_, aux_v_, aux_a_, rec_g_, rec_h_, A.nrec, types_, ntypes, A.nsbs,
        npad_, step_, seis_);
    SMX3_HIP(hipGetLastError());
  }

  // Uses the geometry re-formed from the control nodes when the device
  // reproduces every stored per-point value bit for bit (checked once here);
  // otherwise the element passes keep reading the stored values.
  void try_geometry_from_nodes() {
    const HostAssembly3D &A = *asm_;
    nodes_ = upload<float>(A.nodes);
    float *gf = nullptr, *xi = nullptr;
    unsigned char *bad = nullptr;
    const size_t n = size_t(A.nspec) * NPT3;
    SMX3_HIP(hipMalloc(&gf, A.geom.size() * sizeof(float)));
    SMX3_HIP(hipMalloc(&xi, NGLL * sizeof(float)));
    SMX3_HIP(hipMalloc(&bad, n));
    SMX3_HIP(hipMemcpyAsync(gf, A.geom.data(), A.geom.size() * sizeof(float),
                            hipMemcpyHostToDevice, stream_));
    SMX3_HIP(hipMemcpyAsync(xi, A.quad.xi, NGLL * sizeof(float),
                            hipMemcpyHostToDevice, stream_));
    k3_geom_check<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_>>>(
        nodes_, gf, A.nspec, xi, bad);
    SMX3_HIP(hipGetLastError());
    std::vector<unsigned char> h(n);
    SMX3_HIP(hipMemcpyAsync(h.data(), bad, n, hipMemcpyDeviceToHost, stream_));
    SMX3_HIP(hipStreamSynchronize(stream_));
    (void)hipFree(gf);
    (void)hipFree(xi);
    (void)hipFree(bad);
    size_t nbad = 0;
    for (unsigned char b : h)
      nbad += b;
    geom_mismatch_ = nbad;
    if (nbad == 0)
      gmode_ = GEOM_NODES;
  }

  // One forward step: displacement seismograms (u~ of the step), the tile
  // pass, the shared points, velocity / acceleration seismograms, and the
  // step counter.
  template <bool FINAL> void step() {
    // without velocity / acceleration samples nothing after the shared-point
    // pass reads the step counter, so that pass moves it on
    const bool advance = !aux_v_ && asm_->nsh_total > 0;
    launch_seis<false>();
    launch_tile<FINAL ? M3_FINAL : M3_STEP>(A_, Rec3<real>{});
    launch_shared<FINAL ? M3_FINAL : M3_STEP>(Rec3<real>{}, advance);
    launch_seis<true>();
    if (!advance) {
      k3_step_add<<<1, 64, 0, stream_>>>(step_, 1);
      SMX3_HIP(hipGetLastError());
    }
  }

  void build_graph() {
    if (graph_exec_)
      return;
    SMX3_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      for (int s = 0; s < gsteps_; ++s)
        step<false>();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g)
        (void)hipGraphDestroy(g);
      throw;
    }
    SMX3_HIP(hipStreamEndCapture(stream_, &graph_));
    SMX3_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  }

  // rounds of four elements (512 threads) in FP32; one element per round in
  // FP64, whose LDS staging is twice as large
  static constexpr int kRoundWidth = sizeof(real) == 4 ? 4 : 1;
  std::unique_ptr<HostAssembly3D> asm_;
  int gsteps_ = kGraphSteps3;
  TileInfo3 *tiles_ = nullptr;
  uint16_t *lidx_ = nullptr;
  int *round_slots_ = nullptr, *sh_list_ = nullptr, *sh_poff_ = nullptr,
      *sh_pidx_ = nullptr, *color_slots_ = nullptr, *tsrc_li_ = nullptr,
      *tsrc_id_ = nullptr;
  real *tsrc_val_ = nullptr, *part_ = nullptr, *aux_v_ = nullptr,
       *aux_a_ = nullptr;
  long npart_ = 1;
  size_t tile_lds_ = 0, adj_lds_ = 0;
  int nsh_max_ = 0, nround_max_ = 0; // over the tiles (k3_tile's LDS tables)
  bool adj_tiles_ = false; // adjoint as tiles (else colour passes)
  bool need_disp_ = false;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<void *> allocs_;
  long npad_ = 0;
  int nst_ = 1;
  int *ibool_ = nullptr, *sslot_ = nullptr,
      *rec_g_ = nullptr, *step_ = nullptr;
  int gmode_ = GEOM_POINT;
  size_t geom_mismatch_ = 0;
  float *nodes_ = nullptr;
  Quad3<float> xi3_{}; // GLL nodes in w[] (re-formed geometry)
  real *geom_ = nullptr, *mat_ = nullptr, *stacey_ = nullptr, *minv_ = nullptr,
       *stf_ = nullptr, *rec_h_ = nullptr, *U_ = nullptr,
       *V_ = nullptr, *A_ = nullptr, *seis_ = nullptr, *scratch_a_ = nullptr;
  Quad3<real> quad_;
  Types3 types_;
  real dt_ = 0, dto2_ = 0, dt2o2_ = 0;
  bool initialized_ = false, have_stats_ = false;
  double stiff_ms_ = 0;
  int launches_ = 0;
  // adjoint
  int ckpt_every_ = 0, ckpt_cap_ = 0, rec_every_ = 0, nrin_ = 0;
  long nbn_ = 0, nbn_alloc_ = 1;
  std::vector<int> bidx_h_;
  int *bidx_ = nullptr, *bnode_ = nullptr, *adj_step_ = nullptr,
      *rin_node_ = nullptr, *rin_off_ = nullptr, *rin_r_ = nullptr;
  real *ckpt_ = nullptr, *rec_u_ = nullptr, *rec_a_ = nullptr,
       *rec_vb_ = nullptr, *ub_ = nullptr, *vb_ = nullptr, *rb_ = nullptr,
       *kacc_ = nullptr, *racc_ = nullptr, *gd_ = nullptr, *ge_ = nullptr,
       *t9_ = nullptr, *d3_ = nullptr, *bmat_ = nullptr, *gsrc_ = nullptr,
       *rin_h_ = nullptr;
  hipEvent_t adj_ev_[4] = {};
  bool have_gradient_ = false, have_adj_stats_ = false, grad_valid_ = false;
  std::vector<double> grad_;
};

std::unique_ptr<Solver3DBase> make_solver3d(std::unique_ptr<HostAssembly3D> a,
                                            int precision, int graph_steps,
                                            bool nodes_geometry) {
  if (graph_steps < 1 || graph_steps > 256)
    throw SmxError(SMX_ERR_ARG, "graph_steps must lie in [1, 256]");
  if (precision == 4)
    return std::make_unique<Solver3D<float>>(std::move(a), graph_steps,
                                             nodes_geometry);
  if (precision == 8)
    return std::make_unique<Solver3D<double>>(std::move(a), graph_steps,
                                              nodes_geometry);
  throw SmxError(SMX_ERR_ARG, "precision must be 4 or 8");
}

} // namespace smx
