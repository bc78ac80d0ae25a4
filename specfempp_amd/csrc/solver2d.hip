// gfx950 kernels and the device-resident solver for the 2D elastic isotropic
// P-SV time step.
//
// Kernels (each cites the reference dispatch it replaces):
//   k_stiffness   compute_stiffness_interaction (include/kokkos_kernels/impl/
//                 compute_stiffness_interaction.tpp:140-247): gather, GLL
//                 gradient, stress, J*sigma*dxi/dx, weighted divergence,
//                 Stacey traction, scatter-add into the acceleration.
//   k_sources     compute_source_interaction (compute_source_interaction.tpp:17-104)
//   k_predictor   newmark predictor (core/specfem/timescheme/newmark.tpp:67-128)
//   k_divide      divide_mass_matrix (include/kokkos_kernels/impl/divide_mass_matrix.tpp:13-65)
//   k_corrector   newmark corrector (newmark.tpp:13-62)
//   k_step_node   fused divide + corrector + next-step predictor (no
//                 reference counterpart: the three range kernels in one pass)
//   k_seismo      compute_seismograms (include/kokkos_kernels/impl/compute_seismogram.tpp:19-171)
//   k_mass/k_invert  compute_mass_matrix + invert_mass_matrix
//                 (compute_mass_matrix.tpp:23-122, invert_mass_matrix.tpp:13-55)
//
// Stiffness launch geometry: a 320-thread block (5 wavefronts) owns 64
// consecutive elements; wavefront w handles GLL column ix = w of all 64
// elements, lane l handles element tile*64 + l.  Element arrays are stored
// [point][element] so every per-point load of a wavefront is one contiguous
// 256 B (FP32) segment.  The element's displacement and the xi-direction
// stress integrand pass between the five column wavefronts through LDS
// ([component][point][lane], conflict-free); the gamma-direction terms stay
// in registers of the owning column.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "solver2d.hpp"

namespace smx {

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
};

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

constexpr int STIFF_THREADS = NGLL * 64;
constexpr int NODE_THREADS = 256;

// ------------------------------------------------------------------------
// Stiffness + Stacey + assembly.  Arithmetic follows the reference order:
// gradient (include/algorithms/gradient.hpp:40-82), stress
// (include/medium/dim2/elastic/isotropic/stress.hpp:38-78), F = J sigma
// dxi/dx (core/specfem/point/stress.hpp:166-183), divergence
// (include/algorithms/divergence.hpp:40-74), negation, Stacey
// (include/boundary_conditions/stacey/stacey.tpp:46-80), atomic add.
// One launch covers one element colour (slots [ebase, ebase + ecount)); no
// two elements of a colour share a GLL point, so the scatter into A is a
// plain read-modify-write: no atomics, and a fixed, run-to-run identical
// summation order (colour by colour).
template <typename real>
__global__ void __launch_bounds__(STIFF_THREADS)
    k_stiffness(const int *__restrict__ ibool, const real *__restrict__ geom,
                const real *__restrict__ mat, const int *__restrict__ sslot,
                const real *__restrict__ stacey, const real *__restrict__ U,
                const real *__restrict__ V, real *__restrict__ A, int ebase,
                int ecount, int ld, int sld, QuadArgs<real> q) {
  using V2 = typename Vec2<real>::T;
  __shared__ real su[2][NPT][64];
  __shared__ real sf[2][NPT][64];
  const int lane = threadIdx.x & 63;
  const int ix = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int local = blockIdx.x * 64 + lane;
  const int e = ebase + local;
  const bool valid = local < ecount;
  const size_t pld = size_t(NPT) * ld;
  const V2 *__restrict__ U2 = reinterpret_cast<const V2 *>(U);
  const V2 *__restrict__ V2p = reinterpret_cast<const V2 *>(V);
  V2 *__restrict__ A2 = reinterpret_cast<V2 *>(A);

  real hx[NGLL];  // l'_l(xi_ix), used by the xi-gradient
  real hxt[NGLL]; // l'_ix(xi_l), used by the xi-divergence
#pragma unroll
  for (int l = 0; l < NGLL; ++l) {
    hx[l] = q.hp[ix][l];
    hxt[l] = q.hp[l][ix];
  }
  const real wix = q.w[ix];

  int g[NGLL];
  real ux[NGLL], uz[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    const int p = iz * NGLL + ix;
    g[iz] = ibool[size_t(p) * ld + e];
    const V2 uu = U2[g[iz]];
    ux[iz] = uu.x;
    uz[iz] = uu.y;
    su[0][p][lane] = ux[iz];
    su[1][p][lane] = uz[iz];
  }
  __syncthreads();

  real f1x[NGLL], f1z[NGLL];
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    const int p = iz * NGLL + ix;
    real dxi_x = 0, dxi_z = 0, dga_x = 0, dga_z = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      dxi_x += hx[l] * su[0][iz * NGLL + l][lane];
      dga_x += q.hp[iz][l] * ux[l];
      dxi_z += hx[l] * su[1][iz * NGLL + l][lane];
      dga_z += q.hp[iz][l] * uz[l];
    }
    const size_t o = size_t(p) * ld + e;
    const real xix = geom[o], xiz = geom[pld + o], gax = geom[2 * pld + o],
               gaz = geom[3 * pld + o], jac = geom[4 * pld + o];
    const real l2m = mat[o], mu = mat[pld + o];
    const real lam = l2m - real(2.0) * mu;
    const real du_xx = xix * dxi_x + gax * dga_x;
    const real du_xz = xiz * dxi_x + gaz * dga_x;
    const real du_zx = xix * dxi_z + gax * dga_z;
    const real du_zz = xiz * dxi_z + gaz * dga_z;
    const real sxx = l2m * du_xx + lam * du_zz;
    const real szz = l2m * du_zz + lam * du_xx;
    const real sxz = mu * (du_xz + du_zx);
    sf[0][p][lane] = jac * (sxx * xix + sxz * xiz);
    sf[1][p][lane] = jac * (sxz * xix + szz * xiz);
    f1x[iz] = jac * (sxx * gax + sxz * gaz);
    f1z[iz] = jac * (sxz * gax + szz * gaz);
  }
  __syncthreads();

  const int slot = valid ? sslot[e] : -1;
#pragma unroll
  for (int iz = 0; iz < NGLL; ++iz) {
    real t1x = 0, t1z = 0, t2x = 0, t2z = 0;
#pragma unroll
    for (int l = 0; l < NGLL; ++l) {
      t1x += sf[0][iz * NGLL + l][lane] * hxt[l] * q.w[l];
      t1z += sf[1][iz * NGLL + l][lane] * hxt[l] * q.w[l];
      t2x += f1x[l] * q.hp[l][iz] * q.w[l];
      t2z += f1z[l] * q.hp[l][iz] * q.w[l];
    }
    real ax = q.w[iz] * t1x + wix * t2x;
    real az = q.w[iz] * t1z + wix * t2z;
    ax *= real(-1.0);
    az *= real(-1.0);
    if (slot >= 0) {
      const int p = iz * NGLL + ix;
      const size_t so = size_t(p) * sld + slot;
      const size_t sp = size_t(NPT) * sld;
      const real n0 = stacey[so], n1 = stacey[sp + so],
                 nrm = stacey[2 * sp + so], wgt = stacey[3 * sp + so],
                 rvp = stacey[4 * sp + so], rvs = stacey[5 * sp + so];
      const V2 vv = V2p[g[iz]];
      const real vx = vv.x, vz = vv.y;
      const real vn = vx * n0 + vz * n1;
      const real fx = ((vn * n0 / (nrm * nrm)) * (rvp - rvs)) + vx * rvs;
      const real fz = ((vn * n1 / (nrm * nrm)) * (rvp - rvs)) + vz * rvs;
      ax += real(-1.0) * fx * nrm * wgt;
      az += real(-1.0) * fz * nrm * wgt;
    }
    if (valid) {
      V2 acc = A2[g[iz]];
      acc.x += ax;
      acc.y += az;
      A2[g[iz]] = acc;
    }
  }
}

// ------------------------------------------------------------------------
// Point sources: A[ibool] += stf(step) * h_xi h_gamma f
// (include/medium/dim2/elastic/isotropic/source.hpp:36-59).
template <typename real>
__global__ void k_sources(real *__restrict__ A, const real *__restrict__ stf,
                          const real *__restrict__ src_array,
                          const int *__restrict__ src_ispec,
                          const int *__restrict__ ibool, int ld, int nsrc,
                          const int *__restrict__ step_ptr, int istep) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = t / NPT, p = t % NPT;
  if (s >= nsrc)
    return;
  const int step = step_ptr ? *step_ptr : istep;
  const int g = ibool[size_t(p) * ld + src_ispec[s]];
#pragma unroll
  for (int c = 0; c < 2; ++c)
    unsafeAtomicAdd(&A[2 * size_t(g) + c],
                    stf[(size_t(step) * nsrc + s) * 2 + c] *
                        src_array[(size_t(s) * 2 + c) * NPT + p]);
}

// ------------------------------------------------------------------------
// Newmark pieces on [nglob][2] arrays, 16 B per thread per array.
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
    k_divide(real *__restrict__ A, const real *__restrict__ Minv, long n2) {
  using VT = typename Vec16<real>::T;
  constexpr int W = Vec16<real>::W;
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i * W >= n2)
    return;
  real a[W], m[W];
  *reinterpret_cast<VT *>(a) = reinterpret_cast<const VT *>(A)[i];
  *reinterpret_cast<VT *>(m) = reinterpret_cast<const VT *>(Minv)[i];
#pragma unroll
  for (int k = 0; k < W; ++k)
    a[k] = a[k] * m[k];
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

// Divide + corrector of step n, then (if do_predict) the predictor of step
// n+1, in one pass; thread 0 advances the device step counter.
template <typename real>
__global__ void __launch_bounds__(NODE_THREADS)
    k_step_node(real *__restrict__ U, real *__restrict__ V,
                real *__restrict__ A, const real *__restrict__ Minv, long n2,
                real dt, real dto2, real dt2o2, int do_predict,
                int *__restrict__ step_ptr) {
  using VT = typename Vec16<real>::T;
  constexpr int W = Vec16<real>::W;
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0 && step_ptr)
    *step_ptr += 1;
  if (i * W >= n2)
    return;
  real u[W], v[W], a[W], m[W];
  *reinterpret_cast<VT *>(v) = reinterpret_cast<const VT *>(V)[i];
  *reinterpret_cast<VT *>(a) = reinterpret_cast<const VT *>(A)[i];
  *reinterpret_cast<VT *>(m) = reinterpret_cast<const VT *>(Minv)[i];
#pragma unroll
  for (int k = 0; k < W; ++k)
    divide_correct(a[k], v[k], m[k], dto2);
  if (do_predict) {
    *reinterpret_cast<VT *>(u) = reinterpret_cast<const VT *>(U)[i];
#pragma unroll
    for (int k = 0; k < W; ++k)
      predict(u[k], v[k], a[k], dt, dto2, dt2o2);
    reinterpret_cast<VT *>(U)[i] = *reinterpret_cast<VT *>(u);
  }
  reinterpret_cast<VT *>(V)[i] = *reinterpret_cast<VT *>(v);
  reinterpret_cast<VT *>(A)[i] = *reinterpret_cast<VT *>(a);
}

// ------------------------------------------------------------------------
// Seismograms: s = sum_{iz,ix} h(iz,ix) F[ibool] (interpolate_function,
// include/algorithms/interpolate.hpp).  raw != 0 reads the acceleration
// before division by the mass matrix and the velocity before the corrector
// (the fused step records between stiffness and the node pass) and applies
// the same divide/correct arithmetic per point.
struct SeisTypes {
  int t[4];
};

template <typename real>
__global__ void k_seismo(const real *__restrict__ U, const real *__restrict__ V,
                         const real *__restrict__ A,
                         const real *__restrict__ Minv, int raw, real dto2,
                         const int *__restrict__ rec_ispec,
                         const real *__restrict__ rec_h,
                         const int *__restrict__ ibool, int ld, int nrec,
                         SeisTypes types, int ntypes, int nsbs,
                         const int *__restrict__ step_ptr, int isig_direct,
                         real *__restrict__ out) {
  int isig = isig_direct;
  if (step_ptr) {
    const int step = *step_ptr;
    if (step % nsbs != 0)
      return;
    isig = step / nsbs;
  }
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = t % nrec, it = t / nrec;
  if (it >= ntypes)
    return;
  const int type = types.t[it];
  const int e = rec_ispec[r];
  real sx = 0, sz = 0;
  for (int p = 0; p < NPT; ++p) {
    const int g = ibool[size_t(p) * ld + e];
    const real h = rec_h[size_t(r) * NPT + p];
    real fx, fz;
    if (type == 0) {
      fx = U[2 * size_t(g)];
      fz = U[2 * size_t(g) + 1];
    } else if (!raw) {
      const real *F = type == 1 ? V : A;
      fx = F[2 * size_t(g)];
      fz = F[2 * size_t(g) + 1];
    } else {
      real ax = A[2 * size_t(g)], az = A[2 * size_t(g) + 1];
      real vx = V[2 * size_t(g)], vz = V[2 * size_t(g) + 1];
      divide_correct(ax, vx, Minv[2 * size_t(g)], dto2);
      divide_correct(az, vz, Minv[2 * size_t(g) + 1], dto2);
      fx = type == 1 ? vx : ax;
      fz = type == 1 ? vz : az;
    }
    sx += h * fx;
    sz += h * fz;
  }
  const size_t o = ((size_t(isig) * ntypes + it) * nrec + r) * 2;
  out[o] = sx;
  out[o + 1] = sz;
}

// ------------------------------------------------------------------------
// Mass matrix: M[ibool] += rho (w_z w_x J) + Stacey term with the velocity
// set to (-dt/2, -dt/2) (include/boundary_conditions/stacey/stacey.tpp:817-853).
// One launch per element colour (plain read-modify-write, as k_stiffness).
template <typename real>
__global__ void k_mass(real *__restrict__ M, const int *__restrict__ ibool,
                       const real *__restrict__ geom,
                       const real *__restrict__ kmr,
                       const int *__restrict__ sslot,
                       const real *__restrict__ stacey, int ebase, int ecount,
                       int ld, int sld, QuadArgs<real> q, real dt) {
  const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(NPT) * ecount)
    return;
  const int p = int(t / ecount), e = ebase + int(t % ecount);
  const size_t pld = size_t(NPT) * ld;
  const size_t o = size_t(p) * ld + e;
  const int iz = p / NGLL, ix = p % NGLL;
  const real rho = kmr[2 * pld + o];
  const real jac = geom[4 * pld + o];
  const real f = q.w[iz] * q.w[ix] * jac;
  real m0 = rho * f, m1 = rho * f;
  const int slot = sslot[e];
  if (slot >= 0) {
    const size_t so = size_t(p) * sld + slot, sp = size_t(NPT) * sld;
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
  const int g = ibool[o];
  M[2 * size_t(g)] += m0;
  M[2 * size_t(g) + 1] += m1;
}

__global__ void k_set_step(int *__restrict__ step_ptr, int value) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *step_ptr = value;
}

template <typename real>
__global__ void k_invert(real *__restrict__ M, long n) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    M[i] = real(1.0) / M[i];
}

// ========================================================================
template <typename real> class Solver2D final : public Solver2DBase {
public:
  Solver2D(std::unique_ptr<HostAssembly2D> a) : asm_(std::move(a)) {
    const HostAssembly2D &A = *asm_;
    SMX_HIP(hipGetDevice(&device_));
    SMX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    SMX_HIP(hipEventCreate(&ev0_));
    SMX_HIP(hipEventCreate(&ev1_));
    nglob_pad_ = (A.nglob + 255) / 256 * 256;
    n2_ = 2L * nglob_pad_;
    ibool_ = upload(A.ibool);
    geom_ = upload_real(A.geom);
    mat_ = upload_real(A.lam2mu_mu);
    kmr_ = upload_real(A.kappa_mu_rho);
    sslot_ = upload(A.stacey_slot);
    stacey_ = upload_real(A.stacey);
    src_ispec_ = upload(A.src_ispec);
    src_array_ = upload_real(A.src_array);
    stf_ = upload_real(A.stf);
    rec_ispec_ = upload(A.rec_ispec);
    rec_h_ = upload_real(A.rec_h);
    minv_ = alloc<real>(n2_);
    SMX_HIP(hipMalloc(&step_, sizeof(int)));
    allocs_.push_back(step_);
    const size_t nseis = size_t(std::max(1, A.nsig)) * A.seis_types.size() *
                         std::max(1, A.nrec) * 2;
    seis_ = alloc<real>(std::max<size_t>(nseis, 2));
    for (int i = 0; i < NGLL; ++i) {
      quad_.w[i] = A.quad.w[i];
      for (int j = 0; j < NGLL; ++j)
        quad_.hp[i][j] = A.quad.hprime[i][j];
    }
    for (int k = 0; k < 4; ++k)
      stypes_.t[k] = k < int(A.seis_types.size()) ? A.seis_types[k] : 0;
    if (A.seis_types.size() > 4)
      throw SmxError(SMX_ERR_ARG, "at most 4 seismogram types");
    dt_ = A.dt;
    dto2_ = A.dt_over2;
    dt2o2_ = A.dt2_over2;
    field(SMX_FIELD_FORWARD);
    SMX_HIP(hipStreamSynchronize(stream_));
  }

  ~Solver2D() override {
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

  void initialize() override {
    const HostAssembly2D &A = *asm_;
    real *M = minv_;
    SMX_HIP(hipMemsetAsync(M, 0, n2_ * sizeof(real), stream_));
    for (size_t c = 0; c < A.color_count.size(); ++c) {
      const int cnt = A.color_count[c];
      if (cnt == 0)
        continue;
      const long n = long(NPT) * cnt;
      k_mass<real><<<dim3((n + 255) / 256), dim3(256), 0, stream_>>>(
          M, ibool_, geom_, kmr_, sslot_, stacey_, A.color_offset[c], cnt,
          A.nspec_pad, A.nstacey_pad, quad_, real(A.dt));
      SMX_HIP(hipGetLastError());
    }
    const long m = 2L * A.nglob;
    k_invert<real><<<dim3((m + 255) / 256), dim3(256), 0, stream_>>>(M, m);
    SMX_HIP(hipGetLastError());
    SMX_HIP(hipStreamSynchronize(stream_));
    initialized_ = true;
  }

  int predictor(int f) override {
    Field &F = field(f);
    launch_predictor(F, dt_, dto2_, dt2o2_);
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
    launch_sources(F, nullptr, istep);
    launch_stiffness(F.u, F.v, F.a);
    k_divide<real><<<node_grid(), dim3(NODE_THREADS), 0, stream_>>>(
        F.a, minv_, n2_);
    SMX_HIP(hipGetLastError());
    return A.nspec;
  }

  void compute_seismograms(int f, int isig) override {
    const HostAssembly2D &A = *asm_;
    if (isig < 0 || isig >= A.nsig)
      throw SmxError(SMX_ERR_ARG, "seismogram step outside [0, nsig)");
    Field &F = field(f);
    launch_seismo(F, /*raw=*/0, nullptr, isig);
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
    launch_predictor(F, dt_, dto2_, dt2o2_);
    const int body = nstep - 1; // steps followed by a fused predictor
    const int reps = body / kGraphSteps;
    if (reps > 0) {
      build_graph(F);
      for (int r = 0; r < reps; ++r)
        SMX_HIP(hipGraphLaunch(graph_exec_, stream_));
    }
    for (int s = reps * kGraphSteps; s < body; ++s)
      fused_step(F, /*do_predict=*/1);
    fused_step(F, /*do_predict=*/0);
    SMX_HIP(hipEventRecord(ev1_, stream_));
    last_launches_ = nstep;
    have_stats_ = true;
  }

  void synchronize() override { SMX_HIP(hipStreamSynchronize(stream_)); }

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

  double time_stiffness(int n) override {
    if (n < 1)
      throw SmxError(SMX_ERR_ARG, "n must be positive");
    Field &F = field(SMX_FIELD_FORWARD);
    if (!scratch_a_)
      scratch_a_ = alloc<real>(n2_);
    launch_stiffness(F.u, F.v, scratch_a_); // warm-up
    SMX_HIP(hipEventRecord(ev0_, stream_));
    for (int i = 0; i < n; ++i)
      launch_stiffness(F.u, F.v, scratch_a_);
    SMX_HIP(hipEventRecord(ev1_, stream_));
    SMX_HIP(hipEventSynchronize(ev1_));
    float t = 0;
    SMX_HIP(hipEventElapsedTime(&t, ev0_, ev1_));
    last_stiff_ms_ = double(t) / n;
    have_stats_ = false;
    return last_stiff_ms_;
  }

private:
  struct Field {
    real *u = nullptr, *v = nullptr, *a = nullptr;
  };
  static constexpr int kGraphSteps = 8;

  template <typename T> T *alloc(size_t n) {
    void *p = nullptr;
    SMX_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    allocs_.push_back(p);
    SMX_HIP(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), stream_));
    return static_cast<T *>(p);
  }
  int *upload(const std::vector<int32_t> &h) {
    int *d = alloc<int>(h.size());
    SMX_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(int),
                           hipMemcpyHostToDevice, stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
    return d;
  }
  real *upload_real(const std::vector<float> &h) {
    real *d = alloc<real>(h.size());
    std::vector<real> tmp(h.begin(), h.end());
    SMX_HIP(hipMemcpyAsync(d, tmp.data(), tmp.size() * sizeof(real),
                           hipMemcpyHostToDevice, stream_));
    SMX_HIP(hipStreamSynchronize(stream_));
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

  void launch_predictor(Field &F, real dt, real dto2, real dt2o2) {
    k_predictor<real><<<node_grid(), dim3(NODE_THREADS), 0, stream_>>>(
        F.u, F.v, F.a, n2_, dt, dto2, dt2o2);
    SMX_HIP(hipGetLastError());
  }
  void launch_sources(Field &F, const int *step_ptr, int istep) {
    const HostAssembly2D &A = *asm_;
    if (A.nsrc == 0)
      return;
    const int n = A.nsrc * NPT;
    k_sources<real><<<dim3((n + 127) / 128), dim3(128), 0, stream_>>>(
        F.a, stf_, src_array_, src_ispec_, ibool_, A.nspec_pad, A.nsrc,
        step_ptr, istep);
    SMX_HIP(hipGetLastError());
  }
  void launch_stiffness(const real *U, const real *V, real *Aout) {
    const HostAssembly2D &A = *asm_;
    for (size_t c = 0; c < A.color_count.size(); ++c) {
      const int cnt = A.color_count[c];
      if (cnt == 0)
        continue;
      k_stiffness<real><<<dim3((cnt + 63) / 64), dim3(STIFF_THREADS), 0,
                           stream_>>>(ibool_, geom_, mat_, sslot_, stacey_, U,
                                      V, Aout, A.color_offset[c], cnt,
                                      A.nspec_pad, A.nstacey_pad, quad_);
      SMX_HIP(hipGetLastError());
    }
  }
  void launch_seismo(Field &F, int raw, const int *step_ptr, int isig) {
    const HostAssembly2D &A = *asm_;
    const int ntypes = int(A.seis_types.size());
    if (A.nrec == 0 || ntypes == 0)
      return;
    const int n = A.nrec * ntypes;
    k_seismo<real><<<dim3((n + 127) / 128), dim3(128), 0, stream_>>>(
        F.u, F.v, F.a, minv_, raw, dto2_, rec_ispec_, rec_h_, ibool_,
        A.nspec_pad, A.nrec, stypes_, ntypes, A.nsbs, step_ptr, isig, seis_);
    SMX_HIP(hipGetLastError());
  }
  // One fused step: sources + stiffness, seismograms from the raw
  // acceleration, node pass (divide + corrector [+ next predictor]).
  void fused_step(Field &F, int do_predict) {
    launch_sources(F, step_, 0);
    launch_stiffness(F.u, F.v, F.a);
    launch_seismo(F, /*raw=*/1, step_, 0);
    k_step_node<real><<<node_grid(), dim3(NODE_THREADS), 0, stream_>>>(
        F.u, F.v, F.a, minv_, n2_, dt_, dto2_, dt2o2_, do_predict, step_);
    SMX_HIP(hipGetLastError());
  }
  void build_graph(Field &F) {
    if (graph_exec_)
      return;
    SMX_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      for (int s = 0; s < kGraphSteps; ++s)
        fused_step(F, 1);
    } catch (...) {
      hipGraph_t g;
      (void)hipStreamEndCapture(stream_, &g);
      throw;
    }
    SMX_HIP(hipStreamEndCapture(stream_, &graph_));
    SMX_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  }

  std::unique_ptr<HostAssembly2D> asm_;
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<void *> allocs_;
  long nglob_pad_ = 0, n2_ = 0;
  int *ibool_ = nullptr, *sslot_ = nullptr, *src_ispec_ = nullptr,
      *rec_ispec_ = nullptr, *step_ = nullptr;
  real *geom_ = nullptr, *mat_ = nullptr, *kmr_ = nullptr, *stacey_ = nullptr,
       *src_array_ = nullptr, *stf_ = nullptr, *rec_h_ = nullptr,
       *minv_ = nullptr, *seis_ = nullptr, *scratch_a_ = nullptr;
  Field fields_[3];
  QuadArgs<real> quad_;
  SeisTypes stypes_;
  real dt_ = 0, dto2_ = 0, dt2o2_ = 0;
  bool initialized_ = false, have_stats_ = false;
  double last_stiff_ms_ = 0;
  int last_launches_ = 0;
};

std::unique_ptr<Solver2DBase> make_solver2d(std::unique_ptr<HostAssembly2D> a,
                                            int precision) {
  if (precision == 4)
    return std::make_unique<Solver2D<float>>(std::move(a));
  if (precision == 8)
    return std::make_unique<Solver2D<double>>(std::move(a));
  throw SmxError(SMX_ERR_ARG, "precision must be 4 or 8");
}

} // namespace smx
