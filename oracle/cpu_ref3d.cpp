// ORACLE — test infrastructure only.
//
// CPU restatement of SPECFEM++'s 3D elastic isotropic forward path (hex8
// meshes, NGLL = 5), written to check the HIP implementation in
// specfempp_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this code, and only as the checker / baseline
// timer.  The product never links it.
//
// Routines follow the reference (file:line cited at each) with its float /
// double promotions; the FP64 instantiation is the same algorithm in double.
// The reference's 3D path has no absorbing boundary (its dim3 stiffness is
// instantiated only for BOUNDARY_TAG(NONE),
// src/kokkos_kernels/impl/compute_stiffness_interaction.cpp:53-63); the
// Stacey terms here carry the 2D formula over to faces and are this build's
// own design, pinned by closed forms rather than reference data.
//
// Parity pins (tests/test_oracle3d.py):
//   * SPECFEM3D golden traces shipped with the reference
//     (HomogeneousHalfspaceSmallNoABCForceSource), reference metric <= 1e-3
//   * rigid motion gives zero force, K symmetric, sum(M) = rho * volume,
//     affine Jacobian, moment-tensor source array moments
//   * Stacey: normal-incidence impedance of a face, energy decay
//   * discrete adjoint (tests/test_oracle3d_adjoint.py): FP64 central
//     differences of a waveform misfit along random model directions (error
//     scaling as the step squared), zero source -> zero gradient, linearity
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

#include "gll_common.hpp"

constexpr int NP3 = NGLL * NGLL * NGLL;

inline int pidx(int iz, int iy, int ix) { return (iz * NGLL + iy) * NGLL + ix; }

// Hex8 shape functions and derivatives: computed in double (the reference's
// literals promote), stored in real
// (core/specfem/shape_function/dim3/shape_functions3d.cpp:9-23,
//  shape_function_derivatives3d.cpp:9-40).
template <typename real>
void shape8(real xi, real eta, real ga, real *N) {
  const double x = xi, e = eta, g = ga;
  N[0] = 0.125 * (1.0 - x) * (1.0 - e) * (1.0 - g);
  N[1] = 0.125 * (1.0 + x) * (1.0 - e) * (1.0 - g);
  N[2] = 0.125 * (1.0 + x) * (1.0 + e) * (1.0 - g);
  N[3] = 0.125 * (1.0 - x) * (1.0 + e) * (1.0 - g);
  N[4] = 0.125 * (1.0 - x) * (1.0 - e) * (1.0 + g);
  N[5] = 0.125 * (1.0 + x) * (1.0 - e) * (1.0 + g);
  N[6] = 0.125 * (1.0 + x) * (1.0 + e) * (1.0 + g);
  N[7] = 0.125 * (1.0 - x) * (1.0 + e) * (1.0 + g);
}

template <typename real>
void dshape8(real xi, real eta, real ga, real d[3][8]) {
  const double x = xi, e = eta, g = ga;
  d[0][0] = -0.125 * (1.0 - e) * (1.0 - g);
  d[0][1] = 0.125 * (1.0 - e) * (1.0 - g);
  d[0][2] = 0.125 * (1.0 + e) * (1.0 - g);
  d[0][3] = -0.125 * (1.0 + e) * (1.0 - g);
  d[0][4] = -0.125 * (1.0 - e) * (1.0 + g);
  d[0][5] = 0.125 * (1.0 - e) * (1.0 + g);
  d[0][6] = 0.125 * (1.0 + e) * (1.0 + g);
  d[0][7] = -0.125 * (1.0 + e) * (1.0 + g);
  d[1][0] = -0.125 * (1.0 - x) * (1.0 - g);
  d[1][1] = -0.125 * (1.0 + x) * (1.0 - g);
  d[1][2] = 0.125 * (1.0 + x) * (1.0 - g);
  d[1][3] = 0.125 * (1.0 - x) * (1.0 - g);
  d[1][4] = -0.125 * (1.0 - x) * (1.0 + g);
  d[1][5] = -0.125 * (1.0 + x) * (1.0 + g);
  d[1][6] = 0.125 * (1.0 + x) * (1.0 + g);
  d[1][7] = 0.125 * (1.0 - x) * (1.0 + g);
  d[2][0] = -0.125 * (1.0 - x) * (1.0 - e);
  d[2][1] = -0.125 * (1.0 + x) * (1.0 - e);
  d[2][2] = -0.125 * (1.0 + x) * (1.0 + e);
  d[2][3] = -0.125 * (1.0 - x) * (1.0 + e);
  d[2][4] = 0.125 * (1.0 - x) * (1.0 - e);
  d[2][5] = 0.125 * (1.0 + x) * (1.0 - e);
  d[2][6] = 0.125 * (1.0 + x) * (1.0 + e);
  d[2][7] = 0.125 * (1.0 - x) * (1.0 + e);
}

template <typename real> struct Jac3 {
  // inverse map: xi_{x,y,z}, eta_{x,y,z}, gamma_{x,y,z}; jacobian
  real xix, xiy, xiz, etax, etay, etaz, gax, gay, gaz, jac;
};

// core/specfem/jacobian/dim3/jacobian.cpp:27-74 (accumulation in real)
template <typename real>
Jac3<real> jacobian3(const real *cx, const real *cy, const real *cz,
                     const real d[3][8]) {
  real xxi = 0, yxi = 0, zxi = 0, xeta = 0, yeta = 0, zeta = 0, xga = 0,
       yga = 0, zga = 0;
  for (int n = 0; n < 8; ++n) {
    xxi += d[0][n] * cx[n];
    yxi += d[0][n] * cy[n];
    zxi += d[0][n] * cz[n];
    xeta += d[1][n] * cx[n];
    yeta += d[1][n] * cy[n];
    zeta += d[1][n] * cz[n];
    xga += d[2][n] * cx[n];
    yga += d[2][n] * cy[n];
    zga += d[2][n] * cz[n];
  }
  Jac3<real> J;
  J.jac = xxi * (yeta * zga - yga * zeta) - xeta * (yxi * zga - yga * zxi) +
          xga * (yxi * zeta - yeta * zxi);
  J.xix = (yeta * zga - yga * zeta) / J.jac;
  J.xiy = (xga * zeta - xeta * zga) / J.jac;
  J.xiz = (xeta * yga - xga * yeta) / J.jac;
  J.etax = (yga * zxi - yxi * zga) / J.jac;
  J.etay = (xxi * zga - xga * zxi) / J.jac;
  J.etaz = (xga * yxi - xxi * yga) / J.jac;
  J.gax = (yxi * zeta - yeta * zxi) / J.jac;
  J.gay = (xeta * zxi - xxi * zeta) / J.jac;
  J.gaz = (xxi * yeta - xeta * yxi) / J.jac;
  return J;
}

template <typename real> struct Oracle3D {
  GLL<real> gll;
  int nspec = 0, nglob = 0, nstep = 0;
  std::vector<real> cx, cy, cz;     // [nspec][8]
  std::vector<int> ibool;           // [nspec][125]
  std::vector<Jac3<real>> jac;      // [nspec][125]
  std::vector<real> l2m, mu, rho;   // [nspec][125]
  std::vector<real> minv;           // [nglob][3]
  std::vector<real> u, v, a;        // [nglob][3]
  // Stacey per local point (0 weight = none): normal (3), |n|, weight, P, S
  std::vector<std::array<real, 7>> st;
  std::vector<uint8_t> elem_st;
  // sources: element, array [3][125]; stf [nstep][nsrc]
  int nsrc = 0;
  std::vector<int> src_e;
  std::vector<real> src_arr, stf;
  double t0 = 0;
  // receivers: element, interpolant [125]
  int nrec = 0;
  std::vector<int> rec_e;
  std::vector<real> rec_h;
  real dt = 0, dto2 = 0, dt2o2 = 0;

  // ------------------------------------------------------------------ set-up
  void build(int nspec_, int nnodes, const double *coord, const int *knods,
             const float *kappa, const float *mu_, const float *rho_, int nabs,
             const int *abs_e, const int *abs_f) {
    nspec = nspec_;
    cx.resize(size_t(nspec) * 8);
    cy.resize(cx.size());
    cz.resize(cx.size());
    for (int e = 0; e < nspec; ++e)
      for (int n = 0; n < 8; ++n) {
        const int id = knods[e * 8 + n];
        if (id < 0 || id >= nnodes)
          throw std::runtime_error("knods out of range");
        cx[e * 8 + n] = static_cast<real>(coord[3 * size_t(id)]);
        cy[e * 8 + n] = static_cast<real>(coord[3 * size_t(id) + 1]);
        cz[e * 8 + n] = static_cast<real>(coord[3 * size_t(id) + 2]);
      }
    // Jacobian at every GLL point
    jac.resize(size_t(nspec) * NP3);
    real dN[NP3][3][8];
    for (int iz = 0; iz < NGLL; ++iz)
      for (int iy = 0; iy < NGLL; ++iy)
        for (int ix = 0; ix < NGLL; ++ix)
          dshape8(gll.xi[ix], gll.xi[iy], gll.xi[iz], dN[pidx(iz, iy, ix)]);
    for (int e = 0; e < nspec; ++e)
      for (int p = 0; p < NP3; ++p) {
        jac[size_t(e) * NP3 + p] =
            jacobian3(&cx[e * 8], &cy[e * 8], &cz[e * 8], dN[p]);
        if (!(jac[size_t(e) * NP3 + p].jac > 0))
          throw std::runtime_error("non-positive Jacobian");
      }
    // global numbering: GLL points at equal coordinates (rounded to 1e-6 of
    // the element size) are one point; ids in order of first touch
    number_points();
    // material, as the reference forms it (medium/dim2/elastic/isotropic/
    // properties.hpp:71-91 shared by dim3): lambda+2mu = kappa + 4/3 mu
    l2m.resize(size_t(nspec) * NP3);
    mu.resize(l2m.size());
    rho.resize(l2m.size());
    for (size_t i = 0; i < l2m.size(); ++i) {
      l2m[i] = static_cast<real>(kappa[i]) +
               static_cast<real>(4.0 / 3.0) * static_cast<real>(mu_[i]);
      mu[i] = mu_[i];
      rho[i] = rho_[i];
    }
    // absorbing faces: 1 bottom (gamma=-1), 2 right (xi=+1), 3 top, 4 left,
    // 5 front (eta=-1), 6 back (include/enumerations/dim3/mesh_entities.hpp)
    st.assign(size_t(nspec) * NP3, {0, 0, 0, 1, 0, 0, 0});
    elem_st.assign(nspec, 0);
    for (int k = 0; k < nabs; ++k) {
      const int e = abs_e[k], f = abs_f[k];
      if (e < 0 || e >= nspec || f < 1 || f > 6)
        throw std::runtime_error("invalid absorbing face");
      elem_st[e] = 1;
      for (int iz = 0; iz < NGLL; ++iz)
        for (int iy = 0; iy < NGLL; ++iy)
          for (int ix = 0; ix < NGLL; ++ix) {
            const bool on = (f == 1 && iz == 0) || (f == 3 && iz == NGLL - 1) ||
                            (f == 4 && ix == 0) || (f == 2 && ix == NGLL - 1) ||
                            (f == 5 && iy == 0) || (f == 6 && iy == NGLL - 1);
            if (!on)
              continue;
            const size_t i = size_t(e) * NP3 + pidx(iz, iy, ix);
            const Jac3<real> &J = jac[i];
            real n[3], w;
            if (f == 2 || f == 4) {
              const real s = f == 2 ? real(1) : real(-1);
              n[0] = s * J.xix * J.jac;
              n[1] = s * J.xiy * J.jac;
              n[2] = s * J.xiz * J.jac;
              w = gll.w[iy] * gll.w[iz];
            } else if (f == 5 || f == 6) {
              const real s = f == 6 ? real(1) : real(-1);
              n[0] = s * J.etax * J.jac;
              n[1] = s * J.etay * J.jac;
              n[2] = s * J.etaz * J.jac;
              w = gll.w[ix] * gll.w[iz];
            } else {
              const real s = f == 3 ? real(1) : real(-1);
              n[0] = s * J.gax * J.jac;
              n[1] = s * J.gay * J.jac;
              n[2] = s * J.gaz * J.jac;
              w = gll.w[ix] * gll.w[iy];
            }
            const real nrm = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            const real P = std::sqrt(l2m[i] / rho[i]) * rho[i];
            const real S = std::sqrt(rho[i] * mu[i]);
            st[i] = {n[0], n[1], n[2], nrm, w, P, S};
          }
    }
  }

  void number_points() {
    ibool.assign(size_t(nspec) * NP3, -1);
    std::unordered_map<uint64_t, int> ids;
    ids.reserve(size_t(nspec) * 70);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t k = 0; k < cx.size(); ++k) {
      const double c[3] = {cx[k], cy[k], cz[k]};
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], c[d]);
        hi[d] = std::max(hi[d], c[d]);
      }
    }
    double vol = 1;
    for (int d = 0; d < 3; ++d)
      vol *= std::max(hi[d] - lo[d], 1e-30);
    const double q = 1e-4 * std::cbrt(vol / nspec); // quantum
    int next = 0;
    for (int e = 0; e < nspec; ++e)
      for (int iz = 0; iz < NGLL; ++iz)
        for (int iy = 0; iy < NGLL; ++iy)
          for (int ix = 0; ix < NGLL; ++ix) {
            double N[8];
            shape8<double>(double(gll.xi[ix]), double(gll.xi[iy]),
                           double(gll.xi[iz]), N);
            double x = 0, y = 0, z = 0;
            for (int n = 0; n < 8; ++n) {
              x += N[n] * cx[e * 8 + n];
              y += N[n] * cy[e * 8 + n];
              z += N[n] * cz[e * 8 + n];
            }
            const uint64_t kx = uint64_t(std::llround((x - lo[0]) / q)) & 0x1fffff;
            const uint64_t ky = uint64_t(std::llround((y - lo[1]) / q)) & 0x1fffff;
            const uint64_t kz = uint64_t(std::llround((z - lo[2]) / q)) & 0x1fffff;
            const uint64_t key = (kx << 42) | (ky << 21) | kz;
            auto it = ids.emplace(key, next);
            if (it.second)
              ++next;
            ibool[size_t(e) * NP3 + pidx(iz, iy, ix)] = it.first->second;
          }
    nglob = next;
  }

  // Mass with the Stacey term at v = (-dt/2, -dt/2, -dt/2), inverted
  // (compute_mass_matrix.tpp:23-122, medium/dim3/elastic/isotropic/
  // mass_matrix.tpp, stacey.tpp:817-853 carried to faces).
  void mass() {
    std::vector<real> M(size_t(nglob) * 3, 0);
    for (int e = 0; e < nspec; ++e)
      for (int iz = 0; iz < NGLL; ++iz)
        for (int iy = 0; iy < NGLL; ++iy)
          for (int ix = 0; ix < NGLL; ++ix) {
            const int p = pidx(iz, iy, ix);
            const size_t i = size_t(e) * NP3 + p;
            const real f = gll.w[iz] * gll.w[iy] * gll.w[ix] * jac[i].jac;
            real m[3] = {rho[i] * f, rho[i] * f, rho[i] * f};
            if (elem_st[e] && st[i][4] != real(0)) {
              const real vf = real(-1.0) * dt * real(0.5);
              const real vv[3] = {vf, vf, vf};
              real t[3];
              traction(i, vv, t);
              for (int c = 0; c < 3; ++c)
                m[c] += t[c];
            }
            const int g = ibool[i];
            for (int c = 0; c < 3; ++c)
              M[3 * size_t(g) + c] += m[c];
          }
    minv.resize(M.size());
    for (size_t k = 0; k < M.size(); ++k)
      minv[k] = real(1.0) / M[k];
  }

  // Stacey traction contribution (added to the acceleration accumulator).
  void traction(size_t i, const real *vv, real *t) const {
    const auto &s = st[i];
    const real n0 = s[0], n1 = s[1], n2 = s[2], nrm = s[3], w = s[4],
               P = s[5], S = s[6];
    const real vn = vv[0] * n0 + vv[1] * n1 + vv[2] * n2;
    const real n[3] = {n0, n1, n2};
    for (int c = 0; c < 3; ++c) {
      const real fc = ((vn * n[c] / (nrm * nrm)) * (P - S)) + vv[c] * S;
      t[c] = real(-1.0) * fc * nrm * w;
    }
  }

  // ------------------------------------------------------------- stiffness
  // gradient (algorithms/gradient.hpp dim3), stress
  // (medium/dim3/elastic/isotropic/stress.hpp:46-99), F = J sigma dxi/dx
  // (point/stress.hpp:217-240), divergence (algorithms/divergence.hpp:95-137),
  // then the sign and the Stacey traction.
  void stiffness(const std::vector<real> &U, const std::vector<real> &V,
                 std::vector<real> &A, bool with_stacey = true) const {
    real f[NP3][3];
    real F[NP3][3][3];
    for (int e = 0; e < nspec; ++e) {
      for (int p = 0; p < NP3; ++p) {
        const int g = ibool[size_t(e) * NP3 + p];
        for (int c = 0; c < 3; ++c)
          f[p][c] = U[3 * size_t(g) + c];
      }
      for (int iz = 0; iz < NGLL; ++iz)
        for (int iy = 0; iy < NGLL; ++iy)
          for (int ix = 0; ix < NGLL; ++ix) {
            const int p = pidx(iz, iy, ix);
            const size_t i = size_t(e) * NP3 + p;
            real dxi[3] = {0, 0, 0}, deta[3] = {0, 0, 0}, dga[3] = {0, 0, 0};
            for (int l = 0; l < NGLL; ++l)
              for (int c = 0; c < 3; ++c) {
                dxi[c] += gll.hprime[ix][l] * f[pidx(iz, iy, l)][c];
                deta[c] += gll.hprime[iy][l] * f[pidx(iz, l, ix)][c];
                dga[c] += gll.hprime[iz][l] * f[pidx(l, iy, ix)][c];
              }
            const Jac3<real> &J = jac[i];
            real du[3][3];
            for (int c = 0; c < 3; ++c) {
              du[c][0] = J.xix * dxi[c] + J.etax * deta[c] + J.gax * dga[c];
              du[c][1] = J.xiy * dxi[c] + J.etay * deta[c] + J.gay * dga[c];
              du[c][2] = J.xiz * dxi[c] + J.etaz * deta[c] + J.gaz * dga[c];
            }
            const real L = l2m[i], M = mu[i];
            const real lam = L - real(2.0) * M;
            real T[3][3];
            T[0][0] = L * du[0][0] + lam * (du[1][1] + du[2][2]);
            T[1][1] = L * du[1][1] + lam * (du[0][0] + du[2][2]);
            T[2][2] = L * du[2][2] + lam * (du[0][0] + du[1][1]);
            T[0][1] = T[1][0] = M * (du[0][1] + du[1][0]);
            T[0][2] = T[2][0] = M * (du[0][2] + du[2][0]);
            T[1][2] = T[2][1] = M * (du[1][2] + du[2][1]);
            for (int c = 0; c < 3; ++c) {
              F[p][c][0] = J.jac * (T[c][0] * J.xix + T[c][1] * J.xiy + T[c][2] * J.xiz);
              F[p][c][1] = J.jac * (T[c][0] * J.etax + T[c][1] * J.etay + T[c][2] * J.etaz);
              F[p][c][2] = J.jac * (T[c][0] * J.gax + T[c][1] * J.gay + T[c][2] * J.gaz);
            }
          }
      for (int iz = 0; iz < NGLL; ++iz)
        for (int iy = 0; iy < NGLL; ++iy)
          for (int ix = 0; ix < NGLL; ++ix) {
            const int p = pidx(iz, iy, ix);
            const size_t i = size_t(e) * NP3 + p;
            real t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0}, t3[3] = {0, 0, 0};
            for (int l = 0; l < NGLL; ++l)
              for (int c = 0; c < 3; ++c) {
                t1[c] += F[pidx(iz, iy, l)][c][0] * gll.hprime[l][ix] * gll.w[l];
                t2[c] += F[pidx(iz, l, ix)][c][1] * gll.hprime[l][iy] * gll.w[l];
                t3[c] += F[pidx(l, iy, ix)][c][2] * gll.hprime[l][iz] * gll.w[l];
              }
            real r[3];
            for (int c = 0; c < 3; ++c) {
              r[c] = gll.w[iz] * gll.w[iy] * t1[c] + gll.w[iz] * gll.w[ix] * t2[c] +
                     gll.w[iy] * gll.w[ix] * t3[c];
              r[c] *= real(-1.0);
            }
            const int g = ibool[i];
            if (with_stacey && elem_st[e] && st[i][4] != real(0)) {
              const real vv[3] = {V[3 * size_t(g)], V[3 * size_t(g) + 1],
                                  V[3 * size_t(g) + 2]};
              real t[3];
              traction(i, vv, t);
              for (int c = 0; c < 3; ++c)
                r[c] += t[c];
            }
            for (int c = 0; c < 3; ++c)
              A[3 * size_t(g) + c] += r[c];
          }
    }
  }

  // ------------------------------------------------------ sources/receivers
  // Point location: nearest GLL point, then Newton on the candidate elements
  // sharing one of its corners (src/algorithms/dim3/locate_point_impl.cpp).
  void locate(double xd, double yd, double zd, int &e_out, real lc[3]) const {
    const real x = real(xd), y = real(yd), z = real(zd);
    real best = std::numeric_limits<real>::max();
    int ebest = -1;
    for (int e = 0; e < nspec; ++e)
      for (int iz = 0; iz < NGLL; ++iz)
        for (int iy = 0; iy < NGLL; ++iy)
          for (int ix = 0; ix < NGLL; ++ix) {
            real N[8];
            shape8(gll.xi[ix], gll.xi[iy], gll.xi[iz], N);
            real px = 0, py = 0, pz = 0;
            for (int n = 0; n < 8; ++n) {
              px += N[n] * cx[e * 8 + n];
              py += N[n] * cy[e * 8 + n];
              pz += N[n] * cz[e * 8 + n];
            }
            const real d = (px - x) * (px - x) + (py - y) * (py - y) +
                           (pz - z) * (pz - z);
            if (d < best) {
              best = d;
              ebest = e;
            }
          }
    // candidates: elements sharing a corner with the best one
    std::vector<int> cand = {ebest};
    for (int e = 0; e < nspec; ++e) {
      if (e == ebest)
        continue;
      bool share = false;
      for (int a = 0; a < 8 && !share; ++a)
        for (int b = 0; b < 8 && !share; ++b)
          share = cx[e * 8 + a] == cx[ebest * 8 + b] &&
                  cy[e * 8 + a] == cy[ebest * 8 + b] &&
                  cz[e * 8 + a] == cz[ebest * 8 + b];
      if (share)
        cand.push_back(e);
    }
    real dbest = std::numeric_limits<real>::max();
    e_out = -1;
    for (int e : cand) {
      real xi = 0, eta = 0, ga = 0;
      for (int it = 0; it < 100; ++it) {
        real N[8], d[3][8];
        shape8(xi, eta, ga, N);
        dshape8(xi, eta, ga, d);
        real px = 0, py = 0, pz = 0;
        for (int n = 0; n < 8; ++n) {
          px += N[n] * cx[e * 8 + n];
          py += N[n] * cy[e * 8 + n];
          pz += N[n] * cz[e * 8 + n];
        }
        const Jac3<real> J = jacobian3(&cx[e * 8], &cy[e * 8], &cz[e * 8], d);
        const real dx = x - px, dy = y - py, dz = z - pz;
        const real dxi = J.xix * dx + J.xiy * dy + J.xiz * dz;
        const real deta = J.etax * dx + J.etay * dy + J.etaz * dz;
        const real dga = J.gax * dx + J.gay * dy + J.gaz * dz;
        xi = std::min(real(1.01), std::max(real(-1.01), xi + dxi));
        eta = std::min(real(1.01), std::max(real(-1.01), eta + deta));
        ga = std::min(real(1.01), std::max(real(-1.01), ga + dga));
        if (std::abs(dxi) < 1e-12 && std::abs(deta) < 1e-12 && std::abs(dga) < 1e-12)
          break;
      }
      real N[8];
      shape8(xi, eta, ga, N);
      real px = 0, py = 0, pz = 0;
      for (int n = 0; n < 8; ++n) {
        px += N[n] * cx[e * 8 + n];
        py += N[n] * cy[e * 8 + n];
        pz += N[n] * cz[e * 8 + n];
      }
      const real d = std::sqrt((x - px) * (x - px) + (y - py) * (y - py) +
                               (z - pz) * (z - pz));
      if (d < dbest) {
        dbest = d;
        e_out = e;
        lc[0] = xi;
        lc[1] = eta;
        lc[2] = ga;
      }
    }
    for (int k = 0; k < 3; ++k)
      if (std::fabs(std::fabs(lc[k]) - real(1.01)) < 1e-6)
        throw std::runtime_error("point outside the mesh");
  }

  // Sources: type 0 = force (f = components), 1 = moment tensor
  // (m = Mxx, Myy, Mzz, Mxy, Mxz, Myz); Ricker time function as in 2D
  // (compute_source_array_from_vector.cpp, dim3/impl/
  // compute_source_array_from_tensor.cpp:15-108, ricker.cpp:55-66).
  void set_sources(int n, const double *x, const double *y, const double *z,
                   const int *type, const float *comp6, const float *f0,
                   const float *tshift, const float *factor) {
    nsrc = n;
    src_e.assign(n, 0);
    src_arr.assign(size_t(n) * 3 * NP3, 0);
    std::vector<real> t0s(n);
    real min_t0 = std::numeric_limits<real>::max();
    for (int s = 0; s < n; ++s) {
      const real hdur = 1.0 / f0[s];
      t0s[s] = -static_cast<real>(1.2) * hdur + tshift[s];
      min_t0 = std::min(min_t0, t0s[s]);
    }
    t0 = n ? double(min_t0) : 0.0;
    stf.assign(size_t(nstep) * std::max(1, n), 0);
    const real pi = static_cast<real>(M_PI);
    for (int s = 0; s < n; ++s) {
      const real ts = t0s[s] - min_t0, fs = f0[s];
      for (int it = 0; it < nstep; ++it) {
        const real t = min_t0 + it * dt;
        const real tt = t - ts;
        const real a = pi * pi * fs * fs;
        const real d2 = (1.0 - 2.0 * a * tt * tt) * std::exp(-a * tt * tt);
        stf[size_t(it) * n + s] = factor[s] * d2;
      }
      int e;
      real lc[3];
      locate(x[s], y[s], z[s], e, lc);
      src_e[s] = e;
      real h[3][NGLL], hp[3][NGLL];
      for (int d = 0; d < 3; ++d)
        lagrange_interpolants(lc[d], gll.xi, h[d], hp[d]);
      const float *c6 = comp6 + 6 * size_t(s);
      if (type[s] == 0) {
        for (int iz = 0; iz < NGLL; ++iz)
          for (int iy = 0; iy < NGLL; ++iy)
            for (int ix = 0; ix < NGLL; ++ix) {
              const real hl = h[0][ix] * h[1][iy] * h[2][iz];
              for (int c = 0; c < 3; ++c)
                src_arr[(size_t(s) * 3 + c) * NP3 + pidx(iz, iy, ix)] = hl * c6[c];
            }
      } else {
        // the Jacobian interpolated at the source (algorithms/interpolate.hpp)
        Jac3<real> Js{};
        real *js = &Js.xix;
        for (int iz = 0; iz < NGLL; ++iz)
          for (int iy = 0; iy < NGLL; ++iy)
            for (int ix = 0; ix < NGLL; ++ix) {
              const real hl = h[0][ix] * h[1][iy] * h[2][iz];
              const real *jp = &jac[size_t(e) * NP3 + pidx(iz, iy, ix)].xix;
              for (int k = 0; k < 10; ++k)
                js[k] += hl * jp[k];
            }
        const real M[3][3] = {{c6[0], c6[3], c6[4]},
                              {c6[3], c6[1], c6[5]},
                              {c6[4], c6[5], c6[2]}};
        for (int iz = 0; iz < NGLL; ++iz)
          for (int iy = 0; iy < NGLL; ++iy)
            for (int ix = 0; ix < NGLL; ++ix) {
              const real dx = (hp[0][ix] * Js.xix) * h[1][iy] * h[2][iz] +
                              h[0][ix] * (hp[1][iy] * Js.etax) * h[2][iz] +
                              h[0][ix] * h[1][iy] * (hp[2][iz] * Js.gax);
              const real dy = (hp[0][ix] * Js.xiy) * h[1][iy] * h[2][iz] +
                              h[0][ix] * (hp[1][iy] * Js.etay) * h[2][iz] +
                              h[0][ix] * h[1][iy] * (hp[2][iz] * Js.gay);
              const real dz = (hp[0][ix] * Js.xiz) * h[1][iy] * h[2][iz] +
                              h[0][ix] * (hp[1][iy] * Js.etaz) * h[2][iz] +
                              h[0][ix] * h[1][iy] * (hp[2][iz] * Js.gaz);
              for (int c = 0; c < 3; ++c)
                src_arr[(size_t(s) * 3 + c) * NP3 + pidx(iz, iy, ix)] =
                    M[c][0] * dx + M[c][1] * dy + M[c][2] * dz;
            }
      }
    }
  }

  void set_receivers(int n, const double *x, const double *y, const double *z) {
    nrec = n;
    rec_e.assign(n, 0);
    rec_h.assign(size_t(n) * NP3, 0);
    for (int r = 0; r < n; ++r) {
      int e;
      real lc[3];
      locate(x[r], y[r], z[r], e, lc);
      rec_e[r] = e;
      real h[3][NGLL];
      for (int d = 0; d < 3; ++d)
        lagrange_interpolants(lc[d], gll.xi, h[d], (real *)nullptr);
      for (int iz = 0; iz < NGLL; ++iz)
        for (int iy = 0; iy < NGLL; ++iy)
          for (int ix = 0; ix < NGLL; ++ix)
            rec_h[size_t(r) * NP3 + pidx(iz, iy, ix)] = h[0][ix] * h[1][iy] * h[2][iz];
    }
  }

  void add_sources(int it, std::vector<real> &A) const {
    for (int s = 0; s < nsrc; ++s)
      for (int p = 0; p < NP3; ++p) {
        const int g = ibool[size_t(src_e[s]) * NP3 + p];
        for (int c = 0; c < 3; ++c)
          A[3 * size_t(g) + c] +=
              src_arr[(size_t(s) * 3 + c) * NP3 + p] * stf[size_t(it) * nsrc + s];
      }
  }

  // ------------------------------------------------------------ time loop
  void step(int it) {
    const size_t n = u.size();
    for (size_t k = 0; k < n; ++k) {
      u[k] += dt * v[k] + dt2o2 * a[k];
      v[k] += dto2 * a[k];
      a[k] = 0;
    }
    add_sources(it, a);
    stiffness(u, v, a);
    for (size_t k = 0; k < n; ++k) {
      a[k] = a[k] * minv[k];
      v[k] += dto2 * a[k];
    }
  }

  // seismograms [nstep][ntypes][nrec][3] of types 0 d, 1 v, 2 a
  void run(int nsteps, int ntypes, const int *types, real *out) {
    u.assign(size_t(nglob) * 3, 0);
    v.assign(u.size(), 0);
    a.assign(u.size(), 0);
    for (int it = 0; it < nsteps; ++it) {
      step(it);
      for (int k = 0; k < ntypes; ++k) {
        const std::vector<real> &F = types[k] == 0 ? u : (types[k] == 1 ? v : a);
        for (int r = 0; r < nrec; ++r) {
          real s[3] = {0, 0, 0};
          for (int p = 0; p < NP3; ++p) {
            const int g = ibool[size_t(rec_e[r]) * NP3 + p];
            const real hw = rec_h[size_t(r) * NP3 + p];
            for (int c = 0; c < 3; ++c)
              s[c] += hw * F[3 * size_t(g) + c];
          }
          for (int c = 0; c < 3; ++c)
            out[((size_t(it) * ntypes + k) * nrec + r) * 3 + c] = s[c];
        }
      }
    }
  }

  // ------------------------------------------- discrete adjoint (checker) --
  // Exact adjoint of step() above (not a reference routine: the reference's
  // 3D path has no adjoint).  Forward step k maps (u, v, a)_k to
  // (u, v, a)_{k+1}:
  //   u~ = u + dt v + dt^2/2 a,  v~ = v + dt/2 a,
  //   a' = W (f_k - K u~ - B v~)  with W = 1/M' (M' = M + the Stacey term),
  //   u' = u~,  v' = v~ + dt/2 a'.
  // Stored per step: u_{k+1} (= u~_k), v~_k (before the corrector), a_{k+1}.
  std::vector<std::vector<real>> st_u, st_v, st_a;

  void forward_store(int nsteps, real *seis /* [nsteps][nrec][3] */) {
    u.assign(size_t(nglob) * 3, 0);
    v.assign(u.size(), 0);
    a.assign(u.size(), 0);
    st_u.assign(nsteps, {});
    st_v.assign(nsteps, {});
    st_a.assign(nsteps, {});
    const size_t n = u.size();
    for (int it = 0; it < nsteps; ++it) {
      for (size_t k = 0; k < n; ++k) {
        u[k] += dt * v[k] + dt2o2 * a[k];
        v[k] += dto2 * a[k];
        a[k] = 0;
      }
      add_sources(it, a);
      stiffness(u, v, a);
      for (size_t k = 0; k < n; ++k)
        a[k] = a[k] * minv[k];
      st_u[it] = u;
      st_v[it] = v;
      st_a[it] = a;
      for (size_t k = 0; k < n; ++k)
        v[k] += dto2 * a[k];
      for (int r = 0; r < nrec; ++r) {
        real sr[3] = {0, 0, 0};
        for (int p = 0; p < NP3; ++p) {
          const int g = ibool[size_t(rec_e[r]) * NP3 + p];
          const real hw = rec_h[size_t(r) * NP3 + p];
          for (int c = 0; c < 3; ++c)
            sr[c] += hw * u[3 * size_t(g) + c];
        }
        for (int c = 0; c < 3; ++c)
          seis[(size_t(it) * nrec + r) * 3 + c] = sr[c];
      }
    }
  }

  // Displacement gradient du[p][c][d] = d u_c / d x_d of element e.
  void element_gradient(int e, const std::vector<real> &U, real du[NP3][3][3]) const {
    real f[NP3][3];
    for (int p = 0; p < NP3; ++p) {
      const int g = ibool[size_t(e) * NP3 + p];
      for (int c = 0; c < 3; ++c)
        f[p][c] = U[3 * size_t(g) + c];
    }
    for (int iz = 0; iz < NGLL; ++iz)
      for (int iy = 0; iy < NGLL; ++iy)
        for (int ix = 0; ix < NGLL; ++ix) {
          const int p = pidx(iz, iy, ix);
          const Jac3<real> &J = jac[size_t(e) * NP3 + p];
          real dxi[3] = {0, 0, 0}, deta[3] = {0, 0, 0}, dga[3] = {0, 0, 0};
          for (int l = 0; l < NGLL; ++l)
            for (int c = 0; c < 3; ++c) {
              dxi[c] += gll.hprime[ix][l] * f[pidx(iz, iy, l)][c];
              deta[c] += gll.hprime[iy][l] * f[pidx(iz, l, ix)][c];
              dga[c] += gll.hprime[iz][l] * f[pidx(l, iy, ix)][c];
            }
          for (int c = 0; c < 3; ++c) {
            du[p][c][0] = J.xix * dxi[c] + J.etax * deta[c] + J.gax * dga[c];
            du[p][c][1] = J.xiy * dxi[c] + J.etay * deta[c] + J.gay * dga[c];
            du[p][c][2] = J.xiz * dxi[c] + J.etaz * deta[c] + J.gaz * dga[c];
          }
        }
  }

  // y += B x at the absorbing points (B v is minus the Stacey traction).
  void apply_B(const std::vector<real> &x, std::vector<real> &y) const {
    for (int e = 0; e < nspec; ++e) {
      if (!elem_st[e])
        continue;
      for (int p = 0; p < NP3; ++p) {
        const size_t i = size_t(e) * NP3 + p;
        if (st[i][4] == real(0))
          continue;
        const int g = ibool[i];
        const real xv[3] = {x[3 * size_t(g)], x[3 * size_t(g) + 1], x[3 * size_t(g) + 2]};
        real t[3];
        traction(i, xv, t);
        for (int c = 0; c < 3; ++c)
          y[3 * size_t(g) + c] -= t[c];
      }
    }
  }

  // Gradient of sum_k chi_k(u_{k+1}) with respect to (kappa, mu, rho) at
  // every local point, given dchi/ds_k = gsrc[k][r][c] for the displacement
  // seismograms (after forward_store).  grad [3][nspec][125].
  void adjoint(int nsteps, const double *gsrc, double *grad) {
    const size_t n3 = size_t(nglob) * 3, nloc = size_t(nspec) * NP3;
    std::vector<real> ub(n3, 0), vb(n3, 0), ab(n3, 0), rb(n3, 0), Kr(n3), Br(n3),
        zero(n3, 0);
    std::vector<double> gk(nloc, 0), gm(nloc, 0), gr(nloc, 0);
    const real c = dto2, h = dt2o2;
    real dr[NP3][3][3], du[NP3][3][3];
    for (int k = nsteps - 1; k >= 0; --k) {
      for (int r = 0; r < nrec; ++r)
        for (int p = 0; p < NP3; ++p) {
          const int g = ibool[size_t(rec_e[r]) * NP3 + p];
          const real hw = rec_h[size_t(r) * NP3 + p];
          for (int cc = 0; cc < 3; ++cc)
            ub[3 * size_t(g) + cc] += hw * real(gsrc[(size_t(k) * nrec + r) * 3 + cc]);
        }
      for (size_t i = 0; i < n3; ++i)
        rb[i] = minv[i] * (ab[i] + c * vb[i]);
      std::fill(Kr.begin(), Kr.end(), real(0));
      stiffness(rb, zero, Kr, false);
      for (auto &x : Kr)
        x = -x;
      std::fill(Br.begin(), Br.end(), real(0));
      apply_B(rb, Br);
      const std::vector<real> &U1 = st_u[k], &V0 = st_v[k], &A1 = st_a[k];
      for (int e = 0; e < nspec; ++e) {
        element_gradient(e, rb, dr);
        element_gradient(e, U1, du);
        for (int iz = 0; iz < NGLL; ++iz)
          for (int iy = 0; iy < NGLL; ++iy)
            for (int ix = 0; ix < NGLL; ++ix) {
              const int p = pidx(iz, iy, ix);
              const size_t i = size_t(e) * NP3 + p;
              const size_t g = size_t(ibool[i]);
              const double wj = double(gll.w[iz]) * gll.w[iy] * gll.w[ix] * jac[i].jac;
              const double divr = double(dr[p][0][0]) + dr[p][1][1] + dr[p][2][2];
              const double divu = double(du[p][0][0]) + du[p][1][1] + du[p][2][2];
              double ee = 0;
              for (int a1 = 0; a1 < 3; ++a1)
                for (int b1 = 0; b1 < 3; ++b1)
                  ee += 0.25 * (double(dr[p][a1][b1]) + dr[p][b1][a1]) *
                        (double(du[p][a1][b1]) + du[p][b1][a1]);
              double ra = 0;
              for (int cc = 0; cc < 3; ++cc)
                ra += double(rb[3 * g + cc]) * A1[3 * g + cc];
              gk[i] -= wj * divr * divu;
              gm[i] -= wj * (2.0 * ee - 2.0 / 3.0 * divr * divu);
              gr[i] -= wj * ra;
              if (elem_st[e] && st[i][4] != real(0)) {
                const auto &sp = st[i];
                const double nv[3] = {sp[0], sp[1], sp[2]};
                const double nrm = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
                const double nh[3] = {nv[0] / nrm, nv[1] / nrm, nv[2] / nrm};
                const double wt = sp[4];
                const double s1 = nh[0] + nh[1] + nh[2];
                double rn = 0, vn = 0, rv = 0, tp = 0, ts = 0;
                for (int cc = 0; cc < 3; ++cc) {
                  const double r0 = rb[3 * g + cc], v0 = V0[3 * g + cc],
                               a0 = A1[3 * g + cc];
                  rn += r0 * nh[cc];
                  vn += v0 * nh[cc];
                  rv += r0 * v0;
                  tp += r0 * a0 * nh[cc] * s1;
                  ts += r0 * a0 * (1.0 - nh[cc] * s1);
                }
                const double dP = -(nrm * wt * rn * vn + 0.5 * dt * nrm * wt * tp);
                const double dS = -(nrm * wt * (rv - rn * vn) + 0.5 * dt * nrm * wt * ts);
                const double L = l2m[i], m = mu[i], rh = rho[i];
                const double P = std::sqrt(rh * L), S = std::sqrt(rh * m);
                gk[i] += dP * rh / (2.0 * P);
                gm[i] += dP * (4.0 / 3.0) * rh / (2.0 * P);
                gr[i] += dP * L / (2.0 * P);
                if (S > 0) {
                  gm[i] += dS * rh / (2.0 * S);
                  gr[i] += dS * m / (2.0 * S);
                }
              }
            }
      }
      for (size_t i = 0; i < n3; ++i) {
        const real uu = ub[i] - Kr[i];
        const real vv = vb[i] - Br[i];
        ub[i] = uu;
        vb[i] = vv + dt * uu;
        ab[i] = c * vv + h * uu;
      }
    }
    for (size_t i = 0; i < nloc; ++i) {
      grad[i] = gk[i];
      grad[nloc + i] = gm[i];
      grad[2 * nloc + i] = gr[i];
    }
  }
};

struct Handle3 {
  Oracle3D<float> *f = nullptr;
  Oracle3D<double> *d = nullptr;
};
std::string g_err3;

} // namespace

extern "C" {

const char *oracle3d_last_error() { return g_err3.c_str(); }

void *oracle3d_create(int precision, int nspec, int nnodes, const double *coord,
                      const int *knods, const float *kappa, const float *mu,
                      const float *rho, int nabs, const int *abs_e,
                      const int *abs_f, float dt, int nstep, int nsrc,
                      const double *sx, const double *sy, const double *sz,
                      const int *stype, const float *comp6, const float *f0,
                      const float *tshift, const float *factor, int nrec,
                      const double *rx, const double *ry, const double *rz) {
  auto setup = [&](auto *o) {
    o->nstep = nstep;
    o->dt = dt;
    o->dto2 = dt / 2.0;
    o->dt2o2 = dt * dt / 2.0;
    o->build(nspec, nnodes, coord, knods, kappa, mu, rho, nabs, abs_e, abs_f);
    o->mass();
    o->set_sources(nsrc, sx, sy, sz, stype, comp6, f0, tshift, factor);
    o->set_receivers(nrec, rx, ry, rz);
  };
  auto *h = new Handle3;
  try {
    if (precision == 8) {
      h->d = new Oracle3D<double>;
      setup(h->d);
    } else {
      h->f = new Oracle3D<float>;
      setup(h->f);
    }
    return h;
  } catch (const std::exception &e) {
    g_err3 = e.what();
    delete h->f;
    delete h->d;
    delete h;
    return nullptr;
  }
}

void oracle3d_destroy(void *hv) {
  auto *h = static_cast<Handle3 *>(hv);
  delete h->f;
  delete h->d;
  delete h;
}

int oracle3d_nglob(void *hv) {
  auto *h = static_cast<Handle3 *>(hv);
  return h->f ? h->f->nglob : h->d->nglob;
}

double oracle3d_t0(void *hv) {
  auto *h = static_cast<Handle3 *>(hv);
  return h->f ? h->f->t0 : h->d->t0;
}

// Seismograms [nsteps][ntypes][nrec][3] as double.
int oracle3d_run(void *hv, int nsteps, int ntypes, const int *types,
                 double *out) {
  auto *h = static_cast<Handle3 *>(hv);
  try {
    auto go = [&](auto *o) {
      const size_t n = size_t(nsteps) * ntypes * o->nrec * 3;
      std::vector<decltype(o->dt)> buf(n);
      o->run(nsteps, ntypes, types, buf.data());
      for (size_t k = 0; k < n; ++k)
        out[k] = buf[k];
    };
    if (h->f)
      go(h->f);
    else
      go(h->d);
    return 0;
  } catch (const std::exception &e) {
    g_err3 = e.what();
    return -1;
  }
}

// Arrays for closed-form checks: 0 minv [nglob][3], 1 ibool [nspec][125],
// 2 jacobian [nspec][125], 3 source array [nsrc][3][125], 4 receiver
// interpolants [nrec][125], 5 u, 6 v, 7 a (current fields), 8 inverse
// Jacobian [nloc][9].  Returns the
// length; copies when out is non-null.
long oracle3d_get(void *hv, int what, double *out) {
  auto *h = static_cast<Handle3 *>(hv);
  auto get = [&](auto *o) -> long {
    std::vector<double> v;
    switch (what) {
    case 0: v.assign(o->minv.begin(), o->minv.end()); break;
    case 1: v.assign(o->ibool.begin(), o->ibool.end()); break;
    case 2:
      for (auto &j : o->jac) v.push_back(j.jac);
      break;
    case 3: v.assign(o->src_arr.begin(), o->src_arr.end()); break;
    case 4: v.assign(o->rec_h.begin(), o->rec_h.end()); break;
    case 5: v.assign(o->u.begin(), o->u.end()); break;
    case 6: v.assign(o->v.begin(), o->v.end()); break;
    case 7: v.assign(o->a.begin(), o->a.end()); break;
    case 8: // inverse-Jacobian terms [nloc][9] (xi, eta, gamma) x (x, y, z)
      for (auto &j : o->jac)
        for (double t : {j.xix, j.xiy, j.xiz, j.etax, j.etay, j.etaz, j.gax,
                         j.gay, j.gaz})
          v.push_back(t);
      break;
    default: return -1;
    }
    if (out)
      std::copy(v.begin(), v.end(), out);
    return long(v.size());
  };
  return h->f ? get(h->f) : get(h->d);
}

// a <- -K u (+ Stacey with v) for given u, v ([nglob][3]).
int oracle3d_apply_stiffness(void *hv, const double *u, const double *v,
                             int with_stacey) {
  auto *h = static_cast<Handle3 *>(hv);
  auto app = [&](auto *o) {
    const size_t n = size_t(o->nglob) * 3;
    o->u.assign(u, u + n);
    o->v.assign(v, v + n);
    o->a.assign(n, 0);
    o->stiffness(o->u, o->v, o->a, with_stacey != 0);
    return 0;
  };
  return h->f ? app(h->f) : app(h->d);
}

// Forward run from rest storing every step's state for the adjoint;
// displacement seismograms [nsteps][nrec][3].
int oracle3d_forward_store(void *hv, int nsteps, double *seis) {
  auto *h = static_cast<Handle3 *>(hv);
  try {
    auto go = [&](auto *o) {
      const size_t n = size_t(nsteps) * o->nrec * 3;
      std::vector<decltype(o->dt)> buf(std::max<size_t>(n, 1));
      o->forward_store(nsteps, buf.data());
      for (size_t k = 0; k < n; ++k)
        seis[k] = buf[k];
    };
    if (h->f)
      go(h->f);
    else
      go(h->d);
    return 0;
  } catch (const std::exception &e) {
    g_err3 = e.what();
    return -1;
  }
}

// Discrete-adjoint gradient for dchi/ds = gsrc[nsteps][nrec][3] (after
// oracle3d_forward_store with the same nsteps); grad [3][nspec][125] =
// dchi / d(kappa, mu, rho) per local point.
int oracle3d_adjoint(void *hv, int nsteps, const double *gsrc, double *grad) {
  auto *h = static_cast<Handle3 *>(hv);
  try {
    auto go = [&](auto *o) {
      if (int(o->st_u.size()) != nsteps)
        throw std::runtime_error("run oracle3d_forward_store with the same nsteps first");
      o->adjoint(nsteps, gsrc, grad);
    };
    if (h->f)
      go(h->f);
    else
      go(h->d);
    return 0;
  } catch (const std::exception &e) {
    g_err3 = e.what();
    return -1;
  }
}

} // extern "C"
