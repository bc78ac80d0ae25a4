// ORACLE — test infrastructure only.
//
// CPU restatement of SPECFEM++'s 2D elastic isotropic P-SV time-stepping path
// and of its acoustic isotropic (potential) counterpart, written to check the
// HIP implementation in specfempp_amd/.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code,
// and only as the checker / baseline timer.  The product never links it.
//
// Every routine follows the reference's arithmetic in the reference's order
// (FP32 when real=float, with the same float/double promotions), citing the
// reference file:line it restates.  The FP64 instantiation (real=double) is the
// same algorithm carried in double precision.
//
// Parity pins (see tests/test_oracle_*.py):
//   * GLL nodes/weights/h'       closed forms + polynomial exactness
//   * Jacobian, mass matrix      affine closed forms, sum(M) = rho * area
//   * stress / Stacey formulas   closed forms (normal incidence impedance)
//   * whole forward path         SPECFEM2D golden traces shipped with the
//                                reference (HomogeneousElasticDomain and
//                                HomogeneousElasticDomainStaceyDirichletBC),
//                                reference metric <= 1e-3
//   * acoustic forward path      SPECFEM2D golden traces shipped with the
//                                reference (HomogeneousAcousticDomain: Dirac
//                                source, free surface on all sides, d/v/a/p;
//                                ...DirichletBC and ...CompositeStaceyDirichletBC:
//                                four Ricker sources), reference metric; plus
//                                closed forms (mass = area / kappa, Dirichlet
//                                points stay at rest without sources there)
//   * moment tensors, Heaviside  SPECFEM2D golden traces of the reference's
//                                HomogeneousElasticDomainMomentTensor (d, v, a)
//                                and ...MomentTensorHeaviside (d, v; its
//                                acceleration is not pinned by data at 1e-3),
//                                planted sign / transposition of the tensor
//                                fail them; the Heaviside table against
//                                scipy's erf
//   * coupled elastic-acoustic   SPECFEM2D golden traces of the reference's
//                                two coupled fixtures (water above the solid,
//                                restated from its Par_file; water below, from
//                                the mesher's database), reference metric
//   * Frechet kernels            "parity unpinned" by reference data (no
//                                reference fixture holds kernel values); pinned
//                                by finite differences of the misfit instead.
#include <algorithm>
#include <chrono>
#include <omp.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

#include "gll_common.hpp"

// --------------------------------------------------- shape functions ----
// core/specfem/shape_function/dim2/shape_functions2d.cpp (values) and
// shape_function_derivatives2d.cpp:7-75 (derivatives), 4- and 9-node quads.
template <typename real>
void shape_functions(real xi, real gamma, int ngnod, real *N) {
  if (ngnod == 4) {
    N[0] = 0.25 * (xi - 1) * (gamma - 1);
    N[1] = -0.25 * (xi + 1) * (gamma - 1);
    N[2] = 0.25 * (xi + 1) * (gamma + 1);
    N[3] = -0.25 * (xi - 1) * (gamma + 1);
  } else {
    N[0] = 0.25 * (xi - 1) * xi * (gamma - 1) * gamma;
    N[1] = 0.25 * (xi + 1) * xi * (gamma - 1) * gamma;
    N[2] = 0.25 * (xi + 1) * xi * (gamma + 1) * gamma;
    N[3] = 0.25 * (xi - 1) * xi * (gamma + 1) * gamma;
    N[4] = 0.5 * (1 - xi * xi) * (gamma - 1) * gamma;
    N[5] = 0.5 * (xi + 1) * xi * (1 - gamma * gamma);
    N[6] = 0.5 * (1 - xi * xi) * (gamma + 1) * gamma;
    N[7] = 0.5 * (xi - 1) * xi * (1 - gamma * gamma);
    N[8] = (1 - xi * xi) * (1 - gamma * gamma);
  }
}

template <typename real>
void shape_derivatives(real xi, real gamma, int ngnod, real *dxi, real *dga) {
  if (ngnod == 4) {
    const real sp = xi + 1, sm = xi - 1, tp = gamma + 1, tm = gamma - 1;
    dxi[0] = 0.25 * tm;
    dxi[1] = -0.25 * tm;
    dxi[2] = 0.25 * tp;
    dxi[3] = -0.25 * tp;
    dga[0] = 0.25 * sm;
    dga[1] = -0.25 * sp;
    dga[2] = 0.25 * sp;
    dga[3] = -0.25 * sm;
  } else {
    const real sp = xi + 1.0, sm = xi - 1.0, tp = gamma + 1.0,
               tm = gamma - 1.0, s2 = xi * 2.0, t2 = gamma * 2.0,
               ss = xi * xi, tt = gamma * gamma, st = xi * gamma;
    dxi[0] = 0.25 * tm * gamma * (s2 - 1.0);
    dxi[1] = 0.25 * tm * gamma * (s2 + 1.0);
    dxi[2] = 0.25 * tp * gamma * (s2 + 1.0);
    dxi[3] = 0.25 * tp * gamma * (s2 - 1.0);
    dga[0] = 0.25 * sm * xi * (t2 - 1.0);
    dga[1] = 0.25 * sp * xi * (t2 - 1.0);
    dga[2] = 0.25 * sp * xi * (t2 + 1.0);
    dga[3] = 0.25 * sm * xi * (t2 + 1.0);
    dxi[4] = -1.0 * st * tm;
    dxi[5] = 0.5 * (1.0 - tt) * (s2 + 1.0);
    dxi[6] = -1.0 * st * tp;
    dxi[7] = 0.5 * (1.0 - tt) * (s2 - 1.0);
    dga[4] = 0.5 * (1.0 - ss) * (t2 - 1.0);
    dga[5] = -1.0 * st * sp;
    dga[6] = 0.5 * (1.0 - ss) * (t2 + 1.0);
    dga[7] = -1.0 * st * sm;
    dxi[8] = -1.0 * s2 * (1.0 - tt);
    dga[8] = -1.0 * t2 * (1.0 - ss);
  }
}

template <typename real> struct Jac {
  real xix, gammax, xiz, gammaz, jacobian;
};

// core/specfem/jacobian/dim2/jacobian.cpp:25-51
template <typename real>
Jac<real> compute_jacobian(const real *cx, const real *cz, int ngnod,
                           const real *dxi, const real *dga) {
  real xxi = 0.0, zxi = 0.0, xgamma = 0.0, zgamma = 0.0;
  for (int in = 0; in < ngnod; in++) {
    xxi += dxi[in] * cx[in];
    zxi += dxi[in] * cz[in];
    xgamma += dga[in] * cx[in];
    zgamma += dga[in] * cz[in];
  }
  const real jacobian = xxi * zgamma - xgamma * zxi;
  Jac<real> J;
  J.xix = zgamma / jacobian;
  J.gammax = -zxi / jacobian;
  J.xiz = -xgamma / jacobian;
  J.gammaz = xxi / jacobian;
  J.jacobian = jacobian;
  return J;
}

// core/specfem/jacobian/dim2/jacobian.cpp:4-22
template <typename real>
void compute_location(const real *cx, const real *cz, int ngnod, real xi,
                      real gamma, real &x, real &z) {
  real N[9];
  shape_functions<real>(xi, gamma, ngnod, N);
  x = 0.0;
  z = 0.0;
  for (int in = 0; in < ngnod; in++) {
    x += N[in] * cx[in];
    z += N[in] * cz[in];
  }
}

// Edge types (include/enumerations/dim2/mesh_entities.hpp:29-38).
enum EdgeType { BOTTOM = 1, RIGHT = 2, TOP = 3, LEFT = 4 };

// core/specfem/assembly/boundaries/dim2/impl/utilities.hpp:7-20
bool on_edge(int type, int iz, int ix) {
  return (type == TOP && iz == NGLL - 1) || (type == BOTTOM && iz == 0) ||
         (type == LEFT && ix == 0) || (type == RIGHT && ix == NGLL - 1);
}

} // namespace

// =========================================================== Oracle2D ====
template <typename real> struct Oracle2D {
  // mesh
  int nspec = 0, ngnod = 0, nglob = 0;
  GLL<real> gll;
  std::vector<real> cx, cz;        // control nodes [nspec][ngnod]
  std::vector<int> ibool;          // [nspec][NGLL][NGLL]  (iz, ix)
  std::vector<real> gx, gz;        // GLL point coordinates [nspec][NGLL][NGLL]
  std::vector<Jac<real>> jac;      // [nspec][NGLL][NGLL]
  std::vector<real> kappa, mu, rho; // [nspec][NGLL][NGLL]
  // Stacey: per element flag + per point (tag, normal, weight)
  std::vector<char> elem_stacey;
  std::vector<char> pt_stacey;
  std::vector<real> pt_n0, pt_n1, pt_w;
  std::vector<int> order; // element processing order: none group, then stacey
  // acoustic medium (every element has mu == 0): potential in component 0 of
  // the [nglob][2] fields; per-point acoustic free-surface (Dirichlet) tag
  bool acoustic = false;
  std::vector<char> pt_free;
  // fields [nglob][2]
  std::vector<real> u, v, a, minv;
  // Frechet kernels (rho, mu, kappa) per local point
  std::vector<real> k_rho, k_mu, k_kappa;
  // sources
  int nsrc = 0, nstep = 0;
  std::vector<int> src_ispec;
  std::vector<real> src_array; // [nsrc][2][NGLL][NGLL]
  std::vector<real> stf;       // [nstep][nsrc][2]
  real t0 = 0;
  // receivers
  int nrec = 0;
  std::vector<int> rec_ispec;
  std::vector<real> rec_h; // [nrec][NGLL][NGLL]
  std::vector<real> rec_xi, rec_gamma;
  // time scheme
  real deltat = 0, deltatover2 = 0, deltasquareover2 = 0;

  int idx(int e, int iz, int ix) const { return (e * NGLL + iz) * NGLL + ix; }

  // ------------------------------------------------------------ mesh ----
  void build_mesh(int nspec_, int ngnod_, int nnodes, const double *coord,
                  const int *knods, const float *kappa_, const float *mu_,
                  const float *rho_, int nabs, const int *abs_ispec,
                  const int *abs_type, int nfree = 0,
                  const int *free_ispec = nullptr,
                  const int *free_type = nullptr) {
    nspec = nspec_;
    ngnod = ngnod_;
    if (ngnod != 4 && ngnod != 9)
      throw std::runtime_error("ngnod must be 4 or 9");
    cx.resize(size_t(nspec) * ngnod);
    cz.resize(size_t(nspec) * ngnod);
    for (int e = 0; e < nspec; ++e)
      for (int n = 0; n < ngnod; ++n) {
        const int id = knods[e * ngnod + n];
        if (id < 0 || id >= nnodes)
          throw std::runtime_error("knods out of range");
        cx[e * ngnod + n] = static_cast<real>(static_cast<float>(coord[2 * id]));
        cz[e * ngnod + n] =
            static_cast<real>(static_cast<float>(coord[2 * id + 1]));
      }

    // Global numbering from the corner topology: corners are shared through
    // their control-node ids, edge-interior points through the (sorted)
    // corner pair and their position along the edge, element-interior
    // points are unique.  Numbers are assigned on first encounter.
    const size_t nloc = size_t(nspec) * NGLL * NGLL;
    ibool.assign(nloc, -1);
    std::map<int, int> corner_id;
    std::map<std::pair<int, int>, std::vector<int>> edge_ids;
    int next = 0;
    auto corner = [&](int node) {
      auto it = corner_id.find(node);
      if (it != corner_id.end())
        return it->second;
      corner_id[node] = next;
      return next++;
    };
    auto edge_pt = [&](int ca, int cb, int k) { // k in 1..NGLL-2 from ca
      std::pair<int, int> key(std::min(ca, cb), std::max(ca, cb));
      const int pos = (ca < cb) ? k : (NGLL - 1 - k);
      auto &vec = edge_ids[key];
      if (vec.empty())
        vec.assign(NGLL, -1);
      if (vec[pos] < 0)
        vec[pos] = next++;
      return vec[pos];
    };
    for (int e = 0; e < nspec; ++e) {
      const int c0 = knods[e * ngnod + 0], c1 = knods[e * ngnod + 1],
                c2 = knods[e * ngnod + 2], c3 = knods[e * ngnod + 3];
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const bool bz = iz == 0, tz = iz == NGLL - 1, lx = ix == 0,
                     rx = ix == NGLL - 1;
          int g;
          if (bz && lx)
            g = corner(c0);
          else if (bz && rx)
            g = corner(c1);
          else if (tz && rx)
            g = corner(c2);
          else if (tz && lx)
            g = corner(c3);
          else if (bz)
            g = edge_pt(c0, c1, ix);
          else if (rx)
            g = edge_pt(c1, c2, iz);
          else if (tz)
            g = edge_pt(c3, c2, ix);
          else if (lx)
            g = edge_pt(c0, c3, iz);
          else
            g = next++;
          ibool[idx(e, iz, ix)] = g;
        }
    }
    nglob = next;

    // Geometry (core/specfem/assembly/jacobian_matrix/dim2/jacobian_matrix.cpp:28-93)
    jac.resize(nloc);
    gx.resize(nloc);
    gz.resize(nloc);
    for (int e = 0; e < nspec; ++e)
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          real dxi[9], dga[9];
          shape_derivatives<real>(gll.xi[ix], gll.xi[iz], ngnod, dxi, dga);
          jac[idx(e, iz, ix)] = compute_jacobian<real>(
              &cx[e * ngnod], &cz[e * ngnod], ngnod, dxi, dga);
          compute_location<real>(&cx[e * ngnod], &cz[e * ngnod], ngnod,
                                 gll.xi[ix], gll.xi[iz], gx[idx(e, iz, ix)],
                                 gz[idx(e, iz, ix)]);
        }

    kappa.resize(nloc);
    mu.resize(nloc);
    rho.resize(nloc);
    for (size_t i = 0; i < nloc; ++i) {
      kappa[i] = kappa_[i];
      mu[i] = mu_[i];
      rho[i] = rho_[i];
    }
    // acoustic elements: mu == 0 at every point (a material with cs = 0,
    // src/io/mesh/impl/fortran/dim2/read_material_properties.cpp:178-204)
    {
      int nac = 0;
      for (int e = 0; e < nspec; ++e) {
        bool fluid = true;
        for (int p = 0; p < NGLL * NGLL; ++p)
          fluid = fluid && mu[size_t(e) * NGLL * NGLL + p] == 0;
        nac += fluid;
      }
      if (nac != 0 && nac != nspec)
        throw std::runtime_error("mixed acoustic/elastic meshes are not restated");
      acoustic = nac == nspec;
    }
    // acoustic free surface: per-point tags from the (element, entity) list
    // (core/specfem/assembly/boundaries/dim2/impl/acoustic_free_surface.tpp:
    // 60-94, utilities.hpp:6-20)
    pt_free.assign(nloc, 0);
    for (int k = 0; k < nfree; ++k) {
      const int e = free_ispec[k], t = free_type[k];
      if (e < 0 || e >= nspec || t < 1 || t > 8)
        throw std::runtime_error("free-surface entity out of range");
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const bool bz = iz == 0, tz = iz == NGLL - 1, lx = ix == 0,
                     rx = ix == NGLL - 1;
          if ((t <= 4 && on_edge(t, iz, ix)) || (t == 5 && bz && lx) ||
              (t == 6 && bz && rx) || (t == 7 && tz && rx) ||
              (t == 8 && tz && lx))
            pt_free[idx(e, iz, ix)] = 1;
        }
    }

    // Stacey data: one (normal, weight) per GLL point; the last listed edge
    // covering a point wins
    // (core/specfem/assembly/boundaries/dim2/impl/stacey.tpp:131-167 and
    //  .../utilities.hpp:22-61).
    elem_stacey.assign(nspec, 0);
    pt_stacey.assign(nloc, 0);
    pt_n0.assign(nloc, 0);
    pt_n1.assign(nloc, 0);
    pt_w.assign(nloc, 0);
    // stacey.tpp groups the edge list by element (std::map over ispec) and
    // walks each element's edges in list order.
    std::map<int, std::vector<int>> by_elem;
    for (int i = 0; i < nabs; ++i) {
      if (abs_ispec[i] < 0 || abs_ispec[i] >= nspec)
        throw std::runtime_error("absorbing edge element out of range");
      by_elem[abs_ispec[i]].push_back(abs_type[i]);
    }
    for (auto &kv : by_elem) {
      const int e = kv.first;
      elem_stacey[e] = 1;
      for (int type : kv.second)
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            if (!on_edge(type, iz, ix))
              continue;
            const Jac<real> &J = jac[idx(e, iz, ix)];
            real n0, n1, wgt;
            // core/specfem/point/jacobian_matrix.hpp:424-449
            if (type == LEFT) {
              n0 = static_cast<real>(-1.0) * J.xix * J.jacobian;
              n1 = static_cast<real>(-1.0) * J.xiz * J.jacobian;
              wgt = gll.w[iz];
            } else if (type == RIGHT) {
              n0 = J.xix * J.jacobian;
              n1 = J.xiz * J.jacobian;
              wgt = gll.w[iz];
            } else if (type == TOP) {
              n0 = J.gammax * J.jacobian;
              n1 = J.gammaz * J.jacobian;
              wgt = gll.w[ix];
            } else if (type == BOTTOM) {
              n0 = static_cast<real>(-1.0) * J.gammax * J.jacobian;
              n1 = static_cast<real>(-1.0) * J.gammaz * J.jacobian;
              wgt = gll.w[ix];
            } else {
              throw std::runtime_error("unknown edge type");
            }
            pt_stacey[idx(e, iz, ix)] = 1;
            pt_n0[idx(e, iz, ix)] = n0;
            pt_n1[idx(e, iz, ix)] = n1;
            pt_w[idx(e, iz, ix)] = wgt;
          }
    }
    // Element groups in the order domain_kernels visits them: boundary tag
    // none first, then stacey (include/kokkos_kernels/domain_kernels.hpp:112-126).
    order.clear();
    for (int e = 0; e < nspec; ++e)
      if (!elem_stacey[e])
        order.push_back(e);
    for (int e = 0; e < nspec; ++e)
      if (elem_stacey[e])
        order.push_back(e);

    u.assign(size_t(nglob) * 2, 0);
    v.assign(size_t(nglob) * 2, 0);
    a.assign(size_t(nglob) * 2, 0);
    minv.assign(size_t(nglob) * 2, 0);
    k_rho.assign(nloc, 0);
    k_mu.assign(nloc, 0);
    k_kappa.assign(nloc, 0);
  }

  // include/medium/dim2/elastic/isotropic/properties.hpp:71-91
  real lambdaplus2mu(int i) const {
    return kappa[i] + (static_cast<real>(4.0 / 3.0)) * mu[i];
  }
  real lambda(int i) const {
    return lambdaplus2mu(i) - (static_cast<real>(2.0)) * mu[i];
  }
  real rho_vp(int i) const {
    return std::sqrt(lambdaplus2mu(i) / rho[i]) * rho[i];
  }
  real rho_vs(int i) const { return std::sqrt(rho[i] * mu[i]); }

  // include/medium/dim2/acoustic/isotropic/properties.hpp:55-70,
  // material.hpp:88-91
  real rho_inverse(int i) const { return static_cast<real>(1.0) / rho[i]; }
  real rho_vpinverse(int i) const {
    const real kappa_inverse = static_cast<real>(1.0) / kappa[i];
    return std::sqrt(rho_inverse(i) * kappa_inverse);
  }
  // include/boundary_conditions/stacey/stacey.tpp:293-321 (acoustic)
  void stacey_traction_acoustic(int i, real vel, real &traction) const {
    const real nrm = std::sqrt(pt_n0[i] * pt_n0[i] + pt_n1[i] * pt_n1[i]);
    const real factor = pt_w[i] * nrm;
    traction += static_cast<real>(-1.0) * factor * rho_vpinverse(i) * vel;
  }

  // include/boundary_conditions/stacey/stacey.tpp:46-80
  void stacey_traction(int i, const real *vel, real *traction) const {
    const real vn = vel[0] * pt_n0[i] + vel[1] * pt_n1[i];
    const real dn[2] = {pt_n0[i], pt_n1[i]};
    const real jacobian1d = std::sqrt(dn[0] * dn[0] + dn[1] * dn[1]);
    const real rvp = rho_vp(i), rvs = rho_vs(i);
    real factor[2];
    for (int c = 0; c < 2; ++c)
      factor[c] = ((vn * dn[c] / (jacobian1d * jacobian1d)) * (rvp - rvs)) +
                  vel[c] * rvs;
    traction[0] += static_cast<real>(-1.0) * factor[0] * jacobian1d * pt_w[i];
    traction[1] += static_cast<real>(-1.0) * factor[1] * jacobian1d * pt_w[i];
  }

  // ----------------------------------------------------- mass matrix ----
  // include/kokkos_kernels/impl/compute_mass_matrix.tpp:95-121,
  // include/boundary_conditions/stacey/stacey.tpp:817-853,
  // include/kokkos_kernels/impl/invert_mass_matrix.tpp:41-52
  void initialize(real dt) {
    std::fill(minv.begin(), minv.end(), real(0));
    for (int e : order)
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const int i = idx(e, iz, ix);
          const real prod = gll.w[iz] * gll.w[ix];
          if (acoustic) {
            // mass_inverse component 1/kappa (acoustic/isotropic/mass_matrix.tpp)
            real m = static_cast<real>(1.0) / kappa[i];
            m *= prod * jac[i].jacobian;
            if (elem_stacey[e] && pt_stacey[i])
              stacey_traction_acoustic(
                  i, static_cast<real>(-1.0) * dt * static_cast<real>(0.5), m);
            minv[2 * ibool[i]] += m;
            minv[2 * ibool[i] + 1] += static_cast<real>(1.0);
            continue;
          }
          real m[2] = {rho[i], rho[i]};
          const real f = prod * jac[i].jacobian;
          m[0] *= f;
          m[1] *= f;
          if (elem_stacey[e] && pt_stacey[i]) {
            const real vf = static_cast<real>(-1.0) * dt *
                            static_cast<real>(0.5);
            const real vel[2] = {vf, vf};
            stacey_traction(i, vel, m);
          }
          const int g = ibool[i];
          minv[2 * g] += m[0];
          minv[2 * g + 1] += m[1];
        }
    for (auto &x : minv)
      x = static_cast<real>(1.0) / x;
  }

  // ------------------------------------------------------- time scheme ----
  // core/specfem/timescheme/newmark.tpp:67-128 (predictor)
  void predictor(std::vector<real> &U, std::vector<real> &V,
                 std::vector<real> &A, real dt, real dto2, real dt2o2) {
    for (size_t k = 0; k < U.size(); ++k) {
      U[k] += dt * V[k] + dt2o2 * A[k];
      V[k] += dto2 * A[k];
      A[k] = 0;
    }
  }
  // core/specfem/timescheme/newmark.tpp:13-62 (corrector)
  void corrector(std::vector<real> &V, const std::vector<real> &A, real dto2) {
    for (size_t k = 0; k < V.size(); ++k)
      V[k] += dto2 * A[k];
  }
  // include/kokkos_kernels/impl/divide_mass_matrix.tpp:13-65
  void divide_mass(std::vector<real> &A) {
    for (size_t k = 0; k < A.size(); ++k)
      A[k] = A[k] * minv[k];
  }

  // -------------------------------------------------------- stiffness ----
  // Element gradient of one field (include/algorithms/gradient.hpp:40-82).
  void element_gradient(int e, const std::vector<real> &U, real du[NGLL][NGLL][2][2],
                        bool add_second = false,
                        const std::vector<real> *U2 = nullptr) const {
    real f[NGLL][NGLL][2];
    for (int iz = 0; iz < NGLL; ++iz)
      for (int ix = 0; ix < NGLL; ++ix) {
        const int g = ibool[idx(e, iz, ix)];
        f[iz][ix][0] = U[2 * g];
        f[iz][ix][1] = U[2 * g + 1];
      }
    real f2[NGLL][NGLL][2];
    if (U2)
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const int g = ibool[idx(e, iz, ix)];
          f2[iz][ix][0] = (*U2)[2 * g];
          f2[iz][ix][1] = (*U2)[2 * g + 1];
        }
    for (int iz = 0; iz < NGLL; ++iz)
      for (int ix = 0; ix < NGLL; ++ix) {
        real dxi[2] = {0, 0}, dga[2] = {0, 0};
        for (int l = 0; l < NGLL; ++l)
          for (int c = 0; c < 2; ++c) {
            dxi[c] += gll.hprime[ix][l] * f[iz][l][c];
            dga[c] += gll.hprime[iz][l] * f[l][ix][c];
          }
        if (add_second) {
          // the two-field overload keeps accumulating into the same buffers
          // (include/algorithms/gradient.hpp:302-316)
          for (int l = 0; l < NGLL; ++l)
            for (int c = 0; c < 2; ++c) {
              dxi[c] += gll.hprime[ix][l] * f2[iz][l][c];
              dga[c] += gll.hprime[iz][l] * f2[l][ix][c];
            }
        }
        const Jac<real> &J = jac[idx(e, iz, ix)];
        for (int c = 0; c < 2; ++c) {
          du[iz][ix][c][0] = J.xix * dxi[c] + J.gammax * dga[c];
          du[iz][ix][c][1] = J.xiz * dxi[c] + J.gammaz * dga[c];
        }
      }
  }

  // include/kokkos_kernels/impl/compute_stiffness_interaction.tpp:140-247
  void stiffness(const std::vector<real> &U, const std::vector<real> &V,
                 std::vector<real> &A, bool apply_stacey = true) {
    for (int e : order)
      element_force(e, U, V, A, apply_stacey);
  }

  // CPU-baseline timing only (never a checker): the same element work over
  // colour groups (no two elements of a group share a GLL point), each
  // group split across OpenMP threads.  The summation order differs from
  // stiffness(), so results agree to rounding, not bit for bit.
  std::vector<std::vector<int>> colour_groups;
  void build_colour_groups() {
    if (!colour_groups.empty())
      return;
    std::vector<uint32_t> used(size_t(nglob), 0u);
    for (int e : order) {
      uint32_t m = 0;
      for (int p = 0; p < NGLL * NGLL; ++p)
        m |= used[ibool[idx(e, p / NGLL, p % NGLL)]];
      int c = 0;
      while (c < 32 && ((m >> c) & 1u))
        ++c;
      if (c >= 32)
        throw std::runtime_error("colouring needs more than 32 colours");
      if (int(colour_groups.size()) <= c)
        colour_groups.resize(c + 1);
      colour_groups[c].push_back(e);
      for (int p = 0; p < NGLL * NGLL; ++p)
        used[ibool[idx(e, p / NGLL, p % NGLL)]] |= 1u << c;
    }
  }
  void stiffness_parallel(const std::vector<real> &U, const std::vector<real> &V,
                          std::vector<real> &A) {
    build_colour_groups();
    for (const auto &grp : colour_groups) {
      const int n = int(grp.size());
#pragma omp parallel for schedule(static)
      for (int k = 0; k < n; ++k)
        element_force(grp[k], U, V, A, true);
    }
  }
  void forward_step_parallel(int istep) {
    const long n2 = long(u.size());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n2; ++i) {
      u[i] += deltat * v[i] + deltasquareover2 * a[i];
      v[i] += deltatover2 * a[i];
      a[i] = 0;
    }
    add_sources(istep, a);
    stiffness_parallel(u, v, a);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n2; ++i) {
      a[i] = a[i] * minv[i];
      v[i] += deltatover2 * a[i];
    }
  }

  // Acoustic element: T = rho^-1 grad chi (acoustic/isotropic/stress.hpp:
  // 15-56), F = J T . dxi/dx, weighted divergence, Stacey, then the free
  // surface zeroes the tagged points (dirichlet.hpp:20-45; composite points
  // take the Stacey term first, composite_stacey_dirichlet.tpp:11-45).
  void element_force_acoustic(int e, const std::vector<real> &U,
                              const std::vector<real> &V, std::vector<real> &A,
                              bool apply_stacey) {
    real du[NGLL][NGLL][2][2];
    element_gradient(e, U, du);
    real F[NGLL][NGLL][2];
    for (int iz = 0; iz < NGLL; ++iz)
      for (int ix = 0; ix < NGLL; ++ix) {
        const int i = idx(e, iz, ix);
        const real ri = rho_inverse(i);
        const real T[2] = {ri * du[iz][ix][0][0], ri * du[iz][ix][0][1]};
        const Jac<real> &J = jac[i];
        F[iz][ix][0] = J.jacobian * (T[0] * J.xix + T[1] * J.xiz);
        F[iz][ix][1] = J.jacobian * (T[0] * J.gammax + T[1] * J.gammaz);
      }
    for (int iz = 0; iz < NGLL; ++iz)
      for (int ix = 0; ix < NGLL; ++ix) {
        real t1 = 0, t2 = 0;
        for (int l = 0; l < NGLL; ++l) {
          t1 += F[iz][l][0] * gll.hprime[l][ix] * gll.w[l];
          t2 += F[l][ix][1] * gll.hprime[l][iz] * gll.w[l];
        }
        real acc = gll.w[iz] * t1 + gll.w[ix] * t2;
        acc *= static_cast<real>(-1.0);
        const int i = idx(e, iz, ix);
        const int g = ibool[i];
        if (apply_stacey && elem_stacey[e] && pt_stacey[i])
          stacey_traction_acoustic(i, V[2 * g], acc);
        if (pt_free[i])
          acc = 0;
        A[2 * g] += acc;
      }
  }

  void element_force(int e, const std::vector<real> &U, const std::vector<real> &V,
                     std::vector<real> &A, bool apply_stacey) {
    if (acoustic) {
      element_force_acoustic(e, U, V, A, apply_stacey);
      return;
    }
    {
      real du[NGLL][NGLL][2][2];
      element_gradient(e, U, du);
      real F[NGLL][NGLL][2][2];
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const int i = idx(e, iz, ix);
          // include/medium/dim2/elastic/isotropic/stress.hpp:38-78
          const real l2m = lambdaplus2mu(i), lam = lambda(i);
          const real sxx = l2m * du[iz][ix][0][0] + lam * du[iz][ix][1][1];
          const real szz = l2m * du[iz][ix][1][1] + lam * du[iz][ix][0][0];
          const real sxz = mu[i] * (du[iz][ix][0][1] + du[iz][ix][1][0]);
          const real T[2][2] = {{sxx, sxz}, {sxz, szz}};
          // core/specfem/point/stress.hpp:166-183
          const Jac<real> &J = jac[i];
          for (int c = 0; c < 2; ++c) {
            F[iz][ix][c][0] =
                J.jacobian * (T[c][0] * J.xix + T[c][1] * J.xiz);
            F[iz][ix][c][1] =
                J.jacobian * (T[c][0] * J.gammax + T[c][1] * J.gammaz);
          }
        }
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          // include/algorithms/divergence.hpp:40-74
          real t1[2] = {0, 0}, t2[2] = {0, 0};
          for (int l = 0; l < NGLL; ++l) {
            for (int c = 0; c < 2; ++c)
              t1[c] += F[iz][l][c][0] * gll.hprime[l][ix] * gll.w[l];
            for (int c = 0; c < 2; ++c)
              t2[c] += F[l][ix][c][1] * gll.hprime[l][iz] * gll.w[l];
          }
          real acc[2];
          for (int c = 0; c < 2; ++c)
            acc[c] = gll.w[iz] * t1[c] + gll.w[ix] * t2[c];
          for (int c = 0; c < 2; ++c)
            acc[c] *= static_cast<real>(-1.0);
          const int i = idx(e, iz, ix);
          const int g = ibool[i];
          if (apply_stacey && elem_stacey[e] && pt_stacey[i]) {
            const real vel[2] = {V[2 * g], V[2 * g + 1]};
            stacey_traction(i, vel, acc);
          }
          A[2 * g] += acc[0];
          A[2 * g + 1] += acc[1];
        }
    }
  }

  // ------------------------------------------------------------ sources ----
  // include/kokkos_kernels/impl/compute_source_interaction.tpp:84-101,
  // include/medium/dim2/elastic/isotropic/source.hpp:36-59
  void add_sources(int istep, std::vector<real> &A, real scale = 1) {
    if (acoustic) { // -stf L / kappa (acoustic/isotropic/source.hpp:36-70)
      for (int s = 0; s < nsrc; ++s)
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            const int i = idx(src_ispec[s], iz, ix);
            const int g = ibool[i];
            A[2 * g] += scale * (static_cast<real>(-1.0) *
                                 stf[(size_t(istep) * nsrc + s) * 2] *
                                 src_array[((s * 2) * NGLL + iz) * NGLL + ix]) /
                        kappa[i];
          }
      return;
    }
    for (int s = 0; s < nsrc; ++s)
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const int g = ibool[idx(src_ispec[s], iz, ix)];
          for (int c = 0; c < 2; ++c)
            A[2 * g + c] +=
                scale * stf[(size_t(istep) * nsrc + s) * 2 + c] *
                src_array[((s * 2 + c) * NGLL + iz) * NGLL + ix];
        }
  }

  // -------------------------------------------------------- locate point ----
  // src/algorithms/dim2/locate_point_impl.cpp:13-45 (rough), :47-85
  // (candidates), :245-283 (Newton), :285-345 (best candidate).
  void locate(real x, real z, int &ispec_out, real &xi_out, real &ga_out) {
    real dist_min = std::numeric_limits<real>::max();
    int ispec_sel = -1;
    for (int e = 0; e < nspec; ++e)
      for (int j = 0; j < NGLL; ++j)
        for (int i = 0; i < NGLL; ++i) {
          const real dx = x - gx[idx(e, j, i)], dz = z - gz[idx(e, j, i)];
          const real d = std::sqrt(dx * dx + dz * dz);
          if (d < dist_min) {
            dist_min = d;
            ispec_sel = e;
          }
        }
    std::vector<int> cand{ispec_sel};
    const int gc[4] = {ibool[idx(ispec_sel, 0, 0)],
                       ibool[idx(ispec_sel, NGLL - 1, 0)],
                       ibool[idx(ispec_sel, 0, NGLL - 1)],
                       ibool[idx(ispec_sel, NGLL - 1, NGLL - 1)]};
    for (int e = 0; e < nspec; ++e) {
      if (e == ispec_sel)
        continue;
      for (int j : {0, NGLL - 1})
        for (int i : {0, NGLL - 1}) {
          const int g = ibool[idx(e, j, i)];
          if (std::find(gc, gc + 4, g) != gc + 4)
            if (cand.back() != e)
              cand.push_back(e);
        }
    }
    real final_dist = std::numeric_limits<real>::max();
    ispec_out = -1;
    xi_out = -9999.0;
    ga_out = -9999.0;
    for (int e : cand) {
      real xi = 0.0, ga = 0.0;
      const real *ex = &cx[e * ngnod], *ez = &cz[e * ngnod];
      for (int it = 0; it < 100; ++it) {
        real lx, lz;
        compute_location<real>(ex, ez, ngnod, xi, ga, lx, lz);
        real dxi_s[9], dga_s[9];
        shape_derivatives<real>(xi, ga, ngnod, dxi_s, dga_s);
        const Jac<real> J = compute_jacobian<real>(ex, ez, ngnod, dxi_s, dga_s);
        const real dx = -(lx - x), dz = -(lz - z);
        const real dxi = J.xix * dx + J.xiz * dz;
        const real dga = J.gammax * dx + J.gammaz * dz;
        xi += dxi;
        ga += dga;
        if (xi > 1.01) xi = 1.01;
        if (xi < -1.01) xi = -1.01;
        if (ga > 1.01) ga = 1.01;
        if (ga < -1.01) ga = -1.01;
        if (std::abs(dxi) < 1e-12 && std::abs(dga) < 1e-12)
          break;
      }
      real lx, lz;
      compute_location<real>(ex, ez, ngnod, xi, ga, lx, lz);
      const real d = std::sqrt((x - lx) * (x - lx) + (z - lz) * (z - lz));
      if (d < final_dist) {
        ispec_out = e;
        xi_out = xi;
        ga_out = ga;
        final_dist = d;
      }
    }
    if (ispec_out < 0 || std::fabs(std::fabs(xi_out) - 1.01) < 1e-6 ||
        std::fabs(std::fabs(ga_out) - 1.01) < 1e-6)
      throw std::runtime_error("point lies outside the mesh");
  }

  // -------------------------------------------------------- sources set ----
  // Ricker STF (core/specfem/source_time_functions/ricker.cpp:9-66,
  // impl/time_functions.hpp:77-83), t0 rule (src/io/sources/dim2/read_sources.cpp:142-172),
  // source array (core/specfem/assembly/compute_source_array/dim2/impl/compute_source_array_from_vector.cpp:14-60,
  // core/specfem/source/dim2/vector_source/force_source.cpp:22-73).
  // Dirac (kind 1): factor * gaussian(t - tshift, f0) with f0 = 1 / (10 dt)
  // (core/specfem/source_time_functions/dirac.cpp:7-52,
  //  impl/time_functions.hpp:9-14).
  // Heaviside (kind 2, f0 holds the half duration): hdur raised to 5 dt,
  // t0 = -2 hdur + tshift, then hdur / 1.628 and factor * (1 + erf(t / hdur))
  // / 2 (core/specfem/source_time_functions/heaviside.cpp:10-78,
  //  impl/time_functions.hpp:132-136, setup/constants.hpp.in:52).
  // Moment tensors (src_type 1, moment = Mxx, Mzz, Mxz): the element's
  // Jacobian interpolated at the source, then per point
  // [Mxx Mxz; Mxz Mzz] (d/dx, d/dz)(h_xi h_gamma)
  // (core/specfem/assembly/compute_source_array/dim2/impl/
  //  compute_source_array_from_tensor.cpp:15-77,
  //  core/specfem/source/dim2/tensor_source/moment_tensor_source.cpp:21-34,
  //  include/algorithms/interpolate.hpp:73-94).
  std::vector<int> stf_kind;
  std::vector<real> f0_used;
  void set_sources(int nsrc_, const double *x, const double *z,
                   const float *angle, const float *f0, const float *tshift,
                   const float *factor, real dt, int nstep_,
                   const int *kinds = nullptr, const real *t0_all = nullptr,
                   const int *src_type = nullptr, const float *moment = nullptr) {
    nsrc = nsrc_;
    nstep = nstep_;
    src_ispec.assign(nsrc, 0);
    src_array.assign(size_t(nsrc) * 2 * NGLL * NGLL, 0);
    stf_kind.assign(nsrc, 0);
    f0_used.assign(nsrc, 0);
    std::vector<real> t0s(nsrc), tsh(nsrc);
    real min_t0 = std::numeric_limits<real>::max();
    for (int s = 0; s < nsrc; ++s) {
      stf_kind[s] = kinds ? kinds[s] : 0;
      if (stf_kind[s] < 0 || stf_kind[s] > 2)
        throw std::runtime_error("unknown source time function");
      if (stf_kind[s] == 2) {
        real hd = static_cast<real>(f0[s]);
        if (hd < 5.0 * dt)
          hd = 5.0 * dt;
        t0s[s] = -static_cast<real>(2.0) * hd + static_cast<real>(tshift[s]);
        f0_used[s] = hd / static_cast<real>(1.628); // the adjusted half duration
      } else {
        f0_used[s] = stf_kind[s] == 1 ? static_cast<real>(1.0 / (10.0 * dt))
                                      : static_cast<real>(f0[s]);
        const real f = f0_used[s];
        const real hdur = 1.0 / f;
        t0s[s] = -static_cast<real>(1.2) * hdur + static_cast<real>(tshift[s]);
      }
      tsh[s] = tshift[s];
      if (t0s[s] < min_t0)
        min_t0 = t0s[s];
    }
    // the start time is the minimum over every source of the run, which a
    // sub-mesh of a coupled run receives from outside
    if (t0_all)
      min_t0 = *t0_all;
    for (int s = 0; s < nsrc; ++s)
      tsh[s] = t0s[s] - min_t0;
    t0 = min_t0;
    stf.assign(size_t(nstep) * nsrc * 2, 0);
    const real pi = static_cast<real>(M_PI);
    for (int s = 0; s < nsrc; ++s) {
      for (int i = 0; i < nstep; ++i) {
        const real t = t0 + i * dt;
        const real tt = t - tsh[s];
        const real fs = f0_used[s];
        const real aa = pi * pi * fs * fs;
        real val;
        if (stf_kind[s] == 2) {
          const real hv = 0.5 * (1.0 + std::erf(tt / fs));
          val = static_cast<real>(factor[s]) * hv;
        } else if (stf_kind[s] == 1) {
          const real gs = -1.0 * std::exp(-aa * tt * tt) / (2.0 * aa);
          val = static_cast<real>(factor[s]) * gs;
        } else {
          const real d2 = (1.0 - 2.0 * aa * tt * tt) * std::exp(-aa * tt * tt);
          val = static_cast<real>(factor[s]) * d2;
        }
        stf[(size_t(i) * nsrc + s) * 2 + 0] = val;
        stf[(size_t(i) * nsrc + s) * 2 + 1] = val;
      }
      int ispec;
      real xi, ga;
      locate(static_cast<real>(static_cast<float>(x[s])),
             static_cast<real>(static_cast<float>(z[s])), ispec, xi, ga);
      src_ispec[s] = ispec;
      real hxi[NGLL], hga[NGLL], hpxi[NGLL], hpga[NGLL];
      lagrange_interpolants<real>(xi, gll.xi, hxi, hpxi);
      lagrange_interpolants<real>(ga, gll.xi, hga, hpga);
      if (src_type && src_type[s] == 1) {
        if (acoustic)
          throw std::runtime_error("moment-tensor sources need elastic elements");
        Jac<real> d{0, 0, 0, 0, 0};
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            const real hl = hxi[ix] * hga[iz];
            const Jac<real> &J = jac[idx(ispec, iz, ix)];
            d.xix += hl * J.xix;
            d.xiz += hl * J.xiz;
            d.gammax += hl * J.gammax;
            d.gammaz += hl * J.gammaz;
          }
        const real mxx = moment[3 * s], mzz = moment[3 * s + 1],
                   mxz = moment[3 * s + 2];
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            const real dx = (hpxi[ix] * d.xix) * hga[iz] + hxi[ix] * (hpga[iz] * d.gammax);
            const real dz = (hpxi[ix] * d.xiz) * hga[iz] + hxi[ix] * (hpga[iz] * d.gammaz);
            src_array[((s * 2 + 0) * NGLL + iz) * NGLL + ix] = mxx * dx + mxz * dz;
            src_array[((s * 2 + 1) * NGLL + iz) * NGLL + ix] = mxz * dx + mzz * dz;
          }
        continue;
      }
      const real ang = static_cast<real>(angle[s]) * pi /
                       static_cast<real>(180.0);
      real fv[2] = {std::sin(ang), std::cos(ang)};
      if (acoustic) { // force vector (1) (force_source.cpp:36-40)
        fv[0] = 1;
        fv[1] = 0;
      }
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const real hl = hxi[ix] * hga[iz];
          for (int c = 0; c < 2; ++c)
            src_array[((s * 2 + c) * NGLL + iz) * NGLL + ix] = hl * fv[c];
        }
    }
  }

  // core/specfem/assembly/receivers/dim2/receivers.cpp:63-104
  void set_receivers(int nrec_, const double *x, const double *z) {
    nrec = nrec_;
    rec_ispec.assign(nrec, 0);
    rec_h.assign(size_t(nrec) * NGLL * NGLL, 0);
    rec_xi.assign(nrec, 0);
    rec_gamma.assign(nrec, 0);
    for (int r = 0; r < nrec; ++r) {
      int ispec;
      real xi, ga;
      locate(static_cast<real>(static_cast<float>(x[r])),
             static_cast<real>(static_cast<float>(z[r])), ispec, xi, ga);
      rec_ispec[r] = ispec;
      rec_xi[r] = xi;
      rec_gamma[r] = ga;
      real hxi[NGLL], hga[NGLL];
      lagrange_interpolants<real>(xi, gll.xi, hxi, nullptr);
      lagrange_interpolants<real>(ga, gll.xi, hga, nullptr);
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix)
          rec_h[(size_t(r) * NGLL + iz) * NGLL + ix] = hxi[ix] * hga[iz];
    }
  }

  // include/kokkos_kernels/impl/compute_seismogram.tpp:117-166
  // (types: 0 displacement, 1 velocity, 2 acceleration)
  void seismograms(const std::vector<real> &U, const std::vector<real> &V,
                   const std::vector<real> &A, int ntypes, const int *types,
                   real *out /* [ntypes][nrec][2] */) const {
    if (acoustic) {
      seismograms_acoustic(U, V, A, ntypes, types, out);
      return;
    }
    for (int t = 0; t < ntypes; ++t) {
      if (types[t] == 3) {
        seismograms_elastic_pressure(U, out + size_t(t) * nrec * 2);
        continue;
      }
      const std::vector<real> &F = types[t] == 0 ? U : (types[t] == 1 ? V : A);
      for (int r = 0; r < nrec; ++r) {
        real s[2] = {0, 0};
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            const int g = ibool[idx(rec_ispec[r], iz, ix)];
            const real h = rec_h[(size_t(r) * NGLL + iz) * NGLL + ix];
            s[0] += h * F[2 * g];
            s[1] += h * F[2 * g + 1];
          }
        out[(size_t(t) * nrec + r) * 2 + 0] = s[0];
        out[(size_t(t) * nrec + r) * 2 + 1] = s[1];
      }
    }
  }

  // Elastic P-SV pressure: p = -(lambda + 2/3 mu)(du_x/dx + du_z/dz) at the
  // receiver element's points, formed in double from the FP32 lambda, mu and
  // gradient, stored in FP32, then interpolated
  // (medium/dim2/elastic/isotropic/wavefield.hpp:57-94,
  // compute_seismogram.tpp:117-166).  Only the first component is a
  // seismogram (one PRE trace); the second is written 0.
  void seismograms_elastic_pressure(const std::vector<real> &U,
                                    real *out /* [nrec][2] */) const {
    for (int r = 0; r < nrec; ++r) {
      const int e = rec_ispec[r];
      real du[NGLL][NGLL][2][2];
      element_gradient(e, U, du);
      real s = 0;
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const int i = idx(e, iz, ix);
          const double coef = static_cast<double>(lambda(i)) +
                              (2.0 / 3.0) * static_cast<double>(mu[i]);
          const real p = static_cast<real>(
              -1.0 * (coef * static_cast<double>(du[iz][ix][0][0] + du[iz][ix][1][1])));
          const real h = rec_h[(size_t(r) * NGLL + iz) * NGLL + ix];
          s += h * p;
        }
      out[size_t(r) * 2 + 0] = s;
      out[size_t(r) * 2 + 1] = 0;
    }
  }

  // Acoustic receivers: rho^-1 grad of the potential (d), its velocity (v)
  // or acceleration (a), or -acceleration for pressure (type 3), at the
  // element's points, then interpolated (acoustic/isotropic/wavefield.hpp:
  // 12-97, compute_seismogram.tpp:117-166).
  void seismograms_acoustic(const std::vector<real> &U, const std::vector<real> &V,
                            const std::vector<real> &A, int ntypes,
                            const int *types, real *out) const {
    for (int t = 0; t < ntypes; ++t) {
      const std::vector<real> &F = types[t] == 0 ? U : (types[t] == 1 ? V : A);
      for (int r = 0; r < nrec; ++r) {
        const int e = rec_ispec[r];
        real du[NGLL][NGLL][2][2];
        if (types[t] != 3)
          element_gradient(e, F, du);
        real s[2] = {0, 0};
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            const int i = idx(e, iz, ix);
            const real h = rec_h[(size_t(r) * NGLL + iz) * NGLL + ix];
            if (types[t] == 3) {
              s[0] += h * (static_cast<real>(-1.0) * F[2 * ibool[i]]);
            } else {
              const real ri = rho_inverse(i);
              s[0] += h * (ri * du[iz][ix][0][0]);
              s[1] += h * (ri * du[iz][ix][0][1]);
            }
          }
        out[(size_t(t) * nrec + r) * 2 + 0] = s[0];
        out[(size_t(t) * nrec + r) * 2 + 1] = s[1];
      }
    }
  }

  // ------------------------------------------------------- time loop ----
  // core/specfem/solver/time_marching.tpp:33-116 for a pure elastic P-SV
  // mesh: predictor -> sources -> stiffness -> divide -> corrector ->
  // seismograms.
  void set_timescheme(real dt) {
    deltat = dt;
    deltatover2 = dt / 2.0;
    deltasquareover2 = dt * dt / 2.0;
  }
  void forward_step(int istep) {
    predictor(u, v, a, deltat, deltatover2, deltasquareover2);
    add_sources(istep, a);
    stiffness(u, v, a);
    divide_mass(a);
    corrector(v, a, deltatover2);
  }

  // ------------------------------------------- discrete adjoint (checker) --
  // Exact adjoint of the time step above (not a reference routine: the
  // reference reconstructs the forward field backwards and accumulates the
  // Tromp et al. (2005) kernels; see ref_kernels below for that formula).
  // Forward step k maps (u, v, a)_k to (u, v, a)_{k+1}:
  //   u~ = u + dt v + dt^2/2 a,  v~ = v + dt/2 a,
  //   a' = W (f_k - K u~ - B v~)  with W = 1/M' (M' = M + dt/2 rowsum B),
  //   u' = u~,  v' = v~ + dt/2 a'.
  // Stored per step: u_{k+1}, v~_k (before the corrector), a_{k+1}.
  std::vector<std::vector<real>> st_u, st_v, st_a;

  void forward_store(int nsteps, real *seis /* [nsteps][nrec][2] */) {
    std::fill(u.begin(), u.end(), real(0));
    std::fill(v.begin(), v.end(), real(0));
    std::fill(a.begin(), a.end(), real(0));
    st_u.assign(nsteps, {});
    st_v.assign(nsteps, {});
    st_a.assign(nsteps, {});
    const int types[1] = {0};
    for (int k = 0; k < nsteps; ++k) {
      predictor(u, v, a, deltat, deltatover2, deltasquareover2);
      add_sources(k, a);
      stiffness(u, v, a);
      divide_mass(a);
      st_u[k] = u;
      st_v[k] = v;
      st_a[k] = a;
      corrector(v, a, deltatover2);
      seismograms(u, v, a, 1, types, seis + size_t(k) * nrec * 2);
    }
  }

  // y += B x at the absorbing points (B v is minus the Stacey traction).
  void apply_B(const std::vector<real> &x, std::vector<real> &y) const {
    for (int e : order) {
      if (!elem_stacey[e])
        continue;
      for (int p = 0; p < NGLL * NGLL; ++p) {
        const int i = e * NGLL * NGLL + p;
        if (!pt_stacey[i])
          continue;
        const int g = ibool[i];
        real t[2] = {0, 0};
        const real xv[2] = {x[2 * g], x[2 * g + 1]};
        stacey_traction(i, xv, t);
        y[2 * g] -= t[0];
        y[2 * g + 1] -= t[1];
      }
    }
  }

  // Gradient of sum_k chi_k(u_{k+1}) w.r.t. (kappa, mu, rho) at every local
  // point, given dchi/ds_k = gsrc[k][r][c] for the displacement seismograms.
  // refk (optional, [6][nloc]): the reference's kernel formula evaluated on
  // this adjoint's fields with the adjoint displacement s+ = rb / dt and the
  // forward displacement s = u~ (compute_frechet_derivatives,
  // include/medium/dim2/elastic/isotropic/frechet_derivative.hpp:98-178, as
  // called by compute_material_derivatives.tpp:112-151):
  //   [0..2] kappa, mu, rho kernels with the reference's two quirks -- the
  //          "backward" gradient is grad(s+ + s) (the two-field gradient
  //          reuses its accumulators, include/algorithms/gradient.hpp:302-316)
  //          and kappa' = (lambda+2mu) - mu = kappa + mu/3 (line 77);
  //   [3..5] the same formula with grad(s) and kappa.
  // The density term uses s+ . a' (forward acceleration), the time-by-parts
  // form of the reference's adjoint-acceleration . backward-displacement.
  void adjoint(int nsteps, const double *gsrc, double *grad /* [3][nloc] */,
               double *refk = nullptr) {
    const size_t n2 = size_t(nglob) * 2, nloc = size_t(nspec) * NGLL * NGLL;
    std::vector<real> ub(n2, 0), vb(n2, 0), ab(n2, 0), rb(n2, 0), Kr(n2),
        Br(n2), zero(n2, 0);
    std::vector<double> gk(nloc, 0), gm(nloc, 0), gr(nloc, 0);
    const real c = deltatover2, h = deltasquareover2, dt = deltat;
    for (int k = nsteps - 1; k >= 0; --k) {
      // misfit source: ub += H^T g_k
      for (int r = 0; r < nrec; ++r)
        for (int p = 0; p < NGLL * NGLL; ++p) {
          const int g = ibool[rec_ispec[r] * NGLL * NGLL + p];
          const real hw = rec_h[size_t(r) * NGLL * NGLL + p];
          for (int cc = 0; cc < 2; ++cc)
            ub[2 * g + cc] += hw * real(gsrc[(size_t(k) * nrec + r) * 2 + cc]);
        }
      for (size_t i = 0; i < n2; ++i)
        rb[i] = minv[i] * (ab[i] + c * vb[i]);
      // K rb (no boundary) and B rb
      std::fill(Kr.begin(), Kr.end(), real(0));
      stiffness(rb, zero, Kr, false);
      for (auto &x : Kr)
        x = -x;
      std::fill(Br.begin(), Br.end(), real(0));
      apply_B(rb, Br);
      // parameter gradient of step k
      const std::vector<real> &U1 = st_u[k], &V0 = st_v[k], &A1 = st_a[k];
      for (int e = 0; e < nspec; ++e) {
        real dr[NGLL][NGLL][2][2], du[NGLL][NGLL][2][2];
        element_gradient(e, rb, dr);
        element_gradient(e, U1, du);
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            const int i = idx(e, iz, ix);
            const int g = ibool[i];
            const double wj = double(gll.w[iz]) * gll.w[ix] * jac[i].jacobian;
            const double divr = double(dr[iz][ix][0][0]) + dr[iz][ix][1][1];
            const double divu = double(du[iz][ix][0][0]) + du[iz][ix][1][1];
            const double exz_r = 0.5 * (double(dr[iz][ix][0][1]) + dr[iz][ix][1][0]);
            const double exz_u = 0.5 * (double(du[iz][ix][0][1]) + du[iz][ix][1][0]);
            const double ee = double(dr[iz][ix][0][0]) * du[iz][ix][0][0] +
                              double(dr[iz][ix][1][1]) * du[iz][ix][1][1] +
                              2.0 * exz_r * exz_u;
            gk[i] -= wj * divr * divu;
            gm[i] -= wj * (2.0 * ee - 2.0 / 3.0 * divr * divu);
            gr[i] -= wj * (double(rb[2 * g]) * A1[2 * g] +
                           double(rb[2 * g + 1]) * A1[2 * g + 1]);
            if (refk) {
              double ad[2][2], bq[2][2], bc[2][2];
              for (int c = 0; c < 2; ++c)
                for (int d = 0; d < 2; ++d) {
                  ad[c][d] = double(dr[iz][ix][c][d]) / double(dt);
                  bc[c][d] = du[iz][ix][c][d];
                  bq[c][d] = ad[c][d] + bc[c][d];
                }
              auto point_kernels = [&](const double (&b)[2][2], double kap,
                                       double *out3) {
                const double ad_xz = 0.5 * (ad[1][0] + ad[0][1]);
                const double b_xz = 0.5 * (b[1][0] + b[0][1]);
                const double k_kl = (ad[0][0] + ad[1][1]) * (b[0][0] + b[1][1]);
                const double m_kl = ad[0][0] * b[0][0] + ad[1][1] * b[1][1] +
                                    2.0 * ad_xz * b_xz - 1.0 / 3.0 * k_kl;
                const double r_kl = (double(rb[2 * g]) * A1[2 * g] +
                                     double(rb[2 * g + 1]) * A1[2 * g + 1]) /
                                    double(dt);
                out3[0] = -kap * double(dt) * k_kl;
                out3[1] = -2.0 * double(mu[i]) * double(dt) * m_kl;
                out3[2] = -double(rho[i]) * double(dt) * r_kl;
              };
              const double l2m = double(kappa[i]) + 4.0 / 3.0 * double(mu[i]);
              double q[3], c3[3];
              point_kernels(bq, l2m - double(mu[i]), q);
              point_kernels(bc, double(kappa[i]), c3);
              for (int t = 0; t < 3; ++t) {
                refk[t * nloc + i] += q[t];
                refk[(3 + t) * nloc + i] += c3[t];
              }
            }
            if (elem_stacey[e] && pt_stacey[i]) {
              const double n0 = pt_n0[i], n1 = pt_n1[i];
              const double nrm = std::sqrt(n0 * n0 + n1 * n1);
              const double nh0 = n0 / nrm, nh1 = n1 / nrm, wt = pt_w[i];
              const double r0 = rb[2 * g], r1 = rb[2 * g + 1];
              const double v0 = V0[2 * g], v1 = V0[2 * g + 1];
              const double a0 = A1[2 * g], a1 = A1[2 * g + 1];
              const double rn = r0 * nh0 + r1 * nh1, vn = v0 * nh0 + v1 * nh1;
              const double s1 = nh0 + nh1;
              const double dP = -(nrm * wt * rn * vn +
                                  0.5 * dt * nrm * wt *
                                      (r0 * a0 * nh0 * s1 + r1 * a1 * nh1 * s1));
              const double dS =
                  -(nrm * wt * (r0 * v0 + r1 * v1 - rn * vn) +
                    0.5 * dt * nrm * wt *
                        (r0 * a0 * (1.0 - nh0 * s1) + r1 * a1 * (1.0 - nh1 * s1)));
              const double l2m = double(kappa[i]) + 4.0 / 3.0 * double(mu[i]);
              const double rh = rho[i], m = mu[i];
              const double P = std::sqrt(rh * l2m), S = std::sqrt(rh * m);
              gk[i] += dP * rh / (2.0 * P);
              gm[i] += dP * (4.0 / 3.0) * rh / (2.0 * P) + dS * rh / (2.0 * S);
              gr[i] += dP * l2m / (2.0 * P) + dS * m / (2.0 * S);
            }
          }
      }
      // adjoint state update
      for (size_t i = 0; i < n2; ++i) {
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

// ================================================ coupled (Coupled2D) ====
// Elastic-acoustic coupling on a conforming interface, restated as two
// single-medium oracles (the reference keeps one field per medium) plus the
// interface terms of include/medium/dim2/coupling_terms/
// {acoustic_elastic,elastic_acoustic}.hpp:
//   acoustic self point:  chi_dd += w (n . u_elastic)        (u after the predictor)
//   elastic self point:   a_c    += (w n_c) chi_dd_acoustic   (chi_dd of this step)
// with n the self element's edge normal (compute_normal, not unit length)
// and w the GLL weight along the edge (core/specfem/assembly/
// conforming_interfaces/dim2/impl/interface_container.tpp:20-80); the
// acoustic term is zeroed at free-surface points (compute_coupling.tpp:
// 70-90).  The step order is time_marching.tpp:33-77: both predictors, the
// acoustic update (coupling, sources, stiffness, divide) and corrector, then
// the elastic update and corrector.
template <typename real> struct Coupled2D {
  Oracle2D<real> el, ac;
  struct Ent {
    int i_self;  // local point of the self sub-oracle
    int g_other; // global point of the other sub-oracle
    real factor, n0, n1;
  };
  std::vector<Ent> ac_ent, el_ent;
  int nrec = 0;
  std::vector<int> rec_med, rec_sub; // station -> (0 elastic / 1 acoustic, index)
  real t0 = 0;
  // test knob (planted mistakes on the CPU): scales of the two coupling
  // terms, 1 for the restatement
  real ac_scale = 1, el_scale = 1;

  void build(int nspec, int ngnod, int nnodes, const double *coord,
             const int *knods, const float *kappa, const float *mu,
             const float *rho, int nabs, const int *abs_ispec,
             const int *abs_type, int nfree, const int *free_ispec,
             const int *free_type, real dt, int nstep, int nsrc,
             const double *sx, const double *sz, const float *angle,
             const float *f0, const float *tshift, const float *factor,
             const int *kinds, int nrec_, const double *rx, const double *rz) {
    const int NP = NGLL * NGLL;
    std::vector<int> medium(nspec), sub(nspec);
    std::vector<int> knods_m[2];
    std::vector<float> mat_m[2][3];
    int count[2] = {0, 0};
    for (int e = 0; e < nspec; ++e) {
      bool fluid = true;
      for (int p = 0; p < NP; ++p)
        fluid = fluid && mu[size_t(e) * NP + p] == 0.0f;
      const int m = fluid ? 1 : 0;
      medium[e] = m;
      sub[e] = count[m]++;
      for (int n = 0; n < ngnod; ++n)
        knods_m[m].push_back(knods[size_t(e) * ngnod + n]);
      for (int p = 0; p < NP; ++p) {
        mat_m[m][0].push_back(kappa[size_t(e) * NP + p]);
        mat_m[m][1].push_back(mu[size_t(e) * NP + p]);
        mat_m[m][2].push_back(rho[size_t(e) * NP + p]);
      }
    }
    if (count[0] == 0 || count[1] == 0)
      throw std::runtime_error("a coupled run needs both media");
    std::vector<int> abs_e[2], abs_t[2], fr_e, fr_t;
    for (int i = 0; i < nabs; ++i) {
      const int m = medium[abs_ispec[i]];
      abs_e[m].push_back(sub[abs_ispec[i]]);
      abs_t[m].push_back(abs_type[i]);
    }
    for (int i = 0; i < nfree; ++i) {
      if (medium[free_ispec[i]] != 1)
        throw std::runtime_error("free-surface entity on an elastic element");
      fr_e.push_back(sub[free_ispec[i]]);
      fr_t.push_back(free_type[i]);
    }
    el.build_mesh(count[0], ngnod, nnodes, coord, knods_m[0].data(),
                  mat_m[0][0].data(), mat_m[0][1].data(), mat_m[0][2].data(),
                  int(abs_e[0].size()), abs_e[0].data(), abs_t[0].data());
    ac.build_mesh(count[1], ngnod, nnodes, coord, knods_m[1].data(),
                  mat_m[1][0].data(), mat_m[1][1].data(), mat_m[1][2].data(),
                  int(abs_e[1].size()), abs_e[1].data(), abs_t[1].data(),
                  int(fr_e.size()), fr_e.data(), fr_t.data());
    el.set_timescheme(dt);
    ac.set_timescheme(dt);

    // medium of a point: the sub-mesh whose best candidate lies closest
    // (a point inside one medium falls outside the other sub-mesh)
    auto medium_of = [&](double x, double z) {
      real d[2];
      Oracle2D<real> *o[2] = {&el, &ac};
      for (int m = 0; m < 2; ++m) {
        try {
          int e;
          real xi, ga;
          o[m]->locate(static_cast<real>(static_cast<float>(x)),
                       static_cast<real>(static_cast<float>(z)), e, xi, ga);
          real lx, lz;
          compute_location<real>(&o[m]->cx[e * ngnod], &o[m]->cz[e * ngnod],
                                 ngnod, xi, ga, lx, lz);
          d[m] = std::hypot(lx - static_cast<real>(static_cast<float>(x)),
                            lz - static_cast<real>(static_cast<float>(z)));
        } catch (const std::exception &) {
          d[m] = std::numeric_limits<real>::max();
        }
      }
      if (d[0] == std::numeric_limits<real>::max() &&
          d[1] == std::numeric_limits<real>::max())
        throw std::runtime_error("point lies outside the mesh");
      return d[1] < d[0] ? 1 : 0;
    };
    // sources per medium; one start time for the whole run
    std::vector<int> src_m[2];
    real t0_all = std::numeric_limits<real>::max();
    for (int k = 0; k < nsrc; ++k) {
      const int kind = kinds ? kinds[k] : 0;
      const real f = kind == 1 ? static_cast<real>(1.0 / (10.0 * dt))
                               : static_cast<real>(f0[k]);
      const real hdur = 1.0 / f;
      const real t0k =
          -static_cast<real>(1.2) * hdur + static_cast<real>(tshift[k]);
      t0_all = std::min(t0_all, t0k);
      src_m[medium_of(sx[k], sz[k])].push_back(k);
    }
    if (nsrc == 0)
      t0_all = 0;
    t0 = t0_all;
    Oracle2D<real> *o[2] = {&el, &ac};
    for (int m = 0; m < 2; ++m) {
      std::vector<double> x, z;
      std::vector<float> an, ff, ts, fa;
      std::vector<int> kd;
      for (int k : src_m[m]) {
        x.push_back(sx[k]);
        z.push_back(sz[k]);
        an.push_back(angle[k]);
        ff.push_back(f0[k]);
        ts.push_back(tshift[k]);
        fa.push_back(factor[k]);
        kd.push_back(kinds ? kinds[k] : 0);
      }
      o[m]->set_sources(int(x.size()), x.data(), z.data(), an.data(),
                        ff.data(), ts.data(), fa.data(), dt, nstep, kd.data(),
                        &t0_all);
    }
    // receivers per medium
    nrec = nrec_;
    rec_med.assign(nrec, 0);
    rec_sub.assign(nrec, 0);
    std::vector<double> rxm[2], rzm[2];
    for (int r = 0; r < nrec; ++r) {
      const int m = medium_of(rx[r], rz[r]);
      rec_med[r] = m;
      rec_sub[r] = int(rxm[m].size());
      rxm[m].push_back(rx[r]);
      rzm[m].push_back(rz[r]);
    }
    for (int m = 0; m < 2; ++m)
      o[m]->set_receivers(int(rxm[m].size()), rxm[m].data(), rzm[m].data());
    el.initialize(dt);
    ac.initialize(dt);
    build_interface(nspec, ngnod, knods, medium, sub);
  }

  // Interface edges: an acoustic and an elastic element sharing an edge
  // (same two corner nodes); the points pair by position along the edge.
  void build_interface(int nspec, int ngnod, const int *knods,
                       const std::vector<int> &medium,
                       const std::vector<int> &sub) {
    // edge types 1 bottom, 2 right, 3 top, 4 left; corners (start, end)
    // in the direction the edge's GLL index increases
    const int cs[5][2] = {{0, 0}, {0, 1}, {1, 2}, {3, 2}, {0, 3}};
    auto pt = [](int type, int k, int &iz, int &ix) {
      iz = type == 1 ? 0 : (type == 3 ? NGLL - 1 : k);
      ix = type == 4 ? 0 : (type == 2 ? NGLL - 1 : k);
    };
    struct Side {
      int e, type, start;
    };
    std::map<std::pair<int, int>, std::vector<Side>> edges;
    for (int e = 0; e < nspec; ++e)
      for (int t = 1; t <= 4; ++t) {
        const int a = knods[size_t(e) * ngnod + cs[t][0]],
                  b = knods[size_t(e) * ngnod + cs[t][1]];
        edges[{std::min(a, b), std::max(a, b)}].push_back({e, t, a});
      }
    for (auto &kv : edges) {
      if (kv.second.size() != 2 ||
          medium[kv.second[0].e] == medium[kv.second[1].e])
        continue;
      for (int self = 0; self < 2; ++self) {
        const Side &S = kv.second[self], &O = kv.second[1 - self];
        const bool self_ac = medium[S.e] == 1;
        Oracle2D<real> &so = self_ac ? ac : el;
        Oracle2D<real> &oo = self_ac ? el : ac;
        for (int k = 0; k < NGLL; ++k) {
          int iz, ix, jz, jx;
          pt(S.type, k, iz, ix);
          pt(O.type, O.start == S.start ? k : NGLL - 1 - k, jz, jx);
          const int i = so.idx(sub[S.e], iz, ix);
          const Jac<real> &J = so.jac[i];
          real n0, n1, w;
          // core/specfem/point/jacobian_matrix.hpp:424-449 (compute_normal)
          if (S.type == 4) {
            n0 = static_cast<real>(-1.0) * J.xix * J.jacobian;
            n1 = static_cast<real>(-1.0) * J.xiz * J.jacobian;
            w = so.gll.w[iz];
          } else if (S.type == 2) {
            n0 = J.xix * J.jacobian;
            n1 = J.xiz * J.jacobian;
            w = so.gll.w[iz];
          } else if (S.type == 3) {
            n0 = J.gammax * J.jacobian;
            n1 = J.gammaz * J.jacobian;
            w = so.gll.w[ix];
          } else {
            n0 = static_cast<real>(-1.0) * J.gammax * J.jacobian;
            n1 = static_cast<real>(-1.0) * J.gammaz * J.jacobian;
            w = so.gll.w[ix];
          }
          const int g = oo.ibool[oo.idx(sub[O.e], jz, jx)];
          (self_ac ? ac_ent : el_ent).push_back({i, g, w, n0, n1});
        }
      }
    }
  }

  void step(int istep) {
    el.predictor(el.u, el.v, el.a, el.deltat, el.deltatover2,
                 el.deltasquareover2);
    ac.predictor(ac.u, ac.v, ac.a, ac.deltat, ac.deltatover2,
                 ac.deltasquareover2);
    // acoustic: coupling (elastic displacement), sources, stiffness, divide
    for (const Ent &t : ac_ent) {
      if (ac.pt_free[t.i_self])
        continue;
      const real c = t.factor * (t.n0 * el.u[2 * t.g_other] +
                                 t.n1 * el.u[2 * t.g_other + 1]);
      ac.a[2 * ac.ibool[t.i_self]] += ac_scale * c;
    }
    ac.add_sources(istep, ac.a);
    ac.stiffness(ac.u, ac.v, ac.a);
    ac.divide_mass(ac.a);
    ac.corrector(ac.v, ac.a, ac.deltatover2);
    // elastic: coupling (acoustic acceleration of this step), sources, ...
    for (const Ent &t : el_ent) {
      const int g = el.ibool[t.i_self];
      const real chi = el_scale * ac.a[2 * t.g_other];
      el.a[2 * g] += t.factor * t.n0 * chi;
      el.a[2 * g + 1] += t.factor * t.n1 * chi;
    }
    el.add_sources(istep, el.a);
    el.stiffness(el.u, el.v, el.a);
    el.divide_mass(el.a);
    el.corrector(el.v, el.a, el.deltatover2);
  }

  void seismograms(int ntypes, const int *types, real *out) const {
    std::vector<real> b[2];
    const Oracle2D<real> *o[2] = {&el, &ac};
    for (int m = 0; m < 2; ++m) {
      b[m].assign(size_t(ntypes) * std::max(1, o[m]->nrec) * 2, 0);
      o[m]->seismograms(o[m]->u, o[m]->v, o[m]->a, ntypes, types, b[m].data());
    }
    for (int t = 0; t < ntypes; ++t)
      for (int r = 0; r < nrec; ++r)
        for (int c = 0; c < 2; ++c) {
          const int m = rec_med[r];
          out[(size_t(t) * nrec + r) * 2 + c] =
              b[m][(size_t(t) * o[m]->nrec + rec_sub[r]) * 2 + c];
        }
  }
};

// ===================================================== C interface ======
struct OracleHandle {
  int precision; // 4 or 8
  Oracle2D<float> *f;
  Oracle2D<double> *d;
};

static std::string g_err;

extern "C" {

const char *oracle2d_last_error() { return g_err.c_str(); }

// oracle2d_create with acoustic free-surface entities (nfree, element,
// type 1..8) and source time-function kinds (0 Ricker, 1 Dirac; NULL: Ricker).
// As oracle2d_create_ex, with source kinds (0 point force, 1 moment tensor;
// NULL: all forces) and moment tensors [nsrc][3] (Mxx, Mzz, Mxz).
void *oracle2d_create_mt(int precision, int nspec, int ngnod, int nnodes,
                         const double *coord, const int *knods,
                         const float *kappa, const float *mu, const float *rho,
                         int nabs, const int *abs_ispec, const int *abs_type,
                         int nfree, const int *free_ispec, const int *free_type,
                         float dt, int nstep, int nsrc, const double *sx,
                         const double *sz, const float *angle, const float *f0,
                         const float *tshift, const float *factor,
                         const int *stf_kind, int nrec, const double *rx,
                         const double *rz, const int *src_type,
                         const float *moment);

void *oracle2d_create_ex(int precision, int nspec, int ngnod, int nnodes,
                         const double *coord, const int *knods,
                         const float *kappa, const float *mu, const float *rho,
                         int nabs, const int *abs_ispec, const int *abs_type,
                         int nfree, const int *free_ispec, const int *free_type,
                         float dt, int nstep, int nsrc, const double *sx,
                         const double *sz, const float *angle, const float *f0,
                         const float *tshift, const float *factor,
                         const int *stf_kind, int nrec, const double *rx,
                         const double *rz) {
  return oracle2d_create_mt(precision, nspec, ngnod, nnodes, coord, knods,
                            kappa, mu, rho, nabs, abs_ispec, abs_type, nfree,
                            free_ispec, free_type, dt, nstep, nsrc, sx, sz,
                            angle, f0, tshift, factor, stf_kind, nrec, rx, rz,
                            nullptr, nullptr);
}

void *oracle2d_create_mt(int precision, int nspec, int ngnod, int nnodes,
                         const double *coord, const int *knods,
                         const float *kappa, const float *mu, const float *rho,
                         int nabs, const int *abs_ispec, const int *abs_type,
                         int nfree, const int *free_ispec, const int *free_type,
                         float dt, int nstep, int nsrc, const double *sx,
                         const double *sz, const float *angle, const float *f0,
                         const float *tshift, const float *factor,
                         const int *stf_kind, int nrec, const double *rx,
                         const double *rz, const int *src_type,
                         const float *moment) {
  try {
    auto *h = new OracleHandle{precision, nullptr, nullptr};
    auto setup = [&](auto *o) {
      using real = std::remove_reference_t<decltype(o->deltat)>;
      o->build_mesh(nspec, ngnod, nnodes, coord, knods, kappa, mu, rho, nabs,
                    abs_ispec, abs_type, nfree, free_ispec, free_type);
      o->set_timescheme(static_cast<real>(dt));
      o->set_sources(nsrc, sx, sz, angle, f0, tshift, factor,
                     static_cast<real>(dt), nstep, stf_kind, nullptr, src_type,
                     moment);
      o->set_receivers(nrec, rx, rz);
      o->initialize(static_cast<real>(dt));
    };
    if (precision == 8) {
      h->d = new Oracle2D<double>();
      setup(h->d);
    } else {
      h->f = new Oracle2D<float>();
      setup(h->f);
    }
    return h;
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

void *oracle2d_create(int precision, int nspec, int ngnod, int nnodes,
                      const double *coord, const int *knods,
                      const float *kappa, const float *mu, const float *rho,
                      int nabs, const int *abs_ispec, const int *abs_type,
                      float dt, int nstep, int nsrc, const double *sx,
                      const double *sz, const float *angle, const float *f0,
                      const float *tshift, const float *factor, int nrec,
                      const double *rx, const double *rz) {
  return oracle2d_create_ex(precision, nspec, ngnod, nnodes, coord, knods,
                            kappa, mu, rho, nabs, abs_ispec, abs_type, 0,
                            nullptr, nullptr, dt, nstep, nsrc, sx, sz, angle,
                            f0, tshift, factor, nullptr, nrec, rx, rz);
}

// Coupled elastic-acoustic run (every argument as oracle2d_create_ex).
void *oracle2d_coupled_create(int precision, int nspec, int ngnod, int nnodes,
                              const double *coord, const int *knods,
                              const float *kappa, const float *mu,
                              const float *rho, int nabs, const int *abs_ispec,
                              const int *abs_type, int nfree,
                              const int *free_ispec, const int *free_type,
                              float dt, int nstep, int nsrc, const double *sx,
                              const double *sz, const float *angle,
                              const float *f0, const float *tshift,
                              const float *factor, const int *stf_kind,
                              int nrec, const double *rx, const double *rz) {
  try {
    if (precision == 8) {
      auto *c = new Coupled2D<double>();
      c->build(nspec, ngnod, nnodes, coord, knods, kappa, mu, rho, nabs,
               abs_ispec, abs_type, nfree, free_ispec, free_type, dt, nstep,
               nsrc, sx, sz, angle, f0, tshift, factor, stf_kind, nrec, rx,
               rz);
      return new std::pair<int, void *>(8, c);
    }
    auto *c = new Coupled2D<float>();
    c->build(nspec, ngnod, nnodes, coord, knods, kappa, mu, rho, nabs,
             abs_ispec, abs_type, nfree, free_ispec, free_type, dt, nstep, nsrc,
             sx, sz, angle, f0, tshift, factor, stf_kind, nrec, rx, rz);
    return new std::pair<int, void *>(4, c);
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

double oracle2d_coupled_t0(void *hv) {
  auto *h = static_cast<std::pair<int, void *> *>(hv);
  return h->first == 8 ? static_cast<Coupled2D<double> *>(h->second)->t0
                       : static_cast<Coupled2D<float> *>(h->second)->t0;
}

// Runs nstep steps from istep0, seismograms every nsbs steps into
// out[nsig][ntypes][nrec][2]; returns the number of samples.
int oracle2d_coupled_run(void *hv, int istep0, int nstep, int nsbs, int ntypes,
                         const int *types, double *out) {
  auto *h = static_cast<std::pair<int, void *> *>(hv);
  auto run = [&](auto *c) {
    using real = std::remove_reference_t<decltype(c->t0)>;
    std::vector<real> buf(size_t(ntypes) * std::max(1, c->nrec) * 2);
    int isig = 0;
    for (int istep = istep0; istep < istep0 + nstep; ++istep) {
      c->step(istep);
      if (istep % nsbs == 0) {
        c->seismograms(ntypes, types, buf.data());
        for (size_t k = 0; k < size_t(ntypes) * c->nrec * 2; ++k)
          out[size_t(isig) * ntypes * c->nrec * 2 + k] = buf[k];
        ++isig;
      }
    }
    return isig;
  };
  try {
    return h->first == 8 ? run(static_cast<Coupled2D<double> *>(h->second))
                         : run(static_cast<Coupled2D<float> *>(h->second));
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Planted mistakes (tests only): scale the fluid-side and solid-side
// coupling terms.
void oracle2d_coupled_scale(void *hv, double ac_scale, double el_scale) {
  auto *h = static_cast<std::pair<int, void *> *>(hv);
  if (h->first == 8) {
    auto *c = static_cast<Coupled2D<double> *>(h->second);
    c->ac_scale = ac_scale;
    c->el_scale = el_scale;
  } else {
    auto *c = static_cast<Coupled2D<float> *>(h->second);
    c->ac_scale = static_cast<float>(ac_scale);
    c->el_scale = static_cast<float>(el_scale);
  }
}

void oracle2d_coupled_destroy(void *hv) {
  auto *h = static_cast<std::pair<int, void *> *>(hv);
  if (h->first == 8)
    delete static_cast<Coupled2D<double> *>(h->second);
  else
    delete static_cast<Coupled2D<float> *>(h->second);
  delete h;
}

void oracle2d_destroy(void *hv) {
  auto *h = static_cast<OracleHandle *>(hv);
  delete h->f;
  delete h->d;
  delete h;
}

int oracle2d_nglob(void *hv) {
  auto *h = static_cast<OracleHandle *>(hv);
  return h->f ? h->f->nglob : h->d->nglob;
}

double oracle2d_t0(void *hv) {
  auto *h = static_cast<OracleHandle *>(hv);
  return h->f ? h->f->t0 : h->d->t0;
}

// Runs nstep forward steps from the current state, sampling seismograms every
// nsbs steps into out[nsig][ntypes][nrec][2] (as double).
int oracle2d_run_forward(void *hv, int istep0, int nstep, int nsbs, int ntypes,
                         const int *types, double *out) {
  auto *h = static_cast<OracleHandle *>(hv);
  auto run = [&](auto *o) {
    using real = std::remove_reference_t<decltype(o->deltat)>;
    std::vector<real> buf(size_t(ntypes) * o->nrec * 2);
    int isig = 0;
    for (int istep = istep0; istep < istep0 + nstep; ++istep) {
      o->forward_step(istep);
      if (istep % nsbs == 0) {
        o->seismograms(o->u, o->v, o->a, ntypes, types, buf.data());
        for (size_t k = 0; k < buf.size(); ++k)
          out[size_t(isig) * buf.size() + k] = buf[k];
        ++isig;
      }
    }
    return isig;
  };
  try {
    return h->f ? run(h->f) : run(h->d);
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// CPU-baseline timing: nsteps of the colour-parallel forward step on
// `nthreads` OpenMP threads from the current state; returns the threads
// used and the wall seconds of the loop.
int oracle2d_time_parallel(void *hv, int istep0, int nsteps, int nthreads,
                           double *seconds) {
  auto *h = static_cast<OracleHandle *>(hv);
  try {
    omp_set_num_threads(nthreads);
    int used = 1;
#pragma omp parallel
    {
#pragma omp single
      used = omp_get_num_threads();
    }
    auto go = [&](auto *o) {
      o->build_colour_groups();
      const auto t0 = std::chrono::steady_clock::now();
      for (int k = 0; k < nsteps; ++k)
        o->forward_step_parallel(istep0 + k);
      *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    if (h->f)
      go(h->f);
    else
      go(h->d);
    return used;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Copies assembled arrays out (as double) for unit checks.
//   what: 0 minv [nglob*2], 1 u, 2 v, 3 a, 4 ibool [nspec*25] (as double),
//         5 jacobian [nspec*25], 6 xix, 7 xiz, 8 gammax, 9 gammaz,
//         10 gll xi [5], 11 gll w [5], 12 hprime [25], 13 stf [nstep*nsrc*2],
//         14 source array [nsrc*2*25], 15 receiver h [nrec*25],
//         16 source ispec, 17 receiver ispec, 18 receiver xi, 19 receiver gamma,
//         20 stacey point flag [nspec*25], 21 stacey n0, 22 stacey n1, 23 stacey w
long oracle2d_get(void *hv, int what, double *out) {
  auto *h = static_cast<OracleHandle *>(hv);
  auto get = [&](auto *o) -> long {
    std::vector<double> tmp;
    auto put = [&](const auto &vec) {
      for (auto x : vec)
        tmp.push_back(static_cast<double>(x));
    };
    switch (what) {
    case 0: put(o->minv); break;
    case 1: put(o->u); break;
    case 2: put(o->v); break;
    case 3: put(o->a); break;
    case 4: put(o->ibool); break;
    case 5: for (auto &j : o->jac) tmp.push_back(j.jacobian); break;
    case 6: for (auto &j : o->jac) tmp.push_back(j.xix); break;
    case 7: for (auto &j : o->jac) tmp.push_back(j.xiz); break;
    case 8: for (auto &j : o->jac) tmp.push_back(j.gammax); break;
    case 9: for (auto &j : o->jac) tmp.push_back(j.gammaz); break;
    case 10: for (int i = 0; i < NGLL; ++i) tmp.push_back(o->gll.xi[i]); break;
    case 11: for (int i = 0; i < NGLL; ++i) tmp.push_back(o->gll.w[i]); break;
    case 12:
      for (int i = 0; i < NGLL; ++i)
        for (int j = 0; j < NGLL; ++j)
          tmp.push_back(o->gll.hprime[i][j]);
      break;
    case 13: put(o->stf); break;
    case 14: put(o->src_array); break;
    case 15: put(o->rec_h); break;
    case 16: put(o->src_ispec); break;
    case 17: put(o->rec_ispec); break;
    case 18: put(o->rec_xi); break;
    case 19: put(o->rec_gamma); break;
    case 20: put(o->pt_stacey); break;
    case 21: put(o->pt_n0); break;
    case 22: put(o->pt_n1); break;
    case 23: put(o->pt_w); break;
    default: return -1;
    }
    if (out)
      std::memcpy(out, tmp.data(), tmp.size() * sizeof(double));
    return static_cast<long>(tmp.size());
  };
  return h->f ? get(h->f) : get(h->d);
}

// Overwrites a field (1 u, 2 v, 3 a) from double values [nglob*2].
int oracle2d_set_field(void *hv, int what, const double *in) {
  auto *h = static_cast<OracleHandle *>(hv);
  auto set = [&](auto *o) {
    // 13: the source time function table [nstep][nsrc][2] (checks of the
    // table's own rounding)
    auto &F = what == 13 ? o->stf : (what == 1 ? o->u : (what == 2 ? o->v : o->a));
    for (size_t k = 0; k < F.size(); ++k)
      F[k] = in[k];
    return 0;
  };
  return h->f ? set(h->f) : set(h->d);
}

// Forward run from rest storing every step's state for the adjoint;
// displacement seismograms into seis[nsteps][nrec][2].
int oracle2d_forward_store(void *hv, int nsteps, double *seis) {
  auto *h = static_cast<OracleHandle *>(hv);
  auto run = [&](auto *o) {
    using real = std::remove_reference_t<decltype(o->deltat)>;
    std::vector<real> buf(size_t(nsteps) * o->nrec * 2);
    o->forward_store(nsteps, buf.data());
    for (size_t k = 0; k < buf.size(); ++k)
      seis[k] = buf[k];
    return 0;
  };
  try {
    return h->f ? run(h->f) : run(h->d);
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Discrete-adjoint gradient of a seismogram functional with
// dchi/ds = gsrc[nsteps][nrec][2] (after oracle2d_forward_store); grad is
// [3][nspec*25] for (kappa, mu, rho).
int oracle2d_adjoint(void *hv, int nsteps, const double *gsrc, double *grad) {
  auto *h = static_cast<OracleHandle *>(hv);
  try {
    if (h->f)
      h->f->adjoint(nsteps, gsrc, grad);
    else
      h->d->adjoint(nsteps, gsrc, grad);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// The reference's kernel formula on the adjoint's fields (see
// Oracle2D::adjoint): refk [6][nspec*25] = quirky (kappa, mu, rho) and
// corrected (kappa, mu, rho) kernels; grad as oracle2d_adjoint.
int oracle2d_ref_kernels(void *hv, int nsteps, const double *gsrc,
                         double *grad, double *refk) {
  auto *h = static_cast<OracleHandle *>(hv);
  try {
    const size_t nloc = size_t(h->f ? h->f->nspec : h->d->nspec) * 25;
    std::fill(refk, refk + 6 * nloc, 0.0);
    if (h->f)
      h->f->adjoint(nsteps, gsrc, grad, refk);
    else
      h->d->adjoint(nsteps, gsrc, grad, refk);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// One stiffness evaluation: a <- -K u (+ Stacey with v) on the current
// fields, without sources, mass division or time stepping.  Used to pin the
// operator against closed forms.
int oracle2d_apply_stiffness(void *hv, int with_stacey) {
  auto *h = static_cast<OracleHandle *>(hv);
  auto app = [&](auto *o) {
    std::fill(o->a.begin(), o->a.end(), 0);
    o->stiffness(o->u, o->v, o->a, with_stacey != 0);
    return 0;
  };
  return h->f ? app(h->f) : app(h->d);
}

} // extern "C"
