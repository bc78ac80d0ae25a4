// Host-side assembly of the 2D elastic P-SV problem (see assembly2d.hpp).
//
// All floating-point work here is FP32 with the promotions the reference's
// host code uses; this translation unit is compiled with -ffp-contract=off so
// that no multiply-add is fused (the reference's x86-64 host build has none).
#include "assembly2d.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <map>
#include <unordered_map>

namespace smx {

namespace {

// Legendre value and derivative by recurrence
// (src/quadrature/gll/gll_library.cpp:11-33, 53-93).
float legendre(float z, int n) {
  float p1 = 1.0, p2 = z, p3 = p2;
  for (int k = 1; k < n; k++) {
    const float fk = static_cast<float>(k);
    p3 = ((2.0 * fk + 1.0) * z * p2 - fk * p1) / (fk + 1.0);
    p1 = p2;
    p2 = p3;
  }
  return p3;
}

float legendre_deriv(float z, int n) {
  float p1 = 1.0, p2 = z, p1d = 0.0, p2d = 1.0, p3d = 1.0;
  for (int k = 1; k < n; k++) {
    const float fk = static_cast<float>(k);
    const float p3 = ((2.0 * fk + 1.0) * z * p2 - fk * p1) / (fk + 1.0);
    p3d = ((2.0 * fk + 1.0) * p2 + (2.0 * fk + 1.0) * z * p2d - fk * p1d) /
          (fk + 1.0);
    p1 = p2;
    p2 = p3;
    p1d = p2d;
    p2d = p3d;
  }
  return p3d;
}

// Lagrange interpolants and their derivatives through the GLL nodes at xi
// (src/quadrature/gll/lagrange_poly.cpp:11-43: h = prod / den, h' = sum of
// the partial products / den).
void interpolants_d(float xi, const float *nodes, float *h, float *hp) {
  for (int d = 0; d < NGLL; d++) {
    float num = 1.0, den = 1.0, sum = 0.0;
    for (int i = 0; i < NGLL; i++) {
      if (i == d)
        continue;
      num = num * (xi - nodes[i]);
      den = den * (nodes[d] - nodes[i]);
      float part = 1.0;
      for (int j = 0; j < NGLL; j++)
        if (j != d && j != i)
          part = part * (xi - nodes[j]);
      sum = sum + part;
    }
    const float inv = 1.0 / den;
    h[d] = num * inv;
    hp[d] = sum * inv;
  }
}

// Lagrange interpolants through the GLL nodes at xi
// (src/quadrature/gll/lagrange_poly.cpp:11-43).
void interpolants(float xi, const float *nodes, float *h) {
  for (int d = 0; d < NGLL; d++) {
    float num = 1.0, den = 1.0;
    for (int i = 0; i < NGLL; i++) {
      if (i == d)
        continue;
      num = num * (xi - nodes[i]);
      den = den * (nodes[d] - nodes[i]);
    }
    const float inv = 1.0 / den;
    h[d] = num * inv;
  }
}

// Shape functions and derivatives of 4/9-node quads
// (core/specfem/shape_function/dim2/shape_functions2d.cpp:7-32,
//  shape_function_derivatives2d.cpp:7-75).  Each value is formed in double
// from float inputs and rounded to float, as the reference's std::vector<float>
// results are.
void shape(float xi, float ga, int ngnod, float *N) {
  if (ngnod == 4) {
    N[0] = 0.25 * (xi - 1) * (ga - 1);
    N[1] = -0.25 * (xi + 1) * (ga - 1);
    N[2] = 0.25 * (xi + 1) * (ga + 1);
    N[3] = -0.25 * (xi - 1) * (ga + 1);
    return;
  }
  N[0] = 0.25 * (xi - 1) * xi * (ga - 1) * ga;
  N[1] = 0.25 * (xi + 1) * xi * (ga - 1) * ga;
  N[2] = 0.25 * (xi + 1) * xi * (ga + 1) * ga;
  N[3] = 0.25 * (xi - 1) * xi * (ga + 1) * ga;
  N[4] = 0.5 * (1 - xi * xi) * (ga - 1) * ga;
  N[5] = 0.5 * (xi + 1) * xi * (1 - ga * ga);
  N[6] = 0.5 * (1 - xi * xi) * (ga + 1) * ga;
  N[7] = 0.5 * (xi - 1) * xi * (1 - ga * ga);
  N[8] = (1 - xi * xi) * (1 - ga * ga);
}

void shape_deriv(float xi, float ga, int ngnod, float *dxi, float *dga) {
  if (ngnod == 4) {
    const float sp = xi + 1, sm = xi - 1, tp = ga + 1, tm = ga - 1;
    dxi[0] = 0.25 * tm;
    dxi[1] = -0.25 * tm;
    dxi[2] = 0.25 * tp;
    dxi[3] = -0.25 * tp;
    dga[0] = 0.25 * sm;
    dga[1] = -0.25 * sp;
    dga[2] = 0.25 * sp;
    dga[3] = -0.25 * sm;
    return;
  }
  const float sp = xi + 1.0, sm = xi - 1.0, tp = ga + 1.0, tm = ga - 1.0;
  const float s2 = xi * 2.0, t2 = ga * 2.0, ss = xi * xi, tt = ga * ga,
              st = xi * ga;
  dxi[0] = 0.25 * tm * ga * (s2 - 1.0);
  dxi[1] = 0.25 * tm * ga * (s2 + 1.0);
  dxi[2] = 0.25 * tp * ga * (s2 + 1.0);
  dxi[3] = 0.25 * tp * ga * (s2 - 1.0);
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

struct Geo {
  float xix, xiz, gammax, gammaz, jac;
};

// core/specfem/jacobian/dim2/jacobian.cpp:25-51
Geo jacobian_at(const float *cx, const float *cz, int ngnod, const float *dxi,
                const float *dga) {
  float xxi = 0.0, zxi = 0.0, xga = 0.0, zga = 0.0;
  for (int n = 0; n < ngnod; n++) {
    xxi += dxi[n] * cx[n];
    zxi += dxi[n] * cz[n];
    xga += dga[n] * cx[n];
    zga += dga[n] * cz[n];
  }
  Geo g;
  g.jac = xxi * zga - xga * zxi;
  g.xix = zga / g.jac;
  g.gammax = -zxi / g.jac;
  g.xiz = -xga / g.jac;
  g.gammaz = xxi / g.jac;
  return g;
}

void location_at(const float *cx, const float *cz, int ngnod, float xi,
                 float ga, float &x, float &z) {
  float N[9];
  shape(xi, ga, ngnod, N);
  x = 0.0;
  z = 0.0;
  for (int n = 0; n < ngnod; n++) {
    x += N[n] * cx[n];
    z += N[n] * cz[n];
  }
}

// Nearest-GLL-point search over a bucket grid.  Reproduces the reference's
// brute-force scan (src/algorithms/dim2/locate_point_impl.cpp:13-45): the
// smallest distance wins, ties go to the earliest (element, iz, ix) in scan
// order.
class PointGrid {
public:
  PointGrid(const std::vector<float> &px, const std::vector<float> &pz)
      : px_(px), pz_(pz) {
    const size_t n = px.size();
    xmin_ = zmin_ = std::numeric_limits<float>::max();
    float xmax = -xmin_, zmax = -zmin_;
    for (size_t i = 0; i < n; ++i) {
      xmin_ = std::min(xmin_, px[i]);
      xmax = std::max(xmax, px[i]);
      zmin_ = std::min(zmin_, pz[i]);
      zmax = std::max(zmax, pz[i]);
    }
    const double area =
        std::max(double(xmax - xmin_), 1e-6) * std::max(double(zmax - zmin_), 1e-6);
    cell_ = static_cast<float>(std::sqrt(area / std::max<double>(n / 4.0, 1.0)));
    if (!(cell_ > 0))
      cell_ = 1.0f;
    nxc_ = std::max(1, int((xmax - xmin_) / cell_) + 1);
    nzc_ = std::max(1, int((zmax - zmin_) / cell_) + 1);
    start_.assign(size_t(nxc_) * nzc_ + 1, 0);
    std::vector<int> cellof(n);
    for (size_t i = 0; i < n; ++i) {
      cellof[i] = cell_index(px[i], pz[i]);
      start_[cellof[i] + 1]++;
    }
    for (size_t c = 0; c < size_t(nxc_) * nzc_; ++c)
      start_[c + 1] += start_[c];
    items_.resize(n);
    std::vector<int> fill(start_.begin(), start_.end() - 1);
    for (size_t i = 0; i < n; ++i)
      items_[fill[cellof[i]]++] = static_cast<int>(i);
  }

  // Returns the scan-order index of the nearest point.
  long nearest(float x, float z) const {
    const int cx = clampi(int((x - xmin_) / cell_), nxc_),
              cz = clampi(int((z - zmin_) / cell_), nzc_);
    float best = std::numeric_limits<float>::max();
    long best_i = -1;
    const int maxr = std::max(nxc_, nzc_);
    for (int r = 0; r <= maxr; ++r) {
      for (int jz = cz - r; jz <= cz + r; ++jz) {
        if (jz < 0 || jz >= nzc_)
          continue;
        for (int jx = cx - r; jx <= cx + r; ++jx) {
          if (jx < 0 || jx >= nxc_)
            continue;
          if (std::max(std::abs(jx - cx), std::abs(jz - cz)) != r)
            continue;
          const int c = jz * nxc_ + jx;
          for (int k = start_[c]; k < start_[c + 1]; ++k) {
            const int i = items_[k];
            const float dx = x - px_[i], dz = z - pz_[i];
            const float d = std::sqrt(dx * dx + dz * dz);
            if (d < best || (d == best && i < best_i)) {
              best = d;
              best_i = i;
            }
          }
        }
      }
      // Every point outside rings 0..r is farther than (r) cells from the
      // query's cell boundary; stop once that bound exceeds the best distance
      // with a safety margin.
      if (best_i >= 0 && double(r) * cell_ > double(best) * 1.0001 + 1e-6)
        break;
    }
    return best_i;
  }

private:
  int cell_index(float x, float z) const {
    return clampi(int((z - zmin_) / cell_), nzc_) * nxc_ +
           clampi(int((x - xmin_) / cell_), nxc_);
  }
  static int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }
  const std::vector<float> &px_, &pz_;
  float xmin_, zmin_, cell_;
  int nxc_, nzc_;
  std::vector<int> start_, items_;
};

} // namespace

Quadrature::Quadrature() {
  // Lobatto nodes and weights of degree 4 (closed forms of the roots
  // src/quadrature/gll/gll_library.cpp:102-147 computes).
  const double s = std::sqrt(3.0 / 7.0);
  const double xd[NGLL] = {-1.0, -s, 0.0, s, 1.0};
  const double wd[NGLL] = {0.1, 49.0 / 90.0, 32.0 / 45.0, 49.0 / 90.0, 0.1};
  for (int i = 0; i < NGLL; ++i) {
    xi[i] = static_cast<float>(xd[i]);
    w[i] = static_cast<float>(wd[i]);
  }
  // h'(i,j) = l'_j(xi_i) (src/quadrature/gll/lagrange_poly.cpp:161-190).
  const int n = NGLL - 1;
  const float fn = static_cast<float>(n);
  for (int i = 0; i < NGLL; i++)
    for (int j = 0; j < NGLL; j++) {
      if (i == 0 && j == 0)
        hprime[i][j] = -1.0 * fn * (fn + 1.0) * 0.25;
      else if (i == n && j == n)
        hprime[i][j] = 1.0 * fn * (fn + 1.0) * 0.25;
      else if (i == j)
        hprime[i][j] = 0.0;
      else
        hprime[i][j] = legendre(xi[i], n) /
                           (legendre(xi[j], n) * (xi[i] - xi[j])) +
                       (1.0 - xi[i] * xi[i]) * legendre_deriv(xi[i], n) /
                           (fn * (fn + 1.0) * legendre(xi[j], n) *
                            (xi[i] - xi[j]) * (xi[i] - xi[j]));
    }
}

void refresh_boundary_matrices(HostAssembly2D &A);
static void build_shared_csr(HostAssembly2D &A);
static void finish_assembly(HostAssembly2D &A);

float source_start_time(int stf_kind, float f0, float tshift, double dt) {
  const float dtf = static_cast<float>(dt);
  if (stf_kind == SMX_STF_HEAVISIDE) {
    float hd = f0;
    if (!(hd > 0))
      throw SmxError(SMX_ERR_ARG, "Heaviside half duration must be positive");
    if (hd < 5.0 * dtf)
      hd = 5.0 * dtf;
    return -static_cast<float>(2.0) * hd + tshift;
  }
  if (stf_kind != SMX_STF_RICKER && stf_kind != SMX_STF_DIRAC)
    throw SmxError(SMX_ERR_ARG, "unknown source time function");
  const float fv = stf_kind == SMX_STF_DIRAC ? static_cast<float>(1.0 / (10.0 * dtf)) : f0;
  if (!(fv > 0))
    throw SmxError(SMX_ERR_ARG, "source f0 must be positive");
  const float hdur = 1.0 / fv;
  return -static_cast<float>(1.2) * hdur + tshift;
}

void assemble2d(const smx_mesh2d &m, const smx_sources2d &src,
                const smx_receivers2d &rec, double dt_in, int nstep,
                HostAssembly2D &A, const AssemblyOptions &opt) {
  const int tile_size = opt.tile_elems;
  if (tile_size < 1 || tile_size > ELEM_TILE)
    throw SmxError(SMX_ERR_ARG, "tile_elems must lie in [1, 64]");
  if (m.nspec <= 0 || (m.ngnod != 4 && m.ngnod != 9) || m.nnodes <= 0 ||
      !m.coord || !m.knods || !m.kappa || !m.mu || !m.rho)
    throw SmxError(SMX_ERR_ARG, "invalid mesh description");
  if (m.nabs < 0 || (m.nabs > 0 && (!m.abs_ispec || !m.abs_type)))
    throw SmxError(SMX_ERR_ARG, "invalid absorbing edge list");
  if (nstep <= 0 || !(dt_in > 0))
    throw SmxError(SMX_ERR_ARG, "dt and nstep must be positive");
  if (src.nsrc < 0 || rec.nrec < 0 || rec.ntypes < 0 ||
      rec.nstep_between_samples <= 0)
    throw SmxError(SMX_ERR_ARG, "invalid source/receiver description");

  const int nspec = m.nspec, ngnod = m.ngnod;
  A.nspec = nspec;
  const Quadrature &q = A.quad;

  // ---- control nodes (type_real coordinates)
  std::vector<float> cx(size_t(nspec) * ngnod), cz(size_t(nspec) * ngnod);
  for (int e = 0; e < nspec; ++e)
    for (int n = 0; n < ngnod; ++n) {
      const int id = m.knods[size_t(e) * ngnod + n];
      if (id < 0 || id >= m.nnodes)
        throw SmxError(SMX_ERR_ARG, "knods out of range");
      cx[size_t(e) * ngnod + n] = static_cast<float>(m.coord[2 * size_t(id)]);
      cz[size_t(e) * ngnod + n] = static_cast<float>(m.coord[2 * size_t(id) + 1]);
    }

  // ---- element colouring: two elements of one colour share no GLL point.
  // Conforming quads share points only through their corner nodes, so a
  // greedy pass over the corner control nodes (per-node mask of the colours
  // already touching it) is sufficient.  The colours order the deterministic
  // LDS assembly inside a tile and the mass-matrix scatter.
  std::vector<int> ecolor_mesh(nspec);
  {
    std::vector<uint32_t> node_mask(m.nnodes, 0u);
    int ncolors = 0;
    for (int e = 0; e < nspec; ++e) {
      const int *kn = &m.knods[size_t(e) * ngnod];
      uint32_t used = 0;
      for (int c = 0; c < 4; ++c)
        used |= node_mask[kn[c]];
      int col = 0;
      while (col < 32 && ((used >> col) & 1u))
        ++col;
      if (col >= 32)
        throw SmxError(SMX_ERR_MESH, "element colouring needs more than 32 colours");
      ecolor_mesh[e] = col;
      ncolors = std::max(ncolors, col + 1);
      for (int c = 0; c < 4; ++c)
        node_mask[kn[c]] |= 1u << col;
    }
    A.nelem_colors = ncolors;
  }

  // ---- provisional numbering from corner topology (first touch in element
  // order); renumbered below once the tiles are known
  std::vector<int> ib0(size_t(nspec) * NPT);
  int nglob0 = 0;
  {
    std::unordered_map<int, int> corner_ids;
    std::unordered_map<uint64_t, int> edge_base; // sorted corner pair -> first id
    corner_ids.reserve(size_t(nspec) * 2);
    edge_base.reserve(size_t(nspec) * 3);
    int next = 0;
    auto corner = [&](int node) {
      auto it = corner_ids.emplace(node, next);
      if (it.second)
        ++next;
      return it.first->second;
    };
    // an edge owns NGLL-2 consecutive ids, ordered from its lower-id corner
    auto edge = [&](int a, int b, int k) {
      const uint64_t key = (uint64_t(uint32_t(std::min(a, b))) << 32) |
                           uint32_t(std::max(a, b));
      auto it = edge_base.emplace(key, next);
      if (it.second)
        next += NGLL - 2;
      const int pos = (a < b) ? k : (NGLL - 1 - k);
      return it.first->second + pos - 1;
    };
    for (int e = 0; e < nspec; ++e) {
      const int *kn = &m.knods[size_t(e) * ngnod];
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const bool b = iz == 0, t = iz == NGLL - 1, l = ix == 0,
                     r = ix == NGLL - 1;
          int g;
          if (b && l) g = corner(kn[0]);
          else if (b && r) g = corner(kn[1]);
          else if (t && r) g = corner(kn[2]);
          else if (t && l) g = corner(kn[3]);
          else if (b) g = edge(kn[0], kn[1], ix);
          else if (r) g = edge(kn[1], kn[2], iz);
          else if (t) g = edge(kn[3], kn[2], ix);
          else if (l) g = edge(kn[0], kn[3], iz);
          else g = next++;
          ib0[size_t(e) * NPT + iz * NGLL + ix] = g;
        }
    }
    nglob0 = next;
  }

  // ---- tiles: elements sorted along a Morton (Z-order) curve of their
  // centroids on a grid of the mean element size, cut into groups of
  // tile_size (<= ELEM_TILE).  On a structured box 64 yields 8x8-element
  // squares, 32 yields 8x4 and 16 yields 4x4.
  std::vector<std::vector<int>> tile_elems;
  {
    std::vector<float> ccx(nspec), ccz(nspec);
    float xmin = std::numeric_limits<float>::max(), zmin = xmin;
    float xmax = -xmin, zmax = -zmin;
    for (int e = 0; e < nspec; ++e) {
      float sx = 0, sz = 0;
      for (int c = 0; c < 4; ++c) {
        sx += cx[size_t(e) * ngnod + c];
        sz += cz[size_t(e) * ngnod + c];
      }
      ccx[e] = 0.25f * sx;
      ccz[e] = 0.25f * sz;
      xmin = std::min(xmin, ccx[e]);
      xmax = std::max(xmax, ccx[e]);
      zmin = std::min(zmin, ccz[e]);
      zmax = std::max(zmax, ccz[e]);
    }
    // mean element size from the total element area
    double area = 0.0;
    for (int e = 0; e < nspec; ++e) {
      const float *x = &cx[size_t(e) * ngnod], *z = &cz[size_t(e) * ngnod];
      area += 0.5 * std::fabs(double(x[0]) * z[1] - double(x[1]) * z[0] +
                              double(x[1]) * z[2] - double(x[2]) * z[1] +
                              double(x[2]) * z[3] - double(x[3]) * z[2] +
                              double(x[3]) * z[0] - double(x[0]) * z[3]);
    }
    const double h = std::sqrt(area / nspec);
    auto spread = [](uint32_t v) {
      uint64_t x = v & 0xffffffffu;
      x = (x | (x << 16)) & 0x0000ffff0000ffffull;
      x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
      x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
      x = (x | (x << 2)) & 0x3333333333333333ull;
      x = (x | (x << 1)) & 0x5555555555555555ull;
      return x;
    };
    std::vector<std::pair<uint64_t, int>> key(nspec);
    for (int e = 0; e < nspec; ++e) {
      const uint32_t gx = uint32_t(std::max(0.0, std::floor((ccx[e] - xmin) / h + 0.5)));
      const uint32_t gz = uint32_t(std::max(0.0, std::floor((ccz[e] - zmin) / h + 0.5)));
      key[e] = {spread(gx) | (spread(gz) << 1), e};
    }
    std::sort(key.begin(), key.end());
    for (int i = 0; i < nspec; i += tile_size) {
      std::vector<int> t;
      for (int k = i; k < std::min(nspec, i + tile_size); ++k)
        t.push_back(key[k].second);
      tile_elems.push_back(std::move(t));
    }
  }
  const int ntiles = static_cast<int>(tile_elems.size());

  // ---- node ownership: a node touched by one tile only is exclusive to it.
  // Final numbering: each tile's exclusive nodes contiguously, tile after
  // tile (Morton order), then every shared node.
  std::vector<int> new_id(nglob0, -1);
  std::vector<std::vector<int>> tile_shared(ntiles); // provisional ids
  {
    std::vector<int> touch(nglob0, 0), last(nglob0, -1);
    for (int t = 0; t < ntiles; ++t)
      for (int e : tile_elems[t])
        for (int p = 0; p < NPT; ++p) {
          const int g = ib0[size_t(e) * NPT + p];
          if (last[g] != t) {
            last[g] = t;
            touch[g]++;
          }
        }
    int next = 0;
    std::vector<int> nex(ntiles, 0);
    A.tiles.assign(ntiles, TileInfo{});
    for (int t = 0; t < ntiles; ++t) {
      A.tiles[t].ex_off = next;
      for (int e : tile_elems[t])
        for (int p = 0; p < NPT; ++p) {
          const int g = ib0[size_t(e) * NPT + p];
          if (touch[g] == 1 && new_id[g] < 0)
            new_id[g] = next++;
        }
      A.tiles[t].nex = next - A.tiles[t].ex_off;
    }
    A.nex_total = next;
    std::fill(last.begin(), last.end(), -1);
    for (int t = 0; t < ntiles; ++t)
      for (int e : tile_elems[t])
        for (int p = 0; p < NPT; ++p) {
          const int g = ib0[size_t(e) * NPT + p];
          if (touch[g] > 1) {
            if (new_id[g] < 0)
              new_id[g] = next++;
            if (last[g] != t) {
              last[g] = t;
              tile_shared[t].push_back(g);
            }
          }
        }
    A.nglob = next;
    A.nshared = next - A.nex_total;
    if (A.nglob != nglob0)
      throw SmxError(SMX_ERR_MESH, "numbering lost nodes");
  }

  // ---- launch order: Morton order (one launch covers every tile; points
  // shared between tiles are summed afterwards from per-tile partials)
  std::vector<int> tile_order(ntiles);
  for (int t = 0; t < ntiles; ++t)
    tile_order[t] = t;

  // ---- slots: launch-order tile index * ELEM_TILE + position in the tile
  A.ntiles = ntiles;
  A.nspec_pad = ntiles * ELEM_TILE;
  const int ld = A.nspec_pad;
  A.elem_slot.assign(nspec, 0);
  A.ibool.assign(size_t(NPT) * ld, 0);
  A.lidx.assign(size_t(NPT) * ld, 0);
  A.ecolor.assign(ld, 0xff);
  {
    std::vector<TileInfo> ordered(ntiles);
    std::vector<int> local(A.nglob, -1);
    for (int k = 0; k < ntiles; ++k) {
      const int t = tile_order[k];
      TileInfo info = A.tiles[t];
      info.ecount = static_cast<int>(tile_elems[t].size());
      info.sh_off = static_cast<int>(A.sh_list.size());
      info.nsh = static_cast<int>(tile_shared[t].size());
      for (int i = 0; i < info.nsh; ++i) {
        const int g = new_id[tile_shared[t][i]];
        A.sh_list.push_back(g);
        local[g] = info.nex + i;
      }
      if (info.nex + info.nsh > NLOC_MAX)
        throw SmxError(SMX_ERR_MESH, "tile has too many GLL points");
      for (int j = 0; j < info.ecount; ++j) {
        const int e = tile_elems[t][j];
        const int s = k * ELEM_TILE + j;
        A.elem_slot[e] = s;
        A.ecolor[s] = static_cast<uint8_t>(ecolor_mesh[e]);
        for (int p = 0; p < NPT; ++p) {
          const int g = new_id[ib0[size_t(e) * NPT + p]];
          A.ibool[size_t(p) * ld + s] = g;
          const int li = g < A.nex_total ? g - info.ex_off : local[g];
          if (li < 0 || li >= info.nex + info.nsh)
            throw SmxError(SMX_ERR_MESH, "tile-local numbering failed");
          A.lidx[size_t(p) * ld + s] = static_cast<uint16_t>(li);
        }
      }
      for (int i = 0; i < info.nsh; ++i)
        local[A.sh_list[info.sh_off + i]] = -1;
      ordered[k] = info;
    }
    A.tiles = std::move(ordered);
  }
  build_shared_csr(A);
  const std::vector<int> &slot_of = A.elem_slot;

  // ---- geometry (core/specfem/assembly/jacobian_matrix/dim2/jacobian_matrix.cpp:28-93)
  float dN[NPT][2][9];
  for (int iz = 0; iz < NGLL; ++iz)
    for (int ix = 0; ix < NGLL; ++ix)
      shape_deriv(q.xi[ix], q.xi[iz], ngnod, dN[iz * NGLL + ix][0],
                  dN[iz * NGLL + ix][1]);
  A.geom.assign(size_t(5) * NPT * ld, 0.0f);
  std::vector<float> px(size_t(nspec) * NPT), pz(size_t(nspec) * NPT);
  for (int e = 0; e < nspec; ++e)
    for (int p = 0; p < NPT; ++p) {
      const Geo g = jacobian_at(&cx[size_t(e) * ngnod], &cz[size_t(e) * ngnod],
                                ngnod, dN[p][0], dN[p][1]);
      if (!(g.jac > 1e-10f))
        throw SmxError(SMX_ERR_MESH, "non-positive Jacobian in element " +
                                         std::to_string(e));
      const size_t o = size_t(p) * ld + slot_of[e];
      A.geom[0 * size_t(NPT) * ld + o] = g.xix;
      A.geom[1 * size_t(NPT) * ld + o] = g.xiz;
      A.geom[2 * size_t(NPT) * ld + o] = g.gammax;
      A.geom[3 * size_t(NPT) * ld + o] = g.gammaz;
      A.geom[4 * size_t(NPT) * ld + o] = g.jac;
      location_at(&cx[size_t(e) * ngnod], &cz[size_t(e) * ngnod], ngnod,
                  q.xi[p % NGLL], q.xi[p / NGLL], px[size_t(e) * NPT + p],
                  pz[size_t(e) * NPT + p]);
    }

  // ---- element-centre geometry of affine elements (the five terms agree at
  // every GLL point up to FP32 rounding: relative 4e-7 of the element's
  // largest term)
  A.geom_e.assign(size_t(5) * ld, 0.0f);
  A.elem_affine.assign(ld, 0);
  if (ngnod == 4) {
    A.nodes.assign(size_t(8) * ld, 0.0f);
    for (int e = 0; e < nspec; ++e)
      for (int n = 0; n < 4; ++n) {
        A.nodes[size_t(n) * ld + slot_of[e]] = cx[size_t(e) * 4 + n];
        A.nodes[size_t(4 + n) * ld + slot_of[e]] = cz[size_t(e) * 4 + n];
      }
  }
  {
    float c0[9], c1[9];
    shape_deriv(0.0f, 0.0f, ngnod, c0, c1);
    for (int e = 0; e < nspec; ++e) {
      const int sl = slot_of[e];
      const Geo gc = jacobian_at(&cx[size_t(e) * ngnod], &cz[size_t(e) * ngnod],
                                 ngnod, c0, c1);
      const float cv[5] = {gc.xix, gc.xiz, gc.gammax, gc.gammaz, gc.jac};
      float scale_d = std::max(std::max(std::fabs(cv[0]), std::fabs(cv[1])),
                               std::max(std::fabs(cv[2]), std::fabs(cv[3])));
      bool ok = true;
      for (int p = 0; p < NPT && ok; ++p)
        for (int c = 0; c < 5 && ok; ++c) {
          const float v = A.geom[(c * size_t(NPT) + p) * ld + sl];
          const float scale = c < 4 ? scale_d : std::fabs(cv[4]);
          ok = std::fabs(v - cv[c]) <= 4e-7f * scale;
        }
      A.elem_affine[sl] = ok ? 1 : 0;
      for (int c = 0; c < 5; ++c)
        A.geom_e[c * size_t(ld) + sl] = cv[c];
    }
  }

  // ---- material (include/medium/dim2/elastic/isotropic/properties.hpp:71-91)
  A.lam2mu_mu.assign(size_t(2) * NPT * ld, 0.0f);
  A.kappa_mu_rho.assign(size_t(3) * NPT * ld, 0.0f);
  auto l2m_of = [](float kappa, float mu) {
    return kappa + (static_cast<float>(4.0 / 3.0)) * mu;
  };
  // Medium: an element is acoustic when mu == 0 at all its points
  // (src/io/mesh/impl/fortran/dim2/read_material_properties.cpp:178-204).
  {
    int nac = 0;
    for (int e = 0; e < nspec; ++e) {
      bool fluid = true;
      for (int p = 0; p < NPT; ++p)
        fluid = fluid && m.mu[size_t(e) * NPT + p] == 0.0f;
      nac += fluid;
    }
    if (nac != 0 && nac != nspec)
      throw SmxError(SMX_ERR_MESH,
                     "mixed acoustic / elastic meshes need the elastic-acoustic "
                     "coupling, which is not built");
    A.medium = nac == nspec ? 1 : 0;
  }
  const bool acoustic = A.medium == 1;
  for (int e = 0; e < nspec; ++e)
    for (int p = 0; p < NPT; ++p) {
      const size_t src_i = size_t(e) * NPT + p, o = size_t(p) * ld + slot_of[e];
      const float kappa = m.kappa[src_i], mu = m.mu[src_i], rho = m.rho[src_i];
      if (!(rho > 0) || !(mu >= 0) || !(kappa > 0))
        throw SmxError(SMX_ERR_ARG, "non-physical material at element " +
                                        std::to_string(e));
      if (acoustic) {
        // include/medium/dim2/acoustic/isotropic/material.hpp:88-91
        A.lam2mu_mu[o] = static_cast<float>(1.0) / rho;
        A.lam2mu_mu[size_t(NPT) * ld + o] = kappa;
      } else {
        A.lam2mu_mu[o] = l2m_of(kappa, mu);
        A.lam2mu_mu[size_t(NPT) * ld + o] = mu;
      }
      A.kappa_mu_rho[o] = kappa;
      A.kappa_mu_rho[size_t(NPT) * ld + o] = mu;
      A.kappa_mu_rho[2 * size_t(NPT) * ld + o] = rho;
    }

  // ---- acoustic free surface (core/specfem/assembly/boundaries/dim2/impl/
  // acoustic_free_surface.tpp:8-96; is_on_boundary, utilities.hpp:6-20)
  if (m.nfree < 0 || (m.nfree > 0 && (!m.free_ispec || !m.free_type)))
    throw SmxError(SMX_ERR_ARG, "invalid free-surface list");
  if (m.nfree > 0 && !acoustic)
    throw SmxError(SMX_ERR_MESH,
                   "acoustic free-surface entities on an elastic mesh");
  if (acoustic) {
    A.dmask.assign(ld, 0u);
    for (int i = 0; i < m.nfree; ++i) {
      const int e = m.free_ispec[i], t = m.free_type[i];
      if (e < 0 || e >= nspec || t < 1 || t > 8)
        throw SmxError(SMX_ERR_ARG, "invalid free-surface entity");
      uint32_t mask = 0;
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const bool bz = iz == 0, tz = iz == NGLL - 1, lx = ix == 0,
                     rx = ix == NGLL - 1;
          const bool on = (t == 1 && bz) || (t == 2 && rx) || (t == 3 && tz) ||
                          (t == 4 && lx) || (t == 5 && bz && lx) ||
                          (t == 6 && bz && rx) || (t == 7 && tz && rx) ||
                          (t == 8 && tz && lx);
          if (on)
            mask |= 1u << (iz * NGLL + ix);
        }
      A.dmask[slot_of[e]] |= mask;
    }
  }

  // ---- Stacey (core/specfem/assembly/boundaries/dim2/impl/stacey.tpp:15-176):
  // elements in ascending order, each element's edges in list order, the last
  // edge covering a point sets its normal and weight.
  std::map<int, std::vector<int>> edges_of;
  for (int i = 0; i < m.nabs; ++i) {
    const int e = m.abs_ispec[i], t = m.abs_type[i];
    if (e < 0 || e >= nspec || t < 1 || t > 4)
      throw SmxError(SMX_ERR_ARG, "invalid absorbing edge");
    edges_of[e].push_back(t);
  }
  A.stacey_slot.assign(ld, -1);
  A.nstacey = static_cast<int>(edges_of.size());
  A.nstacey_pad = std::max(ELEM_TILE, (A.nstacey + ELEM_TILE - 1) / ELEM_TILE *
                                          ELEM_TILE);
  const int sld = A.nstacey_pad;
  A.stacey.assign(size_t(6) * NPT * sld, 0.0f);
  {
    int slot = 0;
    for (auto &kv : edges_of) {
      const int e = kv.first;
      A.stacey_slot[slot_of[e]] = slot;
      // points not on an absorbing edge keep a unit normal and zero weight,
      // which makes their traction contribution an exact -0.
      for (int p = 0; p < NPT; ++p) {
        A.stacey[(0 * size_t(NPT) + p) * sld + slot] = 1.0f;
        A.stacey[(2 * size_t(NPT) + p) * sld + slot] = 1.0f;
      }
      for (int t : kv.second)
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            const bool on = (t == 3 && iz == NGLL - 1) || (t == 1 && iz == 0) ||
                            (t == 4 && ix == 0) || (t == 2 && ix == NGLL - 1);
            if (!on)
              continue;
            const int p = iz * NGLL + ix;
            const size_t o = size_t(p) * ld + slot_of[e];
            const float xix = A.geom[o], xiz = A.geom[size_t(NPT) * ld + o],
                        gax = A.geom[2 * size_t(NPT) * ld + o],
                        gaz = A.geom[3 * size_t(NPT) * ld + o],
                        jac = A.geom[4 * size_t(NPT) * ld + o];
            float n0, n1, wgt;
            // core/specfem/point/jacobian_matrix.hpp:424-449
            if (t == 4) {
              n0 = static_cast<float>(-1.0) * xix * jac;
              n1 = static_cast<float>(-1.0) * xiz * jac;
              wgt = q.w[iz];
            } else if (t == 2) {
              n0 = xix * jac;
              n1 = xiz * jac;
              wgt = q.w[iz];
            } else if (t == 3) {
              n0 = gax * jac;
              n1 = gaz * jac;
              wgt = q.w[ix];
            } else {
              n0 = static_cast<float>(-1.0) * gax * jac;
              n1 = static_cast<float>(-1.0) * gaz * jac;
              wgt = q.w[ix];
            }
            const float kappa = A.kappa_mu_rho[o],
                        mu = A.kappa_mu_rho[size_t(NPT) * ld + o],
                        rho = A.kappa_mu_rho[2 * size_t(NPT) * ld + o];
            float rvp, rvs;
            if (acoustic) {
              // rho_vpinverse (acoustic/isotropic/properties.hpp:62-70)
              const float rho_inv = static_cast<float>(1.0) / rho;
              const float kappa_inv = static_cast<float>(1.0) / kappa;
              rvp = std::sqrt(rho_inv * kappa_inv);
              rvs = 0.0f;
              (void)mu;
            } else {
              const float l2m = l2m_of(kappa, mu);
              rvp = std::sqrt(l2m / rho) * rho;
              rvs = std::sqrt(rho * mu);
            }
            const float nrm = std::sqrt(n0 * n0 + n1 * n1);
            A.stacey[(0 * size_t(NPT) + p) * sld + slot] = n0;
            A.stacey[(1 * size_t(NPT) + p) * sld + slot] = n1;
            A.stacey[(2 * size_t(NPT) + p) * sld + slot] = nrm;
            A.stacey[(3 * size_t(NPT) + p) * sld + slot] = wgt;
            A.stacey[(4 * size_t(NPT) + p) * sld + slot] = rvp;
            A.stacey[(5 * size_t(NPT) + p) * sld + slot] = rvs;
          }
      ++slot;
    }
  }

  // ---- point location (src/algorithms/dim2/locate_point_impl.cpp:245-345)
  PointGrid grid(px, pz);
  // corner id -> elements touching it, ascending
  std::unordered_map<int, std::vector<int>> elems_at_corner;
  const int corner_p[4] = {0, (NGLL - 1) * NGLL, NGLL - 1, NPT - 1};
  for (int e = 0; e < nspec; ++e)
    for (int c = 0; c < 4; ++c)
      elems_at_corner[A.ibool[size_t(corner_p[c]) * ld + slot_of[e]]].push_back(e);
  auto locate = [&](double xd, double zd, int &ispec_out, float &xi_out,
                    float &ga_out) {
    const float x = static_cast<float>(xd), z = static_cast<float>(zd);
    const long best = grid.nearest(x, z);
    if (best < 0)
      throw SmxError(SMX_ERR_MESH, "empty mesh");
    const int guess = static_cast<int>(best / NPT);
    std::vector<int> cand;
    for (int c = 0; c < 4; ++c)
      for (int e : elems_at_corner[A.ibool[size_t(corner_p[c]) * ld + slot_of[guess]]])
        if (e != guess)
          cand.push_back(e);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    cand.insert(cand.begin(), guess);
    float final_dist = std::numeric_limits<float>::max();
    ispec_out = -1;
    xi_out = ga_out = -9999.0f;
    for (int e : cand) {
      const float *ex = &cx[size_t(e) * ngnod], *ez = &cz[size_t(e) * ngnod];
      float xi = 0.0f, ga = 0.0f;
      for (int it = 0; it < 100; ++it) {
        float lx, lz, dxs[9], dgs[9];
        location_at(ex, ez, ngnod, xi, ga, lx, lz);
        shape_deriv(xi, ga, ngnod, dxs, dgs);
        const Geo g = jacobian_at(ex, ez, ngnod, dxs, dgs);
        const float dx = -(lx - x), dz = -(lz - z);
        const float dxi = g.xix * dx + g.xiz * dz;
        const float dga = g.gammax * dx + g.gammaz * dz;
        xi += dxi;
        ga += dga;
        xi = std::min(1.01f, std::max(-1.01f, xi));
        ga = std::min(1.01f, std::max(-1.01f, ga));
        if (std::abs(dxi) < 1e-12 && std::abs(dga) < 1e-12)
          break;
      }
      float lx, lz;
      location_at(ex, ez, ngnod, xi, ga, lx, lz);
      const float d = std::sqrt((x - lx) * (x - lx) + (z - lz) * (z - lz));
      if (d < final_dist) {
        final_dist = d;
        ispec_out = e;
        xi_out = xi;
        ga_out = ga;
      }
    }
    if (ispec_out < 0 || std::fabs(std::fabs(xi_out) - 1.01f) < 1e-6 ||
        std::fabs(std::fabs(ga_out) - 1.01f) < 1e-6)
      throw SmxError(SMX_ERR_MESH, "point (" + std::to_string(xd) + ", " +
                                       std::to_string(zd) +
                                       ") lies outside the mesh");
  };

  // ---- sources: t0 rule (src/io/sources/dim2/read_sources.cpp:142-172),
  // Ricker table (core/specfem/source_time_functions/ricker.cpp:55-66,
  // impl/time_functions.hpp:77-83), array = h_xi h_gamma f
  // (compute_source_array_from_vector.cpp:14-60, force_source.cpp:22-73).
  const float dt = static_cast<float>(dt_in);
  A.dt = dt;
  A.dt_over2 = dt / 2.0;
  A.dt2_over2 = dt * dt / 2.0;
  A.nsrc = src.nsrc;
  A.nstep = nstep;
  A.src_ispec.assign(std::max(1, src.nsrc), 0);
  A.src_elem.assign(std::max(1, src.nsrc), 0);
  A.src_array.assign(size_t(std::max(1, src.nsrc)) * 2 * NPT, 0.0f);
  A.stf.assign(size_t(nstep) * std::max(1, src.nsrc) * 2, 0.0f);
  {
    std::vector<float> t0s(src.nsrc);
    float min_t0 = std::numeric_limits<float>::max();
    // Dirac: f0 = 1 / (10 dt) in float (dirac.cpp:23-35); both kinds use
    // t0 = -1.2 hdur + tshift, hdur = 1 / f0 (ricker.cpp:17-20, dirac.cpp:13-16)
    std::vector<int> kind(src.nsrc, SMX_STF_RICKER);
    std::vector<float> f0v(src.nsrc);
    const float dtf = static_cast<float>(dt_in);
    for (int s = 0; s < src.nsrc; ++s) {
      kind[s] = src.stf ? src.stf[s] : SMX_STF_RICKER;
      t0s[s] = source_start_time(kind[s], src.f0[s], src.tshift[s], dt_in);
      if (kind[s] == SMX_STF_HEAVISIDE) { // the adjusted half duration
        float hd = src.f0[s];
        if (hd < 5.0 * dtf)
          hd = 5.0 * dtf;
        f0v[s] = hd / static_cast<float>(1.628);
      } else {
        f0v[s] = kind[s] == SMX_STF_DIRAC ? static_cast<float>(1.0 / (10.0 * dtf))
                                          : src.f0[s];
      }
      min_t0 = std::min(min_t0, t0s[s]);
    }
    if (opt.has_t0) // one medium of a coupled run: the run's start time
      min_t0 = static_cast<float>(opt.t0);
    A.t0 = (src.nsrc || opt.has_t0) ? min_t0 : 0.0;
    const float t0 = (src.nsrc || opt.has_t0) ? min_t0 : 0.0f;
    const float pi = static_cast<float>(M_PI);
    for (int s = 0; s < src.nsrc; ++s) {
      const float tshift = t0s[s] - min_t0;
      const float f0 = f0v[s];
      for (int i = 0; i < nstep; ++i) {
        const float t = t0 + i * dt;
        const float tt = t - tshift;
        const float a = pi * pi * f0 * f0;
        float val;
        if (kind[s] == SMX_STF_HEAVISIDE) {
          // factor * (1 + erf(t / hdur)) / 2 (impl/time_functions.hpp:132-136)
          const float hv = 0.5 * (1.0 + std::erf(tt / f0));
          val = src.factor[s] * hv;
        } else if (kind[s] == SMX_STF_DIRAC) {
          // factor * gaussian(t - tshift, f0) (impl/time_functions.hpp:9-14)
          const float g = -1.0 * std::exp(-a * tt * tt) / (2.0 * a);
          val = src.factor[s] * g;
        } else {
          const float d2 = (1.0 - 2.0 * a * tt * tt) * std::exp(-a * tt * tt);
          val = src.factor[s] * d2;
        }
        A.stf[(size_t(i) * src.nsrc + s) * 2 + 0] = val;
        A.stf[(size_t(i) * src.nsrc + s) * 2 + 1] = val;
      }
      int e;
      float xi, ga;
      locate(src.x[s], src.z[s], e, xi, ga);
      A.src_ispec[s] = slot_of[e];
      A.src_elem[s] = e;
      float hx[NGLL], hg[NGLL];
      interpolants(xi, q.xi, hx);
      interpolants(ga, q.xi, hg);
      const int skind = src.kind ? src.kind[s] : SMX_SRC_FORCE;
      if (skind != SMX_SRC_FORCE && skind != SMX_SRC_MOMENT)
        throw SmxError(SMX_ERR_ARG, "unknown source kind");
      if (skind == SMX_SRC_MOMENT) {
        // moment tensor (core/specfem/assembly/compute_source_array/dim2/
        // impl/compute_source_array_from_tensor.cpp:15-77): the element's
        // Jacobian interpolated at the source (include/algorithms/
        // interpolate.hpp:73-94), then [Mxx Mxz; Mxz Mzz] grad(h_xi h_gamma)
        if (acoustic)
          throw SmxError(SMX_ERR_ARG, "moment-tensor sources need elastic elements");
        if (!src.moment)
          throw SmxError(SMX_ERR_ARG, "moment-tensor source without a moment tensor");
        float hpx[NGLL], hpg[NGLL];
        interpolants_d(xi, q.xi, hx, hpx);
        interpolants_d(ga, q.xi, hg, hpg);
        const int sl = slot_of[e];
        const size_t pl = size_t(NPT) * ld;
        float dxix = 0, dxiz = 0, dgax = 0, dgaz = 0;
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            const float hl = hx[ix] * hg[iz];
            const size_t o = size_t(iz * NGLL + ix) * ld + sl;
            dxix += hl * A.geom[o];
            dxiz += hl * A.geom[pl + o];
            dgax += hl * A.geom[2 * pl + o];
            dgaz += hl * A.geom[3 * pl + o];
          }
        const float mxx = src.moment[3 * s], mzz = src.moment[3 * s + 1],
                    mxz = src.moment[3 * s + 2];
        for (int iz = 0; iz < NGLL; ++iz)
          for (int ix = 0; ix < NGLL; ++ix) {
            const float dsx = (hpx[ix] * dxix) * hg[iz] + hx[ix] * (hpg[iz] * dgax);
            const float dsz = (hpx[ix] * dxiz) * hg[iz] + hx[ix] * (hpg[iz] * dgaz);
            A.src_array[(size_t(s) * 2 + 0) * NPT + iz * NGLL + ix] = mxx * dsx + mxz * dsz;
            A.src_array[(size_t(s) * 2 + 1) * NPT + iz * NGLL + ix] = mxz * dsx + mzz * dsz;
          }
        continue;
      }
      const float ang = src.angle[s] * pi / static_cast<float>(180.0);
      float f[2] = {std::sin(ang), std::cos(ang)};
      if (acoustic) { // force vector (1) (force_source.cpp:36-40)
        f[0] = 1.0f;
        f[1] = 0.0f;
      }
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const float hl = hx[ix] * hg[iz];
          for (int c = 0; c < 2; ++c)
            A.src_array[(size_t(s) * 2 + c) * NPT + iz * NGLL + ix] = hl * f[c];
        }
    }
  }

  // ---- receivers (core/specfem/assembly/receivers/dim2/receivers.cpp:63-104)
  A.nrec = rec.nrec;
  A.nsbs = rec.nstep_between_samples;
  A.nsig = (nstep + A.nsbs - 1) / A.nsbs;
  A.seis_types.assign(rec.types, rec.types + rec.ntypes);
  for (int t : A.seis_types)
    if (t < 0 || t > 3)
      throw SmxError(SMX_ERR_ARG, "unsupported seismogram type");
  A.rec_ispec.assign(std::max(1, rec.nrec), 0);
  A.rec_elem.assign(std::max(1, rec.nrec), 0);
  A.rec_xi.assign(std::max(1, rec.nrec), 0.0f);
  A.rec_gamma.assign(std::max(1, rec.nrec), 0.0f);
  A.rec_h.assign(size_t(std::max(1, rec.nrec)) * NPT, 0.0f);
  for (int r = 0; r < rec.nrec; ++r) {
    int e;
    float xi, ga;
    locate(rec.x[r], rec.z[r], e, xi, ga);
    A.rec_ispec[r] = slot_of[e];
    A.rec_elem[r] = e;
    A.rec_xi[r] = xi;
    A.rec_gamma[r] = ga;
    float hx[NGLL], hg[NGLL];
    interpolants(xi, q.xi, hx);
    interpolants(ga, q.xi, hg);
    for (int iz = 0; iz < NGLL; ++iz)
      for (int ix = 0; ix < NGLL; ++ix)
        A.rec_h[size_t(r) * NPT + iz * NGLL + ix] = hx[ix] * hg[iz];
  }

  // ---- per-tile source entries: every point of a source's element lives in
  // that element's tile, which injects the term into its LDS accumulator
  {
    std::vector<std::array<int, 3>> ent; // (tile, source, point)
    for (int s = 0; s < src.nsrc; ++s)
      for (int p = 0; p < NPT; ++p)
        ent.push_back({A.src_ispec[s] / ELEM_TILE, s, p});
    std::stable_sort(ent.begin(), ent.end(),
                     [](const auto &a, const auto &b) { return a[0] < b[0]; });
    for (auto &t : A.tiles) {
      t.src_off = 0;
      t.nsrc = 0;
    }
    for (size_t i = 0; i < ent.size(); ++i) {
      const int k = ent[i][0], s = ent[i][1], p = ent[i][2];
      TileInfo &t = A.tiles[k];
      if (t.nsrc == 0)
        t.src_off = static_cast<int>(i);
      t.nsrc++;
      A.src_li.push_back(A.lidx[size_t(p) * ld + A.src_ispec[s]]);
      A.src_idx.push_back(s);
      A.src_L.push_back(A.src_array[(size_t(s) * 2 + 0) * NPT + p]);
      if (acoustic) // -stf L / kappa (acoustic/isotropic/source.hpp:60-66)
        A.src_L.push_back(A.kappa_mu_rho[size_t(p) * ld + A.src_ispec[s]]);
      else
        A.src_L.push_back(A.src_array[(size_t(s) * 2 + 1) * NPT + p]);
    }
  }

  // ---- receiver entries (receiver r, point p) -> entry r*NPT + p.  Entries
  // on a tile's exclusive node are evaluated by that tile; entries on shared
  // nodes by the seismogram pass that runs before the shared-node update,
  // except when an elastic mesh records displacement (and pressure) only:
  // the tile that owns the receiver's element has every point's
  // displacement in hand, so it evaluates all 25 entries.
  {
    bool tile_all = !acoustic;
    for (int t : A.seis_types)
      tile_all = tile_all && (t == SMX_SEIS_DISPLACEMENT || t == SMX_SEIS_PRESSURE);
    const int nent = A.nrec * NPT;
    A.rent_g.assign(std::max(1, nent), 0);
    A.rent_shared.assign(std::max(1, nent), -1);
    A.rent_h.assign(std::max(1, nent), 0.0f);
    std::vector<std::array<int, 3>> ent; // (tile, local node, entry)
    for (int r = 0; r < A.nrec; ++r)
      for (int p = 0; p < NPT; ++p) {
        const int s = A.rec_ispec[r];
        const int g = A.ibool[size_t(p) * ld + s];
        const int id = r * NPT + p;
        A.rent_g[id] = g;
        A.rent_h[id] = A.rec_h[size_t(r) * NPT + p];
        if (g >= A.nex_total && !tile_all)
          A.rent_shared[id] = g - A.nex_total;
        else
          ent.push_back({s / ELEM_TILE, A.lidx[size_t(p) * ld + s], id});
      }
    std::stable_sort(ent.begin(), ent.end(),
                     [](const auto &a, const auto &b) { return a[0] < b[0]; });
    for (auto &t : A.tiles) {
      t.rec_off = 0;
      t.nrec = 0;
    }
    for (size_t i = 0; i < ent.size(); ++i) {
      TileInfo &t = A.tiles[ent[i][0]];
      if (t.nrec == 0)
        t.rec_off = static_cast<int>(i);
      t.nrec++;
      A.trec_li.push_back(ent[i][1]);
      A.trec_entry.push_back(ent[i][2]);
    }
  }

  // ---- acoustic receiver weights: the reference interpolates
  // rho^-1 grad f at the element's points (acoustic/isotropic/
  // wavefield.hpp:74-93, compute_seismogram.tpp:150-160); that is linear in
  // the 25 nodal values, w_q = sum_p h_p rho^-1_p dgrad_p / df_q.
  if (acoustic) {
    const int nent = A.nrec * NPT;
    A.rent_w.assign(2 * size_t(std::max(1, nent)), 0.0f);
    A.rent_w64.assign(2 * size_t(std::max(1, nent)), 0.0);
    const size_t pld = size_t(NPT) * ld;
    for (int r = 0; r < A.nrec; ++r) {
      const int sl = A.rec_ispec[r];
      double w[NPT][2] = {};
      for (int iz = 0; iz < NGLL; ++iz)
        for (int ix = 0; ix < NGLL; ++ix) {
          const int p = iz * NGLL + ix;
          const size_t o = size_t(p) * ld + sl;
          const double h = A.rec_h[size_t(r) * NPT + p];
          const double rinv = 1.0 / double(A.kappa_mu_rho[2 * pld + o]);
          const double xix = A.geom[o], xiz = A.geom[pld + o],
                       gax = A.geom[2 * pld + o], gaz = A.geom[3 * pld + o];
          for (int l = 0; l < NGLL; ++l) {
            const double dxi = q.hprime[ix][l], dga = q.hprime[iz][l];
            w[iz * NGLL + l][0] += h * rinv * xix * dxi;
            w[iz * NGLL + l][1] += h * rinv * xiz * dxi;
            w[l * NGLL + ix][0] += h * rinv * gax * dga;
            w[l * NGLL + ix][1] += h * rinv * gaz * dga;
          }
        }
      for (int p = 0; p < NPT; ++p)
        for (int c = 0; c < 2; ++c) {
          A.rent_w64[2 * (size_t(r) * NPT + p) + c] = w[p][c];
          A.rent_w[2 * (size_t(r) * NPT + p) + c] = static_cast<float>(w[p][c]);
        }
    }
  }

  refresh_compact(A);
  A.shot_elem_off = {0, A.nspec};
  A.shot_src_off = {0, A.nsrc};
  A.shot_rec_off = {0, A.nrec};
  A.shot_tile_off = {0, A.ntiles};
  A.shot_t0 = {A.t0};
  finish_assembly(A);
}


// Shared point j -> its partial slots in sh_list order (tile order), the
// fixed summation order of the shared pass; and the largest tile.
static void build_shared_csr(HostAssembly2D &A) {
  A.shp_off.assign(A.nshared + 1, 0);
  for (int g : A.sh_list)
    A.shp_off[g - A.nex_total + 1]++;
  for (int j = 0; j < A.nshared; ++j)
    A.shp_off[j + 1] += A.shp_off[j];
  A.shp_pos.assign(A.sh_list.size(), 0);
  std::vector<int> fill(A.shp_off.begin(), A.shp_off.end() - 1);
  for (size_t k = 0; k < A.sh_list.size(); ++k)
    A.shp_pos[fill[A.sh_list[k] - A.nex_total]++] = static_cast<int>(k);
  int nmax = 0;
  for (const auto &t : A.tiles)
    nmax = std::max(nmax, t.nex + t.nsh);
  A.nloc_max = nmax;
}

// Lists derived from the numbering, receivers and Stacey arrays: receiver
// nodes for the adjoint source, absorbing points and nodes.
static void finish_assembly(HostAssembly2D &A) {
  const int ld = A.nspec_pad, sld = A.nstacey_pad;
  A.rin_node.clear();
  A.rin_off.clear();
  A.rin_ent.clear();
  A.rin_excl.clear();
  A.bpt_slot.clear();
  A.bpt_p.clear();
  A.bpt_bslot.clear();
  A.bpt_g.clear();
  A.bnode_g.clear();
  // ---- receiver entries grouped by node (adjoint source injection in a
  // fixed order)
  {
    std::map<int, std::vector<int>> by_node;
    for (int id = 0; id < A.nrec * NPT; ++id)
      by_node[A.rent_g[id]].push_back(id);
    A.rin_off.assign(1, 0);
    for (auto &kv : by_node) {
      A.rin_node.push_back(kv.first);
      for (int id : kv.second)
        A.rin_ent.push_back(id);
      A.rin_off.push_back(static_cast<int>(A.rin_ent.size()));
    }
    A.sh_rin.assign(std::max(1, A.nshared), -1);
    for (size_t j = 0; j < A.rin_node.size(); ++j) {
      const int g = A.rin_node[j];
      if (g >= A.nex_total)
        A.sh_rin[g - A.nex_total] = static_cast<int>(j);
      else
        A.rin_excl.push_back(static_cast<int>(j));
    }
  }

  // ---- absorbing points and nodes
  {
    std::map<int, int> node_index;
    for (int s = 0; s < ld; ++s) {
      const int bs = A.stacey_slot[s];
      if (bs < 0)
        continue;
      for (int p = 0; p < NPT; ++p) {
        if (A.stacey[(3 * size_t(NPT) + p) * sld + bs] == 0.0f)
          continue; // not on an absorbing edge
        const int g = A.ibool[size_t(p) * ld + s];
        A.bpt_slot.push_back(s);
        A.bpt_p.push_back(p);
        A.bpt_bslot.push_back(bs);
        A.bpt_g.push_back(g);
        if (node_index.emplace(g, static_cast<int>(A.bnode_g.size())).second)
          A.bnode_g.push_back(g);
      }
    }
    refresh_boundary_matrices(A);
  }
}

// Model-derived coefficients in double for FP64 contexts: lambda+2mu at every
// point and the Stacey impedances at absorbing points, from the FP32 model.
static void refresh_fp64(HostAssembly2D &A) {
  const int ld = A.nspec_pad, sld = A.nstacey_pad;
  const size_t pld = size_t(NPT) * ld, sp = size_t(NPT) * sld;
  A.lam2mu64.assign(pld, 0.0);
  A.imp64.assign(2 * sp, 0.0);
  if (A.medium == 1) { // acoustic: 1/rho and 1/(rho vp) in double
    for (size_t o = 0; o < pld; ++o)
      A.lam2mu64[o] = A.kappa_mu_rho[2 * pld + o] > 0
                          ? 1.0 / double(A.kappa_mu_rho[2 * pld + o])
                          : 0.0;
    for (size_t k = 0; k < A.bpt_g.size(); ++k) {
      const size_t o = size_t(A.bpt_p[k]) * ld + A.bpt_slot[k];
      const size_t so = size_t(A.bpt_p[k]) * sld + A.bpt_bslot[k];
      const double kappa = A.kappa_mu_rho[o], rho = A.kappa_mu_rho[2 * pld + o];
      A.imp64[so] = std::sqrt(1.0 / (rho * kappa));
    }
    return;
  }
  for (size_t o = 0; o < pld; ++o)
    A.lam2mu64[o] = double(A.kappa_mu_rho[o]) +
                    4.0 / 3.0 * double(A.kappa_mu_rho[pld + o]);
  for (size_t k = 0; k < A.bpt_g.size(); ++k) {
    const size_t o = size_t(A.bpt_p[k]) * ld + A.bpt_slot[k];
    const size_t so = size_t(A.bpt_p[k]) * sld + A.bpt_bslot[k];
    const double mu = A.kappa_mu_rho[pld + o], rho = A.kappa_mu_rho[2 * pld + o];
    A.imp64[so] = std::sqrt(A.lam2mu64[o] / rho) * rho;
    A.imp64[sp + so] = std::sqrt(rho * mu);
  }
}

void refresh_boundary_matrices(HostAssembly2D &A) {
  refresh_fp64(A);
  const int sld = A.nstacey_pad;
  std::unordered_map<int, int> index;
  for (size_t j = 0; j < A.bnode_g.size(); ++j)
    index[A.bnode_g[j]] = static_cast<int>(j);
  A.bsum.assign(3 * std::max<size_t>(1, A.bnode_g.size()), 0.0);
  A.bsum64.assign(A.bsum.size(), 0.0);
  const size_t sp = size_t(NPT) * sld;
  for (size_t k = 0; k < A.bpt_g.size(); ++k) {
    const int p = A.bpt_p[k], bs = A.bpt_bslot[k];
    const size_t so = size_t(p) * sld + bs;
    const double n0 = A.stacey[so], n1 = A.stacey[sp + so],
                 nrm = A.stacey[2 * sp + so], w = A.stacey[3 * sp + so];
    const double h0 = n0 / nrm, h1 = n1 / nrm, f = nrm * w;
    const int j = index[A.bpt_g[k]];
    const double imp[2][2] = {{A.stacey[4 * sp + so], A.stacey[5 * sp + so]},
                              {A.imp64[so], A.imp64[sp + so]}};
    std::vector<double> *out[2] = {&A.bsum, &A.bsum64};
    for (int v = 0; v < 2; ++v) {
      const double P = imp[v][0], S = imp[v][1];
      std::vector<double> &acc = *out[v];
      acc[3 * j + 0] += f * ((P - S) * h0 * h0 + S);
      acc[3 * j + 1] += f * ((P - S) * h0 * h1);
      acc[3 * j + 2] += f * ((P - S) * h1 * h1 + S);
    }
  }
}

void set_model2d(HostAssembly2D &A, const float *kappa, const float *mu,
                 const float *rho) {
  const int ld = A.nspec_pad, sld = A.nstacey_pad;
  const size_t pld = size_t(NPT) * ld;
  auto l2m_of = [](float k, float m) {
    return k + (static_cast<float>(4.0 / 3.0)) * m;
  };
  for (int e = 0; e < A.nspec; ++e) // validate before changing anything
    for (int p = 0; p < NPT; ++p) {
      const size_t i = size_t(e) * NPT + p;
      if (!(rho[i] > 0) || !(mu[i] >= 0) || !(kappa[i] > 0))
        throw SmxError(SMX_ERR_ARG, "non-physical material at element " +
                                        std::to_string(e));
    }
  for (int e = 0; e < A.nspec; ++e)
    for (int p = 0; p < NPT; ++p) {
      const size_t i = size_t(e) * NPT + p;
      const size_t o = size_t(p) * ld + A.elem_slot[e];
      A.lam2mu_mu[o] = l2m_of(kappa[i], mu[i]);
      A.lam2mu_mu[pld + o] = mu[i];
      A.kappa_mu_rho[o] = kappa[i];
      A.kappa_mu_rho[pld + o] = mu[i];
      A.kappa_mu_rho[2 * pld + o] = rho[i];
    }
  // Stacey impedances (include/medium/dim2/elastic/isotropic/properties.hpp:77-87)
  const size_t sp = size_t(NPT) * sld;
  for (size_t k = 0; k < A.bpt_g.size(); ++k) {
    const int p = A.bpt_p[k], bs = A.bpt_bslot[k], s = A.bpt_slot[k];
    const size_t o = size_t(p) * ld + s;
    const float kk = A.kappa_mu_rho[o], mm = A.kappa_mu_rho[pld + o],
                rr = A.kappa_mu_rho[2 * pld + o];
    const float l2m = l2m_of(kk, mm);
    A.stacey[4 * sp + size_t(p) * sld + bs] = std::sqrt(l2m / rr) * rr;
    A.stacey[5 * sp + size_t(p) * sld + bs] = std::sqrt(rr * mm);
  }
  refresh_boundary_matrices(A);
  refresh_compact(A);
}

void refresh_compact(HostAssembly2D &A) {
  const int ld = A.nspec_pad;
  const size_t pld = size_t(NPT) * ld;
  A.mat_e.assign(size_t(2) * ld, 0.0f);
  std::vector<uint8_t> mat_const(ld, 0);
  for (int sl = 0; sl < ld; ++sl) {
    if (A.ecolor[sl] == 0xff)
      continue;
    bool same = true;
    for (int c = 0; c < 2 && same; ++c)
      for (int p = 1; p < NPT && same; ++p)
        same = A.lam2mu_mu[c * pld + size_t(p) * ld + sl] ==
               A.lam2mu_mu[c * pld + sl];
    mat_const[sl] = same;
    for (int c = 0; c < 2; ++c)
      A.mat_e[c * size_t(ld) + sl] = A.lam2mu_mu[c * pld + sl];
  }
  for (int t = 0; t < A.ntiles; ++t) {
    bool g = true, m = true;
    for (int j = 0; j < A.tiles[t].ecount; ++j) {
      const int sl = t * ELEM_TILE + j;
      g = g && A.elem_affine[sl];
      m = m && mat_const[sl];
    }
    // the geometry source chosen by the solver does not depend on the model
    A.tiles[t].flags = (g ? TILE_GEOM_ELEM : 0) | (m ? TILE_MAT_ELEM : 0) |
                       (A.tiles[t].flags & TILE_GEOM_NODES);
  }
}

void merge_assemblies(std::vector<HostAssembly2D> &shots, HostAssembly2D &M) {
  if (shots.empty())
    throw SmxError(SMX_ERR_ARG, "no shots");
  for (const HostAssembly2D &S : shots)
    if (S.medium != 0)
      throw SmxError(SMX_ERR_ARG, "batched shots are built for elastic meshes");
  const HostAssembly2D &S0 = shots[0];
  for (const auto &S : shots)
    if (S.nstep != S0.nstep || S.dt != S0.dt || S.nsbs != S0.nsbs ||
        S.seis_types != S0.seis_types)
      throw SmxError(SMX_ERR_ARG,
                     "shots must share dt, nstep and the seismogram set-up");
  M = HostAssembly2D{};
  M.quad = S0.quad;
  M.nstep = S0.nstep;
  M.dt = S0.dt;
  M.dt_over2 = S0.dt_over2;
  M.dt2_over2 = S0.dt2_over2;
  M.nsbs = S0.nsbs;
  M.nsig = S0.nsig;
  M.seis_types = S0.seis_types;
  M.t0 = S0.t0;
  // totals and per-shot bases
  const size_t ns = shots.size();
  std::vector<int> tile_b(ns), ex_b(ns), sh_b(ns), shl_b(ns), el_b(ns),
      src_b(ns), rec_b(ns), st_b(ns), sent_b(ns), rent_b(ns);
  int ntiles = 0, nex = 0, nsh = 0, nshl = 0, nel = 0, nsrc = 0, nrec = 0,
      nst = 0, nsent = 0, nrent = 0;
  for (size_t k = 0; k < ns; ++k) {
    const HostAssembly2D &S = shots[k];
    tile_b[k] = ntiles, ex_b[k] = nex, sh_b[k] = nsh, shl_b[k] = nshl;
    el_b[k] = nel, src_b[k] = nsrc, rec_b[k] = nrec, st_b[k] = nst;
    sent_b[k] = nsent, rent_b[k] = nrent;
    ntiles += S.ntiles, nex += S.nex_total, nsh += S.nshared;
    nshl += int(S.sh_list.size()), nel += S.nspec, nsrc += S.nsrc;
    nrec += S.nrec, nst += S.nstacey, nsent += int(S.src_li.size());
    nrent += int(S.trec_li.size());
  }
  M.nspec = nel;
  M.ntiles = ntiles;
  M.nspec_pad = ntiles * ELEM_TILE;
  M.nex_total = nex;
  M.nshared = nsh;
  M.nglob = nex + nsh;
  M.nsrc = nsrc;
  M.nrec = nrec;
  M.nstacey = nst;
  M.nstacey_pad = std::max(ELEM_TILE, (nst + ELEM_TILE - 1) / ELEM_TILE * ELEM_TILE);
  const int ld = M.nspec_pad, sld = M.nstacey_pad;
  const size_t pld = size_t(NPT) * ld, sp = size_t(NPT) * sld;
  M.ecolor.assign(ld, 0xff);
  M.elem_slot.assign(nel, 0);
  M.ibool.assign(pld, 0);
  M.lidx.assign(pld, 0);
  M.geom.assign(5 * pld, 0.0f);
  M.lam2mu_mu.assign(2 * pld, 0.0f);
  M.kappa_mu_rho.assign(3 * pld, 0.0f);
  M.geom_e.assign(size_t(5) * ld, 0.0f);
  M.elem_affine.assign(ld, 0);
  bool all_q4 = true;
  for (const auto &S : shots)
    all_q4 = all_q4 && !S.nodes.empty();
  if (all_q4)
    M.nodes.assign(size_t(8) * ld, 0.0f);
  M.stacey_slot.assign(ld, -1);
  M.stacey.assign(6 * sp, 0.0f);
  M.src_ispec.assign(std::max(1, nsrc), 0);
  M.src_elem.assign(std::max(1, nsrc), 0);
  M.src_array.assign(size_t(std::max(1, nsrc)) * 2 * NPT, 0.0f);
  M.stf.assign(size_t(M.nstep) * std::max(1, nsrc) * 2, 0.0f);
  M.rec_ispec.assign(std::max(1, nrec), 0);
  M.rec_elem.assign(std::max(1, nrec), 0);
  M.rec_xi.assign(std::max(1, nrec), 0.0f);
  M.rec_gamma.assign(std::max(1, nrec), 0.0f);
  M.rec_h.assign(size_t(std::max(1, nrec)) * NPT, 0.0f);
  M.rent_g.assign(std::max(1, nrec * NPT), 0);
  M.rent_shared.assign(std::max(1, nrec * NPT), -1);
  M.rent_h.assign(std::max(1, nrec * NPT), 0.0f);
  M.shot_elem_off.assign(1, 0);
  M.shot_src_off.assign(1, 0);
  M.shot_rec_off.assign(1, 0);
  M.shot_tile_off.assign(1, 0);
  for (size_t k = 0; k < ns; ++k) {
    const HostAssembly2D &S = shots[k];
    const int sld0 = S.nspec_pad, ssl = S.nstacey_pad;
    const int slot_b = tile_b[k] * ELEM_TILE;
    auto remap = [&](int g) {
      return g < S.nex_total ? ex_b[k] + g : nex + sh_b[k] + (g - S.nex_total);
    };
    M.nloc_max = std::max(M.nloc_max, S.nloc_max);
    M.nelem_colors = std::max(M.nelem_colors, S.nelem_colors);
    for (TileInfo t : S.tiles) {
      t.ex_off += ex_b[k];
      t.sh_off += shl_b[k];
      t.src_off += sent_b[k];
      t.rec_off += rent_b[k];
      M.tiles.push_back(t);
    }
    for (int g : S.sh_list)
      M.sh_list.push_back(remap(g));
    for (int e = 0; e < S.nspec; ++e)
      M.elem_slot[el_b[k] + e] = slot_b + S.elem_slot[e];
    for (int sl = 0; sl < sld0; ++sl) {
      const int d = slot_b + sl;
      M.ecolor[d] = S.ecolor[sl];
      M.elem_affine[d] = S.elem_affine[sl];
      if (all_q4)
        for (int c = 0; c < 8; ++c)
          M.nodes[c * size_t(ld) + d] = S.nodes[c * size_t(sld0) + sl];
      for (int c = 0; c < 5; ++c)
        M.geom_e[c * size_t(ld) + d] = S.geom_e[c * size_t(sld0) + sl];
      M.stacey_slot[d] = S.stacey_slot[sl] >= 0 ? st_b[k] + S.stacey_slot[sl] : -1;
      for (int p = 0; p < NPT; ++p) {
        const size_t so = size_t(p) * sld0 + sl, dO = size_t(p) * ld + d;
        M.ibool[dO] = S.ecolor[sl] == 0xff ? 0 : remap(S.ibool[so]);
        M.lidx[dO] = S.lidx[so];
        for (int c = 0; c < 5; ++c)
          M.geom[c * pld + dO] = S.geom[c * size_t(NPT) * sld0 + so];
        for (int c = 0; c < 2; ++c)
          M.lam2mu_mu[c * pld + dO] = S.lam2mu_mu[c * size_t(NPT) * sld0 + so];
        for (int c = 0; c < 3; ++c)
          M.kappa_mu_rho[c * pld + dO] = S.kappa_mu_rho[c * size_t(NPT) * sld0 + so];
      }
    }
    for (int b = 0; b < S.nstacey; ++b)
      for (int c = 0; c < 6; ++c)
        for (int p = 0; p < NPT; ++p)
          M.stacey[(c * size_t(NPT) + p) * sld + st_b[k] + b] =
              S.stacey[(c * size_t(NPT) + p) * ssl + b];
    for (int s = 0; s < S.nsrc; ++s) {
      const int d = src_b[k] + s;
      M.src_ispec[d] = slot_b + S.src_ispec[s];
      M.src_elem[d] = el_b[k] + S.src_elem[s];
      for (int i = 0; i < 2 * NPT; ++i)
        M.src_array[size_t(d) * 2 * NPT + i] = S.src_array[size_t(s) * 2 * NPT + i];
      for (int it = 0; it < M.nstep; ++it)
        for (int c = 0; c < 2; ++c)
          M.stf[(size_t(it) * nsrc + d) * 2 + c] =
              S.stf[(size_t(it) * S.nsrc + s) * 2 + c];
    }
    for (size_t i = 0; i < S.src_li.size(); ++i) {
      M.src_li.push_back(S.src_li[i]);
      M.src_idx.push_back(src_b[k] + S.src_idx[i]);
      M.src_L.push_back(S.src_L[2 * i]);
      M.src_L.push_back(S.src_L[2 * i + 1]);
    }
    for (int r = 0; r < S.nrec; ++r) {
      const int d = rec_b[k] + r;
      M.rec_ispec[d] = slot_b + S.rec_ispec[r];
      M.rec_elem[d] = el_b[k] + S.rec_elem[r];
      M.rec_xi[d] = S.rec_xi[r];
      M.rec_gamma[d] = S.rec_gamma[r];
      for (int p = 0; p < NPT; ++p) {
        const int id = d * NPT + p, sid = r * NPT + p;
        M.rec_h[id] = S.rec_h[sid];
        const int g = remap(S.rent_g[sid]);
        M.rent_g[id] = g;
        M.rent_shared[id] = S.rent_shared[sid] >= 0 ? g - nex : -1;
        M.rent_h[id] = S.rent_h[sid];
      }
    }
    for (size_t i = 0; i < S.trec_li.size(); ++i) {
      M.trec_li.push_back(S.trec_li[i]);
      M.trec_entry.push_back(rec_b[k] * NPT + S.trec_entry[i]);
    }
    M.shot_elem_off.push_back(el_b[k] + S.nspec);
    M.shot_src_off.push_back(src_b[k] + S.nsrc);
    M.shot_rec_off.push_back(rec_b[k] + S.nrec);
    M.shot_tile_off.push_back(tile_b[k] + S.ntiles);
    M.shot_t0.push_back(S.t0);
  }
  build_shared_csr(M);
  finish_assembly(M);
  refresh_compact(M);
}

} // namespace smx
