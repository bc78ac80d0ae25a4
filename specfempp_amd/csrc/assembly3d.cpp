// Host assembly of the 3D elastic isotropic problem (see assembly3d.hpp).
#include "assembly3d.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <map>
#include <string>
#include <unordered_map>

namespace smx {

namespace {

inline int pidx(int iz, int iy, int ix) { return (iz * NGLL + iy) * NGLL + ix; }

// Hex8 shape functions / derivatives, evaluated in double and rounded to
// float (core/specfem/shape_function/dim3/shape_functions3d.cpp:9-23,
// shape_function_derivatives3d.cpp:9-40).
void hex8(double x, double e, double g, float *N) {
  N[0] = static_cast<float>(0.125 * (1.0 - x) * (1.0 - e) * (1.0 - g));
  N[1] = static_cast<float>(0.125 * (1.0 + x) * (1.0 - e) * (1.0 - g));
  N[2] = static_cast<float>(0.125 * (1.0 + x) * (1.0 + e) * (1.0 - g));
  N[3] = static_cast<float>(0.125 * (1.0 - x) * (1.0 + e) * (1.0 - g));
  N[4] = static_cast<float>(0.125 * (1.0 - x) * (1.0 - e) * (1.0 + g));
  N[5] = static_cast<float>(0.125 * (1.0 + x) * (1.0 - e) * (1.0 + g));
  N[6] = static_cast<float>(0.125 * (1.0 + x) * (1.0 + e) * (1.0 + g));
  N[7] = static_cast<float>(0.125 * (1.0 - x) * (1.0 + e) * (1.0 + g));
}

void dhex8(double x, double e, double g, float d[3][8]) {
  const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
  const double se[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
  const double sg[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  for (int n = 0; n < 8; ++n) {
    d[0][n] = static_cast<float>(0.125 * sx[n] * (1.0 + se[n] * e) * (1.0 + sg[n] * g));
    d[1][n] = static_cast<float>(0.125 * se[n] * (1.0 + sx[n] * x) * (1.0 + sg[n] * g));
    d[2][n] = static_cast<float>(0.125 * sg[n] * (1.0 + sx[n] * x) * (1.0 + se[n] * e));
  }
}

struct G3 {
  float t[10]; // xix, xiy, xiz, etax, etay, etaz, gax, gay, gaz, J
};

// core/specfem/jacobian/dim3/jacobian.cpp:27-74, accumulation in float
G3 jac3(const float *cx, const float *cy, const float *cz, const float d[3][8]) {
  float xxi = 0, yxi = 0, zxi = 0, xet = 0, yet = 0, zet = 0, xga = 0, yga = 0,
        zga = 0;
  for (int n = 0; n < 8; ++n) {
    xxi += d[0][n] * cx[n];
    yxi += d[0][n] * cy[n];
    zxi += d[0][n] * cz[n];
    xet += d[1][n] * cx[n];
    yet += d[1][n] * cy[n];
    zet += d[1][n] * cz[n];
    xga += d[2][n] * cx[n];
    yga += d[2][n] * cy[n];
    zga += d[2][n] * cz[n];
  }
  G3 g;
  const float J = xxi * (yet * zga - yga * zet) - xet * (yxi * zga - yga * zxi) +
                  xga * (yxi * zet - yet * zxi);
  g.t[0] = (yet * zga - yga * zet) / J;
  g.t[1] = (xga * zet - xet * zga) / J;
  g.t[2] = (xet * yga - xga * yet) / J;
  g.t[3] = (yga * zxi - yxi * zga) / J;
  g.t[4] = (xxi * zga - xga * zxi) / J;
  g.t[5] = (xga * yxi - xxi * yga) / J;
  g.t[6] = (yxi * zet - yet * zxi) / J;
  g.t[7] = (xet * zxi - xxi * zet) / J;
  g.t[8] = (xxi * yet - xet * yxi) / J;
  g.t[9] = J;
  return g;
}

uint64_t spread3(uint32_t v) { // 21 bits -> every third bit
  uint64_t x = v & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// Material (medium/dim2/elastic/isotropic/properties.hpp:71-91, shared by
// dim3): lambda+2mu = kappa + 4/3 mu, in float and in double.  kappa, mu, rho
// are [nspec][125] in mesh element order; everything is validated before A
// changes.
void material3d(HostAssembly3D &A, const float *kappa, const float *mu_in,
                const float *rho) {
  const int nspec = A.nspec, ld = nspec;
  for (size_t i = 0; i < size_t(nspec) * NPT3; ++i)
    if (!(rho[i] > 0) || !(mu_in[i] >= 0) || !(kappa[i] > 0))
      throw SmxError(SMX_ERR_ARG, "non-physical material at element " +
                                      std::to_string(i / NPT3));
  A.kappa_mu_rho.assign(size_t(3) * NPT3 * ld, 0.0f);
  std::vector<float> mpt(size_t(2) * NPT3 * ld);
  std::vector<double> mpt64(size_t(2) * NPT3 * ld);
  bool mat_const = true;
  for (int e = 0; e < nspec; ++e) {
    const int s = A.elem_slot[e];
    for (int p = 0; p < NPT3; ++p) {
      const size_t i = size_t(e) * NPT3 + p;
      const size_t o3 = size_t(s) * 3 * NPT3 + p, o2 = size_t(s) * 2 * NPT3 + p;
      const float k = kappa[i], mu = mu_in[i], r = rho[i];
      A.kappa_mu_rho[o3] = k;
      A.kappa_mu_rho[o3 + NPT3] = mu;
      A.kappa_mu_rho[o3 + 2 * NPT3] = r;
      mpt[o2] = k + static_cast<float>(4.0 / 3.0) * mu;
      mpt[o2 + NPT3] = mu;
      mpt64[o2] = double(k) + 4.0 / 3.0 * double(mu);
      mpt64[o2 + NPT3] = mu;
      if (mpt[o2] != mpt[o2 - p] || mu != mpt[o2 - p + NPT3])
        mat_const = false;
    }
  }
  A.mat_elem = mat_const;
  if (mat_const) {
    A.mat.assign(size_t(2) * ld, 0.0f);
    A.mat64.assign(size_t(2) * ld, 0.0);
    for (int c = 0; c < 2; ++c)
      for (int s = 0; s < ld; ++s) {
        A.mat[size_t(s) * 2 + c] = mpt[(size_t(s) * 2 + c) * NPT3];
        A.mat64[size_t(s) * 2 + c] = mpt64[(size_t(s) * 2 + c) * NPT3];
      }
  } else {
    A.mat = std::move(mpt);
    A.mat64 = std::move(mpt64);
  }
}

// Stacey impedances rho vp = sqrt((lambda+2mu)/rho) rho and rho vs =
// sqrt(rho mu) at the absorbing face points (medium/dim2/elastic/isotropic/
// properties.hpp:77-87), in float and, for FP64 contexts, in double.
void impedances3d(HostAssembly3D &A) {
  for (int s = 0; s < A.nspec; ++s) {
    const int b = A.stacey_slot[s];
    if (b < 0)
      continue;
    for (int p = 0; p < NPT3; ++p) {
      if (A.stacey[(size_t(b) * 7 + 4) * NPT3 + p] == 0.0f)
        continue;
      const size_t o3 = size_t(s) * 3 * NPT3 + p;
      const float kk = A.kappa_mu_rho[o3], mm = A.kappa_mu_rho[o3 + NPT3],
                  rr = A.kappa_mu_rho[o3 + 2 * NPT3];
      const float l2m = kk + static_cast<float>(4.0 / 3.0) * mm;
      A.stacey[(size_t(b) * 7 + 5) * NPT3 + p] = std::sqrt(l2m / rr) * rr;
      A.stacey[(size_t(b) * 7 + 6) * NPT3 + p] = std::sqrt(rr * mm);
      const double l2m64 = double(kk) + 4.0 / 3.0 * double(mm);
      A.stacey64[(size_t(b) * 2) * NPT3 + p] = std::sqrt(l2m64 / rr) * rr;
      A.stacey64[(size_t(b) * 2 + 1) * NPT3 + p] = std::sqrt(double(rr) * mm);
    }
  }
}

// Mass with the Stacey term at v = (-dt/2, -dt/2, -dt/2) (compute_mass_matrix.
// tpp:23-122, stacey.tpp:817-853 carried to faces), summed in double,
// inverted per component.
void mass3d(HostAssembly3D &A) {
  const Quadrature &q = A.quad;
  const float dt = A.dt;
  std::vector<double> M(size_t(3) * A.nglob, 0.0), M64(M.size(), 0.0);
  for (int s = 0; s < A.nspec; ++s) {
    const int b = A.stacey_slot[s];
    for (int iz = 0; iz < NGLL; ++iz)
      for (int iy = 0; iy < NGLL; ++iy)
        for (int ix = 0; ix < NGLL; ++ix) {
          const int p = pidx(iz, iy, ix);
          const float f = q.w[iz] * q.w[iy] * q.w[ix] * A.jac_pt[size_t(s) * NPT3 + p];
          const float rr = A.kappa_mu_rho[(size_t(s) * 3 + 2) * NPT3 + p];
          double add[3] = {double(rr * f), double(rr * f), double(rr * f)};
          double add64[3] = {double(rr) * f, double(rr) * f, double(rr) * f};
          if (b >= 0 && A.stacey[(size_t(b) * 7 + 4) * NPT3 + p] != 0.0f) {
            const float *S = A.stacey.data();
            auto at = [&](int c) { return S[(size_t(b) * 7 + c) * NPT3 + p]; };
            const float nn[3] = {at(0), at(1), at(2)}, nrm = at(3), w = at(4);
            const float vf = -1.0f * dt * 0.5f;
            const float vn = vf * nn[0] + vf * nn[1] + vf * nn[2];
            const double P64 = A.stacey64[(size_t(b) * 2) * NPT3 + p],
                         S64 = A.stacey64[(size_t(b) * 2 + 1) * NPT3 + p];
            for (int c = 0; c < 3; ++c) {
              const float fc = ((vn * nn[c] / (nrm * nrm)) * (at(5) - at(6))) + vf * at(6);
              add[c] += double(-1.0f * fc * nrm * w);
              const double fc64 =
                  ((double(vn) * nn[c] / (double(nrm) * nrm)) * (P64 - S64)) + double(vf) * S64;
              add64[c] += -fc64 * nrm * w;
            }
          }
          const int g = A.ibool[size_t(s) * NPT3 + p];
          for (int c = 0; c < 3; ++c) {
            M[size_t(c) * A.nglob + g] += add[c];
            M64[size_t(c) * A.nglob + g] += add64[c];
          }
        }
  }
  A.minv.resize(M.size());
  A.minv64.resize(M.size());
  for (size_t k = 0; k < M.size(); ++k) {
    A.minv[k] = static_cast<float>(1.0 / M[k]);
    A.minv64[k] = 1.0 / M64[k];
  }
}

// Consistency of the tile layout (every index the tile kernels use stays
// in bounds and maps back to the global numbering); throws on violation.
void validate_tiles(const HostAssembly3D &A) {
  auto fail = [](const std::string &w) {
    throw SmxError(SMX_ERR_MESH, "3D tile layout: " + w);
  };
  const int RW = A.round_width;
  if (int(A.round_slots.size()) % RW)
    fail("round table size");
  const int nrounds = int(A.round_slots.size()) / RW;
  std::vector<int> seen_slot(A.nspec, 0), stamp(A.nglob, -1);
  long long ex_sum = 0;
  for (int t = 0; t < A.ntiles; ++t) {
    const TileInfo3 &T = A.tiles[t];
    const int nloc = T.nex + T.nsh;
    if (T.slot0 < 0 || T.nelem < 1 || T.slot0 + T.nelem > A.nspec ||
        nloc > A.nloc_max || T.ex_off < 0 || T.ex_off + T.nex > A.nex_total ||
        T.sh_off < 0 || T.sh_off + T.nsh > A.npart || T.round_off < 0 ||
        T.round_off + T.nround > nrounds || T.src_off < 0 ||
        T.src_off + T.nsrc > int(A.tsrc_id.size()))
      fail("tile " + std::to_string(t) + " header out of range");
    ex_sum += T.nex;
    for (int sl = T.slot0; sl < T.slot0 + T.nelem; ++sl)
      for (int p = 0; p < NPT3; ++p) {
        const int l = A.lidx[size_t(sl) * NPT3 + p];
        const int g = A.ibool[size_t(sl) * NPT3 + p];
        if (l >= nloc)
          fail("local id out of range");
        const int gl = l < T.nex ? T.ex_off + l : A.sh_list[T.sh_off + l - T.nex];
        if (gl != g || g < 0 || g >= A.nglob)
          fail("local id does not map back to the global id");
      }
    for (int r = 0; r < T.nround; ++r) {
      for (int k = 0; k < RW; ++k) {
        const int sl = A.round_slots[size_t(T.round_off + r) * RW + k];
        if (sl < 0)
          continue;
        if (sl < T.slot0 || sl >= T.slot0 + T.nelem)
          fail("round slot outside its tile");
        seen_slot[sl]++;
        for (int p = 0; p < NPT3; ++p) {
          const int g = A.ibool[size_t(sl) * NPT3 + p];
          if (stamp[g] == t * 4096 + r)
            fail("two elements of a round share a point");
          stamp[g] = t * 4096 + r;
        }
      }
    }
    for (int k = T.src_off; k < T.src_off + T.nsrc; ++k)
      if (A.tsrc_li[k] < 0 || A.tsrc_li[k] >= nloc || A.tsrc_id[k] < 0 ||
          A.tsrc_id[k] >= std::max(1, A.nsrc))
        fail("source entry out of range");
  }
  for (int sl = 0; sl < A.nspec; ++sl)
    if (seen_slot[sl] != 1)
      fail("slot not in exactly one round");
  if (ex_sum != A.nex_total || A.nex_total + A.nsh_total != A.nglob)
    fail("exclusive / shared counts");
  if (int(A.sh_poff.size()) != A.nsh_total + 1 || A.sh_poff[A.nsh_total] != A.npart)
    fail("shared CSR size");
  std::vector<int> used(A.npart, 0);
  for (int s = 0; s < A.nsh_total; ++s)
    for (int k = A.sh_poff[s]; k < A.sh_poff[s + 1]; ++k) {
      const int q = A.sh_pidx[k];
      if (q < 0 || q >= A.npart || A.sh_list[q] != A.nex_total + s)
        fail("partial slot of the wrong point");
      used[q]++;
    }
  for (int q = 0; q < A.npart; ++q)
    if (used[q] != 1)
      fail("partial slot not summed exactly once");
}

} // namespace

void set_model3d(HostAssembly3D &A, const float *kappa, const float *mu,
                 const float *rho) {
  if (!kappa || !mu || !rho)
    throw SmxError(SMX_ERR_ARG, "null model");
  material3d(A, kappa, mu, rho);
  impedances3d(A);
  mass3d(A);
}

void assemble3d(const smx_mesh3d &m, const smx_sources3d &src,
                const smx_receivers3d &rec, double dt_in, int nstep,
                HostAssembly3D &A, const AssemblyOptions3D &opt) {
  if (opt.tile_elems < 1 || opt.tile_elems > 64 || opt.round_width < 1 ||
      opt.round_width > 4)
    throw SmxError(SMX_ERR_ARG, "3D tile options out of range");
  if (m.nspec <= 0 || m.nnodes <= 0 || !m.coord || !m.knods || !m.kappa ||
      !m.mu || !m.rho)
    throw SmxError(SMX_ERR_ARG, "invalid 3D mesh description");
  if (m.nabs < 0 || (m.nabs > 0 && (!m.abs_ispec || !m.abs_face)))
    throw SmxError(SMX_ERR_ARG, "invalid absorbing face list");
  if (nstep <= 0 || !(dt_in > 0))
    throw SmxError(SMX_ERR_ARG, "dt and nstep must be positive");
  if (src.nsrc < 0 || rec.nrec < 0 || rec.ntypes < 0 ||
      rec.nstep_between_samples <= 0)
    throw SmxError(SMX_ERR_ARG, "invalid source/receiver description");
  const int nspec = m.nspec;
  A.nspec = nspec;
  const Quadrature &q = A.quad;

  std::vector<float> cx(size_t(nspec) * 8), cy(cx.size()), cz(cx.size());
  for (int e = 0; e < nspec; ++e)
    for (int n = 0; n < 8; ++n) {
      const int id = m.knods[size_t(e) * 8 + n];
      if (id < 0 || id >= m.nnodes)
        throw SmxError(SMX_ERR_ARG, "knods out of range");
      cx[e * 8 + n] = static_cast<float>(m.coord[3 * size_t(id)]);
      cy[e * 8 + n] = static_cast<float>(m.coord[3 * size_t(id) + 1]);
      cz[e * 8 + n] = static_cast<float>(m.coord[3 * size_t(id) + 2]);
    }

  // ---- colouring by corner nodes (elements sharing a GLL point share a
  // control node on a conforming hex8 mesh)
  std::vector<int> color(nspec);
  {
    std::vector<uint32_t> mask(m.nnodes, 0u);
    int nc = 0;
    for (int e = 0; e < nspec; ++e) {
      uint32_t used = 0;
      for (int n = 0; n < 8; ++n)
        used |= mask[m.knods[size_t(e) * 8 + n]];
      int c = 0;
      while (c < 32 && ((used >> c) & 1u))
        ++c;
      if (c >= 32)
        throw SmxError(SMX_ERR_MESH, "element colouring needs more than 32 colours");
      color[e] = c;
      nc = std::max(nc, c + 1);
      for (int n = 0; n < 8; ++n)
        mask[m.knods[size_t(e) * 8 + n]] |= 1u << c;
    }
    A.ncolors = nc;
  }

  // ---- Morton order of the element centroids; tiles of consecutive
  // elements, ordered by colour inside a tile
  std::vector<int> morton(nspec);
  {
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    std::vector<std::array<float, 3>> c(nspec);
    for (int e = 0; e < nspec; ++e) {
      float s[3] = {0, 0, 0};
      for (int n = 0; n < 8; ++n) {
        s[0] += cx[e * 8 + n];
        s[1] += cy[e * 8 + n];
        s[2] += cz[e * 8 + n];
      }
      for (int d = 0; d < 3; ++d) {
        c[e][d] = s[d] * 0.125f;
        lo[d] = std::min(lo[d], c[e][d]);
        hi[d] = std::max(hi[d], c[e][d]);
      }
    }
    // grid spacing: the edge of an element of average volume, from the
    // control nodes' bounding box (on a structured box the centroids then
    // sit exactly on the grid, and tiles of 32 are 4x4x2 blocks)
    double vol = 1;
    {
      const std::vector<float> *cs[3] = {&cx, &cy, &cz};
      for (int d = 0; d < 3; ++d) {
        const auto mm = std::minmax_element(cs[d]->begin(), cs[d]->end());
        vol *= std::max(double(*mm.second) - double(*mm.first), 1e-6);
      }
    }
    const double h = std::max(std::cbrt(vol / nspec), 1e-6);
    std::vector<std::pair<uint64_t, int>> key(nspec);
    for (int e = 0; e < nspec; ++e) {
      uint32_t g[3];
      for (int d = 0; d < 3; ++d)
        g[d] = uint32_t(std::max(0.0, std::floor((c[e][d] - lo[d]) / h + 0.5)));
      key[e] = {spread3(g[0]) | (spread3(g[1]) << 1) | (spread3(g[2]) << 2), e};
    }
    std::sort(key.begin(), key.end());
    for (int k = 0; k < nspec; ++k)
      morton[k] = key[k].second;
  }
  A.elem_slot.assign(nspec, 0);
  const int TE = opt.tile_elems, RW = opt.round_width;
  A.tile_elems = TE;
  A.round_width = RW;
  A.ntiles = (nspec + TE - 1) / TE;
  A.tiles.assign(A.ntiles, TileInfo3{});
  std::vector<int> slot_color(nspec);
  for (int t = 0; t < A.ntiles; ++t) {
    const int k0 = t * TE, k1 = std::min(nspec, k0 + TE);
    std::vector<int> el(morton.begin() + k0, morton.begin() + k1);
    std::stable_sort(el.begin(), el.end(),
                     [&](int a, int b) { return color[a] < color[b]; });
    TileInfo3 &T = A.tiles[t];
    T.slot0 = k0;
    T.nelem = k1 - k0;
    T.round_off = int(A.round_slots.size()) / RW;
    for (size_t i = 0; i < el.size(); ++i) {
      A.elem_slot[el[i]] = k0 + int(i);
      slot_color[k0 + i] = color[el[i]];
    }
    // rounds: up to RW elements of one colour
    for (size_t i = 0; i < el.size();) {
      size_t j = i;
      while (j < el.size() && j - i < size_t(RW) && color[el[j]] == color[el[i]])
        ++j;
      for (int k = 0; k < RW; ++k)
        A.round_slots.push_back(i + k < j ? k0 + int(i) + k : -1);
      ++T.nround;
      i = j;
    }
  }
  A.color_off.assign(A.ncolors + 1, 0);
  A.color_slots.assign(nspec, 0);
  {
    std::vector<int> count(A.ncolors, 0);
    for (int sl = 0; sl < nspec; ++sl)
      count[slot_color[sl]]++;
    for (int c = 0; c < A.ncolors; ++c)
      A.color_off[c + 1] = A.color_off[c] + count[c];
    std::vector<int> fill(A.color_off.begin(), A.color_off.end() - 1);
    for (int sl = 0; sl < nspec; ++sl)
      A.color_slots[fill[slot_color[sl]]++] = sl;
  }
  const int ld = nspec;
  A.nodes.assign(size_t(24) * ld, 0.0f);
  for (int e = 0; e < nspec; ++e)
    for (int n = 0; n < 8; ++n) {
      const size_t o = size_t(A.elem_slot[e]) * 24;
      A.nodes[o + n] = cx[e * 8 + n];
      A.nodes[o + 8 + n] = cy[e * 8 + n];
      A.nodes[o + 16 + n] = cz[e * 8 + n];
    }

  // ---- geometry at the GLL points
  float dN[NPT3][3][8];
  for (int iz = 0; iz < NGLL; ++iz)
    for (int iy = 0; iy < NGLL; ++iy)
      for (int ix = 0; ix < NGLL; ++ix)
        dhex8(q.xi[ix], q.xi[iy], q.xi[iz], dN[pidx(iz, iy, ix)]);
  float dC[3][8];
  dhex8(0.0, 0.0, 0.0, dC);
  std::vector<float> gpt(size_t(10) * NPT3 * ld);
  std::vector<float> gel(size_t(10) * ld);
  bool all_affine = true;
  for (int e = 0; e < nspec; ++e) {
    const int s = A.elem_slot[e];
    const G3 gc = jac3(&cx[e * 8], &cy[e * 8], &cz[e * 8], dC);
    float scale_d = 0;
    for (int k = 0; k < 9; ++k)
      scale_d = std::max(scale_d, std::fabs(gc.t[k]));
    for (int k = 0; k < 10; ++k)
      gel[size_t(s) * 10 + k] = gc.t[k];
    for (int p = 0; p < NPT3; ++p) {
      const G3 g = jac3(&cx[e * 8], &cy[e * 8], &cz[e * 8], dN[p]);
      if (!(g.t[9] > 0))
        throw SmxError(SMX_ERR_MESH,
                       "non-positive Jacobian in element " + std::to_string(e));
      for (int k = 0; k < 10; ++k) {
        gpt[(size_t(s) * 10 + k) * NPT3 + p] = g.t[k];
        const float sc = k < 9 ? scale_d : std::fabs(gc.t[9]);
        if (std::fabs(g.t[k] - gc.t[k]) > 4e-7f * sc)
          all_affine = false;
      }
    }
  }
  A.geom_elem = all_affine;
  A.geom = all_affine ? std::move(gel) : std::move(gpt);
  A.jac_pt.resize(size_t(NPT3) * ld); // J per point for the mass matrix
  for (int s = 0; s < ld; ++s)
    for (int p = 0; p < NPT3; ++p)
      A.jac_pt[size_t(s) * NPT3 + p] =
          all_affine ? A.geom[size_t(s) * 10 + 9]
                     : A.geom[(size_t(s) * 10 + 9) * NPT3 + p];

  // ---- numbering: equal GLL coordinates are one point; ids by first touch
  // in Morton element order
  A.ibool.assign(size_t(NPT3) * ld, -1);
  {
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
    const double qn = 1e-4 * std::cbrt(vol / nspec);
    std::unordered_map<uint64_t, int> ids;
    ids.reserve(size_t(nspec) * 70);
    int next = 0;
    float Nn[NPT3][8];
    for (int iz = 0; iz < NGLL; ++iz)
      for (int iy = 0; iy < NGLL; ++iy)
        for (int ix = 0; ix < NGLL; ++ix)
          hex8(q.xi[ix], q.xi[iy], q.xi[iz], Nn[pidx(iz, iy, ix)]);
    for (int e : morton) {
      const int s = A.elem_slot[e];
      for (int p = 0; p < NPT3; ++p) {
        double x = 0, y = 0, z = 0;
        for (int n = 0; n < 8; ++n) {
          x += double(Nn[p][n]) * cx[e * 8 + n];
          y += double(Nn[p][n]) * cy[e * 8 + n];
          z += double(Nn[p][n]) * cz[e * 8 + n];
        }
        const uint64_t kx = uint64_t(std::llround((x - lo[0]) / qn)) & 0x1fffff;
        const uint64_t ky = uint64_t(std::llround((y - lo[1]) / qn)) & 0x1fffff;
        const uint64_t kz = uint64_t(std::llround((z - lo[2]) / qn)) & 0x1fffff;
        auto it = ids.emplace((kx << 42) | (ky << 21) | kz, next);
        if (it.second)
          ++next;
        A.ibool[size_t(s) * NPT3 + p] = it.first->second;
      }
    }
    A.nglob = next;
  }

  // ---- exclusive / shared points: renumber tile by tile (each tile's
  // exclusive points contiguous, in first-touch order), then the shared
  // points; tile-local ids (exclusive first, then the tile's shared points
  // in first-touch order); one partial slot per (tile, shared point)
  {
    const int ng = A.nglob;
    std::vector<int> owner(ng, -1); // tile, or -2 when shared
    for (int t = 0; t < A.ntiles; ++t) {
      const TileInfo3 &T = A.tiles[t];
      for (size_t i = size_t(T.slot0) * NPT3; i < size_t(T.slot0 + T.nelem) * NPT3; ++i) {
        const int g = A.ibool[i];
        owner[g] = owner[g] == -1 ? t : (owner[g] == t ? t : -2);
      }
    }
    std::vector<int> nid(ng, -1);
    int next = 0;
    for (int t = 0; t < A.ntiles; ++t) {
      TileInfo3 &T = A.tiles[t];
      T.ex_off = next;
      for (size_t i = size_t(T.slot0) * NPT3; i < size_t(T.slot0 + T.nelem) * NPT3; ++i) {
        const int g = A.ibool[i];
        if (owner[g] == t && nid[g] < 0)
          nid[g] = next++;
      }
      T.nex = next - T.ex_off;
    }
    A.nex_total = next;
    for (size_t i = 0; i < A.ibool.size(); ++i) {
      const int g = A.ibool[i];
      if (owner[g] == -2 && nid[g] < 0)
        nid[g] = next++;
    }
    if (next != ng)
      throw SmxError(SMX_ERR_MESH, "3D point renumbering is inconsistent");
    A.nsh_total = ng - A.nex_total;
    A.lidx.assign(size_t(nspec) * NPT3, 0);
    A.sh_list.clear();
    std::vector<int> stamp(ng, -1), lid(ng, 0);
    A.nloc_max = 0;
    for (int t = 0; t < A.ntiles; ++t) {
      TileInfo3 &T = A.tiles[t];
      T.sh_off = int(A.sh_list.size());
      int k = 0;
      for (size_t i = size_t(T.slot0) * NPT3; i < size_t(T.slot0 + T.nelem) * NPT3; ++i) {
        const int g = A.ibool[i];
        int l;
        if (owner[g] == t) {
          l = nid[g] - T.ex_off;
        } else {
          if (stamp[g] != t) {
            stamp[g] = t;
            lid[g] = T.nex + k++;
            A.sh_list.push_back(nid[g]);
          }
          l = lid[g];
        }
        if (l < 0 || l > 65535)
          throw SmxError(SMX_ERR_MESH, "tile has more than 65536 points");
        A.lidx[i] = static_cast<uint16_t>(l);
      }
      T.nsh = k;
      A.nloc_max = std::max(A.nloc_max, T.nex + T.nsh);
    }
    A.npart = int(A.sh_list.size());
    // shared point -> its partial slots, in tile order
    A.sh_poff.assign(A.nsh_total + 1, 0);
    for (int g : A.sh_list)
      A.sh_poff[g - A.nex_total + 1]++;
    for (int k = 0; k < A.nsh_total; ++k)
      A.sh_poff[k + 1] += A.sh_poff[k];
    A.sh_pidx.assign(A.npart, 0);
    {
      std::vector<int> fill(A.sh_poff.begin(), A.sh_poff.end() - 1);
      for (int k = 0; k < A.npart; ++k)
        A.sh_pidx[fill[A.sh_list[k] - A.nex_total]++] = k;
    }
    for (auto &g : A.ibool)
      g = nid[g];
  }

  material3d(A, m.kappa, m.mu, m.rho);

  // ---- absorbing faces (this build's 3D Stacey: the 2D point formula on
  // face points, normal n = +-J grad(xi_d), weight = the two in-face GLL
  // weights)
  std::map<int, std::vector<int>> faces_of;
  for (int k = 0; k < m.nabs; ++k) {
    const int e = m.abs_ispec[k], f = m.abs_face[k];
    if (e < 0 || e >= nspec || f < 1 || f > 6)
      throw SmxError(SMX_ERR_ARG, "invalid absorbing face");
    faces_of[e].push_back(f);
  }
  A.stacey_slot.assign(ld, -1);
  A.nstacey = static_cast<int>(faces_of.size());
  const int nst = std::max(1, A.nstacey);
  A.stacey.assign(size_t(7) * NPT3 * nst, 0.0f);
  A.stacey64.assign(size_t(2) * NPT3 * nst, 0.0);
  {
    int b = 0;
    for (auto &kv : faces_of) {
      const int e = kv.first, s = A.elem_slot[e];
      A.stacey_slot[s] = b;
      for (int p = 0; p < NPT3; ++p) {
        A.stacey[(size_t(b) * 7 + 0) * NPT3 + p] = 1.0f;
        A.stacey[(size_t(b) * 7 + 3) * NPT3 + p] = 1.0f;
      }
      for (int f : kv.second)
        for (int iz = 0; iz < NGLL; ++iz)
          for (int iy = 0; iy < NGLL; ++iy)
            for (int ix = 0; ix < NGLL; ++ix) {
              const bool on = (f == 1 && iz == 0) || (f == 3 && iz == NGLL - 1) ||
                              (f == 4 && ix == 0) || (f == 2 && ix == NGLL - 1) ||
                              (f == 5 && iy == 0) || (f == 6 && iy == NGLL - 1);
              if (!on)
                continue;
              const int p = pidx(iz, iy, ix);
              const G3 g = jac3(&cx[e * 8], &cy[e * 8], &cz[e * 8], dN[p]);
              int d;
              float sg, w;
              if (f == 2 || f == 4) {
                d = 0, sg = f == 2 ? 1.0f : -1.0f, w = q.w[iy] * q.w[iz];
              } else if (f == 5 || f == 6) {
                d = 1, sg = f == 6 ? 1.0f : -1.0f, w = q.w[ix] * q.w[iz];
              } else {
                d = 2, sg = f == 3 ? 1.0f : -1.0f, w = q.w[ix] * q.w[iy];
              }
              const float n0 = sg * g.t[3 * d] * g.t[9];
              const float n1 = sg * g.t[3 * d + 1] * g.t[9];
              const float n2 = sg * g.t[3 * d + 2] * g.t[9];
              const float vals[5] = {n0, n1, n2,
                                     std::sqrt(n0 * n0 + n1 * n1 + n2 * n2), w};
              for (int c = 0; c < 5; ++c)
                A.stacey[(size_t(b) * 7 + c) * NPT3 + p] = vals[c];
            }
      ++b;
    }
  }

  // ---- time scheme constants (newmark.hpp:131-135)
  const float dt = static_cast<float>(dt_in);
  A.dt = dt;
  A.dt_over2 = dt / 2.0;
  A.dt2_over2 = dt * dt / 2.0;

  impedances3d(A);
  mass3d(A);

  // ---- point location: candidate elements from a grid of element bounding
  // boxes, Newton in each (src/algorithms/dim3/locate_point_impl.cpp), the
  // closest converged point wins (ties: lowest element index)
  float blo[3] = {1e30f, 1e30f, 1e30f}, bhi[3] = {-1e30f, -1e30f, -1e30f};
  for (size_t k = 0; k < cx.size(); ++k) {
    blo[0] = std::min(blo[0], cx[k]), bhi[0] = std::max(bhi[0], cx[k]);
    blo[1] = std::min(blo[1], cy[k]), bhi[1] = std::max(bhi[1], cy[k]);
    blo[2] = std::min(blo[2], cz[k]), bhi[2] = std::max(bhi[2], cz[k]);
  }
  const int gdim = std::max(1, std::min(256, int(std::cbrt(double(nspec)) + 1)));
  std::vector<std::vector<int>> cells(size_t(gdim) * gdim * gdim);
  auto cell_of = [&](float v, int d) {
    const float w = std::max(bhi[d] - blo[d], 1e-6f);
    return std::min(gdim - 1, std::max(0, int((v - blo[d]) / w * gdim)));
  };
  for (int e = 0; e < nspec; ++e) {
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int n = 0; n < 8; ++n) {
      const float c[3] = {cx[e * 8 + n], cy[e * 8 + n], cz[e * 8 + n]};
      for (int d = 0; d < 3; ++d)
        lo[d] = std::min(lo[d], c[d]), hi[d] = std::max(hi[d], c[d]);
    }
    for (int a = cell_of(lo[0], 0); a <= cell_of(hi[0], 0); ++a)
      for (int b = cell_of(lo[1], 1); b <= cell_of(hi[1], 1); ++b)
        for (int c = cell_of(lo[2], 2); c <= cell_of(hi[2], 2); ++c)
          cells[(size_t(a) * gdim + b) * gdim + c].push_back(e);
  }
  auto locate = [&](double xd, double yd, double zd, int &e_out, float lc[3]) {
    const float x = float(xd), y = float(yd), z = float(zd);
    const auto &cand = cells[(size_t(cell_of(x, 0)) * gdim + cell_of(y, 1)) * gdim +
                             cell_of(z, 2)];
    float best = std::numeric_limits<float>::max();
    e_out = -1;
    for (int e : cand) {
      float xi = 0, et = 0, ga = 0;
      for (int it = 0; it < 100; ++it) {
        float N[8], d[3][8];
        hex8(xi, et, ga, N);
        dhex8(xi, et, ga, d);
        float px = 0, py = 0, pz = 0;
        for (int n = 0; n < 8; ++n) {
          px += N[n] * cx[e * 8 + n];
          py += N[n] * cy[e * 8 + n];
          pz += N[n] * cz[e * 8 + n];
        }
        const G3 g = jac3(&cx[e * 8], &cy[e * 8], &cz[e * 8], d);
        const float dx = x - px, dy = y - py, dz = z - pz;
        const float dxi = g.t[0] * dx + g.t[1] * dy + g.t[2] * dz;
        const float det = g.t[3] * dx + g.t[4] * dy + g.t[5] * dz;
        const float dga = g.t[6] * dx + g.t[7] * dy + g.t[8] * dz;
        xi = std::min(1.01f, std::max(-1.01f, xi + dxi));
        et = std::min(1.01f, std::max(-1.01f, et + det));
        ga = std::min(1.01f, std::max(-1.01f, ga + dga));
        if (std::fabs(dxi) < 1e-12f && std::fabs(det) < 1e-12f && std::fabs(dga) < 1e-12f)
          break;
      }
      float N[8];
      hex8(xi, et, ga, N);
      float px = 0, py = 0, pz = 0;
      for (int n = 0; n < 8; ++n) {
        px += N[n] * cx[e * 8 + n];
        py += N[n] * cy[e * 8 + n];
        pz += N[n] * cz[e * 8 + n];
      }
      const float dd = std::sqrt((x - px) * (x - px) + (y - py) * (y - py) +
                                 (z - pz) * (z - pz));
      if (dd < best || (dd == best && e < e_out)) {
        best = dd;
        e_out = e;
        lc[0] = xi, lc[1] = et, lc[2] = ga;
      }
    }
    if (e_out < 0 || std::fabs(std::fabs(lc[0]) - 1.01f) < 1e-6f ||
        std::fabs(std::fabs(lc[1]) - 1.01f) < 1e-6f ||
        std::fabs(std::fabs(lc[2]) - 1.01f) < 1e-6f)
      throw SmxError(SMX_ERR_MESH, "point (" + std::to_string(xd) + ", " +
                                       std::to_string(yd) + ", " +
                                       std::to_string(zd) + ") lies outside the mesh");
  };
  auto lagrange = [&](float xi, float *h, float *hp) {
    // src/quadrature/gll/lagrange_poly.cpp:11-43
    for (int dgr = 0; dgr < NGLL; ++dgr) {
      float p1 = 1.0f, p2 = 1.0f, sum = 0.0f;
      const float x0 = q.xi[dgr];
      for (int i = 0; i < NGLL; ++i) {
        if (i == dgr)
          continue;
        p1 = p1 * (xi - q.xi[i]);
        p2 = p2 * (x0 - q.xi[i]);
        float p3 = 1.0f;
        for (int j = 0; j < NGLL; ++j)
          if (j != dgr && j != i)
            p3 = p3 * (xi - q.xi[j]);
        sum = sum + p3;
      }
      const float inv = 1.0f / p2;
      h[dgr] = p1 * inv;
      hp[dgr] = sum * inv;
    }
  };

  // ---- sources: t0 rule and Ricker table as in 2D (read_sources.cpp,
  // ricker.cpp:55-66); arrays from vector / tensor (compute_source_array/
  // dim3/impl/compute_source_array_from_{vector,tensor}.cpp)
  A.nsrc = src.nsrc;
  A.nstep = nstep;
  A.src_slot.assign(std::max(1, src.nsrc), 0);
  A.src_array.assign(size_t(std::max(1, src.nsrc)) * 3 * NPT3, 0.0f);
  A.stf.assign(size_t(nstep) * std::max(1, src.nsrc), 0.0f);
  {
    std::vector<float> t0s(src.nsrc);
    float min_t0 = std::numeric_limits<float>::max();
    for (int s = 0; s < src.nsrc; ++s) {
      if (!(src.f0[s] > 0))
        throw SmxError(SMX_ERR_ARG, "source f0 must be positive");
      if (src.kind[s] != 0 && src.kind[s] != 1)
        throw SmxError(SMX_ERR_ARG, "source kind must be 0 (force) or 1 (moment tensor)");
      const float hdur = 1.0 / src.f0[s];
      t0s[s] = -static_cast<float>(1.2) * hdur + src.tshift[s];
      min_t0 = std::min(min_t0, t0s[s]);
    }
    A.t0 = src.nsrc ? min_t0 : 0.0;
    const float pi = static_cast<float>(M_PI);
    for (int s = 0; s < src.nsrc; ++s) {
      const float ts = t0s[s] - min_t0, f0 = src.f0[s];
      for (int it = 0; it < nstep; ++it) {
        const float t = min_t0 + it * dt;
        const float tt = t - ts;
        const float a = pi * pi * f0 * f0;
        const float d2 = (1.0 - 2.0 * a * tt * tt) * std::exp(-a * tt * tt);
        A.stf[size_t(it) * src.nsrc + s] = src.factor[s] * d2;
      }
      int e;
      float lc[3];
      locate(src.x[s], src.y[s], src.z[s], e, lc);
      A.src_slot[s] = A.elem_slot[e];
      float h[3][NGLL], hp[3][NGLL];
      for (int d = 0; d < 3; ++d)
        lagrange(lc[d], h[d], hp[d]);
      const float *c6 = src.comp + 6 * size_t(s);
      float *L = &A.src_array[size_t(s) * 3 * NPT3];
      if (src.kind[s] == 0) {
        for (int iz = 0; iz < NGLL; ++iz)
          for (int iy = 0; iy < NGLL; ++iy)
            for (int ix = 0; ix < NGLL; ++ix) {
              const float hl = h[0][ix] * h[1][iy] * h[2][iz];
              for (int c = 0; c < 3; ++c)
                L[size_t(c) * NPT3 + pidx(iz, iy, ix)] = hl * c6[c];
            }
      } else {
        float J[10] = {0};
        for (int iz = 0; iz < NGLL; ++iz)
          for (int iy = 0; iy < NGLL; ++iy)
            for (int ix = 0; ix < NGLL; ++ix) {
              const float hl = h[0][ix] * h[1][iy] * h[2][iz];
              const G3 g = jac3(&cx[e * 8], &cy[e * 8], &cz[e * 8],
                                dN[pidx(iz, iy, ix)]);
              for (int k = 0; k < 10; ++k)
                J[k] += hl * g.t[k];
            }
        const float M[3][3] = {{c6[0], c6[3], c6[4]},
                               {c6[3], c6[1], c6[5]},
                               {c6[4], c6[5], c6[2]}};
        for (int iz = 0; iz < NGLL; ++iz)
          for (int iy = 0; iy < NGLL; ++iy)
            for (int ix = 0; ix < NGLL; ++ix) {
              float dd[3];
              for (int j = 0; j < 3; ++j)
                dd[j] = (hp[0][ix] * J[j]) * h[1][iy] * h[2][iz] +
                        h[0][ix] * (hp[1][iy] * J[3 + j]) * h[2][iz] +
                        h[0][ix] * h[1][iy] * (hp[2][iz] * J[6 + j]);
              for (int c = 0; c < 3; ++c)
                L[size_t(c) * NPT3 + pidx(iz, iy, ix)] =
                    M[c][0] * dd[0] + M[c][1] * dd[1] + M[c][2] * dd[2];
            }
      }
    }
  }

  // ---- per-tile source entries (every point of each source element, in
  // source order)
  {
    std::vector<std::vector<int>> of_tile(A.ntiles);
    for (int s2 = 0; s2 < src.nsrc; ++s2)
      of_tile[A.src_slot[s2] / TE].push_back(s2);
    A.tsrc_li.clear();
    A.tsrc_id.clear();
    A.tsrc_val.clear();
    for (int t = 0; t < A.ntiles; ++t) {
      TileInfo3 &T = A.tiles[t];
      T.src_off = int(A.tsrc_id.size());
      for (int s2 : of_tile[t])
        for (int p = 0; p < NPT3; ++p) {
          A.tsrc_li.push_back(A.lidx[size_t(A.src_slot[s2]) * NPT3 + p]);
          A.tsrc_id.push_back(s2);
          for (int c = 0; c < 3; ++c)
            A.tsrc_val.push_back(A.src_array[(size_t(s2) * 3 + c) * NPT3 + p]);
        }
      T.nsrc = int(A.tsrc_id.size()) - T.src_off;
    }
  }

  validate_tiles(A);

  // ---- receivers (core/specfem/assembly/receivers/dim3/receivers.cpp)
  A.nrec = rec.nrec;
  A.nsbs = rec.nstep_between_samples;
  A.nsig = (nstep + A.nsbs - 1) / A.nsbs;
  A.seis_types.assign(rec.types, rec.types + rec.ntypes);
  for (int t : A.seis_types)
    if (t < 0 || t > 2)
      throw SmxError(SMX_ERR_ARG, "unsupported seismogram type");
  if (A.seis_types.size() > 4)
    throw SmxError(SMX_ERR_ARG, "at most 4 seismogram types");
  A.rec_g.assign(size_t(std::max(1, rec.nrec)) * NPT3, 0);
  A.rec_h.assign(A.rec_g.size(), 0.0f);
  for (int r = 0; r < rec.nrec; ++r) {
    int e;
    float lc[3];
    locate(rec.x[r], rec.y[r], rec.z[r], e, lc);
    float h[3][NGLL], hp[3][NGLL];
    for (int d = 0; d < 3; ++d)
      lagrange(lc[d], h[d], hp[d]);
    const int s = A.elem_slot[e];
    for (int iz = 0; iz < NGLL; ++iz)
      for (int iy = 0; iy < NGLL; ++iy)
        for (int ix = 0; ix < NGLL; ++ix) {
          const int p = pidx(iz, iy, ix);
          A.rec_g[size_t(r) * NPT3 + p] = A.ibool[size_t(s) * NPT3 + p];
          A.rec_h[size_t(r) * NPT3 + p] = h[0][ix] * h[1][iy] * h[2][iz];
        }
  }
}

} // namespace smx
