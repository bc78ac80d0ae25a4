// Host assembly of coupled elastic-acoustic 2D meshes.
//
// The reference keeps one field per medium and couples them on conforming
// interfaces (core/specfem/assembly/conforming_interfaces/dim2, include/
// kokkos_kernels/impl/compute_coupling.tpp:20-95).  Here the mesh is split by
// medium, each part is assembled as a single-medium mesh (its own numbering,
// tiles and boundaries), and every GLL point of an interface edge gets an
// entry on each side:
//   acoustic side (self = fluid element):  chi_dd += w (n . u)      (n, w of the
//     fluid element's edge; none at acoustic free-surface points)
//   elastic side  (self = solid element):  a += (w n) chi_dd        (n, w of the
//     solid element's edge)
// with n = compute_normal(edge) of the self element (not unit length) and w
// the GLL weight along the edge (interface_container.tpp:40-75).
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <map>
#include <string>

#include "coupled2d.hpp"

namespace smx {

namespace {

bool element_is_fluid(const smx_mesh2d &m, int e) {
  for (int p = 0; p < NPT; ++p)
    if (m.mu[size_t(e) * NPT + p] != 0.0f)
      return false;
  return true;
}

// Point in the element's corner quadrilateral (counter-clockwise corners
// 0..3), with a tolerance relative to the element's size.
bool in_quad(const smx_mesh2d &m, int e, double x, double z) {
  double cx[4], cz[4];
  for (int c = 0; c < 4; ++c) {
    const int n = m.knods[size_t(e) * m.ngnod + c];
    cx[c] = m.coord[2 * size_t(n)];
    cz[c] = m.coord[2 * size_t(n) + 1];
  }
  const double size = std::fabs(cx[1] - cx[0]) + std::fabs(cz[1] - cz[0]) +
                      std::fabs(cx[3] - cx[0]) + std::fabs(cz[3] - cz[0]);
  const double tol = 1e-9 * size * size;
  for (int c = 0; c < 4; ++c) {
    const int d = (c + 1) % 4;
    const double cross =
        (cx[d] - cx[c]) * (z - cz[c]) - (cz[d] - cz[c]) * (x - cx[c]);
    if (cross < -tol)
      return false;
  }
  return true;
}

struct Part {
  std::vector<int> elems; // full-mesh element ids
  std::vector<int> knods, abs_e, abs_t, fr_e, fr_t;
  std::vector<float> kappa, mu, rho;
  std::vector<double> sx, sz, rx, rz;
  std::vector<float> ang, f0, tsh, fac, mom;
  std::vector<int> kind, skind;
};

} // namespace

bool mesh_is_coupled2d(const smx_mesh2d &m) {
  if (m.nspec <= 0 || !m.mu)
    return false;
  bool fluid = false, solid = false;
  for (int e = 0; e < m.nspec && !(fluid && solid); ++e)
    (element_is_fluid(m, e) ? fluid : solid) = true;
  return fluid && solid;
}

void assemble_coupled2d(const smx_mesh2d &m, const smx_sources2d &src,
                        const smx_receivers2d &rec, double dt, int nstep,
                        const AssemblyOptions &opt, CoupledAssembly2D &out) {
  if (!m.coord || !m.knods || !m.kappa || !m.mu || !m.rho ||
      (m.ngnod != 4 && m.ngnod != 9))
    throw SmxError(SMX_ERR_ARG, "invalid mesh description");
  const int nspec = m.nspec, ngnod = m.ngnod;
  std::vector<int> medium(nspec), sub(nspec);
  Part part[2];
  for (int e = 0; e < nspec; ++e) {
    const int md = element_is_fluid(m, e) ? 1 : 0;
    medium[e] = md;
    sub[e] = static_cast<int>(part[md].elems.size());
    Part &P = part[md];
    P.elems.push_back(e);
    for (int n = 0; n < ngnod; ++n)
      P.knods.push_back(m.knods[size_t(e) * ngnod + n]);
    for (int p = 0; p < NPT; ++p) {
      P.kappa.push_back(m.kappa[size_t(e) * NPT + p]);
      P.mu.push_back(m.mu[size_t(e) * NPT + p]);
      P.rho.push_back(m.rho[size_t(e) * NPT + p]);
    }
  }
  for (int i = 0; i < m.nabs; ++i) {
    const int e = m.abs_ispec[i];
    if (e < 0 || e >= nspec)
      throw SmxError(SMX_ERR_ARG, "invalid absorbing edge");
    part[medium[e]].abs_e.push_back(sub[e]);
    part[medium[e]].abs_t.push_back(m.abs_type[i]);
  }
  for (int i = 0; i < m.nfree; ++i) {
    const int e = m.free_ispec[i];
    if (e < 0 || e >= nspec)
      throw SmxError(SMX_ERR_ARG, "invalid free-surface entity");
    if (medium[e] != 1)
      throw SmxError(SMX_ERR_MESH, "acoustic free-surface entity on an elastic element");
    part[1].fr_e.push_back(sub[e]);
    part[1].fr_t.push_back(m.free_type[i]);
  }

  // medium of a point: the solid if a solid element contains it, else the
  // fluid (a point inside one medium lies outside the other's elements)
  auto medium_of = [&](double x, double z) {
    int found = -1;
    for (int e = 0; e < nspec; ++e)
      if (in_quad(m, e, x, z)) {
        if (medium[e] == 0)
          return 0;
        found = 1;
      }
    if (found < 0)
      throw SmxError(SMX_ERR_MESH, "point (" + std::to_string(x) + ", " +
                                       std::to_string(z) +
                                       ") lies outside the mesh");
    return found;
  };

  // one start time for the whole run (src/io/sources/dim2/read_sources.cpp:142-172)
  float t0_all = std::numeric_limits<float>::max();
  for (int s = 0; s < src.nsrc; ++s) {
    const int kind = src.stf ? src.stf[s] : SMX_STF_RICKER;
    t0_all = std::min(t0_all, source_start_time(kind, src.f0[s], src.tshift[s], dt));
    Part &P = part[medium_of(src.x[s], src.z[s])];
    const int sk = src.kind ? src.kind[s] : SMX_SRC_FORCE;
    if (sk == SMX_SRC_MOMENT && !src.moment) // as assemble2d
      throw SmxError(SMX_ERR_ARG, "moment-tensor source without moment components");
    P.skind.push_back(sk);
    for (int c = 0; c < 3; ++c)
      P.mom.push_back(sk == SMX_SRC_MOMENT ? src.moment[3 * s + c] : 0.0f);
    P.sx.push_back(src.x[s]);
    P.sz.push_back(src.z[s]);
    P.ang.push_back(src.angle[s]);
    P.f0.push_back(src.f0[s]);
    P.tsh.push_back(src.tshift[s]);
    P.fac.push_back(src.factor[s]);
    P.kind.push_back(kind);
  }
  if (src.nsrc == 0)
    t0_all = 0.0f;
  out.rec_med.assign(rec.nrec, 0);
  out.rec_sub.assign(rec.nrec, 0);
  for (int r = 0; r < rec.nrec; ++r) {
    const int md = medium_of(rec.x[r], rec.z[r]);
    out.rec_med[r] = md;
    out.rec_sub[r] = static_cast<int>(part[md].rx.size());
    part[md].rx.push_back(rec.x[r]);
    part[md].rz.push_back(rec.z[r]);
  }

  AssemblyOptions o = opt;
  o.has_t0 = true;
  o.t0 = t0_all;
  std::unique_ptr<HostAssembly2D> A[2];
  for (int md = 0; md < 2; ++md) {
    Part &P = part[md];
    smx_mesh2d sm{};
    sm.nspec = static_cast<int>(P.elems.size());
    sm.ngnod = ngnod;
    sm.nnodes = m.nnodes;
    sm.coord = m.coord;
    sm.knods = P.knods.data();
    sm.kappa = P.kappa.data();
    sm.mu = P.mu.data();
    sm.rho = P.rho.data();
    sm.nabs = static_cast<int>(P.abs_e.size());
    sm.abs_ispec = P.abs_e.data();
    sm.abs_type = P.abs_t.data();
    sm.nfree = static_cast<int>(P.fr_e.size());
    sm.free_ispec = P.fr_e.data();
    sm.free_type = P.fr_t.data();
    smx_sources2d ss{};
    ss.nsrc = static_cast<int>(P.sx.size());
    ss.x = P.sx.data();
    ss.z = P.sz.data();
    ss.angle = P.ang.data();
    ss.f0 = P.f0.data();
    ss.tshift = P.tsh.data();
    ss.factor = P.fac.data();
    ss.stf = P.kind.data();
    ss.kind = P.skind.data();
    ss.moment = P.mom.data();
    smx_receivers2d rr{};
    rr.nrec = static_cast<int>(P.rx.size());
    rr.x = P.rx.data();
    rr.z = P.rz.data();
    rr.ntypes = rec.ntypes;
    rr.types = rec.types;
    rr.nstep_between_samples = rec.nstep_between_samples;
    A[md] = std::make_unique<HostAssembly2D>();
    assemble2d(sm, ss, rr, dt, nstep, *A[md], o);
  }
  HostAssembly2D &E = *A[0], &F = *A[1];

  // ---- interface entries
  const int cs[5][2] = {{0, 0}, {0, 1}, {1, 2}, {3, 2}, {0, 3}};
  auto edge_point = [](int type, int k, int &iz, int &ix) {
    iz = type == 1 ? 0 : (type == 3 ? NGLL - 1 : k);
    ix = type == 4 ? 0 : (type == 2 ? NGLL - 1 : k);
  };
  struct Side {
    int e, type, start;
  };
  std::map<std::pair<int, int>, std::vector<Side>> edges;
  for (int e = 0; e < nspec; ++e)
    for (int t = 1; t <= 4; ++t) {
      const int a = m.knods[size_t(e) * ngnod + cs[t][0]],
                b = m.knods[size_t(e) * ngnod + cs[t][1]];
      edges[{std::min(a, b), std::max(a, b)}].push_back({e, t, a});
    }
  struct Ent {
    int tile, li, idx;
    float w[3];
  };
  std::vector<Ent> ent[2];
  std::map<int, int> iface_k; // acoustic global id -> interface index
  std::vector<int> iface_g;
  const Quadrature &q = E.quad;
  for (const auto &kv : edges) {
    if (kv.second.size() != 2 ||
        medium[kv.second[0].e] == medium[kv.second[1].e])
      continue;
    for (int self = 0; self < 2; ++self) {
      const Side &S = kv.second[self], &Oth = kv.second[1 - self];
      const int md = medium[S.e];
      HostAssembly2D &SA = *A[md], &OA = *A[1 - md];
      const int slot = SA.elem_slot[sub[S.e]], oslot = OA.elem_slot[sub[Oth.e]];
      const int ld = SA.nspec_pad, old = OA.nspec_pad;
      const size_t pld = size_t(NPT) * ld;
      for (int k = 0; k < NGLL; ++k) {
        int iz, ix, jz, jx;
        edge_point(S.type, k, iz, ix);
        edge_point(Oth.type, Oth.start == S.start ? k : NGLL - 1 - k, jz, jx);
        const int p = iz * NGLL + ix;
        const size_t o = size_t(p) * ld + slot;
        if (md == 1 && ((SA.dmask[slot] >> p) & 1u))
          continue; // free-surface point: the coupling term is zeroed
        const float xix = SA.geom[o], xiz = SA.geom[pld + o],
                    gax = SA.geom[2 * pld + o], gaz = SA.geom[3 * pld + o],
                    jac = SA.geom[4 * pld + o];
        float n0, n1, w;
        // core/specfem/point/jacobian_matrix.hpp:424-449 (compute_normal)
        if (S.type == 4) {
          n0 = static_cast<float>(-1.0) * xix * jac;
          n1 = static_cast<float>(-1.0) * xiz * jac;
          w = q.w[iz];
        } else if (S.type == 2) {
          n0 = xix * jac;
          n1 = xiz * jac;
          w = q.w[iz];
        } else if (S.type == 3) {
          n0 = gax * jac;
          n1 = gaz * jac;
          w = q.w[ix];
        } else {
          n0 = static_cast<float>(-1.0) * gax * jac;
          n1 = static_cast<float>(-1.0) * gaz * jac;
          w = q.w[ix];
        }
        const int g_other = OA.ibool[size_t(jz * NGLL + jx) * old + oslot];
        Ent en;
        en.tile = slot / ELEM_TILE;
        en.li = SA.lidx[o];
        if (md == 1) { // fluid self: chi_dd += w (n . u_solid)
          en.idx = g_other;
          en.w[0] = w;
          en.w[1] = n0;
          en.w[2] = n1;
        } else { // solid self: a += (w n) chi_dd
          auto it = iface_k.find(g_other);
          if (it == iface_k.end()) {
            it = iface_k.emplace(g_other, static_cast<int>(iface_g.size())).first;
            iface_g.push_back(g_other);
          }
          en.idx = it->second;
          en.w[0] = w * n0;
          en.w[1] = w * n1;
          en.w[2] = 0.0f;
        }
        ent[md].push_back(en);
      }
    }
  }
  for (int md = 0; md < 2; ++md) {
    HostAssembly2D &SA = *A[md];
    std::stable_sort(ent[md].begin(), ent[md].end(),
                     [](const Ent &a, const Ent &b) { return a.tile < b.tile; });
    for (auto &t : SA.tiles)
      t.cpl_off = t.ncpl = 0;
    for (size_t i = 0; i < ent[md].size(); ++i) {
      const Ent &en = ent[md][i];
      TileInfo &t = SA.tiles[en.tile];
      if (t.ncpl == 0)
        t.cpl_off = static_cast<int>(i);
      t.ncpl++;
      SA.cpl_li.push_back(en.li);
      SA.cpl_idx.push_back(en.idx);
      for (int c = 0; c < 3; ++c)
        SA.cpl_w.push_back(en.w[c]);
    }
  }
  // fluid interface points whose new acceleration the solid reads: tiles
  // store their exclusive ones, the shared-point pass the others
  F.n_iface = E.n_iface = static_cast<int>(iface_g.size());
  F.sh_iface.assign(std::max(1, F.nshared), -1);
  {
    std::vector<std::array<int, 3>> co; // (tile, li, k)
    for (int k = 0; k < static_cast<int>(iface_g.size()); ++k) {
      const int g = iface_g[k];
      if (g >= F.nex_total) {
        F.sh_iface[g - F.nex_total] = k;
        continue;
      }
      // tiles hold their exclusive points contiguously, in tile order
      int lo = 0, hi = F.ntiles - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (F.tiles[mid].ex_off <= g)
          lo = mid;
        else
          hi = mid - 1;
      }
      const TileInfo &t = F.tiles[lo];
      if (g < t.ex_off || g >= t.ex_off + t.nex)
        throw SmxError(SMX_ERR_MESH, "interface point without an owning tile");
      co.push_back({lo, g - t.ex_off, k});
    }
    std::stable_sort(co.begin(), co.end(),
                     [](const auto &a, const auto &b) { return a[0] < b[0]; });
    for (auto &t : F.tiles)
      t.cout_off = t.ncout = 0;
    for (size_t i = 0; i < co.size(); ++i) {
      TileInfo &t = F.tiles[co[i][0]];
      if (t.ncout == 0)
        t.cout_off = static_cast<int>(i);
      t.ncout++;
      F.cout_li.push_back(co[i][1]);
      F.cout_k.push_back(co[i][2]);
    }
  }

  // ---- the summary the C ABI reports
  HostAssembly2D &Sm = out.summary;
  Sm.medium = 2;
  Sm.nspec = nspec;
  Sm.nglob = E.nglob + F.nglob;
  Sm.nsig = E.nsig;
  Sm.nrec = rec.nrec;
  Sm.nsbs = E.nsbs;
  Sm.nstep = nstep;
  Sm.seis_types.assign(rec.types, rec.types + rec.ntypes);
  Sm.t0 = t0_all;
  Sm.ntiles = E.ntiles + F.ntiles;
  Sm.nex_total = E.nex_total + F.nex_total;
  Sm.nshared = E.nshared + F.nshared;
  Sm.sh_list.assign(E.sh_list.size() + F.sh_list.size(), 0);
  Sm.tiles = E.tiles;
  Sm.tiles.insert(Sm.tiles.end(), F.tiles.begin(), F.tiles.end());
  Sm.shot_elem_off = {0, nspec};
  Sm.shot_rec_off = {0, rec.nrec};
  Sm.shot_tile_off = {0, Sm.ntiles};
  Sm.shot_t0 = {t0_all};
  Sm.n_iface = static_cast<int>(iface_g.size());
  out.el = std::move(A[0]);
  out.ac = std::move(A[1]);
}

} // namespace smx
