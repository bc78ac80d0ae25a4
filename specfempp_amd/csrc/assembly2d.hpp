// Host-side assembly of the 2D elastic P-SV problem into the device layout.
//
// Restates specfem::assembly::assembly<dim2> (core/specfem/assembly/assembly/
// dim2/assembly.cpp:6-93) for one medium (elastic P-SV, isotropic): GLL
// quadrature, Jacobian from the control nodes, global numbering, Stacey
// normals/weights, source arrays + source time functions and receiver
// interpolants, all in FP32 with the reference's operation order.
//
// Device layout.  Elements are grouped into tiles of up to ELEM_TILE
// elements (Morton order of the element centroids; 8x8 squares on a
// structured box); one workgroup owns one tile.  A GLL point touched by a
// single tile is "exclusive" to it; the final numbering stores each tile's
// exclusive points contiguously, tile after tile, followed by all points
// shared between tiles.  Per-element arrays are
// structure-of-arrays [point][slot] with slot = tile * ELEM_TILE + position,
// so one wavefront's 64 lanes read 64 consecutive elements.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "specfempp_amd.h"

namespace smx {

constexpr int NGLL = 5;
constexpr int NPT = NGLL * NGLL;  // GLL points per element
constexpr int ELEM_TILE = 64;     // elements per tile (= lanes of a wavefront)
constexpr int NLOC_MAX = 1600;    // GLL points of a tile (64 * 25 at most)

struct SmxError : std::runtime_error {
  int code;
  SmxError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// FP32 GLL quadrature for NGLL=5 (nodes, weights, h'(i,j) = l'_j(xi_i)).
struct Quadrature {
  float xi[NGLL];
  float w[NGLL];
  float hprime[NGLL][NGLL];
  Quadrature();
};

// One tile (layout shared with the device code).
struct TileInfo {
  int32_t ecount;  // elements in the tile
  int32_t ex_off;  // first exclusive point (global id)
  int32_t nex;     // exclusive points (local ids 0..nex-1)
  int32_t sh_off;  // first entry in sh_list
  int32_t nsh;     // shared points (local ids nex..nex+nsh-1)
  int32_t src_off; // source entries
  int32_t nsrc;
  int32_t rec_off; // receiver entries on exclusive points
  int32_t nrec;
  int32_t flags;   // TILE_GEOM_ELEM | TILE_MAT_ELEM
  int32_t cpl_off; // coupling entries (coupled meshes; 0 otherwise)
  int32_t ncpl;
  int32_t cout_off; // acoustic tiles: interface points whose new
  int32_t ncout;    // acceleration the elastic update reads
};

// Tile flags: every element of the tile has element-constant geometry
// (affine element: the five Jacobian terms agree at all 25 points to within
// FP32 rounding) / element-constant lambda+2mu and mu; the kernels then read
// the per-element arrays geom_e / mat_e instead of the per-point ones.
constexpr int32_t TILE_GEOM_ELEM = 1;
constexpr int32_t TILE_MAT_ELEM = 2;
// The tile re-forms its per-point geometry from the elements' four control
// nodes (set by the solver after a bitwise check on the device).
constexpr int32_t TILE_GEOM_NODES = 4;

struct HostAssembly2D {
  int nspec = 0, nspec_pad = 0, nglob = 0;
  int nex_total = 0, nshared = 0; // exclusive points are ids [0, nex_total)
  Quadrature quad;

  // Tiles in launch order (one workgroup each, all in one launch).
  int ntiles = 0, nloc_max = 0;
  std::vector<TileInfo> tiles;
  // Global ids of each tile's shared points; a tile writes its partial sum
  // for sh_list[k] to slot k of the partials buffer.
  std::vector<int32_t> sh_list;
  // Shared point j (global id nex_total + j) sums the partial slots
  // shp_pos[shp_off[j] .. shp_off[j+1]) in that (tile) order.
  std::vector<int32_t> shp_off, shp_pos;

  // Element colouring (no two elements of a colour share a point).
  int nelem_colors = 0;
  std::vector<uint8_t> ecolor;  // [nspec_pad], 0xff for padding slots
  std::vector<int> elem_slot;   // mesh element -> slot

  // Medium shared by every element: 0 elastic P-SV, 1 acoustic.  Acoustic
  // meshes reuse the elastic arrays with these meanings: lam2mu_mu =
  // (1/rho, kappa), Stacey slot 4 = 1/(rho vp) (slot 5 zero), src_L pairs =
  // (L, kappa at the point), stf pairs equal.
  int medium = 0;
  // Acoustic: 25-bit mask per slot of the points tagged acoustic free
  // surface (Dirichlet) in that element (empty for elastic meshes).
  std::vector<uint32_t> dmask;
  // Acoustic receiver entries: (wx, wz) per entry id such that the x / z
  // seismogram of a potential f is sum_q w_q f_q, i.e. the interpolated
  // rho^-1 grad f; [nrec*NPT][2] in FP32 and FP64.
  std::vector<float> rent_w;
  std::vector<double> rent_w64;

  // Coupling with the other medium of a coupled mesh (empty otherwise),
  // entries in tile order (TileInfo cpl_off / ncpl), added to the tile's
  // accumulator with the sources:
  //   elastic tile:  (li, k, w n0, w n1):   a += (w n) chi_dd[k]
  //   acoustic tile: (li, g, w, n0, n1):    chi_dd += w (n . u_elastic[g])
  // cpl_w holds 3 values per entry ((w n0, w n1, 0) or (w, n0, n1)).
  // Acoustic tiles also store the new acceleration of their exclusive
  // interface points (cout_li, cout_k; TileInfo cout_off / ncout) and the
  // shared-point pass that of shared ones (sh_iface[j] = k or -1); k indexes
  // the n_iface acoustic interface points.
  std::vector<int32_t> cpl_li, cpl_idx;
  std::vector<float> cpl_w;
  std::vector<int32_t> cout_li, cout_k, sh_iface;
  int n_iface = 0;

  std::vector<int32_t> ibool;   // [NPT][nspec_pad] global ids
  std::vector<uint16_t> lidx;   // [NPT][nspec_pad] tile-local ids
  // [5][NPT][nspec_pad]: xix, xiz, gammax, gammaz, jacobian
  std::vector<float> geom;
  // [2][NPT][nspec_pad]: lambda+2mu, mu
  std::vector<float> lam2mu_mu;
  // [3][NPT][nspec_pad]: kappa, mu, rho (model parameters)
  std::vector<float> kappa_mu_rho;
  // Per-element copies for tiles flagged element-constant: geometry
  // [5][nspec_pad] (evaluated at the element centre) and (lambda+2mu, mu)
  // [2][nspec_pad]; elem_affine marks slots whose geometry qualifies.
  std::vector<float> geom_e, mat_e;
  std::vector<uint8_t> elem_affine;
  // Control nodes of 4-node elements [8][nspec_pad] (x0..x3, z0..z3), FP32;
  // empty for 9-node meshes.
  std::vector<float> nodes;

  // Stacey: slot -> boundary slot (-1 if none); per slot [6][NPT][nst_pad]
  // (n0, n1, |n|, edge weight (0 if the point is not on an absorbing edge),
  // rho*vp, rho*vs).
  std::vector<int32_t> stacey_slot;
  int nstacey = 0, nstacey_pad = 0;
  std::vector<float> stacey;

  // Sources: element slot (src_ispec) and mesh element (src_elem), array
  // [nsrc][2][NPT], time function [nstep][nsrc][2]; per-tile entries
  // (tile-local point, source, L_x, L_z).
  int nsrc = 0, nstep = 0;
  std::vector<int32_t> src_ispec, src_elem;
  std::vector<float> src_array;
  std::vector<float> stf;
  std::vector<int32_t> src_li, src_idx;
  std::vector<float> src_L;
  double t0 = 0.0;

  // Receivers: slot, mesh element, xi, gamma, interpolant [nrec][NPT].
  int nrec = 0;
  std::vector<int32_t> rec_ispec, rec_elem;
  std::vector<float> rec_xi, rec_gamma;
  std::vector<float> rec_h;
  // Receiver entries id = r*NPT + p: global point, shared index (-1 if the
  // point is exclusive), weight; per-tile entries on exclusive points.
  std::vector<int32_t> rent_g, rent_shared;
  std::vector<float> rent_h;
  std::vector<int32_t> trec_li, trec_entry;
  std::vector<int32_t> seis_types;
  int nsbs = 1, nsig = 0;

  // Adjoint support.  Receiver points as a node list: node rin_node[j]
  // collects the entries rin_ent[rin_off[j] .. rin_off[j+1]) (ascending).
  std::vector<int32_t> rin_node, rin_off, rin_ent;
  // Receiver node of each shared point (index into rin_node, -1 if none) and
  // the receiver nodes that are exclusive points.
  std::vector<int32_t> sh_rin, rin_excl;
  // Absorbing GLL points (slot, point, boundary slot, global node), and the
  // absorbing nodes with the 2x2 boundary matrix summed over their points
  // (bsum: xx, xz, zz, in double from the FP32 Stacey arrays), so that B v at
  // node g is bsum(g) v(g).
  std::vector<int32_t> bpt_slot, bpt_p, bpt_bslot, bpt_g;
  std::vector<int32_t> bnode_g;
  std::vector<double> bsum;
  // The model-derived coefficients formed in double (for FP64 contexts):
  // lambda+2mu [NPT][nspec_pad], Stacey impedances [2][NPT][nstacey_pad]
  // (rho vp, rho vs) and the boundary matrices built from them.
  std::vector<double> lam2mu64, imp64, bsum64;

  // Time scheme constants as the reference forms them (newmark.hpp:131-135).
  float dt = 0, dt_over2 = 0, dt2_over2 = 0;

  // Batches of independent shots: shot s owns mesh elements
  // [shot_elem_off[s], shot_elem_off[s+1]), sources, receivers and tiles
  // likewise; shot_t0[s] is its time origin.  A single mesh is one shot.
  std::vector<int32_t> shot_elem_off, shot_src_off, shot_rec_off, shot_tile_off;
  std::vector<double> shot_t0;
};

// Assembly choices that change the layout, not the results' meaning.
// Start time of one source's time function before the run's shift
// (src/io/sources/dim2/read_sources.cpp:142-172): -1.2 / f0 + tshift for
// Ricker and Dirac (f0 = 1 / (10 dt)), -2 max(hdur, 5 dt) + tshift for
// Heaviside; throws on an unknown kind or a non-positive f0 / hdur.
float source_start_time(int stf_kind, float f0, float tshift, double dt);

struct AssemblyOptions {
  int tile_elems = ELEM_TILE; // elements per tile (1..64)
  // Start time of the run when the sources given are only part of it (one
  // medium of a coupled mesh): the minimum t0 over every source.
  bool has_t0 = false;
  double t0 = 0.0;
};

// Builds the assembly; throws SmxError on invalid input.
void assemble2d(const smx_mesh2d &mesh, const smx_sources2d &sources,
                const smx_receivers2d &receivers, double dt, int nstep,
                HostAssembly2D &out, const AssemblyOptions &opt = {});

// Concatenates independently assembled shots (same dt, nstep, seismogram
// types and sampling) into one assembly whose tiles are the shots' tiles in
// shot order: one launch then advances every shot.  Every shot keeps its own
// numbering order, tiling and summation order, so its results are those of
// its own context bit for bit.
void merge_assemblies(std::vector<HostAssembly2D> &shots, HostAssembly2D &out);

// Recomputes the per-element material copy and the tile flags from the
// per-point arrays (after assembly or a model change).
void refresh_compact(HostAssembly2D &A);

// Replaces the material at every GLL point (arrays [nspec][5][5] in mesh
// element order) and refreshes every quantity derived from it (lambda+2mu,
// mu, the Stacey impedances and boundary matrices).
void set_model2d(HostAssembly2D &A, const float *kappa, const float *mu,
                 const float *rho);

} // namespace smx
