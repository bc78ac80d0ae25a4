// Host-side assembly of the 3D elastic isotropic problem (hex8, NGLL = 5)
// into the device layout.
//
// Restates specfem::assembly::assembly<dim3> (core/specfem/assembly/assembly/
// dim3/assembly.cpp) for one elastic isotropic medium: GLL quadrature,
// Jacobian from the control nodes (core/specfem/jacobian/dim3/jacobian.cpp),
// global numbering, mass matrix, force and moment-tensor source arrays
// (compute_source_array/dim3/impl/*), receiver interpolants, plus absorbing
// faces (this build's 3D Stacey condition; the reference's 3D path has none).
//
// Device layout.  Elements are sorted along a Morton curve of their
// centroids and cut into tiles of `tile_elems` consecutive elements (4x4x2
// blocks on a structured box); inside a tile the elements are ordered by
// colour (no two elements of a colour share a GLL point; 8 colours on a
// structured box); slot = position in that order.  A tile's elements of one
// colour are processed in rounds of up to `round_width` elements.
// Per-element arrays are element-major, [slot][term][point] (point p =
// (iz*5 + iy)*5 + ix), so that the 125 threads of an element read contiguous
// bytes.  A GLL point touched by one tile only is exclusive to it; global
// points are numbered tile by tile (each tile's exclusive points
// contiguous), then the shared points.  Each tile numbers its points
// locally (exclusive first, then its shared points) and owns one partial
// slot per shared point.  Fields are three component planes [3][nglob_pad].
#pragma once

#include <cstdint>
#include <vector>

#include "assembly2d.hpp" // SmxError, Quadrature, NGLL
#include "specfempp_amd.h"

namespace smx {

constexpr int NPT3 = NGLL * NGLL * NGLL;

// One tile of the device layout (slots [slot0, slot0 + nelem)).
struct TileInfo3 {
  int32_t slot0, nelem;
  int32_t ex_off, nex;   // exclusive points: global ids ex_off .. ex_off+nex-1
  int32_t sh_off, nsh;   // shared points: sh_list / partial slots sh_off ..
  int32_t round_off, nround; // rounds: round_slots[(round_off + r) * rw + k]
  int32_t src_off, nsrc;     // source entries
};

struct AssemblyOptions3D {
  int tile_elems = 32;  // elements per tile (1 .. 64)
  int round_width = 4;  // elements processed together in a tile (1 .. 4)
};

struct HostAssembly3D {
  int nspec = 0, nglob = 0;
  Quadrature quad;
  // global colours (adjoint colour passes): slots color_slots[color_off[c]
  // .. color_off[c+1]), ascending
  int ncolors = 0;
  std::vector<int32_t> color_off, color_slots;
  std::vector<int32_t> elem_slot; // mesh element -> slot
  std::vector<int32_t> ibool;     // [nspec][NPT3] global ids
  // tiles
  int tile_elems = 32, round_width = 4, ntiles = 0;
  int nex_total = 0, nsh_total = 0, npart = 0, nloc_max = 0;
  std::vector<TileInfo3> tiles;
  std::vector<uint16_t> lidx;       // [nspec][NPT3] tile-local ids
  std::vector<int32_t> sh_list;     // [npart] global id of each partial slot
  std::vector<int32_t> sh_poff;     // [nsh_total + 1] shared point -> partials
  std::vector<int32_t> sh_pidx;     // partial slots, in tile order
  std::vector<int32_t> round_slots; // [nrounds][round_width], -1 = none
  // per-tile point sources: tile-local point, source, array value [3]
  std::vector<int32_t> tsrc_li, tsrc_id;
  std::vector<float> tsrc_val;
  // geometry: per element [nspec][10] (xix, xiy, xiz, etax, etay, etaz,
  // gammax, gammay, gammaz, J) when every element is affine (geom_elem),
  // else per point [nspec][10][NPT3]
  bool geom_elem = false;
  std::vector<float> geom;
  std::vector<float> jac_pt; // J per GLL point [nspec][NPT3] (mass, gradient)
  // control-node coordinates [nspec][24] (x0..x7, y0..y7, z0..z7) in FP32:
  // the kernels can re-form the per-point geometry from them
  std::vector<float> nodes;
  // moduli (lambda+2mu, mu): per element [nspec][2] when constant inside
  // every element (mat_elem), else per point [nspec][2][NPT3]
  bool mat_elem = false;
  std::vector<float> mat;
  std::vector<float> kappa_mu_rho; // [nspec][3][NPT3] model
  // FP64 copies formed in double (for FP64 contexts)
  std::vector<double> mat64;
  // absorbing faces: slot -> boundary slot (-1); [nst][7][NPT3] (n0, n1,
  // n2, |n|, face weight (0 off the face), rho vp, rho vs)
  std::vector<int32_t> stacey_slot;
  int nstacey = 0;
  std::vector<float> stacey;
  std::vector<double> stacey64; // impedances formed in double: [nst][2][NPT3]
  // mass matrix inverse [3][nglob] (with the Stacey term), and in double
  std::vector<float> minv;
  std::vector<double> minv64;
  // sources: slot, array [nsrc][3][NPT3], time function [nstep][nsrc]
  int nsrc = 0, nstep = 0;
  std::vector<int32_t> src_slot;
  std::vector<float> src_array, stf;
  double t0 = 0.0;
  // receivers: interpolation entries [nrec][NPT3] (global id, weight)
  int nrec = 0, nsbs = 1, nsig = 0;
  std::vector<int32_t> rec_g;
  std::vector<float> rec_h;
  std::vector<int32_t> seis_types;
  float dt = 0, dt_over2 = 0, dt2_over2 = 0;
};

void assemble3d(const smx_mesh3d &mesh, const smx_sources3d &sources,
                const smx_receivers3d &receivers, double dt, int nstep,
                HostAssembly3D &out,
                const AssemblyOptions3D &opt = AssemblyOptions3D());

// Replaces the material ([nspec][125] each, mesh element order) of an
// assembled problem: moduli, Stacey impedances and mass.  A is unchanged
// when the model is rejected.
void set_model3d(HostAssembly3D &A, const float *kappa, const float *mu,
                 const float *rho);

} // namespace smx
