// Host-side assembly of the 2D elastic P-SV problem into the device layout.
//
// Restates specfem::assembly::assembly<dim2> (core/specfem/assembly/assembly/
// dim2/assembly.cpp:6-93) for one medium (elastic P-SV, isotropic): GLL
// quadrature, Jacobian from the control nodes, global numbering, Stacey
// normals/weights, source arrays + source time functions and receiver
// interpolants, all in FP32 with the reference's operation order.  The output
// is laid out for the gfx950 kernels: element-fastest structure-of-arrays
// ([point][element], element stride padded to a multiple of 64) so that one
// wavefront's 64 lanes read 64 consecutive elements.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "specfempp_amd.h"

namespace smx {

constexpr int NGLL = 5;
constexpr int NPT = NGLL * NGLL; // GLL points per element
constexpr int ELEM_TILE = 64;    // elements per wavefront

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

struct HostAssembly2D {
  int nspec = 0, nspec_pad = 0, nglob = 0;
  Quadrature quad;

  // Element slots: elements are grouped by colour (no two elements of a
  // colour share a GLL point), each colour padded to whole wavefronts.
  // Colour c occupies slots [color_offset[c], color_offset[c] +
  // color_count[c]); elem_slot maps the mesh's element index to its slot.
  std::vector<int> color_offset, color_count, elem_slot;

  // [NPT][nspec_pad] (point p = iz*NGLL + ix), indexed by slot
  std::vector<int32_t> ibool;
  // [5][NPT][nspec_pad]: xix, xiz, gammax, gammaz, jacobian
  std::vector<float> geom;
  // [2][NPT][nspec_pad]: lambda+2mu, mu
  std::vector<float> lam2mu_mu;
  // [3][NPT][nspec_pad]: kappa, mu, rho (model parameters, kept for kernels
  // that need them outside the stiffness operator)
  std::vector<float> kappa_mu_rho;

  // Stacey: element -> boundary slot (-1 if none); per slot [6][NPT][nst_pad]
  // (n0, n1, |n|, edge weight (0 if the point is not on an absorbing edge),
  // rho*vp, rho*vs).
  std::vector<int32_t> stacey_slot;
  int nstacey = 0, nstacey_pad = 0;
  std::vector<float> stacey;

  // Sources: element slot (src_ispec) and mesh element (src_elem), array
  // [nsrc][2][NPT], time function [nstep][nsrc][2].
  int nsrc = 0, nstep = 0;
  std::vector<int32_t> src_ispec, src_elem;
  std::vector<float> src_array;
  std::vector<float> stf;
  double t0 = 0.0;

  // Receivers: element, xi, gamma, interpolant [nrec][NPT].
  int nrec = 0;
  std::vector<int32_t> rec_ispec, rec_elem; // slot, mesh element
  std::vector<float> rec_xi, rec_gamma;
  std::vector<float> rec_h;
  std::vector<int32_t> seis_types;
  int nsbs = 1, nsig = 0;

  // Time scheme constants as the reference forms them (newmark.hpp:131-135).
  float dt = 0, dt_over2 = 0, dt2_over2 = 0;
};

// Builds the assembly; throws SmxError on invalid input.
void assemble2d(const smx_mesh2d &mesh, const smx_sources2d &sources,
                const smx_receivers2d &receivers, double dt, int nstep,
                HostAssembly2D &out);

} // namespace smx
