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
// Device layout.  Elements are grouped by colour (no two elements of a
// colour share a GLL point; 8 colours on a structured hex box), each colour
// contiguous and in Morton order of the element centroids; slot = position
// in that order.  Per-element arrays are element-major, [slot][term][point]
// (point p = (iz*5 + iy)*5 + ix), so that the 125 threads of an element's
// workgroup read contiguous bytes.  Global points are numbered by first touch
// in Morton element order.  Fields are three component planes
// [3][nglob_pad].
#pragma once

#include <cstdint>
#include <vector>

#include "assembly2d.hpp" // SmxError, Quadrature, NGLL
#include "specfempp_amd.h"

namespace smx {

constexpr int NPT3 = NGLL * NGLL * NGLL;

struct HostAssembly3D {
  int nspec = 0, nglob = 0;
  Quadrature quad;
  // colours: slots [color_off[c], color_off[c+1])
  int ncolors = 0;
  std::vector<int32_t> color_off;
  std::vector<int32_t> elem_slot; // mesh element -> slot
  std::vector<int32_t> ibool;     // [nspec][NPT3] global ids
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
                HostAssembly3D &out);

// Replaces the material ([nspec][125] each, mesh element order) of an
// assembled problem: moduli, Stacey impedances and mass.  A is unchanged
// when the model is rejected.
void set_model3d(HostAssembly3D &A, const float *kappa, const float *mu,
                 const float *rho);

} // namespace smx
