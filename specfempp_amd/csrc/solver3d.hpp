// Device-resident 3D elastic solver: the object behind a 3D smx_ctx.
#pragma once

#include <memory>

#include "assembly3d.hpp"

namespace smx {

class Solver3DBase {
public:
  virtual ~Solver3DBase() = default;
  virtual int precision() const = 0;
  virtual const HostAssembly3D &assembly() const = 0;
  virtual void initialize() = 0;
  // Fused, graph-replayed forward steps istep0 .. istep0+nstep-1 (the last
  // one ends with the corrected velocity and acceleration stored).
  virtual void run_forward(int istep0, int nstep) = 0;
  virtual void synchronize() = 0;
  virtual void get_seismograms(void *host) = 0;   // [nsig][ntypes][nrec][3]
  virtual void get_field(int which, void *host) = 0;       // [nglob][3]
  virtual void set_field(int which, const void *host) = 0; // [nglob][3]
  virtual void last_run_stats(double *ms, double *stiffness_ms,
                              int *stiffness_launches) = 0;
  // Mean duration (ms) of one step's element passes (all colours).
  virtual double time_stiffness(int n) = 0;
  // Geometry source of the element passes (0 per point, 1 per element,
  // 2 re-formed from the control nodes) and, for 2, the number of stored
  // values the device did not reproduce at set-up (then 0 is used).
  virtual void geometry_mode(int *mode, long long *mismatches) const = 0;

  // ---- discrete adjoint (same contract as the 2D solver's)
  // Replaces the material ([nspec][125] each, mesh element order).
  virtual void set_model(const float *kappa, const float *mu,
                         const float *rho) = 0;
  // All nstep steps from rest, keeping (u~, v~) every `every` steps.
  virtual void run_forward_checkpointed(int every) = 0;
  // Adjoint run for dchi/ds = gsrc [nstep][nrec][3] (context precision), s
  // the displacement seismogram of each step; recomputes each segment from
  // its checkpoint (the forward fields are left at some segment's end).
  virtual void run_adjoint(const void *gsrc) = 0;
  // dchi / d(kappa, mu, rho) [3][nspec][125], mesh element order.
  virtual void get_gradient(double *out) = 0;
  virtual void last_adjoint_stats(double *forward_ms, double *adjoint_ms) = 0;
};

// graph_steps: time steps per replayed HIP graph; nodes_geometry: re-form
// per-point geometry from the control nodes when bitwise reproducible.
std::unique_ptr<Solver3DBase> make_solver3d(std::unique_ptr<HostAssembly3D> a,
                                            int precision, int graph_steps = 8,
                                            bool nodes_geometry = false);

} // namespace smx
