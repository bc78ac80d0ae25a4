// Device-resident 2D elastic P-SV solver: the object behind an smx_ctx.
#pragma once

#include <functional>
#include <memory>

#include "assembly2d.hpp"

namespace smx {

// Precision-independent interface; Solver2D<float> and Solver2D<double>
// implement it in solver2d.hip.
class Solver2DBase {
public:
  virtual ~Solver2DBase() = default;
  virtual int precision() const = 0;
  virtual const HostAssembly2D &assembly() const = 0;

  virtual void initialize() = 0;
  virtual int predictor(int field) = 0;
  virtual int corrector(int field) = 0;
  virtual int update_wavefields(int field, int istep) = 0;
  // The medium a single-medium solver advances (SMX_MEDIUM_*); a coupled
  // driver returns -1 and dispatches the tagged operators per medium.
  virtual int medium() const = 0;
  // Medium-tagged operators (time_marching.tpp:33-84): the DOFs / elements
  // of `medium` updated, 0 when this context has no element of it.
  virtual int predictor_medium(int field, int m) {
    return m == medium() ? predictor(field) : 0;
  }
  virtual int corrector_medium(int field, int m) {
    return m == medium() ? corrector(field) : 0;
  }
  virtual int update_wavefields_medium(int field, int m, int istep) {
    return m == medium() ? update_wavefields(field, istep) : 0;
  }
  // Degrees of freedom of the context: nglob x components summed over its
  // media (assembly::get_total_degrees_of_freedom).
  virtual long total_dofs() const {
    return long(assembly().nglob) * (medium() == SMX_MEDIUM_ACOUSTIC ? 1 : 2);
  }
  // Runs fn with the solver's launches enqueued on `stream` (a coupled
  // driver orders both media's operators on its own stream).
  virtual void on_stream(void *stream, const std::function<void()> &fn) = 0;
  // update_wavefields on caller-held device fields [nglob][2] (u, v read;
  // a := (a + f - K u - B v) / M) with the shared-point partials in the
  // library's fixed device scratch array: allocates nothing.
  virtual int update_wavefields_noalloc(const void *, const void *, void *,
                                        int) {
    throw SmxError(SMX_ERR_STATE,
                   "smx_update_wavefields_noalloc: 2D elastic contexts only");
  }
  virtual void compute_seismograms(int field, int isig) = 0;
  virtual void run_forward(int istep0, int nstep) = 0;
  // Time steps per replayed graph and occupancy variant of the time-step
  // tile kernel (0 default, 6, 7, 8 waves per SIMD), for the autotuner.
  virtual void set_launch(int, int) {
    throw SmxError(SMX_ERR_STATE,
                   "smx_ctx_set_launch: 2D elastic contexts only");
  }
  virtual void synchronize() = 0;

  virtual void get_seismograms(void *host) = 0;
  virtual void get_field(int field, int which, void *host) = 0;
  virtual void set_field(int field, int which, const void *host) = 0;
  virtual void *field_ptr(int field, int which) = 0;
  virtual void last_run_stats(double *ms, double *stiffness_ms,
                              int *stiffness_launches) = 0;
  // Launches only the stiffness kernel n times into a scratch acceleration
  // and returns the mean duration in ms (HIP events on the context stream).
  virtual double time_stiffness(int n) = 0;

  // Replaces the material ([nspec][5][5] per parameter, mesh element order)
  // and rebuilds the mass matrix; earlier checkpoints are dropped.
  virtual void set_model(const float *kappa, const float *mu,
                         const float *rho) = 0;
  // set_model / get_gradient with device arrays (2D FP32 contexts): the
  // model [nspec][5][5] x3 (float) is read, and the gradient [3][nspec][5][5]
  // (double) written, in device memory; no host copy of either.
  virtual void set_model_device(const float *, const float *, const float *) {
    throw SmxError(SMX_ERR_STATE,
                   "device-memory models are built for 2D elastic FP32 contexts");
  }
  virtual void get_gradient_device(double *) {
    throw SmxError(SMX_ERR_STATE,
                   "device-memory gradients are built for 2D elastic FP32 contexts");
  }
  // Brings the host copy of the model-derived arrays up to date after a
  // device-memory model (no-op otherwise).
  virtual void sync_host_model() {}
  // Forward run of all nstep steps from rest (seismograms as run_forward),
  // keeping the state every `every` steps for the adjoint.
  virtual void run_forward_checkpointed(int every) = 0;
  // Discrete-adjoint gradient of a functional of the displacement
  // seismograms with dchi/ds = gsrc[nstep][nrec][2] (host, context
  // precision), using the checkpoints of the last checkpointed run.
  virtual void run_adjoint(const void *gsrc) = 0;
  // Gradient [3][nspec][5][5] (kappa, mu, rho) as double.
  virtual void get_gradient(double *host) = 0;
  virtual void last_adjoint_stats(double *forward_ms, double *adjoint_ms) = 0;
  // Geometry source of the tile kernels: 0 per point, 1 per element (some
  // tiles), 2 re-formed from the control nodes; mismatches as in 3D.
  virtual void geometry_mode(int *mode, long long *mismatches) const = 0;

  // Coupled meshes: one solver per medium, driven step by step on the
  // driver's HIP stream (passed as void*).  bind_coupling gives the elastic
  // displacement (acoustic solver) and the buffer of the fluid's interface
  // accelerations (both); coupling_field is this solver's displacement.
  virtual void bind_coupling(void *, void *) {
    throw SmxError(SMX_ERR_ARG, "this context does not take part in coupling");
  }
  virtual void *coupling_field() { return nullptr; }
  virtual void enqueue_begin(void *, int) {
    throw SmxError(SMX_ERR_ARG, "this context does not take part in coupling");
  }
  virtual void enqueue_step(void *, bool) {
    throw SmxError(SMX_ERR_ARG, "this context does not take part in coupling");
  }
};

// Swaps a solver's stream for the duration of a scope (coupled drivers
// enqueue a medium's kernels on their own stream).
template <typename S> struct StreamSwapT {
  S &slot;
  S saved;
  StreamSwapT(S &s, S with) : slot(s), saved(s) { slot = with; }
  ~StreamSwapT() { slot = saved; }
};

// graph_steps: time steps per replayed HIP graph; nodes_geometry: re-form
// per-point geometry from the control nodes when bitwise reproducible;
// waves_per_eu: occupancy variant of the FP32 time-step tile kernel (0 the
// default, 6, 7 or 8 waves per SIMD; elastic contexts only).
std::unique_ptr<Solver2DBase> make_solver2d(std::unique_ptr<HostAssembly2D> a,
                                            int precision, int graph_steps = 8,
                                            bool nodes_geometry = false,
                                            int waves_per_eu = 0);
// Throws SMX_ERR_ARG unless graph_steps and waves_per_eu are valid for a
// context of this precision.
void check_launch(int graph_steps, int waves_per_eu, int precision);

// The acoustic solver (acoustic2d.hip); make_solver2d dispatches to it for
// assemblies whose medium is acoustic.
std::unique_ptr<Solver2DBase>
make_acoustic_solver2d(std::unique_ptr<HostAssembly2D> a, int precision,
                       int graph_steps);

} // namespace smx
