// Header-only C++ classes with the names and signatures of SPECFEM++'s
// time-stepping operators, forwarding to the C ABI of specfempp_amd.h.  A
// solver written against the reference's operator surface switches by
// constructing these from an smx_ctx instead of from an assembly:
//
//   specfem::element::medium_tag            include/enumerations/medium.hpp:18-28
//   specfem::wavefield::simulation_field    include/enumerations/wavefield.hpp
//   specfem::kokkos_kernels::domain_kernels include/kokkos_kernels/domain_kernels.hpp:37-213
//   specfem::time_scheme::time_scheme       core/specfem/timescheme/timescheme.hpp:16-201
//   specfem::time_scheme::newmark<..., forward>
//                                           core/specfem/timescheme/newmark.hpp:106-231
//   specfem::solver::time_marching<forward, dim2, NGLL>::run
//                                           core/specfem/solver/time_marching.tpp:8-131
//   specfem::solver::time_marching<combined, dim2, NGLL>::run
//                                           core/specfem/solver/time_marching.tpp:133-264
//   specfem::kokkos_kernels::frechet_kernels
//                                           include/kokkos_kernels/frechet_kernels.hpp:19-61
//
// The combined (adjoint) simulation is replaced as a whole: the reference
// reconstructs the forward field backwards next to the adjoint field and
// accumulates the kernels step by step in frechet_kernels::compute_derivatives;
// this library keeps forward checkpoints, recomputes each segment and runs the
// exact discrete adjoint, with the kernel accumulation fused into each adjoint
// step.  So the gradient's drop-in unit is time_marching<combined>::run, and
// frechet_kernels here is the holder of the accumulated kernels of that run.
//
// Errors throw std::runtime_error with smx_last_error(), as the reference's
// operators throw on invalid state.
#pragma once

#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "specfempp_amd.h"

namespace specfempp_amd {

inline void check(int status) {
  if (status != SMX_OK)
    throw std::runtime_error(std::string("specfempp_amd: ") + smx_last_error());
}

// specfem::element::medium_tag, same order and names.
enum class medium_tag : int {
  elastic_psv = SMX_MEDIUM_ELASTIC_PSV,
  elastic_sh = SMX_MEDIUM_ELASTIC_SH,
  elastic_psv_t = SMX_MEDIUM_ELASTIC_PSV_T,
  acoustic = SMX_MEDIUM_ACOUSTIC,
  poroelastic = SMX_MEDIUM_POROELASTIC,
  electromagnetic_te = SMX_MEDIUM_ELECTROMAGNETIC_TE,
  elastic = SMX_MEDIUM_ELASTIC,
  elastic_spin = SMX_MEDIUM_ELASTIC_SPIN,
  electromagnetic = SMX_MEDIUM_ELECTROMAGNETIC
};

// specfem::wavefield::simulation_field (the fields an operator acts on).
enum class simulation_field : int {
  forward = SMX_FIELD_FORWARD,
  adjoint = SMX_FIELD_ADJOINT,
  backward = SMX_FIELD_BACKWARD
};

// domain_kernels<WavefieldType, dim2, NGLL>: the operators of one field.
template <simulation_field WavefieldType, int NGLL = 5> class domain_kernels {
  static_assert(NGLL == 5, "the kernels are built for NGLL = 5 (GLL4)");

public:
  static constexpr simulation_field wavefield = WavefieldType;
  static constexpr int ngll = NGLL;

  explicit domain_kernels(smx_ctx *ctx) : ctx_(ctx) {}

  // Mass matrix (+ Stacey term) and its inverse.  dt is fixed when the
  // context is created (the Stacey mass term uses it).
  void initialize(const float & /*dt*/) { check(smx_initialize(ctx_)); }

  // Coupling, sources, stiffness and the divide by the mass matrix of one
  // medium; returns the elements of that medium updated (0 if none).
  template <medium_tag medium> int update_wavefields(const int istep) {
    int n = 0;
    check(smx_update_wavefields_medium(ctx_, int(WavefieldType), int(medium),
                                       istep, &n));
    return n;
  }

  void compute_seismograms(const int &isig_step) {
    check(smx_compute_seismograms(ctx_, int(WavefieldType), isig_step));
  }

private:
  smx_ctx *ctx_;
};

namespace impl {
// One (istep, dt) pair of the time loop (timescheme/impl/time_iterators.hpp).
struct time_step {
  int istep;
  float dt;
};

// Steps start, start + inc, ... up to (excluding) end.
class step_range {
public:
  class iterator {
  public:
    using iterator_category = std::input_iterator_tag;
    using value_type = time_step;
    using difference_type = int;
    using pointer = const time_step *;
    using reference = time_step;
    iterator(int istep, int inc, float dt) : istep_(istep), inc_(inc), dt_(dt) {}
    time_step operator*() const { return {istep_, dt_}; }
    iterator &operator++() {
      istep_ += inc_;
      return *this;
    }
    bool operator!=(const iterator &o) const { return istep_ != o.istep_; }
    bool operator==(const iterator &o) const { return istep_ == o.istep_; }

  private:
    int istep_, inc_;
    float dt_;
  };
  step_range(int start, int end, int inc, float dt)
      : start_(start), end_(end), inc_(inc), dt_(dt) {}
  iterator begin() const { return iterator(start_, inc_, dt_); }
  iterator end() const { return iterator(end_, inc_, dt_); }

private:
  int start_, end_, inc_;
  float dt_;
};
} // namespace impl

// time_scheme: step iteration and seismogram cadence; the phases are the
// derived scheme's.
class time_scheme {
public:
  time_scheme(const int nstep, const int nstep_between_samples, const float dt)
      : nstep_(nstep), nsbs_(nstep_between_samples), dt_(dt) {}
  virtual ~time_scheme() = default;

  impl::step_range iterate_forward() const { return {0, nstep_, 1, dt_}; }
  impl::step_range iterate_backward() const { return {nstep_ - 1, -1, -1, dt_}; }

  int get_max_timestep() const { return nstep_; }
  void increment_seismogram_step() { ++isig_; }
  bool compute_seismogram(const int istep) const { return istep % nsbs_ == 0; }
  int get_seismogram_step() const { return isig_; }
  int get_max_seismogram_step() const { return nstep_ / nsbs_; }
  int get_nstep_between_samples() const { return nsbs_; }

  virtual int apply_predictor_phase_forward(const medium_tag tag) = 0;
  virtual int apply_corrector_phase_forward(const medium_tag tag) = 0;
  virtual int apply_predictor_phase_backward(const medium_tag tag) = 0;
  virtual int apply_corrector_phase_backward(const medium_tag tag) = 0;
  virtual float get_timestep() const = 0;

private:
  int nstep_, nsbs_;
  float dt_;
  int isig_ = 0;
};

// newmark<AssemblyFields, forward>: predictor / corrector of one medium
// return the degrees of freedom updated (0 for a medium the mesh does not
// hold); a forward scheme updates no backward field (newmark.hpp:131-135).
class newmark : public time_scheme {
public:
  newmark(smx_ctx *ctx, const int nstep, const int nstep_between_samples,
          const float dt)
      : time_scheme(nstep, nstep_between_samples, dt), ctx_(ctx), dt_(dt) {}

  int apply_predictor_phase_forward(const medium_tag tag) override {
    int n = 0;
    check(smx_predictor_medium(ctx_, SMX_FIELD_FORWARD, int(tag), &n));
    return n;
  }
  int apply_corrector_phase_forward(const medium_tag tag) override {
    int n = 0;
    check(smx_corrector_medium(ctx_, SMX_FIELD_FORWARD, int(tag), &n));
    return n;
  }
  int apply_predictor_phase_backward(const medium_tag) override { return 0; }
  int apply_corrector_phase_backward(const medium_tag) override { return 0; }
  float get_timestep() const override { return dt_; }

private:
  smx_ctx *ctx_;
  float dt_;
};

// The total degrees of freedom the loop checks its counts against
// (assembly::get_total_degrees_of_freedom).
inline int total_degrees_of_freedom(smx_ctx *ctx) {
  long long n = 0;
  check(smx_ctx_degrees_of_freedom(ctx, &n));
  return int(n);
}

inline int total_number_of_elements(smx_ctx *ctx) {
  int nspec = 0;
  check(smx_ctx_info(ctx, &nspec, nullptr, nullptr, nullptr));
  return nspec;
}

// time_marching<forward, dim2, NGLL>: the forward loop, operator by operator
// in the reference's order, with its degree-of-freedom and element checks.
template <int NGLL = 5> class time_marching_forward {
public:
  time_marching_forward(smx_ctx *ctx, time_scheme &scheme,
                        domain_kernels<simulation_field::forward, NGLL> &kernels)
      : ctx_(ctx), time_scheme_(&scheme), kernels(kernels) {}

  void run() {
    constexpr auto acoustic = medium_tag::acoustic;
    constexpr auto elastic = medium_tag::elastic;
    constexpr auto elastic_psv = medium_tag::elastic_psv;
    constexpr auto elastic_sh = medium_tag::elastic_sh;
    constexpr auto poroelastic = medium_tag::poroelastic;
    constexpr auto elastic_psv_t = medium_tag::elastic_psv_t;
    time_scheme *ts = time_scheme_;
    kernels.initialize(ts->get_timestep());
    const int total_dofs = 2 * total_degrees_of_freedom(ctx_);
    const int total_elements = total_number_of_elements(ctx_);
    for (const auto step : ts->iterate_forward()) {
      const int istep = step.istep;
      int dofs = 0, elements = 0;
      dofs += ts->apply_predictor_phase_forward(acoustic);
      dofs += ts->apply_predictor_phase_forward(elastic);
      dofs += ts->apply_predictor_phase_forward(elastic_psv);
      dofs += ts->apply_predictor_phase_forward(elastic_sh);
      dofs += ts->apply_predictor_phase_forward(poroelastic);
      dofs += ts->apply_predictor_phase_forward(elastic_psv_t);
      elements += kernels.template update_wavefields<acoustic>(istep);
      dofs += ts->apply_corrector_phase_forward(acoustic);
      elements += kernels.template update_wavefields<elastic>(istep);
      elements += kernels.template update_wavefields<elastic_psv>(istep);
      elements += kernels.template update_wavefields<elastic_sh>(istep);
      elements += kernels.template update_wavefields<elastic_psv_t>(istep);
      dofs += ts->apply_corrector_phase_forward(elastic);
      dofs += ts->apply_corrector_phase_forward(elastic_psv);
      dofs += ts->apply_corrector_phase_forward(elastic_sh);
      dofs += ts->apply_corrector_phase_forward(elastic_psv_t);
      elements += kernels.template update_wavefields<poroelastic>(istep);
      dofs += ts->apply_corrector_phase_forward(poroelastic);
      if (ts->compute_seismogram(istep)) {
        kernels.compute_seismograms(ts->get_seismogram_step());
        ts->increment_seismogram_step();
      }
      if (dofs != total_dofs)
        throw std::runtime_error(
            "The time loop has not updated all the degrees of freedom. Only " +
            std::to_string(dofs) + " out of " + std::to_string(total_dofs) +
            " degrees of freedom have been updated.");
      if (elements != total_elements)
        throw std::runtime_error(
            "The time loop has not updated all the elements. Only " +
            std::to_string(elements) + " out of " +
            std::to_string(total_elements) + " elements have been updated.");
    }
    check(smx_synchronize(ctx_));
  }

private:
  smx_ctx *ctx_;
  time_scheme *time_scheme_;
  domain_kernels<simulation_field::forward, NGLL> &kernels;
};

// frechet_kernels<dim2, NGLL>: the kappa, mu and rho kernels of the last
// combined run, [3][nspec][5][5] in double (d chi / d parameter at every
// GLL point).
template <int NGLL = 5> class frechet_kernels {
public:
  explicit frechet_kernels(smx_ctx *ctx) : ctx_(ctx) {}
  std::vector<double> derivatives() const {
    int nspec = 0;
    check(smx_ctx_info(ctx_, &nspec, nullptr, nullptr, nullptr));
    std::vector<double> g(size_t(3) * nspec * NGLL * NGLL);
    check(smx_get_gradient(ctx_, g.data()));
    return g;
  }

private:
  smx_ctx *ctx_;
};

// time_marching<combined, dim2, NGLL>: the adjoint simulation for the
// adjoint source dchi/ds = adjoint_source[nstep][nrec][2] (the context's
// precision), as one run: forward with checkpoints every `every` steps, then
// the segments recomputed newest first with the adjoint steps and the kernel
// accumulation on the device.
template <int NGLL = 5> class time_marching_combined {
public:
  time_marching_combined(smx_ctx *ctx, const void *adjoint_source, int every)
      : ctx_(ctx), source_(adjoint_source), every_(every) {}
  void run() {
    check(smx_run_forward_checkpointed(ctx_, every_));
    check(smx_run_adjoint(ctx_, source_));
  }
  frechet_kernels<NGLL> kernels() const { return frechet_kernels<NGLL>(ctx_); }

private:
  smx_ctx *ctx_;
  const void *source_;
  int every_;
};

} // namespace specfempp_amd
