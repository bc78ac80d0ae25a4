// Header-only C++ classes with the names and signatures of SPECFEM++'s
// time-stepping operators, forwarding to the C ABI of specfempp_amd.h.  A
// solver written against the reference's operator surface switches by
// constructing these from an smx_ctx instead of from an assembly.
//
//   specfem::kokkos_kernels::domain_kernels  (include/kokkos_kernels/domain_kernels.hpp:37-213)
//   specfem::time_scheme::newmark<..., forward>  (core/specfem/timescheme/newmark.hpp:106-231)
//
// Errors throw std::runtime_error with smx_last_error(), as the reference's
// operators throw on invalid state.
#pragma once

#include <stdexcept>
#include <string>

#include "specfempp_amd.h"

namespace specfempp_amd {

inline void check(int status) {
  if (status != SMX_OK)
    throw std::runtime_error(std::string("specfempp_amd: ") + smx_last_error());
}

// domain_kernels<forward, dim2, 5> for the elastic P-SV medium.
class domain_kernels {
public:
  explicit domain_kernels(smx_ctx *ctx, int field = SMX_FIELD_FORWARD)
      : ctx_(ctx), field_(field) {}

  // Mass matrix (+ Stacey term) and its inverse.  dt is fixed when the
  // context is created (the Stacey mass term uses it).
  void initialize(float /*dt*/) { check(smx_initialize(ctx_)); }

  // Sources, stiffness, divide by the mass matrix; returns elements updated.
  int update_wavefields(int istep) {
    int n = 0;
    check(smx_update_wavefields(ctx_, field_, istep, &n));
    return n;
  }

  void compute_seismograms(int isig_step) {
    check(smx_compute_seismograms(ctx_, field_, isig_step));
  }

private:
  smx_ctx *ctx_;
  int field_;
};

// newmark<AssemblyFields, forward>: predictor / corrector return the degrees
// of freedom updated, as the reference's do.
class newmark {
public:
  newmark(smx_ctx *ctx, int nstep, int nstep_between_samples, float dt)
      : ctx_(ctx), nstep_(nstep), nsbs_(nstep_between_samples), dt_(dt) {}

  int apply_predictor_phase_forward() {
    int n = 0;
    check(smx_predictor(ctx_, SMX_FIELD_FORWARD, &n));
    return n;
  }
  int apply_corrector_phase_forward() {
    int n = 0;
    check(smx_corrector(ctx_, SMX_FIELD_FORWARD, &n));
    return n;
  }
  bool compute_seismogram(int istep) const { return istep % nsbs_ == 0; }
  int get_seismogram_step() const { return isig_; }
  void increment_seismogram_step() { ++isig_; }
  float get_timestep() const { return dt_; }
  int get_max_timestep() const { return nstep_; }

private:
  smx_ctx *ctx_;
  int nstep_, nsbs_;
  float dt_;
  int isig_ = 0;
};

} // namespace specfempp_amd
