// Driver of coupled elastic-acoustic 2D runs: an acoustic and an elastic
// solver advanced together on one HIP stream.  Per step (core/specfem/solver/
// time_marching.tpp:33-77): the fluid's fused pass (coupling from the solid's
// predicted displacement, sources, stiffness, divide, corrector, next
// predictor; it also stores the new accelerations of its interface points),
// then the solid's fused pass (coupling from those accelerations, sources,
// stiffness, divide, corrector, next predictor).  Steps are graph-replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "coupled2d.hpp"

namespace smx {

#define CP_HIP(call)                                                           \
  do {                                                                         \
    hipError_t err_ = (call);                                                  \
    if (err_ != hipSuccess)                                                    \
      throw SmxError(SMX_ERR_HIP, std::string(#call) + ": " +                  \
                                      hipGetErrorString(err_));                \
  } while (0)

namespace {

class CoupledSolver2D final : public Solver2DBase {
public:
  CoupledSolver2D(std::unique_ptr<CoupledAssembly2D> a, int graph_steps)
      : ca_(std::move(a)), gsteps_(graph_steps) {
    CP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    CP_HIP(hipEventCreate(&ev0_));
    CP_HIP(hipEventCreate(&ev1_));
    const int nif = std::max(1, ca_->summary.n_iface);
    CP_HIP(hipMalloc(&chi_dd_, size_t(nif) * sizeof(float)));
    CP_HIP(hipMemset(chi_dd_, 0, size_t(nif) * sizeof(float)));
    el_ = make_solver2d(std::move(ca_->el), 4, graph_steps);
    ac_ = make_solver2d(std::move(ca_->ac), 4, graph_steps);
    ac_->bind_coupling(el_->coupling_field(), chi_dd_);
    el_->bind_coupling(nullptr, chi_dd_);
  }

  ~CoupledSolver2D() override {
    if (graph_exec_)
      (void)hipGraphExecDestroy(graph_exec_);
    if (graph_)
      (void)hipGraphDestroy(graph_);
    (void)hipStreamSynchronize(stream_);
    el_.reset();
    ac_.reset();
    (void)hipFree(chi_dd_);
    (void)hipEventDestroy(ev0_);
    (void)hipEventDestroy(ev1_);
    (void)hipStreamDestroy(stream_);
  }

  int precision() const override { return 4; }
  const HostAssembly2D &assembly() const override { return ca_->summary; }

  void initialize() override {
    ac_->initialize();
    el_->initialize();
    build_graph(); // outside every run_forward call
    CP_HIP(hipGraphUpload(graph_exec_, stream_));
    CP_HIP(hipStreamSynchronize(stream_));
    initialized_ = true;
  }

  int medium() const override { return -1; }
  long total_dofs() const override {
    return el_->total_dofs() + ac_->total_dofs();
  }
  void on_stream(void *, const std::function<void()> &) override {
    throw SmxError(SMX_ERR_STATE, "a coupled driver runs on its own stream");
  }

  // The untagged operators would have to advance both media at once, which
  // the reference's loop never asks: its operators name one medium each.
  int predictor(int) override { return untagged(); }
  int corrector(int) override { return untagged(); }
  int update_wavefields(int, int) override { return untagged(); }

  // Medium-tagged operators (time_marching.tpp:33-84): each runs on the
  // medium's solver, enqueued on this driver's stream, so the reference's
  // order predictor(acoustic), predictor(elastic_psv), update<acoustic>,
  // corrector(acoustic), update<elastic_psv>, corrector(elastic_psv) holds
  // the coupling: the fluid update reads the solid's predicted
  // displacement, the solid update the fluid's new acceleration on the
  // interface.
  int predictor_medium(int f, int m) override {
    return on_medium(f, m, [&](Solver2DBase &s) { return s.predictor(f); });
  }
  int corrector_medium(int f, int m) override {
    return on_medium(f, m, [&](Solver2DBase &s) { return s.corrector(f); });
  }
  int update_wavefields_medium(int f, int m, int istep) override {
    if (!initialized_)
      throw SmxError(SMX_ERR_STATE, "smx_initialize has not been called");
    return on_medium(f, m,
                     [&](Solver2DBase &s) { return s.update_wavefields(f, istep); });
  }
  // Both media's receivers (compute_seismogram.tpp:19-171 runs per medium).
  void compute_seismograms(int f, int isig) override {
    forward_only(f);
    for (Solver2DBase *s : {ac_.get(), el_.get()})
      if (s->assembly().nrec > 0)
        s->on_stream(stream_, [&] { s->compute_seismograms(f, isig); });
  }

  void run_forward(int istep0, int nstep) override {
    if (!initialized_)
      throw SmxError(SMX_ERR_STATE, "smx_initialize has not been called");
    const HostAssembly2D &A = ca_->summary;
    if (istep0 < 0 || nstep < 1 || istep0 + nstep > A.nstep)
      throw SmxError(SMX_ERR_ARG, "steps outside [0, nstep)");
    CP_HIP(hipEventRecord(ev0_, stream_));
    ac_->enqueue_begin(stream_, istep0);
    el_->enqueue_begin(stream_, istep0);
    const int body = nstep - 1;
    const int reps = body / gsteps_;
    if (reps > 0) {
      build_graph();
      for (int r = 0; r < reps; ++r)
        CP_HIP(hipGraphLaunch(graph_exec_, stream_));
    }
    for (int s = reps * gsteps_; s < body; ++s)
      step(false);
    step(true);
    CP_HIP(hipEventRecord(ev1_, stream_));
    launches_ = nstep;
    have_stats_ = true;
  }

  void synchronize() override { CP_HIP(hipStreamSynchronize(stream_)); }

  // [nsig][ntypes][nrec][2] in station order, gathered from both media
  void get_seismograms(void *host) override {
    synchronize();
    const HostAssembly2D &A = ca_->summary;
    const int nt = int(A.seis_types.size()), ns = A.nsig;
    const int nr[2] = {el_->assembly().nrec, ac_->assembly().nrec};
    std::vector<float> part[2];
    Solver2DBase *s[2] = {el_.get(), ac_.get()};
    for (int m = 0; m < 2; ++m) {
      part[m].assign(size_t(ns) * nt * std::max(1, nr[m]) * 2, 0.0f);
      if (nr[m] > 0)
        s[m]->get_seismograms(part[m].data());
    }
    float *out = static_cast<float *>(host);
    for (int i = 0; i < ns; ++i)
      for (int t = 0; t < nt; ++t)
        for (int r = 0; r < A.nrec; ++r) {
          const int m = ca_->rec_med[r], k = ca_->rec_sub[r];
          for (int c = 0; c < 2; ++c)
            out[((size_t(i) * nt + t) * A.nrec + r) * 2 + c] =
                part[m][((size_t(i) * nt + t) * nr[m] + k) * 2 + c];
        }
  }

  void get_field(int, int, void *) override { no_fields(); }
  void set_field(int, int, const void *) override { no_fields(); }
  void *field_ptr(int, int) override {
    no_fields();
    return nullptr;
  }

  void last_run_stats(double *ms, double *stiff_ms, int *launches) override {
    float t = 0;
    if (have_stats_) {
      CP_HIP(hipEventSynchronize(ev1_));
      CP_HIP(hipEventElapsedTime(&t, ev0_, ev1_));
    }
    if (ms)
      *ms = t;
    if (stiff_ms)
      *stiff_ms = 0;
    if (launches)
      *launches = launches_;
  }
  double time_stiffness(int) override {
    throw SmxError(SMX_ERR_ARG,
                   "time the media of a coupled mesh on single-medium meshes");
  }
  void set_model(const float *, const float *, const float *) override {
    throw SmxError(SMX_ERR_ARG, "set_model is built for elastic meshes only");
  }
  void run_forward_checkpointed(int) override {
    throw SmxError(SMX_ERR_ARG, "the adjoint is built for elastic meshes only");
  }
  void run_adjoint(const void *) override {
    throw SmxError(SMX_ERR_ARG, "the adjoint is built for elastic meshes only");
  }
  void get_gradient(double *) override {
    throw SmxError(SMX_ERR_ARG, "the adjoint is built for elastic meshes only");
  }
  void last_adjoint_stats(double *f, double *a) override {
    if (f)
      *f = 0;
    if (a)
      *a = 0;
  }
  void geometry_mode(int *mode, long long *mismatches) const override {
    if (mode)
      *mode = 0;
    if (mismatches)
      *mismatches = 0;
  }

private:
  int untagged() {
    throw SmxError(SMX_ERR_STATE,
                   "a coupled mesh advances one medium per operator: use the "
                   "medium-tagged operators (smx_*_medium) or smx_run_forward");
  }
  static void forward_only(int f) {
    if (f != SMX_FIELD_FORWARD)
      throw SmxError(SMX_ERR_ARG, "coupled meshes hold the forward field only");
  }
  template <typename Fn> int on_medium(int f, int m, Fn fn) {
    forward_only(f);
    Solver2DBase *s = m == SMX_MEDIUM_ELASTIC_PSV
                          ? el_.get()
                          : (m == SMX_MEDIUM_ACOUSTIC ? ac_.get() : nullptr);
    if (!s)
      return 0;
    int n = 0;
    s->on_stream(stream_, [&] { n = fn(*s); });
    return n;
  }
  void no_fields() {
    throw SmxError(SMX_ERR_STATE,
                   "a coupled mesh holds one field per medium; they are not "
                   "exposed through the single-field calls");
  }
  void step(bool final) {
    ac_->enqueue_step(stream_, final);
    el_->enqueue_step(stream_, final);
  }
  void build_graph() {
    if (graph_exec_)
      return;
    CP_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      for (int s = 0; s < gsteps_; ++s)
        step(false);
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(stream_, &tmp);
      if (tmp)
        (void)hipGraphDestroy(tmp);
      throw;
    }
    CP_HIP(hipStreamEndCapture(stream_, &graph_));
    CP_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  }

  std::unique_ptr<CoupledAssembly2D> ca_;
  std::unique_ptr<Solver2DBase> el_, ac_;
  int gsteps_ = 8;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  float *chi_dd_ = nullptr;
  bool initialized_ = false, have_stats_ = false;
  int launches_ = 0;
};

} // namespace

std::unique_ptr<Solver2DBase>
make_coupled_solver2d(std::unique_ptr<CoupledAssembly2D> a, int precision,
                      int graph_steps) {
  if (precision != 4)
    throw SmxError(SMX_ERR_ARG, "coupled meshes run in FP32 (precision 4)");
  if (graph_steps < 1 || graph_steps > 256)
    throw SmxError(SMX_ERR_ARG, "graph_steps must lie in [1, 256]");
  return std::make_unique<CoupledSolver2D>(std::move(a), graph_steps);
}

} // namespace smx
