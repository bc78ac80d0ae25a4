// C ABI of libspecfempp_amd (declared in include/specfempp_amd.h).  Every
// entry point converts exceptions into a status code and a per-thread
// message; none lets a C++ exception cross the boundary.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "coupled2d.hpp"
#include "solver2d.hpp"
#include "solver3d.hpp"
#include "specfempp_amd.h"

struct smx_ctx {
  std::unique_ptr<smx::Solver2DBase> solver;   // 2D contexts
  std::unique_ptr<smx::Solver3DBase> solver3d; // 3D contexts
};

namespace {

thread_local std::string g_last_error;

template <typename F> int guarded(F &&f) {
  try {
    f();
    return SMX_OK;
  } catch (const smx::SmxError &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return SMX_ERR_ARG;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return SMX_ERR_ARG;
  }
}

smx::Solver2DBase &S(smx_ctx *ctx) {
  if (ctx && ctx->solver3d)
    throw smx::SmxError(SMX_ERR_STATE, "operation not available for 3D contexts");
  if (!ctx || !ctx->solver)
    throw smx::SmxError(SMX_ERR_ARG, "null context");
  return *ctx->solver;
}

// The 3D solver of a context, or null for a 2D one.
smx::Solver3DBase *S3(smx_ctx *ctx) {
  if (!ctx)
    throw smx::SmxError(SMX_ERR_ARG, "null context");
  return ctx->solver3d.get();
}

} // namespace

extern "C" {

const char *smx_last_error(void) { return g_last_error.c_str(); }

const char *smx_version(void) { return "specfempp_amd 0.1.0 (gfx950)"; }

int smx_ctx_create2d(const smx_mesh2d *mesh, const smx_sources2d *sources,
                     const smx_receivers2d *receivers, double dt, int nstep,
                     int precision, smx_ctx **out) {
  return guarded([&] {
    if (!mesh || !sources || !receivers || !out)
      throw smx::SmxError(SMX_ERR_ARG, "null argument");
    *out = nullptr;
    auto ctx = std::make_unique<smx_ctx>();
    if (smx::mesh_is_coupled2d(*mesh)) {
      auto c = std::make_unique<smx::CoupledAssembly2D>();
      smx::assemble_coupled2d(*mesh, *sources, *receivers, dt, nstep, {}, *c);
      ctx->solver = smx::make_coupled_solver2d(std::move(c), precision, 8);
    } else {
      auto a = std::make_unique<smx::HostAssembly2D>();
      smx::assemble2d(*mesh, *sources, *receivers, dt, nstep, *a);
      ctx->solver = smx::make_solver2d(std::move(a), precision);
    }
    *out = ctx.release();
  });
}

namespace {
smx::AssemblyOptions assembly_options(const smx_options *o) {
  smx::AssemblyOptions a;
  if (o && o->tile_elems)
    a.tile_elems = o->tile_elems;
  return a;
}
int graph_steps(const smx_options *o) {
  return (o && o->graph_steps) ? o->graph_steps : 8;
}
bool geometry_from_nodes(const smx_options *o) {
  return o && o->geometry_from_nodes;
}
int waves_per_eu(const smx_options *o) { return o ? o->waves_per_eu : 0; }
} // namespace

int smx_ctx_create2d_opts(const smx_mesh2d *mesh, const smx_sources2d *sources,
                          const smx_receivers2d *receivers, double dt,
                          int nstep, int precision, const smx_options *opts,
                          smx_ctx **out) {
  return guarded([&] {
    if (!mesh || !sources || !receivers || !out)
      throw smx::SmxError(SMX_ERR_ARG, "null argument");
    *out = nullptr;
    auto ctx = std::make_unique<smx_ctx>();
    if (smx::mesh_is_coupled2d(*mesh)) {
      auto c = std::make_unique<smx::CoupledAssembly2D>();
      if (waves_per_eu(opts) != 0)
        throw smx::SmxError(SMX_ERR_ARG, "waves_per_eu variants exist for "
                                         "single-medium elastic contexts only");
      smx::assemble_coupled2d(*mesh, *sources, *receivers, dt, nstep,
                              assembly_options(opts), *c);
      ctx->solver = smx::make_coupled_solver2d(std::move(c), precision,
                                               graph_steps(opts));
    } else {
      auto a = std::make_unique<smx::HostAssembly2D>();
      smx::assemble2d(*mesh, *sources, *receivers, dt, nstep, *a,
                      assembly_options(opts));
      ctx->solver = smx::make_solver2d(std::move(a), precision,
                                       graph_steps(opts),
                                       geometry_from_nodes(opts),
                                       waves_per_eu(opts));
    }
    *out = ctx.release();
  });
}

int smx_ctx_create2d_batch(int nshots, const smx_mesh2d *meshes,
                           const smx_sources2d *sources,
                           const smx_receivers2d *receivers, double dt,
                           int nstep, int precision, const smx_options *opts,
                           smx_ctx **out) {
  return guarded([&] {
    if (nshots < 1 || !meshes || !sources || !receivers || !out)
      throw smx::SmxError(SMX_ERR_ARG, "null argument or no shots");
    *out = nullptr;
    // shots are assembled independently on host threads, then merged
    std::vector<smx::HostAssembly2D> shots(nshots);
    std::vector<std::string> errors(nshots);
    std::vector<int> codes(nshots, 0);
    const smx::AssemblyOptions aopt = assembly_options(opts);
    const int nthreads = std::max(
        1, std::min<int>(nshots, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
      pool.emplace_back([&, t] {
        for (int k = t; k < nshots; k += nthreads) {
          try {
            smx::assemble2d(meshes[k], sources[k], receivers[k], dt, nstep,
                            shots[k], aopt);
          } catch (const smx::SmxError &e) {
            codes[k] = e.code;
            errors[k] = e.what();
          } catch (const std::exception &e) {
            codes[k] = SMX_ERR_ARG;
            errors[k] = e.what();
          }
        }
      });
    for (auto &th : pool)
      th.join();
    for (int k = 0; k < nshots; ++k)
      if (codes[k])
        throw smx::SmxError(codes[k], "shot " + std::to_string(k) + ": " + errors[k]);
    auto a = std::make_unique<smx::HostAssembly2D>();
    smx::merge_assemblies(shots, *a);
    shots.clear();
    auto ctx = std::make_unique<smx_ctx>();
    ctx->solver = smx::make_solver2d(std::move(a), precision, graph_steps(opts),
                                      geometry_from_nodes(opts),
                                      waves_per_eu(opts));
    *out = ctx.release();
  });
}

int smx_ctx_shots(const smx_ctx *ctx, int *nshots, int *elem_off, int *rec_off,
                  int *tile_off) {
  return guarded([&] {
    const auto &A = S(const_cast<smx_ctx *>(ctx)).assembly();
    const int n = static_cast<int>(A.shot_elem_off.size()) - 1;
    if (nshots) *nshots = n;
    for (int k = 0; k <= n; ++k) {
      if (elem_off) elem_off[k] = A.shot_elem_off[k];
      if (rec_off) rec_off[k] = A.shot_rec_off[k];
      if (tile_off) tile_off[k] = A.shot_tile_off[k];
    }
  });
}

int smx_ctx_create3d(const smx_mesh3d *mesh, const smx_sources3d *sources,
                     const smx_receivers3d *receivers, double dt, int nstep,
                     int precision, smx_ctx **out) {
  return smx_ctx_create3d_opts(mesh, sources, receivers, dt, nstep, precision,
                               nullptr, out);
}

int smx_ctx_create3d_opts(const smx_mesh3d *mesh, const smx_sources3d *sources,
                          const smx_receivers3d *receivers, double dt,
                          int nstep, int precision, const smx_options *opts,
                          smx_ctx **out) {
  return guarded([&] {
    if (!mesh || !sources || !receivers || !out)
      throw smx::SmxError(SMX_ERR_ARG, "null argument");
    *out = nullptr;
    if (waves_per_eu(opts) != 0)
      throw smx::SmxError(SMX_ERR_ARG,
                          "waves_per_eu variants exist for 2D elastic contexts only");
    // tiles of 32 elements in rounds of 4 (FP32); 8 elements, one per round,
    // in FP64 (twice the LDS per value)
    smx::AssemblyOptions3D ao;
    ao.tile_elems = (opts && opts->tile_elems) ? opts->tile_elems
                                               : (precision == 8 ? 8 : 32);
    ao.round_width = precision == 8 ? 1 : 4;
    auto a = std::make_unique<smx::HostAssembly3D>();
    smx::assemble3d(*mesh, *sources, *receivers, dt, nstep, *a, ao);
    auto ctx = std::make_unique<smx_ctx>();
    ctx->solver3d = smx::make_solver3d(std::move(a), precision, graph_steps(opts),
                                       geometry_from_nodes(opts));
    *out = ctx.release();
  });
}

int smx_ctx_destroy(smx_ctx *ctx) {
  return guarded([&] { delete ctx; });
}

int smx_ctx_components(const smx_ctx *ctx, int *ncomp) {
  return guarded([&] {
    if (!ncomp) throw smx::SmxError(SMX_ERR_ARG, "null output");
    if (S3(const_cast<smx_ctx *>(ctx))) {
      *ncomp = 3;
      return;
    }
    const int md = S(const_cast<smx_ctx *>(ctx)).assembly().medium;
    *ncomp = md == 1 ? 1 : (md == 2 ? 0 : 2); // 0: one field per medium
  });
}

int smx_ctx_info(const smx_ctx *ctx, int *nspec, int *nglob, int *nsig,
                 int *nrec) {
  return guarded([&] {
    if (auto *s3 = S3(const_cast<smx_ctx *>(ctx))) {
      const auto &A3 = s3->assembly();
      if (nspec) *nspec = A3.nspec;
      if (nglob) *nglob = A3.nglob;
      if (nsig) *nsig = A3.nsig;
      if (nrec) *nrec = A3.nrec;
      return;
    }
    const auto &A = S(const_cast<smx_ctx *>(ctx)).assembly();
    if (nspec) *nspec = A.nspec;
    if (nglob) *nglob = A.nglob;
    if (nsig) *nsig = A.nsig;
    if (nrec) *nrec = A.nrec;
  });
}

int smx_ctx_layout(const smx_ctx *ctx, int *ntiles, int *nexclusive,
                   int *nshared, int *npartials) {
  return guarded([&] {
    if (auto *s3 = S3(const_cast<smx_ctx *>(ctx))) {
      const auto &B = s3->assembly();
      if (ntiles) *ntiles = B.ntiles;
      if (nexclusive) *nexclusive = B.nex_total;
      if (nshared) *nshared = B.nsh_total;
      if (npartials) *npartials = B.npart;
      return;
    }
    const auto &A = S(const_cast<smx_ctx *>(ctx)).assembly();
    if (ntiles) *ntiles = A.ntiles;
    if (nexclusive) *nexclusive = A.nex_total;
    if (nshared) *nshared = A.nshared;
    if (npartials) *npartials = static_cast<int>(A.sh_list.size());
  });
}

int smx_ctx_compact(const smx_ctx *ctx, int *geom_elements, int *mat_elements) {
  return guarded([&] {
    if (auto *s3 = S3(const_cast<smx_ctx *>(ctx))) {
      const auto &B = s3->assembly();
      if (geom_elements) *geom_elements = B.geom_elem ? B.nspec : 0;
      if (mat_elements) *mat_elements = B.mat_elem ? B.nspec : 0;
      return;
    }
    S(const_cast<smx_ctx *>(ctx)).sync_host_model();
    const auto &A = S(const_cast<smx_ctx *>(ctx)).assembly();
    int g = 0, m = 0;
    for (const auto &t : A.tiles) {
      if (t.flags & smx::TILE_GEOM_ELEM) g += t.ecount;
      if (t.flags & smx::TILE_MAT_ELEM) m += t.ecount;
    }
    if (geom_elements) *geom_elements = g;
    if (mat_elements) *mat_elements = m;
  });
}

int smx_ctx_geometry(const smx_ctx *ctx, int *mode, long long *mismatches) {
  return guarded([&] {
    auto *c = const_cast<smx_ctx *>(ctx);
    if (auto *s3 = S3(c))
      return s3->geometry_mode(mode, mismatches);
    S(c).geometry_mode(mode, mismatches);
  });
}

double smx_ctx_t0(const smx_ctx *ctx) {
  if (ctx && ctx->solver3d)
    return ctx->solver3d->assembly().t0;
  if (!ctx || !ctx->solver)
    return 0.0;
  return ctx->solver->assembly().t0;
}

int smx_initialize(smx_ctx *ctx) {
  return guarded([&] {
    if (auto *s3 = S3(ctx))
      return s3->initialize();
    S(ctx).initialize();
  });
}

int smx_predictor(smx_ctx *ctx, int field, int *dofs) {
  return guarded([&] {
    const int n = S(ctx).predictor(field);
    if (dofs) *dofs = n;
  });
}

int smx_corrector(smx_ctx *ctx, int field, int *dofs) {
  return guarded([&] {
    const int n = S(ctx).corrector(field);
    if (dofs) *dofs = n;
  });
}

int smx_update_wavefields(smx_ctx *ctx, int field, int istep, int *nelem) {
  return guarded([&] {
    const int n = S(ctx).update_wavefields(field, istep);
    if (nelem) *nelem = n;
  });
}

int smx_ctx_degrees_of_freedom(const smx_ctx *ctx, long long *dofs) {
  return guarded([&] {
    if (!dofs)
      throw smx::SmxError(SMX_ERR_ARG, "null output");
    if (S3(const_cast<smx_ctx *>(ctx)))
      throw smx::SmxError(SMX_ERR_STATE, "operation not available for 3D contexts");
    *dofs = S(const_cast<smx_ctx *>(ctx)).total_dofs();
  });
}

int smx_predictor_medium(smx_ctx *ctx, int field, int medium, int *dofs) {
  return guarded([&] {
    const int n = S(ctx).predictor_medium(field, medium);
    if (dofs) *dofs = n;
  });
}

int smx_corrector_medium(smx_ctx *ctx, int field, int medium, int *dofs) {
  return guarded([&] {
    const int n = S(ctx).corrector_medium(field, medium);
    if (dofs) *dofs = n;
  });
}

int smx_update_wavefields_medium(smx_ctx *ctx, int field, int medium,
                                 int istep, int *nelem) {
  return guarded([&] {
    const int n = S(ctx).update_wavefields_medium(field, medium, istep);
    if (nelem) *nelem = n;
  });
}

int smx_update_wavefields_noalloc(smx_ctx *ctx, const void *u, const void *v,
                                  void *a, int istep, int *nelem) {
  return guarded([&] {
    const int n = S(ctx).update_wavefields_noalloc(u, v, a, istep);
    if (nelem) *nelem = n;
  });
}

int smx_compute_seismograms(smx_ctx *ctx, int field, int isig) {
  return guarded([&] { S(ctx).compute_seismograms(field, isig); });
}

int smx_run_forward(smx_ctx *ctx, int istep0, int nstep) {
  return guarded([&] {
    if (auto *s3 = S3(ctx))
      return s3->run_forward(istep0, nstep);
    S(ctx).run_forward(istep0, nstep);
  });
}

int smx_ctx_set_launch(smx_ctx *ctx, int graph_steps, int waves_per_eu) {
  return guarded([&] {
    if (S3(ctx))
      throw smx::SmxError(SMX_ERR_STATE,
                          "smx_ctx_set_launch: 2D elastic contexts only");
    S(ctx).set_launch(graph_steps, waves_per_eu);
  });
}

int smx_synchronize(smx_ctx *ctx) {
  return guarded([&] {
    if (auto *s3 = S3(ctx))
      return s3->synchronize();
    S(ctx).synchronize();
  });
}

int smx_get_seismograms(smx_ctx *ctx, void *host_out) {
  return guarded([&] {
    if (!host_out) throw smx::SmxError(SMX_ERR_ARG, "null output");
    if (auto *s3 = S3(ctx))
      return s3->get_seismograms(host_out);
    S(ctx).get_seismograms(host_out);
  });
}

int smx_get_field(smx_ctx *ctx, int field, int which, void *host_out) {
  return guarded([&] {
    if (!host_out) throw smx::SmxError(SMX_ERR_ARG, "null output");
    if (auto *s3 = S3(ctx)) {
      if (field != SMX_FIELD_FORWARD)
        throw smx::SmxError(SMX_ERR_ARG, "3D contexts hold the forward field only");
      return s3->get_field(which, host_out);
    }
    S(ctx).get_field(field, which, host_out);
  });
}

int smx_set_field(smx_ctx *ctx, int field, int which, const void *host_in) {
  return guarded([&] {
    if (!host_in) throw smx::SmxError(SMX_ERR_ARG, "null input");
    if (auto *s3 = S3(ctx)) {
      if (field != SMX_FIELD_FORWARD)
        throw smx::SmxError(SMX_ERR_ARG, "3D contexts hold the forward field only");
      return s3->set_field(which, host_in);
    }
    S(ctx).set_field(field, which, host_in);
  });
}

int smx_get_ibool(smx_ctx *ctx, int *host_out) {
  return guarded([&] {
    if (!host_out) throw smx::SmxError(SMX_ERR_ARG, "null output");
    const auto &A = S(ctx).assembly();
    if (A.medium == 2)
      throw smx::SmxError(SMX_ERR_STATE,
                          "a coupled mesh numbers each medium separately");
    for (int e = 0; e < A.nspec; ++e)
      for (int p = 0; p < smx::NPT; ++p)
        host_out[size_t(e) * smx::NPT + p] =
            A.ibool[size_t(p) * A.nspec_pad + A.elem_slot[e]];
  });
}

int smx_field_device_ptr(smx_ctx *ctx, int field, int which, void **dptr) {
  return guarded([&] {
    if (!dptr) throw smx::SmxError(SMX_ERR_ARG, "null output");
    *dptr = S(ctx).field_ptr(field, which);
  });
}

int smx_last_run_stats(smx_ctx *ctx, double *ms, double *stiffness_ms,
                       int *stiffness_launches) {
  return guarded(
      [&] {
        if (auto *s3 = S3(ctx))
          return s3->last_run_stats(ms, stiffness_ms, stiffness_launches);
        S(ctx).last_run_stats(ms, stiffness_ms, stiffness_launches);
      });
}

int smx_time_stiffness(smx_ctx *ctx, int nlaunch, double *avg_ms) {
  return guarded([&] {
    auto *s3 = S3(ctx);
    const double t = s3 ? s3->time_stiffness(nlaunch) : S(ctx).time_stiffness(nlaunch);
    if (avg_ms) *avg_ms = t;
  });
}

int smx_set_model(smx_ctx *ctx, const float *kappa, const float *mu,
                  const float *rho) {
  return guarded([&] {
    if (!kappa || !mu || !rho) throw smx::SmxError(SMX_ERR_ARG, "null model");
    if (auto *s3 = S3(ctx))
      s3->set_model(kappa, mu, rho);
    else
      S(ctx).set_model(kappa, mu, rho);
  });
}

int smx_set_model_device(smx_ctx *ctx, const float *kappa, const float *mu,
                         const float *rho) {
  return guarded([&] {
    if (!kappa || !mu || !rho) throw smx::SmxError(SMX_ERR_ARG, "null model");
    S(ctx).set_model_device(kappa, mu, rho);
  });
}

int smx_get_gradient_device(smx_ctx *ctx, double *device_out) {
  return guarded([&] {
    if (!device_out) throw smx::SmxError(SMX_ERR_ARG, "null output");
    S(ctx).get_gradient_device(device_out);
  });
}

int smx_run_forward_checkpointed(smx_ctx *ctx, int every) {
  return guarded([&] {
    if (auto *s3 = S3(ctx))
      s3->run_forward_checkpointed(every);
    else
      S(ctx).run_forward_checkpointed(every);
  });
}

int smx_run_adjoint(smx_ctx *ctx, const void *gsrc) {
  return guarded([&] {
    if (auto *s3 = S3(ctx)) {
      if (!gsrc && s3->assembly().nrec > 0)
        throw smx::SmxError(SMX_ERR_ARG, "null adjoint source");
      s3->run_adjoint(gsrc);
      return;
    }
    if (!gsrc && S(ctx).assembly().nrec > 0)
      throw smx::SmxError(SMX_ERR_ARG, "null adjoint source");
    S(ctx).run_adjoint(gsrc);
  });
}

int smx_get_gradient(smx_ctx *ctx, double *host_out) {
  return guarded([&] {
    if (!host_out) throw smx::SmxError(SMX_ERR_ARG, "null output");
    if (auto *s3 = S3(ctx))
      s3->get_gradient(host_out);
    else
      S(ctx).get_gradient(host_out);
  });
}

int smx_last_adjoint_stats(smx_ctx *ctx, double *forward_ms,
                           double *adjoint_ms) {
  return guarded([&] {
    if (auto *s3 = S3(ctx))
      s3->last_adjoint_stats(forward_ms, adjoint_ms);
    else
      S(ctx).last_adjoint_stats(forward_ms, adjoint_ms);
  });
}

} // extern "C"
