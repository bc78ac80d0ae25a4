/*
 * specfempp_amd — MI355X (gfx950) implementation of SPECFEM++'s elastic
 * isotropic time-stepping path (2D P-SV; 3D hex8), its discrete adjoint and
 * batched multi-shot runs, and the 2D acoustic isotropic time step, exposed
 * through a plain C ABI.
 *
 * Every entry point returns 0 on success and a negative status on error;
 * smx_last_error() returns the message of the last failure on the calling
 * thread.  A context owns its device buffers and one HIP stream; it is not
 * thread-safe (use one context per host thread).  All host arrays passed in
 * are copied; none is retained after the call returns.
 *
 * Citations are to the reference interface each entry point replaces
 * (paths relative to the SPECFEM++ source tree).
 */
#ifndef SPECFEMPP_AMD_H
#define SPECFEMPP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define SMX_OK 0
#define SMX_ERR_ARG -1     /* invalid argument / inconsistent sizes */
#define SMX_ERR_HIP -2     /* HIP runtime error */
#define SMX_ERR_STATE -3   /* call not valid in the context's state */
#define SMX_ERR_MESH -4    /* mesh / point location error */

/* Wavefields (include/enumerations/wavefield.hpp simulation_field). */
#define SMX_FIELD_FORWARD 0
#define SMX_FIELD_ADJOINT 1
#define SMX_FIELD_BACKWARD 2

/* Medium tags, in the order of specfem::element::medium_tag
 * (include/enumerations/medium.hpp:18-28).  This library builds elastic_psv
 * and acoustic in 2D (and elastic in 3D, through smx_run_forward); the
 * medium-tagged operators update nothing and return 0 for any other tag and
 * for a built tag without elements in the context's mesh, as the
 * reference's loop expects of media a mesh does not hold
 * (core/specfem/solver/time_marching.tpp:33-84). */
#define SMX_MEDIUM_ELASTIC_PSV 0
#define SMX_MEDIUM_ELASTIC_SH 1
#define SMX_MEDIUM_ELASTIC_PSV_T 2
#define SMX_MEDIUM_ACOUSTIC 3
#define SMX_MEDIUM_POROELASTIC 4
#define SMX_MEDIUM_ELECTROMAGNETIC_TE 5
#define SMX_MEDIUM_ELASTIC 6
#define SMX_MEDIUM_ELASTIC_SPIN 7
#define SMX_MEDIUM_ELECTROMAGNETIC 8

/* Seismogram types (specfem::wavefield::type). */
#define SMX_SEIS_DISPLACEMENT 0
#define SMX_SEIS_VELOCITY 1
#define SMX_SEIS_ACCELERATION 2
/* 2D only: -chi_dd in acoustic elements (acoustic/isotropic/wavefield.hpp);
 * -(lambda + 2/3 mu)(du_x/dx + du_z/dz) in elastic P-SV elements
 * (elastic/isotropic/wavefield.hpp:57-94); component 0 holds it, 1 is 0 */
#define SMX_SEIS_PRESSURE 3

/* Source time functions (core/specfem/source_time_functions). */
#define SMX_STF_RICKER 0
#define SMX_STF_DIRAC 1 /* Gaussian of f0 = 1 / (10 dt) (dirac.cpp:23-35) */
/* factor (1 + erf(t / (hdur / 1.628))) / 2, hdur raised to 5 dt, start time
 * -2 hdur + tshift (heaviside.cpp:10-78) */
#define SMX_STF_HEAVISIDE 2

/* 2D source kinds. */
#define SMX_SRC_FORCE 0
#define SMX_SRC_MOMENT 1

/* Field components selectable by smx_get_field / smx_set_field. */
#define SMX_U 0
#define SMX_V 1
#define SMX_A 2
#define SMX_MASS_INVERSE 3

typedef struct smx_ctx smx_ctx;

/* Mesh: what the reference reads from a mesh database
 * (src/io/mesh/impl/fortran/dim2/mesh.cpp:26-215) plus the per-GLL-point
 * material of assembly::properties (core/specfem/assembly/properties).
 * Edge types follow include/enumerations/dim2/mesh_entities.hpp:29-38
 * (1 bottom, 2 right, 3 top, 4 left).
 *
 * Elements whose 25 points all have mu == 0 are acoustic (a material with
 * cs = 0, src/io/mesh/impl/fortran/dim2/read_material_properties.cpp:
 * 178-204).  A mesh with both media is coupled on its conforming
 * elastic-acoustic edges (include/kokkos_kernels/impl/compute_coupling.tpp,
 * include/medium/dim2/coupling_terms): such a context runs in FP32 through
 * smx_run_forward only and exposes no single field (smx_ctx_components
 * reports 0).  Acoustic free-surface (Dirichlet) entities come from the
 * database's acoustic-surface records (read_boundaries.cpp:276-300): edge
 * types 1..4 as above, corners 5 bottom-left, 6 bottom-right, 7 top-right,
 * 8 top-left. */
typedef struct {
  int nspec;            /* number of spectral elements */
  int ngnod;            /* control nodes per element: 4 or 9 */
  int nnodes;           /* number of control nodes */
  const double *coord;  /* [nnodes][2] (x, z) */
  const int *knods;     /* [nspec][ngnod] control-node ids */
  const float *kappa;   /* [nspec][5][5] bulk modulus at GLL points (iz, ix) */
  const float *mu;      /* [nspec][5][5] shear modulus */
  const float *rho;     /* [nspec][5][5] density */
  int nabs;             /* absorbing (Stacey) edges */
  const int *abs_ispec; /* [nabs] element of each edge */
  const int *abs_type;  /* [nabs] edge type 1..4 */
  int nfree;             /* acoustic free-surface entities (0 if none) */
  const int *free_ispec; /* [nfree] element */
  const int *free_type;  /* [nfree] entity type 1..8 */
} smx_mesh2d;

/* Point forces (core/specfem/source/dim2/vector_source/force_source.cpp)
 * and moment tensors (core/specfem/source/dim2/tensor_source/
 * moment_tensor_source.cpp, elastic elements only) with a Ricker, Dirac or
 * Heaviside time function (core/specfem/source_time_functions/ricker.cpp,
 * dirac.cpp, heaviside.cpp).  In acoustic elements the force has one
 * component and the angle is unused (force_source.cpp:36-40). */
typedef struct {
  int nsrc;
  const double *x, *z;  /* [nsrc] location */
  const float *angle;   /* [nsrc] degrees clockwise from +z (forces) */
  const float *f0;      /* [nsrc] dominant frequency (Ricker); half duration
                           hdur in seconds (Heaviside); unused (Dirac) */
  const float *tshift;  /* [nsrc] */
  const float *factor;  /* [nsrc] amplitude */
  const int *stf;       /* [nsrc] SMX_STF_*, or NULL for all Ricker */
  const int *kind;      /* [nsrc] SMX_SRC_*, or NULL for all point forces */
  const float *moment;  /* [nsrc][3] Mxx, Mzz, Mxz of SMX_SRC_MOMENT sources
                           (may be NULL when there are none) */
} smx_sources2d;

/* Stations and the seismograms recorded at them
 * (core/specfem/assembly/receivers/dim2/receivers.cpp:11-148). */
typedef struct {
  int nrec;
  const double *x, *z;  /* [nrec] */
  int ntypes;           /* seismogram types recorded */
  const int *types;     /* [ntypes] SMX_SEIS_* */
  int nstep_between_samples;
} smx_receivers2d;

/* 3D hex8 mesh (NGLL = 5).  Control nodes [nnodes][3] (x, y, z); element
 * connectivity [nspec][8] in the reference's hex8 order (z- face
 * counter-clockwise from (-,-,-), then the z+ face,
 * core/specfem/shape_function/dim3/shape_functions3d.cpp); material per GLL
 * point [nspec][5][5][5] (iz, iy, ix); absorbing faces (element, face with
 * bottom=1 (z-), right=2 (x+), top=3 (z+), left=4 (x-), front=5 (y-),
 * back=6 (y+), include/enumerations/dim3/mesh_entities.hpp). */
typedef struct smx_mesh3d {
  int nspec;
  int nnodes;
  const double *coord;
  const int *knods;
  const float *kappa;
  const float *mu;
  const float *rho;
  int nabs;
  const int *abs_ispec;
  const int *abs_face;
} smx_mesh3d;

/* 3D point sources with a Ricker time function: kind 0 = force (comp[0..2]
 * = fx, fy, fz), kind 1 = moment tensor (comp = Mxx, Myy, Mzz, Mxy, Mxz,
 * Myz); comp is [nsrc][6]. */
typedef struct smx_sources3d {
  int nsrc;
  const double *x;
  const double *y;
  const double *z;
  const int *kind;
  const float *comp;
  const float *f0;
  const float *tshift;
  const float *factor;
} smx_sources3d;

typedef struct smx_receivers3d {
  int nrec;
  const double *x;
  const double *y;
  const double *z;
  int ntypes;
  const int *types; /* [ntypes] SMX_SEIS_* */
  int nstep_between_samples;
} smx_receivers3d;

/* Last error message of the calling thread (never NULL). */
const char *smx_last_error(void);

/* Library version string. */
const char *smx_version(void);

/* Builds the assembly on the host (GLL quadrature, Jacobian, global
 * numbering, Stacey normals/weights, source arrays and time functions,
 * receiver interpolants) and uploads it.  precision is 4 (FP32, the
 * reference's type_real) or 8 (FP64 twin of the same kernels).
 * Replaces specfem::assembly::assembly<dim2>::assembly
 * (core/specfem/assembly/assembly/dim2/assembly.cpp:6-93) together with the
 * domain_kernels / newmark constructors
 * (include/kokkos_kernels/domain_kernels.hpp:58-59,
 *  core/specfem/timescheme/newmark.hpp:131-135). */
int smx_ctx_create2d(const smx_mesh2d *mesh, const smx_sources2d *sources,
                     const smx_receivers2d *receivers, double dt, int nstep,
                     int precision, smx_ctx **out);

/* Layout choices (results keep their meaning; a zero field takes the
 * default).  Chosen per mesh by specfempp_amd.autotune. */
typedef struct smx_options {
  int tile_elems;  /* elements per tile, 1..64 (default 64) */
  int graph_steps; /* time steps per replayed HIP graph, 1..256 (default 8) */
  /* 1: tiles re-form per-point geometry from the element control nodes when
   * the device reproduces every stored value bit for bit (fewer HBM bytes,
   * more registers); 0 (default): read the stored per-point geometry. */
  int geometry_from_nodes;
  /* Occupancy variant of the FP32 time-step tile kernel of 2D elastic
   * contexts: the fewest waves per SIMD its register allocation must allow
   * (6, 7 or 8; 0 = the default, 7).  Results are bitwise the same for
   * every value; only the speed changes.  Nonzero values are an error for
   * FP64, acoustic, coupled and 3D contexts. */
  int waves_per_eu;
} smx_options;

/* Changes the time steps per replayed HIP graph (1..256) and the
 * occupancy variant (smx_options.waves_per_eu) of a 2D elastic context,
 * single-shot or batched, at any point between runs: the step graph is
 * captured and uploaded again here (on an initialised context), so the next
 * smx_run_forward pays for neither.  Results are bitwise unchanged.  No
 * reference counterpart (Kokkos chooses its launch shapes itself); the
 * layout autotuner uses it to time candidates on one assembly. */
int smx_ctx_set_launch(smx_ctx *ctx, int graph_steps, int waves_per_eu);

/* smx_ctx_create2d with layout options (opts may be NULL). */
int smx_ctx_create2d_opts(const smx_mesh2d *mesh, const smx_sources2d *sources,
                          const smx_receivers2d *receivers, double dt,
                          int nstep, int precision, const smx_options *opts,
                          smx_ctx **out);

/* One context for nshots independent shots (meshes of any sizes, each with
 * its own sources, receivers and medium; one dt, nstep and seismogram
 * set-up).  Every operator and smx_run_forward then advance all shots with
 * one launch per kernel per step; seismograms hold the shots' receivers in
 * shot order.  Each shot's results equal those of its own context (same
 * options) bit for bit.  No reference counterpart: the reference runs one
 * simulation per process (core/specfem/solver/time_marching.tpp:8-131). */
int smx_ctx_create2d_batch(int nshots, const smx_mesh2d *meshes,
                           const smx_sources2d *sources,
                           const smx_receivers2d *receivers, double dt,
                           int nstep, int precision, const smx_options *opts,
                           smx_ctx **out);

/* Shot boundaries of a context: element, receiver and tile offsets (arrays
 * of nshots+1, each may be NULL). */
int smx_ctx_shots(const smx_ctx *ctx, int *nshots, int *elem_off, int *rec_off,
                  int *tile_off);

/* A 3D elastic context (hex8, NGLL = 5).  Supported operations:
 * smx_initialize, smx_run_forward, smx_synchronize, smx_get_seismograms
 * ([nsig][ntypes][nrec][3]), smx_get_field / smx_set_field ([nglob][3]),
 * smx_ctx_info, smx_last_run_stats, smx_time_stiffness, smx_ctx_geometry,
 * the adjoint path (smx_set_model, smx_run_forward_checkpointed,
 * smx_run_adjoint, smx_get_gradient, smx_last_adjoint_stats) and
 * smx_ctx_destroy; others return SMX_ERR_STATE.  Replaces assembly<dim3> + the dim3
 * instantiations of domain_kernels (src/kokkos_kernels/impl/
 * compute_stiffness_interaction.cpp:53-63) and time_marching<forward, dim3>.
 * Absorbing faces are this build's own 3D Stacey condition. */
int smx_ctx_create3d(const smx_mesh3d *mesh, const smx_sources3d *sources,
                     const smx_receivers3d *receivers, double dt, int nstep,
                     int precision, smx_ctx **out);

/* smx_ctx_create3d with options (opts may be NULL; tile_elems is ignored). */
int smx_ctx_create3d_opts(const smx_mesh3d *mesh, const smx_sources3d *sources,
                          const smx_receivers3d *receivers, double dt,
                          int nstep, int precision, const smx_options *opts,
                          smx_ctx **out);

/* Frees every device buffer and the stream of the context. */
int smx_ctx_destroy(smx_ctx *ctx);

/* Sizes: nspec, nglob, number of seismogram samples, receivers. */
int smx_ctx_info(const smx_ctx *ctx, int *nspec, int *nglob, int *nsig,
                 int *nrec);

/* Field components per GLL point: 2 (2D elastic P-SV), 1 (2D acoustic: the
 * potential), 3 (3D elastic), 0 for a coupled mesh (one field per medium,
 * not exposed) — the medium's `components` attribute
 * (include/enumerations/medium.hpp:98,162, attributes<dim, medium>::components).
 * smx_get_field / smx_set_field / smx_field_device_ptr arrays are
 * [nglob][ncomp]. */
int smx_ctx_components(const smx_ctx *ctx, int *ncomp);

/* Device layout of the assembly (2D or 3D): element tiles, GLL points
 * exclusive to one tile, points shared between tiles, and per-tile partial
 * slots of shared points.  No reference counterpart (the reference's layout is Kokkos
 * views); used to state the kernels' algorithmic byte counts. */
int smx_ctx_layout(const smx_ctx *ctx, int *ntiles, int *nexclusive,
                   int *nshared, int *npartials);

/* Start time t0 of the simulation (src/io/sources/dim2/read_sources.cpp:142-172). */
double smx_ctx_t0(const smx_ctx *ctx);

/* Elements whose tiles read element-constant geometry / moduli (one value
 * per element instead of one per GLL point; see the assembly's tile flags). */
int smx_ctx_compact(const smx_ctx *ctx, int *geom_elements, int *mat_elements);

/* Where the stiffness kernels take the geometry from: 0 stored per GLL point,
 * 1 stored per element (affine elements), 2 re-formed on the device from the
 * control nodes with the host's FP32 operations (used only when a set-up
 * check reproduced every stored value bit for bit; `mismatches` counts the
 * values it did not).  A measurement aid with no reference counterpart. */
int smx_ctx_geometry(const smx_ctx *ctx, int *mode, long long *mismatches);

/* Mass matrix + Stacey mass term, then its inverse.
 * Replaces domain_kernels::initialize(dt)
 * (include/kokkos_kernels/domain_kernels.hpp:151-179). */
int smx_initialize(smx_ctx *ctx);

/* Newmark predictor u += dt v + dt^2/2 a; v += dt/2 a; a = 0.  *dofs gets the
 * number of degrees of freedom updated (nglob * 2).
 * Replaces newmark::apply_predictor_phase_forward/backward
 * (core/specfem/timescheme/newmark.tpp:152-174, 233-256). */
int smx_predictor(smx_ctx *ctx, int field, int *dofs);

/* Newmark corrector v += dt/2 a.
 * Replaces newmark::apply_corrector_phase_forward/backward
 * (core/specfem/timescheme/newmark.tpp:130-150, 189-210). */
int smx_corrector(smx_ctx *ctx, int field, int *dofs);

/* Source injection, stiffness (gradient, stress, divergence, Stacey) with
 * assembly into the acceleration, then division by the mass matrix.
 * *nelem gets the number of elements updated (nspec).
 * Replaces domain_kernels::update_wavefields<elastic_psv>(istep)
 * (include/kokkos_kernels/domain_kernels.hpp:72-141). */
int smx_update_wavefields(smx_ctx *ctx, int field, int istep, int *nelem);

/* smx_update_wavefields on caller-held device fields instead of the
 * context's: u and v ([nglob][2], context precision) are read, a is updated
 * in place (a := (a + f - K u - B v) / M for step istep).  No counterpart in
 * the reference: a variant for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  It calls no
 * allocator; the partial sums of points shared between tiles live in a
 * fixed 16 MiB device array of the library, one per GPU (up to 2,097,152
 * partial slots in FP32, 1,048,576 in FP64; larger meshes are refused with
 * SMX_ERR_ARG).  Every call on a GPU uses that GPU's one array, so calls
 * that overlap on one GPU (two contexts, two threads) race; nothing in the
 * array needs to start at zero.  2D elastic contexts only (SMX_ERR_STATE
 * otherwise); the context must be initialized. */
int smx_update_wavefields_noalloc(smx_ctx *ctx, const void *u, const void *v,
                                  void *a, int istep, int *nelem);

/* The operators of one medium, as the reference's loop calls them
 * (core/specfem/solver/time_marching.tpp:33-84):
 *   newmark::apply_predictor_phase_forward(tag)  (newmark.tpp:152-174)
 *   newmark::apply_corrector_phase_forward(tag)  (newmark.tpp:130-150)
 *   domain_kernels::update_wavefields<tag>(istep) (domain_kernels.hpp:72-141)
 * *dofs / *nelem get the degrees of freedom / elements of that medium
 * updated (0 for a medium the mesh does not hold).  On a coupled mesh each
 * call advances one medium's field, and called in the reference's order
 * (predictors, update<acoustic>, corrector(acoustic), update<elastic_psv>,
 * corrector(elastic_psv)) they couple the media as the fused step does:
 * update<acoustic> reads the solid's predicted displacement on the
 * interface, update<elastic_psv> the fluid's new acceleration. */
int smx_predictor_medium(smx_ctx *ctx, int field, int medium, int *dofs);
/* Degrees of freedom of a 2D context: nglob x components summed over its
 * media (assembly::get_total_degrees_of_freedom, which the reference's loop
 * checks the operators' counts against, time_marching.tpp:22-23,97-115). */
int smx_ctx_degrees_of_freedom(const smx_ctx *ctx, long long *dofs);
int smx_corrector_medium(smx_ctx *ctx, int field, int medium, int *dofs);
int smx_update_wavefields_medium(smx_ctx *ctx, int field, int medium,
                                 int istep, int *nelem);

/* Seismograms of the current fields into sample isig (both media of a
 * coupled mesh).
 * Replaces domain_kernels::compute_seismograms(isig)
 * (include/kokkos_kernels/domain_kernels.hpp:188-203). */
int smx_compute_seismograms(smx_ctx *ctx, int field, int isig);

/* nstep forward steps from istep0 on the fused path (stiffness + sources,
 * fused divide/corrector/predictor, seismograms every
 * nstep_between_samples), replayed from a HIP graph.  Returns after the work
 * is enqueued; smx_synchronize waits for it.
 * Replaces time_marching<forward>::run's loop
 * (core/specfem/solver/time_marching.tpp:33-116). */
int smx_run_forward(smx_ctx *ctx, int istep0, int nstep);

/* Waits for all work queued on the context's stream. */
int smx_synchronize(smx_ctx *ctx);

/* Copies seismograms [nsig][ntypes][nrec][2] to host (element size 4 or 8
 * bytes = the context's precision).  Replaces the receivers'
 * seismogram_components host sync (newmark_tests.cpp:185). */
int smx_get_seismograms(smx_ctx *ctx, void *host_out);

/* Copies one field array [nglob][ncomp] of `field` to / from host
 * (element size = the context's precision). */
int smx_get_field(smx_ctx *ctx, int field, int which, void *host_out);
int smx_set_field(smx_ctx *ctx, int field, int which, const void *host_in);

/* Global numbering [nspec][5][5] (iz, ix) the context assembled
 * (assembly::mesh index_mapping, core/specfem/assembly/mesh/dim2/impl/assemble.tpp). */
int smx_get_ibool(smx_ctx *ctx, int *host_out);

/* Device pointer of a field array [nglob][ncomp] (for zero-copy interop). */
int smx_field_device_ptr(smx_ctx *ctx, int field, int which, void **dptr);

/* Elapsed milliseconds of the last smx_run_forward, measured with HIP events
 * on the context's stream, and the number of stiffness launches it made. */
int smx_last_run_stats(smx_ctx *ctx, double *ms, double *stiffness_ms,
                       int *stiffness_launches);

/* Launches the stiffness kernel alone nlaunch times on the context's stream
 * (into a scratch acceleration, leaving the fields untouched) and returns the
 * mean launch duration in ms measured with HIP events.  Measurement aid with
 * no reference counterpart (the reference reports Kokkos-Tools region times
 * for compute_stiffness_interaction.tpp:121,250). */
int smx_time_stiffness(smx_ctx *ctx, int nlaunch, double *avg_ms);

/* ---- adjoint / Frechet path ------------------------------------------ */

/* Replaces the material at every GLL point (kappa, mu, rho, each
 * [nspec][5][5] in mesh element order; [nspec][5][5][5] for a 3D context),
 * refreshes every derived array
 * (lambda+2mu, the Stacey impedances) and rebuilds the mass matrix.
 * Checkpoints of an earlier run are dropped.  Replaces reassembling
 * specfem::assembly::properties (core/specfem/assembly/properties) for a new
 * model. */
int smx_set_model(smx_ctx *ctx, const float *kappa, const float *mu,
                  const float *rho);

/* smx_set_model with kappa, mu, rho in device memory ([nspec][5][5] each,
 * mesh element order, FP32), for callers whose model already lives on the
 * GPU (a torch.autograd.Function over CUDA tensors): validated and turned
 * into every derived array on the device, bitwise as smx_set_model does on
 * the host; the arrays must stay valid until the call returns.  2D elastic
 * FP32 single-mesh contexts (SMX_ERR_STATE otherwise).  No reference
 * counterpart beyond smx_set_model's. */
int smx_set_model_device(smx_ctx *ctx, const float *kappa, const float *mu,
                         const float *rho);

/* smx_get_gradient into device memory: [3][nspec][5][5] double, bitwise as
 * the host path forms it.  2D elastic FP32 contexts. */
int smx_get_gradient_device(smx_ctx *ctx, double *device_out);

/* Forward run of all nstep steps from rest, with seismograms as
 * smx_run_forward, keeping (u, v) in device memory every `every` steps.
 * Replaces the forward half of a combined simulation, which writes the last
 * wavefield and the Stacey boundary history instead
 * (include/io/wavefield/writer.tpp:114-200). */
int smx_run_forward_checkpointed(smx_ctx *ctx, int every);

/* Discrete-adjoint gradient of a functional chi of the displacement
 * seismograms, given dchi/ds = gsrc[nstep][nrec][2] ([nstep][nrec][3] for a
 * 3D context, whose forward fields the recompute overwrites) (element size = the
 * context's precision; row k is the seismogram after step k).  Recomputes
 * each checkpoint segment (newest first), recording its states, then runs
 * the adjoint steps of the segment backwards, accumulating the kappa, mu and
 * rho gradients.  Needs a preceding smx_run_forward_checkpointed for the
 * current model; returns after the run.  Replaces the combined loop
 * time_marching<combined>::run (core/specfem/solver/time_marching.tpp:133-264)
 * with frechet_kernels::compute_derivatives
 * (include/kokkos_kernels/frechet_kernels.hpp:39-53). */
int smx_run_adjoint(smx_ctx *ctx, const void *gsrc);

/* Gradient [3][nspec][5][5] (d chi / d kappa, d mu, d rho at every GLL
 * point, mesh element order; [3][nspec][5][5][5] in 3D) of the last
 * smx_run_adjoint, as double.
 * Replaces reading assembly::kernels (core/specfem/assembly/kernels). */
int smx_get_gradient(smx_ctx *ctx, double *host_out);

/* Elapsed milliseconds (HIP events) of the last checkpointed forward run and
 * of the last adjoint run (recompute included). */
int smx_last_adjoint_stats(smx_ctx *ctx, double *forward_ms,
                           double *adjoint_ms);

#ifdef __cplusplus
}
#endif

#endif /* SPECFEMPP_AMD_H */
