"""Host-side mirror of the reference's time-stepping operator surface.

Class and method names follow the C++ solver so a driver written against
SPECFEM++ reads the same:

* ``DomainKernels``  — specfem::kokkos_kernels::domain_kernels
  (include/kokkos_kernels/domain_kernels.hpp:37-213)
* ``Newmark``        — specfem::time_scheme::newmark<..., forward>
  (core/specfem/timescheme/newmark.hpp:106-231, timescheme.hpp:16-201)
* ``TimeMarching``   — specfem::solver::time_marching<forward, dim2, 5>
  (core/specfem/solver/time_marching.tpp:8-131)

Every operator runs on the GPU through libspecfempp_amd (C ABI); the host
only sequences calls.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import SMX_A, SMX_FIELD_FORWARD, SMX_MASS_INVERSE, SMX_U, SMX_V, check


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


def _structs(mesh, sources, receivers, seismo_types, nsbs, keep):
    """ctypes descriptions of one mesh / source set / receiver set; the
    arrays they point to are appended to ``keep``."""
    mesh.validate()

    def arr(x, dt_):
        a = np.ascontiguousarray(np.asarray(x, dtype=dt_))
        keep.append(a)
        return a

    m = _lib.SmxMesh2D(mesh.nspec, mesh.ngnod, mesh.nnodes,
                       _p(arr(mesh.coord, np.float64)), _p(arr(mesh.knods, np.int32)),
                       _p(arr(mesh.kappa, np.float32)), _p(arr(mesh.mu, np.float32)),
                       _p(arr(mesh.rho, np.float32)), len(mesh.abs_ispec),
                       _p(arr(mesh.abs_ispec, np.int32)), _p(arr(mesh.abs_type, np.int32)),
                       len(mesh.free_ispec), _p(arr(mesh.free_ispec, np.int32)),
                       _p(arr(mesh.free_type, np.int32)))
    s = _lib.SmxSources2D(len(sources), _p(arr(sources.x, np.float64)),
                          _p(arr(sources.z, np.float64)), _p(arr(sources.angle, np.float32)),
                          _p(arr(sources.f0, np.float32)), _p(arr(sources.tshift, np.float32)),
                          _p(arr(sources.factor, np.float32)), _p(arr(sources.kinds(), np.int32)),
                          _p(arr(_source_kinds(sources), np.int32)),
                          _p(arr(_moments(sources), np.float32)))
    types = arr([_lib.SEIS_TYPES[t] for t in seismo_types], np.int32)
    r = _lib.SmxReceivers2D(len(receivers), _p(arr(receivers.x, np.float64)),
                            _p(arr(receivers.z, np.float64)), len(types), _p(types), int(nsbs))
    return m, s, r


def _source_kinds(sources):
    f = getattr(sources, "source_kinds", None)
    return f() if f else np.zeros(len(sources.x), np.int32)


def _moments(sources):
    f = getattr(sources, "moments", None)
    return f() if f else np.zeros((len(sources.x), 3), np.float32)


def _options(options):
    if options is None:
        return None
    o = _lib.SmxOptions(int(options.get("tile_elems", 0)), int(options.get("graph_steps", 0)),
                        int(bool(options.get("geometry_from_nodes", False))),
                        int(options.get("waves_per_eu", 0)))
    return ctypes.byref(o)


class Context:
    """An smx_ctx: the device-resident assembly, fields and seismograms.

    ``options`` (dict, optional): ``tile_elems`` (elements per tile, 1..64),
    ``graph_steps`` (time steps per replayed graph) and ``waves_per_eu``
    (occupancy variant of the FP32 time-step tile kernel: 0 default, 6, 7,
    8), as chosen by ``specfempp_amd.autotune``, and ``geometry_from_nodes``
    (re-form the per-point geometry from the control nodes when the device
    reproduces it bit for bit; off by default).  ``Context.batch`` builds one
    context over many independent shots."""

    def __init__(self, mesh, sources, receivers, dt, nstep, *, nsbs=1,
                 seismo_types=("displacement",), precision=4, options=None, _batch=None):
        lib = _lib.load()
        keep = []
        h = ctypes.c_void_p()
        if _batch is None:
            m, s, r = _structs(mesh, sources, receivers, seismo_types, nsbs, keep)
            check(lib.smx_ctx_create2d_opts(ctypes.byref(m), ctypes.byref(s), ctypes.byref(r),
                                            float(dt), int(nstep), int(precision),
                                            _options(options), ctypes.byref(h)))
        else:
            parts = [_structs(c.mesh, c.sources, c.receivers, seismo_types, nsbs, keep)
                     for c in _batch]
            n = len(parts)
            M = (_lib.SmxMesh2D * n)(*[p[0] for p in parts])
            S = (_lib.SmxSources2D * n)(*[p[1] for p in parts])
            R = (_lib.SmxReceivers2D * n)(*[p[2] for p in parts])
            check(lib.smx_ctx_create2d_batch(n, M, S, R, float(dt), int(nstep), int(precision),
                                             _options(options), ctypes.byref(h)))
        del keep
        self.lib, self.h = lib, h
        self.precision = precision
        self.options = dict(options or {})
        self.dtype = np.float32 if precision == 4 else np.float64
        self.dt, self.nstep, self.nsbs = float(dt), int(nstep), int(nsbs)
        self.seismo_types = tuple(seismo_types)
        ns, ng, nsig, nr = (ctypes.c_int() for _ in range(4))
        check(lib.smx_ctx_info(h, ctypes.byref(ns), ctypes.byref(ng), ctypes.byref(nsig),
                               ctypes.byref(nr)))
        self.nspec, self.nglob, self.nsig, self.nrec = ns.value, ng.value, nsig.value, nr.value
        nc = ctypes.c_int()
        check(lib.smx_ctx_components(h, ctypes.byref(nc)))
        self.ncomp = nc.value  # 2 elastic P-SV, 1 acoustic potential
        self.t0 = lib.smx_ctx_t0(h)
        nt, nx, nsh, npart = (ctypes.c_int() for _ in range(4))
        check(lib.smx_ctx_layout(h, ctypes.byref(nt), ctypes.byref(nx), ctypes.byref(nsh),
                                 ctypes.byref(npart)))
        gel, mel = ctypes.c_int(), ctypes.c_int()
        check(lib.smx_ctx_compact(h, ctypes.byref(gel), ctypes.byref(mel)))
        self.layout = {"ntiles": nt.value, "nexclusive": nx.value, "nshared": nsh.value,
                       "npartials": npart.value, "geom_elem_elements": gel.value,
                       "mat_elem_elements": mel.value}
        nsh_ = ctypes.c_int()
        check(lib.smx_ctx_shots(h, ctypes.byref(nsh_), None, None, None))
        k = nsh_.value
        eo, ro, to = (np.zeros(k + 1, np.int32) for _ in range(3))
        check(lib.smx_ctx_shots(h, ctypes.byref(nsh_), _p(eo), _p(ro), _p(to)))
        self.nshots = k
        self.shot_elem_off, self.shot_rec_off, self.shot_tile_off = eo, ro, to

    @classmethod
    def batch(cls, configs, *, precision=4, options=None):
        """One context advancing every Config2D of ``configs`` (same dt,
        nstep, seismogram types and sampling) with one launch per kernel per
        step."""
        c0 = configs[0]
        for c in configs:
            if (c.dt, c.nstep, c.nsbs, tuple(c.seismo_types)) != \
                    (c0.dt, c0.nstep, c0.nsbs, tuple(c0.seismo_types)):
                raise ValueError("batched shots must share dt, nstep and the seismogram set-up")
        return cls(None, None, None, c0.dt, c0.nstep, nsbs=c0.nsbs,
                   seismo_types=c0.seismo_types, precision=precision, options=options,
                   _batch=list(configs))

    def shot_seismograms(self, seis=None):
        """Per-shot views (nsig, ntypes, nrec_shot, 2) of the seismograms."""
        seis = self.seismograms() if seis is None else seis
        o = self.shot_rec_off
        return [seis[:, :, o[k]:o[k + 1]] for k in range(self.nshots)]

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.smx_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- raw operators --------------------------------------------------
    def initialize(self):
        check(self.lib.smx_initialize(self.h))

    def synchronize(self):
        check(self.lib.smx_synchronize(self.h))

    def run_forward(self, istep0, nstep):
        check(self.lib.smx_run_forward(self.h, int(istep0), int(nstep)))

    def seismograms(self):
        """(nsig, ntypes, nrec, 2) array of the recorded seismograms."""
        out = np.empty((self.nsig, len(self.seismo_types), self.nrec, 2), dtype=self.dtype)
        check(self.lib.smx_get_seismograms(self.h, _p(out)))
        return out

    def get_field(self, which, field=SMX_FIELD_FORWARD):
        idx = {"u": SMX_U, "v": SMX_V, "a": SMX_A, "minv": SMX_MASS_INVERSE}[which]
        out = np.empty((self.nglob, max(self.ncomp, 1)), dtype=self.dtype)
        check(self.lib.smx_get_field(self.h, field, idx, _p(out)))
        return out if self.ncomp > 1 else out[:, 0]

    def set_field(self, which, values, field=SMX_FIELD_FORWARD):
        idx = {"u": SMX_U, "v": SMX_V, "a": SMX_A}[which]
        v = np.ascontiguousarray(np.asarray(values, dtype=self.dtype)
                                 .reshape(self.nglob, self.ncomp))
        check(self.lib.smx_set_field(self.h, field, idx, _p(v)))

    def ibool(self):
        """(nspec, 5, 5) global numbering."""
        out = np.empty((self.nspec, 5, 5), dtype=np.int32)
        check(self.lib.smx_get_ibool(self.h, _p(out)))
        return out

    def last_run_ms(self):
        ms, sms, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(self.lib.smx_last_run_stats(self.h, ctypes.byref(ms), ctypes.byref(sms),
                                          ctypes.byref(n)))
        return ms.value

    def set_launch(self, graph_steps, waves_per_eu=0):
        """New graph length and tile-kernel occupancy variant (smx_ctx_set_launch);
        the step graph is re-captured here, results are bitwise unchanged."""
        check(self.lib.smx_ctx_set_launch(self.h, int(graph_steps), int(waves_per_eu)))

    def time_stiffness(self, nlaunch=20):
        t = ctypes.c_double()
        check(self.lib.smx_time_stiffness(self.h, int(nlaunch), ctypes.byref(t)))
        return t.value

    # ---- adjoint / Frechet path ------------------------------------------

    def set_model(self, kappa, mu, rho):
        """Replaces the material (each (nspec, 5, 5)) and rebuilds the mass."""
        k, m, r = (np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(self.nspec, 5, 5))
                   for x in (kappa, mu, rho))
        check(self.lib.smx_set_model(self.h, _p(k), _p(m), _p(r)))

    def set_model_device(self, kappa_ptr, mu_ptr, rho_ptr):
        """set_model from device arrays (addresses of contiguous float32
        [nspec][5][5] arrays in GPU memory, e.g. CUDA tensors' data_ptr());
        2D elastic FP32 contexts."""
        check(self.lib.smx_set_model_device(self.h, ctypes.c_void_p(kappa_ptr),
                                            ctypes.c_void_p(mu_ptr), ctypes.c_void_p(rho_ptr)))

    def gradient_device(self, out_ptr):
        """The gradient written into device memory at out_ptr: contiguous
        float64 (3, nspec, 5, 5)."""
        check(self.lib.smx_get_gradient_device(self.h, ctypes.c_void_p(out_ptr)))

    def run_forward_checkpointed(self, every=100):
        """All nstep steps from rest, keeping the state every ``every`` steps."""
        check(self.lib.smx_run_forward_checkpointed(self.h, int(every)))

    def run_adjoint(self, gsrc):
        """Adjoint run for dchi/ds = gsrc (nstep, nrec, 2), s the displacement
        seismogram after each step."""
        g = np.ascontiguousarray(np.asarray(gsrc, dtype=self.dtype).reshape(self.nstep, self.nrec, 2))
        check(self.lib.smx_run_adjoint(self.h, _p(g)))

    def gradient(self):
        """(3, nspec, 5, 5) float64: d chi / d (kappa, mu, rho) per GLL point."""
        out = np.empty((3, self.nspec, 5, 5), dtype=np.float64)
        check(self.lib.smx_get_gradient(self.h, _p(out)))
        return out

    def last_adjoint_ms(self):
        """(checkpointed forward ms, adjoint ms incl. recompute) of the last runs."""
        f, a = ctypes.c_double(), ctypes.c_double()
        check(self.lib.smx_last_adjoint_stats(self.h, ctypes.byref(f), ctypes.byref(a)))
        return f.value, a.value


class DomainKernels:
    """specfem::kokkos_kernels::domain_kernels<forward, dim2, 5>
    (include/kokkos_kernels/domain_kernels.hpp:37-213)."""

    def __init__(self, ctx: Context, field=SMX_FIELD_FORWARD):
        self.ctx, self.field = ctx, field

    def initialize(self, dt=None):
        # dt is fixed at assembly time (the Stacey mass term uses it)
        self.ctx.initialize()

    def update_wavefields(self, istep, medium=None) -> int:
        """update_wavefields<medium>(istep): the elements of ``medium``
        updated (0 for a medium the mesh does not hold).  Without a medium:
        the context's single medium."""
        n = ctypes.c_int()
        if medium is None:
            check(self.ctx.lib.smx_update_wavefields(self.ctx.h, self.field, int(istep),
                                                     ctypes.byref(n)))
        else:
            check(self.ctx.lib.smx_update_wavefields_medium(
                self.ctx.h, self.field, _lib.MEDIUM_TAGS[medium], int(istep), ctypes.byref(n)))
        return n.value

    def compute_seismograms(self, isig):
        check(self.ctx.lib.smx_compute_seismograms(self.ctx.h, self.field, int(isig)))


class Newmark:
    """Forward Newmark time scheme (core/specfem/timescheme/newmark.hpp,
    time_scheme interface timescheme.hpp:16-201)."""

    def __init__(self, ctx: Context, field=SMX_FIELD_FORWARD):
        self.ctx, self.field = ctx, field
        self.nstep, self.nsbs = ctx.nstep, ctx.nsbs
        self.isig = 0

    def iterate_forward(self):
        for istep in range(self.nstep):
            yield istep, self.ctx.dt

    def apply_predictor_phase_forward(self, medium=None) -> int:
        n = ctypes.c_int()
        if medium is None:
            check(self.ctx.lib.smx_predictor(self.ctx.h, self.field, ctypes.byref(n)))
        else:
            check(self.ctx.lib.smx_predictor_medium(self.ctx.h, self.field,
                                                    _lib.MEDIUM_TAGS[medium], ctypes.byref(n)))
        return n.value

    def apply_corrector_phase_forward(self, medium=None) -> int:
        n = ctypes.c_int()
        if medium is None:
            check(self.ctx.lib.smx_corrector(self.ctx.h, self.field, ctypes.byref(n)))
        else:
            check(self.ctx.lib.smx_corrector_medium(self.ctx.h, self.field,
                                                    _lib.MEDIUM_TAGS[medium], ctypes.byref(n)))
        return n.value

    def compute_seismogram(self, istep) -> bool:
        return istep % self.nsbs == 0

    def get_seismogram_step(self) -> int:
        return self.isig

    def increment_seismogram_step(self):
        self.isig += 1

    def get_timestep(self) -> float:
        return self.ctx.dt

    def get_max_timestep(self) -> int:
        return self.nstep

    def get_max_seismogram_step(self) -> int:
        return self.ctx.nsig


class TimeMarching:
    """specfem::solver::time_marching<forward, dim2, 5>.

    ``run(fused=False)`` calls the medium-tagged operators one by one in the
    reference loop's order (core/specfem/solver/time_marching.tpp:33-116),
    for every medium tag, with the same DOF / element-count checks; it runs
    single-medium and coupled meshes.  ``run(fused=True)`` (default)
    enqueues the fused, graph-replayed step."""

    PREDICTED = ("acoustic", "elastic", "elastic_psv", "elastic_sh", "poroelastic",
                 "elastic_psv_t")
    SOLID = ("elastic", "elastic_psv", "elastic_sh", "elastic_psv_t")

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.kernels = DomainKernels(ctx)
        self.time_scheme = Newmark(ctx)

    def total_degrees_of_freedom(self) -> int:
        n = ctypes.c_longlong()
        check(self.ctx.lib.smx_ctx_degrees_of_freedom(self.ctx.h, ctypes.byref(n)))
        return n.value

    def run(self, fused=True):
        ctx = self.ctx
        self.kernels.initialize(ctx.dt)
        if fused:
            ctx.run_forward(0, ctx.nstep)
            ctx.synchronize()
            return
        ts, k = self.time_scheme, self.kernels
        total_dofs = 2 * self.total_degrees_of_freedom()
        for istep, _ in ts.iterate_forward():
            dofs = sum(ts.apply_predictor_phase_forward(m) for m in self.PREDICTED)
            nelem = k.update_wavefields(istep, "acoustic")
            dofs += ts.apply_corrector_phase_forward("acoustic")
            nelem += sum(k.update_wavefields(istep, m) for m in self.SOLID)
            dofs += sum(ts.apply_corrector_phase_forward(m) for m in self.SOLID)
            nelem += k.update_wavefields(istep, "poroelastic")
            dofs += ts.apply_corrector_phase_forward("poroelastic")
            if ts.compute_seismogram(istep):
                k.compute_seismograms(ts.get_seismogram_step())
                ts.increment_seismogram_step()
            if dofs != total_dofs:
                raise RuntimeError(f"The time loop has not updated all the degrees of "
                                   f"freedom. Only {dofs} out of {total_dofs}.")
            if nelem != ctx.nspec:
                raise RuntimeError(f"The time loop has not updated all the elements. "
                                   f"Only {nelem} out of {ctx.nspec}.")
        ctx.synchronize()


def simulate(config, *, precision=4, fused=True, nstep=None, options=None):
    """Runs a forward simulation of a Config2D; returns (seismograms, ctx)."""
    n = config.nstep if nstep is None else int(nstep)
    ctx = Context(config.mesh, config.sources, config.receivers, config.dt, n,
                  nsbs=config.nsbs, seismo_types=config.seismo_types, precision=precision,
                  options=options)
    TimeMarching(ctx).run(fused=fused)
    return ctx.seismograms(), ctx


def simulate_batch(configs, *, precision=4, options=None):
    """Runs every Config2D of ``configs`` in one batched context; returns
    (list of per-shot seismograms (nsig, ntypes, nrec_shot, 2), ctx)."""
    ctx = Context.batch(configs, precision=precision, options=options)
    ctx.initialize()
    ctx.run_forward(0, ctx.nstep)
    ctx.synchronize()
    return ctx.shot_seismograms(), ctx
