"""Host-side mirror of the 3D forward path (time_marching<forward, dim3>
with the dim3 domain_kernels, core/specfem/solver/time_marching.tpp:8-131).

Every operator runs on the GPU through libspecfempp_amd (C ABI,
``smx_ctx_create3d``); the host only sequences calls.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import SMX_A, SMX_FIELD_FORWARD, SMX_MASS_INVERSE, SMX_U, SMX_V, check
from .solver import _options


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


class Context3D:
    """A 3D smx_ctx: hex8 mesh, sources, receivers and fields on the GPU.

    ``options`` (dict, optional): ``graph_steps`` and ``geometry_from_nodes``
    as for the 2D ``Context``."""

    def __init__(self, mesh, sources, receivers, dt, nstep, *, nsbs=1,
                 seismo_types=("displacement",), precision=4, options=None):
        lib = _lib.load()
        mesh.validate()
        keep = []

        def arr(x, dt_):
            a = np.ascontiguousarray(np.asarray(x, dtype=dt_))
            keep.append(a)
            return a

        m = _lib.SmxMesh3D(mesh.nspec, mesh.nnodes, _p(arr(mesh.coord, np.float64)),
                           _p(arr(mesh.knods, np.int32)), _p(arr(mesh.kappa, np.float32)),
                           _p(arr(mesh.mu, np.float32)), _p(arr(mesh.rho, np.float32)),
                           len(mesh.abs_ispec), _p(arr(mesh.abs_ispec, np.int32)),
                           _p(arr(mesh.abs_face, np.int32)))
        s = _lib.SmxSources3D(len(sources), _p(arr(sources.x, np.float64)),
                              _p(arr(sources.y, np.float64)), _p(arr(sources.z, np.float64)),
                              _p(arr(sources.kind, np.int32)), _p(arr(sources.comp, np.float32)),
                              _p(arr(sources.f0, np.float32)), _p(arr(sources.tshift, np.float32)),
                              _p(arr(sources.factor, np.float32)))
        types = arr([_lib.SEIS_TYPES[t] for t in seismo_types], np.int32)
        r = _lib.SmxReceivers3D(len(receivers), _p(arr(receivers.x, np.float64)),
                                _p(arr(receivers.y, np.float64)), _p(arr(receivers.z, np.float64)),
                                len(types), _p(types), int(nsbs))
        h = ctypes.c_void_p()
        check(lib.smx_ctx_create3d_opts(ctypes.byref(m), ctypes.byref(s), ctypes.byref(r),
                                        float(dt), int(nstep), int(precision),
                                        _options(options), ctypes.byref(h)))
        del keep
        self.lib, self.h = lib, h
        self.precision = precision
        self.dtype = np.float32 if precision == 4 else np.float64
        self.dt, self.nstep, self.nsbs = float(dt), int(nstep), int(nsbs)
        self.seismo_types = tuple(seismo_types)
        ns, ng, nsig, nr = (ctypes.c_int() for _ in range(4))
        check(lib.smx_ctx_info(h, ctypes.byref(ns), ctypes.byref(ng), ctypes.byref(nsig),
                               ctypes.byref(nr)))
        self.nspec, self.nglob, self.nsig, self.nrec = ns.value, ng.value, nsig.value, nr.value
        self.t0 = lib.smx_ctx_t0(h)
        nt, nx, nsh, npart = (ctypes.c_int() for _ in range(4))
        check(lib.smx_ctx_layout(h, ctypes.byref(nt), ctypes.byref(nx), ctypes.byref(nsh),
                                 ctypes.byref(npart)))
        gel, mel = ctypes.c_int(), ctypes.c_int()
        check(lib.smx_ctx_compact(h, ctypes.byref(gel), ctypes.byref(mel)))
        self.layout = {"ntiles": nt.value, "nexclusive": nx.value, "nshared": nsh.value,
                       "npartials": npart.value, "geom_elem_elements": gel.value,
                       "mat_elem_elements": mel.value}
        gm, mis = ctypes.c_int(), ctypes.c_longlong()
        check(lib.smx_ctx_geometry(h, ctypes.byref(gm), ctypes.byref(mis)))
        self.geometry = {"mode": ("point", "element", "nodes")[gm.value],
                         "mismatches": mis.value}

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.smx_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self):
        check(self.lib.smx_initialize(self.h))

    def run_forward(self, istep0, nstep):
        check(self.lib.smx_run_forward(self.h, int(istep0), int(nstep)))

    def synchronize(self):
        check(self.lib.smx_synchronize(self.h))

    def seismograms(self):
        """(nsig, ntypes, nrec, 3) array of the recorded seismograms."""
        out = np.empty((self.nsig, len(self.seismo_types), self.nrec, 3), dtype=self.dtype)
        check(self.lib.smx_get_seismograms(self.h, _p(out)))
        return out

    def get_field(self, which):
        idx = {"u": SMX_U, "v": SMX_V, "a": SMX_A, "minv": SMX_MASS_INVERSE}[which]
        out = np.empty((self.nglob, 3), dtype=self.dtype)
        check(self.lib.smx_get_field(self.h, SMX_FIELD_FORWARD, idx, _p(out)))
        return out

    def set_field(self, which, values):
        idx = {"u": SMX_U, "v": SMX_V, "a": SMX_A}[which]
        v = np.ascontiguousarray(np.asarray(values, dtype=self.dtype).reshape(self.nglob, 3))
        check(self.lib.smx_set_field(self.h, SMX_FIELD_FORWARD, idx, _p(v)))

    def last_run_ms(self):
        ms, sms, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(self.lib.smx_last_run_stats(self.h, ctypes.byref(ms), ctypes.byref(sms),
                                          ctypes.byref(n)))
        return ms.value

    def time_stiffness(self, nlaunch=10):
        t = ctypes.c_double()
        check(self.lib.smx_time_stiffness(self.h, int(nlaunch), ctypes.byref(t)))
        return t.value

    # ---- adjoint (same contract as the 2D Context)
    ncomp = 3

    def set_model(self, kappa, mu, rho):
        """Replaces the material (each (nspec, 5, 5, 5)) and rebuilds the mass."""
        k, m, r = (np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(self.nspec, 125))
                   for x in (kappa, mu, rho))
        check(self.lib.smx_set_model(self.h, _p(k), _p(m), _p(r)))

    def run_forward_checkpointed(self, every=100):
        """All nstep steps from rest, keeping the state every ``every`` steps."""
        check(self.lib.smx_run_forward_checkpointed(self.h, int(every)))

    def run_adjoint(self, gsrc):
        """Adjoint run for dchi/ds = gsrc (nstep, nrec, 3), s the displacement
        seismogram after each step.  The recompute overwrites the forward
        fields (the seismograms stay)."""
        g = np.ascontiguousarray(np.asarray(gsrc, dtype=self.dtype).reshape(self.nstep, self.nrec, 3))
        check(self.lib.smx_run_adjoint(self.h, _p(g)))

    def gradient(self):
        """(3, nspec, 5, 5, 5) float64: d chi / d (kappa, mu, rho) per GLL point."""
        out = np.empty((3, self.nspec, 5, 5, 5), dtype=np.float64)
        check(self.lib.smx_get_gradient(self.h, _p(out)))
        return out

    def last_adjoint_ms(self):
        """(checkpointed forward ms, adjoint ms incl. recompute) of the last runs."""
        f, a = ctypes.c_double(), ctypes.c_double()
        check(self.lib.smx_last_adjoint_stats(self.h, ctypes.byref(f), ctypes.byref(a)))
        return f.value, a.value


def simulate3d(config, *, precision=4, nstep=None, options=None):
    """Runs a forward simulation of a Config3D; returns (seismograms, ctx)."""
    n = config.nstep if nstep is None else int(nstep)
    ctx = Context3D(config.mesh, config.sources, config.receivers, config.dt, n,
                    nsbs=config.nsbs, seismo_types=config.seismo_types, precision=precision,
                    options=options)
    ctx.initialize()
    ctx.run_forward(0, n)
    ctx.synchronize()
    return ctx.seismograms(), ctx
