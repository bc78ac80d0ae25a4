"""ORACLE — test infrastructure only.

ctypes wrapper over oracle/build/liboracle2d.so, the CPU restatement of the
reference's 2D elastic P-SV forward path (oracle/cpu_ref2d.cpp).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker / CPU baseline; the specfempp_amd
package never does.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle2d.so")
_lib = None

SEIS_TYPES = {"displacement": 0, "velocity": 1, "acceleration": 2, "pressure": 3}


def build(force: bool = False) -> str:
    """Compiles the oracle with make (g++, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH) or not os.path.exists(
            os.path.join(_HERE, "build", "liboracle3d.so")):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    i, f, d = ctypes.c_int, ctypes.c_float, ctypes.c_double
    lib.oracle2d_create.restype = P
    lib.oracle2d_create.argtypes = [i, i, i, i, P, P, P, P, P, i, P, P, f, i, i, P, P, P, P, P, P, i, P, P]
    lib.oracle2d_create_ex.restype = P
    lib.oracle2d_create_ex.argtypes = [i, i, i, i, P, P, P, P, P, i, P, P, i, P, P, f, i, i, P, P,
                                       P, P, P, P, P, i, P, P]
    lib.oracle2d_create_mt.restype = P
    lib.oracle2d_create_mt.argtypes = lib.oracle2d_create_ex.argtypes + [P, P]
    lib.oracle2d_destroy.argtypes = [P]
    lib.oracle2d_nglob.argtypes = [P]
    lib.oracle2d_t0.argtypes = [P]
    lib.oracle2d_t0.restype = d
    lib.oracle2d_run_forward.argtypes = [P, i, i, i, i, P, P]
    lib.oracle2d_get.argtypes = [P, i, P]
    lib.oracle2d_get.restype = ctypes.c_long
    lib.oracle2d_set_field.argtypes = [P, i, P]
    lib.oracle2d_apply_stiffness.argtypes = [P, i]
    lib.oracle2d_forward_store.argtypes = [P, i, P]
    lib.oracle2d_adjoint.argtypes = [P, i, P, P]
    lib.oracle2d_ref_kernels.argtypes = [P, i, P, P, P]
    lib.oracle2d_time_parallel.argtypes = [P, i, i, i, P]
    lib.oracle2d_coupled_create.restype = P
    lib.oracle2d_coupled_create.argtypes = lib.oracle2d_create_ex.argtypes
    lib.oracle2d_coupled_t0.restype = d
    lib.oracle2d_coupled_t0.argtypes = [P]
    lib.oracle2d_coupled_run.argtypes = [P, i, i, i, i, P, P]
    lib.oracle2d_coupled_destroy.argtypes = [P]
    lib.oracle2d_coupled_scale.argtypes = [P, d, d]
    lib.oracle2d_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


GET = {"minv": 0, "u": 1, "v": 2, "a": 3, "ibool": 4, "jacobian": 5, "xix": 6, "xiz": 7,
       "gammax": 8, "gammaz": 9, "xi": 10, "w": 11, "hprime": 12, "stf": 13,
       "source_array": 14, "receiver_h": 15, "source_ispec": 16, "receiver_ispec": 17,
       "receiver_xi": 18, "receiver_gamma": 19, "stacey_flag": 20, "stacey_n0": 21,
       "stacey_n1": 22, "stacey_w": 23}


class Oracle2D:
    """One forward simulation on the CPU (precision 4 = FP32, 8 = FP64).
    A mesh with both elastic and acoustic elements runs as the coupled
    restatement (run() and t0 only)."""

    def __init__(self, mesh, sources, receivers, dt, nstep, precision=4):
        lib = _load()
        self._keep = []

        def arr(x, dt_):
            a = np.ascontiguousarray(np.asarray(x, dtype=dt_))
            self._keep.append(a)
            return a

        coord = arr(mesh.coord, np.float64)
        knods = arr(mesh.knods, np.int32)
        kappa, mu, rho = (arr(mesh.kappa, np.float32), arr(mesh.mu, np.float32),
                          arr(mesh.rho, np.float32))
        ai, at = arr(mesh.abs_ispec, np.int32), arr(mesh.abs_type, np.int32)
        sx, sz = arr(sources.x, np.float64), arr(sources.z, np.float64)
        ang, f0 = arr(sources.angle, np.float32), arr(sources.f0, np.float32)
        tsh, fac = arr(sources.tshift, np.float32), arr(sources.factor, np.float32)
        rx, rz = arr(receivers.x, np.float64), arr(receivers.z, np.float64)
        fi = arr(getattr(mesh, "free_ispec", np.zeros(0)), np.int32)
        ft = arr(getattr(mesh, "free_type", np.zeros(0)), np.int32)
        kinds = arr(sources.kinds() if hasattr(sources, "kinds") else np.zeros(len(sx)),
                    np.int32)
        self.nrec = len(rx)
        self.nstep = int(nstep)
        self.nloc = mesh.nspec * 25
        fluid = (np.asarray(mesh.mu).reshape(mesh.nspec, -1) == 0).all(axis=1)
        self.acoustic = bool(fluid.all())
        self.coupled = bool(fluid.any() and not fluid.all())
        args = (precision, mesh.nspec, mesh.ngnod, mesh.nnodes,
                _ptr(coord), _ptr(knods), _ptr(kappa), _ptr(mu),
                _ptr(rho), len(ai), _ptr(ai), _ptr(at), len(fi),
                _ptr(fi), _ptr(ft), float(dt), int(nstep), len(sx),
                _ptr(sx), _ptr(sz), _ptr(ang), _ptr(f0), _ptr(tsh),
                _ptr(fac), _ptr(kinds), self.nrec, _ptr(rx), _ptr(rz))
        stype = getattr(sources, "source_kinds", lambda: None)()
        if stype is not None and stype.any():
            if self.coupled:
                raise ValueError("the coupled oracle takes point forces only")
            self.h = lib.oracle2d_create_mt(*args, _ptr(arr(stype, np.int32)),
                                            _ptr(arr(sources.moments(), np.float32)))
        else:
            create = lib.oracle2d_coupled_create if self.coupled else lib.oracle2d_create_ex
            self.h = create(*args)
        if not self.h:
            raise RuntimeError("oracle2d_create failed: " + lib.oracle2d_last_error().decode())
        self.lib = lib
        self.istep = 0

    def __del__(self):
        if getattr(self, "h", None):
            if self.coupled:
                self.lib.oracle2d_coupled_destroy(self.h)
            else:
                self.lib.oracle2d_destroy(self.h)
            self.h = None

    def plant_coupling_scales(self, fluid_side, solid_side):
        """Tests only: scale the two coupling terms of a coupled run (a
        planted mistake on the CPU)."""
        self.lib.oracle2d_coupled_scale(self.h, float(fluid_side), float(solid_side))

    @property
    def nglob(self):
        return self.lib.oracle2d_nglob(self.h)

    @property
    def t0(self):
        if self.coupled:
            return self.lib.oracle2d_coupled_t0(self.h)
        return self.lib.oracle2d_t0(self.h)

    def get(self, what):
        n = self.lib.oracle2d_get(self.h, GET[what], None)
        out = np.empty(n, dtype=np.float64)
        self.lib.oracle2d_get(self.h, GET[what], _ptr(out))
        return out

    def set_field(self, what, values):
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).ravel())
        self.lib.oracle2d_set_field(self.h, {"u": 1, "v": 2, "a": 3, "stf": 13}[what], _ptr(v))

    def apply_stiffness(self, with_stacey=True):
        self.lib.oracle2d_apply_stiffness(self.h, int(with_stacey))
        return self.get("a")

    def time_parallel(self, nstep, threads, istep0=0):
        """CPU-baseline timing only: wall seconds of `nstep` colour-parallel
        forward steps on `threads` OpenMP threads; returns (seconds, threads
        used)."""
        sec = np.zeros(1, dtype=np.float64)
        used = self.lib.oracle2d_time_parallel(self.h, int(istep0), int(nstep), int(threads),
                                               _ptr(sec))
        if used < 0:
            raise RuntimeError(self.lib.oracle2d_last_error().decode())
        return float(sec[0]), int(used)

    def forward_store(self, nstep=None):
        """Forward run from rest keeping every state for the adjoint;
        returns displacement seismograms (nstep, nrec, 2)."""
        n = self.nstep if nstep is None else int(nstep)
        out = np.zeros((n, self.nrec, 2), dtype=np.float64)
        if self.lib.oracle2d_forward_store(self.h, n, _ptr(out)) != 0:
            raise RuntimeError(self.lib.oracle2d_last_error().decode())
        self._stored = n
        return out

    def adjoint(self, gsrc):
        """Discrete-adjoint gradient (3, nspec, 5, 5) = d chi / d (kappa, mu,
        rho) for dchi/ds = gsrc (nstep, nrec, 2), after forward_store."""
        g = np.ascontiguousarray(np.asarray(gsrc, dtype=np.float64))
        n = g.shape[0]
        nloc = self.nloc
        out = np.zeros(3 * nloc, dtype=np.float64)
        if self.lib.oracle2d_adjoint(self.h, n, _ptr(g), _ptr(out)) != 0:
            raise RuntimeError(self.lib.oracle2d_last_error().decode())
        return out.reshape(3, -1, 5, 5)

    def ref_kernels(self, gsrc):
        """(gradient (3, nspec, 5, 5), kernels (6, nspec, 5, 5)): the
        discrete-adjoint gradient and the reference's Frechet-kernel formula
        on the same fields -- [0:3] with the reference's quirks (grad(s+ + s),
        kappa + mu/3), [3:6] corrected -- after forward_store."""
        g = np.ascontiguousarray(np.asarray(gsrc, dtype=np.float64))
        grad = np.zeros(3 * self.nloc, dtype=np.float64)
        k = np.zeros(6 * self.nloc, dtype=np.float64)
        if self.lib.oracle2d_ref_kernels(self.h, g.shape[0], _ptr(g), _ptr(grad), _ptr(k)) != 0:
            raise RuntimeError(self.lib.oracle2d_last_error().decode())
        return grad.reshape(3, -1, 5, 5), k.reshape(6, -1, 5, 5)

    def run(self, nstep=None, nsbs=1, types=("displacement",)):
        """Advances nstep steps; returns seismograms (nsig, ntypes, nrec, 2)."""
        nstep = self.nstep - self.istep if nstep is None else int(nstep)
        t = np.array([SEIS_TYPES[s] for s in types], dtype=np.int32)
        nsig = (nstep + nsbs - 1) // nsbs
        out = np.zeros((nsig, len(t), self.nrec, 2), dtype=np.float64)
        run = self.lib.oracle2d_coupled_run if self.coupled else self.lib.oracle2d_run_forward
        got = run(self.h, self.istep, nstep, nsbs, len(t), _ptr(t), _ptr(out))
        if got < 0:
            raise RuntimeError(self.lib.oracle2d_last_error().decode())
        self.istep += nstep
        return out[:got]


# ---------------------------------------------------------------- 3D oracle
_LIB3_PATH = os.path.join(_HERE, "build", "liboracle3d.so")
_lib3 = None


def _load3():
    global _lib3
    if _lib3 is not None:
        return _lib3
    if not os.path.exists(_LIB3_PATH):
        build(force=True)
    lib = ctypes.CDLL(_LIB3_PATH)
    P = ctypes.c_void_p
    i, f, d = ctypes.c_int, ctypes.c_float, ctypes.c_double
    lib.oracle3d_create.restype = P
    lib.oracle3d_create.argtypes = [i, i, i, P, P, P, P, P, i, P, P, f, i, i, P, P, P, P, P,
                                    P, P, P, i, P, P, P]
    lib.oracle3d_destroy.argtypes = [P]
    lib.oracle3d_nglob.argtypes = [P]
    lib.oracle3d_t0.argtypes = [P]
    lib.oracle3d_t0.restype = d
    lib.oracle3d_run.argtypes = [P, i, i, P, P]
    lib.oracle3d_get.argtypes = [P, i, P]
    lib.oracle3d_get.restype = ctypes.c_long
    lib.oracle3d_apply_stiffness.argtypes = [P, P, P, i]
    lib.oracle3d_forward_store.argtypes = [P, i, P]
    lib.oracle3d_adjoint.argtypes = [P, i, P, P]
    lib.oracle3d_last_error.restype = ctypes.c_char_p
    _lib3 = lib
    return lib


GET3 = {"minv": 0, "ibool": 1, "jacobian": 2, "source_array": 3, "receiver_h": 4,
        "u": 5, "v": 6, "a": 7, "inverse_jacobian": 8}


class Oracle3D:
    """One 3D forward simulation on the CPU (precision 4 = FP32, 8 = FP64)."""

    def __init__(self, mesh, sources, receivers, dt, nstep, precision=4):
        lib = _load3()
        self._keep = []

        def arr(x, dt_):
            a = np.ascontiguousarray(np.asarray(x, dtype=dt_))
            self._keep.append(a)
            return a

        self.nrec = len(receivers.x)
        self.nstep = int(nstep)
        self.h = lib.oracle3d_create(
            precision, mesh.nspec, mesh.nnodes, _ptr(arr(mesh.coord, np.float64)),
            _ptr(arr(mesh.knods, np.int32)), _ptr(arr(mesh.kappa, np.float32)),
            _ptr(arr(mesh.mu, np.float32)), _ptr(arr(mesh.rho, np.float32)),
            len(mesh.abs_ispec), _ptr(arr(mesh.abs_ispec, np.int32)),
            _ptr(arr(mesh.abs_face, np.int32)), float(dt), int(nstep), len(sources.x),
            _ptr(arr(sources.x, np.float64)), _ptr(arr(sources.y, np.float64)),
            _ptr(arr(sources.z, np.float64)), _ptr(arr(sources.kind, np.int32)),
            _ptr(arr(sources.comp, np.float32)), _ptr(arr(sources.f0, np.float32)),
            _ptr(arr(sources.tshift, np.float32)), _ptr(arr(sources.factor, np.float32)),
            self.nrec, _ptr(arr(receivers.x, np.float64)), _ptr(arr(receivers.y, np.float64)),
            _ptr(arr(receivers.z, np.float64)))
        if not self.h:
            raise RuntimeError("oracle3d_create failed: " + lib.oracle3d_last_error().decode())
        self.lib = lib

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle3d_destroy(self.h)
            self.h = None

    @property
    def nglob(self):
        return self.lib.oracle3d_nglob(self.h)

    @property
    def t0(self):
        return self.lib.oracle3d_t0(self.h)

    def get(self, what):
        n = self.lib.oracle3d_get(self.h, GET3[what], None)
        out = np.empty(n, dtype=np.float64)
        self.lib.oracle3d_get(self.h, GET3[what], _ptr(out))
        return out

    def apply_stiffness(self, u, v=None, with_stacey=True):
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).ravel())
        v = np.zeros_like(u) if v is None else np.ascontiguousarray(
            np.asarray(v, dtype=np.float64).ravel())
        self.lib.oracle3d_apply_stiffness(self.h, _ptr(u), _ptr(v), int(with_stacey))
        return self.get("a").reshape(-1, 3)

    def forward_store(self, nstep=None):
        """Displacement seismograms (nstep, nrec, 3) from rest, keeping every
        step's state for ``adjoint``."""
        n = self.nstep if nstep is None else int(nstep)
        out = np.zeros((n, self.nrec, 3), dtype=np.float64)
        if self.lib.oracle3d_forward_store(self.h, n, _ptr(out)) != 0:
            raise RuntimeError(self.lib.oracle3d_last_error().decode())
        self._stored = n
        return out

    def adjoint(self, gsrc):
        """Discrete-adjoint gradient (3, nspec, 5, 5, 5) of a functional with
        d chi / d s = gsrc (nstep, nrec, 3), after ``forward_store``."""
        g = np.ascontiguousarray(np.asarray(gsrc, dtype=np.float64))
        n = g.shape[0]
        nspec = len(self.get("jacobian")) // 125
        out = np.zeros((3, nspec, 5, 5, 5), dtype=np.float64)
        if self.lib.oracle3d_adjoint(self.h, n, _ptr(g), _ptr(out)) != 0:
            raise RuntimeError(self.lib.oracle3d_last_error().decode())
        return out

    def run(self, nstep=None, types=("displacement",)):
        """Seismograms (nstep, ntypes, nrec, 3) from rest."""
        n = self.nstep if nstep is None else int(nstep)
        t = np.array([SEIS_TYPES[s] for s in types], dtype=np.int32)
        out = np.zeros((n, len(t), self.nrec, 3), dtype=np.float64)
        if self.lib.oracle3d_run(self.h, n, len(t), _ptr(t), _ptr(out)) != 0:
            raise RuntimeError(self.lib.oracle3d_last_error().decode())
        return out
