"""smx_update_wavefields_noalloc against smx_update_wavefields.

The no-alloc entry point runs the same two launches as the operator on
caller-held device fields, with the shared-point partials in the library's
fixed device array instead of the context's buffer.  Its output must equal
the operator's bit for bit: on two calls in a row (a 120x90 layered box,
then a smaller 20x15 box, both FP32, random fields and a step with its
sources active), through the C ABI.  A mesh whose partials exceed the
array is refused with SMX_ERR_ARG before anything is launched."""

import ctypes
import sys

import numpy as np
import pytest

from specfempp_amd import _lib
from specfempp_amd.config import homogeneous_box, layered_box
from specfempp_amd.solver import Context

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _both(cfg, seed, istep):
    import torch
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    ctx.initialize()
    rng = np.random.default_rng(seed)
    u, v, a = (rng.standard_normal((ctx.nglob, 2)).astype(np.float32) * s
               for s in (1e-3, 1e-2, 1.0))
    for name, f in (("u", u), ("v", v), ("a", a)):
        ctx.set_field(name, f)
    n_op = ctypes.c_int()
    _lib.check(ctx.lib.smx_update_wavefields(ctx.h, _lib.SMX_FIELD_FORWARD, istep,
                                             ctypes.byref(n_op)))
    want = ctx.get_field("a")
    tu, tv, ta = (torch.from_numpy(f).to("cuda:0") for f in (u, v, a))
    torch.cuda.synchronize()
    n_na = ctypes.c_int()
    _lib.check(ctx.lib.smx_update_wavefields_noalloc(
        ctx.h, ctypes.c_void_p(tu.data_ptr()), ctypes.c_void_p(tv.data_ptr()),
        ctypes.c_void_p(ta.data_ptr()), istep, ctypes.byref(n_na)))
    ctx.synchronize()
    got = ta.cpu().numpy()
    # the caller's inputs are left as they were
    assert np.array_equal(tu.cpu().numpy(), u) and np.array_equal(tv.cpu().numpy(), v)
    return ctx, n_op.value, n_na.value, want, got


def test_noalloc_matches_operator_on_two_calls():
    big = layered_box(nx=120, nz=90, nstep=40, nsrc=3, nrec=10, seed=1001)
    small = homogeneous_box(nx=20, nz=15, nstep=40, nrec=4)
    for cfg, seed in ((big, 1), (small, 2)):
        _say(f"no-alloc vs operator: {cfg.name} ({cfg.mesh.nspec} elements), istep 5")
        ctx, n_op, n_na, want, got = _both(cfg, seed, istep=5)
        assert n_op == n_na == ctx.nspec
        assert np.isfinite(want).all() and np.abs(want).max() > 0
        assert np.array_equal(got, want)
        ctx.close()


def test_noalloc_refuses_a_mesh_beyond_the_scratch():
    """900x700 elements in FP64 need about 1.2M partial slots of 16 B; the
    16 MiB array holds 1,048,576."""
    cfg = homogeneous_box(nx=900, nz=700, nstep=4, nrec=2)
    _say("no-alloc size refusal: 900x700 FP64")
    ctx = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, precision=8)
    ctx.initialize()
    assert ctx.layout["npartials"] > (16 << 20) // 16
    # the context's own fields: valid device arrays of the right size
    ptrs = []
    for which in (_lib.SMX_U, _lib.SMX_V, _lib.SMX_A):
        p = ctypes.c_void_p()
        _lib.check(ctx.lib.smx_field_device_ptr(ctx.h, _lib.SMX_FIELD_FORWARD, which,
                                                ctypes.byref(p)))
        ptrs.append(p)
    rc = ctx.lib.smx_update_wavefields_noalloc(ctx.h, *ptrs, 0, None)
    assert rc == _lib.SMX_ERR_ARG
    assert b"no-alloc scratch" in ctx.lib.smx_last_error()
    ctx.close()
