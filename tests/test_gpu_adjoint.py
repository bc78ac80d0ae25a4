"""GPU adjoint path (through the C ABI) against the oracle's discrete adjoint
and against finite differences of the GPU's own FP64 misfit.

Tolerances (relative L2 over each of the kappa / mu / rho gradients):
* FP64 GPU vs FP64 oracle: 5e-5.  The GPU assembles its coefficients
  (Jacobian, moduli, impedances) in FP32 as the reference does, so even in
  FP64 arithmetic it solves a problem whose coefficients differ from the FP64
  oracle's by ~1e-7 relative (its seismograms differ by ~2e-6).
* FP32 GPU vs FP64 oracle: 2e-4 (FP32 forward states and accumulation).
* Finite differences (fourth-order central stencil on the FP64 GPU misfit,
  relative step 1e-3, directions quantized so that every stencil point is an
  exact FP32 model): 1e-5 for the FP64 gradient, 1e-3 for the FP32
  gradient.
"""

import dataclasses
import sys

import numpy as np
import pytest

from specfempp_amd import _lib
from specfempp_amd.adjoint import Seismograms, exact_direction, l2_misfit, misfit_and_gradient
from specfempp_amd.config import smooth_perturbation
from specfempp_amd.solver import Context
from tests._cases import adjoint_small, perturbed_model, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ctx(cfg, precision):
    c = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, precision=precision)
    c.initialize()
    return c


def _model(cfg):
    return (cfg.mesh.kappa, cfg.mesh.mu, cfg.mesh.rho)


def _gsrc(cfg, seed=7):
    rng = np.random.default_rng(seed)
    t = np.arange(cfg.nstep) * cfg.dt
    w = np.sin(2 * np.pi * 40.0 * t)[:, None, None]
    return rng.standard_normal((cfg.nstep, len(cfg.receivers), 2)) * w * 1e-9


@pytest.fixture(scope="module")
def oracle_grad(oracle_lib):
    cfg = adjoint_small(nstep=300)
    o = oracle_lib.Oracle2D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, precision=8)
    s = o.forward_store()
    gs = _gsrc(cfg)
    return cfg, s, gs, o.adjoint(gs)


@pytest.mark.parametrize("precision,tol", [(8, 5e-5), (4, 2e-4)])
def test_adjoint_gradient_matches_oracle(oracle_grad, precision, tol):
    cfg, s_ref, gs, g_ref = oracle_grad
    _say(f"adjoint vs oracle: 12x9 box, 300 steps, precision {precision}")
    c = _ctx(cfg, precision)
    c.run_forward_checkpointed(every=64)
    c.synchronize()
    e_s = rel_l2(c.seismograms()[:, 0], s_ref)
    _say(f"  seismogram rel L2 vs oracle: {e_s:.3e}")
    assert e_s < 1e-5
    c.run_adjoint(gs.astype(c.dtype))
    g = c.gradient()
    errs = [rel_l2(g[i], g_ref[i]) for i in range(3)]
    _say(f"  gradient rel L2 (kappa, mu, rho): {errs}")
    assert max(errs) < tol, errs


def test_adjoint_independent_of_checkpoint_interval(oracle_grad):
    cfg, _, gs, _ = oracle_grad
    _say("adjoint: checkpoint intervals 1, 7, 64, 300 (FP32)")
    c = _ctx(cfg, 4)
    grads = []
    for every in (64, 1, 7, 300):
        c.run_forward_checkpointed(every=every)
        c.run_adjoint(gs.astype(np.float32))
        grads.append(c.gradient())
    for g in grads[1:]:
        for i in range(3):
            assert rel_l2(g[i], grads[0][i]) < 1e-6


def test_checkpointed_forward_seismograms_are_the_forward_ones():
    cfg = adjoint_small(nstep=300)
    _say("checkpointed forward vs plain forward (FP32, bitwise)")
    c = _ctx(cfg, 4)
    c.run_forward(0, cfg.nstep)
    c.synchronize()
    s0 = c.seismograms().copy()
    c.run_forward_checkpointed(every=50)
    c.synchronize()
    assert np.array_equal(c.seismograms(), s0)


def test_adjoint_zero_and_linear(oracle_grad):
    cfg, _, gs, _ = oracle_grad
    _say("adjoint: zero source, linearity (FP64)")
    c = _ctx(cfg, 8)
    c.run_forward_checkpointed(every=100)
    c.run_adjoint(np.zeros_like(gs))
    assert np.all(c.gradient() == 0.0)
    c.run_adjoint(gs)
    g1 = c.gradient()
    c.run_adjoint(2.0 * gs)
    g2 = c.gradient()
    np.testing.assert_allclose(g2, 2.0 * g1, rtol=1e-12, atol=1e-12 * np.abs(g1).max())


def _misfit(c, model, obs):
    c.set_model(*model)
    c.run_forward_checkpointed(every=100)
    c.synchronize()
    return l2_misfit(c.seismograms()[:, 0], obs, c.dt)


def test_adjoint_finite_differences():
    cfg = adjoint_small(nstep=300)
    _say("adjoint vs finite differences of the FP64 GPU misfit, 6 directions")
    c64, c32 = _ctx(cfg, 8), _ctx(cfg, 4)
    c64.set_model(*perturbed_model(cfg))
    c64.run_forward_checkpointed(every=100)
    c64.synchronize()
    obs = c64.seismograms()[:, 0].copy()
    m0 = _model(cfg)
    _, g64 = misfit_and_gradient(c64, m0, obs, every=100)
    _, g32 = misfit_and_gradient(c32, m0, obs.astype(np.float32), every=100)
    rng = np.random.default_rng(21)
    base = [x.astype(np.float64) for x in m0]
    for k, (which, smooth) in enumerate(((0, False), (1, False), (2, False), (None, False),
                                         (None, True), (None, True))):
        dirs = [np.zeros_like(x) for x in base]
        for i in ([which] if which is not None else [0, 1, 2]):
            f = smooth_perturbation(cfg.mesh, rng) if smooth else rng.standard_normal(base[i].shape)
            dirs[i] = exact_direction(m0[i], f * base[i] * 1e-3)

        def chi(t):
            return _misfit(c64, [(x + t * d).astype(np.float32) for x, d in zip(base, dirs)],
                           obs)
        fd = (8.0 * (chi(1) - chi(-1)) - (chi(2) - chi(-2))) / 12.0
        a64 = sum(float(np.sum(g * d)) for g, d in zip(g64, dirs))
        a32 = sum(float(np.sum(g * d)) for g, d in zip(g32, dirs))
        e64, e32 = abs(a64 - fd) / abs(fd), abs(a32 - fd) / abs(fd)
        _say(f"  direction {k} ({which}, smooth={smooth}): FD {fd:.6e}  fp64 err {e64:.2e}  fp32 err {e32:.2e}")
        assert e64 < 1e-5
        assert e32 < 1e-3


def test_adjoint_with_subsampled_seismograms():
    """Seismograms kept every third step (301 steps, so the last sample is
    the step-300 state): the adjoint source enters only at the sampled steps
    and the misfit weighs samples by dt * 3.  Checked against fourth-order
    finite differences of the FP64 GPU misfit in three directions."""
    from specfempp_amd.adjoint import forward_seismograms
    cfg = adjoint_small(nstep=301)
    _say("adjoint with nsbs=3 vs finite differences of the FP64 GPU misfit, 3 directions")
    c64 = Context(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, nsbs=3,
                  precision=8)
    c64.initialize()
    assert c64.nsbs == 3 and c64.nsig < cfg.nstep // 2
    c64.set_model(*perturbed_model(cfg))
    c64.run_forward_checkpointed(every=100)
    c64.synchronize()
    obs = c64.seismograms()[:, 0].copy()
    m0 = _model(cfg)
    _, g64 = misfit_and_gradient(c64, m0, obs, every=100)
    rng = np.random.default_rng(33)
    base = [x.astype(np.float64) for x in m0]
    for k, which in enumerate((0, 1, None)):
        dirs = [np.zeros_like(x) for x in base]
        for i in ([which] if which is not None else [0, 1, 2]):
            dirs[i] = exact_direction(m0[i], rng.standard_normal(base[i].shape) * base[i] * 1e-3)

        def chi(t):
            s = forward_seismograms(c64, *[(x + t * d).astype(np.float32)
                                           for x, d in zip(base, dirs)], every=100)
            return l2_misfit(s, obs, c64.dt * 3)
        fd = (8.0 * (chi(1) - chi(-1)) - (chi(2) - chi(-2))) / 12.0
        a64 = sum(float(np.sum(g * d)) for g, d in zip(g64, dirs))
        e64 = abs(a64 - fd) / abs(fd)
        _say(f"  direction {k} ({which}): FD {fd:.6e}  fp64 err {e64:.2e}")
        assert e64 < 1e-5


def test_autograd_function_matches_direct_gradient():
    torch = pytest.importorskip("torch")
    cfg = adjoint_small(nstep=200)
    _say("torch.autograd.Function: L2 misfit backward (FP32)")
    c = _ctx(cfg, 4)
    c.set_model(*perturbed_model(cfg))
    c.run_forward_checkpointed(every=50)
    c.synchronize()
    obs = c.seismograms()[:, 0].copy()
    m0 = _model(cfg)
    chi, g = misfit_and_gradient(c, m0, obs, every=50)
    k, m, r = (torch.tensor(x, requires_grad=True) for x in m0)
    s = Seismograms.apply(k, m, r, c, 50)
    loss = 0.5 * cfg.dt * ((s.double() - torch.from_numpy(obs).double()) ** 2).sum()
    loss.backward()
    assert abs(loss.item() - chi) <= 1e-12 * abs(chi) + 1e-30
    for t, ref in zip((k, m, r), g):
        assert rel_l2(t.grad.numpy(), ref) < 1e-6


def test_adjoint_error_paths():
    cfg = adjoint_small(nstep=50)
    c = _ctx(cfg, 4)
    with pytest.raises(RuntimeError, match="checkpointed"):
        c.run_adjoint(np.zeros((cfg.nstep, len(cfg.receivers), 2)))
    c.run_forward_checkpointed(every=10)
    c.set_model(*_model(cfg))  # a new model drops the checkpoints
    with pytest.raises(RuntimeError, match="checkpointed"):
        c.run_adjoint(np.zeros((cfg.nstep, len(cfg.receivers), 2)))
    with pytest.raises(RuntimeError, match="positive"):
        c.run_forward_checkpointed(every=0)
    bad = [x.copy() for x in _model(cfg)]
    bad[2][0, 0, 0] = -1.0
    with pytest.raises(RuntimeError, match="non-physical"):
        c.set_model(*bad)
    assert _lib.load().smx_run_adjoint(None, None) != 0


def _compact(ctx):
    import ctypes
    g, m = ctypes.c_int(), ctypes.c_int()
    _lib.check(ctx.lib.smx_ctx_compact(ctx.h, ctypes.byref(g), ctypes.byref(m)))
    return g.value, m.value


def test_device_model_and_gradient_match_the_host_path(oracle_grad):
    """smx_set_model_device and smx_get_gradient_device (CUDA tensors in and
    out) against smx_set_model and smx_get_gradient, bit for bit: the
    seismograms, the gradient read either way, and the element-constant
    moduli flags, for a model varying inside elements and one constant per
    element; a non-physical value is refused naming its element and leaves
    the model as it was."""
    import torch
    cfg, _, gs, _ = oracle_grad
    _say("device-memory model and gradient vs the host path, 12x9 box")
    varying = perturbed_model(cfg)
    const = tuple(np.ascontiguousarray(np.broadcast_to(x[:, :1, :1], x.shape))
                  for x in varying)
    for name, model in (("varying", varying), ("element-constant", const)):
        ch = _ctx(cfg, 4)
        ch.set_model(*model)
        ch.run_forward_checkpointed(every=64)
        ch.synchronize()
        s_h = ch.seismograms()
        ch.run_adjoint(gs.astype(np.float32))
        g_h = ch.gradient()
        cd = _ctx(cfg, 4)
        tm = [torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0") for x in model]
        torch.cuda.synchronize()
        cd.set_model_device(*[t.data_ptr() for t in tm])
        cd.run_forward_checkpointed(every=64)
        cd.synchronize()
        s_d = cd.seismograms()
        cd.run_adjoint(gs.astype(np.float32))
        gdev = torch.empty((3, cfg.mesh.nspec, 5, 5), dtype=torch.float64, device="cuda:0")
        cd.gradient_device(gdev.data_ptr())
        g_d = gdev.cpu().numpy()
        _say(f"  {name}: seismograms equal {np.array_equal(s_h, s_d)}, gradient equal "
             f"{np.array_equal(g_h, g_d)}, flags {_compact(ch)} / {_compact(cd)}")
        assert np.array_equal(s_h, s_d)
        assert np.array_equal(g_h, g_d)
        assert np.array_equal(g_h, cd.gradient())  # the host copy brought up to date
        assert _compact(ch) == _compact(cd)
    bad = [t.clone() for t in tm]
    bad[2].view(-1)[7 * 25 + 3] = -1.0
    torch.cuda.synchronize()
    with pytest.raises(_lib.SmxError, match="element 7"):
        cd.set_model_device(*[t.data_ptr() for t in bad])
    cd.run_forward_checkpointed(every=64)
    cd.synchronize()
    assert np.array_equal(cd.seismograms(), s_d)


def test_autograd_on_cuda_tensors_matches_host_tensors(oracle_grad):
    """The autograd function with CUDA tensors (model and gradient kept in
    device memory) gives the host-tensor result bit for bit."""
    import torch
    cfg, _, _, _ = oracle_grad
    _say("autograd: CUDA tensors vs host tensors")
    model = perturbed_model(cfg)
    w = torch.from_numpy(np.random.default_rng(3).standard_normal(
        (cfg.nstep, len(cfg.receivers), 2)))
    out = {}
    for dev in ("cpu", "cuda:0"):
        c = _ctx(cfg, 4)
        ts = [torch.from_numpy(np.ascontiguousarray(x)).to(dev).requires_grad_(True)
              for x in model]
        s = Seismograms.apply(*ts, c, 64)
        loss = (s.double() * w.to(dev)).sum()
        loss.backward()
        out[dev] = (s.detach().cpu().numpy(), [t.grad.cpu().numpy() for t in ts])
    assert np.array_equal(out["cpu"][0], out["cuda:0"][0])
    for a, b in zip(out["cpu"][1], out["cuda:0"][1]):
        assert np.array_equal(a, b)
