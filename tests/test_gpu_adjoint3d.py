"""3D GPU adjoint path (through the C ABI) against the 3D oracle's discrete
adjoint and against finite differences of the GPU's own FP64 misfit.

Case: 4x3x3 hex8 box of 100 m, free top, Stacey on five faces, moment-tensor
source, four receivers, smooth +-3% heterogeneity, 120 steps.

Tolerances (relative L2 over each of the kappa / mu / rho gradients):
* FP64 GPU vs FP64 oracle: 5e-5 (the GPU forms its Jacobian, moduli and
  impedances in FP32 as the reference does; the oracle's FP64 instantiation
  forms them in double).
* FP32 GPU vs FP64 oracle: 2e-4 (FP32 states and accumulation).
* Finite differences (fourth-order central stencil on the FP64 GPU misfit,
  relative step 1e-3, directions quantized so that every stencil point is an
  exact FP32 model): 1e-5 for the FP64 gradient, 1e-3 for the FP32 gradient.
"""

import sys

import numpy as np
import pytest

from specfempp_amd.adjoint import Seismograms, exact_direction, l2_misfit, misfit_and_gradient
from specfempp_amd.mesh3d import gll_coordinates3d
from specfempp_amd.solver3d import Context3D
from tests._cases import adjoint_small3d, perturbed_model, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ctx(cfg, precision):
    c = Context3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep, precision=precision)
    c.initialize()
    return c


def _model(cfg):
    return (cfg.mesh.kappa, cfg.mesh.mu, cfg.mesh.rho)


def _gsrc(cfg, seed=7):
    rng = np.random.default_rng(seed)
    t = np.arange(cfg.nstep) * cfg.dt
    w = np.sin(2 * np.pi * 20.0 * t)[:, None, None]
    return rng.standard_normal((cfg.nstep, len(cfg.receivers.x), 3)) * w * 1e-9


@pytest.fixture(scope="module")
def oracle_grad(oracle_lib):
    cfg = adjoint_small3d(nstep=120)
    o = oracle_lib.Oracle3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                            precision=8)
    s = o.forward_store()
    gs = _gsrc(cfg)
    return cfg, s, gs, o.adjoint(gs)


@pytest.mark.parametrize("precision,tol", [(8, 5e-5), (4, 2e-4)])
def test_3d_adjoint_gradient_matches_oracle(oracle_grad, precision, tol):
    cfg, s_ref, gs, g_ref = oracle_grad
    _say(f"3D adjoint vs oracle: 4x3x3 box, 120 steps, precision {precision}")
    c = _ctx(cfg, precision)
    c.run_forward_checkpointed(every=50)
    c.synchronize()
    e_s = rel_l2(c.seismograms()[:, 0], s_ref)
    _say(f"  seismogram rel L2 vs oracle: {e_s:.3e}")
    assert e_s < 1e-5
    c.run_adjoint(gs.astype(c.dtype))
    g = c.gradient()
    errs = [rel_l2(g[i], g_ref[i]) for i in range(3)]
    _say(f"  gradient rel L2 (kappa, mu, rho): {errs}")
    assert max(errs) < tol, errs


def test_3d_adjoint_independent_of_checkpoint_interval(oracle_grad):
    cfg, _, gs, _ = oracle_grad
    _say("3D adjoint: checkpoint intervals 50, 1, 7, 120 (FP64 bitwise, FP32 1e-6)")
    for precision in (8, 4):
        c = _ctx(cfg, precision)
        grads = []
        for every in (50, 1, 7, 120):
            c.run_forward_checkpointed(every=every)
            c.run_adjoint(gs.astype(c.dtype))
            grads.append(c.gradient())
        for g in grads[1:]:
            if precision == 8:
                assert np.array_equal(g, grads[0])
            else:
                for i in range(3):
                    assert rel_l2(g[i], grads[0][i]) < 1e-6


def test_3d_checkpointed_forward_seismograms_are_the_forward_ones():
    cfg = adjoint_small3d(nstep=120)
    _say("3D checkpointed forward vs plain forward (FP32, bitwise)")
    c = _ctx(cfg, 4)
    c.run_forward(0, cfg.nstep)
    c.synchronize()
    s0 = c.seismograms().copy()
    c.run_forward_checkpointed(every=25)
    c.synchronize()
    assert np.array_equal(c.seismograms(), s0)


def test_3d_set_model_equals_fresh_context():
    cfg = adjoint_small3d(nstep=60)
    _say("3D set_model vs a context assembled with that model (bitwise)")
    true = perturbed_model(cfg)
    c = _ctx(cfg, 4)
    c.set_model(*true)
    c.run_forward(0, cfg.nstep)
    c.synchronize()
    import dataclasses
    mesh = dataclasses.replace(cfg.mesh, kappa=true[0], mu=true[1], rho=true[2])
    fresh = Context3D(mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep)
    fresh.initialize()
    fresh.run_forward(0, cfg.nstep)
    fresh.synchronize()
    assert np.array_equal(c.seismograms(), fresh.seismograms())
    assert np.array_equal(c.get_field("minv"), fresh.get_field("minv"))


def test_3d_adjoint_zero_and_linear(oracle_grad):
    cfg, _, gs, _ = oracle_grad
    _say("3D adjoint: zero source, linearity (FP64)")
    c = _ctx(cfg, 8)
    c.run_forward_checkpointed(every=40)
    c.run_adjoint(np.zeros_like(gs))
    assert np.all(c.gradient() == 0.0)
    c.run_adjoint(gs)
    g1 = c.gradient()
    c.run_adjoint(2.0 * gs)
    g2 = c.gradient()
    np.testing.assert_allclose(g2, 2.0 * g1, rtol=1e-12, atol=1e-12 * np.abs(g1).max())


def _misfit(c, model, obs):
    c.set_model(*model)
    c.run_forward_checkpointed(every=40)
    c.synchronize()
    return l2_misfit(c.seismograms()[:, 0], obs, c.dt)


def _smooth(mesh, rng):
    xyz = gll_coordinates3d(mesh)
    lo, hi = xyz.reshape(-1, 3).min(0), xyz.reshape(-1, 3).max(0)
    u = (xyz - lo) / (hi - lo)
    k = rng.uniform(0.5, 2.0, 3)
    ph = rng.uniform(0, 2 * np.pi, 3)
    return (np.cos(np.pi * k[0] * u[..., 0] + ph[0]) * np.cos(np.pi * k[1] * u[..., 1] + ph[1])
            * np.cos(np.pi * k[2] * u[..., 2] + ph[2]))


def test_3d_adjoint_finite_differences():
    cfg = adjoint_small3d(nstep=120)
    _say("3D adjoint vs finite differences of the FP64 GPU misfit, 6 directions")
    c64, c32 = _ctx(cfg, 8), _ctx(cfg, 4)
    c64.set_model(*perturbed_model(cfg))
    c64.run_forward_checkpointed(every=40)
    c64.synchronize()
    obs = c64.seismograms()[:, 0].copy()
    m0 = _model(cfg)
    _, g64 = misfit_and_gradient(c64, m0, obs, every=40)
    _, g32 = misfit_and_gradient(c32, m0, obs.astype(np.float32), every=40)
    rng = np.random.default_rng(21)
    base = [x.astype(np.float64) for x in m0]
    for k, (which, smooth) in enumerate(((0, False), (1, False), (2, False), (None, False),
                                         (None, True), (None, True))):
        dirs = [np.zeros_like(x) for x in base]
        for i in ([which] if which is not None else [0, 1, 2]):
            f = _smooth(cfg.mesh, rng) if smooth else rng.standard_normal(base[i].shape)
            dirs[i] = exact_direction(m0[i], f * base[i] * 1e-3)

        def chi(t):
            return _misfit(c64, [(x + t * d).astype(np.float32) for x, d in zip(base, dirs)],
                           obs)
        fd = (8.0 * (chi(1) - chi(-1)) - (chi(2) - chi(-2))) / 12.0
        a64 = sum(float(np.sum(g * d)) for g, d in zip(g64, dirs))
        a32 = sum(float(np.sum(g * d)) for g, d in zip(g32, dirs))
        e64, e32 = abs(a64 - fd) / abs(fd), abs(a32 - fd) / abs(fd)
        _say(f"  direction {k} ({which}, smooth={smooth}): FD {fd:.6e}  "
             f"fp64 err {e64:.2e}  fp32 err {e32:.2e}")
        assert e64 < 1e-5
        assert e32 < 1e-3


def test_3d_autograd_function_matches_direct_gradient():
    torch = pytest.importorskip("torch")
    cfg = adjoint_small3d(nstep=80)
    _say("3D torch.autograd.Function: L2 misfit backward (FP32)")
    c = _ctx(cfg, 4)
    c.set_model(*perturbed_model(cfg))
    c.run_forward_checkpointed(every=20)
    c.synchronize()
    obs = c.seismograms()[:, 0].copy()
    m0 = _model(cfg)
    chi, g = misfit_and_gradient(c, m0, obs, every=20)
    k, m, r = (torch.tensor(x, requires_grad=True) for x in m0)
    s = Seismograms.apply(k, m, r, c, 20)
    loss = 0.5 * cfg.dt * ((s.double() - torch.from_numpy(obs).double()) ** 2).sum()
    loss.backward()
    assert abs(loss.item() - chi) <= 1e-12 * abs(chi) + 1e-30
    for t, ref in zip((k, m, r), g):
        assert rel_l2(t.grad.numpy(), ref) < 1e-6


def test_3d_adjoint_error_paths():
    cfg = adjoint_small3d(nstep=30)
    c = _ctx(cfg, 4)
    with pytest.raises(RuntimeError, match="checkpointed"):
        c.run_adjoint(np.zeros((cfg.nstep, len(cfg.receivers.x), 3)))
    c.run_forward_checkpointed(every=10)
    c.set_model(*_model(cfg))  # a new model drops the checkpoints
    with pytest.raises(RuntimeError, match="checkpointed"):
        c.run_adjoint(np.zeros((cfg.nstep, len(cfg.receivers.x), 3)))
    with pytest.raises(RuntimeError, match="positive"):
        c.run_forward_checkpointed(every=0)
    bad = [x.copy() for x in _model(cfg)]
    bad[2][0, 0, 0, 0] = -1.0
    with pytest.raises(RuntimeError, match="non-physical"):
        c.set_model(*bad)


def test_3d_tile_sizes_agree(oracle_grad):
    """Forward seismograms and adjoint gradients do not depend on the tile
    layout beyond rounding: FP64 with tiles of 8, 1 and 4 elements; FP32 with
    32, 8 and 64 (the 64-element tile of this mesh needs more LDS than the
    adjoint tile pass may use on a 64 KiB device, so it also exercises the
    colour-pass adjoint)."""
    cfg, _, gs, _ = oracle_grad
    for precision, sizes, ts, tg in ((8, (8, 1, 4), 1e-12, 1e-10), (4, (32, 8, 64), 1e-6, 1e-5)):
        _say(f"3D tile layouts {sizes}, precision {precision}: forward and gradient agree")
        res = []
        for te in sizes:
            c = Context3D(cfg.mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                          precision=precision, options={"tile_elems": te})
            c.initialize()
            c.run_forward_checkpointed(every=40)
            c.synchronize()
            s = c.seismograms().copy()
            c.run_adjoint(gs.astype(c.dtype))
            res.append((s, c.gradient()))
        for s, g in res[1:]:
            assert rel_l2(s, res[0][0]) < ts
            for i in range(3):
                assert rel_l2(g[i], res[0][1][i]) < tg
