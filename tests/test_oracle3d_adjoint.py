"""The 3D oracle's discrete adjoint, pinned by mathematics (the reference has
no 3D adjoint and ships no kernel values):

* central finite differences of the L2 waveform misfit in FP64 (relative
  steps 1e-4 and 1e-5), in random directions of kappa, mu and rho separately and
  jointly, on a box with Stacey on five faces and a moment-tensor source (a
  dropped Stacey or mass term, a wrong sign or a transposed index fails
  these); the error must shrink as the step squared;
* zero adjoint source gives a zero gradient, and the gradient is linear in the
  adjoint source.
"""

import dataclasses

import numpy as np
import pytest

from specfempp_amd.adjoint import l2_misfit
from tests._cases import adjoint_small3d, perturbed_model


def _oracle(oracle_lib, cfg, model=None, precision=8):
    mesh = cfg.mesh
    if model is not None:
        mesh = dataclasses.replace(mesh, kappa=model[0], mu=model[1], rho=model[2])
    return oracle_lib.Oracle3D(mesh, cfg.sources, cfg.receivers, cfg.dt, cfg.nstep,
                               precision=precision)


def _seis(oracle_lib, cfg, model):
    return _oracle(oracle_lib, cfg, model).forward_store()


@pytest.fixture(scope="module")
def case(oracle_lib):
    cfg = adjoint_small3d(nstep=120)
    obs = _seis(oracle_lib, cfg, perturbed_model(cfg))
    o = _oracle(oracle_lib, cfg)
    s = o.forward_store()
    grad = o.adjoint(cfg.dt * (s - obs))
    return cfg, obs, grad


def _fd(oracle_lib, cfg, obs, m0, dirs):
    mp = [(x + d).astype(np.float32) for x, d in zip(m0, dirs)]
    mm = [(x - d).astype(np.float32) for x, d in zip(m0, dirs)]
    chip = l2_misfit(_seis(oracle_lib, cfg, mp), obs, cfg.dt)
    chim = l2_misfit(_seis(oracle_lib, cfg, mm), obs, cfg.dt)
    # the model is FP32: the exactly representable +- perturbations
    delta = [p.astype(np.float64) - m.astype(np.float64) for p, m in zip(mp, mm)]
    return chip - chim, delta


def test_oracle3d_adjoint_finite_differences(oracle_lib, case):
    cfg, obs, grad = case
    m0 = [x.astype(np.float64) for x in (cfg.mesh.kappa, cfg.mesh.mu, cfg.mesh.rho)]
    rng = np.random.default_rng(11)
    worst = 0.0
    for which in (0, 1, 2, None):
        base = [rng.standard_normal(x.shape) * x for x in m0]
        errs = []
        for step in (1e-4, 1e-5):
            dirs = [np.zeros_like(x) for x in m0]
            for i in ([which] if which is not None else [0, 1, 2]):
                dirs[i] = base[i] * step
            fd, delta = _fd(oracle_lib, cfg, obs, m0, dirs)
            an = sum(float(np.sum(g * d)) for g, d in zip(grad, delta))
            errs.append(abs(fd - an) / abs(fd))
        worst = max(worst, errs[1])
        print(f"parameter {which}: FD error {errs[0]:.2e} (step 1e-4), {errs[1]:.2e} (1e-5)")
        assert errs[1] < 3e-7, (which, errs)
        # truncation error: shrinks ~100x for a 10x smaller step
        assert errs[1] < errs[0] / 20 or errs[1] < 1e-9, (which, errs)
    print("3D oracle adjoint vs FD: worst relative error", worst)


def test_oracle3d_adjoint_stacey_terms_matter(oracle_lib, case):
    """A direction supported on the absorbing-face elements only: the
    gradient there includes the impedance terms of the Stacey damping and
    mass, and must still match finite differences."""
    cfg, obs, grad = case
    nx, ny, nz = cfg.mesh.shape
    e = np.arange(cfg.mesh.nspec)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    edge = (ex == 0) | (ex == nx - 1) | (ey == 0) | (ey == ny - 1) | (ez == 0)
    m0 = [x.astype(np.float64) for x in (cfg.mesh.kappa, cfg.mesh.mu, cfg.mesh.rho)]
    rng = np.random.default_rng(5)
    dirs = [rng.standard_normal(x.shape) * x * 1e-5 * edge[:, None, None, None] for x in m0]
    fd, delta = _fd(oracle_lib, cfg, obs, m0, dirs)
    an = sum(float(np.sum(g * d)) for g, d in zip(grad, delta))
    print("Stacey-face direction: FD", fd, "adjoint", an)
    assert abs(fd - an) <= 3e-7 * abs(fd), (fd, an)


def test_oracle3d_adjoint_zero_and_linear(oracle_lib, case):
    cfg, _, _ = case
    o = _oracle(oracle_lib, cfg)
    o.forward_store()
    rng = np.random.default_rng(3)
    gsrc = rng.standard_normal((cfg.nstep, len(cfg.receivers.x), 3)) * 1e-12
    g1 = o.adjoint(gsrc)
    g2 = o.adjoint(2.0 * gsrc)
    assert np.all(o.adjoint(np.zeros_like(gsrc)) == 0.0)
    np.testing.assert_allclose(g2, 2.0 * g1, rtol=1e-12, atol=1e-30 + 1e-12 * np.abs(g1).max())
