"""Adjoint (Frechet) path: seismograms as a differentiable function of the
material, and the L2 waveform misfit.

The reference computes material kernels with a combined simulation that
re-runs the forward field backwards in time next to the adjoint field
(core/specfem/solver/time_marching.tpp:133-264) and accumulates them in
frechet_kernels::compute_derivatives
(include/kokkos_kernels/frechet_kernels.hpp:39-53,
include/kokkos_kernels/impl/compute_material_derivatives.tpp:14-153).  Here the
gradient is the exact discrete adjoint of the GPU time step: the forward run
keeps its state every ``every`` steps, the adjoint run recomputes each
segment (newest first) and steps the adjoint state backwards through it, all
in device memory.  The gradient is d chi / d (kappa, mu, rho) at every GLL
point, i.e. the reference's kernels already integrated against the
quadrature weights and Jacobian.

``Seismograms`` wraps this as a ``torch.autograd.Function``: its forward
returns the displacement seismograms, its backward turns the incoming
d loss / d seismogram into the adjoint source.  Every function here takes a
2D ``Context`` (2 components, material (nspec, 5, 5)) or a 3D ``Context3D``
(3 components, material (nspec, 5, 5, 5)).
"""

from __future__ import annotations

import numpy as np

from .solver import Context

# Context or Context3D: both expose ncomp, nspec, nrec, nsig, nstep, nsbs, dt,
# seismo_types and the adjoint methods.


def l2_misfit(s, d, dt):
    """chi = 1/2 dt sum (s - d)^2 over steps, receivers and components."""
    r = np.asarray(s, dtype=np.float64) - np.asarray(d, dtype=np.float64)
    return 0.5 * dt * float(np.sum(r * r))


def _displacement_index(sim: Context) -> int:
    if "displacement" not in sim.seismo_types:
        raise ValueError("the context must record displacement seismograms")
    return sim.seismo_types.index("displacement")


def forward_seismograms(sim: Context, kappa, mu, rho, every=100):
    """Sets the model, runs the checkpointed forward and returns the
    displacement seismograms (nsig, nrec, ncomp)."""
    sim.set_model(kappa, mu, rho)
    return _checkpointed_seismograms(sim, every)


def _checkpointed_seismograms(sim, every):
    sim.run_forward_checkpointed(every)
    sim.synchronize()
    sim._model_version = getattr(sim, "_model_version", 0) + 1
    return sim.seismograms()[:, _displacement_index(sim)]


def _adjoint_source(sim, dchi_ds):
    g = np.asarray(dchi_ds, dtype=np.float64)
    nc = sim.ncomp
    if g.shape != (sim.nsig, sim.nrec, nc):
        raise ValueError(f"dchi/ds must have shape {(sim.nsig, sim.nrec, nc)}, got {g.shape}")
    gsrc = np.zeros((sim.nstep, sim.nrec, nc), dtype=np.float64)
    gsrc[::sim.nsbs][:sim.nsig] = g
    return gsrc


def adjoint_gradient(sim: Context, dchi_ds):
    """Gradient (3, nspec, 5, 5[, 5]) for dchi/ds = dchi_ds (nsig, nrec,
    ncomp), after forward_seismograms on the same context.  Seismogram i was
    recorded after step i * nsbs; the other steps get no adjoint source."""
    sim.run_adjoint(_adjoint_source(sim, dchi_ds))
    return sim.gradient()


def misfit_and_gradient(sim: Context, model, observed, every=100):
    """L2 waveform misfit against ``observed`` (nsig, nrec, ncomp) and its
    gradient with respect to ``model`` = (kappa, mu, rho)."""
    s = forward_seismograms(sim, *model, every=every)
    dt_rec = sim.dt * sim.nsbs
    chi = l2_misfit(s, observed, dt_rec)
    grad = adjoint_gradient(sim, dt_rec * (s.astype(np.float64) - observed))
    return chi, grad


try:  # torch is plumbing for the autograd interface, not a compute path
    import torch

    def _on_device(sim, t):
        """Whether the model and gradient can stay in device memory: CUDA
        tensors and a 2D elastic FP32 single-mesh context."""
        return (t.is_cuda and isinstance(sim, Context) and sim.precision == 4
                and sim.ncomp == 2 and getattr(sim, "nshots", 1) == 1)

    class Seismograms(torch.autograd.Function):
        """s = Seismograms.apply(kappa, mu, rho, sim, every): displacement
        seismograms (nsig, nrec, ncomp) of context ``sim`` for the material
        tensors (nspec, 5, 5[, 5]); differentiable in kappa, mu and rho.
        CUDA tensors on a 2D elastic FP32 context stay on the GPU: the model
        is read from, and the gradient written to, device memory."""

        @staticmethod
        def forward(ctx, kappa, mu, rho, sim, every=100):
            ctx.device_path = _on_device(sim, kappa)
            if ctx.device_path:
                model = [t.detach().to(torch.float32).contiguous() for t in (kappa, mu, rho)]
                torch.cuda.current_stream(kappa.device).synchronize()
                sim.set_model_device(*[t.data_ptr() for t in model])
                s = _checkpointed_seismograms(sim, every)
            else:
                host = [t.detach().to("cpu", torch.float32).numpy() for t in (kappa, mu, rho)]
                s = forward_seismograms(sim, *host, every=every)
            ctx.sim = sim
            ctx.version = sim._model_version
            ctx.like = kappa
            return torch.from_numpy(np.array(s)).to(kappa.device)

        @staticmethod
        def backward(ctx, grad_s):
            sim = ctx.sim
            if sim._model_version != ctx.version:
                raise RuntimeError("the context ran another model since this forward; "
                                   "its checkpoints are gone")
            gs = grad_s.detach().to("cpu", torch.float64).numpy()
            like = ctx.like
            if ctx.device_path:
                sim.run_adjoint(_adjoint_source(sim, gs))
                g = torch.empty((3,) + tuple(like.shape), dtype=torch.float64,
                                device=like.device)
                sim.gradient_device(g.data_ptr())
                out = [g[i].to(like.dtype) for i in range(3)]
            else:
                g = adjoint_gradient(sim, gs)
                out = [torch.from_numpy(g[i]).to(device=like.device, dtype=like.dtype)
                       for i in range(3)]
            return out[0], out[1], out[2], None, None

except ImportError:  # pragma: no cover - torch is part of the image
    Seismograms = None


def exact_direction(model32, direction, steps=(1, -1, 2, -2)):
    """Quantizes `direction` (float64, same shape as the FP32 array
    `model32`) to multiples of the FP32 spacing of each value, so that the
    model points model + t * direction (t in `steps`) are exactly
    representable in FP32 and lie on one line: finite differences of a
    function of the FP32 model then see no rounding of the step."""
    m32 = np.asarray(model32, dtype=np.float32)
    m = m32.astype(np.float64)
    u = np.spacing(np.abs(m32)).astype(np.float64)
    dq = np.round(np.asarray(direction, dtype=np.float64) / u) * u
    for t in steps:
        p = m + t * dq
        if not np.array_equal(p.astype(np.float32).astype(np.float64), p):
            raise ValueError("the step leaves the FP32 binade of some values; use a smaller step")
    return dq
