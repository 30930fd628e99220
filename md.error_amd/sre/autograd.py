"""torch.autograd bridge to the engine's hand-written adjoint kernel.

``LJEnergyForces.apply(eps, sigma, rc, engine, tail)`` returns the per-replica
energies E (R,) and the forces F (N, 3) (float64, on the GPU) of the engine's
current positions; the backward pass calls ``sre_adjoint`` with
dL/dE and dL/dF, which recomputes the pair terms from the neighbor list.

Differentiable definition (the cutoff is a step function of rc): the cutoff
mask and the neighbor list are held fixed at the forward pass, so d/d rc acts
through the tail correction only (its closed form is smooth in rc).
"""
import numpy as np
import torch


class LJEnergyForces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, sigma, rc, engine, tail=True):
        """eps, sigma: (nt, nt) tensors; rc: the rc(z) bins of every replica,
        concatenated (None keeps the engine's tables)."""
        nt = engine.n_types
        engine.set_pair_params(eps.detach().cpu().double().numpy().reshape(nt, nt),
                               sigma.detach().cpu().double().numpy().reshape(nt, nt))
        n_rc = 0
        if rc is not None:
            tabs = engine._rc_layout
            flat = rc.detach().cpu().double().numpy().ravel()
            off = 0
            for r, shape in enumerate(tabs):
                m = int(np.prod(shape))
                engine.set_rc_table(flat[off:off + m].reshape(shape), replica=r)
                off += m
            n_rc = off
            engine.rebuild()
        engine.compute(stats=True, tail=tail)
        F = torch.empty((engine.n_atoms, 3), dtype=torch.float64, device="cuda")
        engine.forces_to_device(F)
        engine.synchronize()
        E = torch.from_numpy(engine.stats()[:, 0].copy()).to("cuda")
        ctx.engine = engine
        ctx.tail = tail
        ctx.n_rc = n_rc
        ctx.rc_given = rc is not None
        ctx.dtypes = (eps.dtype, sigma.dtype, rc.dtype if rc is not None else None)
        return E, F

    @staticmethod
    def backward(ctx, gE, gF):
        e = ctx.engine
        dLdE = (gE.detach().double().cpu().numpy() if gE is not None
                else np.zeros(e.n_replicas))
        dLdF = gF.detach().double().contiguous() if gF is not None else None
        ge, gs, grc = e.adjoint(dLdE=dLdE, dLdF=dLdF, tail=ctx.tail,
                                n_rc=ctx.n_rc if ctx.rc_given else None)
        g_eps = torch.from_numpy(ge).to(ctx.dtypes[0])
        g_sig = torch.from_numpy(gs).to(ctx.dtypes[1])
        g_rc = torch.from_numpy(grc).to(ctx.dtypes[2]) if ctx.rc_given else None
        return g_eps, g_sig, g_rc, None, None


def lj_energy_forces(engine, eps, sigma, rc=None, tail=True):
    """Functional form; ``engine`` must have positions bound and a list built."""
    return LJEnergyForces.apply(eps, sigma, rc, engine, tail)
