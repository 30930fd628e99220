"""torch.autograd bridge to the engine's hand-written adjoint kernel.

``LJEnergyForces.apply(eps, sigma, rc, engine, tail)`` returns the per-replica
energies E (R,) and the forces F (N, 3) (float64, on the GPU) of the engine's
current positions; the backward pass calls ``sre_adjoint`` with
dL/dE and dL/dF, which recomputes the pair terms from the neighbor list.

Differentiable definition (the cutoff is a step function of rc): the cutoff
mask and the neighbor list are held fixed at the forward pass, so d/d rc acts
through the tail correction only (its closed form is smooth in rc).

Pair parameters: the force field has one parameter per unordered type pair,
and the engine takes symmetric matrices.  The forward is defined on the
symmetrised matrices (eps + eps^T) / 2 and (sigma + sigma^T) / 2, so it is a
function of all n_types^2 entries and its gradient is the ordinary one: the
derivative with respect to a shared off-diagonal parameter is split equally
between the two entries (a symmetric matrix built from one parameter per
pair therefore receives the whole derivative, as torch sums both entries).
"""
import numpy as np
import torch


def _sym(m):
    a = m.detach().cpu().double().numpy()
    return 0.5 * (a + a.T)


class LJEnergyForces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, sigma, rc, engine, tail=True):
        """eps, sigma: (nt, nt) tensors; rc: the rc(z) bins of every replica,
        concatenated (None keeps the engine's tables)."""
        nt = engine.n_types
        engine.set_pair_params(_sym(eps.reshape(nt, nt)), _sym(sigma.reshape(nt, nt)))
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
        ctx.shapes = (eps.shape, sigma.shape)
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
        # sre_adjoint returns the derivative with respect to each shared pair
        # parameter in both of its entries; the symmetrised forward splits it
        half = np.where(np.eye(e.n_types, dtype=bool), 1.0, 0.5)
        g_eps = torch.from_numpy(ge * half).to(ctx.dtypes[0]).reshape(ctx.shapes[0])
        g_sig = torch.from_numpy(gs * half).to(ctx.dtypes[1]).reshape(ctx.shapes[1])
        g_rc = torch.from_numpy(grc).to(ctx.dtypes[2]) if ctx.rc_given else None
        return g_eps, g_sig, g_rc, None, None


def lj_energy_forces(engine, eps, sigma, rc=None, tail=True):
    """Functional form; ``engine`` must have positions bound and a list built."""
    return LJEnergyForces.apply(eps, sigma, rc, engine, tail)


def pair_matrix(values, n_types):
    """Symmetric (n_types, n_types) tensor from one value per unordered type
    pair, in the order (0,0), (0,1), ..., (0,nt-1), (1,1), ... (differentiable:
    the gradient of each value is the derivative for that pair)."""
    rows = []
    k = 0
    idx = {}
    for a in range(n_types):
        for b in range(a, n_types):
            idx[(a, b)] = idx[(b, a)] = k
            k += 1
    assert values.numel() == k
    for a in range(n_types):
        rows.append(torch.stack([values[idx[(a, b)]] for b in range(n_types)]))
    return torch.stack(rows)
