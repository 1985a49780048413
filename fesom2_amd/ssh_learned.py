"""Training data and training loop of the learned SSH preconditioner.

The network, its kernels and the flexible CG live in the library
(fgpu_ssh_nn_* / fgpu_ssh_solve_learned); this module only prepares right-hand
sides, solves them with the plain diagonally preconditioned CG on the device
(the pairs the network learns from) and calls the device training loop.
"""
from __future__ import annotations

import numpy as np

from . import meshgen as mg


def rhs_samples(mesh, n, seed=0):
    """Seeded SSH right-hand sides: area-weighted smooth fields of a few scales
    plus white noise, the mix the divergence of a transport field produces."""
    rng = np.random.default_rng(seed + 7000)
    x = mesh.coord_nod2D[:, 0]
    y = mesh.coord_nod2D[:, 1]
    L = mesh.cyclic_length if mesh.cyclic_length > 0 else max(np.ptp(x), 1e-12)
    Ly = max(np.ptp(y), 1e-12)
    y0 = y.min()
    area = mesh.areasvol[:, 0] / np.mean(mesh.areasvol[:, 0])
    out = np.empty((n, mesh.nod2D))
    for s in range(n):
        kind = s % 3
        if kind == 0:
            f = rng.normal(0.0, 1.0, mesh.nod2D)
        else:
            kmax = 4 if kind == 1 else 16
            f = mg._smooth_field(x, y - y0, L, Ly, rng, nmodes=24, kmax=kmax)
            f = f / max(np.std(f), 1e-30) + 0.1 * rng.normal(0.0, 1.0, mesh.nod2D)
        out[s] = f * area
    return out


def _potential_flow(mesh, rng, nfields=3):
    """Seeded element velocity fields u = grad(phi) of smooth nodal potentials,
    (nfields, elem2D, 2): the divergent part a momentum tendency carries."""
    x = mesh.coord_nod2D[:, 0]
    y = mesh.coord_nod2D[:, 1]
    L = mesh.cyclic_length if mesh.cyclic_length > 0 else max(np.ptp(x), 1e-12)
    Ly = max(np.ptp(y), 1e-12)
    gs = mesh.gradient_sca
    en = mesh.elem2D_nodes
    out = np.empty((nfields, mesh.elem2D, 2))
    for k in range(nfields):
        phi = mg._smooth_field(x, y - y.min(), L, Ly, rng, nmodes=24, kmax=3 + 2 * k)
        pe = phi[en]
        g = np.stack([(gs[:, 0:3] * pe).sum(1), (gs[:, 3:6] * pe).sum(1)], 1)
        out[k] = g / max(np.abs(g).max(), 1e-300)
    return out


def model_rhs_sequence(ctx, mesh, state, nsteps=2000, every=10, seed=0, umax=0.01, soltol=1e-12,
                       maxiter=20000, basis=None):
    """SSH right-hand sides and solutions from a run of the plain solver.

    Every step sets a slowly turning momentum tendency uv_rhs (three seeded
    potential-flow fields with phases drifting at their own rates, scaled to
    umax m/s per step, with the vertical profile of the state's flow), forms
    the right-hand side on the device with compute_ssh_rhs_ale (linfs,
    alpha = 1, the resident uv of the state, ssh_rhs_old = the previous
    step's) and solves it with the plain diagonally preconditioned CG,
    warm-started from the previous step's solution as the model does. Every
    `every`-th step's (rhs, solution) is kept.
    basis (default: meshes over 2e7 element layers): at alpha = 1 the
    right-hand side is affine in uv_rhs and does not read ssh_rhs_old
    (oce_ale.F90:1803-1938), so each step's is formed on the host from the
    device's right-hand sides of the resident flow alone and of each tendency
    field, instead of uploading a full uv_rhs every step; it agrees with the
    direct evaluation to rounding.
    Returns (rhs, sol), each (nsteps // every, nod2D)."""
    rng = np.random.default_rng(seed + 9100)
    fields = _potential_flow(mesh, rng)
    rate = rng.uniform(0.002, 0.02, fields.shape[0])
    phase = rng.uniform(0.0, 2.0 * np.pi, fields.shape[0])
    # vertical profile of the generated flow: its column shape, normalised
    prof = np.abs(state.uv).sum(2)
    prof = prof / np.maximum(prof.max(1, keepdims=True), 1e-300)
    ctx.set_uv(state.uv)
    ctx.ssh_solve_preconditioner(2)
    x = np.zeros(mesh.nod2D)
    old = np.zeros(mesh.nod2D)
    keep_b, keep_x = [], []
    if basis is None:
        basis = mesh.elem2D * (mesh.nl - 1) > 2e7
    if basis:
        r0 = ctx.compute_ssh_rhs_ale(np.zeros_like(state.uv), old)
        rk = [ctx.compute_ssh_rhs_ale(umax * fields[k][:, None, :] * prof[:, :, None], old) - r0
              for k in range(fields.shape[0])]
    for s in range(nsteps):
        c = np.cos(rate * s + phase)
        if basis:
            rhs = r0 + sum(c[k] * rk[k] for k in range(len(rk)))
        else:
            tend = np.tensordot(c, fields, 1)                          # (elem2D, 2)
            uv_rhs = umax * tend[:, None, :] * prof[:, :, None]
            rhs = ctx.compute_ssh_rhs_ale(uv_rhs, old)
        x, st = ctx.ssh_solve_cg(x, rhs, soltol=soltol, maxiter=maxiter)
        if not st["converged"]:
            raise RuntimeError(f"plain CG did not converge at step {s}: {st}")
        old = rhs
        if (s + 1) % every == 0:
            keep_b.append(rhs.copy())
            keep_x.append(x.copy())
    return np.stack(keep_b), np.stack(keep_x)


def solve_pairs(ctx, rhs, soltol=1e-12, maxiter=20000):
    """Solutions of A x = b with the plain diagonally preconditioned CG (device)."""
    ctx.ssh_solve_preconditioner(2)
    xs = np.empty_like(rhs)
    for s in range(rhs.shape[0]):
        x, st = ctx.ssh_solve_cg(np.zeros(rhs.shape[1]), rhs[s], soltol=soltol, maxiter=maxiter)
        if not st["converged"]:
            raise RuntimeError(f"plain CG did not converge on training sample {s}: {st}")
        xs[s] = x
    return xs


def residual_pairs(ctx, rhs, sol, ks=(0, 10, 30, 60, 100)):
    """Residual / error pairs along plain CG runs: after k iterations from x0 = 0,
    r_k = b - A x_k is a right-hand side whose solution is e_k = x - x_k. These
    are the vectors a preconditioner meets inside CG."""
    import scipy.sparse as sp
    rp, ci, v, _ = ctx.ssh_get_matrix()
    A = sp.csr_matrix((v, ci - 1, rp - 1), shape=(len(rp) - 1, len(rp) - 1))
    ctx.ssh_solve_preconditioner(2)
    rs, es = [], []
    for s in range(rhs.shape[0]):
        for k in ks:
            if k == 0:
                xk = np.zeros(rhs.shape[1])
            else:
                xk, _ = ctx.ssh_solve_cg(np.zeros(rhs.shape[1]), rhs[s], soltol=1e-14, maxiter=k)
            e = sol[s] - xk
            if not np.any(e):
                continue
            rs.append(rhs[s] - A @ xk)
            es.append(e)
    return np.stack(rs), np.stack(es)


def model_training_set(ctx, mesh, state, nsteps=2000, every=20, seed=1, ks=(0, 5, 10, 20, 40, 80, 150)):
    """Training pairs of the learned solver from the model's own steps: the
    right-hand sides and solutions of `nsteps` steps of the plain solver
    (model_rhs_sequence, every `every`-th step kept), each expanded into the
    residual/error pairs plain CG meets on it after ks iterations
    (residual_pairs; k = 0 is the pair itself). On the 126,858-node operator,
    100 kept steps give 700 pairs; 40 epochs on them bring flexible CG from
    215 iterations (diagonal) to about 125 on held-out steps of another run."""
    b, x = model_rhs_sequence(ctx, mesh, state, nsteps=nsteps, every=every, seed=seed)
    return residual_pairs(ctx, b, x, ks=ks)


def train(ctx, rhs, sol, epochs=20, batch=8, lr=2e-3, seed=0):
    """Initialise the network and train it on the device; returns the loss per epoch."""
    ctx.nn_init(seed)
    return ctx.nn_train(rhs, sol, epochs=epochs, batch=batch, lr=lr)
