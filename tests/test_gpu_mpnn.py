"""Learned SSH preconditioner on the GPU against its CPU reference
(oracle/mpnn_ref.py) and against plain diagonally preconditioned CG.

Tolerances: the device network computes in bf16 with FP32 accumulation on
MFMA; the reference rounds to bf16 at the same points but sums in another
order, so single bf16 roundings can differ by one unit (2^-8 relative) and
propagate. Forward outputs are compared to 2e-2 relative (L2), gradients by
cosine similarity >= 0.995 and 5e-2 relative L2; an MFMA operand-layout error
gives O(1) differences. The FP64 solve itself is held to the reference's
stopping rule (||b - A x|| / ||b|| < 1e-10).
"""
import numpy as np
import pytest

from fesom2_amd import meshgen as mg
from fesom2_amd import ssh_learned as sl
from oracle import mpnn_ref as nn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(c0, c0_oracle):
    from fesom2_amd import gpu
    mesh, state, _ = c0
    ctx = gpu.Context(0)
    ctx.mesh_init(mesh)
    ctx.init_stiff_mat_ale(state.dt)
    ctx.ssh_solve_preconditioner(2)
    rp, ci, v, _ = ctx.ssh_get_matrix()
    g = nn.Graph(rp - 1, ci - 1, v)
    yield mesh, state, ctx, g
    ctx.close()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _random_params(seed):
    p = nn.init_params(seed)
    rng = np.random.default_rng(seed + 1)
    p["w_out"] = rng.normal(0, 0.2, nn.C).astype(np.float32)
    p["b_out"] = np.array([0.05], np.float32)
    for l in range(1, nn.NLAYERS + 1):
        p[f"b{l}"] = rng.normal(0, 0.1, nn.C).astype(np.float32)
    return p


def test_param_count(setup):
    ctx = setup[2]
    assert ctx.nn_param_count() == nn.n_params()


def test_untrained_is_jacobi(setup):
    mesh, state, ctx, g = setup
    ctx.nn_init(3)
    r = np.random.default_rng(0).normal(size=mesh.nod2D)
    print("[gpu] nn apply, untrained", flush=True)
    z = ctx.nn_apply(r)
    np.testing.assert_allclose(z, r / g.diag, rtol=1e-13)


def test_forward_vs_reference(setup):
    mesh, state, ctx, g = setup
    p = _random_params(4)
    ctx.nn_set_weights(nn.flatten(p))
    r = np.random.default_rng(1).normal(size=mesh.nod2D)
    print("[gpu] nn forward vs reference", flush=True)
    z = ctx.nn_apply(r)
    zr, acts = nn.forward(g, p, r)
    y_dev = (z / g.isq - g.isq * r) / acts["s"]
    assert _rel(y_dev, acts["y"].astype(np.float64)) < 2e-2
    assert _rel(z, zr) < 2e-2


def test_gradient_vs_reference(setup):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    mesh, state, ctx, g = setup
    p = _random_params(5)
    ctx.nn_set_weights(nn.flatten(p))
    bs = sl.rhs_samples(mesh, 3, seed=11)
    A = sp.csr_matrix((g.values.copy(), g.colind.copy(), g.rowptr.copy()), shape=(g.n, g.n)).tocsc()
    xs = np.stack([spl.spsolve(A, b) for b in bs])
    print("[gpu] nn loss/gradient vs reference", flush=True)
    loss, grad = ctx.nn_loss_grad(bs, xs)
    lref, gref = nn.loss_and_grad(g, p, list(bs), list(xs))
    assert abs(loss - lref) <= 1e-2 * abs(lref)
    gd = nn.unflatten(grad.astype(np.float64))
    for k in gref:
        a, b = gd[k].ravel(), gref[k].ravel()
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-300))
        assert cos >= 0.995, (k, cos)
        assert _rel(a, b) < 5e-2, (k, _rel(a, b))


def test_training_and_learned_solve(setup):
    """Trained on residual/error pairs of plain-CG runs: the loss falls, the
    learned FCG converges to the reference tolerance without the fallback, with
    fewer iterations than the untrained (= diagonal) network, and its solution
    matches plain CG's."""
    import scipy.sparse as sp
    mesh, state, ctx, g = setup
    rhs = sl.rhs_samples(mesh, 24, seed=1)
    sol = sl.solve_pairs(ctx, rhs)
    rs, es = sl.residual_pairs(ctx, rhs[:16], sol[:16])
    print(f"[gpu] nn training on {len(rs)} residual pairs", flush=True)
    hist = sl.train(ctx, rs, es, epochs=30, batch=8, lr=2e-3)
    print("[gpu] loss per epoch", np.round(hist, 4).tolist(), flush=True)
    A = sp.csr_matrix((g.values.copy(), g.colind.copy(), g.rowptr.copy()), shape=(g.n, g.n))
    theta = ctx.nn_get_weights()
    res = []
    for s in range(16, 24):
        b = rhs[s]
        x, st = ctx.ssh_solve_learned(np.zeros(mesh.nod2D), b, soltol=1e-10, maxiter=2000)
        xp, stp = ctx.ssh_solve_cg(np.zeros(mesh.nod2D), b, soltol=1e-10, maxiter=5000)
        res.append((x, st, xp, stp, b))
    ctx.nn_init(0)   # untrained network: the diagonal preconditioner inside FCG
    its_untrained = [ctx.ssh_solve_learned(np.zeros(mesh.nod2D), rhs[s], soltol=1e-10, maxiter=5000)[1]["iters"]
                     for s in range(16, 24)]
    ctx.nn_set_weights(theta)
    its_trained = [r[1]["iters"] for r in res]
    print(f"[gpu] held-out iterations: trained FCG {its_trained}, untrained FCG {its_untrained}, "
          f"plain CG {[r[3]['iters'] for r in res]}; diagonal steps {[r[1]['nnegrz'] for r in res]}, "
          f"fallback {[r[1]['fallback_iters'] for r in res]}", flush=True)
    assert hist[-1] < 0.8 * hist[0]
    for x, st, xp, stp, b in res:
        assert st["converged"] and st["fallback_iters"] == 0, st
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10 * 1.01
        assert np.linalg.norm(x - xp) / np.linalg.norm(xp) < 1e-6
    assert np.mean(its_trained) < np.mean(its_untrained)


def test_fallback_at_cap(setup):
    """An FCG cap of 3 iterations hands over to plain CG, which converges."""
    import scipy.sparse as sp
    mesh, state, ctx, g = setup
    b = mg.ssh_rhs(mesh, 5)
    x, st = ctx.ssh_solve_learned(np.zeros(mesh.nod2D), b, soltol=1e-10, maxiter=3, fallback_maxiter=5000)
    print(f"[gpu] fallback: {st}", flush=True)
    assert st["fallback_iters"] > 0 and st["converged"]
    A = sp.csr_matrix((g.values.copy(), g.colind.copy(), g.rowptr.copy()), shape=(g.n, g.n))
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10 * 1.01
