"""CPU checks of the learned-preconditioner reference (oracle/mpnn_ref.py):
bf16 rounding, the zero-output network equals the diagonal preconditioner,
analytic gradients equal central finite differences (FP64 mode), and a few
Adam steps on a small mesh reduce the loss."""
import numpy as np
import pytest

from fesom2_amd import meshgen as mg
from oracle import mpnn_ref as nn
from oracle.oracle import Oracle


@pytest.fixture(scope="module")
def small_graph():
    mesh = mg.build_mesh(nx=16, ny=12, dx=20e3, jitter=0.2, seed=5, cartesian=True, nlayers=8,
                         hmax=3000.0, min_layers=3, max_layers=8, topo="bumps", bump_height=1000.0)
    rp, ci, v = Oracle(mesh).stiff_matrix(900.0)
    return mesh, nn.Graph(rp, ci, v)


def _pairs(mesh, g, n, seed):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    A = sp.csr_matrix((g.values.copy(), g.colind.copy(), g.rowptr.copy()), shape=(g.n, g.n))
    rng = np.random.default_rng(seed)
    bs = [rng.normal(0, 1, g.n) for _ in range(n)]
    return bs, [spl.spsolve(A.tocsc(), b) for b in bs]


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38, 1e-40], np.float32)
    r = nn.bf16(x)
    # ties to even at 1+2^-8 (-> 1.0) and 1+3*2^-8 (-> 1+2^-6); 1+2^-9 rounds down
    assert list(r[:5]) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5]
    assert np.all(nn.bf16(r) == r)


def test_zero_output_is_jacobi(small_graph):
    mesh, g = small_graph
    p = nn.init_params(1)
    r = np.random.default_rng(0).normal(size=g.n)
    z, _ = nn.forward(g, p, r)
    np.testing.assert_allclose(z, r / g.diag, rtol=1e-13)


def test_gradient_finite_differences(small_graph):
    mesh, g = small_graph
    p = nn.init_params(2)
    rng = np.random.default_rng(3)
    for k in p:   # nonzero output weights so every layer receives a gradient
        p[k] = (p[k] + rng.normal(0, 0.3, p[k].shape)).astype(np.float64)
    bs, xs = _pairs(mesh, g, 2, 4)
    L0, gr = nn.loss_and_grad(g, p, bs, xs, exact=True)
    theta = nn.flatten(p).astype(np.float64)
    gflat = nn.flatten(gr)
    for _ in range(6):
        d = rng.normal(size=theta.size)
        eps = 1e-7   # small enough that no ReLU input crosses zero between the two evaluations
        Lp, _ = nn.loss_and_grad(g, nn.unflatten(theta + eps * d), bs, xs, exact=True)
        Lm, _ = nn.loss_and_grad(g, nn.unflatten(theta - eps * d), bs, xs, exact=True)
        fd = (Lp - Lm) / (2 * eps)
        an = float(gflat @ d)
        assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-3), (fd, an)


def test_adam_reduces_loss(small_graph):
    mesh, g = small_graph
    p = nn.init_params(0)
    bs, xs = _pairs(mesh, g, 4, 7)
    L0, _ = nn.loss_and_grad(g, p, bs, xs)
    st = {}
    for _ in range(30):
        L, gr = nn.loss_and_grad(g, p, bs, xs)
        nn.adam_step(p, gr, st, 3e-3)
    L1, _ = nn.loss_and_grad(g, p, bs, xs)
    assert L1 < 0.9 * L0, (L0, L1)
