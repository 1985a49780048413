"""Pinning the oracle's diffusion routines outside the implicit vertical solve
(no GPU): diff_part_hor_redi (horizontal and Redi diffusion),
diff_ver_part_expl_ale (explicit vertical diffusion, i_vert_diff off),
diff_ver_part_redi_expl and the Redi K33 terms of diff_ver_part_impl_ale.

Each check compares with a value fixed independently of the oracle's code:
* the dual-cell flux of a linear field, with the dual segments rebuilt from the
  node coordinates;
* closed forms for linear vertical profiles;
* column budgets;
* the Redi identity K33 = Kv + Ki S3^2 for uniform slopes.
The last test plants each mistake (29-33 in fesom_oracle.c) and checks that
one of these pins fails.
"""
import dataclasses

import numpy as np
import pytest

from fesom2_amd import meshgen as mg
from oracle import oracle as orc

R = mg.R_EARTH


@pytest.fixture(scope="module")
def small():
    mesh = mg.build_mesh(nx=20, ny=14, dx=20e3, jitter=0.2, seed=5, cartesian=True, nlayers=10,
                         hmax=2000.0, min_layers=4, max_layers=10, topo="bumps", bump_height=900.0)
    st = mg.build_state(mesh, seed=5, ntr=2, dt=900.0, dx=20e3)
    return mesh, st


def _stretched(mesh, seed=3):
    """The same mesh with layer interfaces at non-uniform depths (zbar_3d_n,
    Z_3d_n mid-layer), and random node thicknesses hnode_new."""
    rng = np.random.default_rng(seed)
    L = mesh.nl - 1
    dz = 50.0 * 1.35 ** np.arange(L)
    zb = np.concatenate([[0.0], -np.cumsum(dz)])
    zbar = np.tile(zb, (mesh.nod2D, 1))
    Z = np.tile(0.5 * (zb[:-1] + zb[1:]), (mesh.nod2D, 1))
    hn = mesh.hnode_new * rng.uniform(0.6, 1.6, mesh.hnode_new.shape)
    return dataclasses.replace(mesh, zbar_3d_n=zbar, Z_3d_n=Z, hnode_new=hn)


def _seam_nodes(mesh):
    """Nodes of an element that straddles the periodic seam (a linear field in x
    is discontinuous there)."""
    x = mesh.coord_nod2D[mesh.elem2D_nodes, 0]
    bad = (x.max(1) - x.min(1)) > mesh.cyclic_length / 4
    out = np.zeros(mesh.nod2D, bool)
    out[mesh.elem2D_nodes[bad].ravel()] = True
    return out


def _dual_normals(mesh):
    """Per edge and side (el1, el2): the dual segment from the edge midpoint to
    the element centroid, rotated a quarter turn to point from node 1's cell
    into node 2's, in metres, from the node coordinates (cartesian mesh)."""
    c = mesh.coord_nod2D
    L = mesh.cyclic_length
    n1 = mesh.edges[:, 0]

    def rel(b):
        return np.stack([mg._trim_cyclic(c[b, 0] - c[n1, 0], L) * R, (c[b, 1] - c[n1, 1]) * R], -1)

    p2 = rel(mesh.edges[:, 1])
    mid = 0.5 * p2
    nu = np.zeros((mesh.edge2D, 2, 2))
    for j in range(2):
        el = mesh.edge_tri[:, j]
        ok = el >= 0
        en = mesh.elem2D_nodes[np.where(ok, el, 0)]
        cen = (rel(en[:, 0]) + rel(en[:, 1]) + rel(en[:, 2])) / 3.0
        s = cen - mid
        v = np.stack([s[:, 1], -s[:, 0]], -1)
        sgn = np.sign((v * p2).sum(-1))
        nu[:, j] = np.where(ok[:, None], v * sgn[:, None], 0.0)
    return nu


def _present(mesh, el, k):
    return (el >= 0) & (mesh.ulevels[np.maximum(el, 0)] <= k) & (mesh.nlevels[np.maximum(el, 0)] - 1 >= k)


def _dual_flux_divergence(mesh, Ki, F, helem, dt):
    """del_ttf of a constant flux density F (vector per level, Ki-scaled by the
    edge's mean Ki) through each node's dual cell: sum over its edges of
    +-Kh h F.nu dt / areasvol, h the layer thickness of the element(s) present
    (the mean of both where both are)."""
    nu = _dual_normals(mesh)
    n1, n2 = mesh.edges[:, 0], mesh.edges[:, 1]
    el1, el2 = mesh.edge_tri[:, 0], mesh.edge_tri[:, 1]
    out = np.zeros((mesh.nod2D, mesh.nl - 1))
    for k in range(1, mesh.nl):
        Fk = F[k - 1]
        p1, p2 = _present(mesh, el1, k), _present(mesh, el2, k)
        h1 = helem[el1, k - 1]
        h2 = helem[np.maximum(el2, 0), k - 1]
        f1, f2 = nu[:, 0] @ Fk, nu[:, 1] @ Fk
        Kh = (Ki[n1, k - 1] + Ki[n2, k - 1]) / 2.0
        fl = np.where(p1 & p2, 0.5 * (h1 + h2) * (f1 + f2), np.where(p1, h1 * f1, np.where(p2, h2 * f2, 0.0)))
        fl = Kh * fl * dt
        on = p1 | p2
        np.add.at(out[:, k - 1], n1[on], fl[on] / mesh.areasvol[n1[on], k - 1])
        np.add.at(out[:, k - 1], n2[on], -fl[on] / mesh.areasvol[n2[on], k - 1])
    return out


# ---------------------------------------------------------------- checks
def check_hor_diff_linear(mesh, st):
    """Ki diffusion of T = G.x: the flux K G through the dual segments rebuilt
    from the coordinates, with random Ki per node and level and random element
    thicknesses; away from the seam."""
    rng = np.random.default_rng(21)
    helem = mesh.helem * rng.uniform(0.5, 1.5, mesh.helem.shape)
    m2 = dataclasses.replace(mesh, helem=helem)
    o = orc.Oracle(m2)
    Ki = rng.uniform(50.0, 500.0, (mesh.nod2D, mesh.nl - 1))
    G = np.array([3e-5, -7e-5])
    x, y = mesh.coord_nod2D[:, 0] * R, mesh.coord_nod2D[:, 1] * R
    T = np.repeat((G[0] * x + G[1] * y)[:, None], mesh.nl - 1, 1)
    got = o.hor_diff(dataclasses.replace(st, Ki=Ki), T)
    exp = _dual_flux_divergence(mesh, Ki, np.tile(G, (mesh.nl - 1, 1)), helem, st.dt)
    keep = ~_seam_nodes(mesh)
    scale = np.abs(exp[keep]).max()
    return scale > 0 and np.abs(got[keep] - exp[keep]).max() <= 1e-9 * scale


def check_hor_diff_conserves(mesh, st):
    """Every edge moves as much into one node as out of the other: the
    areasvol-weighted sum of each level's increments vanishes."""
    rng = np.random.default_rng(22)
    o = orc.Oracle(mesh)
    Ki = rng.uniform(50.0, 500.0, (mesh.nod2D, mesh.nl - 1))
    T = rng.normal(10.0, 2.0, (mesh.nod2D, mesh.nl - 1))
    got = o.hor_diff(dataclasses.replace(st, Ki=Ki), T)
    tot = (got * mesh.areasvol[:, :-1]).sum(0)
    return np.abs(tot).max() <= 1e-12 * np.abs(got * mesh.areasvol[:, :-1]).sum()


def check_hor_redi_shift(mesh, st):
    """Redi with uniform slopes S and a vertical gradient tr_z = g at both
    interfaces of a layer: the flux K (grad T + S g) of a horizontally uniform T
    equals the plain flux of the linear field with gradient g S."""
    rng = np.random.default_rng(23)
    o = orc.Oracle(mesh)
    L = mesh.nl - 1
    g = 2e-3
    T = np.zeros((mesh.nod2D, L))
    T[:, 0] = 15.0
    dz = mesh.zbar[:-1] - mesh.zbar[1:]   # nominal thicknesses: T uniform on a level, cavity or not
    for k in range(1, L):   # tr_z = (T(k-1) - T(k)) / (h(k-1)/2 + h(k)/2) = g
        T[:, k] = T[:, k - 1] - g * 0.5 * (dz[k - 1] + dz[k])
    S = np.zeros((mesh.nod2D, L, 3))
    S[..., 0], S[..., 1], S[..., 2] = 0.02, -0.035, 0.01
    Ki = np.full((mesh.nod2D, L), 300.0)
    got = o.hor_diff(dataclasses.replace(st, Ki=Ki, slope_tapered=S, redi=True), T)
    exp = _dual_flux_divergence(mesh, Ki, np.tile(g * S[0, 0, :2], (L, 1)), mesh.helem, st.dt)
    # layers whose node and edge neighbours have interior interfaces above and below
    nb = [set() for _ in range(mesh.nod2D)]
    for a, b in mesh.edges:
        nb[a].add(b)
        nb[b].add(a)
    bad = 0
    checked = 0
    scale = np.abs(exp).max()
    for n in range(mesh.nod2D):
        kmin = max(mesh.ulevels_nod2D_max[m] for m in nb[n] | {n}) + 1
        kmax = min(mesh.nlevels_nod2D[m] for m in nb[n] | {n}) - 2
        for k in range(kmin, kmax + 1):
            checked += 1
            if abs(got[n, k - 1] - exp[n, k - 1]) > 1e-9 * scale:
                bad += 1
    return checked > 100 and bad == 0


def check_expl_vdiff_linear(mesh, st):
    """T = a Z on stretched layers, uniform Kv: the flux through every
    interior interface k is Kv a area_k (the exact gradient of the linear
    profile), the surface interface carries the forcing and the bottom none."""
    m2 = _stretched(mesh)
    o = orc.Oracle(m2)
    a = 1e-3
    kv = 3e-4
    Kv = np.full(st.Kv.shape, kv)
    hf = np.linspace(-200.0, 300.0, mesh.nod2D)
    st2 = dataclasses.replace(st, Kv=Kv, heat_flux=hf)
    T = a * m2.Z_3d_n
    got = o.expl_vdiff(st2, T, ID=1)
    dt = st.dt
    for n in range(mesh.nod2D):
        u, l = mesh.ulevels_nod2D[n], mesh.nlevels_nod2D[n] - 1
        dzb = m2.zbar_3d_n[n, :-1] - m2.zbar_3d_n[n, 1:]
        av = mesh.areasvol[n]
        flux = np.zeros(mesh.nl + 1)     # at interfaces 1..nl
        flux[u] = -hf[n] / 4.2e6
        flux[u + 1:l + 1] = kv * a * mesh.area[n, u:l]
        exp = np.zeros(mesh.nl - 1)
        ks = np.arange(u, l + 1) - 1
        exp[ks] = (flux[ks + 1] - flux[ks + 2]) / dzb[ks] * dt / av[ks]
        if not np.allclose(got[n, ks], exp[ks], rtol=1e-10, atol=1e-10 * np.abs(exp).max()):
            return False
    return True


def check_expl_vdiff_budget(mesh, st):
    """Column budget of the explicit vertical diffusion: sum over layers of
    del_ttf areasvol (zbar_k - zbar_k+1) / dt equals the surface flux."""
    m2 = _stretched(mesh)
    o = orc.Oracle(m2)
    rng = np.random.default_rng(24)
    T = rng.normal(35.0, 0.5, (mesh.nod2D, mesh.nl - 1))
    Kv = 10 ** rng.uniform(-5, -2, st.Kv.shape)
    vs = rng.normal(0.0, 1e-5, mesh.nod2D)
    st2 = dataclasses.replace(st, Kv=Kv, virtual_salt=vs, relax_salt=np.zeros(mesh.nod2D),
                              real_salt_flux=np.zeros(mesh.nod2D))
    got = o.expl_vdiff(st2, T, ID=2)
    dzb = m2.zbar_3d_n[:, :-1] - m2.zbar_3d_n[:, 1:]
    terms = got * mesh.areasvol[:, :-1] * dzb / st.dt
    col = terms.sum(1)
    # the interior fluxes (area-weighted) cancel; round-off scales with them
    tol = 1e-13 * np.abs(terms).sum(1) + 1e-13 * np.abs(vs)
    return bool(np.all(np.abs(col - vs) <= tol))


def check_redi_expl_linear(mesh, st):
    """Explicit Redi vertical flux of T = G.x with uniform slopes and Ki at
    nodes whose elements reach their full depth: the node gradient is G, every
    interior flux is Ki (S.G) area, so the surface layer loses and the bottom
    layer gains Ki (S.G) area dt / areasvol."""
    o = orc.Oracle(mesh)
    L = mesh.nl - 1
    G = np.array([4e-5, 2.5e-5])
    x, y = mesh.coord_nod2D[:, 0] * R, mesh.coord_nod2D[:, 1] * R
    T = np.repeat((G[0] * x + G[1] * y)[:, None], L, 1)
    S = np.zeros((mesh.nod2D, L, 3))
    S[..., 0], S[..., 1] = 0.015, -0.02
    Ki = np.full((mesh.nod2D, L), 250.0)
    got = o.redi_expl(dataclasses.replace(st, Ki=Ki, slope_tapered=S, redi=True), T)
    full = (mesh.nlevels_nod2D_min == mesh.nlevels_nod2D) & ~_seam_nodes(mesh)
    sg = 250.0 * (0.015 * G[0] - 0.02 * G[1])
    checked = 0
    for n in np.nonzero(full)[0]:
        u, l = mesh.ulevels_nod2D[n], mesh.nlevels_nod2D[n] - 1
        av = mesh.areasvol[n]
        exp = np.zeros(L)
        exp[u - 1] = -sg * mesh.area[n, u] * st.dt / av[u - 1]
        exp[l - 1] = sg * mesh.area[n, l - 1] * st.dt / av[l - 1]
        ks = np.arange(u, l + 1) - 1
        if not np.allclose(got[n, ks], exp[ks], rtol=1e-9, atol=1e-9 * abs(exp).max()):
            return False
        checked += 1
    return checked > 20


def check_redi_k33(mesh, st):
    """Implicit diffusion with Redi and uniform Ki S3^2 = k equals the plain
    implicit diffusion with Kv + k (the interface interpolation weights sum to
    one on any layer thicknesses)."""
    m2 = _stretched(mesh)
    o = orc.Oracle(m2)
    rng = np.random.default_rng(25)
    L = mesh.nl - 1
    T = rng.normal(10.0, 2.0, (mesh.nod2D, L))
    Kv = 10 ** rng.uniform(-5, -3, st.Kv.shape)
    S = np.zeros((mesh.nod2D, L, 3))
    S[..., 2] = 0.02
    Ki = np.full((mesh.nod2D, L), 2.0)
    k33 = 2.0 * 0.02 * 0.02
    T1, T2 = T.copy(), T.copy()
    o.vdiff(dataclasses.replace(st, Kv=Kv, Ki=Ki, slope_tapered=S, redi=True), T1, ID=301)
    o.vdiff(dataclasses.replace(st, Kv=Kv + k33), T2, ID=301)
    return np.allclose(T1, T2, rtol=1e-12, atol=1e-12) and not np.allclose(T1, T, rtol=1e-6)


# ---------------------------------------------------------------- tests
def test_hor_diff_linear_field(small):
    assert check_hor_diff_linear(*small)


def test_hor_diff_conservation(small):
    assert check_hor_diff_conserves(*small)


def test_hor_redi_slope_shift(small):
    assert check_hor_redi_shift(*small)


def test_expl_vdiff_linear_and_budget(small):
    assert check_expl_vdiff_linear(*small)
    assert check_expl_vdiff_budget(*small)


def test_redi_expl_linear(small):
    assert check_redi_expl_linear(*small)


def test_redi_k33_identity(small):
    assert check_redi_k33(*small)


MUTATIONS = {
    29: check_expl_vdiff_linear,
    30: check_hor_diff_linear,
    31: check_hor_diff_linear,
    32: check_redi_expl_linear,
    33: check_redi_k33,
}


@pytest.mark.parametrize("mut", sorted(MUTATIONS))
def test_planted_mistake_is_caught(small, mut):
    assert MUTATIONS[mut](*small), "check must pass on the faithful oracle"
    orc.set_mutation(mut)
    try:
        caught = not MUTATIONS[mut](*small)
    except Exception:
        caught = True
    finally:
        orc.set_mutation(0)
    assert caught, f"mutation {mut} not caught"
