"""Pinning the oracle's restatement of the steps either side of the transport
(no GPU): vert_vel_ale with compute_CFLz and compute_Wvel_split, and
compute_ssh_rhs_ale (oce_ale.F90:1803-1938, 2107-3049).

Each check is fixed by the mathematics of the finite-volume divergence, not by
the oracle's code:
* a spatially uniform flow has zero divergence on every closed dual cell;
* a flow v = s (y - y0) sampled at element centroids has constant divergence
  s, and the median-dual flux of such a linear field is exact, so at nodes
  away from the walls W(k)·area(k) equals -s times the water volume below
  interface k (W is positive upward);
* every edge moves transport from one node to the other, so the node sums of
  W·area vanish level by level;
* the generator's flow has zero column-integrated divergence, so the surface
  W vanishes;
* no flow crosses the sea floor, so CFL_z at the bottom interface is zero;
* the W split caps the explicit part's Courant number at wsplit_maxcfl;
* the SSH right-hand side of alpha·(u + u_rhs) equals the column-integrated
  divergence, which is W·area at the surface;
* under zstar the column thickness changes by hbar - hbar_old, and with no
  horizontal divergence and a water flux balancing the ssh change the
  surface W vanishes;
* update_stiff_mat_ale adds the stiffness operator of the ssh change: it is
  symmetric, annihilates constants, and equals the stiffness part of
  init_stiff_mat_ale (pinned in test_oracle_pins.py) on a mesh whose element
  depths are -dhe.
The last test plants the mistakes 19-23 of fesom_oracle.c and checks that one
of these pins fails.
"""
import dataclasses

import numpy as np
import pytest

from fesom2_amd import meshgen as mg
from oracle import oracle as orc
from oracle.oracle import Oracle

R = mg.R_EARTH


@pytest.fixture(scope="module")
def bumps():
    mesh = mg.build_mesh(nx=20, ny=14, dx=20e3, jitter=0.2, seed=5, cartesian=True, nlayers=10,
                         hmax=2000.0, min_layers=4, max_layers=10, topo="bumps", bump_height=900.0)
    st = mg.build_state(mesh, seed=5, ntr=2, dt=900.0, dx=20e3)
    return mesh, st, Oracle(mesh)


@pytest.fixture(scope="module")
def flat():
    mesh = mg.build_mesh(nx=20, ny=14, dx=20e3, jitter=0.2, seed=6, cartesian=True, nlayers=10,
                         hmax=2000.0, min_layers=4, max_layers=10, topo="flat")
    st = mg.build_state(mesh, seed=6, ntr=2, dt=900.0, dx=20e3)
    return mesh, st, Oracle(mesh)


def _wet_elem(mesh):
    k = np.arange(1, mesh.nl)
    return (k[None, :] >= mesh.ulevels[:, None]) & (k[None, :] <= (mesh.nlevels - 1)[:, None])


def _closed_nodes(mesh):
    """Nodes whose dual cell is closed: no incident boundary edge."""
    bnd = mesh.edge_tri[:, 1] < 0
    open_ = np.zeros(mesh.nod2D, bool)
    open_[mesh.edges[bnd].ravel()] = True
    return ~open_


def _open(mesh):
    """Nodes with a free surface: not under an ice-shelf cavity."""
    return mesh.ulevels_nod2D == 1


def _elem_centroid_y(mesh):
    return mesh.coord_nod2D[mesh.elem2D_nodes, 1].mean(1) * R


def _transport_scale(mesh, uv):
    """Typical magnitude of one edge-face transport |u| dx h."""
    return float(np.abs(uv).max() * np.abs(mesh.edge_cross_dxdy).max() * mesh.helem.max())


# ----------------------------------------------------------------- checks
def check_uniform_flow(mesh, st, o):
    uv = np.zeros_like(st.uv)
    wet = _wet_elem(mesh)
    uv[..., 0] = np.where(wet, 0.3, 0.0)
    uv[..., 1] = np.where(wet, -0.2, 0.0)
    r = o.vert_vel_ale(uv, st.dt)
    area = np.where(mesh.area > 0, mesh.area, 1.0)
    wa = r["w"] * area
    scale = _transport_scale(mesh, uv)
    closed = _closed_nodes(mesh)
    return bool(np.abs(wa[closed]).max() < 1e-9 * scale)


def check_divergent_flow(mesh, st, o):
    """v = s (y - yc): divergence s everywhere; interior nodes away from the
    walls carry W(k)·area(k) = -s·(volume below interface k)."""
    s = 1e-6
    yc = _elem_centroid_y(mesh)
    y0 = 0.5 * (yc.min() + yc.max())
    wet = _wet_elem(mesh)
    uv = np.zeros_like(st.uv)
    uv[..., 1] = np.where(wet, s * (yc - y0)[:, None], 0.0)
    r = o.vert_vel_ale(uv, st.dt)
    ny = mesh.coord_nod2D[:, 1] * R
    lo, hi = ny.min(), ny.max()
    inner = _closed_nodes(mesh) & (ny > lo + 0.2 * (hi - lo)) & (ny < hi - 0.2 * (hi - lo))
    ok = True
    for n in np.where(inner)[0]:
        k0, k1 = mesh.ulevels_nod2D[n], mesh.nlevels_nod2D[n] - 1
        for k in range(k0, k1 + 1):
            vol_below = float(np.sum(mesh.hnode[n, k - 1:k1] * mesh.areasvol[n, k - 1:k1]))
            expect = -s * vol_below
            got = r["w"][n, k - 1] * mesh.area[n, k - 1]
            if not (got < 0 and abs(got - expect) <= 1e-9 * abs(expect)):
                ok = False
        if r["w"][n, k1] != 0.0:   # bottom interface
            ok = False
    return ok


def check_level_sums_vanish(mesh, st, o):
    rng = np.random.default_rng(3)
    wet = _wet_elem(mesh)
    uv = np.where(wet[..., None], rng.normal(0, 0.3, st.uv.shape), 0.0)
    # undo the division by area and the cumulative sum: layer divergences
    r = o.vert_vel_ale(uv, st.dt)
    wa = r["w"] * mesh.area
    div = wa[:, :-1] - wa[:, 1:]
    k = np.arange(1, mesh.nl)
    div = np.where(k[None, :] >= mesh.ulevels_nod2D[:, None], div, 0.0)   # wet layers only
    scale = _transport_scale(mesh, uv)
    return bool(np.abs(div.sum(0)).max() < 1e-9 * scale)


def check_column_flow_surface(mesh, st, o):
    r = o.vert_vel_ale(st.uv, st.dt)
    top = r["w"][np.arange(mesh.nod2D), mesh.ulevels_nod2D - 1]
    return bool(np.abs(top).max() < 1e-9 * np.abs(r["w"]).max())


def check_cfl(mesh, st, o):
    r = o.vert_vel_ale(st.uv, st.dt)
    n = np.arange(mesh.nod2D)
    bottom = r["cfl_z"][n, mesh.nlevels_nod2D - 1]
    top = r["cfl_z"][n, mesh.ulevels_nod2D - 1]
    top_expect = np.abs(r["w"][n, mesh.ulevels_nod2D - 1] * st.dt / mesh.hnode_new[n, mesh.ulevels_nod2D - 1])
    return bool(np.all(bottom == 0.0) and np.all(r["cfl_z"] >= 0.0) and np.allclose(top, top_expect, rtol=1e-14, atol=0)
                and r["cflmax"] == r["cfl_z"].max())


def check_wsplit(mesh, st, o):
    r0 = o.vert_vel_ale(st.uv, st.dt)
    cap = 0.3 * r0["cflmax"]
    r = o.vert_vel_ale(st.uv, st.dt, use_wsplit=True, wsplit_maxcfl=cap)
    w, we, wi, cfl = r["w"], r["w_e"], r["w_i"], r["cfl_z"]
    over = cfl > cap
    if not over.any():
        return False
    ok = np.allclose(we + wi, w, rtol=1e-14, atol=1e-30)
    # explicit Courant number at the cap where the total exceeded it
    ratio = np.where(over & (w != 0), we / np.where(w != 0, w, 1.0), 1.0)
    ok &= np.allclose((ratio * cfl)[over & (w != 0)], cap, rtol=1e-12)
    ok &= bool(np.all(wi[~over] == 0.0)) and bool(np.all(we[~over] == w[~over]))
    return bool(ok)


def check_ssh_rhs_column_divergence(mesh, st, o):
    rng = np.random.default_rng(4)
    wet = _wet_elem(mesh)
    uv = np.where(wet[..., None], rng.normal(0, 0.3, st.uv.shape), 0.0)
    rhs = o.compute_ssh_rhs_ale(uv, np.zeros_like(uv), np.zeros(mesh.nod2D), alpha=1.0)
    r = o.vert_vel_ale(uv, st.dt)
    n = np.arange(mesh.nod2D)
    k = mesh.ulevels_nod2D - 1
    col = r["w"][n, k] * mesh.area[n, k]
    scale = _transport_scale(mesh, uv)
    return bool(np.abs(rhs - col).max() < 1e-9 * scale)


def check_ssh_rhs_alpha_terms(mesh, st, o):
    rng = np.random.default_rng(5)
    wet = _wet_elem(mesh)
    uv = np.where(wet[..., None], rng.normal(0, 0.3, st.uv.shape), 0.0)
    old = rng.normal(0, 1e3, mesh.nod2D)
    base = o.compute_ssh_rhs_ale(uv, np.zeros_like(uv), np.zeros(mesh.nod2D), alpha=1.0)
    # alpha (u + u_rhs) with u_rhs = u and alpha = 1/2 scales by exact powers of two
    half = o.compute_ssh_rhs_ale(uv, uv, np.zeros(mesh.nod2D), alpha=0.5)
    ok = np.array_equal(half, base)
    with_old = o.compute_ssh_rhs_ale(uv, uv, old, alpha=0.5)
    ok &= np.allclose(with_old - half, 0.5 * old, rtol=1e-12, atol=1e-9 * np.abs(base).max())
    # full free surface: the water flux enters through the surface areasvol;
    # under a cavity (rigid ice-shelf lid) neither it nor ssh_rhs_old enters,
    # unless use_cavity_fw2press adds the freshwater as pressure
    wf = rng.normal(0, 1e-6, mesh.nod2D)
    fs = o.compute_ssh_rhs_ale(uv, uv, old, alpha=0.5, which_ale=2, water_flux=wf)
    fw = o.compute_ssh_rhs_ale(uv, uv, old, alpha=0.5, which_ale=2, water_flux=wf, use_cavity_fw2press=True)
    k = mesh.ulevels_nod2D - 1
    op = _open(mesh)
    wfa = -0.5 * wf * mesh.areasvol[np.arange(mesh.nod2D), k]
    tol = 1e-9 * np.abs(base).max()
    ok &= np.allclose((fs - with_old)[op], wfa[op], rtol=1e-9, atol=tol)
    ok &= np.array_equal(fw[op], fs[op])
    ok &= np.array_equal(fs[~op], half[~op])
    ok &= np.allclose((fw - half)[~op], wfa[~op], rtol=1e-9, atol=tol)
    return bool(ok)


def check_hbar_transport(mesh, st, o):
    """compute_hbar_ale: the surface rises by the column convergence, i.e.
    (hbar - hbar_old) areasvol / dt equals the pinned surface W area of
    vert_vel_ale, and ssh_rhs_old equals compute_ssh_rhs_ale with alpha = 1
    and no uv_rhs."""
    rng = np.random.default_rng(7)
    wet = _wet_elem(mesh)
    uv = np.where(wet[..., None], rng.normal(0, 0.3, st.uv.shape), 0.0)
    h0 = rng.normal(0, 0.5, mesh.nod2D)
    r = o.compute_hbar_ale(uv, st.dt, h0)
    n = np.arange(mesh.nod2D)
    k = mesh.ulevels_nod2D - 1
    col = o.vert_vel_ale(uv, st.dt)["w"][n, k] * mesh.area[n, k]
    scale = _transport_scale(mesh, uv)
    op = _open(mesh)
    ok = np.array_equal(r["hbar_old"], h0)
    ok &= np.abs(((r["hbar"] - h0) * mesh.areasvol[n, k] / st.dt - col)[op]).max() < 1e-9 * scale
    ok &= np.array_equal(r["hbar"][~op], h0[~op])   # the ice-shelf lid holds the surface
    rhs = o.compute_ssh_rhs_ale(uv, np.zeros_like(uv), np.zeros(mesh.nod2D), alpha=1.0)
    ok &= np.allclose(r["ssh_rhs_old"], rhs, rtol=1e-13, atol=1e-12 * scale)
    return bool(ok)


def check_hbar_freshwater(mesh, st, o):
    """At rest under the full free surface, a freshwater flux moves the
    surface by -water_flux dt at every node, the element thickness change is
    the mean of its nodes', and the areasvol-weighted volume change is
    -dt sum(water_flux areasvol)."""
    rng = np.random.default_rng(8)
    wf = rng.normal(0, 1e-5, mesh.nod2D)
    h0 = rng.normal(0, 0.5, mesh.nod2D)
    r = o.compute_hbar_ale(np.zeros_like(st.uv), st.dt, h0, which_ale=2, water_flux=wf)
    dh = r["hbar"] - h0
    op = _open(mesh)
    expect = np.where(op, -wf * st.dt, 0.0)   # cavity nodes keep hbar
    ok = np.allclose(dh, expect, rtol=1e-12, atol=1e-15)
    eo = mesh.ulevels == 1                    # elements under the shelf get no dhe
    ok &= np.allclose(r["dhe"], np.where(eo, expect[mesh.elem2D_nodes].mean(1), 0.0), rtol=1e-12, atol=1e-15)
    n = np.arange(mesh.nod2D)
    av = mesh.areasvol[n, mesh.ulevels_nod2D - 1]
    ok &= np.isclose((dh * av).sum(), -st.dt * (wf * av)[op].sum(), rtol=1e-10)
    return bool(ok)


def check_zstar(mesh, st, o):
    rng = np.random.default_rng(6)
    dh = rng.normal(0, 0.05, mesh.nod2D)
    wf = -dh / st.dt
    hbar_old = rng.normal(0, 0.5, mesh.nod2D)
    hbar = hbar_old + dh
    uv = np.zeros_like(st.uv)
    r = o.vert_vel_ale(uv, st.dt, which_ale=2, hbar=hbar, hbar_old=hbar_old, water_flux=wf)
    op = _open(mesh)
    ok = np.allclose((r["hnode_new"].sum(1) - mesh.hnode.sum(1))[op], dh[op], rtol=1e-9, atol=1e-12)
    ok &= np.array_equal(r["hnode_new"][~op], mesh.hnode[~op])   # cavity columns stay linfs
    top = r["w"][np.arange(mesh.nod2D), mesh.ulevels_nod2D - 1]
    ok &= np.abs(top).max() < 1e-12 * np.abs(wf).max() + 1e-18
    return bool(ok)


def check_update_stiff(mesh, st, o):
    import scipy.sparse as sp
    rng = np.random.default_rng(7)
    dhe = rng.normal(0.0, 0.4, mesh.elem2D)
    rp, ci, v = o.stiff_matrix(st.dt)
    upd = o.update_stiff_matrix(st.dt, dhe, rp, ci, np.zeros_like(v))
    # stiffness part of init_stiff_mat_ale on element depths -dhe
    m2 = dataclasses.replace(mesh, zbar_e_bot=mesh.zbar_e_srf - dhe)
    rp2, ci2, v2 = Oracle(m2).stiff_matrix(st.dt)
    if not (np.array_equal(rp, rp2) and np.array_equal(ci, ci2)):
        return False
    diag = rp[:-1]
    off = np.ones(len(v), bool)
    off[diag] = False
    # (zbar_e_srf - dhe) - zbar_e_srf rounds where the element lies under a cavity
    ok = np.allclose(upd[off], v2[off], rtol=0, atol=1e-13 * np.abs(v2).max())
    av = np.where(_open(mesh), mesh.areasvol[np.arange(mesh.nod2D), mesh.ulevels_nod2D - 1] / st.dt, 0.0)
    ok &= np.allclose(upd[diag], v2[diag] - av, rtol=0, atol=1e-12 * np.abs(v2).max())
    A = sp.csr_matrix((upd, ci, rp), shape=(mesh.nod2D, mesh.nod2D))
    scale = np.abs(upd).max()
    ok &= abs(A - A.T).max() < 1e-12 * scale
    ok &= np.abs(A @ np.ones(mesh.nod2D)).max() < 1e-12 * scale
    # additive: the update starts from the assembled values
    full = o.update_stiff_matrix(st.dt, dhe, rp, ci, v)
    ok &= np.allclose(full, v + upd, rtol=1e-14, atol=1e-12 * scale)
    return bool(ok)


# ------------------------------------------------------------------ tests
def test_uniform_flow_has_no_divergence(flat):
    assert check_uniform_flow(*flat)


def test_divergent_flow_sinks(flat):
    assert check_divergent_flow(*flat)


def test_level_sums_vanish(bumps):
    assert check_level_sums_vanish(*bumps)


def test_column_flow_leaves_surface_at_rest(bumps):
    assert check_column_flow_surface(*bumps)


def test_cfl_z(bumps):
    assert check_cfl(*bumps)


def test_wsplit_caps_explicit_courant(bumps):
    assert check_wsplit(*bumps)


def test_ssh_rhs_is_column_divergence(bumps):
    assert check_ssh_rhs_column_divergence(*bumps)


def test_ssh_rhs_alpha_and_free_surface_terms(bumps):
    assert check_ssh_rhs_alpha_terms(*bumps)


def test_zstar_thickness_and_surface(flat):
    assert check_zstar(*flat)


def test_update_stiff_mat(bumps):
    assert check_update_stiff(*bumps)


def test_hbar_transport(bumps):
    assert check_hbar_transport(*bumps)


def test_hbar_freshwater(bumps):
    assert check_hbar_freshwater(*bumps)


def _zlevel(mesh, st, dh, hnode=None, cfl_prev=None, min_hnode=0.5, lz=4):
    """zlevel vert_vel_ale at rest (uv = 0, no water flux) for a surface change dh."""
    import dataclasses
    m2 = mesh if hnode is None else dataclasses.replace(mesh, hnode=hnode, hnode_new=hnode.copy())
    o = Oracle(m2)
    z = np.zeros(mesh.nod2D)
    r = o.vert_vel_ale(np.zeros_like(st.uv), st.dt, which_ale=1, hbar=dh, hbar_old=z, water_flux=z,
                       cfl_prev=cfl_prev, min_hnode=min_hnode, lzstar_lev=lz)
    return m2, r


def _zlevel_laws(m2, st, r, dh):
    """Continuity (each layer's change = -(W top - W bottom) dt) and
    conservation (the layer changes add up to dh) at every open-ocean node."""
    dthk = r["hnode_new"] - m2.hnode
    w = r["w"]
    div = -(w[:, :-1] - w[:, 1:]) * st.dt
    scale = np.abs(dh).max()
    op = _open(m2)
    ok = np.allclose(dthk[op], div[op], rtol=0, atol=1e-12 * scale)
    ok &= np.allclose(dthk.sum(1)[op], dh[op], rtol=0, atol=1e-12 * scale)
    ok &= not np.any(dthk[~op])   # cavity columns stay linfs
    return bool(ok)


def check_zlevel(mesh, st, o):
    """zlevel: a small change goes into the surface layer; a fall that would
    thin it below min_hnode of its nominal thickness is spread over the top
    layers, none thinner than min_hnode nominal; layers whose previous CFL_z
    reached 0.95 keep their thickness; a rise first refills thinned layers,
    none above nominal. Continuity and conservation hold in every case."""
    nominal = mesh.zbar[:-1] - mesh.zbar[1:]
    h1 = mesh.hnode[:, 0]
    op = _open(mesh)
    ok = True
    # small change: surface layer only
    small = np.linspace(-0.1, 0.1, mesh.nod2D) * h1
    m2, r = _zlevel(mesh, st, small)
    ok &= _zlevel_laws(m2, st, r, small)
    ok &= np.allclose((r["hnode_new"][:, 0] - mesh.hnode[:, 0])[op], small[op], rtol=1e-13, atol=0)
    # large fall: spread, never below min_hnode of nominal
    big = -0.8 * h1
    m2, r = _zlevel(mesh, st, big)
    ok &= _zlevel_laws(m2, st, r, big)
    ok &= bool(np.all(r["hnode_new"][op, :4] >= 0.5 * nominal[None, :4] * (1 - 1e-12)))
    ok &= bool(np.all(r["hnode_new"][op, 1] < mesh.hnode[op, 1]))
    # previous CFL_z at the surface interface >= 0.95: the surface layer stays
    cfl = np.zeros((mesh.nod2D, mesh.nl))
    cfl[:, 0] = 0.97
    m2, r = _zlevel(mesh, st, big, cfl_prev=cfl)
    ok &= _zlevel_laws(m2, st, r, big)
    ok &= bool(np.all(r["hnode_new"][op, 0] == mesh.hnode[op, 0]))
    # rise over thinned layers 2 and 3: refilled from below, none above nominal
    hn = mesh.hnode.copy()
    hn[:, 1] = 0.6 * nominal[1]
    hn[:, 2] = 0.6 * nominal[2]
    rise = 0.3 * nominal[0] * np.ones(mesh.nod2D)
    m2, r = _zlevel(mesh, st, rise, hnode=hn)
    ok &= _zlevel_laws(m2, st, r, rise)
    ok &= bool(np.all(r["hnode_new"][op, 1:4] <= nominal[None, 1:4] * (1 + 1e-12)))
    ok &= bool(np.all(r["hnode_new"][op, 2] > hn[op, 2]))
    return bool(ok)


def test_zlevel_redistribution(flat):
    assert check_zlevel(*flat)


def test_generator_w_agrees(bumps):
    """The state generator's own W (numpy, another summation order) agrees."""
    mesh, st, o = bumps
    r = o.vert_vel_ale(st.uv, st.dt)
    w = r["w"].copy()
    w[np.arange(mesh.nod2D), mesh.ulevels_nod2D - 1] = 0.0   # the generator zeroes the surface W
    assert np.abs(w - st.w).max() <= 1e-12 * np.abs(st.w).max()


MUTATIONS = {
    19: lambda b, f: check_uniform_flow(*f),
    20: lambda b, f: check_divergent_flow(*f),
    21: lambda b, f: check_ssh_rhs_column_divergence(*b),
    22: lambda b, f: check_cfl(*b),
    23: lambda b, f: check_update_stiff(*b),
    34: lambda b, f: check_hbar_transport(*b),
    35: lambda b, f: check_hbar_freshwater(*b),
    36: lambda b, f: check_zlevel(*f),
    37: lambda b, f: check_zlevel(*f),
}


@pytest.mark.parametrize("mut", sorted(MUTATIONS))
def test_planted_mistake_is_caught(bumps, flat, mut):
    assert MUTATIONS[mut](bumps, flat), "check must pass on the faithful oracle"
    orc.set_mutation(mut)
    try:
        caught = not MUTATIONS[mut](bumps, flat)
    except Exception:
        caught = True
    finally:
        orc.set_mutation(0)
    assert caught, f"mutation {mut} not caught"
