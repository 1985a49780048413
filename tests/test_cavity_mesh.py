"""Ice-shelf cavities (no GPU): the generated cavity levels follow the
reference's rules, and the oracle's pins hold on a mesh whose first wet level
varies from column to column.

Under a cavity the surface is a rigid ice-shelf lid (oce_ale.F90): no mass
term on the SSH diagonal (:1579-1581), no ssh_rhs_old or freshwater term in
the free-surface RHS unless use_cavity_fw2press (:1913-1926), hbar and dhe
kept (compute_hbar_ale), and the zstar / zlevel stretching skips the column
(:2356-2361, :2544-2553). Surface fluxes enter the first wet layer. The pins of
the open-ocean meshes are reused with those rules; planted mistakes in the
cavity branches must each fail one of them.

Edges whose elements differ in first and last wet level use all five level
segments of the horizontal flux (oce_adv_tra_hor.F90:166-249: one element
wet above the shared top, both, one below the shared bottom). A constant
tracer stays constant only if each segment takes exactly its elements'
transports, with and without the FCT limiter; planted segment mistakes must
fail that.
"""
import dataclasses

import numpy as np
import pytest

from fesom2_amd import meshgen as mg
from oracle import oracle as orc
from oracle.oracle import Oracle
from tests import test_oracle_diff_pins as F
from tests import test_oracle_dyn_pins as D
from tests import test_oracle_pins as P


@pytest.fixture(scope="module")
def cav():
    mesh = mg.build_mesh(nx=20, ny=14, dx=20e3, jitter=0.2, seed=5, cartesian=True, nlayers=10,
                         hmax=2000.0, min_layers=4, max_layers=10, topo="bumps", bump_height=900.0,
                         cavity=dict(frac=0.4, draft=600.0))
    st = mg.build_state(mesh, seed=5, ntr=2, dt=900.0, dx=20e3)
    assert mesh.ulevels.max() >= 3 and np.any(mesh.ulevels_nod2D != mesh.ulevels_nod2D_max)
    return mesh, st


def _level_rules(mesh):
    """fvom_init.F90:1127-1274: at least three wet layers per element and
    node, node levels the min / max over their elements, every wet prism at a
    cavity level with two wet neighbours, every node layer covered, twice
    where the node's elements leave a gap."""
    ul, nl_ = mesh.ulevels.astype(int), mesh.nlevels.astype(int)
    en = mesh.elem2D_nodes
    ok = bool(np.all(ul >= 1) and np.all(nl_ - ul >= 3))
    ok &= bool(np.all(mesh.nlevels_nod2D - mesh.ulevels_nod2D >= 3))
    umin = np.full(mesh.nod2D, 10 ** 6)
    umax = np.zeros(mesh.nod2D, int)
    np.minimum.at(umin, en.ravel(), np.repeat(ul, 3))
    np.maximum.at(umax, en.ravel(), np.repeat(ul, 3))
    ok &= np.array_equal(umin, mesh.ulevels_nod2D) and np.array_equal(umax, mesh.ulevels_nod2D_max)
    nb = mesh.elem_neighbors
    for nz in range(1, int(ul.max()) + 1):
        wet = (ul <= nz) & (nz < nl_)
        nbw = (nb >= 0) & wet[np.maximum(nb, 0)]
        ok &= bool(np.all(nbw.sum(1)[wet] >= 2))
    for n in range(mesh.nod2D):
        els = mesh.nod_in_elem2D[n, :mesh.nod_in_elem2D_num[n]]
        need = 2 if nl_[els].min() < ul[els].max() else 1
        for nz in range(mesh.ulevels_nod2D[n], mesh.nlevels_nod2D[n]):
            cnt = int(np.sum((ul[els] <= nz) & (nz < nl_[els])))
            ok &= cnt >= need and mesh.area[n, nz - 1] > 0
    return bool(ok)


def test_cavity_levels_follow_the_reference_rules(cav):
    assert _level_rules(cav[0])
    mesh, _, _ = mg.make_config("c0cav")
    assert mesh.ulevels.max() >= 4 and _level_rules(mesh)


def test_cavity_thicknesses_and_depths(cav):
    """Dry layers above the cavity have no thickness; Z_3d_n and zbar_3d_n
    hold the nominal levels there (oce_ale.F90:340-373)."""
    mesh = cav[0]
    k = np.arange(1, mesh.nl)
    above = k[None, :] < mesh.ulevels_nod2D[:, None]
    assert above.any()
    assert not np.any(mesh.hnode[above])
    Z = np.broadcast_to(mesh.Z[None, :], above.shape)
    assert np.array_equal(mesh.Z_3d_n[above], Z[above])
    e_above = k[None, :] < mesh.ulevels[:, None]
    assert not np.any(mesh.helem[e_above])
    assert np.array_equal(mesh.zbar_e_srf, mesh.zbar[mesh.ulevels - 1])


def check_surface_budget(mesh, st):
    """At rest (u = W = 0), one step changes each tracer's content by the
    surface flux through the first wet layer's area: -dt Q / (rho cp) for
    temperature, dt (virtual_salt + relax_salt) for salinity (linfs)."""
    rng = np.random.default_rng(29)
    N2 = mesh.nod2D
    hf = rng.normal(0.0, 100.0, N2)
    vs = rng.normal(0.0, 1e-5, N2)
    rs = rng.normal(0.0, 1e-5, N2)
    z = np.zeros_like(st.w)
    Kv = 10.0 ** rng.uniform(-5, -2, st.Kv.shape)
    st2 = dataclasses.replace(st, uv=np.zeros_like(st.uv), w=z, w_e=z, w_i=z, Kv=Kv, heat_flux=hf,
                              virtual_salt=vs, relax_salt=rs)
    v, vo = st.values.copy(), st.valuesold.copy()
    Oracle(mesh).step(st2, v, vo)
    vol = mesh.areasvol[:, :mesh.nl - 1] * mesh.hnode
    top = mesh.areasvol[np.arange(N2), mesh.ulevels_nod2D - 1]
    exp = [(-st.dt * hf / 4.2e6 * top).sum(), (st.dt * (vs + rs) * top).sum()]
    ok = True
    for i in range(2):
        got = (vol * (v[i] - st.values[i])).sum()
        ok &= abs(got - exp[i]) <= 1e-13 * np.abs(vol * st.values[i]).sum()
    return bool(ok)


def check_meltwater_budget(mesh, st):
    """Ice-shelf water (ID 501) at rest: one step changes its content by
    dt water_flux (T - 1) areasvol at the first wet layer of the cavity
    columns only, T the tracer there before the step."""
    rng = np.random.default_rng(31)
    N2 = mesh.nod2D
    wf = rng.normal(0.0, 1e-5, N2)
    z = np.zeros_like(st.w)
    st3 = mg.build_state(mesh, seed=5, ntr=3, dt=st.dt, dx=20e3)
    st3 = dataclasses.replace(st3, ids=np.array([1, 2, 501], np.int32), uv=np.zeros_like(st3.uv), w=z, w_e=z,
                              w_i=z, water_flux=wf)
    v, vo = st3.values.copy(), st3.valuesold.copy()
    Oracle(mesh).step(st3, v, vo)
    n = np.arange(N2)
    top = mesh.ulevels_nod2D - 1
    cav = mesh.ulevels_nod2D > 1
    exp = (st3.dt * wf * (st3.values[2][n, top] - 1.0) * mesh.areasvol[n, top])[cav].sum()
    vol = mesh.areasvol[:, :mesh.nl - 1] * mesh.hnode
    got = (vol * (v[2] - st3.values[2])).sum()
    return bool(abs(got - exp) <= 1e-12 * abs(exp) + 1e-13 * np.abs(vol * st3.values[2]).sum())


def test_meltwater_budget_under_cavity(cav):
    assert check_meltwater_budget(*cav)


def forcing_tracer_state(mesh, st, hosing_depth, seed=37):
    """Six tracers with the IDs 101, 102, 103, 104, 304 and 2, at rest, with
    random per-node fluxes prec_rain, wiso_flux_oce(:,1..3), hosing_flux and a
    random per-layer hosing_flux3D."""
    rng = np.random.default_rng(seed)
    N2, L = mesh.nod2D, mesh.nl - 1
    z = np.zeros_like(st.w)
    st6 = mg.build_state(mesh, seed=6, ntr=6, dt=st.dt, dx=20e3)
    return dataclasses.replace(
        st6, ids=np.array([101, 102, 103, 104, 304, 2], np.int32), uv=np.zeros_like(st6.uv), w=z, w_e=z, w_i=z,
        prec_rain=rng.normal(0.0, 1e-5, N2), wiso_flux_oce=rng.normal(0.0, 1e-5, (3, N2)),
        hosing_flux=rng.normal(0.0, 1e-5, N2), hosing_flux3D=rng.normal(0.0, 1e-5, (N2, L)),
        hosing_depth=hosing_depth)


def check_forcing_tracer_budget(mesh, st):
    """bc_surface of the IDs 101-104 and 304 (oce_ale_tracer.F90:2059-2079) at
    rest: one step changes each tracer's content by dt * flux * areasvol at
    the first wet layer, with flux prec_rain (101), wiso_flux_oce column
    ID-101 (102-104) or hosing_flux (304); in depth mode the hosing tracer
    takes dt * hosing_flux3D * areasvol at every wet layer
    (oce_ale_tracer.F90:1374-1380). The implicit diffusion moves content
    only within a column."""
    n = np.arange(mesh.nod2D)
    top = mesh.ulevels_nod2D - 1
    L = mesh.nl - 1
    vol = mesh.areasvol[:, :L] * mesh.hnode
    k = np.arange(1, L + 1)
    wet = (k[None, :] >= mesh.ulevels_nod2D[:, None]) & (k[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])
    ok = True
    for depth in (False, True):
        s6 = forcing_tracer_state(mesh, st, depth)
        v, vo = s6.values.copy(), s6.valuesold.copy()
        Oracle(mesh).step(s6, v, vo)
        terms = [s6.dt * f * mesh.areasvol[n, top] for f in
                 (s6.prec_rain, s6.wiso_flux_oce[0], s6.wiso_flux_oce[1], s6.wiso_flux_oce[2])]
        if depth:
            terms.append(np.where(wet, s6.dt * s6.hosing_flux3D * mesh.areasvol[:, :L], 0.0))
        else:
            terms.append(s6.dt * s6.hosing_flux * mesh.areasvol[n, top])
        for i, t in enumerate(terms):
            got = (vol * (v[i] - s6.values[i])).sum()
            ok &= abs(got - t.sum()) <= 1e-12 * np.abs(t).sum() + 1e-13 * np.abs(vol * s6.values[i]).sum()
    return bool(ok)


def test_forcing_tracer_budget_under_cavity(cav):
    assert check_forcing_tracer_budget(*cav)


def test_forcing_tracer_needs_its_flux(cav):
    """A tracer with ID 101 and no prec_rain is refused, like an unknown ID."""
    mesh, st = cav
    s6 = dataclasses.replace(forcing_tracer_state(mesh, st, False), prec_rain=None)
    v, vo = s6.values.copy(), s6.valuesold.copy()
    with pytest.raises(Exception):
        Oracle(mesh).step(s6, v, vo)


def check_constant_unlimited(mesh, st):
    """A constant stays constant without the FCT limiter too (MFCT/QR4C and
    MUSCL/PPM, lim NON): the horizontal flux of every segment of an edge,
    top-only, shared and bottom-only levels, balances the W that continuity
    built from the same element transports."""
    v0 = np.where(P._wet(mesh)[None], 5.0, 0.0) * np.ones((st.ntr, 1, 1))
    ok = True
    for sch in (dict(lim="NON"), dict(hor="MUSCL", ver="PPM", lim="NON")):
        v = P._step(mesh, st, 2, values=v0, schemes=sch)
        w = P._wet(mesh)
        ok &= max(np.max(np.abs(v[i][w] - 5.0)) for i in range(st.ntr)) < 1e-12 * 5.0
    return bool(ok)


def _segments_present(mesh):
    """Edges whose two elements differ in first (segments A/B) and last
    (segments D/E) wet level, so that every segment of the flux is in use."""
    el1, el2 = mesh.edge_tri[:, 0], mesh.edge_tri[:, 1]
    two = el2 >= 0
    e2 = np.maximum(el2, 0)
    top = two & (mesh.ulevels[el1] != mesh.ulevels[e2])
    bot = two & (mesh.nlevels[el1] != mesh.nlevels[e2])
    a = top & (mesh.ulevels[el1] < mesh.ulevels[e2])
    b = top & (mesh.ulevels[el1] > mesh.ulevels[e2])
    d = bot & (mesh.nlevels[el1] > mesh.nlevels[e2])
    e = bot & (mesh.nlevels[el1] < mesh.nlevels[e2])
    return all(x.any() for x in (a, b, d, e))


def test_all_flux_segments_in_use(cav):
    assert _segments_present(cav[0])


def test_constant_preserved_without_limiter_under_cavity(cav):
    assert check_constant_unlimited(*cav)


def check_cavity_stiffness(mesh, st):
    """The pinned matrix properties, and cavity rows with no mass term: their
    row sums vanish."""
    o = Oracle(mesh)
    rp, ci, v = o.stiff_matrix(st.dt)
    rs = np.add.reduceat(v, rp[:-1])
    cavn = mesh.ulevels_nod2D > 1
    return P.check_stiffness(mesh, st) and bool(np.abs(rs[cavn]).max() < 1e-12 * np.abs(v).max())


_TRACER = ["check_constant", "check_conservation", "check_fct_bounds", "check_vdiff_fv", "check_vert_impl",
           "check_vdiff_wimpl", "check_mfct_linear", "check_qr4c_cubic", "check_ppm_linear", "check_cdiff_linear",
           "check_fct_padding"]
_DYN = ["check_level_sums_vanish", "check_column_flow_surface", "check_cfl", "check_wsplit",
        "check_ssh_rhs_column_divergence", "check_ssh_rhs_alpha_terms", "check_hbar_transport",
        "check_hbar_freshwater", "check_zstar", "check_update_stiff", "check_zlevel"]
_DIFF = ["check_hor_diff_linear", "check_hor_diff_conserves", "check_hor_redi_shift", "check_expl_vdiff_linear",
         "check_expl_vdiff_budget", "check_redi_expl_linear", "check_redi_k33"]


@pytest.mark.parametrize("name", _TRACER)
def test_tracer_pins_under_cavity(cav, name):
    assert getattr(P, name)(*cav)


@pytest.mark.parametrize("name", _DYN)
def test_dynamics_pins_under_cavity(cav, name):
    mesh, st = cav
    assert getattr(D, name)(mesh, st, Oracle(mesh))


@pytest.mark.parametrize("name", _DIFF)
def test_diffusion_pins_under_cavity(cav, name):
    assert getattr(F, name)(*cav)


def test_surface_flux_budget_under_cavity(cav):
    assert check_surface_budget(*cav)


def test_ssh_under_cavity(cav):
    assert check_cavity_stiffness(*cav)
    assert P.check_cg(*cav)


MUTATIONS = {
    38: lambda c: check_cavity_stiffness(*c),
    39: lambda c: D.check_zstar(c[0], c[1], Oracle(c[0])),
    40: lambda c: D.check_ssh_rhs_alpha_terms(c[0], c[1], Oracle(c[0])),
    41: lambda c: check_surface_budget(*c),
    42: lambda c: P.check_constant(*c),
    43: lambda c: check_constant_unlimited(*c),
    44: lambda c: P.check_constant(*c),
    46: lambda c: check_meltwater_budget(*c),
    47: lambda c: check_forcing_tracer_budget(*c),
    48: lambda c: check_forcing_tracer_budget(*c),
}


@pytest.mark.parametrize("mut", sorted(MUTATIONS))
def test_planted_cavity_mistake_is_caught(cav, mut):
    """Each planted mistake in a cavity branch of the oracle fails its check."""
    assert MUTATIONS[mut](cav), "check must pass on the faithful oracle"
    orc.set_mutation(mut)
    try:
        caught = not MUTATIONS[mut](cav)
    except Exception:
        caught = True
    finally:
        orc.set_mutation(0)
    assert caught, f"mutation {mut} not caught"
