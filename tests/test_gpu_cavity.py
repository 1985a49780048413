"""The HIP path on a mesh under an ice-shelf cavity, against the oracle.

The c0 channel with an ice shelf over its southern third: elements start at
levels 1-6 (ulevels), nodes at the front have ulevels_nod2D below
ulevels_nod2D_max, and surface fluxes enter at the first wet level. Every
transport scheme, the diffusion terms, the steps either side (vert_vel_ale
linfs/zstar/zlevel, compute_ssh_rhs_ale, compute_hbar_ale, the stiffness
matrix update) and the SSH solve run on it; tracer, W and matrix values are
compared bit for bit, the CG solution to the solver's tolerance."""
import dataclasses

import numpy as np
import pytest

from fesom2_amd import meshgen as mg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cav():
    from oracle.oracle import Oracle
    mesh, state, _ = mg.make_config("c0cav")
    rng = np.random.default_rng(41)
    N2 = mesh.nod2D
    state = dataclasses.replace(state, heat_flux=rng.normal(0.0, 40.0, N2), water_flux=rng.normal(0.0, 1e-6, N2),
                                virtual_salt=rng.normal(0.0, 1e-6, N2), relax_salt=rng.normal(0.0, 1e-6, N2),
                                real_salt_flux=rng.normal(0.0, 1e-6, N2))
    assert mesh.ulevels.max() >= 4 and np.any(mesh.ulevels_nod2D != mesh.ulevels_nod2D_max)
    return mesh, state, Oracle(mesh)


def _wet(mesh):
    k = np.arange(1, mesh.nl)
    return (k[None, :] >= mesh.ulevels_nod2D[:, None]) & (k[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])


_SCHEMES = {
    "default": dict(),
    "upw1-fct": dict(hor="UPW1", ver="UPW1", lim="FCT"),
    "mfct-qr4c-non": dict(hor="MFCT", ver="QR4C", lim="NON", ph=0.5, pv=0.0),
    "cdiff-fct": dict(hor="MFCT", ver="CDIFF", lim="FCT"),
    "ppm-fct": dict(hor="MFCT", ver="PPM", lim="FCT"),
    "muscl-ppm-non": dict(hor="MUSCL", ver="PPM", lim="NON", ph=0.5),
}


@pytest.mark.parametrize("path", ["fused", "staged"])
@pytest.mark.parametrize("sch", list(_SCHEMES), ids=list(_SCHEMES))
def test_cavity_step_bitwise(cav, monkeypatch, path, sch):
    """Three batched steps under the cavity equal the oracle's
    solve_tracers_ale bit for bit."""
    from fesom2_amd import gpu
    mesh, state, o = cav
    schemes = _SCHEMES[sch]
    if path == "staged":
        monkeypatch.setenv("FGPU_STEP", "staged")
    print(f"[gpu] cavity {sch} -> 3 {path} steps", flush=True)
    ctx = gpu.context_for(mesh, state, schemes=schemes)
    try:
        ctx.solve_tracers_ale(state.dt, 3)
        out = [ctx.tracer_download(i + 1) for i in range(state.ntr)]
    finally:
        ctx.close()
    v, vo = state.values.copy(), state.valuesold.copy()
    for _ in range(3):
        o.step(state, v, vo, schemes=schemes)
    wet = _wet(mesh)
    for i in range(state.ntr):
        np.testing.assert_array_equal(out[i][0][wet], v[i][wet], err_msg=f"tracer {i + 1}")
        np.testing.assert_array_equal(out[i][1][wet], vo[i][wet], err_msg=f"tracer {i + 1} valuesold")


@pytest.mark.parametrize("path", ["fused", "staged"])
def test_cavity_meltwater_tracer_bitwise(cav, monkeypatch, path):
    """A third tracer with ID 501 (ice-shelf water, fed by the water flux under
    the cavity only): three steps equal the oracle's bit for bit."""
    from fesom2_amd import gpu
    mesh, state, o = cav
    _, st3, _ = mg.make_config("c0cav", ntr=3)
    st3 = dataclasses.replace(st3, ids=np.array([1, 2, 501], np.int32), heat_flux=state.heat_flux,
                              water_flux=state.water_flux, virtual_salt=state.virtual_salt,
                              relax_salt=state.relax_salt, real_salt_flux=state.real_salt_flux)
    if path == "staged":
        monkeypatch.setenv("FGPU_STEP", "staged")
    print(f"[gpu] cavity meltwater tracer -> 3 {path} steps", flush=True)
    ctx = gpu.context_for(mesh, st3)
    try:
        ctx.solve_tracers_ale(st3.dt, 3)
        out = [ctx.tracer_download(i + 1)[0] for i in range(st3.ntr)]
    finally:
        ctx.close()
    v, vo = st3.values.copy(), st3.valuesold.copy()
    for _ in range(3):
        o.step(st3, v, vo)
    wet = _wet(mesh)
    for i in range(st3.ntr):
        np.testing.assert_array_equal(out[i][wet], v[i][wet], err_msg=f"tracer {i + 1}")


def test_cavity_adv_stage_bitwise(cav):
    """init_tracers_AB, do_oce_adv_tra and diff_tracers_ale one by one: the
    low-order solution, the horizontal and vertical fluxes and the tracer."""
    from fesom2_amd import gpu
    mesh, state, o = cav
    print("[gpu] cavity adv stage", flush=True)
    ctx = gpu.context_for(mesh, state)
    try:
        ctx.init_tracers_AB(1)
        ctx.do_oce_adv_tra(state.dt, 1)
        L = mesh.nl - 1
        lo = ctx.work(1, "fct_LO", (mesh.nod2D, L))
        av = ctx.work(1, "adv_flux_ver", (mesh.nod2D, mesh.nl))
        ah = ctx.work(1, "adv_flux_hor", (mesh.edge2D, L))
        ctx.diff_tracers_ale(state.dt, 1)
        gv, _ = ctx.tracer_download(1)
    finally:
        ctx.close()
    v, vo = state.values[0].copy(), state.valuesold[0].copy()
    ref = o.adv_stage(state, v, vo, ID=int(state.ids[0]))
    wet = _wet(mesh)
    np.testing.assert_array_equal(lo[wet], ref["fct_LO"][wet])
    ki = np.arange(1, mesh.nl + 1)
    wi = (ki[None, :] >= mesh.ulevels_nod2D[:, None]) & (ki[None, :] <= mesh.nlevels_nod2D[:, None])
    np.testing.assert_array_equal(av[wi], ref["adv_flux_ver"][wi])
    nu, nl = mg.edge_level_range(mesh)
    k = np.arange(1, mesh.nl)
    em = (k[None, :] >= nu[:, None]) & (k[None, :] <= nl[:, None])
    np.testing.assert_array_equal(ah[em], ref["adv_flux_hor"][em])
    o.diff_stage(state)
    np.testing.assert_array_equal(gv[wet], v[wet])


@pytest.mark.parametrize("path", ["fused", "staged"])
def test_cavity_diffusion_bitwise(cav, monkeypatch, path):
    """Ki, Redi and the explicit vertical diffusion (i_vert_diff off for the
    second tracer) under the cavity, two steps."""
    from fesom2_amd import gpu
    mesh, state, o = cav
    rng = np.random.default_rng(42)
    L = mesh.nl - 1
    S = np.zeros((mesh.nod2D, L, 3))
    S[..., 0] = rng.normal(0.0, 0.01, (mesh.nod2D, L))
    S[..., 1] = rng.normal(0.0, 0.01, (mesh.nod2D, L))
    S[..., 2] = rng.normal(0.0, 0.05, (mesh.nod2D, L))
    st2 = dataclasses.replace(state, Ki=rng.uniform(20.0, 400.0, (mesh.nod2D, L)), slope_tapered=S, redi=True)
    schemes = [dict(), dict(i_vert_diff=False)]
    if path == "staged":
        monkeypatch.setenv("FGPU_STEP", "staged")
    print(f"[gpu] cavity diffusion -> 2 {path} steps", flush=True)
    ctx = gpu.context_for(mesh, st2, schemes=schemes)
    try:
        ctx.solve_tracers_ale(state.dt, 2)
        out = [ctx.tracer_download(i + 1)[0] for i in range(state.ntr)]
    finally:
        ctx.close()
    v, vo = state.values.copy(), state.valuesold.copy()
    for _ in range(2):
        o.step(st2, v, vo, schemes=schemes)
    wet = _wet(mesh)
    for i in range(state.ntr):
        np.testing.assert_array_equal(out[i][wet], v[i][wet], err_msg=f"tracer {i + 1}")


def _vv_compare(dev, ref):
    for k in ("w", "w_e", "w_i", "cfl_z", "hnode_new"):
        np.testing.assert_array_equal(dev[k], ref[k], err_msg=k)


@pytest.mark.parametrize("which", ["linfs", "zstar", "zlevel"])
def test_cavity_vert_vel_bitwise(cav, which):
    """vert_vel_ale under the cavity: the zstar stretching and the zlevel
    surface-layer change skip the cavity columns as the reference does."""
    from fesom2_amd import gpu
    mesh, state, o = cav
    rng = np.random.default_rng(43)
    hbar_old = rng.normal(0.0, 0.3, mesh.nod2D)
    hbar = hbar_old + rng.normal(0.0, 0.05, mesh.nod2D) * mesh.hnode[np.arange(mesh.nod2D), mesh.ulevels_nod2D - 1]
    wf = state.water_flux
    code = {"linfs": 0, "zlevel": 1, "zstar": 2}[which]
    print(f"[gpu] cavity vert_vel_ale {which}", flush=True)
    ctx = gpu.context_for(mesh, state)
    try:
        ctx.set_uv(state.uv)
        if which == "linfs":
            ctx.vert_vel_ale(state.dt)
        else:
            ctx.set_ale_config(mesh.zbar, 0.5, 4)
            ctx.vert_vel_ale(state.dt, which_ale=which, hbar=hbar, hbar_old=hbar_old)
        dev = ctx.vert_vel_download()
    finally:
        ctx.close()
    if which == "linfs":
        ref = o.vert_vel_ale(state.uv, state.dt)
    else:
        ref = o.vert_vel_ale(state.uv, state.dt, which_ale=code, hbar=hbar, hbar_old=hbar_old, water_flux=wf,
                             min_hnode=0.5, lzstar_lev=4)
    _vv_compare(dev, ref)


@pytest.mark.parametrize("which", ["linfs", "zstar"])
def test_cavity_hbar_and_ssh_rhs_bitwise(cav, which):
    """compute_hbar_ale (cavity nodes keep hbar), compute_ssh_rhs_ale with and
    without the cavity freshwater pressure term, and update_stiff_mat_ale from
    the resident dhe."""
    from fesom2_amd import gpu
    mesh, state, o = cav
    rng = np.random.default_rng(44)
    h0 = rng.normal(0.0, 0.3, mesh.nod2D)
    uvr = rng.normal(0.0, 0.05, state.uv.shape)
    wf = state.water_flux
    code = {"linfs": 0, "zstar": 2}[which]
    print(f"[gpu] cavity compute_hbar_ale / compute_ssh_rhs_ale {which}", flush=True)
    ctx = gpu.context_for(mesh, state)
    try:
        ctx.set_uv(state.uv)
        ctx.init_stiff_mat_ale(state.dt)
        hb = ctx.compute_hbar_ale(state.dt, which_ale=which, hbar=h0)
        rhs = ctx.compute_ssh_rhs_ale(uvr, None, alpha=0.6, which_ale=which)
        rhs_fw = ctx.compute_ssh_rhs_ale(uvr, hb["ssh_rhs_old"], alpha=0.6, which_ale=which, use_cavity_fw2press=True)
        ctx.update_stiff_mat_ale(None, state.dt)
        _, _, dv, _ = ctx.ssh_get_matrix()
    finally:
        ctx.close()
    ref = o.compute_hbar_ale(state.uv, state.dt, h0, which_ale=code, water_flux=wf)
    for key in ("hbar", "hbar_old", "ssh_rhs_old", "dhe"):
        np.testing.assert_array_equal(hb[key], ref[key], err_msg=key)
    cav_nodes = mesh.ulevels_nod2D > 1
    np.testing.assert_array_equal(ref["hbar"][cav_nodes], h0[cav_nodes])
    np.testing.assert_array_equal(rhs, o.compute_ssh_rhs_ale(state.uv, uvr, ref["ssh_rhs_old"], alpha=0.6,
                                                             which_ale=code, water_flux=wf))
    np.testing.assert_array_equal(rhs_fw, o.compute_ssh_rhs_ale(state.uv, uvr, ref["ssh_rhs_old"], alpha=0.6,
                                                                which_ale=code, water_flux=wf,
                                                                use_cavity_fw2press=True))
    rp, ci, v = o.stiff_matrix(state.dt)
    np.testing.assert_array_equal(dv, o.update_stiff_matrix(state.dt, ref["dhe"], rp, ci, v))


def test_cavity_ssh_solve(cav):
    """Stiffness matrix and preconditioner bit for bit; CG to 1e-10 with the
    residual below tolerance and the oracle's iteration count within +-2."""
    import scipy.sparse as sp
    from fesom2_amd import gpu
    mesh, state, o = cav
    print("[gpu] cavity SSH matrix, preconditioner, CG", flush=True)
    ctx = gpu.Context(0)
    try:
        ctx.mesh_init(mesh)
        ctx.init_stiff_mat_ale(state.dt)
        rp, ci, v, _ = ctx.ssh_get_matrix()
        ctx.ssh_solve_preconditioner(2)
        _, _, _, pr = ctx.ssh_get_matrix(with_pr=True)
        b = mg.ssh_rhs(mesh, 5)
        x, st = ctx.ssh_solve_cg(np.zeros(mesh.nod2D), b, soltol=1e-10, maxiter=5000)
    finally:
        ctx.close()
    orp, oci, ov = o.stiff_matrix(state.dt)
    np.testing.assert_array_equal(rp - 1, orp)
    np.testing.assert_array_equal(ci - 1, oci)
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(pr, o.preconditioner(orp.copy(), oci.copy(), ov.copy(), 2))
    ox, ost = o.cg(orp, oci, ov, o.preconditioner(orp, oci, ov, 2), b, soltol=1e-10, maxiter=5000)
    A = sp.csr_matrix((ov.copy(), oci.copy(), orp.copy()), shape=(mesh.nod2D, mesh.nod2D))
    assert st["converged"] and ost["converged"]
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10 * 1.01
    assert abs(st["iters"] - ost["iters"]) <= 2


@pytest.mark.parametrize("path", ["fused", "staged"])
@pytest.mark.parametrize("depth", [False, True], ids=["hosing_flux", "hosing_flux3D"])
def test_cavity_forcing_tracers_bitwise(cav, monkeypatch, path, depth):
    """Tracers with the IDs 101 (prec_rain), 102-104 (wiso_flux_oce) and 304
    (hosing, per node or at the first wet layer of hosing_flux3D) next to
    salinity, with the flow on: three steps equal the oracle's bit for bit."""
    from fesom2_amd import gpu
    mesh, state, o = cav
    rng = np.random.default_rng(43)
    N2, L = mesh.nod2D, mesh.nl - 1
    _, st6, _ = mg.make_config("c0cav", ntr=6)
    st6 = dataclasses.replace(
        st6, ids=np.array([101, 102, 103, 104, 304, 2], np.int32), heat_flux=state.heat_flux,
        water_flux=state.water_flux, virtual_salt=state.virtual_salt, relax_salt=state.relax_salt,
        real_salt_flux=state.real_salt_flux, prec_rain=rng.normal(0.0, 1e-5, N2),
        wiso_flux_oce=rng.normal(0.0, 1e-5, (3, N2)), hosing_flux=rng.normal(0.0, 1e-5, N2),
        hosing_flux3D=rng.normal(0.0, 1e-5, (N2, L)), hosing_depth=depth)
    if path == "staged":
        monkeypatch.setenv("FGPU_STEP", "staged")
    print(f"[gpu] cavity IDs 101-104/304 depth={depth} -> 3 {path} steps", flush=True)
    ctx = gpu.context_for(mesh, st6)
    try:
        ctx.solve_tracers_ale(st6.dt, 3)
        out = [ctx.tracer_download(i + 1)[0] for i in range(st6.ntr)]
    finally:
        ctx.close()
    v, vo = st6.values.copy(), st6.valuesold.copy()
    for _ in range(3):
        o.step(st6, v, vo)
    wet = _wet(mesh)
    for i in range(st6.ntr):
        np.testing.assert_array_equal(out[i][wet], v[i][wet], err_msg=f"tracer {i + 1}")


def test_forcing_tracer_without_flux_is_refused(cav):
    """A tracer with ID 101 whose prec_rain was never set fails the step with
    FGPU_ERR_STATE before any launch."""
    from fesom2_amd import gpu
    mesh, state, _ = cav
    _, st1, _ = mg.make_config("c0cav", ntr=2)
    st1 = dataclasses.replace(st1, ids=np.array([1, 101], np.int32))
    ctx = gpu.context_for(mesh, st1)
    try:
        with pytest.raises(gpu.FesomGpuError) as e:
            ctx.solve_tracers_ale(st1.dt, 1)
        assert e.value.code == -5 and "prec_rain" in str(e.value)
    finally:
        ctx.close()
