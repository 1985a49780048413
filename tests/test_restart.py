"""Derived-type binary restarts (fesom2_amd/restart.py) against the
reference's own restart I/O.

oracle/_ref/ref_restart_io is built with flang from the reference's
MOD_MESH.F90, MOD_TRACER.F90 and MOD_DYN.F90 (and their binary-array helpers):
it writes and reads t_mesh, t_tracer and t_dynamics records through the
types' own derived-type I/O, as write_all_bin_restarts and
read_all_bin_restarts do. Checked here: the files this package writes are
byte-identical to the reference's; each side reads the other's files back to
the same arrays; a state loaded from restart files steps exactly like the
state it was written from. The tests that need oracle/_ref skip where the
reference tree (and so that build) is absent.
"""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

from fesom2_amd import meshgen as mg
from fesom2_amd import restart as rs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_restart_io")


def _need_ref():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref/ref_restart_io is not built (the reference tree is absent here)")


@pytest.fixture(scope="module")
def case():
    from oracle.oracle import Oracle
    mesh, st, _ = mg.make_config("c0cav", ntr=3)
    rng = np.random.default_rng(17)
    st = dataclasses.replace(st, w_i=rng.normal(0.0, 1e-6, st.w.shape), use_wsplit=True)
    rp, ci, val = Oracle(mesh).stiff_matrix(st.dt)
    cfl = rng.uniform(0.0, 1.0, st.w.shape)
    return mesh, st, (rp, ci, val), cfl


def _fields(mesh, st, ssh, cfl, path, schemes, maxiter=3000, soltol=1e-10, wsplit_maxcfl=0.7):
    """The input stream of ref_restart_io (see its header), Fortran array order."""
    i32 = lambda a: np.ascontiguousarray(a, "<i4").tobytes()
    f64 = lambda a: np.ascontiguousarray(a, "<f8").tobytes()
    rp, ci, val = ssh
    parts = [i32([mesh.nod2D, mesh.elem2D, mesh.edge2D, mesh.nl, mesh.kmax, len(ci), st.ntr, mesh.edge2D_in]),
             f64([mesh.areasvol[:, 0].sum(), mesh.area[:, 0].sum()])]
    rec = rs.mesh_record(mesh, ssh)
    for k in ("elem2D_nodes", "edges", "edge_tri"):
        parts.append(i32(rec[k]))
    parts += [f64(rec["elem_area"]), f64(rec["edge_dxdy"]), f64(rec["edge_cross_dxdy"]), f64(rec["elem_cos"]),
              i32(rec["nod_in_elem2D"]), i32(rec["nod_in_elem2D_num"]), f64(rec["depth"]), f64(rec["gradient_sca"]),
              f64(rec["zbar"]), f64(rec["Z"])]
    for k in ("ulevels", "ulevels_nod2D", "ulevels_nod2D_max", "nlevels", "nlevels_nod2D", "nlevels_nod2D_min"):
        parts.append(i32(rec[k]))
    parts += [f64(rec["area"]), f64(rec["areasvol"]), i32(rec["ssh_stiff%rowptr"]), i32(rec["ssh_stiff%colind"]),
              f64(rec["ssh_stiff%values"])]
    for k in ("hnode", "hnode_new", "zbar_3d_n", "Z_3d_n", "helem", "zbar_n_bot", "zbar_e_bot", "zbar_e_srf"):
        parts.append(f64(rec[k]))
    for i in range(st.ntr):
        s = schemes[i]
        parts.append(i32([st.ids[i], 1]))
        parts.append(f64([s["ph"], s["pv"]]))
        parts.append("".join(x.ljust(20) for x in (s["hor"], s["ver"], s["lim"])).encode())
        parts += [f64(st.values[i]), f64(st.valuesold[i]), f64(np.zeros_like(st.values[i]))]
    parts += [f64(st.uv), f64(st.w), f64(st.w_e), f64(st.w_i), f64(cfl)]
    parts += [i32([1 if st.use_wsplit else 0, maxiter]), f64([wsplit_maxcfl, soltol])]
    with open(path, "wb") as f:
        for p in parts:
            f.write(p)


SCHEMES = [dict(hor="MFCT", ver="QR4C", lim="FCT", ph=0.0, pv=1.0), dict(hor="UPW1", ver="PPM", lim="NON", ph=1.0,
                                                                         pv=0.5),
           dict(hor="MUSCL", ver="CDIFF", lim="FCT", ph=0.25, pv=1.0)]


def _records(mesh, st, ssh, cfl):
    """The records holding exactly what ref_restart_io fills in."""
    full = rs.mesh_record(mesh, ssh)
    keep = ["nod2D", "edge2D", "edge2D_in", "elem2D", "nl", "ocean_area", "ocean_areawithcav", "elem2D_nodes",
            "edges", "edge_tri", "elem_area", "edge_dxdy", "edge_cross_dxdy", "elem_cos", "nod_in_elem2D",
            "nod_in_elem2D_num", "depth", "gradient_sca", "zbar", "Z", "ulevels", "ulevels_nod2D",
            "ulevels_nod2D_max", "nlevels", "nlevels_nod2D", "nlevels_nod2D_min", "area", "areasvol", "hnode",
            "hnode_new", "zbar_3d_n", "Z_3d_n", "helem", "zbar_n_bot", "zbar_e_bot", "zbar_e_srf"]
    mrec = {k: full[k] for k in keep}
    mrec.update({k: full[k] for k in full if k.startswith("ssh_stiff%")})
    trec = rs.tracer_record(st, SCHEMES)
    drec = rs.dyn_record(st)
    drec.update(cfl_z=cfl, wsplit_maxcfl=0.7, solverinfo=dict(maxiter=3000, soltol=1e-10))
    return mrec, trec, drec


def test_written_files_equal_the_references(case, tmp_path):
    _need_ref()
    mesh, st, ssh, cfl = case
    fields = str(tmp_path / "fields.bin")
    _fields(mesh, st, ssh, cfl, fields, SCHEMES)
    dref = tmp_path / "ref"
    dref.mkdir()
    r = subprocess.run([REF, "write", fields, str(dref)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    mrec, trec, drec = _records(mesh, st, ssh, cfl)
    mine = tmp_path / "mine"
    rs.write_restart(str(mine), mesh_rec=mrec, tracer_rec=trec, dyn_rec=drec)
    for name in ("t_mesh.0", "t_tracer.0", "t_dynamics.0"):
        a = (dref / name).read_bytes()
        b = (mine / name).read_bytes()
        assert len(a) == len(b), name
        assert a == b, name


def test_reads_the_references_files(case, tmp_path):
    _need_ref()
    mesh, st, ssh, cfl = case
    fields = str(tmp_path / "fields.bin")
    _fields(mesh, st, ssh, cfl, fields, SCHEMES)
    r = subprocess.run([REF, "write", fields, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = rs.read_restart(str(tmp_path))
    m, t, d = got["mesh"], got["tracers"], got["dynamics"]
    exp = rs.mesh_record(mesh, ssh)
    for k, v in exp.items():
        if k in ("elem_neighbors",):
            continue
        np.testing.assert_array_equal(np.asarray(m[k]), np.asarray(v), err_msg=k)
    assert m["elem_edges"] is None and m["coriolis"] is None
    assert t["num_tracers"] == st.ntr
    for i, dd in enumerate(t["data"]):
        np.testing.assert_array_equal(dd["values"], st.values[i])
        np.testing.assert_array_equal(dd["valuesold"][:, :, 0], st.valuesold[i])
        assert dd["ID"] == st.ids[i] and dd["tra_adv_hor"].strip() == SCHEMES[i]["hor"]
        assert dd["tra_adv_ver"].strip() == SCHEMES[i]["ver"] and dd["tra_adv_ph"] == SCHEMES[i]["ph"]
        assert dd["i_vert_diff"] is True
    for k in ("uv", "w", "w_e", "w_i"):
        np.testing.assert_array_equal(d[k], getattr(st, k))
    np.testing.assert_array_equal(d["cfl_z"], cfl)
    assert d["use_wsplit"] and d["wsplit_maxcfl"] == 0.7 and not d["Fer_GM"]
    assert d["solverinfo"]["maxiter"] == 3000 and d["solverinfo"]["soltol"] == 1e-10
    assert d["opt_visc"] == 5 and d["visc_gamma0"] == float(np.float32(0.03))


def test_reference_reads_our_files(case, tmp_path):
    _need_ref()
    mesh, st, ssh, cfl = case
    mrec, trec, drec = _records(mesh, st, ssh, cfl)
    rs.write_restart(str(tmp_path), mesh_rec=mrec, tracer_rec=trec, dyn_rec=drec)
    back = str(tmp_path / "back.bin")
    r = subprocess.run([REF, "read", str(tmp_path), back], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    exp = str(tmp_path / "exp.bin")
    _fields(mesh, st, ssh, cfl, exp, SCHEMES)
    assert open(back, "rb").read() == open(exp, "rb").read()


def test_restart_round_trip_steps_alike(case, tmp_path):
    """mesh_from_restart / state_from_restart: the oracle's step on the state
    read back equals its step on the original, bit for bit."""
    from oracle.oracle import Oracle
    mesh, st, ssh, cfl = case
    st = dataclasses.replace(st, w_i=np.zeros_like(st.w_i), use_wsplit=False)
    rs.write_restart(str(tmp_path), mesh_rec=rs.mesh_record(mesh, ssh), tracer_rec=rs.tracer_record(st),
                     dyn_rec=rs.dyn_record(st))
    got = rs.read_restart(str(tmp_path))
    mesh2 = rs.mesh_from_restart(got["mesh"], mesh.coord_nod2D, mesh.cyclic_length, mesh.cartesian)
    forcing = dict(heat_flux=st.heat_flux, water_flux=st.water_flux, virtual_salt=st.virtual_salt,
                   relax_salt=st.relax_salt, real_salt_flux=st.real_salt_flux)
    st2, schemes = rs.state_from_restart(mesh2, got["tracers"], got["dynamics"], st.Kv, st.dt, forcing)
    assert schemes[0]["hor"] == "MFCT" and schemes[0]["lim"] == "FCT"
    a, ao = st.values.copy(), st.valuesold.copy()
    b, bo = st2.values.copy(), st2.valuesold.copy()
    for _ in range(2):
        Oracle(mesh).step(st, a, ao)
        Oracle(mesh2).step(st2, b, bo)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ao, bo)


def test_record_framing(tmp_path):
    """One record, 4-byte markers; a payload beyond one subrecord splits into
    subrecords with negative markers where more follow, and reads back whole."""
    p = str(tmp_path / "r.bin")
    rs.write_record(p, b"abcdef")
    raw = open(p, "rb").read()
    assert raw[:4] == (6).to_bytes(4, "little") and raw[-4:] == (6).to_bytes(4, "little")
    assert rs.read_record(p) == b"abcdef"
    old = rs._MAX_SUB
    rs._MAX_SUB = 4
    try:
        rs.write_record(p, b"0123456789")
    finally:
        rs._MAX_SUB = old
    raw = open(p, "rb").read()
    assert int.from_bytes(raw[:4], "little", signed=True) == -4
    assert rs.read_record(p) == b"0123456789"
    with open(p, "ab") as f:
        f.write(b"x")
    with pytest.raises(ValueError):
        rs.read_record(p)


@pytest.mark.gpu
def test_gpu_step_from_restart(case, tmp_path):
    """The fused step on a mesh and state read from restart files equals the
    oracle's on the originals, bit for bit."""
    from fesom2_amd import gpu
    from oracle.oracle import Oracle
    mesh, st, ssh, cfl = case
    st = dataclasses.replace(st, w_i=np.zeros_like(st.w_i), use_wsplit=False)
    rs.write_restart(str(tmp_path), mesh_rec=rs.mesh_record(mesh, ssh), tracer_rec=rs.tracer_record(st),
                     dyn_rec=rs.dyn_record(st))
    got = rs.read_restart(str(tmp_path))
    mesh2 = rs.mesh_from_restart(got["mesh"], mesh.coord_nod2D, mesh.cyclic_length, mesh.cartesian)
    forcing = dict(heat_flux=st.heat_flux, water_flux=st.water_flux, virtual_salt=st.virtual_salt,
                   relax_salt=st.relax_salt, real_salt_flux=st.real_salt_flux)
    st2, _ = rs.state_from_restart(mesh2, got["tracers"], got["dynamics"], st.Kv, st.dt, forcing)
    print("[gpu] step from restart files", flush=True)
    ctx = gpu.context_for(mesh2, st2)
    try:
        ctx.solve_tracers_ale(st2.dt, 2)
        out = [ctx.tracer_download(i + 1)[0] for i in range(st2.ntr)]
    finally:
        ctx.close()
    v, vo = st.values.copy(), st.valuesold.copy()
    o = Oracle(mesh)
    for _ in range(2):
        o.step(st, v, vo)
    k = np.arange(1, mesh.nl)
    wet = (k[None, :] >= mesh.ulevels_nod2D[:, None]) & (k[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])
    for i in range(st.ntr):
        np.testing.assert_array_equal(out[i][wet], v[i][wet])
