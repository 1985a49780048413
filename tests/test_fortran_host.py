"""Fortran host side (fesom2_amd/fortran): the ISO_C_BINDING modules and the case
driver built with flang, linked against libfesom_gpu.so.

CPU: the driver resolves the library and reads a case file written by
tests/fortran_case.py with the sizes and sums numpy sees. GPU: the driver runs
the reference-shaped per-tracer loop (init_tracers_AB, do_oce_adv_tra,
diff_tracers_ale) for one step and the batched solve_tracers_ale for two more;
the tracers must equal three oracle steps bit for bit, and solve_ssh_ale must
reach the 1e-10 residual.
"""
import os
import subprocess

import numpy as np
import pytest

from fesom2_amd import build as fbuild
from fesom2_amd import meshgen as mg
from tests import fortran_case as fc

DRIVER = fbuild.FORTRAN_DRIVER


def _need_driver():
    if not os.path.exists(DRIVER):
        if fbuild.flang() is None:
            pytest.skip("no Fortran compiler in this image: the driver is not built")
        pytest.fail("Fortran driver not built: run __graft_entry__.build()")


def test_driver_links_library():
    _need_driver()
    r = subprocess.run([DRIVER, "--abi"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "fesom_gpu library version 1" in r.stdout


def test_case_file_read_by_fortran(c0, tmp_path):
    _need_driver()
    mesh, state, _ = c0
    rhs = mg.ssh_rhs(mesh)
    path = str(tmp_path / "case.bin")
    fc.write_case(path, mesh, state, rhs)
    r = subprocess.run([DRIVER, "--read", path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    si = [int(x) for x in out["sums_int"]]
    assert si[0] == int((mesh.edges + 1).sum())
    assert si[1] == int(np.where(mesh.nod_in_elem2D >= 0, mesh.nod_in_elem2D + 1, 0).sum())
    assert si[2] == int(state.ids.sum())
    sr = [float(x) for x in out["sums_real"]]
    np.testing.assert_allclose(sr, [state.values.sum(), state.uv.sum(), rhs.sum()], rtol=1e-12)
    last = [float(x) for x in out["last"]]
    assert last == [float(mesh.gradient_sca[-1, 5]), float(state.Kv[-1, -1])]


def _diffusion_state(c0_oracle, state, mesh):
    """The c0 state with W split by the oracle's vert_vel_ale, Ki, and Redi slopes."""
    import dataclasses
    cap = 0.3 * c0_oracle.vert_vel_ale(state.uv, state.dt)["cflmax"]
    ref = c0_oracle.vert_vel_ale(state.uv, state.dt, use_wsplit=True, wsplit_maxcfl=cap)
    rng = np.random.default_rng(41)
    L = mesh.nl - 1
    S = rng.normal(0.0, 0.01, (mesh.nod2D, L, 3))
    return dataclasses.replace(state, w=ref["w"], w_e=ref["w_e"], w_i=ref["w_i"], use_wsplit=True,
                               Ki=rng.uniform(20.0, 400.0, (mesh.nod2D, L)), slope_tapered=S, redi=True)


def test_case_file_with_diffusion_read_by_fortran(c0, c0_oracle, tmp_path):
    _need_driver()
    mesh, state, _ = c0
    st2 = _diffusion_state(c0_oracle, state, mesh)
    path = str(tmp_path / "case.bin")
    fc.write_case(path, mesh, st2, mg.ssh_rhs(mesh))
    r = subprocess.run([DRIVER, "--read", path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    assert [int(x) for x in out["flags"]] == [1, 1, 1, 2]
    np.testing.assert_allclose(float(out["ki"][0]), st2.Ki.sum(), rtol=1e-12)
    assert float(out["slope"][0]) == float(st2.slope_tapered[-1, -1, 2])


def _older(mesh, state):
    k = np.arange(1, mesh.nl)
    wet = (k[None, :] >= mesh.ulevels_nod2D[:, None]) & (k[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])
    rng = np.random.default_rng(52)
    return wet, np.where(wet, state.valuesold[0] + rng.normal(0.0, 0.05, state.valuesold[0].shape), 0.0)


def test_case_file_with_ab3_read_by_fortran(c0, tmp_path):
    _need_driver()
    mesh, state, _ = c0
    _, older = _older(mesh, state)
    path = str(tmp_path / "case.bin")
    fc.write_case(path, mesh, state, mg.ssh_rhs(mesh), older=older)
    r = subprocess.run([DRIVER, "--read", path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    assert [int(x) for x in out["flags"]] == [0, 0, 0, 3]


@pytest.mark.gpu
def test_fortran_host_ab3(c0, c0_oracle, tmp_path):
    """Tracer 1 runs AB3 through the Fortran host: gpu_tracer_put/get with the
    reference's valuesold(2, nl-1, nod2D), the per-tracer loop and the batched
    step; both tracers and both old levels equal three oracle steps."""
    _need_driver()
    mesh, state, _ = c0
    wet, older = _older(mesh, state)
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "result.bin")
    fc.write_case(case, mesh, state, mg.ssh_rhs(mesh), nsingle=1, nbatched=2, older=older)
    print("[gpu] fortran driver c0, tracer 1 AB3: 1 per-tracer + 2 batched steps", flush=True)
    r = subprocess.run([DRIVER, case, res], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    out = fc.read_result(res, mesh, state.ntr, ab3=True)
    v = state.values.copy()
    vos = [np.ascontiguousarray(np.stack([state.valuesold[0], older], axis=-1)), state.valuesold[1].copy()]
    for _ in range(3):
        c0_oracle.step(state, v, vos, schemes=[dict(AB_order=3), dict()])
    for i in range(state.ntr):
        np.testing.assert_array_equal(out["values"][i][wet], v[i][wet], err_msg=f"tracer {i + 1}")
    np.testing.assert_array_equal(out["valuesold"][0][wet], vos[0][..., 0][wet])
    np.testing.assert_array_equal(out["older"][wet], vos[0][..., 1][wet])
    np.testing.assert_array_equal(out["valuesold"][1][wet], vos[1][wet])


@pytest.mark.gpu
def test_fortran_host_diffusion_wsplit(c0, c0_oracle, tmp_path):
    """The driver with Ki, Redi and a split W (gpu_set_ki, use_wsplit in
    gpu_set_state and do_oce_adv_tra): 1 per-tracer + 2 batched steps equal
    three oracle steps bit for bit."""
    _need_driver()
    mesh, state, _ = c0
    st2 = _diffusion_state(c0_oracle, state, mesh)
    rhs = mg.ssh_rhs(mesh)
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "result.bin")
    fc.write_case(case, mesh, st2, rhs, nsingle=1, nbatched=2, precond_variant=1)
    print("[gpu] fortran driver c0 with Ki, Redi, W split", flush=True)
    r = subprocess.run([DRIVER, case, res], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    out = fc.read_result(res, mesh, state.ntr)
    v = state.values.copy()
    vo = state.valuesold.copy()
    for _ in range(3):
        c0_oracle.step(st2, v, vo)
    L = mesh.nl - 1
    k = np.arange(1, L + 1)
    wet = (k[None, :] >= mesh.ulevels_nod2D[:, None]) & (k[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])
    for i in range(state.ntr):
        np.testing.assert_array_equal(out["values"][i][wet], v[i][wet])
        np.testing.assert_array_equal(out["valuesold"][i][wet], vo[i][wet])


@pytest.mark.gpu
def test_fortran_host_c0(c0, c0_oracle, tmp_path):
    _need_driver()
    import scipy.sparse as sp
    mesh, state, _ = c0
    rhs = mg.ssh_rhs(mesh)
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "result.bin")
    fc.write_case(case, mesh, state, rhs, nsingle=1, nbatched=2, precond_variant=1)
    print("[gpu] fortran driver c0: 1 per-tracer + 2 batched steps, vert_vel_ale, compute_ssh_rhs_ale, "
          "solve_ssh_ale", flush=True)
    r = subprocess.run([DRIVER, case, res], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    out = fc.read_result(res, mesh, state.ntr)
    v = state.values.copy()
    vo = state.valuesold.copy()
    for _ in range(3):
        c0_oracle.step(state, v, vo)
    L = mesh.nl - 1
    k = np.arange(1, L + 1)
    wet = (k[None, :] >= mesh.ulevels_nod2D[:, None]) & (k[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])
    for i in range(state.ntr):
        np.testing.assert_array_equal(out["values"][i][wet], v[i][wet])
        np.testing.assert_array_equal(out["valuesold"][i][wet], vo[i][wet])
        assert abs(out["content_after"][i] - out["content_before"][i]) < 1e-13 * abs(out["content_before"][i])
    assert out["converged"] and out["nbreakdown"] == 0 and out["nnegrz"] == 0
    # vert_vel_ale and compute_ssh_rhs_ale through the Fortran host routines
    ref = c0_oracle.vert_vel_ale(state.uv, state.dt)
    for key in ("w", "w_e", "w_i", "cfl_z"):
        np.testing.assert_array_equal(out["dyn"][key], ref[key], err_msg=key)
    np.testing.assert_array_equal(out["dyn"]["ssh_rhs"], c0_oracle.compute_ssh_rhs_ale(state.uv, state.uv, rhs,
                                                                                        alpha=0.5))
    orp, oci, ov = c0_oracle.stiff_matrix(state.dt)
    A = sp.csr_matrix((ov.copy(), oci.copy(), orp.copy()), shape=(mesh.nod2D, mesh.nod2D))
    assert np.linalg.norm(rhs - A @ out["d_eta"]) / np.linalg.norm(rhs) < 1e-10 * 1.01


def test_fortran_interfaces_cover_the_header():
    """Every entry point of include/fesom_gpu.h has a bind(C) interface."""
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "fesom_gpu.h")).read()
    decl = set(re.findall(r"\b(fgpu_[A-Za-z0-9_]+)\s*\(", hdr))
    f90 = open(os.path.join(root, "fesom2_amd", "fortran", "fesom_gpu_c.f90")).read()
    bound = set(re.findall(r'bind\(C, name="(fgpu_[A-Za-z0-9_]+)"\)', f90))
    assert decl == bound, (sorted(decl - bound), sorted(bound - decl))
