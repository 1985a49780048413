"""Case files of the Fortran host driver (fesom2_amd/fortran/fgpu_case_driver.f90).

write_case() serialises a meshgen Mesh/State in the driver's stream layout
(Fortran array order = the C order of the numpy arrays, 1-based connectivity);
read_result() parses what the driver writes back.
"""
from __future__ import annotations

import numpy as np

MAGIC = 1179076949


def write_case(path, mesh, state, ssh_rhs, nsingle=1, nbatched=2, precond_variant=1, maxiter=2000,
               soltol=1e-10, g=9.81, older=None):
    """older (nod2D, nl-1): tracer 1 then runs AB3 with old levels
    state.valuesold[0] (newer) and older."""
    i32 = lambda a: np.ascontiguousarray(a, dtype="<i4")
    f64 = lambda a: np.ascontiguousarray(a, dtype="<f8")
    hdr = np.zeros(16, dtype="<i4")
    hdr[:12] = [MAGIC, 1, mesh.nod2D, mesh.elem2D, mesh.edge2D, mesh.nl, mesh.kmax, state.ntr,
                nsingle, nbatched, precond_variant, maxiter]
    ki = getattr(state, "Ki", None)
    redi = bool(getattr(state, "redi", False))
    hdr[12:16] = [ki is not None, redi, bool(getattr(state, "use_wsplit", False)), 3 if older is not None else 2]
    parts = [hdr, f64([state.dt, mesh.cyclic_length, soltol, g]),
             f64(mesh.coord_nod2D), i32(mesh.elem2D_nodes + 1), i32(mesh.edges + 1),
             i32(np.where(mesh.edge_tri >= 0, mesh.edge_tri + 1, -999)),
             i32(np.where(mesh.nod_in_elem2D >= 0, mesh.nod_in_elem2D + 1, 0)), i32(mesh.nod_in_elem2D_num),
             i32(mesh.nlevels), i32(mesh.ulevels), i32(mesh.nlevels_nod2D), i32(mesh.ulevels_nod2D),
             i32(mesh.nlevels_nod2D_min), i32(mesh.ulevels_nod2D_max),
             f64(mesh.elem_area), f64(mesh.elem_cos), f64(mesh.edge_dxdy), f64(mesh.edge_cross_dxdy),
             f64(mesh.gradient_sca),
             f64(mesh.area), f64(mesh.areasvol), f64(mesh.hnode), f64(mesh.hnode_new), f64(mesh.helem),
             f64(mesh.zbar_3d_n), f64(mesh.Z_3d_n), f64(mesh.zbar_n_bot), f64(mesh.zbar_e_bot),
             f64(mesh.zbar_e_srf),
             f64(state.uv), f64(state.w), f64(state.w_e), f64(state.w_i), f64(state.Kv),
             f64(state.heat_flux), f64(state.water_flux), f64(state.virtual_salt), f64(state.relax_salt),
             f64(state.real_salt_flux),
             i32(state.ids), f64(state.values), f64(state.valuesold), f64(ssh_rhs)]
    if ki is not None:
        parts.append(f64(ki))
    if redi:
        parts.append(f64(state.slope_tapered))
    if older is not None:
        parts.append(f64(older))
    with open(path, "wb") as f:
        for p in parts:
            f.write(p.tobytes())
    return sum(p.nbytes for p in parts)


def read_result(path, mesh, ntr, ab3=False):
    L = mesh.nl - 1
    raw = open(path, "rb").read()
    st = np.frombuffer(raw, dtype="<i4", count=8)
    off = 32
    r = np.frombuffer(raw, dtype="<f8", offset=off)
    resid, rtol = r[0], r[1]
    c0 = r[2:2 + ntr]
    c1 = r[2 + ntr:2 + 2 * ntr]
    o = 2 + 2 * ntr
    n3 = ntr * mesh.nod2D * L
    values = r[o:o + n3].reshape(ntr, mesh.nod2D, L)
    valuesold = r[o + n3:o + 2 * n3].reshape(ntr, mesh.nod2D, L)
    d_eta = r[o + 2 * n3:o + 2 * n3 + mesh.nod2D]
    assert d_eta.size == mesh.nod2D, "truncated result file"
    o = o + 2 * n3 + mesh.nod2D
    ni = mesh.nod2D * mesh.nl
    dyn = {}
    for k in ("w", "w_e", "w_i", "cfl_z"):
        dyn[k] = r[o:o + ni].reshape(mesh.nod2D, mesh.nl).copy()
        o += ni
    dyn["ssh_rhs"] = r[o:o + mesh.nod2D].copy()
    assert dyn["ssh_rhs"].size == mesh.nod2D, "truncated result file"
    o += mesh.nod2D
    older = None
    if ab3:
        older = r[o:o + mesh.nod2D * L].reshape(mesh.nod2D, L).copy()
        assert older.size == mesh.nod2D * L, "truncated result file"
    return dict(older=older, iters=int(st[0]), converged=bool(st[1]), nbreakdown=int(st[2]), nnegrz=int(st[3]),
                nonconv=int(st[4]), resid=float(resid), rtol=float(rtol), content_before=c0.copy(),
                content_after=c1.copy(), values=values.copy(), valuesold=valuesold.copy(), d_eta=d_eta.copy(),
                dyn=dyn)
