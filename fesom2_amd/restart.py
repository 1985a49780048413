"""FESOM2's derived-type binary restarts of t_mesh, t_tracer and t_dyn.

write_all_bin_restarts / read_all_bin_restarts (io_restart_derivedtype.F90:
37-147, 153-234) store each derived type of a rank as one unformatted
sequential Fortran record, ``t_mesh.<rank>``, ``t_tracer.<rank>``,
``t_dynamics.<rank>``, through the types' user-defined derived-type I/O. A
child data transfer starts no new record, so the whole type is one record
whose payload is the sequence of items the type's write procedure lists:

* scalars as they are (int32, real64, logical as int32 0/1, character(n));
* every allocatable array through write_bin_array
  (MOD_WRITE_BINARY_ARRAYS.F90): its extents as int32, then its elements in
  Fortran order; an unallocated array is its extents, all zero.

The item order is that of WRITE_T_MESH (MOD_MESH.F90:179-274),
WRITE_T_TRACER(_DATA, _WORK) (MOD_TRACER.F90:112-228) and
WRITE_T_DYN(_WORK) / WRITE_T_SOLVERINFO (MOD_DYN.F90:212-346); the tables
below follow them item by item, since the files are only readable in that
order. The record is framed by 4-byte length markers; records of 2 GiB and
more are split into subrecords whose leading markers are negative when more
follow, the convention gfortran uses.

Arrays come back as numpy arrays in C order with the reversed shape, which is
the Fortran array's memory image: a Fortran (nl-1, nod2D) array is numpy
(nod2D, nl-1), as everywhere in this package.

Coordinates are not part of the t_mesh record; mesh_from_restart takes them
from the mesh files (meshio) or the caller.
"""
from __future__ import annotations

import os
import struct
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import meshgen as mg

# item kinds: i4 / f8 / l4 scalars, c20 character(20), i1..i4 / r1..r4 arrays
MESH_ITEMS: List[Tuple[str, str]] = [
    ("nod2D", "i4"), ("ocean_area", "f8"), ("ocean_areawithcav", "f8"), ("edge2D", "i4"), ("edge2D_in", "i4"),
    ("elem2D", "i4"), ("elem2D_nodes", "i2"), ("edges", "i2"), ("edge_tri", "i2"), ("elem_edges", "i2"),
    ("elem_area", "r1"), ("edge_dxdy", "r2"), ("edge_cross_dxdy", "r2"), ("elem_cos", "r1"),
    ("metric_factor", "r1"), ("elem_neighbors", "i2"), ("nod_in_elem2D", "i2"), ("x_corners", "r2"),
    ("y_corners", "r2"), ("nod_in_elem2D_num", "i1"), ("depth", "r1"), ("gradient_vec", "r2"),
    ("gradient_sca", "r2"), ("bc_index_nod2D", "i1"), ("nl", "i4"), ("zbar", "r1"), ("Z", "r1"),
    ("elem_depth", "r1"), ("ulevels", "i1"), ("ulevels_nod2D", "i1"), ("ulevels_nod2D_max", "i1"),
    ("nlevels", "i1"), ("nlevels_nod2D", "i1"), ("nlevels_nod2D_min", "i1"), ("area", "r2"), ("area_inv", "r2"),
    ("areasvol", "r2"), ("areasvol_inv", "r2"), ("mesh_resolution", "r1"), ("cavity_flag_n", "i1"),
    ("cavity_flag_e", "i1"), ("cavity_depth", "r1"), ("cavity_nrst_cavlpnt_xyz", "r2"),
    ("ssh_stiff%dim", "i4"), ("ssh_stiff%nza", "i4"), ("ssh_stiff%rowptr", "i1"), ("ssh_stiff%colind", "i1"),
    ("ssh_stiff%values", "r1"), ("ssh_stiff%colind_loc", "i1"), ("ssh_stiff%rowptr_loc", "i1"),
    ("lump2d_south", "r1"), ("lump2d_north", "r1"), ("ind_south", "i1"), ("ind_north", "i1"), ("nn_size", "i4"),
    ("nn_num", "i1"), ("nn_pos", "i2"), ("hnode", "r2"), ("hnode_new", "r2"), ("zbar_3d_n", "r2"),
    ("Z_3d_n", "r2"), ("Z_3d_n_ib", "r2"), ("helem", "r2"), ("bottom_elem_thickness", "r1"),
    ("bottom_node_thickness", "r1"), ("dhe", "r1"), ("hbar", "r1"), ("hbar_old", "r1"), ("zbar_n_bot", "r1"),
    ("zbar_e_bot", "r1"), ("zbar_n_srf", "r1"), ("zbar_e_srf", "r1"), ("coriolis", "r1"), ("coriolis_node", "r1"),
]

TRACER_DATA_ITEMS = [
    ("values", "r2"), ("valuesold", "r3"), ("valuesAB", "r2"), ("smooth_bh_tra", "l4"), ("gamma0_tra", "f8"),
    ("gamma1_tra", "f8"), ("gamma2_tra", "f8"), ("i_vert_diff", "l4"), ("tra_adv_hor", "c20"),
    ("tra_adv_ver", "c20"), ("tra_adv_lim", "c20"), ("tra_adv_ph", "f8"), ("tra_adv_pv", "f8"), ("ID", "i4"),
]
TRACER_WORK_ITEMS = [
    ("del_ttf", "r2"), ("del_ttf_advhoriz", "r2"), ("del_ttf_advvert", "r2"), ("dvd_trflx_hor", "r3"),
    ("dvd_trflx_ver", "r3"), ("fct_LO", "r2"), ("adv_flux_hor", "r2"), ("adv_flux_ver", "r2"),
    ("fct_ttf_max", "r2"), ("fct_ttf_min", "r2"), ("fct_plus", "r2"), ("fct_minus", "r2"),
    ("nboundary_lay", "i1"), ("edge_up_dn_tri", "i2"), ("edge_up_dn_grad", "r3"),
]

DYN_HEAD_ITEMS = [
    ("opt_visc", "i4"), ("visc_gamma0", "f8"), ("visc_gamma1", "f8"), ("visc_gamma2", "f8"),
    ("visc_easybsreturn", "f8"), ("use_ivertvisc", "l4"), ("momadv_opt", "i4"), ("use_freeslip", "l4"),
    ("use_wsplit", "l4"), ("wsplit_maxcfl", "f8"), ("use_ssh_se_subcycl", "l4"),
]
SOLVERINFO_ITEMS = [
    ("ident", "i4"), ("maxiter", "i4"), ("restart", "i4"), ("fillin", "i4"), ("lutype", "i4"), ("droptol", "f8"),
    ("soltol", "f8"), ("rr", "r1"), ("zz", "r1"), ("pp", "r1"), ("App", "r1"),
]
DYN_WORK_ITEMS = [("uvnode_rhs", "r3"), ("u_c", "r2"), ("v_c", "r2"), ("u_b", "r2"), ("v_b", "r2")]
DYN_ARRAY_ITEMS = [
    ("uv", "r3"), ("uv_rhs", "r3"), ("uv_rhsAB", "r4"), ("uvnode", "r3"), ("w", "r2"), ("w_e", "r2"),
    ("w_i", "r2"), ("cfl_z", "r2"),
]
DYN_FER_ITEMS = [("fer_w", "r2"), ("fer_uv", "r3")]
DYN_SE_ITEMS = [
    ("se_uvh", "r3"), ("se_uvBT_rhs", "r2"), ("se_uvBT_4AB", "r2"), ("se_uvBT", "r2"), ("se_uvBT_theta", "r2"),
    ("se_uvBT_mean", "r2"), ("se_uvBT_12", "r2"), ("se_uvBT_stab_hvisc", "r2"), ("se_uvBT_stab_bdrag", "r1"),
]


def _f32(x):
    """A default-real literal given to a real64 component, as the reference's
    type defaults are (e.g. visc_gamma0 = 0.03 in MOD_DYN.F90)."""
    return float(np.float32(x))


# the reference's component defaults (MOD_DYN.F90:13-170, MOD_TRACER.F90:9-27),
# for records written from scratch
DYN_DEFAULTS = dict(opt_visc=5, visc_gamma0=_f32(0.03), visc_gamma1=_f32(0.1), visc_gamma2=_f32(0.285),
                    visc_easybsreturn=1.5, use_ivertvisc=True, momadv_opt=2, use_freeslip=False, use_wsplit=False,
                    wsplit_maxcfl=1.0, use_ssh_se_subcycl=False)
SOLVERINFO_DEFAULTS = dict(ident=1, maxiter=2000, restart=15, fillin=3, lutype=2, droptol=_f32(1.e-8),
                           soltol=_f32(1e-5))
TRACER_DATA_DEFAULTS = dict(smooth_bh_tra=False, gamma0_tra=0.0, gamma1_tra=0.0, gamma2_tra=0.0,
                            i_vert_diff=False, tra_adv_ph=1.0, tra_adv_pv=1.0)

_MAX_SUB = 2 ** 31 - 9   # largest payload of one (sub)record


# ------------------------------------------------------------------ framing
def read_record(path: str) -> bytes:
    """Payload of the one record of a sequential unformatted file."""
    raw = open(path, "rb").read()
    out = []
    pos = 0
    while True:
        if pos + 4 > len(raw):
            raise ValueError(f"{path}: truncated record marker")
        (m,) = struct.unpack_from("<i", raw, pos)
        n = abs(m)
        if pos + 8 + n > len(raw):
            raise ValueError(f"{path}: record of {n} bytes runs past the end of the file")
        out.append(raw[pos + 4:pos + 4 + n])
        (tail,) = struct.unpack_from("<i", raw, pos + 4 + n)
        if abs(tail) != n:
            raise ValueError(f"{path}: record markers {m} and {tail} differ")
        pos += 8 + n
        if m >= 0:
            break
    if pos != len(raw):
        raise ValueError(f"{path}: {len(raw) - pos} bytes after the record (a restart holds one record)")
    return b"".join(out)


def write_record(path: str, payload: bytes) -> None:
    with open(path, "wb") as f:
        if len(payload) <= _MAX_SUB:
            f.write(struct.pack("<i", len(payload)))
            f.write(payload)
            f.write(struct.pack("<i", len(payload)))
            return
        pos = 0
        first = True
        while pos < len(payload):
            n = min(_MAX_SUB, len(payload) - pos)
            more = pos + n < len(payload)
            lead = -n if more else n
            trail = n if first else -n
            f.write(struct.pack("<i", lead))
            f.write(payload[pos:pos + n])
            f.write(struct.pack("<i", trail))
            pos += n
            first = False


# ------------------------------------------------------------------ items
class _Payload:
    def __init__(self, data: bytes):
        self.b = memoryview(data)
        self.p = 0

    def take(self, n):
        if self.p + n > len(self.b):
            raise ValueError("record ends before the last item of its type")
        v = self.b[self.p:self.p + n]
        self.p += n
        return v

    def item(self, kind):
        if kind == "i4":
            return struct.unpack("<i", self.take(4))[0]
        if kind == "f8":
            return struct.unpack("<d", self.take(8))[0]
        if kind == "l4":
            return struct.unpack("<i", self.take(4))[0] != 0
        if kind.startswith("c"):
            return bytes(self.take(int(kind[1:]))).decode("ascii")
        nd = int(kind[1])
        ext = struct.unpack(f"<{nd}i", self.take(4 * nd))
        if ext[0] == 0:
            return None
        dt = "<i4" if kind[0] == "i" else "<f8"
        cnt = int(np.prod(ext))
        a = np.frombuffer(self.take(cnt * np.dtype(dt).itemsize), dtype=dt)
        return a.reshape(ext[::-1]).astype(np.int32 if kind[0] == "i" else np.float64)

    def done(self):
        return self.p == len(self.b)


def _put(out: list, kind: str, v) -> None:
    if kind == "i4":
        out.append(struct.pack("<i", int(v)))
    elif kind == "f8":
        out.append(struct.pack("<d", float(v)))
    elif kind == "l4":
        out.append(struct.pack("<i", 1 if v else 0))
    elif kind.startswith("c"):
        n = int(kind[1:])
        s = (v or "").encode("ascii")[:n]
        out.append(s + b" " * (n - len(s)))
    else:
        nd = int(kind[1])
        if v is None:
            out.append(struct.pack(f"<{nd}i", *([0] * nd)))
            return
        a = np.asarray(v)
        if a.ndim != nd:
            raise ValueError(f"{nd}-D array expected, got shape {a.shape}")
        out.append(struct.pack(f"<{nd}i", *a.shape[::-1]))
        out.append(np.ascontiguousarray(a, dtype="<i4" if kind[0] == "i" else "<f8").tobytes())


def _parse(p: _Payload, items, rec: dict):
    for name, kind in items:
        rec[name] = p.item(kind)


def _emit(out: list, items, rec: dict, defaults: Optional[dict] = None):
    for name, kind in items:
        if name in rec:
            v = rec[name]
        elif defaults and name in defaults:
            v = defaults[name]
        elif kind in ("i4", "f8", "l4"):
            v = 0
        else:
            v = None
        _put(out, kind, v)


# ------------------------------------------------------------------ t_mesh
def read_t_mesh(path: str) -> Dict:
    p = _Payload(read_record(path))
    rec: Dict = {}
    _parse(p, MESH_ITEMS, rec)
    if not p.done():
        raise ValueError(f"{path}: bytes left after the last item of t_mesh")
    return rec


def write_t_mesh(path: str, rec: Dict) -> None:
    out: list = []
    _emit(out, MESH_ITEMS, rec)
    write_record(path, b"".join(out))


# ---------------------------------------------------------------- t_tracer
def read_t_tracer(path: str) -> Dict:
    p = _Payload(read_record(path))
    n = p.item("i4")
    data = []
    for _ in range(n):
        d: Dict = {}
        _parse(p, TRACER_DATA_ITEMS, d)
        data.append(d)
    work: Dict = {}
    _parse(p, TRACER_WORK_ITEMS, work)
    if not p.done():
        raise ValueError(f"{path}: bytes left after the last item of t_tracer")
    return dict(num_tracers=n, data=data, work=work)


def write_t_tracer(path: str, rec: Dict) -> None:
    out: list = []
    _put(out, "i4", rec["num_tracers"])
    for d in rec["data"]:
        _emit(out, TRACER_DATA_ITEMS, d, TRACER_DATA_DEFAULTS)
    _emit(out, TRACER_WORK_ITEMS, rec.get("work", {}))
    write_record(path, b"".join(out))


# ---------------------------------------------------------------- t_dyn
def read_t_dyn(path: str, fer_gm: Optional[bool] = None) -> Dict:
    """fer_gm: o_PARAM Fer_GM of the run that wrote the file, which decides
    whether fer_w and fer_uv follow cfl_z; None infers it from what is left
    of the record."""
    p = _Payload(read_record(path))
    rec: Dict = {}
    _parse(p, DYN_HEAD_ITEMS, rec)
    si: Dict = {}
    _parse(p, SOLVERINFO_ITEMS, si)
    rec["solverinfo"] = si
    work: Dict = {}
    _parse(p, DYN_WORK_ITEMS, work)
    rec["work"] = work
    _parse(p, DYN_ARRAY_ITEMS, rec)
    if fer_gm is None:
        fer_gm = _remaining_has_fer(p) if rec["use_ssh_se_subcycl"] else not p.done()
    if fer_gm:
        _parse(p, DYN_FER_ITEMS, rec)
    if rec["use_ssh_se_subcycl"]:
        _parse(p, DYN_SE_ITEMS, rec)
    rec["Fer_GM"] = bool(fer_gm)
    if not p.done():
        raise ValueError(f"{path}: bytes left after the last item of t_dyn (Fer_GM={fer_gm})")
    return rec


def _remaining_has_fer(p: _Payload) -> bool:
    """With split-explicit subcycling on, the GM arrays are present when
    parsing them and the nine subcycling arrays uses up the record exactly."""
    save = p.p
    try:
        for _, kind in DYN_FER_ITEMS + DYN_SE_ITEMS:
            p.item(kind)
        return p.done()
    except ValueError:
        return False
    finally:
        p.p = save


def write_t_dyn(path: str, rec: Dict, fer_gm: Optional[bool] = None) -> None:
    out: list = []
    _emit(out, DYN_HEAD_ITEMS, rec, DYN_DEFAULTS)
    _emit(out, SOLVERINFO_ITEMS, rec.get("solverinfo", {}), SOLVERINFO_DEFAULTS)
    _emit(out, DYN_WORK_ITEMS, rec.get("work", {}))
    _emit(out, DYN_ARRAY_ITEMS, rec)
    if fer_gm if fer_gm is not None else rec.get("Fer_GM", False):
        _emit(out, DYN_FER_ITEMS, rec)
    if rec.get("use_ssh_se_subcycl", False):
        _emit(out, DYN_SE_ITEMS, rec)
    write_record(path, b"".join(out))


# ------------------------------------------------------- to and from meshgen
def mesh_record(mesh: mg.Mesh, ssh: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None) -> Dict:
    """The t_mesh items of a meshgen Mesh, 1-based as the reference holds them;
    ssh = (rowptr, colind, values) 0-based, the assembled stiffness matrix."""
    one = lambda a: np.asarray(a, np.int32) + 1
    rec = dict(nod2D=mesh.nod2D, edge2D=mesh.edge2D, edge2D_in=mesh.edge2D_in, elem2D=mesh.elem2D, nl=mesh.nl,
               ocean_area=float(mesh.areasvol[:, 0].sum()), ocean_areawithcav=float(mesh.area[:, 0].sum()),
               elem2D_nodes=one(mesh.elem2D_nodes), edges=one(mesh.edges),
               edge_tri=np.where(mesh.edge_tri >= 0, mesh.edge_tri + 1, 0).astype(np.int32),
               elem_area=mesh.elem_area, edge_dxdy=mesh.edge_dxdy, edge_cross_dxdy=mesh.edge_cross_dxdy,
               elem_cos=mesh.elem_cos,
               elem_neighbors=np.where(mesh.elem_neighbors >= 0, mesh.elem_neighbors + 1, 0).astype(np.int32),
               nod_in_elem2D=np.where(mesh.nod_in_elem2D >= 0, mesh.nod_in_elem2D + 1, 0).astype(np.int32),
               nod_in_elem2D_num=np.asarray(mesh.nod_in_elem2D_num, np.int32), depth=mesh.depth,
               gradient_sca=mesh.gradient_sca, zbar=mesh.zbar, Z=mesh.Z,
               ulevels=np.asarray(mesh.ulevels, np.int32), ulevels_nod2D=np.asarray(mesh.ulevels_nod2D, np.int32),
               ulevels_nod2D_max=np.asarray(mesh.ulevels_nod2D_max, np.int32),
               nlevels=np.asarray(mesh.nlevels, np.int32), nlevels_nod2D=np.asarray(mesh.nlevels_nod2D, np.int32),
               nlevels_nod2D_min=np.asarray(mesh.nlevels_nod2D_min, np.int32), area=mesh.area,
               areasvol=mesh.areasvol, hnode=mesh.hnode, hnode_new=mesh.hnode_new, zbar_3d_n=mesh.zbar_3d_n,
               Z_3d_n=mesh.Z_3d_n, helem=mesh.helem, zbar_n_bot=mesh.zbar_n_bot, zbar_e_bot=mesh.zbar_e_bot,
               zbar_e_srf=mesh.zbar_e_srf)
    if ssh is not None:
        rp, ci, val = ssh
        rec.update({"ssh_stiff%dim": mesh.nod2D, "ssh_stiff%nza": len(ci),
                    "ssh_stiff%rowptr": one(rp), "ssh_stiff%colind": one(ci), "ssh_stiff%values": val,
                    "ssh_stiff%rowptr_loc": one(rp), "ssh_stiff%colind_loc": one(ci)})
    return rec


def mesh_from_restart(rec: Dict, coord_nod2D: np.ndarray, cyclic_length: float, cartesian: bool) -> mg.Mesh:
    """A meshgen Mesh from a t_mesh record. coord_nod2D (nod2D, 2) radians and
    cyclic_length (radians) are not in the record: take them from the mesh
    files (meshio) or the run's configuration."""
    zero = lambda a: np.asarray(a, np.int32) - 1
    need = [k for k in ("elem2D_nodes", "edges", "edge_tri", "nod_in_elem2D", "nlevels", "area", "hnode",
                        "Z_3d_n", "zbar_3d_n", "helem", "gradient_sca", "edge_cross_dxdy") if rec.get(k) is None]
    if need:
        raise ValueError(f"t_mesh record lacks {need}")
    N2 = rec["nod2D"]
    coord = np.asarray(coord_nod2D, np.float64)
    if coord.shape != (N2, 2):
        raise ValueError(f"coord_nod2D has shape {coord.shape}, the record has {N2} nodes")
    et = np.asarray(rec["edge_tri"])
    nie = np.asarray(rec["nod_in_elem2D"])
    nb = rec.get("elem_neighbors")
    return mg.Mesh(
        nod2D=N2, elem2D=rec["elem2D"], edge2D=rec["edge2D"], edge2D_in=rec["edge2D_in"], nl=rec["nl"],
        cartesian=bool(cartesian), cyclic_length=float(cyclic_length), coord_nod2D=coord,
        elem2D_nodes=zero(rec["elem2D_nodes"]), edges=zero(rec["edges"]),
        edge_tri=np.where(et > 0, et - 1, -1).astype(np.int32),
        nod_in_elem2D=np.where(nie > 0, nie - 1, -1).astype(np.int32),
        nod_in_elem2D_num=np.asarray(rec["nod_in_elem2D_num"], np.int32),
        elem_neighbors=(np.where(nb > 0, nb - 1, -1).astype(np.int32) if nb is not None
                        else np.full((rec["elem2D"], 3), -1, np.int32)),
        zbar=rec["zbar"], Z=rec["Z"], depth=rec["depth"], nlevels=np.asarray(rec["nlevels"], np.int32),
        ulevels=np.asarray(rec["ulevels"], np.int32), nlevels_nod2D=np.asarray(rec["nlevels_nod2D"], np.int32),
        ulevels_nod2D=np.asarray(rec["ulevels_nod2D"], np.int32),
        nlevels_nod2D_min=np.asarray(rec["nlevels_nod2D_min"], np.int32),
        ulevels_nod2D_max=np.asarray(rec["ulevels_nod2D_max"], np.int32), elem_area=rec["elem_area"],
        area=rec["area"], areasvol=rec["areasvol"], elem_cos=rec["elem_cos"], edge_dxdy=rec["edge_dxdy"],
        edge_cross_dxdy=rec["edge_cross_dxdy"], gradient_sca=rec["gradient_sca"], zbar_e_bot=rec["zbar_e_bot"],
        zbar_e_srf=rec["zbar_e_srf"], zbar_n_bot=rec["zbar_n_bot"], zbar_3d_n=rec["zbar_3d_n"],
        Z_3d_n=rec["Z_3d_n"], hnode=rec["hnode"], hnode_new=rec["hnode_new"], helem=rec["helem"])


def tracer_record(state: mg.State, schemes=None) -> Dict:
    """t_tracer items of a meshgen State (AB2: valuesold(1, nl-1, nod2D))."""
    data = []
    for i in range(state.ntr):
        sch = dict(hor="MFCT", ver="QR4C", lim="FCT", ph=0.0, pv=1.0)
        if schemes:
            sch.update(schemes[i] if isinstance(schemes, (list, tuple)) else schemes)
        data.append(dict(values=state.values[i], valuesold=state.valuesold[i][:, :, None],
                         valuesAB=np.zeros_like(state.values[i]), i_vert_diff=True, tra_adv_hor=sch["hor"],
                         tra_adv_ver=sch["ver"], tra_adv_lim=sch["lim"], tra_adv_ph=sch["ph"],
                         tra_adv_pv=sch["pv"], ID=int(state.ids[i])))
    return dict(num_tracers=state.ntr, data=data, work={})


def dyn_record(state: mg.State) -> Dict:
    return dict(uv=state.uv, w=state.w, w_e=state.w_e, w_i=state.w_i,
                use_wsplit=bool(getattr(state, "use_wsplit", False)))


def state_from_restart(mesh: mg.Mesh, trec: Dict, drec: Dict, Kv: np.ndarray, dt: float,
                       forcing: Optional[Dict[str, np.ndarray]] = None):
    """A meshgen State and the tracers' schemes from t_tracer and t_dynamics
    records. Kv (o_ARRAYS) and the surface fluxes are not in these records;
    forcing maps heat_flux, water_flux, virtual_salt, relax_salt,
    real_salt_flux to (nod2D,) arrays (zero where absent)."""
    ntr = trec["num_tracers"]
    vals, olds, schemes = [], [], []
    for d in trec["data"]:
        vo = d["valuesold"]
        if vo is None or vo.shape[-1] != 1:
            raise ValueError("an AB2 tracer (valuesold(1, nl-1, nod2D)) is expected")
        vals.append(d["values"])
        olds.append(vo[:, :, 0])
        schemes.append(dict(hor=d["tra_adv_hor"].strip(), ver=d["tra_adv_ver"].strip(),
                            lim=d["tra_adv_lim"].strip(), ph=d["tra_adv_ph"], pv=d["tra_adv_pv"],
                            i_vert_diff=d["i_vert_diff"]))
    N2 = mesh.nod2D
    f = forcing or {}
    z = np.zeros(N2)
    st = mg.State(ids=np.array([d["ID"] for d in trec["data"]], np.int32), values=np.stack(vals),
                  valuesold=np.stack(olds), uv=drec["uv"], w=drec["w"], w_e=drec["w_e"], w_i=drec["w_i"],
                  Kv=np.asarray(Kv, np.float64), dt=float(dt), heat_flux=f.get("heat_flux", z),
                  water_flux=f.get("water_flux", z), virtual_salt=f.get("virtual_salt", z),
                  relax_salt=f.get("relax_salt", z), real_salt_flux=f.get("real_salt_flux", z),
                  use_wsplit=bool(drec.get("use_wsplit", False)))
    assert st.ntr == ntr
    return st, schemes


def rank_suffix(rank: int = 0, npes: int = 1) -> str:
    """The rank as mpirank_to_txt writes it: zero-padded to the number of
    digits of npes (fortran_utils.F90:45-56)."""
    return str(rank).zfill(len(str(npes)) if npes > 0 else 1)


def write_restart(directory: str, rank: int = 0, npes: int = 1, mesh_rec: Optional[Dict] = None,
                  tracer_rec: Optional[Dict] = None, dyn_rec: Optional[Dict] = None) -> None:
    """The files write_all_bin_restarts writes for these types on one rank:
    t_mesh.<rank>, t_tracer.<rank>, t_dynamics.<rank>."""
    os.makedirs(directory, exist_ok=True)
    r = rank_suffix(rank, npes)
    if mesh_rec is not None:
        write_t_mesh(os.path.join(directory, f"t_mesh.{r}"), mesh_rec)
    if tracer_rec is not None:
        write_t_tracer(os.path.join(directory, f"t_tracer.{r}"), tracer_rec)
    if dyn_rec is not None:
        write_t_dyn(os.path.join(directory, f"t_dynamics.{r}"), dyn_rec)


def read_restart(directory: str, rank: int = 0, npes: int = 1, fer_gm: Optional[bool] = None) -> Dict:
    """The t_mesh, t_tracer and t_dynamics records of one rank that exist in
    the directory."""
    r = rank_suffix(rank, npes)
    out = {}
    for key, name, fn in (("mesh", "t_mesh", read_t_mesh), ("tracers", "t_tracer", read_t_tracer),
                          ("dynamics", "t_dynamics", lambda pth: read_t_dyn(pth, fer_gm))):
        pth = os.path.join(directory, f"{name}.{r}")
        if os.path.exists(pth):
            out[key] = fn(pth)
    return out
