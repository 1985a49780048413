"""ctypes binding of libfesom_gpu.so (the C ABI declared in include/fesom_gpu.h).

There is no fallback: if the library or a GPU is missing, loading or creating a
context raises. Arrays are numpy, C-contiguous, with the memory image of the
reference's Fortran arrays (level index fastest); connectivity handed to the
library is 1-based, exactly as the Fortran derived types hold it.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfesom_gpu.so")

dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int)
vp = C.c_void_p


class FesomGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fesom_gpu error {code}: {msg}")
        self.code = code


class CgStats(C.Structure):
    _fields_ = [("iters_last", C.c_int), ("converged", C.c_int), ("nbreakdown", C.c_int),
                ("nnegrz", C.c_int), ("resid_last", C.c_double), ("rtol_last", C.c_double),
                ("ms", C.c_double)]

    def as_dict(self):
        return dict(iters=self.iters_last, converged=bool(self.converged), nbreakdown=self.nbreakdown,
                    nnegrz=self.nnegrz, resid=self.resid_last, rtol=self.rtol_last, ms=self.ms)


# (name, argtypes); every function returns int status
_SIGS = {
    "fgpu_create": [C.POINTER(vp), C.c_int],
    "fgpu_destroy": [vp],
    "fgpu_synchronize": [vp],
    "fgpu_mesh_init": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, dp, ip, ip, ip, ip, ip,
                       ip, ip, ip, ip, ip, ip, dp, dp, dp, dp, dp],
    "fgpu_set_vertical": [vp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp],
    "fgpu_set_velocity": [vp, dp, dp, dp, dp, C.c_int],
    "fgpu_set_kv": [vp, dp],
    "fgpu_set_ki": [vp, dp, dp, C.c_int],
    "fgpu_set_surface_forcing": [vp, dp, dp, dp, dp, dp, C.c_double],
    "fgpu_set_surface_forcing_tracers": [vp, dp, dp, dp, dp, C.c_int],
    "fgpu_tracer_set_valuesAB": [vp, C.c_int, dp],
    "fgpu_ssh_set_preconditioner": [vp, C.c_int, dp],
    "fgpu_set_uv": [vp, dp],
    "fgpu_vert_vel_ale": [vp, C.c_double, C.c_int, dp, dp, C.c_int, C.c_double, dp, ip],
    "fgpu_vert_vel_download": [vp, dp, dp, dp, dp, dp],
    "fgpu_compute_ssh_rhs_ale": [vp, dp, dp, C.c_double, C.c_int, C.c_int, dp],
    "fgpu_compute_hbar_ale": [vp, C.c_double, C.c_int, dp, dp, dp, dp, dp],
    "fgpu_set_ale_config": [vp, dp, C.c_double, C.c_int],
    "fgpu_tracers_init": [vp, C.c_int],
    "fgpu_tracer_config": [vp, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_double,
                           C.c_double, C.c_int],
    "fgpu_tracer_upload": [vp, C.c_int, dp, dp],
    "fgpu_tracer_download": [vp, C.c_int, dp, dp],
    "fgpu_tracer_work_download": [vp, C.c_int, C.c_char_p, dp, C.c_long],
    "fgpu_edge_up_dn_tri": [vp, ip],
    "fgpu_init_tracers_AB": [vp, C.c_int],
    "fgpu_do_oce_adv_tra": [vp, C.c_double, C.c_int],
    "fgpu_diff_tracers_ale": [vp, C.c_double, C.c_int],
    "fgpu_diff_ver_part_impl_ale": [vp, C.c_double, C.c_int],
    "fgpu_solve_tracers_ale": [vp, C.c_double, C.c_int],
    "fgpu_solve_tracers_ale_timed": [vp, C.c_double, C.c_int, dp],
    "fgpu_profile_enable": [vp, C.c_int],
    "fgpu_profile_count": [vp],
    "fgpu_profile_read": [vp, C.c_int, C.c_char_p, C.c_int, dp, ip],
    "fgpu_tracer_content": [vp, C.c_int, dp],
    "fgpu_autotune_step": [vp, C.c_double, C.c_int, ip, dp],
    "fgpu_step_shape": [vp, ip, ip, ip],
    "fgpu_step_variant": [vp, ip, ip, ip, ip],
    "fgpu_set_step_shape": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "fgpu_init_stiff_mat_ale": [vp, C.c_double, C.c_double, C.c_double, C.c_double],
    "fgpu_update_stiff_mat_ale": [vp, dp, C.c_double, C.c_double, C.c_double, C.c_double],
    "fgpu_ssh_set_matrix": [vp, C.c_int, C.c_int, ip, ip, dp],
    "fgpu_ssh_get_matrix": [vp, ip, ip, dp, dp, ip],
    "fgpu_ssh_solve_preconditioner": [vp, C.c_int],
    "fgpu_ssh_solve_cg": [vp, dp, dp, C.c_double, C.c_int, C.POINTER(CgStats)],
    "fgpu_ssh_upload": [vp, dp, dp],
    "fgpu_ssh_solve_resident": [vp, C.c_double, C.c_int, C.c_int, C.POINTER(CgStats)],
    "fgpu_ssh_download": [vp, dp],
    "fgpu_ssh_solve_cg_noalloc": [vp, vp, vp, C.c_double, C.c_int, C.POINTER(CgStats)],
    "fgpu_ssh_nn_param_count": [],
    "fgpu_ssh_nn_init": [vp, C.c_ulonglong],
    "fgpu_ssh_nn_set_weights": [vp, fp, C.c_int],
    "fgpu_ssh_nn_get_weights": [vp, fp, C.c_int],
    "fgpu_ssh_nn_apply": [vp, dp, dp],
    "fgpu_ssh_nn_loss_grad": [vp, dp, dp, C.c_int, dp, fp],
    "fgpu_ssh_nn_train": [vp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_double, dp],
    "fgpu_ssh_solve_learned": [vp, dp, dp, C.c_double, C.c_int, C.c_int, C.POINTER(CgStats), ip],
    "fgpu_ssh_solve_learned_resident": [vp, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(CgStats), ip],
}
EXPORTED = sorted(list(_SIGS) + ["fgpu_last_error", "fgpu_version"])

_lib = None
_live = weakref.WeakSet()   # open contexts, closed by an exit hook before the HIP runtime tears down


def _close_live_contexts():
    for ctx in list(_live):
        try:
            ctx.close()
        except Exception:
            pass


def load(path: Optional[str] = None):
    """Load the library (raises if it is not built). The first load registers
    an exit hook that closes every context still open, while the HIP runtime
    is alive; a context collected later, during interpreter shutdown, releases
    nothing itself."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FesomGpuError(-100, f"{path} is not built; run fesom2_amd/build.py (or __graft_entry__.build())")
    L = C.CDLL(path)
    L.fgpu_last_error.restype = C.c_char_p
    L.fgpu_last_error.argtypes = []
    L.fgpu_version.restype = C.c_int
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    atexit.register(_close_live_contexts)
    return L


def _check(rc):
    if rc < 0:
        raise FesomGpuError(rc, load().fgpu_last_error().decode())
    return rc


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "float64 C-contiguous array expected"
    return a.ctypes.data_as(dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous, "int32 C-contiguous array expected"
    return a.ctypes.data_as(ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """One FESOM2 partition resident on one GPU."""

    def __init__(self, device: int = 0):
        L = load()
        h = vp()
        _check(L.fgpu_create(C.byref(h), int(device)))
        self.h = h
        self.L = L
        _live.add(self)
        self.ntr = 0
        self.nl = 0
        self.N2 = 0
        self.F2 = 0

    def close(self):
        if self.h:
            self.L.fgpu_destroy(self.h)
            self.h = None
        _live.discard(self)

    def __del__(self):
        if sys.is_finalizing():   # the exit hook has closed it, or the runtime may be gone
            return
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ mesh & fields
    def mesh_init(self, m):
        """Upload a meshgen.Mesh (0-based numpy connectivity -> 1-based)."""
        self._mesh_keep = [
            _f64(m.coord_nod2D), _i32(m.elem2D_nodes + 1), _i32(m.edges + 1),
            _i32(np.where(m.edge_tri >= 0, m.edge_tri + 1, -999)),
            _i32(np.where(m.nod_in_elem2D >= 0, m.nod_in_elem2D + 1, 0)), _i32(m.nod_in_elem2D_num),
            _i32(m.nlevels), _i32(m.ulevels), _i32(m.nlevels_nod2D), _i32(m.ulevels_nod2D),
            _i32(m.nlevels_nod2D_min), _i32(m.ulevels_nod2D_max), _f64(m.elem_area), _f64(m.elem_cos),
            _f64(m.edge_dxdy), _f64(m.edge_cross_dxdy), _f64(m.gradient_sca)]
        k = self._mesh_keep
        _check(self.L.fgpu_mesh_init(self.h, m.nod2D, m.elem2D, m.edge2D, m.nl, m.kmax, m.cyclic_length,
                                     _d(k[0]), _i(k[1]), _i(k[2]), _i(k[3]), _i(k[4]), _i(k[5]), _i(k[6]),
                                     _i(k[7]), _i(k[8]), _i(k[9]), _i(k[10]), _i(k[11]), _d(k[12]),
                                     _d(k[13]), _d(k[14]), _d(k[15]), _d(k[16])))
        self.nl = m.nl
        self.N2 = m.nod2D
        self.F2 = m.elem2D
        self.set_vertical(m)

    def set_vertical(self, m):
        arrs = [_f64(a) for a in (m.area, m.areasvol, m.hnode, m.hnode_new, m.helem, m.zbar_3d_n, m.Z_3d_n,
                                  m.zbar_n_bot, m.zbar_e_bot, m.zbar_e_srf)]
        _check(self.L.fgpu_set_vertical(self.h, *[_d(a) for a in arrs]))

    def set_velocity(self, uv, w, w_e, w_i, use_wsplit=False):
        a = [_f64(x) for x in (uv, w, w_e, w_i)]
        _check(self.L.fgpu_set_velocity(self.h, *[_d(x) for x in a], int(bool(use_wsplit))))

    def set_kv(self, Kv):
        _check(self.L.fgpu_set_kv(self.h, _d(_f64(Kv))))

    def set_ki(self, Ki, slope_tapered=None, redi=False):
        """Ki (N2, nl-1), slope_tapered (N2, nl-1, 3), Redi; Ki None means Ki = 0."""
        ki = None if Ki is None else _f64(Ki)
        sl = None if slope_tapered is None else _f64(slope_tapered)
        _check(self.L.fgpu_set_ki(self.h, None if ki is None else _d(ki), None if sl is None else _d(sl),
                                  int(bool(redi))))

    def set_surface_forcing(self, heat_flux, water_flux, virtual_salt, relax_salt, real_salt_flux, is_nonlinfs=0.0):
        a = [_f64(x) for x in (heat_flux, water_flux, virtual_salt, relax_salt, real_salt_flux)]
        _check(self.L.fgpu_set_surface_forcing(self.h, *[_d(x) for x in a], float(is_nonlinfs)))

    def set_state(self, st):
        """Velocities, Kv and surface forcing of a meshgen.State."""
        self.set_velocity(st.uv, st.w, st.w_e, st.w_i, use_wsplit=getattr(st, "use_wsplit", False))
        self.set_kv(st.Kv)
        self.set_surface_forcing(st.heat_flux, st.water_flux, st.virtual_salt, st.relax_salt, st.real_salt_flux)
        if getattr(st, "Ki", None) is not None:
            self.set_ki(st.Ki, getattr(st, "slope_tapered", None), getattr(st, "redi", False))
        extra = [getattr(st, f, None) for f in ("prec_rain", "wiso_flux_oce", "hosing_flux", "hosing_flux3D")]
        if any(a is not None for a in extra):
            self.set_surface_forcing_tracers(*extra, hosing_depth=getattr(st, "hosing_depth", False))

    def set_surface_forcing_tracers(self, prec_rain=None, wiso_flux_oce=None, hosing_flux=None, hosing_flux3D=None,
                                    hosing_depth=False):
        """Fluxes of the tracer IDs 101 (prec_rain, N2), 102-104 (wiso_flux_oce, (3, N2): the Fortran
        (N2, 3) array) and 304 (hosing_flux, N2; hosing_flux3D, (N2, nl-1))."""
        keep = [None if a is None else _f64(a) for a in (prec_rain, wiso_flux_oce, hosing_flux, hosing_flux3D)]
        _check(self.L.fgpu_set_surface_forcing_tracers(self.h, *[_d(a) for a in keep], int(bool(hosing_depth))))

    # ------------------------------------------- dynamics either side of the transport
    ALE = {"linfs": 0, "zlevel": 1, "zstar": 2}

    def set_uv(self, uv):
        _check(self.L.fgpu_set_uv(self.h, _d(_f64(uv))))

    def vert_vel_ale(self, dt, which_ale="linfs", hbar=None, hbar_old=None, use_wsplit=False, wsplit_maxcfl=1.0):
        """W, w_e, w_i, CFL_z on the device from the resident uv; returns (cflmax, nan_columns)."""
        cmax = C.c_double()
        nan = C.c_int()
        hb = None if hbar is None else _f64(hbar)
        ho = None if hbar_old is None else _f64(hbar_old)
        _check(self.L.fgpu_vert_vel_ale(self.h, float(dt), self.ALE[which_ale], _d(hb) if hb is not None else None,
                                        _d(ho) if ho is not None else None, int(bool(use_wsplit)),
                                        float(wsplit_maxcfl), C.byref(cmax), C.byref(nan)))
        return cmax.value, nan.value

    def vert_vel_download(self):
        NL = self.nl
        out = {k: np.zeros((self.N2, NL)) for k in ("w", "w_e", "w_i", "cfl_z")}
        out["hnode_new"] = np.zeros((self.N2, NL - 1))
        _check(self.L.fgpu_vert_vel_download(self.h, _d(out["w"]), _d(out["w_e"]), _d(out["w_i"]),
                                             _d(out["cfl_z"]), _d(out["hnode_new"])))
        return out

    def compute_ssh_rhs_ale(self, uv_rhs, ssh_rhs_old, alpha=1.0, which_ale="linfs", use_cavity_fw2press=False):
        """ssh_rhs_old None: the device's, from compute_hbar_ale."""
        out = np.zeros(self.N2)
        old = None if ssh_rhs_old is None else _f64(ssh_rhs_old)
        _check(self.L.fgpu_compute_ssh_rhs_ale(self.h, _d(_f64(uv_rhs)), None if old is None else _d(old),
                                               float(alpha), self.ALE[which_ale], int(bool(use_cavity_fw2press)),
                                               _d(out)))
        return out

    def set_ale_config(self, zbar, min_hnode=0.5, lzstar_lev=4):
        """zlevel settings: mesh%zbar (nl), min_hnode, lzstar_lev."""
        _check(self.L.fgpu_set_ale_config(self.h, _d(_f64(zbar)), float(min_hnode), int(lzstar_lev)))

    def compute_hbar_ale(self, dt, which_ale="linfs", hbar=None):
        """compute_hbar_ale; hbar None continues from the device's. Returns
        dict(hbar, hbar_old, ssh_rhs_old, dhe); all four also stay on the device."""
        h = None if hbar is None else _f64(hbar)
        out = dict(hbar=np.zeros(self.N2), hbar_old=np.zeros(self.N2), ssh_rhs_old=np.zeros(self.N2),
                   dhe=np.zeros(self.F2))
        _check(self.L.fgpu_compute_hbar_ale(self.h, float(dt), self.ALE[which_ale], None if h is None else _d(h),
                                            _d(out["hbar"]), _d(out["hbar_old"]), _d(out["ssh_rhs_old"]),
                                            _d(out["dhe"])))
        return out

    # ------------------------------------------------------------ tracers
    def tracers_init(self, ntr):
        _check(self.L.fgpu_tracers_init(self.h, int(ntr)))
        self.ntr = ntr
        self._abo = [2] * ntr

    def tracer_config(self, tr_num, ID, AB_order=2, hor="MFCT", ver="QR4C", lim="FCT", ph=0.0, pv=1.0,
                      i_vert_diff=True):
        _check(self.L.fgpu_tracer_config(self.h, int(tr_num), int(ID), int(AB_order), hor.encode(), ver.encode(),
                                         lim.encode(), float(ph), float(pv), int(bool(i_vert_diff))))
        self._abo[tr_num - 1] = AB_order

    def tracer_upload(self, tr_num, values, valuesold):
        """values (nod2D, nl-1); valuesold (nod2D, nl-1) for AB_order 2 or
        (nod2D, nl-1, 2) for AB_order 3, the newer level first."""
        v, vo = _f64(values), _f64(valuesold)
        n3 = self.N2 * (self.nl - 1)
        abo = self._abo[tr_num - 1]
        if v.size != n3 or vo.size != (abo - 1) * n3:
            raise ValueError(f"tracer {tr_num}: values must hold {n3} and valuesold {(abo - 1) * n3} numbers "
                             f"(AB_order {abo}), got {v.size} and {vo.size}")
        _check(self.L.fgpu_tracer_upload(self.h, int(tr_num), _d(v), _d(vo)))

    def tracer_download(self, tr_num):
        L = self.nl - 1
        v = np.zeros((self.N2, L))
        abo = self._abo[tr_num - 1]
        vo = np.zeros((self.N2, L) if abo == 2 else (self.N2, L, abo - 1))
        _check(self.L.fgpu_tracer_download(self.h, int(tr_num), _d(v), _d(vo)))
        return v, vo

    def work(self, tr_num, name, shape):
        out = np.zeros(shape)
        _check(self.L.fgpu_tracer_work_download(self.h, int(tr_num), name.encode(), _d(out), int(out.size)))
        return out

    def edge_up_dn_tri(self, E2):
        out = np.zeros((E2, 2), np.int32)
        _check(self.L.fgpu_edge_up_dn_tri(self.h, _i(out)))
        return out

    # ------------------------------------------------------------ entry points
    def init_tracers_AB(self, tr_num):
        _check(self.L.fgpu_init_tracers_AB(self.h, int(tr_num)))

    def do_oce_adv_tra(self, dt, tr_num):
        _check(self.L.fgpu_do_oce_adv_tra(self.h, float(dt), int(tr_num)))

    def diff_tracers_ale(self, dt, tr_num):
        _check(self.L.fgpu_diff_tracers_ale(self.h, float(dt), int(tr_num)))

    def diff_ver_part_impl_ale(self, dt, tr_num):
        _check(self.L.fgpu_diff_ver_part_impl_ale(self.h, float(dt), int(tr_num)))

    def solve_tracers_ale(self, dt, nsteps=1):
        _check(self.L.fgpu_solve_tracers_ale(self.h, float(dt), int(nsteps)))

    def solve_tracers_ale_timed(self, dt, nsteps):
        ms = C.c_double()
        _check(self.L.fgpu_solve_tracers_ale_timed(self.h, float(dt), int(nsteps), C.byref(ms)))
        return ms.value

    def profile_enable(self, on=True):
        _check(self.L.fgpu_profile_enable(self.h, int(bool(on))))

    def profile(self):
        n = _check(self.L.fgpu_profile_count(self.h))
        out = {}
        for i in range(n):
            name = C.create_string_buffer(64)
            ms = C.c_double()
            cnt = C.c_int()
            _check(self.L.fgpu_profile_read(self.h, i, name, 64, C.byref(ms), C.byref(cnt)))
            out[name.value.decode()] = (ms.value, cnt.value)
        return out

    def autotune_step(self, dt, nsteps=5):
        """Pick the fastest launch shape of the fused step; returns (tracers per block, ms/step)."""
        tpb = C.c_int()
        ms = C.c_double()
        _check(self.L.fgpu_autotune_step(self.h, float(dt), int(nsteps), C.byref(tpb), C.byref(ms)))
        return tpb.value, ms.value

    def step_shape(self):
        """(tracers per block, tracers per Thomas chunk, columns per block) of the fused step."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(self.L.fgpu_step_shape(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_step_shape(self, tpb, tch, cpb, kd=0, kc=0, ka=1, kb=1):
        """Fix (tracers per block, Thomas chunk, columns per block, last-kernel variant,
        tracers per block of the (tracer, level) bound kernel, 0 for the per-column one,
        levels per lane of the edge-flux and of the LO kernel, 1 or 2)."""
        _check(self.L.fgpu_set_step_shape(self.h, int(tpb), int(tch), int(cpb), int(kd), int(kc), int(ka), int(kb)))

    def step_variant(self):
        """Last kernel of the plain step: 0 lock-step, 1 with a sweep wave, 2 one wave per column."""
        return self.step_variants()[0]

    def step_variants(self):
        """(last-kernel variant, tracers per block of the (tracer, level) bound kernel or 0,
        levels per lane of the edge-flux kernel, levels per lane of the LO kernel)."""
        v, kc, ka, kb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(self.L.fgpu_step_variant(self.h, C.byref(v), C.byref(kc), C.byref(ka), C.byref(kb)))
        return v.value, kc.value, ka.value, kb.value

    def tracer_content(self, tr_num):
        v = C.c_double()
        _check(self.L.fgpu_tracer_content(self.h, int(tr_num), C.byref(v)))
        return v.value

    def synchronize(self):
        _check(self.L.fgpu_synchronize(self.h))

    # ------------------------------------------------------------ SSH
    def init_stiff_mat_ale(self, dt, g=9.81, alpha=1.0, theta=1.0):
        _check(self.L.fgpu_init_stiff_mat_ale(self.h, float(dt), float(g), float(alpha), float(theta)))

    def update_stiff_mat_ale(self, dhe, dt, g=9.81, alpha=1.0, theta=1.0):
        """dhe None: the device's, from compute_hbar_ale."""
        d = None if dhe is None else _f64(dhe)
        _check(self.L.fgpu_update_stiff_mat_ale(self.h, None if d is None else _d(d), float(dt), float(g),
                                                float(alpha), float(theta)))

    def ssh_set_matrix(self, rowptr1, colind1, values):
        rowptr1, colind1 = _i32(rowptr1), _i32(colind1)
        _check(self.L.fgpu_ssh_set_matrix(self.h, len(rowptr1) - 1, len(colind1), _i(rowptr1), _i(colind1),
                                          _d(_f64(values))))

    def ssh_get_matrix(self, with_pr=False):
        nza = C.c_int()
        _check(self.L.fgpu_ssh_get_matrix(self.h, None, None, None, None, C.byref(nza)))
        rp = np.zeros(self.N2 + 1, np.int32)
        ci = np.zeros(nza.value, np.int32)
        v = np.zeros(nza.value)
        pr = np.zeros(nza.value) if with_pr else None
        _check(self.L.fgpu_ssh_get_matrix(self.h, _i(rp), _i(ci), _d(v), _d(pr) if with_pr else None,
                                          C.byref(nza)))
        return rp, ci, v, pr

    def ssh_solve_preconditioner(self, variant):
        _check(self.L.fgpu_ssh_solve_preconditioner(self.h, int(variant)))

    def ssh_solve_cg(self, x, rhs, soltol=1e-10, maxiter=2000):
        x = _f64(x).copy()
        st = CgStats()
        _check(self.L.fgpu_ssh_solve_cg(self.h, _d(x), _d(_f64(rhs)), float(soltol), int(maxiter), C.byref(st)))
        return x, st.as_dict()

    def ssh_upload(self, x0, rhs):
        _check(self.L.fgpu_ssh_upload(self.h, _d(_f64(x0)), _d(_f64(rhs))))

    def ssh_solve_resident(self, soltol=1e-10, maxiter=2000, nrep=1):
        st = CgStats()
        _check(self.L.fgpu_ssh_solve_resident(self.h, float(soltol), int(maxiter), int(nrep), C.byref(st)))
        return st.as_dict()

    def ssh_solve_cg_noalloc(self, x_ptr, rhs_ptr, soltol=1e-10, maxiter=2000):
        """CG on the caller's device vectors (addresses of nod2D doubles on this
        context's device; x in/out); the work vectors come from the library's
        static device area, so the call allocates nothing."""
        st = CgStats()
        _check(self.L.fgpu_ssh_solve_cg_noalloc(self.h, C.c_void_p(int(x_ptr)), C.c_void_p(int(rhs_ptr)),
                                                float(soltol), int(maxiter), C.byref(st)))
        return st.as_dict()

    def ssh_download(self):
        x = np.zeros(self.N2)
        _check(self.L.fgpu_ssh_download(self.h, _d(x)))
        return x

    # ------------------------------------------------------------ learned preconditioner
    def nn_param_count(self):
        return _check(self.L.fgpu_ssh_nn_param_count())

    def nn_init(self, seed=0):
        _check(self.L.fgpu_ssh_nn_init(self.h, int(seed)))

    def nn_set_weights(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        _check(self.L.fgpu_ssh_nn_set_weights(self.h, th.ctypes.data_as(fp), int(th.size)))

    def nn_get_weights(self):
        th = np.zeros(self.nn_param_count(), np.float32)
        _check(self.L.fgpu_ssh_nn_get_weights(self.h, th.ctypes.data_as(fp), int(th.size)))
        return th

    def nn_apply(self, r):
        z = np.zeros(self.N2)
        _check(self.L.fgpu_ssh_nn_apply(self.h, _d(_f64(r)), _d(z)))
        return z

    def nn_loss_grad(self, bs, xs):
        b = _f64(np.asarray(bs))
        x = _f64(np.asarray(xs))
        loss = C.c_double()
        g = np.zeros(self.nn_param_count(), np.float32)
        _check(self.L.fgpu_ssh_nn_loss_grad(self.h, _d(b), _d(x), int(b.shape[0]), C.byref(loss),
                                            g.ctypes.data_as(fp)))
        return loss.value, g

    def nn_train(self, bs, xs, epochs=10, batch=8, lr=1e-3):
        b = _f64(np.asarray(bs))
        x = _f64(np.asarray(xs))
        hist = np.zeros(int(epochs))
        _check(self.L.fgpu_ssh_nn_train(self.h, _d(b), _d(x), int(b.shape[0]), int(epochs), int(batch),
                                        float(lr), _d(hist)))
        return hist

    def ssh_solve_learned(self, x, rhs, soltol=1e-10, maxiter=500, fallback_maxiter=2000):
        x = _f64(x).copy()
        st = CgStats()
        fb = C.c_int()
        _check(self.L.fgpu_ssh_solve_learned(self.h, _d(x), _d(_f64(rhs)), float(soltol), int(maxiter),
                                             int(fallback_maxiter), C.byref(st), C.byref(fb)))
        d = st.as_dict()
        d["fallback_iters"] = fb.value
        return x, d

    def ssh_solve_learned_resident(self, soltol=1e-10, maxiter=500, fallback_maxiter=2000, nrep=1):
        st = CgStats()
        fb = C.c_int()
        _check(self.L.fgpu_ssh_solve_learned_resident(self.h, float(soltol), int(maxiter), int(fallback_maxiter),
                                                      int(nrep), C.byref(st), C.byref(fb)))
        d = st.as_dict()
        d["fallback_iters"] = fb.value
        return d


def context_for(mesh, state=None, device=0, schemes=None, valuesold=None):
    """Context with mesh, state and tracers uploaded (helper for tests/bench).
    ``valuesold`` (one array per tracer) replaces the state's old levels, as an
    AB_order 3 tracer needs two of them."""
    ctx = Context(device)
    ctx.mesh_init(mesh)
    if state is not None:
        ctx.set_state(state)
        ctx.tracers_init(state.ntr)
        for i in range(state.ntr):
            sch = dict(hor="MFCT", ver="QR4C", lim="FCT", ph=0.0, pv=1.0)
            if schemes:
                sch.update(schemes[i] if isinstance(schemes, (list, tuple)) else schemes)
            ctx.tracer_config(i + 1, int(state.ids[i]), **sch)
            vo = state.valuesold[i] if valuesold is None else valuesold[i]
            ctx.tracer_upload(i + 1, state.values[i], vo)
    return ctx
