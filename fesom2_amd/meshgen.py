"""Seeded synthetic FESOM2 meshes and ocean states for the tracer/SSH hot path.

Everything here produces the *inputs* that FESOM2 hands to the hot path
(the ``t_mesh`` / ``t_tracer`` / ``t_dyn`` components), following the
reference's own construction rules:

* edges / edge_tri ordering and orientation: ``fvom_init.F90:500-607``
  (inner edges first, then boundary edges with ``edge_tri(2) <= 0``;
  ``edge_tri(1)`` lies to the left of ``edges(1)->edges(2)``);
* clockwise elements: ``oce_mesh.F90:1721-1742`` (``test_tri``);
* level numbers: ``fvom_init.F90:849-995`` (first ``Z(nz) < dmean``, minimum
  level count, the ">= 2 open neighbours" sweep, ``nlevels_nod2D = max``);
* ``nlevels_nod2D_min`` / ``ulevels_nod2D_max``: ``oce_mesh.F90:1675-1680``;
* areas: ``oce_mesh.F90:2223-2298`` (``mesh_areas``);
* ``edge_dxdy``, ``edge_cross_dxdy``, ``elem_cos``, ``gradient_sca``:
  ``oce_mesh.F90:2523-2618`` (``mesh_auxiliary_arrays``), with ``edge_center`` /
  ``elem_center`` of ``fvom_init.F90:1475-1518``;
* full-cell bottoms and linear free surface thicknesses:
  ``oce_ale.F90:495-501,607-613,818-864``;
* W from the discrete continuity equation: ``oce_ale.F90:2166-2306``
  (``vert_vel_ale``).

Arrays are numpy, C order with the *last* axis fastest, so ``T[n, k]`` has the
same memory image as Fortran ``T(k, n)`` (level fastest). Connectivity is
0-based here; ``-1`` marks a missing second element of a boundary edge.
Level numbers (``nlevels`` etc.) keep the reference's 1-based meaning:
a node has wet layers ``ulevels_nod2D .. nlevels_nod2D-1``.

Generated meshes are jittered triangulations of a band that is periodic in x:
rows of nodes with equilateral-like spacing, neighbouring rows zipped into
triangles. ``cartesian=True`` gives a flat channel (coordinates in metres /
r_earth, ``elem_cos = 1``); ``cartesian=False`` gives a latitude band on the
sphere whose row lengths follow cos(latitude) ("global-like").
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, Optional

import numpy as np

R_EARTH = 6367500.0          # oce_modules.F90:29
G = 9.81                     # oce_modules.F90:28
EPSILON_AB = 0.1             # oce_modules.F90:105
VCPW = 4.2e6                 # oce_modules.F90:31
PI = 3.141592653589793238462643383279502884


@dataclasses.dataclass
class Mesh:
    """Subset of ``t_mesh`` (MOD_MESH.F90:22-175) used by the hot path."""
    nod2D: int
    elem2D: int
    edge2D: int
    edge2D_in: int
    nl: int
    cartesian: bool
    cyclic_length: float
    coord_nod2D: np.ndarray          # (N2, 2) radians
    elem2D_nodes: np.ndarray         # (F2, 3) int32, 0-based, clockwise
    edges: np.ndarray                # (E2, 2) int32, 0-based
    edge_tri: np.ndarray             # (E2, 2) int32, 0-based, -1 = none
    nod_in_elem2D: np.ndarray        # (N2, kmax) int32, 0-based, ascending, -1 padded
    nod_in_elem2D_num: np.ndarray    # (N2,) int32
    elem_neighbors: np.ndarray       # (F2, 3) int32, -1 = none
    zbar: np.ndarray                 # (nl,)  interfaces, 0 .. -Hmax
    Z: np.ndarray                    # (nl-1,) mid levels
    depth: np.ndarray                # (N2,) negative below sea level
    nlevels: np.ndarray              # (F2,) 1-based level number (layers 1..nlevels-1)
    ulevels: np.ndarray              # (F2,)
    nlevels_nod2D: np.ndarray        # (N2,)
    ulevels_nod2D: np.ndarray        # (N2,)
    nlevels_nod2D_min: np.ndarray    # (N2,)
    ulevels_nod2D_max: np.ndarray    # (N2,)
    elem_area: np.ndarray            # (F2,) m^2
    area: np.ndarray                 # (N2, nl) m^2
    areasvol: np.ndarray             # (N2, nl) m^2
    elem_cos: np.ndarray             # (F2,)
    edge_dxdy: np.ndarray            # (E2, 2) radians
    edge_cross_dxdy: np.ndarray      # (E2, 4) metres
    gradient_sca: np.ndarray         # (F2, 6) 1/m
    zbar_e_bot: np.ndarray           # (F2,)
    zbar_e_srf: np.ndarray           # (F2,)
    zbar_n_bot: np.ndarray           # (N2,)
    zbar_3d_n: np.ndarray            # (N2, nl)
    Z_3d_n: np.ndarray               # (N2, nl-1)
    hnode: np.ndarray                # (N2, nl-1)
    hnode_new: np.ndarray            # (N2, nl-1)
    helem: np.ndarray                # (F2, nl-1)

    @property
    def kmax(self) -> int:
        return int(self.nod_in_elem2D.shape[1])

    @property
    def n3(self) -> int:
        """Wet 3-D nodes, sum over columns of nlevels_nod2D - ulevels_nod2D."""
        return int(np.sum(self.nlevels_nod2D - self.ulevels_nod2D))

    @property
    def e3(self) -> int:
        """Active edge-levels (the nu12..nl12 ranges of the edge scatters)."""
        nu, nl = edge_level_range(self)
        return int(np.sum(nl - nu + 1))

    @property
    def f3(self) -> int:
        return int(np.sum(self.nlevels - self.ulevels))


@dataclasses.dataclass
class State:
    """Tracer state and dynamics inputs of one step (t_tracer / t_dyn subset)."""
    ids: np.ndarray                  # (ntr,) tracer IDs (1=T, 2=S, 301=passive)
    values: np.ndarray               # (ntr, N2, nl-1)
    valuesold: np.ndarray            # (ntr, N2, nl-1)  AB2: valuesold(1,:,:)
    uv: np.ndarray                   # (F2, nl-1, 2)
    w: np.ndarray                    # (N2, nl)
    w_e: np.ndarray                  # (N2, nl)
    w_i: np.ndarray                  # (N2, nl)
    Kv: np.ndarray                   # (N2, nl)
    dt: float
    heat_flux: np.ndarray            # (N2,)
    water_flux: np.ndarray           # (N2,)
    virtual_salt: np.ndarray         # (N2,)
    relax_salt: np.ndarray           # (N2,)
    real_salt_flux: np.ndarray       # (N2,)
    use_wsplit: bool = False         # w_i carries an implicit part (dynamics%use_wsplit)
    Ki: Optional[np.ndarray] = None  # (N2, nl-1) horizontal/isoneutral diffusivity; None: Ki = 0
    slope_tapered: Optional[np.ndarray] = None  # (N2, nl-1, 3) neutral slopes, read when redi
    redi: bool = False               # Redi (oce_modules.F90:75)
    # surface fluxes of the tracer IDs 101-104 and 304 (bc_surface, oce_ale_tracer.F90:2059-2079)
    prec_rain: Optional[np.ndarray] = None      # (N2,) ID 101
    wiso_flux_oce: Optional[np.ndarray] = None  # (3, N2): Fortran wiso_flux_oce(N2, 3), IDs 102-104
    hosing_flux: Optional[np.ndarray] = None    # (N2,) ID 304
    hosing_flux3D: Optional[np.ndarray] = None  # (N2, nl-1) ID 304 when hosing_depth
    hosing_depth: bool = False                  # use_hosing .and. hosing_mode == 'depth'

    @property
    def ntr(self) -> int:
        return int(self.values.shape[0])


# ----------------------------------------------------------------------------
# helpers
# ----------------------------------------------------------------------------

def _trim_cyclic(b, L):
    """gen_modules_rotate_grid.F90:204-211."""
    b = np.where(b > L / 2.0, b - L, b)
    b = np.where(b < -L / 2.0, b + L, b)
    return b


def edge_level_range(mesh: Mesh):
    """1-based level range nu12..nl12 of the LO/FCT edge scatters
    (oce_adv_tra_driver.F90:142-156, the max/min rule)."""
    el1 = mesh.edge_tri[:, 0]
    el2 = mesh.edge_tri[:, 1]
    nl1 = mesh.nlevels[el1] - 1
    nu1 = mesh.ulevels[el1]
    has2 = el2 >= 0
    nl2 = np.where(has2, mesh.nlevels[np.maximum(el2, 0)] - 1, 0)
    nu2 = np.where(has2, mesh.ulevels[np.maximum(el2, 0)], 0)
    nl12 = np.maximum(nl1, nl2)
    nu12 = np.where(nu2 > 0, np.minimum(nu1, nu2), nu1)
    return nu12, nl12


def _zip_rows(a_idx, xa, b_idx, xb, L):
    """Zip two periodic rows of nodes (sorted by x in [0, L)) into triangles.

    Walks both rows in x and emits one triangle per advance, so a strip
    between rows of m_a and m_b nodes gives m_a + m_b triangles.
    """
    ma, mb = len(a_idx), len(b_idx)
    # rotate row b so that its first node is the one nearest a[0]
    d = _trim_cyclic(xb - xa[0], L)
    s = int(np.argmin(np.abs(d)))
    b_idx = np.roll(b_idx, -s)
    xb = np.roll(xb, -s)
    xb = xa[0] + _trim_cyclic(xb - xa[0], L)        # unwrap around xa[0]
    xb = xb[0] + np.mod(xb - xb[0], L)              # monotone from xb[0]
    xa_u = xa[0] + np.mod(xa - xa[0], L)
    # advance keys: position of the node we advance onto
    ka = np.concatenate([xa_u[1:], [xa_u[0] + L]])
    kb = np.concatenate([xb[1:], [xb[0] + L]])
    keys = np.concatenate([ka, kb])
    typ = np.concatenate([np.zeros(ma, np.int8), np.ones(mb, np.int8)])
    order = np.lexsort((typ, keys))
    typ = typ[order]
    ia = np.concatenate([[0], np.cumsum(typ == 0)])[:-1]   # state before event
    ib = np.concatenate([[0], np.cumsum(typ == 1)])[:-1]
    a_cur = a_idx[ia % ma]
    b_cur = b_idx[ib % mb]
    a_nxt = a_idx[(ia + 1) % ma]
    b_nxt = b_idx[(ib + 1) % mb]
    tri = np.where((typ == 0)[:, None],
                   np.stack([a_cur, a_nxt, b_cur], 1),
                   np.stack([a_cur, b_nxt, b_cur], 1))
    return tri


def _build_rows(row_counts, row_y, L, dx_rad, jitter, rng, cartesian):
    """Node coordinates row by row with uniform jitter of +-jitter*dx."""
    xs, ys, idx = [], [], []
    start = 0
    ny = len(row_counts)
    for j, (m, y) in enumerate(zip(row_counts, row_y)):
        sp = L / m
        off = 0.5 * sp if (j % 2) else 0.0
        x = off + sp * np.arange(m)
        # jitter in units of the local spacing (metres on the sphere)
        scale_x = sp
        jx = rng.uniform(-jitter, jitter, m) * scale_x
        if 0 < j < ny - 1:
            jy = rng.uniform(-jitter, jitter, m) * dx_rad
        else:
            jy = np.zeros(m)
        x = np.mod(x + jx, L)
        order = np.argsort(x, kind="stable")
        x = x[order]
        xs.append(x)
        ys.append(y + jy)
        idx.append(np.arange(start, start + m))
        start += m
    return xs, ys, idx


def _stretched_levels(nlayers, hmax, dz0):
    """Interfaces 0 > zbar(2) > ... > -hmax with geometrically growing dz."""
    if dz0 * nlayers >= hmax:
        dz = np.full(nlayers, hmax / nlayers)
    else:
        lo, hi = 1.0, 2.0
        for _ in range(200):
            r = 0.5 * (lo + hi)
            s = dz0 * (r ** nlayers - 1) / (r - 1)
            if s > hmax:
                hi = r
            else:
                lo = r
        r = 0.5 * (lo + hi)
        dz = dz0 * r ** np.arange(nlayers)
        dz *= hmax / dz.sum()
    zbar = np.concatenate([[0.0], -np.cumsum(dz)])
    zbar[-1] = -hmax
    return zbar


def _smooth_field(x, y, L, Ly, rng, nmodes=24, kmax=6):
    """Smooth random field, periodic in x, from a few Fourier modes."""
    f = np.zeros_like(x)
    for _ in range(nmodes):
        kx = rng.integers(1, kmax + 1)
        ky = rng.uniform(0.5, kmax)
        ph1, ph2 = rng.uniform(0, 2 * PI, 2)
        amp = rng.normal() / (kx * kx + ky * ky) ** 0.5
        f += amp * np.cos(2 * PI * kx * x / L + ph1) * np.cos(PI * ky * y / Ly + ph2)
    return f


# ----------------------------------------------------------------------------
# mesh construction
# ----------------------------------------------------------------------------

def build_mesh(*, nx: int = 60, ny: int = 50, dx: float = 20e3,
               jitter: float = 0.2, seed: int = 0, cartesian: bool = True,
               lat_range=(-70.0, 75.0), target_nodes: Optional[int] = None,
               nlayers: int = 24, hmax: float = 4000.0, dz0: Optional[float] = None,
               min_layers: int = 10, max_layers: Optional[int] = None,
               topo: str = "bumps", bump_height: float = 1500.0,
               open_neighbour_sweep: bool = True, cavity: Optional[dict] = None) -> Mesh:
    """Jittered periodic band triangulation plus levels and geometry.

    ``cartesian=True``: nx x ny rows, spacing dx (metres), walls at the first
    and last row, periodic in x with cyclic_length = nx*dx/r_earth.
    ``cartesian=False``: latitude band, rows of cos(lat)-scaled length, spacing
    dx at the equator; with ``target_nodes`` the spacing is tuned and the last
    row lengths adjusted so that exactly ``target_nodes`` nodes result.
    ``cavity=dict(frac=f, draft=d)`` puts an ice shelf over the southern
    fraction f of the band: its draft falls from d metres at the wall to zero
    at the front, with a seeded wave along x so that the first wet level
    varies along the front as well.
    """
    rng = np.random.default_rng(seed)
    if cartesian:
        L = nx * dx / R_EARTH
        dy = dx * math.sqrt(3.0) / 2.0
        row_y = np.arange(ny) * dy / R_EARTH
        row_counts = np.full(ny, nx, dtype=np.int64)
        dx_rad = dx / R_EARTH
    else:
        L = 2.0 * PI
        lat0, lat1 = np.deg2rad(lat_range[0]), np.deg2rad(lat_range[1])

        def rows_for(dxm):
            dyr = dxm * math.sqrt(3.0) / 2.0 / R_EARTH
            nrow = int(round((lat1 - lat0) / dyr)) + 1
            ys = np.linspace(lat0, lat1, nrow)
            cnt = np.maximum(8, np.round(L * np.cos(ys) * R_EARTH / dxm)).astype(np.int64)
            return ys, cnt

        if target_nodes is not None:
            lo, hi = 100.0, 2e6
            for _ in range(80):
                mid = math.sqrt(lo * hi)
                ys, cnt = rows_for(mid)
                if cnt.sum() > target_nodes:
                    lo = mid
                else:
                    hi = mid
            dx = hi
            row_y, row_counts = rows_for(dx)
            diff = int(target_nodes - row_counts.sum())
            # distribute the remainder over the longest rows
            order = np.argsort(-row_counts)
            i = 0
            while diff != 0:
                step = 1 if diff > 0 else -1
                row_counts[order[i % len(order)]] += step
                diff -= step
                i += 1
        else:
            row_y, row_counts = rows_for(dx)
        dx_rad = dx / R_EARTH
    xs, ys, idx = _build_rows(row_counts, row_y, L, dx_rad, jitter, rng, cartesian)
    coord = np.stack([np.concatenate(xs), np.concatenate(ys)], 1)
    N2 = coord.shape[0]

    tris = [_zip_rows(idx[j], xs[j], idx[j + 1], xs[j + 1], L) for j in range(len(idx) - 1)]
    elem = np.concatenate(tris).astype(np.int64)

    # --- vertical grid and topography ---------------------------------------
    nl = nlayers + 1
    if dz0 is None:
        zbar = _stretched_levels(nlayers, hmax, hmax / nlayers)
    else:
        zbar = _stretched_levels(nlayers, hmax, dz0)
    Zmid = 0.5 * (zbar[:-1] + zbar[1:])
    Ly = coord[:, 1].max() - coord[:, 1].min()
    y0 = coord[:, 1].min()
    if topo == "bumps":
        depth = np.full(N2, hmax)
        nb = 8
        for _ in range(nb):
            cxb = rng.uniform(0, L)
            cyb = rng.uniform(y0, y0 + Ly)
            rad = rng.uniform(0.08, 0.2) * min(L, Ly if Ly > 0 else L)
            hgt = rng.uniform(0.3, 1.0) * bump_height
            ddx = _trim_cyclic(coord[:, 0] - cxb, L) * (np.cos(coord[:, 1]) if not cartesian else 1.0)
            ddy = coord[:, 1] - cyb
            depth -= hgt * np.exp(-(ddx * ddx + ddy * ddy) / (2 * rad * rad))
        depth = np.maximum(depth, hmax - bump_height)
    elif topo == "flat":
        depth = np.full(N2, hmax)
    else:  # "shelf": wide range of depths down to a few layers
        f = _smooth_field(coord[:, 0], coord[:, 1] - y0, L, max(Ly, 1e-12), rng)
        f = (f - f.min()) / max(f.max() - f.min(), 1e-30)
        depth = 30.0 + (hmax - 30.0) * f ** 0.6
    depth = -depth   # negative below sea level (aux3d.out convention)
    cavity_depth = None
    if cavity is not None:
        frac = float(cavity.get("frac", 0.3))
        dmax = float(cavity.get("draft", 600.0))
        yr = (coord[:, 1] - y0) / max(Ly, 1e-30)
        ph = rng.uniform(0.0, 2.0 * PI)
        front = frac * (1.0 + 0.3 * np.sin(2.0 * PI * coord[:, 0] / L * 3.0 + ph))
        cavity_depth = -dmax * np.clip(1.0 - yr / front, 0.0, 1.0)

    return assemble_mesh(coord, elem, L=L, cartesian=cartesian, zbar=zbar, depth=depth, min_layers=min_layers,
                         max_layers=max_layers, open_neighbour_sweep=open_neighbour_sweep,
                         cavity_depth=cavity_depth)


def build_edges(coord, elem, L):
    """Edges and their two triangles from a triangulation (fvom_init.F90:500-607).

    Returns 0-based (edges, edge_tri, edge2D_in): inner edges first, then
    boundary edges (edge_tri(2) = -1), each inner/boundary group ordered by its
    node pair; edge_tri(1) lies to the left of edges(1) -> edges(2), and on a
    boundary edge the nodes are swapped until it does.
    """
    N2 = coord.shape[0]
    F2 = elem.shape[0]
    he_a = np.concatenate([elem[:, 0], elem[:, 1], elem[:, 2]])
    he_b = np.concatenate([elem[:, 1], elem[:, 2], elem[:, 0]])
    he_e = np.concatenate([np.arange(F2)] * 3)
    lo_n = np.minimum(he_a, he_b)
    hi_n = np.maximum(he_a, he_b)
    key = lo_n * N2 + hi_n
    order = np.lexsort((he_e, key))
    key_s = key[order]
    e_s = he_e[order]
    uniq, first, counts = np.unique(key_s, return_index=True, return_counts=True)
    assert counts.max() <= 2
    el1 = e_s[first]
    el2 = np.where(counts == 2, e_s[np.minimum(first + 1, len(e_s) - 1)], -1)
    n1 = (uniq // N2).astype(np.int64)
    n2 = (uniq % N2).astype(np.int64)
    inner = counts == 2
    perm = np.concatenate([np.nonzero(inner)[0], np.nonzero(~inner)[0]])
    n1, n2, el1, el2 = n1[perm], n2[perm], el1[perm], el2[perm]
    E2_in = int(inner.sum())

    def elem_center(ei):
        ax = coord[elem[ei], 0]
        amin = ax.min(axis=1, keepdims=True)
        ax = np.where(ax - amin > L / 2.0, ax - L, ax)
        x = (ax[:, 0] + ax[:, 1] + ax[:, 2]) / 3.0
        yy = coord[elem[ei], 1]
        y = (yy[:, 0] + yy[:, 1] + yy[:, 2]) / 3.0
        return x, y

    # orientation: edge_tri(1) to the left of n1 -> n2 (fvom_init.F90:579-607)
    cx, cy = elem_center(el1)
    xc0 = _trim_cyclic(cx - coord[n1, 0], L)
    xc1 = cy - coord[n1, 1]
    xe0 = _trim_cyclic(coord[n2, 0] - coord[n1, 0], L)
    xe1 = coord[n2, 1] - coord[n1, 1]
    right = xc0 * xe1 - xc1 * xe0 > 0.0
    swap_el = right & (el2 >= 0)
    swap_nd = right & (el2 < 0)
    el1[swap_el], el2[swap_el] = el2[swap_el].copy(), el1[swap_el].copy()
    n1[swap_nd], n2[swap_nd] = n2[swap_nd].copy(), n1[swap_nd].copy()
    edges = np.stack([n1, n2], 1).astype(np.int32)
    edge_tri = np.stack([el1, el2], 1).astype(np.int32)

    return edges, edge_tri, E2_in


def assemble_mesh(coord, elem, *, L, cartesian, zbar, depth, min_layers=10, max_layers=None,
                  open_neighbour_sweep=True, edges=None, edge_tri=None, nlevels=None,
                  cavity_depth=None) -> Mesh:
    """Connectivity, levels and geometry of a triangulation.

    coord (N2, 2) in radians, elem (F2, 3) 0-based, L the cyclic length in
    radians, zbar (nl,) the interfaces, depth (N2,) negative below sea level.
    edges/edge_tri (0-based, -1 = none) and nlevels (1-based per element) may be
    given, as the reference's mesh files hold them; otherwise they are built.
    cavity_depth (N2,), negative under an ice shelf and 0 elsewhere, sets the
    first wet level of each element (see ``cavity_levels``).
    """
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    elem = np.array(elem, dtype=np.int64)
    N2 = coord.shape[0]
    # clockwise orientation (oce_mesh.F90:1721-1742 test_tri)
    a = coord[elem[:, 0]]
    b = _trim_cyclic(coord[elem[:, 1], 0] - a[:, 0], L), coord[elem[:, 1], 1] - a[:, 1]
    c = _trim_cyclic(coord[elem[:, 2], 0] - a[:, 0], L), coord[elem[:, 2], 1] - a[:, 1]
    r = b[0] * c[1] - b[1] * c[0]
    sw = r > 0
    elem[sw, 1], elem[sw, 2] = elem[sw, 2].copy(), elem[sw, 1].copy()
    F2 = elem.shape[0]

    if edges is None:
        edges, edge_tri, E2_in = build_edges(coord, elem, L)
    else:
        edges = np.ascontiguousarray(edges, dtype=np.int32)
        edge_tri = np.ascontiguousarray(edge_tri, dtype=np.int32)
        E2_in = int((edge_tri[:, 1] >= 0).sum())
    n1 = edges[:, 0].astype(np.int64)
    n2 = edges[:, 1].astype(np.int64)
    el1 = edge_tri[:, 0].astype(np.int64)
    el2 = edge_tri[:, 1].astype(np.int64)
    E2 = len(n1)

    def elem_center(ei):
        ax = coord[elem[ei], 0]
        amin = ax.min(axis=1, keepdims=True)
        ax = np.where(ax - amin > L / 2.0, ax - L, ax)
        x = (ax[:, 0] + ax[:, 1] + ax[:, 2]) / 3.0
        yy = coord[elem[ei], 1]
        y = (yy[:, 0] + yy[:, 1] + yy[:, 2]) / 3.0
        return x, y

    # element neighbours across edges
    elem_nb = -np.ones((F2, 3), np.int64)
    fill = np.zeros(F2, np.int64)
    for (ea, eb) in ((el1, el2), (el2, el1)):
        m = (ea >= 0) & (eb >= 0)
        ea_m, eb_m = ea[m], eb[m]
        o = np.argsort(ea_m, kind="stable")
        ea_m, eb_m = ea_m[o], eb_m[o]
        # position within each element's list
        starts = np.searchsorted(ea_m, ea_m, side="left")
        pos = np.arange(len(ea_m)) - starts + fill[ea_m]
        elem_nb[ea_m, pos] = eb_m
        np.add.at(fill, ea_m, 1)

    # nod_in_elem2D: ascending elements per node
    ne_node = elem.reshape(-1)
    ne_elem = np.repeat(np.arange(F2), 3)
    o = np.lexsort((ne_elem, ne_node))
    ne_node, ne_elem = ne_node[o], ne_elem[o]
    nnum = np.bincount(ne_node, minlength=N2)
    kmax = int(nnum.max())
    starts = np.concatenate([[0], np.cumsum(nnum)[:-1]])
    pos = np.arange(len(ne_node)) - starts[ne_node]
    nie = -np.ones((N2, kmax), np.int64)
    nie[ne_node, pos] = ne_elem

    nl = len(zbar)
    Zmid = 0.5 * (zbar[:-1] + zbar[1:])
    if nlevels is None:
        dmean = (depth[elem[:, 0]] + depth[elem[:, 1]] + depth[elem[:, 2]]) / 3.0
        thers = min_layers + 1
        deeper = Zmid[None, :] < dmean[:, None]
        has = deeper.any(axis=1)
        nlev = np.where(has, np.argmax(deeper, axis=1) + 1, nl)
        nlev = np.where(dmean >= 0, thers, nlev)
        nlev = np.maximum(nlev, thers)
        if max_layers is not None:
            nlev = np.minimum(nlev, max_layers + 1)
        if open_neighbour_sweep:
            # fvom_init.F90:888-930, swept in parallel instead of in element order
            for nz in range(thers + 1, nl + 1):
                for _ in range(1000):
                    nbl = np.where(elem_nb >= 0, nlev[np.maximum(elem_nb, 0)], 0)
                    open_cnt = ((elem_nb >= 0) & (nbl >= nz)).sum(axis=1)
                    bad = (nlev >= nz) & (open_cnt < 2)
                    if not bad.any():
                        break
                    nlev = np.where(bad, nz - 1, nlev)
        nlevels = nlev.astype(np.int32)
    else:
        nlevels = np.ascontiguousarray(nlevels, dtype=np.int32)
    nln = np.zeros(N2, np.int64)
    np.maximum.at(nln, elem.reshape(-1), np.repeat(nlevels, 3))
    nlevels_nod2D = nln.astype(np.int32)
    nlmin = np.full(N2, 10 ** 9, np.int64)
    np.minimum.at(nlmin, elem.reshape(-1), np.repeat(nlevels, 3))
    nlevels_nod2D_min = nlmin.astype(np.int32)
    if cavity_depth is None:
        ulevels = np.ones(F2, np.int32)
        ulevels_nod2D = np.ones(N2, np.int32)
        ulevels_nod2D_max = np.ones(N2, np.int32)
    else:
        ulevels, ulevels_nod2D, ulevels_nod2D_max = cavity_levels(
            elem, elem_nb, nie, zbar, nlevels, nlevels_nod2D, np.asarray(cavity_depth, np.float64))

    # --- horizontal geometry ------------------------------------------------
    # mesh_areas (oce_mesh.F90:2223-2236): areas in rad^2, scaled later
    ay = (coord[elem[:, 0], 1] + coord[elem[:, 1], 1] + coord[elem[:, 2], 1]) / 3.0
    ay = np.ones_like(ay) if cartesian else np.cos(ay)
    a0 = _trim_cyclic(coord[elem[:, 1], 0] - coord[elem[:, 0], 0], L) * ay
    a1 = coord[elem[:, 1], 1] - coord[elem[:, 0], 1]
    b0 = _trim_cyclic(coord[elem[:, 2], 0] - coord[elem[:, 0], 0], L) * ay
    b1 = coord[elem[:, 2], 1] - coord[elem[:, 0], 1]
    elem_area_rad = 0.5 * np.abs(a0 * b1 - b0 * a1)

    # area(nz,n): sum over nod_in_elem2D in ascending order (oce_mesh.F90:2279-2298)
    area_rad = np.zeros((N2, nl))
    lvl = np.arange(1, nl + 1)
    for k in range(kmax):
        e = nie[:, k]
        ok = e >= 0
        ee = np.maximum(e, 0)
        mask = ok[:, None] & (lvl[None, :] >= ulevels[ee][:, None]) & (lvl[None, :] <= (nlevels[ee] - 1)[:, None])
        area_rad = area_rad + np.where(mask, (elem_area_rad[ee] / 3.0)[:, None], 0.0)
    areasvol_rad = np.zeros((N2, nl))
    wet_nodes = (lvl[None, :] >= ulevels_nod2D[:, None]) & (lvl[None, :] <= (nlevels_nod2D - 1)[:, None])
    areasvol_rad[wet_nodes] = area_rad[wet_nodes]
    elem_area = elem_area_rad * R_EARTH * R_EARTH
    area = area_rad * R_EARTH * R_EARTH
    areasvol = areasvol_rad * R_EARTH * R_EARTH

    # mesh_auxiliary_arrays (oce_mesh.F90:2523-2618)
    ecx, ecy = elem_center(np.arange(F2))
    elem_cos = np.ones(F2) if cartesian else np.cos(ecy)
    edge_dxdy = np.stack([_trim_cyclic(coord[n2, 0] - coord[n1, 0], L),
                          coord[n2, 1] - coord[n1, 1]], 1)
    # edge_center (fvom_init.F90:1475-1493)
    ax_ = coord[edges[:, 0], 0].copy()
    bx_ = coord[edges[:, 1], 0].copy()
    m1 = ax_ - bx_ > L / 2.0
    ax_[m1] -= L
    m2 = ax_ - bx_ < -L / 2.0
    bx_[m2] -= L
    ecen_x = 0.5 * (ax_ + bx_)
    ecen_y = 0.5 * (coord[edges[:, 0], 1] + coord[edges[:, 1], 1])
    ecd = np.zeros((E2, 4))
    for i in range(2):
        el = edge_tri[:, i]
        ok = el >= 0
        ee = np.maximum(el, 0)
        bx = _trim_cyclic(ecx[ee] - ecen_x, L) * elem_cos[ee] * R_EARTH
        by = (ecy[ee] - ecen_y) * R_EARTH
        ecd[:, 2 * i] = np.where(ok, bx, 0.0)
        ecd[:, 2 * i + 1] = np.where(ok, by, 0.0)
    dX31 = elem_cos * _trim_cyclic(coord[elem[:, 2], 0] - coord[elem[:, 0], 0], L)
    dX21 = elem_cos * _trim_cyclic(coord[elem[:, 1], 0] - coord[elem[:, 0], 0], L)
    dY31 = coord[elem[:, 2], 1] - coord[elem[:, 0], 1]
    dY21 = coord[elem[:, 1], 1] - coord[elem[:, 0], 1]
    dfac = -0.5 * R_EARTH / elem_area
    gsca = np.stack([(-dY31 + dY21) * dfac, dY31 * dfac, -dY21 * dfac,
                     (dX31 - dX21) * dfac, -dX31 * dfac, dX21 * dfac], 1)

    # --- thicknesses (linfs, full cells) ------------------------------------
    zbar_e_bot = zbar[nlevels - 1]
    zbar_e_srf = zbar[ulevels - 1]
    zbar_n_bot = zbar[nlevels_nod2D - 1]
    zbar_3d_n = np.zeros((N2, nl))
    wet_if = lvl[None, :] <= nlevels_nod2D[:, None]
    zbar_3d_n[wet_if] = np.broadcast_to(zbar[None, :], (N2, nl))[wet_if]
    # oce_ale.F90:340-373: Z and zbar above the cavity too, zero below the bottom
    Z_3d_n = np.zeros((N2, nl - 1))
    wet_l = wet_nodes[:, : nl - 1]
    to_bot = lvl[None, : nl - 1] <= (nlevels_nod2D - 1)[:, None]
    Z_3d_n[to_bot] = np.broadcast_to(Zmid[None, :], (N2, nl - 1))[to_bot]
    dzl = zbar[:-1] - zbar[1:]
    hnode = np.where(wet_l, dzl[None, :], 0.0)
    lvl_l = np.arange(1, nl)
    wet_e = (lvl_l[None, :] >= ulevels[:, None]) & (lvl_l[None, :] <= (nlevels - 1)[:, None])
    helem = np.where(wet_e, dzl[None, :], 0.0)

    return Mesh(
        nod2D=N2, elem2D=F2, edge2D=E2, edge2D_in=E2_in, nl=nl, cartesian=cartesian,
        cyclic_length=float(L), coord_nod2D=np.ascontiguousarray(coord),
        elem2D_nodes=elem.astype(np.int32), edges=edges, edge_tri=edge_tri,
        nod_in_elem2D=nie.astype(np.int32), nod_in_elem2D_num=nnum.astype(np.int32),
        elem_neighbors=elem_nb.astype(np.int32), zbar=zbar, Z=Zmid, depth=depth,
        nlevels=nlevels, ulevels=ulevels, nlevels_nod2D=nlevels_nod2D,
        ulevels_nod2D=ulevels_nod2D, nlevels_nod2D_min=nlevels_nod2D_min,
        ulevels_nod2D_max=ulevels_nod2D_max, elem_area=elem_area, area=area,
        areasvol=areasvol, elem_cos=elem_cos, edge_dxdy=np.ascontiguousarray(edge_dxdy),
        edge_cross_dxdy=ecd, gradient_sca=np.ascontiguousarray(gsca),
        zbar_e_bot=zbar_e_bot, zbar_e_srf=zbar_e_srf, zbar_n_bot=zbar_n_bot,
        zbar_3d_n=zbar_3d_n, Z_3d_n=Z_3d_n, hnode=hnode, hnode_new=hnode.copy(),
        helem=helem)


def cavity_levels(elem, elem_nb, nie, zbar, nlevels, nlevels_nod2D, cavity_depth):
    """First wet level under an ice shelf: ulevels, ulevels_nod2D, ulevels_nod2D_max.

    Per element (fvom_init.F90:1062-1091): the mean cavity depth of the three
    nodes; ulevels = 2 under any draft, else 1, and the first mid level deeper
    than the draft or with at most three levels left above the bottom wins.
    Then, level by level, a wet prism with fewer than two wet neighbours is
    pushed one level down (fvom_init.F90:1127-1220, swept in parallel instead
    of in element order; the branch that would leave fewer than three layers
    is never taken here and raises instead). A node layer between the deepest
    top and the shallowest bottom of its elements that fewer than two elements
    cover opens the nearest covering tops (fvom_init.F90:1312-1408).
    ulevels_nod2D is the minimum over the node's elements (:1230-1240),
    ulevels_nod2D_max the maximum (oce_mesh.F90:1675-1680); the checks of
    fvom_init.F90:1243-1274 raise.
    """
    F2 = elem.shape[0]
    N2 = nie.shape[0]
    nl = len(zbar)
    Zmid = 0.5 * (zbar[:-1] + zbar[1:])
    dmean = (cavity_depth[elem[:, 0]] + cavity_depth[elem[:, 1]] + cavity_depth[elem[:, 2]]) / 3.0
    nlev = nlevels.astype(np.int64)
    nzs = np.arange(1, nl)
    hit = (Zmid[None, :] < dmean[:, None]) | (nlev[:, None] - nzs[None, :] <= 3)
    hit &= nzs[None, :] <= (nlev - 1)[:, None]
    ul = np.where(hit.any(axis=1), np.argmax(hit, axis=1) + 1, np.where(dmean < 0.0, 2, 1)).astype(np.int64)
    for _outer in range(10):
        for nz in range(1, int(ul.max()) + 1):
            for _ in range(1000):
                ok = elem_nb >= 0
                nb = np.maximum(elem_nb, 0)
                nb_open = ok & (ul[nb] <= nz) & (nlev[nb] > nz)
                wet = (nz >= ul) & (nz < nlev)
                iso = wet & (nb_open.sum(axis=1) <= 1)
                if not iso.any():
                    break
                if np.any(iso & (nlev - (nz + 1) < 3)):
                    raise ValueError("cavity: an isolated prism would keep fewer than three layers")
                ul = np.where(iso, nz + 1, ul)
        changed = False
        nmin = np.full(N2, nl, np.int64)
        umax = np.zeros(N2, np.int64)
        np.minimum.at(nmin, elem.reshape(-1), np.repeat(nlev, 3))
        np.maximum.at(umax, elem.reshape(-1), np.repeat(ul, 3))
        for n in np.nonzero(nmin < umax)[0]:   # only where the elements leave a gap
            els = nie[n][nie[n] >= 0]
            cnt = np.zeros(nl + 1, np.int64)
            for e in els:
                cnt[ul[e]:nlev[e]] += 1
            for nz in range(int(ul[els].min()), int(nlev[els].max())):
                if cnt[nz] <= 1:
                    above = [(int(ul[e]) - nz, j) for j, e in enumerate(els) if ul[e] > nz]
                    above.sort()
                    for _, j in above[: 2 - int(cnt[nz])]:
                        ul[els[j]] = nz
                    changed = True
                    break
        if not changed:
            break
    else:
        raise ValueError("cavity: levels did not converge")
    uln = np.full(N2, nl, np.int64)
    np.minimum.at(uln, elem.reshape(-1), np.repeat(ul, 3))
    ulx = np.zeros(N2, np.int64)
    np.maximum.at(ulx, elem.reshape(-1), np.repeat(ul, 3))
    if np.any(nlev - ul < 3) or np.any(nlevels_nod2D - uln < 3):
        raise ValueError("cavity: fewer than three wet layers")
    return ul.astype(np.int32), uln.astype(np.int32), ulx.astype(np.int32)


# ----------------------------------------------------------------------------
# state
# ----------------------------------------------------------------------------

def vert_vel_ale(mesh: Mesh, uv: np.ndarray) -> np.ndarray:
    """W from the discrete continuity equation (oce_ale.F90:2166-2306).

    Edge transports through the two dual-cell faces of each edge are summed
    into the nodes (in edge order), accumulated from the bottom up and
    divided by the interface area.
    """
    N2, nl = mesh.nod2D, mesh.nl
    W = np.zeros((N2, nl))
    e1 = mesh.edge_tri[:, 0]
    e2 = mesh.edge_tri[:, 1]
    lvl = np.arange(1, nl)
    for i, el in enumerate((e1, e2)):
        ok = el >= 0
        ee = np.maximum(el, 0)
        dX = mesh.edge_cross_dxdy[:, 2 * i]
        dY = mesh.edge_cross_dxdy[:, 2 * i + 1]
        c1 = (uv[ee, :, 1] * dX[:, None] - uv[ee, :, 0] * dY[:, None]) * mesh.helem[ee]
        if i == 1:
            c1 = -c1
        mask = ok[:, None] & (lvl[None, :] >= mesh.ulevels[ee][:, None]) & (lvl[None, :] <= (mesh.nlevels[ee] - 1)[:, None])
        c1 = np.where(mask, c1, 0.0)
        # sequential-in-edge-order accumulation is not needed for a generator
        np.add.at(W[:, : nl - 1], mesh.edges[:, 0], c1)
        np.add.at(W[:, : nl - 1], mesh.edges[:, 1], -c1)
    # cumulative sum from the bottom (oce_ale.F90:2282)
    for n_lvl in range(nl - 2, -1, -1):
        wet = (n_lvl + 1) <= (mesh.nlevels_nod2D - 1)
        W[:, n_lvl] = np.where(wet, W[:, n_lvl] + W[:, n_lvl + 1], W[:, n_lvl])
    wet_if = (np.arange(1, nl + 1)[None, :] >= mesh.ulevels_nod2D[:, None]) & \
             (np.arange(1, nl + 1)[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])
    W = np.where(wet_if, W / np.where(mesh.area > 0, mesh.area, 1.0), 0.0)
    return W


def build_state(mesh: Mesh, *, seed: int = 0, ntr: int = 2, dt: float = 900.0,
                umax: float = 0.5, kv: float = 1e-5, kv_range=None,
                zero_surface_w: bool = True, passive_radius_cells: float = 5.0,
                dx: Optional[float] = None) -> State:
    """Seeded T/S (+ passive blobs), divergence-free velocity, W, Kv.

    T = 20 - 15 d/H + N(0, 0.1) and S = 35 + N(0, 0.05) at wet 3-D nodes (d the
    depth of the mid level, H the deepest interface). Passive tracers (ID 301)
    are Gaussian blobs of amplitude 1 and radius ``passive_radius_cells`` mesh
    spacings. The horizontal velocity is u = rot(psi) of a seeded random
    streamfunction that vanishes on both walls, given a per-element vertical
    profile s_k with sum_k s_k helem_k = 1, so the column transport is rot(psi)
    and the surface W is zero up to round-off (then set to zero exactly).
    """
    rng = np.random.default_rng(seed + 1000)
    N2, F2, nl = mesh.nod2D, mesh.elem2D, mesh.nl
    L = mesh.cyclic_length
    H = -mesh.zbar[-1]
    wet = (np.arange(1, nl)[None, :] >= mesh.ulevels_nod2D[:, None]) & \
          (np.arange(1, nl)[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])
    d = -mesh.Z[None, :]
    ids = [1, 2] + [301] * max(0, ntr - 2)
    ids = np.array(ids[:ntr], np.int32)
    vals = np.zeros((ntr, N2, nl - 1))
    for t in range(ntr):
        if ids[t] == 1:
            v = 20.0 - 15.0 * d / H + rng.normal(0.0, 0.1, (N2, nl - 1))
        elif ids[t] == 2:
            v = 35.0 + rng.normal(0.0, 0.05, (N2, nl - 1))
        else:
            cx = rng.uniform(0, L)
            ymin, ymax = mesh.coord_nod2D[:, 1].min(), mesh.coord_nod2D[:, 1].max()
            cy = rng.uniform(ymin, ymax)
            cz = rng.uniform(0.1, 0.7) * H
            spacing = (dx / R_EARTH) if dx is not None else math.sqrt(
                np.mean(mesh.elem_area) * 2.0 / math.sqrt(3.0) * 2.0) / R_EARTH
            rad = passive_radius_cells * spacing
            cosl = 1.0 if mesh.cartesian else np.cos(mesh.coord_nod2D[:, 1])
            ddx = _trim_cyclic(mesh.coord_nod2D[:, 0] - cx, L) * cosl
            ddy = mesh.coord_nod2D[:, 1] - cy
            r2 = (ddx * ddx + ddy * ddy) / (rad * rad)
            vz = np.exp(-((d - cz) / (0.2 * H)) ** 2)
            v = np.exp(-0.5 * r2)[:, None] * vz
        vals[t] = np.where(wet, v, 0.0)

    # streamfunction at nodes, zero on both walls
    y = mesh.coord_nod2D[:, 1]
    ymin, ymax = y.min(), y.max()
    s = (y - ymin) / max(ymax - ymin, 1e-30)
    psi = _smooth_field(mesh.coord_nod2D[:, 0], y - ymin, L, max(ymax - ymin, 1e-30), rng, nmodes=16, kmax=4)
    psi = psi * np.sin(PI * s) ** 2
    # element transport = rot psi with gradient_sca: u = -dpsi/dy, v = dpsi/dx
    en = mesh.elem2D_nodes
    gx = (mesh.gradient_sca[:, 0] * psi[en[:, 0]] + mesh.gradient_sca[:, 1] * psi[en[:, 1]]
          + mesh.gradient_sca[:, 2] * psi[en[:, 2]])
    gy = (mesh.gradient_sca[:, 3] * psi[en[:, 0]] + mesh.gradient_sca[:, 4] * psi[en[:, 1]]
          + mesh.gradient_sca[:, 5] * psi[en[:, 2]])
    U = -gy
    V = gx
    # vertical profile with sum_k s_k helem_k = 1
    zmid = -mesh.Z
    prof = (1.0 + 0.6 * np.cos(PI * zmid / H))[None, :] * np.ones((F2, 1))
    wet_e = (np.arange(1, nl)[None, :] >= mesh.ulevels[:, None]) & \
            (np.arange(1, nl)[None, :] <= (mesh.nlevels - 1)[:, None])
    prof = np.where(wet_e, prof, 0.0)
    norm = (prof * mesh.helem).sum(axis=1)
    sk = prof / norm[:, None]
    uv = np.zeros((F2, nl - 1, 2))
    uv[:, :, 0] = U[:, None] * sk
    uv[:, :, 1] = V[:, None] * sk
    vmax = np.max(np.hypot(uv[:, :, 0], uv[:, :, 1]))
    if vmax > 0:
        uv *= umax / vmax
    W = vert_vel_ale(mesh, uv)
    if zero_surface_w:
        W[np.arange(N2), mesh.ulevels_nod2D - 1] = 0.0
    if kv_range is not None:
        lo, hi = np.log10(kv_range[0]), np.log10(kv_range[1])
        Kv = 10.0 ** rng.uniform(lo, hi, (N2, nl))
    else:
        Kv = np.full((N2, nl), kv)
    z = np.zeros(N2)
    return State(ids=ids, values=vals, valuesold=vals.copy(), uv=uv, w=W, w_e=W.copy(),
                 w_i=np.zeros_like(W), Kv=Kv, dt=float(dt), heat_flux=z.copy(),
                 water_flux=z.copy(), virtual_salt=z.copy(), relax_salt=z.copy(),
                 real_salt_flux=z.copy())


def ssh_rhs(mesh: Mesh, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed + 2000)
    return rng.normal(0.0, 1.0, mesh.nod2D)


# ----------------------------------------------------------------------------
# the named workloads
# ----------------------------------------------------------------------------

CONFIGS: Dict[str, dict] = {
    # 60x50 periodic channel, dx=20 km, 24 layers, T/S, dt=900 s, 10 steps
    "c0": dict(mesh=dict(nx=60, ny=50, dx=20e3, jitter=0.2, seed=0, cartesian=True,
                         nlayers=24, hmax=4000.0, min_layers=10, max_layers=24,
                         topo="bumps", bump_height=1500.0),
               state=dict(ntr=2, dt=900.0, kv=1e-5), steps=10),
    # c0 under an ice shelf over the southern third: first wet levels up to 5
    "c0cav": dict(mesh=dict(nx=60, ny=50, dx=20e3, jitter=0.2, seed=0, cartesian=True,
                            nlayers=24, hmax=4000.0, min_layers=10, max_layers=24,
                            topo="bumps", bump_height=1500.0,
                            cavity=dict(frac=0.33, draft=800.0)),
                  state=dict(ntr=2, dt=900.0, kv=1e-5), steps=10),
    # 40x30 channel with 70 layers (8-70 per column, stretched from 5 m): the
    # 128-lane kernels of the fused step on a mesh the oracle runs in seconds
    "c0deep": dict(mesh=dict(nx=40, ny=30, dx=20e3, jitter=0.2, seed=9, cartesian=True,
                             nlayers=70, hmax=6250.0, dz0=5.0, min_layers=5, max_layers=70,
                             topo="bumps", bump_height=6200.0),
                   state=dict(ntr=12, dt=600.0, kv_range=(1e-5, 1e-2)), steps=2),
    # 30x24 channel with 90 layers: the last kernel's long LDS rows (columns of
    # more than 72 interfaces)
    "c0deep90": dict(mesh=dict(nx=30, ny=24, dx=20e3, jitter=0.2, seed=11, cartesian=True,
                               nlayers=90, hmax=6250.0, dz0=5.0, min_layers=5, max_layers=90,
                               topo="bumps", bump_height=6200.0),
                     state=dict(ntr=12, dt=600.0, kv_range=(1e-5, 1e-2)), steps=2),
    # 126,858-node latitude band, 47 layers (PI-like stretching), T/S, dt=1800 s
    "c1": dict(mesh=dict(cartesian=False, target_nodes=126858, jitter=0.2, seed=1,
                         nlayers=47, hmax=6250.0, dz0=10.0, min_layers=5,
                         max_layers=47, topo="shelf"),
               state=dict(ntr=2, dt=1800.0, kv=1e-5), steps=100),
    # c1 mesh with 34 tracers and log-uniform Kv
    "c2": dict(mesh="c1", state=dict(ntr=34, dt=1800.0, kv_range=(1e-5, 1e-2)), steps=100),
    # 1.3 M nodes, 70 layers, 10 tracers, dt=600 s
    "c4": dict(mesh=dict(cartesian=False, target_nodes=1300000, jitter=0.2, seed=4,
                         nlayers=70, hmax=6250.0, dz0=5.0, min_layers=5,
                         max_layers=70, topo="shelf"),
               state=dict(ntr=10, dt=600.0, kv_range=(1e-5, 1e-2)), steps=50),
}


def make_config(name: str, *, ntr: Optional[int] = None, seed: int = 0):
    cfg = CONFIGS[name]
    mcfg = cfg["mesh"]
    if isinstance(mcfg, str):
        mcfg = CONFIGS[mcfg]["mesh"]
    mesh = build_mesh(**mcfg)
    st = dict(cfg["state"])
    if ntr is not None:
        st["ntr"] = ntr
    dx = mcfg.get("dx") if mcfg.get("cartesian", True) else None
    state = build_state(mesh, seed=seed, dx=dx, **st)
    return mesh, state, cfg["steps"]
