"""FESOM2 ASCII mesh files -> Mesh.

The reference reads a partitioned mesh from plain text files
(oce_mesh.F90 read_mesh; written by fvom_init.F90):
  nod2d.out   nod2D, then "index lon lat flag" per node (degrees)
  elem2d.out  elem2D, then the three node indices per element
  aux3d.out   nl, the nl interface depths zbar, then one depth per node
  edges.out / edge_tri.out   the two nodes / two triangles of every edge
                             (edge_tri(2) <= 0 on a boundary edge)
  elvls.out / nlvls.out      nlevels per element / per node
This module parses those files and builds the Mesh the hot path uses, with
the connectivity and levels taken from the files as they are and the
geometry computed from the (optionally rotated) coordinates.
"""
from __future__ import annotations

import os

import numpy as np

from . import meshgen as mg


def read_fesom_mesh(path):
    """Arrays of a FESOM2 mesh directory (1-based indices, as in the files)."""
    def ints(name, skip=0):
        return np.loadtxt(os.path.join(path, name), skiprows=skip, dtype=np.int64)
    nod = np.loadtxt(os.path.join(path, "nod2d.out"), skiprows=1)
    aux = np.loadtxt(os.path.join(path, "aux3d.out"))
    nl = int(aux[0])
    return dict(coord=nod[:, 1:3].astype(np.float64), node_flag=nod[:, 3].astype(np.int32),
                elem=ints("elem2d.out", 1).astype(np.int32), edges=ints("edges.out").astype(np.int32),
                edge_tri=ints("edge_tri.out").astype(np.int32), zbar=aux[1:1 + nl].astype(np.float64),
                depth=aux[1 + nl:].astype(np.float64), nlvls=ints("nlvls.out").astype(np.int32),
                elvls=ints("elvls.out").astype(np.int32))


def _rotation(alpha, beta, gamma):
    """Geographic -> rotated Cartesian frame for Euler angles (degrees): about z
    by alpha, about the new x by beta, about the new z by gamma
    (gen_modules_rotate_grid.F90:30-53)."""
    def rz(t):
        c, s = np.cos(t), np.sin(t)
        return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])

    def rx(t):
        c, s = np.cos(t), np.sin(t)
        return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])

    a, b, g = np.deg2rad([alpha, beta, gamma])
    return rz(g) @ rx(b) @ rz(a)


def geo_to_rotated(lon, lat, euler=(50.0, 15.0, -90.0)):
    """Rotated lon/lat (radians) of geographic lon/lat (radians), as the
    reference's g2r (gen_modules_rotate_grid.F90:89-120) for force_rotation."""
    M = _rotation(*euler)
    v = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])
    r = M @ v
    rlat = np.arcsin(np.clip(r[2], -1.0, 1.0))
    rlon = np.where((r[0] == 0.0) & (r[1] == 0.0), 0.0, np.arctan2(r[1], r[0]))
    return rlon, rlat


def mesh_from_arrays(a, cyclic_length_deg=360.0, cartesian=False, rotate=None):
    """Mesh from the arrays of read_fesom_mesh (or a fixture holding the same keys).
    rotate: Euler angles (degrees) of the reference's force_rotation, or None."""
    rad = np.pi / 180.0
    coord = np.asarray(a["coord"], np.float64) * rad
    if rotate is not None:
        coord = np.stack(geo_to_rotated(coord[:, 0], coord[:, 1], rotate), 1)
    elem = np.asarray(a["elem"], np.int64) - 1
    edges = np.asarray(a["edges"], np.int64) - 1
    et = np.asarray(a["edge_tri"], np.int64)
    edge_tri = np.where(et > 0, et - 1, -1)
    depth = np.asarray(a["depth"], np.float64)
    depth = np.where(depth > 0, -depth, depth)   # negative below sea level (fvom_init.F90:830-836)
    zbar = np.asarray(a["zbar"], np.float64)
    zbar = np.where(zbar > 0, -zbar, zbar)
    return mg.assemble_mesh(coord, elem, L=cyclic_length_deg * rad, cartesian=cartesian, zbar=zbar, depth=depth,
                            edges=edges, edge_tri=edge_tri, nlevels=np.asarray(a["elvls"]))


def load_fesom_mesh(path, cyclic_length_deg=360.0, cartesian=False, rotate=None):
    return mesh_from_arrays(read_fesom_mesh(path), cyclic_length_deg, cartesian, rotate)
