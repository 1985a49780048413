"""Write the reference's own test meshes as compact fixtures (data only).

Reads the ASCII mesh files that FESOM2 ships for its tests
(test/meshes/{soufflet,pi}: nod2d.out, elem2d.out, edges.out, edge_tri.out,
aux3d.out, nlvls.out, elvls.out) and stores them, unchanged, as arrays in
tests/golden/mesh_<name>.npz. Run in a container that has the reference tree:

    python tests/golden/make_mesh_fixtures.py /root/reference
"""
import os
import sys

import numpy as np


def read_mesh(d):
    nod = np.loadtxt(os.path.join(d, "nod2d.out"), skiprows=1)
    elem = np.loadtxt(os.path.join(d, "elem2d.out"), skiprows=1, dtype=np.int64)
    edges = np.loadtxt(os.path.join(d, "edges.out"), dtype=np.int64)
    edge_tri = np.loadtxt(os.path.join(d, "edge_tri.out"), dtype=np.int64)
    aux = np.loadtxt(os.path.join(d, "aux3d.out"))
    nl = int(aux[0])
    zbar = aux[1:1 + nl]
    depth = aux[1 + nl:]
    nlvls = np.loadtxt(os.path.join(d, "nlvls.out"), dtype=np.int64)
    elvls = np.loadtxt(os.path.join(d, "elvls.out"), dtype=np.int64)
    return dict(coord=nod[:, 1:3].astype(np.float64), node_flag=nod[:, 3].astype(np.int32),
                elem=elem.astype(np.int32), edges=edges.astype(np.int32), edge_tri=edge_tri.astype(np.int32),
                zbar=zbar.astype(np.float64), depth=depth.astype(np.float64),
                nlvls=nlvls.astype(np.int32), elvls=elvls.astype(np.int32))


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    for name in ("soufflet", "pi"):
        arrays = read_mesh(os.path.join(ref_root, "test", "meshes", name))
        np.savez_compressed(os.path.join(here, f"mesh_{name}.npz"), **arrays)
        print(name, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
