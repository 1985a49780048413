"""The OpenMP build of the oracle (the multi-core CPU baseline of bench.py)
against the serial build: column, edge and element loops run in parallel and
the edge->node scatters become per-node gathers over ascending edges, so the
results must be bit-identical. Checked on the channel, the cavity channel and
the 70-layer channel, with the default schemes and with a mix."""
import dataclasses

import numpy as np
import pytest

from fesom2_amd import meshgen as mg
from oracle.oracle import Oracle


def _wet(mesh):
    k = np.arange(1, mesh.nl)
    return (k[None, :] >= mesh.ulevels_nod2D[:, None]) & (k[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])


def _steps(mesh, st, omp, schemes=None, nsteps=2):
    o = Oracle(mesh, omp=omp)
    v, vo = st.values.copy(), st.valuesold.copy()
    for _ in range(nsteps):
        o.step(st, v, vo, schemes=schemes)
    return v, vo


@pytest.mark.parametrize("name", ["c0", "c0cav", "c0deep"])
def test_omp_build_is_bitwise(name):
    mesh, st, _ = mg.make_config(name, ntr=4)
    rng = np.random.default_rng(3)
    st = dataclasses.replace(st, heat_flux=rng.normal(0.0, 40.0, mesh.nod2D),
                             virtual_salt=rng.normal(0.0, 1e-6, mesh.nod2D))
    w = _wet(mesh)
    for schemes in (None, [dict(), dict(), dict(hor="UPW1", ver="UPW1", lim="FCT"),
                           dict(hor="MUSCL", ver="QR4C", lim="NON", ph=0.5)]):
        a, ao = _steps(mesh, st, False, schemes)
        b, bo = _steps(mesh, st, True, schemes)
        for i in range(st.ntr):
            np.testing.assert_array_equal(a[i][w], b[i][w])
            np.testing.assert_array_equal(ao[i][w], bo[i][w])


def test_omp_build_work_arrays(c0):
    """Every work array of do_oce_adv_tra, per stage."""
    mesh, st, _ = c0
    outs = []
    for omp in (False, True):
        o = Oracle(mesh, omp=omp)
        v, vo = st.values[0].copy(), st.valuesold[0].copy()
        outs.append(o.adv_stage(st, v, vo, ID=1))
    for k in outs[0]:
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)
