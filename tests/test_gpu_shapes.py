"""Every launch shape of the fused step, and tracers far from zero, against
the oracle (bit for bit).

The autotuner picks tracers per block (1, 2, 4, 8 or all), tracers per Thomas
chunk (2, 4, 5 or 8) and columns per block of the last kernel (1, 4 or 8 for
columns of at most 64 levels; 1 to 4 for deeper ones of at most 72
interfaces, whose LDS rows are short; 1 or 2 beyond). Each choice
instantiates different kernels, so each is run here with twelve tracers, on
the 24-layer channel (64-lane kernels, one level per lane in the last kernel),
on a 70-layer channel (128-lane kernels, two levels per lane in the last
kernel, short rows) and on a 90-layer one (long rows), two steps each, with
each variant of the last kernel (FGPU_STEP_KD: lock-step, with a sweep wave,
one wave per column).

Two of the twelve tracers sit far from zero (about +2100 and -1500, as DIC
or alkalinity do): the FCT bounds of a node next to an element without the
level take that element's +-1e3 padding (oce_adv_tra_fct.F90:148-178), which
only binds for such values. The oracle with the padding left out (planted
mistake 49) gives a different result on these states, which the last test
shows on the CPU side of this file.
"""
import numpy as np
import pytest

from fesom2_amd import meshgen as mg

pytestmark = pytest.mark.gpu

TPB = [1, 2, 4, 8, 12]
TCH = [2, 4, 5, 8]
CPB = [1, 4, 8]
KD = ["lockstep", "sweepwave", "wavelocal"]   # variants of the last kernel (FGPU_STEP_KD)
KD_ID = {"lockstep": 0, "sweepwave": 1, "wavelocal": 2}
SHORT_NL = 72   # the last kernel's short LDS rows (3 or 4 columns per block) up to this many interfaces


def _wet(mesh):
    k = np.arange(1, mesh.nl)
    return (k[None, :] >= mesh.ulevels_nod2D[:, None]) & (k[None, :] <= (mesh.nlevels_nod2D - 1)[:, None])


def offset_state(name, seed=3):
    """Twelve tracers: T, S, ten passive blobs; the third offset to ~2100 and
    the fourth to ~-1500, with small noise."""
    mesh, _, _ = mg.make_config(name)
    cfgs = dict(mg.CONFIGS[name]["state"])
    cfgs.update(ntr=12)
    st = mg.build_state(mesh, seed=seed, dx=20e3, **cfgs)
    rng = np.random.default_rng(seed + 17)
    w = _wet(mesh)
    v = st.values.copy()
    v[2] = np.where(w, v[2] + 2100.0 + rng.normal(0.0, 2.0, v[2].shape), 0.0)
    v[3] = np.where(w, v[3] - 1500.0 + rng.normal(0.0, 2.0, v[3].shape), 0.0)
    import dataclasses
    st = dataclasses.replace(st, values=v, valuesold=np.where(w[None], v - 0.01 * np.sign(v), 0.0))
    return mesh, st


def _oracle_steps(mesh, st, nsteps, schemes=None):
    from oracle.oracle import Oracle
    v, vo = st.values.copy(), st.valuesold.copy()
    o = Oracle(mesh)
    for _ in range(nsteps):
        o.step(st, v, vo, schemes=schemes)
    return v, vo


@pytest.fixture(scope="module")
def chan():
    mesh, st = offset_state("c0")
    return mesh, st, _oracle_steps(mesh, st, 2)


@pytest.fixture(scope="module")
def deep():
    mesh, st = offset_state("c0deep")
    assert mesh.nl - 1 > 64 and (mesh.nlevels_nod2D - 1).min() < 20
    return mesh, st, _oracle_steps(mesh, st, 2)


@pytest.fixture(scope="module")
def deep90():
    mesh, st = offset_state("c0deep90")
    assert mesh.nl > SHORT_NL and (mesh.nlevels_nod2D - 1).min() < 20
    return mesh, st, _oracle_steps(mesh, st, 2)


def _run(mesh, st, monkeypatch, tpb=None, tch=None, cpb=None, staged=False, nsteps=2, schemes=None, kd=None,
         ka=None, kb=None):
    from fesom2_amd import gpu
    for k, v in (("FGPU_STEP_TPB", tpb), ("FGPU_STEP_TCH", tch), ("FGPU_STEP_CPB", cpb), ("FGPU_STEP_KD", kd),
                 ("FGPU_STEP_KA", ka), ("FGPU_STEP_KB", kb)):
        if v is None:
            monkeypatch.delenv(k, raising=False)
        else:
            monkeypatch.setenv(k, str(v))
    if staged:
        monkeypatch.setenv("FGPU_STEP", "staged")
    else:
        monkeypatch.delenv("FGPU_STEP", raising=False)
    ctx = gpu.context_for(mesh, st, schemes=schemes)
    try:
        ctx.solve_tracers_ale(st.dt, nsteps)
        shape = ctx.step_shape()
        if kd is not None:
            assert ctx.step_variants()[0] == KD_ID[kd]
        if ka is not None:
            assert ctx.step_variants()[2] == ka
        if kb is not None:
            assert ctx.step_variants()[3] == kb
        out = [ctx.tracer_download(i + 1) for i in range(st.ntr)]
    finally:
        ctx.close()
    return out, shape


def _compare(mesh, out, ref):
    w = _wet(mesh)
    v, vo = ref
    for i, (gv, gvo) in enumerate(out):
        np.testing.assert_array_equal(gv[w], v[i][w], err_msg=f"tracer {i + 1}")
        np.testing.assert_array_equal(gvo[w], vo[i][w], err_msg=f"tracer {i + 1} valuesold")


def _launched(tch, cpb, deep, kd, short=True):
    """The (Thomas chunk, columns per block) the last kernel runs with for a
    requested shape: several columns per block take chunks of 2, or 4 at four
    64-lane columns or two deep columns; several deep columns need a
    one-wave-per-column kernel, which takes 3 or 4 of them with the short LDS
    rows (at most SHORT_NL interfaces), else 2."""
    if not deep:
        if cpb == 8:
            return 2, 8
        if cpb == 4:
            return (4 if tch == 4 else 2), 4
        return tch, 1
    if kd != "lockstep" and cpb >= 3 and short:
        return 2, min(cpb, 4)
    if kd != "lockstep" and cpb >= 2:
        return (4 if tch == 4 else 2), 2
    return tch, 1


@pytest.mark.parametrize("kd", KD)
@pytest.mark.parametrize("cpb", CPB)
@pytest.mark.parametrize("tch", TCH)
@pytest.mark.parametrize("tpb", TPB)
def test_fused_shapes_64_lanes(chan, monkeypatch, tpb, tch, cpb, kd):
    mesh, st, ref = chan
    print(f"[gpu] c0, 12 tracers, shape tpb={tpb} tch={tch} cpb={cpb} K_D {kd}", flush=True)
    out, shape = _run(mesh, st, monkeypatch, tpb, tch, cpb, kd=kd)
    assert shape == (tpb, *_launched(tch, cpb, False, kd))
    _compare(mesh, out, ref)


@pytest.mark.parametrize("kd", KD)
@pytest.mark.parametrize("cpb", [1, 2, 3, 4])
@pytest.mark.parametrize("tch", TCH)
@pytest.mark.parametrize("tpb", TPB)
def test_fused_shapes_128_lanes(deep, monkeypatch, tpb, tch, cpb, kd):
    mesh, st, ref = deep
    assert mesh.nl <= SHORT_NL
    print(f"[gpu] 70 layers, 12 tracers, shape tpb={tpb} tch={tch} cpb={cpb} K_D {kd}", flush=True)
    out, shape = _run(mesh, st, monkeypatch, tpb, tch, cpb, kd=kd)
    assert shape == (tpb, *_launched(tch, cpb, True, kd))
    _compare(mesh, out, ref)


@pytest.mark.parametrize("kd", KD)
@pytest.mark.parametrize("cpb", [1, 2, 4])
@pytest.mark.parametrize("tch", TCH)
@pytest.mark.parametrize("tpb", [1, 4, 12])
def test_fused_shapes_long_rows(deep90, monkeypatch, tpb, tch, cpb, kd):
    """90 layers: the last kernel's LDS rows for up to 128 levels (at most two
    columns per block), and two-level K_A on an even layer count."""
    mesh, st, ref = deep90
    print(f"[gpu] 90 layers, 12 tracers, shape tpb={tpb} tch={tch} cpb={cpb} K_D {kd}", flush=True)
    out, shape = _run(mesh, st, monkeypatch, tpb, tch, cpb, kd=kd, ka=2 if tpb == 12 else 1)
    assert shape == (tpb, *_launched(tch, cpb, True, kd, short=False))
    _compare(mesh, out, ref)


@pytest.mark.parametrize("kc", [1, 2, 4])
@pytest.mark.parametrize("which", ["chan", "deep"])
def test_flat_bounds_kernel(chan, deep, monkeypatch, which, kc):
    """K_C over (tracer, level) items (FGPU_STEP_KC: tracers per block), with
    each K_D variant, bit for bit."""
    mesh, st, ref = chan if which == "chan" else deep
    for kd in KD:
        print(f"[gpu] flat K_C {which} kc={kc} K_D {kd}", flush=True)
        monkeypatch.setenv("FGPU_STEP_KC", str(kc))
        out, _ = _run(mesh, st, monkeypatch, 12, 2, 2 if which == "deep" else 4, kd=kd)
        _compare(mesh, out, ref)
    monkeypatch.delenv("FGPU_STEP_KC", raising=False)


@pytest.mark.parametrize("kd", KD)
@pytest.mark.parametrize("which", ["chan", "deep"])
def test_edge_kernel_two_levels(chan, deep, monkeypatch, which, kd):
    """K_A with two adjacent levels per lane (FGPU_STEP_KA=2: 16-byte loads
    and stores, one wave per edge; both meshes have an even number of
    layers), before each K_D variant and both K_C variants, bit for bit."""
    mesh, st, ref = chan if which == "chan" else deep
    assert (mesh.nl - 1) % 2 == 0
    for kc in (0, 4):
        print(f"[gpu] two-level K_A {which} K_D {kd} kc={kc}", flush=True)
        monkeypatch.setenv("FGPU_STEP_KC", str(kc))
        out, _ = _run(mesh, st, monkeypatch, 12, 2, 2 if which == "deep" else 4, kd=kd, ka=2)
        _compare(mesh, out, ref)
    monkeypatch.delenv("FGPU_STEP_KC", raising=False)


MIXED_QR4C_UPW1 = [dict(), dict(), dict(hor="UPW1", ver="UPW1", lim="FCT"), dict(hor="MFCT", ver="QR4C", lim="NON", ph=0.5, pv=0.0),
                   dict(hor="MUSCL", ver="QR4C", lim="FCT"), dict(hor="UPW1", ver="UPW1", lim="NON"), dict(ph=0.5),
                   dict(hor="MUSCL", ver="UPW1", lim="NON", ph=0.5), dict(pv=0.5), dict(), dict(hor="MFCT", ver="UPW1", lim="FCT"),
                   dict(AB_order=2)]


@pytest.mark.parametrize("kd", KD)
@pytest.mark.parametrize("which", ["chan", "deep", "deep90"])
def test_bound_kernels_two_levels(chan, deep, deep90, monkeypatch, which, kd):
    """K_A and K_B with two adjacent levels per lane (FGPU_STEP_KA=2,
    FGPU_STEP_KB=2: 16-byte accesses, one wave per column or edge), with the
    default schemes and with tracers of different QR4C/UPW1 schemes side by
    side, before each K_D variant, bit for bit."""
    mesh, st, ref = {"chan": chan, "deep": deep, "deep90": deep90}[which]
    assert (mesh.nl - 1) % 2 == 0
    cpb = 4 if which == "chan" else 2
    print(f"[gpu] two-level K_A and K_B {which} K_D {kd}", flush=True)
    out, _ = _run(mesh, st, monkeypatch, 12, 2, cpb, kd=kd, ka=2, kb=2)
    _compare(mesh, out, ref)
    ref2 = _oracle_steps(mesh, st, 2, schemes=MIXED_QR4C_UPW1)
    print(f"[gpu] two-level K_A and K_B {which} K_D {kd}, mixed QR4C/UPW1 schemes", flush=True)
    out, _ = _run(mesh, st, monkeypatch, 4, 4, cpb, kd=kd, ka=2, kb=2, schemes=MIXED_QR4C_UPW1)
    _compare(mesh, out, ref2)


@pytest.mark.parametrize("which", ["chan", "deep"])
def test_staged_offsets(chan, deep, monkeypatch, which):
    mesh, st, ref = chan if which == "chan" else deep
    print(f"[gpu] staged {which}, 12 tracers", flush=True)
    out, _ = _run(mesh, st, monkeypatch, staged=True)
    _compare(mesh, out, ref)


def test_autotuned_deep(deep, monkeypatch):
    """The autotuner on the 70-layer mesh restores the state it timed on and
    its pick stays bit-identical."""
    from fesom2_amd import gpu
    mesh, st, ref = deep
    for k in ("FGPU_STEP_TPB", "FGPU_STEP_TCH", "FGPU_STEP_CPB", "FGPU_STEP", "FGPU_STEP_KD"):
        monkeypatch.delenv(k, raising=False)
    print("[gpu] autotune 70 layers", flush=True)
    ctx = gpu.context_for(mesh, st)
    try:
        ctx.autotune_step(st.dt, 2)
        ctx.solve_tracers_ale(st.dt, 2)
        out = [ctx.tracer_download(i + 1) for i in range(st.ntr)]
    finally:
        ctx.close()
    _compare(mesh, out, ref)


MIXED = [dict(), dict(), dict(), dict(), dict(hor="UPW1", ver="UPW1", lim="FCT"),
         dict(hor="MFCT", ver="QR4C", lim="NON", ph=0.5, pv=0.0), dict(hor="MUSCL", ver="QR4C", lim="FCT"),
         dict(hor="MFCT", ver="CDIFF", lim="FCT"), dict(hor="MFCT", ver="PPM", lim="FCT"),
         dict(hor="UPW1", ver="PPM", lim="NON"), dict(ph=0.5), dict(hor="MUSCL", ver="PPM", lim="NON", ph=0.5)]


@pytest.mark.parametrize("shape", [(1, 2, 1), (4, 4, 4), (12, 8, 1), (8, 2, 8)], ids=str)
@pytest.mark.parametrize("which", ["chan", "deep"])
def test_mixed_schemes_shapes(chan, deep, monkeypatch, which, shape):
    """Twelve tracers with different schemes in one block (the per-tracer
    branches of every kernel), on both meshes."""
    mesh, st, _ = chan if which == "chan" else deep
    ref = _oracle_steps(mesh, st, 2, schemes=MIXED)
    print(f"[gpu] mixed schemes {which} shape={shape}", flush=True)
    out, _ = _run(mesh, st, monkeypatch, *shape, schemes=MIXED)
    _compare(mesh, out, ref)


@pytest.mark.parametrize("ka", [1, 2])
@pytest.mark.parametrize("kd", ["lockstep", "sweepwave"])
@pytest.mark.parametrize("which", ["chan", "deep"])
def test_mixed_schemes_flat_bounds(chan, deep, monkeypatch, which, kd, ka):
    """The (tracer, level) K_C with tracers of different schemes in one block
    (FCT and unlimited ones side by side in a chunk), before each K_D, after
    both K_A variants (UPW1, MFCT and MUSCL edges, with and without FCT)."""
    mesh, st, _ = chan if which == "chan" else deep
    ref = _oracle_steps(mesh, st, 2, schemes=MIXED)
    monkeypatch.setenv("FGPU_STEP_KC", "4")
    print(f"[gpu] mixed schemes {which} (tracer, level) K_C, K_D {kd}, K_A {ka} levels per lane", flush=True)
    out, _ = _run(mesh, st, monkeypatch, 12, 4, 2 if which == "deep" else 4, kd=kd, schemes=MIXED, ka=ka)
    _compare(mesh, out, ref)
    monkeypatch.delenv("FGPU_STEP_KC", raising=False)


@pytest.mark.parametrize("path", ["fused", "staged"])
def test_two_layer_columns(monkeypatch, path):
    """A channel whose shallowest columns have two layers: the second and the
    last-but-one QR4C rows coincide there and the reference applies that row
    twice (oce_adv_tra_ver.F90:392-400); the Thomas solve has two rows. Three
    steps, twelve tracers, bit for bit."""
    mesh = mg.build_mesh(nx=20, ny=14, dx=20e3, jitter=0.2, seed=5, cartesian=True, nlayers=10, hmax=2000.0,
                         min_layers=2, max_layers=10, topo="bumps", bump_height=1900.0)
    assert np.any(mesh.nlevels_nod2D - mesh.ulevels_nod2D == 2)
    st = mg.build_state(mesh, seed=5, ntr=12, dt=900.0, dx=20e3)
    ref = _oracle_steps(mesh, st, 3)
    print(f"[gpu] two-layer columns, {path}", flush=True)
    out, _ = _run(mesh, st, monkeypatch, staged=(path == "staged"), nsteps=3)
    _compare(mesh, out, ref)
