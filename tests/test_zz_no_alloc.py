"""fgpu_ssh_solve_cg_noalloc: the SSH CG with its work vectors in the library's
static device area, on the caller's device vectors, against fgpu_ssh_solve_cg
(same kernels and iteration order: the same x bit for bit, the same iteration
count). Two calls in a row, the second on a smaller system, so the second
reads nothing the first left in the static area; then a warm start and the
sparse preconditioner (variant 1) on the small system."""
import ctypes as C

import numpy as np
import pytest

from fesom2_amd import meshgen as mg

pytestmark = pytest.mark.gpu


class _Dev:
    """Device vectors through the HIP runtime the library itself links (the
    test allocates them; the entry point under test allocates nothing)."""

    def __init__(self):
        from fesom2_amd import gpu
        gpu.load()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self, *ps):
        for p in ps:
            self.hip.hipFree(p)


def _solve_both(ctx, mesh, rhs, x0, variant):
    dev = _Dev()
    ctx.ssh_solve_preconditioner(variant)
    x_ref, st_ref = ctx.ssh_solve_cg(x0, rhs, soltol=1e-10, maxiter=5000)
    xd, bd = dev.put(x0), dev.put(rhs)
    try:
        st = ctx.ssh_solve_cg_noalloc(xd.value, bd.value, soltol=1e-10, maxiter=5000)
        x = dev.get(xd, mesh.nod2D)
        b = dev.get(bd, mesh.nod2D)
    finally:
        dev.free(xd, bd)
    print(f"[gpu] noalloc n={mesh.nod2D} variant {variant}: {st['iters']} its (allocating entry "
          f"{st_ref['iters']} its)", flush=True)
    assert st["converged"] and st_ref["converged"]
    assert st["iters"] == st_ref["iters"]
    np.testing.assert_array_equal(x, x_ref)
    np.testing.assert_array_equal(b, rhs)   # rhs is read only
    return x


def test_noalloc_cg_matches_the_allocating_entry():
    from fesom2_amd import gpu
    big, bstate, _ = mg.make_config("c1")
    small, sstate, _ = mg.make_config("c0")
    assert small.nod2D < big.nod2D <= 1500000   # FGPU_NOALLOC_MAX_ROWS
    for mesh, state in ((big, bstate), (small, sstate)):
        print(f"[gpu] noalloc CG on N2={mesh.nod2D}", flush=True)
        ctx = gpu.context_for(mesh, state)
        try:
            ctx.init_stiff_mat_ale(state.dt)
            rhs = mg.ssh_rhs(mesh)
            x = _solve_both(ctx, mesh, rhs, np.zeros(mesh.nod2D), 2)
            if mesh is small:
                rng = np.random.default_rng(5)
                _solve_both(ctx, mesh, rhs * 1.5, x + rng.normal(0.0, 1e-3, mesh.nod2D), 2)
                _solve_both(ctx, mesh, rhs, np.zeros(mesh.nod2D), 1)
        finally:
            ctx.close()


def test_noalloc_rejects_null_vectors():
    from fesom2_amd import gpu
    mesh, state, _ = mg.make_config("c0")
    ctx = gpu.context_for(mesh, state)
    try:
        ctx.init_stiff_mat_ale(state.dt)
        ctx.ssh_solve_preconditioner(2)
        with pytest.raises(gpu.FesomGpuError):
            ctx.ssh_solve_cg_noalloc(0, 0)
    finally:
        ctx.close()
