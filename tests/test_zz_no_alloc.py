"""ofea_solve_noalloc: the (block) Jacobi PCG of ofea_solve on device vectors
with every buffer of the run in the library's static device array.

Checked against ofea_solve with the same method on the same system: the
kernels, grids and control flow are the same, so the iterates must agree bit
for bit (the bar asserted is 1e-13 relative, far inside the 1e-8 displacement
bar, and equal iteration counts).  Two calls in a row, the second on a smaller
system, so a stale scratch array from the first call would show; then the
error paths (unknown method, system above the limit, no assembly).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import openfea_amd as ofa
from openfea_amd import _capi as A, meshgen as mg

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


class DevBuf:
    """Caller-owned device memory from the HIP runtime the library uses (the
    caller may allocate; the entry point under test does not)."""
    _hip = None

    def __init__(self, host):
        if DevBuf._hip is None:
            DevBuf._hip = C.CDLL("libamdhip64.so")
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = C.c_void_p()
        assert DevBuf._hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.host.nbytes)) == 0
        assert DevBuf._hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.host.nbytes),
                                     1) == 0  # hipMemcpyHostToDevice

    def get(self):
        out = np.empty_like(self.host)
        assert DevBuf._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.ptr:
            DevBuf._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _noalloc(ctx, method, b, rtol):
    db, dx = DevBuf(b), DevBuf(np.full(ctx.n, np.nan))
    try:
        it, rel = C.c_int(), C.c_double()
        rc = A.lib().ofea_solve_noalloc(ctx._h, method, db.ptr, dx.ptr, rtol, 100000, C.byref(it), C.byref(rel))
        return rc, dx.get(), it.value, rel.value
    finally:
        db.free()
        dx.free()


@pytest.mark.parametrize("method", [A.PCG_JACOBI, A.PCG_BLOCK_JACOBI])
def test_noalloc_matches_solve(method):
    rng = np.random.default_rng(5)
    with ofa.Context(0) as ctx:
        # first call on the larger system, second on a smaller one
        for dims in ((16, 8, 8), (7, 4, 3)):
            say(f"noalloc method {method}: Kuhn tets {dims}")
            m = mg.kuhn_tets(*dims, jitter=0.15, seed=11)
            resort = mg.free_index(m.nd, m.fix_set)
            ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
            ctx.assemble(mg.E_STEEL, mg.NU_STEEL, copy=False)
            b = rng.normal(size=ctx.n)
            xs, its, rels, ok = ctx.solve(b, method=method, rtol=1e-10)
            assert ok
            rc, xn, itn, reln = _noalloc(ctx, method, b, 1e-10)
            assert rc == A.OFEA_OK, A.lib().ofea_last_error()
            assert itn == its and reln <= 1e-10
            err = np.linalg.norm(xn - xs) / np.linalg.norm(xs)
            say(f"  n {ctx.n}: iterations {itn} / {its}, |x_noalloc - x_solve| / |x_solve| = {err:.3e}")
            assert err <= 1e-13


def test_noalloc_errors():
    L = A.lib()
    with ofa.Context(0) as ctx:
        say("noalloc errors: no mesh, unknown method, limit")
        d = DevBuf(np.zeros(8))
        dp = d.ptr
        assert L.ofea_solve_noalloc(ctx._h, A.PCG_JACOBI, dp, dp, 1e-8, 10, None, None) == A.OFEA_ERR_STATE
        m = mg.kuhn_tets(4, 2, 2, seed=1)
        ctx.set_mesh(m.xyz, m.conn, m.etype, mg.free_index(m.nd, m.fix_set))
        ctx.assemble(mg.E_STEEL, mg.NU_STEEL, copy=False)
        assert L.ofea_solve_noalloc(ctx._h, A.PCG_TWO_LEVEL, dp, dp, 1e-8, 10, None, None) == A.OFEA_ERR_ARG
        assert L.ofea_solve_noalloc(ctx._h, A.PCG_JACOBI, None, dp, 1e-8, 10, None, None) == A.OFEA_ERR_ARG
        # 45 x 45 x 45 cells: 93,150 free nodes, 279,450 DOFs > 262,144
        big = mg.kuhn_tets(45, 45, 45, seed=1)
        ctx.set_mesh(big.xyz, big.conn, big.etype, mg.free_index(big.nd, big.fix_set))
        assert ctx.n > 262144
        ctx.assemble(mg.E_STEEL, mg.NU_STEEL, copy=False)
        assert L.ofea_solve_noalloc(ctx._h, A.PCG_JACOBI, dp, dp, 1e-8, 10, None, None) == A.OFEA_ERR_LIMIT
        assert b"262144" in L.ofea_last_error()
        d.free()
