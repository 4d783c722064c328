"""GPU context: argument, state and limit errors of the C ABI are reported
as status codes with a message (host-side checks, nothing reaches a kernel),
and the context stays usable afterwards."""
import ctypes as C
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import openfea_amd as ofa
from openfea_amd import _capi as A, meshgen as mg

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def fan_mesh(ntet):
    """ntet tets sharing node 1, each with three nodes of its own: node 1
    couples to 3 ntet neighbours."""
    xyz = [[0.0, 0.0, 0.0]]
    conn = []
    for t in range(ntet):
        a = 2.0 * np.pi * t / ntet
        base = len(xyz) + 1
        xyz += [[np.cos(a), np.sin(a), 0.0], [np.cos(a + 0.01), np.sin(a + 0.01), 0.0], [0.0, 0.0, 1.0 + t]]
        conn.append([1, base, base + 1, base + 2, 0, 0, 0, 0])
    return f64(xyz), i32(conn), i32(np.zeros(ntet))


def test_api_errors_and_recovery():
    L = A.lib()
    ctx = ofa.Context(0)
    h = ctx._h
    try:
        say("errors: state before a mesh")
        b = f64(np.ones(3))
        x = np.zeros(3)
        it, rel = C.c_int(), C.c_double()
        assert L.ofea_solve(h, A.PCG_JACOBI, p(b), p(x), 1e-8, 100, C.byref(it), C.byref(rel)) == A.OFEA_ERR_STATE
        assert L.ofea_last_error()
        m = mg.kuhn_tets(4, 2, 2, jitter=0.1, seed=2)
        resort = i32(mg.free_index(m.nd, m.fix_set))
        xyz, conn, et = f64(m.xyz), i32(m.conn), i32(m.etype)
        say("errors: bad mesh arguments")
        bad_et = et.copy()
        bad_et[3] = 5
        assert L.ofea_set_mesh(h, m.nd, p(xyz), len(et), p(conn), p(bad_et), p(resort)) == A.OFEA_ERR_ARG
        bad_rs = resort.copy()
        bad_rs[-1] = m.nd + 3
        assert L.ofea_set_mesh(h, m.nd, p(xyz), len(et), p(conn), p(et), p(bad_rs)) == A.OFEA_ERR_ARG
        assert L.ofea_set_mesh(h, 0, p(xyz), len(et), p(conn), p(et), p(resort)) == A.OFEA_ERR_ARG
        say("errors: a node with more neighbours than the gather plan holds")
        fx, fc, fe = fan_mesh(100)
        fr = i32(np.arange(len(fx)))
        assert L.ofea_set_mesh(h, len(fx), p(fx), len(fe), p(fc), p(fe), p(fr)) == A.OFEA_ERR_LIMIT
        assert b"neighbour" in L.ofea_last_error()
        say("errors: after a failed set_mesh the context holds no mesh")
        nn = C.c_int(-1)
        assert L.ofea_pattern(h, C.byref(nn), None, None, None) == A.OFEA_ERR_STATE
        assert L.ofea_assemble(h, mg.E_STEEL, mg.NU_STEEL, None) == A.OFEA_ERR_STATE
        y = np.zeros(3)
        assert L.ofea_spmv(h, p(b), p(y)) == A.OFEA_ERR_STATE
        assert L.ofea_solve(h, A.PCG_JACOBI, p(b), p(x), 1e-8, 100, None, None) == A.OFEA_ERR_STATE
        uu = np.zeros(3)
        assert L.ofea_nl_solve(h, mg.E_STEEL, mg.NU_STEEL, p(b), 1, 1e-8, 5, A.PCG_JACOBI, 1e-10, p(uu),
                               None, None, None) == A.OFEA_ERR_STATE
        assert L.ofea_nl_assemble(h, mg.E_STEEL, mg.NU_STEEL, p(uu), None, None) == A.OFEA_ERR_STATE
        assert L.ofea_assemble_batch(h, 1, p(b), mg.NU_STEEL, None) == A.OFEA_ERR_STATE
        say("errors: state and arguments on a valid mesh")
        # through the wrapper, so that its sizes (n, nnz) follow this mesh
        ctx.set_mesh(xyz, conn, et, resort)
        n = ctx.n
        assert n == 3 * int((resort >= 0).sum()) and ctx.nnz > 0
        b = f64(np.random.default_rng(1).normal(size=n))
        x = np.zeros(n)
        assert L.ofea_solve(h, A.PCG_JACOBI, p(b), p(x), 1e-8, 100, None, None) == A.OFEA_ERR_STATE
        assert L.ofea_assemble(h, mg.E_STEEL, 0.5, None) == A.OFEA_ERR_ARG
        assert L.ofea_assemble(h, mg.E_STEEL, mg.NU_STEEL, None) == A.OFEA_OK
        assert L.ofea_solve(h, A.PCG_JACOBI, None, p(x), 1e-8, 100, None, None) == A.OFEA_ERR_STATE  # no RHS yet
        assert L.ofea_solve(h, A.PCG_JACOBI, p(b), p(x), 0.0, 100, None, None) == A.OFEA_ERR_ARG
        assert L.ofea_solve(h, 42, p(b), p(x), 1e-8, 100, None, None) == A.OFEA_ERR_ARG
        for meth in (A.PCG_LEARNED, A.PCG_LEARNED_TWO_LEVEL):
            assert L.ofea_solve(h, meth, p(b), p(x), 1e-8, 100, None, None) == A.OFEA_ERR_STATE
        say("errors: the context still solves correctly")
        cp, ri = ctx.pattern()
        val = ctx.assemble(mg.E_STEEL, mg.NU_STEEL)
        xd = spla.spsolve(sp.csc_matrix((val, ri, cp), shape=(n, n)), b)
        for meth in (A.PCG_BLOCK_JACOBI, A.PCG_TWO_LEVEL):
            xs, it2, rel2, ok = ctx.solve(b, method=meth, rtol=1e-11)
            assert ok and np.linalg.norm(xs - xd) / np.linalg.norm(xd) < 1e-8
    finally:
        ctx.close()
