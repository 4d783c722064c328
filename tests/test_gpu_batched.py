"""GPU parity of the batched reanalysis path (many variants, many load cases).

Bars: each variant's FP64 stiffness within 1e-12 of max|K_v| entrywise of the
oracle's fill with the same per-element moduli; each multi-RHS PCG solution
within 1e-8 (relative 2-norm) of a direct solve of the oracle's K_v at
rtol 1e-10, with its true FP64 residual <= rtol.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import rel_max
from oracle import oracle as O
import openfea_amd as ofa
from openfea_amd import _capi as A, meshgen as mg

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-batched] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ofa.Context(0)
    yield c
    c.close()


def csc(cp, ri, v):
    return sp.csc_matrix((v, ri, cp), shape=(len(cp) - 1, len(cp) - 1))


def setup(ctx, m):
    resort = mg.free_index(m.nd, m.fix_set)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    return resort, cp, ri


def loads(m, resort, nvar, nrhs):
    B = np.zeros((nvar, nrhs, 3 * int(resort.max() + 1)))
    for v in range(nvar):
        for k in range(nrhs):
            B[v, k] = mg.free_rhs(mg.face_traction(m, mg.load_case(v, k)), resort)
    return B


@pytest.mark.parametrize("kind,nvar", [("hex", 5), ("tet", 3), ("hex", 64)])
def test_assemble_batch_matches_oracle(ctx, kind, nvar):
    m = mg.hex_block(6, 4, 3, jitter=0.2, seed=11) if kind == "hex" else mg.kuhn_tets(5, 4, 3, jitter=0.2, seed=5)
    if nvar == 64:
        m = mg.hex_block(4, 3, 3, jitter=0.1, seed=2)
    say(f"assemble_batch {kind} ne={m.ne} nvar={nvar}")
    resort, cp, ri = setup(ctx, m)
    Ev = np.stack([mg.design_field(m, v) for v in range(nvar)])
    vals = ctx.assemble_batch(Ev, mg.NU_STEEL)
    for v in sorted({0, nvar // 2, nvar - 1}):
        ref = O.fill_csc(m.conn, m.etype, m.xyz, resort, 0.0, mg.NU_STEEL, cp, ri, E_elem=Ev[v])
        assert rel_max(vals[v], ref) < 1e-12, v
    # deterministic: a second launch gives the same bits, also from the resident moduli
    assert np.array_equal(ctx.assemble_batch(Ev, mg.NU_STEEL), vals)
    assert np.array_equal(ctx.assemble_batch(None, mg.NU_STEEL, nvar=nvar), vals)


def test_assemble_batch_limits(ctx):
    m = mg.hex_block(3, 2, 2)
    setup(ctx, m)
    with pytest.raises(A.OfeaError) as e:
        ctx.assemble_batch(np.ones((65, m.ne)), 0.3)
    assert e.value.code == A.OFEA_ERR_LIMIT
    # mixed C3D4 + C3D8 meshes are rejected before any launch
    t = mg.kuhn_tets(2, 2, 2)
    h = mg.hex_block(2, 2, 2)
    conn = np.vstack([h.conn, t.conn])
    etype = np.concatenate([h.etype, t.etype])
    ctx.set_mesh(h.xyz, conn, etype, mg.free_index(h.nd, h.fix_set))
    with pytest.raises(A.OfeaError) as e:
        ctx.assemble_batch(np.ones((2, len(etype))), 0.3)
    assert e.value.code == A.OFEA_ERR_ARG


@pytest.mark.parametrize("nrhs", [1, 2, 8])
def test_solve_batch_matches_direct(ctx, nrhs):
    m = mg.hex_block(8, 4, 4, jitter=0.1, seed=3)
    say(f"solve_batch hex ne={m.ne} nrhs={nrhs}")
    resort, cp, ri = setup(ctx, m)
    nvar = 3
    Ev = np.stack([mg.design_field(m, v) for v in range(nvar)])
    ctx.assemble_batch(Ev, mg.NU_STEEL, copy=False)
    B = loads(m, resort, nvar, nrhs)
    X, it, rr, ok = ctx.solve_batch(B, rtol=1e-10)
    assert ok and np.all(rr <= 1e-10) and np.all(it > 0)
    for v in range(nvar):
        K = csc(cp, ri, O.fill_csc(m.conn, m.etype, m.xyz, resort, 0.0, mg.NU_STEEL, cp, ri, E_elem=Ev[v]))
        lu = spla.splu(K)
        for k in range(nrhs):
            xd = lu.solve(B[v, k])
            assert np.linalg.norm(X[v, k] - xd) / np.linalg.norm(xd) < 1e-8, (v, k)


def test_solve_batch_independent_systems(ctx):
    """Systems stop on their own: a zero load stops at once with x = 0, a
    scaled load gives the scaled solution, and results do not depend on the
    other systems of the batch."""
    m = mg.hex_block(6, 3, 3, jitter=0.1, seed=8)
    resort, cp, ri = setup(ctx, m)
    Ev = np.stack([mg.design_field(m, v) for v in range(2)])
    ctx.assemble_batch(Ev, mg.NU_STEEL, copy=False)
    B = loads(m, resort, 2, 4)
    B[0, 1] = 0.0
    B[1, 2] = 1e-6 * B[1, 0]
    X, it, rr, ok = ctx.solve_batch(B, rtol=1e-10)
    assert ok
    assert it[0, 1] == 0 and np.all(X[0, 1] == 0.0) and rr[0, 1] == 0.0
    assert np.linalg.norm(X[1, 2] - 1e-6 * X[1, 0]) <= 1e-8 * np.linalg.norm(1e-6 * X[1, 0])
    # the same systems one at a time
    X1, it1, _, ok1 = ctx.solve_batch(B[:, :1].copy(), rtol=1e-10)
    assert ok1
    for v in range(2):
        assert np.linalg.norm(X1[v, 0] - X[v, 0]) <= 2e-8 * np.linalg.norm(X[v, 0])
    g, ms = ctx.batch_tiles(4)
    assert g in (16, 32, 64) and ms > 0.0


def test_autotune_keeps_results(ctx):
    m, resort, F = mg.cantilever_c3d8(10, 5, 5)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    ctx.assemble(mg.E_STEEL, mg.NU_STEEL, copy=False)
    x0, it0, _, ok0 = ctx.solve(F, A.PCG_BLOCK_JACOBI, rtol=1e-10)
    best, times = ctx.autotune()
    say(f"autotune best={best} times={times}")
    assert best in (4, 8, 16, 32, 64) and np.all(times > 0)
    assert times[[4, 8, 16, 32, 64].index(best)] == times.min()
    x1, it1, _, ok1 = ctx.solve(F, A.PCG_BLOCK_JACOBI, rtol=1e-10)
    assert ok0 and ok1
    assert np.linalg.norm(x1 - x0) <= 1e-9 * np.linalg.norm(x0)


@pytest.mark.parametrize("nrhs", [1, 8])
def test_solve_batch_two_level_matches_direct(ctx, nrhs):
    """The multi-RHS solver with the per-variant coarse space (ofea_solve_batch_pc,
    OFEA_PCG_TWO_LEVEL): solutions within 1e-8 of direct solves of the oracle's
    K_v, and fewer iterations than the same batch with block Jacobi alone."""
    m = mg.hex_block(16, 6, 6, jitter=0.1, seed=4)
    say(f"solve_batch two-level hex ne={m.ne} nrhs={nrhs}")
    resort, cp, ri = setup(ctx, m)
    nvar = 3
    Ev = np.stack([mg.design_field(m, v) for v in range(nvar)])
    ctx.assemble_batch(Ev, mg.NU_STEEL, copy=False)
    B = loads(m, resort, nvar, nrhs)
    X, it, rr, ok = ctx.solve_batch(B, rtol=1e-10, method=A.PCG_TWO_LEVEL)
    assert ok and np.all(rr <= 1e-10) and np.all(it > 0)
    _, it_bj, _, ok_bj = ctx.solve_batch(B, rtol=1e-10, method=A.PCG_BLOCK_JACOBI)
    assert ok_bj
    say(f"iterations two-level {it.min()}..{it.max()} block Jacobi {it_bj.min()}..{it_bj.max()}")
    assert it.max() < it_bj.min()
    for v in range(nvar):
        K = csc(cp, ri, O.fill_csc(m.conn, m.etype, m.xyz, resort, 0.0, mg.NU_STEEL, cp, ri, E_elem=Ev[v]))
        lu = spla.splu(K)
        for k in range(nrhs):
            xd = lu.solve(B[v, k])
            assert np.linalg.norm(X[v, k] - xd) / np.linalg.norm(xd) < 1e-8, (v, k)
    # the single-system two-level solver on the same K_v still works afterwards
    ctx.set_elem_E(Ev[1])
    ctx.assemble(0.0, mg.NU_STEEL, copy=False)
    ctx.set_rhs(B[1, 0])
    x1, _, _, ok1 = ctx.solve(B[1, 0], method=A.PCG_TWO_LEVEL, rtol=1e-10)
    ctx.set_elem_E(None)
    assert ok1 and np.linalg.norm(x1 - X[1, 0]) <= 1e-8 * np.linalg.norm(X[1, 0])


def test_solve_batch_two_level_independent_systems(ctx):
    """Two-level batch: a zero load stops at once with x = 0, a scaled load
    gives the scaled solution, a system alone gives the batch's solution."""
    m = mg.hex_block(10, 4, 4, jitter=0.1, seed=9)
    resort, cp, ri = setup(ctx, m)
    Ev = np.stack([mg.design_field(m, v) for v in range(2)])
    ctx.assemble_batch(Ev, mg.NU_STEEL, copy=False)
    B = loads(m, resort, 2, 4)
    B[0, 1] = 0.0
    B[1, 2] = 1e-6 * B[1, 0]
    X, it, rr, ok = ctx.solve_batch(B, rtol=1e-10, method=A.PCG_TWO_LEVEL)
    assert ok
    assert it[0, 1] == 0 and np.all(X[0, 1] == 0.0) and rr[0, 1] == 0.0
    assert np.linalg.norm(X[1, 2] - 1e-6 * X[1, 0]) <= 1e-8 * np.linalg.norm(1e-6 * X[1, 0])
    X1, _, _, ok1 = ctx.solve_batch(B[:, :1].copy(), rtol=1e-10, method=A.PCG_TWO_LEVEL)
    assert ok1
    for v in range(2):
        assert np.linalg.norm(X1[v, 0] - X[v, 0]) <= 2e-8 * np.linalg.norm(X[v, 0])
    with pytest.raises(A.OfeaError) as e:
        ctx.solve_batch(B, rtol=1e-10, method=A.PCG_JACOBI)
    assert e.value.code == A.OFEA_ERR_ARG
