"""GPU checks at the full sizes of the five benchmark configurations, through
properties that do not depend on size (the element-level parity runs on
small meshes in the other files).

  * C3D4 60x30x30 Kuhn (324,000 tets): the CSC pattern equals the oracle's
    bit for bit and its nnz the closed form; K x equals the oracle's
    matrix-free sum of Ke x_e (1e-12); K is symmetric; two-level PCG reaches
    a true residual of 1e-10 checked by an independent SpMV;
  * C3D8 100x50x50 explicit (250,000 hexes): two FP64 steps equal the
    oracle's reference recurrence (1e-9), and FP32 stays within 1e-5 of FP64;
  * 64 variants of C3D8 40x20x20: batched values equal the oracle's
    per-element-E fill for the first, middle and last variant (1e-12); all 512
    multi-RHS solves converge, two checked by an independent SpMV;
  * C3D8 80x40x40 TL-SVK (128,000 hexes): K_T(0) = K, f_int(0) = 0; K_T is
    symmetric; central differences of f_int reproduce K_T x.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel_max
from oracle import oracle as O
import openfea_amd as ofa
from openfea_amd import _capi as A, meshgen as mg

pytestmark = pytest.mark.gpu
E, NU = mg.E_STEEL, mg.NU_STEEL


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ofa.Context(0)
    yield c
    c.close()


def csc(cp, ri, v):
    return sp.csc_matrix((v, ri, cp), shape=(len(cp) - 1, len(cp) - 1))


def test_cfg2_full(ctx):
    m, resort, F = mg.cantilever_c3d4()
    say(f"cfg2 full size ne={m.ne}")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    assert ctx.nnz == 7451442
    cpo, rio = O.build_csc(m.conn, m.etype, resort, ctx.n)
    assert np.array_equal(cp, cpo) and np.array_equal(ri, rio)
    val = ctx.assemble(E, NU)
    x = np.random.default_rng(3).normal(size=ctx.n)
    assert rel_max(csc(cp, ri, val) @ x, O.matfree_apply(m.conn, m.etype, m.xyz, resort, E, NU, x, ctx.n)) < 1e-12
    K = csc(cp, ri, val)
    assert abs(K - K.T).max() <= 1e-12 * np.abs(val).max()
    u, it, rel, ok = ctx.solve(F, method=A.PCG_TWO_LEVEL, rtol=1e-10, maxit=100000)
    assert ok and np.linalg.norm(F - K @ u) / np.linalg.norm(F) <= 1e-10 * 1.01, (it, rel)


def test_cfg3_full_explicit(ctx):
    m = mg.hex_block(100, 50, 50)
    say(f"cfg3 full size ne={m.ne}")
    ctx.set_mesh(m.xyz, m.conn, m.etype, mg.free_index(m.nd, m.fix_set))
    mass = ctx.lumped_mass(mg.RHO_STEEL)
    fixed = np.zeros(m.nd, dtype=np.int32)
    fixed[m.fix_set - 1] = 1
    v0 = mg.initial_velocity(m)
    f = np.zeros((m.nd, 3))
    Ke = O.c3d8_ke(m.xyz[m.conn[0] - 1], E, NU)
    me = O.c3d8_mass(m.xyz[m.conn[0] - 1], mg.RHO_STEEL)
    dt = 0.9 * 2.0 / np.sqrt(np.linalg.eigvals(Ke / np.repeat(me, 3)[:, None]).real.max())
    u64 = ctx.explicit_run(E, NU, mass, f, v0, fixed, dt, 2, precision=A.FP64)
    uo = O.explicit_run(m.conn, m.etype, m.xyz, E, NU, mass, f, v0, fixed, dt, 2)
    assert np.linalg.norm(u64 - uo) / np.linalg.norm(uo) < 1e-9
    u64 = ctx.explicit_run(E, NU, mass, f, v0, fixed, dt, 200, precision=A.FP64)
    u32 = ctx.explicit_run(E, NU, mass, f, v0, fixed, dt, 200, precision=A.FP32)
    assert np.linalg.norm(u32 - u64) / np.linalg.norm(u64) < 1e-5


def test_cfg4_full_batched(ctx):
    nvar, nrhs = 64, 8
    m = mg.hex_block(40, 20, 20)
    resort = mg.free_index(m.nd, m.fix_set)
    say(f"cfg4 full size ne={m.ne} x {nvar} variants x {nrhs} cases")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    Ev = np.stack([mg.design_field(m, v) for v in range(nvar)])
    vals = ctx.assemble_batch(Ev, NU)
    for v in (0, nvar // 2, nvar - 1):
        ref = O.fill_csc(m.conn, m.etype, m.xyz, resort, 0.0, NU, cp, ri, E_elem=Ev[v])
        assert rel_max(vals[v], ref) < 1e-12, v
    B = mg.batched_loads(m, resort, nvar, nrhs)
    X, it, rr, ok = ctx.solve_batch(B, rtol=1e-8)
    assert ok and rr.max() <= 1e-8
    for v, k in ((0, 0), (nvar - 1, nrhs - 1)):
        K = csc(cp, ri, vals[v])
        assert np.linalg.norm(B[v, k] - K @ X[v, k]) / np.linalg.norm(B[v, k]) <= 1.01e-8


def test_cfg5_full_tl(ctx):
    m = mg.hex_block(80, 40, 40)
    resort = mg.free_index(m.nd, m.fix_set)
    say(f"cfg5 full size ne={m.ne}")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    K = ctx.assemble(E, NU)
    K0, f0 = ctx.nl_assemble(E, NU, np.zeros(ctx.n))
    assert rel_max(K0, K) < 1e-12 and np.abs(f0).max() <= 1e-12 * np.abs(K).max()
    rng = np.random.default_rng(9)
    u = rng.normal(size=ctx.n) * 1e-2
    x = rng.normal(size=ctx.n) * 1e-2
    KT, _ = ctx.nl_assemble(E, NU, u)
    A_ = csc(cp, ri, KT)
    assert abs(A_ - A_.T).max() <= 1e-12 * np.abs(KT).max()
    eps = 1e-3
    _, fp = ctx.nl_assemble(E, NU, u + eps * x)
    _, fm = ctx.nl_assemble(E, NU, u - eps * x)
    assert rel_max((fp - fm) / (2 * eps), A_ @ x) < 1e-6
