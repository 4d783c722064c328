"""GPU: learned block-Jacobi correction against its float64 oracle.

Bars: forward blocks within 1e-5 (max-norm relative) of the oracle evaluated
at the same FP32 parameters; loss within 1e-4 and gradient within 2e-3
(max-norm relative, FP32 training arithmetic); zero head == block Jacobi;
training lowers the loss; learned PCG reaches the direct solution within 1e-8.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import rel_max
from oracle import learned_oracle as LO
from oracle import oracle as O
import openfea_amd as ofa
from openfea_amd import _capi as A, learned as L, meshgen as mg

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    m, resort, F = mg.cantilever_c3d8(6, 3, 3)
    ctx = ofa.Context(0)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    val = ctx.assemble(mg.E_STEEL, mg.NU_STEEL)
    g = LO.Graph(cp, ri, val)
    yield ctx, g, F, cp, ri, val
    ctx.close()


def random_params(seed, head=0.2):
    rng = np.random.default_rng(seed)
    d = LO.unpack(rng.normal(0, 0.3, LO.NPARAM))
    d["Wo"] *= head
    d["bo"] *= head
    return LO.pack(d).astype(np.float32)


def test_forward_matches_oracle(setup):
    ctx, g, F, *_ = setup
    say("learned forward")
    ln = L.Learned(ctx, seed=3)
    assert L.nparam() == LO.NPARAM
    blocks, feats = ln.blocks()
    assert rel_max(blocks, np.linalg.inv(g.D)) < 1e-12          # zero head: block Jacobi
    assert rel_max(feats, g.features()) < 1e-6
    p = random_params(5)
    ln.set_params(p)
    blocks, _ = ln.blocks()
    M, _ = LO.forward(g, p.astype(np.float64))
    assert rel_max(blocks, M) < 1e-5


def test_loss_grad_matches_oracle(setup):
    ctx, g, F, *_ = setup
    say("learned loss/grad")
    ln = L.Learned(ctx, seed=3)
    p = random_params(6)
    ln.set_params(p)
    rng = np.random.default_rng(1)
    E = rng.normal(size=(4, g.n))
    R = np.array([g.K @ e for e in E])
    ln.set_samples(E, R)
    loss, grad = ln.loss_grad()
    lo, go = LO.loss_grad(g, p.astype(np.float64), E, R)
    say(f"loss gpu {loss:.6e} oracle {lo:.6e}")
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert rel_max(grad, go) < 2e-3


def test_training_and_learned_pcg(setup):
    ctx, g, F, cp, ri, val = setup
    say("learned training")
    rng = np.random.default_rng(2)
    rhs = [F] + [ctx.spmv(rng.normal(size=ctx.n)) for _ in range(2)]
    E, R = L.collect_pairs(ctx, rhs, snapshots=(4, 8, 16, 32))
    ln = L.Learned(ctx, seed=1)
    ln.set_samples(E, R)
    hist = ln.train(steps=60, lr=3e-3)
    say(f"loss {hist[0]:.4e} -> {hist[-1]:.4e}")
    assert hist[-1] < hist[0]
    x, it, rel, ok = ctx.solve(F, method=A.PCG_LEARNED, rtol=1e-10)
    xd = spla.spsolve(sp.csc_matrix((val, ri, cp)), F)
    assert ok and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    _, it_bj, _, _ = ctx.solve(F, method=A.PCG_BLOCK_JACOBI, rtol=1e-10)
    say(f"iterations learned {it} vs block Jacobi {it_bj}")
    x2, it2, rel2, ok2 = ctx.solve(F, method=A.PCG_LEARNED_MIXED, rtol=1e-10)
    assert ok2 and np.linalg.norm(x2 - xd) / np.linalg.norm(xd) < 1e-8
    # the two-level variant: trained on the two-level solver's own iterates with
    # the coarse correction removed from the targets
    E2, R2 = L.collect_pairs(ctx, rhs, snapshots=(4, 8, 16, 32), two_level=True)
    ln2 = L.Learned(ctx, seed=1)
    ln2.set_samples(E2, R2)
    hist2 = ln2.train(steps=60, lr=3e-3)
    assert hist2[-1] < hist2[0]
    _, it_tl, _, _ = ctx.solve(F, method=A.PCG_TWO_LEVEL, rtol=1e-10)
    for meth in (A.PCG_LEARNED_TWO_LEVEL, A.PCG_LEARNED_TWO_LEVEL_MIXED):
        x3, it3, rel3, ok3 = ctx.solve(F, method=meth, rtol=1e-10)
        say(f"learned two-level method {meth}: {it3} iterations (two-level block Jacobi {it_tl}); "
            f"loss {hist2[0]:.4e} -> {hist2[-1]:.4e}")
        assert ok3 and np.linalg.norm(x3 - xd) / np.linalg.norm(xd) < 1e-8


def test_bf16_inference_on_matrix_cores(setup):
    """BF16 inference (v_mfma_f32_16x16x16_bf16) against the oracle with the
    hidden layer's inputs and weights rounded to BF16: 2e-4 and at least ten
    times closer than to the FP32 forward (which pins the operand layout);
    PCG with the BF16 blocks still reaches the direct solution."""
    ctx, g, F, cp, ri, val = setup
    say("learned BF16 inference")
    ln = L.Learned(ctx, seed=3)
    p = random_params(7, head=0.5)
    ln.set_params(p)
    b32, _ = ln.blocks()
    ln.precision(A.BF16)
    b16, _ = ln.blocks()
    Me, _ = LO.forward(g, p.astype(np.float64), bf16_hidden=True)
    M32, _ = LO.forward(g, p.astype(np.float64))
    e16, e32 = rel_max(b16, Me), rel_max(b16, M32)
    say(f"bf16 blocks: {e16:.2e} from the BF16 oracle, {e32:.2e} from FP32")
    assert e16 < 2e-4 and 10 * e16 < e32
    assert rel_max(b32, M32) < 1e-5
    x, it, rel, ok = ctx.solve(F, method=A.PCG_LEARNED, rtol=1e-10, maxit=100000)
    xd = spla.spsolve(sp.csc_matrix((val, ri, cp), shape=(len(cp) - 1,) * 2), F)
    assert ok and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    ln.precision(A.FP32)
    assert rel_max(ln.blocks()[0], b32) == 0.0


def test_newton_learned_judged_by_block_jacobi(setup):
    """Newton with a learned method whose network is far from block Jacobi
    (large random head): plain block-Jacobi PCG judges it on the first linear
    solve and serves the rest of the run, so the total PCG count stays near
    the plain run's and the solution equals the oracle's Newton solution."""
    ctx, g, F, cp, ri, val = setup
    m, resort, _ = mg.cantilever_c3d8(6, 3, 3)
    say("newton: learned method judged by block Jacobi")
    load = 40 * F
    u_bj, nt_bj, lt_bj, rel_bj, ok_bj = ctx.nl_solve(mg.E_STEEL, mg.NU_STEEL, load, n_incr=3, tol=1e-10,
                                                     max_newton=20, method=A.PCG_BLOCK_JACOBI, inner_rtol=1e-11)
    ln = L.Learned(ctx, seed=3)
    ln.set_params(random_params(5, head=3.0))
    u_l, nt_l, lt_l, rel_l, ok_l = ctx.nl_solve(mg.E_STEEL, mg.NU_STEEL, load, n_incr=3, tol=1e-10,
                                                max_newton=20, method=A.PCG_LEARNED, inner_rtol=1e-11)
    uo, _ = O.nl_newton(m.conn, m.etype, m.xyz, resort, mg.E_STEEL, mg.NU_STEEL, load, cp, ri, n_incr=3, tol=1e-10)
    say(f"PCG iterations: block Jacobi {lt_bj}, learned (judged) {lt_l}; Newton {nt_bj} / {nt_l}")
    assert ok_bj and ok_l and rel_l <= 1e-10
    assert np.linalg.norm(u_l - uo) / np.linalg.norm(uo) < 1e-8
    assert lt_l <= 2 * lt_bj, (lt_l, lt_bj)
    ctx.assemble(mg.E_STEEL, mg.NU_STEEL)  # leave the linear K resident for later tests
