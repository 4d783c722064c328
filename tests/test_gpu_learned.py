"""GPU: the learned factorized correction of block Jacobi against its float64
oracle, and what the trained network gains.

Model (learned.hip): M = G^T G, G_ii = L_i^-1 (D_i = L_i L_i^T), G_is =
-alpha_is L_i^-1 K_is D_s^-1 for lower neighbours s < i, alpha_is =
2 sigmoid(u_i + v_s) from a 2-layer node-graph network (hidden 64).  Loss
tr(G K G^T) / n (Kaporin's condition number of G K G^T up to a constant).

Bars: the network outputs within 1e-5 of the oracle's at the same FP32
parameters (FP32 network arithmetic); the solver's factor G, stored in BF16,
within one BF16 rounding of the oracle's (2^-8 relative, plus 1e-6 of the
largest entry for entries that sit near a rounding tie); alpha -> 0 gives
the block-Jacobi factor; loss within 1e-6 and gradient within 2e-3 (FP32
backward pass); BF16 inference of the hidden layer within 2e-4 of the oracle
with BF16-rounded hidden inputs and ten times closer to it than to FP32.
The gain: a network trained on config-4 style systems (C3D8 blocks with
per-element modulus fields, four design variants) needs fewer PCG iterations
and less time than block-Jacobi PCG (the same solver without the network, its
fallback and judge) on meshes it was not trained on, a larger C3D8 variant
and a jittered C3D4 block, and reaches the direct solution within 1e-8.
"""
import sys
import time

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


def oracle_factor(g, p, bf16_hidden=False):
    slots, c = LO.forward(g, np.asarray(p, dtype=np.float64), bf16_hidden)
    return np.concatenate([s.reshape(-1) for s in slots]), c["O"]


def check_bf16_factor(G, ref):
    """G (BF16 entries widened) against the float64 factor: one BF16 rounding"""
    tol = 2.0 ** -8 * np.abs(ref) + 1e-6 * np.abs(ref).max()
    bad = np.abs(G - ref) > tol
    assert not bad.any(), (int(bad.sum()), float(np.abs(G - ref).max()))
    assert np.array_equal(G, LO.bf16(G))  # the entries are BF16 values


def test_factor_matches_oracle(setup):
    ctx, g, F, *_ = setup
    say("learned factor")
    ln = L.Learned(ctx, seed=3)
    assert L.nparam() == LO.NPARAM
    G, feats = ln.factor()
    assert rel_max(feats, g.features()) < 1e-6
    ref, Oref = oracle_factor(g, ln.get_params())
    assert rel_max(ln.outputs(), Oref) < 1e-5
    check_bf16_factor(G, ref)
    d = LO.unpack(ln.get_params().astype(np.float64))
    d["bo"][0] = -60.0  # alpha -> 0: G = block-diagonal L^-1, M = block Jacobi
    ln.set_params(LO.pack(d).astype(np.float32))
    G0, _ = ln.factor()
    Gd = np.concatenate([np.stack([g.Linv[i] if s == i else np.zeros((3, 3)) for s in g.nbrs[i]]).reshape(-1)
                         for i in range(g.N)])
    check_bf16_factor(G0, Gd)
    p = random_params(5)
    ln.set_params(p)
    G1, _ = ln.factor()
    ref1, Oref1 = oracle_factor(g, p)
    assert rel_max(ln.outputs(), Oref1) < 1e-5
    check_bf16_factor(G1, ref1)


def test_loss_grad_matches_oracle(setup):
    ctx, g, F, *_ = setup
    say("learned loss/grad")
    ln = L.Learned(ctx, seed=3)
    d = LO.unpack(ln.get_params().astype(np.float64))
    d["bo"][0] = -60.0
    ln.set_params(LO.pack(d).astype(np.float32))
    l0, _ = ln.loss_grad()
    assert abs(l0 - 1.0) < 1e-12  # alpha = 0: tr(L^-1 D L^-T) = 3 per node
    p = random_params(6)
    ln.set_params(p)
    loss, grad = ln.loss_grad()
    lo, go = LO.loss_grad(g, p.astype(np.float64))
    say(f"loss gpu {loss:.9f} oracle {lo:.9f}")
    assert abs(loss - lo) <= 1e-6 * abs(lo)
    assert rel_max(grad, go) < 2e-3


def test_bf16_inference_on_matrix_cores(setup):
    """BF16 inference (v_mfma_f32_16x16x16_bf16) against the oracle with the
    hidden layer's inputs and weights rounded to BF16; PCG with the BF16
    network still reaches the direct solution."""
    ctx, g, F, cp, ri, val = setup
    say("learned BF16 inference")
    ln = L.Learned(ctx, seed=3)
    p = random_params(7, head=0.5)
    ln.set_params(p)
    o32 = ln.outputs().copy()
    ln.precision(A.BF16)
    o16 = ln.outputs().copy()
    _, O16 = oracle_factor(g, p, bf16_hidden=True)
    _, O32 = oracle_factor(g, p)
    e16, e32 = rel_max(o16, O16), rel_max(o16, O32)
    say(f"bf16 outputs: {e16:.2e} from the BF16 oracle, {e32:.2e} from FP32")
    assert e16 < 2e-4 and 10 * e16 < e32
    x, it, rel, ok = ctx.solve(F, method=A.PCG_LEARNED, rtol=1e-10, maxit=100000)
    xd = spla.spsolve(sp.csc_matrix((val, ri, cp), shape=(len(cp) - 1,) * 2), F)
    assert ok and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    ln.precision(A.FP32)
    assert np.array_equal(ln.outputs(), o32)


@pytest.mark.parametrize("method", [A.PCG_LEARNED, A.PCG_LEARNED_MIXED, A.PCG_LEARNED_TWO_LEVEL,
                                    A.PCG_LEARNED_TWO_LEVEL_MIXED, A.PCG_LEARNED_TWO_LEVEL_MIXED32])
def test_learned_methods_match_direct(setup, method):
    """Every learned method (FP64, FP32 values, two-level, FP32 values and
    vectors with the coarse space) with a non-trivial network reaches the
    direct solution within 1e-8 at rtol 1e-10 (true FP64 residual)."""
    ctx, g, F, cp, ri, val = setup
    ln = L.Learned(ctx, seed=4)
    ln.set_params(random_params(6))
    ctx.set_rhs(F)
    x, it, rel, ok = ctx.solve(F, method=method, rtol=1e-10, maxit=100000)
    xd = spla.spsolve(sp.csc_matrix((val, ri, cp), shape=(len(cp) - 1,) * 2), F)
    say(f"learned method {method}: {it} iterations, relres {rel:.2e}")
    assert ok and rel <= 1e-10
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8


def timed_solve(ctx, F, method, reps=3):
    best = None
    for _ in range(reps):
        ctx.sync()
        t = time.perf_counter()
        x, it, rel, ok = ctx.solve(F, method=method, rtol=1e-8, maxit=100000)
        el = time.perf_counter() - t
        best = el if best is None else min(best, el)
    return x, it, ok, best


def train_on_variants(steps=150, seed=1):
    """A network trained on config-4 style systems: a 16x8x8 C3D8 block with
    the per-element modulus fields of design variants 0..3, one context each."""
    m = mg.hex_block(16, 8, 8)
    resort = mg.free_index(m.nd, m.fix_set)
    ctxs, lns = [], []
    for v in range(4):
        c = ofa.Context(0)
        c.set_mesh(m.xyz, m.conn, m.etype, resort)
        c.set_elem_E(mg.design_field(m, v))
        c.assemble(0.0, mg.NU_STEEL, copy=False)
        ctxs.append(c)
        lns.append(L.Learned(c, seed=seed))
    hist = L.train_multi(lns, steps=steps, lr=1e-2)
    p = lns[0].get_params()
    for c in ctxs:
        c.close()
    return p, hist


def _alpha_one(p):
    """parameters whose outputs give alpha = 1 at every pair (head zero): the
    untrained first-order rule, reported beside the trained network"""
    d = LO.unpack(np.asarray(p, dtype=np.float64))
    d["Wo"][:] = 0.0
    d["bo"][:] = 0.0
    return LO.pack(d).astype(np.float32)


def test_trained_network_gains_on_held_out_meshes():
    say("learned: training on config-4 style systems")
    p, hist = train_on_variants()
    say(f"  loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    assert hist[-1] < hist[0]
    cases = []
    m = mg.hex_block(24, 12, 12)
    resort = mg.free_index(m.nd, m.fix_set)
    cases.append(("C3D8 24x12x12, design variant 7", m, resort, mg.batched_loads(m, resort, 8, 1)[7, 0],
                  mg.design_field(m, 7)))
    m2, r2, F2 = mg.cantilever_c3d4(24, 12, 12, seed=11)
    cases.append(("C3D4 24x12x12 jittered", m2, r2, F2, None))
    for name, m, resort, F, Ee in cases:
        with ofa.Context(0) as ctx:
            ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
            cp, ri = ctx.pattern()
            ctx.set_elem_E(Ee)
            val = ctx.assemble(0.0 if Ee is not None else mg.E_STEEL, mg.NU_STEEL)
            ln = L.Learned(ctx, seed=1)
            ln.set_params(p)
            loss_t, _ = ln.loss_grad()
            xb, itb, okb, tb = timed_solve(ctx, F, A.PCG_BLOCK_JACOBI)
            xl, itl, okl, tl = timed_solve(ctx, F, A.PCG_LEARNED)
            ln.set_params(_alpha_one(p))
            loss_1, _ = ln.loss_grad()
            _, it1, _, t1 = timed_solve(ctx, F, A.PCG_LEARNED)
            a = 2.0 / (1.0 + np.exp(-(ln.outputs()[:, 0].mean())))
            xd = spla.spsolve(sp.csc_matrix((val, ri, cp), shape=(ctx.n, ctx.n)), F)
            err = np.linalg.norm(xl - xd) / np.linalg.norm(xd)
            say(f"  {name}: block Jacobi {itb} it {tb * 1e3:.2f} ms; trained net {itl} it {tl * 1e3:.2f} ms "
                f"(loss {loss_t:.4f}); alpha = 1 everywhere: {it1} it {t1 * 1e3:.2f} ms (loss {loss_1:.4f}); "
                f"|x - x_direct| / |x_direct| = {err:.1e}")
            assert okb and okl and err < 1e-8
            assert itl < itb and tl < tb


def test_newton_learned_judged_by_block_jacobi(setup):
    """Newton with a learned method whose network is far from any sensible
    setting (large random head): block-Jacobi PCG judges it on the first linear
    solve; the solution equals the oracle's Newton solution either way."""
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
