"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars (stated per test): CSC pattern bit-exact; FP64 stiffness values within
1e-12 of max|K| entrywise; PCG displacements within 1e-8 (relative 2-norm) of
a direct solve; explicit FP64 within 1e-9, FP32 within 2e-3 (relative 2-norm)
of the FP64 oracle; TL tangent and internal force within 1e-12.
"""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import GOLDEN, rel_max
from oracle import oracle as O
import openfea_amd as ofa
from openfea_amd import _capi as A, meshgen as mg

pytestmark = pytest.mark.gpu


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


MESHES = {
    "hex_cfg1_20x10x10": lambda: mg.hex_block(20, 10, 10),
    "hex_jitter_6x4x3": lambda: mg.hex_block(6, 4, 3, jitter=0.2, seed=7),
    "tet_kuhn_jitter_8x5x4": lambda: mg.kuhn_tets(8, 5, 4, jitter=0.2, seed=20261203),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_assembly_matches_oracle(ctx, name):
    m = MESHES[name]()
    resort = mg.free_index(m.nd, m.fix_set)
    say(f"assemble {name} ne={m.ne}")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    cpo, rio = O.build_csc(m.conn, m.etype, resort, ctx.n)
    assert np.array_equal(cp, cpo) and np.array_equal(ri, rio)
    val = ctx.assemble(mg.E_STEEL, mg.NU_STEEL)
    ref = O.fill_csc(m.conn, m.etype, m.xyz, resort, mg.E_STEEL, mg.NU_STEEL, cp, ri)
    assert rel_max(val, ref) < 1e-12
    # bitwise deterministic (no atomics)
    assert np.array_equal(ctx.assemble(mg.E_STEEL, mg.NU_STEEL), val)


def test_assembly_mixed_inverted_and_per_element_E(ctx):
    d = ofa.read_inp(os.path.join(GOLDEN, "mesh_o_cae.inp"))  # every tet inverted
    resort, _ = ofa.build_free_index(d.nd, d.fix_set)
    say("assemble mesh_o_cae (inverted tets)")
    ctx.set_mesh(d.xyz, d.conn, d.etype, resort)
    cp, ri = ctx.pattern()
    val = ctx.assemble(2.1e5, 0.3)
    assert rel_max(val, O.fill_csc(d.conn, d.etype, d.xyz, resort, 2.1e5, 0.3, cp, ri)) < 1e-12
    # mixed hex + tet mesh with per-element E
    h = mg.hex_block(3, 2, 2, jitter=0.15, seed=1)
    t = mg.kuhn_tets(3, 2, 2, jitter=0.15, seed=1)
    xyz = h.xyz
    conn = np.vstack([h.conn, t.conn])
    etype = np.concatenate([h.etype, t.etype])
    t.xyz = h.xyz  # same node grid (jitter draws are identical for the same seed)
    resort = mg.free_index(len(xyz), h.fix_set)
    Ee = np.random.default_rng(3).uniform(1e5, 3e5, len(etype))
    say("assemble mixed hex+tet with per-element E")
    ctx.set_mesh(xyz, conn, etype, resort)
    ctx.set_elem_E(Ee)
    cp, ri = ctx.pattern()
    val = ctx.assemble(0.0, 0.3)
    ref = O.fill_csc(conn, etype, xyz, resort, 0.0, 0.3, cp, ri, E_elem=Ee)
    assert rel_max(val, ref) < 1e-12


def test_spmv_matches_oracle(ctx):
    m = mg.kuhn_tets(6, 4, 4, jitter=0.2, seed=2)
    resort = mg.free_index(m.nd, m.fix_set)
    say("spmv tet 6x4x4")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    val = ctx.assemble(mg.E_STEEL, mg.NU_STEEL)
    x = np.random.default_rng(0).normal(size=ctx.n)
    assert rel_max(ctx.spmv(x), O.spmv_csc(cp, ri, val, x)) < 1e-13


@pytest.mark.parametrize("method", [A.PCG_JACOBI, A.PCG_BLOCK_JACOBI, A.PCG_MIXED, A.PCG_MIXED32,
                                    A.PCG_TWO_LEVEL, A.PCG_TWO_LEVEL_MIXED, A.PCG_TWO_LEVEL_MIXED32])
def test_pcg_cfg1_matches_direct_solve(ctx, method):
    m, resort, F = mg.cantilever_c3d8()
    say(f"pcg method={method} cfg1 20x10x10")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    val = ctx.assemble(mg.E_STEEL, mg.NU_STEEL)
    x, it, rel, ok = ctx.solve(F, method=method, rtol=1e-10, maxit=200000)
    assert ok and rel <= 1e-10, (it, rel)
    xd = spla.spsolve(csc(cp, ri, val), F)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    # the oracle's own PCG takes a comparable number of iterations (Jacobi only)
    if method == A.PCG_JACOBI:
        _, it_o, _ = O.pcg_jacobi(cp, ri, val, F, rtol=1e-10)
        assert abs(it - it_o) <= max(10, 0.05 * it_o)
    if method == A.PCG_TWO_LEVEL:  # the coarse correction must pay for itself in iterations
        _, it_bj, _, _ = ctx.solve(F, method=A.PCG_BLOCK_JACOBI, rtol=1e-10, maxit=200000)
        say(f"two-level {it} vs block Jacobi {it_bj} iterations")
        assert it < it_bj


TWO_LEVEL_MESHES = {
    # one aggregate (nc = 6, a single Cholesky panel)
    "hex_2x1x1": lambda: mg.hex_block(2, 1, 1, jitter=0.1, seed=3),
    # a few aggregates, jittered tets
    "tet_5x3x3": lambda: mg.kuhn_tets(5, 3, 3, jitter=0.2, seed=5),
    # long thin block: many slabs of one part
    "hex_40x2x2": lambda: mg.hex_block(40, 2, 2, jitter=0.1, seed=6),
}


@pytest.mark.parametrize("name", list(TWO_LEVEL_MESHES))
def test_two_level_small_and_thin_meshes(ctx, name):
    """Two-level PCG (coarse set-up on the device) reaches the direct solution
    on meshes with one aggregate, a few, and a long row of slabs; the
    aggregates cover the free nodes with all but the last of equal size."""
    from openfea_amd import learned as L

    m = TWO_LEVEL_MESHES[name]()
    resort = mg.free_index(m.nd, m.fix_set)
    say(f"two-level on {name} ne={m.ne}")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    val = ctx.assemble(mg.E_STEEL, mg.NU_STEEL)
    F = ctx.spmv(np.random.default_rng(8).normal(size=ctx.n))
    xd = spla.spsolve(csc(cp, ri, val), F)
    for meth in (A.PCG_TWO_LEVEL, A.PCG_TWO_LEVEL_MIXED, A.PCG_TWO_LEVEL_MIXED32):
        # rtol 1e-12: the thin block's condition number turns a 1e-10 residual into ~4e-8 of error
        x, it, rel, ok = ctx.solve(F, method=meth, rtol=1e-12, maxit=100000)
        assert ok and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8, (meth, it, rel)
    agg, nagg = L.coarse_aggregates(ctx)
    counts = np.bincount(agg, minlength=nagg)
    assert counts.sum() == ctx.n // 3 and len(set(counts[:-1].tolist())) <= 1


def test_coarse_correction_matches_numpy(ctx):
    """Z A_c^-1 Z^T r of the two-level methods against an independent numpy
    construction from the library's aggregates: rigid-body modes about each
    aggregate's centroid, A_c = Z^T K Z.  The aggregates hold exactly
    per * ra nodes (per = ceil(nfree/512) nodes per CG range, ra =
    max(2, ceil(64/per)) ranges) except the last.  The library inverts A_c in
    FP64 and applies it in FP32: bar 1e-5 relative."""
    from openfea_amd import learned as L

    m, resort, F = mg.cantilever_c3d8(12, 4, 4)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    val = ctx.assemble(mg.E_STEEL, mg.NU_STEEL)
    K = csc(cp, ri, val)
    nf = ctx.n // 3
    X = np.zeros((nf, 3))
    X[resort[resort >= 0]] = m.xyz[resort >= 0]
    per = -(-nf // 512)
    ra = max(2, -(-64 // per))  # ranges per aggregate: at least 2 and 64 nodes
    agg, nagg = L.coarse_aggregates(ctx)
    counts = np.bincount(agg, minlength=nagg)
    assert nagg == -(-nf // (per * ra)) and np.all(counts[:-1] == per * ra) and 0 < counts[-1] <= per * ra
    cen = np.stack([np.bincount(agg, weights=X[:, c]) / np.bincount(agg) for c in range(3)], axis=1)
    d = (X - cen[agg]).astype(np.float32).astype(np.float64)
    Z = np.zeros((ctx.n, 6 * nagg))
    for i in range(nf):
        a = agg[i]
        dx, dy, dz = d[i]
        Z[3 * i:3 * i + 3, 6 * a:6 * a + 6] = [[1, 0, 0, 0, dz, -dy], [0, 1, 0, -dz, 0, dx], [0, 0, 1, dy, -dx, 0]]
    Ac = Z.T @ (K @ Z)
    r = np.random.default_rng(4).normal(size=ctx.n)
    # a rotation about the line of a collinear aggregate is a zero column of Z:
    # A_c is then singular and that component of y does not enter Z y
    ref = Z @ np.linalg.lstsq(Ac, Z.T @ r, rcond=1e-13)[0]
    out = L.coarse_apply(ctx, r)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-5


def test_coarse_aggregates_compact_and_effective(ctx):
    """The two-level aggregates are compact parts of the node cloud (their
    bounding boxes are not much longer than wide, where consecutive node
    ranges would give strips through the whole block) and the coarse space
    cuts the block-Jacobi iteration count several times over."""
    from openfea_amd import learned as L

    m, resort, F = mg.cantilever_c3d4(24, 12, 12)
    say("two-level aggregates, C3D4 24x12x12")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    ctx.assemble(mg.E_STEEL, mg.NU_STEEL)
    agg, nagg = L.coarse_aggregates(ctx)
    X = np.zeros((ctx.n // 3, 3))
    X[resort[resort >= 0]] = m.xyz[resort >= 0]
    aspect = []
    for a in range(nagg - 1):
        ext = np.ptp(X[agg == a], axis=0) + 1.0  # +h: a one-layer part has thickness h
        aspect.append(ext.max() / ext.min())
    assert np.median(aspect) <= 2.5, np.median(aspect)
    _, it_bj, _, ok1 = ctx.solve(F, method=A.PCG_BLOCK_JACOBI, rtol=1e-8)
    _, it_tl, _, ok2 = ctx.solve(F, method=A.PCG_TWO_LEVEL, rtol=1e-8)
    say(f"block Jacobi {it_bj} it, two-level {it_tl} it, {nagg} aggregates")
    assert ok1 and ok2 and it_tl * 4 <= it_bj, (it_bj, it_tl)


def test_two_level_tet_and_rhs_reuse(ctx):
    """Two-level PCG on a jittered tet mesh, two right-hand sides in a row
    (the coarse space is built once per K), and after a reassembly with
    another modulus field (the coarse space is rebuilt)."""
    m, resort, F = mg.cantilever_c3d4(12, 6, 6, seed=3)
    say("two-level tet 12x6x6")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    rng = np.random.default_rng(5)
    for E_elem in (None, rng.uniform(1e5, 3e5, m.ne)):
        ctx.set_elem_E(E_elem)
        val = ctx.assemble(mg.E_STEEL, mg.NU_STEEL)
        lu = spla.splu(csc(cp, ri, val))
        for b in (F, rng.normal(size=ctx.n)):
            for method in (A.PCG_TWO_LEVEL, A.PCG_TWO_LEVEL_MIXED):
                x, it, rel, ok = ctx.solve(b, method=method, rtol=1e-10)
                xd = lu.solve(b)
                assert ok and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8, (method, it, rel)
    ctx.set_elem_E(None)


def test_pipeline_reference_deck_vs_pardiso_golden(ctx):
    d = ofa.read_inp(os.path.join(GOLDEN, "mesh_o.inp"))
    d.resort, nfree = ofa.build_free_index(d.nd, d.fix_set)
    d.F = ofa.build_single_load(3 * nfree, d.resort, d.load_set, d.load_dof, d.load_value)
    say("drop-in pipeline mesh_o.inp")
    assam = ofa.AssambleStiffness(ctx)
    assam.build_CSR(d)
    assam.fill_CSR_sparse_mat(d, 2.1e5, 0.3)
    assert ofa.solution_api(assam, d, "GPU_PCG", rtol=1e-10)
    assert not ofa.solution_api(assam, d, "no_such_solver")
    u = ofa.reset_displacement(d.resort, d.single_dis_vec_).reshape(-1, 3)
    ref = np.load(os.path.join(GOLDEN, "cadcae_o_u.npz"))["u_o"]
    assert np.all(np.abs(u - ref) <= 6e-6 * np.abs(ref) + 1e-9)


def test_lumped_mass_matches_oracle(ctx):
    m = mg.hex_block(4, 3, 2, jitter=0.2, seed=4)
    say("lumped mass")
    ctx.set_mesh(m.xyz, m.conn, m.etype, mg.free_index(m.nd, m.fix_set))
    mass = ctx.lumped_mass(mg.RHO_STEEL)
    assert rel_max(mass, O.lumped_mass(m.conn, m.etype, m.xyz, mg.RHO_STEEL)) < 1e-13


@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_explicit_matches_oracle(ctx, kind):
    m = mg.hex_block(8, 3, 3, jitter=0.1, seed=2) if kind == "hex" else mg.kuhn_tets(6, 3, 3, jitter=0.1, seed=2)
    say(f"explicit {kind}")
    ctx.set_mesh(m.xyz, m.conn, m.etype, mg.free_index(m.nd, m.fix_set))
    mass = O.lumped_mass(m.conn, m.etype, m.xyz, mg.RHO_STEEL)
    fixed = np.zeros(m.nd, dtype=np.int32)
    fixed[m.fix_set - 1] = 1
    v0 = mg.initial_velocity(m)
    fext = mg.face_traction(m, (0.0, 0.0, -1.0))
    # 0.5 x the element-wise stability limit of the unit cells
    dt, nsteps = 5e-8, 120
    ref = O.explicit_run(m.conn, m.etype, m.xyz, mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, nsteps)
    u64 = ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, nsteps, precision=A.FP64)
    u32 = ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, nsteps, precision=A.FP32)
    e64 = np.linalg.norm(u64 - ref) / np.linalg.norm(ref)
    e32 = np.linalg.norm(u32 - ref) / np.linalg.norm(ref)
    say(f"explicit {kind}: rel err fp64 {e64:.3e} fp32 {e32:.3e}")
    assert e64 < 1e-9
    assert e32 < 2e-3


@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_tl_tangent_and_force_match_oracle(ctx, kind):
    m = mg.hex_block(5, 3, 3, jitter=0.15, seed=8) if kind == "hex" else mg.kuhn_tets(4, 3, 3, jitter=0.15, seed=8)
    resort = mg.free_index(m.nd, m.fix_set)
    say(f"tl {kind}")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    u = np.random.default_rng(1).normal(0, 0.02, ctx.n)
    val, f = ctx.nl_assemble(mg.E_STEEL, mg.NU_STEEL, u)
    vo, fo = O.nl_fill(m.conn, m.etype, m.xyz, resort, mg.E_STEEL, mg.NU_STEEL, u, cp, ri)
    assert rel_max(val, vo) < 1e-12
    assert rel_max(f, fo) < 1e-12


@pytest.mark.parametrize("method,inner", [(A.PCG_BLOCK_JACOBI, 1e-11), (A.PCG_BLOCK_JACOBI, -1e-11),
                                          (A.PCG_TWO_LEVEL, -1e-11)])
def test_newton_matches_oracle(ctx, method, inner):
    """Load-stepped Newton against the oracle's (direct inner solves): with a
    fixed inner tolerance and with the adaptive (Eisenstat-Walker) one."""
    m, resort, F = mg.cantilever_c3d8(6, 3, 3)
    say(f"newton tl 6x3x3 method {method} inner {inner}")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = ctx.pattern()
    load = 40 * F
    u, nt, lt, rel, ok = ctx.nl_solve(mg.E_STEEL, mg.NU_STEEL, load, n_incr=3, tol=1e-10, max_newton=20,
                                      method=method, inner_rtol=inner)
    uo, it_o = O.nl_newton(m.conn, m.etype, m.xyz, resort, mg.E_STEEL, mg.NU_STEEL, load, cp, ri, n_incr=3, tol=1e-10)
    say(f"newton: gpu {nt} iterations ({lt} PCG), oracle {it_o}; rel {rel:.2e}")
    assert ok and rel <= 1e-10
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8
