"""GPU parity of the node-based smoothed FEM through the C ABI.

Bars: smoothed pattern bit-exact with the oracle; values within 1e-12 of
max|K| entrywise (the device sums the expanded form, the oracle the
reference's six-term loop); bitwise deterministic; PCG solutions within 1e-8
of a direct solve of the oracle's K.
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
E0, NU0 = 2.1e5, 0.3


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ofa.Context(0)
    yield c
    c.close()


def mesh_o():
    d = ofa.read_inp(os.path.join(GOLDEN, "mesh_o.inp"))
    return d.xyz, d.conn, d.etype, mg.free_index(d.nd, np.sort(d.fix_set))


CASES = {
    "kuhn_jitter_6x3x3": lambda: (lambda m: (m.xyz, m.conn, m.etype, mg.free_index(m.nd, m.fix_set)))(
        mg.kuhn_tets(6, 3, 3, jitter=0.2, seed=9)),
    "mesh_o": mesh_o,
}


@pytest.mark.parametrize("name", list(CASES))
def test_sfem_matches_oracle(ctx, name):
    xyz, conn, etype, resort = CASES[name]()
    say(f"sfem {name} ne={len(etype)}")
    ctx.set_mesh(xyz, conn, etype, resort)
    ctx.sfem_setup()
    cp, ri = ctx.pattern()
    cpo, rio = O.sfem_build_csc(conn, xyz, resort)
    assert np.array_equal(cp, cpo) and np.array_equal(ri, rio)
    val = ctx.sfem_assemble(E0, NU0)
    ref = O.sfem_fill(conn, xyz, resort, E0, NU0, cpo, rio)
    assert rel_max(val, ref) < 1e-12
    assert np.array_equal(ctx.sfem_assemble(E0, NU0), val)
    n = len(cp) - 1
    rng = np.random.default_rng(1)
    b = rng.normal(size=n)
    xd = spla.spsolve(sp.csc_matrix((ref, rio, cpo), shape=(n, n)), b)
    for method in (A.PCG_BLOCK_JACOBI, A.PCG_TWO_LEVEL):
        x, it, rel, ok = ctx.solve(b, method=method, rtol=1e-11)
        assert ok and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8, (method, it, rel)
    # the FEM entry points refuse the smoothed context until the mesh is set again
    with pytest.raises(A.OfeaError):
        ctx.assemble(E0, NU0)
    ctx.set_mesh(xyz, conn, etype, resort)
    ctx.assemble(E0, NU0, copy=False)


def test_sfem_pipeline_deck(tmp_path):
    say("implict_SFEManalysis mesh_o.inp")
    d = ofa.read_inp(os.path.join(GOLDEN, "mesh_o.inp"))
    u = ofa.implict_SFEManalysis(d, str(tmp_path / "sfem.vtk"), E=E0, nu=NU0, rtol=1e-10)
    xyz, conn, etype, resort = mesh_o()
    cp, ri = O.sfem_build_csc(conn, xyz, resort)
    n = len(cp) - 1
    F = ofa.build_single_load(n, resort, d.load_set, d.load_dof, d.load_value)
    x = spla.spsolve(sp.csc_matrix((O.sfem_fill(conn, xyz, resort, E0, NU0, cp, ri), ri, cp), shape=(n, n)), F)
    uo = O.reset_displacement(resort, x).reshape(-1, 3)
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-7  # cond(K) ~ 1e6 on this deck
