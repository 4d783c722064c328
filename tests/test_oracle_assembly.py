"""Pins the oracle's pattern, fill, BC, solve and explicit routines (CPU only).

  * the whole linear pipeline (reader -> BCs -> pattern -> fill -> direct
    solve -> reset_displacement) against the reference's own PARDISO output
    for sample_model/CA/mesh_o.inp (golden, 6 significant digits) and its
    inverted twin mesh_o_cae.inp (signed detJ => exactly -u);
  * pattern: brute force from the node adjacency, and the closed-form nnz of
    the structured hex and Kuhn tet grids;
  * fill: K x == sum_e Ke x_e (matrix-free identity) and symmetry;
  * PCG: converges to the direct solution;
  * explicit: the reference's incremental -Ke du recurrence equals the
    direct recurrence with f = -K u, and the central-difference shadow energy
    is conserved.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import GOLDEN, rel_max
from oracle import oracle as O
from openfea_amd import meshgen as mg


def parse_inp_minimal(path):
    """Independent minimal parse of the reference deck dialect (for the test only)."""
    lines = open(path).read().splitlines()
    xyz, conn, sets, cload = [], [], {}, None
    mode = None
    for i, ln in enumerate(lines):
        s = ln.strip()
        if s.startswith("*"):
            mode = None
            if s.startswith("*Node"):
                mode = "node"
            elif s.startswith("*Element"):
                mode = "elem"
            elif "nset=Set-" in s:
                mode = s.split("nset=")[1].strip()
                sets[mode] = []
            elif s.startswith("*Cload"):
                p = lines[i + 1].split(",")
                cload = (int(p[1]), float(p[2]))
            continue
        if mode == "node":
            xyz.append([float(t) for t in s.split(",")[1:4]])
        elif mode == "elem":
            conn.append([int(t) for t in s.split(",")[1:]])
        elif mode in sets:
            sets[mode].extend(int(t) for t in s.replace(",", " ").split())
    return np.array(xyz), np.array(conn), sets, cload


def csc_matrix(col_ptr, row_idx, val):
    n = len(col_ptr) - 1
    return sp.csc_matrix((val, row_idx, col_ptr), shape=(n, n))


@pytest.mark.parametrize("deck,key,sign", [("mesh_o.inp", "u_o", 1.0), ("mesh_o_cae.inp", "u_o_cae", 1.0)])
def test_pipeline_matches_reference_pardiso(deck, key, sign):
    xyz, tets, sets, (dof, val) = parse_inp_minimal(os.path.join(GOLDEN, deck))
    nd = len(xyz)
    conn = np.zeros((len(tets), 8), dtype=np.int32)
    conn[:, :4] = tets
    etype = np.zeros(len(tets), dtype=np.int32)
    resort, nfree = O.free_index(nd, np.array(sets["Set-fix"]))
    n = 3 * nfree
    F = O.single_load(n, resort, np.array(sets["Set-load"]), dof, val)
    cp, ri = O.build_csc(conn, etype, resort, n)
    K = O.fill_csc(conn, etype, xyz, resort, 2.1e5, 0.3, cp, ri)  # material of src/main.cpp:121
    u = spla.spsolve(csc_matrix(cp, ri, K), F)
    uf = O.reset_displacement(resort, u).reshape(-1, 3)
    ref = np.load(os.path.join(GOLDEN, "cadcae_o_u.npz"))[key]
    # the reference printed 6 significant digits
    assert rel_max(sign * uf, ref) < 2e-5
    assert np.all(np.abs(uf - ref) <= 6e-6 * np.abs(ref) + 1e-9)


def test_inverted_deck_gives_negated_solution():
    ref = np.load(os.path.join(GOLDEN, "cadcae_o_u.npz"))
    assert np.array_equal(ref["u_o"], -ref["u_o_cae"])


def brute_pattern(conn, etype, resort):
    n = 3 * int(resort.max() + 1)
    cols = [set() for _ in range(n)]
    for e in range(len(etype)):
        nn = 8 if etype[e] == 1 else 4
        dofs = [3 * resort[conn[e, a] - 1] + c for a in range(nn) for c in range(3) if resort[conn[e, a] - 1] >= 0]
        for r in dofs:
            for c in dofs:
                cols[c].add(r)
    cp = np.zeros(n + 1, dtype=np.int64)
    ri = []
    for j in range(n):
        rows = sorted(cols[j])
        ri.extend(rows)
        cp[j + 1] = cp[j] + len(rows)
    return cp, np.array(ri)


@pytest.mark.parametrize("mk", [lambda: mg.hex_block(3, 2, 2), lambda: mg.kuhn_tets(2, 2, 3, jitter=0.2, seed=4)])
def test_pattern_brute_force(mk):
    m = mk()
    resort = mg.free_index(m.nd, m.fix_set)
    cp, ri = O.build_csc(m.conn, m.etype, resort, 3 * int(resort.max() + 1))
    bcp, bri = brute_pattern(m.conn, m.etype, resort)
    assert np.array_equal(cp, bcp) and np.array_equal(ri, bri)


def hex_nnz(nx, ny, nz):
    return 9 * (3 * nx - 2) * (3 * ny + 1) * (3 * nz + 1)


def kuhn_nnz(nx, ny, nz):
    # offsets d in {0} U {+-(subset sums of e1,e2,e3)} minus the clamped x=0 plane
    offs = {(0, 0, 0)}
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                if (a, b, c) != (0, 0, 0):
                    offs.add((a, b, c))
                    offs.add((-a, -b, -c))
    tot = 0
    for dx, dy, dz in offs:
        tot += max(0, nx - abs(dx)) * max(0, ny + 1 - abs(dy)) * max(0, nz + 1 - abs(dz))
    return 9 * tot


@pytest.mark.parametrize("dims", [(4, 2, 2), (5, 3, 2), (20, 10, 10)])
def test_pattern_nnz_closed_forms(dims):
    for m, f in ((mg.hex_block(*dims), hex_nnz), (mg.kuhn_tets(*dims), kuhn_nnz)):
        resort = mg.free_index(m.nd, m.fix_set)
        _, ri = O.build_csc(m.conn, m.etype, resort, 3 * int(resort.max() + 1))
        assert len(ri) == f(*dims)


def test_fill_matrix_free_identity_and_symmetry():
    rng = np.random.default_rng(2)
    m = mg.kuhn_tets(3, 2, 2, jitter=0.2, seed=9)
    h = mg.hex_block(2, 2, 2, jitter=0.2, seed=3)
    for mesh in (m, h):
        resort = mg.free_index(mesh.nd, mesh.fix_set)
        n = 3 * int(resort.max() + 1)
        cp, ri = O.build_csc(mesh.conn, mesh.etype, resort, n)
        Ee = rng.uniform(1e5, 3e5, mesh.ne)
        val = O.fill_csc(mesh.conn, mesh.etype, mesh.xyz, resort, 0.0, 0.3, cp, ri, E_elem=Ee)
        K = csc_matrix(cp, ri, val)
        x = rng.normal(size=n)
        y = O.matfree_apply(mesh.conn, mesh.etype, mesh.xyz, resort, 0.0, 0.3, x, n, E_elem=Ee)
        assert rel_max(K @ x, y) < 1e-13
        assert rel_max(O.spmv_csc(cp, ri, val, x), K @ x) < 1e-14
        assert abs(K - K.T).max() < 1e-12 * abs(K).max()


def test_pcg_converges_to_direct_solution():
    m, resort, F = mg.cantilever_c3d8(6, 3, 3)
    n = len(F)
    cp, ri = O.build_csc(m.conn, m.etype, resort, n)
    val = O.fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri)
    x, it, rr = O.pcg_jacobi(cp, ri, val, F, rtol=1e-12)
    xd = spla.spsolve(csc_matrix(cp, ri, val), F)
    assert rr <= 1e-12 and 0 < it < n
    assert rel_max(x, xd) < 1e-9


def test_explicit_incremental_equals_direct_and_conserves_energy():
    m = mg.hex_block(4, 2, 2, jitter=0.1, seed=5)
    rho = mg.RHO_STEEL
    mass = O.lumped_mass(m.conn, m.etype, m.xyz, rho)
    assert np.isclose(mass.sum(), rho * 4 * 2 * 2, rtol=1e-13)
    fixed = np.zeros(m.nd, dtype=np.int32)
    fixed[m.fix_set - 1] = 1
    v0 = mg.initial_velocity(m)
    fext = np.zeros((m.nd, 3))
    dt, nsteps = 2e-8, 40
    u = O.explicit_run(m.conn, m.etype, m.xyz, 210000.0, 0.3, mass, fext, v0, fixed, dt, nsteps)
    # direct recurrence on all DOFs with K from the pattern of an unclamped mesh
    allfree = np.arange(m.nd, dtype=np.int32)
    cp, ri = O.build_csc(m.conn, m.etype, allfree, 3 * m.nd)
    K = csc_matrix(cp, ri, O.fill_csc(m.conn, m.etype, m.xyz, allfree, 210000.0, 0.3, cp, ri))
    M = np.repeat(mass, 3)
    fx = np.repeat(fixed, 3).astype(bool)
    v = v0.ravel() + 0.0 * dt / 2
    ud = np.zeros(3 * m.nd)
    for _ in range(nsteps):
        ud = ud + v * dt
        ud[fx] = 0.0
        v = v - (K @ ud) / M * dt
    assert rel_max(u, ud) < 1e-12
    # the C form of the direct recurrence (used as the long-run reference):
    # equal to the incremental loop, and the same bits on 1 and 4 threads
    fixed3 = np.repeat(fixed, 3).astype(np.int32)
    minv = 1.0 / M
    kv = O.fill_csc(m.conn, m.etype, m.xyz, allfree, 210000.0, 0.3, cp, ri)
    uc1 = O.explicit_recurrence(cp, ri, kv, minv, fixed3, v0.ravel(), dt, nsteps, threads=1)
    uc4 = O.explicit_recurrence(cp, ri, kv, minv, fixed3, v0.ravel(), dt, nsteps, threads=4)
    assert rel_max(uc1, u) < 1e-12 and np.array_equal(uc1, uc4)
    # shadow energy of central differences, free DOFs only: 1/2 v+ M v- + 1/2 u K u
    Kf = K[~fx][:, ~fx]
    Mf = M[~fx]
    v = v0.ravel()[~fx]
    uu = np.zeros(Mf.size)
    energies = []
    for _ in range(nsteps):
        uu = uu + v * dt
        vn = v - (Kf @ uu) / Mf * dt
        energies.append(0.5 * np.sum(Mf * v * vn) + 0.5 * uu @ (Kf @ uu))
        v = vn
    energies = np.array(energies)
    assert np.abs(energies - energies[0]).max() < 1e-10 * abs(energies[0])


def test_newton_small_load_matches_linear_and_converges():
    import scipy.sparse.linalg as spla2
    m, resort, F = mg.cantilever_c3d8(4, 2, 2)
    n = len(F)
    cp, ri = O.build_csc(m.conn, m.etype, resort, n)
    # strains ~1e-5: the Green-Lagrange strain (F^T F - I)/2 of the reference
    # formulation cancels to ~1e-16/1e-5 relative, so Newton tolerance 1e-9
    eps = 1e-2
    u_nl, it = O.nl_newton(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, eps * F, cp, ri, n_incr=1, tol=1e-9)
    K = csc_matrix(cp, ri, O.fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri))
    u_lin = spla2.spsolve(K, F)
    assert np.linalg.norm(u_nl / eps - u_lin) / np.linalg.norm(u_lin) < 1e-3   # geometric effect is O(strain)
    # a large load: quadratic convergence to a tight residual within a few steps per increment
    u_big, it_big = O.nl_newton(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, 50 * F, cp, ri, n_incr=2)
    _, fint = O.nl_fill(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, u_big, cp, ri)
    assert np.linalg.norm(50 * F - fint) <= 1e-10 * np.linalg.norm(50 * F)
    assert it_big <= 16
