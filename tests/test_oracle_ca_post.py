"""Pins the oracle's CA reanalysis and nodal strain/stress routines, and the
host-side CA reader, CA displacement mapping and VTK writer (CPU only).

  * CA: the whole reanalysis of sample_model/CA (mesh_o.inp, 384 tets of
    CA_info.txt deleted) against the reference's own output cadcae_ca.vtk
    (6 significant digits).  The reference driver solves K0 with
    "Pardiso_class" (src/cae.cpp:274), so the SuperLU factors that
    ca_build_rom's series needs (src/ca_analysis.cpp:160) were never made,
    superlu_solution_next returns false without writing (phase_ = 0,
    src/solver/solver_superlu.cpp:76-109) and every series vector stays 0:
    the printed result is the one-vector Rayleigh-Ritz solution, i.e. CA with
    n_basis = 1, except at free DOFs 1-3: Eigen's JacobiSVD of [u0, 0, 0, 0]
    (float, thin U, QR preconditioner, src/ca_analysis.cpp:174-197) returns
    u0/|u0| plus three Householder columns e_j - tau v v_j, which let those
    DOFs relax; restating that span reproduces every printed value;
  * CA with the series: converges to the direct solve of the modified
    system; exact (1e-8) once the Krylov space is complete (one deleted tet:
    rank Delta K = 6, n_basis = 7);
  * Delta K: deleting every element gives -K0; a move to the same place
    gives 0; add-then-delete of one element gives 0;
  * nodal strain: the pipeline's S_* fields against cadcae_o.vtk; linear
    displacement fields give the exact constant strain on distorted tets and
    hexes (patch test); stress = Lame form of the strain.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import GOLDEN, rel_max
from oracle import oracle as O
from openfea_amd import meshgen as mg

E0, NU0 = 2.1e5, 0.3  # material of src/main.cpp:121


def parse_ca_info_minimal(path):
    """Independent minimal parse of the CA change file (test only)."""
    nodes, mods, idref, mode = [], [], [], None
    for ln in open(path).read().splitlines():
        if ln.startswith("*"):
            mode = {"*nodes_all": "n", "*eles_change": "e", "*idref2new": "i"}.get(ln.split(",")[0])
            if ln.startswith("*End File"):
                break
            continue
        if mode == "n":
            nodes.append([float(t.strip(",")) for t in ln.split()])
        elif mode == "e":
            mods.append([int(t) for t in ln.split(",")])
        elif mode == "i":
            idref.append(int(ln))
    return np.array(nodes), np.array(mods, dtype=np.int32), np.array(idref, dtype=np.int32)


def load_sample():
    """mesh_o.inp through the oracle: K0, its direct solution and the CA change list."""
    from test_oracle_assembly import parse_inp_minimal

    xyz, tets, sets, (dof, val) = parse_inp_minimal(os.path.join(GOLDEN, "mesh_o.inp"))
    nd = len(xyz)
    conn = np.zeros((len(tets), 8), dtype=np.int32)
    conn[:, :4] = tets
    etype = np.zeros(len(tets), dtype=np.int32)
    resort, nfree = O.free_index(nd, np.array(sets["Set-fix"]))
    n = 3 * nfree
    F = O.single_load(n, resort, np.array(sets["Set-load"]), dof, val)
    cp, ri = O.build_csc(conn, etype, resort, n)
    K0 = O.fill_csc(conn, etype, xyz, resort, E0, NU0, cp, ri)
    u0 = spla.splu(sp.csc_matrix((K0, ri, cp), shape=(n, n))).solve(F)
    nodes, mods, idref = parse_ca_info_minimal(os.path.join(GOLDEN, "CA_info.txt"))
    cm = np.zeros((len(mods), 8), dtype=np.int32)
    cm[:, :4] = mods[:, 1:5]
    return dict(xyz=xyz, conn=conn, etype=etype, resort=resort, n=n, F=F, cp=cp, ri=ri, K0=K0, u0=u0,
                nodes=nodes, kind=mods[:, 0].copy(), cm=cm, idref=idref)


@pytest.fixture(scope="module")
def sample():
    return load_sample()


def to_new_model(resort, u, idref, nd_new):
    full = O.reset_displacement(resort, u).reshape(-1, 3)
    out = np.zeros((nd_new, 3))
    for i, j in enumerate(idref):
        if j >= 0:
            out[j] = full[i]
    return out


def test_ca_one_vector_matches_reference_output(sample):
    s = sample
    assert np.array_equal(s["nodes"], s["xyz"])  # the sample deletes only: nodes_all == mesh_o nodes
    assert set(s["kind"]) == {0}
    dK = O.ca_fill(s["kind"], s["cm"], np.zeros(len(s["kind"]), np.int32), s["nodes"], s["xyz"], s["resort"], E0, NU0,
                   s["cp"], s["ri"])
    u, m = O.ca_solve(s["cp"], s["ri"], s["K0"], dK, s["u0"], s["F"], 1)
    assert m == 1
    ref = np.load(os.path.join(GOLDEN, "cadcae_ca_u.npz"))["u_ca"]
    got = to_new_model(s["resort"], u, s["idref"], len(ref))
    assert rel_max(got, ref) < 2e-5
    # entrywise, away from free DOFs 1-3 (the Householder columns of the reference's degenerate SVD)
    inv = {j: i for i, j in enumerate(s["idref"]) if j >= 0}
    mask = np.ones_like(ref, dtype=bool)
    for j, i in inv.items():
        r = s["resort"][i]
        for c in range(3):
            if r >= 0 and 1 <= 3 * r + c <= 3:
                mask[j, c] = False
    assert mask.sum() == ref.size - 3
    tol = 1e-5 * np.abs(ref) + 2e-6 * np.abs(ref).max()
    assert np.all(np.abs(got - ref)[mask] <= tol[mask])
    # the reference's own span {u0, H e_1, H e_2, H e_3} (H: Householder reflection with H u0 = beta e_0)
    u0 = s["u0"]
    n = s["n"]
    beta = -np.sign(u0[0]) * np.linalg.norm(u0)
    v = u0 / (u0[0] - beta)
    v[0] = 1.0
    tau = (beta - u0[0]) / beta
    cols = [u0]
    for j in (1, 2, 3):
        e = np.zeros(n)
        e[j] = 1.0
        cols.append(e - tau * v * v[j])
    U, _ = np.linalg.qr(np.array(cols).T)
    Kc = sp.csc_matrix((s["K0"] + dK, s["ri"], s["cp"]), shape=(n, n))
    uh = U @ np.linalg.solve(U.T @ (Kc @ U), U.T @ s["F"])
    goth = to_new_model(s["resort"], uh, s["idref"], len(ref))
    assert np.all(np.abs(goth - ref) <= tol)
    # with the series the answer moves by ~1%: the reference output is the one-vector case
    u4, _ = O.ca_solve(s["cp"], s["ri"], s["K0"], dK, s["u0"], s["F"], 4)
    assert rel_max(to_new_model(s["resort"], u4, s["idref"], len(ref)), ref) > 5e-3


def exact_modified(cp, ri, K, F):
    """Direct solve of K u = F on the DOFs that keep stiffness (deleted
    elements leave orphan nodes with empty columns)."""
    n = len(cp) - 1
    A = sp.csc_matrix((K, ri, cp), shape=(n, n))
    act = np.abs(A.diagonal()) > 0
    u = np.zeros(n)
    u[act] = spla.spsolve(A[act][:, act].tocsc(), F[act])
    return u


def test_ca_series_converges_to_direct_solve(sample):
    s = sample
    dK = O.ca_fill(s["kind"], s["cm"], np.zeros(len(s["kind"]), np.int32), s["nodes"], s["xyz"], s["resort"], E0, NU0,
                   s["cp"], s["ri"])
    ue = exact_modified(s["cp"], s["ri"], s["K0"] + dK, s["F"])
    n = s["n"]
    act = np.abs(sp.csc_matrix((s["K0"] + dK, s["ri"], s["cp"]), shape=(n, n)).diagonal()) > 0
    errs = []
    for nb in (1, 2, 4, 8):
        u, _ = O.ca_solve(s["cp"], s["ri"], s["K0"], dK, s["u0"], s["F"], nb)
        # DOFs of nodes that keep an element (the removed nodes have no stiffness left)
        errs.append(np.linalg.norm((u - ue)[act]) / np.linalg.norm(ue[act]))
    assert errs[0] > 1e-3 and errs[-1] < errs[0] * 1e-2
    assert all(b <= a * 1.01 for a, b in zip(errs, errs[1:]))


def test_ca_single_deletion_exact():
    m, resort, F = mg.cantilever_c3d4(6, 3, 3, seed=5)
    n = 3 * int(resort.max() + 1)
    cp, ri = O.build_csc(m.conn, m.etype, resort, n)
    K0 = O.fill_csc(m.conn, m.etype, m.xyz, resort, E0, NU0, cp, ri)
    u0 = spla.spsolve(sp.csc_matrix((K0, ri, cp), shape=(n, n)), F)
    # an interior tet away from the clamp: all 12 DOFs free, rank Delta K = 6
    e = len(m.etype) // 2
    assert all(resort[m.conn[e, :4] - 1] >= 0)
    cm = np.zeros((1, 8), dtype=np.int32)
    cm[0, :4] = m.conn[e, :4] - 1
    dK = O.ca_fill([0], cm, [0], m.xyz, m.xyz, resort, E0, NU0, cp, ri)
    Kc = sp.csc_matrix((K0 + dK, ri, cp), shape=(n, n))
    ue = spla.spsolve(Kc, F)
    u, k = O.ca_solve(cp, ri, K0, dK, u0, F, 7)
    assert np.linalg.norm(u - ue) / np.linalg.norm(ue) < 1e-8
    u3, _ = O.ca_solve(cp, ri, K0, dK, u0, F, 3)
    assert np.linalg.norm(u3 - ue) / np.linalg.norm(ue) > 1e-10  # not exact before the space is complete


def test_ca_fill_closed_forms(sample):
    s = sample
    ne = len(s["etype"])
    allk = np.zeros(ne, np.int32)
    cm = np.zeros((ne, 8), np.int32)
    cm[:, :4] = s["conn"][:, :4] - 1
    dK = O.ca_fill(allk, cm, np.zeros(ne, np.int32), s["xyz"], s["xyz"], s["resort"], E0, NU0, s["cp"], s["ri"])
    assert rel_max(dK, -s["K0"]) < 1e-13
    one = cm[:7]
    z = O.ca_fill(np.ones(7, np.int32), one, np.zeros(7, np.int32), s["xyz"], s["xyz"], s["resort"], E0, NU0, s["cp"],
                  s["ri"])
    assert np.abs(z).max() == 0.0
    z2 = O.ca_fill([2, 0], np.repeat(one[:1], 2, 0), [0, 0], s["xyz"], s["xyz"], s["resort"], E0, NU0, s["cp"], s["ri"])
    assert np.abs(z2).max() == 0.0
    # moving a node changes exactly the blocks of the elements that touch it
    xm = s["xyz"].copy()
    xm[cm[0, 0]] += 0.05
    d = O.ca_fill([1], one[:1], [0], xm, s["xyz"], s["resort"], E0, NU0, s["cp"], s["ri"])
    Kn = O.fill_csc(s["conn"][:1], s["etype"][:1], xm, s["resort"], E0, NU0, s["cp"], s["ri"])
    Ko = O.fill_csc(s["conn"][:1], s["etype"][:1], s["xyz"], s["resort"], E0, NU0, s["cp"], s["ri"])
    assert rel_max(d, Kn - Ko) < 1e-12
    with pytest.raises(ValueError):  # couples two nodes that share no element
        far = np.zeros((1, 8), np.int32)
        far[0, :4] = [0, 1, 2, len(s["xyz"]) - 1]
        O.ca_fill([2], far, [0], s["xyz"], s["xyz"], s["resort"], E0, NU0, s["cp"], s["ri"])


def test_nodal_strain_matches_reference_output(sample):
    s = sample
    u_full = O.reset_displacement(s["resort"], s["u0"])
    ev = O.post_elem(s["conn"], s["etype"], s["xyz"], u_full, E0, NU0, 0)
    nv = O.post_node(s["conn"], s["etype"], len(s["xyz"]), ev)
    ref = np.load(os.path.join(GOLDEN, "cadcae_o_fields.npz"))["S"]
    for c in range(7):
        scale = np.abs(ref[c]).max()
        assert np.all(np.abs(nv[c] - ref[c]) <= 1e-5 * np.abs(ref[c]) + 1e-5 * scale), c


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_patch_linear_field_exact(kind):
    rng = np.random.default_rng(3)
    m, _, _ = (mg.cantilever_c3d4(3, 2, 2, seed=9) if kind == "tet" else mg.cantilever_c3d8(3, 2, 2))
    xyz = m.xyz.copy()
    if kind == "hex":  # distort interior-free coordinates a little (still valid)
        xyz += 0.08 * rng.uniform(-1, 1, xyz.shape)
    A = rng.normal(size=(3, 3)) * 1e-3
    u = xyz @ A.T
    eps = [A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[1, 2] + A[2, 1], A[0, 2] + A[2, 0]]
    ev = O.post_elem(m.conn, m.etype, xyz, u.ravel(), E0, NU0, 0)
    for c in range(6):
        assert np.allclose(ev[c], eps[c], rtol=1e-10, atol=1e-16)
    nv = O.post_node(m.conn, m.etype, len(xyz), ev)
    assert np.allclose(nv[:6], np.array(eps)[:, None], rtol=1e-10, atol=1e-16)
    # stress: Lame form, sigma_ii = lam tr + 2 mu eps_ii, tau = mu gamma
    lam = E0 * NU0 / ((1 + NU0) * (1 - 2 * NU0))
    mu = E0 / (2 * (1 + NU0))
    tr = eps[0] + eps[1] + eps[2]
    sig = [lam * tr + 2 * mu * eps[i] for i in range(3)] + [mu * eps[i] for i in range(3, 6)]
    sv = O.post_elem(m.conn, m.etype, xyz, u.ravel(), E0, NU0, 1)
    for c in range(6):
        assert np.allclose(sv[c], sig[c], rtol=1e-9, atol=1e-12)
    vm = np.sqrt(0.5 * ((sig[0] - sig[1]) ** 2 + (sig[1] - sig[2]) ** 2 + (sig[2] - sig[0]) ** 2) +
                 3 * (sig[3] ** 2 + sig[4] ** 2 + sig[5] ** 2))
    assert np.allclose(sv[6], vm, rtol=1e-9)


# ---------------------------------------------------------------- host drop-in (no device)
def test_ca_reader_and_mapping(sample):
    import openfea_amd as ofa

    info = ofa.read_ca_info(os.path.join(GOLDEN, "CA_info.txt"))
    s = sample
    assert np.array_equal(info.xyz, s["nodes"])
    assert np.array_equal(info.kind, s["kind"])
    assert np.array_equal(info.conn[:, :4], s["cm"][:, :4]) and not info.conn[:, 4:].any()
    assert np.array_equal(info.idref2new, s["idref"])
    assert (info.etype == 0).all()
    new = ofa.read_inp(os.path.join(GOLDEN, "mesh_m.inp"))
    assert new.nd == 525 and new.ne == 1920 and info.idref2new.max() == new.nd - 1
    u = np.random.default_rng(1).normal(size=s["n"])
    got = ofa.reset_ca_displacement(s["resort"], u, info.idref2new, new.nd).reshape(-1, 3)
    assert np.array_equal(got, to_new_model(s["resort"], u, s["idref"], new.nd))


def read_vtk(path):
    lines = open(path).read().splitlines()
    sec = {ln.split()[1]: k + 2 for k, ln in enumerate(lines) if ln.startswith("SCALARS ")}
    return lines, sec


def test_vtk_writer_reference_dialect(sample, tmp_path):
    import openfea_amd as ofa

    s = sample
    deck = ofa.read_inp(os.path.join(GOLDEN, "mesh_o.inp"))
    u_full = O.reset_displacement(s["resort"], s["u0"])
    ev = O.post_elem(s["conn"], s["etype"], s["xyz"], u_full, E0, NU0, 0)
    nv = O.post_node(s["conn"], s["etype"], len(s["xyz"]), ev)
    path = str(tmp_path / "out.vtk")
    ofa.export_dis_2_vtk(path, deck, u_full, 1.0, nv)
    lines, sec = read_vtk(path)
    nd, ne = len(s["xyz"]), len(s["etype"])
    assert lines[:6] == ["# vtk DataFile Version 3.0", "The density field of the optimized results", "ASCII",
                         "DATASET UNSTRUCTURED_GRID", "", f"POINTS\t{nd}\tdouble"]
    pts = np.array([[float(t) for t in lines[6 + i].split("\t\t")] for i in range(nd)])
    assert np.allclose(pts, s["xyz"] + u_full.reshape(-1, 3), rtol=1e-5, atol=1e-5)
    k = 6 + nd
    assert lines[k] == f"CELLS\t{ne}\t{5 * ne}"
    golden = np.load(os.path.join(GOLDEN, "cadcae_o_fields.npz"))
    cells = np.array([[int(t) for t in lines[k + 1 + e].split("\t")] for e in range(ne)])
    assert (cells[:, 0] == 4).all() and np.array_equal(cells[:, 1:], golden["cells"])
    assert lines[k + 1 + ne] == f"CELL_TYPES\t\t{ne}"
    assert np.array_equal([int(x) for x in lines[k + 2 + ne:k + 2 + 2 * ne]], golden["cell_types"])
    assert lines[k + 2 + 2 * ne:k + 5 + 2 * ne] == ["", f"POINT_DATA\t{nd}", "SCALARS u_x double 1"]
    names = ["u_x", "u_y", "u_z", "u_magnitude", "S_xx", "S_yy", "S_zz", "S_xy", "S_yz", "S_xz", "S_Mises"]
    assert list(sec) == names
    for t, name in enumerate(names):
        assert lines[sec[name] - 1] == f"LOOKUP_TABLE  table{t + 1}"
    ref_u = np.load(os.path.join(GOLDEN, "cadcae_o_u.npz"))["u_o"]
    for c, name in enumerate(names[:3]):
        vals = np.array([float(x) for x in lines[sec[name]:sec[name] + nd]])
        assert np.all(np.abs(vals - ref_u[:, c]) <= 1e-5 * np.abs(ref_u[:, c]) + 1e-9)
    for c, name in enumerate(names[4:]):
        vals = np.array([float(x) for x in lines[sec[name]:sec[name] + nd]])
        ref = golden["S"][c]
        assert np.all(np.abs(vals - ref) <= 1e-5 * np.abs(ref) + 1e-5 * np.abs(ref).max())
    # 6 significant digits, as the reference's default stream formatting
    assert lines[sec["u_z"]] == f"{u_full[2]:.6g}"
