"""GPU: the drop-in implicit pipeline driven by decks.

  * sample_model/CA/mesh_o.inp end to end (reader -> BCs -> pattern ->
    assembly -> PCG -> reset_displacement -> nodal strain -> VTK) against
    the reference's own output file cadcae_o.vtk: cells and types exact,
    u and S_* fields to its 6 printed digits;
  * a deck written with the keyword extensions (two materials, TRVEC
    traction on an element surface, *Boundary ENCASTRE) against the oracle:
    per-element-E assembly and a direct solve, 1e-8.
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
from openfea_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_mesh_o_pipeline_writes_reference_vtk(tmp_path):
    from test_oracle_ca_post import read_vtk

    say("implict_analysis mesh_o.inp -> VTK")
    d = ofa.read_inp(os.path.join(GOLDEN, "mesh_o.inp"))
    path = str(tmp_path / "cadcae_o.vtk")
    u, fields = ofa.implict_analysis(d, path, E=2.1e5, nu=0.3, rtol=1e-10)
    lines, sec = read_vtk(path)
    nd, ne = d.nd, d.ne
    g = np.load(os.path.join(GOLDEN, "cadcae_o_fields.npz"))
    k = 6 + nd
    cells = np.array([[int(t) for t in lines[k + 1 + e].split("\t")] for e in range(ne)])
    assert np.array_equal(cells[:, 1:], g["cells"])
    ref_u = np.load(os.path.join(GOLDEN, "cadcae_o_u.npz"))["u_o"]
    for c, name in enumerate(("u_x", "u_y", "u_z")):
        vals = np.array([float(x) for x in lines[sec[name]:sec[name] + nd]])
        assert np.all(np.abs(vals - ref_u[:, c]) <= 6e-6 * np.abs(ref_u[:, c]) + 1e-9)
    for c, name in enumerate(("S_xx", "S_yy", "S_zz", "S_xy", "S_yz", "S_xz", "S_Mises")):
        vals = np.array([float(x) for x in lines[sec[name]:sec[name] + nd]])
        ref = g["S"][c]
        assert np.all(np.abs(vals - ref) <= 1e-5 * np.abs(ref) + 1e-5 * np.abs(ref).max())


@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_extension_deck_matches_oracle(tmp_path, kind):
    m = mg.hex_block(6, 3, 3, jitter=0.1, seed=4) if kind == "hex" else mg.kuhn_tets(6, 3, 3, jitter=0.15, seed=4)
    say(f"extension deck {kind}")
    c = mg.element_centroids(m)
    soft = c[:, 0] > 3.0
    path = str(tmp_path / "x.inp")
    mg.write_inp(path, m, traction=(0.5, 0.0, -8.0), cloads=[(m.nd, 2, 40.0)],
                 materials=[(~soft, 2.1e5, 0.3, 7.85e-9), (soft, 7.0e4, 0.3, 2.7e-9)])
    d = ofa.read_inp(path)
    u, fields = ofa.implict_analysis(d, str(tmp_path / "x.vtk"), rtol=1e-11)
    resort = mg.free_index(m.nd, m.fix_set)
    n = 3 * int(resort.max() + 1)
    cp, ri = O.build_csc(m.conn, m.etype, resort, n)
    K = O.fill_csc(m.conn, m.etype, m.xyz, resort, 0.0, 0.3, cp, ri, E_elem=np.where(soft, 7.0e4, 2.1e5))
    f = mg.face_traction(m, (0.5, 0.0, -8.0))
    f[m.nd - 1, 1] += 40.0
    x = spla.spsolve(sp.csc_matrix((K, ri, cp), shape=(n, n)), mg.free_rhs(f, resort))
    uo = O.reset_displacement(resort, x).reshape(-1, 3)
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8
    nv = O.post_node(m.conn, m.etype, m.nd, O.post_elem(m.conn, m.etype, m.xyz, uo.ravel(), 2.1e5, 0.3, 0))
    assert rel_max(fields, nv) < 1e-7
