"""Host logic of libofea_hip.so and its C ABI (CPU only: no kernel is launched).

  * the library loads and exports every function declared in include/*.h;
  * the native .inp reader against an independent minimal parse of the
    reference's sample decks;
  * free-DOF numbering, single-DOF load vector and the CSC pattern against the
    oracle, bit-exact;
  * the mesh generators' geometry (positive orientation, volumes, tractions).
"""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, REPO
from oracle import oracle as O
import openfea_amd as ofa
from openfea_amd import _capi, meshgen as mg
from test_oracle_assembly import parse_inp_minimal


def header_functions():
    names = set()
    for fn in os.listdir(os.path.join(REPO, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(REPO, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(ofea_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = _capi.lib()
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert set(_capi.SIGNATURES) == declared
    assert L.ofea_version().decode().startswith("openfea_amd")


def test_null_arguments_are_errors_not_crashes():
    L = _capi.lib()
    assert L.ofea_inp_open(b"/nonexistent.inp", None) == _capi.OFEA_ERR_ARG
    import ctypes as C
    h = C.c_void_p()
    assert L.ofea_inp_open(b"/nonexistent/deck.inp", C.byref(h)) == _capi.OFEA_ERR_IO
    assert "cannot open" in L.ofea_last_error().decode()
    assert L.ofea_build_single_load(3, None, None, 0, 7, 1.0, None) == _capi.OFEA_ERR_ARG


@pytest.mark.parametrize("deck", ["mesh_o.inp", "mesh_o_cae.inp"])
def test_reader_matches_minimal_parse(deck):
    path = os.path.join(GOLDEN, deck)
    d = ofa.read_inp(path)
    xyz, tets, sets, (dof, val) = parse_inp_minimal(path)
    assert d.nd == 625 and d.ne == 2304
    assert np.array_equal(d.xyz, xyz)
    assert np.array_equal(d.conn[:, :4], tets) and not d.conn[:, 4:].any()
    assert np.all(d.etype == 0)
    assert np.array_equal(d.fix_set, sets["Set-fix"]) and np.array_equal(d.load_set, sets["Set-load"])
    assert (d.load_dof, d.load_value) == (dof, val) == (3, -10.0)
    assert np.isnan(d.E)  # the deck has no *Elastic line


def test_reader_material_time_and_mixed_elements(tmp_path):
    deck = tmp_path / "mixed.inp"
    deck.write_text("""*Heading
*Node
1, 0., 0., 0.
2, 1., 0., 0.
3, 1., 1., 0.
4, 0., 1., 0.
5, 0., 0., 1.
6, 1., 0., 1.
7, 1., 1., 1.
8, 0., 1., 1.
9, 0.5, 0.5, 2.
*Element, type=C3D8\r
1, 1, 2, 3, 4, 5, 6, 7, 8
*Element, type=C3D4
2, 5, 6, 7, 9
*Nset
1, 9, 1
*Elset
1, 2, 1
*Nset, nset=Set-fix
1, 2, 3, 4
*Nset, nset=Set-load
9,
*End Assembly
*Material
*Density
7.85e-09,
*Elastic
210000., 0.3
*Step
*Dynamic, Explicit
, 0.001
*Cload
Set-load, 2, 5.5
""")
    d = ofa.read_inp(str(deck))
    assert (d.nd, d.ne) == (9, 2)
    assert list(d.etype) == [1, 0]
    assert list(d.conn[1, :4]) == [5, 6, 7, 9]
    assert list(d.fix_set) == [1, 2, 3, 4] and list(d.load_set) == [9]
    assert (d.load_dof, d.load_value) == (2, 5.5)
    assert (d.E, d.nu, d.rho) == (210000.0, 0.3, 7.85e-09)
    assert (d.dt, d.t_total) == (0.0, 0.001)


def test_free_index_and_load_match_oracle():
    rng = np.random.default_rng(0)
    nd = 200
    fix = rng.choice(np.arange(1, nd + 1), 37, replace=False)
    r1, n1 = ofa.build_free_index(nd, fix)
    r2, n2 = O.free_index(nd, fix)
    assert n1 == n2 == nd - 37 and np.array_equal(r1, r2)
    load = rng.choice(np.where(r1 >= 0)[0] + 1, 12, replace=False)
    F1 = ofa.build_single_load(3 * n1, r1, load, 2, -3.5)
    F2 = O.single_load(3 * n1, r1, load, 2, -3.5)
    assert np.array_equal(F1, F2) and np.count_nonzero(F1) == 12
    u = rng.normal(size=3 * n1)
    assert np.array_equal(ofa.reset_displacement(r1, u), O.reset_displacement(r1, u))


def test_explicit_load_truncates_the_value_to_an_int():
    """src/cae.cpp:449-456: int force_value = load_value_, assigned (not
    added) at DOF load_dof of every load node, clamped nodes included."""
    from openfea_amd import fea

    f = fea.build_explicit_load(6, [2, 5, 5], 3, -10.9)
    want = np.zeros(18)
    want[[3 * 1 + 2, 3 * 4 + 2]] = -10.0
    assert np.array_equal(f, want)
    assert np.array_equal(fea.build_explicit_load(3, [1], 1, 2.999), np.array([2.0] + [0.0] * 8))
    assert np.array_equal(fea.build_explicit_load(3, [3], 2, 0.7), np.zeros(9))
    L = _capi.lib()
    out = np.zeros(9)
    assert L.ofea_build_explicit_load(3, _capi.ptr(_capi.i32([4])), 1, 1, 1.0, _capi.ptr(out)) == _capi.OFEA_ERR_ARG
    assert L.ofea_build_explicit_load(3, _capi.ptr(_capi.i32([1])), 1, 4, 1.0, _capi.ptr(out)) == _capi.OFEA_ERR_ARG
    assert L.ofea_build_explicit_load(3, _capi.ptr(_capi.i32([1])), 1, 1, 3e9, _capi.ptr(out)) == _capi.OFEA_ERR_ARG


@pytest.mark.parametrize("mk", [lambda: mg.hex_block(5, 3, 2, jitter=0.2, seed=1),
                                lambda: mg.kuhn_tets(4, 3, 3, jitter=0.2, seed=2),
                                lambda: mg.hex_block(20, 10, 10)])
def test_build_csr_bit_exact_vs_oracle(mk):
    m = mk()
    resort = mg.free_index(m.nd, m.fix_set)
    cp, ri = ofa.build_csr(m.conn, m.etype, resort)
    cpo, rio = O.build_csc(m.conn, m.etype, resort, len(cp) - 1)
    assert np.array_equal(cp, cpo) and np.array_equal(ri, rio)


def test_build_csr_mixed_deck_with_reader_numbering():
    d = ofa.read_inp(os.path.join(GOLDEN, "mesh_o.inp"))
    resort, nfree = ofa.build_free_index(d.nd, d.fix_set)
    cp, ri = ofa.build_csr(d.conn, d.etype, resort)
    cpo, rio = O.build_csc(d.conn, d.etype, resort, 3 * nfree)
    assert np.array_equal(cp, cpo) and np.array_equal(ri, rio)


def test_meshgen_geometry():
    m = mg.kuhn_tets(6, 3, 3, jitter=0.2, seed=20261203)
    X = m.xyz[m.conn[:, :4] - 1]
    vol = np.einsum("ij,ij->i", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0])) / 6
    assert vol.min() > 0 and np.isclose(vol.sum(), 6 * 3 * 3)
    f = mg.face_traction(m, (0, 0, -5.0))
    assert np.isclose(f[:, 2].sum(), -5.0 * 3 * 3) and np.abs(f[:, :2]).max() == 0
    h = mg.hex_block(4, 2, 2)
    f = mg.face_traction(h, (0, 0, -10.0))
    on = np.isclose(h.xyz[:, 0], 4.0)
    vals = sorted(set(np.round(f[on, 2], 12)))
    assert vals == [-10.0, -5.0, -2.5]   # interior, edge, corner: -10 h^2 {1, 1/2, 1/4}
    Ee = mg.design_field(mg.hex_block(40, 20, 20), 7)
    assert Ee.min() >= 0.5 * 210000 - 1e-6 and Ee.max() <= 1.5 * 210000 + 1e-6
