"""Keyword-level deck extensions of the host reader (no device needed).

The reference reads one *Cload line, Set-fix and one material
(src/data2cae.cpp:315-446); the extensions read what the benchmark
configurations need natively: face tractions and pressures (*Dsload on an
element *Surface), several point loads, *Boundary types, initial velocities
and per-region materials.  Pinned by:
  * the side-band arrays the synthetic workloads use (consistent face loads
    of meshgen.face_traction, meshgen.initial_velocity): a deck written with
    the same data reads back to the same arrays;
  * closed forms: the pressure resultant on the x = L face is -p A e_x;
    loads on one node add; generate ranges;
  * the reference's own deck mesh_o.inp: the *Cload extension equals
    build_single_load on its Set-load nodes.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
import openfea_amd as ofa
from openfea_amd import meshgen as mg


@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_deck_round_trip(tmp_path, kind):
    m = mg.hex_block(4, 3, 2) if kind == "hex" else mg.kuhn_tets(4, 3, 2, jitter=0.15, seed=3)
    t = (0.3, -1.0, -5.0)
    v = mg.initial_velocity(m)
    regions = [(np.arange(m.ne) % 3 == 0, 2.1e5, 0.3, 7.85e-9), (np.arange(m.ne) % 3 != 0, 7.0e4, 0.3, 2.7e-9)]
    path = str(tmp_path / "deck.inp")
    mg.write_inp(path, m, traction=t, cloads=[(5, 1, 2.5), (5, 1, 1.0), (7, 3, -4.0)], velocity=v, materials=regions)
    d = ofa.read_inp(path)
    assert d.ext_error == ""
    assert np.array_equal(d.xyz, m.xyz) and np.array_equal(d.conn, m.conn) and np.array_equal(d.etype, m.etype)
    assert np.array_equal(d.bc_fixed, m.fix_set) and d.bc_partial == 0 and d.bc_nonzero == 0
    f = mg.face_traction(m, t)
    f[4, 0] += 3.5  # two *Cload lines on node 5 add
    f[6, 2] -= 4.0
    assert np.allclose(d.f_nodal, f, rtol=0, atol=1e-13 * np.abs(f).max())
    assert np.array_equal(d.v0, v)
    E = np.where(regions[0][0], 2.1e5, 7.0e4)
    assert np.array_equal(d.E_elem, E) and np.all(d.nu_elem == 0.3)
    assert np.array_equal(d.rho_elem, np.where(regions[0][0], 7.85e-9, 2.7e-9))
    # the reference-dialect fields of the same deck: Set-fix, first *Cload line (Set-load, 3, 0.), last *Elastic
    assert np.array_equal(np.sort(d.fix_set), m.fix_set)
    assert d.load_dof == 3 and d.load_value == 0.0 and d.E == 7.0e4


def test_pressure_resultant(tmp_path):
    for m in (mg.hex_block(3, 2, 2, h=0.5), mg.kuhn_tets(3, 2, 2, h=0.5)):
        path = str(tmp_path / "p.inp")
        mg.write_inp(path, m, pressure=2.0)
        d = ofa.read_inp(path)
        nx, ny, nz, h = m.dims
        area = ny * h * nz * h
        # pressure pushes into the face: resultant -p A along the outward normal +x
        assert np.allclose(d.f_nodal.sum(0), [-2.0 * area, 0.0, 0.0], atol=1e-12)
        tip = np.abs(m.xyz[:, 0] - nx * h) < 1e-12
        assert np.all(d.f_nodal[~tip] == 0.0)


def test_config1_load_expressed_natively(tmp_path):
    m, resort, F = mg.cantilever_c3d8()
    path = str(tmp_path / "cfg1.inp")
    mg.write_inp(path, m, traction=(0.0, 0.0, -10.0))
    d = ofa.read_inp(path)
    r, nfree = ofa.build_free_index(d.nd, d.bc_fixed)
    assert np.array_equal(r, resort)
    Fd = mg.free_rhs(d.f_nodal, r)
    assert np.allclose(Fd, F, rtol=0, atol=1e-13) and np.isclose(Fd.sum(), -1000.0)


@pytest.mark.parametrize("bc,nfixed,partial_per_node", [("ENCASTRE", 1, 0), ("PINNED", 1, 0), ("XSYMM", 0, 1),
                                                        ("1, 2", 0, 2), ("2, 3, 0.", 0, 2), ("1, 3", 1, 0)])
def test_boundary_kinds(tmp_path, bc, nfixed, partial_per_node):
    m = mg.hex_block(2, 2, 2)
    path = str(tmp_path / "b.inp")
    mg.write_inp(path, m, boundary=bc)
    d = ofa.read_inp(path)
    nf = len(m.fix_set)
    assert len(d.bc_fixed) == nfixed * nf and d.bc_partial == partial_per_node * nf and d.bc_nonzero == 0
    mg.write_inp(path, m, boundary="3, 3, 0.5")
    d = ofa.read_inp(path)
    assert d.bc_nonzero == nf and d.bc_partial == nf


def test_generate_and_set_references(tmp_path):
    m = mg.hex_block(2, 1, 1)
    path = str(tmp_path / "g.inp")
    mg.write_inp(path, m)
    text = open(path).read().replace("*End Assembly", "*Nset, nset=Gen, generate\n2, 10, 4\n*Nset, nset=Both\nGen, 12\n"
                                     "*End Assembly")
    text = text.replace("*Cload\nSet-load, 3, 0.", "*Cload\nBoth, 2, 1.5\nGen, 2, 1.0")
    open(path, "w").write(text)
    d = ofa.read_inp(path)
    f = np.zeros((m.nd, 3))
    f[[1, 5, 9], 1] += 2.5
    f[11, 1] += 1.5
    assert np.array_equal(d.f_nodal, f)


def test_bad_extension_keeps_reference_fields(tmp_path):
    m = mg.hex_block(2, 1, 1)
    path = str(tmp_path / "bad.inp")
    mg.write_inp(path, m, traction=(0, 0, -1))
    open(path, "a").write("*Dsload\nNo-such-surface, P, 1.\n")
    d = ofa.read_inp(path)
    assert "No-such-surface" in d.ext_error and d.f_nodal is None
    assert np.array_equal(d.xyz, m.xyz) and len(d.fix_set) == len(m.fix_set)


def test_reference_deck_cload_matches_single_load():
    d = ofa.read_inp(os.path.join(GOLDEN, "mesh_o.inp"))
    assert d.ext_error == "" and len(d.bc_fixed) == 0  # the reference clamps Set-fix by name, no *Boundary
    r, nfree = ofa.build_free_index(d.nd, d.fix_set)
    F = ofa.build_single_load(3 * nfree, r, d.load_set, d.load_dof, d.load_value)
    assert np.array_equal(mg.free_rhs(d.f_nodal, r), F)
    assert np.all(np.isnan(d.E_elem))  # no *Solid Section in the deck


def _part_deck(m):
    """The mesh m as one Abaqus part with one instance; sets are part-local
    and the step names them instance-qualified (the layout Abaqus/CAE writes,
    as in the reference's sample_model/NLFEA.inp)."""
    lines = ["*Heading", "*Part, name=Beam", "*Node"]
    lines += [f"{i + 1}, {x:.17g}, {y:.17g}, {z:.17g}" for i, (x, y, z) in enumerate(m.xyz)]
    lines.append("*Element, type=C3D8")
    lines += [f"{e + 1}, " + ", ".join(str(int(v)) for v in m.conn[e]) for e in range(m.ne)]
    # the reference dialect takes the counts from the first *Nset / *Elset (generate) lines
    lines += ["*Nset, nset=All-nodes, internal, generate", f"1, {m.nd}, 1", "*Elset, elset=All, generate",
              f"1, {m.ne}, 1"]
    lines += ["*Nset, nset=Set-fix"] + [", ".join(str(i + 1) for i in m.fix_set)]
    tip = np.flatnonzero(np.abs(m.xyz[:, 0] - m.xyz[:, 0].max()) < 1e-12)
    lines += ["*Nset, nset=Set-load"] + [", ".join(str(i + 1) for i in tip)]
    lines += ["*Solid Section, elset=All, material=Steel", ",",
              "*End Part", "*Assembly, name=Assembly", "*Instance, name=Beam-1, part=Beam", "*End Instance",
              "*Nset, nset=Corner, instance=Beam-1", "1", "*End Assembly",
              "*Material, name=Steel", "*Elastic", "210000., 0.3", "*Step, name=Step-1", "*Static",
              "*Boundary", "Beam-1.Set-fix, ENCASTRE", "*Cload", "Beam-1.Set-load, 2, -4.", "Beam-1.3, 3, 1.5",
              "Corner, 1, 2.", "*End Step"]
    return "\n".join(lines) + "\n", tip


def test_instance_qualified_sets(tmp_path):
    m = mg.hex_block(3, 2, 2)
    text, tip = _part_deck(m)
    path = str(tmp_path / "part.inp")
    open(path, "w").write(text)
    d = ofa.read_inp(path)
    assert d.ext_error == ""
    assert np.array_equal(np.sort(d.bc_fixed), np.sort(m.fix_set + 1))
    f = np.zeros((m.nd, 3))
    f[tip, 1] = -4.0
    f[2, 2] += 1.5  # instance-qualified node label
    f[0, 0] += 2.0  # assembly-level set of the instance
    assert np.array_equal(d.f_nodal, f)
    assert np.all(d.E_elem == 210000.0) and np.all(d.nu_elem == 0.3)
    # an instance-qualified name of a set the part does not hold stays an error
    open(path, "w").write(text.replace("Beam-1.Set-fix, ENCASTRE", "Beam-1.Set-none, ENCASTRE"))
    assert "Set-none" in ofa.read_inp(path).ext_error


REF_SAMPLES = "/root/reference/sample_model"


@pytest.mark.skipif(not os.path.isdir(REF_SAMPLES), reason="reference checkout not present")
@pytest.mark.parametrize("deck,load", [("NLFEA.inp", -1000000.0), ("L_NL/nlcae.inp", -10.0)])
def test_reference_part_decks(deck, load):
    """The reference's own Abaqus/CAE decks (one part, one instance) load their
    *Boundary (Part-1-1.Set-fix, ENCASTRE) and *Cload (Part-1-1.Set-load, 2, F)
    lines through the instance-qualified names."""
    d = ofa.read_inp(os.path.join(REF_SAMPLES, deck))
    assert d.ext_error == ""
    assert np.array_equal(np.sort(d.bc_fixed), np.sort(d.fix_set))  # the part's Set-fix, 1-based ids
    f = np.zeros((d.nd, 3))
    f[d.load_set - 1, 1] = load
    assert np.array_equal(d.f_nodal, f)
    assert d.load_dof == 2 and d.load_value == load
    assert np.all(d.E_elem == d.E) and np.all(d.nu_elem == d.nu)
