"""GPU run of the C++ host layer (include/openfea_amd_cae.hpp): the
reference's whole linear static pipeline on its sample deck mesh_o.inp,
written in the reference's call order (tests/cpp/implicit_pipeline.cpp).

Bars: CSC pattern bit-exact and values within 1e-12 of max|K| (oracle);
displacements within 6 significant digits of the reference's PARDISO result
(its cadcae_o.vtk); nodal S_* fields within 1e-5 of the same file; an
unknown solver string returns false; the VTK file carries the same numbers.
"""
import os
import sys

import numpy as np
import pytest

import cpp_layer as CL
from conftest import GOLDEN, rel_max
from oracle import oracle as O
import openfea_amd as ofa

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["GPU_PCG", "GPU_PCG_TWO_LEVEL_MIXED"])
def test_cpp_pipeline_reference_deck(solver, tmp_path):
    path = os.path.join(GOLDEN, "mesh_o.inp")
    print(f"[gpu] C++ host layer pipeline {solver}", flush=True)
    sys.stdout.flush()
    pre = tmp_path / "out"
    r = CL.run(path, pre, solver)
    assert r.returncode == 0, r.stdout + r.stderr
    s = CL.summary(r.stdout)
    assert s["unknown_rejected"] == "1" and float(s["relres"]) <= 1e-10
    d = ofa.read_inp(path)
    ro, nfree = O.free_index(d.nd, d.fix_set)
    cp, ri = O.build_csc(d.conn, d.etype, ro, 3 * nfree)
    assert np.array_equal(CL.load(pre, "col", np.int32), cp)
    assert np.array_equal(CL.load(pre, "row", np.int32), ri)
    ref = O.fill_csc(d.conn, d.etype, d.xyz, ro, 2.1e5, 0.3, cp, ri)
    assert rel_max(CL.load(pre, "val", np.float64), ref) < 1e-12
    u = CL.load(pre, "u", np.float64).reshape(-1, 3)
    u_ref = np.load(os.path.join(GOLDEN, "cadcae_o_u.npz"))["u_o"]
    assert np.all(np.abs(u - u_ref) <= 6e-6 * np.abs(u_ref) + 1e-9)
    S = CL.load(pre, "S", np.float64).reshape(-1, 7).T
    S_ref = np.load(os.path.join(GOLDEN, "cadcae_o_fields.npz"))["S"]
    for c in range(7):
        assert np.all(np.abs(S[c] - S_ref[c]) <= 1e-5 * np.abs(S_ref[c]) + 1e-5 * np.abs(S_ref[c]).max()), c
    lines = open(f"{pre}.vtk").read().splitlines()
    k = lines.index("SCALARS S_Mises double 1") + 2
    mises = np.array([float(x) for x in lines[k:k + d.nd]])
    assert np.allclose(mises, S[6], rtol=1e-5, atol=1e-12)
