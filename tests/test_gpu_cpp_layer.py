"""GPU run of the C++ host layer (include/openfea_amd_cae.hpp): the
reference's whole linear static pipeline on its sample deck mesh_o.inp,
written in the reference's call order (tests/cpp/implicit_pipeline.cpp),
and its CAE_process driver (tests/cpp/cae_process.cpp) for the NLFEA Newton
loop (loop count equal to the oracle's, u within 1e-8) and the explicit loop
(step count equal, last output within 1e-9 of the oracle, one VTK file per
output).

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


def test_cpp_cae_process_nlfea_newton(tmp_path):
    path = os.path.join(GOLDEN, "mesh_o.inp")
    print("[gpu] C++ CAE_process NLFEA", flush=True)
    pre = tmp_path / "nl"
    r = CL.run_cae(path, pre, "nl", "GPU_PCG_TWO_LEVEL")
    assert r.returncode == 0, r.stdout + r.stderr
    loops = [ln for ln in r.stdout.splitlines() if ln.startswith("this is loop ")]
    d = ofa.read_inp(path)
    ro, nfree = O.free_index(d.nd, d.fix_set)
    cp, ri = O.build_csc(d.conn, d.etype, ro, 3 * nfree)
    F = O.single_load(3 * nfree, ro, d.load_set, d.load_dof, d.load_value)
    u_ref, loops_ref, _, hist = O.nl_newton_ref(d.conn, d.etype, d.xyz, ro, 2.1e5, 0.3, F, cp, ri)
    assert len(loops) == loops_ref, r.stdout
    for ln, h in zip(loops, hist):
        assert abs(float(ln.rsplit(" ", 1)[1]) / h - 1.0) < 1e-4
    u = CL.load(pre, "u", np.float64)
    assert rel_max(u, O.reset_displacement(ro, u_ref).ravel()) < 1e-8
    assert os.path.exists(f"{pre}.vtk")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cpp_cae_process_explicit(precision, tmp_path):
    path = os.path.join(GOLDEN, "mesh_o.inp")
    print(f"[gpu] C++ CAE_process explicit {precision}", flush=True)
    d = ofa.read_inp(path)
    dt0 = 0.9 * O.crit_timestep(d.conn, d.etype, d.xyz, 2.1e5, 0.3)
    total = 30.4 * dt0
    pre = tmp_path / "ex"
    r = CL.run_cae(path, pre, "explicit", repr(total), "0", precision)
    assert r.returncode == 0, r.stdout + r.stderr
    s = CL.summary(r.stdout)
    ro, _ = O.free_index(d.nd, d.fix_set)
    fixed = (ro < 0).astype(np.int32)
    mass = O.lumped_mass(d.conn, d.etype, d.xyz, 7800.0)
    fext = np.zeros(3 * d.nd)
    fext[3 * (np.asarray(d.load_set) - 1) + d.load_dof - 1] = int(d.load_value)
    ref = O.explicit_timed(d.conn, d.etype, d.xyz, 2.1e5, 0.3, mass, fext, fixed, total, 0.0)
    for k in range(11):
        assert os.path.exists(f"{pre}_{k}.vtk")
    u = CL.load(pre, "u", np.float64)
    err = np.linalg.norm(u - ref["frames"][-1]) / np.linalg.norm(ref["frames"][-1])
    if precision == "fp64":
        assert int(s["steps"]) == ref["steps"] and err < 1e-9
    else:
        assert err < 1e-4
