"""CPU tests of the C++ host layer (include/openfea_amd_cae.hpp) through its
driver tests/cpp/implicit_pipeline: the reference's pre_info + set_BCs calls
on its own sample decks against the oracle, and the loud failure of the
device path where there is no GPU."""
import os

import numpy as np
import pytest

import cpp_layer as CL
from conftest import GOLDEN
from oracle import oracle as O
import openfea_amd as ofa


@pytest.mark.parametrize("deck", ["mesh_o.inp", "mesh_o_cae.inp"])
def test_reader_bcs_and_pattern_match_oracle(deck, tmp_path):
    path = os.path.join(GOLDEN, deck)
    pre = tmp_path / "out"
    r = CL.run(path, pre, "--host-only")
    assert r.returncode == 0, r.stderr
    s = CL.summary(r.stdout)
    d = ofa.read_inp(path)
    assert int(s["nd"]) == d.nd and int(s["ne"]) == len(d.etype)
    # no *Elastic in these decks: the material of src/main.cpp:45 is kept
    assert float(s["E"]) == 2.1e5 and float(s["nu"]) == 0.3
    assert int(s["load_dof"]) == d.load_dof and float(s["load_value"]) == d.load_value
    resort = CL.load(pre, "resort", np.int32)
    ro, nfree = O.free_index(d.nd, d.fix_set)
    assert np.array_equal(resort, ro)
    assert np.array_equal(CL.load(pre, "fixsorted", np.int32), np.sort(d.fix_set))
    n = 3 * nfree
    F = CL.load(pre, "load", np.float64)
    assert np.array_equal(F, O.single_load(n, ro, d.load_set, d.load_dof, d.load_value))
    cp, ri = O.build_csc(d.conn, d.etype, ro, n)
    assert int(s["n"]) == n and int(s["nnz"]) == len(ri)
    assert np.array_equal(CL.load(pre, "col", np.int32), cp)
    assert np.array_equal(CL.load(pre, "row", np.int32), ri)


def test_missing_deck_exits_with_message(tmp_path):
    r = CL.run(str(tmp_path / "no_such.inp"), tmp_path / "out", "--host-only")
    assert r.returncode == 1 and "openfea_amd: ReadInfo" in r.stderr


def test_device_path_fails_loudly_without_gpu(tmp_path):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = CL.run(os.path.join(GOLDEN, "mesh_o.inp"), tmp_path / "out")
    assert r.returncode == 1 and "build_CSR" in r.stderr


@pytest.mark.parametrize("mode", [("nl",), ("explicit", "1e-3")])
def test_cae_process_device_paths_fail_loudly_without_gpu(mode, tmp_path):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = CL.run_cae(os.path.join(GOLDEN, "mesh_o.inp"), tmp_path / "out", *mode)
    where = "build_CSR" if mode[0] == "nl" else "explicit_analysis"
    assert r.returncode == 1 and f"openfea_amd: {where}" in r.stderr, r.stdout + r.stderr
