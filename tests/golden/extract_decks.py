"""Writes tests/golden/ref_decks.npz: the meshes and boundary data of three of
the reference's own sample decks, read by this package's host reader (the
reference's .inp dialect), for the GPU tests that run them through
set_mesh / assemble / solve on a machine without the reference checkout.

  sample_model/test/implicite_test_C3D4_C3D8.inp  mixed C3D4 + C3D8
  sample_model/NLFEA.inp                          69,376 C3D4 (one part)
  sample_model/CA/CA_zongliang.inp                50,797 C3D4

Stored per deck (prefix = deck key): xyz (float64 nd x 3), conn (int32, ne x
8 for mixed decks, ne x 4 for all-C3D4 decks), etype (int8), fix_set and
load_set (int32, 1-based), load_dof, load_value, E, nu.  Data only: no
reference source is stored.

Run from the repository root where /root/reference exists:
    python tests/golden/extract_decks.py
"""
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)
REF = "/root/reference/sample_model"
DECKS = {"mixed": "test/implicite_test_C3D4_C3D8.inp", "nlfea": "NLFEA.inp", "ca": "CA/CA_zongliang.inp"}


def main():
    import openfea_amd as ofa

    out = {}
    for key, rel in DECKS.items():
        d = ofa.read_inp(os.path.join(REF, rel))
        conn = d.conn if np.any(d.etype == 1) else d.conn[:, :4]
        out.update({f"{key}_xyz": d.xyz, f"{key}_conn": conn.astype(np.int32), f"{key}_etype": d.etype.astype(np.int8),
                    f"{key}_fix_set": d.fix_set.astype(np.int32), f"{key}_load_set": d.load_set.astype(np.int32),
                    f"{key}_load": np.array([d.load_dof, d.load_value, d.E, d.nu])})
        print(key, d.nd, "nodes", len(d.etype), "elements", "types", np.bincount(d.etype, minlength=2), flush=True)
    path = os.path.join(REPO, "tests", "golden", "ref_decks.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
