"""The reference's own element code as a CPU assembly (oracle/_ref, built
from the reference checkout where it exists): it equals the oracle's fill
within 1e-13 and gives the same bits on 1 and 4 threads.  It is the CPU
baseline of the benchmark; skipped where oracle/_ref was not built."""
import numpy as np
import pytest

from conftest import rel_max
from oracle import oracle as O
from openfea_amd import meshgen as mg

pytestmark = pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built (no reference checkout)")


@pytest.mark.parametrize("kind", ["tet", "hex", "mixed"])
def test_reference_fill_equals_oracle(kind):
    if kind == "tet":
        m, resort, F = mg.cantilever_c3d4(6, 3, 3)
    elif kind == "hex":
        m, resort, F = mg.cantilever_c3d8(6, 3, 3)
    else:  # hexes and tets on one node grid (identical jitter draws for one seed)
        h = mg.hex_block(4, 3, 3, jitter=0.15, seed=1)
        t = mg.kuhn_tets(4, 3, 3, jitter=0.15, seed=1)
        m = h
        m.conn = np.vstack([h.conn, t.conn])
        m.etype = np.concatenate([h.etype, t.etype])
        resort = mg.free_index(m.nd, h.fix_set)
    n = 3 * int(resort.max() + 1)
    cp, ri = O.build_csc(m.conn, m.etype, resort, n)
    ref = O.fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri)
    v1, t1 = O.ref_fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri, threads=1)
    v4, t4 = O.ref_fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri, threads=4)
    assert t1 == 1 and t4 == 4
    assert rel_max(v1, ref) < 1e-13
    assert np.array_equal(v1, v4)
    Ee = np.random.default_rng(2).uniform(1e5, 3e5, len(m.etype))
    ve, _ = O.ref_fill_csc(m.conn, m.etype, m.xyz, resort, 0.0, 0.3, cp, ri, threads=2, E_elem=Ee)
    assert rel_max(ve, O.fill_csc(m.conn, m.etype, m.xyz, resort, 0.0, 0.3, cp, ri, E_elem=Ee)) < 1e-13
