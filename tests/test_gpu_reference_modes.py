"""The reference's own loop semantics on the GPU, against the oracle:
* ofea_explicit_analysis: the explicit loop of src/cae.cpp:420-574 with its
  time control (automatic step recomputed from Ke at the current coordinates
  every step, CheckTime clamps onto the outputs and the total, the
  int-truncated single load); FP64 frames within 1e-9 of the oracle, the same
  output indices and times, the same step count; FP32 within 1e-4;
* ofea_nl_solve_ref: the Newton loop of src/cae.cpp:209-253 with its stopping
  rule ||r||_2 / (3 nfree) <= res_lmit; the same loop count as the oracle's
  direct-solve loop and u within 1e-8."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, rel_max
from oracle import oracle as O
import openfea_amd as ofa
from openfea_amd import _capi as A, fea, meshgen as mg

pytestmark = pytest.mark.gpu

E, NU, RHO_REF = 210000.0, 0.3, 7800.0  # the reference's explicit example (src/main.cpp:155)


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ofa.Context(0)
    yield c
    c.close()


def mixed_clamped(seed=4):
    h = mg.hex_block(6, 3, 3, jitter=0.12, seed=seed)
    t = mg.kuhn_tets(6, 3, 3, jitter=0.12, seed=seed)
    m = h
    m.conn = np.vstack([h.conn, t.conn])
    m.etype = np.concatenate([h.etype, t.etype])
    fixed = np.zeros(m.nd, dtype=np.int32)
    fixed[h.fix_set - 1] = 1
    return m, fixed


def frames_close(a, b):
    return max(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300) for x, y in zip(a, b))


@pytest.mark.parametrize("kind", ["hex", "mixed"])
def test_explicit_analysis_auto_step_matches_oracle(ctx, kind):
    if kind == "hex":
        m = mg.hex_block(8, 3, 3, jitter=0.1, seed=2)
        fixed = np.zeros(m.nd, dtype=np.int32)
        fixed[m.fix_set - 1] = 1
    else:
        m, fixed = mixed_clamped()
    say(f"explicit_analysis auto {kind} nd={m.nd} ne={len(m.etype)}")
    resort = mg.free_index(m.nd, np.flatnonzero(fixed) + 1)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    mass = ctx.lumped_mass(RHO_REF)
    load = np.flatnonzero(m.xyz[:, 0] == m.xyz[:, 0].max()) + 1
    fext = fea.build_explicit_load(m.nd, load, 3, -250.7)
    assert set(np.unique(fext)) <= {0.0, -250.0}
    dt0 = 0.9 * O.crit_timestep(m.conn, m.etype, m.xyz, E, NU)
    total = 61.7 * dt0
    ref = O.explicit_timed(m.conn, m.etype, m.xyz, E, NU, mass, fext, fixed, total, 0.0, output_gap=10)
    assert ref["complete"]
    g64 = ctx.explicit_analysis(E, NU, mass, fext, fixed, total, 0.0, 10, precision=A.FP64)
    say(f"  steps gpu {g64['steps']} oracle {ref['steps']}, frames {len(g64['index'])}")
    assert g64["steps"] == ref["steps"]
    assert np.array_equal(g64["index"], ref["idx"]) and np.array_equal(g64["time"], ref["t"])
    e64 = frames_close(g64["frames"][1:], ref["frames"][1:])
    assert not g64["frames"][0].any()
    g32 = ctx.explicit_analysis(E, NU, mass, fext, fixed, total, 0.0, 10, precision=A.FP32)
    e32 = frames_close(g32["frames"][1:], ref["frames"][1:])
    say(f"  frames rel err fp64 {e64:.3e} fp32 {e32:.3e}")
    assert e64 < 1e-9
    assert np.array_equal(g32["index"], ref["idx"]) and e32 < 1e-4
    # a second run on the same context (graph reused) gives the same bits
    again = ctx.explicit_analysis(E, NU, mass, fext, fixed, total, 0.0, 10, precision=A.FP64)
    assert np.array_equal(again["frames"], g64["frames"])


def test_explicit_analysis_fixed_step(ctx):
    m = mg.hex_block(8, 3, 3, jitter=0.1, seed=2)
    fixed = np.zeros(m.nd, dtype=np.int32)
    fixed[m.fix_set - 1] = 1
    say("explicit_analysis fixed step")
    ctx.set_mesh(m.xyz, m.conn, m.etype, mg.free_index(m.nd, m.fix_set))
    mass = O.lumped_mass(m.conn, m.etype, m.xyz, mg.RHO_STEEL)
    v0 = mg.initial_velocity(m)
    fext = mg.face_traction(m, (0.0, 0.0, -1.0))
    dt, nsteps = 2.0 ** -25, 120  # 3e-8: below the stability limit of the unit steel cells
    # one output at the total: the loop is the fixed-step run
    g = ctx.explicit_analysis(E, NU, mass, fext, fixed, nsteps * dt, dt, 1, v0=v0, precision=A.FP64)
    run = ctx.explicit_run(E, NU, mass, fext, v0, fixed, dt, nsteps, precision=A.FP64)
    assert g["steps"] == nsteps and list(g["index"]) == [0, 1] and list(g["time"]) == [0.0, nsteps * dt]
    # the same loop; the tiled fixed-step run adds each node's element forces
    # in another fixed order (corner passes, then tiles) than the time-controlled
    # one (ascending elements), so the two agree to rounding
    assert rel_max(g["frames"][1], run) < 1e-12
    ref = O.explicit_timed(m.conn, m.etype, m.xyz, E, NU, mass, fext, fixed, nsteps * dt, dt, output_gap=1, v0=v0)
    assert rel_max(g["frames"][1], ref["frames"][1]) < 1e-9
    # a step that is clamped onto the outputs (0.1 total / 1.3 dt is not whole)
    total = 97.0 * dt
    g = ctx.explicit_analysis(E, NU, mass, fext, fixed, total, 1.3 * dt, 4, v0=v0, precision=A.FP64,
                              max_steps=5000)
    ref = O.explicit_timed(m.conn, m.etype, m.xyz, E, NU, mass, fext, fixed, total, 1.3 * dt, output_gap=4, v0=v0,
                           max_steps=5000)
    say(f"  clamped fixed step: gpu steps {g['steps']} oracle {ref['steps']} complete {ref['complete']}")
    assert np.array_equal(g["index"], ref["idx"]) and np.array_equal(g["time"], ref["t"])
    assert g["steps"] == ref["steps"]
    assert frames_close(g["frames"][1:], ref["frames"][1:]) < 1e-9


def test_explicit_analysis_limits_return_errors(ctx):
    """The reference's fixed step that lands on an output time sticks at 0
    (src/explicit_tools.cpp:33-46): max_steps ends it with OFEA_ERR_LIMIT;
    more saves than max_frames likewise."""
    m = mg.hex_block(4, 2, 2, jitter=0.1, seed=5)
    fixed = np.zeros(m.nd, dtype=np.int32)
    fixed[m.fix_set - 1] = 1
    say("explicit_analysis limits")
    ctx.set_mesh(m.xyz, m.conn, m.etype, mg.free_index(m.nd, m.fix_set))
    mass = O.lumped_mass(m.conn, m.etype, m.xyz, mg.RHO_STEEL)
    fext = np.zeros((m.nd, 3))
    fext[-1] = (0.0, 0.0, -3.0)
    dt = 2.0 ** -26
    with pytest.raises(A.OfeaError, match="max_steps"):
        ctx.explicit_analysis(E, NU, mass, fext, fixed, 40 * dt, dt, 4, max_steps=30, max_frames=64)
    with pytest.raises(A.OfeaError, match="max_frames"):
        ctx.explicit_analysis(E, NU, mass, fext, fixed, 40 * dt, dt, 1, max_frames=1)
    with pytest.raises(A.OfeaError):
        ctx.explicit_analysis(E, NU, mass, fext, fixed, 40 * dt, -dt, 1)
    # the context still works
    g = ctx.explicit_analysis(E, NU, mass, fext, fixed, 40 * dt, dt, 1)
    assert g["steps"] == 40


def test_explicit_analysis_reference_deck(tmp_path):
    """mesh_o.inp through the reference's explicit pipeline: Set-fix clamped,
    the Cload on Set-load int-truncated, outputs <path>_<k>.vtk."""
    deck = fea.read_inp(os.path.join(GOLDEN, "mesh_o.inp"))
    say(f"explicit_analysis mesh_o nd={deck.nd}")
    dt0 = 0.9 * O.crit_timestep(deck.conn, deck.etype, deck.xyz, 2.1e5, 0.3)
    total = 30.4 * dt0
    out = fea.explicit_analysis(deck, str(tmp_path / "run"), E=2.1e5, nu=0.3, rho=RHO_REF, time_total=total,
                                time_step=0.0)
    assert list(out["index"]) == list(range(11))
    for k in range(11):
        assert (tmp_path / f"run_{k}.vtk").exists()
    resort = mg.free_index(deck.nd, deck.fix_set)
    fixed = (resort < 0).astype(np.int32)
    mass = O.lumped_mass(deck.conn, deck.etype, deck.xyz, RHO_REF)
    fext = np.zeros(3 * deck.nd)
    fext[3 * (np.asarray(deck.load_set) - 1) + deck.load_dof - 1] = int(deck.load_value)
    ref = O.explicit_timed(deck.conn, deck.etype, deck.xyz, 2.1e5, 0.3, mass, fext, fixed, total, 0.0)
    assert out["steps"] == ref["steps"] and np.array_equal(out["time"], ref["t"])
    assert frames_close(out["frames"][1:], ref["frames"][1:]) < 1e-9


@pytest.mark.parametrize("method", [A.PCG_BLOCK_JACOBI, A.PCG_TWO_LEVEL])
def test_nl_solve_ref_matches_oracle(ctx, method):
    m, resort, F = mg.cantilever_c3d8(10, 4, 4)
    n = len(F)
    say(f"nl_solve_ref method {method} n={n}")
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    cp, ri = O.build_csc(m.conn, m.etype, resort, n)
    Fb = 40.0 * F
    u_ref, loops_ref, res_ref, hist = O.nl_newton_ref(m.conn, m.etype, m.xyz, resort, E, NU, Fb, cp, ri,
                                                      res_lmit=1e-6)
    u, loops, pcg, res, ok = ctx.nl_solve_ref(E, NU, Fb, 1e-6, 100, method, 1e-12)
    say(f"  loops gpu {loops} oracle {loops_ref} (history {['%.2e' % h for h in hist]}), pcg {pcg}")
    assert ok and loops == loops_ref
    assert abs(res - res_ref) <= 1e-3 * res_ref + 1e-10 * np.linalg.norm(Fb) / n
    assert rel_max(u, u_ref) < 1e-8
    # a loop cap
    u2, loops2, _, res2, ok2 = ctx.nl_solve_ref(E, NU, Fb, 1e-6, 1, method, 1e-12)
    assert loops2 == 1 and not ok2 and abs(res2 / hist[0] - 1.0) < 1e-12
