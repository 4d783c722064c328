"""Pins the float64 oracle of the learned preconditioner (CPU only).

  * zero head => the blocks are exactly the block-Jacobi inverses D_r^-1;
  * the blocks are SPD for arbitrary parameters;
  * the analytic gradient equals central finite differences of the loss
    (every parameter group, random parameters with a non-zero head).
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import learned_oracle as LO
from openfea_amd import meshgen as mg


@pytest.fixture(scope="module")
def graph():
    m = mg.hex_block(3, 2, 2, jitter=0.15, seed=3)
    resort = mg.free_index(m.nd, m.fix_set)
    cp, ri = O.build_csc(m.conn, m.etype, resort, 3 * int(resort.max() + 1))
    val = O.fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri)
    return LO.Graph(cp, ri, val)


def random_params(seed, head_scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, LO.NPARAM)
    d = LO.unpack(p)
    d["Wo"] *= head_scale
    d["bo"] *= head_scale
    return LO.pack(d)


def test_zero_head_is_block_jacobi(graph):
    p = LO.unpack(random_params(1))
    p["Wo"][:] = 0
    p["bo"][:] = 0
    M, _ = LO.forward(graph, LO.pack(p))
    assert np.allclose(M, np.linalg.inv(graph.D), rtol=1e-12, atol=0)


def test_blocks_spd(graph):
    M, _ = LO.forward(graph, random_params(2, head_scale=0.5))
    assert np.allclose(M, M.transpose(0, 2, 1), rtol=1e-12)
    ev = np.linalg.eigvalsh(M)
    assert np.all(ev.min(axis=1) > 1e-10 * ev.max(axis=1))


def test_gradient_finite_differences(graph):
    rng = np.random.default_rng(7)
    E = rng.normal(size=(3, graph.n))
    R = np.array([graph.K @ e for e in E])
    p = random_params(4)
    loss, g = LO.loss_grad(graph, p, E, R)
    assert loss > 0
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in LO.SIZES])
    for gi in range(len(LO.SIZES)):
        idx = rng.choice(np.arange(offs[gi], offs[gi + 1]), size=min(4, offs[gi + 1] - offs[gi]), replace=False)
        for i in idx:
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (LO.loss_grad(graph, pp, E, R)[0] - LO.loss_grad(graph, pm, E, R)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-5 * np.abs(g).max(), (LO.SIZES[gi][0], i, fd, g[i])


def test_bf16_rounding():
    # BF16 keeps 7 mantissa bits: 1 + 2^-7 is exact; 1 + 2^-8 and 1 + 1.5 * 2^-7 are ties that go to the even mantissa
    x = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 1.5 * 2 ** -7, 1.0 + 2 ** -9, -3.0, 0.0], dtype=np.float32)
    assert np.array_equal(LO.bf16(x), [1.0, 1.0 + 2 ** -7, 1.0, 1.0 + 2 ** -6, 1.0, -3.0, 0.0])
    r = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.all(np.abs(LO.bf16(r) - r) <= 2 ** -8 * np.abs(r))
