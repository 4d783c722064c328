"""Learned factorized correction of block Jacobi: Python driver.

The network, its forward/backward kernels and the Adam loop live in the native
library (learned.hip, ofea_learned_*).  The preconditioner is M = G^T G with
G_ii = L_i^-1 (D_i = L_i L_i^T) and G_is = -alpha_is L_i^-1 K_is D_s^-1 for the
lower neighbours s < i, alpha_is = 2 sigmoid(u_i + v_s) from the network.
Training minimises tr(G K G^T) / n over one or several resident systems (for
example the design variants of one mesh): with G's diagonal blocks fixed that
is Kaporin's condition number of G K G^T up to a constant, and it tracks the
PCG iteration count (it rises where a too-large alpha creates small
eigenvalues that losses on sampled error vectors do not see).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A
from ._capi import check, lib, ptr


def nparam() -> int:
    return lib().ofea_learned_nparam()


def coarse_apply(ctx, r):
    """Z A_c^-1 Z^T r of the two-level methods for the context's K."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros_like(r)
    check(lib().ofea_coarse_apply(ctx._h, ptr(r), ptr(out)))
    return out


def coarse_aggregates(ctx):
    """Aggregate of every free node of the two-level coarse space (and their count)."""
    agg = np.zeros(ctx.n // 3, dtype=np.int32)
    nagg = lib().ofea_coarse_aggregates(ctx._h, ptr(agg))
    if nagg < 0:
        check(nagg)
    return agg, nagg


def coarse_builds(ctx):
    """Coarse inverses A_c^-1 built since the mesh was set."""
    n = lib().ofea_coarse_builds(ctx._h)
    if n < 0:
        check(n)
    return n


class Learned:
    """Handle on the learned preconditioner of one context."""

    def __init__(self, ctx, seed: int = 0):
        self.ctx = ctx
        check(lib().ofea_learned_init(ctx._h, seed))

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().ofea_learned_set_params(self.ctx._h, ptr(p), len(p)))

    def get_params(self):
        p = np.zeros(nparam(), dtype=np.float32)
        check(lib().ofea_learned_get_params(self.ctx._h, ptr(p), len(p)))
        return p

    def loss_grad(self):
        """tr(G K G^T) / n for the resident K and its gradient"""
        loss = C.c_double()
        g = np.zeros(nparam(), dtype=np.float32)
        check(lib().ofea_learned_loss_grad(self.ctx._h, C.byref(loss), ptr(g)))
        return loss.value, g

    def train(self, steps=200, lr=2e-3):
        hist = np.zeros(max(steps, 1))
        check(lib().ofea_learned_train(self.ctx._h, steps, lr, ptr(hist)))
        return hist[:steps]

    def factor(self):
        """G on K's node-block layout (nnz values: per free node i and block
        slot k a row-major 3x3: G_is for s < i, G_ii, (G_si)^T for s > i) and
        the node features (nfree x 9)."""
        n = self.ctx.n // 3
        b = np.zeros(self.ctx.nnz)
        f = np.zeros(9 * n, dtype=np.float32)
        check(lib().ofea_learned_blocks(self.ctx._h, ptr(b), ptr(f)))
        return b, f.reshape(n, 9)

    def outputs(self):
        """the network's outputs (u, v) per free node (nfree x 2), as they serve the solver"""
        uv = np.zeros(self.ctx.n // 3 * 2, dtype=np.float32)
        check(lib().ofea_learned_outputs(self.ctx._h, ptr(uv)))
        return uv.reshape(-1, 2)

    def precision(self, prec):
        """Inference precision of the hidden layer (A.FP32 or A.BF16 on the matrix cores)."""
        check(lib().ofea_learned_precision(self.ctx._h, prec))

    def refresh(self):
        check(lib().ofea_learned_refresh(self.ctx._h))


def train_multi(learned, steps=200, lr=2e-3):
    """Adam on the mean loss of several contexts' resident systems (Learned
    handles on one device); the parameters start from the first handle's and
    end in all of them.  Returns the loss history."""
    hs = (C.c_void_p * len(learned))(*[ln.ctx._h for ln in learned])
    hist = np.zeros(max(steps, 1))
    check(lib().ofea_learned_train_multi(hs, len(learned), steps, lr, ptr(hist)))
    return hist[:steps]
