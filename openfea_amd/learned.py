"""Learned block-Jacobi correction: training pairs and training driver.

The network, its forward/backward kernels and the Adam loop live in the native
library (learned.hip, ofea_learned_*).  This module collects the training pairs
from the method's own solves and drives the training:

  * x* is solved to a tight tolerance with block-Jacobi PCG;
  * the iterates x_k of a block-Jacobi PCG run after k = 8, 16, 32, ... steps give
    errors e_k = x* - x_k and residuals r_k = K e_k (one SpMV each, on the GPU);
  * ofea_learned_train minimises the energy-norm error left by one
    preconditioned Richardson step over those pairs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A
from ._capi import check, lib, ptr


def nparam() -> int:
    return lib().ofea_learned_nparam()


def collect_pairs(ctx, rhs_list, snapshots=(8, 16, 32, 64, 128, 256), rtol_star=1e-11, two_level=False):
    """(E, R) arrays (S x n) from the solver's own PCG iterates on the resident K.

    Block-Jacobi PCG by default.  With two_level the iterates come from the
    two-level solver and the targets are the errors left after its coarse
    correction, e - Z A_c^-1 Z^T r: the network then learns the local part
    of the two-level preconditioner."""
    method = A.PCG_TWO_LEVEL if two_level else A.PCG_BLOCK_JACOBI
    E, R = [], []
    for b in rhs_list:
        xs, _, _, _ = ctx.solve(b, method=method, rtol=rtol_star, maxit=200000)
        for k in snapshots:
            xk, it, _, ok = ctx.solve(b, method=method, rtol=1e-30, maxit=k)
            e = xs - xk
            if not np.any(e):
                continue
            r = ctx.spmv(e)
            if two_level:
                e = e - coarse_apply(ctx, r)
            E.append(e)
            R.append(r)
    return np.ascontiguousarray(np.array(E)), np.ascontiguousarray(np.array(R))


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

    def set_samples(self, E, R):
        E = np.ascontiguousarray(E, dtype=np.float64)
        R = np.ascontiguousarray(R, dtype=np.float64)
        check(lib().ofea_learned_set_samples(self.ctx._h, ptr(E), ptr(R), len(E)))
        self._keep = (E, R)

    def loss_grad(self):
        loss = C.c_double()
        g = np.zeros(nparam(), dtype=np.float32)
        check(lib().ofea_learned_loss_grad(self.ctx._h, C.byref(loss), ptr(g)))
        return loss.value, g

    def train(self, steps=200, lr=2e-3):
        hist = np.zeros(max(steps, 1))
        check(lib().ofea_learned_train(self.ctx._h, steps, lr, ptr(hist)))
        return hist[:steps]

    def blocks(self):
        n = self.ctx.n // 3
        b = np.zeros(9 * n)
        f = np.zeros(9 * n, dtype=np.float32)
        check(lib().ofea_learned_blocks(self.ctx._h, ptr(b), ptr(f)))
        return b.reshape(n, 3, 3), f.reshape(n, 9)

    def precision(self, prec):
        """Inference precision of the hidden layer (A.FP32 or A.BF16 on the matrix cores)."""
        check(lib().ofea_learned_precision(self.ctx._h, prec))

    def refresh(self):
        check(lib().ofea_learned_refresh(self.ctx._h))
