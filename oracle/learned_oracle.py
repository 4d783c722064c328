"""Float64 numpy restatement of the learned block-Jacobi correction.

TEST INFRASTRUCTURE ONLY (the checker of openfea_amd/csrc/learned.hip).  The
learned preconditioner has no counterpart in the reference; this oracle is
pinned by finite differences of its own loss (tests/test_learned_oracle.py)
and by the identity "zero head == block Jacobi".

Layout: parameters as in learned.hip (W1s, W1n, b1, W2s, W2n, b2, Wo, bo);
K given by the CSC arrays of the free DOFs (symmetric, node-blocked).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

F, H, O = 9, 64, 6
SIZES = [("W1s", (H, F)), ("W1n", (H, F)), ("b1", (H,)), ("W2s", (H, H)), ("W2n", (H, H)), ("b2", (H,)),
         ("Wo", (O, H)), ("bo", (O,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


class Graph:
    """Free-node adjacency, diagonal blocks and K from the CSC arrays."""

    def __init__(self, col_ptr, row_idx, val):
        n = len(col_ptr) - 1
        self.n, self.N = n, n // 3
        self.K = sp.csc_matrix((val, row_idx, col_ptr), shape=(n, n)).tocsr()
        cp = np.asarray(col_ptr)
        deg = (cp[3::3][: self.N] - cp[0:-1:3][: self.N]) // 9  # 3 columns of 3*deg rows
        self.deg = deg.astype(np.int64)
        self.nbrs = []
        D = np.zeros((self.N, 3, 3))
        offn = np.zeros(self.N)
        dn = np.zeros(self.N)
        for r in range(self.N):
            rows = row_idx[cp[3 * r]:cp[3 * r] + 3 * deg[r]:3] // 3
            self.nbrs.append(np.asarray(rows))
            for k, s in enumerate(rows):
                blk = np.array([[val[cp[3 * r + c] + 3 * k + d] for d in range(3)] for c in range(3)])
                f = np.linalg.norm(blk)
                if s == r:
                    D[r] = blk
                    dn[r] = f
                else:
                    offn[r] += f
        self.D = D
        self.offn, self.dn = offn, dn

    def features(self):
        D = self.D
        t = np.trace(D, axis1=1, axis2=2) / 3.0
        X = np.zeros((self.N, F))
        X[:, 0], X[:, 1], X[:, 2] = D[:, 0, 0] / t, D[:, 1, 1] / t, D[:, 2, 2] / t
        X[:, 3], X[:, 4], X[:, 5] = D[:, 0, 1] / t, D[:, 1, 2] / t, D[:, 0, 2] / t
        lt = np.log(t)
        X[:, 6] = lt - lt.mean()
        X[:, 7] = self.deg / 27.0
        X[:, 8] = np.where((self.deg > 1) & (self.dn > 0), self.offn / np.maximum(self.deg - 1, 1) / self.dn, 0.0)
        return X

    def agg(self, Hm):
        """mean over neighbours s != r"""
        out = np.zeros_like(Hm)
        for r in range(self.N):
            s = self.nbrs[r][self.nbrs[r] != r]
            if len(s):
                out[r] = Hm[s].mean(axis=0)
        return out

    def agg_T(self, Y):
        """adjoint of agg: sum over r in N(s) of Y[r] / (deg_r - 1)"""
        out = np.zeros_like(Y)
        for r in range(self.N):
            s = self.nbrs[r][self.nbrs[r] != r]
            if len(s):
                out[s] += Y[r] / len(s)
        return out

    def rinv(self):
        R = np.linalg.cholesky(self.D).transpose(0, 2, 1)  # D = R^T R, R upper
        return np.linalg.inv(R)


def lmat(o):
    L = np.zeros((len(o), 3, 3))
    L[:, 0, 0], L[:, 1, 1], L[:, 2, 2] = np.exp(o[:, 0]), np.exp(o[:, 1]), np.exp(o[:, 2])
    L[:, 1, 0], L[:, 2, 1], L[:, 2, 0] = o[:, 3], o[:, 4], o[:, 5]
    return L


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(g: Graph, p, bf16_hidden=False):
    """Blocks M (n x 3 x 3) and the activations.  bf16_hidden rounds the
    hidden layer's inputs and weights to BF16 (products accumulated exactly),
    the arithmetic of the library's BF16 inference."""
    P = unpack(p)
    X = g.features()
    A0 = g.agg(X)
    Z1 = X @ P["W1s"].T + A0 @ P["W1n"].T + P["b1"]
    H1 = np.maximum(Z1, 0)
    A1 = g.agg(H1)
    if bf16_hidden:
        Z2 = bf16(H1) @ bf16(P["W2s"]).T + bf16(A1) @ bf16(P["W2n"]).T + P["b2"]
    else:
        Z2 = H1 @ P["W2s"].T + A1 @ P["W2n"].T + P["b2"]
    H2 = np.maximum(Z2, 0)
    Oo = H2 @ P["Wo"].T + P["bo"]
    A = g.rinv()
    B = A @ lmat(Oo)
    M = B @ B.transpose(0, 2, 1)
    cache = dict(X=X, A0=A0, Z1=Z1, H1=H1, A1=A1, Z2=Z2, H2=H2, O=Oo, A=A)
    return M, cache


def apply_blocks(M, r):
    return np.einsum("nij,nj->ni", M, r.reshape(-1, 3)).ravel()


def loss_grad(g: Graph, p, E, R):
    """loss = (1/S) sum_k (e-Mr)^T K (e-Mr) / (e^T r) and its gradient."""
    P = unpack(p)
    M, c = forward(g, p)
    S = len(E)
    G = np.zeros((g.N, 3, 3))
    loss = 0.0
    for e, r in zip(E, R):
        w = 1.0 / (S * (e @ r))
        d = e - apply_blocks(M, r)
        q = g.K @ d
        loss += w * (d @ q)
        G -= 2.0 * w * np.einsum("ni,nj->nij", q.reshape(-1, 3), r.reshape(-1, 3))
    A = c["A"]
    L = lmat(c["O"])
    Hm = A.transpose(0, 2, 1) @ G @ A
    dL = (Hm + Hm.transpose(0, 2, 1)) @ L
    dO = np.stack([dL[:, 0, 0] * L[:, 0, 0], dL[:, 1, 1] * L[:, 1, 1], dL[:, 2, 2] * L[:, 2, 2],
                   dL[:, 1, 0], dL[:, 2, 1], dL[:, 2, 0]], axis=1)
    gr = {"Wo": dO.T @ c["H2"], "bo": dO.sum(0)}
    dZ2 = (dO @ P["Wo"]) * (c["Z2"] > 0)
    gr["W2s"], gr["W2n"], gr["b2"] = dZ2.T @ c["H1"], dZ2.T @ c["A1"], dZ2.sum(0)
    dH1 = dZ2 @ P["W2s"] + g.agg_T(dZ2 @ P["W2n"])
    dZ1 = dH1 * (c["Z1"] > 0)
    gr["W1s"], gr["W1n"], gr["b1"] = dZ1.T @ c["X"], dZ1.T @ c["A0"], dZ1.sum(0)
    return loss, pack(gr)
