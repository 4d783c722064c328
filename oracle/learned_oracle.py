"""Float64 numpy restatement of the learned factorized block-Jacobi correction.

TEST INFRASTRUCTURE ONLY (the checker of openfea_amd/csrc/learned.hip and of
the factorized preconditioner passes of solver.hip).  The learned
preconditioner has no counterpart in the reference; this oracle is pinned by
finite differences of its own loss (tests/test_learned_oracle.py), by the
identities "alpha = 0 is block Jacobi" and "the loss is 1 at alpha = 0", by
M = G^T G being SPD, by the loss equalling tr(G K G^T) / n computed from dense
matrices, and by the per-row minimiser of the loss solving its normal
equations.

Model.  D_i = L_i L_i^T is the 3x3 diagonal block of free node i (L_i lower
Cholesky), B_i = D_i.  G is block lower-triangular on K's node-block pattern:
    G_ii = L_i^-1,   G_is = -alpha_is L_i^-1 K_is B_s^-1   (s < i a neighbour),
    alpha_is = 2 sigmoid(u_i + v_s),  (u_i, v_i) = the network's two outputs,
and the preconditioner is M = G^T G (alpha = 0 gives block Jacobi D^-1).
Network: 9 node features, two node-graph layers (hidden 64, mean aggregation
over the neighbours), linear head 64 -> 2.

Loss: tr(G K G^T) / n.  G's diagonal blocks do not depend on alpha, so
det(G K G^T) = det(K) prod det(L_i)^-2 is fixed and the loss is Kaporin's
condition number of G K G^T up to a constant factor (the quality measure of
factorized sparse approximate inverses).  Row i contributes
T_i = sum_{s,t in P_i} tr(G_is K_st G_it^T), P_i = {s <= i}; with
W_s = (K G^T)[s, i] = sum_t K_st G_it^T, dT_i / dalpha_is = -2 tr(C_is W_s)
(C_is = L_i^-1 K_is B_s^-1).

Layout: parameters W1s, W1n, b1, W2s, W2n, b2, Wo, bo (as learned.hip).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

F, H, O = 9, 64, 2
SIZES = [("W1s", (H, F)), ("W1n", (H, F)), ("b1", (H,)), ("W2s", (H, H)), ("W2n", (H, H)), ("b2", (H,)),
         ("Wo", (O, H)), ("bo", (O,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)
HEAD_BIAS0 = 0.0  # initial bias of u: alpha = 2 sigmoid(0) = 1, the first-order rule


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
    """Free-node adjacency, the 3x3 blocks of K and K itself from the CSC arrays."""

    def __init__(self, col_ptr, row_idx, val):
        n = len(col_ptr) - 1
        self.n, self.N = n, n // 3
        self.K = sp.csc_matrix((val, row_idx, col_ptr), shape=(n, n)).tocsr()
        cp = np.asarray(col_ptr)
        deg = (cp[3::3][: self.N] - cp[0:-1:3][: self.N]) // 9  # 3 columns of 3*deg rows
        self.deg = deg.astype(np.int64)
        self.nbrs = []
        self.blk = []  # per node: deg x 3 x 3, blk[i][k][c][d] = K[3i + c, 3 s_k + d]
        D = np.zeros((self.N, 3, 3))
        offn = np.zeros(self.N)
        dn = np.zeros(self.N)
        for r in range(self.N):
            rows = row_idx[cp[3 * r]:cp[3 * r] + 3 * deg[r]:3] // 3
            self.nbrs.append(np.asarray(rows))
            B = np.zeros((deg[r], 3, 3))
            for k, s in enumerate(rows):
                blk = np.array([[val[cp[3 * r + c] + 3 * k + d] for d in range(3)] for c in range(3)])
                B[k] = blk
                f = np.linalg.norm(blk)
                if s == r:
                    D[r] = blk
                    dn[r] = f
                else:
                    offn[r] += f
            self.blk.append(B)
        self.D = D
        self.offn, self.dn = offn, dn
        self.Linv = np.linalg.inv(np.linalg.cholesky(D))  # D = L L^T
        self.Binv = np.linalg.inv(D)

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

    def lower_edges(self):
        """(i, k, s) for every neighbour s < i of every free node i"""
        return [(i, k, int(s)) for i in range(self.N) for k, s in enumerate(self.nbrs[i]) if s < i]

    def block(self, s, t):
        """K_st (zero when s and t are not neighbours)"""
        return self.K[3 * s:3 * s + 3, 3 * t:3 * t + 3].toarray()

    def C(self, i, k, s):
        """C_is = L_i^-1 K_is B_s^-1 (G_is = -alpha_is C_is)"""
        return self.Linv[i] @ self.blk[i][k] @ self.Binv[s]


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def network(g: Graph, p, bf16_hidden=False):
    """Head outputs (N x 2) and the activations.  bf16_hidden rounds the
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
    return Oo, dict(X=X, A0=A0, Z1=Z1, H1=H1, A1=A1, Z2=Z2, H2=H2, O=Oo)


def alpha(Oo, i, s):
    return 2.0 * sigmoid(Oo[i, 0] + Oo[s, 1])


def factor(g: Graph, Oo=None, alphas=None):
    """G (sparse, n x n) and its blocks on K's node-block layout: per node i,
    deg x 3 x 3 with slot k = G_is (s < i), G_ii (s == i), (G_si)^T (s > i).
    alpha_is from the network outputs Oo, or from alphas[(i, s)]."""
    al = (lambda i, s: alphas[(i, s)]) if alphas is not None else (lambda i, s: alpha(Oo, i, s))
    rows, cols, vals = [], [], []
    slots = []
    for i in range(g.N):
        S = np.zeros((g.deg[i], 3, 3))
        for k, s in enumerate(g.nbrs[i]):
            if s == i:
                S[k] = g.Linv[i]
            elif s < i:
                S[k] = -al(i, s) * g.C(i, k, s)
            else:  # (G_si)^T = -alpha_si (L_s^-1 K_si B_i^-1)^T = -alpha_si B_i^-1 K_is L_s^-T
                S[k] = -al(s, i) * g.Binv[i] @ g.blk[i][k] @ g.Linv[s].T
            if s <= i:
                for c in range(3):
                    for d in range(3):
                        rows.append(3 * i + c)
                        cols.append(3 * s + d)
                        vals.append(S[k][c, d])
        slots.append(S)
    G = sp.csr_matrix((vals, (rows, cols)), shape=(g.n, g.n))
    return G, slots


def trace_loss(g: Graph, G):
    """tr(G K G^T) / n"""
    return float((G @ g.K).multiply(G).sum()) / g.n


def edge_grad(g: Graph, G):
    """d loss / d alpha_is for every lower pair: -2 tr(C_is W_s) / n, W_s = (K G^T)[s, i]"""
    KGt = (g.K @ G.T).tocsr()
    out = {}
    for (i, k, s) in g.lower_edges():
        W = KGt[3 * s:3 * s + 3, 3 * i:3 * i + 3].toarray()
        out[(i, s)] = -2.0 * np.trace(g.C(i, k, s) @ W) / g.n
    return out


def forward(g: Graph, p, bf16_hidden=False):
    Oo, cache = network(g, p, bf16_hidden)
    G, slots = factor(g, Oo)
    cache["G"] = G
    return slots, cache


def apply_M(G, r):
    return G.T @ (G @ r)


def backward(g: Graph, P, c, dO):
    """parameter gradient from dloss/dO (N x 2) through the head and the two layers"""
    gr = {"Wo": dO.T @ c["H2"], "bo": dO.sum(0)}
    dZ2 = (dO @ P["Wo"]) * (c["Z2"] > 0)
    gr["W2s"], gr["W2n"], gr["b2"] = dZ2.T @ c["H1"], dZ2.T @ c["A1"], dZ2.sum(0)
    dH1 = dZ2 @ P["W2s"] + g.agg_T(dZ2 @ P["W2n"])
    dZ1 = dH1 * (c["Z1"] > 0)
    gr["W1s"], gr["W1n"], gr["b1"] = dZ1.T @ c["X"], dZ1.T @ c["A0"], dZ1.sum(0)
    return pack(gr)


def loss_grad(g: Graph, p):
    """loss = tr(G K G^T) / n and its gradient with respect to the parameters."""
    P = unpack(p)
    Oo, c = network(g, p)
    G, _ = factor(g, Oo)
    dO = np.zeros((g.N, 2))
    for (i, s), da in edge_grad(g, G).items():
        a = alpha(Oo, i, s)
        t = da * a * (1.0 - 0.5 * a)  # d alpha / d x = 2 sigmoid (1 - sigmoid) = a (1 - a / 2)
        dO[i, 0] += t
        dO[s, 1] += t
    return trace_loss(g, G), backward(g, P, c, dO)


def row_optimal_alpha(g: Graph, i):
    """The alphas of row i's lower pairs that minimise T_i: the solution of
    A a = b with A_st = tr(C_is K_st C_it^T), b_s = tr(C_is K_si L_i^-T)."""
    low = [(k, int(s)) for k, s in enumerate(g.nbrs[i]) if s < i]
    Cs = [g.C(i, k, s) for k, s in low]
    b = np.array([np.trace(C @ g.block(s, i) @ g.Linv[i].T) for C, (k, s) in zip(Cs, low)])
    A = np.zeros((len(low), len(low)))
    for a_, (Ca, (ka, sa)) in enumerate(zip(Cs, low)):
        for b_, (Cb, (kb, sb)) in enumerate(zip(Cs, low)):
            A[a_, b_] = np.trace(Ca @ g.block(sa, sb) @ Cb.T)
    return [s for _, s in low], np.linalg.solve(A, b)
