// learned.hip -- learned factorized correction of the block-Jacobi preconditioner.
//
// No counterpart in the reference (its solvers are direct, src/linear_solution.cpp).
// With D_i = L_i L_i^T the 3x3 diagonal block of free node i (L_i lower
// Cholesky), the preconditioner is M = G^T G for the block lower-triangular
//   G_ii = L_i^-1,   G_is = -alpha_is L_i^-1 K_is D_s^-1   (s < i a neighbour of i),
// the first step of a factorized sparse approximate inverse on K's node-block
// pattern.  alpha = 0 is block Jacobi; a 2-layer node-graph network sets
//   alpha_is = 2 sigmoid(u_i + v_s),  (u_i, v_i) = its two outputs at node i.
// M is SPD for any parameters (G is triangular with nonsingular diagonal
// blocks), so PCG stays a valid CG; plain block Jacobi is the fallback and
// judge.  Applying M costs one pass over each half of G (BF16) per iteration
// (the update kernel's y recursion and k_fac_upper in solver.hip).
//
//   features x_r (9): D_r / t_r (6 entries), log t_r - mean log t, deg_r / 27,
//                     sum_{s!=r} ||K_rs||_F / ((deg_r - 1) ||D_r||_F);  t_r = tr(D_r)/3
//   layer l:  h' = relu(W_self h_r + W_nbr mean_{s in N(r), s != r} h_s + b)   (9 -> 64 -> 64)
//   head:     (u_r, v_r) = W_o h2_r + b_o   (W_o, b_o start at zero: alpha = 1,
//             the untrained first-order rule)
//
// Training loss: tr(G K G^T) / n on the resident K.  G's diagonal blocks do
// not depend on alpha, so det(G K G^T) is fixed and the loss is Kaporin's
// condition number of G K G^T up to a constant factor, the quality measure of
// factorized sparse approximate inverses; it is 1 at alpha = 0 and, unlike a
// loss on sampled error vectors, it rises when alpha creates small
// eigenvalues.  Row i contributes T_i = sum_{s,t <= i} tr(G_is K_st G_it^T);
// with W_s = sum_{t <= i} K_st G_it^T, dT_i / dalpha_is = -2 tr(C_is W_s)
// (C_is = L_i^-1 K_is D_s^-1), then dalpha/dx = alpha (1 - alpha / 2) and the
// network's backward pass.  Network arithmetic FP32; G, the loss and the
// per-pair derivatives FP64; the solver's copy of G BF16.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "elem_device.h"
#include "ofea_internal.h"
#include "solver_ws.h"

namespace ofea {

constexpr int LN_F = 9;     // input features
constexpr int LN_H = 64;    // hidden width
constexpr int LN_O = 2;     // head outputs (u, v)
constexpr float LN_HEAD_BIAS0 = 0.0f;  // initial u: alpha = 2 sigmoid(0) = 1, the first-order rule
constexpr int LN_WG = 256;

// parameter layout (floats)
constexpr int OFF_W1S = 0;
constexpr int OFF_W1N = OFF_W1S + LN_H * LN_F;
constexpr int OFF_B1 = OFF_W1N + LN_H * LN_F;
constexpr int OFF_W2S = OFF_B1 + LN_H;
constexpr int OFF_W2N = OFF_W2S + LN_H * LN_H;
constexpr int OFF_B2 = OFF_W2N + LN_H * LN_H;
constexpr int OFF_WO = OFF_B2 + LN_H;
constexpr int OFF_BO = OFF_WO + LN_O * LN_H;
constexpr int LN_NPARAM = OFF_BO + LN_O;

struct LearnedWS {
    int nfree = 0;
    std::vector<float> params;  // host copy
    float *d_params = nullptr, *d_grad = nullptr;
    float *X = nullptr, *A0 = nullptr;              // N x 9
    float *Z1 = nullptr, *H1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *H2 = nullptr;  // N x 64
    float *O = nullptr;                             // N x 2
    double *Rinv = nullptr;                         // N x 9 (upper-triangular R^-1, D = R^T R, row-major)
    float *dO = nullptr;                            // N x 2
    float *dZ2 = nullptr, *dH1 = nullptr, *dZ1 = nullptr, *T1 = nullptr;  // N x 64
    float *partial = nullptr;                       // weight-gradient partials
    int npartial_rows = 0;
    int gen = 0;                                    // incremented by every forward pass that serves the solver
    double *Trow = nullptr;                         // N: T_i of the loss
    double *dA = nullptr;                           // lower blocks of G: dloss/dalpha_is (0 on the diagonal)
    double *tpart = nullptr;                        // partial sums of T
    int feat_kver = -1;  // K version the features were computed from
    bool bf16 = false;   // inference of layer 2 on the matrix cores in BF16 (training stays FP32)
};

static LearnedWS *lws(ofea_ctx *c) { return (LearnedWS *)c->learned; }

void learned_free(ofea_ctx *c)
{
    LearnedWS *w = lws(c);
    if (!w) return;
    for (void *p : {(void *)w->d_params, (void *)w->d_grad, (void *)w->X, (void *)w->A0, (void *)w->Z1, (void *)w->H1,
                    (void *)w->A1, (void *)w->Z2, (void *)w->H2, (void *)w->O, (void *)w->Rinv, (void *)w->dO,
                    (void *)w->dZ2, (void *)w->dH1, (void *)w->dZ1, (void *)w->T1, (void *)w->partial,
                    (void *)w->Trow, (void *)w->dA, (void *)w->tpart})
        if (p) (void)hipFree(p);
    delete w;
    c->learned = nullptr;
}

// ------------------------------------------------------------------ kernels
// Features and the Cholesky inverse of each node's diagonal block.
__global__ void k_ln_features(int nfree, const int *__restrict__ node_blk, const int *__restrict__ bcol,
                              const double *__restrict__ val, const double *__restrict__ dblk, float *__restrict__ X,
                              double *__restrict__ Rinv, double *__restrict__ logt)
{
    const int r = blockIdx.x * LN_WG + threadIdx.x;
    if (r >= nfree) return;
    const double *D = dblk + 9 * r;
    const double t = (D[0] + D[4] + D[8]) / 3.0;
    const double it = 1.0 / t;
    float *x = X + LN_F * r;
    x[0] = (float)(D[0] * it);
    x[1] = (float)(D[4] * it);
    x[2] = (float)(D[8] * it);
    x[3] = (float)(D[1] * it);
    x[4] = (float)(D[5] * it);
    x[5] = (float)(D[2] * it);
    logt[r] = log(t);
    const int nb0 = node_blk[r], deg = node_blk[r + 1] - nb0;
    x[7] = (float)(deg / 27.0);
    double off = 0.0, dn = 0.0;
    for (int k = 0; k < deg; ++k) {
        double f = 0.0;
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d) {
                const double v = val[9 * (long long)nb0 + c * 3 * deg + 3 * k + d];
                f += v * v;
            }
        if (bcol[nb0 + k] == r) dn = sqrt(f);
        else off += sqrt(f);
    }
    x[8] = (float)(deg > 1 && dn > 0.0 ? off / ((deg - 1) * dn) : 0.0);
    // D = R^T R (R upper), store R^-1 row-major
    const double R00 = sqrt(D[0]);
    const double R01 = D[1] / R00, R02 = D[2] / R00;
    const double R11 = sqrt(D[4] - R01 * R01);
    const double R12 = (D[5] - R01 * R02) / R11;
    const double R22 = sqrt(D[8] - R02 * R02 - R12 * R12);
    double *A = Rinv + 9 * r;
    A[0] = 1.0 / R00;
    A[1] = -R01 / (R00 * R11);
    A[2] = (R01 * R12 - R02 * R11) / (R00 * R11 * R22);
    A[3] = 0.0;
    A[4] = 1.0 / R11;
    A[5] = -R12 / (R11 * R22);
    A[6] = 0.0;
    A[7] = 0.0;
    A[8] = 1.0 / R22;
}

// mean of log t over the free nodes (one workgroup, fixed order) -> x[6]
__global__ void k_ln_center(int nfree, const double *__restrict__ logt, float *__restrict__ X)
{
    __shared__ double sm[LN_WG / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < nfree; i += LN_WG) s += logt[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int k = 0; k < LN_WG / 64; ++k) tot += sm[k];
    const double mean = tot / nfree;
    for (int i = threadIdx.x; i < nfree; i += LN_WG) X[LN_F * i + 6] = (float)(logt[i] - mean);
}

// Hout[r] = mean over neighbours s != r of Hin[s]  (C channels).  With
// transpose != 0 it computes the adjoint, sum over s in N(r) of Hin[s]/(deg_s - 1),
// which equals the transpose because the adjacency is symmetric.
__global__ void k_ln_agg(int nfree, int C, const int *__restrict__ node_blk, const int *__restrict__ bcol,
                         const float *__restrict__ Hin, float *__restrict__ Hout, int transpose)
{
    const long long idx = (long long)blockIdx.x * LN_WG + threadIdx.x;
    if (idx >= (long long)nfree * C) return;
    const int r = (int)(idx / C), ch = (int)(idx % C);
    const int nb0 = node_blk[r], deg = node_blk[r + 1] - nb0;
    float acc = 0.0f;
    for (int k = 0; k < deg; ++k) {
        const int s = bcol[nb0 + k];
        if (s == r) continue;
        float v = Hin[(long long)s * C + ch];
        if (transpose) {
            const int ds = node_blk[s + 1] - node_blk[s] - 1;
            v = ds > 0 ? v / (float)ds : 0.0f;
        }
        acc += v;
    }
    if (!transpose) acc = deg > 1 ? acc / (float)(deg - 1) : 0.0f;
    Hout[idx] = acc;
}

// Z = W_s Hs + W_n Ha + b; H = relu(Z).  Thread per (node, out channel).
__global__ void k_ln_dense(int nfree, int Cin, int Cout, const float *__restrict__ Hs, const float *__restrict__ Ha,
                           const float *__restrict__ Ws, const float *__restrict__ Wn, const float *__restrict__ b,
                           float *__restrict__ Z, float *__restrict__ H)
{
    const long long idx = (long long)blockIdx.x * LN_WG + threadIdx.x;
    if (idx >= (long long)nfree * Cout) return;
    const int r = (int)(idx / Cout), o = (int)(idx % Cout);
    const float *hs = Hs + (long long)r * Cin, *ha = Ha + (long long)r * Cin;
    const float *ws = Ws + o * Cin, *wn = Wn + o * Cin;
    float acc = b[o];
    for (int i = 0; i < Cin; ++i) acc += ws[i] * hs[i] + wn[i] * ha[i];
    Z[idx] = acc;
    H[idx] = acc > 0.0f ? acc : 0.0f;
}

// BF16 with round-to-nearest-even (finite inputs)
__device__ inline short to_bf16(float f)
{
    unsigned u = __builtin_bit_cast(unsigned, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

typedef short ln_short4 __attribute__((ext_vector_type(4)));
typedef float ln_float4 __attribute__((ext_vector_type(4)));

// Inference form of k_ln_dense for CIN = COUT = 64 on the matrix cores:
// Z = [Hs | Ha] [Ws | Wn]^T + b (K = 128), BF16 operands, FP32 accumulation,
// v_mfma_f32_16x16x16_bf16.  Four waves per workgroup, each a 16-node x
// 64-channel tile (four 16x16 accumulators, eight K steps).  Operand layout
// of the instruction: lane l supplies A[l % 16][4 (l / 16) + i] and
// B[4 (l / 16) + i][l % 16] (i = 0..3) and receives D[4 (l / 16) + i][l % 16].
// The weights are converted to BF16 once per workgroup into LDS.
constexpr int LN_MF_WG = 256;
__global__ __launch_bounds__(LN_MF_WG) void k_ln_dense_mfma(int nfree, const float *__restrict__ Hs,
                                                            const float *__restrict__ Ha, const float *__restrict__ Ws,
                                                            const float *__restrict__ Wn, const float *__restrict__ b,
                                                            float *__restrict__ Z, float *__restrict__ H)
{
    constexpr int CIN = LN_H, COUT = LN_H, K = 2 * CIN;
    __shared__ short s_w[COUT][K + 4];  // +4: rows start on different banks
    for (int i = threadIdx.x; i < COUT * K; i += LN_MF_WG) {
        const int o = i / K, k = i - o * K;
        s_w[o][k] = to_bf16(k < CIN ? Ws[o * CIN + k] : Wn[o * CIN + k - CIN]);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = (blockIdx.x * (LN_MF_WG / 64) + wave) * 16;
    const int m = lane & 15, kq = lane >> 4;
    const int node = row0 + m;
    ln_float4 acc[COUT / 16];
#pragma unroll
    for (int t = 0; t < COUT / 16; ++t) acc[t] = ln_float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < K; kk += 16) {
        const int k0 = kk + 4 * kq;  // never straddles the Hs | Ha boundary (CIN % 4 == 0)
        const float *src = k0 < CIN ? Hs + (size_t)node * CIN + k0 : Ha + (size_t)node * CIN + (k0 - CIN);
        ln_short4 a = {0, 0, 0, 0};
        if (node < nfree) {
            const float4 v = *reinterpret_cast<const float4 *>(src);
            a = ln_short4{to_bf16(v.x), to_bf16(v.y), to_bf16(v.z), to_bf16(v.w)};
        }
#pragma unroll
        for (int t = 0; t < COUT / 16; ++t) {
            const short *wr = &s_w[16 * t + m][k0];
            const ln_short4 bv = {wr[0], wr[1], wr[2], wr[3]};
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < COUT / 16; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = row0 + 4 * kq + i, o = 16 * t + m;
            if (r < nfree) {
                const float z = acc[t][i] + b[o];
                Z[(size_t)r * COUT + o] = z;
                H[(size_t)r * COUT + o] = z > 0.0f ? z : 0.0f;
            }
        }
}

// head: (u, v) = Wo h2 + bo.  Thread per node.
__global__ void k_ln_head(int nfree, const float *__restrict__ H2, const float *__restrict__ Wo,
                          const float *__restrict__ bo, float *__restrict__ O)
{
    const int r = blockIdx.x * LN_WG + threadIdx.x;
    if (r >= nfree) return;
    const float *h = H2 + (long long)r * LN_H;
    for (int j = 0; j < LN_O; ++j) {
        float acc = bo[j];
        for (int i = 0; i < LN_H; ++i) acc += Wo[j * LN_H + i] * h[i];
        O[LN_O * r + j] = acc;
    }
}

__device__ inline double ln_alpha(float u, float v) { return 2.0 / (1.0 + exp(-(double)u - (double)v)); }

// 3x3 products (row-major)
__device__ inline void mm3(const double A[9], const double B[9], double C[9])
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// node data through an optional map (reordered rows: map[i] = the caller's node)
struct FacNodes {
    const double *Rinv;  // 9 per node, R^-1 with D = R^T R (L^-1 = (R^-1)^T)
    const double *Binv;  // 9 per node, D^-1
    const float *O;      // 2 per node, (u, v)
    const int *map;
    __device__ int at(int i) const { return map ? map[i] : i; }
    __device__ void linv(int i, double L[9]) const
    {
        const double *A = Rinv + 9 * (size_t)at(i);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) L[3 * a + b] = A[3 * b + a];
    }
    __device__ void binv(int i, double B[9]) const
    {
        const double *A = Binv + 9 * (size_t)at(i);
#pragma unroll
        for (int e = 0; e < 9; ++e) B[e] = A[e];
    }
    __device__ double alpha(int i, int s) const { return ln_alpha(O[2 * at(i)], O[2 * at(s) + 1]); }
};

// K_is of node i's block slot k ([c][d] = K[3i + c][3s + d]) from K's node-block layout
__device__ inline void k_block(const double *val, int nb0, int deg, int k, double Kb[9])
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) Kb[3 * c + d] = val[9 * (long long)nb0 + c * 3 * deg + 3 * k + d];
}

// G as the two half matrices of the solver (FacView in solver.hip, BF16): the
// blocks of K's row i in K's order, s <= i into row i of lo, s >= i into row
// i of up, each in the column-of-rows layout (entry (c, 3 k + d) of a row of
// nb blocks at 9 ptr[i] + 3 nb c + 3 k + d):
//   lo: s < i: G_is = -alpha_is L_i^-1 K_is D_s^-1;  s == i: L_i^-1;
//   up: s > i: (G_si)^T = -alpha_si D_i^-1 K_is L_s^-T (K_si = K_is^T);  s == i: L_i^-T.
// Thread per node (its blocks in a loop).
__global__ void k_fac_build(int nfree, const int *__restrict__ node_blk, const int *__restrict__ bcol,
                            const double *__restrict__ val, FacNodes nd, const int *__restrict__ lo_ptr,
                            const int *__restrict__ up_ptr, unsigned short *__restrict__ lo,
                            unsigned short *__restrict__ up)
{
    const int i = blockIdx.x * LN_WG + threadIdx.x;
    if (i >= nfree) return;
    const int nb0 = node_blk[i], deg = node_blk[i + 1] - nb0;
    const int l0 = lo_ptr[i], nl = lo_ptr[i + 1] - l0, u0 = up_ptr[i], nu = up_ptr[i + 1] - u0;
    unsigned short *lrow = lo + 9 * (size_t)l0, *urow = up + 9 * (size_t)u0;
    auto put = [](unsigned short *row, int nb, int kk, const double G[9]) {  // FP64 -> FP32 -> BF16
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) row[3 * nb * c + 3 * kk + d] = (unsigned short)to_bf16((float)G[3 * c + d]);
    };
    double Li[9], Bi[9];
    nd.linv(i, Li);
    nd.binv(i, Bi);
    int kl = 0, ku = 0;
    for (int k = 0; k < deg; ++k) {
        const int s = bcol[nb0 + k];
        double G[9];
        if (s == i) {
            put(lrow, nl, kl++, Li);
#pragma unroll
            for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
                for (int b2 = 0; b2 < 3; ++b2) G[3 * a2 + b2] = Li[3 * b2 + a2];
            put(urow, nu, ku++, G);
            continue;
        }
        double Kb[9], T[9], X[9];
        k_block(val, nb0, deg, k, Kb);
        if (s < i) {
            double Bs[9];
            nd.binv(s, Bs);
            mm3(Li, Kb, T);
            mm3(T, Bs, X);
            const double a = nd.alpha(i, s);
#pragma unroll
            for (int e = 0; e < 9; ++e) G[e] = -a * X[e];
            put(lrow, nl, kl++, G);
        } else {
            double Ls[9], LsT[9];
            nd.linv(s, Ls);
#pragma unroll
            for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
                for (int b2 = 0; b2 < 3; ++b2) LsT[3 * a2 + b2] = Ls[3 * b2 + a2];
            mm3(Bi, Kb, T);
            mm3(T, LsT, X);
            const double a = nd.alpha(s, i);
#pragma unroll
            for (int e = 0; e < 9; ++e) G[e] = -a * X[e];
            put(urow, nu, ku++, G);
        }
    }
}

// Loss rows, one wave per free node i of K's rows (sorted, so the blocks
// s <= i of row i are its first np slots, the diagonal last): lane k < np
// forms G_ik (and C_ik) of slot k into LDS, then W_k = sum_m K_{s_k s_m}
// G_{i s_m}^T (K_{s_k s_m} found by binary search in row s_k; zero when the
// two are not neighbours), T_i = sum_k tr(G_ik W_k) and
// dA[lo_ptr[i] + k] = -2 tr(C_ik W_k) / n for s_k < i.  Sums over the lanes
// in a fixed order.
constexpr int KAP_MAXP = 64;  // lower blocks per row (checked on the host)
__global__ __launch_bounds__(LN_WG) void k_kap_rows(int nfree, const int *__restrict__ node_blk,
                                                   const int *__restrict__ bcol, const double *__restrict__ val,
                                                   FacNodes nd, const int *__restrict__ lo_ptr, double inv_n,
                                                   double *__restrict__ Trow, double *__restrict__ dA)
{
    __shared__ double sG[LN_WG / 64][KAP_MAXP][9];
    __shared__ int sS[LN_WG / 64][KAP_MAXP];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * (LN_WG / 64) + wv;  // one row per wave, no loop: every wave reaches the barrier
    const bool act = i < nfree;
    int np = 0, nb0 = 0, deg = 0;
    double C[9], G[9];
    int s = -1;
    if (act) {
        np = lo_ptr[i + 1] - lo_ptr[i];
        nb0 = node_blk[i];
        deg = node_blk[i + 1] - nb0;
        if (lane < np) {
            s = bcol[nb0 + lane];
            double Li[9];
            nd.linv(i, Li);
            if (s == i) {
#pragma unroll
                for (int e = 0; e < 9; ++e) G[e] = Li[e];
            } else {
                double Kb[9], T[9], Bs[9];
                k_block(val, nb0, deg, lane, Kb);
                nd.binv(s, Bs);
                mm3(Li, Kb, T);
                mm3(T, Bs, C);
                const double a = nd.alpha(i, s);
#pragma unroll
                for (int e = 0; e < 9; ++e) G[e] = -a * C[e];
            }
#pragma unroll
            for (int e = 0; e < 9; ++e) sG[wv][lane][e] = G[e];
            sS[wv][lane] = s;
        }
    }
    __syncthreads();
    double t = 0.0;
    if (act && lane < np) {
        const int b0 = node_blk[s], dg = node_blk[s + 1] - b0;
        double W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int m = 0; m < np; ++m) {
            const int tm = sS[wv][m];
            int lo = 0, hi = dg - 1, q = -1;  // slot of column tm in row s
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1, cm = bcol[b0 + mid];
                if (cm == tm) {
                    q = mid;
                    break;
                }
                if (cm < tm) lo = mid + 1;
                else hi = mid - 1;
            }
            if (q < 0) continue;
            double Kst[9];
            k_block(val, b0, dg, q, Kst);
            const double *Gm = sG[wv][m];
#pragma unroll
            for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
                for (int b2 = 0; b2 < 3; ++b2)  // W[a][b] += sum_c K_st[a][c] G_it[b][c]
                    W[3 * a2 + b2] += Kst[3 * a2] * Gm[3 * b2] + Kst[3 * a2 + 1] * Gm[3 * b2 + 1] +
                                      Kst[3 * a2 + 2] * Gm[3 * b2 + 2];
        }
        double gw = 0.0, cw = 0.0;
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
            for (int b2 = 0; b2 < 3; ++b2) {
                gw += G[3 * a2 + b2] * W[3 * b2 + a2];
                cw += (s == i ? 0.0 : C[3 * a2 + b2]) * W[3 * b2 + a2];
            }
        t = gw;
        dA[lo_ptr[i] + lane] = s == i ? 0.0 : -2.0 * cw * inv_n;
    }
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if (act && lane == 0) Trow[i] = t;
}

// dO = dloss/d(u, v): u_r collects its lower pairs (r, s), v_r the pairs
// (i, r) of its upper neighbours i > r (their lower slot for r found by
// binary search in row i), each times dalpha/dx = alpha (1 - alpha / 2).
// Thread per node, fixed order.
__global__ void k_kap_dO(int nfree, const int *__restrict__ lo_ptr, const int *__restrict__ lo_col,
                         const int *__restrict__ up_ptr, const int *__restrict__ up_col, FacNodes nd,
                         const double *__restrict__ dA, float *__restrict__ dO)
{
    const int r = blockIdx.x * LN_WG + threadIdx.x;
    if (r >= nfree) return;
    double du = 0.0, dv = 0.0;
    for (int q = lo_ptr[r]; q < lo_ptr[r + 1]; ++q) {
        const int s = lo_col[q];
        if (s == r) continue;
        const double a = nd.alpha(r, s);
        du += dA[q] * a * (1.0 - 0.5 * a);
    }
    for (int q = up_ptr[r]; q < up_ptr[r + 1]; ++q) {
        const int i = up_col[q];
        if (i == r) continue;
        int lo = lo_ptr[i], hi = lo_ptr[i + 1] - 1, k = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1, cm = lo_col[mid];
            if (cm == r) {
                k = mid;
                break;
            }
            if (cm < r) lo = mid + 1;
            else hi = mid - 1;
        }
        if (k < 0) continue;  // the pattern is symmetric: not reached
        const double a = nd.alpha(i, r);
        dv += dA[k] * a * (1.0 - 0.5 * a);
    }
    dO[2 * r] = (float)du;
    dO[2 * r + 1] = (float)dv;
}

// partial sums of T over fixed chunks (one workgroup each)
__global__ void k_kap_sum(int nfree, const double *__restrict__ Trow, double *__restrict__ part)
{
    __shared__ double sm[LN_WG / 64];
    double a = 0.0;
    for (int i = blockIdx.x * LN_WG + threadIdx.x; i < nfree; i += gridDim.x * LN_WG) a += Trow[i];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < LN_WG / 64; ++w) t += sm[w];
        part[blockIdx.x] = t;
    }
}

// dH[r][i] = sum_o W[o][i] dZ[r][o]  (then optionally masked by relu'(Zmask))
__global__ void k_ln_back_in(int nfree, int Cout, int Cin, const float *__restrict__ dZ, const float *__restrict__ W,
                             const float *__restrict__ Zmask, float *__restrict__ dH)
{
    const long long idx = (long long)blockIdx.x * LN_WG + threadIdx.x;
    if (idx >= (long long)nfree * Cin) return;
    const int r = (int)(idx / Cin), i = (int)(idx % Cin);
    const float *dz = dZ + (long long)r * Cout;
    float acc = 0.0f;
    for (int o = 0; o < Cout; ++o) acc += W[o * Cin + i] * dz[o];
    if (Zmask && Zmask[idx] <= 0.0f) acc = 0.0f;
    dH[idx] = acc;
}

// a += b (elementwise), optional relu mask from Z
__global__ void k_ln_add_mask(long long n, float *__restrict__ a, const float *__restrict__ b,
                              const float *__restrict__ Zmask)
{
    const long long i = (long long)blockIdx.x * LN_WG + threadIdx.x;
    if (i >= n) return;
    float v = a[i] + b[i];
    if (Zmask && Zmask[i] <= 0.0f) v = 0.0f;
    a[i] = v;
}

// Weight-gradient partials: for a chunk of LN_WG nodes, P[o][i] = sum_r dZ[r][o] In[r][i]
// (and the bias sum in column Cin).  One workgroup per chunk, thread per (o, i) tile entry.
__global__ void k_ln_wgrad(int nfree, int Cout, int Cin, const float *__restrict__ dZ, const float *__restrict__ In,
                           float *__restrict__ part)
{
    const int chunk = blockIdx.x;
    const int r0 = chunk * LN_WG, r1 = min(r0 + LN_WG, nfree);
    const int ncol = Cin + 1;
    float *dst = part + (long long)chunk * Cout * ncol;
    for (int e = threadIdx.x; e < Cout * ncol; e += LN_WG) {
        const int o = e / ncol, i = e % ncol;
        float acc = 0.0f;
        for (int r = r0; r < r1; ++r) acc += dZ[(long long)r * Cout + o] * (i < Cin ? In[(long long)r * Cin + i] : 1.0f);
        dst[e] = acc;
    }
}

// sum of the chunk partials (fixed order) into the gradient slots
__global__ void k_ln_wgrad_sum(int nchunks, int Cout, int Cin, const float *__restrict__ part, float *__restrict__ gW,
                               float *__restrict__ gb)
{
    const int ncol = Cin + 1;
    const int e = blockIdx.x * LN_WG + threadIdx.x;
    if (e >= Cout * ncol) return;
    float acc = 0.0f;
    for (int c = 0; c < nchunks; ++c) acc += part[(long long)c * Cout * ncol + e];
    const int o = e / ncol, i = e % ncol;
    if (i < Cin) {
        if (gW) gW[o * Cin + i] = acc;
    } else if (gb) {
        gb[o] = acc;
    }
}

// ------------------------------------------------------------------ host
static unsigned grid_for(long long n) { return (unsigned)((n + LN_WG - 1) / LN_WG); }

static int ensure_lws(ofea_ctx *c)
{
    const int N = c->P.nfree;
    LearnedWS *w = lws(c);
    if (w && w->nfree == N) return OFEA_OK;
    learned_free(c);
    w = new LearnedWS();
    c->learned = w;
    w->nfree = N;
    w->params.assign(LN_NPARAM, 0.0f);
    const size_t Nb = (size_t)std::max(N, 1);
    OFEA_HIP(hipMalloc((void **)&w->d_params, sizeof(float) * LN_NPARAM));
    OFEA_HIP(hipMalloc((void **)&w->d_grad, sizeof(float) * LN_NPARAM));
    OFEA_HIP(hipMalloc((void **)&w->X, sizeof(float) * Nb * LN_F));
    OFEA_HIP(hipMalloc((void **)&w->A0, sizeof(float) * Nb * LN_F));
    for (float **p : {&w->Z1, &w->H1, &w->A1, &w->Z2, &w->H2, &w->dZ2, &w->dH1, &w->dZ1, &w->T1})
        OFEA_HIP(hipMalloc((void **)p, sizeof(float) * Nb * LN_H));
    OFEA_HIP(hipMalloc((void **)&w->O, sizeof(float) * Nb * LN_O));
    OFEA_HIP(hipMalloc((void **)&w->dO, sizeof(float) * Nb * LN_O));
    OFEA_HIP(hipMalloc((void **)&w->Rinv, sizeof(double) * Nb * 9));
    const int nchunks = (N + LN_WG - 1) / LN_WG;
    w->npartial_rows = nchunks;
    OFEA_HIP(hipMalloc((void **)&w->partial, sizeof(float) * (size_t)std::max(nchunks, 1) * LN_H * (LN_H + 1)));
    return OFEA_OK;
}

int solver_prepare(ofea_ctx *ctx);

static int ln_features(ofea_ctx *c)
{
    LearnedWS *w = lws(c);
    ofea_solver_ws *s = c->ws;
    const int N = w->nfree;
    double *logt = nullptr;
    OFEA_HIP(hipMalloc((void **)&logt, sizeof(double) * std::max(N, 1)));
    hipLaunchKernelGGL(k_ln_features, dim3(grid_for(N)), dim3(LN_WG), 0, c->stream, N, c->d_node_blk, c->d_bcol,
                       c->d_val, s->dblk, w->X, w->Rinv, logt);
    hipLaunchKernelGGL(k_ln_center, dim3(1), dim3(LN_WG), 0, c->stream, N, logt, w->X);
    hipLaunchKernelGGL(k_ln_agg, dim3(grid_for((long long)N * LN_F)), dim3(LN_WG), 0, c->stream, N, LN_F,
                       c->d_node_blk, c->d_bcol, w->X, w->A0, 0);
    hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipFree(logt);
    if (e != hipSuccess) return fail(OFEA_ERR_HIP, hipGetErrorString(e));
    return OFEA_OK;
}

// forward pass with the device parameters: the head outputs (u, v) per node.
// infer: the outputs serve the solver (BF16 layer 2 when set; the generation
// counter tells the solver to rebuild G); training always runs FP32.
static int ln_forward(ofea_ctx *c, bool infer)
{
    LearnedWS *w = lws(c);
    const int N = w->nfree;
    const float *P = w->d_params;
    hipLaunchKernelGGL(k_ln_dense, dim3(grid_for((long long)N * LN_H)), dim3(LN_WG), 0, c->stream, N, LN_F, LN_H, w->X,
                       w->A0, P + OFF_W1S, P + OFF_W1N, P + OFF_B1, w->Z1, w->H1);
    hipLaunchKernelGGL(k_ln_agg, dim3(grid_for((long long)N * LN_H)), dim3(LN_WG), 0, c->stream, N, LN_H,
                       c->d_node_blk, c->d_bcol, w->H1, w->A1, 0);
    if (infer && w->bf16)
        hipLaunchKernelGGL(k_ln_dense_mfma, dim3((N + 63) / 64), dim3(LN_MF_WG), 0, c->stream, N, w->H1, w->A1,
                           P + OFF_W2S, P + OFF_W2N, P + OFF_B2, w->Z2, w->H2);
    else
        hipLaunchKernelGGL(k_ln_dense, dim3(grid_for((long long)N * LN_H)), dim3(LN_WG), 0, c->stream, N, LN_H, LN_H,
                           w->H1, w->A1, P + OFF_W2S, P + OFF_W2N, P + OFF_B2, w->Z2, w->H2);
    hipLaunchKernelGGL(k_ln_head, dim3(grid_for(N)), dim3(LN_WG), 0, c->stream, N, w->H2, P + OFF_WO, P + OFF_BO, w->O);
    OFEA_HIP(hipGetLastError());
    if (infer) w->gen++;
    return OFEA_OK;
}

static FacNodes fac_nodes(ofea_ctx *c, const int *map)
{
    LearnedWS *w = lws(c);
    return FacNodes{w->Rinv, c->ws->minv, w->O, map};
}

// G into the half matrices lo / up for the context's K (view 0: rows of K;
// view 1: the reordered rows of the two-level methods, through perm)
static int build_factor(ofea_ctx *c, int view, unsigned short *lo, unsigned short *up)
{
    ofea_solver_ws *s = c->ws;
    const int N = lws(c)->nfree;
    int rc = fac_prepare(c, view);
    if (rc) return rc;
    const FacMat &f = s->fac[view];
    if (view == 1)
        hipLaunchKernelGGL(k_fac_build, dim3(grid_for(N)), dim3(LN_WG), 0, c->stream, N, s->nb_p, s->bcol_p, s->val_p,
                           fac_nodes(c, s->perm), f.lo_ptr, f.up_ptr, lo, up);
    else
        hipLaunchKernelGGL(k_fac_build, dim3(grid_for(N)), dim3(LN_WG), 0, c->stream, N, c->d_node_blk, c->d_bcol,
                           c->d_val, fac_nodes(c, nullptr), f.lo_ptr, f.up_ptr, lo, up);
    OFEA_HIP(hipGetLastError());
    return OFEA_OK;
}

static int ln_prepare(ofea_ctx *c)
{
    if (!c->assembled) return fail(OFEA_ERR_STATE, "ofea_assemble first");
    int rc = solver_prepare(c);
    if (rc) return rc;
    if ((rc = ensure_lws(c))) return rc;
    LearnedWS *w = lws(c);
    if (w->feat_kver != c->kver) {
        if ((rc = ln_features(c))) return rc;
        w->feat_kver = c->kver;
    }
    return OFEA_OK;
}

static void wgrad(ofea_ctx *c, int Cout, int Cin, const float *dZ, const float *In, float *gW, float *gb)
{
    LearnedWS *w = lws(c);
    const int N = w->nfree, nchunks = w->npartial_rows;
    hipLaunchKernelGGL(k_ln_wgrad, dim3(nchunks), dim3(LN_WG), 0, c->stream, N, Cout, Cin, dZ, In, w->partial);
    hipLaunchKernelGGL(k_ln_wgrad_sum, dim3(grid_for((long long)Cout * (Cin + 1))), dim3(LN_WG), 0, c->stream, nchunks,
                       Cout, Cin, w->partial, gW, gb);
}

// Buffers of the loss for the resident pattern (view 0 of the factor's index
// arrays): rows with more than KAP_MAXP lower blocks, or K rows that are not
// sorted, are refused.
static int kap_prepare(ofea_ctx *c)
{
    LearnedWS *w = lws(c);
    if (w->Trow) return OFEA_OK;
    int rc = fac_prepare(c, 0);
    if (rc) return rc;
    const int N = w->nfree;
    const std::vector<int> &nb = c->P.node_blk, &bc = c->P.bcol;
    for (int i = 0; i < N; ++i) {
        int low = 0;
        for (int q = nb[i]; q < nb[i + 1]; ++q) {
            if (q > nb[i] && bc[q] <= bc[q - 1]) return fail(OFEA_ERR_STATE, "learned loss: unsorted pattern row");
            low += bc[q] <= i;
        }
        if (low > KAP_MAXP) return fail(OFEA_ERR_LIMIT, "learned loss: a node has more than 64 lower neighbours");
    }
    const FacMat &f = c->ws->fac[0];
    OFEA_HIP(hipMalloc((void **)&w->Trow, sizeof(double) * std::max(N, 1)));
    OFEA_HIP(hipMalloc((void **)&w->dA, sizeof(double) * (size_t)std::max(f.nlo, 1LL)));
    OFEA_HIP(hipMalloc((void **)&w->tpart, sizeof(double) * 64));
    return OFEA_OK;
}

// loss tr(G K G^T) / n and its gradient (into d_grad) at the device parameters
static int ln_loss_grad(ofea_ctx *c, double *loss)
{
    LearnedWS *w = lws(c);
    const int N = w->nfree;
    int rc = kap_prepare(c);
    if (rc) return rc;
    if ((rc = ln_forward(c, false))) return rc;
    const FacMat &f = c->ws->fac[0];
    const double inv_n = 1.0 / (3.0 * std::max(N, 1));
    const FacNodes nd = fac_nodes(c, nullptr);
    hipLaunchKernelGGL(k_kap_rows, dim3((N + LN_WG / 64 - 1) / (LN_WG / 64)), dim3(LN_WG), 0, c->stream, N,
                       c->d_node_blk, c->d_bcol, c->d_val, nd, f.lo_ptr, inv_n, w->Trow, w->dA);
    const int nbs = std::max(1, std::min(64, (int)grid_for(N)));
    hipLaunchKernelGGL(k_kap_sum, dim3(nbs), dim3(LN_WG), 0, c->stream, N, w->Trow, w->tpart);
    hipLaunchKernelGGL(k_kap_dO, dim3(grid_for(N)), dim3(LN_WG), 0, c->stream, N, f.lo_ptr, f.lo_col, f.up_ptr,
                       f.up_col, nd, w->dA, w->dO);
    double tp[64];
    OFEA_HIP(hipMemcpyAsync(tp, w->tpart, sizeof(double) * nbs, hipMemcpyDeviceToHost, c->stream));
    float *g = w->d_grad;
    const float *P = w->d_params;
    // head: dWo, dbo; dH2 -> dZ2 (relu mask)
    wgrad(c, LN_O, LN_H, w->dO, w->H2, g + OFF_WO, g + OFF_BO);
    hipLaunchKernelGGL(k_ln_back_in, dim3(grid_for((long long)N * LN_H)), dim3(LN_WG), 0, c->stream, N, LN_O, LN_H,
                       w->dO, P + OFF_WO, w->Z2, w->dZ2);
    // layer 2: dW2s, dW2n, db2; dH1 = W2s^T dZ2 + agg^T(W2n^T dZ2), masked by relu'(Z1)
    wgrad(c, LN_H, LN_H, w->dZ2, w->H1, g + OFF_W2S, g + OFF_B2);
    wgrad(c, LN_H, LN_H, w->dZ2, w->A1, g + OFF_W2N, nullptr);
    hipLaunchKernelGGL(k_ln_back_in, dim3(grid_for((long long)N * LN_H)), dim3(LN_WG), 0, c->stream, N, LN_H, LN_H,
                       w->dZ2, P + OFF_W2N, (const float *)nullptr, w->T1);
    hipLaunchKernelGGL(k_ln_agg, dim3(grid_for((long long)N * LN_H)), dim3(LN_WG), 0, c->stream, N, LN_H,
                       c->d_node_blk, c->d_bcol, w->T1, w->dZ1, 1);
    hipLaunchKernelGGL(k_ln_back_in, dim3(grid_for((long long)N * LN_H)), dim3(LN_WG), 0, c->stream, N, LN_H, LN_H,
                       w->dZ2, P + OFF_W2S, (const float *)nullptr, w->dH1);
    hipLaunchKernelGGL(k_ln_add_mask, dim3(grid_for((long long)N * LN_H)), dim3(LN_WG), 0, c->stream,
                       (long long)N * LN_H, w->dH1, w->dZ1, w->Z1);
    // layer 1: dW1s, dW1n, db1 (dZ1 now in dH1)
    wgrad(c, LN_H, LN_F, w->dH1, w->X, g + OFF_W1S, g + OFF_B1);
    wgrad(c, LN_H, LN_F, w->dH1, w->A0, g + OFF_W1N, nullptr);
    OFEA_HIP(hipGetLastError());
    OFEA_HIP(hipStreamSynchronize(c->stream));
    double L = 0.0;
    for (int k = 0; k < nbs; ++k) L += tp[k];
    if (!std::isfinite(L)) return fail(OFEA_ERR_STATE, "learned loss is not finite");
    if (loss) *loss = L * inv_n;
    return OFEA_OK;
}

static int upload_params(ofea_ctx *c)
{
    LearnedWS *w = lws(c);
    OFEA_HIP(hipMemcpyAsync(w->d_params, w->params.data(), sizeof(float) * LN_NPARAM, hipMemcpyHostToDevice, c->stream));
    return OFEA_OK;
}

// deterministic initial parameters: scaled uniform from a 64-bit LCG; head zero
static void init_params(std::vector<float> &p, unsigned long long seed)
{
    unsigned long long s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto uni = [&]() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)((s >> 11) & ((1ULL << 53) - 1)) / (double)(1ULL << 53) * 2.0 - 1.0;
    };
    std::fill(p.begin(), p.end(), 0.0f);
    const double a1 = std::sqrt(3.0 / (2 * LN_F)), a2 = std::sqrt(3.0 / (2 * LN_H));
    for (int i = 0; i < 2 * LN_H * LN_F; ++i) p[OFF_W1S + i] = (float)(a1 * uni());
    for (int i = 0; i < 2 * LN_H * LN_H; ++i) p[OFF_W2S + i] = (float)(a2 * uni());
    // head zero: alpha = 2 sigmoid(0) = 1 at every pair, the first-order rule
    p[OFF_BO] = LN_HEAD_BIAS0;
}

int learned_init(ofea_ctx *c, unsigned long long seed)
{
    int rc = ln_prepare(c);
    if (rc) return rc;
    init_params(lws(c)->params, seed);
    if ((rc = upload_params(c))) return rc;
    return ln_forward(c, true);
}

int learned_set_params(ofea_ctx *c, const float *params, int nparam)
{
    if (nparam != LN_NPARAM) return fail(OFEA_ERR_ARG, "parameter count mismatch");
    int rc = ln_prepare(c);
    if (rc) return rc;
    std::memcpy(lws(c)->params.data(), params, sizeof(float) * LN_NPARAM);
    if ((rc = upload_params(c))) return rc;
    return ln_forward(c, true);
}

int learned_get_params(ofea_ctx *c, float *params, int nparam)
{
    if (nparam != LN_NPARAM) return fail(OFEA_ERR_ARG, "parameter count mismatch");
    if (!lws(c)) return fail(OFEA_ERR_STATE, "learned preconditioner not initialised");
    std::memcpy(params, lws(c)->params.data(), sizeof(float) * LN_NPARAM);
    return OFEA_OK;
}

int learned_nparam() { return LN_NPARAM; }

int learned_loss_grad(ofea_ctx *c, double *loss, float *grad)
{
    LearnedWS *w = lws(c);
    if (!w) return fail(OFEA_ERR_STATE, "learned preconditioner not initialised");
    int rc = ln_loss_grad(c, loss);
    if (rc) return rc;
    if (grad) {
        OFEA_HIP(hipMemcpyAsync(grad, w->d_grad, sizeof(float) * LN_NPARAM, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
    }
    return OFEA_OK;
}

// Adam on the host over the tiny parameter vector, on the mean loss of the
// contexts' resident systems; gradients from each context's device.  The
// parameters start from the first context's and end in every context.
int learned_train(ofea_ctx *const *cs, int nctx, int steps, float lr, double *loss_hist)
{
    for (int k = 0; k < nctx; ++k)
        if (!lws(cs[k])) return fail(OFEA_ERR_STATE, "learned preconditioner not initialised (ofea_learned_init)");
    std::vector<float> p = lws(cs[0])->params, g(LN_NPARAM), gs(LN_NPARAM), m(LN_NPARAM, 0.0f), v(LN_NPARAM, 0.0f);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    auto put = [&](ofea_ctx *c, bool infer) -> int {
        std::memcpy(lws(c)->params.data(), p.data(), sizeof(float) * LN_NPARAM);
        int rc = upload_params(c);
        if (!rc && infer) rc = ln_forward(c, true);
        return rc;
    };
    for (int t = 1; t <= steps; ++t) {
        double loss = 0.0;
        std::fill(gs.begin(), gs.end(), 0.0f);
        for (int k = 0; k < nctx; ++k) {
            double lk = 0.0;
            int rc = put(cs[k], false);
            if (!rc) rc = learned_loss_grad(cs[k], &lk, g.data());
            if (rc) return rc;
            loss += lk / nctx;
            for (int i = 0; i < LN_NPARAM; ++i) gs[i] += g[i] / nctx;
        }
        if (loss_hist) loss_hist[t - 1] = loss;
        const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
        for (int i = 0; i < LN_NPARAM; ++i) {
            if (!std::isfinite(gs[i])) return fail(OFEA_ERR_STATE, "non-finite gradient during training");
            m[i] = b1 * m[i] + (1.0f - b1) * gs[i];
            v[i] = b2 * v[i] + (1.0f - b2) * gs[i] * gs[i];
            p[i] -= lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
        }
    }
    for (int k = 0; k < nctx; ++k) {
        int rc = put(cs[k], true);
        if (rc) return rc;
        OFEA_HIP(hipStreamSynchronize(cs[k]->stream));
    }
    return OFEA_OK;
}

// the factor G on K's node-block layout (nnz values, FP64 copies of the BF16
// entries the solver applies; per block slot k of row i a row-major 3x3:
// G_is for s < i, G_ii, (G_si)^T for s > i) and the node features
int learned_blocks(ofea_ctx *c, double *blocks, float *features)
{
    LearnedWS *w = lws(c);
    if (!w || !c->ws) return fail(OFEA_ERR_STATE, "learned preconditioner not initialised");
    const size_t N = w->nfree;
    if (blocks) {
        int rc = learned_factor(c, false);
        if (rc) return rc;
        const FacMat &f = c->ws->fac[0];
        std::vector<unsigned short> lo(9 * (size_t)f.nlo), up(9 * (size_t)f.nup);
        std::vector<int> lp(N + 1), upp(N + 1);
        OFEA_HIP(hipMemcpyAsync(lo.data(), f.lo, sizeof(unsigned short) * lo.size(), hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipMemcpyAsync(up.data(), f.up, sizeof(unsigned short) * up.size(), hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipMemcpyAsync(lp.data(), f.lo_ptr, sizeof(int) * (N + 1), hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipMemcpyAsync(upp.data(), f.up_ptr, sizeof(int) * (N + 1), hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
        const std::vector<int> &nb = c->P.node_blk, &bc = c->P.bcol;
        for (size_t i = 0; i < N; ++i) {
            const int nl = lp[i + 1] - lp[i], nu = upp[i + 1] - upp[i];
            int kl = 0, ku = 0;
            for (int q = nb[i]; q < nb[i + 1]; ++q) {
                const size_t s = (size_t)bc[q];
                // the s <= i half holds G_is (G_ii on the diagonal), the s > i half (G_si)^T
                const bool low = s <= i;
                const unsigned short *row = low ? &lo[9 * (size_t)lp[i]] : &up[9 * (size_t)upp[i]];
                const int nbk = low ? nl : nu, kk = low ? kl++ : ku;
                if (s >= i) ++ku;  // the diagonal also has a slot in the upper half
                for (int cc = 0; cc < 3; ++cc)
                    for (int d = 0; d < 3; ++d) {
                        const unsigned bits = (unsigned)row[3 * nbk * cc + 3 * kk + d] << 16;
                        float fv;
                        std::memcpy(&fv, &bits, sizeof fv);
                        blocks[9 * (size_t)q + 3 * cc + d] = fv;
                    }
            }
        }
    }
    if (features) {
        OFEA_HIP(hipMemcpyAsync(features, w->X, sizeof(float) * LN_F * N, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
    }
    return OFEA_OK;
}

// the network's outputs (u, v) per free node, as they serve the solver
int learned_outputs(ofea_ctx *c, float *uv)
{
    LearnedWS *w = lws(c);
    if (!w || w->gen == 0) return fail(OFEA_ERR_STATE, "learned preconditioner not set (ofea_learned_*)");
    OFEA_HIP(hipMemcpyAsync(uv, w->O, sizeof(float) * LN_O * (size_t)w->nfree, hipMemcpyDeviceToHost, c->stream));
    OFEA_HIP(hipStreamSynchronize(c->stream));
    return OFEA_OK;
}

// the solver's G for the resident K (rebuilt when K or the outputs changed)
int learned_factor(ofea_ctx *c, bool permuted)
{
    LearnedWS *w = lws(c);
    if (!w || w->gen == 0) return fail(OFEA_ERR_STATE, "learned preconditioner not set (ofea_learned_*)");
    const int view = permuted ? 1 : 0;
    int rc = fac_prepare(c, view);
    if (rc) return rc;
    FacMat &f = c->ws->fac[view];
    if (f.kver == c->kver && f.gen == w->gen) return OFEA_OK;
    if ((rc = build_factor(c, view, f.lo, f.up))) return rc;
    f.kver = c->kver;
    f.gen = w->gen;
    return fac_tune(c, view);
}

int learned_precision(ofea_ctx *c, int precision)
{
    if (precision != OFEA_FP32 && precision != OFEA_BF16)
        return fail(OFEA_ERR_ARG, "inference precision must be OFEA_FP32 or OFEA_BF16");
    LearnedWS *w = lws(c);
    if (!w) return fail(OFEA_ERR_STATE, "learned preconditioner not initialised");
    w->bf16 = precision == OFEA_BF16;
    int rc = ln_prepare(c);
    if (rc) return rc;
    if ((rc = upload_params(c))) return rc;
    return ln_forward(c, true);
}

// re-run the network on the resident K (after a new assembly)
int learned_refresh(ofea_ctx *c)
{
    if (!lws(c)) return fail(OFEA_ERR_STATE, "learned preconditioner not initialised");
    int rc = ln_prepare(c);
    if (rc) return rc;
    if ((rc = upload_params(c))) return rc;
    return ln_forward(c, true);
}

}  // namespace ofea
