// learned.hip -- learned correction of the block-Jacobi preconditioner.
//
// No counterpart in the reference (its solvers are direct, src/linear_solution.cpp).
// A 2-layer node-graph network over the free-node adjacency of K maps local
// features of each node to a lower-triangular L_r; the preconditioner block
// is M_r = R_r^-1 (L_r L_r^T) R_r^-T with D_r = R_r^T R_r the node's 3x3
// diagonal block of K.  M_r is SPD for any parameters, and L_r = I gives
// block Jacobi exactly (the output head starts at zero).  PCG stays a valid
// CG with a fixed SPD preconditioner; plain block-Jacobi PCG is the fallback.
//
//   features x_r (9): D_r / t_r (6 entries), log t_r - mean log t, deg_r / 27,
//                     sum_{s!=r} ||K_rs||_F / ((deg_r - 1) ||D_r||_F);  t_r = tr(D_r)/3
//   layer l:  h' = relu(W_self h_r + W_nbr mean_{s in N(r), s != r} h_s + b)   (9 -> 64 -> 64)
//   head:     o_r = W_o h2_r + b_o (6);  L = [[e^o0,0,0],[o3,e^o1,0],[o5,o4,e^o2]]
//
// Training loss over S pairs (e_k, r_k = K e_k) from the solver's own runs:
//   loss = (1/S) sum_k (e_k - M r_k)^T K (e_k - M r_k) / (e_k^T r_k)
// i.e. the energy-norm error left by one preconditioned Richardson step.
// Its gradient needs one SpMV per sample; everything else is node-local or a
// neighbour gather.  All training arithmetic is FP32 (M and the SpMV in FP64).
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
constexpr int LN_O = 6;     // head outputs
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
    float *O = nullptr;                             // N x 6
    double *Rinv = nullptr;                         // N x 9 (upper-triangular R^-1, row-major)
    float *dO = nullptr;                            // N x 6
    float *dZ2 = nullptr, *dH1 = nullptr, *dZ1 = nullptr, *T1 = nullptr;  // N x 64
    double *G = nullptr;                            // N x 9 gradient of the loss wrt M_r
    float *partial = nullptr;                       // weight-gradient partials
    int npartial_rows = 0;
    // samples
    int S = 0;
    double *E = nullptr, *R = nullptr, *Dv = nullptr, *Q = nullptr;  // n x S each
    double *wsmp = nullptr;                         // S: 1/(S e_k^T r_k)
    double *loss_part = nullptr;
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
                    (void *)w->dZ2, (void *)w->dH1, (void *)w->dZ1, (void *)w->T1, (void *)w->G, (void *)w->partial,
                    (void *)w->E, (void *)w->R, (void *)w->Dv, (void *)w->Q, (void *)w->wsmp, (void *)w->loss_part})
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

// head: o = Wo h2 + bo; M = A L L^T A^T (FP64).  Thread per node.
__device__ inline void ln_Lmat(const float o[6], double L[9])
{
    L[0] = exp((double)o[0]);
    L[1] = 0.0;
    L[2] = 0.0;
    L[3] = o[3];
    L[4] = exp((double)o[1]);
    L[5] = 0.0;
    L[6] = o[5];
    L[7] = o[4];
    L[8] = exp((double)o[2]);
}

__global__ void k_ln_head(int nfree, const float *__restrict__ H2, const float *__restrict__ Wo,
                          const float *__restrict__ bo, const double *__restrict__ Rinv, float *__restrict__ O,
                          double *__restrict__ M)
{
    const int r = blockIdx.x * LN_WG + threadIdx.x;
    if (r >= nfree) return;
    float o[6];
    const float *h = H2 + (long long)r * LN_H;
    for (int j = 0; j < LN_O; ++j) {
        float acc = bo[j];
        for (int i = 0; i < LN_H; ++i) acc += Wo[j * LN_H + i] * h[i];
        o[j] = acc;
        O[LN_O * r + j] = acc;
    }
    double L[9], B[9];
    ln_Lmat(o, L);
    const double *A = Rinv + 9 * r;
    // B = A L ; M = B B^T
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[3 * i + j] = A[3 * i] * L[j] + A[3 * i + 1] * L[3 + j] + A[3 * i + 2] * L[6 + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            M[9 * r + 3 * i + j] = B[3 * i] * B[3 * j] + B[3 * i + 1] * B[3 * j + 1] + B[3 * i + 2] * B[3 * j + 2];
}

// d_k = e_k - M r_k (node-local), all samples.  Thread per (node, sample).
__global__ void k_ln_residual(int nfree, int S, const double *__restrict__ M, const double *__restrict__ E,
                              const double *__restrict__ R, double *__restrict__ Dv)
{
    const long long idx = (long long)blockIdx.x * LN_WG + threadIdx.x;
    if (idx >= (long long)nfree * S) return;
    const int k = (int)(idx / nfree), r = (int)(idx % nfree);
    const long long base = (long long)k * 3 * nfree + 3 * r;
    const double *m = M + 9 * r;
    for (int c = 0; c < 3; ++c)
        Dv[base + c] = E[base + c] - (m[3 * c] * R[base] + m[3 * c + 1] * R[base + 1] + m[3 * c + 2] * R[base + 2]);
}

// loss partials: sum_k w_k d_k . q_k ; and the gradient wrt M_r:
// G_r = -2 sum_k w_k q_{k,r} r_{k,r}^T.  Thread per node.
__global__ void k_ln_lossgrad(int nfree, int S, const double *__restrict__ Dv, const double *__restrict__ Q,
                              const double *__restrict__ R, const double *__restrict__ wsmp, double *__restrict__ G,
                              double *__restrict__ loss_part)
{
    __shared__ double sm[LN_WG / 64];
    const int r = blockIdx.x * LN_WG + threadIdx.x;
    double l = 0.0;
    if (r < nfree) {
        double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < S; ++k) {
            const long long base = (long long)k * 3 * nfree + 3 * r;
            const double w = wsmp[k];
            l += w * (Dv[base] * Q[base] + Dv[base + 1] * Q[base + 1] + Dv[base + 2] * Q[base + 2]);
            for (int c = 0; c < 3; ++c)
                for (int d = 0; d < 3; ++d) g[3 * c + d] -= 2.0 * w * Q[base + c] * R[base + d];
        }
        for (int i = 0; i < 9; ++i) G[9 * r + i] = g[i];
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < LN_WG / 64; ++k) t += sm[k];
        loss_part[blockIdx.x] = t;
    }
}

// dO from G through M = A L L^T A^T and the exp diagonal.  Thread per node.
__global__ void k_ln_head_back(int nfree, const double *__restrict__ G, const double *__restrict__ Rinv,
                               const float *__restrict__ O, float *__restrict__ dO)
{
    const int r = blockIdx.x * LN_WG + threadIdx.x;
    if (r >= nfree) return;
    const double *A = Rinv + 9 * r, *Gr = G + 9 * r;
    float o[6];
    for (int j = 0; j < 6; ++j) o[j] = O[LN_O * r + j];
    double L[9], H[9], T[9], dL[9];
    ln_Lmat(o, L);
    // H = A^T G A
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int a = 0; a < 3; ++a) acc += A[3 * a + i] * Gr[3 * a + j];
            T[3 * i + j] = acc;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int a = 0; a < 3; ++a) acc += T[3 * i + a] * A[3 * a + j];
            H[3 * i + j] = acc;
        }
    // dL = (H + H^T) L
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int a = 0; a < 3; ++a) acc += (H[3 * i + a] + H[3 * a + i]) * L[3 * a + j];
            dL[3 * i + j] = acc;
        }
    float *d = dO + LN_O * r;
    d[0] = (float)(dL[0] * L[0]);
    d[1] = (float)(dL[4] * L[4]);
    d[2] = (float)(dL[8] * L[8]);
    d[3] = (float)dL[3];
    d[4] = (float)dL[7];
    d[5] = (float)dL[6];
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
    OFEA_HIP(hipMalloc((void **)&w->G, sizeof(double) * Nb * 9));
    const int nchunks = (N + LN_WG - 1) / LN_WG;
    w->npartial_rows = nchunks;
    OFEA_HIP(hipMalloc((void **)&w->partial, sizeof(float) * (size_t)std::max(nchunks, 1) * LN_H * (LN_H + 1)));
    OFEA_HIP(hipMalloc((void **)&w->loss_part, sizeof(double) * (size_t)std::max(nchunks, 1)));
    return OFEA_OK;
}

int solver_prepare(ofea_ctx *ctx);
int solver_spmv(ofea_ctx *ctx, const double *d_x, double *d_y);

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

// forward pass with the device parameters; M blocks (FP64) into learned64.
// infer: the network produces the solver's blocks (BF16 layer 2 when set);
// training always runs FP32.
static int ln_forward(ofea_ctx *c, bool infer)
{
    LearnedWS *w = lws(c);
    ofea_solver_ws *s = c->ws;
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
    if (!s->learned64) OFEA_HIP(hipMalloc((void **)&s->learned64, sizeof(double) * 9 * (size_t)std::max(N, 1)));
    hipLaunchKernelGGL(k_ln_head, dim3(grid_for(N)), dim3(LN_WG), 0, c->stream, N, w->H2, P + OFF_WO, P + OFF_BO,
                       w->Rinv, w->O, s->learned64);
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

// loss and gradient at the device parameters for the resident samples
static int ln_loss_grad(ofea_ctx *c, double *loss)
{
    LearnedWS *w = lws(c);
    const int N = w->nfree, S = w->S, n = 3 * N;
    int rc = ln_forward(c, false);
    if (rc) return rc;
    ofea_solver_ws *s = c->ws;
    hipLaunchKernelGGL(k_ln_residual, dim3(grid_for((long long)N * S)), dim3(LN_WG), 0, c->stream, N, S, s->learned64,
                       w->E, w->R, w->Dv);
    for (int k = 0; k < S; ++k)
        if ((rc = solver_spmv(c, w->Dv + (size_t)k * n, w->Q + (size_t)k * n))) return rc;
    hipLaunchKernelGGL(k_ln_lossgrad, dim3(grid_for(N)), dim3(LN_WG), 0, c->stream, N, S, w->Dv, w->Q, w->R, w->wsmp,
                       w->G, w->loss_part);
    hipLaunchKernelGGL(k_ln_head_back, dim3(grid_for(N)), dim3(LN_WG), 0, c->stream, N, w->G, w->Rinv, w->O, w->dO);
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
    if (loss) {
        std::vector<double> lp(std::max(w->npartial_rows, 1));
        OFEA_HIP(hipMemcpyAsync(lp.data(), w->loss_part, sizeof(double) * w->npartial_rows, hipMemcpyDeviceToHost,
                                c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
        double t = 0.0;
        for (double v : lp) t += v;
        *loss = t;
    }
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
    // head (OFF_WO .. end) stays zero: L = I, block Jacobi
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

// samples: E, R host arrays (S x n, sample-major); R = K E is not recomputed
int learned_set_samples(ofea_ctx *c, const double *E, const double *R, int S)
{
    int rc = ln_prepare(c);
    if (rc) return rc;
    LearnedWS *w = lws(c);
    const size_t n = 3 * (size_t)w->nfree;
    // the buffers below are freed and refilled by blocking copies on the null
    // stream: kernels of the context's stream must not still read them
    OFEA_HIP(hipStreamSynchronize(c->stream));
    for (double **p : {&w->E, &w->R, &w->Dv, &w->Q, &w->wsmp})
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    w->S = S;
    if (S <= 0) return fail(OFEA_ERR_ARG, "need at least one sample");
    OFEA_HIP(hipMalloc((void **)&w->E, sizeof(double) * n * S));
    OFEA_HIP(hipMalloc((void **)&w->R, sizeof(double) * n * S));
    OFEA_HIP(hipMalloc((void **)&w->Dv, sizeof(double) * n * S));
    OFEA_HIP(hipMalloc((void **)&w->Q, sizeof(double) * n * S));
    OFEA_HIP(hipMalloc((void **)&w->wsmp, sizeof(double) * S));
    std::vector<double> wk(S);
    for (int k = 0; k < S; ++k) {
        double er = 0.0;
        for (size_t i = 0; i < n; ++i) er += E[k * n + i] * R[k * n + i];
        if (!(er > 0.0)) return fail(OFEA_ERR_ARG, "sample with e^T r <= 0 (r must be K e)");
        wk[k] = 1.0 / (S * er);
    }
    OFEA_HIP(hipMemcpy(w->E, E, sizeof(double) * n * S, hipMemcpyHostToDevice));
    OFEA_HIP(hipMemcpy(w->R, R, sizeof(double) * n * S, hipMemcpyHostToDevice));
    OFEA_HIP(hipMemcpy(w->wsmp, wk.data(), sizeof(double) * S, hipMemcpyHostToDevice));
    return OFEA_OK;
}

int learned_loss_grad(ofea_ctx *c, double *loss, float *grad)
{
    LearnedWS *w = lws(c);
    if (!w || w->S == 0) return fail(OFEA_ERR_STATE, "no training samples (ofea_learned_set_samples)");
    int rc = ln_loss_grad(c, loss);
    if (rc) return rc;
    if (grad) {
        OFEA_HIP(hipMemcpyAsync(grad, w->d_grad, sizeof(float) * LN_NPARAM, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
    }
    return OFEA_OK;
}

// Adam on the host over the tiny parameter vector; gradients from the device.
int learned_train(ofea_ctx *c, int steps, float lr, double *loss_hist)
{
    LearnedWS *w = lws(c);
    if (!w || w->S == 0) return fail(OFEA_ERR_STATE, "no training samples (ofea_learned_set_samples)");
    std::vector<float> g(LN_NPARAM), m(LN_NPARAM, 0.0f), v(LN_NPARAM, 0.0f);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int t = 1; t <= steps; ++t) {
        double loss = 0.0;
        int rc = learned_loss_grad(c, &loss, g.data());
        if (rc) return rc;
        if (loss_hist) loss_hist[t - 1] = loss;
        const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
        for (int i = 0; i < LN_NPARAM; ++i) {
            if (!std::isfinite(g[i])) return fail(OFEA_ERR_STATE, "non-finite gradient during training");
            m[i] = b1 * m[i] + (1.0f - b1) * g[i];
            v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
            w->params[i] -= lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
        }
        if ((rc = upload_params(c))) return rc;
    }
    int rc = ln_forward(c, true);
    if (rc) return rc;
    OFEA_HIP(hipStreamSynchronize(c->stream));
    return OFEA_OK;
}

// current preconditioner blocks (FP64, 9 per free node) and node features
int learned_blocks(ofea_ctx *c, double *blocks, float *features)
{
    LearnedWS *w = lws(c);
    if (!w || !c->ws || !c->ws->learned64) return fail(OFEA_ERR_STATE, "learned preconditioner not initialised");
    const size_t N = w->nfree;
    if (blocks)
        OFEA_HIP(hipMemcpyAsync(blocks, c->ws->learned64, sizeof(double) * 9 * N, hipMemcpyDeviceToHost, c->stream));
    if (features) OFEA_HIP(hipMemcpyAsync(features, w->X, sizeof(float) * LN_F * N, hipMemcpyDeviceToHost, c->stream));
    OFEA_HIP(hipStreamSynchronize(c->stream));
    return OFEA_OK;
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
