// assemble.hip -- FP64 stiffness assembly into the reference's CSC arrays.
//
// Replaces assamble_stiffness::fill_CSR_sparse_mat (src/assemble.cpp:95-151)
// with the element stiffness of build_ele_stiff_mat_gauss
// (src/Elements/C3D8_Gauss.cpp:80-108) and build_ele_stiff_mat_hammer
// (src/Elements/C3D4_hammer.cpp:61-73).
//
// Design (column-owner gather, no atomics):
//   * one 256-thread workgroup owns a chunk of consecutive free nodes, i.e.
//     the CSC columns 3r..3r+2 of each node r of the chunk;
//   * phase 1 computes, for every (owned node r, incident element e) pair,
//     the element data needed for row block a = local index of r:
//       C3D4: the constant gradients g_b (4x3) and lambda*V, mu*V;
//       C3D8: one lane per (pair, Gauss point) computes the point's gradients
//             and the 8 blocks K_ab at that point; an 8-lane xor-shuffle
//             reduce-scatter leaves lane (pair, b) holding the Gauss sum K_ab;
//   * phase 2 gives one lane per 3x3 block (r, k) of the chunk's columns; it
//     sums its contributors (precomputed list, ascending element order, the
//     reference's summation order) from LDS and writes the 9 CSC values once.
// For isotropic D = [lambda, mu] the block is
//   K_ab[c][d] = w (lambda g_a[c] g_b[d] + mu g_a[d] g_b[c] + mu delta_cd g_a.g_b),
// algebraically equal to B_a^T D B_b, and the CSC slot (row 3s+d, col 3r+c)
// receives K_ab[c][d] (src/assemble.cpp:128-146 adds Ke(mm, nn) at row dof[nn],
// column dof[mm]).
#include "elem_device.h"
#include "ofea_internal.h"

namespace ofea {

constexpr int ASM_WG = 256;
constexpr int TET_CAP = 256;  // pairs per tet chunk: one per thread
constexpr int HEX_CAP = 32;   // pairs per hex chunk: 8 Gauss-point lanes each

// Phase 2, shared by both element types: one lane per block (r, k).
template <class BlockSource>
__device__ inline void gather_blocks(int b0, int b1, const int *__restrict__ blk_rk, const int *__restrict__ contrib_ptr,
                                     const uint16_t *__restrict__ contrib, const int *__restrict__ col_ptr,
                                     double *__restrict__ val, int accumulate, const BlockSource &src)
{
    for (int bk = b0 + (int)threadIdx.x; bk < b1; bk += ASM_WG) {
        int rk = blk_rk[bk];
        int r = rk >> 8, k = rk & 255;
        double acc[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] = 0.0;
        int t1 = contrib_ptr[bk + 1];
        for (int t = contrib_ptr[bk]; t < t1; ++t) src.add(contrib[t], acc);
        int base = col_ptr[3 * r];
        int stride = col_ptr[3 * r + 1] - base;  // 3 * deg(r)
        double *dst = val + base + 3 * k;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double v = acc[3 * c + d];
                if (accumulate) v += dst[c * stride + d];
                dst[c * stride + d] = v;
            }
    }
}

// ------------------------------------------------------------------ C3D4
struct TetSource {
    const double (*g)[12];
    const double (*w)[2];
    const int *a;
    __device__ void add(int ci, double acc[9]) const
    {
        int lp = ci >> 2, b = ci & 3, aa = a[lp];
        double ga0 = g[lp][3 * aa], ga1 = g[lp][3 * aa + 1], ga2 = g[lp][3 * aa + 2];
        double gb0 = g[lp][3 * b], gb1 = g[lp][3 * b + 1], gb2 = g[lp][3 * b + 2];
        double wl = w[lp][0], wm = w[lp][1];
        double ga[3] = {ga0, ga1, ga2}, gb[3] = {gb0, gb1, gb2};
        double dm = wm * (ga0 * gb0 + ga1 * gb1 + ga2 * gb2);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) acc[3 * c + d] += wl * ga[c] * gb[d] + wm * ga[d] * gb[c] + (c == d ? dm : 0.0);
    }
};

__global__ __launch_bounds__(ASM_WG) void k_assemble_tet(const double *__restrict__ xyz, const int *__restrict__ conn,
                                                         const double *__restrict__ Eel, double E, double nu,
                                                         const int *__restrict__ chunk_pair,
                                                         const int *__restrict__ chunk_blk, const int *__restrict__ pair,
                                                         const int *__restrict__ blk_rk,
                                                         const int *__restrict__ contrib_ptr,
                                                         const uint16_t *__restrict__ contrib,
                                                         const int *__restrict__ col_ptr, double *__restrict__ val,
                                                         int accumulate, int nchunks)
{
    __shared__ double s_g[TET_CAP][12];
    __shared__ double s_w[TET_CAP][2];
    __shared__ int s_a[TET_CAP];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int p0 = chunk_pair[chunk], p1 = chunk_pair[chunk + 1];
    double lam0, mu0;
    lame<double>(E, nu, lam0, mu0);
    for (int p = p0 + (int)threadIdx.x; p < p1; p += ASM_WG) {
        int pk = pair[p];
        int el = pk >> 3, lp = p - p0;
        const int *cn = conn + 4 * el;
        double X[12];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            int node = cn[b];
            X[3 * b] = xyz[3 * node];
            X[3 * b + 1] = xyz[3 * node + 1];
            X[3 * b + 2] = xyz[3 * node + 2];
        }
        double g[3][4];
        double detJ = tet_grads(X, g);
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) s_g[lp][3 * b + c] = g[c][b];
        double lam = lam0, mu = mu0;
        if (Eel) lame<double>(Eel[el], nu, lam, mu);
        double V = detJ / 6.0;  // signed (src/Elements/C3D4.cpp:45-54 does not abs())
        s_w[lp][0] = lam * V;
        s_w[lp][1] = mu * V;
        s_a[lp] = pk & 7;
    }
    __syncthreads();
    TetSource src{s_g, s_w, s_a};
    gather_blocks(chunk_blk[chunk], chunk_blk[chunk + 1], blk_rk, contrib_ptr, contrib, col_ptr, val, accumulate, src);
}

// ------------------------------------------------------------------ C3D8
struct HexSource {
    const double (*blk)[9];
    __device__ void add(int ci, double acc[9]) const
    {
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += blk[ci][i];
    }
};

__global__ __launch_bounds__(ASM_WG) void k_assemble_hex(const double *__restrict__ xyz, const int *__restrict__ conn,
                                                         const double *__restrict__ Eel, double E, double nu,
                                                         const int *__restrict__ chunk_pair,
                                                         const int *__restrict__ chunk_blk, const int *__restrict__ pair,
                                                         const int *__restrict__ blk_rk,
                                                         const int *__restrict__ contrib_ptr,
                                                         const uint16_t *__restrict__ contrib,
                                                         const int *__restrict__ col_ptr, double *__restrict__ val,
                                                         int accumulate, int nchunks)
{
    __shared__ double s_blk[HEX_CAP * 8][9];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int p0 = chunk_pair[chunk], p1 = chunk_pair[chunk + 1];
    if (p1 > p0) {  // uniform over the workgroup
        const int lp = threadIdx.x >> 3, gp = threadIdx.x & 7;
        const int p = min(p0 + lp, p1 - 1);  // idle groups recompute the last pair, store nothing
        const int pk = pair[p];
        const int el = pk >> 3, a = pk & 7;
        double X[24];
        const int *cn = conn + 8 * el;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            int node = cn[b];
            X[3 * b] = xyz[3 * node];
            X[3 * b + 1] = xyz[3 * node + 1];
            X[3 * b + 2] = xyz[3 * node + 2];
        }
        double g[3][8];
        double detJ = hex_point<double>(X, gp, g);
        double lam, mu;
        lame<double>(Eel ? Eel[el] : E, nu, lam, mu);
        const double wl = lam * detJ, wm = mu * detJ;  // Gauss weight 1 (src/Elements/C3D8_Gauss.cpp:93)
        // row a gradient, selected without dynamic register indexing
        double ga[3];
        pick_col<double, 8>(g, a, ga);
        auto blk = [&](int b, double out[9]) {
            double gb0 = g[0][b], gb1 = g[1][b], gb2 = g[2][b];
            double gb[3] = {gb0, gb1, gb2};
            double dm = wm * (ga[0] * gb0 + ga[1] * gb1 + ga[2] * gb2);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    out[3 * c + d] = wl * ga[c] * gb[d] + wm * ga[d] * gb[c] + (c == d ? dm : 0.0);
        };
        double mine[9];
        reduce_scatter8<9, double>(gp, blk, mine);
        if (p0 + lp < p1) {
#pragma unroll
            for (int i = 0; i < 9; ++i) s_blk[lp * 8 + gp][i] = mine[i];
        }
    }
    __syncthreads();
    HexSource src{s_blk};
    gather_blocks(chunk_blk[chunk], chunk_blk[chunk + 1], blk_rk, contrib_ptr, contrib, col_ptr, val, accumulate, src);
}

int launch_assemble(ofea_ctx *ctx, double E, double nu, double *d_out, bool zero_first)
{
    bool mixed = ctx->has_type[0] && ctx->has_type[1];
    if (zero_first || mixed) OFEA_HIP(hipMemsetAsync(d_out, 0, sizeof(double) * ctx->P.nnz, ctx->stream));
    int acc = (zero_first || mixed) ? 1 : 0;
    for (int t = 0; t < 2; ++t) {
        if (!ctx->has_type[t]) continue;
        const DevPlan &dp = ctx->dp[t];
        if (dp.nchunks == 0) continue;
        if (t == OFEA_C3D4)
            hipLaunchKernelGGL(k_assemble_tet, dim3(dp.nchunks), dim3(ASM_WG), 0, ctx->stream, ctx->d_xyz, dp.conn, dp.E,
                               E, nu, dp.chunk_pair, dp.chunk_blk, dp.pair, dp.blk_rk, dp.contrib_ptr, dp.contrib,
                               ctx->d_col_ptr, d_out, acc, dp.nchunks);
        else
            hipLaunchKernelGGL(k_assemble_hex, dim3(dp.nchunks), dim3(ASM_WG), 0, ctx->stream, ctx->d_xyz, dp.conn, dp.E,
                               E, nu, dp.chunk_pair, dp.chunk_blk, dp.pair, dp.blk_rk, dp.contrib_ptr, dp.contrib,
                               ctx->d_col_ptr, d_out, acc, dp.nchunks);
        OFEA_HIP(hipGetLastError());
    }
    return OFEA_OK;
}

int plan_caps(int type) { return type == OFEA_C3D8 ? HEX_CAP : TET_CAP; }

}  // namespace ofea
