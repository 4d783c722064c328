// nl.hip -- total-Lagrangian St.Venant-Kirchhoff tangent and internal force.
//
// Reference: build_GL_strain_gauss / build_ele_nl_stiff_mat_gauss
// (src/Elements/C3D8_Gauss.cpp:110-266), build_ele_nl_stiff_mat_hammer
// (src/Elements/C3D4_hammer.cpp:75-183), assembled by
// assamble_nl_stiffness::fill_CSR_sparse_mat (src/nl_assemble.cpp:94-166).
//
// With F = I + grad u, E = (F^T F - I)/2, S = lambda tr(E) I + 2 mu E (the
// reference's D applied to Voigt E with doubled shears), h_a = F g_a, the
// reference's B_L^T D B_L + B_NL^T Sigma B_NL block and B_L^T S force are
//   K_ab[c][d] = w (lambda h_a[c] h_b[d] + mu (F F^T)_cd g_a.g_b + mu h_a[d] h_b[c]
//                   + delta_cd g_a^T S g_b)
//   f_a = w F S g_a
// The kernels reuse the column-owner gather plan of the linear assembly.
#include <algorithm>
#include <cmath>
#include <vector>

#include "elem_device.h"
#include "ofea_internal.h"

namespace ofea {

constexpr int NL_WG = 256;
constexpr int NL_TET_CAP = PLAN_TET_CAP;
constexpr int NL_HEX_CAP = PLAN_HEX_CAP;

struct NLPoint {
    double F[3][3], S[3][3], FFt[3][3];
};

__device__ inline void nl_state(const double H[3][3], double lam, double mu, NLPoint &P)
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int J = 0; J < 3; ++J) P.F[c][J] = H[c][J] + (c == J ? 1.0 : 0.0);
    double E[3][3];
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
        for (int J = 0; J < 3; ++J) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) acc += P.F[c][I] * P.F[c][J];
            E[I][J] = 0.5 * (acc - (I == J ? 1.0 : 0.0));
        }
    double tr = E[0][0] + E[1][1] + E[2][2];
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
        for (int J = 0; J < 3; ++J) P.S[I][J] = 2.0 * mu * E[I][J] + (I == J ? lam * tr : 0.0);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) P.FFt[c][d] = P.F[c][0] * P.F[d][0] + P.F[c][1] * P.F[d][1] + P.F[c][2] * P.F[d][2];
}

__device__ inline void nl_block(const NLPoint &P, const double ga[3], const double gb[3], double wl, double wm, double w,
                                double out[9])
{
    double ha[3], hb[3], Sgb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ha[c] = P.F[c][0] * ga[0] + P.F[c][1] * ga[1] + P.F[c][2] * ga[2];
        hb[c] = P.F[c][0] * gb[0] + P.F[c][1] * gb[1] + P.F[c][2] * gb[2];
        Sgb[c] = P.S[c][0] * gb[0] + P.S[c][1] * gb[1] + P.S[c][2] * gb[2];
    }
    double gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
    double gSg = w * (ga[0] * Sgb[0] + ga[1] * Sgb[1] + ga[2] * Sgb[2]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            out[3 * c + d] = wl * ha[c] * hb[d] + wm * P.FFt[c][d] * gg + wm * ha[d] * hb[c] + (c == d ? gSg : 0.0);
}

// element displacement from the free-DOF vector (clamped DOFs are 0,
// src/nl_assemble.cpp:116-131)
template <int NN>
__device__ inline void gather_u(const int *cn, const int *__restrict__ resort, const double *__restrict__ u, double U[3 * NN])
{
#pragma unroll
    for (int b = 0; b < NN; ++b) {
        int r = resort[cn[b]];
#pragma unroll
        for (int c = 0; c < 3; ++c) U[3 * b + c] = r >= 0 ? u[3 * r + c] : 0.0;
    }
}

template <class Source>
__device__ inline void nl_gather_blocks(int b0, int b1, const int *__restrict__ blk_rk,
                                        const int *__restrict__ contrib_ptr, const uint16_t *__restrict__ contrib,
                                        const int *__restrict__ col_ptr, double *__restrict__ val, int accumulate,
                                        const Source &src)
{
    for (int bk = b0 + (int)threadIdx.x; bk < b1; bk += NL_WG) {
        int rk = blk_rk[bk];
        int r = rk >> 8, k = rk & 255;
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        int t1 = contrib_ptr[bk + 1];
        for (int t = contrib_ptr[bk]; t < t1; ++t) src.add(contrib[t], acc);
        int base = col_ptr[3 * r];
        int stride = col_ptr[3 * r + 1] - base;
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

// f_int of the chunk's nodes: sum of the pair forces in ascending element order
template <class ForceFn>
__device__ inline void nl_gather_force(int chunk, const int *__restrict__ chunk_node, const int *__restrict__ node_pair,
                                       int p0, const ForceFn &pf, double *__restrict__ fint, int accumulate)
{
    const int r0 = chunk_node[chunk], r1 = chunk_node[chunk + 1];
    for (int r = r0 + (int)threadIdx.x; r < r1; r += NL_WG) {
        double f[3] = {0.0, 0.0, 0.0};
        for (int p = node_pair[r]; p < node_pair[r + 1]; ++p) pf(p - p0, f);
        if (accumulate) {
            f[0] += fint[3 * r];
            f[1] += fint[3 * r + 1];
            f[2] += fint[3 * r + 2];
        }
        fint[3 * r] = f[0];
        fint[3 * r + 1] = f[1];
        fint[3 * r + 2] = f[2];
    }
}

// ------------------------------------------------------------------ C3D8
struct NLHexSource {
    const double (*blk)[9];
    __device__ void add(int ci, double acc[9]) const
    {
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += blk[ci][i];
    }
};

__global__ __launch_bounds__(NL_WG) void k_nl_hex(const double *__restrict__ xyz, const int *__restrict__ conn,
                                                  const int *__restrict__ resort, const double *__restrict__ u, double lam,
                                                  double mu, const int *__restrict__ chunk_pair,
                                                  const int *__restrict__ chunk_blk, const int *__restrict__ chunk_node,
                                                  const int *__restrict__ node_pair, const int *__restrict__ pair,
                                                  const int *__restrict__ blk_rk, const int *__restrict__ contrib_ptr,
                                                  const uint16_t *__restrict__ contrib, const int *__restrict__ col_ptr,
                                                  double *__restrict__ val, double *__restrict__ fint, int accumulate,
                                                  int nchunks)
{
    __shared__ double s_blk[NL_HEX_CAP * 8][9];
    __shared__ double s_f[NL_HEX_CAP][3];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int p0 = chunk_pair[chunk], p1 = chunk_pair[chunk + 1];
    if (p1 > p0) {
        const int lp = threadIdx.x >> 3, gp = threadIdx.x & 7;
        const int p = min(p0 + lp, p1 - 1);
        const int pk = pair[p];
        const int el = pk >> 3, a = pk & 7;
        const int *cn = conn + 8 * el;
        double X[24], U[24];
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) X[3 * b + c] = xyz[3 * cn[b] + c];
        gather_u<8>(cn, resort, u, U);
        double g[3][8];
        const double w = hex_point<double>(X, gp, g);  // weight 1
        double H[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int J = 0; J < 3; ++J) {
                double acc = 0.0;
#pragma unroll
                for (int b = 0; b < 8; ++b) acc += U[3 * b + c] * g[J][b];
                H[c][J] = acc;
            }
        NLPoint P;
        nl_state(H, lam, mu, P);
        double ga[3];
        pick_col<double, 8>(g, a, ga);
        const double wl = w * lam, wm = w * mu;
        auto blk = [&](int b, double out[9]) {
            double gb[3] = {g[0][b], g[1][b], g[2][b]};
            nl_block(P, ga, gb, wl, wm, w, out);
        };
        double mine[9];
        reduce_scatter8<9, double>(gp, blk, mine);
        // f_a = w F S g_a, summed over the 8 Gauss points of the group
        double fa[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int J = 0; J < 3; ++J) {
                double SgJ = P.S[J][0] * ga[0] + P.S[J][1] * ga[1] + P.S[J][2] * ga[2];
                acc += P.F[c][J] * SgJ;
            }
            fa[c] = group8_sum(w * acc);
        }
        if (p0 + lp < p1) {
#pragma unroll
            for (int i = 0; i < 9; ++i) s_blk[lp * 8 + gp][i] = mine[i];
            if (gp == 0) {
                s_f[lp][0] = fa[0];
                s_f[lp][1] = fa[1];
                s_f[lp][2] = fa[2];
            }
        }
    }
    __syncthreads();
    NLHexSource src{s_blk};
    nl_gather_blocks(chunk_blk[chunk], chunk_blk[chunk + 1], blk_rk, contrib_ptr, contrib, col_ptr, val, accumulate, src);
    auto pf = [&](int lp, double f[3]) {
        f[0] += s_f[lp][0];
        f[1] += s_f[lp][1];
        f[2] += s_f[lp][2];
    };
    nl_gather_force(chunk, chunk_node, node_pair, p0, pf, fint, accumulate);
}

// ------------------------------------------------------------------ C3D4
__device__ inline void load_tet_state(const double *st, NLPoint &P)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P.F[i][j] = st[3 * i + j];
    P.S[0][0] = st[9];
    P.S[1][1] = st[10];
    P.S[2][2] = st[11];
    P.S[0][1] = P.S[1][0] = st[12];
    P.S[1][2] = P.S[2][1] = st[13];
    P.S[0][2] = P.S[2][0] = st[14];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) P.FFt[c][d] = P.F[c][0] * P.F[d][0] + P.F[c][1] * P.F[d][1] + P.F[c][2] * P.F[d][2];
}

struct NLTetSource {
    const double (*g)[12];   // slot j: gradient of local node (a + j) & 3
    const double (*st)[16];  // F (9), S (xx yy zz xy yz xz), w
    double lam, mu;
    __device__ void add(int ci, double acc[9]) const
    {
        int lp = ci >> 2, j = ci & 3;
        NLPoint P;
        load_tet_state(st[lp], P);
        double w = st[lp][15];
        double ga[3] = {g[lp][0], g[lp][1], g[lp][2]};
        double gb[3] = {g[lp][3 * j], g[lp][3 * j + 1], g[lp][3 * j + 2]};
        double out[9];
        nl_block(P, ga, gb, w * lam, w * mu, w, out);
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += out[i];
    }
};

__global__ __launch_bounds__(NL_WG) void k_nl_tet(const double *__restrict__ xyz, const int *__restrict__ conn,
                                                  const int *__restrict__ resort, const double *__restrict__ u, double lam,
                                                  double mu, const int *__restrict__ chunk_pair,
                                                  const int *__restrict__ chunk_blk, const int *__restrict__ chunk_node,
                                                  const int *__restrict__ node_pair, const int *__restrict__ pair,
                                                  const int *__restrict__ blk_rk, const int *__restrict__ contrib_ptr,
                                                  const uint16_t *__restrict__ contrib, const int *__restrict__ col_ptr,
                                                  double *__restrict__ val, double *__restrict__ fint, int accumulate,
                                                  int nchunks)
{
    __shared__ double s_g[NL_TET_CAP][12];
    __shared__ double s_st[NL_TET_CAP][16];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int p0 = chunk_pair[chunk], p1 = chunk_pair[chunk + 1];
    for (int p = p0 + (int)threadIdx.x; p < p1; p += NL_WG) {
        const int pk = pair[p], el = pk >> 3, a = pk & 7, lp = p - p0;
        const int *cn = conn + 4 * el;
        double X[12], U[12];
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) X[3 * b + c] = xyz[3 * cn[b] + c];
        gather_u<4>(cn, resort, u, U);
        double g[3][4];
        const double w = tet_grads<double>(X, g) / 6.0;  // Hammer weight 1/6
        double H[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int J = 0; J < 3; ++J) H[c][J] = U[c] * g[J][0] + U[3 + c] * g[J][1] + U[6 + c] * g[J][2] + U[9 + c] * g[J][3];
        NLPoint P;
        nl_state(H, lam, mu, P);
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) s_g[lp][3 * ((b - a) & 3) + c] = g[c][b];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) s_st[lp][3 * i + j] = P.F[i][j];
        s_st[lp][9] = P.S[0][0];
        s_st[lp][10] = P.S[1][1];
        s_st[lp][11] = P.S[2][2];
        s_st[lp][12] = P.S[0][1];
        s_st[lp][13] = P.S[1][2];
        s_st[lp][14] = P.S[0][2];
        s_st[lp][15] = w;
    }
    __syncthreads();
    NLTetSource src{s_g, s_st, lam, mu};
    nl_gather_blocks(chunk_blk[chunk], chunk_blk[chunk + 1], blk_rk, contrib_ptr, contrib, col_ptr, val, accumulate, src);
    // f_a = w F S g_a from the staged point state
    auto pf = [&](int lp, double f[3]) {
        NLPoint P;
        load_tet_state(s_st[lp], P);
        const double ga[3] = {s_g[lp][0], s_g[lp][1], s_g[lp][2]};
        const double w = s_st[lp][15];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int J = 0; J < 3; ++J) acc += P.F[c][J] * (P.S[J][0] * ga[0] + P.S[J][1] * ga[1] + P.S[J][2] * ga[2]);
            f[c] += w * acc;
        }
    };
    nl_gather_force(chunk, chunk_node, node_pair, p0, pf, fint, accumulate);
}

int nl_plan_cap(int type) { return type == OFEA_C3D8 ? NL_HEX_CAP : NL_TET_CAP; }

// lag_coarse: a Newton tangent whose two-level solve may keep the coarse
// inverse of the previous tangent (solver_invalidate_values)
static int nl_assemble_dev(ofea_ctx *c, double E, double nu, bool lag_coarse = false)
{
    const int n = c->P.n;
    const bool mixed = c->has_type[0] && c->has_type[1];
    if (mixed) {
        OFEA_HIP(hipMemsetAsync(c->d_val, 0, sizeof(double) * c->P.nnz, c->stream));
        OFEA_HIP(hipMemsetAsync(c->d_fint, 0, sizeof(double) * std::max(n, 1), c->stream));
    }
    double lam, mu;
    lame<double>(E, nu, lam, mu);
    for (int t = 0; t < 2; ++t) {
        if (!c->has_type[t]) continue;
        const DevPlan &d = c->dp[t];
        if (d.nchunks == 0) continue;
        auto kern = t == OFEA_C3D8 ? k_nl_hex : k_nl_tet;
        hipLaunchKernelGGL(kern, dim3(d.nchunks), dim3(NL_WG), 0, c->stream, c->d_xyz, d.conn, c->d_resort, c->d_u, lam,
                           mu, d.chunk_pair, d.chunk_blk, d.chunk_node, d.node_pair, d.pair, d.blk_rk, d.contrib_ptr,
                           d.contrib, c->d_col_ptr, c->d_val, c->d_fint, mixed ? 1 : 0, d.nchunks);
        OFEA_HIP(hipGetLastError());
    }
    c->assembled = true;
    c->kver++;
    if (lag_coarse) solver_invalidate_values(c);
    else solver_invalidate(c);
    return OFEA_OK;
}

static int nl_buffers(ofea_ctx *c)
{
    const int n = c->P.n;
    if (!c->d_u) OFEA_HIP(hipMalloc((void **)&c->d_u, sizeof(double) * std::max(n, 1)));
    if (!c->d_fint) OFEA_HIP(hipMalloc((void **)&c->d_fint, sizeof(double) * std::max(n, 1)));
    return OFEA_OK;
}

int nl_assemble(ofea_ctx *c, double E, double nu, const double *u, double *nz_val, double *f_int)
{
    const int n = c->P.n;
    int rc = nl_buffers(c);
    if (rc) return rc;
    OFEA_HIP(hipMemcpyAsync(c->d_u, u, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if ((rc = nl_assemble_dev(c, E, nu))) return rc;
    if (nz_val) OFEA_HIP(hipMemcpyAsync(nz_val, c->d_val, sizeof(double) * c->P.nnz, hipMemcpyDeviceToHost, c->stream));
    if (f_int) OFEA_HIP(hipMemcpyAsync(f_int, c->d_fint, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    OFEA_HIP(hipStreamSynchronize(c->stream));
    return OFEA_OK;
}

// r = lam F - f_int (into the solver's RHS buffer) with per-block partial
// sums of r.r and (lam F).(lam F)
constexpr int NLR_WG = 256;
__global__ void k_nl_residual(int n, double lam, const double *__restrict__ F, const double *__restrict__ fint,
                              double *__restrict__ r, double *__restrict__ part)
{
    __shared__ double s1[NLR_WG / 64], s2[NLR_WG / 64];
    const int i = blockIdx.x * NLR_WG + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (i < n) {
        const double f = lam * F[i];
        const double v = f - fint[i];
        r[i] = v;
        a = v * v;
        b = f * f;
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if ((threadIdx.x & 63) == 0) {
        s1[threadIdx.x >> 6] = a;
        s2[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0;
        for (int k = 0; k < NLR_WG / 64; ++k) {
            ta += s1[k];
            tb += s2[k];
        }
        part[2 * blockIdx.x] = ta;
        part[2 * blockIdx.x + 1] = tb;
    }
}

__global__ void k_axpy(int n, double a, const double *__restrict__ x, double *__restrict__ y)
{
    const int i = blockIdx.x * NLR_WG + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

int learned_refresh(ofea_ctx *c);

// Linear solves of a Newton run, K_T du = r with the resident K_T and rhs.
// A learned method is judged by the same solver with plain block Jacobi: on
// the first linear solve the plain solver runs, then the learned one limited
// to the plain iteration count; the learned method serves the rest of the run
// only if it converged in fewer iterations, and the plain solver takes over
// whenever it fails later.  The solution is left in c->d_x.
struct NewtonSolves {
    int use;         // method in use
    bool judged;     // the learned method has been judged (or none is used)
    double *xs;      // plain solution kept while the learned one is judged (n)
    int linear = 0;  // PCG iterations of every solve

    static bool is_learned(int m)
    {
        return m == OFEA_PCG_LEARNED || m == OFEA_PCG_LEARNED_MIXED || m == OFEA_PCG_LEARNED_TWO_LEVEL ||
               m == OFEA_PCG_LEARNED_TWO_LEVEL_MIXED || m == OFEA_PCG_LEARNED_TWO_LEVEL_MIXED32;
    }
    // the same solver with block Jacobi in place of the learned blocks
    static int plain_of(int m)
    {
        return m == OFEA_PCG_LEARNED_MIXED             ? OFEA_PCG_MIXED
               : m == OFEA_PCG_LEARNED_TWO_LEVEL       ? OFEA_PCG_TWO_LEVEL
               : m == OFEA_PCG_LEARNED_TWO_LEVEL_MIXED ? OFEA_PCG_TWO_LEVEL_MIXED
               : m == OFEA_PCG_LEARNED_TWO_LEVEL_MIXED32 ? OFEA_PCG_TWO_LEVEL_MIXED32
                                                       : OFEA_PCG_BLOCK_JACOBI;
    }
    NewtonSolves(int method, double *buf) : use(method), judged(!is_learned(method)), xs(buf) {}
    // The two-level coarse inverse is rebuilt for a tangent only when the
    // last solve needed 25% more iterations per decade of residual reduction
    // than the first solve after the previous rebuild; otherwise the next
    // tangent keeps it (its values, block preconditioners and learned factor
    // are always the tangent's own).
    double base_rate = -1.0, last_rate = 0.0;
    bool lag_coarse()
    {
        if (base_rate >= 0.0 && last_rate <= 1.25 * base_rate) return true;
        base_rate = -1.0;  // the next solve, on a rebuilt coarse space, sets the base
        return false;
    }

    int solve(ofea_ctx *c, int n, double inner_rtol)
    {
        c->rhs_set = true;
        int lit = 0;
        double lrel = 0.0;
        int m = use;
        if (is_learned(m) && (!c->learned || learned_refresh(c) != OFEA_OK)) m = use = plain_of(m);
        int src;
        if (is_learned(m) && !judged) {
            judged = true;
            int pit = 0;
            src = solver_solve(c, plain_of(m), inner_rtol, 20000, &pit, &lrel);
            linear += pit;
            if (src == OFEA_OK) {
                if (hipMemcpyAsync(xs, c->d_x, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
                    return fail(OFEA_ERR_HIP, "solution copy failed");
                const int lsrc = solver_solve(c, m, inner_rtol, pit, &lit, &lrel);
                if (lsrc < 0 && lsrc != OFEA_ERR_STATE) {  // OFEA_ERR_STATE: breakdown of the learned PCG
                    src = lsrc;
                } else {
                    linear += lit;
                    if (lsrc != OFEA_OK || lit >= pit) {
                        use = plain_of(m);
                        if (hipMemcpyAsync(c->d_x, xs, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream) !=
                            hipSuccess)
                            return fail(OFEA_ERR_HIP, "solution copy failed");
                    }
                }
            }
        } else {
            src = solver_solve(c, m, inner_rtol, 20000, &lit, &lrel);
            linear += lit;
            if (src != OFEA_OK && is_learned(m)) {  // learned path failed: the plain solver takes over
                use = plain_of(m);
                src = solver_solve(c, use, inner_rtol, 20000, &lit, &lrel);
                linear += lit;
            }
        }
        last_rate = lit / std::max(0.5, -std::log10(inner_rtol));
        if (base_rate < 0.0) base_rate = last_rate;
        return src < 0 ? src : OFEA_OK;
    }
};

// r = lam F - f_int into c->d_rhs; returns ||r||^2 and ||lam F||^2 (FP64,
// partial sums added on the host in a fixed order).
static int nl_residual(ofea_ctx *c, int n, double lam, const double *dF, double *part, std::vector<double> &hp,
                       double &rr, double &ff)
{
    const int nb = (n + NLR_WG - 1) / NLR_WG;
    hipLaunchKernelGGL(k_nl_residual, dim3(nb), dim3(NLR_WG), 0, c->stream, n, lam, dF, c->d_fint, c->d_rhs, part);
    if (hipMemcpyAsync(hp.data(), part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(OFEA_ERR_HIP, "residual copy failed");
    rr = 0.0;
    ff = 0.0;
    for (int b = 0; b < nb; ++b) {
        rr += hp[2 * b];
        ff += hp[2 * b + 1];
    }
    return OFEA_OK;
}

// Device buffers of a Newton run: the load, residual partials, the kept plain solution.
struct NewtonBuffers {
    double *dF = nullptr, *part = nullptr, *xs = nullptr;
    std::vector<double> hp;
    int alloc(int n)
    {
        const int nb = (n + NLR_WG - 1) / NLR_WG;
        hp.assign(2 * std::max(nb, 1), 0.0);
        OFEA_HIP(hipMalloc((void **)&dF, sizeof(double) * std::max(n, 1)));
        OFEA_HIP(hipMalloc((void **)&part, sizeof(double) * 2 * std::max(nb, 1)));
        OFEA_HIP(hipMalloc((void **)&xs, sizeof(double) * std::max(n, 1)));
        return OFEA_OK;
    }
    ~NewtonBuffers()
    {
        (void)hipFree(dF);
        (void)hipFree(part);
        (void)hipFree(xs);
    }
};

// The Newton drivers let tangents keep an earlier coarse inverse; on the way
// out (any path) the resident K_T gets its own on the next solve.
struct FullInvalidateOnExit {
    ofea_ctx *c;
    ~FullInvalidateOnExit() { solver_invalidate(c); }
};

// Load-stepped Newton on the total-Lagrangian St.Venant-Kirchhoff problem.
// Reference loop: src/cae.cpp:209-253 (full load in one step, r = F - f_int,
// K_T du = r, u += du, stop at ||r||_2/(3n) < 1e-6 or 100 loops; that rule is
// nl_solve_ref below).  Here the load is ramped in n_incr increments
// lam_k = k/n_incr and each increment iterates until
// ||lam F - f_int|| <= tol ||lam F|| (FP64 residual) or max_newton
// iterations; the linear solves are GPU PCG (method, inner_rtol; learned
// methods judged as NewtonSolves describes).
int nl_solve(ofea_ctx *c, double E, double nu, const double *F, int n_incr, double tol, int max_newton, int method,
             double inner_rtol, double *u_out, int *newton_total, int *linear_total, double *relres)
{
    const int n = c->P.n;
    int rc = nl_buffers(c);
    if (rc) return rc;
    NewtonBuffers B;
    if ((rc = B.alloc(n))) return rc;
    const int nb = (n + NLR_WG - 1) / NLR_WG;
    OFEA_HIP(hipMemcpyAsync(B.dF, F, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    OFEA_HIP(hipMemsetAsync(c->d_u, 0, sizeof(double) * std::max(n, 1), c->stream));
    NewtonSolves S(method, B.xs);
    FullInvalidateOnExit guard{c};
    int newton = 0;
    double rel = 0.0;
    // inner_rtol < 0: inexact Newton with the forcing terms of Eisenstat and
    // Walker (choice 2: eta_k = 0.9 (|r_k| / |r_k-1|)^2, kept at least
    // 0.9 eta_k-1^2 when that exceeds 0.1), at most 0.1, never tighter than
    // half of what the outer tolerance needs nor than -inner_rtol
    const bool adaptive = inner_rtol < 0.0;
    const double eta_min = adaptive ? -inner_rtol : inner_rtol;
    for (int k = 1; k <= n_incr; ++k) {
        const double lam = (double)k / n_incr;
        bool conv = false;
        double eta_prev = 0.0, rn_prev = 0.0;
        for (int it = 0; it <= max_newton; ++it) {
            if ((rc = nl_assemble_dev(c, E, nu, S.lag_coarse()))) break;
            double rr = 0.0, ff = 0.0;
            if ((rc = nl_residual(c, n, lam, B.dF, B.part, B.hp, rr, ff))) break;
            rel = ff > 0.0 ? std::sqrt(rr / ff) : std::sqrt(rr);
            if (rel <= tol) {
                conv = true;
                break;
            }
            if (it == max_newton) break;
            double eta = eta_min;
            if (adaptive) {
                const double rn = std::sqrt(rr);
                eta = 0.1;
                if (it > 0 && rn_prev > 0.0) {
                    const double q = rn / rn_prev;
                    eta = 0.9 * q * q;
                    const double sg = 0.9 * eta_prev * eta_prev;
                    if (sg > 0.1) eta = std::max(eta, sg);
                }
                eta = std::min(eta, 0.1);
                if (rn > 0.0 && ff > 0.0) eta = std::max(eta, 0.5 * tol * std::sqrt(ff) / rn);
                eta = std::max(eta, eta_min);
                eta_prev = eta;
                rn_prev = rn;
            }
            if ((rc = S.solve(c, n, eta))) break;
            hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(NLR_WG), 0, c->stream, n, 1.0, c->d_x, c->d_u);
            newton++;
        }
        if (rc) break;
        if (!conv) {
            rc = OFEA_NOT_CONVERGED;
            break;
        }
    }
    if (rc < 0) return rc;
    if (u_out) {
        OFEA_HIP(hipMemcpyAsync(u_out, c->d_u, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
    }
    if (newton_total) *newton_total = newton;
    if (linear_total) *linear_total = S.linear;
    if (relres) *relres = rel;
    return rc;
}

// The reference's Newton loop as it stands (src/cae.cpp:209-253): res_norm
// starts at 1; while res_norm > res_limit and loop < max_loop: K_T, f_int at
// u; r = F - f_int; K_T du = r; u += du; res_norm = ||r||_2 / (3 nfree) of
// that r (the residual before the update); loop++.  The full load at once,
// no increments; the last du is applied after the residual test passes.
// Linear solves: GPU PCG (method, inner_rtol) in place of PARDISO.
int nl_solve_ref(ofea_ctx *c, double E, double nu, const double *F, double res_limit, int max_loop, int method,
                 double inner_rtol, double *u_out, int *loops, int *linear_total, double *res_norm_out)
{
    const int n = c->P.n;
    int rc = nl_buffers(c);
    if (rc) return rc;
    NewtonBuffers B;
    if ((rc = B.alloc(n))) return rc;
    const int nb = (n + NLR_WG - 1) / NLR_WG;
    OFEA_HIP(hipMemcpyAsync(B.dF, F, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    OFEA_HIP(hipMemsetAsync(c->d_u, 0, sizeof(double) * std::max(n, 1), c->stream));
    NewtonSolves S(method, B.xs);
    FullInvalidateOnExit guard{c};
    double res_norm = 1.;
    int loop = 0;
    while (res_norm > res_limit && loop < max_loop) {
        if ((rc = nl_assemble_dev(c, E, nu, S.lag_coarse()))) return rc;
        double rr = 0.0, ff = 0.0;
        if ((rc = nl_residual(c, n, 1.0, B.dF, B.part, B.hp, rr, ff))) return rc;
        if ((rc = S.solve(c, n, inner_rtol))) return rc;
        hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(NLR_WG), 0, c->stream, n, 1.0, c->d_x, c->d_u);
        res_norm = std::sqrt(rr) / n;
        loop++;
    }
    if (u_out) {
        OFEA_HIP(hipMemcpyAsync(u_out, c->d_u, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
    }
    if (loops) *loops = loop;
    if (linear_total) *linear_total = S.linear;
    if (res_norm_out) *res_norm_out = res_norm;
    return res_norm > res_limit ? OFEA_NOT_CONVERGED : OFEA_OK;
}

}  // namespace ofea
