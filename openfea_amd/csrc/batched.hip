// batched.hip -- many design variants and load cases per launch, and the
// SpMV tile autotuner.
//
// No direct counterpart in the reference (one material, one load, direct
// solves; src/cae.cpp:198-301).  The reference's building blocks are reused:
// the variant stiffness is fill_CSR_sparse_mat (src/assemble.cpp:95-151) with a
// per-element modulus, and each load case is one solution_api solve.
//
//  * ofea_assemble_batch: one launch assembles nvar variants K_v =
//    sum_e E_v[e] Khat_e, where Khat_e is the unit-modulus element matrix (D is
//    linear in E at fixed nu).  The element blocks are computed once per
//    (node, element) pair, then each gathered 3x3 block is scaled by the
//    variants' moduli (staged in LDS) and written for every variant: the
//    kernel is bound by the nvar * nnz * 8 bytes it must write.
//  * ofea_solve_batch: for each variant, nrhs independent PCG systems share
//    every read of K_v (vectors interleaved [node][dof][rhs]); each system has
//    its own scalars and stops on its own.  The matrix is reached through a
//    device pointer slot, so one captured hipGraph serves all variants.
//  * ofea_autotune: times the SpMV lane-group size on the resident matrix and
//    keeps the fastest for the solvers.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "elem_device.h"
#include "ofea_internal.h"
#include "solver_ws.h"

namespace ofea {

constexpr int BA_WG = 256;
constexpr int BA_HEX_CAP = PLAN_HEX_CAP;
constexpr int BA_TET_CAP = PLAN_TET_CAP;
constexpr int BA_VMAX = 64;  // variants per launch (LDS staging of the moduli)
constexpr int BA_VC = 4;     // variants accumulated per pass over the contributors

// ------------------------------------------------------------------ batched assembly
template <class UnitBlock>
__device__ inline void gather_blocks_batch(int b0, int b1, const int *__restrict__ blk_rk,
                                           const int *__restrict__ contrib_ptr, const uint16_t *__restrict__ contrib,
                                           const int *__restrict__ col_ptr, double *__restrict__ valb, long long nnz,
                                           int nvar, int nn, const double (*sE)[BA_VMAX], const UnitBlock &unit)
{
    // one task per (block, group of BA_VC variants), blocks fastest: every
    // lane of the workgroup has work although a chunk has fewer blocks than
    // lanes, and neighbouring lanes still write neighbouring blocks
    const int nblk = b1 - b0;
    const int ntask = nblk * ((nvar + BA_VC - 1) / BA_VC);
    for (int task = (int)threadIdx.x; task < ntask; task += BA_WG) {
        const int grp = task / nblk;
        const int bk = b0 + (task - grp * nblk);
        const int v0 = grp * BA_VC;
        const int rk = blk_rk[bk];
        const int r = rk >> 8, k = rk & 255;
        const int t0 = contrib_ptr[bk], t1 = contrib_ptr[bk + 1];
        const int base = col_ptr[3 * r];
        const int stride = col_ptr[3 * r + 1] - base;
        {
            double acc[BA_VC][9];
#pragma unroll
            for (int v = 0; v < BA_VC; ++v)
#pragma unroll
                for (int i = 0; i < 9; ++i) acc[v][i] = 0.0;
            for (int t = t0; t < t1; ++t) {
                const int ci = contrib[t];
                const int lp = ci / nn;
                double blk[9];
                unit(ci, blk);
#pragma unroll
                for (int v = 0; v < BA_VC; ++v) {
                    const double e = (v0 + v < nvar) ? sE[lp][v0 + v] : 0.0;
#pragma unroll
                    for (int i = 0; i < 9; ++i) acc[v][i] += e * blk[i];
                }
            }
#pragma unroll
            for (int v = 0; v < BA_VC; ++v) {
                if (v0 + v >= nvar) break;
                double *dst = valb + (long long)(v0 + v) * nnz + base + 3 * k;
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d) dst[c * stride + d] = acc[v][3 * c + d];
            }
        }
    }
}

__global__ __launch_bounds__(BA_WG) void k_assemble_hex_batch(
    const double *__restrict__ xyz, const int *__restrict__ conn, int nvar, const double *__restrict__ Evar, double nu,
    const int *__restrict__ chunk_pair, const int *__restrict__ chunk_blk, const int *__restrict__ pair,
    const int *__restrict__ blk_rk, const int *__restrict__ contrib_ptr, const uint16_t *__restrict__ contrib,
    const int *__restrict__ col_ptr, double *__restrict__ valb, long long nnz, int nchunks)
{
    __shared__ double s_blk[BA_HEX_CAP * 8][9];
    __shared__ double s_E[BA_HEX_CAP][BA_VMAX];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int p0 = chunk_pair[chunk], p1 = chunk_pair[chunk + 1];
    if (p1 > p0) {
        const int lp = threadIdx.x >> 3, gp = threadIdx.x & 7;
        const int p = min(p0 + lp, p1 - 1);
        const int pk = pair[p];
        const int el = pk >> 3, a = pk & 7;
        double X[24];
        const int *cn = conn + 8 * el;
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) X[3 * b + c] = xyz[3 * cn[b] + c];
        double g[3][8];
        const double detJ = hex_point<double>(X, gp, g);
        double lam, mu;
        lame<double>(1.0, nu, lam, mu);
        const double wl = lam * detJ, wm = mu * detJ;
        double ga[3];
        pick_col<double, 8>(g, a, ga);
        auto blk = [&](int b, double out[9]) {
            const double gb[3] = {g[0][b], g[1][b], g[2][b]};
            const double dm = wm * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
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
            for (int v = gp; v < nvar; v += 8) s_E[lp][v] = Evar[(long long)el * nvar + v];
        }
    }
    __syncthreads();
    auto unit = [&](int ci, double out[9]) {
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = s_blk[ci][i];
    };
    gather_blocks_batch(chunk_blk[chunk], chunk_blk[chunk + 1], blk_rk, contrib_ptr, contrib, col_ptr, valb, nnz, nvar,
                        8, s_E, unit);
}

__global__ __launch_bounds__(BA_WG) void k_assemble_tet_batch(
    const double *__restrict__ xyz, const int *__restrict__ conn, int nvar, const double *__restrict__ Evar, double nu,
    const int *__restrict__ chunk_pair, const int *__restrict__ chunk_blk, const int *__restrict__ pair,
    const int *__restrict__ blk_rk, const int *__restrict__ contrib_ptr, const uint16_t *__restrict__ contrib,
    const int *__restrict__ col_ptr, double *__restrict__ valb, long long nnz, int nchunks)
{
    // 256 pairs x 64 moduli would not fit beside the gradients in LDS: the
    // moduli are read from global memory (the element's row of nvar values).
    __shared__ double s_g[BA_TET_CAP][12];  // slot j: gradient of local node (a + j) & 3
    __shared__ double s_w[BA_TET_CAP][2];
    __shared__ int s_el[BA_TET_CAP];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int p0 = chunk_pair[chunk], p1 = chunk_pair[chunk + 1];
    double lam, mu;
    lame<double>(1.0, nu, lam, mu);
    for (int p = p0 + (int)threadIdx.x; p < p1; p += BA_WG) {
        const int pk = pair[p], el = pk >> 3, a = pk & 7, lp = p - p0;
        const int *cn = conn + 4 * el;
        double X[12];
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) X[3 * b + c] = xyz[3 * cn[b] + c];
        double g[3][4];
        const double V = tet_grads<double>(X, g) / 6.0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) s_g[lp][3 * ((b - a) & 3) + c] = g[c][b];
        s_w[lp][0] = lam * V;
        s_w[lp][1] = mu * V;
        s_el[lp] = el;
    }
    __syncthreads();
    // tasks (block, group of BA_VC variants), blocks fastest (as gather_blocks_batch)
    const int b0 = chunk_blk[chunk], nblk = chunk_blk[chunk + 1] - b0;
    const int ntask = nblk * ((nvar + BA_VC - 1) / BA_VC);
    for (int task = (int)threadIdx.x; task < ntask; task += BA_WG) {
        const int grp = task / nblk;
        const int bk = b0 + (task - grp * nblk);
        const int v0 = grp * BA_VC;
        const int rk = blk_rk[bk];
        const int r = rk >> 8, k = rk & 255;
        const int t0 = contrib_ptr[bk], t1 = contrib_ptr[bk + 1];
        const int base = col_ptr[3 * r];
        const int stride = col_ptr[3 * r + 1] - base;
        {
            double acc[BA_VC][9];
#pragma unroll
            for (int v = 0; v < BA_VC; ++v)
#pragma unroll
                for (int i = 0; i < 9; ++i) acc[v][i] = 0.0;
            for (int t = t0; t < t1; ++t) {
                const int ci = contrib[t], lp = ci >> 2, j = ci & 3;
                const double ga[3] = {s_g[lp][0], s_g[lp][1], s_g[lp][2]};
                const double gb[3] = {s_g[lp][3 * j], s_g[lp][3 * j + 1], s_g[lp][3 * j + 2]};
                const double wl = s_w[lp][0], wm = s_w[lp][1];
                const double dm = wm * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
                double blk[9];
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        blk[3 * c + d] = wl * ga[c] * gb[d] + wm * ga[d] * gb[c] + (c == d ? dm : 0.0);
                const double *Ee = Evar + (long long)s_el[lp] * nvar;
#pragma unroll
                for (int v = 0; v < BA_VC; ++v) {
                    const double e = (v0 + v < nvar) ? Ee[v0 + v] : 0.0;
#pragma unroll
                    for (int i = 0; i < 9; ++i) acc[v][i] += e * blk[i];
                }
            }
#pragma unroll
            for (int v = 0; v < BA_VC; ++v) {
                if (v0 + v >= nvar) break;
                double *dst = valb + (long long)(v0 + v) * nnz + base + 3 * k;
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d) dst[c * stride + d] = acc[v][3 * c + d];
            }
        }
    }
}

__global__ void k_transpose_moduli(int ne, int nvar, const double *__restrict__ Ev, double *__restrict__ Et)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;  // output index e * nvar + v
    if (i >= (long long)ne * nvar) return;
    const int e = (int)(i / nvar), v = (int)(i - (long long)e * nvar);
    Et[i] = Ev[(size_t)v * ne + e];
}

struct BatchWS {
    int nvar = 0;
    long long nnz = 0;
    double *valb = nullptr;  // nvar * nnz
    double *Evar[2] = {nullptr, nullptr};
    double *Eraw = nullptr;  // E_var as given, [variant][element]
    size_t Eraw_n = 0;
    int E_nvar = 0;
    // multi-RHS CG
    int nrhs = 0;
    double *B = nullptr, *X = nullptr, *R = nullptr, *Z = nullptr, *PA = nullptr, *PB = nullptr, *Q = nullptr,
           *Y = nullptr;
    float *minv = nullptr;  // inverse 3x3 diagonal blocks of the current variant, FP32
    double *part = nullptr;    // partial sums
    void *ms = nullptr;        // MScalars (device)
    void *ms_host = nullptr;   // pinned
    const double **valp = nullptr;  // device slot with the current variant's values
    // two-level: per-range partials of Z^T r, Z^T r per aggregate and y = A_c^-1 Z^T r (all systems)
    double *part_co = nullptr, *gco = nullptr, *yco = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_nrhs = 0, graph_group = 0, graph_co = -1;
    const void *graph_key = nullptr;  // pattern the captured graph reads
    int mb_group[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // tuned lane group per nrhs
    float mb_time[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
};

static BatchWS *bws(ofea_ctx *c) { return (BatchWS *)c->batch_ws; }

void batch_free(ofea_ctx *c)
{
    BatchWS *w = bws(c);
    if (!w) return;
    if (w->exec) (void)hipGraphExecDestroy(w->exec);
    if (w->graph) (void)hipGraphDestroy(w->graph);
    for (void *p : {(void *)w->valb, (void *)w->Evar[0], (void *)w->Evar[1], (void *)w->Eraw, (void *)w->B, (void *)w->X, (void *)w->R,
                    (void *)w->Z, (void *)w->PA, (void *)w->PB, (void *)w->Q, (void *)w->Y, (void *)w->minv,
                    (void *)w->part, w->ms, (void *)w->valp, (void *)w->part_co, (void *)w->gco, (void *)w->yco})
        if (p) (void)hipFree(p);
    if (w->ms_host) (void)hipHostFree(w->ms_host);
    delete w;
    c->batch_ws = nullptr;
}

static BatchWS *ensure_bws(ofea_ctx *c)
{
    if (!bws(c)) c->batch_ws = new BatchWS();
    return bws(c);
}

int assemble_batch(ofea_ctx *c, int nvar, const double *E_var, double nu, double *out)
{
    if (nvar < 1) return fail(OFEA_ERR_ARG, "nvar must be >= 1");
    if (nvar > BA_VMAX) return fail(OFEA_ERR_LIMIT, "at most 64 variants per batched assembly call");
    BatchWS *w = ensure_bws(c);
    const long long nnz = c->P.nnz;
    const bool mixed = c->has_type[0] && c->has_type[1];
    if (mixed) return fail(OFEA_ERR_ARG, "batched assembly supports single-type (all C3D4 or all C3D8) meshes");
    if (w->nvar != nvar || w->nnz != nnz) {
        if (w->valb) (void)hipFree(w->valb);
        w->valb = nullptr;
        OFEA_HIP(hipMalloc((void **)&w->valb, sizeof(double) * (size_t)nvar * std::max(nnz, 1LL)));
        w->nvar = nvar;
        w->nnz = nnz;
    }
    // moduli: E_var [variant][element] is copied as given and transposed on the
    // device to [element][variant]; NULL keeps the moduli of the previous call
    // (single-type mesh: type-local element index == global index)
    const int t_used = c->has_type[OFEA_C3D8] ? OFEA_C3D8 : OFEA_C3D4;
    const size_t ne = (size_t)c->ne;
    if (E_var) {
        if (w->Eraw_n != ne * nvar) {
            for (double **pp : {&w->Eraw, &w->Evar[0], &w->Evar[1]})
                if (*pp) {
                    (void)hipFree(*pp);
                    *pp = nullptr;
                }
            OFEA_HIP(hipMalloc((void **)&w->Eraw, sizeof(double) * std::max<size_t>(ne * nvar, 1)));
            OFEA_HIP(hipMalloc((void **)&w->Evar[t_used], sizeof(double) * std::max<size_t>(ne * nvar, 1)));
            w->Eraw_n = ne * nvar;
        }
        OFEA_HIP(hipMemcpyAsync(w->Eraw, E_var, sizeof(double) * ne * nvar, hipMemcpyHostToDevice, c->stream));
        const long long tot = (long long)ne * nvar;
        hipLaunchKernelGGL(k_transpose_moduli, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, (int)ne,
                           nvar, (const double *)w->Eraw, w->Evar[t_used]);
        OFEA_HIP(hipGetLastError());
        w->E_nvar = nvar;
    } else if (w->E_nvar != nvar || !w->Evar[t_used]) {
        return fail(OFEA_ERR_STATE, "E_var == NULL needs an earlier call with the same nvar");
    }
    for (int t = 0; t < 2; ++t) {
        if (!c->has_type[t]) continue;
        const DevPlan &d = c->dp[t];
        if (d.nchunks == 0) continue;
        auto kern = t == OFEA_C3D8 ? k_assemble_hex_batch : k_assemble_tet_batch;
        hipLaunchKernelGGL(kern, dim3(d.nchunks), dim3(BA_WG), 0, c->stream, c->d_xyz, d.conn, nvar, w->Evar[t], nu,
                           d.chunk_pair, d.chunk_blk, d.pair, d.blk_rk, d.contrib_ptr, d.contrib, c->d_col_ptr,
                           w->valb, nnz, d.nchunks);
        OFEA_HIP(hipGetLastError());
    }
    if (out) {
        OFEA_HIP(hipMemcpyAsync(out, w->valb, sizeof(double) * (size_t)nvar * nnz, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
    }
    return OFEA_OK;
}

// ------------------------------------------------------------------ multi-RHS PCG
// NR systems of one variant.  Vectors are interleaved [dof][system], so the
// NR lanes that share a matrix entry read NR consecutive doubles, and each
// matrix value is loaded once for all systems.  A workgroup lane owns one
// system k = lane % NR; the per-system scalars live in LDS.
constexpr int MB_NB = 512;    // workgroups of the CG kernels
constexpr int MB_WG = 1024;   // threads per workgroup
constexpr int MB_NW = MB_WG / 64;
constexpr int MB_NRMAX = 8;   // systems per variant
constexpr int MB_BATCH = 16;  // iterations per captured graph
constexpr int MB_UNR = 4;     // matrix entries per lane whose loads the SpMV issues together

struct MScalars {
    double tol2[MB_NRMAX];
    double rr[MB_NRMAX];
    int done[MB_NRMAX];  // 0 running, 1 converged, 2 maxit, 3 breakdown
    int it[MB_NRMAX];
    int par0, it0, maxit, all_done;
};

// partial-sum layout (doubles): rz[2][NRMAX][NB], rr[NRMAX][NB], pq[NRMAX][NB]
__device__ inline double *mb_rz(double *part, int par, int k) { return part + ((size_t)par * MB_NRMAX + k) * MB_NB; }
__device__ inline double *mb_rr(double *part, int k) { return part + ((size_t)2 * MB_NRMAX + k) * MB_NB; }
__device__ inline double *mb_pq(double *part, int k) { return part + ((size_t)3 * MB_NRMAX + k) * MB_NB; }

// s_out[q] = sum over the MB_NB partials of quantity q (q < nq).  Wave w
// handles q = w, w + 16, ...; lane l adds partials l, l + 64, ... and a fixed
// xor tree finishes, so every workgroup gets bit-identical sums.
template <class Src>
__device__ inline void mb_sum_partials(int nq, Src src, double *s_out)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int q = wave; q < nq; q += MB_NW) {
        const double *p = src(q);
        double v = 0.0;
        for (int i = lane; i < MB_NB; i += 64) v += p[i];  // written by an earlier kernel on the stream
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) s_out[q] = v;
    }
    __syncthreads();
}

// per-system workgroup sums of v (thread's system = threadIdx.x % NR) into s_red[q0 + k]
template <int NR>
__device__ inline void mb_wg_sum(double v, double (*s_tmp)[MB_NRMAX], double *dst_k)
{
    for (int off = 32; off >= NR; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) < NR) s_tmp[threadIdx.x >> 6][threadIdx.x & 63] = v;
    __syncthreads();
    if (threadIdx.x < NR) {
        double t = 0.0;
        for (int w = 0; w < MB_NW; ++w) t += s_tmp[w][threadIdx.x];
        dst_k[threadIdx.x] = t;
    }
    __syncthreads();
}

__device__ inline void mb_range(int n, int &i0, int &i1, int *range = nullptr)
{
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int per = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    i0 = min(b * per, n);
    i1 = min(i0 + per, n);
    if (range) *range = b;
}

// per-system workgroup sums of NV values per thread (thread's system =
// threadIdx.x % NR) into dst[q * NR + k]; one fixed order
template <int NR, int NV>
__device__ inline void mb_wg_sum_n(double (&v)[NV], double (*s_tmp)[NV][MB_NRMAX], double *dst)
{
#pragma unroll
    for (int q = 0; q < NV; ++q)
        for (int off = 32; off >= NR; off >>= 1) v[q] += __shfl_xor(v[q], off);
    if ((threadIdx.x & 63) < NR)
#pragma unroll
        for (int q = 0; q < NV; ++q) s_tmp[threadIdx.x >> 6][q][threadIdx.x & 63] = v[q];
    __syncthreads();
    if ((int)threadIdx.x < NV * NR) {
        const int q = threadIdx.x / NR, k = threadIdx.x % NR;
        double t = 0.0;
        for (int w = 0; w < MB_NW; ++w) t += s_tmp[w][q][k];
        dst[threadIdx.x] = t;
    }
    __syncthreads();
}

// Z^T r contribution of one node (3 translations, 3 rotations about the
// aggregate centroid; d = node position minus centroid)
__device__ inline void mb_ztr(const float *__restrict__ d, int i, const double r[3], double t[6])
{
    const double dx = d[3 * i], dy = d[3 * i + 1], dz = d[3 * i + 2];
    t[0] = r[0];
    t[1] = r[1];
    t[2] = r[2];
    t[3] = dy * r[2] - dz * r[1];
    t[4] = dz * r[0] - dx * r[2];
    t[5] = dx * r[1] - dy * r[0];
}

// Iteration start: per-system status from the previous update, then
// p_k = z_k + beta_k p_k (stopped systems: p = 0), q_k = K p_k, <p_k, q_k>.
template <int NR, int G>
__global__ __launch_bounds__(MB_WG) void k_mb_spmv(int bi, int nfree, const int *__restrict__ node_blk,
                                                   const int *__restrict__ bcol, const double *const *valp,
                                                   const double *__restrict__ Z, double *__restrict__ PA,
                                                   double *__restrict__ PB, double *__restrict__ Q, double *part,
                                                   MScalars *ms)
{
    static_assert(G % NR == 0 && 64 % G == 0, "lane group must hold whole systems and divide a wave");
    __shared__ double s_sc[3 * NR];
    __shared__ double s_beta[NR];
    __shared__ int s_act[NR + 1];
    __shared__ double s_tmp[MB_NW][MB_NRMAX];
    __shared__ double s_dot[NR];
    if (threadIdx.x == 0) s_act[NR] = ms->all_done;
    __syncthreads();
    if (s_act[NR]) return;
    const int par = ms->par0 ^ (bi & 1);
    mb_sum_partials(3 * NR, [&](int q) -> const double * {
        return q < NR ? mb_rz(part, par, q) : (q < 2 * NR ? mb_rz(part, par ^ 1, q - NR) : mb_rr(part, q - 2 * NR));
    }, s_sc);
    if (threadIdx.x < NR) {
        const int k = threadIdx.x;
        const double rz = s_sc[k], rzp = s_sc[NR + k], rr = s_sc[2 * NR + k];
        const int it = ms->it0 + bi;
        int dn = ms->done[k];
        if (!dn) {
            if (!isfinite(rz) || !isfinite(rr)) dn = 3;
            else if (rr <= ms->tol2[k]) dn = 1;
            else if (it >= ms->maxit) dn = 2;
            if (dn && blockIdx.x == 0) {
                ms->done[k] = dn;
                ms->it[k] = it;
                ms->rr[k] = rr;
            }
        }
        s_act[k] = !dn;
        s_beta[k] = dn ? 0.0 : rz / rzp;
    }
    __syncthreads();
    bool any = false;
#pragma unroll
    for (int k = 0; k < NR; ++k) any = any || s_act[k];
    if (!any) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ms->all_done = 1;
        return;
    }
    constexpr int GJ = G / NR, NPW = MB_WG / G;
    const int lane = threadIdx.x % G, k = lane % NR, jl = lane / NR;
    const bool act = s_act[k];
    const double beta = s_beta[k];
    const double *val = *valp;
    const double *p_old = par ? PB : PA;
    double *p_new = par ? PA : PB;
    int r0, r1;
    mb_range(nfree, r0, r1);
    double dot = 0.0;
    for (int r = r0 + (int)threadIdx.x / G; r < r1; r += NPW) {
        const int nb0 = node_blk[r], len = 3 * (node_blk[r + 1] - nb0);
        const double *v0 = val + 9 * (long long)nb0;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        // MB_UNR entries per pass with every load of the pass issued first
        // (absent entries read entry jl and add nothing)
        for (int jb = jl; jb < len; jb += MB_UNR * GJ) {
            int col[MB_UNR];
#pragma unroll
            for (int u = 0; u < MB_UNR; ++u) {
                const int j = jb + u * GJ < len ? jb + u * GJ : jl;
                const int kk = j / 3;
                col[u] = 3 * bcol[nb0 + kk] + (j - 3 * kk);
            }
            double zv[MB_UNR], pv[MB_UNR], m0[MB_UNR], m1[MB_UNR], m2[MB_UNR];
#pragma unroll
            for (int u = 0; u < MB_UNR; ++u) {
                const int j = jb + u * GJ < len ? jb + u * GJ : jl;
                const long long xi = (long long)col[u] * NR + k;
                zv[u] = Z[xi];
                pv[u] = p_old[xi];
                m0[u] = v0[j];
                m1[u] = v0[len + j];
                m2[u] = v0[2 * len + j];
            }
#pragma unroll
            for (int u = 0; u < MB_UNR; ++u) {
                const double pe = (act && jb + u * GJ < len) ? zv[u] + beta * pv[u] : 0.0;
                a0 += m0[u] * pe;
                a1 += m1[u] * pe;
                a2 += m2[u] * pe;
            }
        }
        for (int off = G >> 1; off >= NR; off >>= 1) {
            a0 += __shfl_xor(a0, off);
            a1 += __shfl_xor(a1, off);
            a2 += __shfl_xor(a2, off);
        }
        if (jl == 0) {
            const double av[3] = {a0, a1, a2};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const long long oi = (long long)(3 * r + c) * NR + k;
                const double pn = act ? Z[oi] + beta * p_old[oi] : 0.0;
                p_new[oi] = pn;
                Q[oi] = av[c];
                dot += av[c] * pn;
            }
        }
    }
    mb_wg_sum<NR>(dot, s_tmp, s_dot);
    if (threadIdx.x < NR) mb_pq(part, threadIdx.x)[blockIdx.x] = s_dot[threadIdx.x];
}

// alpha_k = <r,z>/<p,q>; x += alpha p; r -= alpha q; partials of <r,r>.
// Block Jacobi (CO false): z = M r (FP32 blocks) and partials of <r,z>.
// Two-level (CO true): per-range partials of Z^T r instead; k_mb_zapply
// forms z after the coarse solve.  A finished system keeps x and r and gets
// z = 0, so the next p is 0.  Thread (node slot, system k = tid % NR).
template <int NR, bool CO>
__global__ __launch_bounds__(MB_WG) void k_mb_update(int bi, int nfree, double *__restrict__ X, double *__restrict__ R,
                                                     double *__restrict__ Z, const double *__restrict__ PA,
                                                     const double *__restrict__ PB, const double *__restrict__ Q,
                                                     const float *__restrict__ minv, const float *__restrict__ cd,
                                                     double *__restrict__ part_co, double *part, MScalars *ms)
{
    __shared__ double s_sc[2 * NR];
    __shared__ double s_alpha[NR];
    __shared__ int s_act[NR];
    __shared__ int s_stop;
    __shared__ double s_tmp[MB_NW][6][MB_NRMAX];
    __shared__ double s_out[6 * NR];
    if (threadIdx.x == 0) s_stop = ms->all_done;
    __syncthreads();
    if (s_stop) return;
    const int par = ms->par0 ^ (bi & 1);
    mb_sum_partials(2 * NR, [&](int q) -> const double * { return q < NR ? mb_pq(part, q) : mb_rz(part, par, q - NR); },
                    s_sc);
    if (threadIdx.x < NR) {
        const int k = threadIdx.x;
        bool act = !ms->done[k];
        const double pq = s_sc[k];
        double alpha = 0.0;
        if (act) {
            if (!(pq > 0.0) || !isfinite(pq)) {
                if (blockIdx.x == 0) ms->done[k] = 3;
                act = false;
            } else {
                alpha = s_sc[NR + k] / pq;
            }
        }
        s_alpha[k] = alpha;
        s_act[k] = act;
    }
    __syncthreads();
    const int k = threadIdx.x % NR;
    const double alpha = s_alpha[k];
    const bool act = s_act[k];
    const double *P = par ? PA : PB;
    int i0, i1, rg;
    mb_range(nfree, i0, i1, &rg);
    double rz = 0.0, rr = 0.0;
    double g[6] = {0, 0, 0, 0, 0, 0};
    for (int i = i0 + (int)threadIdx.x / NR; i < i1; i += MB_WG / NR) {
        double rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long o = (long long)(3 * i + c) * NR + k;
            rv[c] = R[o];
            if (act) {
                X[o] += alpha * P[o];
                rv[c] -= alpha * Q[o];
                R[o] = rv[c];
            }
            rr += rv[c] * rv[c];
        }
        if constexpr (CO) {
            double t[6];
            mb_ztr(cd, i, rv, t);
#pragma unroll
            for (int m = 0; m < 6; ++m) g[m] += t[m];
        } else {
            const float *m = minv + 9 * i;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double zv = (double)m[3 * c] * rv[0] + (double)m[3 * c + 1] * rv[1] + (double)m[3 * c + 2] * rv[2];
                zv = act ? zv : 0.0;
                Z[(long long)(3 * i + c) * NR + k] = zv;
                rz += rv[c] * zv;
            }
        }
    }
    if constexpr (CO) {
        double v[6] = {g[0], g[1], g[2], g[3], g[4], g[5]};
        mb_wg_sum_n<NR, 6>(v, s_tmp, s_out);
        if ((int)threadIdx.x < 6 * NR) part_co[(size_t)rg * 6 * MB_NRMAX + threadIdx.x] = s_out[threadIdx.x];
        double v1[1] = {rr};
        mb_wg_sum_n<NR, 1>(v1, reinterpret_cast<double (*)[1][MB_NRMAX]>(&s_tmp[0][0][0]), s_out);
        if (threadIdx.x < NR) mb_rr(part, threadIdx.x)[blockIdx.x] = s_out[threadIdx.x];
    } else {
        double v[2] = {rz, rr};
        mb_wg_sum_n<NR, 2>(v, reinterpret_cast<double (*)[2][MB_NRMAX]>(&s_tmp[0][0][0]), s_out);
        if (threadIdx.x < NR) {
            mb_rz(part, par ^ 1, threadIdx.x)[blockIdx.x] = s_out[threadIdx.x];
            mb_rr(part, threadIdx.x)[blockIdx.x] = s_out[NR + threadIdx.x];
        }
    }
}

// Two-level: Z^T r per aggregate from the per-range partials, in range order
// (thread per (aggregate, mode, system)): g[a][m][k]
__global__ void k_mb_cred(int nagg, int ra, int nranges, int nr, const double *__restrict__ part_co,
                          double *__restrict__ gco, const MScalars *ms)
{
    if (ms->all_done) return;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nagg * 6 * nr) return;
    const int a = t / (6 * nr), mk = t - a * 6 * nr, m = mk / nr, k = mk - m * nr;
    double s = 0.0;
    for (int b = a * ra; b < min((a + 1) * ra, nranges); ++b) s += part_co[(size_t)b * 6 * MB_NRMAX + m * nr + k];
    gco[((size_t)a * 6 + m) * MB_NRMAX + k] = s;
}

// y = A_c^-1 g for every system: workgroup per CY_ROWS rows of A_c^-1 (FP32),
// thread (system k, column slice); column sums in a fixed order
constexpr int CY_ROWS = 6;
template <int NR>
__global__ __launch_bounds__(256) void k_mb_cy(int nc, const float *__restrict__ ainv, const double *__restrict__ gco,
                                               double *__restrict__ yco, const MScalars *ms)
{
    constexpr int NS = 256 / NR;  // column slices
    __shared__ double s_acc[NS][CY_ROWS][NR];
    if (ms->all_done) return;
    const int row0 = blockIdx.x * CY_ROWS;
    const int k = threadIdx.x % NR, sl = threadIdx.x / NR;
    double acc[CY_ROWS];
#pragma unroll
    for (int r = 0; r < CY_ROWS; ++r) acc[r] = 0.0;
    // 4 columns per pass, loads first (absent columns read column sl and add nothing)
    for (int cb = sl; cb < nc; cb += 4 * NS) {
        double gv[4];
        float av[4][CY_ROWS];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int col = cb + u * NS < nc ? cb + u * NS : sl;
            gv[u] = cb + u * NS < nc ? gco[(size_t)col * MB_NRMAX + k] : 0.0;
#pragma unroll
            for (int r = 0; r < CY_ROWS; ++r) av[u][r] = ainv[(size_t)min(row0 + r, nc - 1) * nc + col];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < CY_ROWS; ++r) acc[r] += (double)av[u][r] * gv[u];
    }
#pragma unroll
    for (int r = 0; r < CY_ROWS; ++r) s_acc[sl][r][k] = acc[r];
    __syncthreads();
    if ((int)threadIdx.x < CY_ROWS * NR) {
        const int r = threadIdx.x / NR, kk = threadIdx.x % NR;
        double t = 0.0;
        for (int j = 0; j < NS; ++j) t += s_acc[j][r][kk];
        if (row0 + r < nc) yco[(size_t)(row0 + r) * MB_NRMAX + kk] = t;
    }
}

// Two-level: z = M r + Z y_a for the range's aggregate a; partials of <r,z>
// into parity slot (par0 ^ (bi & 1)) ^ 1, or at the start (bi < 0) into
// slot par0 with slot par0 ^ 1 summing to +inf (the first beta is 0).
// Finished systems get z = 0.
template <int NR>
__global__ __launch_bounds__(MB_WG) void k_mb_zapply(int bi, int nfree, int ra, const double *__restrict__ R,
                                                     double *__restrict__ Z, const float *__restrict__ minv,
                                                     const float *__restrict__ cd, const double *__restrict__ yco,
                                                     double *part, const MScalars *ms)
{
    __shared__ double s_y[6][NR];
    __shared__ int s_act[NR];
    __shared__ double s_tmp[MB_NW][1][MB_NRMAX];
    __shared__ double s_out[NR];
    if (ms->all_done) return;
    int i0, i1, rg;
    mb_range(nfree, i0, i1, &rg);
    const int a = rg / ra;
    if ((int)threadIdx.x < 6 * NR) s_y[threadIdx.x / NR][threadIdx.x % NR] = yco[((size_t)a * 6 + threadIdx.x / NR) * MB_NRMAX + threadIdx.x % NR];
    if (threadIdx.x < NR) s_act[threadIdx.x] = !ms->done[threadIdx.x];
    __syncthreads();
    const int k = threadIdx.x % NR;
    const bool act = s_act[k];
    const double y0 = s_y[0][k], y1 = s_y[1][k], y2 = s_y[2][k], y3 = s_y[3][k], y4 = s_y[4][k], y5 = s_y[5][k];
    double rz = 0.0;
    for (int i = i0 + (int)threadIdx.x / NR; i < i1; i += MB_WG / NR) {
        double rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) rv[c] = R[(long long)(3 * i + c) * NR + k];
        const float *m = minv + 9 * i;
        const double dx = cd[3 * i], dy = cd[3 * i + 1], dz = cd[3 * i + 2];
        const double zc[3] = {y0 + y4 * dz - y5 * dy, y1 + y5 * dx - y3 * dz, y2 + y3 * dy - y4 * dx};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double zv = (double)m[3 * c] * rv[0] + (double)m[3 * c + 1] * rv[1] + (double)m[3 * c + 2] * rv[2] + zc[c];
            zv = act ? zv : 0.0;
            Z[(long long)(3 * i + c) * NR + k] = zv;
            rz += rv[c] * zv;
        }
    }
    double v[1] = {rz};
    mb_wg_sum_n<NR, 1>(v, s_tmp, s_out);
    if (threadIdx.x < NR) {
        const int p0 = ms->par0;
        if (bi < 0) {
            mb_rz(part, p0, threadIdx.x)[blockIdx.x] = s_out[threadIdx.x];
            mb_rz(part, p0 ^ 1, threadIdx.x)[blockIdx.x] = blockIdx.x == 0 ? INFINITY : 0.0;
        } else {
            mb_rz(part, (p0 ^ (bi & 1)) ^ 1, threadIdx.x)[blockIdx.x] = s_out[threadIdx.x];
        }
    }
}

__global__ void k_mb_batch_end(MScalars *ms, int nbatch)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && !ms->all_done) ms->it0 += nbatch;
}

// start: x = 0, p = 0, r = b, partials of <r,r>; block Jacobi: z = M r and
// partials <r,z> (slot par0; slot par0^1 sums to +inf so the first beta is 0);
// two-level: per-range partials of Z^T r (k_mb_zapply forms z).
template <int NR, bool CO>
__global__ __launch_bounds__(MB_WG) void k_mb_start(int nfree, const double *__restrict__ B, double *__restrict__ X,
                                                    double *__restrict__ R, double *__restrict__ Z,
                                                    double *__restrict__ PA, double *__restrict__ PB,
                                                    const float *__restrict__ minv, const float *__restrict__ cd,
                                                    double *__restrict__ part_co, double *part, const MScalars *ms)
{
    __shared__ double s_tmp[MB_NW][6][MB_NRMAX];
    __shared__ double s_out[6 * NR];
    const int par = ms->par0;
    const int k = threadIdx.x % NR;
    int i0, i1, rg;
    mb_range(nfree, i0, i1, &rg);
    double rz = 0.0, rr = 0.0;
    double g[6] = {0, 0, 0, 0, 0, 0};
    for (int i = i0 + (int)threadIdx.x / NR; i < i1; i += MB_WG / NR) {
        double rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long o = (long long)(3 * i + c) * NR + k;
            rv[c] = B[o];
            R[o] = rv[c];
            X[o] = 0.0;
            PA[o] = 0.0;
            PB[o] = 0.0;
            rr += rv[c] * rv[c];
        }
        if constexpr (CO) {
            double t[6];
            mb_ztr(cd, i, rv, t);
#pragma unroll
            for (int m = 0; m < 6; ++m) g[m] += t[m];
        } else {
            const float *m = minv + 9 * i;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double zv = (double)m[3 * c] * rv[0] + (double)m[3 * c + 1] * rv[1] + (double)m[3 * c + 2] * rv[2];
                Z[(long long)(3 * i + c) * NR + k] = zv;
                rz += rv[c] * zv;
            }
        }
    }
    if constexpr (CO) {
        double v[6] = {g[0], g[1], g[2], g[3], g[4], g[5]};
        mb_wg_sum_n<NR, 6>(v, s_tmp, s_out);
        if ((int)threadIdx.x < 6 * NR) part_co[(size_t)rg * 6 * MB_NRMAX + threadIdx.x] = s_out[threadIdx.x];
        double v1[1] = {rr};
        mb_wg_sum_n<NR, 1>(v1, reinterpret_cast<double (*)[1][MB_NRMAX]>(&s_tmp[0][0][0]), s_out);
        if (threadIdx.x < NR) mb_rr(part, threadIdx.x)[blockIdx.x] = s_out[threadIdx.x];
    } else {
        double v[2] = {rz, rr};
        mb_wg_sum_n<NR, 2>(v, reinterpret_cast<double (*)[2][MB_NRMAX]>(&s_tmp[0][0][0]), s_out);
        if (threadIdx.x < NR) {
            mb_rz(part, par, threadIdx.x)[blockIdx.x] = s_out[threadIdx.x];
            mb_rr(part, threadIdx.x)[blockIdx.x] = s_out[NR + threadIdx.x];
            mb_rz(part, par ^ 1, threadIdx.x)[blockIdx.x] = blockIdx.x == 0 ? INFINITY : 0.0;
        }
    }
}

// Y = K_v X for NR interleaved vectors (true residuals and tile timing)
template <int NR, int G>
__global__ __launch_bounds__(MB_WG) void k_mb_apply(int nfree, const int *__restrict__ node_blk,
                                                    const int *__restrict__ bcol, const double *const *valp,
                                                    const double *__restrict__ Xv, double *__restrict__ Yv)
{
    static_assert(G % NR == 0 && 64 % G == 0, "lane group must hold whole systems and divide a wave");
    constexpr int GJ = G / NR, NPW = MB_WG / G;
    const int lane = threadIdx.x % G, k = lane % NR, jl = lane / NR;
    const double *val = *valp;
    int r0, r1;
    mb_range(nfree, r0, r1);
    for (int r = r0 + (int)threadIdx.x / G; r < r1; r += NPW) {
        const int nb0 = node_blk[r], len = 3 * (node_blk[r + 1] - nb0);
        const double *v0 = val + 9 * (long long)nb0;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int j = jl; j < len; j += GJ) {
            const int kk = j / 3, d = j - 3 * kk;
            const double xv = Xv[(long long)(3 * bcol[nb0 + kk] + d) * NR + k];
            a0 += v0[j] * xv;
            a1 += v0[len + j] * xv;
            a2 += v0[2 * len + j] * xv;
        }
        for (int off = G >> 1; off >= NR; off >>= 1) {
            a0 += __shfl_xor(a0, off);
            a1 += __shfl_xor(a1, off);
            a2 += __shfl_xor(a2, off);
        }
        if (jl == 0) {
            Yv[(long long)(3 * r) * NR + k] = a0;
            Yv[(long long)(3 * r + 1) * NR + k] = a1;
            Yv[(long long)(3 * r + 2) * NR + k] = a2;
        }
    }
}

// 3x3 diagonal blocks of one variant, inverted in FP64 and stored in FP32 (thread per node)
__global__ void k_mb_precond(int nfree, const int *__restrict__ node_blk, const int *__restrict__ bcol,
                             const double *const *valp, float *__restrict__ minv)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nfree) return;
    const double *val = *valp;
    const int nb0 = node_blk[r], deg = node_blk[r + 1] - nb0;
    int k = 0;
    while (k < deg && bcol[nb0 + k] != r) ++k;
    double A[9];
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) A[3 * c + d] = k < deg ? val[9 * (long long)nb0 + c * 3 * deg + 3 * k + d] : 0.0;
    double M[9];
    inv3(A, M);
    for (int i = 0; i < 9; ++i) minv[9 * r + i] = (float)M[i];
}

constexpr int MB_GROUPS[3] = {16, 32, 64};

// The system the multi-RHS solver works on: K's pattern in the caller's node
// order (block Jacobi), or the two-level order with the variant's coarse
// space (the values then come from coarse_for_values).
struct MbSys {
    const int *node_blk = nullptr, *bcol = nullptr;
    const CoarseVariant *co = nullptr;
};

template <int NR>
static void mb_launch_apply(ofea_ctx *c, BatchWS *w, const MbSys &S, int group, const double *x, double *y)
{
    const int nf = c->P.nfree;
    switch (group) {
    case 32:
        hipLaunchKernelGGL((k_mb_apply<NR, 32>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, nf, S.node_blk, S.bcol,
                           w->valp, x, y);
        break;
    case 64:
        hipLaunchKernelGGL((k_mb_apply<NR, 64>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, nf, S.node_blk, S.bcol,
                           w->valp, x, y);
        break;
    default:
        hipLaunchKernelGGL((k_mb_apply<NR, 16>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, nf, S.node_blk, S.bcol,
                           w->valp, x, y);
    }
}

// two-level: Z^T r per aggregate, y = A_c^-1 Z^T r, z = M r + Z y (bi < 0: the start)
template <int NR>
static void mb_launch_coarse(ofea_ctx *c, BatchWS *w, const MbSys &S, int bi)
{
    const CoarseVariant &co = *S.co;
    const int nf = c->P.nfree, nc = 6 * co.nagg;
    const int nthr = co.nagg * 6 * NR;
    hipLaunchKernelGGL(k_mb_cred, dim3((nthr + 255) / 256), dim3(256), 0, c->stream, co.nagg, co.ra, MB_NB, NR,
                       (const double *)w->part_co, w->gco, (const MScalars *)w->ms);
    hipLaunchKernelGGL((k_mb_cy<NR>), dim3((nc + CY_ROWS - 1) / CY_ROWS), dim3(256), 0, c->stream, nc, co.ainv,
                       (const double *)w->gco, w->yco, (const MScalars *)w->ms);
    hipLaunchKernelGGL((k_mb_zapply<NR>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, bi, nf, co.ra, (const double *)w->R,
                       w->Z, (const float *)w->minv, co.d, (const double *)w->yco, w->part, (const MScalars *)w->ms);
}

template <int NR, int G>
static void mb_capture_batch(ofea_ctx *c, BatchWS *w, const MbSys &S)
{
    const int nf = c->P.nfree;
    const float *cd = S.co ? S.co->d : nullptr;
    for (int bi = 0; bi < MB_BATCH; ++bi) {
        hipLaunchKernelGGL((k_mb_spmv<NR, G>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, bi, nf, S.node_blk, S.bcol,
                           w->valp, w->Z, w->PA, w->PB, w->Q, w->part, (MScalars *)w->ms);
        if (S.co) {
            hipLaunchKernelGGL((k_mb_update<NR, true>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, bi, nf, w->X, w->R,
                               w->Z, w->PA, w->PB, w->Q, w->minv, cd, w->part_co, w->part, (MScalars *)w->ms);
            mb_launch_coarse<NR>(c, w, S, bi);
        } else {
            hipLaunchKernelGGL((k_mb_update<NR, false>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, bi, nf, w->X, w->R,
                               w->Z, w->PA, w->PB, w->Q, w->minv, cd, w->part_co, w->part, (MScalars *)w->ms);
        }
    }
    hipLaunchKernelGGL(k_mb_batch_end, dim3(1), dim3(64), 0, c->stream, (MScalars *)w->ms, MB_BATCH);
}

// Lane-group size of the multi-RHS SpMV for NR systems: each candidate is
// timed on the resident variant (5 launches after a warm-up) and the fastest
// is kept for this context and NR.
template <int NR>
static int mb_tune(ofea_ctx *c, BatchWS *w, const MbSys &S)
{
    if (w->mb_group[NR]) return OFEA_OK;
    float best_t = 1e30f;
    int best = 16;
    for (int g : MB_GROUPS) {
        mb_launch_apply<NR>(c, w, S, g, w->B, w->Y);
        OFEA_HIP(hipEventRecord(c->ev0, c->stream));
        for (int i = 0; i < 5; ++i) mb_launch_apply<NR>(c, w, S, g, w->B, w->Y);
        OFEA_HIP(hipEventRecord(c->ev1, c->stream));
        OFEA_HIP(hipEventSynchronize(c->ev1));
        OFEA_HIP(hipGetLastError());
        float t = 0.f;
        OFEA_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
        if (t < best_t) {
            best_t = t;
            best = g;
        }
    }
    w->mb_group[NR] = best;
    w->mb_time[NR] = best_t / 5.f;
    return OFEA_OK;
}

template <int NR>
static int mb_solve_variant(ofea_ctx *c, BatchWS *w, const MbSys &S, double rtol, int maxit,
                            const std::vector<double> &bb, int *iters, double *relres)
{
    const int nf = c->P.nfree;
    int rc = mb_tune<NR>(c, w, S);
    if (rc) return rc;
    const int group = w->mb_group[NR];
    MScalars &Sc = *(MScalars *)w->ms_host;
    std::memset(&Sc, 0, sizeof(Sc));
    for (int k = 0; k < NR; ++k) Sc.tol2[k] = 0.25 * rtol * rtol * bb[k];  // recursive target 0.5 rtol
    Sc.maxit = maxit;
    OFEA_HIP(hipMemcpyAsync(w->ms, &Sc, sizeof(Sc), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_mb_precond, dim3((nf + 255) / 256), dim3(256), 0, c->stream, nf, S.node_blk, S.bcol, w->valp,
                       w->minv);
    const float *cd = S.co ? S.co->d : nullptr;
    if (S.co) {
        hipLaunchKernelGGL((k_mb_start<NR, true>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, nf, w->B, w->X, w->R, w->Z,
                           w->PA, w->PB, w->minv, cd, w->part_co, w->part, (const MScalars *)w->ms);
        mb_launch_coarse<NR>(c, w, S, -1);
    } else {
        hipLaunchKernelGGL((k_mb_start<NR, false>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, nf, w->B, w->X, w->R,
                           w->Z, w->PA, w->PB, w->minv, cd, w->part_co, w->part, (const MScalars *)w->ms);
    }
    OFEA_HIP(hipGetLastError());
    auto direct_batch = [&]() {
        switch (group) {
        case 32: mb_capture_batch<NR, 32>(c, w, S); break;
        case 64: mb_capture_batch<NR, 64>(c, w, S); break;
        default: mb_capture_batch<NR, 16>(c, w, S); break;
        }
    };
    // OFEA_NO_GRAPH=1 launches the batches directly (for tools that cannot follow graph launches)
    static const bool no_graph = std::getenv("OFEA_NO_GRAPH") && std::getenv("OFEA_NO_GRAPH")[0] == '1';
    const int co_key = S.co ? 6 * S.co->nagg : 0;
    if (!no_graph && (!w->exec || w->graph_nrhs != NR || w->graph_group != group || w->graph_co != co_key ||
                      w->graph_key != (const void *)S.node_blk)) {
        if (w->exec) (void)hipGraphExecDestroy(w->exec);
        if (w->graph) (void)hipGraphDestroy(w->graph);
        w->exec = nullptr;
        w->graph = nullptr;
        OFEA_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        direct_batch();
        OFEA_HIP(hipStreamEndCapture(c->stream, &w->graph));
        OFEA_HIP(hipGraphInstantiate(&w->exec, w->graph, nullptr, nullptr, 0));
        w->graph_nrhs = NR;
        w->graph_group = group;
        w->graph_co = co_key;
        w->graph_key = S.node_blk;
    }
    for (;;) {
        if (no_graph) {
            direct_batch();
            OFEA_HIP(hipGetLastError());
        } else {
            OFEA_HIP(hipGraphLaunch(w->exec, c->stream));
        }
        OFEA_HIP(hipMemcpyAsync(&Sc, w->ms, sizeof(Sc), hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
        if (Sc.all_done) break;
    }
    // true FP64 residuals b - K x of every system
    mb_launch_apply<NR>(c, w, S, group, w->X, w->Y);
    OFEA_HIP(hipGetLastError());
    const size_t nn = (size_t)3 * nf * NR;
    std::vector<double> y(nn), b(nn);
    OFEA_HIP(hipMemcpyAsync(y.data(), w->Y, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    OFEA_HIP(hipMemcpyAsync(b.data(), w->B, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    OFEA_HIP(hipStreamSynchronize(c->stream));
    bool all_ok = true;
    for (int k = 0; k < NR; ++k) {
        double rr = 0.0;
        for (size_t i = 0; i < (size_t)3 * nf; ++i) {
            const double d = b[i * NR + k] - y[i * NR + k];
            rr += d * d;
        }
        const double rel = bb[k] > 0.0 ? std::sqrt(rr / bb[k]) : 0.0;
        if (relres) relres[k] = rel;
        if (iters) iters[k] = Sc.done[k] ? Sc.it[k] : maxit;
        if (Sc.done[k] == 3) return fail(OFEA_ERR_STATE, "PCG breakdown in a batched system");
        all_ok = all_ok && rel <= rtol;
    }
    return all_ok ? OFEA_OK : OFEA_NOT_CONVERGED;
}

int solve_batch(ofea_ctx *c, int nvar, int nrhs, const double *B, double *X, double rtol, int maxit, int precond,
                int *iters, double *relres)
{
    BatchWS *w = bws(c);
    if (!w || !w->valb || w->nvar < nvar) return fail(OFEA_ERR_STATE, "ofea_assemble_batch first (with >= nvar variants)");
    if (nrhs != 1 && nrhs != 2 && nrhs != 4 && nrhs != 8) return fail(OFEA_ERR_ARG, "nrhs must be 1, 2, 4 or 8");
    if (precond != OFEA_PCG_BLOCK_JACOBI && precond != OFEA_PCG_TWO_LEVEL)
        return fail(OFEA_ERR_ARG, "batched preconditioner must be OFEA_PCG_BLOCK_JACOBI or OFEA_PCG_TWO_LEVEL");
    const bool two = precond == OFEA_PCG_TWO_LEVEL;
    const int nf = c->P.nfree, n = c->P.n;
    const size_t nn = (size_t)n * nrhs;
    if (w->nrhs != nrhs) {
        for (double **p : {&w->B, &w->X, &w->R, &w->Z, &w->PA, &w->PB, &w->Q, &w->Y})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
        for (double **p : {&w->B, &w->X, &w->R, &w->Z, &w->PA, &w->PB, &w->Q, &w->Y})
            OFEA_HIP(hipMalloc((void **)p, sizeof(double) * std::max<size_t>(nn, 1)));
        w->nrhs = nrhs;
        if (w->exec) (void)hipGraphExecDestroy(w->exec);
        if (w->graph) (void)hipGraphDestroy(w->graph);
        w->exec = nullptr;
        w->graph = nullptr;
    }
    if (!w->minv) OFEA_HIP(hipMalloc((void **)&w->minv, sizeof(float) * 9 * (size_t)std::max(nf, 1)));
    if (!w->part) OFEA_HIP(hipMalloc((void **)&w->part, sizeof(double) * 4 * MB_NRMAX * MB_NB));
    if (!w->ms) OFEA_HIP(hipMalloc(&w->ms, sizeof(MScalars)));
    if (!w->ms_host) OFEA_HIP(hipHostMalloc(&w->ms_host, sizeof(MScalars), hipHostMallocDefault));
    if (!w->valp) OFEA_HIP(hipMalloc((void **)&w->valp, sizeof(double *)));
    if (two && !w->part_co) {
        OFEA_HIP(hipMalloc((void **)&w->part_co, sizeof(double) * 6 * MB_NRMAX * MB_NB));
        OFEA_HIP(hipMalloc((void **)&w->gco, sizeof(double) * CG_NC_MAX * MB_NRMAX));
        OFEA_HIP(hipMalloc((void **)&w->yco, sizeof(double) * CG_NC_MAX * MB_NRMAX));
    }
    std::vector<double> bi(nn), xo(nn), bb(nrhs);
    int status = OFEA_OK;
    CoarseVariant co;
    MbSys S;
    S.node_blk = c->d_node_blk;
    S.bcol = c->d_bcol;
    // A variant keeps the coarse inverse built for an earlier one until its
    // predecessor's slowest system took 25% more iterations than the slowest
    // of the first variant after the last build.
    int base_it = -1, last_it = 0;
    std::vector<int> itl(nrhs);
    for (int v = 0; v < nvar; ++v) {
        const double *pv = w->valb + (size_t)v * w->nnz;
        if (two) {
            // the variant's coarse space; the solve runs on its reordered copy
            const bool keep = base_it >= 0 && 4 * last_it <= 5 * base_it;
            if (!keep) base_it = -1;
            int rc = coarse_for_values(c, pv, co, keep);
            if (rc) return rc;
            if (6 * co.nagg > CG_NC_MAX) return fail(OFEA_ERR_LIMIT, "coarse space larger than the batched buffers");
            S.node_blk = co.nb_p;
            S.bcol = co.bcol_p;
            S.co = &co;
            pv = co.val_p;
        }
        OFEA_HIP(hipMemcpyAsync(w->valp, &pv, sizeof(pv), hipMemcpyHostToDevice, c->stream));
        // interleave the right-hand sides of variant v: [dof][system], nodes in
        // the order the solver works in
        for (int k = 0; k < nrhs; ++k) {
            const double *src = B + ((size_t)v * nrhs + k) * n;
            double s = 0.0;
            for (int i = 0; i < nf; ++i) {
                const int j = two ? co.iperm[i] : i;
                for (int d = 0; d < 3; ++d) {
                    const double x = src[3 * i + d];
                    bi[(size_t)(3 * j + d) * nrhs + k] = x;
                    s += x * x;
                }
            }
            bb[k] = s;
        }
        OFEA_HIP(hipMemcpyAsync(w->B, bi.data(), sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));  // pv and bi are pageable host memory reused next variant
        int rc;
        int *itv = iters ? iters + (size_t)v * nrhs : itl.data();
        double *rv = relres ? relres + (size_t)v * nrhs : nullptr;
        switch (nrhs) {
        case 1: rc = mb_solve_variant<1>(c, w, S, rtol, maxit, bb, itv, rv); break;
        case 2: rc = mb_solve_variant<2>(c, w, S, rtol, maxit, bb, itv, rv); break;
        case 4: rc = mb_solve_variant<4>(c, w, S, rtol, maxit, bb, itv, rv); break;
        default: rc = mb_solve_variant<8>(c, w, S, rtol, maxit, bb, itv, rv); break;
        }
        if (rc < 0) return rc;
        if (rc != OFEA_OK) status = rc;
        last_it = *std::max_element(itv, itv + nrhs);
        if (base_it < 0) base_it = last_it;
        if (X) {
            OFEA_HIP(hipMemcpyAsync(xo.data(), w->X, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
            OFEA_HIP(hipStreamSynchronize(c->stream));
            for (int k = 0; k < nrhs; ++k) {
                double *dst = X + ((size_t)v * nrhs + k) * n;
                for (int i = 0; i < nf; ++i) {
                    const int j = two ? co.iperm[i] : i;
                    for (int d = 0; d < 3; ++d) dst[3 * i + d] = xo[(size_t)(3 * j + d) * nrhs + k];
                }
            }
        }
    }
    return status;
}

int batch_tiles(ofea_ctx *c, int nrhs, int *group, float *ms)
{
    BatchWS *w = bws(c);
    if (!w || nrhs < 1 || nrhs > MB_NRMAX || !w->mb_group[nrhs]) return fail(OFEA_ERR_STATE, "no batched solve with this nrhs yet");
    if (group) *group = w->mb_group[nrhs];
    if (ms) *ms = w->mb_time[nrhs];
    return OFEA_OK;
}

// ------------------------------------------------------------------ autotuner
int solver_spmv_group_time(ofea_ctx *c, int group, int reps, float *ms);

int autotune(ofea_ctx *c, int *best_group, float *times)
{
    if (!c->assembled) return fail(OFEA_ERR_STATE, "ofea_assemble first");
    const int groups[5] = {4, 8, 16, 32, 64};
    int best = 16;
    float best_t = 1e30f;
    for (int i = 0; i < 5; ++i) {
        float t = 0.f;
        int rc = solver_spmv_group_time(c, groups[i], 20, &t);
        if (rc) return rc;
        if (times) times[i] = t;
        if (t < best_t) {
            best_t = t;
            best = groups[i];
        }
    }
    // the CG kernels instantiate 4..64; the multi-RHS kernels 8, 16, 32
    c->ws->spmv_group = best;
    solver_drop_graphs(c);
    if (best_group) *best_group = best;
    return OFEA_OK;
}

}  // namespace ofea
