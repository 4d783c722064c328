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
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_nrhs = 0, graph_group = 0;
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
                    (void *)w->part, w->ms, (void *)w->valp})
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

__device__ inline void mb_range(int n, int &i0, int &i1)
{
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int per = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    i0 = min(b * per, n);
    i1 = min(i0 + per, n);
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
        for (int j = jl; j < len; j += GJ) {
            const int kk = j / 3, d = j - 3 * kk;
            const long long xi = (long long)(3 * bcol[nb0 + kk] + d) * NR + k;
            const double pv = act ? Z[xi] + beta * p_old[xi] : 0.0;
            a0 += v0[j] * pv;
            a1 += v0[len + j] * pv;
            a2 += v0[2 * len + j] * pv;
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

// alpha_k = <r,z>/<p,q>; x += alpha p; r -= alpha q; z = M r (FP32 blocks);
// partials of <r,z> and <r,r>.  Thread (node slot, system k = tid % NR).
template <int NR>
__global__ __launch_bounds__(MB_WG) void k_mb_update(int bi, int nfree, double *__restrict__ X, double *__restrict__ R,
                                                     double *__restrict__ Z, const double *__restrict__ PA,
                                                     const double *__restrict__ PB, const double *__restrict__ Q,
                                                     const float *__restrict__ minv, double *part, MScalars *ms)
{
    __shared__ double s_sc[2 * NR];
    __shared__ double s_alpha[NR];
    __shared__ int s_stop;
    __shared__ double s_tmp[MB_NW][MB_NRMAX];
    __shared__ double s_out[2 * NR];
    if (threadIdx.x == 0) s_stop = ms->all_done;
    __syncthreads();
    if (s_stop) return;
    const int par = ms->par0 ^ (bi & 1);
    mb_sum_partials(2 * NR, [&](int q) -> const double * { return q < NR ? mb_pq(part, q) : mb_rz(part, par, q - NR); },
                    s_sc);
    if (threadIdx.x < NR) {
        const int k = threadIdx.x;
        const bool act = !ms->done[k];
        const double pq = s_sc[k];
        double alpha = 0.0;
        if (act) {
            if (!(pq > 0.0) || !isfinite(pq)) {
                if (blockIdx.x == 0) ms->done[k] = 3;
            } else {
                alpha = s_sc[NR + k] / pq;
            }
        }
        s_alpha[k] = alpha;
    }
    __syncthreads();
    const int k = threadIdx.x % NR;
    const double alpha = s_alpha[k];
    const double *P = par ? PA : PB;
    int i0, i1;
    mb_range(nfree, i0, i1);
    double rz = 0.0, rr = 0.0;
    for (int i = i0 + (int)threadIdx.x / NR; i < i1; i += MB_WG / NR) {
        const float *m = minv + 9 * i;
        double rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long o = (long long)(3 * i + c) * NR + k;
            X[o] += alpha * P[o];
            rv[c] = R[o] - alpha * Q[o];
            R[o] = rv[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double zv = (double)m[3 * c] * rv[0] + (double)m[3 * c + 1] * rv[1] + (double)m[3 * c + 2] * rv[2];
            Z[(long long)(3 * i + c) * NR + k] = zv;
            rz += rv[c] * zv;
            rr += rv[c] * rv[c];
        }
    }
    mb_wg_sum<NR>(rz, s_tmp, s_out);
    mb_wg_sum<NR>(rr, s_tmp, s_out + NR);
    if (threadIdx.x < NR) {
        mb_rz(part, par ^ 1, threadIdx.x)[blockIdx.x] = s_out[threadIdx.x];
        mb_rr(part, threadIdx.x)[blockIdx.x] = s_out[NR + threadIdx.x];
    }
}

__global__ void k_mb_batch_end(MScalars *ms, int nbatch)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && !ms->all_done) ms->it0 += nbatch;
}

// start: x = 0, p = 0, r = b, z = M r, partials <r,z> (slot par0) and <r,r>;
// slot par0^1 sums to +inf so the first beta is 0.
template <int NR>
__global__ __launch_bounds__(MB_WG) void k_mb_start(int nfree, const double *__restrict__ B, double *__restrict__ X,
                                                    double *__restrict__ R, double *__restrict__ Z,
                                                    double *__restrict__ PA, double *__restrict__ PB,
                                                    const float *__restrict__ minv, double *part, const MScalars *ms)
{
    __shared__ double s_tmp[MB_NW][MB_NRMAX];
    __shared__ double s_out[2 * NR];
    const int par = ms->par0;
    const int k = threadIdx.x % NR;
    int i0, i1;
    mb_range(nfree, i0, i1);
    double rz = 0.0, rr = 0.0;
    for (int i = i0 + (int)threadIdx.x / NR; i < i1; i += MB_WG / NR) {
        const float *m = minv + 9 * i;
        double rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long o = (long long)(3 * i + c) * NR + k;
            rv[c] = B[o];
            R[o] = rv[c];
            X[o] = 0.0;
            PA[o] = 0.0;
            PB[o] = 0.0;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double zv = (double)m[3 * c] * rv[0] + (double)m[3 * c + 1] * rv[1] + (double)m[3 * c + 2] * rv[2];
            Z[(long long)(3 * i + c) * NR + k] = zv;
            rz += rv[c] * zv;
            rr += rv[c] * rv[c];
        }
    }
    mb_wg_sum<NR>(rz, s_tmp, s_out);
    mb_wg_sum<NR>(rr, s_tmp, s_out + NR);
    if (threadIdx.x < NR) {
        mb_rz(part, par, threadIdx.x)[blockIdx.x] = s_out[threadIdx.x];
        mb_rr(part, threadIdx.x)[blockIdx.x] = s_out[NR + threadIdx.x];
        mb_rz(part, par ^ 1, threadIdx.x)[blockIdx.x] = blockIdx.x == 0 ? INFINITY : 0.0;
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

template <int NR>
static void mb_launch_apply(ofea_ctx *c, BatchWS *w, int group, const double *x, double *y)
{
    const int nf = c->P.nfree;
    switch (group) {
    case 32:
        hipLaunchKernelGGL((k_mb_apply<NR, 32>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, nf, c->d_node_blk, c->d_bcol,
                           w->valp, x, y);
        break;
    case 64:
        hipLaunchKernelGGL((k_mb_apply<NR, 64>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, nf, c->d_node_blk, c->d_bcol,
                           w->valp, x, y);
        break;
    default:
        hipLaunchKernelGGL((k_mb_apply<NR, 16>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, nf, c->d_node_blk, c->d_bcol,
                           w->valp, x, y);
    }
}

template <int NR, int G>
static void mb_capture_batch(ofea_ctx *c, BatchWS *w)
{
    const int nf = c->P.nfree;
    for (int bi = 0; bi < MB_BATCH; ++bi) {
        hipLaunchKernelGGL((k_mb_spmv<NR, G>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, bi, nf, c->d_node_blk, c->d_bcol,
                           w->valp, w->Z, w->PA, w->PB, w->Q, w->part, (MScalars *)w->ms);
        hipLaunchKernelGGL((k_mb_update<NR>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, bi, nf, w->X, w->R, w->Z, w->PA,
                           w->PB, w->Q, w->minv, w->part, (MScalars *)w->ms);
    }
    hipLaunchKernelGGL(k_mb_batch_end, dim3(1), dim3(64), 0, c->stream, (MScalars *)w->ms, MB_BATCH);
}

// Lane-group size of the multi-RHS SpMV for NR systems: each candidate is
// timed on the resident variant (5 launches after a warm-up) and the fastest
// is kept for this context and NR.
template <int NR>
static int mb_tune(ofea_ctx *c, BatchWS *w)
{
    if (w->mb_group[NR]) return OFEA_OK;
    float best_t = 1e30f;
    int best = 16;
    for (int g : MB_GROUPS) {
        mb_launch_apply<NR>(c, w, g, w->B, w->Y);
        OFEA_HIP(hipEventRecord(c->ev0, c->stream));
        for (int i = 0; i < 5; ++i) mb_launch_apply<NR>(c, w, g, w->B, w->Y);
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
static int mb_solve_variant(ofea_ctx *c, BatchWS *w, double rtol, int maxit, const std::vector<double> &bb, int *iters,
                            double *relres)
{
    const int nf = c->P.nfree;
    int rc = mb_tune<NR>(c, w);
    if (rc) return rc;
    const int group = w->mb_group[NR];
    MScalars &S = *(MScalars *)w->ms_host;
    std::memset(&S, 0, sizeof(S));
    for (int k = 0; k < NR; ++k) S.tol2[k] = 0.25 * rtol * rtol * bb[k];  // recursive target 0.5 rtol
    S.maxit = maxit;
    OFEA_HIP(hipMemcpyAsync(w->ms, &S, sizeof(S), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_mb_precond, dim3((nf + 255) / 256), dim3(256), 0, c->stream, nf, c->d_node_blk, c->d_bcol,
                       w->valp, w->minv);
    hipLaunchKernelGGL((k_mb_start<NR>), dim3(MB_NB), dim3(MB_WG), 0, c->stream, nf, w->B, w->X, w->R, w->Z, w->PA,
                       w->PB, w->minv, w->part, (const MScalars *)w->ms);
    OFEA_HIP(hipGetLastError());
    auto direct_batch = [&]() {
        switch (group) {
        case 32: mb_capture_batch<NR, 32>(c, w); break;
        case 64: mb_capture_batch<NR, 64>(c, w); break;
        default: mb_capture_batch<NR, 16>(c, w); break;
        }
    };
    // OFEA_NO_GRAPH=1 launches the batches directly (for tools that cannot follow graph launches)
    static const bool no_graph = std::getenv("OFEA_NO_GRAPH") && std::getenv("OFEA_NO_GRAPH")[0] == '1';
    if (!no_graph && (!w->exec || w->graph_nrhs != NR || w->graph_group != group)) {
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
    }
    for (;;) {
        if (no_graph) {
            direct_batch();
            OFEA_HIP(hipGetLastError());
        } else {
            OFEA_HIP(hipGraphLaunch(w->exec, c->stream));
        }
        OFEA_HIP(hipMemcpyAsync(&S, w->ms, sizeof(S), hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
        if (S.all_done) break;
    }
    // true FP64 residuals b - K x of every system
    mb_launch_apply<NR>(c, w, group, w->X, w->Y);
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
        if (iters) iters[k] = S.done[k] ? S.it[k] : maxit;
        if (S.done[k] == 3) return fail(OFEA_ERR_STATE, "PCG breakdown in a batched system");
        all_ok = all_ok && rel <= rtol;
    }
    return all_ok ? OFEA_OK : OFEA_NOT_CONVERGED;
}

int solve_batch(ofea_ctx *c, int nvar, int nrhs, const double *B, double *X, double rtol, int maxit, int *iters,
                double *relres)
{
    BatchWS *w = bws(c);
    if (!w || !w->valb || w->nvar < nvar) return fail(OFEA_ERR_STATE, "ofea_assemble_batch first (with >= nvar variants)");
    if (nrhs != 1 && nrhs != 2 && nrhs != 4 && nrhs != 8) return fail(OFEA_ERR_ARG, "nrhs must be 1, 2, 4 or 8");
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
    std::vector<double> bi(nn), xo(nn), bb(nrhs);
    int status = OFEA_OK;
    for (int v = 0; v < nvar; ++v) {
        const double *pv = w->valb + (size_t)v * w->nnz;
        OFEA_HIP(hipMemcpyAsync(w->valp, &pv, sizeof(pv), hipMemcpyHostToDevice, c->stream));
        // interleave the right-hand sides of variant v: [dof][system]
        for (int k = 0; k < nrhs; ++k) {
            const double *src = B + ((size_t)v * nrhs + k) * n;
            double s = 0.0;
            for (int i = 0; i < n; ++i) {
                bi[(size_t)i * nrhs + k] = src[i];
                s += src[i] * src[i];
            }
            bb[k] = s;
        }
        OFEA_HIP(hipMemcpyAsync(w->B, bi.data(), sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));  // pv and bi are pageable host memory reused next variant
        int rc;
        int *itv = iters ? iters + (size_t)v * nrhs : nullptr;
        double *rv = relres ? relres + (size_t)v * nrhs : nullptr;
        switch (nrhs) {
        case 1: rc = mb_solve_variant<1>(c, w, rtol, maxit, bb, itv, rv); break;
        case 2: rc = mb_solve_variant<2>(c, w, rtol, maxit, bb, itv, rv); break;
        case 4: rc = mb_solve_variant<4>(c, w, rtol, maxit, bb, itv, rv); break;
        default: rc = mb_solve_variant<8>(c, w, rtol, maxit, bb, itv, rv); break;
        }
        if (rc < 0) return rc;
        if (rc != OFEA_OK) status = rc;
        if (X) {
            OFEA_HIP(hipMemcpyAsync(xo.data(), w->X, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
            OFEA_HIP(hipStreamSynchronize(c->stream));
            for (int k = 0; k < nrhs; ++k) {
                double *dst = X + ((size_t)v * nrhs + k) * n;
                for (int i = 0; i < n; ++i) dst[i] = xo[(size_t)i * nrhs + k];
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
