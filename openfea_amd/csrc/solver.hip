// solver.hip -- GPU preconditioned conjugate gradients beside the reference's
// direct solvers (solution_api, src/linear_solution.cpp:17-64).
//
// Matrix: the resident CSC values read node-block-wise.  The three columns of
// free node r are contiguous and share one row pattern (rows 3s+d of the
// sorted neighbours s), so y[3r+c] = sum_k,d val[col(3r+c) + 3k + d] x[3 s_k + d]
// needs the neighbour list s_k (bcol, one int per 3x3 block) instead of
// row_idx: 8.44 bytes per stored value in FP64, 4.44 in FP32.  Because K is
// symmetric this product is K x = K^T x.
//
// PCG iteration = 3 kernels (SpMV + <p,q>, update + <r,z>,<r,r>, direction),
// scalars live on the device and are finalised by the last workgroup of each
// reduction (fixed-order sum => deterministic).  Batches of iterations are
// captured in a hipGraph; kernels return at once after convergence, and the
// host reads the flag between graph launches.
#include <cmath>
#include <vector>

#include "ofea_internal.h"

namespace ofea {

constexpr int RED_WG = 256;   // threads per workgroup of vector kernels
constexpr int CG_BATCH = 16;  // iterations per graph launch

enum Scalar { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_TOL2, S_NSCAL };

struct DevScalars {
    double s[S_NSCAL];
    int done;  // 0 running, 1 converged, 2 maxit, 3 breakdown
    int it;
    int maxit;
    unsigned int ticket;
};

template <class T>
struct VecSet {
    T *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *x = nullptr;
};

}  // namespace ofea

struct ofea_solver_ws {
    int n = 0, nfree = 0;
    int spmv_group = 16;        // lanes per node row group (autotuned)
    int spmv_nodes_per_wg = 16; // = 256 / group
    int nblk_spmv = 0, nblk_vec = 0, nblk_node = 0;
    double *partials = nullptr;  // 3 * max blocks
    ofea::DevScalars *sc = nullptr;
    ofea::DevScalars *sc_host = nullptr;  // pinned
    ofea::VecSet<double> v64;
    ofea::VecSet<float> v32;
    double *r64 = nullptr;    // true residual for refinement
    double *dinv = nullptr;   // point Jacobi, n
    double *minv = nullptr;   // 3x3 block inverses, 9 * nfree
    float *minv32 = nullptr;  // FP32 copy of the preconditioner blocks
    float *val32 = nullptr;   // FP32 values, nnz
    bool val32_valid = false, precond_valid = false;
    hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace ofea {

__device__ inline int xcd_remap_s(int i, int n)
{
    int q = n >> 3, rem = n & 7, x = i & 7, j = i >> 3;
    return x < rem ? x * (q + 1) + j : rem * (q + 1) + (x - rem) * q + j;
}

template <class T>
__device__ inline T wave_group_sum(T v, int group)
{
    for (int off = group >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Block-wide sum of NV doubles; result valid in thread 0.
template <int NV>
__device__ inline void block_sum(double v[NV])
{
    __shared__ double sm[NV][RED_WG / 64];
    for (int i = 0; i < NV; ++i) {
        double t = v[i];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
        v[i] = t;
    }
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0)
        for (int i = 0; i < NV; ++i) sm[i][w] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; ++i) {
            double t = 0.0;
            for (int k = 0; k < RED_WG / 64; ++k) t += sm[i][k];
            v[i] = t;
        }
}

// Writes the block partials; returns true in thread 0 of the last block to
// finish, which then sees every partial.
template <int NV>
__device__ inline bool publish_partials(double v[NV], double *partials, DevScalars *sc)
{
    __shared__ bool last;
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
        for (int i = 0; i < NV; ++i) partials[i * gridDim.x + blockIdx.x] = v[i];
        __threadfence();
        unsigned int t = atomicAdd(&sc->ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    return last;
}

template <int NV>
__device__ inline void final_sum(const double *partials, int nb, double out[NV])
{
    // thread 0 of the last block, fixed order; the partials of other CUs are
    // read past this CU's L1 (agent-scope acquire, then agent-scope loads)
    __threadfence();
    for (int i = 0; i < NV; ++i) {
        double t = 0.0;
        for (int b = 0; b < nb; ++b)
            t += __hip_atomic_load(&partials[i * nb + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out[i] = t;
    }
}

// y = K x (+ optional <x, y> partial for PCG).  One group of G lanes per free node.
template <class TV, class T, int G, bool DOT>
__global__ __launch_bounds__(RED_WG) void k_spmv(int nfree, const int *__restrict__ node_blk,
                                                 const int *__restrict__ bcol, const TV *__restrict__ val,
                                                 const T *__restrict__ x, T *__restrict__ y, double *partials,
                                                 DevScalars *sc)
{
    if (DOT && sc->done) return;
    constexpr int NPW = RED_WG / G;
    const int blk = xcd_remap_s(blockIdx.x, gridDim.x);
    const int r = blk * NPW + (int)threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double dot = 0.0;
    if (r < nfree) {
        const int nb0 = node_blk[r], deg = node_blk[r + 1] - nb0;
        const int len = 3 * deg;
        const TV *v0 = val + 9 * (long long)nb0;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int j = lane; j < len; j += G) {
            int kk = j / 3, d = j - 3 * kk;
            int s = bcol[nb0 + kk];
            double xv = (double)x[3 * s + d];
            a0 += (double)v0[j] * xv;
            a1 += (double)v0[len + j] * xv;
            a2 += (double)v0[2 * len + j] * xv;
        }
        a0 = wave_group_sum(a0, G);
        a1 = wave_group_sum(a1, G);
        a2 = wave_group_sum(a2, G);
        if (lane < 3) {
            double a = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
            y[3 * r + lane] = (T)a;
            if (DOT) dot = a * (double)x[3 * r + lane];
        }
    }
    if (DOT) {
        double v[1] = {dot};
        if (publish_partials<1>(v, partials, sc)) {
            double t[1];
            final_sum<1>(partials, gridDim.x, t);
            sc->s[S_PQ] = t[0];
            if (!(t[0] > 0.0) || !isfinite(t[0])) sc->done = 3;  // breakdown: K p not positive
            else sc->s[S_ALPHA] = sc->s[S_RZ] / t[0];
            sc->ticket = 0;
        }
    }
}

// preconditioner application on node r: z_r = M_r r_r
template <class T, int PC>
__device__ inline void apply_pc(int r, const T *rr, T *zz, const double *dinv, const T *minv)
{
    if (PC == 0) {
        for (int c = 0; c < 3; ++c) zz[c] = (T)((double)rr[c] * dinv[3 * r + c]);
    } else {
        const T *m = minv + 9 * r;
        for (int c = 0; c < 3; ++c) zz[c] = m[3 * c] * rr[0] + m[3 * c + 1] * rr[1] + m[3 * c + 2] * rr[2];
    }
}

// x += alpha p; r -= alpha q; z = M r; partial <r,z>, <r,r>.  Thread per node.
template <class T, int PC>
__global__ __launch_bounds__(RED_WG) void k_update(int nfree, T *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
                                                   const T *__restrict__ p, const T *__restrict__ q,
                                                   const double *__restrict__ dinv, const T *__restrict__ minv,
                                                   double *partials, DevScalars *sc)
{
    if (sc->done) return;
    const double alpha = sc->s[S_ALPHA];
    const int i = blockIdx.x * RED_WG + threadIdx.x;
    double rz = 0.0, rr2 = 0.0;
    if (i < nfree) {
        T rv[3], zv[3];
        for (int c = 0; c < 3; ++c) {
            int k = 3 * i + c;
            x[k] = (T)((double)x[k] + alpha * (double)p[k]);
            rv[c] = (T)((double)r[k] - alpha * (double)q[k]);
            r[k] = rv[c];
        }
        apply_pc<T, PC>(i, rv, zv, dinv, minv);
        for (int c = 0; c < 3; ++c) {
            z[3 * i + c] = zv[c];
            rz += (double)rv[c] * (double)zv[c];
            rr2 += (double)rv[c] * (double)rv[c];
        }
    }
    double v[2] = {rz, rr2};
    if (publish_partials<2>(v, partials, sc)) {
        double t[2];
        final_sum<2>(partials, gridDim.x, t);
        double rz_old = sc->s[S_RZ];
        sc->s[S_BETA] = t[0] / rz_old;
        sc->s[S_RZ] = t[0];
        sc->s[S_RR] = t[1];
        int it = sc->it + 1;
        sc->it = it;
        if (t[1] <= sc->s[S_TOL2]) sc->done = 1;
        else if (it >= sc->maxit) sc->done = 2;
        else if (!isfinite(t[0])) sc->done = 3;
        sc->ticket = 0;
    }
}

template <class T>
__global__ __launch_bounds__(RED_WG) void k_direction(int n, T *__restrict__ p, const T *__restrict__ z, DevScalars *sc)
{
    if (sc->done) return;
    const double beta = sc->s[S_BETA];
    int i = blockIdx.x * RED_WG + threadIdx.x;
    if (i < n) p[i] = (T)((double)z[i] + beta * (double)p[i]);
}

// Start of a PCG run on residual r (already set): z = M r, p = z, <r,z>, <r,r>.
template <class T, int PC>
__global__ __launch_bounds__(RED_WG) void k_cg_start(int nfree, const T *__restrict__ r, T *__restrict__ z,
                                                     T *__restrict__ p, const double *__restrict__ dinv,
                                                     const T *__restrict__ minv, double *partials, DevScalars *sc,
                                                     double tol2_rel, int maxit)
{
    const int i = blockIdx.x * RED_WG + threadIdx.x;
    double rz = 0.0, rr2 = 0.0;
    if (i < nfree) {
        T rv[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]}, zv[3];
        apply_pc<T, PC>(i, rv, zv, dinv, minv);
        for (int c = 0; c < 3; ++c) {
            z[3 * i + c] = zv[c];
            p[3 * i + c] = zv[c];
            rz += (double)rv[c] * (double)zv[c];
            rr2 += (double)rv[c] * (double)rv[c];
        }
    }
    double v[2] = {rz, rr2};
    if (publish_partials<2>(v, partials, sc)) {
        double t[2];
        final_sum<2>(partials, gridDim.x, t);
        sc->s[S_RZ] = t[0];
        sc->s[S_RR] = t[1];
        sc->s[S_TOL2] = tol2_rel * t[1];
        sc->it = 0;
        sc->maxit = maxit;
        sc->done = (t[1] == 0.0) ? 1 : 0;
        sc->ticket = 0;
    }
}

// r = b - y (FP64) and <r,r>, <b,b> into s[S_RR], s[S_BB]
__global__ __launch_bounds__(RED_WG) void k_residual(int n, const double *__restrict__ b, const double *__restrict__ y,
                                                     double *__restrict__ r, double *partials, DevScalars *sc)
{
    int i = blockIdx.x * RED_WG + threadIdx.x;
    double rr = 0.0, bb = 0.0;
    if (i < n) {
        double v = b[i] - y[i];
        r[i] = v;
        rr = v * v;
        bb = b[i] * b[i];
    }
    double v[2] = {rr, bb};
    if (publish_partials<2>(v, partials, sc)) {
        double t[2];
        final_sum<2>(partials, gridDim.x, t);
        sc->s[S_RR] = t[0];
        sc->s[S_BB] = t[1];
        sc->ticket = 0;
    }
}

template <class A, class B>
__global__ void k_convert(int n, const A *__restrict__ a, B *__restrict__ b)
{
    int i = blockIdx.x * RED_WG + threadIdx.x;
    if (i < n) b[i] = (B)a[i];
}

__global__ void k_add_to(int n, double *__restrict__ x, const float *__restrict__ e)
{
    int i = blockIdx.x * RED_WG + threadIdx.x;
    if (i < n) x[i] += (double)e[i];
}

template <class T>
__global__ void k_fill(int n, T *__restrict__ a, T v)
{
    int i = blockIdx.x * RED_WG + threadIdx.x;
    if (i < n) a[i] = v;
}

// Diagonal data of the preconditioners: point Jacobi 1/K_ii and the inverse of
// each free node's 3x3 diagonal block.  Thread per free node.
__global__ void k_precond_setup(int nfree, const int *__restrict__ node_blk, const int *__restrict__ bcol,
                                const double *__restrict__ val, double *__restrict__ dinv, double *__restrict__ minv,
                                float *__restrict__ minv32)
{
    int r = blockIdx.x * RED_WG + threadIdx.x;
    if (r >= nfree) return;
    int nb0 = node_blk[r], deg = node_blk[r + 1] - nb0;
    int k = 0;
    while (k < deg && bcol[nb0 + k] != r) ++k;
    double A[9];
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) A[3 * c + d] = k < deg ? val[9 * (long long)nb0 + c * 3 * deg + 3 * k + d] : 0.0;
    for (int c = 0; c < 3; ++c) dinv[3 * r + c] = 1.0 / A[4 * c];
    // block stored as [col c][row d]; K is symmetric so it is its own transpose
    double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8], c02 = A[3] * A[7] - A[4] * A[6];
    double id = 1.0 / (A[0] * c00 + A[1] * c01 + A[2] * c02);
    double M[9] = {c00 * id, (A[2] * A[7] - A[1] * A[8]) * id, (A[1] * A[5] - A[2] * A[4]) * id,
                   c01 * id, (A[0] * A[8] - A[2] * A[6]) * id, (A[2] * A[3] - A[0] * A[5]) * id,
                   c02 * id, (A[1] * A[6] - A[0] * A[7]) * id, (A[0] * A[4] - A[1] * A[3]) * id};
    for (int i = 0; i < 9; ++i) {
        minv[9 * r + i] = M[i];
        minv32[9 * r + i] = (float)M[i];
    }
}

// ------------------------------------------------------------------ host side
static int ensure_ws(ofea_ctx *ctx)
{
    ofea_solver_ws *&w = ctx->ws;
    if (w && w->n == ctx->P.n) return OFEA_OK;
    solver_free(ctx);
    w = new ofea_solver_ws();
    w->n = ctx->P.n;
    w->nfree = ctx->P.nfree;
    int n = w->n, nf = w->nfree;
    w->nblk_vec = (n + RED_WG - 1) / RED_WG;
    w->nblk_node = (nf + RED_WG - 1) / RED_WG;
    int maxb = (nf + 1) / 1 + 64;  // spmv with group 1 would be the most blocks; bounded below
    OFEA_HIP(hipMalloc((void **)&w->partials, sizeof(double) * 3 * (size_t)maxb));
    OFEA_HIP(hipMalloc((void **)&w->sc, sizeof(DevScalars)));
    OFEA_HIP(hipMemset(w->sc, 0, sizeof(DevScalars)));
    OFEA_HIP(hipHostMalloc((void **)&w->sc_host, sizeof(DevScalars), hipHostMallocDefault));
    size_t nb = (size_t)std::max(n, 1);
    for (double **p : {&w->v64.r, &w->v64.z, &w->v64.p, &w->v64.q, &w->r64})
        OFEA_HIP(hipMalloc((void **)p, sizeof(double) * nb));
    for (float **p : {&w->v32.r, &w->v32.z, &w->v32.p, &w->v32.q, &w->v32.x})
        OFEA_HIP(hipMalloc((void **)p, sizeof(float) * nb));
    OFEA_HIP(hipMalloc((void **)&w->dinv, sizeof(double) * nb));
    OFEA_HIP(hipMalloc((void **)&w->minv, sizeof(double) * 9 * (size_t)std::max(nf, 1)));
    OFEA_HIP(hipMalloc((void **)&w->minv32, sizeof(float) * 9 * (size_t)std::max(nf, 1)));
    OFEA_HIP(hipMalloc((void **)&w->val32, sizeof(float) * (size_t)std::max(ctx->P.nnz, 1LL)));
    w->spmv_nodes_per_wg = RED_WG / w->spmv_group;
    w->nblk_spmv = (nf + w->spmv_nodes_per_wg - 1) / w->spmv_nodes_per_wg;
    return OFEA_OK;
}

void solver_free(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    if (!w) return;
    for (int i = 0; i < 4; ++i) {
        if (w->exec[i]) hipGraphExecDestroy(w->exec[i]);
        if (w->graph[i]) hipGraphDestroy(w->graph[i]);
    }
    for (double *p : {w->partials, w->v64.r, w->v64.z, w->v64.p, w->v64.q, w->r64, w->dinv, w->minv})
        if (p) hipFree(p);
    for (float *p : {w->v32.r, w->v32.z, w->v32.p, w->v32.q, w->v32.x, w->minv32, w->val32})
        if (p) hipFree(p);
    if (w->sc) hipFree(w->sc);
    if (w->sc_host) hipHostFree(w->sc_host);
    delete w;
    ctx->ws = nullptr;
}

void solver_invalidate(ofea_ctx *ctx)
{
    if (ctx->ws) ctx->ws->precond_valid = false;
}

template <class TV, class T, bool DOT>
static void launch_spmv_t(ofea_ctx *ctx, int group, const TV *val, const T *x, T *y)
{
    ofea_solver_ws *w = ctx->ws;
    int npw = RED_WG / group;
    int nb = (w->nfree + npw - 1) / npw;
    if (nb == 0) return;
#define OFEA_SPMV_CASE(G)                                                                                        \
    case G:                                                                                                      \
        hipLaunchKernelGGL((k_spmv<TV, T, G, DOT>), dim3(nb), dim3(RED_WG), 0, ctx->stream, w->nfree,          \
                           ctx->d_node_blk, ctx->d_bcol, val, x, y, w->partials, w->sc);                        \
        break;
    switch (group) {
        OFEA_SPMV_CASE(4)
        OFEA_SPMV_CASE(8)
        OFEA_SPMV_CASE(16)
        OFEA_SPMV_CASE(32)
        default:
        OFEA_SPMV_CASE(64)
    }
#undef OFEA_SPMV_CASE
}

int solver_spmv(ofea_ctx *ctx, const double *d_x, double *d_y)
{
    int rc = ensure_ws(ctx);
    if (rc) return rc;
    launch_spmv_t<double, double, false>(ctx, ctx->ws->spmv_group, ctx->d_val, d_x, d_y);
    OFEA_HIP(hipGetLastError());
    return OFEA_OK;
}

static int precond_setup(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    if (w->precond_valid) return OFEA_OK;
    if (w->nblk_node)
        hipLaunchKernelGGL(k_precond_setup, dim3(w->nblk_node), dim3(RED_WG), 0, ctx->stream, w->nfree,
                           ctx->d_node_blk, ctx->d_bcol, ctx->d_val, w->dinv, w->minv, w->minv32);
    long long nnz = ctx->P.nnz;
    if (nnz)
        hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((nnz + RED_WG - 1) / RED_WG)), dim3(RED_WG), 0,
                           ctx->stream, (int)nnz, ctx->d_val, w->val32);
    OFEA_HIP(hipGetLastError());
    w->precond_valid = true;
    return OFEA_OK;
}

// One CG iteration (three kernels) on precision T with preconditioner PC.
template <class TV, class T, int PC>
static void cg_iteration(ofea_ctx *ctx, const TV *val, VecSet<T> &v, const T *minv)
{
    ofea_solver_ws *w = ctx->ws;
    launch_spmv_t<TV, T, true>(ctx, w->spmv_group, val, v.p, v.q);
    hipLaunchKernelGGL((k_update<T, PC>), dim3(w->nblk_node), dim3(RED_WG), 0, ctx->stream, w->nfree, v.x, v.r, v.z,
                       v.p, v.q, w->dinv, minv, w->partials, w->sc);
    hipLaunchKernelGGL((k_direction<T>), dim3(w->nblk_vec), dim3(RED_WG), 0, ctx->stream, w->n, v.p, v.z, w->sc);
}

template <class TV, class T, int PC>
static int run_cg(ofea_ctx *ctx, int slot, const TV *val, VecSet<T> &v, const T *minv, double tol_rel, int maxit,
                  int *iters)
{
    ofea_solver_ws *w = ctx->ws;
    // start: r already holds the residual, x the initial guess
    hipLaunchKernelGGL((k_cg_start<T, PC>), dim3(w->nblk_node), dim3(RED_WG), 0, ctx->stream, w->nfree, v.r, v.z, v.p,
                       w->dinv, minv, w->partials, w->sc, tol_rel * tol_rel, maxit);
    OFEA_HIP(hipGetLastError());
    // (re)capture the batch graph for this vector set / precision / preconditioner
    if (!w->exec[slot]) {
        OFEA_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < CG_BATCH; ++i) cg_iteration<TV, T, PC>(ctx, val, v, minv);
        OFEA_HIP(hipStreamEndCapture(ctx->stream, &w->graph[slot]));
        OFEA_HIP(hipGraphInstantiate(&w->exec[slot], w->graph[slot], nullptr, nullptr, 0));
    }
    for (;;) {
        OFEA_HIP(hipGraphLaunch(w->exec[slot], ctx->stream));
        OFEA_HIP(hipMemcpyAsync(w->sc_host, w->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
        OFEA_HIP(hipStreamSynchronize(ctx->stream));
        if (w->sc_host->done) break;
    }
    if (iters) *iters = w->sc_host->it;
    if (w->sc_host->done == 3) return fail(OFEA_ERR_STATE, "PCG breakdown (matrix or preconditioner not SPD)");
    return OFEA_OK;
}

// r64 = b - K x (FP64), relres = ||r|| / ||b||
static int true_residual(ofea_ctx *ctx, const double *b, const double *x, double *relres)
{
    ofea_solver_ws *w = ctx->ws;
    launch_spmv_t<double, double, false>(ctx, w->spmv_group, ctx->d_val, x, w->v64.q);
    hipLaunchKernelGGL(k_residual, dim3(w->nblk_vec), dim3(RED_WG), 0, ctx->stream, w->n, b, w->v64.q, w->r64,
                       w->partials, w->sc);
    OFEA_HIP(hipGetLastError());
    OFEA_HIP(hipMemcpyAsync(w->sc_host, w->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
    OFEA_HIP(hipStreamSynchronize(ctx->stream));
    double bb = w->sc_host->s[S_BB], rr = w->sc_host->s[S_RR];
    *relres = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    return OFEA_OK;
}

int solver_solve(ofea_ctx *ctx, int method, double rtol, int maxit, int *iters, double *relres)
{
    int rc = ensure_ws(ctx);
    if (rc) return rc;
    rc = precond_setup(ctx);
    if (rc) return rc;
    ofea_solver_ws *w = ctx->ws;
    const int n = w->n;
    const unsigned gv = w->nblk_vec;
    double *x = ctx->d_x, *b = ctx->d_rhs;
    int total_it = 0;
    double rel = 1.0;
    if (n == 0) {
        if (iters) *iters = 0;
        if (relres) *relres = 0.0;
        return OFEA_OK;
    }
    hipLaunchKernelGGL((k_fill<double>), dim3(gv), dim3(RED_WG), 0, ctx->stream, n, x, 0.0);
    if (method == OFEA_PCG_JACOBI || method == OFEA_PCG_BLOCK_JACOBI) {
        // FP64 PCG on r = b; restart from the true residual if the recursive
        // residual drifted from it (at most 4 restarts)
        hipLaunchKernelGGL((k_convert<double, double>), dim3(gv), dim3(RED_WG), 0, ctx->stream, n, b, w->v64.r);
        w->v64.x = x;
        double tol = rtol;
        for (int restart = 0; restart < 5; ++restart) {
            int it = 0;
            int left = maxit - total_it;
            if (left <= 0) break;
            if (method == OFEA_PCG_JACOBI)
                rc = run_cg<double, double, 0>(ctx, 0, ctx->d_val, w->v64, w->minv, tol, left, &it);
            else
                rc = run_cg<double, double, 1>(ctx, 1, ctx->d_val, w->v64, w->minv, tol, left, &it);
            total_it += it;
            if (rc) return rc;
            rc = true_residual(ctx, b, x, &rel);
            if (rc) return rc;
            if (rel <= rtol) break;
            // continue from the true residual with a tighter recursive target
            hipLaunchKernelGGL((k_convert<double, double>), dim3(gv), dim3(RED_WG), 0, ctx->stream, n, w->r64,
                               w->v64.r);
            // the next run starts from ||r|| = rel ||b||; aim below rtol ||b||
            tol = std::min(1.0, 0.5 * rtol / rel);
        }
    } else if (method == OFEA_PCG_MIXED || method == OFEA_PCG_LEARNED) {
        if (method == OFEA_PCG_LEARNED && !ctx->learned)
            return fail(OFEA_ERR_STATE, "learned preconditioner not set (ofea_learned_*)");
        // FP64 defect correction around an FP32 block-Jacobi PCG
        const float *M32 = w->minv32;
        hipLaunchKernelGGL((k_convert<double, double>), dim3(gv), dim3(RED_WG), 0, ctx->stream, n, b, w->r64);
        rc = true_residual(ctx, b, x, &rel);
        if (rc) return rc;
        const double inner_tol = 1e-3;
        for (int outer = 0; outer < 30 && rel > rtol; ++outer) {
            int left = maxit - total_it;
            if (left <= 0) break;
            hipLaunchKernelGGL((k_convert<double, float>), dim3(gv), dim3(RED_WG), 0, ctx->stream, n, w->r64, w->v32.r);
            hipLaunchKernelGGL((k_fill<float>), dim3(gv), dim3(RED_WG), 0, ctx->stream, n, w->v32.x, 0.0f);
            int it = 0;
            double tol_in = std::max(inner_tol, 0.5 * rtol / rel);
            rc = run_cg<float, float, 1>(ctx, 2, w->val32, w->v32, M32, tol_in, left, &it);
            total_it += it;
            if (rc) return rc;
            hipLaunchKernelGGL(k_add_to, dim3(gv), dim3(RED_WG), 0, ctx->stream, n, x, w->v32.x);
            rc = true_residual(ctx, b, x, &rel);
            if (rc) return rc;
        }
    } else {
        return fail(OFEA_ERR_ARG, "unknown solver method");
    }
    if (iters) *iters = total_it;
    if (relres) *relres = rel;
    return rel <= rtol ? OFEA_OK : OFEA_NOT_CONVERGED;
}

}  // namespace ofea
