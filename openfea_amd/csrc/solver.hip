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
// PCG iteration = 2 kernels: (p = z + beta p folded into the SpMV, q = K p,
// <p,q>) and (x += alpha p, r -= alpha q, z = M r, <r,z>, <r,r>).  Both run on
// a fixed grid; each workgroup reduces the previous kernel's per-workgroup
// partials itself in a fixed order (deterministic; the two iteration kernels
// use no atomics).  The stand-alone kernels outside the iteration (the SpMV
// with a dot product and the true residual) find their last workgroup by one
// atomicAdd ticket per workgroup; that workgroup then sums the partials in a
// fixed order, so their results are deterministic as well.  Batches of
// iterations are captured in a hipGraph; kernels return at once after
// convergence and the host reads the scalars between graph launches.
//
// Precision: the plain solvers run CG in FP64.  The mixed solver runs one
// continuous CG in FP32 (FP32 values, FP32 vectors) while accumulating x in
// FP64, and replaces its recursive residual by the true FP64 residual
// b - K64 x whenever the residual has dropped by 10x since the last
// replacement ("reliable updates"); convergence is only accepted on the true
// FP64 residual.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "elem_device.h"
#include "ofea_internal.h"
#include "solver_ws.h"

namespace ofea {

constexpr int RED_WG = 256;   // threads per workgroup of vector kernels
constexpr int CG_BATCH = 16;  // iterations per graph launch
constexpr int CG_TAIL = 4;    // iterations per graph launch near convergence

// Cross-lane moves for the reductions, on the VALU (DPP) and the scalar
// read-lane path: a shuffle through ds_bpermute costs LDS issue slots in every
// wave, which at the end of a memory-bound kernel is exposed time.
template <int CTRL>
__device__ inline double dpp_mov(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ inline double read_lane(double x, int l)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Sum over each row of 16 lanes, in every lane of the row: pairs (quad_perm
// [1,0,3,2]), quads ([2,3,0,1]), half-row mirror, row mirror.  Each step adds
// two values that the partner lane adds in the other order, so all lanes of
// the row hold the same bits.
__device__ inline double row16_sum(double x)
{
    x += dpp_mov<0xB1>(x);
    x += dpp_mov<0x4E>(x);
    x += dpp_mov<0x141>(x);
    x += dpp_mov<0x140>(x);
    return x;
}

// Sum over the 64 lanes of a wave, the same value in every lane (fixed order).
__device__ inline double wave_sum(double x)
{
    x = row16_sum(x);
    return (read_lane(x, 0) + read_lane(x, 16)) + (read_lane(x, 32) + read_lane(x, 48));
}

// Sum over aligned groups of G lanes, in every lane of the group: DPP steps
// within a 16-lane row, shuffles only across rows (G = 32, 64).
template <int G>
__device__ inline double group_sum(double v)
{
    static_assert(G >= 4 && G <= 64 && (G & (G - 1)) == 0, "group of 4..64 lanes");
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    if constexpr (G >= 8) v += dpp_mov<0x141>(v);
    if constexpr (G >= 16) v += dpp_mov<0x140>(v);
    if constexpr (G >= 32) v += __shfl_xor(v, 16);
    if constexpr (G >= 64) v += __shfl_xor(v, 32);
    return v;
}

// Block-wide sum of NV doubles in a fixed tree; result valid in thread 0.
template <int NV>
__device__ inline void block_sum(double v[NV])
{
    __shared__ double sm[NV][RED_WG / 64];
    for (int i = 0; i < NV; ++i) {
        double t = v[i];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
        v[i] = t;
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
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

// Grid-wide sum of NV doubles: every block publishes its partial, the last
// block to arrive sums all partials (all its threads, strided, then a fixed
// tree).  Returns true in thread 0 of that last block, with out[] the totals.
template <int NV>
__device__ inline bool grid_sum(double v[NV], double *partials, DevScalars *sc, double out[NV])
{
    __shared__ int s_last;
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
        for (int i = 0; i < NV; ++i) partials[i * gridDim.x + blockIdx.x] = v[i];
        __threadfence();
        unsigned int t = atomicAdd(&sc->ticket, 1u);
        s_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last) return false;  // uniform over the block
    __threadfence();
    const int nb = gridDim.x;
    for (int i = 0; i < NV; ++i) {
        double t = 0.0;
        for (int b = threadIdx.x; b < nb; b += RED_WG)
            t += __hip_atomic_load(&partials[i * nb + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out[i] = t;
    }
    block_sum<NV>(out);
    if (threadIdx.x == 0) sc->ticket = 0;
    return threadIdx.x == 0;
}

// y = K x (+ optional <x, y> for PCG).  One group of G lanes per free node.
template <class TV, class T, int G, bool DOT>
__global__ __launch_bounds__(RED_WG) void k_spmv(int nfree, const int *__restrict__ node_blk,
                                                 const int *__restrict__ bcol, const TV *__restrict__ val,
                                                 const T *__restrict__ x, T *__restrict__ y, double *partials,
                                                 DevScalars *sc)
{
    if (DOT && sc->done) return;
    constexpr int NPW = RED_WG / G;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
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
        a0 = group_sum<G>(a0);
        a1 = group_sum<G>(a1);
        a2 = group_sum<G>(a2);
        if (lane < 3) {
            double a = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
            y[3 * r + lane] = (T)a;
            if (DOT) dot = a * (double)x[3 * r + lane];
        }
    }
    if (DOT) {
        double v[1] = {dot}, t[1];
        if (grid_sum<1>(v, partials, sc, t)) {
            sc->s[S_PQ] = t[0];
            if (!(t[0] > 0.0) || !isfinite(t[0])) sc->done = 3;  // breakdown: K or M not SPD
            else sc->s[S_ALPHA] = sc->s[S_RZ] / t[0];
        }
    }
}

// ---------------------------------------------------------------- CG kernels
// The CG kernels run on a fixed grid of CG_NB workgroups of CG_WG threads
// (two 16-wave workgroups per CU), each owning a contiguous node range.  A
// reduction writes one partial per workgroup; every workgroup of the next
// kernel sums the CG_NB partials itself, in the same fixed order, so all of
// them derive identical scalars without atomics or a last-block pass.  The
// batch index bi (kernel argument) and the host-set base values it0/par0 give
// the iteration number and the parity of the ping-pong buffers.
constexpr int CG_NB = 512;
constexpr int CG_WG = 1024;

struct CGParts {
    double *rz;  // 2 * CG_NB: <r,z> of the current and the previous iteration (by parity)
    double *rr;  // CG_NB: <r,r>
    double *pq;  // CG_NB: <p,Kp>
};

// Sum over the WG threads of a workgroup (fixed order), result in every
// thread: wave sums, one LDS slot per wave, then lanes 0..NW-1 of every wave
// read the NW wave sums and reduce them as a 16-lane row.
template <int NV, int WG = CG_WG>
__device__ inline void cg_block_sum(double v[NV])
{
    constexpr int NW = WG / 64;
    static_assert(NW >= 1 && NW <= 16, "at most 16 waves");
    __shared__ double sm[NV][NW];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    __syncthreads();  // an earlier use of sm by this workgroup is complete
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) sm[i][threadIdx.x >> 6] = v[i];
    __syncthreads();
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double t = l < NW ? sm[i][l] : 0.0;
        v[i] = read_lane(row16_sum(t), 0);
    }
}

// Total of N doubles at col (LDS or global), in every lane of the calling
// wave: lane l adds entries l, l + 64, ... in that order, then a wave sum.
template <int N>
__device__ inline double col_sum(const double *col)
{
    const int l = threadIdx.x & 63;
    double t = 0.0;
    if constexpr (N >= 64) {
        static_assert(N % 64 == 0, "whole waves");
#pragma unroll
        for (int k = 0; k < N / 64; ++k) t += col[l + 64 * k];
    } else {
        t = l < N ? col[l] : 0.0;
    }
    return wave_sum(t);
}

// Totals of NV partial arrays (CG_NB entries each), same value in every
// thread: wave w < NV sums array w, the totals pass through LDS (one call
// per kernel).  The partials were written by an earlier kernel on the same
// stream; the kernel boundary makes them visible, so plain (L2-cached) loads
// suffice.
template <int NV, int WG = CG_WG>
__device__ inline void red_sums(const double *const arr[NV], double out[NV])
{
    static_assert(NV <= WG / 64, "one wave per array");
    __shared__ double s_tot[NV];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i)
        if (w == i) {
            const double t = col_sum<CG_NB>(arr[i]);
            if ((threadIdx.x & 63) == 0) s_tot[i] = t;
        }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) out[i] = s_tot[i];
}

// One partial per workgroup and array (thread 0 writes).
template <int NV>
__device__ inline void put_partials(double v[NV], double *const dst[NV])
{
    cg_block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) dst[i][blockIdx.x] = v[i];
}

// contiguous range [i0, i1) of n items for this workgroup; consecutive ranges
// go to the same XCD
__device__ inline void cg_range(int n, int &i0, int &i1)
{
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int per = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    i0 = min(b * per, n);
    i1 = min(i0 + per, n);
}

// Iteration start: convergence / reliable-update / maxit checks on the
// residual of the previous update, then p_new = z + beta p_old folded into
// q = K p_new, and the partial <p_new, q>.  With the coarse space (CO) the
// workgroup also writes its range's partial of Z^T q, from which the update
// kernel advances Z^T r without another pass over r.
template <class TV, class T, int G, bool CO>
__global__ __launch_bounds__(CG_WG) void k_cg_spmv(int bi, int nfree, const int *__restrict__ node_blk,
                                                   const int *__restrict__ bcol, const TV *__restrict__ val,
                                                   const T *__restrict__ z, T *__restrict__ pA, T *__restrict__ pB,
                                                   T *__restrict__ q, CGParts P, CoarseArgs C, DevScalars *sc)
{
    const int done0 = sc->done, par = sc->par0 ^ (bi & 1);
    const T *p_old = par ? pB : pA;
    T *p_new = par ? pA : pB;
    constexpr int NPW = CG_WG / G;
    const int lane = threadIdx.x % G;
    int r0, r1;
    cg_range(nfree, r0, r1);
    // the lane's first matrix column entry and vector values do not depend on
    // beta: load them before the reduction so that the two latencies overlap
    int r = r0 + (int)threadIdx.x / G;
    int nb0 = 0, len = 0;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, zf = 0.0, pf = 0.0;
    float dr[3] = {0.f, 0.f, 0.f};  // coarse offsets of node r (CO), loaded a pass ahead of their use
    if (r < r1) {
        if constexpr (CO)
#pragma unroll
            for (int c = 0; c < 3; ++c) dr[c] = C.d[3 * r + c];
        nb0 = node_blk[r];
        len = 3 * (node_blk[r + 1] - nb0);
        if (lane < len) {
            const int kk = lane / 3, d = lane - 3 * kk;
            const int s = bcol[nb0 + kk];
            const TV *v0 = val + 9 * (long long)nb0;
            m0 = (double)v0[lane];
            m1 = (double)v0[len + lane];
            m2 = (double)v0[2 * len + lane];
            zf = (double)z[3 * s + d];
            pf = (double)p_old[3 * s + d];
        }
    }
    if (done0) return;
    double red[3];
    {
        const double *const arr[3] = {P.rz + par * CG_NB, P.rz + (par ^ 1) * CG_NB, P.rr};
        red_sums<3>(arr, red);
    }
    const double rz = red[0], rzp = red[1], rr = red[2];
    const int it = sc->it0 + bi;
    int dn = 0;
    if (!isfinite(rz) || !isfinite(rr)) dn = 3;
    else if (rr <= sc->s[S_TOL2]) dn = 1;
    else if (rr <= sc->s[S_RU2]) dn = 4;
    else if (it >= sc->maxit) dn = 2;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->it = it;
        sc->s[S_RR] = rr;
        sc->s[S_RZ] = rz;
        if (dn) {
            sc->done = dn;
            sc->bi_done = bi;
        }
    }
    if (dn) return;
    const double beta = rz / rzp;  // rzp = inf at a start: beta = 0
    // dot: lane c < 3 of a node group accumulates q_c p_c; zq[u]: lane c
    // accumulates component c + u G (< 6) of Z^T q
    constexpr int ZC = G >= 6 ? 1 : 2;
    double dot = 0.0, zq[ZC];
#pragma unroll
    for (int u = 0; u < ZC; ++u) zq[u] = 0.0;
    for (bool first = true; r < r1; r += NPW, first = false) {
        if (!first) {
            nb0 = node_blk[r];
            len = 3 * (node_blk[r + 1] - nb0);
            if constexpr (CO)
#pragma unroll
                for (int c = 0; c < 3; ++c) dr[c] = C.d[3 * r + c];
        }
        const TV *v0 = val + 9 * (long long)nb0;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        int j = lane;
        if (first) {
            if (j < len) {
                const double pv = (double)(T)(zf + beta * pf);
                a0 += m0 * pv;
                a1 += m1 * pv;
                a2 += m2 * pv;
            }
            j += G;
        }
        for (; j < len; j += G) {
            const int kk = j / 3, d = j - 3 * kk;
            const int s = bcol[nb0 + kk];
            const double pv = (double)(T)((double)z[3 * s + d] + beta * (double)p_old[3 * s + d]);
            a0 += (double)v0[j] * pv;
            a1 += (double)v0[len + j] * pv;
            a2 += (double)v0[2 * len + j] * pv;
        }
        a0 = group_sum<G>(a0);
        a1 = group_sum<G>(a1);
        a2 = group_sum<G>(a2);
        if (lane < 3) {
            const double a = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
            const int k = 3 * r + lane;
            const T pn = (T)((double)z[k] + beta * (double)p_old[k]);
            p_new[k] = pn;
            q[k] = (T)a;
            dot += a * (double)pn;
        }
        if constexpr (CO) {
            // Z_r^T q_r = (q_r, d_r x q_r) from the stored (T) values
            const double q0 = (double)(T)a0, q1 = (double)(T)a1, q2 = (double)(T)a2;
            const double d0 = dr[0], d1 = dr[1], d2 = dr[2];
#pragma unroll
            for (int u = 0; u < ZC; ++u) {
                const int c = lane + u * G;
                if (c < 6)
                    zq[u] += c == 0 ? q0
                           : c == 1 ? q1
                           : c == 2 ? q2
                           : c == 3 ? d1 * q2 - d2 * q1
                           : c == 4 ? d2 * q0 - d0 * q2
                                    : d0 * q1 - d1 * q0;
            }
        }
    }
    // Partials: every node group stores its lane values in LDS (column c, row
    // = group); wave 0 sums the three <p,q> columns, wave 1 + c column c of
    // Z^T q, and lane 0 of each writes the workgroup's partial.
    {
        constexpr int NQ = CO ? 9 : 3;
        __shared__ double s_q[NQ][NPW];
        const int gi = (int)threadIdx.x / G, w = (int)threadIdx.x >> 6;
        if (lane < 3) s_q[lane][gi] = dot;
        if constexpr (CO)
#pragma unroll
            for (int u = 0; u < ZC; ++u)
                if (lane + u * G < 6) s_q[3 + lane + u * G][gi] = zq[u];
        __syncthreads();
        if (w == 0) {
            const double t = (col_sum<NPW>(s_q[0]) + col_sum<NPW>(s_q[1])) + col_sum<NPW>(s_q[2]);
            if (threadIdx.x == 0) P.pq[blockIdx.x] = t;
        } else if (CO && w <= 6) {
            const double t = col_sum<NPW>(s_q[2 + w]);
            if ((threadIdx.x & 63) == 0) C.zq[(w - 1) * CG_NB + xcd_remap(blockIdx.x, gridDim.x)] = t;
        }
    }
}

// preconditioner application on node r: z_r = M_r r_r (point or 3x3 block)
template <class T, int PC>
__device__ inline void apply_pc(int r, const T *rr, T *zz, const double *dinv, const T *minv)
{
    if (PC == 0) {
        for (int c = 0; c < 3; ++c) zz[c] = (T)((double)rr[c] * dinv[3 * r + c]);
    } else if (PC == 1) {
        const T *m = minv + 9 * r;
        for (int c = 0; c < 3; ++c) zz[c] = m[3 * c] * rr[0] + m[3 * c + 1] * rr[1] + m[3 * c + 2] * rr[2];
    } else {  // factorized correction: applied by the k_fac passes
        for (int c = 0; c < 3; ++c) zz[c] = T(0);
    }
}

// alpha = <r,z>/<p,q>; x += alpha p (x in XT); r -= alpha q; z = M r;
// partials <r,z> (into the other parity slot) and <r,r>.  Thread per node.
// Per-workgroup partials <r,z>, <r,r> (at blockIdx) and, with the coarse
// space, Z^T r of the workgroup's node range (at its range index).
template <bool CO, int WG>
__device__ inline void put_rz_rr_g(double rzn, double rr2, const double g[6], double *rz_dst, double *rr_dst,
                                   const CoarseArgs &C)
{
    if constexpr (CO) {
        double v[8] = {rzn, rr2, g[0], g[1], g[2], g[3], g[4], g[5]};
        cg_block_sum<8, WG>(v);
        if (threadIdx.x == 0) {
            rz_dst[blockIdx.x] = v[0];
            rr_dst[blockIdx.x] = v[1];
            const int b = xcd_remap(blockIdx.x, gridDim.x);
#pragma unroll
            for (int c = 0; c < 6; ++c) C.g[c * CG_NB + b] = v[2 + c];
        }
    } else {
        double v[2] = {rzn, rr2};
        cg_block_sum<2, WG>(v);
        if (threadIdx.x == 0) {
            rz_dst[blockIdx.x] = v[0];
            rr_dst[blockIdx.x] = v[1];
        }
    }
}

// g += Z_i^T r_i = (r_i, d_i x r_i)
template <class T>
__device__ inline void add_zt_r(const CoarseArgs &C, int i, const T rv[3], double g[6])
{
    const double dx = C.d[3 * i], dy = C.d[3 * i + 1], dz = C.d[3 * i + 2];
    const double r0 = rv[0], r1 = rv[1], r2 = rv[2];
    g[0] += r0;
    g[1] += r1;
    g[2] += r2;
    g[3] += dy * r2 - dz * r1;
    g[4] += dz * r0 - dx * r2;
    g[5] += dx * r1 - dy * r0;
}

// The learned factor G of M = G^T G as two half block-matrices (rows of G
// and rows of G^T) in the column-of-rows layout of K's values: row i's nb
// blocks start at 9 ptr[i], entry (c, 3 k + d) at 9 ptr[i] + 3 nb c + 3 k + d.
struct FacView {
    const int *lo_ptr = nullptr, *lo_col = nullptr, *up_ptr = nullptr, *up_col = nullptr;
    const unsigned short *lo = nullptr, *up = nullptr;  // BF16
};

__device__ inline double bf16_val(unsigned short h) { return (double)__uint_as_float((unsigned)h << 16); }

// Row i of a half matrix times x, by the GL lanes of an aligned group (lane j
// takes columns j, j + GL, ... of the row's 3 nb columns); every lane of the
// group returns the three sums.  FU columns per lane and round: the column
// indices, then the gathered x values, then the products, so that the loads
// of a round are in flight together.
template <class T, int GL>
__device__ inline void fac_row(int i, int lane, const int *__restrict__ ptr, const int *__restrict__ col,
                               const unsigned short *__restrict__ v, const T *__restrict__ x, double &a0, double &a1,
                               double &a2)
{
    constexpr int FU = 4;
    const int b0 = ptr[i], len = 3 * (ptr[i + 1] - b0);
    const unsigned short *v0 = v + 9 * (size_t)b0;
    a0 = a1 = a2 = 0.0;
    for (int j0 = lane; j0 < len; j0 += FU * GL) {
        int sx[FU];
        double xv[FU];
#pragma unroll
        for (int u = 0; u < FU; ++u) {
            const int j = j0 + u * GL, kk = j / 3;
            sx[u] = j < len ? 3 * col[b0 + kk] + (j - 3 * kk) : -1;
        }
#pragma unroll
        for (int u = 0; u < FU; ++u) xv[u] = sx[u] >= 0 ? (double)x[sx[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < FU; ++u) {
            const int j = j0 + u * GL;
            if (j < len) {
                a0 += bf16_val(v0[j]) * xv[u];
                a1 += bf16_val(v0[len + j]) * xv[u];
                a2 += bf16_val(v0[2 * len + j]) * xv[u];
            }
        }
    }
    a0 = group_sum<GL>(a0);
    a1 = group_sum<GL>(a1);
    a2 = group_sum<GL>(a2);
}

// The update kernels run 512-thread workgroups on the CG grid: a node range
// rarely exceeds 512 nodes, and the smaller workgroup keeps the whole grid
// resident at up to 80 VGPRs.
constexpr int UP_WG = 512;

// PC 2 (factorized correction M = G^T G): the kernel also advances y = G r by
// y -= alpha w, w = G q (q is complete after the SpMV kernel): w of each chunk
// of up to UP_WG nodes of the range is gathered into LDS by groups of FG
// lanes, before the reduction for the first chunk.  G^T y follows in k_fac_upper.
template <class T, class XT, int PC, bool CO>
__global__ __launch_bounds__(UP_WG) void k_cg_update(int bi, int nfree, XT *__restrict__ x, T *__restrict__ r,
                                                     T *__restrict__ z, const T *__restrict__ pA,
                                                     const T *__restrict__ pB, const T *__restrict__ q,
                                                     const double *__restrict__ dinv, const T *__restrict__ minv,
                                                     T *__restrict__ gy, FacView F, CGParts P, CoarseArgs C,
                                                     DevScalars *sc)
{
    const int done0 = sc->done, par = sc->par0 ^ (bi & 1);
    const T *p = par ? pA : pB;  // p_new of this iteration
    int i0, i1;
    cg_range(nfree, i0, i1);
    constexpr int FG = 8;
    __shared__ double s_w[PC == 2 ? 3 : 1][PC == 2 ? UP_WG : 1];
    auto fac_gather = [&](int c0) {  // w = G q for nodes c0 .. c0 + UP_WG - 1 of the range
        if constexpr (PC == 2) {
            const int lane = threadIdx.x % FG, cnt = min(UP_WG, i1 - c0);
            for (int nn = (int)threadIdx.x / FG; nn < cnt; nn += UP_WG / FG) {
                double a0, a1, a2;
                fac_row<T, FG>(c0 + nn, lane, F.lo_ptr, F.lo_col, F.lo, q, a0, a1, a2);
                if (lane < 3) s_w[lane][nn] = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
            }
        }
    };
    if (!done0) fac_gather(i0);
    // the node data do not depend on alpha: issue all loads before the reduction
    int i = i0 + (int)threadIdx.x;
    T pv[3], qv[3], rv[3], mv[PC == 0 ? 3 : 9], gyv[3];
    XT xv[3];
    float dv[3];
    auto load_node = [&](int ii) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            pv[c] = p[3 * ii + c];
            qv[c] = q[3 * ii + c];
            rv[c] = r[3 * ii + c];
            xv[c] = x[3 * ii + c];
        }
        if constexpr (PC == 2)
#pragma unroll
            for (int c = 0; c < 3; ++c) gyv[c] = gy[3 * ii + c];
        if constexpr (PC == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) mv[c] = (T)dinv[3 * ii + c];
        } else if constexpr (PC == 1) {
#pragma unroll
            for (int e = 0; e < 9; ++e) mv[e] = minv[9 * (size_t)ii + e];
        }
        if constexpr (CO) {
#pragma unroll
            for (int c = 0; c < 3; ++c) dv[c] = C.d[3 * ii + c];
        }
    };
    if (i < i1) load_node(i);
    // Coarse part (CO): y = A_c^-1 Z^T r of this workgroup's aggregate ag follows
    // the recursion y_new = y_old - alpha A_c^-1 Z^T q (r_new = r - alpha q).
    // Thread k (< 3 UP_WG) holds coarse columns k, k + UP_WG, k + 2 UP_WG: the
    // summed Z^T q partials of the column's aggregate and the six A_c^-1
    // entries of rows 6 ag .. 6 ag + 5.  The six products are reduced together
    // with <p,q> and <r,z>.  Both workgroups of an aggregate sum in the same
    // order and hold the same y; the first one stores it for the next iteration.
    constexpr int KC = CO ? 3 : 1;
    static_assert(!CO || CG_NC_MAX <= KC * UP_WG, "coarse columns per thread");
    double zqs[KC], y[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    float am[KC][6];
    int ag = 0;
    bool own_agg = false;
    if constexpr (CO) {
        const int b = xcd_remap(blockIdx.x, gridDim.x);
        ag = b / C.ra;
        own_agg = ag < C.nagg;  // trailing workgroups without nodes may lie past the last aggregate
        const int nc = 6 * C.nagg;
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
            const int k = threadIdx.x + kk * UP_WG;
            zqs[kk] = 0.0;
#pragma unroll
            for (int m = 0; m < 6; ++m) am[kk][m] = 0.f;
            if (own_agg && k < nc) {
#pragma unroll
                for (int m = 0; m < 6; ++m) am[kk][m] = C.ainv[(size_t)(6 * ag + m) * nc + k];
                const int aa = k / 6, c = k - 6 * aa;
                const double *zp = C.zq + c * CG_NB + aa * C.ra;  // written by the SpMV kernel
                const int nb = min(C.ra, CG_NB - aa * C.ra);
                for (int b2 = 0; b2 < nb; ++b2) zqs[kk] += zp[b2];
            }
        }
        if (own_agg)
#pragma unroll
            for (int m = 0; m < 6; ++m) y[m] = C.ya[par * CG_NC_MAX + 6 * ag + m];
    }
    if (done0) return;
    // wave 0: <p,q>, wave 1: <r,z> from the partial arrays; CO: every thread
    // stores its six A_c^-1 Z^T q products (column m, row = thread) and wave
    // 2 + m sums column m.  Totals pass through LDS to every thread.
    constexpr int NT = CO ? 8 : 2;
    static_assert(NT <= UP_WG / 64, "one wave per total");
    __shared__ double s_tot[NT];
    __shared__ double s_col[CO ? 6 : 2][UP_WG];  // CO products, then the <r,z>, <r,r> partials
    const int w = (int)threadIdx.x >> 6;
    if constexpr (CO) {
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            double t = 0.0;
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) t += (double)am[kk][m] * zqs[kk];
            s_col[m][threadIdx.x] = t;
        }
    }
    if (w < 2) {
        const double t = col_sum<CG_NB>(w == 0 ? P.pq : P.rz + par * CG_NB);
        if ((threadIdx.x & 63) == 0) s_tot[w] = t;
    }
    __syncthreads();
    if constexpr (CO) {
        if (w >= 2) {
            const double t = col_sum<UP_WG>(s_col[w - 2]);
            if ((threadIdx.x & 63) == 0) s_tot[w] = t;
        }
        __syncthreads();
    }
    const double pq = s_tot[0], rz = s_tot[1];
    double yq[6];
    if constexpr (CO)
#pragma unroll
        for (int m = 0; m < 6; ++m) yq[m] = s_tot[2 + m];
    if (!(pq > 0.0) || !isfinite(pq)) {  // breakdown: K or M not SPD
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->done = 3;
            sc->bi_done = bi;
        }
        return;
    }
    const double alpha = rz / pq;
    if constexpr (CO) {
#pragma unroll
        for (int m = 0; m < 6; ++m) y[m] -= alpha * yq[m];
        if (own_agg && xcd_remap(blockIdx.x, gridDim.x) % C.ra == 0 && threadIdx.x < 6) {
            double t = y[0];
#pragma unroll
            for (int m = 1; m < 6; ++m) t = (int)threadIdx.x == m ? y[m] : t;
            C.ya[(par ^ 1) * CG_NC_MAX + 6 * ag + threadIdx.x] = t;
        }
    }
    double rzn = 0.0, rr2 = 0.0;
    // chunks of UP_WG nodes (ranges longer than one workgroup take several);
    // every thread runs every chunk, for the barriers around the PC 2 gather
    for (int c0 = i0; c0 < i1; c0 += UP_WG, i += UP_WG) {
        if (c0 > i0) {
            if constexpr (PC == 2) {
                __syncthreads();  // s_w of the previous chunk has been read
                fac_gather(c0);
                __syncthreads();
            }
            if (i < i1) load_node(i);
        }
        if (i >= i1) continue;
        T zv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int k = 3 * i + c;
            x[k] = (XT)((double)xv[c] + alpha * (double)pv[c]);
            rv[c] = (T)((double)rv[c] - alpha * (double)qv[c]);
            r[k] = rv[c];
        }
        if constexpr (PC == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) zv[c] = rv[c] * mv[c];
        } else if constexpr (PC == 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) zv[c] = mv[3 * c] * rv[0] + mv[3 * c + 1] * rv[1] + mv[3 * c + 2] * rv[2];
        } else {  // factorized correction: y -= alpha G q here, z = G^T y in k_fac_upper
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gy[3 * i + c] = (T)((double)gyv[c] - alpha * s_w[c][i - c0]);
                zv[c] = T(0);
            }
        }
        if constexpr (CO) {  // z_i += (Z y)_i = t + theta x d_i of this workgroup's aggregate
            const double d0 = dv[0], d1 = dv[1], d2 = dv[2];
            const double zc[3] = {y[0] + y[4] * d2 - y[5] * d1, y[1] + y[5] * d0 - y[3] * d2,
                                  y[2] + y[3] * d1 - y[4] * d0};
#pragma unroll
            for (int c = 0; c < 3; ++c) zv[c] = (T)((double)zv[c] + zc[c]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (PC != 2 || CO) z[3 * i + c] = zv[c];
            rzn += (double)rv[c] * (double)zv[c];
            rr2 += (double)rv[c] * (double)rv[c];
        }
    }
    // partials <r,z> (wave 0) and <r,r> (wave 1) through LDS; the CO columns
    // of s_col were last read before the second barrier above
    s_col[0][threadIdx.x] = rzn;
    s_col[1][threadIdx.x] = rr2;
    __syncthreads();
    if (w < 2) {
        const double t = col_sum<UP_WG>(s_col[w]);
        if (threadIdx.x == 0 && PC != 2) P.rz[(par ^ 1) * CG_NB + blockIdx.x] = t;  // PC 2: k_fac_upper
        if (threadIdx.x == 64) P.rr[blockIdx.x] = t;
    }
}

// Coarse correction after a start or residual replacement (k_replace), one
// workgroup per aggregate a: sums the per-range partials of Z^T r (all
// aggregates, same order in every workgroup), forms y_a = rows 6a..6a+5 of
// A_c^-1 (Z^T r), adds (Z y)_i = t_a + theta_a x d_i to z_i for the nodes of
// aggregate a (z then holds the full two-level M r) and adds the aggregate's
// <r, Z y> to one <r, z> partial of slot par0.  y is stored (slot par0) as
// the start of the update kernels' recursion.
template <class T>
__global__ __launch_bounds__(CG_WG) void k_coarse(int nfree, int nagg, const float *__restrict__ ainv,
                                                  const T *__restrict__ r, T *__restrict__ z, double *rz_part,
                                                  CoarseArgs C, DevScalars *sc)
{
    // every load that does not depend on the scalars is issued before the
    // done-check, so their latencies overlap
    const int done = sc->done, slot = sc->par0;
    const int nc = 6 * nagg;  // <= 6 CG_NB / CG_AGG <= 2 CG_WG: at most two coarse columns per thread
    const int a = blockIdx.x;
    const int i0 = min(a * C.ra * C.per, nfree), i1 = min(i0 + C.ra * C.per, nfree);
    double y[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
        const int k = threadIdx.x + kk * CG_WG;
        if (k < nc) {
            float am[6];
#pragma unroll
            for (int i2 = 0; i2 < 6; ++i2) am[i2] = ainv[(size_t)(6 * a + i2) * nc + k];
            const int aa = k / 6, c = k - 6 * aa;
            const double *gp = C.g + c * CG_NB + aa * C.ra;  // written by the previous kernel
            const int nb = min(C.ra, CG_NB - aa * C.ra);
            double gk = 0.0;
            for (int b = 0; b < nb; ++b) gk += gp[b];
#pragma unroll
            for (int i2 = 0; i2 < 6; ++i2) y[i2] += (double)am[i2] * gk;
        }
    }
    int i = i0 + (int)threadIdx.x;
    double dv[3] = {0.0, 0.0, 0.0}, zv[3] = {0.0, 0.0, 0.0}, rv[3] = {0.0, 0.0, 0.0};
    if (i < i1)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dv[c] = C.d[3 * i + c];
            zv[c] = (double)z[3 * i + c];
            rv[c] = (double)r[3 * i + c];
        }
    if (done) return;
    cg_block_sum<6>(y);
    if (threadIdx.x < 6) {  // y of aggregate a for the update kernels' recursion
        double t = y[0];
#pragma unroll
        for (int m = 1; m < 6; ++m) t = (int)threadIdx.x == m ? y[m] : t;
        C.ya[slot * CG_NC_MAX + 6 * a + threadIdx.x] = t;
    }
    double rzc = 0.0;
    for (; i < i1; i += CG_WG) {
        if (i >= i0 + CG_WG)  // aggregates longer than the workgroup
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                dv[c] = C.d[3 * i + c];
                zv[c] = (double)z[3 * i + c];
                rv[c] = (double)r[3 * i + c];
            }
        const double zc[3] = {y[0] + y[4] * dv[2] - y[5] * dv[1], y[1] + y[5] * dv[0] - y[3] * dv[2],
                              y[2] + y[3] * dv[1] - y[4] * dv[0]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            z[3 * i + c] = (T)(zv[c] + zc[c]);
            rzc += rv[c] * zc[c];
        }
    }
    double v[1] = {rzc};
    cg_block_sum<1>(v);
    // the local part of <r, z> of this slot is complete (previous kernel); entry a
    // gets the coarse part of aggregate a added, by this workgroup only
    if (threadIdx.x == 0) rz_part[slot * CG_NB + a] += v[0];
}

// advances the host-visible base iteration after a full batch
__global__ void k_cg_batch_end(DevScalars *sc, int nbatch)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && !sc->done) sc->it0 += nbatch;
}

// stopping targets from ||b||^2 (written by k_residual): tol2 = rtol^2 ||b||^2,
// reliable-update level ru^2 ||b||^2 (ru2 < 0: none); b = 0 ends the run at once
__global__ void k_set_tol(DevScalars *sc, double rtol2, double ru2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double bb = sc->s[S_BB];
    sc->s[S_TOL2] = rtol2 * bb;
    sc->s[S_RU2] = ru2 > 0.0 ? ru2 * bb : -1.0;
    if (bb == 0.0) sc->done = 1;
}

// Residual replacement / start: r = (T) r64, z = M r; partials <r,z> into
// the slot of parity par0 and <r,r>.  At a start (restart != 0) the other
// slot is set to +inf, so the next iteration takes beta = 0 (p = z).  At a
// reliable update the other slot keeps <r,z> of the previous iteration, so
// beta = <r,z>_new / <r,z>_prev and the search direction is kept.
template <class T, int PC, bool CO>
__global__ __launch_bounds__(UP_WG) void k_replace(int nfree, const double *__restrict__ r64, T *__restrict__ r,
                                                   T *__restrict__ z, int restart, const double *__restrict__ dinv,
                                                   const T *__restrict__ minv, CGParts P, CoarseArgs C,
                                                   DevScalars *sc)
{
    const int par = sc->par0;
    int i0, i1;
    cg_range(nfree, i0, i1);
    double rz = 0.0, rr2 = 0.0, g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = i0 + (int)threadIdx.x; i < i1; i += UP_WG) {
        T rv[3] = {(T)r64[3 * i], (T)r64[3 * i + 1], (T)r64[3 * i + 2]}, zv[3];
        apply_pc<T, PC>(i, rv, zv, dinv, minv);
        for (int c = 0; c < 3; ++c) {
            r[3 * i + c] = rv[c];
            if (PC != 2) z[3 * i + c] = zv[c];
            rz += (double)rv[c] * (double)zv[c];
            rr2 += (double)rv[c] * (double)rv[c];
        }
        if constexpr (CO) add_zt_r(C, i, rv, g);
    }
    // PC 2: the <r, z> partials of this slot come from k_fac_upper (a scratch slot here)
    put_rz_rr_g<CO, UP_WG>(rz, rr2, g, PC == 2 ? P.pq : P.rz + par * CG_NB, P.rr, C);
    if (restart && threadIdx.x == 0) P.rz[(par ^ 1) * CG_NB + blockIdx.x] = blockIdx.x == 0 ? INFINITY : 0.0;
}

// ---------------------------------------------------------------- factorized correction
// k_fac_lower: y = G r (rows of the lower half matrix), at starts and residual
// replacements; inside the iteration y follows y -= alpha G q in the update
// kernel.  k_fac_upper: z (+)= G^T y (rows of the upper half matrix) and the
// per-workgroup partial <r, z> into the CG slot the update kernel would have
// written.  Both run on the CG grid with GL lanes per node and return at once
// after convergence.
template <class T, int GL>
__global__ __launch_bounds__(CG_WG) void k_fac_lower(int nfree, FacView F, const T *__restrict__ r,
                                                     T *__restrict__ y, DevScalars *sc)
{
    if (sc && sc->done) return;
    constexpr int NPW = CG_WG / GL;
    const int lane = threadIdx.x % GL;
    int r0, r1;
    cg_range(nfree, r0, r1);
    for (int i = r0 + (int)threadIdx.x / GL; i < r1; i += NPW) {
        double a0, a1, a2;
        fac_row<T, GL>(i, lane, F.lo_ptr, F.lo_col, F.lo, r, a0, a1, a2);
        if (lane < 3) y[3 * i + lane] = (T)(lane == 0 ? a0 : (lane == 1 ? a1 : a2));
    }
}

// slot of the <r, z> partials: mode 0 after an update (iteration bi: the slot
// the update kernel writes, par ^ 1), mode 1 after k_replace (par0); mode 2
// writes no partials (training passes)
template <class T, int GL>
__global__ __launch_bounds__(CG_WG) void k_fac_upper(int bi, int mode, int add, int nfree, FacView F,
                                                     const T *__restrict__ y, const T *__restrict__ r,
                                                     T *__restrict__ z, CGParts P, DevScalars *sc)
{
    if (sc && sc->done) return;
    constexpr int NPW = CG_WG / GL;
    const int lane = threadIdx.x % GL;
    int r0, r1;
    cg_range(nfree, r0, r1);
    double rz = 0.0;
    for (int i = r0 + (int)threadIdx.x / GL; i < r1; i += NPW) {
        double a0, a1, a2;
        fac_row<T, GL>(i, lane, F.up_ptr, F.up_col, F.up, y, a0, a1, a2);
        if (lane < 3) {
            const int k = 3 * i + lane;
            double zv = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
            if (add) zv += (double)z[k];
            const T zt = (T)zv;
            z[k] = zt;
            rz += (double)r[k] * (double)zt;
        }
    }
    if (mode == 2) return;
    double v[1] = {rz};
    cg_block_sum<1>(v);
    if (threadIdx.x == 0) {
        const int par = sc->par0 ^ (bi & 1);
        const int slot = mode == 0 ? (par ^ 1) : sc->par0;
        P.rz[slot * CG_NB + blockIdx.x] = v[0];
    }
}

// r64 = b - y, <r64,r64> -> S_TRR, <b,b> -> S_BB
__global__ __launch_bounds__(RED_WG) void k_residual(int n, const double *__restrict__ b, const double *__restrict__ y,
                                                     double *__restrict__ r, double *partials, DevScalars *sc)
{
    const int i = blockIdx.x * RED_WG + threadIdx.x;
    double rr = 0.0, bb = 0.0;
    if (i < n) {
        double v = b[i] - y[i];
        r[i] = v;
        rr = v * v;
        bb = b[i] * b[i];
    }
    double v[2] = {rr, bb}, t[2];
    if (grid_sum<2>(v, partials, sc, t)) {
        sc->s[S_TRR] = t[0];
        sc->s[S_BB] = t[1];
    }
}

template <class A, class B>
__global__ void k_convert(long long n, const A *__restrict__ a, B *__restrict__ b)
{
    long long i = (long long)blockIdx.x * RED_WG + threadIdx.x;
    if (i < n) b[i] = (B)a[i];
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
                                float *__restrict__ minv32, double *__restrict__ dblk)
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
    if (dblk)
        for (int i = 0; i < 9; ++i) dblk[9 * r + i] = A[i];
    // the block is stored [col c][row d]; K is symmetric, so it is its own transpose
    double M[9];
    inv3(A, M);
    for (int i = 0; i < 9; ++i) {
        minv[9 * r + i] = M[i];
        if (minv32) minv32[9 * r + i] = (float)M[i];
    }
}

// Reordered system of the two-level methods: the node rows of K copied into
// the new node order (64 lanes per node; a row keeps its block order, its
// neighbour indices are renumbered on the host).
// bcol of the reordered pattern: node row i of the new order is row perm[i]
// of K with its block order kept and each neighbour renumbered by iperm
__global__ __launch_bounds__(RED_WG) void k_perm_bcol(int nfree, const int *__restrict__ perm,
                                                      const int *__restrict__ iperm, const int *__restrict__ node_blk,
                                                      const int *__restrict__ bcol, const int *__restrict__ nb_p,
                                                      int *__restrict__ bcol_p)
{
    const int i = blockIdx.x * RED_WG + (int)threadIdx.x;
    if (i >= nfree) return;
    const int o = perm[i], o0 = node_blk[o], deg = node_blk[o + 1] - o0, q0 = nb_p[i];
    for (int k = 0; k < deg; ++k) bcol_p[q0 + k] = iperm[bcol[o0 + k]];
}

__global__ __launch_bounds__(RED_WG) void k_perm_vals(int nfree, const int *__restrict__ perm,
                                                      const int *__restrict__ node_blk, const int *__restrict__ nb_p,
                                                      const double *__restrict__ val, double *__restrict__ val_p)
{
    const int node = blockIdx.x * (RED_WG / 64) + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (node >= nfree) return;
    const int old = perm[node];
    const long long s0 = 9LL * node_blk[old], d0 = 9LL * nb_p[node];
    const int len = 9 * (node_blk[old + 1] - node_blk[old]);
    for (int i = lane; i < len; i += 64) val_p[d0 + i] = val[s0 + i];
}

// dst[NV i + c] = src[NV map[i] + c] for the nfree nodes (vectors NV = 3,
// preconditioner blocks NV = 9)
template <int NV>
__global__ __launch_bounds__(RED_WG) void k_gather_nodes(int nfree, const int *__restrict__ map,
                                                         const double *__restrict__ src, double *__restrict__ dst)
{
    const long long i = (long long)blockIdx.x * RED_WG + threadIdx.x;
    if (i >= (long long)NV * nfree) return;
    const int node = (int)(i / NV), c = (int)(i - (long long)node * NV);
    dst[i] = src[(long long)NV * map[node] + c];
}

// ------------------------------------------------------------------ host side
static int ensure_ws(ofea_ctx *ctx)
{
    ofea_solver_ws *&w = ctx->ws;
    if (w && w->n == ctx->P.n && w->nnz == ctx->P.nnz) return OFEA_OK;
    solver_free(ctx);
    w = new ofea_solver_ws();
    w->n = ctx->P.n;
    w->nfree = ctx->P.nfree;
    w->nnz = ctx->P.nnz;
    const int n = w->n, nf = w->nfree;
    w->nblk_vec = (n + RED_WG - 1) / RED_WG;
    w->nblk_node = (nf + RED_WG - 1) / RED_WG;
    const size_t maxb = (size_t)nf + 64;  // SpMV with 4 lanes per node has the largest grid
    OFEA_HIP(hipMalloc((void **)&w->partials, sizeof(double) * 2 * maxb));
    OFEA_HIP(hipMalloc((void **)&w->part_rz, sizeof(double) * 2 * CG_NB));
    OFEA_HIP(hipMalloc((void **)&w->part_rr, sizeof(double) * CG_NB));
    OFEA_HIP(hipMalloc((void **)&w->part_pq, sizeof(double) * CG_NB));
    OFEA_HIP(hipMalloc((void **)&w->sc, sizeof(DevScalars)));
    OFEA_HIP(hipMemset(w->sc, 0, sizeof(DevScalars)));
    // pinned: the scalar blocks, the coarse factorization's flag, the
    // per-aggregate results of the coarse block lists
    OFEA_HIP(hipHostMalloc((void **)&w->sc_host,
                           (1 + NQ_PIPE) * sizeof(DevScalars) + 64 + sizeof(int4) * (CG_NB / CG_AGG),
                           hipHostMallocDefault));
    w->co_flag_host = reinterpret_cast<int *>(w->sc_host + 1 + NQ_PIPE);
    *w->co_flag_host = 0;
    w->co_lst_host = reinterpret_cast<int4 *>(reinterpret_cast<char *>(w->co_flag_host) + 64);
    OFEA_HIP(hipEventCreateWithFlags(&w->co_ev, hipEventDisableTiming));
    for (auto &e : w->pipe_ev) OFEA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const size_t nb = (size_t)std::max(n, 1), nfb = (size_t)std::max(nf, 1);
    for (double **p : {&w->v64.r, &w->v64.z, &w->v64.p, &w->v64.p2, &w->v64.q, &w->r64, &w->y64})
        OFEA_HIP(hipMalloc((void **)p, sizeof(double) * nb));
    for (float **p : {&w->v32.r, &w->v32.z, &w->v32.p, &w->v32.p2, &w->v32.q})
        OFEA_HIP(hipMalloc((void **)p, sizeof(float) * nb));
    OFEA_HIP(hipMalloc((void **)&w->dinv, sizeof(double) * nb));
    OFEA_HIP(hipMalloc((void **)&w->minv, sizeof(double) * 9 * nfb));
    OFEA_HIP(hipMalloc((void **)&w->minv32, sizeof(float) * 9 * nfb));
    OFEA_HIP(hipMalloc((void **)&w->dblk, sizeof(double) * 9 * nfb));
    OFEA_HIP(hipMalloc((void **)&w->val32, sizeof(float) * (size_t)std::max(ctx->P.nnz, 1LL)));
    return OFEA_OK;
}

void solver_free(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    if (!w) return;
    solver_drop_graphs(ctx);
    for (double *p : {w->part_rz, w->part_rr, w->part_pq, w->partials, w->v64.r, w->v64.z, w->v64.p, w->v64.p2, w->v64.q, w->r64, w->y64, w->dinv, w->minv,
                      w->dblk})
        if (p) (void)hipFree(p);
    for (float *p : {w->v32.r, w->v32.z, w->v32.p, w->v32.p2, w->v32.q, w->v32.y, w->minv32, w->val32})
        if (p) (void)hipFree(p);
    if (w->v64.y) (void)hipFree(w->v64.y);
    for (FacMat &f : w->fac)
        for (void *p : {(void *)f.lo_ptr, (void *)f.lo_col, (void *)f.up_ptr, (void *)f.up_col, (void *)f.lo, (void *)f.up})
            if (p) (void)hipFree(p);
    if (w->co_arena) (void)hipFree(w->co_arena);  // every two-level buffer
    if (w->sc) (void)hipFree(w->sc);
    if (w->sc_host) (void)hipHostFree(w->sc_host);
    if (w->co_ev) (void)hipEventDestroy(w->co_ev);
    for (auto &e : w->pipe_ev)
        if (e) (void)hipEventDestroy(e);
    delete w;
    ctx->ws = nullptr;
}

// the captured PCG graphs (both lengths), e.g. after the lane group changed
void solver_drop_graphs(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    if (!w) return;
    for (int i = 0; i < NSLOT; ++i) {
        w->runs[i] = 0;
        for (hipGraphExec_t *e : {&w->exec[i], &w->exec_tail[i]})
            if (*e) {
                (void)hipGraphExecDestroy(*e);
                *e = nullptr;
            }
        for (hipGraph_t *g : {&w->graph[i], &w->graph_tail[i]})
            if (*g) {
                (void)hipGraphDestroy(*g);
                *g = nullptr;
            }
    }
}

void solver_invalidate(ofea_ctx *ctx)
{
    if (ctx->ws) {
        ctx->ws->precond_valid = false;
        ctx->ws->co_valid = false;
    }
}

void solver_invalidate_values(ofea_ctx *ctx)
{
    if (ctx->ws) {
        ctx->ws->precond_valid = false;
        if (ctx->ws->co_valid) ctx->ws->co_vals_stale = true;
    }
}

// The buffers one PCG run works in: the context's solver work space, or the
// static scratch of ofea_solve_noalloc.
struct PcgBufs {
    int nfree = 0, n = 0, nblk_vec = 0, spmv_group = 16;
    double *r64 = nullptr, *y64 = nullptr;  // true residual, K x
    double *partials = nullptr;             // 2 * nblk_vec (true residual)
    CGParts P{};
    CoarseArgs co{};
    int co_nagg = 0;
    const float *co_ainv = nullptr;
    DevScalars *sc = nullptr;
};

static PcgBufs ws_bufs(ofea_solver_ws *w)
{
    PcgBufs B;
    B.nfree = w->nfree;
    B.n = w->n;
    B.nblk_vec = w->nblk_vec;
    B.spmv_group = w->spmv_group;
    B.r64 = w->r64;
    B.y64 = w->y64;
    B.partials = w->partials;
    B.P = CGParts{w->part_rz, w->part_rr, w->part_pq};
    B.co = w->co;
    B.co_nagg = w->co_nagg;
    B.co_ainv = w->co_ainv;
    B.sc = w->sc;
    return B;
}

template <class TV, class T, bool DOT>
static void launch_spmv_b(hipStream_t st, int nfree, double *partials, DevScalars *sc, int group,
                          const int *node_blk, const int *bcol, const TV *val, const T *x, T *y)
{
    int npw = RED_WG / group;
    int nb = (nfree + npw - 1) / npw;
    if (nb == 0) return;
#define OFEA_SPMV_CASE(G)                                                                                        \
    case G:                                                                                                      \
        hipLaunchKernelGGL((k_spmv<TV, T, G, DOT>), dim3(nb), dim3(RED_WG), 0, st, nfree, node_blk, bcol, val, x, \
                           y, partials, sc);                                                                     \
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

template <class TV, class T, bool DOT>
static void launch_spmv_p(ofea_ctx *ctx, int group, const int *node_blk, const int *bcol, const TV *val, const T *x,
                          T *y)
{
    ofea_solver_ws *w = ctx->ws;
    launch_spmv_b<TV, T, DOT>(ctx->stream, w->nfree, w->partials, w->sc, group, node_blk, bcol, val, x, y);
}

template <class TV, class T, bool DOT>
static void launch_spmv_t(ofea_ctx *ctx, int group, const TV *val, const T *x, T *y)
{
    launch_spmv_p<TV, T, DOT>(ctx, group, ctx->d_node_blk, ctx->d_bcol, val, x, y);
}

// The system a PCG run works on: the context's K and vectors, or the
// reordered copy of the two-level methods.
struct SysView {
    const int *node_blk, *bcol;
    const double *val64;  // FP64 values (true residual)
    const double *dinv;   // point-Jacobi inverse diagonal
    const double *b;      // right-hand side
    double *x;            // solution
    FacView fac;          // learned factor G on this pattern (PC 2)
    int fac_group;        // lanes per node of its passes
};

static FacView fac_view(const FacMat &f)
{
    return {f.lo_ptr, f.lo_col, f.up_ptr, f.up_col, f.lo, f.up};
}

static SysView ctx_view(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    return {ctx->d_node_blk, ctx->d_bcol, ctx->d_val, w->dinv,          ctx->d_rhs,
            ctx->d_x,        fac_view(w->fac[0]), w->fac[0].group ? w->fac[0].group : 8};
}

static SysView perm_view(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    return {w->nb_p, w->bcol_p, w->val_p, w->dinv_p, w->b_p, w->x_p, fac_view(w->fac[1]),
            w->fac[1].group ? w->fac[1].group : 8};
}

#define OFEA_GROUP_SWITCH(group, CASE) \
    switch (group) {                   \
        CASE(4)                        \
        CASE(8)                        \
        CASE(16)                       \
        CASE(32)                       \
        default:                       \
        CASE(64)                       \
    }

// y = G r (k_fac_lower), z (+)= G^T y (k_fac_upper); mode as in k_fac_upper
template <class T>
static void launch_fac_lower(hipStream_t st, int group, int nfree, const FacView &F, const T *r, T *y,
                             DevScalars *sc)
{
#define OFEA_FACL_CASE(GL)                                                                                 \
    case GL:                                                                                               \
        hipLaunchKernelGGL((k_fac_lower<T, GL>), dim3(CG_NB), dim3(CG_WG), 0, st, nfree, F, r, y, sc); \
        break;
    OFEA_GROUP_SWITCH(group, OFEA_FACL_CASE)
#undef OFEA_FACL_CASE
}

template <class T>
static void launch_fac_upper(hipStream_t st, int group, int bi, int mode, int add, int nfree, const FacView &F,
                             const T *y, const T *r, T *z, CGParts P, DevScalars *sc)
{
#define OFEA_FACU_CASE(GL)                                                                                       \
    case GL:                                                                                                     \
        hipLaunchKernelGGL((k_fac_upper<T, GL>), dim3(CG_NB), dim3(CG_WG), 0, st, bi, mode, add, nfree, F, y, r, \
                           z, P, sc);                                                                            \
        break;
    OFEA_GROUP_SWITCH(group, OFEA_FACU_CASE)
#undef OFEA_FACU_CASE
}

// Index arrays of the learned factor's half matrices for view 0 (K's rows) or
// 1 (the two-level order), from the host pattern; allocates the value arrays
// and the y vectors of the iteration.  Built once per mesh.
int fac_prepare(ofea_ctx *ctx, int view)
{
    ofea_solver_ws *w = ctx->ws;
    FacMat &f = w->fac[view];
    if (f.lo_ptr) return OFEA_OK;
    const int nf = w->nfree;
    const std::vector<int> &nb = ctx->P.node_blk;
    const std::vector<int> &bc = ctx->P.bcol;
    std::vector<int> perm;
    if (view == 1) {  // row i of the two-level order is K's row perm[i], columns renumbered by iperm
        perm.resize(nf);
        for (int o = 0; o < nf; ++o) perm[w->iperm_host[o]] = o;
    }
    std::vector<int> lp(nf + 1, 0), up(nf + 1, 0), lc, uc;
    lc.reserve((size_t)(nb[nf] + nf) / 2 + 1);
    uc.reserve((size_t)(nb[nf] + nf) / 2 + 1);
    for (int i = 0; i < nf; ++i) {
        const int o = view == 1 ? perm[i] : i;
        for (int q = nb[o]; q < nb[o + 1]; ++q) {
            const int s = view == 1 ? w->iperm_host[bc[q]] : bc[q];
            if (s <= i) lc.push_back(s);
            if (s >= i) uc.push_back(s);
        }
        lp[i + 1] = (int)lc.size();
        up[i + 1] = (int)uc.size();
    }
    f.nlo = (long long)lc.size();
    f.nup = (long long)uc.size();
    OFEA_HIP(hipMalloc((void **)&f.lo_ptr, sizeof(int) * (nf + 1)));
    OFEA_HIP(hipMalloc((void **)&f.up_ptr, sizeof(int) * (nf + 1)));
    OFEA_HIP(hipMalloc((void **)&f.lo_col, sizeof(int) * std::max<size_t>(lc.size(), 1)));
    OFEA_HIP(hipMalloc((void **)&f.up_col, sizeof(int) * std::max<size_t>(uc.size(), 1)));
    OFEA_HIP(hipMalloc((void **)&f.lo, sizeof(unsigned short) * 9 * std::max<size_t>(lc.size(), 1)));
    OFEA_HIP(hipMalloc((void **)&f.up, sizeof(unsigned short) * 9 * std::max<size_t>(uc.size(), 1)));
    OFEA_HIP(hipMemcpy(f.lo_ptr, lp.data(), sizeof(int) * (nf + 1), hipMemcpyHostToDevice));
    OFEA_HIP(hipMemcpy(f.up_ptr, up.data(), sizeof(int) * (nf + 1), hipMemcpyHostToDevice));
    if (!lc.empty()) OFEA_HIP(hipMemcpy(f.lo_col, lc.data(), sizeof(int) * lc.size(), hipMemcpyHostToDevice));
    if (!uc.empty()) OFEA_HIP(hipMemcpy(f.up_col, uc.data(), sizeof(int) * uc.size(), hipMemcpyHostToDevice));
    const size_t nbv = (size_t)std::max(w->n, 1);
    if (!w->v64.y) OFEA_HIP(hipMalloc((void **)&w->v64.y, sizeof(double) * nbv));
    if (!w->v32.y) OFEA_HIP(hipMalloc((void **)&w->v32.y, sizeof(float) * nbv));
    f.kver = f.gen = -1;
    return OFEA_OK;
}

// Lanes per node of the factor passes for view v: k_fac_upper timed for 4, 8
// and 16 lanes on the built factor (y and z of the FP64 vector set as
// scratch, before any solve uses them), the fastest kept.
int fac_tune(ofea_ctx *ctx, int view)
{
    ofea_solver_ws *w = ctx->ws;
    FacMat &f = w->fac[view];
    if (f.group) return OFEA_OK;
    const FacView F = fac_view(f);
    const CGParts P{nullptr, nullptr, nullptr};
    const unsigned gv = (unsigned)w->nblk_vec;
    if (gv) hipLaunchKernelGGL((k_fill<double>), dim3(gv), dim3(RED_WG), 0, ctx->stream, w->n, w->v64.y, 1.0);
    float best = -1.f;
    for (int g : {4, 8, 16}) {
        launch_fac_upper<double>(ctx->stream, g, 0, 2, 0, w->nfree, F, w->v64.y, w->v64.y, w->v64.z, P, nullptr);
        OFEA_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        for (int rep = 0; rep < 5; ++rep)
            launch_fac_upper<double>(ctx->stream, g, 0, 2, 0, w->nfree, F, w->v64.y, w->v64.y, w->v64.z, P, nullptr);
        OFEA_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        OFEA_HIP(hipEventSynchronize(ctx->ev1));
        OFEA_HIP(hipGetLastError());
        float t = 0.f;
        OFEA_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        if (best < 0.f || t < best) {
            best = t;
            f.group = g;
        }
    }
    return OFEA_OK;
}

int solver_spmv(ofea_ctx *ctx, const double *d_x, double *d_y)
{
    int rc = ensure_ws(ctx);
    if (rc) return rc;
    launch_spmv_t<double, double, false>(ctx, ctx->ws->spmv_group, ctx->d_val, d_x, d_y);
    OFEA_HIP(hipGetLastError());
    return OFEA_OK;
}

int solver_staging(ofea_ctx *ctx, double **d_x, double **d_y)
{
    int rc = ensure_ws(ctx);
    if (rc) return rc;
    *d_x = ctx->ws->r64;  // both are recomputed at the start of every solve
    *d_y = ctx->ws->y64;
    return OFEA_OK;
}

int solver_spmv_vals(ofea_ctx *ctx, const double *d_vals, const double *d_x, double *d_y)
{
    int rc = ensure_ws(ctx);
    if (rc) return rc;
    launch_spmv_t<double, double, false>(ctx, ctx->ws->spmv_group, d_vals, d_x, d_y);
    OFEA_HIP(hipGetLastError());
    return OFEA_OK;
}

int solver_spmv_group_time(ofea_ctx *ctx, int group, int reps, float *ms)
{
    int rc = ensure_ws(ctx);
    if (rc) return rc;
    ofea_solver_ws *w = ctx->ws;
    const int gv = (w->n + RED_WG - 1) / RED_WG;
    if (gv == 0 || reps < 1) return fail(OFEA_ERR_ARG, "empty system or reps < 1");
    hipLaunchKernelGGL((k_fill<double>), dim3(gv), dim3(RED_WG), 0, ctx->stream, w->n, w->r64, 1.0);
    launch_spmv_t<double, double, false>(ctx, group, ctx->d_val, w->r64, w->y64);  // warm-up
    OFEA_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < reps; ++i) launch_spmv_t<double, double, false>(ctx, group, ctx->d_val, w->r64, w->y64);
    OFEA_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    OFEA_HIP(hipEventSynchronize(ctx->ev1));
    OFEA_HIP(hipGetLastError());
    float t = 0.f;
    OFEA_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    *ms = t / reps;
    return OFEA_OK;
}

int solver_spmv32(ofea_ctx *ctx, const float *d_x, float *d_y)
{
    launch_spmv_t<float, float, false>(ctx, ctx->ws->spmv_group, ctx->ws->val32, d_x, d_y);
    OFEA_HIP(hipGetLastError());
    return OFEA_OK;
}

int solver_prepare(ofea_ctx *ctx)
{
    int rc = ensure_ws(ctx);
    if (rc) return rc;
    ofea_solver_ws *w = ctx->ws;
    if (w->precond_valid) return OFEA_OK;
    if (w->nblk_node)
        hipLaunchKernelGGL(k_precond_setup, dim3(w->nblk_node), dim3(RED_WG), 0, ctx->stream, w->nfree,
                           ctx->d_node_blk, ctx->d_bcol, ctx->d_val, w->dinv, w->minv, w->minv32, w->dblk);
    const long long nnz = ctx->P.nnz;
    if (nnz)
        hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((nnz + RED_WG - 1) / RED_WG)), dim3(RED_WG), 0,
                           ctx->stream, nnz, ctx->d_val, w->val32);
    OFEA_HIP(hipGetLastError());
    w->precond_valid = true;
    return OFEA_OK;
}

// The graph of one batch: nbatch (even) iterations (2 kernels each, 3 with
// the factorized correction) + batch end.
template <class TV, class T, class XT, int PC, bool CO>
static void cg_batch(ofea_ctx *ctx, const PcgBufs &w, const SysView &S, const TV *val, VecSet<T> &v, XT *x,
                     const T *minv, int nbatch = CG_BATCH)
{
    const CGParts P = w.P;
    for (int bi = 0; bi < nbatch; ++bi) {
#define OFEA_CG_CASE(G)                                                                                          \
    case G:                                                                                                      \
        hipLaunchKernelGGL((k_cg_spmv<TV, T, G, CO>), dim3(CG_NB), dim3(CG_WG), 0, ctx->stream, bi, w.nfree,  \
                           S.node_blk, S.bcol, val, v.z, v.p, v.p2, v.q, P, w.co, w.sc);                        \
        break;
        switch (w.spmv_group) {
            OFEA_CG_CASE(4)
            OFEA_CG_CASE(8)
            OFEA_CG_CASE(16)
            OFEA_CG_CASE(32)
            default:
            OFEA_CG_CASE(64)
        }
#undef OFEA_CG_CASE
        hipLaunchKernelGGL((k_cg_update<T, XT, PC, CO>), dim3(CG_NB), dim3(UP_WG), 0, ctx->stream, bi, w.nfree, x,
                           v.r, v.z, v.p, v.p2, v.q, S.dinv, minv, v.y, S.fac, P, w.co, w.sc);
        if constexpr (PC == 2)
            launch_fac_upper<T>(ctx->stream, S.fac_group, bi, 0, CO ? 1 : 0, w.nfree, S.fac, v.y, v.r, v.z, P, w.sc);
    }
    hipLaunchKernelGGL(k_cg_batch_end, dim3(1), dim3(64), 0, ctx->stream, w.sc, nbatch);
}

// true FP64 residual r64 = b - K64 x; ||r64||^2 -> S_TRR, ||b||^2 -> S_BB
static int true_residual(ofea_ctx *ctx, const PcgBufs &w, const SysView &S, const double *b, const double *x)
{
    launch_spmv_b<double, double, false>(ctx->stream, w.nfree, w.partials, w.sc, w.spmv_group, S.node_blk, S.bcol,
                                         S.val64, x, w.y64);
    hipLaunchKernelGGL(k_residual, dim3(w.nblk_vec), dim3(RED_WG), 0, ctx->stream, w.n, b, w.y64, w.r64, w.partials,
                       w.sc);
    OFEA_HIP(hipGetLastError());
    return OFEA_OK;
}

static int read_scalars(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    OFEA_HIP(hipMemcpyAsync(w->sc_host, w->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
    OFEA_HIP(hipStreamSynchronize(ctx->stream));
    return OFEA_OK;
}

static int write_scalars(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    OFEA_HIP(hipMemcpyAsync(w->sc, w->sc_host, sizeof(DevScalars), hipMemcpyHostToDevice, ctx->stream));
    return OFEA_OK;
}

// Residual drop between two replacements of the recursive residual by the
// true FP64 one in the methods with FP32 matrix values or vectors.
constexpr double MIXED_RU = 0.01;

// PCG with reliable updates.  x (FP64, device) starts at 0.  TV/T: matrix and
// vector precision of the iteration; PC: 0 point Jacobi, 1 block
// preconditioner, 2 factorized correction.  ru_factor > 0 requests a
// residual replacement each time the recursive residual has dropped by that
// factor since the last replacement.
//
// Host loop: graphs of CG_BATCH iterations, up to two in flight; the host
// reads each batch's scalars (pinned copy + event) while the next runs.  From
// the recursive residuals of the completed batches it predicts the
// iterations still needed; once they fit in what is in flight plus less than
// a batch, it tops up with graphs of CG_TAIL iterations instead, so that
// little work is queued past convergence.  The stopping targets are set on the
// device from ||b||^2, so the start needs no round trip; at the end the true
// residual is queued behind the drained batches and read with one sync.
template <class TV, class T, int PC, bool CO>
static int pcg_run(ofea_ctx *ctx, const SysView &Sv, int slot, const TV *val, VecSet<T> &v, const T *minv,
                   double rtol, int maxit, double ru_factor, int *iters, double *relres)
{
    ofea_solver_ws *w = ctx->ws;
    const PcgBufs B = ws_bufs(w);
    const double *b = Sv.b;
    double *x = Sv.x;
    const unsigned gv = w->nblk_vec;
    const int n = w->n;
    CGParts P{w->part_rz, w->part_rr, w->part_pq};
    DevScalars &S = *w->sc_host;
    hipLaunchKernelGGL((k_fill<double>), dim3(gv), dim3(RED_WG), 0, ctx->stream, n, x, 0.0);
    hipLaunchKernelGGL((k_fill<T>), dim3(gv), dim3(RED_WG), 0, ctx->stream, n, v.p, T(0));
    hipLaunchKernelGGL((k_fill<T>), dim3(gv), dim3(RED_WG), 0, ctx->stream, n, v.p2, T(0));
    S = DevScalars{};
    S.maxit = maxit;
    if (int rc = write_scalars(ctx)) return rc;
    int rc = true_residual(ctx, B, Sv, b, x);  // r64 = b and ||b||^2
    if (rc) return rc;
    hipLaunchKernelGGL(k_set_tol, dim3(1), dim3(64), 0, ctx->stream, w->sc, rtol * rtol,
                       ru_factor > 0.0 ? ru_factor * ru_factor : -1.0);
    auto start = [&](int restart) {  // r from r64, z = M r and its partials
        hipLaunchKernelGGL((k_replace<T, PC, CO>), dim3(CG_NB), dim3(UP_WG), 0, ctx->stream, w->nfree, w->r64, v.r,
                           v.z, restart, Sv.dinv, minv, P, w->co, w->sc);
        if (PC == 2) {  // y = G r, z = G^T y and <r, z> of slot par0
            launch_fac_lower<T>(ctx->stream, Sv.fac_group, w->nfree, Sv.fac, v.r, v.y, w->sc);
            launch_fac_upper<T>(ctx->stream, Sv.fac_group, 0, 1, 0, w->nfree, Sv.fac, v.y, v.r, v.z, P, w->sc);
        }
        if (CO)
            hipLaunchKernelGGL((k_coarse<T>), dim3(w->co_nagg), dim3(CG_WG), 0, ctx->stream, w->nfree, w->co_nagg,
                               w->co_ainv, v.r, v.z, w->part_rz, w->co, w->sc);
    };
    start(1);
    OFEA_HIP(hipGetLastError());
    // OFEA_NO_GRAPH=1 launches the batches directly (for tools that cannot follow graph launches)
    static const bool no_graph = std::getenv("OFEA_NO_GRAPH") && std::getenv("OFEA_NO_GRAPH")[0] == '1';
    // a slot's graphs are captured on its second solve: the first one (often
    // the only solve of a new K) launches its batches directly and pays no
    // capture and instantiation
    const bool direct = no_graph || (!w->exec[slot] && w->runs[slot]++ == 0);
    if (!w->exec[slot] && !direct) {
        OFEA_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        cg_batch<TV, T, double, PC, CO>(ctx, B, Sv, val, v, x, minv, CG_BATCH);
        OFEA_HIP(hipStreamEndCapture(ctx->stream, &w->graph[slot]));
        OFEA_HIP(hipGraphInstantiate(&w->exec[slot], w->graph[slot], nullptr, nullptr, 0));
        OFEA_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        cg_batch<TV, T, double, PC, CO>(ctx, B, Sv, val, v, x, minv, CG_TAIL);
        OFEA_HIP(hipStreamEndCapture(ctx->stream, &w->graph_tail[slot]));
        OFEA_HIP(hipGraphInstantiate(&w->exec_tail[slot], w->graph_tail[slot], nullptr, nullptr, 0));
    }
    double rel = 1.0;
    int tighten = 0;
    DevScalars *Hb = w->sc_host + 1;
    int qh[NQ_PIPE], qn[NQ_PIPE];  // ring of batches in flight: scalar slot, iterations
    int q0 = 0, qc = 0, next = 0, inflight = 0;
    auto launch_batch = [&](int nit) -> int {
        const int hb = next;
        next = (next + 1) % NQ_PIPE;
        if (direct) {
            cg_batch<TV, T, double, PC, CO>(ctx, B, Sv, val, v, x, minv, nit);
            OFEA_HIP(hipGetLastError());
        } else {
            OFEA_HIP(hipGraphLaunch(nit == CG_BATCH ? w->exec[slot] : w->exec_tail[slot], ctx->stream));
        }
        OFEA_HIP(hipMemcpyAsync(&Hb[hb], w->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
        OFEA_HIP(hipEventRecord(w->pipe_ev[hb], ctx->stream));
        qh[(q0 + qc) % NQ_PIPE] = hb;
        qn[(q0 + qc) % NQ_PIPE] = nit;
        ++qc;
        inflight += nit;
        return OFEA_OK;
    };
    // the prediction below aims at the next stop of either kind (convergence
    // or a residual replacement), so the replacements of the mixed methods
    // waste little of what is in flight either
    const int depth = 2;
    double lr = 0.0;         // log decay of ||r||^2 per iteration, kept across replacements
    double rr_start = -1.0;  // ||r||^2 where a round starts, when the host knows it
    auto top_up = [&](double need) -> int {  // batches until `need` iterations are in flight
        const int want = (int)std::min<double>(need, (double)depth * CG_BATCH);
        while (inflight < want && qc < NQ_PIPE - 1)
            if (int r2 = launch_batch(want - inflight >= CG_BATCH ? CG_BATCH : CG_TAIL)) return r2;
        return OFEA_OK;
    };
    int it_start = 0;  // iteration at which a round starts
    for (;;) {
        double last_rr = rr_start;  // residual at the last completed batch (or the round's start)
        int last_it = it_start;
        {
            const double tgt = std::max(S.s[S_TOL2], S.s[S_RU2]);
            const double need = lr < 0.0 && rr_start > 0.0 && tgt > 0.0 ? std::log(tgt / rr_start) / lr + 1.0 : 1e30;
            if ((rc = top_up(need))) return rc;
            if (qc == 0 && (rc = launch_batch(CG_TAIL))) return rc;
        }
        for (;;) {
            const int hb = qh[q0];
            inflight -= qn[q0];
            q0 = (q0 + 1) % NQ_PIPE;
            --qc;
            OFEA_HIP(hipEventSynchronize(w->pipe_ev[hb]));
            const DevScalars &D = Hb[hb];
            if (D.done) break;
            const double rr = D.s[S_RR], tol2 = std::max(D.s[S_TOL2], D.s[S_RU2]);  // the nearer stop
            if (last_rr < 0.0 && D.s[S_BB] > 0.0) last_rr = D.s[S_BB];  // first round: r_0 = b
            if (last_rr > 0.0 && rr > 0.0 && rr < last_rr && D.it0 > last_it)
                lr = std::log(rr / last_rr) / (D.it0 - last_it);
            last_rr = rr;
            last_it = D.it0;
            // iterations still needed from this point (unknown: a full pipeline)
            const double need = lr < 0.0 && tol2 > 0.0 ? std::log(tol2 / rr) / lr + 1.0 : 1e30;
            if ((rc = top_up(need))) return rc;
            if (qc == 0 && (rc = launch_batch(CG_TAIL))) return rc;  // keep going past a wrong prediction
        }
        // the batches still in flight return at once from every kernel; the
        // true FP64 residual follows them in the stream
        if ((rc = true_residual(ctx, B, Sv, b, x))) return rc;
        OFEA_HIP(hipGetLastError());
        if ((rc = read_scalars(ctx))) return rc;
        q0 = qc = inflight = 0;
        const int done = S.done;
        const double bb = S.s[S_BB];
        if (bb == 0.0) {  // b = 0: x = 0 is exact
            if (iters) *iters = 0;
            if (relres) *relres = 0.0;
            return OFEA_OK;
        }
        if (done == 3) return fail(OFEA_ERR_STATE, "PCG breakdown (matrix or preconditioner not SPD, or NaN)");
        const int it_done = S.it, par_done = S.par0 ^ (S.bi_done & 1);
        {  // the round's own decay, for the next round's first launches
            const double r0 = rr_start > 0.0 ? rr_start : bb;
            if (S.s[S_RR] > 0.0 && S.s[S_RR] < r0 && it_done > it_start) lr = std::log(S.s[S_RR] / r0) / (it_done - it_start);
        }
        rel = std::sqrt(S.s[S_TRR] / bb);
        if (rel <= rtol || done == 2) break;
        if (done == 1 && ++tighten > 8) break;  // recursive and true residual disagree persistently
        if (done == 1) S.s[S_TOL2] *= 0.25;     // recursive residual converged but the true one did not
        // replace the recursive residual, keep the search direction, resume at it_done
        S.par0 = par_done;
        S.it0 = it_done;
        S.done = 0;
        S.ticket = 0;
        S.s[S_RU2] = ru_factor > 0.0 ? ru_factor * ru_factor * S.s[S_TRR] : -1.0;
        rr_start = S.s[S_TRR];
        it_start = it_done;
        if ((rc = write_scalars(ctx))) return rc;
        start(0);
        OFEA_HIP(hipGetLastError());
    }
    if (iters) *iters = S.it;
    if (relres) *relres = rel;
    return rel <= rtol ? OFEA_OK : OFEA_NOT_CONVERGED;
}

// Recursive coordinate bisection of b[0, m) into parts of exactly L nodes
// (the last part may be shorter): split across the longest axis of the set's
// bounding box after the first half of its parts; a part keeps its nodes in
// ascending index order.
// Nodes as (coordinate on the split axis, index) keys: one total order
// (coordinate, then index), kept in contiguous arrays so that the sorts and
// selections below do not chase node indices into X.
struct AxisKey {
    double x;
    int i;
    bool operator<(const AxisKey &o) const { return x < o.x || (x == o.x && i < o.i); }
};

// Cuts b[0..m) into parts of L nodes (the last may be shorter) by recursive
// coordinate bisection across the longest side of each box; each part ends
// sorted by index.  The parts depend only on the total order, not on the
// arrangement nth_element leaves inside each side.
static void rcb_parts(int *b, int m, int L, const double *X, std::vector<AxisKey> &key)
{
    const int parts = (m + L - 1) / L;
    if (parts <= 1) {
        std::sort(b, b + m);
        return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < m; ++i)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], X[3 * (size_t)b[i] + c]);
            hi[c] = std::max(hi[c], X[3 * (size_t)b[i] + c]);
        }
    int ax = 0;
    for (int c = 1; c < 3; ++c)
        if (hi[c] - lo[c] > hi[ax] - lo[ax]) ax = c;
    const int lm = (parts / 2) * L;
    key.resize(m);
    for (int i = 0; i < m; ++i) key[i] = {X[3 * (size_t)b[i] + ax], b[i]};
    std::nth_element(key.begin(), key.begin() + lm, key.begin() + m);
    for (int i = 0; i < m; ++i) b[i] = key[i].i;
    rcb_parts(b, lm, L, X, key);
    rcb_parts(b + lm, m - lm, L, X, key);
}

// runs fn(t) for t in [0, n) on n host threads (t = 0 on the caller's)
template <class F>
static void run_threads(int n, F &&fn)
{
    std::vector<std::thread> pool;
    for (int t = 1; t < n; ++t) pool.emplace_back([&fn, t] { fn(t); });
    fn(0);
    for (auto &th : pool) th.join();
}

static int host_threads(int cap)
{
    return std::max(1, std::min(cap, std::min(16, (int)std::max(1u, std::thread::hardware_concurrency()))));
}

// slab[i] = the slab of free node i, slabs being runs of S nodes (the last may
// be shorter) in the total order (X[3 i + ax], i).  The nodes are binned on
// the axis (the bin is monotone in the coordinate), so a slab boundary falls
// inside at most one bin and only the nodes of such bins are sorted.
static void slab_of_nodes(int nf, const double *X, int ax, double lo, double hi, long long S, std::vector<int> &slab)
{
    constexpr int NBIN = 1 << 14;
    const double scale = hi > lo ? NBIN / (hi - lo) : 0.0;
    slab.resize(nf);
    std::vector<int> cnt(NBIN + 1, 0), cut_of(NBIN, -1);
    for (int i = 0; i < nf; ++i) {
        const double t = (X[3 * (size_t)i + ax] - lo) * scale;
        const int b = t >= NBIN ? NBIN - 1 : (t > 0.0 ? (int)t : 0);
        slab[i] = b;  // the bin, for now
        cnt[b + 1]++;
    }
    for (int b = 0; b < NBIN; ++b) cnt[b + 1] += cnt[b];  // cnt[b]: nodes before bin b
    int ncut = 0;
    for (long long k = S; k < nf; k += S) {
        const int b = (int)(std::upper_bound(cnt.begin(), cnt.end(), (int)k) - cnt.begin()) - 1;
        if (cnt[b] < k && cut_of[b] < 0) cut_of[b] = ncut++;
    }
    std::vector<std::vector<AxisKey>> inbin(ncut);
    std::vector<int> cut_bin(ncut);
    for (int b = 0; b < NBIN; ++b)
        if (cut_of[b] >= 0) cut_bin[cut_of[b]] = b;
    for (int i = 0; i < nf; ++i) {
        const int b = slab[i];
        if (cut_of[b] >= 0) {
            const double x = X[3 * (size_t)i + ax];
            inbin[cut_of[b]].push_back({x == 0.0 ? 0.0 : x, i});  // -0 and +0 compare equal, as for <
        } else {
            slab[i] = (int)(cnt[b] / S);  // the whole bin lies in one slab
        }
    }
    for (int j = 0; j < ncut; ++j) {
        std::vector<AxisKey> &k = inbin[j];
        std::sort(k.begin(), k.end());
        for (size_t r = 0; r < k.size(); ++r) slab[k[r].i] = (int)((cnt[cut_bin[j]] + (long long)r) / S);
    }
}

// ---- coarse-space kernels (device set-up of A_c^-1 for each K) ----
// column q of Z at a node with offset dv from its aggregate centroid:
// translations e_q (q < 3), rotations R e_(q-3) with R theta = theta x d
__device__ inline void zcol(const float *dv, int q, double z[3])
{
    const double x = dv[0], y = dv[1], w = dv[2];
    z[0] = q == 0 ? 1.0 : q == 4 ? w : q == 5 ? -y : 0.0;
    z[1] = q == 1 ? 1.0 : q == 3 ? -w : q == 5 ? x : 0.0;
    z[2] = q == 2 ? 1.0 : q == 3 ? y : q == 4 ? -x : 0.0;
}

constexpr int CA_MAXCH = 160;  // chunks, and coupled aggregates, per aggregate (k_co_lists checks)

// ---- device part of the coarse structure (once per mesh) ----
// free-node coordinates in the caller's free numbering
__global__ __launch_bounds__(RED_WG) void k_co_xfree(int nd, const int *__restrict__ resort,
                                                     const double *__restrict__ xyz, double *__restrict__ xf)
{
    const int n = blockIdx.x * RED_WG + (int)threadIdx.x;
    if (n >= nd) return;
    const int r = resort[n];
    if (r < 0) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) xf[3 * (size_t)r + c] = xyz[3 * (size_t)n + c];
}

// iperm, and the row lengths of the renumbered pattern in nb_p[i + 1]
__global__ __launch_bounds__(RED_WG) void k_co_rows(int nf, const int *__restrict__ perm,
                                                    const int *__restrict__ node_blk, int *__restrict__ iperm,
                                                    int *__restrict__ nb_p)
{
    const int i = blockIdx.x * RED_WG + (int)threadIdx.x;
    if (i == 0) nb_p[0] = 0;
    if (i >= nf) return;
    const int o = perm[i];
    iperm[o] = i;
    nb_p[i + 1] = node_blk[o + 1] - node_blk[o];
}

// In-place inclusive scan of a[0 .. n) (integers, exact): k_scan_tiles scans
// each tile of SC_TILE entries (rounds of SC_WG: wave scans, then the wave
// totals) and stores the tile's total; k_scan_add adds the totals of the
// earlier tiles to each tile.
constexpr int SC_WG = 1024, SC_TILE = 4 * SC_WG;
__device__ inline int block_scan_incl(int v, int *s_w)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int dd = 1; dd < 64; dd <<= 1) {
        const int y = __shfl_up(v, dd);
        if (lane >= dd) v += y;
    }
    if (lane == 63) s_w[wv] = v;
    __syncthreads();
    int off = 0;
    for (int w2 = 0; w2 < wv; ++w2) off += s_w[w2];
    __syncthreads();  // s_w is reused by the next call
    return v + off;
}

__global__ __launch_bounds__(SC_WG) void k_scan_tiles(int n, int *__restrict__ a, int *__restrict__ tot)
{
    __shared__ int s_w[SC_WG / 64];
    __shared__ int s_carry;
    const int t0 = blockIdx.x * SC_TILE;
    if (threadIdx.x == 0) s_carry = 0;
    for (int u = 0; u < SC_TILE / SC_WG; ++u) {
        const int j = t0 + u * SC_WG + (int)threadIdx.x;
        const int v = j < n ? a[j] : 0;
        __syncthreads();  // s_carry of the previous round
        const int carry = s_carry;
        const int incl = block_scan_incl(v, s_w);
        if (j < n) a[j] = carry + incl;
        if (threadIdx.x == SC_WG - 1) s_carry = carry + incl;
    }
    __syncthreads();
    if (threadIdx.x == 0) tot[blockIdx.x] = s_carry;
}

__global__ __launch_bounds__(SC_WG) void k_scan_add(int n, int *__restrict__ a, const int *__restrict__ tot)
{
    __shared__ int s_off;
    if (threadIdx.x == 0) {
        int off = 0;
        for (int t = 0; t < (int)blockIdx.x; ++t) off += tot[t];
        s_off = off;
    }
    __syncthreads();
    const int off = s_off;
    for (int j = blockIdx.x * SC_TILE + (int)threadIdx.x; j < min(n, (int)(blockIdx.x + 1) * SC_TILE); j += SC_WG)
        a[j] += off;
}

// per aggregate a: the centroid of its nodes (a fixed-order sum) and the
// offsets d_i = x_i - centroid in FP32, as the kernels apply Z
constexpr int CD_WG = 256;
__global__ __launch_bounds__(CD_WG) void k_co_offsets(int nf, int L, const int *__restrict__ perm,
                                                      const double *__restrict__ xf, float *__restrict__ d)
{
    __shared__ double s_c[3][CD_WG / 64];
    const int a = blockIdx.x, i0 = a * L, i1 = min(nf, i0 + L), t = threadIdx.x;
    double c3[3] = {0.0, 0.0, 0.0};
    for (int i = i0 + t; i < i1; i += CD_WG)
#pragma unroll
        for (int c = 0; c < 3; ++c) c3[c] += xf[3 * (size_t)perm[i] + c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double v = c3[c];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((t & 63) == 0) s_c[c][t >> 6] = v;
    }
    __syncthreads();
    double cen[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double v = 0.0;
        for (int w = 0; w < CD_WG / 64; ++w) v += s_c[c][w];
        cen[c] = v / max(i1 - i0, 1);
    }
    for (int i = i0 + t; i < i1; i += CD_WG)
#pragma unroll
        for (int c = 0; c < 3; ++c) d[3 * (size_t)i + c] = (float)(xf[3 * (size_t)perm[i] + c] - cen[c]);
}

// Block lists of A_c for aggregate b (one workgroup; the blocks of b's nodes
// are [q0, q1) of the renumbered pattern): the coupled aggregates a
// ("slots", ascending), the blocks grouped by the slot of their column's
// aggregate -- in node and block order within a slot, written to b's own
// range of sl_blk as (node << 8) | slot in the node's row -- and the slot
// lists cut into at most CA_MAXCH chunks (chunk size 64 or more, doubled
// until the chunks fit).  lst[b] = (slots, chunks, largest |a - b|); slots >
// CA_MAXCH means the lists do not fit (the host reports it).
constexpr int CL_WG = 1024;
constexpr int CL_NAGG = CG_NB / CG_AGG;  // aggregates at most
__global__ __launch_bounds__(CL_WG) void k_co_lists(int nf, int L, int nagg, const int *__restrict__ nb_p,
                                                    const int *__restrict__ bcol_p, int *__restrict__ sl_blk,
                                                    int *__restrict__ sl_part, int *__restrict__ sl_cptr,
                                                    int *__restrict__ ch_beg, int4 *__restrict__ lst)
{
    __shared__ int s_flag[CL_NAGG], s_slot[CL_NAGG];
    __shared__ int s_cnt[CA_MAXCH], s_base[CA_MAXCH + 1], s_run[CA_MAXCH];
    __shared__ int s_wc[CL_WG / 64][CA_MAXCH];
    __shared__ int s_ns, s_dmax;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int i0 = b * L, i1 = min(nf, i0 + L), q0 = nb_p[i0], q1 = nb_p[i1];
    for (int t = tid; t < CL_NAGG; t += CL_WG) s_flag[t] = 0;
    for (int t = tid; t < CA_MAXCH; t += CL_WG) s_cnt[t] = 0;
    __syncthreads();
    for (int q = q0 + tid; q < q1; q += CL_WG) s_flag[bcol_p[q] / L] = 1;
    __syncthreads();
    if (wv == 0) {  // slots: ranks of the flagged aggregates, CL_NAGG / 64 per lane
        constexpr int U = CL_NAGG / 64;
        int f[U], c = 0, dmax = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int a = U * lane + u;
            f[u] = a < nagg ? s_flag[a] : 0;
            c += f[u];
            if (f[u]) dmax = max(dmax, abs(a - b));
        }
        int incl = c;
        for (int dd = 1; dd < 64; dd <<= 1) {
            const int y = __shfl_up(incl, dd);
            if (lane >= dd) incl += y;
        }
        for (int o = 32; o > 0; o >>= 1) dmax = max(dmax, __shfl_xor(dmax, o));
        int base = incl - c;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int a = U * lane + u;
            s_slot[a] = f[u] ? base : -1;
            if (f[u]) {
                if (base < CA_MAXCH) sl_part[b * CA_MAXCH + base] = a;
                ++base;
            }
        }
        if (lane == 63) {
            s_ns = incl;
            s_dmax = dmax;
        }
    }
    __syncthreads();
    const int ns = s_ns;
    if (ns > CA_MAXCH) {
        if (tid == 0) lst[b] = make_int4(ns, 0, s_dmax, 0);
        return;
    }
    for (int q = q0 + tid; q < q1; q += CL_WG) atomicAdd(&s_cnt[s_slot[bcol_p[q] / L]], 1);  // counts: order-free
    __syncthreads();
    if (tid == 0) {  // slot starts, chunks (ns <= CA_MAXCH: a short serial loop)
        int run = 0;
        for (int t = 0; t < ns; ++t) {
            s_base[t] = run;
            s_run[t] = run;
            run += s_cnt[t];
        }
        s_base[ns] = run;
        int cs = max(64, (q1 - q0 + CA_MAXCH - 1) / CA_MAXCH);
        for (;; cs *= 2) {
            int nch = 0;
            for (int t = 0; t < ns; ++t) nch += (s_cnt[t] + cs - 1) / cs;
            if (nch <= CA_MAXCH) break;
        }
        const int cb = b * (CA_MAXCH + 1);
        int c = 0;
        for (int t = 0; t < ns; ++t) {
            sl_cptr[cb + t] = c;
            for (int e = s_base[t]; e < s_base[t + 1]; e += cs) ch_beg[cb + c++] = q0 + e;
        }
        sl_cptr[cb + ns] = c;
        ch_beg[cb + c] = q1;
        lst[b] = make_int4(ns, c, s_dmax, 0);
    }
    // stable scatter, CL_WG consecutive blocks at a time: a block's place in
    // its slot = the slot's running base + the same-slot blocks of the
    // earlier waves + those of the earlier lanes of its own wave
    for (int c0 = q0; c0 < q1; c0 += CL_WG) {
        for (int t = tid; t < (CL_WG / 64) * ns; t += CL_WG) s_wc[t / ns][t % ns] = 0;
        __syncthreads();
        const int q = c0 + tid;
        const bool act = q < q1;
        const int sl = act ? s_slot[bcol_p[q] / L] : -1;
        int rank = 0;
        unsigned long long todo = __ballot(act);
        while (todo) {  // one pass per distinct slot of the wave
            const int leader = __ffsll((long long)todo) - 1;
            const int s1 = __shfl(sl, leader);
            const unsigned long long m = __ballot(act && sl == s1);
            if (act && sl == s1) rank = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == leader) s_wc[wv][s1] = __popcll(m);
            todo &= ~m;
        }
        __syncthreads();
        if (act) {
            int pos = s_run[sl] + rank;
            for (int w2 = 0; w2 < wv; ++w2) pos += s_wc[w2][sl];
            int lo = i0, hi = i1 - 1;  // row node of block q: nb_p[r] <= q < nb_p[r + 1]
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (nb_p[mid] <= q) lo = mid;
                else hi = mid - 1;
            }
            sl_blk[q0 + pos] = (lo << 8) | (q - nb_p[lo]);
        }
        __syncthreads();
        for (int t = tid; t < ns; t += CL_WG) {
            int add = 0;
            for (int w2 = 0; w2 < CL_WG / 64; ++w2) add += s_wc[w2][t];
            s_run[t] += add;
        }
        __syncthreads();
    }
}

// A_c = Z^T K Z on the reordered K: workgroup b owns the column blocks of
// aggregate b.  The blocks (r, k) of b's nodes are listed per coupled
// aggregate a ("slot") and the lists are cut into chunks.  One wave per
// chunk: lane l takes the chunk's blocks l, l + 64, ..., forms Z_s^T K_sr Z_r
// (6 x 6; s the block's column node in a, r its row node in b) and sums them,
// and the 36 sums are added over the wave by a fixed butterfly into LDS.  Then
// one thread per (slot, p, q) adds the slot's chunk sums in order and writes
// A[6a+p][6b+q] (fixed order, no atomics).
constexpr int CA_WG = 512;  // 256 VGPRs per lane: the 36 sums and the block stay in registers
__global__ __launch_bounds__(CA_WG) void k_coarse_ac(int nc, const int *__restrict__ nb_p,
                                                     const int *__restrict__ bcol_p, const double *__restrict__ val_p,
                                                     const float *__restrict__ d, const int4 *__restrict__ lst,
                                                     const int *__restrict__ sl_part, const int *__restrict__ sl_cptr,
                                                     const int *__restrict__ ch_beg, const int *__restrict__ sl_blk,
                                                     double *__restrict__ A)
{
    __shared__ double s_part[CA_MAXCH * 36];
    const int b = blockIdx.x;
    const int4 ls = lst[b];  // slots, chunks
    if (ls.x > CA_MAXCH) return;  // lists that do not fit: the host reports it
    const int cb = b * (CA_MAXCH + 1), nch = ls.y;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int ch = wv; ch < nch; ch += CA_WG / 64) {
        double acc[36];
#pragma unroll
        for (int t = 0; t < 36; ++t) acc[t] = 0.0;
        for (int e = ch_beg[cb + ch] + lane; e < ch_beg[cb + ch + 1]; e += 64) {
            const int rk = sl_blk[e];  // (node << 8) | slot in the node's row
            const int r = rk >> 8, k = rk & 255;
            const int nb0 = nb_p[r], deg = nb_p[r + 1] - nb0;
            const int sn = bcol_p[nb0 + k];
            const double *blk = val_p + 9LL * nb0 + 3 * k;  // blk[c * 3 deg + m] = K(3s + m, 3r + c)
            double kb[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int m = 0; m < 3; ++m) kb[m][c] = blk[c * 3 * deg + m];
            const double rx = d[3 * (size_t)r], ry = d[3 * (size_t)r + 1], rz = d[3 * (size_t)r + 2];
            const double sx = d[3 * (size_t)sn], sy = d[3 * (size_t)sn + 1], sz = d[3 * (size_t)sn + 2];
            // T = K_sr Z_r (3 x 6): translations, then the rotation columns
            // (0, -z, y), (z, 0, -x), (-y, x, 0) of the row node's offset
            double T[3][6];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                T[m][0] = kb[m][0];
                T[m][1] = kb[m][1];
                T[m][2] = kb[m][2];
                T[m][3] = -rz * kb[m][1] + ry * kb[m][2];
                T[m][4] = rz * kb[m][0] - rx * kb[m][2];
                T[m][5] = -ry * kb[m][0] + rx * kb[m][1];
            }
            // rows of Z_s^T: the same columns of the column node's offset
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                acc[q] += T[0][q];
                acc[6 + q] += T[1][q];
                acc[12 + q] += T[2][q];
                acc[18 + q] += -sz * T[1][q] + sy * T[2][q];
                acc[24 + q] += sz * T[0][q] - sx * T[2][q];
                acc[30 + q] += -sy * T[0][q] + sx * T[1][q];
            }
        }
#pragma unroll
        for (int t = 0; t < 36; ++t) {
            double v = acc[t];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (lane == 0) s_part[ch * 36 + t] = v;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 36 * ls.x; t += CA_WG) {
        const int sl = t / 36, pq = t % 36;
        double acc = 0.0;
        for (int ch = sl_cptr[cb + sl]; ch < sl_cptr[cb + sl + 1]; ++ch) acc += s_part[ch * 36 + pq];
        A[(size_t)(6 * sl_part[b * CA_MAXCH + sl] + pq / 6) * nc + 6 * b + pq % 6] = acc;
    }
}

// In-place banded Cholesky of the lower band of A (nc x nc row-major,
// half-bandwidth bw; symmetrised beforehand by k_band_sym) in one workgroup.
// The diagonal is first lifted by 1e-12 tr(A)/nc (a Z column that vanishes,
// e.g. a rotation of a collinear aggregate, leaves a zero row).
// Panels of CH_NB columns are factored in LDS (rows of the panel and of the
// band below it), written back, and applied to the trailing band as a rank-
// CH_NB update in 4x4 tiles; a band wider than the LDS panel falls back to
// column-by-column updates in global memory.  Afterwards the lower triangle
// holds L (k_band_upper then copies L^T into the upper triangle).  *flag = 0,
// or j + 1 when the pivot of column j is not positive.
// orders the LDS accesses of one wave's lanes around it (LDS requests of a
// wave complete in issue order)
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int CH_WG = 1024;
constexpr int CH_NB = 16;                       // panel width
constexpr int CH_ROWS = 460;                    // panel rows held in LDS (bw + CH_NB <= CH_ROWS)
constexpr int CH_LD = CH_NB + 1;                // LDS row stride: odd, so a column spreads over the banks
__global__ __launch_bounds__(CH_WG) void k_band_cholesky(int nc, int bw, double *__restrict__ A, int *__restrict__ flag)
{
    __shared__ double s_p[CH_ROWS * CH_LD];  // panel, row-major [row][column], stride CH_LD
    __shared__ double s_red[CH_WG / 64];
    __shared__ double s_piv;
    __shared__ int s_bad;
    const int tid = threadIdx.x;
    double tr = 0.0;
    for (int i = tid; i < nc; i += CH_WG) tr += A[(size_t)i * nc + i];
    for (int o = 32; o > 0; o >>= 1) tr += __shfl_xor(tr, o);
    if ((tid & 63) == 0) s_red[tid >> 6] = tr;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < CH_WG / 64; ++w) t += s_red[w];
        s_piv = 1e-12 * t / nc;
        s_bad = 0;
    }
    __syncthreads();
    const double lift = s_piv;
    for (int i = tid; i < nc; i += CH_WG) A[(size_t)i * nc + i] += lift;
    __syncthreads();
    if (bw + CH_NB <= CH_ROWS) {
        for (int j0 = 0; j0 < nc; j0 += CH_NB) {
            const int nb = min(CH_NB, nc - j0);
            const int H = min(nc, j0 + nb + bw) - j0;  // panel rows (<= CH_ROWS)
            for (int t = tid; t < H * CH_NB; t += CH_WG) {
                const int r = t / CH_NB, c = t % CH_NB;
                s_p[r * CH_LD + c] = (r >= c && c < nb) ? A[(size_t)(j0 + r) * nc + j0 + c] : 0.0;
            }
            __syncthreads();
            if (tid < 64) {
                // the panel's nb x nb diagonal block by wave 0, lane l owning row l in
                // registers (pivots and L[c2][c] broadcast by read_lane)
                double v[CH_NB];
#pragma unroll
                for (int c = 0; c < CH_NB; ++c) v[c] = (tid < nb && c <= tid) ? s_p[tid * CH_LD + c] : 0.0;
#pragma unroll
                for (int c = 0; c < CH_NB; ++c) {
                    if (c < nb) {
                        const double dc = read_lane(v[c], c);
                        const double piv = dc > 0.0 ? sqrt(dc) : 1.0;
                        if (tid == 0 && !(dc > 0.0) && !s_bad) s_bad = j0 + c + 1;
                        if (tid == c) v[c] = piv;
                        if (tid > c) v[c] /= piv;
#pragma unroll
                        for (int c2 = c + 1; c2 < CH_NB; ++c2) {
                            const double lc2 = read_lane(v[c], c2);
                            if (tid >= c2 && c2 < nb) v[c2] -= v[c] * lc2;
                        }
                    }
                }
                if (tid < nb)
#pragma unroll
                    for (int c = 0; c < CH_NB; ++c)
                        if (c <= tid) s_p[tid * CH_LD + c] = v[c];
            }
            __syncthreads();
            if (s_bad) {
                if (tid == 0) *flag = s_bad;
                return;
            }
            // rows below the diagonal block: x L_diag^T = a, one row per thread
            for (int r = nb + tid; r < H; r += CH_WG) {
                double x[CH_NB];
#pragma unroll
                for (int c = 0; c < CH_NB; ++c) x[c] = s_p[r * CH_LD + c];
#pragma unroll
                for (int c = 0; c < CH_NB; ++c) {
                    if (c < nb) {
                        double t = x[c];
#pragma unroll
                        for (int c2 = 0; c2 < c; ++c2) t -= x[c2] * s_p[c * CH_LD + c2];
                        x[c] = t / s_p[c * CH_LD + c];
                    }
                }
#pragma unroll
                for (int c = 0; c < CH_NB; ++c) s_p[r * CH_LD + c] = x[c];
            }
            __syncthreads();
            for (int t = tid; t < H * CH_NB; t += CH_WG) {
                const int r = t / CH_NB, c = t % CH_NB;
                if (r >= c && c < nb) A[(size_t)(j0 + r) * nc + j0 + c] = s_p[r * CH_LD + c];
            }
            // trailing band: rows/columns nb .. H-1 of the panel window, 4x4 tiles of the lower triangle
            const int m = H - nb, T = (m + 3) / 4;
            for (int t = tid; t < T * T; t += CH_WG) {
                const int ti = t / T, tk = t - ti * T;
                if (tk > ti) continue;
                double acc[4][4];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int bq = 0; bq < 4; ++bq) acc[a][bq] = 0.0;
#pragma unroll 4
                for (int c = 0; c < CH_NB; ++c) {  // columns past nb are zero in the panel
                    double li[4], lk[4];
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const int ri = nb + 4 * ti + a, rk = nb + 4 * tk + a;
                        li[a] = ri < H ? s_p[ri * CH_LD + c] : 0.0;
                        lk[a] = rk < H ? s_p[rk * CH_LD + c] : 0.0;
                    }
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int bq = 0; bq < 4; ++bq) acc[a][bq] += li[a] * lk[bq];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int bq = 0; bq < 4; ++bq) {
                        const int i = j0 + nb + 4 * ti + a, k = j0 + nb + 4 * tk + bq;
                        if (i < j0 + H && k <= i) A[(size_t)i * nc + k] -= acc[a][bq];
                    }
            }
            __syncthreads();
        }
    } else {
        double *s_l = s_p;  // CH_ROWS * CH_NB >= CG_NC_MAX
        static_assert(CH_ROWS * CH_LD >= CG_NC_MAX, "column buffer");
        for (int j = 0; j < nc; ++j) {
            const int m = min(nc - 1, j + bw) - j;  // band rows below the diagonal
            if (tid == 0) {
                const double dj = A[(size_t)j * nc + j];
                if (!(dj > 0.0)) s_bad = j + 1;
                s_piv = dj > 0.0 ? sqrt(dj) : 1.0;
                A[(size_t)j * nc + j] = s_piv;
            }
            __syncthreads();
            if (s_bad) {
                if (tid == 0) *flag = s_bad;
                return;
            }
            const double piv = s_piv;
            for (int u = tid; u < m; u += CH_WG) {
                const double l = A[(size_t)(j + 1 + u) * nc + j] / piv;
                A[(size_t)(j + 1 + u) * nc + j] = l;
                s_l[u] = l;
            }
            __syncthreads();
            for (int t = tid; t < m * m; t += CH_WG) {
                const int u = t / m, v = t - u * m;
                if (v <= u) A[(size_t)(j + 1 + u) * nc + j + 1 + v] -= s_l[u] * s_l[v];
            }
            __syncthreads();
        }
    }
    if (tid == 0) *flag = 0;
}

// The same banded Cholesky split over kernels, one panel of CP_NB = 16
// columns per step (32-column panels were slower: the panel's serial part
// grows with the square of the width).  k_chol_lift lifts the diagonal (as
// above) and clears the flag.  Step j (k_chol_step, first column j0) works
// with look-ahead: workgroup 0 applies panel j - 1 to the columns of panel j
// and factors panel j -- its diagonal block and the band rows below it, in
// LDS exactly as k_band_cholesky does -- while the other workgroups apply
// panel j - 1 to the rest of its trailing window (columns from j0 + CP_NB on)
// in 64 x 64 tiles of the lower triangle.  So the critical path per panel is
// one kernel, and the trailing update overlaps the next panel.  Every entry
// receives each panel's update once, as the sum over the panel's columns in
// ascending order, and the panels in order.  Every kernel returns at once
// when an earlier pivot failed (*flag != 0).
__global__ __launch_bounds__(CH_WG) void k_chol_lift(int nc, double *__restrict__ A, int *__restrict__ flag)
{
    __shared__ double s_red[CH_WG / 64];
    const int tid = threadIdx.x;
    double tr = 0.0;
    for (int i = tid; i < nc; i += CH_WG) tr += A[(size_t)i * nc + i];
    for (int o = 32; o > 0; o >>= 1) tr += __shfl_xor(tr, o);
    if ((tid & 63) == 0) s_red[tid >> 6] = tr;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < CH_WG / 64; ++w) t += s_red[w];
    const double lift = 1e-12 * t / nc;
    for (int i = tid; i < nc; i += CH_WG) A[(size_t)i * nc + i] += lift;
    if (tid == 0) *flag = 0;
}

constexpr int CP_NB = 16;          // panel width of the split factorization
constexpr int CP_LD = CP_NB + 1;    // odd LDS row stride
constexpr int CS_T = 64;           // trailing-update tile (32 x 32 threads, 2 x 2 entries each)
static_assert(2 * CS_T <= CH_ROWS, "tile rows in the panel buffer");
__global__ __launch_bounds__(CH_WG) void k_chol_step(int nc, int bw, int j0, double *__restrict__ A,
                                                     int *__restrict__ flag)
{
    constexpr int NB = CP_NB, LD = CP_LD;
    __shared__ double s_p[CH_ROWS * LD];  // workgroup 0: panel j; the others: two row blocks of panel j - 1
    __shared__ double s_q[NB][LD];       // workgroup 0: rows j0 .. j0 + NB - 1 of panel j - 1
    __shared__ double s_rd[NB];          // reciprocals of the panel's pivots
    __shared__ int s_bad;
    const int tid = threadIdx.x;
    if (*flag) return;
    const int jp = j0 - NB;  // first column of panel j - 1 (< 0: none)
    if (blockIdx.x > 0) {    // trailing tiles of panel j - 1, columns [j0 + NB, w1)
        if (jp < 0) return;
        const int w0 = j0 + NB, w1 = min(nc, j0 + bw);
        const int m = w1 - w0;
        if (m <= 0) return;
        const int nt = (m + CS_T - 1) / CS_T, tix = (int)blockIdx.x - 1;
        int ta = (int)((sqrt(8.0 * tix + 1.0) - 1.0) * 0.5);
        while ((ta + 1) * (ta + 2) / 2 <= tix) ++ta;
        while (ta * (ta + 1) / 2 > tix) --ta;
        const int tb = tix - ta * (ta + 1) / 2;
        if (ta >= nt) return;  // the grid is sized for the widest window
        const int i0 = w0 + ta * CS_T, k0 = w0 + tb * CS_T;
        double *s_i = s_p, *s_k = s_p + CS_T * LD;
        for (int t = tid; t < CS_T * NB; t += CH_WG) {
            const int r = t / NB, c = t % NB;
            s_i[r * LD + c] = i0 + r < w1 ? A[(size_t)(i0 + r) * nc + jp + c] : 0.0;
            s_k[r * LD + c] = k0 + r < w1 ? A[(size_t)(k0 + r) * nc + jp + c] : 0.0;
        }
        __syncthreads();
        const int tx = tid & 31, ty = tid >> 5;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int a = 2 * ty + p, bq = 2 * tx + q;
                const int i = i0 + a, k = k0 + bq;
                if (i < w1 && k < w1 && k <= i) {
                    double acc = 0.0;
#pragma unroll
                    for (int c = 0; c < NB; ++c) acc += s_i[a * LD + c] * s_k[bq * LD + c];
                    A[(size_t)i * nc + k] -= acc;
                }
            }
        return;
    }
    if (tid == 0) s_bad = 0;
    const int nb = min(NB, nc - j0);
    const int H = min(nc, j0 + nb + bw) - j0;  // panel rows (<= CH_ROWS)
    // panel j - 1 reaches rows j0 .. j0 + hq - 1; thread t takes half of row
    // t / 2, its row of panel j - 1 loaded with the panel
    const int hq = jp >= 0 ? min(nc, j0 + bw) - j0 : 0;
    static_assert(2 * (CH_ROWS - CP_NB) <= CH_WG, "one row half per thread");
    const bool upd = tid < 2 * hq;
    double lr[NB];
    if (upd)
#pragma unroll
        for (int c = 0; c < NB; ++c) lr[c] = A[(size_t)(j0 + (tid >> 1)) * nc + jp + c];
    for (int t = tid; t < H * NB; t += CH_WG) {
        const int r = t / NB, c = t % NB;
        s_p[r * LD + c] = (r >= c && c < nb) ? A[(size_t)(j0 + r) * nc + j0 + c] : 0.0;
    }
    if (jp >= 0)
        for (int t = tid; t < NB * NB; t += CH_WG) {
            const int r = t / NB, c = t % NB;
            s_q[r][c] = j0 + r < nc ? A[(size_t)(j0 + r) * nc + jp + c] : 0.0;
        }
    __syncthreads();
    if (jp >= 0) {  // panel j - 1 on the columns of panel j
        if (upd) {
            const int r = tid >> 1, c0 = (tid & 1) * (NB / 2);
#pragma unroll
            for (int cc = 0; cc < NB / 2; ++cc) {
                const int c = c0 + cc;
                if (c <= r && c < nb) {
                    double acc = 0.0;
#pragma unroll
                    for (int c2 = 0; c2 < NB; ++c2) acc += lr[c2] * s_q[c][c2];
                    s_p[r * LD + c] -= acc;
                }
            }
        }
        __syncthreads();
    }
    if (tid < 64) {  // the nb x nb diagonal block by wave 0, lane l owning row l
        double v[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) v[c] = (tid < nb && c <= tid) ? s_p[tid * LD + c] : 0.0;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (c < nb) {
                const double dc = read_lane(v[c], c);
                const double piv = dc > 0.0 ? sqrt(dc) : 1.0;
                if (tid == 0 && !(dc > 0.0) && !s_bad) s_bad = j0 + c + 1;
                if (tid == c) v[c] = piv;
                if (tid > c) v[c] /= piv;
#pragma unroll
                for (int c2 = c + 1; c2 < NB; ++c2) {
                    const double lc2 = read_lane(v[c], c2);
                    if (tid >= c2 && c2 < nb) v[c2] -= v[c] * lc2;
                }
            }
        }
        if (tid < nb)
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c <= tid) s_p[tid * LD + c] = v[c];
    }
    __syncthreads();
    if (s_bad) {
        if (tid == 0) *flag = s_bad;
        return;
    }
    if (tid < nb) s_rd[tid] = 1.0 / s_p[tid * LD + tid];
    __syncthreads();
    for (int r = nb + tid; r < H; r += CH_WG) {  // rows below the diagonal block: x L_diag^T = a
        double x[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) x[c] = s_p[r * LD + c];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (c < nb) {
                double t = x[c];
#pragma unroll
                for (int c2 = 0; c2 < c; ++c2) t -= x[c2] * s_p[c * LD + c2];
                x[c] = t * s_rd[c];
            }
        }
#pragma unroll
        for (int c = 0; c < NB; ++c) s_p[r * LD + c] = x[c];
    }
    __syncthreads();
    for (int t = tid; t < H * NB; t += CH_WG) {
        const int r = t / NB, c = t % NB;
        if (r >= c && c < nb) A[(size_t)(j0 + r) * nc + j0 + c] = s_p[r * LD + c];
    }
}

// the band below the diagonal := mean of both triangles (A_c is symmetric up
// to rounding); one thread per band entry
__global__ __launch_bounds__(RED_WG) void k_band_sym(int nc, int bw, double *__restrict__ A)
{
    const long long t = (long long)blockIdx.x * RED_WG + threadIdx.x;
    const int j = (int)(t / bw), i = j + 1 + (int)(t - (long long)j * bw);
    if (j >= nc || i >= nc) return;
    A[(size_t)i * nc + j] = 0.5 * (A[(size_t)i * nc + j] + A[(size_t)j * nc + i]);
}

// L^T into the upper triangle of the band
__global__ __launch_bounds__(RED_WG) void k_band_upper(int nc, int bw, double *__restrict__ A)
{
    const long long t = (long long)blockIdx.x * RED_WG + threadIdx.x;
    const int j = (int)(t / bw), i = j + 1 + (int)(t - (long long)j * bw);
    if (j >= nc || i >= nc) return;
    A[(size_t)j * nc + i] = A[(size_t)i * nc + j];
}

// Column e of A^-1 = L^-T L^-1 e_e by one wave: forward then backward
// substitution over the band (lane-split dot products, shuffle sums),
// written as row e of inv (A^-1 is symmetric).  For bw < 64 INV_U the band
// rows of the next INV_D - 1 steps are in flight while the current one is
// summed (each step waits on the previous one, so a row load must not).
constexpr int INV_WAVES = 4;
constexpr int INV_U = 8;
constexpr int INV_D = 4;  // rows of L in flight per wave
__global__ __launch_bounds__(64 * INV_WAVES) void k_band_inverse(int nc, int bw, const double *__restrict__ L,
                                                                 double *__restrict__ inv)
{
    __shared__ double s_y[INV_WAVES][CG_NC_MAX];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int e = blockIdx.x * INV_WAVES + w;
    if (e >= nc) return;  // whole waves; no workgroup barrier below
    double *y = s_y[w];
    for (int i = lane; i < e; i += 64) y[i] = 0.0;
    wave_sync();
    if (bw < 64 * INV_U) {
        // forward, L y = e_e: row i holds L[i][k], k in [max(e, i - bw), i)
        auto fload = [&](int i, double v[INV_U]) {
            const int lo = max(e, i - bw);
#pragma unroll
            for (int u = 0; u < INV_U; ++u) {
                // unconditional load of a clamped (in-range) index, then a
                // select: loads under branches would make the compiler wait
                // for every load in flight
                const int k = lo + lane + 64 * u;
                const bool ok = i < nc && k < i;
                const double t = L[ok ? (size_t)i * nc + k : 0];
                v[u] = ok ? t : 0.0;
            }
        };
        // rows i + 1 .. i + INV_D - 1 are in flight while row i is reduced
        double buf[INV_D][INV_U], dg[INV_D];
#pragma unroll
        for (int s = 0; s < INV_D; ++s) {
            fload(e + s, buf[s]);
            dg[s] = e + s < nc ? L[(size_t)(e + s) * nc + e + s] : 1.0;
        }
        for (int i0 = e; i0 < nc; i0 += INV_D) {
#pragma unroll
            for (int s = 0; s < INV_D; ++s) {
                const int i = i0 + s;
                if (i < nc) {  // the same for the whole wave
                    const int lo = max(e, i - bw);
                    double part = 0.0;
#pragma unroll
                    for (int u = 0; u < INV_U; ++u) {
                        const int k = lo + lane + 64 * u;
                        part += k < i ? buf[s][u] * y[min(k, nc - 1)] : 0.0;
                    }
                    part = wave_sum(part);
                    const double yi = ((i == e ? 1.0 : 0.0) - part) / dg[s];
                    if (lane == 0) y[i] = yi;
                    wave_sync();
                }
                fload(i + INV_D, buf[s]);
                dg[s] = i + INV_D < nc ? L[(size_t)(i + INV_D) * nc + i + INV_D] : 1.0;
            }
        }
        // backward, L^T x = y: row i of the upper triangle, k in (i, min(nc - 1, i + bw)]
        auto bload = [&](int i, double v[INV_U]) {
            const int hi = min(nc - 1, i + bw);
#pragma unroll
            for (int u = 0; u < INV_U; ++u) {
                const int k = i + 1 + lane + 64 * u;
                const bool ok = i >= 0 && k <= hi;
                const double t = L[ok ? (size_t)i * nc + k : 0];
                v[u] = ok ? t : 0.0;
            }
        };
#pragma unroll
        for (int s = 0; s < INV_D; ++s) {
            bload(nc - 1 - s, buf[s]);
            dg[s] = nc - 1 - s >= 0 ? L[(size_t)(nc - 1 - s) * nc + nc - 1 - s] : 1.0;
        }
        for (int i0 = nc - 1; i0 >= 0; i0 -= INV_D) {
#pragma unroll
            for (int s = 0; s < INV_D; ++s) {
                const int i = i0 - s;
                if (i >= 0) {
                    const int hi = min(nc - 1, i + bw);
                    double part = 0.0;
#pragma unroll
                    for (int u = 0; u < INV_U; ++u) {
                        const int k = i + 1 + lane + 64 * u;
                        part += k <= hi ? buf[s][u] * y[min(k, nc - 1)] : 0.0;
                    }
                    part = wave_sum(part);
                    const double xi = (y[i] - part) / dg[s];
                    wave_sync();
                    if (lane == 0) y[i] = xi;
                    wave_sync();
                }
                bload(i - INV_D, buf[s]);
                dg[s] = i - INV_D >= 0 ? L[(size_t)(i - INV_D) * nc + i - INV_D] : 1.0;
            }
        }
    } else {
        for (int i = e; i < nc; ++i) {  // L y = e_e
            const int lo = max(e, i - bw);
            double part = 0.0;
            for (int k = lo + lane; k < i; k += 64) part += L[(size_t)i * nc + k] * y[k];
            for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
            const double yi = ((i == e ? 1.0 : 0.0) - part) / L[(size_t)i * nc + i];
            if (lane == 0) y[i] = yi;
            wave_sync();
        }
        for (int i = nc - 1; i >= 0; --i) {  // L^T x = y (L^T is the upper triangle)
            const int hi = min(nc - 1, i + bw);
            double part = 0.0;
            for (int k = i + 1 + lane; k <= hi; k += 64) part += L[(size_t)i * nc + k] * y[k];
            for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
            const double xi = (y[i] - part) / L[(size_t)i * nc + i];
            wave_sync();
            if (lane == 0) y[i] = xi;
            wave_sync();
        }
    }
    for (int i = lane; i < nc; i += 64) inv[(size_t)e * nc + i] = y[i];
}

// Y = L^-1 for the band Cholesky factor L (lower triangle of the row-major
// nc x nc array, half-bandwidth bw), one wave per column e, written as row e
// of Yt (Yt[e][i] = Y[i][e]; zero for i < e).  The rows from e on are taken in
// blocks of 64: lane l holds the accumulator of row r0 + l; the block's
// unknowns are found in order (lane s's accumulator is final after step s - 1:
// y_s = t_s / L_ss with the reciprocal formed beforehand, broadcast by
// read_lane, and every later lane subtracts L[r0 + l][r0 + s] y_s, the row
// entries of 16 steps loaded ahead).  Then the block's contribution to each of
// the next bw rows, sum_s L[i][r0 + s] y_s, is one coalesced load of the row's
// 64 entries (lane s holds y_s), LI_R rows at a time, summed over the wave by
// one reduce-scatter; the
// accumulators of those rows wait in a per-wave LDS ring.  Every sum has one
// fixed order.  bw + 64 <= LI_RING (the host checks).
constexpr int LI_WAVES = 4;
constexpr int LI_RING = 512;
constexpr int LI_R = 32;  // rows of the band update in flight (= 64 / 2: one row per lane pair)
__global__ __launch_bounds__(64 * LI_WAVES) void k_band_linv(int nc, int bw, const double *__restrict__ L,
                                                            double *__restrict__ Yt)
{
    __shared__ double s_acc[LI_WAVES][LI_RING];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int e = blockIdx.x * LI_WAVES + w;
    if (e >= nc) return;  // whole waves; no workgroup barrier below
    double *acc = s_acc[w];
    for (int q = lane; q < LI_RING; q += 64) acc[q] = 0.0;
    double *yrow = Yt + (size_t)e * nc;
    for (int q = lane; q < e; q += 64) yrow[q] = 0.0;
    wave_sync();
    for (int r0 = e; r0 < nc; r0 += 64) {
        const int nb = min(64, nc - r0);
        const int i = r0 + lane;
        const bool act = lane < nb;
        double t = act ? acc[i % LI_RING] : 0.0;
        if (i == e) t += 1.0;
        const double rdg = act ? 1.0 / L[(size_t)i * nc + i] : 1.0;
        wave_sync();
        if (act) acc[i % LI_RING] = 0.0;  // the slot serves row i + LI_RING next
        // the diagonal block, 16 steps per round with the row entries loaded first
        double y = 0.0;
        for (int s0 = 0; s0 < nb; s0 += 16) {
            double lv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int k = r0 + s0 + u;
                const bool ok = act && s0 + u < lane && i - k <= bw;
                const double v = L[ok ? (size_t)i * nc + k : 0];
                lv[u] = ok ? v : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int sidx = s0 + u;
                if (sidx < nb) {  // uniform
                    const double ys = read_lane(t, sidx) * read_lane(rdg, sidx);
                    if (lane == sidx) y = ys;
                    t -= lv[u] * ys;  // zero for lanes <= sidx
                }
            }
        }
        if (act) yrow[i] = y;
        // the block's contribution to the rows below it within the band
        const int lo = r0 + nb, hi = min(nc - 1, r0 + nb - 1 + bw);
        const int k = r0 + lane;
        for (int m0 = lo; m0 <= hi; m0 += LI_R) {  // uniform
            double part[LI_R];
#pragma unroll
            for (int u = 0; u < LI_R; ++u) {
                const int i2 = m0 + u;
                const bool ok = act && i2 <= hi && i2 - k <= bw;
                const double v = L[ok ? (size_t)i2 * nc + k : 0];
                part[u] = (ok ? v : 0.0) * y;
            }
            // reduce-scatter over the wave by recursive halving: at lane distance
            // 32, 16, 8, 4, 2 a lane keeps half of its values (the half its lane
            // bit selects) and adds its partner's; the last step adds the pair.
            // Lanes 2u and 2u + 1 end with the total of row m0 + u.
#pragma unroll
            for (int d = 32, h = LI_R / 2; d >= 2; d >>= 1, h >>= 1) {
                const bool up = (lane & d) != 0;
#pragma unroll
                for (int j = 0; j < h; ++j) {
                    const double send = up ? part[j] : part[h + j];
                    const double keep = up ? part[h + j] : part[j];
                    part[j] = keep + __shfl_xor(send, d);
                }
            }
            const double tot = part[0] + __shfl_xor(part[0], 1);
            const int u = lane >> 1;
            if ((lane & 1) == 0 && m0 + u <= hi) acc[(m0 + u) % LI_RING] -= tot;
        }
        wave_sync();
    }
}

// A^-1 = Y^T Y (Y = L^-1, as rows of Yt) in FP32, symmetric: out[a][b] =
// out[b][a] = sum_{i >= max(a, b)} Yt[a][i] Yt[b][i] in FP64, one 32 x 32
// tile (b0 <= a0) per 256-thread workgroup, 2 x 2 outputs per thread, Yt
// staged in LDS 32 columns at a time.
constexpr int IG_T = 32, IG_K = 32;
__global__ __launch_bounds__(256) void k_inv_gemm32(int nc, const double *__restrict__ Yt, float *__restrict__ out)
{
    __shared__ double sa[IG_K][IG_T + 1], sb[IG_K][IG_T + 1];
    const int nt = (nc + IG_T - 1) / IG_T;
    // tile (ta, tb) with tb <= ta from the linear block index
    int ta = (int)((sqrt(8.0 * blockIdx.x + 1.0) - 1.0) * 0.5);
    while ((ta + 1) * (ta + 2) / 2 <= (int)blockIdx.x) ++ta;
    while (ta * (ta + 1) / 2 > (int)blockIdx.x) --ta;
    const int tb = (int)blockIdx.x - ta * (ta + 1) / 2;
    if (ta >= nt) return;
    const int a0 = ta * IG_T, b0 = tb * IG_T;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // outputs a0 + 2 ty + p, b0 + 2 tx + q
    double c[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int i0 = a0; i0 < nc; i0 += IG_K) {  // Yt[a][i] = 0 for i < a0 <= a, and b0 <= a0
        __syncthreads();
        for (int t = threadIdx.x; t < IG_K * IG_T; t += 256) {
            const int r = t / IG_K, k = t % IG_K;  // row of the tile, column i0 + k
            const int i = i0 + k, ra = a0 + r, rb = b0 + r;
            sa[k][r] = (ra < nc && i < nc) ? Yt[(size_t)ra * nc + i] : 0.0;
            sb[k][r] = (rb < nc && i < nc) ? Yt[(size_t)rb * nc + i] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < IG_K; ++k) {
            const double a_0 = sa[k][2 * ty], a_1 = sa[k][2 * ty + 1];
            const double b_0 = sb[k][2 * tx], b_1 = sb[k][2 * tx + 1];
            c[0][0] += a_0 * b_0;
            c[0][1] += a_0 * b_1;
            c[1][0] += a_1 * b_0;
            c[1][1] += a_1 * b_1;
        }
    }
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int a = a0 + 2 * ty + p, b = b0 + 2 * tx + q;
            if (a < nc && b < nc) {
                out[(size_t)a * nc + b] = (float)c[p][q];
                out[(size_t)b * nc + a] = (float)c[p][q];
            }
        }
}

// symmetric FP32 copy of the inverse
__global__ __launch_bounds__(RED_WG) void k_sym32(int nc, const double *__restrict__ inv, float *__restrict__ out)
{
    const long long t = (long long)blockIdx.x * RED_WG + threadIdx.x;
    if (t >= (long long)nc * nc) return;
    const int i = (int)(t / nc), j = (int)(t - (long long)i * nc);
    out[t] = (float)(0.5 * (inv[(size_t)i * nc + j] + inv[(size_t)j * nc + i]));
}

// Coarse space of the two-level methods.  The free nodes are reordered into
// compact parts of ra >= CG_AGG CG node ranges (at least 64 nodes each, so
// that the rotations stay well conditioned); in the new order an aggregate is
// ra consecutive CG ranges, so the CG kernels form Z^T of a vector as
// per-range partials.  Z = rigid-body modes about each aggregate's centroid,
// A_c = Z^T K Z (FP64, 6 nagg square), Cholesky-inverted and stored in FP32.
// coarse_structure (once per mesh and pattern, host): the order, the
// renumbered pattern, the offsets d, the per-aggregate block lists of A_c and
// its bandwidth.  coarse_values (once per K, device): the reordered values,
// their block-Jacobi inverses and FP32 copy, A_c, its Cholesky factor and the
// inverse.
static int coarse_structure(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    const int nf = w->nfree;
    const int per = (nf + CG_NB - 1) / CG_NB;
    const int nranges = (nf + per - 1) / per;
    const int ra = std::max(CG_AGG, (64 + per - 1) / per);
    const int nagg = (nranges + ra - 1) / ra;
    const int L = per * ra;
    const int nblk = ctx->P.node_blk[nf];
    if (nf >= (1 << 23)) return fail(OFEA_ERR_LIMIT, "too many free nodes for the packed coarse block lists");
    static_assert(CL_NAGG == CG_NB / CG_AGG, "aggregate bound");
    // free-node coordinates and their bounding box
    std::vector<double> X((size_t)3 * nf);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < ctx->nd; ++i) {
        const int r = ctx->resort[i];
        if (r < 0) continue;
        for (int c = 0; c < 3; ++c) {
            const double x = ctx->xyz[3 * (size_t)i + c];
            X[3 * (size_t)r + c] = x;
            lo[c] = std::min(lo[c], x);
            hi[c] = std::max(hi[c], x);
        }
    }
    // The order (host): parts of exactly L nodes (the last may be shorter),
    // as slabs of whole parts across the longest axis of the bounding box --
    // about as thick as a cubic part, so A_c stays banded (a part couples only
    // to parts of its own and the neighbouring slabs) -- each slab cut into
    // its parts by recursive coordinate bisection on its own host thread.
    // perm[new] = old.
    const int nparts = (nf + L - 1) / L;
    int ax = 0;
    for (int c = 1; c < 3; ++c)
        if (hi[c] - lo[c] > hi[ax] - lo[ax]) ax = c;
    const double emax = std::max(hi[ax] - lo[ax], 0.0);
    double vol = 1.0;
    for (int c = 0; c < 3; ++c) vol *= std::max(hi[c] - lo[c], 1e-3 * emax);
    const double edge = std::cbrt(vol / std::max(nparts, 1));
    const int slabs = std::max(1, std::min(nparts, (int)std::lround(emax / std::max(edge, 1e-300))));
    const long long slab_nodes = (long long)((nparts + slabs - 1) / slabs) * L;
    const int nslab = nf ? (int)((nf + slab_nodes - 1) / slab_nodes) : 0;
    std::vector<int> perm(nf), iperm(nf);
    if (nparts <= 1) {
        for (int i = 0; i < nf; ++i) perm[i] = i;
    } else {  // slab-major, ascending index inside each slab, then the parts of each slab
        std::vector<int> slab;
        slab_of_nodes(nf, X.data(), ax, lo[ax], hi[ax], slab_nodes, slab);
        std::vector<long long> pos(nslab + 1, 0);
        for (int i = 0; i < nf; ++i) pos[slab[i] + 1]++;
        for (int sl = 0; sl < nslab; ++sl) pos[sl + 1] += pos[sl];
        for (int i = 0; i < nf; ++i) perm[pos[slab[i]]++] = i;
        const int nthr = nf < 16384 ? 1 : host_threads(nslab);
        std::vector<std::vector<AxisKey>> keys(nthr);
        run_threads(nthr, [&](int th) {
            for (int sl = th; sl < nslab; sl += nthr) {
                const long long s0 = sl * slab_nodes;
                rcb_parts(perm.data() + s0, (int)std::min<long long>(slab_nodes, nf - s0), L, X.data(), keys[th]);
            }
        });
    }
    for (int i = 0; i < nf; ++i) iperm[perm[i]] = i;
    // device buffers (sizes depend only on nfree, nnz and the bounds)
    const size_t nfb = (size_t)std::max(nf, 1);
    const long long nnz = ctx->P.nnz;
    static_assert(CG_NC_MAX == 6 * (CG_NB / CG_AGG), "coarse dimension bound");
    if (!w->co_arena) {  // one allocation holds every two-level buffer
        const size_t nbk = (size_t)std::max(nblk, 1), nz = (size_t)std::max(nnz, 1LL);
        const size_t ncm2 = (size_t)CG_NC_MAX * CG_NC_MAX, lch = (size_t)CL_NAGG * (CA_MAXCH + 1);
        struct Part {
            void **p;
            size_t bytes;
        } parts[] = {
            {(void **)&w->co.g, sizeof(double) * 6 * CG_NB},    {(void **)&w->co.zq, sizeof(double) * 6 * CG_NB},
            {(void **)&w->co.ya, sizeof(double) * 2 * CG_NC_MAX}, {(void **)&w->co.d, sizeof(float) * 3 * nfb},
            {(void **)&w->co_ainv, sizeof(float) * ncm2},       {(void **)&w->co_A, sizeof(double) * ncm2},
            {(void **)&w->co_inv, sizeof(double) * ncm2},       {(void **)&w->co_flag, sizeof(int)},
            {(void **)&w->perm, sizeof(int) * nfb},             {(void **)&w->iperm, sizeof(int) * nfb},
            {(void **)&w->nb_p, sizeof(int) * (nfb + 1)},       {(void **)&w->bcol_p, sizeof(int) * nbk},
            {(void **)&w->val_p, sizeof(double) * nz},          {(void **)&w->val32_p, sizeof(float) * nz},
            {(void **)&w->dinv_p, sizeof(double) * 3 * nfb},    {(void **)&w->b_p, sizeof(double) * 3 * nfb},
            {(void **)&w->x_p, sizeof(double) * 3 * nfb},       {(void **)&w->minv_p, sizeof(double) * 9 * nfb},
            {(void **)&w->dblk_p, sizeof(double) * 9 * nfb},    {(void **)&w->minv32_p, sizeof(float) * 9 * nfb},
            {(void **)&w->sl_blk, sizeof(int) * nbk},           {(void **)&w->sl_part, sizeof(int) * CL_NAGG * CA_MAXCH},
            {(void **)&w->sl_cptr, sizeof(int) * lch},          {(void **)&w->ch_beg, sizeof(int) * lch},
            {(void **)&w->co_lst, sizeof(int4) * CL_NAGG},      {(void **)&w->co_xf, sizeof(double) * 3 * nfb},
            {(void **)&w->co_scan, sizeof(int) * (nfb / SC_TILE + 1)}};
        size_t tot = 0;
        for (const Part &q : parts) tot += (q.bytes + 255) & ~(size_t)255;
        char *base = nullptr;
        OFEA_HIP(hipMalloc((void **)&base, tot));
        w->co_arena = base;
        for (const Part &q : parts) {
            *q.p = base;
            base += (q.bytes + 255) & ~(size_t)255;
        }
    }
    // The rest on the device: iperm and the renumbered pattern (nb_p by a
    // scan of the row lengths, bcol_p), the offsets d from the aggregate
    // centroids, and the block lists of A_c.
    hipStream_t st = ctx->stream;
    const unsigned gz6 = (6 * CG_NB + RED_WG - 1) / RED_WG;
    hipLaunchKernelGGL((k_fill<double>), dim3(gz6), dim3(RED_WG), 0, st, 6 * CG_NB, w->co.g, 0.0);
    hipLaunchKernelGGL((k_fill<double>), dim3(gz6), dim3(RED_WG), 0, st, 6 * CG_NB, w->co.zq, 0.0);
    hipLaunchKernelGGL((k_fill<double>), dim3((2 * CG_NC_MAX + RED_WG - 1) / RED_WG), dim3(RED_WG), 0, st,
                       2 * CG_NC_MAX, w->co.ya, 0.0);
    if (nf) {
        OFEA_HIP(hipMemcpyAsync(w->perm, perm.data(), sizeof(int) * nf, hipMemcpyHostToDevice, st));
        const unsigned gn = (unsigned)((ctx->nd + RED_WG - 1) / RED_WG), gf = (unsigned)((nf + RED_WG - 1) / RED_WG);
        hipLaunchKernelGGL(k_co_xfree, dim3(gn), dim3(RED_WG), 0, st, ctx->nd, ctx->d_resort, ctx->d_xyz, w->co_xf);
        hipLaunchKernelGGL(k_co_rows, dim3(gf), dim3(RED_WG), 0, st, nf, w->perm, ctx->d_node_blk, w->iperm, w->nb_p);
        const unsigned ntile = (unsigned)((nf + SC_TILE - 1) / SC_TILE);
        hipLaunchKernelGGL(k_scan_tiles, dim3(ntile), dim3(SC_WG), 0, st, nf, w->nb_p + 1, w->co_scan);
        if (ntile > 1) hipLaunchKernelGGL(k_scan_add, dim3(ntile), dim3(SC_WG), 0, st, nf, w->nb_p + 1, w->co_scan);
        hipLaunchKernelGGL(k_co_offsets, dim3(nagg), dim3(CD_WG), 0, st, nf, L, w->perm, w->co_xf, (float *)w->co.d);
        hipLaunchKernelGGL(k_perm_bcol, dim3(gf), dim3(RED_WG), 0, st, nf, w->perm, w->iperm, ctx->d_node_blk,
                           ctx->d_bcol, w->nb_p, w->bcol_p);
        hipLaunchKernelGGL(k_co_lists, dim3(nagg), dim3(CL_WG), 0, st, nf, L, nagg, w->nb_p, w->bcol_p, w->sl_blk,
                           w->sl_part, w->sl_cptr, w->ch_beg, w->co_lst);
        OFEA_HIP(hipGetLastError());
        OFEA_HIP(hipMemcpyAsync(w->co_lst_host, w->co_lst, sizeof(int4) * nagg, hipMemcpyDeviceToHost, st));
    }
    // the lists' results (bandwidth of A_c, limit) are read by coarse_values
    // once its first kernels are queued
    OFEA_HIP(hipEventRecord(w->co_ev, st));
    w->co_lst_pending = true;
    w->co_ainv_ready = false;
    w->co.per = per;
    w->co.ra = ra;
    w->co_nagg = nagg;
    w->co.nagg = nagg;
    w->co.ainv = w->co_ainv;
    w->co_d_host.clear();  // downloaded by coarse_apply when needed
    w->co_perm_host = std::move(perm);  // the source of a queued copy
    w->iperm_host = std::move(iperm);
    w->co_struct_valid = true;
    return OFEA_OK;
}

// vals reordered into the two-level order (val_p), its block-Jacobi inverses
// and, with fp32, the FP32 copy the mixed two-level methods read
static void coarse_refresh_values(ofea_ctx *ctx, const double *vals, bool fp32)
{
    ofea_solver_ws *w = ctx->ws;
    const int nf = w->nfree;
    const long long nnz = ctx->P.nnz;
    hipStream_t st = ctx->stream;
    if (nf) {
        hipLaunchKernelGGL(k_perm_vals, dim3((nf + RED_WG / 64 - 1) / (RED_WG / 64)), dim3(RED_WG), 0, st, nf,
                           w->perm, ctx->d_node_blk, w->nb_p, vals, w->val_p);
        hipLaunchKernelGGL(k_precond_setup, dim3(w->nblk_node), dim3(RED_WG), 0, st, nf, w->nb_p, w->bcol_p, w->val_p,
                           w->dinv_p, w->minv_p, w->minv32_p, w->dblk_p);
    }
    if (nnz && fp32)
        hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((nnz + RED_WG - 1) / RED_WG)), dim3(RED_WG), 0,
                           st, nnz, w->val_p, w->val32_p);
}

// the coarse factorization's flag, copied to pinned memory after the set-up
static int coarse_check(ofea_ctx *ctx)
{
    ofea_solver_ws *w = ctx->ws;
    OFEA_HIP(hipStreamSynchronize(ctx->stream));
    if (*w->co_flag_host) {
        w->co_valid = false;
        w->co_ainv_ready = false;
        return fail(OFEA_ERR_STATE, "coarse matrix Z^T K Z is not positive definite");
    }
    return OFEA_OK;
}

// A_c^-1 (and the reordered values with their block Jacobi) of the values
// vals: K (the two-level solves) or a batched variant; fp32 adds the FP32
// copy of the reordered values the mixed two-level methods read.
static int coarse_values(ofea_ctx *ctx, const double *vals, bool fp32, bool defer_check = false)
{
    ofea_solver_ws *w = ctx->ws;
    const int nc = 6 * w->co_nagg;
    hipStream_t st = ctx->stream;
    coarse_refresh_values(ctx, vals, fp32);
    if (nc == 0) return OFEA_OK;
    const long long nc2 = (long long)nc * nc;
    const unsigned g2 = (unsigned)((nc2 + RED_WG - 1) / RED_WG);
    hipLaunchKernelGGL((k_fill<double>), dim3(g2), dim3(RED_WG), 0, st, (int)nc2, w->co_A, 0.0);
    hipLaunchKernelGGL(k_coarse_ac, dim3(w->co_nagg), dim3(CA_WG), 0, st, nc, w->nb_p, w->bcol_p, w->val_p, w->co.d,
                       w->co_lst, w->sl_part, w->sl_cptr, w->ch_beg, w->sl_blk, w->co_A);
    if (w->co_lst_pending) {  // the block lists' results, copied while the kernels above run
        OFEA_HIP(hipEventSynchronize(w->co_ev));
        w->co_lst_pending = false;
        int bw = 5;
        for (int b = 0; b < w->co_nagg; ++b) {
            const int4 l = w->co_lst_host[b];
            if (l.x > CA_MAXCH) {
                w->co_struct_valid = false;
                return fail(OFEA_ERR_LIMIT, "an aggregate couples to too many aggregates");
            }
            bw = std::max(bw, 6 * l.z + 5);
        }
        w->co_bw = std::min(bw, nc - 1);
    }
    const unsigned gb = (unsigned)(((long long)nc * w->co_bw + RED_WG - 1) / RED_WG);
    hipLaunchKernelGGL(k_band_sym, dim3(gb), dim3(RED_WG), 0, st, nc, w->co_bw, w->co_A);
    if (w->co_bw + CP_NB <= CH_ROWS) {  // one kernel per CP_NB columns
        const int ntu = (std::max(w->co_bw - CP_NB, 1) + CS_T - 1) / CS_T;
        hipLaunchKernelGGL(k_chol_lift, dim3(1), dim3(CH_WG), 0, st, nc, w->co_A, w->co_flag);
        for (int j0 = 0; j0 < nc; j0 += CP_NB)
            hipLaunchKernelGGL(k_chol_step, dim3(j0 == 0 ? 1 : 1 + ntu * (ntu + 1) / 2), dim3(CH_WG), 0, st, nc,
                               w->co_bw, j0, w->co_A, w->co_flag);
    } else {
        hipLaunchKernelGGL(k_band_cholesky, dim3(1), dim3(CH_WG), 0, st, nc, w->co_bw, w->co_A, w->co_flag);
    }
    const bool blocked = w->co_bw + 64 <= LI_RING;  // blocked inverse: L^-1 by 64-row blocks, then Y^T Y
    if (!blocked) hipLaunchKernelGGL(k_band_upper, dim3(gb), dim3(RED_WG), 0, st, nc, w->co_bw, w->co_A);
    OFEA_HIP(hipGetLastError());
    if (blocked) {
        hipLaunchKernelGGL(k_band_linv, dim3((nc + LI_WAVES - 1) / LI_WAVES), dim3(64 * LI_WAVES), 0, st, nc,
                           w->co_bw, w->co_A, w->co_inv);
        const int nt = (nc + IG_T - 1) / IG_T;
        hipLaunchKernelGGL(k_inv_gemm32, dim3(nt * (nt + 1) / 2), dim3(256), 0, st, nc, w->co_inv, w->co_ainv);
    } else {
        hipLaunchKernelGGL(k_band_inverse, dim3((nc + INV_WAVES - 1) / INV_WAVES), dim3(64 * INV_WAVES), 0, st, nc,
                           w->co_bw, w->co_A, w->co_inv);
        hipLaunchKernelGGL(k_sym32, dim3(g2), dim3(RED_WG), 0, st, nc, w->co_inv, w->co_ainv);
    }
    OFEA_HIP(hipGetLastError());
    w->co_ainv_host.clear();  // downloaded again by coarse_apply when needed
    // the factorization's flag (a failed pivot leaves non-finite values in
    // the inverse, nothing else): read here, or by the caller once the stream
    // has drained (defer_check)
    OFEA_HIP(hipMemcpyAsync(w->co_flag_host, w->co_flag, sizeof(int), hipMemcpyDeviceToHost, st));
    if (!defer_check) return coarse_check(ctx);
    return OFEA_OK;
}

// defer_check: the caller calls coarse_check once the stream has drained
static int coarse_prepare(ofea_ctx *ctx, bool defer_check = false)
{
    ofea_solver_ws *w = ctx->ws;
    if (w->co_valid) {
        if (w->co_vals_stale) {  // K's new values with the coarse inverse of the earlier K
            coarse_refresh_values(ctx, ctx->d_val, true);
            OFEA_HIP(hipGetLastError());
            w->co_vals_stale = false;
        }
        return OFEA_OK;
    }
    w->co_vals_stale = false;
    int rc;
    if (!w->co_struct_valid && (rc = coarse_structure(ctx))) return rc;
    w->co_ainv_ready = false;
    if ((rc = coarse_values(ctx, ctx->d_val, true, defer_check))) return rc;
    w->co_valid = true;
    w->co_ainv_ready = true;  // the deferred check clears it again on failure
    w->co_builds++;
    return OFEA_OK;
}

int coarse_builds(ofea_ctx *ctx) { return ctx->ws ? ctx->ws->co_builds : 0; }

// The two-level coarse space of the values vals (a batched variant, same
// pattern as K): the mesh's aggregate order and block lists (built once per
// mesh), vals reordered into the solver's reordered value array with its
// block Jacobi, and A_c^-1 -- or, with keep_inverse, the inverse built last
// on this mesh (of another variant or of K: still an SPD coarse correction).
// The single-system solver's own coarse space is rebuilt from K on its next
// two-level solve.
int coarse_for_values(ofea_ctx *ctx, const double *vals, CoarseVariant &v, bool keep_inverse)
{
    int rc = solver_prepare(ctx);
    if (rc) return rc;
    ofea_solver_ws *w = ctx->ws;
    if (!w->co_struct_valid && (rc = coarse_structure(ctx))) return rc;
    w->co_valid = false;
    if (keep_inverse && w->co_ainv_ready) {  // the values of vals, the inverse built last on this mesh
        coarse_refresh_values(ctx, vals, false);
        OFEA_HIP(hipGetLastError());
    } else {
        w->co_ainv_ready = false;
        if ((rc = coarse_values(ctx, vals, false))) return rc;
        w->co_ainv_ready = true;
        w->co_builds++;
    }
    v.nb_p = w->nb_p;
    v.bcol_p = w->bcol_p;
    v.val_p = w->val_p;
    v.minv_p = w->minv_p;
    v.d = w->co.d;
    v.ainv = w->co_ainv;
    v.nagg = w->co_nagg;
    v.per = w->co.per;
    v.ra = w->co.ra;
    v.iperm = w->iperm_host.data();
    return OFEA_OK;
}

// out = Z A_c^-1 Z^T r on the host, with the FP32 offsets and inverse the
// device applies (a check of the coarse correction)
int coarse_apply(ofea_ctx *ctx, const double *r, double *out)
{
    int rc = solver_prepare(ctx);
    if (rc) return rc;
    if ((rc = coarse_prepare(ctx))) return rc;
    ofea_solver_ws *w = ctx->ws;
    const int nf = w->nfree, per = w->co.per, nagg = w->co_nagg, nc = 6 * nagg;
    if (w->co_ainv_host.size() != (size_t)nc * nc) {
        w->co_ainv_host.resize((size_t)nc * nc);
        OFEA_HIP(hipMemcpyAsync(w->co_ainv_host.data(), w->co_ainv, sizeof(float) * nc * nc, hipMemcpyDeviceToHost,
                                ctx->stream));
        OFEA_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (w->co_d_host.size() != 3 * (size_t)nf) {
        w->co_d_host.resize(3 * (size_t)nf);
        if (nf)
            OFEA_HIP(hipMemcpyAsync(w->co_d_host.data(), w->co.d, sizeof(float) * 3 * (size_t)nf,
                                    hipMemcpyDeviceToHost, ctx->stream));
        OFEA_HIP(hipStreamSynchronize(ctx->stream));
    }
    const float *d = w->co_d_host.data();  // offsets in the two-level node order
    const int *ip = w->iperm_host.data();
    std::vector<double> g(nc, 0.0), y(nc, 0.0);
    for (int i = 0; i < nf; ++i) {  // i: free node in the caller's numbering
        const int j = ip[i], a = (j / per) / w->co.ra;
        const double dx = d[3 * j], dy = d[3 * j + 1], dz = d[3 * j + 2];
        const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
        g[6 * a] += r0;
        g[6 * a + 1] += r1;
        g[6 * a + 2] += r2;
        g[6 * a + 3] += dy * r2 - dz * r1;
        g[6 * a + 4] += dz * r0 - dx * r2;
        g[6 * a + 5] += dx * r1 - dy * r0;
    }
    for (int i = 0; i < nc; ++i) {
        double t = 0.0;
        for (int k = 0; k < nc; ++k) t += (double)w->co_ainv_host[(size_t)i * nc + k] * g[k];
        y[i] = t;
    }
    for (int i = 0; i < nf; ++i) {
        const int j = ip[i];
        const double *ya = y.data() + 6 * ((j / per) / w->co.ra);
        const double dx = d[3 * j], dy = d[3 * j + 1], dz = d[3 * j + 2];
        out[3 * i] = ya[0] + ya[4] * dz - ya[5] * dy;
        out[3 * i + 1] = ya[1] + ya[5] * dx - ya[3] * dz;
        out[3 * i + 2] = ya[2] + ya[3] * dy - ya[4] * dx;
    }
    return OFEA_OK;
}

// aggregate of every free node (caller's numbering) of the two-level coarse
// space for the resident K; returns the number of aggregates
int coarse_aggregates(ofea_ctx *ctx, int *agg)
{
    int rc = solver_prepare(ctx);
    if (rc) return rc;
    if ((rc = coarse_prepare(ctx))) return rc;
    ofea_solver_ws *w = ctx->ws;
    for (int i = 0; i < w->nfree; ++i) agg[i] = (w->iperm_host[i] / w->co.per) / w->co.ra;
    return w->co_nagg;
}

int solver_solve(ofea_ctx *ctx, int method, double rtol, int maxit, int *iters, double *relres)
{
    int rc = solver_prepare(ctx);
    if (rc) return rc;
    ofea_solver_ws *w = ctx->ws;
    if (w->n == 0) {
        if (iters) *iters = 0;
        if (relres) *relres = 0.0;
        return OFEA_OK;
    }
    const bool learned = method == OFEA_PCG_LEARNED || method == OFEA_PCG_LEARNED_MIXED ||
                         method == OFEA_PCG_LEARNED_TWO_LEVEL || method == OFEA_PCG_LEARNED_TWO_LEVEL_MIXED ||
                         method == OFEA_PCG_LEARNED_TWO_LEVEL_MIXED32;
    if (method >= OFEA_PCG_TWO_LEVEL && method <= OFEA_PCG_LEARNED_TWO_LEVEL_MIXED32) {
        // two-level: the reordered system (b in, x out through the permutation)
        // the set-up is queued without a wait; its flag is checked after the solve
        if ((rc = coarse_prepare(ctx, true))) return rc;
        if (learned && (rc = learned_factor(ctx, true))) {  // G on the reordered rows
            (void)coarse_check(ctx);
            return rc;
        }
        const int nf = w->nfree;
        const unsigned g3 = (unsigned)((3LL * nf + RED_WG - 1) / RED_WG);
        hipLaunchKernelGGL(k_gather_nodes<3>, dim3(g3), dim3(RED_WG), 0, ctx->stream, nf, w->perm, ctx->d_rhs, w->b_p);
        OFEA_HIP(hipGetLastError());
        const SysView S = perm_view(ctx);
        switch (method) {
        case OFEA_PCG_TWO_LEVEL:
            rc = pcg_run<double, double, 1, true>(ctx, S, 6, w->val_p, w->v64, w->minv_p, rtol, maxit, -1.0, iters,
                                                  relres);
            break;
        case OFEA_PCG_TWO_LEVEL_MIXED:
            rc = pcg_run<float, double, 1, true>(ctx, S, 7, w->val32_p, w->v64, w->minv_p, rtol, maxit, MIXED_RU, iters,
                                                 relres);
            break;
        case OFEA_PCG_LEARNED_TWO_LEVEL:
            rc = pcg_run<double, double, 2, true>(ctx, S, 8, w->val_p, w->v64, w->minv_p, rtol, maxit, -1.0, iters,
                                                  relres);
            break;
        case OFEA_PCG_TWO_LEVEL_MIXED32:  // FP32 iteration vectors too: x and the true residual stay FP64
            rc = pcg_run<float, float, 1, true>(ctx, S, 10, w->val32_p, w->v32, w->minv32_p, rtol, maxit, MIXED_RU,
                                                iters, relres);
            break;
        case OFEA_PCG_LEARNED_TWO_LEVEL_MIXED32:
            rc = pcg_run<float, float, 2, true>(ctx, S, 11, w->val32_p, w->v32, w->minv32_p, rtol, maxit, MIXED_RU,
                                                iters, relres);
            break;
        default:
            rc = pcg_run<float, double, 2, true>(ctx, S, 9, w->val32_p, w->v64, w->minv_p, rtol, maxit, MIXED_RU, iters,
                                                 relres);
        }
        if (int rc2 = coarse_check(ctx)) return rc2;
        if (rc != OFEA_OK && rc != OFEA_NOT_CONVERGED) return rc;
        hipLaunchKernelGGL(k_gather_nodes<3>, dim3(g3), dim3(RED_WG), 0, ctx->stream, nf, w->iperm, w->x_p, ctx->d_x);
        OFEA_HIP(hipGetLastError());
        return rc;
    }
    if (learned && (rc = learned_factor(ctx, false))) return rc;
    const SysView S = ctx_view(ctx);
    switch (method) {
    case OFEA_PCG_JACOBI:
        return pcg_run<double, double, 0, false>(ctx, S, 0, ctx->d_val, w->v64, w->minv, rtol, maxit, -1.0, iters,
                                                 relres);
    case OFEA_PCG_BLOCK_JACOBI:
        return pcg_run<double, double, 1, false>(ctx, S, 1, ctx->d_val, w->v64, w->minv, rtol, maxit, -1.0, iters,
                                                 relres);
    case OFEA_PCG_MIXED:
        return pcg_run<float, double, 1, false>(ctx, S, 2, w->val32, w->v64, w->minv, rtol, maxit, MIXED_RU, iters,
                                                relres);
    case OFEA_PCG_LEARNED:
        return pcg_run<double, double, 2, false>(ctx, S, 3, ctx->d_val, w->v64, w->minv, rtol, maxit, -1.0, iters,
                                                 relres);
    case OFEA_PCG_LEARNED_MIXED:
        return pcg_run<float, double, 2, false>(ctx, S, 5, w->val32, w->v64, w->minv, rtol, maxit, MIXED_RU, iters,
                                                relres);
    case OFEA_PCG_MIXED32:
        return pcg_run<float, float, 1, false>(ctx, S, 4, w->val32, w->v32, w->minv32, rtol, maxit, MIXED_RU, iters,
                                               relres);
    default:
        return fail(OFEA_ERR_ARG, "unknown solver method");
    }
}

// ------------------------------------------------------------------ solve without allocation
// ofea_solve_noalloc runs the FP64 point- or block-Jacobi PCG of ofea_solve
// on caller-owned device vectors with every buffer of the run (CG vectors,
// true residual, preconditioner, partial sums, scalars) in one file-scope
// device array, so that the call allocates nothing.  It launches the same
// kernels with the same grids as ofea_solve (batches launched directly, one
// batch in flight), so both give the same iterates bit for bit.
//
// Layout of the array, in doubles: r z p p2 q r64 y64 dinv (NA_N each),
// minv (3 NA_N: 9 per free node), residual partials (2 per 256 DOFs), CG
// partials (4 CG_NB), then the scalar block.
constexpr size_t NA_N = OFEA_NOALLOC_MAX_DOFS;
constexpr size_t NA_VBLK = (NA_N + RED_WG - 1) / RED_WG;
constexpr size_t NA_DOUBLES = 11 * NA_N + 2 * NA_VBLK + 4 * CG_NB;
constexpr size_t NA_SC = (sizeof(DevScalars) + sizeof(double) - 1) / sizeof(double);
static __device__ __align__(256) double scratch[NA_DOUBLES + NA_SC];

// Device address of each GPU's copy of scratch, taken once per device at
// context creation.  In host code the array's name is only a stand-in: every
// use of the array's memory goes through these pointers.
constexpr int NA_MAX_DEV = 64;
static double *g_na_ptr[NA_MAX_DEV] = {};

int noalloc_init(int device)
{
    if (device < 0 || device >= NA_MAX_DEV) return fail(OFEA_ERR_LIMIT, "device index beyond the scratch table");
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (g_na_ptr[device]) return OFEA_OK;
    double *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));  // the caller made `device` current
    if (e != hipSuccess || !ptr)
        return fail(OFEA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_na_ptr[device] = ptr;
    return OFEA_OK;
}

size_t noalloc_bytes() { return sizeof(double) * (NA_DOUBLES + NA_SC); }

int solve_noalloc(ofea_ctx *ctx, int method, const double *d_b, double *d_x, double rtol, int maxit, int *iters,
                  double *relres)
{
    if (method != OFEA_PCG_JACOBI && method != OFEA_PCG_BLOCK_JACOBI)
        return fail(OFEA_ERR_ARG, "ofea_solve_noalloc takes OFEA_PCG_JACOBI or OFEA_PCG_BLOCK_JACOBI");
    if (!ctx->assembled) return fail(OFEA_ERR_STATE, "ofea_assemble first");
    const int n = ctx->P.n, nf = ctx->P.nfree;
    if ((size_t)n > NA_N)
        return fail(OFEA_ERR_LIMIT, "ofea_solve_noalloc holds at most " + std::to_string(NA_N) + " DOFs, the system has " +
                                        std::to_string(n));
    double *base = ctx->device >= 0 && ctx->device < NA_MAX_DEV ? g_na_ptr[ctx->device] : nullptr;
    if (!base) return fail(OFEA_ERR_STATE, "scratch of this device not set up (ofea_create)");
    if (n == 0) {
        if (iters) *iters = 0;
        if (relres) *relres = 0.0;
        return OFEA_OK;
    }
    // carve the array
    VecSet<double> v;
    double *c = base;
    v.r = c, c += NA_N;
    v.z = c, c += NA_N;
    v.p = c, c += NA_N;
    v.p2 = c, c += NA_N;
    v.q = c, c += NA_N;
    PcgBufs B;
    B.r64 = c, c += NA_N;
    B.y64 = c, c += NA_N;
    double *dinv = c;
    c += NA_N;
    double *minv = c;
    c += 3 * NA_N;
    B.partials = c, c += 2 * NA_VBLK;
    B.P = CGParts{c, c + 2 * CG_NB, c + 3 * CG_NB};
    c += 4 * CG_NB;
    B.sc = reinterpret_cast<DevScalars *>(c);
    B.nfree = nf;
    B.n = n;
    B.nblk_vec = (n + RED_WG - 1) / RED_WG;
    B.spmv_group = ctx->ws ? ctx->ws->spmv_group : 16;  // the group ofea_solve would use
    const SysView S{ctx->d_node_blk, ctx->d_bcol, ctx->d_val, dinv, d_b, d_x};
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_precond_setup, dim3((nf + RED_WG - 1) / RED_WG), dim3(RED_WG), 0, st, nf, ctx->d_node_blk,
                       ctx->d_bcol, ctx->d_val, dinv, minv, (float *)nullptr, (double *)nullptr);
    const unsigned gv = B.nblk_vec;
    hipLaunchKernelGGL((k_fill<double>), dim3(gv), dim3(RED_WG), 0, st, n, d_x, 0.0);
    hipLaunchKernelGGL((k_fill<double>), dim3(gv), dim3(RED_WG), 0, st, n, v.p, 0.0);
    hipLaunchKernelGGL((k_fill<double>), dim3(gv), dim3(RED_WG), 0, st, n, v.p2, 0.0);
    // the scalar block keeps the previous call's values: the host rewrites it
    // (ticket included) before any kernel reads it
    DevScalars Sh{};
    Sh.ticket = 0;
    OFEA_HIP(hipMemcpyAsync(B.sc, &Sh, sizeof(DevScalars), hipMemcpyHostToDevice, st));
    auto read_sc = [&]() -> int {
        OFEA_HIP(hipMemcpyAsync(&Sh, B.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, st));
        OFEA_HIP(hipStreamSynchronize(st));
        return OFEA_OK;
    };
    auto write_sc = [&]() -> int {
        OFEA_HIP(hipMemcpyAsync(B.sc, &Sh, sizeof(DevScalars), hipMemcpyHostToDevice, st));
        OFEA_HIP(hipStreamSynchronize(st));  // Sh is a stack value and changes next
        return OFEA_OK;
    };
    const int PCm = method == OFEA_PCG_BLOCK_JACOBI ? 1 : 0;
    auto replace = [&](int restart) {
        if (PCm)
            hipLaunchKernelGGL((k_replace<double, 1, false>), dim3(CG_NB), dim3(UP_WG), 0, st, nf, B.r64, v.r, v.z,
                               restart, dinv, minv, B.P, B.co, B.sc);
        else
            hipLaunchKernelGGL((k_replace<double, 0, false>), dim3(CG_NB), dim3(UP_WG), 0, st, nf, B.r64, v.r, v.z,
                               restart, dinv, minv, B.P, B.co, B.sc);
    };
    auto batch = [&]() {
        if (PCm)
            cg_batch<double, double, double, 1, false>(ctx, B, S, ctx->d_val, v, d_x, minv);
        else
            cg_batch<double, double, double, 0, false>(ctx, B, S, ctx->d_val, v, d_x, minv);
    };
    int rc = true_residual(ctx, B, S, d_b, d_x);  // r64 = b and ||b||^2
    if (rc || (rc = read_sc())) return rc;
    const double bb = Sh.s[S_BB];
    if (bb == 0.0) {
        if (iters) *iters = 0;
        if (relres) *relres = 0.0;
        return OFEA_OK;
    }
    // the same control flow as ofea_solve's FP64 methods (no reliable updates)
    double tol2 = rtol * rtol * bb;
    Sh.par0 = 0;
    Sh.it0 = 0;
    Sh.it = 0;
    Sh.maxit = maxit;
    Sh.done = 0;
    Sh.ticket = 0;
    Sh.s[S_TOL2] = tol2;
    Sh.s[S_RU2] = -1.0;
    if ((rc = write_sc())) return rc;
    replace(1);
    OFEA_HIP(hipGetLastError());
    double rel = 1.0;
    int tighten = 0;
    for (;;) {
        do {
            batch();
            OFEA_HIP(hipGetLastError());
            if ((rc = read_sc())) return rc;
        } while (!Sh.done);
        const int done = Sh.done;
        if (done == 3) return fail(OFEA_ERR_STATE, "PCG breakdown (matrix or preconditioner not SPD, or NaN)");
        const int it_done = Sh.it, par_done = Sh.par0 ^ (Sh.bi_done & 1);
        if ((rc = true_residual(ctx, B, S, d_b, d_x)) || (rc = read_sc())) return rc;
        rel = std::sqrt(Sh.s[S_TRR] / bb);
        if (rel <= rtol || done == 2) break;
        if (done == 1 && ++tighten > 8) break;
        if (done == 1) tol2 *= 0.25;
        Sh.par0 = par_done;
        Sh.it0 = it_done;
        Sh.done = 0;
        Sh.s[S_TOL2] = tol2;
        if ((rc = write_sc())) return rc;
        replace(0);
        OFEA_HIP(hipGetLastError());
    }
    if (iters) *iters = Sh.it;
    if (relres) *relres = rel;
    return rel <= rtol ? OFEA_OK : OFEA_NOT_CONVERGED;
}

// plain streaming copy (16-byte vector loads/stores), grid-stride
__global__ __launch_bounds__(RED_WG) void k_stream_copy(long long n16, const uint4 *__restrict__ src,
                                                        uint4 *__restrict__ dst)
{
    long long i = (long long)blockIdx.x * RED_WG + threadIdx.x;
    const long long stride = (long long)gridDim.x * RED_WG;
    for (; i < n16; i += stride) dst[i] = src[i];
}

int stream_copy(ofea_ctx *ctx, long long bytes, int reps, float *ms)
{
    if (bytes < 16 || reps < 1) return fail(OFEA_ERR_ARG, "bytes >= 16 and reps >= 1 required");
    const long long n16 = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    OFEA_HIP(hipMalloc((void **)&a, n16 * 16));
    if (hipMalloc((void **)&b, n16 * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(OFEA_ERR_HIP, "out of device memory");
    }
    (void)hipMemsetAsync(a, 1, n16 * 16, ctx->stream);
    const unsigned grid = (unsigned)std::min<long long>((n16 + RED_WG - 1) / RED_WG, 256LL * 64);
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(RED_WG), 0, ctx->stream, n16, a, b);
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    for (int i = 0; i < reps; ++i) hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(RED_WG), 0, ctx->stream, n16, a, b);
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    hipError_t e = hipEventSynchronize(ctx->ev1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(OFEA_ERR_HIP, hipGetErrorString(e));
    *ms = t / reps;
    return OFEA_OK;
}

}  // namespace ofea
