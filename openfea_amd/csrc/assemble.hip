// assemble.hip -- FP64 stiffness assembly into the reference's CSC arrays.
//
// Replaces assamble_stiffness::fill_CSR_sparse_mat (src/assemble.cpp:95-151)
// with the element stiffness of build_ele_stiff_mat_gauss
// (src/Elements/C3D8_Gauss.cpp:80-108) and build_ele_stiff_mat_hammer
// (src/Elements/C3D4_hammer.cpp:61-73).
//
// Two C3D4 paths: k_assemble_tet_chunks (the full linear assembly; each chunk
// element is computed once per chunk, see its section) and k_assemble_tet
// (the pair-gather kernel below: sparse plans of the CA update, and meshes
// beyond the chunk plan's capacity).  C3D8 uses k_assemble_hex.
//
// Design of the pair-gather kernels (column-owner gather, no atomics):
//   * one 256-thread workgroup owns a chunk of consecutive free nodes, i.e.
//     the CSC columns 3r..3r+2 of each node r of the chunk;
//   * phase 1 computes, for every (owned node r, incident element e) pair,
//     the element data needed for row block a = local index of r:
//       C3D4: the constant gradients g_b (4x3) and lambda*V, mu*V;
//       C3D8: one lane per (pair, Gauss point) computes the point's gradients
//             and the 8 blocks K_ab at that point; an 8-lane xor-shuffle
//             reduce-scatter leaves lane (pair, b) holding the Gauss sum K_ab;
//   * phase 2 sums, for every 3x3 block (r, k) of the chunk's columns, its
//     contributors (precomputed list, staged in LDS, ascending element order:
//     the reference's summation order) and writes the 9 CSC values once; C3D8
//     uses one lane per block, C3D4 one lane per off-diagonal block and three
//     per diagonal block (see the C3D4 section).
// For isotropic D = [lambda, mu] the block is
//   K_ab[c][d] = w (lambda g_a[c] g_b[d] + mu g_a[d] g_b[c] + mu delta_cd g_a.g_b),
// algebraically equal to B_a^T D B_b, and the CSC slot (row 3s+d, col 3r+c)
// receives K_ab[c][d] (src/assemble.cpp:128-146 adds Ke(mm, nn) at row dof[nn],
// column dof[mm]).
#include "elem_device.h"
#include "ofea_internal.h"

namespace ofea {

constexpr int ASM_WG = 256;
constexpr int TET_CAP = PLAN_TET_CAP;  // pairs per tet chunk: one per thread
constexpr int HEX_CAP = PLAN_HEX_CAP;  // pairs per hex chunk: 8 Gauss-point lanes each

// Phase 2 of the C3D8 kernel: one lane per block (r, k) sums its
// contributors, read from the chunk's contributor list staged in LDS (s_ct,
// entries relative to t0).
template <class BlockSource>
__device__ inline void gather_blocks(int b0, int b1, int t0, const int *__restrict__ blk_rk,
                                     const int *__restrict__ contrib_ptr, const uint16_t *s_ct,
                                     const int *__restrict__ col_ptr, double *__restrict__ val, int accumulate,
                                     const BlockSource &src)
{
    for (int bk = b0 + (int)threadIdx.x; bk < b1; bk += ASM_WG) {
        int rk = blk_rk[bk];
        int r = rk >> 8, k = rk & 255;
        double acc[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] = 0.0;
        int t1 = contrib_ptr[bk + 1] - t0;
        for (int t = contrib_ptr[bk] - t0; t < t1; ++t) src.add(s_ct[t], acc);
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
// Phase 2 is split by cost.  An off-diagonal block (r, s) has the few
// elements that contain the edge r-s as contributors; the diagonal block
// (r, r) has every element of r (about 24 in a Kuhn grid).  Off-diagonal
// blocks get one lane each; each diagonal block gets six lanes, two per row
// c (each half of the node's pairs), which sum the cheap symmetric form
//   K_rr[c][d] = sum_e ((lambda + mu) V g_c g_d + delta_cd mu V g.g)
// in element order and add the two halves.  Poisson's ratio is one value per
// launch, so lambda_e = rho mu_e with rho = 2 nu / (1 - 2 nu) for every
// element and only mu V is kept per pair:
//   K_ab = rho Q + Q^T + I tr Q,  Q = sum_e mu_e V_e g_a (x) g_b.
// The chunk's contributor list is staged in LDS first, so the summation
// loops issue no global loads.
constexpr int TET_CT = 3 * TET_CAP;  // off-diagonal contributor entries of a chunk (<= 3 per pair)
static_assert(TET_CT % 256 == 0, "whole rounds of the workgroup");

// value of lane ^ 1 / lane ^ 2, through DPP quad_perm [1,0,3,2] / [2,3,0,1]
template <int CTRL>
__device__ inline double dpp_quad(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ inline double dpp_xor1(double x) { return dpp_quad<0xB1>(x); }
__device__ inline double dpp_xor2(double x) { return dpp_quad<0x4E>(x); }

// A phase-2 task: an off-diagonal block (r, k) or half of row c of the
// diagonal block of r.  The global loads (the block's chunk-local node and
// contributor range) are issued before phase 1 so that they overlap the
// element computations; the rest comes from the chunk's node tables in LDS.
struct TetTask {
    int kind;          // 0 none, 1 off-diagonal block, 2 diagonal half-row
    int rl, k, c0, c1;  // chunk-local node, block slot k in column r; contributor range (1) or pair range (2)
    int c;             // row of the diagonal block (kind 2)
    int base, stride;  // CSC offset of column 3r and the column length 3 deg(r)
};

// global part of a task (off-diagonal candidates): chunk-local node and
// contributor range
template <bool SPARSE>
__device__ inline void tet_task_load(int task, int nblk, int b0, int t0, const uint8_t *__restrict__ blk_rl,
                                     const int *__restrict__ off_ptr, const int *__restrict__ blk_gid, TetTask &t)
{
    // every field is written on both paths (selects, no partial writes), so
    // the task stays in registers instead of a dynamically indexed stack slot
    const bool off = task < nblk;
    const int bk = b0 + (off ? task : 0);
    int rl1 = 0, k1 = 0, c0 = 0, c1 = 0;
    if (off) {
        rl1 = blk_rl[bk];
        // global block index until resolved: a full plan lists every block in
        // order, a sparse plan (a few modified elements) names each one
        k1 = SPARSE ? blk_gid[bk] : bk;
        c0 = off_ptr[bk] - t0;
        c1 = off_ptr[bk + 1] - t0;
    }
    const int j = (task - nblk) >> 1;  // half-row task 2 (3 node + c) + h
    const int rl2 = off ? 0 : j / 3;
    t.kind = off ? 1 : 2;
    t.rl = off ? rl1 : rl2;
    t.k = k1;
    t.c0 = c0;
    t.c1 = c1;
    t.c = off ? 0 : j - 3 * rl2;
}

// LDS part: node tables s_nb (node_blk of the chunk's nodes, nnode + 1), s_dk
// (diagonal slot), s_np (first pair, chunk-relative, nnode + 1)
__device__ inline void tet_task_resolve(TetTask &t, const int *s_nb, const int *s_dk, const int *s_np)
{
    const int nb = s_nb[t.rl], deg = s_nb[t.rl + 1] - nb;
    t.base = 9 * nb;
    t.stride = 3 * deg;
    if (t.kind == 1) {
        t.k -= nb;
        if (t.k == s_dk[t.rl]) t.kind = 0;  // the diagonal block: its rows are kind-2 tasks
    } else {
        t.k = s_dk[t.rl];
        if (t.k < 0) t.kind = 0;
        t.c0 = s_np[t.rl];
        t.c1 = s_np[t.rl + 1];
    }
}

template <bool SPARSE>
__global__ __launch_bounds__(ASM_WG) void k_assemble_tet(const double *__restrict__ xyz, const int *__restrict__ conn,
                                                         const double *__restrict__ Eel, double E, double nu,
                                                         const int *__restrict__ chunk_pair,
                                                         const int *__restrict__ chunk_blk,
                                                         const int *__restrict__ chunk_node,
                                                         const int *__restrict__ node_pair, const int *__restrict__ pair,
                                                         const uint8_t *__restrict__ blk_rl,
                                                         const int *__restrict__ off_ptr,
                                                         const uint16_t *__restrict__ off_contrib,
                                                         const int *__restrict__ blk_gid,
                                                         const int *__restrict__ node_blk,
                                                         const int *__restrict__ diag_k, double *__restrict__ val,
                                                         int accumulate, int nchunks)
{
    // structure-of-arrays: phase 1 (lane = pair) writes consecutive addresses.
    // Slot j of a pair holds the gradient of local node (a + j) & 3, so slot 0
    // is the owned node a; contributor entries name (pair, slot).
    __shared__ double s_g[12][TET_CAP];  // s_g[3 j + c][pair] = d N_(a+j) / d x_c
    __shared__ double s_w[TET_CAP];      // mu V (signed V)
    __shared__ uint16_t s_ct[TET_CT];
    __shared__ int s_nb[PLAN_CHUNK_NODES + 1], s_dk[PLAN_CHUNK_NODES], s_np[PLAN_CHUNK_NODES + 1];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int p0 = chunk_pair[chunk], p1 = chunk_pair[chunk + 1];
    const int b0 = chunk_blk[chunk], b1 = chunk_blk[chunk + 1];
    const int n0 = chunk_node[chunk], nnode = chunk_node[chunk + 1] - n0;  // <= PLAN_CHUNK_NODES
    const int t0 = off_ptr[b0], nct = off_ptr[b1] - t0;                    // <= 3 (p1 - p0)
    const int nblk = b1 - b0;
    // Staged tables are loaded into registers here and stored to LDS after
    // phase 1: a store right after its load would hold the wave until the load
    // returns, before it has issued the element loads of phase 1.
    const bool node_lane = (int)threadIdx.x <= nnode;
    int nb_v = 0, np_v = 0, dk_v = 0;
    if (node_lane) {
        nb_v = node_blk[n0 + threadIdx.x];
        np_v = node_pair[n0 + threadIdx.x] - p0;
        if ((int)threadIdx.x < nnode) dk_v = diag_k[n0 + threadIdx.x];
    }
    constexpr int CT_PER = TET_CT / ASM_WG;
    uint16_t ct_v[CT_PER];
#pragma unroll
    for (int u = 0; u < CT_PER; ++u) {
        const int t = (int)threadIdx.x + u * ASM_WG;
        ct_v[u] = t < nct ? off_contrib[t0 + t] : (uint16_t)0;
    }
    // Phase-2 lanes: waves 0-2 take the off-diagonal blocks and wave 3 the
    // diagonal half-rows (task 2 (3 node + c) + h), so that no wave runs both
    // loop bodies.
    constexpr int OFF_LANES = ASM_WG - 64;
    const bool diag_wave = (int)threadIdx.x >= OFF_LANES;
    const int lane2 = diag_wave ? (int)threadIdx.x - OFF_LANES : (int)threadIdx.x;
    const int step2 = diag_wave ? 64 : OFF_LANES;
    const int ntask2 = diag_wave ? 6 * nnode : nblk;
    const int task_base = diag_wave ? nblk : 0;
    TetTask first{0, 0, 0, 0, 0, 0, 0, 0};
    if (lane2 < ntask2) tet_task_load<SPARSE>(task_base + lane2, nblk, b0, t0, blk_rl, off_ptr, blk_gid, first);
    double lam0, mu0;
    lame<double>(E, nu, lam0, mu0);
    const double rho = 2.0 * nu / (1.0 - 2.0 * nu);  // lambda / mu
    for (int p = p0 + (int)threadIdx.x; p < p1; p += ASM_WG) {
        const int pk = pair[p];
        const int el = pk >> 3, a = pk & 7, lp = p - p0;
        const int *cn = conn + 4 * el;
        double X[12];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int node = cn[b];
            X[3 * b] = xyz[3 * node];
            X[3 * b + 1] = xyz[3 * node + 1];
            X[3 * b + 2] = xyz[3 * node + 2];
        }
        double g[3][4];
        const double detJ = tet_grads(X, g);
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) s_g[3 * ((b - a) & 3) + c][lp] = g[c][b];
        double lam = lam0, mu = mu0;
        if (Eel) lame<double>(Eel[el], nu, lam, mu);
        const double V = detJ / 6.0;  // signed (src/Elements/C3D4.cpp:45-54 does not abs())
        s_w[lp] = mu * V;
    }
#pragma unroll
    for (int u = 0; u < CT_PER; ++u) {
        const int t = (int)threadIdx.x + u * ASM_WG;
        if (t < nct) s_ct[t] = ct_v[u];
    }
    if (node_lane) {
        s_nb[threadIdx.x] = nb_v;
        s_np[threadIdx.x] = np_v;
        if ((int)threadIdx.x < nnode) s_dk[threadIdx.x] = dk_v;
    }
    __syncthreads();
    for (int t2 = lane2; t2 < ntask2; t2 += step2) {
        TetTask tk = first;
        if (t2 != lane2) tet_task_load<SPARSE>(task_base + t2, nblk, b0, t0, blk_rl, off_ptr, blk_gid, tk);
        tet_task_resolve(tk, s_nb, s_dk, s_np);
        if (tk.kind == 1) {  // off-diagonal block (r, k): K_ab = rho Q + Q^T + I tr(Q)
            double Q[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) Q[i] = 0.0;
#pragma unroll 2
            for (int t = tk.c0; t < tk.c1; ++t) {
                const int ci = s_ct[t];
                const int lp = ci >> 2, j = ci & 3;
                const double wm = s_w[lp];
                const double ga[3] = {s_g[0][lp], s_g[1][lp], s_g[2][lp]};
                const double gb[3] = {s_g[3 * j][lp], s_g[3 * j + 1][lp], s_g[3 * j + 2][lp]};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double mc = wm * ga[c];
#pragma unroll
                    for (int d = 0; d < 3; ++d) Q[3 * c + d] += mc * gb[d];
                }
            }
            const double trQ = Q[0] + Q[4] + Q[8];
            double *dst = val + tk.base + 3 * tk.k;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    double v = rho * Q[3 * c + d] + Q[3 * d + c] + (c == d ? trQ : 0.0);
                    if (accumulate) v += dst[c * tk.stride + d];
                    dst[c * tk.stride + d] = v;
                }
        }
        if (diag_wave) {  // half h of row c of the diagonal block, pairs in element order
            // lanes 2m and 2m + 1 hold the two halves of one row; both run this
            // branch in the same iteration (the task count is even)
            double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, accd = 0.0;
            const int c = tk.c, h = t2 & 1;
            const int mid = (tk.c0 + tk.c1 + 1) >> 1;
            const int l0 = h ? mid : tk.c0, l1 = h ? tk.c1 : mid;
            if (tk.kind == 2) {
#pragma unroll 4
                for (int lp = l0; lp < l1; ++lp) {
                    const double g0 = s_g[0][lp], g1 = s_g[1][lp], g2 = s_g[2][lp];
                    const double wm = s_w[lp];
                    const double sc = wm * (c == 0 ? g0 : (c == 1 ? g1 : g2));
                    acc0 += sc * g0;
                    acc1 += sc * g1;
                    acc2 += sc * g2;
                    accd += wm * (g0 * g0 + g1 * g1 + g2 * g2);
                }
            }
            acc0 += dpp_xor1(acc0);
            acc1 += dpp_xor1(acc1);
            acc2 += dpp_xor1(acc2);
            accd += dpp_xor1(accd);
            if (tk.kind == 2 && h == 0) {
                const double rho1 = rho + 1.0;
                acc0 = rho1 * acc0 + (c == 0 ? accd : 0.0);
                acc1 = rho1 * acc1 + (c == 1 ? accd : 0.0);
                acc2 = rho1 * acc2 + (c == 2 ? accd : 0.0);
                double *dst = val + tk.base + 3 * tk.k + c * tk.stride;
                if (accumulate) {
                    acc0 += dst[0];
                    acc1 += dst[1];
                    acc2 += dst[2];
                }
                dst[0] = acc0;
                dst[1] = acc1;
                dst[2] = acc2;
            }
        }
    }
}

// ------------------------------------------------------------------ C3D4, chunk plan
// The linear C3D4 path (TetChunkPlan, built by build_tet_chunks): one
// workgroup per chunk of consecutive free nodes.  Every per-chunk array has
// a fixed stride, so all loads but the coordinate gather are issued at once.
//   staging: lane i < TC_NTAB loads the coordinates of chunk node i into LDS;
//   phase 1: lane e forms the cofactor gradients of chunk element e,
//            c_1 = e1 x e2, c_2 = e2 x e0, c_3 = e0 x e1, c_0 = -(c_1 + c_2 + c_3)
//            (e_i = x_(i+1) - x_0; grad N_a = c_a / det, det = e0 . c_1) and
//            f = mu / (6 det) into LDS rows (row 3 a + c = c_a[c], row 12 = f);
//            since w g_a (x) g_b = f c_a (x) c_b with w = mu det / 6 the
//            element needs one division and no scaling of its gradients.
//            The row stride TC_RS is a multiple of 16 doubles, so every row of
//            a slot lies in the same bank and the plan orders each lane's
//            contributors so that a quarter wave reads different slots modulo
//            16 at each step; slot TC_ECAP is a zero element;
//   phase 2: lane l sums Q = sum (f c_col) (x) c_row over its contributors
//            (the wave's longest run; shorter runs read the zero element),
//            the lanes of a group add their sums by a fixed DPP tree (only the
//            steps the wave's largest group needs) and the first lane of the
//            block forms K = rho Q + Q^T + I tr Q (a diagonal block has
//            c_col = c_row);
//   output:  every block goes to an LDS image of the chunk's contiguous slice
//            of the CSC values, which is then copied out with 16-byte stores.
// Summation order: each lane's run in the plan's order (its contributors are
// those of src/assemble.cpp:106-149, reordered for the LDS banks), then the
// group tree; every order is fixed, so the values are bitwise reproducible.
constexpr int TC_RS = 256;              // LDS row stride in doubles (rows of slots 0..TC_ECAP)
constexpr int TC_ROWB = 8 * TC_RS;      // bytes per row
constexpr int TC_WOFF = 12 * TC_ROWB;   // byte offset of the f row
static_assert(TC_ECAP <= 255, "element slots are 8-bit");
static_assert(TC_RS > TC_ECAP && TC_RS % 16 == 0, "a multiple of 16 doubles: every row of a slot is in the same LDS bank");
static_assert(9 * TC_WG + 1 <= 13 * TC_RS, "the slice image (at most one block per lane) fits the element rows");
static_assert(3 * TC_NTAB <= 13 * TC_RS, "the staged coordinates fit the element rows");

__device__ inline double lds_f64(const char *base, unsigned off) { return *reinterpret_cast<const double *>(base + off); }

// x + (value of the partner lane) for lanes with do_it, x otherwise; the
// partner is given by a DPP control that stays within a row of 16 lanes
template <int CTRL>
__device__ inline double tree_step(double x, bool do_it)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    const double s = x + __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    return do_it ? s : x;
}

// one cross product: c = u x v
__device__ inline void cross3(const double u[3], const double v[3], double c[3])
{
    c[0] = u[1] * v[2] - u[2] * v[1];
    c[1] = u[2] * v[0] - u[0] * v[2];
    c[2] = u[0] * v[1] - u[1] * v[0];
}

__global__ __launch_bounds__(TC_WG) void k_assemble_tet_chunks(const double *__restrict__ xyz,
                                                               const double *__restrict__ Eel, double fsc,
                                                               double fsc_per_E, double rho,
                                                               const int *__restrict__ hdr,
                                                               const int *__restrict__ ntab,
                                                               const uint32_t *__restrict__ eloc,
                                                               const int *__restrict__ eid,
                                                               const uint4 *__restrict__ task,
                                                               double *__restrict__ val, int accumulate, int nchunks)
{
    __shared__ double s_g[13 * TC_RS];  // staged coordinates, then the element rows, then the slice image
    const char *lb = reinterpret_cast<const char *>(s_g);
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int t = threadIdx.x;
    // every load but the coordinates depends on the chunk index only
    const int b0 = hdr[2 * chunk], nblk = hdr[2 * chunk + 1];
    const int node = t < TC_NTAB ? ntab[(size_t)chunk * TC_NTAB + t] : -1;
    const uint32_t el4 = t < TC_ECAP ? eloc[(size_t)chunk * TC_ECAP + t] : ~0u;
    const int ei = (Eel && t < TC_ECAP) ? eid[(size_t)chunk * TC_ECAP + t] : -1;
    const uint4 tk = task[(size_t)chunk * TC_WG + t];
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    if (node >= 0) {
        const double *p = xyz + 3 * (size_t)node;
        x0 = p[0];
        x1 = p[1];
        x2 = p[2];
    }
    const double fs = ei >= 0 ? Eel[ei] * fsc_per_E : fsc;  // mu / 6
    if (t < TC_NTAB) {
        s_g[t] = x0;
        s_g[TC_NTAB + t] = x1;
        s_g[2 * TC_NTAB + t] = x2;
    }
    __syncthreads();
    const bool has_el = el4 != ~0u;
    double X[12];
    if (has_el) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int sl = (el4 >> (8 * b)) & 255;
            X[3 * b] = s_g[sl];
            X[3 * b + 1] = s_g[TC_NTAB + sl];
            X[3 * b + 2] = s_g[2 * TC_NTAB + sl];
        }
    }
    __syncthreads();  // the coordinates are read: the rows overwrite them
    if (has_el) {
        double e[3][3], c[4][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) e[i][j] = X[3 * (i + 1) + j] - X[j];
        cross3(e[1], e[2], c[1]);
        cross3(e[2], e[0], c[2]);
        cross3(e[0], e[1], c[3]);
#pragma unroll
        for (int j = 0; j < 3; ++j) c[0][j] = -(c[1][j] + c[2][j] + c[3][j]);
        // signed det (src/Elements/C3D4.cpp:45-54 does not abs() the volume)
        const double det = e[0][0] * c[1][0] + e[0][1] * c[1][1] + e[0][2] * c[1][2];
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int j = 0; j < 3; ++j) s_g[(3 * b + j) * TC_RS + t] = c[b][j];
        s_g[12 * TC_RS + t] = fs / det;
    } else if (t == TC_ECAP) {
#pragma unroll
        for (int j = 0; j < 13; ++j) s_g[j * TC_RS + t] = 0.0;
    }
    __syncthreads();
    const uint32_t meta = tk.x;
    const uint32_t cw[3] = {tk.y, tk.z, tk.w};
    const int trip = __builtin_amdgcn_readfirstlane((int)((tk.y >> 12) & 7u));  // the wave's longest run
    double Q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < TC_CPL; ++i) {
        if (i < trip) {
            const unsigned ci = (cw[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
            const unsigned e8 = (ci & 255u) * 8u;
            const unsigned ga = e8 + ((ci >> 8) & 3u) * (3u * TC_ROWB), gb = e8 + ((ci >> 10) & 3u) * (3u * TC_ROWB);
            const double f = lds_f64(lb, e8 + TC_WOFF);
            const double a0 = lds_f64(lb, ga), a1 = lds_f64(lb, ga + TC_ROWB), a2 = lds_f64(lb, ga + 2 * TC_ROWB);
            const double q0 = lds_f64(lb, gb), q1 = lds_f64(lb, gb + TC_ROWB), q2 = lds_f64(lb, gb + 2 * TC_ROWB);
            const double m0 = f * a0, m1 = f * a1, m2 = f * a2;
            Q[0] += m0 * q0;
            Q[1] += m0 * q1;
            Q[2] += m0 * q2;
            Q[3] += m1 * q0;
            Q[4] += m1 * q1;
            Q[5] += m1 * q2;
            Q[6] += m2 * q0;
            Q[7] += m2 * q1;
            Q[8] += m2 * q2;
        }
    }
    // lane groups: aligned, largest first, so lane 0 of a wave has the
    // wave's largest group and the tree runs only the steps it needs
    const int glog = (meta >> 24) & 7;
    const int gmax = __builtin_amdgcn_readfirstlane(glog);
    if (gmax >= 1) {
#pragma unroll
        for (int i = 0; i < 9; ++i) Q[i] = tree_step<0xB1>(Q[i], glog >= 1);  // quad_perm [1,0,3,2]
        if (gmax >= 2) {
#pragma unroll
            for (int i = 0; i < 9; ++i) Q[i] = tree_step<0x4E>(Q[i], glog >= 2);  // quad_perm [2,3,0,1]
            if (gmax >= 3) {
#pragma unroll
                for (int i = 0; i < 9; ++i) Q[i] = tree_step<0x141>(Q[i], glog >= 3);  // row_half_mirror
                if (gmax >= 4) {
#pragma unroll
                    for (int i = 0; i < 9; ++i) Q[i] = tree_step<0x140>(Q[i], glog >= 4);  // row_mirror
                }
            }
        }
    }
    double K[9];
    const double trQ = Q[0] + Q[4] + Q[8];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) K[3 * c + d] = rho * Q[3 * c + d] + Q[3 * d + c] + (c == d ? trQ : 0.0);
    __syncthreads();  // the element rows are no longer read: they become the slice image
    // slice value x sits at s_out[lead + x], so that aligned global pairs are
    // aligned LDS pairs in the copy-out
    const long long v0 = 9LL * b0;
    const int lead = (int)(v0 & 1);
    double *s_out = s_g + lead;
    if ((meta >> 27) & 1u) {
        const int k = meta & 255, nbl = (meta >> 8) & 255, deg = (meta >> 16) & 255;
        double *o = s_out + 9 * nbl + 3 * k;  // column 3r + c holds rows 3s + d at c (3 deg) + 3 k + d
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) o[c * 3 * deg + d] = K[3 * c + d];
    }
    __syncthreads();
    const int len = 9 * nblk;
    double *dst = val + v0;
    if (lead && t == 0) dst[0] = accumulate ? dst[0] + s_out[0] : s_out[0];
    const int npair = (len - lead) >> 1;
    const double2 *s2 = reinterpret_cast<const double2 *>(s_g + 2 * lead);
    double2 *d2 = reinterpret_cast<double2 *>(dst + lead);
    for (int i = t; i < npair; i += TC_WG) {
        double2 v = s2[i];
        if (accumulate) {
            const double2 old = d2[i];
            v.x += old.x;
            v.y += old.y;
        }
        d2[i] = v;
    }
    if (((len - lead) & 1) && t == 0) dst[len - 1] = accumulate ? dst[len - 1] + s_out[len - 1] : s_out[len - 1];
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
    __shared__ uint16_t s_ct[HEX_CAP * 8];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int p0 = chunk_pair[chunk], p1 = chunk_pair[chunk + 1];
    const int b0 = chunk_blk[chunk], b1 = chunk_blk[chunk + 1];
    const int t0 = contrib_ptr[b0], nct = contrib_ptr[b1] - t0;  // <= 8 (p1 - p0) <= ASM_WG
    static_assert(HEX_CAP * 8 <= ASM_WG, "one contributor entry per thread");
    // loaded here, stored to LDS after phase 1 (a store right after its load
    // would hold the wave before it issues the element loads)
    const uint16_t ct_v = (int)threadIdx.x < nct ? contrib[t0 + threadIdx.x] : (uint16_t)0;
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
    if ((int)threadIdx.x < nct) s_ct[threadIdx.x] = ct_v;
    __syncthreads();
    HexSource src{s_blk};
    gather_blocks(b0, b1, t0, blk_rk, contrib_ptr, s_ct, col_ptr, val, accumulate, src);
}

int launch_assemble_plan(ofea_ctx *ctx, int type, const DevPlan &dp, const double *d_xyz, double E, double nu,
                         double *d_out, int accumulate)
{
    if (dp.nchunks == 0) return OFEA_OK;
    const DevTetChunks &tc = ctx->dtc;
    if (type == OFEA_C3D4 && !dp.blk_gid && tc.nchunks > 0 && d_xyz == ctx->d_xyz) {
        // the full linear C3D4 assembly of the context's own mesh: chunk plan
        // mu / 6 for the launch's modulus and per unit modulus (per-element E), and lambda / mu
        double lam1, mu1;
        lame<double>(1.0, nu, lam1, mu1);
        double lamE, muE;
        lame<double>(E, nu, lamE, muE);
        hipLaunchKernelGGL(k_assemble_tet_chunks, dim3(tc.nchunks), dim3(TC_WG), 0, ctx->stream, d_xyz, dp.E,
                           muE / 6.0, mu1 / 6.0, 2.0 * nu / (1.0 - 2.0 * nu), tc.hdr, tc.ntab, tc.eloc, tc.eid,
                           reinterpret_cast<const uint4 *>(tc.task), d_out, accumulate, tc.nchunks);
    } else if (type == OFEA_C3D4 && dp.blk_gid)
        hipLaunchKernelGGL(k_assemble_tet<true>, dim3(dp.nchunks), dim3(ASM_WG), 0, ctx->stream, d_xyz, dp.conn, dp.E,
                           E, nu, dp.chunk_pair, dp.chunk_blk, dp.chunk_node, dp.node_pair, dp.pair, dp.blk_rl,
                           dp.off_ptr, dp.off_contrib, dp.blk_gid, ctx->d_node_blk, ctx->d_diag_k, d_out, accumulate,
                           dp.nchunks);
    else if (type == OFEA_C3D4)
        hipLaunchKernelGGL(k_assemble_tet<false>, dim3(dp.nchunks), dim3(ASM_WG), 0, ctx->stream, d_xyz, dp.conn, dp.E,
                           E, nu, dp.chunk_pair, dp.chunk_blk, dp.chunk_node, dp.node_pair, dp.pair, dp.blk_rl,
                           dp.off_ptr, dp.off_contrib, nullptr, ctx->d_node_blk, ctx->d_diag_k, d_out, accumulate,
                           dp.nchunks);
    else
        hipLaunchKernelGGL(k_assemble_hex, dim3(dp.nchunks), dim3(ASM_WG), 0, ctx->stream, d_xyz, dp.conn, dp.E, E, nu,
                           dp.chunk_pair, dp.chunk_blk, dp.pair, dp.blk_rk, dp.contrib_ptr, dp.contrib,
                           ctx->d_col_ptr, d_out, accumulate, dp.nchunks);
    OFEA_HIP(hipGetLastError());
    return OFEA_OK;
}

int launch_assemble(ofea_ctx *ctx, double E, double nu, double *d_out, bool zero_first)
{
    bool mixed = ctx->has_type[0] && ctx->has_type[1];
    if (zero_first || mixed) OFEA_HIP(hipMemsetAsync(d_out, 0, sizeof(double) * ctx->P.nnz, ctx->stream));
    int acc = (zero_first || mixed) ? 1 : 0;
    for (int t = 0; t < 2; ++t) {
        if (!ctx->has_type[t]) continue;
        int rc = launch_assemble_plan(ctx, t, ctx->dp[t], ctx->d_xyz, E, nu, d_out, acc);
        if (rc) return rc;
    }
    return OFEA_OK;
}

int plan_caps(int type) { return type == OFEA_C3D8 ? HEX_CAP : TET_CAP; }

}  // namespace ofea
