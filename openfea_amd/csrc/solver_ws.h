// solver_ws.h -- PCG work space shared by solver.hip and learned.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace ofea {

enum Scalar { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_TOL2, S_RU2, S_TRR, S_NSCAL };

struct DevScalars {
    double s[S_NSCAL];
    int done;  // 0 running, 1 recursive residual converged, 2 maxit, 3 breakdown, 4 reliable update due
    int it;
    int maxit;
    unsigned int ticket;
    int par0;     // parity of the first iteration of the next batch (set by the host)
    int it0;      // iteration number of the first iteration of the next batch
    int bi_done;  // batch index at which done was set
};

// CG vectors; the search direction is double-buffered (p, p2) because the
// SpMV kernel forms p_new = z + beta p_old on the fly while other workgroups
// still read p_old.
// y = G r of the factorized learned preconditioner (M = G^T G) is carried
// by its own recursion (y -= alpha G q) next to r.
template <class T>
struct VecSet {
    T *r = nullptr, *z = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr, *y = nullptr;
};

// The learned factor G (block lower-triangular on K's node-block pattern) as
// two half block-matrices in the column-of-rows layout of K's values: row i
// of lo holds the blocks G_is, s <= i; row i of up holds (G_si)^T, s >= i
// (the rows of G^T), the diagonal transposed.  Row i's nb blocks start at
// 9 ptr[i]; entry (c, 3 k + d) lies at 9 ptr[i] + 3 nb c + 3 k + d.  Blocks
// keep the order of K's row.  Values in BF16 (the bits of a float's upper
// half, rounded to nearest even).
struct FacMat {
    int *lo_ptr = nullptr, *lo_col = nullptr, *up_ptr = nullptr, *up_col = nullptr;
    unsigned short *lo = nullptr, *up = nullptr;
    long long nlo = 0, nup = 0;  // blocks
    int kver = -1, gen = -1;     // K version and network generation of the values
    int group = 0;               // lanes per node of the factor passes (timed on first use, 0 = not yet)
};

// captured graphs: jacobi, block, mixed, learned, mixed32, learned mixed, and
// the four two-level variants (block, mixed, learned, learned mixed)
constexpr int NSLOT = 12;
// batches of the PCG host loop in flight at most, plus one
constexpr int NQ_PIPE = 4;

// Two-level coarse correction: z = M r + Z A_c^-1 Z^T r, Z = rigid-body modes
// (3 translations, 3 rotations about the centroid) of node aggregates.  An
// aggregate is ra >= CG_AGG consecutive CG ranges (the node ranges of the
// fixed CG grid), so Z^T r is a per-range partial.
constexpr int CG_AGG = 2;
constexpr int CG_NC_MAX = 6 * (512 / CG_AGG);  // coarse dimension bound (512 CG ranges)
struct CoarseArgs {
    const float *d = nullptr;     // 3 * nfree: node position minus its aggregate centroid
    double *g = nullptr;          // 6 * CG_NB: per-range partials of Z^T r, component-major
    double *zq = nullptr;         // 6 * CG_NB: per-range partials of Z^T q (q = K p), component-major
    double *ya = nullptr;         // 2 * CG_NC_MAX: y = A_c^-1 Z^T r per aggregate (6 each), by iteration parity
    const float *ainv = nullptr;  // A_c^-1, (6 nagg)^2 row-major, FP32
    int per = 1;                  // nodes per CG range
    int ra = CG_AGG;              // CG ranges per aggregate
    int nagg = 0;                 // aggregates
};

}  // namespace ofea

struct ofea_solver_ws {
    int n = 0, nfree = 0;
    long long nnz = 0;
    int spmv_group = 16;  // lanes per node in the SpMV (autotuned)
    int nblk_vec = 0, nblk_node = 0;
    double *partials = nullptr;
    double *part_rz = nullptr, *part_rr = nullptr, *part_pq = nullptr;  // CG partials (fixed grid)
    ofea::DevScalars *sc = nullptr;
    ofea::DevScalars *sc_host = nullptr;  // pinned, 1 + NQ_PIPE entries: [0] and the batches in flight
    hipEvent_t pipe_ev[ofea::NQ_PIPE] = {};
    ofea::VecSet<double> v64;
    ofea::VecSet<float> v32;
    double *r64 = nullptr, *y64 = nullptr;  // true residual, K x
    double *dinv = nullptr;                 // point Jacobi, n
    double *minv = nullptr;                 // 3x3 block-Jacobi inverses, 9 * nfree
    float *minv32 = nullptr;                // FP32 copy
    double *dblk = nullptr;                 // 3x3 diagonal blocks of K, 9 * nfree
    float *val32 = nullptr;                 // FP32 values, nnz
    // learned factorized correction M = G^T G for K's rows [0] and the
    // reordered rows of the two-level methods [1]; values rebuilt when K or
    // the network's outputs change
    ofea::FacMat fac[2];
    bool precond_valid = false;
    // two-level coarse space
    ofea::CoarseArgs co;
    int co_nagg = 0;
    float *co_ainv = nullptr;  // (6 nagg)^2, row-major, FP32
    bool co_valid = false;
    bool co_vals_stale = false;  // co_valid, but K's values changed since (solver_invalidate_values)
    int co_builds = 0;           // coarse inverses built
    bool co_ainv_ready = false;  // co_ainv holds an inverse built on this mesh (of K or a batched variant)
    int *co_flag_host = nullptr;  // pinned (inside the sc_host allocation): the factorization's flag
    int4 *co_lst_host = nullptr;  // pinned (same allocation): per-aggregate results of the block lists
    hipEvent_t co_ev = nullptr;   // the copy of those results has landed
    bool co_lst_pending = false;  // the structure was rebuilt and co_lst_host not read yet
    std::vector<int> co_perm_host;
    std::vector<float> co_d_host, co_ainv_host;  // host copies for ofea_coarse_apply
    // The two-level methods run on the free nodes reordered into compact parts
    // (one part = one aggregate = ra consecutive CG ranges of the new order),
    // on a copy of K whose node rows follow that order.
    int *perm = nullptr, *iperm = nullptr;  // nfree: new -> old, old -> new
    int *nb_p = nullptr, *bcol_p = nullptr;  // node_blk and bcol in the new order
    double *val_p = nullptr;                 // K values, node rows in the new order
    float *val32_p = nullptr;
    double *dinv_p = nullptr, *minv_p = nullptr, *dblk_p = nullptr;
    float *minv32_p = nullptr;
    double *b_p = nullptr, *x_p = nullptr;  // right-hand side and solution in the new order
    std::vector<int> iperm_host;
    bool co_struct_valid = false;  // order, pattern and block lists built for this mesh
    int co_bw = 0;                 // half-bandwidth of A_c
    // A_c block lists, built on the device per aggregate b: its slots (coupled
    // aggregates, ascending) at sl_part[b CA_MAXCH], chunk starts of each slot
    // at sl_cptr[b (CA_MAXCH + 1)], chunk starts at ch_beg[b (CA_MAXCH + 1)]
    // (ending with the aggregate's last block), the blocks ((node << 8) | slot
    // in its row) in b's range of sl_blk, and co_lst[b] = (slots, chunks,
    // largest |a - b| of a coupled aggregate a)
    int *sl_part = nullptr, *sl_cptr = nullptr, *ch_beg = nullptr;
    int *sl_blk = nullptr;
    int4 *co_lst = nullptr;
    double *co_xf = nullptr;   // free-node coordinates, caller's free numbering (3 nfree)
    int *co_scan = nullptr;    // tile totals of the row-length scan
    void *co_arena = nullptr;  // device allocation holding every two-level buffer
    double *co_A = nullptr, *co_inv = nullptr;  // A_c (then L and L^T) and its inverse, FP64
    int *co_flag = nullptr;
    hipGraph_t graph[ofea::NSLOT] = {}, graph_tail[ofea::NSLOT] = {};  // CG_BATCH and CG_TAIL iterations
    hipGraphExec_t exec[ofea::NSLOT] = {}, exec_tail[ofea::NSLOT] = {};
    int runs[ofea::NSLOT] = {};  // solves of each slot since its graphs were last dropped
};
