// ofea_internal.h -- shared declarations of libofea_hip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/openfea_amd.h"

namespace ofea {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define OFEA_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t _e = (call);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return ::ofea::fail(OFEA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------- host mesh pieces
// CSC-of-symmetric pattern, identical to assamble_stiffness::build_CSR
// (src/assemble.cpp:19-77).  Built node-block-wise: the three columns of free
// node r hold rows 3s+d for every free neighbour s (sorted) and d = 0..2.
struct Pattern {
    int nfree = 0;                 // free nodes
    int n = 0;                     // 3 * nfree DOFs
    long long nnz = 0;
    std::vector<int> col_ptr;      // n + 1
    std::vector<int> row_idx;      // nnz
    std::vector<int> node_blk;     // nfree + 1: first 3x3 block of node r (= col_ptr[3r] / 9)
    std::vector<int> bcol;         // nnz / 9: neighbour rank of each block, (r, k) order
    std::vector<int> diag_k;       // nfree: k of the diagonal block (r, r) of node r, -1 if none
};

// Column-owner gather plan for one element type.  The free nodes are cut into
// chunks of consecutive ranks; one workgroup owns a chunk's columns, computes
// the element data of every (node, incident element) pair of the chunk and
// gathers each 3x3 block from its contributors in ascending element order
// (the summation order of src/assemble.cpp:106-149).  Every CSC value is
// written once and no atomics are used.
struct GatherPlan {
    int nn = 0;                          // nodes per element (4 or 8)
    int ne = 0;                          // elements of this type
    std::vector<int> elem_global;        // type-local -> global element index
    std::vector<int> conn;               // ne * nn, 0-based node ids
    std::vector<int> chunk_pair;         // nchunks + 1
    std::vector<int> chunk_blk;          // nchunks + 1
    std::vector<int> pair;               // (elem_local << 3) | a, pairs sorted by (node, elem)
    std::vector<int> blk_rk;             // (r << 8) | k per block of the chunk list
    std::vector<int> contrib_ptr;        // blocks + 1
    std::vector<uint16_t> contrib;       // chunk-local pair * nn + slot: C3D8 slot = b, C3D4 slot = (b - a) & 3
    std::vector<int> node_pair;          // nfree + 1: pairs of node r = [node_pair[r], node_pair[r+1])
    std::vector<int> chunk_node;         // nchunks + 1: first free node of each chunk (<= PLAN_CHUNK_NODES nodes)
    // C3D4 linear kernel: contributor lists of the off-diagonal blocks only (its
    // diagonal rows walk the node's pairs), and per block the node's index in
    // its chunk
    std::vector<int> off_ptr;            // blocks + 1
    std::vector<uint16_t> off_contrib;   // as contrib, diagonal blocks empty
    std::vector<uint8_t> blk_rl;         // r - chunk_node[chunk] per block
    std::vector<int> blk_gid;            // sparse plans only: node_blk[r] + k of each listed block
    int nchunks() const { return (int)chunk_pair.size() - 1; }
};

int build_pattern(int nd, int ne, const int *conn, const int *etype, const int *resort, Pattern &P);
// keyword-level deck extensions into D (inp_ext.cpp); D's mesh must be read
int parse_inp_ext(const std::string &path, ofea_deck &D);
// sparse = true leaves out the blocks that no element of the list touches
// (plans over a few modified elements, whose kernels accumulate into K).
int build_gather_plan(int nd, int ne, const int *conn, const int *etype, const int *resort, const Pattern &P,
                      int type, int pair_cap, GatherPlan &G, bool sparse = false);

// Chunk plan of the linear C3D4 assembly (k_assemble_tet_chunks).  A chunk is
// a run of consecutive free nodes, i.e. one contiguous slice of the CSC value
// array.  Every per-chunk array has a fixed stride, so a workgroup finds its
// data from its chunk index alone and the only dependent global load is the
// coordinate gather:
//   ntab  the distinct nodes of the chunk's elements (their coordinates are
//         staged once per chunk in LDS), -1 padded;
//   eloc  per distinct element (ascending) its 4 nodes as 8-bit ntab slots;
//   task  per lane one 16-byte record: a meta word (block slot k in the node's
//         column, the node's first block relative to the chunk, its degree,
//         the lane group and the write flag) and up to TC_CPL contributors
//         (chunk element, slot of the column node, slot of the row node),
//         ordered so that the 16 lanes of a quarter wave read different
//         element slots modulo 16 at each step (LDS banks).  A block with more contributors takes a
//         group of 2, 4, 8 or 16 lanes, each a contiguous run of its list,
//         summed by a fixed DPP tree; groups come first, largest first, so
//         that they are aligned and a wave whose lane 0 has no group has none.
//         Unused contributor entries name a zero element slot (TC_ECAP).
constexpr int TC_WG = 256;    // lanes per chunk workgroup = task records per chunk
constexpr int TC_ECAP = 248;  // distinct elements per chunk; slot TC_ECAP is the zero element
constexpr int TC_NTAB = 160;  // distinct nodes per chunk
constexpr int TC_CPL = 6;     // contributors per lane
constexpr int TC_GMAX = 16;   // lanes per block
struct TetChunkPlan {
    int nchunks = 0;
    std::vector<int> hdr;         // 2 per chunk: first block of the chunk's slice, its block count
    std::vector<int> ntab;        // TC_NTAB per chunk: 0-based node id, -1 = none
    std::vector<uint32_t> eloc;   // TC_ECAP per chunk: ntab slots of nodes 0..3 (8 bits each), ~0u = none
    std::vector<int> eid;         // TC_ECAP per chunk: type-local element index (per-element moduli), -1 = none
    std::vector<uint32_t> task;   // 4 per lane, TC_WG lanes per chunk (meta word, then 6 x 16-bit contributors)
};
// task meta word: k | nbl << 8 | deg << 16 | log2(group) << 24 | write << 27 | count << 28 | diagonal << 31
// contributor: chunk element | slot_col << 8 | slot_row << 10 (the first entry of a lane also
// carries the longest count of the lane's wave in bits 12-14)
// Builds the chunk plan; OFEA_ERR_LIMIT when a node exceeds a chunk's
// capacity (the assembly then keeps the pair-gather kernel).
int build_tet_chunks(int nd, const Pattern &P, const GatherPlan &G, const int *resort, TetChunkPlan &T);

// Pairs per gather chunk, shared by every kernel that walks a GatherPlan:
// C3D4 one pair per thread of a 256-thread workgroup, C3D8 one pair per
// 8-lane Gauss-point group.
constexpr int PLAN_TET_CAP = 256;
constexpr int PLAN_HEX_CAP = 32;
constexpr int PLAN_CHUNK_NODES = 64;  // free nodes per chunk (node tables staged in LDS)

// ---------------------------------------------------------------- device plan copy
struct DevPlan {
    int nn = 0, ne = 0, nchunks = 0, nblk = 0;
    int *conn = nullptr;
    double *E = nullptr;  // per-element modulus (type-local order) or null
    int *chunk_pair = nullptr, *chunk_blk = nullptr, *pair = nullptr, *blk_rk = nullptr, *contrib_ptr = nullptr;
    uint16_t *contrib = nullptr;
    int *node_pair = nullptr, *chunk_node = nullptr;
    int *off_ptr = nullptr;
    uint16_t *off_contrib = nullptr;
    uint8_t *blk_rl = nullptr;
    int *blk_gid = nullptr;  // null for full plans
};
struct DevTetChunks {
    int nchunks = 0;
    int *hdr = nullptr, *ntab = nullptr, *eid = nullptr;
    uint32_t *eloc = nullptr, *task = nullptr;
};

}  // namespace ofea

// --------------------------------------------------------------------- context
struct ofea_deck {
    int nd = 0, ne = 0;
    std::vector<double> xyz;
    std::vector<int> conn, etype, fix_set, load_set;
    int load_dof = 0;
    double load_value = 0.0;
    double E, nu, rho, dt = 0.0, t_total = 0.0;
    // keyword-level extensions (inp_ext.cpp); ext_ok false with ext_error set
    // when a supported keyword block could not be read
    bool ext_ok = false;
    std::string ext_error;
    std::vector<double> f_nodal, v0;               // nd x 3: summed *Cload + *Dsload forces; initial velocity
    std::vector<unsigned char> bc_mask;            // nd: bit d set = DOF d+1 constrained by *Boundary
    int bc_nonzero = 0;                            // constrained (node, range) entries with a nonzero value
    std::vector<double> E_elem, nu_elem, rho_elem;  // ne, NaN without a section
    std::vector<std::pair<std::string, std::string>> sections;  // (elset, material), lower case
};

// CA change file (ReadInfo::CA_read_change, src/data2cae.cpp:519-665)
struct ofea_ca_info {
    std::vector<double> xyz;      // nodes_all x 3
    std::vector<int> kind;        // per modified element: 0 delete, 1 move, 2 add
    std::vector<int> etype;       // per modified element
    std::vector<int> conn;        // 8 per modified element, 0-based node ids
    std::vector<int> idref2new;   // new-model node of each listed node, -1 = removed
};

struct ofea_solver_ws;  // solver work space (solver.hip)

struct ofea_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // mesh (host)
    int nd = 0, ne = 0;
    std::vector<double> xyz;
    std::vector<int> conn, etype, resort;
    ofea::Pattern P;
    ofea::GatherPlan G[2];
    ofea::TetChunkPlan TC;  // linear C3D4 assembly; empty when the mesh exceeds its capacity
    bool has_type[2] = {false, false};
    std::vector<double> E_elem;  // empty = global E
    // device
    double *d_xyz = nullptr;
    int *d_col_ptr = nullptr, *d_node_blk = nullptr, *d_bcol = nullptr, *d_diag_k = nullptr;
    int *d_resort = nullptr;
    double *d_val = nullptr;    // FP64 values, nnz
    ofea::DevPlan dp[2];
    ofea::DevTetChunks dtc;
    bool assembled = false;
    int kver = 0;  // incremented by every assembly of K
    double asm_E = 0.0, asm_nu = 0.0;
    // solver
    ofea_solver_ws *ws = nullptr;
    double *d_rhs = nullptr;
    double *d_x = nullptr;
    bool rhs_set = false;
    // explicit / nl scratch
    void *explicit_ws = nullptr;
    double *d_u = nullptr, *d_fint = nullptr;  // TL displacement / internal force (n)
    // learned preconditioner parameters (device), set by ofea_learned_*
    void *learned = nullptr;
    // batched variants / multi-RHS solves (batched.hip)
    void *batch_ws = nullptr;
    // combined-approximation reanalysis (ca.hip)
    void *ca_ws = nullptr;
    // nodal strain/stress recovery (post.hip)
    void *post_ws = nullptr;
    // node-based smoothed FEM (sfem.hip): once set up, P is the smoothing-domain pattern
    void *sfem_ws = nullptr;
    bool sfem = false;
};

namespace ofea {
// kernels (assemble.hip)
int launch_assemble(ofea_ctx *ctx, double E, double nu, double *d_out, bool zero_first);
// one plan of element type `type` over coordinates d_xyz (any plan built
// against ctx->P); accumulate = 1 adds to d_out instead of overwriting
int launch_assemble_plan(ofea_ctx *ctx, int type, const DevPlan &dp, const double *d_xyz, double E, double nu,
                         double *d_out, int accumulate);
int upload_plan(const GatherPlan &G, DevPlan &d);
void free_plan(DevPlan &d);
// solver (solver.hip)
int solver_spmv(ofea_ctx *ctx, const double *d_x, double *d_y);
// two n-vectors of the solver work space for host<->device staging between
// solves (their contents do not survive a solve)
int solver_staging(ofea_ctx *ctx, double **d_x, double **d_y);
// y = A x for any FP64 value array on the pattern of K
int solver_spmv_vals(ofea_ctx *ctx, const double *d_vals, const double *d_x, double *d_y);
int solver_solve(ofea_ctx *ctx, int method, double rtol, int maxit, int *iters, double *relres);
void solver_free(ofea_ctx *ctx);
// solve without allocation: the static scratch of each device (taken once at
// context creation) and the FP64 (block) Jacobi PCG over it
int noalloc_init(int device);
size_t noalloc_bytes();
int solve_noalloc(ofea_ctx *ctx, int method, const double *d_b, double *d_x, double rtol, int maxit, int *iters,
                  double *relres);
void solver_invalidate(ofea_ctx *ctx);  // K changed: preconditioner and FP32 copy are stale
// K changed a little (the next Newton tangent): the reordered values and the
// block preconditioners are rebuilt on the next two-level solve, the coarse
// inverse A_c^-1 of the earlier K is kept (still an SPD coarse correction)
void solver_invalidate_values(ofea_ctx *ctx);
int coarse_builds(ofea_ctx *ctx);  // A_c^-1 built since the solver work space was created
int stream_copy(ofea_ctx *ctx, long long bytes, int reps, float *ms);
// explicit (explicit.hip)
int explicit_run(ofea_ctx *ctx, double E, double nu, const double *mass, const double *f_ext, const double *v0,
                 const int *fixed_mask, double dt, int nsteps, int precision, double *u_out);
int explicit_analysis(ofea_ctx *c, double E, double nu, const double *mass, const double *f_ext, const double *v0,
                      const int *fixed_mask, double time_total, double time_step, int output_gap, int precision,
                      int max_steps, int max_frames, double *frames, int *frame_index, double *frame_time,
                      int *n_frames, int *n_steps);
int lumped_mass(ofea_ctx *ctx, double rho, double *mass);
void explicit_free(ofea_ctx *ctx);
// learned preconditioner (learned.hip)
void learned_free(ofea_ctx *ctx);
int learned_nparam();
int learned_init(ofea_ctx *c, unsigned long long seed);
int learned_set_params(ofea_ctx *c, const float *params, int nparam);
int learned_get_params(ofea_ctx *c, float *params, int nparam);
int learned_loss_grad(ofea_ctx *c, double *loss, float *grad);
int learned_train(ofea_ctx *const *cs, int nctx, int steps, float lr, double *loss_hist);
int learned_blocks(ofea_ctx *c, double *blocks, float *features);
int learned_outputs(ofea_ctx *c, float *uv);
int learned_refresh(ofea_ctx *c);
int learned_precision(ofea_ctx *c, int precision);
// G of the learned factorized correction for the resident K (reordered rows
// of the two-level methods when permuted), rebuilt when K or the network's
// outputs changed (learned.hip)
int learned_factor(ofea_ctx *c, bool permuted);
// index arrays of the factor's half matrices for view 0 (K's rows) or 1 (the
// two-level order) and the iteration's y vectors (solver.hip)
void solver_drop_graphs(ofea_ctx *ctx);
int fac_prepare(ofea_ctx *ctx, int view);
// lanes per node of the factor passes of a view, timed once its values are built (solver.hip)
int fac_tune(ofea_ctx *ctx, int view);
// nonlinear (nl.hip)
int nl_assemble(ofea_ctx *ctx, double E, double nu, const double *u, double *nz_val, double *f_int);
int nl_solve(ofea_ctx *c, double E, double nu, const double *F, int n_incr, double tol, int max_newton, int method,
             double inner_rtol, double *u_out, int *newton_total, int *linear_total, double *relres);
int nl_solve_ref(ofea_ctx *c, double E, double nu, const double *F, double res_limit, int max_loop, int method,
                 double inner_rtol, double *u_out, int *loops, int *linear_total, double *res_norm_out);
// batched variants (batched.hip)
int assemble_batch(ofea_ctx *c, int nvar, const double *E_var, double nu, double *out);
int solve_batch(ofea_ctx *c, int nvar, int nrhs, const double *B, double *X, double rtol, int maxit, int precond,
                int *iters, double *relres);
int autotune(ofea_ctx *c, int *best_group, float *times);
int batch_tiles(ofea_ctx *c, int nrhs, int *group, float *ms);
void batch_free(ofea_ctx *c);
// combined-approximation reanalysis (ca.hip)
int ca_delta(ofea_ctx *c, double E, double nu, int nmod, const int *kind, const int *conn_mod, const int *etype_mod,
             const double *xyz_m, double *dK_val);
int ca_solve(ofea_ctx *c, int n_basis, const double *u0, const double *F, int method, double rtol, int maxit,
             double *u_out, int *n_used, int *pcg_iters, double *reduced);
void ca_free(ofea_ctx *c);
// node-based smoothed FEM (sfem.hip)
int sfem_setup(ofea_ctx *c);
int sfem_assemble(ofea_ctx *c, double E, double nu, double *nz_val);
void sfem_free(ofea_ctx *c);
// nodal strain / stress recovery (post.hip)
int post_nodal(ofea_ctx *c, double E, double nu, const double *u_full, int quantity, double *elem_out,
               double *node_out);
void post_free(ofea_ctx *c);
// SpMV of K (FP64) with lane-group size `group`, averaged over reps launches (solver.hip)
int solver_spmv_group_time(ofea_ctx *ctx, int group, int reps, float *ms);
// two-level coarse correction Z A_c^-1 Z^T r on the host (solver.hip)
int coarse_apply(ofea_ctx *ctx, const double *r, double *out);
int coarse_aggregates(ofea_ctx *ctx, int *agg);
// The two-level coarse space of a value array on K's pattern (batched
// variants, batched.hip): views into the solver's reordered system.  Node i
// of the caller's numbering is node iperm[i] of the reordered one; reordered
// node j lies in CG range j / per and aggregate (j / per) / ra.
struct CoarseVariant {
    const int *nb_p = nullptr, *bcol_p = nullptr;  // reordered node_blk and bcol
    const double *val_p = nullptr;                 // the values, node rows reordered
    const double *minv_p = nullptr;                // their 3x3 block-Jacobi inverses, FP64
    const float *d = nullptr;                      // 3 per node: position minus its aggregate centroid
    const float *ainv = nullptr;                   // A_c^-1, (6 nagg)^2 row-major
    int nagg = 0, per = 1, ra = 1;
    const int *iperm = nullptr;                    // host: caller's node -> reordered node
};
int coarse_for_values(ofea_ctx *ctx, const double *vals, ofea::CoarseVariant &v, bool keep_inverse = false);

template <class T>
int dev_upload(T *&d, const T *h, size_t count)
{
    if (d) { hipFree(d); d = nullptr; }
    if (count == 0) return OFEA_OK;
    OFEA_HIP(hipMalloc((void **)&d, count * sizeof(T)));
    if (h) OFEA_HIP(hipMemcpy(d, h, count * sizeof(T), hipMemcpyHostToDevice));
    return OFEA_OK;
}
template <class T>
void dev_free(T *&d)
{
    if (d) hipFree(d);
    d = nullptr;
}
}  // namespace ofea
