// capi.cpp -- C ABI entry points of libofea_hip.so that own device state.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ofea_internal.h"

using namespace ofea;

namespace ofea {
int plan_caps(int type);
}

namespace ofea {
void free_plan(DevPlan &d)
{
    dev_free(d.conn);
    dev_free(d.E);
    dev_free(d.chunk_pair);
    dev_free(d.chunk_blk);
    dev_free(d.pair);
    dev_free(d.blk_rk);
    dev_free(d.contrib_ptr);
    dev_free(d.off_ptr);
    dev_free(d.off_contrib);
    dev_free(d.blk_rl);
    dev_free(d.contrib);
    dev_free(d.node_pair);
    dev_free(d.chunk_node);
    dev_free(d.blk_gid);
    d = DevPlan();
}

int upload_plan(const GatherPlan &G, DevPlan &d)
{
    int rc;
    d.nn = G.nn;
    d.ne = G.ne;
    d.nchunks = G.nchunks();
    d.nblk = (int)G.blk_rk.size();
    if ((rc = dev_upload(d.conn, G.conn.data(), std::max<size_t>(G.conn.size(), 1)))) return rc;
    if ((rc = dev_upload(d.chunk_pair, G.chunk_pair.data(), G.chunk_pair.size()))) return rc;
    if ((rc = dev_upload(d.chunk_blk, G.chunk_blk.data(), G.chunk_blk.size()))) return rc;
    if ((rc = dev_upload(d.pair, G.pair.data(), std::max<size_t>(G.pair.size(), 1)))) return rc;
    if ((rc = dev_upload(d.blk_rk, G.blk_rk.data(), std::max<size_t>(G.blk_rk.size(), 1)))) return rc;
    if ((rc = dev_upload(d.contrib_ptr, G.contrib_ptr.data(), G.contrib_ptr.size()))) return rc;
    if ((rc = dev_upload(d.contrib, G.contrib.data(), std::max<size_t>(G.contrib.size(), 1)))) return rc;
    if ((rc = dev_upload(d.node_pair, G.node_pair.data(), G.node_pair.size()))) return rc;
    if ((rc = dev_upload(d.chunk_node, G.chunk_node.data(), G.chunk_node.size()))) return rc;
    if ((rc = dev_upload(d.off_ptr, G.off_ptr.data(), G.off_ptr.size()))) return rc;
    if ((rc = dev_upload(d.off_contrib, G.off_contrib.data(), std::max<size_t>(G.off_contrib.size(), 1)))) return rc;
    if ((rc = dev_upload(d.blk_rl, G.blk_rl.data(), std::max<size_t>(G.blk_rl.size(), 1)))) return rc;
    if (!G.blk_gid.empty() && (rc = dev_upload(d.blk_gid, G.blk_gid.data(), G.blk_gid.size()))) return rc;
    return OFEA_OK;
}
}  // namespace ofea

static void free_mesh(ofea_ctx *c)
{
    learned_free(c);
    batch_free(c);
    ca_free(c);
    post_free(c);
    sfem_free(c);
    c->sfem = false;
    solver_free(c);
    explicit_free(c);
    dev_free(c->d_xyz);
    dev_free(c->d_col_ptr);
    dev_free(c->d_node_blk);
    dev_free(c->d_bcol);
    dev_free(c->d_diag_k);
    dev_free(c->d_resort);
    dev_free(c->d_val);
    dev_free(c->d_rhs);
    dev_free(c->d_x);
    dev_free(c->d_u);
    dev_free(c->d_fint);
    for (auto &d : c->dp) free_plan(d);
    dev_free(c->dtc.hdr);
    dev_free(c->dtc.ntab);
    dev_free(c->dtc.eloc);
    dev_free(c->dtc.eid);
    dev_free(c->dtc.task);
    c->dtc = DevTetChunks();
    c->assembled = false;
    c->rhs_set = false;
}

// forget the host side of the mesh (device buffers are freed by free_mesh)
static void reset_mesh(ofea_ctx *c)
{
    c->nd = c->ne = 0;
    c->xyz.clear();
    c->conn.clear();
    c->etype.clear();
    c->resort.clear();
    c->E_elem.clear();
    c->P = Pattern();
    c->TC = TetChunkPlan();
    for (int t = 0; t < 2; ++t) {
        c->G[t] = GatherPlan();
        c->has_type[t] = false;
    }
    c->kver++;
}

static int upload_mesh(ofea_ctx *c)
{
    int rc;
    if ((rc = dev_upload(c->d_xyz, c->xyz.data(), c->xyz.size()))) return rc;
    if ((rc = dev_upload(c->d_col_ptr, c->P.col_ptr.data(), c->P.col_ptr.size()))) return rc;
    if ((rc = dev_upload(c->d_node_blk, c->P.node_blk.data(), c->P.node_blk.size()))) return rc;
    if ((rc = dev_upload(c->d_bcol, c->P.bcol.data(), c->P.bcol.size()))) return rc;
    if ((rc = dev_upload(c->d_diag_k, c->P.diag_k.data(), std::max<size_t>(c->P.diag_k.size(), 1)))) return rc;
    if ((rc = dev_upload(c->d_resort, c->resort.data(), c->resort.size()))) return rc;
    if ((rc = dev_upload<double>(c->d_val, nullptr, (size_t)std::max(c->P.nnz, 1LL)))) return rc;
    if ((rc = dev_upload<double>(c->d_rhs, nullptr, (size_t)std::max(c->P.n, 1)))) return rc;
    if ((rc = dev_upload<double>(c->d_x, nullptr, (size_t)std::max(c->P.n, 1)))) return rc;
    for (int t = 0; t < 2; ++t) {
        if (!c->has_type[t]) continue;
        if ((rc = upload_plan(c->G[t], c->dp[t]))) return rc;
    }
    if (c->TC.nchunks > 0) {
        const TetChunkPlan &T = c->TC;
        if ((rc = dev_upload(c->dtc.hdr, T.hdr.data(), T.hdr.size()))) return rc;
        if ((rc = dev_upload(c->dtc.ntab, T.ntab.data(), T.ntab.size()))) return rc;
        if ((rc = dev_upload(c->dtc.eloc, T.eloc.data(), T.eloc.size()))) return rc;
        if ((rc = dev_upload(c->dtc.eid, T.eid.data(), T.eid.size()))) return rc;
        if ((rc = dev_upload(c->dtc.task, T.task.data(), T.task.size()))) return rc;
        c->dtc.nchunks = T.nchunks;
    }
    return OFEA_OK;
}

#define CHECK_CTX(c)                                                  \
    do {                                                              \
        if (!(c)) return fail(OFEA_ERR_ARG, "null context");         \
        hipError_t _se = hipSetDevice((c)->device);                   \
        if (_se != hipSuccess)                                        \
            return fail(OFEA_ERR_HIP, hipGetErrorString(_se));        \
    } while (0)

extern "C" {

int ofea_create(ofea_ctx **ctx, int device)
{
    if (!ctx) return fail(OFEA_ERR_ARG, "null ctx pointer");
    *ctx = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(OFEA_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(OFEA_ERR_ARG, "device index out of range");
    OFEA_HIP(hipSetDevice(device));
    int rc = noalloc_init(device);  // the device's static scratch (ofea_solve_noalloc)
    if (rc) return rc;
    auto *c = new ofea_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return fail(OFEA_ERR_HIP, "stream/event creation failed");
    }
    *ctx = c;
    return OFEA_OK;
}

int ofea_destroy(ofea_ctx *c)
{
    if (!c) return OFEA_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    free_mesh(c);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipStreamDestroy(c->stream);
    delete c;
    return OFEA_OK;
}

int ofea_set_mesh(ofea_ctx *c, int nd, const double *xyz, int ne, const int *conn, const int *etype, const int *resort)
{
    CHECK_CTX(c);
    if (nd <= 0 || ne <= 0 || !xyz || !conn || !etype || !resort) return fail(OFEA_ERR_ARG, "bad mesh arguments");
    for (int e = 0; e < ne; ++e)
        if (etype[e] != OFEA_C3D4 && etype[e] != OFEA_C3D8) return fail(OFEA_ERR_ARG, "unknown element type id");
    for (int i = 0; i < nd; ++i)
        if (resort[i] >= nd) return fail(OFEA_ERR_ARG, "resort entry out of range");
    hipStreamSynchronize(c->stream);
    free_mesh(c);
    reset_mesh(c);
    // The host structures are built into locals and the context takes them
    // only when all of them (and the uploads) succeeded: after a failure the
    // context holds no mesh and every later call reports OFEA_ERR_STATE.
    Pattern P;
    GatherPlan G[2];
    TetChunkPlan TC;
    bool has_type[2] = {false, false};
    int rc = build_pattern(nd, ne, conn, etype, resort, P);
    if (rc) return rc;
    for (int t = 0; t < 2; ++t) {
        has_type[t] = std::find(etype, etype + ne, t) != etype + ne;
        if (!has_type[t]) continue;
        rc = build_gather_plan(nd, ne, conn, etype, resort, P, t, plan_caps(t), G[t]);
        if (rc) return rc;
    }
    // chunk plan of the C3D4 assembly; a mesh beyond its capacity keeps the
    // pair-gather kernel (the limit is not an error of set_mesh)
    if (has_type[OFEA_C3D4] && build_tet_chunks(nd, P, G[OFEA_C3D4], resort, TC) != OFEA_OK) TC = TetChunkPlan();
    c->xyz.assign(xyz, xyz + 3 * (size_t)nd);
    c->conn.assign(conn, conn + 8 * (size_t)ne);
    c->etype.assign(etype, etype + ne);
    c->resort.assign(resort, resort + nd);
    c->P = std::move(P);
    c->TC = std::move(TC);
    for (int t = 0; t < 2; ++t) {
        c->G[t] = std::move(G[t]);
        c->has_type[t] = has_type[t];
    }
    rc = upload_mesh(c);
    if (rc) {
        free_mesh(c);
        reset_mesh(c);
        return rc;
    }
    c->nd = nd;
    c->ne = ne;
    return OFEA_OK;
}

int ofea_pattern(ofea_ctx *c, int *n, long long *nnz, int *col_ptr, int *row_idx)
{
    if (!c) return fail(OFEA_ERR_ARG, "null context");
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    if (n) *n = c->P.n;
    if (nnz) *nnz = c->P.nnz;
    if (col_ptr) std::memcpy(col_ptr, c->P.col_ptr.data(), c->P.col_ptr.size() * sizeof(int));
    if (row_idx) std::memcpy(row_idx, c->P.row_idx.data(), c->P.row_idx.size() * sizeof(int));
    return OFEA_OK;
}

int ofea_set_elem_E(ofea_ctx *c, const double *E_elem)
{
    CHECK_CTX(c);
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    hipStreamSynchronize(c->stream);
    for (int t = 0; t < 2; ++t) dev_free(c->dp[t].E);
    c->E_elem.clear();
    if (!E_elem) return OFEA_OK;
    c->E_elem.assign(E_elem, E_elem + c->ne);
    for (int t = 0; t < 2; ++t) {
        if (!c->has_type[t]) continue;
        std::vector<double> loc(c->G[t].ne);
        for (int i = 0; i < c->G[t].ne; ++i) loc[i] = E_elem[c->G[t].elem_global[i]];
        int rc = dev_upload(c->dp[t].E, loc.data(), loc.size());
        if (rc) return rc;
    }
    return OFEA_OK;
}

#define NOT_SFEM(c)                                                                                       \
    do {                                                                                                  \
        if ((c)->sfem) return fail(OFEA_ERR_STATE, "the context holds the smoothed (SFEM) pattern: ofea_set_mesh again"); \
    } while (0)

int ofea_assemble(ofea_ctx *c, double E, double nu, double *nz_val)
{
    CHECK_CTX(c);
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    NOT_SFEM(c);
    if (!(nu > -1.0 && nu < 0.5)) return fail(OFEA_ERR_ARG, "Poisson ratio must lie in (-1, 0.5)");
    int rc = launch_assemble(c, E, nu, c->d_val, false);
    if (rc) return rc;
    c->assembled = true;
    c->kver++;
    c->asm_E = E;
    c->asm_nu = nu;
    solver_invalidate(c);
    if (nz_val) {
        OFEA_HIP(hipMemcpyAsync(nz_val, c->d_val, sizeof(double) * c->P.nnz, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
    }
    return OFEA_OK;
}

int ofea_spmv(ofea_ctx *c, const double *x, double *y)
{
    CHECK_CTX(c);
    if (!c->assembled) return fail(OFEA_ERR_STATE, "ofea_assemble first");
    if (!x || !y) return fail(OFEA_ERR_ARG, "null vector");
    // staged through two n-vectors of the solver work space (allocated once
    // per mesh, reused by every call)
    double *dx = nullptr, *dy = nullptr;
    int rc = solver_staging(c, &dx, &dy);
    if (rc) return rc;
    const size_t n = c->P.n;
    OFEA_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if ((rc = solver_spmv(c, dx, dy))) return rc;
    OFEA_HIP(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    OFEA_HIP(hipStreamSynchronize(c->stream));
    return OFEA_OK;
}

int ofea_set_rhs(ofea_ctx *c, const double *b)
{
    CHECK_CTX(c);
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    if (!b) return fail(OFEA_ERR_ARG, "null rhs");
    OFEA_HIP(hipMemcpyAsync(c->d_rhs, b, sizeof(double) * c->P.n, hipMemcpyHostToDevice, c->stream));
    OFEA_HIP(hipStreamSynchronize(c->stream));
    c->rhs_set = true;
    return OFEA_OK;
}

int ofea_solve(ofea_ctx *c, int method, const double *b, double *x, double rtol, int maxit, int *iters, double *relres)
{
    CHECK_CTX(c);
    if (!c->assembled) return fail(OFEA_ERR_STATE, "ofea_assemble first");
    if (!(rtol > 0.0) || maxit <= 0) return fail(OFEA_ERR_ARG, "rtol must be > 0 and maxit > 0");
    if (b) {
        int rc = ofea_set_rhs(c, b);
        if (rc) return rc;
    }
    if (!c->rhs_set) return fail(OFEA_ERR_STATE, "no right-hand side (ofea_set_rhs)");
    int rc = solver_solve(c, method, rtol, maxit, iters, relres);
    if (rc < 0) return rc;
    if (x) {
        OFEA_HIP(hipMemcpyAsync(x, c->d_x, sizeof(double) * c->P.n, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
    }
    return rc;
}

int ofea_solve_noalloc(ofea_ctx *c, int method, const double *d_b, double *d_x, double rtol, int maxit, int *iters,
                       double *relres)
{
    CHECK_CTX(c);
    if (!d_b || !d_x) return fail(OFEA_ERR_ARG, "null device vector");
    if (!(rtol > 0.0) || maxit <= 0) return fail(OFEA_ERR_ARG, "rtol must be > 0 and maxit > 0");
    return solve_noalloc(c, method, d_b, d_x, rtol, maxit, iters, relres);
}

int ofea_get_solution(ofea_ctx *c, double *x)
{
    CHECK_CTX(c);
    if (!x) return fail(OFEA_ERR_ARG, "null x");
    OFEA_HIP(hipMemcpyAsync(x, c->d_x, sizeof(double) * c->P.n, hipMemcpyDeviceToHost, c->stream));
    OFEA_HIP(hipStreamSynchronize(c->stream));
    return OFEA_OK;
}

int ofea_lumped_mass(ofea_ctx *c, double rho, double *mass)
{
    CHECK_CTX(c);
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    if (!mass) return fail(OFEA_ERR_ARG, "null mass");
    return lumped_mass(c, rho, mass);
}

int ofea_explicit_run(ofea_ctx *c, double E, double nu, const double *mass, const double *f_ext, const double *v0,
                      const int *fixed_mask, double dt, int nsteps, int precision, double *u_out)
{
    CHECK_CTX(c);
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    if (!mass || !f_ext || !fixed_mask || nsteps < 0 || !(dt > 0.0)) return fail(OFEA_ERR_ARG, "bad explicit arguments");
    if (precision != OFEA_FP32 && precision != OFEA_FP64) return fail(OFEA_ERR_ARG, "precision must be FP32 or FP64");
    return explicit_run(c, E, nu, mass, f_ext, v0, fixed_mask, dt, nsteps, precision, u_out);
}

int ofea_explicit_analysis(ofea_ctx *c, double E, double nu, const double *mass, const double *f_ext,
                           const double *v0, const int *fixed_mask, double time_total, double time_step,
                           int output_gap, int precision, int max_steps, int max_frames, double *frames,
                           int *frame_index, double *frame_time, int *n_frames, int *n_steps)
{
    CHECK_CTX(c);
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    if (!mass || !f_ext || !fixed_mask) return fail(OFEA_ERR_ARG, "null explicit input");
    if (!std::isfinite(time_total) || !std::isfinite(time_step) || time_step < 0.0)
        return fail(OFEA_ERR_ARG, "time_total must be finite and time_step finite and >= 0");
    if (output_gap < 1 || max_steps < 1 || max_frames < 1)
        return fail(OFEA_ERR_ARG, "output_gap, max_steps and max_frames must be >= 1");
    if (precision != OFEA_FP32 && precision != OFEA_FP64) return fail(OFEA_ERR_ARG, "precision must be FP32 or FP64");
    if ((double)max_frames * 3.0 * c->nd > 4.0e9) return fail(OFEA_ERR_LIMIT, "max_frames x 3 nd too large");
    return explicit_analysis(c, E, nu, mass, f_ext, v0, fixed_mask, time_total, time_step, output_gap, precision,
                             max_steps, max_frames, frames, frame_index, frame_time, n_frames, n_steps);
}

int ofea_nl_assemble(ofea_ctx *c, double E, double nu, const double *u, double *nz_val, double *f_int)
{
    CHECK_CTX(c);
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    NOT_SFEM(c);
    if (!u) return fail(OFEA_ERR_ARG, "null displacement");
    return nl_assemble(c, E, nu, u, nz_val, f_int);
}

int ofea_nl_solve(ofea_ctx *c, double E, double nu, const double *F, int n_incr, double tol, int max_newton,
                  int method, double inner_rtol, double *u_out, int *newton_iters, int *linear_iters, double *relres)
{
    CHECK_CTX(c);
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    NOT_SFEM(c);
    if (!F || n_incr < 1 || !(tol > 0.0) || max_newton < 1 || !(inner_rtol != 0.0) || !std::isfinite(inner_rtol))
        return fail(OFEA_ERR_ARG, "bad Newton arguments");
    return nl_solve(c, E, nu, F, n_incr, tol, max_newton, method, inner_rtol, u_out, newton_iters, linear_iters,
                    relres);
}

int ofea_nl_solve_ref(ofea_ctx *c, double E, double nu, const double *F, double res_limit, int max_loop, int method,
                      double inner_rtol, double *u_out, int *loops, int *linear_iters, double *res_norm)
{
    CHECK_CTX(c);
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    NOT_SFEM(c);
    if (!F || !(res_limit >= 0.0) || max_loop < 0 || !(inner_rtol > 0.0))
        return fail(OFEA_ERR_ARG, "bad Newton arguments");
    return nl_solve_ref(c, E, nu, F, res_limit, max_loop, method, inner_rtol, u_out, loops, linear_iters, res_norm);
}

int ofea_ca_delta(ofea_ctx *c, double E, double nu, int nmod, const int *kind, const int *conn_mod,
                  const int *etype_mod, const double *xyz_m, double *dK_val)
{
    CHECK_CTX(c);
    NOT_SFEM(c);
    return ca_delta(c, E, nu, nmod, kind, conn_mod, etype_mod, xyz_m, dK_val);
}

int ofea_ca_solve(ofea_ctx *c, int n_basis, const double *u0, const double *F, int method, double rtol, int maxit,
                  double *u_out, int *n_used, int *pcg_iters, double *reduced)
{
    CHECK_CTX(c);
    if (!(rtol > 0.0) || maxit <= 0) return fail(OFEA_ERR_ARG, "rtol must be > 0 and maxit > 0");
    return ca_solve(c, n_basis, u0, F, method, rtol, maxit, u_out, n_used, pcg_iters, reduced);
}

int ofea_sfem_setup(ofea_ctx *c)
{
    CHECK_CTX(c);
    return sfem_setup(c);
}

int ofea_sfem_assemble(ofea_ctx *c, double E, double nu, double *nz_val)
{
    CHECK_CTX(c);
    return sfem_assemble(c, E, nu, nz_val);
}

int ofea_post_nodal(ofea_ctx *c, double E, double nu, const double *u_full, int quantity, double *elem_out,
                    double *node_out)
{
    CHECK_CTX(c);
    return post_nodal(c, E, nu, u_full, quantity, elem_out, node_out);
}

int ofea_learned_nparam(void) { return learned_nparam(); }

int ofea_learned_init(ofea_ctx *c, unsigned long long seed)
{
    CHECK_CTX(c);
    return learned_init(c, seed);
}

int ofea_learned_set_params(ofea_ctx *c, const float *params, int nparam)
{
    CHECK_CTX(c);
    if (!params) return fail(OFEA_ERR_ARG, "null params");
    return learned_set_params(c, params, nparam);
}

int ofea_learned_get_params(ofea_ctx *c, float *params, int nparam)
{
    CHECK_CTX(c);
    if (!params) return fail(OFEA_ERR_ARG, "null params");
    return learned_get_params(c, params, nparam);
}

int ofea_learned_loss_grad(ofea_ctx *c, double *loss, float *grad)
{
    CHECK_CTX(c);
    return learned_loss_grad(c, loss, grad);
}

int ofea_learned_train(ofea_ctx *c, int steps, float lr, double *loss_hist)
{
    CHECK_CTX(c);
    if (steps < 0 || !(lr > 0.0f)) return fail(OFEA_ERR_ARG, "steps >= 0 and lr > 0 required");
    return learned_train(&c, 1, steps, lr, loss_hist);
}

int ofea_learned_train_multi(ofea_ctx *const *ctxs, int nctx, int steps, float lr, double *loss_hist)
{
    if (!ctxs || nctx < 1) return fail(OFEA_ERR_ARG, "need at least one context");
    for (int k = 0; k < nctx; ++k) {
        if (!ctxs[k]) return fail(OFEA_ERR_ARG, "null context");
        if (ctxs[k]->device != ctxs[0]->device) return fail(OFEA_ERR_ARG, "contexts on different devices");
    }
    CHECK_CTX(ctxs[0]);
    if (steps < 0 || !(lr > 0.0f)) return fail(OFEA_ERR_ARG, "steps >= 0 and lr > 0 required");
    return learned_train(ctxs, nctx, steps, lr, loss_hist);
}

int ofea_learned_blocks(ofea_ctx *c, double *blocks, float *features)
{
    CHECK_CTX(c);
    return learned_blocks(c, blocks, features);
}

int ofea_learned_outputs(ofea_ctx *c, float *uv)
{
    CHECK_CTX(c);
    if (!uv) return fail(OFEA_ERR_ARG, "null output");
    return learned_outputs(c, uv);
}

int ofea_learned_precision(ofea_ctx *c, int precision)
{
    CHECK_CTX(c);
    return learned_precision(c, precision);
}

int ofea_learned_refresh(ofea_ctx *c)
{
    CHECK_CTX(c);
    return learned_refresh(c);
}

int ofea_timer_start(ofea_ctx *c)
{
    CHECK_CTX(c);
    OFEA_HIP(hipEventRecord(c->ev0, c->stream));
    return OFEA_OK;
}

int ofea_timer_stop(ofea_ctx *c, float *ms)
{
    CHECK_CTX(c);
    OFEA_HIP(hipEventRecord(c->ev1, c->stream));
    OFEA_HIP(hipEventSynchronize(c->ev1));
    float t = 0.f;
    OFEA_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
    if (ms) *ms = t;
    return OFEA_OK;
}

int ofea_stream_copy(ofea_ctx *c, long long bytes, int reps, float *ms)
{
    CHECK_CTX(c);
    if (!ms) return fail(OFEA_ERR_ARG, "null ms");
    return stream_copy(c, bytes, reps, ms);
}

int ofea_assemble_batch(ofea_ctx *c, int nvar, const double *E_var, double nu, double *nz_val)
{
    CHECK_CTX(c);
    if (!c->d_col_ptr) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    NOT_SFEM(c);
    return assemble_batch(c, nvar, E_var, nu, nz_val);
}

int ofea_solve_batch(ofea_ctx *c, int nvar, int nrhs, const double *B, double *X, double rtol, int maxit, int *iters,
                     double *relres)
{
    CHECK_CTX(c);
    if (!B || nvar < 1 || maxit < 1 || !(rtol > 0.0)) return fail(OFEA_ERR_ARG, "bad B, nvar, maxit or rtol");
    return solve_batch(c, nvar, nrhs, B, X, rtol, maxit, OFEA_PCG_BLOCK_JACOBI, iters, relres);
}

int ofea_solve_batch_pc(ofea_ctx *c, int nvar, int nrhs, const double *B, double *X, double rtol, int maxit,
                        int precond, int *iters, double *relres)
{
    CHECK_CTX(c);
    if (!B || nvar < 1 || maxit < 1 || !(rtol > 0.0)) return fail(OFEA_ERR_ARG, "bad B, nvar, maxit or rtol");
    return solve_batch(c, nvar, nrhs, B, X, rtol, maxit, precond, iters, relres);
}

int ofea_autotune(ofea_ctx *c, int *best_group, float *times)
{
    CHECK_CTX(c);
    return autotune(c, best_group, times);
}

int ofea_batch_tiles(ofea_ctx *c, int nrhs, int *group, float *ms)
{
    CHECK_CTX(c);
    return batch_tiles(c, nrhs, group, ms);
}

int ofea_coarse_apply(ofea_ctx *c, const double *r, double *out)
{
    CHECK_CTX(c);
    if (!c->assembled) return fail(OFEA_ERR_STATE, "ofea_assemble first");
    if (!r || !out) return fail(OFEA_ERR_ARG, "null r or out");
    return coarse_apply(c, r, out);
}

int ofea_coarse_aggregates(ofea_ctx *c, int *agg)
{
    CHECK_CTX(c);
    if (!c->assembled) return fail(OFEA_ERR_STATE, "ofea_assemble first");
    if (!agg) return fail(OFEA_ERR_ARG, "null agg");
    return coarse_aggregates(c, agg);
}

int ofea_coarse_builds(ofea_ctx *c)
{
    CHECK_CTX(c);
    return coarse_builds(c);
}

int ofea_sync(ofea_ctx *c)
{
    CHECK_CTX(c);
    OFEA_HIP(hipStreamSynchronize(c->stream));
    return OFEA_OK;
}

}  // extern "C"
