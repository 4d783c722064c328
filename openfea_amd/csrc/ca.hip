// ca.hip -- combined-approximation (CA) reanalysis on the resident stiffness K0.
//
// Replaces the reanalysis of CAE_process::CA_ReAnalysis (src/cae.cpp:317-376):
//   ca_get_delt_stiffness / ca_fill_CSR_sparse_mat (src/ca_analysis.cpp:18-143)
//     -> ca_delta: Delta K = sum over the modified elements of (+/-) K_e on the
//        pattern of K0, by the column-owner gather kernels of assemble.hip over
//        a plan that holds only the blocks the modified elements touch;
//   ca_build_rom (src/ca_analysis.cpp:146-171)
//     -> the binomial-series basis r_0 = u0, r_i = -K0^-1 (Delta K r_{i-1}),
//        each K0 solve by the device PCG on the resident K0 (its
//        preconditioner is set up once and reused);
//   ca_SVD (src/ca_analysis.cpp:174-197)
//     -> classical Gram-Schmidt with one re-orthogonalisation in FP64 (the
//        reduced solution depends only on the span of the basis; vectors
//        that are numerically dependent on the earlier ones are dropped);
//   ca_solve / ca_reduced_K / ca_reduced_F (src/ca_analysis.cpp:200-309)
//     -> K_r = U^T (K0 + Delta K) U and F_r = U^T F by SpMVs and fixed-order
//        dot products on the device, the m x m system by LU with partial
//        pivoting on the host, u = U y on the device.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ofea_internal.h"

namespace ofea {
int plan_caps(int type);

namespace {

constexpr int CA_WG = 256;
constexpr int CA_NB = 240;    // workgroups of the dot-product kernel
constexpr int CA_MAXB = 32;   // basis vectors

struct CaWs {
    int n = 0;
    long long nnz = 0;
    double *dK = nullptr;      // nnz, Delta K on the pattern of K0
    bool have_dK = false;
    double *basis = nullptr;   // CA_MAXB x n
    double *w = nullptr, *t = nullptr, *F = nullptr;  // n
    double *part = nullptr;    // CA_MAXB x CA_NB dot partials
    double *dots = nullptr;    // CA_MAXB
    double *coef = nullptr;    // CA_MAXB
    double *xyz_ext = nullptr; // 2 nd x 3: modified coordinates, then K0's
};

// part[v][block] = sum over the block's slice of V[v][i] * w[i]; fixed order
__global__ __launch_bounds__(CA_WG) void k_mdot(int n, int nv, const double *__restrict__ V,
                                                const double *__restrict__ w, double *__restrict__ part)
{
    __shared__ double s[CA_WG];
    const int per = (n + gridDim.x - 1) / gridDim.x;
    const int i0 = blockIdx.x * per, i1 = min(n, i0 + per);
    for (int v = 0; v < nv; ++v) {
        const double *x = V + (size_t)v * n;
        double acc = 0.0;
        for (int i = i0 + (int)threadIdx.x; i < i1; i += CA_WG) acc += x[i] * w[i];
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int h = CA_WG / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(size_t)v * gridDim.x + blockIdx.x] = s[0];
        __syncthreads();
    }
}

// one wave per vector: lane l sums partials l, l + 64, ... in order, then a
// fixed xor-shuffle tree; every total has one summation order
__global__ __launch_bounds__(64) void k_mdot_final(int nb, const double *__restrict__ part, double *__restrict__ out)
{
    const int v = blockIdx.x, l = threadIdx.x;
    double acc = 0.0;
    for (int b = l; b < nb; b += 64) acc += part[(size_t)v * nb + b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (l == 0) out[v] = acc;
}

// y = beta * y + sum_v coef[v] V[v]
__global__ __launch_bounds__(CA_WG) void k_combine(int n, int nv, const double *__restrict__ V,
                                                   const double *__restrict__ coef, double beta,
                                                   double *__restrict__ y)
{
    const int i = blockIdx.x * CA_WG + threadIdx.x;
    if (i >= n) return;
    double acc = beta == 0.0 ? 0.0 : beta * y[i];
    for (int v = 0; v < nv; ++v) acc += coef[v] * V[(size_t)v * n + i];
    y[i] = acc;
}

// out = a * x + b * y
__global__ __launch_bounds__(CA_WG) void k_lin2(int n, double a, const double *__restrict__ x, double b,
                                                const double *__restrict__ y, double *__restrict__ out)
{
    const int i = blockIdx.x * CA_WG + threadIdx.x;
    if (i < n) out[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}

void free_ws(CaWs *w)
{
    if (!w) return;
    for (double **p : {&w->dK, &w->basis, &w->w, &w->t, &w->F, &w->part, &w->dots, &w->coef, &w->xyz_ext})
        dev_free(*p);
    delete w;
}

int alloc_ws(CaWs *w)
{
    const size_t n = std::max(w->n, 1);
    OFEA_HIP(hipMalloc((void **)&w->dK, sizeof(double) * std::max(w->nnz, 1LL)));
    OFEA_HIP(hipMalloc((void **)&w->basis, sizeof(double) * n * CA_MAXB));
    OFEA_HIP(hipMalloc((void **)&w->w, sizeof(double) * n));
    OFEA_HIP(hipMalloc((void **)&w->t, sizeof(double) * n));
    OFEA_HIP(hipMalloc((void **)&w->F, sizeof(double) * n));
    OFEA_HIP(hipMalloc((void **)&w->part, sizeof(double) * CA_MAXB * CA_NB));
    OFEA_HIP(hipMalloc((void **)&w->dots, sizeof(double) * CA_MAXB));
    OFEA_HIP(hipMalloc((void **)&w->coef, sizeof(double) * CA_MAXB));
    return OFEA_OK;
}

int ensure(ofea_ctx *c, CaWs *&w)
{
    w = (CaWs *)c->ca_ws;
    if (w && w->n == c->P.n && w->nnz == c->P.nnz) return OFEA_OK;
    free_ws(w);
    c->ca_ws = nullptr;
    w = new CaWs();
    w->n = c->P.n;
    w->nnz = c->P.nnz;
    int rc = alloc_ws(w);
    if (rc) {
        free_ws(w);
        w = nullptr;
        return rc;
    }
    c->ca_ws = w;
    return OFEA_OK;
}

// dots[v] = V[v] . x for v < nv, copied to the host
int mdot(ofea_ctx *c, CaWs *w, const double *V, int nv, const double *x, double *host)
{
    if (nv == 0) return OFEA_OK;
    hipLaunchKernelGGL(k_mdot, dim3(CA_NB), dim3(CA_WG), 0, c->stream, w->n, nv, V, x, w->part);
    hipLaunchKernelGGL(k_mdot_final, dim3(nv), dim3(64), 0, c->stream, CA_NB, w->part, w->dots);
    OFEA_HIP(hipGetLastError());
    OFEA_HIP(hipMemcpyAsync(host, w->dots, sizeof(double) * nv, hipMemcpyDeviceToHost, c->stream));
    OFEA_HIP(hipStreamSynchronize(c->stream));
    return OFEA_OK;
}

// y = beta y + sum coef[v] V[v] (coef on the host)
int combine(ofea_ctx *c, CaWs *w, const double *V, int nv, const double *coef, double beta, double *y)
{
    if (nv > 0) OFEA_HIP(hipMemcpyAsync(w->coef, coef, sizeof(double) * nv, hipMemcpyHostToDevice, c->stream));
    const int g = (w->n + CA_WG - 1) / CA_WG;
    hipLaunchKernelGGL(k_combine, dim3(g), dim3(CA_WG), 0, c->stream, w->n, nv, V, w->coef, beta, y);
    OFEA_HIP(hipGetLastError());
    return OFEA_OK;
}

int lin2(ofea_ctx *c, CaWs *w, double a, const double *x, double b, const double *y, double *out)
{
    const int g = (w->n + CA_WG - 1) / CA_WG;
    hipLaunchKernelGGL(k_lin2, dim3(g), dim3(CA_WG), 0, c->stream, w->n, a, x, b, y, out);
    OFEA_HIP(hipGetLastError());
    return OFEA_OK;
}

// LU with partial pivoting of the m x m row-major A; solves A y = b in place
bool dense_solve(int m, std::vector<double> A, std::vector<double> &b)
{
    for (int k = 0; k < m; ++k) {
        int p = k;
        for (int i = k + 1; i < m; ++i)
            if (std::fabs(A[i * m + k]) > std::fabs(A[p * m + k])) p = i;
        if (A[p * m + k] == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < m; ++j) std::swap(A[k * m + j], A[p * m + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < m; ++i) {
            const double l = A[i * m + k] / A[k * m + k];
            for (int j = k; j < m; ++j) A[i * m + j] -= l * A[k * m + j];
            b[i] -= l * b[k];
        }
    }
    for (int k = m - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < m; ++j) s -= A[k * m + j] * b[j];
        b[k] = s / A[k * m + k];
    }
    return true;
}

}  // namespace

void ca_free(ofea_ctx *c)
{
    free_ws((CaWs *)c->ca_ws);
    c->ca_ws = nullptr;
}

int ca_delta(ofea_ctx *c, double E, double nu, int nmod, const int *kind, const int *conn_mod, const int *etype_mod,
             const double *xyz_m, double *dK_val)
{
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    if (nmod < 0 || (nmod > 0 && (!kind || !conn_mod || !etype_mod))) return fail(OFEA_ERR_ARG, "bad modification list");
    if (!(nu > -1.0 && nu < 0.5)) return fail(OFEA_ERR_ARG, "Poisson ratio must lie in (-1, 0.5)");
    CaWs *w = nullptr;
    int rc = ensure(c, w);
    if (rc) return rc;
    w->have_dK = false;
    const int nd = c->nd;
    // Extended node set: node i carries the modified coordinates, node nd + i
    // the coordinates of K0 (the old side of a moved element); both map to the
    // free rank of node i.
    std::vector<double> xyz_ext(6 * (size_t)nd);
    for (size_t i = 0; i < 3 * (size_t)nd; ++i) {
        xyz_ext[i] = xyz_m ? xyz_m[i] : c->xyz[i];
        xyz_ext[3 * (size_t)nd + i] = c->xyz[i];
    }
    std::vector<int> resort_ext(2 * (size_t)nd);
    for (int i = 0; i < nd; ++i) resort_ext[i] = resort_ext[nd + i] = c->resort[i];
    std::vector<int> conn_ext, etype_ext;
    std::vector<double> E_ext;
    auto add = [&](const int *cn, int type, int offset, double sign) {
        const int nn = type == OFEA_C3D8 ? 8 : 4;
        for (int a = 0; a < 8; ++a) conn_ext.push_back(a < nn ? cn[a] + 1 + offset : 0);
        etype_ext.push_back(type);
        E_ext.push_back(sign * E);
    };
    for (int i = 0; i < nmod; ++i) {
        const int type = etype_mod[i];
        if (type != OFEA_C3D4 && type != OFEA_C3D8) return fail(OFEA_ERR_ARG, "unknown element type id");
        const int nn = type == OFEA_C3D8 ? 8 : 4;
        const int *cn = conn_mod + 8 * (size_t)i;
        for (int a = 0; a < nn; ++a)
            if (cn[a] < 0 || cn[a] >= nd) return fail(OFEA_ERR_ARG, "modified element node out of range");
        switch (kind[i]) {
        case 0: add(cn, type, 0, -1.0); break;  // deleted: -K_e
        case 2: add(cn, type, 0, 1.0); break;   // added: +K_e (its node pairs must be in the pattern of K0)
        case 1:                                 // moved: K_e(new) - K_e(old)
            add(cn, type, 0, 1.0);
            add(cn, type, nd, -1.0);
            break;
        default: return fail(OFEA_ERR_ARG, "modification kind must be 0 (delete), 1 (move) or 2 (add)");
        }
    }
    const int ne_ext = (int)etype_ext.size();
    OFEA_HIP(hipMemsetAsync(w->dK, 0, sizeof(double) * std::max(w->nnz, 1LL), c->stream));
    if (ne_ext > 0) {
        if ((rc = dev_upload(w->xyz_ext, xyz_ext.data(), xyz_ext.size()))) return rc;
        for (int t = 0; t < 2; ++t) {
            if (std::find(etype_ext.begin(), etype_ext.end(), t) == etype_ext.end()) continue;
            GatherPlan G;
            rc = build_gather_plan(2 * nd, ne_ext, conn_ext.data(), etype_ext.data(), resort_ext.data(), c->P, t,
                                   plan_caps(t), G, true);
            if (rc) return rc;
            DevPlan dp;
            rc = upload_plan(G, dp);
            std::vector<double> loc(G.ne);
            for (int i = 0; i < G.ne; ++i) loc[i] = E_ext[G.elem_global[i]];
            if (rc == OFEA_OK) rc = dev_upload(dp.E, loc.data(), loc.size());
            if (rc == OFEA_OK) rc = launch_assemble_plan(c, t, dp, w->xyz_ext, E, nu, w->dK, 1);
            hipError_t se = hipStreamSynchronize(c->stream);  // the plan is freed below
            free_plan(dp);
            if (rc) return rc;
            if (se != hipSuccess) return fail(OFEA_ERR_HIP, hipGetErrorString(se));
        }
    }
    if (dK_val) {
        OFEA_HIP(hipMemcpyAsync(dK_val, w->dK, sizeof(double) * w->nnz, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
    }
    w->have_dK = true;
    return OFEA_OK;
}

int ca_solve(ofea_ctx *c, int n_basis, const double *u0, const double *F, int method, double rtol, int maxit,
             double *u_out, int *n_used, int *pcg_iters, double *reduced)
{
    CaWs *w = (CaWs *)c->ca_ws;
    if (!c->assembled) return fail(OFEA_ERR_STATE, "ofea_assemble (K0) first");
    if (!w || !w->have_dK || w->n != c->P.n) return fail(OFEA_ERR_STATE, "ofea_ca_delta first");
    if (n_basis < 1 || n_basis > CA_MAXB) return fail(OFEA_ERR_ARG, "n_basis must lie in [1, 32]");
    if (!u0 || !F || !u_out) return fail(OFEA_ERR_ARG, "null vector");
    const int n = w->n;
    if (n == 0) return OFEA_OK;
    auto U = [&](int v) { return w->basis + (size_t)v * n; };
    OFEA_HIP(hipMemcpyAsync(U(0), u0, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    OFEA_HIP(hipMemcpyAsync(w->F, F, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    int rc, total_it = 0, status = OFEA_OK;
    double nrm2;
    // series r_i = -K0^-1 (Delta K r_{i-1})  (ca_build_rom, src/ca_analysis.cpp:146-171)
    for (int i = 1; i < n_basis; ++i) {
        if ((rc = solver_spmv_vals(c, w->dK, U(i - 1), w->t))) return rc;
        if ((rc = mdot(c, w, w->t, 1, w->t, &nrm2))) return rc;
        if (nrm2 == 0.0) {  // Delta K r = 0: the series ends here
            if ((rc = lin2(c, w, 0.0, w->t, 0.0, w->t, U(i)))) return rc;
            continue;
        }
        if ((rc = lin2(c, w, -1.0, w->t, 0.0, w->t, c->d_rhs))) return rc;
        c->rhs_set = true;
        int it = 0;
        double rel = 0.0;
        rc = solver_solve(c, method, rtol, maxit, &it, &rel);
        if (rc < 0) return rc;
        if (rc == OFEA_NOT_CONVERGED) status = OFEA_NOT_CONVERGED;
        total_it += it;
        OFEA_HIP(hipMemcpyAsync(U(i), c->d_x, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    // orthonormal basis of the span (classical Gram-Schmidt, twice)
    int m = 0;
    std::vector<double> d(CA_MAXB);
    for (int j = 0; j < n_basis; ++j) {
        if (j != m)
            OFEA_HIP(hipMemcpyAsync(U(m), U(j), sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        double n0;
        if ((rc = mdot(c, w, U(m), 1, U(m), &n0))) return rc;
        if (!(n0 > 0.0)) continue;
        for (int pass = 0; pass < 2 && m > 0; ++pass) {
            if ((rc = mdot(c, w, U(0), m, U(m), d.data()))) return rc;
            for (int k = 0; k < m; ++k) d[k] = -d[k];
            if ((rc = combine(c, w, U(0), m, d.data(), 1.0, U(m)))) return rc;
        }
        double n1;
        if ((rc = mdot(c, w, U(m), 1, U(m), &n1))) return rc;
        if (!(n1 > 1e-20 * n0)) continue;  // numerically inside the span of the earlier vectors
        if ((rc = lin2(c, w, 1.0 / std::sqrt(n1), U(m), 0.0, U(m), U(m)))) return rc;
        ++m;
    }
    if (m == 0) return fail(OFEA_ERR_ARG, "the basis is empty (u0 = 0)");
    // K_r = U^T (K0 + Delta K) U, F_r = U^T F  (ca_reduced_K / ca_reduced_F)
    std::vector<double> Kr((size_t)m * m), Fr(m);
    for (int j = 0; j < m; ++j) {
        if ((rc = solver_spmv_vals(c, c->d_val, U(j), w->w))) return rc;
        if ((rc = solver_spmv_vals(c, w->dK, U(j), w->t))) return rc;
        if ((rc = lin2(c, w, 1.0, w->w, 1.0, w->t, w->w))) return rc;
        if ((rc = mdot(c, w, U(0), m, w->w, d.data()))) return rc;
        for (int i = 0; i < m; ++i) Kr[(size_t)i * m + j] = d[i];
    }
    if ((rc = mdot(c, w, U(0), m, w->F, Fr.data()))) return rc;
    std::vector<double> y = Fr;
    if (!dense_solve(m, Kr, y)) return fail(OFEA_ERR_ARG, "the reduced stiffness is singular");
    if ((rc = combine(c, w, U(0), m, y.data(), 0.0, w->w))) return rc;
    OFEA_HIP(hipMemcpyAsync(u_out, w->w, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    OFEA_HIP(hipStreamSynchronize(c->stream));
    if (n_used) *n_used = m;
    if (pcg_iters) *pcg_iters = total_it;
    if (reduced) {
        std::copy(Kr.begin(), Kr.end(), reduced);
        std::copy(Fr.begin(), Fr.end(), reduced + (size_t)m * m);
        std::copy(y.begin(), y.end(), reduced + (size_t)m * m + m);
    }
    return status;
}

}  // namespace ofea
