// explicit.hip -- lumped mass and the half-step central-difference loop.
//
// Reference: CAE_process::explicit_analysis (src/cae.cpp:420-574) with
// build_ele_mass (src/Elements/C3D8.cpp:58-108, src/Elements/C3D4.cpp:75-100)
// and cal_in_force (src/Elements/C3D8.cpp:158-193).  The reference accumulates
// InForce -= Ke(X0) du every step (never zeroed, src/explicit_tools.cpp:41), so
// InForce_k = -K0 u_k; here f_int = -K0 u_k is evaluated directly, matrix-free,
// which is the same quantity without the running-sum round-off.
//
// One step = two kernels:
//   element phase  lane per (hex, Gauss point) [or thread per tet]: grad u,
//                  sigma = lambda tr(eps) I + 2 mu eps, nodal forces
//                  w detJ sigma g_a; an 8-lane reduce-scatter leaves each lane
//                  the force of one element node -> fe[a][e] (local-node-major,
//                  so that the node phase's gathers of consecutive nodes read
//                  consecutive elements; no atomics);
//   node phase     thread per node: f = -sum fe over incident elements in
//                  ascending order, a = (f + f_ext)/m, v += a dt,
//                  u_next = u + v dt (0 at clamped nodes).
// Runs in FP32 (production) or FP64 (parity) with the same kernels.
#include <algorithm>
#include <vector>

#include "elem_device.h"
#include "ofea_internal.h"

namespace ofea {

constexpr int EX_WG = 256;
constexpr int EX_GRAPH_STEPS = 50;  // steps per captured graph (even)

struct ExplicitWS {
    int nd = 0;
    int *inc_ptr = nullptr, *inc = nullptr;  // node -> (elem_local << 4 | type << 3 | a), ascending elements
    int *conn[2] = {nullptr, nullptr};       // type-local connectivity (0-based), copied from the plans
    int ne[2] = {0, 0};
    void *xyz_t = nullptr;                   // coordinates in the run precision
    void *fe[2] = {nullptr, nullptr};        // element nodal forces, [nn][ne][3]
    void *u[2] = {nullptr, nullptr}, *v = nullptr, *minv = nullptr, *fext = nullptr;
    int *fixed = nullptr;
    int prec = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double g_E = 0, g_nu = 0, g_dt = 0;
};

static ExplicitWS *ews(ofea_ctx *c) { return (ExplicitWS *)c->explicit_ws; }

void explicit_free(ofea_ctx *c)
{
    ExplicitWS *w = ews(c);
    if (!w) return;
    if (w->exec) hipGraphExecDestroy(w->exec);
    if (w->graph) hipGraphDestroy(w->graph);
    for (void *p : {(void *)w->inc_ptr, (void *)w->inc, w->xyz_t, w->fe[0], w->fe[1], w->u[0], w->u[1], w->v, w->minv,
                    w->fext, (void *)w->fixed})
        if (p) hipFree(p);
    delete w;
    c->explicit_ws = nullptr;
}

static int ensure_incidence(ofea_ctx *c)
{
    if (ews(c)) return OFEA_OK;
    auto *w = new ExplicitWS();
    c->explicit_ws = w;
    w->nd = c->nd;
    std::vector<int> ptr(c->nd + 1, 0);
    for (int t = 0; t < 2; ++t) {
        if (!c->has_type[t]) continue;
        const GatherPlan &G = c->G[t];
        for (size_t i = 0; i < G.conn.size(); ++i) ptr[G.conn[i] + 1]++;
    }
    for (int i = 0; i < c->nd; ++i) ptr[i + 1] += ptr[i];
    // ascending global element order per node
    std::vector<int> inc(ptr[c->nd]);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    std::vector<int> local(c->ne, -1);
    for (int t = 0; t < 2; ++t)
        if (c->has_type[t])
            for (int i = 0; i < c->G[t].ne; ++i) local[c->G[t].elem_global[i]] = i;
    for (int e = 0; e < c->ne; ++e) {
        int t = c->etype[e], nn = t == OFEA_C3D8 ? 8 : 4;
        for (int a = 0; a < nn; ++a) {
            int node = c->conn[8 * (size_t)e + a] - 1;
            inc[fill[node]++] = (local[e] << 4) | (t << 3) | a;
        }
    }
    if ((long long)c->ne >= (1LL << 27)) return fail(OFEA_ERR_LIMIT, "too many elements for the packed incidence");
    int rc;
    if ((rc = dev_upload(w->inc_ptr, ptr.data(), ptr.size()))) return rc;
    if ((rc = dev_upload(w->inc, inc.data(), std::max<size_t>(inc.size(), 1)))) return rc;
    for (int t = 0; t < 2; ++t) {
        w->conn[t] = c->dp[t].conn;
        w->ne[t] = c->has_type[t] ? c->G[t].ne : 0;
    }
    return OFEA_OK;
}

// ------------------------------------------------------------------ mass
__global__ void k_mass_hex(int ne, const int *__restrict__ conn, const double *__restrict__ xyz, double rho,
                           double *__restrict__ me)
{
    int e = blockIdx.x * EX_WG + threadIdx.x;
    if (e >= ne) return;
    double X[24];
    for (int b = 0; b < 8; ++b)
        for (int c = 0; c < 3; ++c) X[3 * b + c] = xyz[3 * conn[8 * e + b] + c];
    double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const double q = 0.57735026918962576451;
    for (int gp = 0; gp < 8; ++gp) {
        double g[3][8];
        double detJ = hex_point<double>(X, gp, g);
        double r = hsgn(gp, 0) * q, s = hsgn(gp, 1) * q, t = hsgn(gp, 2) * q;
        double N[8], sumN = 0.0;
        for (int a = 0; a < 8; ++a) {
            N[a] = (1 + hsgn(a, 0) * r) * (1 + hsgn(a, 1) * s) * (1 + hsgn(a, 2) * t) / 8.0;
            sumN += N[a];
        }
        // three rows of node a of the consistent mass, summed, divided by 3
        for (int a = 0; a < 8; ++a) m[a] += N[a] * sumN * detJ * rho;
    }
    for (int a = 0; a < 8; ++a) me[8 * e + a] = m[a];
}

__global__ void k_mass_tet(int ne, const int *__restrict__ conn, const double *__restrict__ xyz, double rho,
                           double *__restrict__ me)
{
    int e = blockIdx.x * EX_WG + threadIdx.x;
    if (e >= ne) return;
    double X[12];
    for (int b = 0; b < 4; ++b)
        for (int c = 0; c < 3; ++c) X[3 * b + c] = xyz[3 * conn[4 * e + b] + c];
    double g[3][4];
    double detJ = tet_grads<double>(X, g);
    double m = fabs(detJ / 6.0) * rho / 4.0;
    for (int a = 0; a < 4; ++a) me[4 * e + a] = m;
}

__global__ void k_mass_gather(int nd, const int *__restrict__ inc_ptr, const int *__restrict__ inc,
                              const double *__restrict__ m4, const double *__restrict__ m8, double *__restrict__ mass)
{
    int i = blockIdx.x * EX_WG + threadIdx.x;
    if (i >= nd) return;
    double s = 0.0;
    for (int t = inc_ptr[i]; t < inc_ptr[i + 1]; ++t) {
        int code = inc[t], el = code >> 4, type = (code >> 3) & 1, a = code & 7;
        s += type ? m8[8 * el + a] : m4[4 * el + a];
    }
    mass[i] = s;
}

int lumped_mass(ofea_ctx *c, double rho, double *mass)
{
    int rc = ensure_incidence(c);
    if (rc) return rc;
    ExplicitWS *w = ews(c);
    double *me[2] = {nullptr, nullptr}, *dm = nullptr;
    for (int t = 0; t < 2; ++t)
        OFEA_HIP(hipMalloc((void **)&me[t], sizeof(double) * std::max(1, w->ne[t] * (t ? 8 : 4))));
    OFEA_HIP(hipMalloc((void **)&dm, sizeof(double) * c->nd));
    if (w->ne[0])
        hipLaunchKernelGGL(k_mass_tet, dim3((w->ne[0] + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream, w->ne[0],
                           w->conn[0], c->d_xyz, rho, me[0]);
    if (w->ne[1])
        hipLaunchKernelGGL(k_mass_hex, dim3((w->ne[1] + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream, w->ne[1],
                           w->conn[1], c->d_xyz, rho, me[1]);
    hipLaunchKernelGGL(k_mass_gather, dim3((c->nd + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream, c->nd, w->inc_ptr,
                       w->inc, me[0], me[1], dm);
    hipError_t e1 = hipGetLastError();
    hipMemcpyAsync(mass, dm, sizeof(double) * c->nd, hipMemcpyDeviceToHost, c->stream);
    hipError_t e2 = hipStreamSynchronize(c->stream);
    hipFree(me[0]);
    hipFree(me[1]);
    hipFree(dm);
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(OFEA_ERR_HIP, "lumped mass kernels failed");
    return OFEA_OK;
}

// ------------------------------------------------------------------ forces
template <class T>
__global__ __launch_bounds__(EX_WG) void k_force_hex(int ne, const int *__restrict__ conn, const T *__restrict__ xyz,
                                                     const T *__restrict__ u, T lam, T mu, T *__restrict__ fe)
{
    const int gid = blockIdx.x * EX_WG + threadIdx.x;
    const int e_raw = gid >> 3, gp = gid & 7;
    const int e = min(e_raw, ne - 1);  // whole 8-lane groups stay in the shuffles
    const int *cn = conn + 8 * e;
    T X[24], U[24];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        int node = cn[b];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            X[3 * b + c] = xyz[3 * node + c];
            U[3 * b + c] = u[3 * node + c];
        }
    }
    T g[3][8];
    T detJ = hex_point<T>(X, gp, g);
    // grad u: H[c][j] = sum_b u_b[c] g[j][b]
    T H[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T acc = T(0);
#pragma unroll
            for (int b = 0; b < 8; ++b) acc += U[3 * b + c] * g[j][b];
            H[c][j] = acc;
        }
    T tr = H[0][0] + H[1][1] + H[2][2];
    T S[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < 3; ++j) S[c][j] = detJ * (mu * (H[c][j] + H[j][c]) + (c == j ? lam * tr : T(0)));
    auto blk = [&](int a, T out[3]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c] = S[c][0] * g[0][a] + S[c][1] * g[1][a] + S[c][2] * g[2][a];
    };
    T f[3];
    reduce_scatter8<3, T>(gp, blk, f);
    if (e_raw < ne) {
        T *dst = fe + 3 * ((size_t)gp * ne + e);
        dst[0] = f[0];
        dst[1] = f[1];
        dst[2] = f[2];
    }
}

template <class T>
__global__ __launch_bounds__(EX_WG) void k_force_tet(int ne, const int *__restrict__ conn, const T *__restrict__ xyz,
                                                     const T *__restrict__ u, T lam, T mu, T *__restrict__ fe)
{
    const int e = blockIdx.x * EX_WG + threadIdx.x;
    if (e >= ne) return;
    const int *cn = conn + 4 * e;
    T X[12], U[12];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            X[3 * b + c] = xyz[3 * cn[b] + c];
            U[3 * b + c] = u[3 * cn[b] + c];
        }
    T g[3][4];
    T V = tet_grads<T>(X, g) / T(6);
    T H[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < 3; ++j) H[c][j] = U[c] * g[j][0] + U[3 + c] * g[j][1] + U[6 + c] * g[j][2] + U[9 + c] * g[j][3];
    T tr = H[0][0] + H[1][1] + H[2][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < 3; ++j) acc += (mu * (H[c][j] + H[j][c]) + (c == j ? lam * tr : T(0))) * g[j][a];
            fe[3 * ((size_t)a * ne + e) + c] = V * acc;
        }
}

template <class T>
__global__ __launch_bounds__(EX_WG) void k_node_update(int nd, const int *__restrict__ inc_ptr, const int *__restrict__ inc,
                                                       int ne4, int ne8, const T *__restrict__ fe4, const T *__restrict__ fe8,
                                                       const T *__restrict__ ucur, T *__restrict__ unext,
                                                       T *__restrict__ v, const T *__restrict__ minv,
                                                       const T *__restrict__ fext, const int *__restrict__ fixed, T dt)
{
    const int i = blockIdx.x * EX_WG + threadIdx.x;
    if (i >= nd) return;
    T f0 = T(0), f1 = T(0), f2 = T(0);
    for (int t = inc_ptr[i]; t < inc_ptr[i + 1]; ++t) {
        int code = inc[t], el = code >> 4, type = (code >> 3) & 1, a = code & 7;
        const T *src = type ? fe8 + 3 * ((size_t)a * ne8 + el) : fe4 + 3 * ((size_t)a * ne4 + el);
        f0 -= src[0];
        f1 -= src[1];
        f2 -= src[2];
    }
    const T mi = minv[i];
    T v0 = v[3 * i] + (f0 + fext[3 * i]) * mi * dt;
    T v1 = v[3 * i + 1] + (f1 + fext[3 * i + 1]) * mi * dt;
    T v2 = v[3 * i + 2] + (f2 + fext[3 * i + 2]) * mi * dt;
    v[3 * i] = v0;
    v[3 * i + 1] = v1;
    v[3 * i + 2] = v2;
    const bool fx = fixed[i] != 0;
    unext[3 * i] = fx ? T(0) : ucur[3 * i] + v0 * dt;
    unext[3 * i + 1] = fx ? T(0) : ucur[3 * i + 1] + v1 * dt;
    unext[3 * i + 2] = fx ? T(0) : ucur[3 * i + 2] + v2 * dt;
}

// v_1/2 = v0 + (f_ext/m) dt/2 (f_int(u0 = 0) = 0, src/cae.cpp:471-482); u_1 = v_1/2 dt
template <class T>
__global__ void k_explicit_init(int nd, T *__restrict__ v, T *__restrict__ u1, const T *__restrict__ minv,
                                const T *__restrict__ fext, const int *__restrict__ fixed, T dt)
{
    const int i = blockIdx.x * EX_WG + threadIdx.x;
    if (i >= nd) return;
    for (int c = 0; c < 3; ++c) {
        T vv = v[3 * i + c] + fext[3 * i + c] * minv[i] * (dt / T(2));
        v[3 * i + c] = vv;
        u1[3 * i + c] = fixed[i] ? T(0) : vv * dt;
    }
}

template <class T>
static void launch_step(ofea_ctx *c, ExplicitWS *w, T lam, T mu, T dt, const T *ucur, T *unext)
{
    if (w->ne[1])
        hipLaunchKernelGGL(k_force_hex<T>, dim3((8 * w->ne[1] + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream,
                           w->ne[1], w->conn[1], (const T *)w->xyz_t, ucur, lam, mu, (T *)w->fe[1]);
    if (w->ne[0])
        hipLaunchKernelGGL(k_force_tet<T>, dim3((w->ne[0] + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream, w->ne[0],
                           w->conn[0], (const T *)w->xyz_t, ucur, lam, mu, (T *)w->fe[0]);
    hipLaunchKernelGGL(k_node_update<T>, dim3((w->nd + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream, w->nd,
                       w->inc_ptr, w->inc, w->ne[0], w->ne[1], (const T *)w->fe[0], (const T *)w->fe[1], ucur, unext,
                       (T *)w->v,
                       (const T *)w->minv, (const T *)w->fext, w->fixed, dt);
}

template <class T>
static int run_t(ofea_ctx *c, double E, double nu, const double *mass, const double *f_ext, const double *v0,
                 const int *fixed_mask, double dt, int nsteps, double *u_out)
{
    ExplicitWS *w = ews(c);
    const int nd = c->nd;
    const size_t n3 = 3 * (size_t)nd;
    for (int i = 0; i < nd; ++i)
        if (!(mass[i] > 0.0)) return fail(OFEA_ERR_ARG, "nodal mass must be positive");
    // The blocking copies below run on the null stream, which the context's
    // non-blocking stream does not order against: a previous run whose last
    // steps still read v, fext or minv must have finished first.
    OFEA_HIP(hipStreamSynchronize(c->stream));
    if (w->prec != (int)sizeof(T)) {
        if (w->exec) hipGraphExecDestroy(w->exec);
        if (w->graph) hipGraphDestroy(w->graph);
        w->exec = nullptr;
        w->graph = nullptr;
        for (void **p : {&w->xyz_t, &w->fe[0], &w->fe[1], &w->u[0], &w->u[1], &w->v, &w->minv, &w->fext})
            if (*p) {
                hipFree(*p);
                *p = nullptr;
            }
        OFEA_HIP(hipMalloc(&w->xyz_t, sizeof(T) * n3));
        OFEA_HIP(hipMalloc(&w->fe[0], sizeof(T) * std::max(1, 12 * w->ne[0])));
        OFEA_HIP(hipMalloc(&w->fe[1], sizeof(T) * std::max(1, 24 * w->ne[1])));
        OFEA_HIP(hipMalloc(&w->u[0], sizeof(T) * n3));
        OFEA_HIP(hipMalloc(&w->u[1], sizeof(T) * n3));
        OFEA_HIP(hipMalloc(&w->v, sizeof(T) * n3));
        OFEA_HIP(hipMalloc(&w->minv, sizeof(T) * nd));
        OFEA_HIP(hipMalloc(&w->fext, sizeof(T) * n3));
        if (!w->fixed) OFEA_HIP(hipMalloc((void **)&w->fixed, sizeof(int) * nd));
        w->prec = (int)sizeof(T);
    }
    // host-side conversion of the inputs to the run precision
    std::vector<T> h(n3);
    for (size_t i = 0; i < n3; ++i) h[i] = (T)c->xyz[i];
    OFEA_HIP(hipMemcpy(w->xyz_t, h.data(), sizeof(T) * n3, hipMemcpyHostToDevice));
    for (size_t i = 0; i < n3; ++i) h[i] = (T)f_ext[i];
    OFEA_HIP(hipMemcpy(w->fext, h.data(), sizeof(T) * n3, hipMemcpyHostToDevice));
    for (size_t i = 0; i < n3; ++i) h[i] = v0 ? (T)v0[i] : T(0);
    OFEA_HIP(hipMemcpy(w->v, h.data(), sizeof(T) * n3, hipMemcpyHostToDevice));
    std::vector<T> mi(nd);
    for (int i = 0; i < nd; ++i) mi[i] = (T)(1.0 / mass[i]);
    OFEA_HIP(hipMemcpy(w->minv, mi.data(), sizeof(T) * nd, hipMemcpyHostToDevice));
    OFEA_HIP(hipMemcpy(w->fixed, fixed_mask, sizeof(int) * nd, hipMemcpyHostToDevice));
    double lam_d, mu_d;
    lame<double>(E, nu, lam_d, mu_d);
    const T lam = (T)lam_d, mu = (T)mu_d, tdt = (T)dt;
    hipLaunchKernelGGL(k_explicit_init<T>, dim3((nd + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream, nd, (T *)w->v,
                       (T *)w->u[0], (const T *)w->minv, (const T *)w->fext, w->fixed, tdt);
    OFEA_HIP(hipGetLastError());
    // graph of EX_GRAPH_STEPS steps starting from buffer 0; parameters baked in
    if (w->exec && (w->g_E != E || w->g_nu != nu || w->g_dt != dt)) {
        hipGraphExecDestroy(w->exec);
        hipGraphDestroy(w->graph);
        w->exec = nullptr;
        w->graph = nullptr;
    }
    int done = 0;
    if (nsteps >= EX_GRAPH_STEPS) {
        if (!w->exec) {
            OFEA_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (int s = 0; s < EX_GRAPH_STEPS; ++s)
                launch_step<T>(c, w, lam, mu, tdt, (const T *)w->u[s & 1], (T *)w->u[(s + 1) & 1]);
            OFEA_HIP(hipStreamEndCapture(c->stream, &w->graph));
            OFEA_HIP(hipGraphInstantiate(&w->exec, w->graph, nullptr, nullptr, 0));
            w->g_E = E;
            w->g_nu = nu;
            w->g_dt = dt;
        }
        while (nsteps - done >= EX_GRAPH_STEPS) {
            OFEA_HIP(hipGraphLaunch(w->exec, c->stream));
            done += EX_GRAPH_STEPS;
        }
    }
    for (; done < nsteps; ++done)
        launch_step<T>(c, w, lam, mu, tdt, (const T *)w->u[done & 1], (T *)w->u[(done + 1) & 1]);
    OFEA_HIP(hipGetLastError());
    if (u_out) {
        // u_nsteps is the buffer the last element phase read
        const T *ures = (const T *)w->u[(nsteps - 1) & 1];
        if (nsteps == 0) {
            std::fill(u_out, u_out + n3, 0.0);
        } else {
            OFEA_HIP(hipMemcpyAsync(h.data(), ures, sizeof(T) * n3, hipMemcpyDeviceToHost, c->stream));
            OFEA_HIP(hipStreamSynchronize(c->stream));
            for (size_t i = 0; i < n3; ++i) u_out[i] = (double)h[i];
        }
    }
    return OFEA_OK;
}

int explicit_run(ofea_ctx *c, double E, double nu, const double *mass, const double *f_ext, const double *v0,
                 const int *fixed_mask, double dt, int nsteps, int precision, double *u_out)
{
    int rc = ensure_incidence(c);
    if (rc) return rc;
    if (precision == OFEA_FP32) return run_t<float>(c, E, nu, mass, f_ext, v0, fixed_mask, dt, nsteps, u_out);
    return run_t<double>(c, E, nu, mass, f_ext, v0, fixed_mask, dt, nsteps, u_out);
}

}  // namespace ofea
