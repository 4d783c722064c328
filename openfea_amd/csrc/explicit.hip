// explicit.hip -- lumped mass and the half-step central-difference loop.
//
// Reference: CAE_process::explicit_analysis (src/cae.cpp:420-574) with
// build_ele_mass (src/Elements/C3D8.cpp:58-108, src/Elements/C3D4.cpp:75-100)
// and cal_in_force (src/Elements/C3D8.cpp:158-193).  The reference accumulates
// InForce -= Ke(X0) du every step (never zeroed, src/explicit_tools.cpp:41), so
// InForce_k = -K0 u_k; here f_int = -K0 u_k is evaluated directly, matrix-free,
// which is the same quantity without the running-sum round-off.
//
// One step = two kernels.  All-hex meshes, fixed step (ofea_explicit_run):
//   k_tile_forces   one workgroup per compact tile of 256 hexes: the tile's
//                   nodes staged in LDS, one lane per hex computes its nodal
//                   forces Ke u_e (moment form of the trilinear element,
//                   hex_moment_forces), the forces are summed in LDS in 8
//                   conflict-free corner passes, and the nodes whose hexes
//                   all lie in the tile get the central-difference update;
//   k_tile_boundary the other nodes add their tiles' partial sums in tile
//                   order and get the same update.
//   The run keeps its node state in the tile order (boundary nodes, then each
//   tile's interior nodes).
// Meshes with tets, tiles that do not qualify, and the reference's time
// control (ofea_explicit_analysis):
//   element phase  one lane per element: the nodal forces Ke u_e stored
//                  local-node-major, fe[a][e], so that a wave's stores of one
//                  local node fill whole lines and the node phase's gathers
//                  of consecutive nodes read consecutive elements; no atomics;
//   node phase     one lane per node: f = -sum fe over its incident elements
//                  in ascending order, a = (f + f_ext)/m, v += a dt,
//                  u_next = u + v dt (0 at clamped nodes).
// Both launches map contiguous element / node ranges to one XCD each.
// Runs in FP32 (production) or FP64 (parity) with the same kernels.
#include <algorithm>
#include <vector>

#include "elem_device.h"
#include "ofea_internal.h"

namespace ofea {

constexpr int EX_WG = 256;
constexpr int EX_GRAPH_STEPS = 50;  // steps per captured graph (even)
constexpr int DT_HEX = 32;          // hexes per workgroup of the critical-step kernel (8 lanes each)
constexpr int DT_LDS = 201;         // doubles per hex in LDS: 8 points x (24 gradients + detJ), +1 pad

// Time control of the reference's loop (src/cae.cpp:420-574), kept on the
// device and advanced by one thread per step (k_time_ctl).  All times FP64.
struct TimeCtl {
    double t, dt, dt_old, total;  // current time, step (DBL_MAX before an automatic update), previous step
    double cv, cu;                // this step's velocity factor (dt_old + dt)/2 and displacement step dt
    int out, gap, auto_dt, steps, max_steps;
    int done;                     // loop ended: kernels of later steps return at once
    int node_go;                  // the node kernel of this step runs
    int save, slot, nsaved, max_frames, overflow;
};

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
    // reference time control (ofea_explicit_analysis)
    TimeCtl *tc = nullptr;                   // device state
    double *dtpart = nullptr;                // per-workgroup critical steps
    void *frames = nullptr;                  // saved displacements, run precision
    int *fidx = nullptr;                     // output index of each save
    double *ft = nullptr;                    // time of each save
    int frames_cap = 0;
    hipGraph_t tgraph = nullptr;
    hipGraphExec_t texec = nullptr;
    double tg_E = 0, tg_nu = 0;
    int tg_auto = -1;
    // element tiles of an all-hex mesh (k_tile_forces + k_tile_boundary); a
    // tiled run keeps its node state in the tile order `perm` (boundary nodes
    // first, then each tile's interior nodes)
    int tiles = 0;                           // 0 not built, 1 built, -1 not applicable
    int ntiles = 0, nbnode = 0, npart = 0;
    int *t_hdr = nullptr;                    // 4 per tile: first interior node, interior count, boundary count, first partial
    int *t_bnode = nullptr;                  // EXT_NT per tile: the tile's boundary nodes (tile order indices)
    uint16_t *t_slot = nullptr;              // 8 per element lane, EX_WG per tile: node-table slots, ~0 = none
    int *b_ptr = nullptr, *b_part = nullptr; // boundary node j: its partial slots, tile order
    void *part = nullptr;                    // 3 per partial, run precision
    std::vector<int> perm;                   // tile order index -> node id
};

// Tiles of the explicit step on all-hex meshes: EX_WG hexes per tile (one
// lane each), compact by recursive bisection of the element centroids, at
// most EXT_NT distinct nodes.  Within a tile no node is corner a of two
// hexes, so the 8 passes (one per corner) of the LDS force sums have no
// conflicts.
constexpr int EXT_NT = 640;

static ExplicitWS *ews(ofea_ctx *c) { return (ExplicitWS *)c->explicit_ws; }

void explicit_free(ofea_ctx *c)
{
    ExplicitWS *w = ews(c);
    if (!w) return;
    if (w->exec) hipGraphExecDestroy(w->exec);
    if (w->graph) hipGraphDestroy(w->graph);
    if (w->texec) hipGraphExecDestroy(w->texec);
    if (w->tgraph) hipGraphDestroy(w->tgraph);
    for (void *p : {(void *)w->inc_ptr, (void *)w->inc, w->xyz_t, w->fe[0], w->fe[1], w->u[0], w->u[1], w->v, w->minv,
                    w->fext, (void *)w->fixed, (void *)w->tc, (void *)w->dtpart, w->frames, (void *)w->fidx,
                    (void *)w->ft, (void *)w->t_hdr, (void *)w->t_bnode, (void *)w->t_slot, (void *)w->b_ptr,
                    (void *)w->b_part, w->part})
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

// Element tiles of an all-hex mesh for the tiled explicit step.  Leaves
// w->tiles = -1 (the two-kernel step is kept) when the mesh has tets, a tile
// touches more than EXT_NT nodes or a node is the same corner of two hexes of
// one tile.  Tiles: the hex centroids cut by recursive bisection across the
// longest extent (ties broken by the other axes, then the index) into pieces
// of EX_WG hexes, all full but the last.  A node whose hexes all lie in one
// tile is that tile's interior node; every other node (a boundary node) gets
// one partial slot per tile that touches it, in tile order.  The run's node
// state is kept in the tile order: the boundary nodes (ascending), then each
// tile's interior nodes, so that both kernels read and write contiguous runs.
static int build_tiles(ofea_ctx *c, ExplicitWS *w)
{
    if (w->tiles) return OFEA_OK;
    w->tiles = -1;
    if (c->has_type[OFEA_C3D4] || !c->has_type[OFEA_C3D8]) return OFEA_OK;
    const GatherPlan &G = c->G[OFEA_C3D8];
    const int ne = G.ne, nd = c->nd;
    if (ne == 0) return OFEA_OK;
    std::vector<double> cen(3 * (size_t)ne, 0.0);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) cen[3 * (size_t)e + k] += 0.125 * c->xyz[3 * (size_t)G.conn[8 * (size_t)e + a] + k];
    std::vector<int> ord(ne);
    for (int e = 0; e < ne; ++e) ord[e] = e;
    std::vector<std::pair<int, int>> stack{{0, ne}}, pieces;
    while (!stack.empty()) {
        const auto [lo, hi] = stack.back();
        stack.pop_back();
        if (hi - lo <= EX_WG) {
            pieces.push_back({lo, hi});
            continue;
        }
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (int i = lo; i < hi; ++i)
            for (int k = 0; k < 3; ++k) {
                mn[k] = std::min(mn[k], cen[3 * (size_t)ord[i] + k]);
                mx[k] = std::max(mx[k], cen[3 * (size_t)ord[i] + k]);
            }
        int ax = 0;
        for (int k = 1; k < 3; ++k)
            if (mx[k] - mn[k] > mx[ax] - mn[ax]) ax = k;
        const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
        const int mid = lo + ((hi - lo) / 2 + EX_WG - 1) / EX_WG * EX_WG;
        std::nth_element(ord.begin() + lo, ord.begin() + mid, ord.begin() + hi, [&](int x, int y) {
            const double *px = &cen[3 * (size_t)x], *py = &cen[3 * (size_t)y];
            if (px[ax] != py[ax]) return px[ax] < py[ax];
            if (px[a1] != py[a1]) return px[a1] < py[a1];
            if (px[a2] != py[a2]) return px[a2] < py[a2];
            return x < y;
        });
        stack.push_back({mid, hi});
        stack.push_back({lo, mid});
    }
    const int nt = (int)pieces.size();
    std::vector<int> total(nd, 0);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 8; ++a) total[G.conn[8 * (size_t)e + a]]++;
    // pass 1: each tile's hexes (ascending) and nodes, interior or boundary
    std::vector<std::vector<int>> tel(nt), tint(nt), tbnd(nt);
    std::vector<int> intile(nd, 0), corner_stamp(8 * (size_t)nd, -1), owner(nd, -1);
    for (int t = 0; t < nt; ++t) {
        std::vector<int> &els = tel[t];
        els.assign(ord.begin() + pieces[t].first, ord.begin() + pieces[t].second);
        std::sort(els.begin(), els.end());
        std::vector<int> nodes;
        for (int e : els)
            for (int a = 0; a < 8; ++a) {
                const int n = G.conn[8 * (size_t)e + a];
                if (corner_stamp[8 * (size_t)n + a] == t) return OFEA_OK;  // a node at one corner of two hexes
                corner_stamp[8 * (size_t)n + a] = t;
                if (intile[n]++ == 0) nodes.push_back(n);
            }
        if ((int)nodes.size() > EXT_NT) return OFEA_OK;
        std::sort(nodes.begin(), nodes.end());
        for (int n : nodes) {
            if (intile[n] == total[n]) {
                tint[t].push_back(n);
                owner[n] = t;
            } else {
                tbnd[t].push_back(n);
            }
            intile[n] = 0;
        }
    }
    // pass 2: the tile order (boundary nodes ascending, then interior nodes tile by tile)
    std::vector<int> perm, iperm(nd, -1);
    for (int n = 0; n < nd; ++n)
        if (owner[n] < 0) perm.push_back(n);
    const int nb = (int)perm.size();
    std::vector<int> hdr(4 * (size_t)nt);
    int npart = 0;
    for (int t = 0; t < nt; ++t) {
        hdr[4 * t] = (int)perm.size();
        hdr[4 * t + 1] = (int)tint[t].size();
        hdr[4 * t + 2] = (int)tbnd[t].size();
        hdr[4 * t + 3] = npart;
        npart += (int)tbnd[t].size();
        perm.insert(perm.end(), tint[t].begin(), tint[t].end());
    }
    for (int j = 0; j < nd; ++j) iperm[perm[j]] = j;
    std::vector<int> t_bnode((size_t)nt * EXT_NT, -1), slot(nd, -1);
    std::vector<uint16_t> t_slot((size_t)nt * EX_WG * 8, 0xFFFFu);
    std::vector<std::vector<int>> parts_of(nb);
    for (int t = 0; t < nt; ++t) {
        const int ni = (int)tint[t].size();
        for (int i = 0; i < ni; ++i) slot[tint[t][i]] = i;
        for (int k = 0; k < (int)tbnd[t].size(); ++k) {
            const int n = tbnd[t][k];
            slot[n] = ni + k;
            t_bnode[(size_t)t * EXT_NT + k] = iperm[n];
            parts_of[iperm[n]].push_back(hdr[4 * t + 3] + k);
        }
        for (int l = 0; l < (int)tel[t].size(); ++l)
            for (int a = 0; a < 8; ++a)
                t_slot[((size_t)t * EX_WG + l) * 8 + a] = (uint16_t)slot[G.conn[8 * (size_t)tel[t][l] + a]];
    }
    std::vector<int> b_ptr{0}, b_part;
    for (int j = 0; j < nb; ++j) {
        b_part.insert(b_part.end(), parts_of[j].begin(), parts_of[j].end());
        b_ptr.push_back((int)b_part.size());
    }
    int rc;
    if ((rc = dev_upload(w->t_hdr, hdr.data(), hdr.size()))) return rc;
    if ((rc = dev_upload(w->t_bnode, t_bnode.data(), t_bnode.size()))) return rc;
    if ((rc = dev_upload(w->t_slot, t_slot.data(), t_slot.size()))) return rc;
    if ((rc = dev_upload(w->b_ptr, b_ptr.data(), b_ptr.size()))) return rc;
    if ((rc = dev_upload(w->b_part, b_part.data(), std::max<size_t>(b_part.size(), 1)))) return rc;
    w->ntiles = nt;
    w->nbnode = nb;
    w->npart = npart;
    w->perm = std::move(perm);
    w->tiles = 1;
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
// Abaqus node of hex corner b = [x > 0] + 2 [y > 0] + 4 [z > 0].
__host__ __device__ constexpr int hex_node_of_corner(int b)
{
    return (b & 4) + ((b & 3) == 0 ? 0 : (b & 3) == 1 ? 1 : (b & 3) == 3 ? 2 : 3);
}

// In place y_m = sum_b prod_{j in m} s_j(b) x_b over the 8 corners (s_j(b) =
// +1 when bit j of b is set, else -1), NV components each: one butterfly
// (lo + hi, hi - lo) per axis.
template <class T, int NV>
__device__ __forceinline__ void wht8(T x[8][NV])
{
#pragma unroll
    for (int j = 1; j < 8; j <<= 1)
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if (!(b & j))
#pragma unroll
                for (int k = 0; k < NV; ++k) {
                    const T lo = x[b][k], hi = x[b | j][k];
                    x[b][k] = lo + hi;
                    x[b | j][k] = hi - lo;
                }
}

// Its transpose, f_b = sum_m prod_{j in m} s_j(b) y_m: per axis (lo - hi, lo + hi).
template <class T, int NV>
__device__ __forceinline__ void wht8_t(T x[8][NV])
{
#pragma unroll
    for (int j = 1; j < 8; j <<= 1)
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if (!(b & j))
#pragma unroll
                for (int k = 0; k < NV; ++k) {
                    const T lo = x[b][k], hi = x[b | j][k];
                    x[b][k] = lo - hi;
                    x[b | j][k] = lo + hi;
                }
}

// Block b of a launch with 8 * q blocks, renumbered so that the blocks the
// hardware places on one XCD (b mod 8) take q consecutive logical blocks:
// each XCD works on one contiguous range of elements or nodes, whose
// gathered neighbours then stay in that XCD's L2.
__device__ __forceinline__ int xcd_block(int q) { return (int)(blockIdx.x & 7) * q + (int)(blockIdx.x >> 3); }

// Hex forces f_a = Ke u_e, one lane per element, in the moment form of the
// trilinear element: with c_m = 1/8 sum_a s_a^m X_a (s_a^m the product of
// node a's corner signs over the axes in m), dX/dxi at a Gauss point is a
// combination of 4 of the c_m, and the same for u; H' = adj(J) dU/dxi is
// detJ grad u, S = mu (H' + H'^T) + lam tr(H') I is detJ sigma, and
// P = S adj(J) / detJ.  The nodal forces sum_p P dN_a/dxi (weights 1) are
// the inverse transform 1/8 sum_m s_a^m A_m of 7 moments A_m accumulated
// over the points, which is Ke u (src/Elements/C3D8_Gauss.cpp:80-108,
// src/Elements/C3D8.cpp:158-193).
// The nodal forces of one hex from its corner-ordered coefficient inputs:
// cX, cU hold 1/8 X and 1/8 u at corner b (node hex_node_of_corner(b)) and
// are transformed in place; on return the force at corner b is 1/8 Am[b].
template <class T>
__device__ __forceinline__ void hex_moment_forces(T cX[8][3], T cU[8][3], T lam, T mu, T Am[8][3])
{
    wht8<T, 3>(cX);
    wht8<T, 3>(cU);
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int k = 0; k < 3; ++k) Am[m][k] = T(0);
    const T g = T(0.57735026918962576451);  // 1/sqrt(3)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        const T xi = T(hsgn(p, 0)) * g, eta = T(hsgn(p, 1)) * g, zeta = T(hsgn(p, 2)) * g;
        T J[3][3], D[3][3];  // J[i][k] = dX_k/dxi_i, D[i][c] = du_c/dxi_i
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            J[0][k] = cX[1][k] + eta * cX[3][k] + zeta * cX[5][k] + (eta * zeta) * cX[7][k];
            J[1][k] = cX[2][k] + xi * cX[3][k] + zeta * cX[6][k] + (xi * zeta) * cX[7][k];
            J[2][k] = cX[4][k] + xi * cX[5][k] + eta * cX[6][k] + (xi * eta) * cX[7][k];
            D[0][k] = cU[1][k] + eta * cU[3][k] + zeta * cU[5][k] + (eta * zeta) * cU[7][k];
            D[1][k] = cU[2][k] + xi * cU[3][k] + zeta * cU[6][k] + (xi * zeta) * cU[7][k];
            D[2][k] = cU[4][k] + xi * cU[5][k] + eta * cU[6][k] + (xi * eta) * cU[7][k];
        }
        T A[3][3];  // adj(J) = detJ J^-1
        A[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        A[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
        A[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        A[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        A[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
        A[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
        A[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        A[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
        A[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const T det = J[0][0] * A[0][0] + J[0][1] * A[1][0] + J[0][2] * A[2][0];
        T Hp[3][3];  // detJ du_c/dx_j
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int j = 0; j < 3; ++j) Hp[c][j] = A[j][0] * D[0][c] + A[j][1] * D[1][c] + A[j][2] * D[2][c];
        const T ltr = lam * (Hp[0][0] + Hp[1][1] + Hp[2][2]);
        T rdet;
        if constexpr (sizeof(T) == 4)
            rdet = __builtin_amdgcn_rcpf(det);  // v_rcp_f32, 1 ulp: within the FP32 path's rounding
        else
            rdet = T(1) / det;
        T S[3][3];  // sigma = (detJ sigma) / detJ
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int j = 0; j < 3; ++j) S[c][j] = rdet * (mu * (Hp[c][j] + Hp[j][c]) + (c == j ? ltr : T(0)));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T P[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) P[i] = S[c][0] * A[0][i] + S[c][1] * A[1][i] + S[c][2] * A[2][i];
            Am[1][c] += P[0];
            Am[2][c] += P[1];
            Am[4][c] += P[2];
            Am[3][c] += eta * P[0] + xi * P[1];
            Am[5][c] += zeta * P[0] + xi * P[2];
            Am[6][c] += zeta * P[1] + eta * P[2];
            Am[7][c] += (eta * zeta) * P[0] + (xi * zeta) * P[1] + (xi * eta) * P[2];
        }
    }
    // f at corner b = 1/8 sum_m s_b^m A_m: the transposed butterfly
    wht8_t<T, 3>(Am);
}

// One lane per hex: its nodal forces (hex_moment_forces) stored fe[a][e].
template <class T>
__global__ __launch_bounds__(EX_WG) void k_force_hex(int ne, int q, const int *__restrict__ conn,
                                                     const T *__restrict__ xyz, const T *__restrict__ u, T lam, T mu,
                                                     T *__restrict__ fe, const int *__restrict__ stop)
{
    if (stop && *stop) return;
    const int e = xcd_block(q) * EX_WG + threadIdx.x;
    if (e >= ne) return;
    const int4 n0 = *reinterpret_cast<const int4 *>(conn + 8 * (size_t)e);
    const int4 n1 = *reinterpret_cast<const int4 *>(conn + 8 * (size_t)e + 4);
    const int nodes[8] = {n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, n1.z, n1.w};
    // coefficients by a 3-stage Walsh-Hadamard butterfly over the corner
    // index b = [x > 0] + 2 [y > 0] + 4 [z > 0] (node hex_corner[b])
    T cX[8][3], cU[8][3];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int node = nodes[hex_node_of_corner(b)];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cX[b][k] = T(0.125) * xyz[3 * node + k];
            cU[b][k] = T(0.125) * u[3 * node + k];
        }
    }
    T Am[8][3];
    hex_moment_forces<T>(cX, cU, lam, mu, Am);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        T *dst = fe + 3 * ((size_t)hex_node_of_corner(b) * ne + e);
        dst[0] = T(0.125) * Am[b][0];
        dst[1] = T(0.125) * Am[b][1];
        dst[2] = T(0.125) * Am[b][2];
    }
}

template <class T>
__global__ __launch_bounds__(EX_WG) void k_force_tet(int ne, int q, const int *__restrict__ conn,
                                                     const T *__restrict__ xyz, const T *__restrict__ u, T lam, T mu,
                                                     T *__restrict__ fe, const int *__restrict__ stop)
{
    if (stop && *stop) return;
    const int e = xcd_block(q) * EX_WG + threadIdx.x;
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

// f -= the forces of node i's elements, in ascending element order.  The
// incidence codes are read 8 at a time and their force loads issued together
// before the sums (absent entries add nothing).
template <class T>
__device__ __forceinline__ void gather_forces(int i, const int *__restrict__ inc_ptr, const int *__restrict__ inc,
                                              int ne4, int ne8, const T *__restrict__ fe4,
                                              const T *__restrict__ fe8, T &f0, T &f1, T &f2)
{
    const int t0 = inc_ptr[i], t1 = inc_ptr[i + 1];
    for (int tb = t0; tb < t1; tb += 8) {
        T g[8][3];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            g[k][0] = g[k][1] = g[k][2] = T(0);
            if (tb + k < t1) {
                const int code = inc[tb + k], el = code >> 4, type = (code >> 3) & 1, a = code & 7;
                const T *src = type ? fe8 + 3 * ((size_t)a * ne8 + el) : fe4 + 3 * ((size_t)a * ne4 + el);
                g[k][0] = src[0];
                g[k][1] = src[1];
                g[k][2] = src[2];
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            f0 -= g[k][0];
            f1 -= g[k][1];
            f2 -= g[k][2];
        }
    }
}

template <class T>
__global__ __launch_bounds__(EX_WG) void k_node_update(int nd, int q, const int *__restrict__ inc_ptr, const int *__restrict__ inc,
                                                       int ne4, int ne8, const T *__restrict__ fe4, const T *__restrict__ fe8,
                                                       const T *__restrict__ ucur, T *__restrict__ unext,
                                                       T *__restrict__ v, const T *__restrict__ minv,
                                                       const T *__restrict__ fext, const int *__restrict__ fixed, T dt)
{
    const int i = xcd_block(q) * EX_WG + threadIdx.x;
    if (i >= nd) return;
    T f0 = T(0), f1 = T(0), f2 = T(0);
    gather_forces<T>(i, inc_ptr, inc, ne4, ne8, fe4, fe8, f0, f1, f2);
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

// Tiled step of an all-hex mesh (build_tiles), one workgroup per tile, node
// state in the tile order:
//   staging: the tile's nodes' X and u go to LDS (its interior nodes are one
//            contiguous run, its boundary nodes are listed), force sums zeroed;
//   forces:  lane l computes the nodal forces of hex l (hex_moment_forces);
//   sums:    8 passes, one per corner b: every hex adds its force at corner b
//            to that node's LDS sum (no node is corner b of two hexes of the
//            tile, so a pass has no conflicts; the order is fixed);
//   nodes:   an interior node is updated here, a = (f + f_ext)/m,
//            v += a dt, u_next = u + v dt (0 at clamped nodes); a boundary
//            node's sum goes to the tile's partial slot for it.
// f = -K u: the sums hold K u and enter with a minus sign.
template <class T>
__device__ __forceinline__ void tile_node_update(int n, T f0, T f1, T f2, T u0, T u1, T u2, T *__restrict__ unext,
                                                 T *__restrict__ v, const T *__restrict__ minv,
                                                 const T *__restrict__ fext, const int *__restrict__ fixed, T dt)
{
    const T mi = minv[n];
    const T v0 = v[3 * (size_t)n] + (f0 + fext[3 * (size_t)n]) * mi * dt;
    const T v1 = v[3 * (size_t)n + 1] + (f1 + fext[3 * (size_t)n + 1]) * mi * dt;
    const T v2 = v[3 * (size_t)n + 2] + (f2 + fext[3 * (size_t)n + 2]) * mi * dt;
    v[3 * (size_t)n] = v0;
    v[3 * (size_t)n + 1] = v1;
    v[3 * (size_t)n + 2] = v2;
    const bool fx = fixed[n] != 0;
    unext[3 * (size_t)n] = fx ? T(0) : u0 + v0 * dt;
    unext[3 * (size_t)n + 1] = fx ? T(0) : u1 + v1 * dt;
    unext[3 * (size_t)n + 2] = fx ? T(0) : u2 + v2 * dt;
}

template <class T>
__global__ __launch_bounds__(EX_WG) void k_tile_forces(int ntiles, const int *__restrict__ t_hdr,
                                                       const int *__restrict__ t_bnode,
                                                       const uint16_t *__restrict__ t_slot,
                                                       const T *__restrict__ xyz, const T *__restrict__ ucur,
                                                       T *__restrict__ unext, T *__restrict__ v,
                                                       const T *__restrict__ minv, const T *__restrict__ fext,
                                                       const int *__restrict__ fixed, T lam, T mu, T dt,
                                                       T *__restrict__ part)
{
    __shared__ T s_x[3][EXT_NT], s_u[3][EXT_NT], s_f[3][EXT_NT];
    const int tile = xcd_remap(blockIdx.x, ntiles);
    const int t = threadIdx.x;
    const int4 h = reinterpret_cast<const int4 *>(t_hdr)[tile];
    const int ibase = h.x, ni = h.y, nn = h.y + h.z, pbase = h.w;
    const uint4 sl = reinterpret_cast<const uint4 *>(t_slot)[(size_t)tile * EX_WG + t];  // 8 x 16-bit slots
    for (int i = t; i < nn; i += EX_WG) {
        const int n = i < ni ? ibase + i : t_bnode[(size_t)tile * EXT_NT + (i - ni)];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s_x[k][i] = xyz[3 * (size_t)n + k];
            s_u[k][i] = ucur[3 * (size_t)n + k];
            s_f[k][i] = T(0);
        }
    }
    __syncthreads();
    const int slots[8] = {(int)(sl.x & 0xFFFFu), (int)(sl.x >> 16), (int)(sl.y & 0xFFFFu), (int)(sl.y >> 16),
                          (int)(sl.z & 0xFFFFu), (int)(sl.z >> 16), (int)(sl.w & 0xFFFFu), (int)(sl.w >> 16)};
    const bool has = slots[0] != 0xFFFF;
    T Am[8][3];
    if (has) {
        T cX[8][3], cU[8][3];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int q = slots[hex_node_of_corner(b)];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                cX[b][k] = T(0.125) * s_x[k][q];
                cU[b][k] = T(0.125) * s_u[k][q];
            }
        }
        hex_moment_forces<T>(cX, cU, lam, mu, Am);
    }
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        if (has) {
            const int q = slots[hex_node_of_corner(b)];
#pragma unroll
            for (int k = 0; k < 3; ++k) s_f[k][q] += T(0.125) * Am[b][k];
        }
        __syncthreads();
    }
    for (int i = t; i < nn; i += EX_WG) {
        const T f0 = -s_f[0][i], f1 = -s_f[1][i], f2 = -s_f[2][i];
        if (i < ni) {
            tile_node_update<T>(ibase + i, f0, f1, f2, s_u[0][i], s_u[1][i], s_u[2][i], unext, v, minv, fext, fixed,
                                dt);
        } else {
            T *p = part + 3 * (size_t)(pbase + i - ni);
            p[0] = f0;
            p[1] = f1;
            p[2] = f2;
        }
    }
}

// Boundary nodes of the tiled step (the first nb nodes of the tile order):
// f = the partial sums of the tiles that touch the node, added in tile order;
// then the node update.
template <class T>
__global__ __launch_bounds__(EX_WG) void k_tile_boundary(int nb, const int *__restrict__ b_ptr,
                                                         const int *__restrict__ b_part, const T *__restrict__ part,
                                                         const T *__restrict__ ucur, T *__restrict__ unext,
                                                         T *__restrict__ v, const T *__restrict__ minv,
                                                         const T *__restrict__ fext, const int *__restrict__ fixed, T dt)
{
    const int n = blockIdx.x * EX_WG + threadIdx.x;
    if (n >= nb) return;
    T f0 = T(0), f1 = T(0), f2 = T(0);
    for (int k = b_ptr[n]; k < b_ptr[n + 1]; ++k) {
        const T *p = part + 3 * (size_t)b_part[k];
        f0 += p[0];
        f1 += p[1];
        f2 += p[2];
    }
    tile_node_update<T>(n, f0, f1, f2, ucur[3 * (size_t)n], ucur[3 * (size_t)n + 1], ucur[3 * (size_t)n + 2], unext, v,
                        minv, fext, fixed, dt);
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

// blocks per XCD for n items of EX_WG per block (the grid is 8 x that)
static int xcd_q(int n) { return ((n + EX_WG - 1) / EX_WG + 7) / 8; }

template <class T>
static void launch_forces(ofea_ctx *c, ExplicitWS *w, T lam, T mu, const T *ucur, const int *stop)
{
    if (w->ne[1]) {
        const int q = xcd_q(w->ne[1]);
        hipLaunchKernelGGL(k_force_hex<T>, dim3(8 * q), dim3(EX_WG), 0, c->stream, w->ne[1], q, w->conn[1],
                           (const T *)w->xyz_t, ucur, lam, mu, (T *)w->fe[1], stop);
    }
    if (w->ne[0]) {
        const int q = xcd_q(w->ne[0]);
        hipLaunchKernelGGL(k_force_tet<T>, dim3(8 * q), dim3(EX_WG), 0, c->stream, w->ne[0], q, w->conn[0],
                           (const T *)w->xyz_t, ucur, lam, mu, (T *)w->fe[0], stop);
    }
}

template <class T>
static void launch_step(ofea_ctx *c, ExplicitWS *w, T lam, T mu, T dt, const T *ucur, T *unext)
{
    if (w->tiles == 1) {
        hipLaunchKernelGGL(k_tile_forces<T>, dim3(w->ntiles), dim3(EX_WG), 0, c->stream, w->ntiles, w->t_hdr,
                           w->t_bnode, w->t_slot, (const T *)w->xyz_t, ucur, unext, (T *)w->v, (const T *)w->minv,
                           (const T *)w->fext, w->fixed, lam, mu, dt, (T *)w->part);
        if (w->nbnode)
            hipLaunchKernelGGL(k_tile_boundary<T>, dim3((w->nbnode + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream,
                               w->nbnode, w->b_ptr, w->b_part, (const T *)w->part, ucur, unext, (T *)w->v,
                               (const T *)w->minv, (const T *)w->fext, w->fixed, dt);
        return;
    }
    launch_forces<T>(c, w, lam, mu, ucur, nullptr);
    hipLaunchKernelGGL(k_node_update<T>, dim3(8 * xcd_q(w->nd)), dim3(EX_WG), 0, c->stream, w->nd, xcd_q(w->nd),
                       w->inc_ptr, w->inc, w->ne[0], w->ne[1], (const T *)w->fe[0], (const T *)w->fe[1], ucur, unext,
                       (T *)w->v, (const T *)w->minv, (const T *)w->fext, w->fixed, dt);
}

static void drop_graphs(ExplicitWS *w)
{
    if (w->exec) hipGraphExecDestroy(w->exec);
    if (w->graph) hipGraphDestroy(w->graph);
    if (w->texec) hipGraphExecDestroy(w->texec);
    if (w->tgraph) hipGraphDestroy(w->tgraph);
    w->exec = nullptr;
    w->graph = nullptr;
    w->texec = nullptr;
    w->tgraph = nullptr;
}

// Buffers in the run precision and the inputs converted to it.
// perm (nullable): the buffers hold node perm[j] at position j (the tiled run's order)
template <class T>
static int upload_inputs(ofea_ctx *c, ExplicitWS *w, const double *mass, const double *f_ext, const double *v0,
                         const int *fixed_mask, const int *perm = nullptr)
{
    const int nd = c->nd;
    const size_t n3 = 3 * (size_t)nd;
    for (int i = 0; i < nd; ++i)
        if (!(mass[i] > 0.0)) return fail(OFEA_ERR_ARG, "nodal mass must be positive");
    // The blocking copies below run on the null stream, which the context's
    // non-blocking stream does not order against: a previous run whose last
    // steps still read v, fext or minv must have finished first.
    OFEA_HIP(hipStreamSynchronize(c->stream));
    if (w->prec != (int)sizeof(T)) {
        drop_graphs(w);
        for (void **p : {&w->xyz_t, &w->fe[0], &w->fe[1], &w->u[0], &w->u[1], &w->v, &w->minv, &w->fext, &w->frames})
            if (*p) {
                hipFree(*p);
                *p = nullptr;
            }
        w->frames_cap = 0;
        OFEA_HIP(hipMalloc(&w->xyz_t, sizeof(T) * n3));
        OFEA_HIP(hipMalloc(&w->fe[0], sizeof(T) * std::max(1, 12 * w->ne[0])));
        OFEA_HIP(hipMalloc(&w->fe[1], sizeof(T) * std::max(1, 24 * w->ne[1])));
        OFEA_HIP(hipMalloc(&w->u[0], sizeof(T) * n3));
        OFEA_HIP(hipMalloc(&w->u[1], sizeof(T) * n3));
        OFEA_HIP(hipMalloc(&w->v, sizeof(T) * n3));
        OFEA_HIP(hipMalloc(&w->minv, sizeof(T) * nd));
        OFEA_HIP(hipMalloc(&w->fext, sizeof(T) * n3));
        if (!w->fixed) OFEA_HIP(hipMalloc((void **)&w->fixed, sizeof(int) * nd));
        if (w->part) {
            hipFree(w->part);
            w->part = nullptr;
        }
        w->prec = (int)sizeof(T);
    }
    if (w->tiles == 1 && !w->part) OFEA_HIP(hipMalloc(&w->part, sizeof(T) * 3 * (size_t)std::max(w->npart, 1)));
    // host-side conversion of the inputs to the run precision (and order)
    auto src = [&](int j) { return perm ? perm[j] : j; };
    std::vector<T> h(n3);
    for (int j = 0; j < nd; ++j)
        for (int k = 0; k < 3; ++k) h[3 * (size_t)j + k] = (T)c->xyz[3 * (size_t)src(j) + k];
    OFEA_HIP(hipMemcpy(w->xyz_t, h.data(), sizeof(T) * n3, hipMemcpyHostToDevice));
    for (int j = 0; j < nd; ++j)
        for (int k = 0; k < 3; ++k) h[3 * (size_t)j + k] = (T)f_ext[3 * (size_t)src(j) + k];
    OFEA_HIP(hipMemcpy(w->fext, h.data(), sizeof(T) * n3, hipMemcpyHostToDevice));
    for (int j = 0; j < nd; ++j)
        for (int k = 0; k < 3; ++k) h[3 * (size_t)j + k] = v0 ? (T)v0[3 * (size_t)src(j) + k] : T(0);
    OFEA_HIP(hipMemcpy(w->v, h.data(), sizeof(T) * n3, hipMemcpyHostToDevice));
    std::vector<T> mi(nd);
    for (int j = 0; j < nd; ++j) mi[j] = (T)(1.0 / mass[src(j)]);
    OFEA_HIP(hipMemcpy(w->minv, mi.data(), sizeof(T) * nd, hipMemcpyHostToDevice));
    std::vector<int> fx(nd);
    for (int j = 0; j < nd; ++j) fx[j] = fixed_mask[src(j)];
    OFEA_HIP(hipMemcpy(w->fixed, fx.data(), sizeof(int) * nd, hipMemcpyHostToDevice));
    return OFEA_OK;
}

template <class T>
static int run_t(ofea_ctx *c, double E, double nu, const double *mass, const double *f_ext, const double *v0,
                 const int *fixed_mask, double dt, int nsteps, double *u_out)
{
    ExplicitWS *w = ews(c);
    const int nd = c->nd;
    const size_t n3 = 3 * (size_t)nd;
    int rc = build_tiles(c, w);
    if (rc) return rc;
    const int *perm = w->tiles == 1 ? w->perm.data() : nullptr;
    if ((rc = upload_inputs<T>(c, w, mass, f_ext, v0, fixed_mask, perm))) return rc;
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
            std::vector<T> h(n3);
            OFEA_HIP(hipMemcpyAsync(h.data(), ures, sizeof(T) * n3, hipMemcpyDeviceToHost, c->stream));
            OFEA_HIP(hipStreamSynchronize(c->stream));
            for (int j = 0; j < nd; ++j)
                for (int k = 0; k < 3; ++k) u_out[3 * (size_t)(perm ? perm[j] : j) + k] = (double)h[3 * (size_t)j + k];
        }
    }
    return OFEA_OK;
}

// ------------------------------------------------------------------ reference time control
// min in the form of update_timestep (src/Elements/C3D8.cpp:215-218): the
// step is lowered only by a smaller value, so a NaN candidate is skipped.
__device__ __forceinline__ double dt_lower(double dt, double c) { return dt > c ? c : dt; }

// Critical step 2 / sqrt(max_v) of each hex at the current coordinates X + u
// (update_timestep, src/Elements/C3D8.cpp:195-219: max_v = largest row sum of
// |Ke|, starting at DBL_MIN), lowered over the workgroup's hexes into
// part[block].  FP64 whatever the run precision.  Lanes (hex, Gauss point)
// first put the point's gradients and detJ in LDS; then lane (hex, node a)
// forms the rows of node a: with A_ab = sum_p detJ_p g_a(p) g_b(p)^T,
// K_ab = lam A_ab + mu A_ab^T + mu tr(A_ab) I.
template <class T>
__global__ __launch_bounds__(EX_WG) void k_dt_hex(int ne, const int *__restrict__ conn, const double *__restrict__ xyz,
                                                  const T *__restrict__ u, double lam, double mu,
                                                  const TimeCtl *__restrict__ tc, double *__restrict__ part)
{
    if (tc->done) return;
    __shared__ double s_g[DT_HEX * DT_LDS];
    __shared__ double s_min[EX_WG / 64];
    const int tid = threadIdx.x, el = tid >> 3, l = tid & 7;
    const int e = blockIdx.x * DT_HEX + el;
    const int ec = min(e, ne - 1);
    double X[24];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int node = conn[8 * ec + b];
#pragma unroll
        for (int k = 0; k < 3; ++k) X[3 * b + k] = xyz[3 * node + k] + (double)u[3 * node + k];
    }
    {
        double g[3][8];
        const double detJ = hex_point<double>(X, l, g);
        double *dst = s_g + el * DT_LDS + l * 25;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int b = 0; b < 8; ++b) dst[8 * k + b] = g[k][b];
        dst[24] = detJ;
    }
    __syncthreads();
    const double *src = s_g + el * DT_LDS;
    const int a = l;
    double wga[8][3];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        const double w = src[25 * p + 24];
#pragma unroll
        for (int k = 0; k < 3; ++k) wga[p][k] = w * src[25 * p + 8 * k + a];
    }
    double rs[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < 8; ++b) {
        double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const double gb0 = src[25 * p + b], gb1 = src[25 * p + 8 + b], gb2 = src[25 * p + 16 + b];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                A[i][0] += wga[p][i] * gb0;
                A[i][1] += wga[p][i] * gb1;
                A[i][2] += wga[p][i] * gb2;
            }
        }
        const double tr = A[0][0] + A[1][1] + A[2][2];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) rs[i] += fabs(lam * A[i][j] + mu * A[j][i] + (i == j ? mu * tr : 0.0));
    }
    double mv = 2.2250738585072014e-308;  // DBL_MIN
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (mv < rs[i]) mv = rs[i];
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        const double other = __shfl_xor(mv, o);
        if (mv < other) mv = other;
    }
    double dt = e < ne ? 2 / sqrt(mv) : 1.7976931348623157e308;
#pragma unroll
    for (int o = 8; o < 64; o <<= 1) dt = dt_lower(dt, __shfl_xor(dt, o));
    if ((tid & 63) == 0) s_min[tid >> 6] = dt;
    __syncthreads();
    if (tid == 0) {
        double m = s_min[0];
        for (int k = 1; k < EX_WG / 64; ++k) m = dt_lower(m, s_min[k]);
        part[blockIdx.x] = m;
    }
}

// The same for tets (src/Elements/C3D4.cpp:140-164), one lane per element:
// K_ab = V (lam g_a g_b^T + mu g_b g_a^T + mu (g_a . g_b) I), V = detJ / 6.
template <class T>
__global__ __launch_bounds__(EX_WG) void k_dt_tet(int ne, const int *__restrict__ conn, const double *__restrict__ xyz,
                                                  const T *__restrict__ u, double lam, double mu,
                                                  const TimeCtl *__restrict__ tc, double *__restrict__ part)
{
    if (tc->done) return;
    __shared__ double s_min[EX_WG / 64];
    const int tid = threadIdx.x, e = blockIdx.x * EX_WG + tid;
    const int ec = min(e, ne - 1);
    double X[12];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int node = conn[4 * ec + b];
#pragma unroll
        for (int k = 0; k < 3; ++k) X[3 * b + k] = xyz[3 * node + k] + (double)u[3 * node + k];
    }
    double g[3][4];
    const double V = tet_grads<double>(X, g) / 6.0;
    double mv = 2.2250738585072014e-308;  // DBL_MIN
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double rs[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const double dot = g[0][a] * g[0][b] + g[1][a] * g[1][b] + g[2][a] * g[2][b];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    rs[i] += fabs(V * (lam * g[i][a] * g[j][b] + mu * g[j][a] * g[i][b] + (i == j ? mu * dot : 0.0)));
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if (mv < rs[i]) mv = rs[i];
    }
    double dt = e < ne ? 2 / sqrt(mv) : 1.7976931348623157e308;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) dt = dt_lower(dt, __shfl_xor(dt, o));
    if ((tid & 63) == 0) s_min[tid >> 6] = dt;
    __syncthreads();
    if (tid == 0) {
        double m = s_min[0];
        for (int k = 1; k < EX_WG / 64; ++k) m = dt_lower(m, s_min[k]);
        part[blockIdx.x] = m;
    }
}

// CheckTime (src/explicit_tools.cpp:48-56)
__device__ __forceinline__ bool check_time(double &dt, double limit)
{
    if (dt > limit) {
        dt = limit;
        return true;
    }
    return false;
}

// One step of the reference's time control, or with init the set-up before
// its loop (src/cae.cpp:457-499).  Loop step (src/cae.cpp:501-570): leave when
// t > total; the step is the automatic one (0.9 x the lowest critical step)
// or the kept fixed one; CheckTime clamps it to the next output time
// (counting the output) and to the total; the node kernel then updates with
// v += a (dt_old + dt)/2, u += v dt; the displacement is saved when t is the
// total or the current output time; after the save at the total the loop
// ends, else t += dt and the step moves to dt_old.
__global__ __launch_bounds__(EX_WG) void k_time_ctl(TimeCtl *__restrict__ tc, const double *__restrict__ part,
                                                    int npart, int init, int *__restrict__ fidx,
                                                    double *__restrict__ ft)
{
    __shared__ double s_min[EX_WG];
    if (tc->done) {
        if (threadIdx.x == 0) tc->node_go = 0;
        return;
    }
    const int tid = threadIdx.x;
    double m = 1.7976931348623157e308;
    if (tc->auto_dt) {
        for (int k = tid; k < npart; k += EX_WG) m = dt_lower(m, part[k]);
        s_min[tid] = m;
        __syncthreads();
        for (int s = EX_WG / 2; s > 0; s >>= 1) {
            if (tid < s) s_min[tid] = dt_lower(s_min[tid], s_min[tid + s]);
            __syncthreads();
        }
        m = s_min[0];
    }
    if (tid != 0) return;
    TimeCtl S = *tc;
    if (init) {
        double dt = S.dt;  // DBL_MAX for the automatic step, else the given one
        if (S.auto_dt) dt = dt_lower(dt, m) * 0.9;
        S.cv = (0.0 + dt) / 2;  // time_step_old is 0 before the loop
        S.cu = dt;
        S.out = 0;
        S.steps = 0;
        S.save = 0;
        S.node_go = 0;
        S.overflow = 0;
        S.nsaved = 1;  // the undeformed state at t = 0 (output 0)
        if (S.max_frames > 0) {
            fidx[0] = 0;
            ft[0] = 0.0;
        }
        S.t = 0.0 + dt;
        S.dt_old = dt;
        S.dt = S.auto_dt ? 1.7976931348623157e308 : dt;
        *tc = S;
        return;
    }
    if (!(S.t <= S.total) || S.steps >= S.max_steps) {
        S.overflow = (S.t <= S.total) ? 1 : 0;
        S.done = 1;
        S.node_go = 0;
        S.save = 0;
        *tc = S;
        return;
    }
    S.steps++;
    double dt = S.dt;
    if (S.auto_dt) dt = dt_lower(dt, m) * 0.9;
    if (check_time(dt, (S.total / S.gap) * (S.out + 1) - S.t)) S.out++;
    check_time(dt, S.total - S.t);
    S.cv = (S.dt_old + dt) / 2;
    S.cu = dt;
    S.node_go = 1;
    S.save = 0;
    if (S.t == S.total || S.t == (S.total / S.gap) * (S.out)) {
        if (S.nsaved < S.max_frames) {
            fidx[S.nsaved] = S.out;
            ft[S.nsaved] = S.t;
            S.save = 1;
            S.slot = S.nsaved;
        }
        S.nsaved++;
    }
    if (S.t == S.total) {
        S.done = 1;
    } else {
        S.t += dt;
        S.dt_old = dt;
        S.dt = S.auto_dt ? 1.7976931348623157e308 : dt;
    }
    *tc = S;
}

// Node phase under the time control: f = -sum fe, v += (f + f_ext)/m cv,
// u_next = u + v cu (0 at clamped nodes); u is saved to its frame slot first
// when this step saves.
template <class T>
__global__ __launch_bounds__(EX_WG) void k_node_update_tc(int nd, int q, const int *__restrict__ inc_ptr,
                                                          const int *__restrict__ inc, int ne4, int ne8,
                                                          const T *__restrict__ fe4, const T *__restrict__ fe8,
                                                          const T *__restrict__ ucur, T *__restrict__ unext,
                                                          T *__restrict__ v, const T *__restrict__ minv,
                                                          const T *__restrict__ fext, const int *__restrict__ fixed,
                                                          const TimeCtl *__restrict__ tc, T *__restrict__ frames)
{
    if (!tc->node_go) return;
    const int i = xcd_block(q) * EX_WG + threadIdx.x;
    if (i >= nd) return;
    const T cv = (T)tc->cv, cu = (T)tc->cu;
    T f0 = T(0), f1 = T(0), f2 = T(0);
    gather_forces<T>(i, inc_ptr, inc, ne4, ne8, fe4, fe8, f0, f1, f2);
    const T u0 = ucur[3 * i], u1 = ucur[3 * i + 1], u2 = ucur[3 * i + 2];
    if (tc->save) {
        T *fr = frames + (size_t)tc->slot * 3 * nd + 3 * i;
        fr[0] = u0;
        fr[1] = u1;
        fr[2] = u2;
    }
    const T mi = minv[i];
    const T v0 = v[3 * i] + (f0 + fext[3 * i]) * mi * cv;
    const T v1 = v[3 * i + 1] + (f1 + fext[3 * i + 1]) * mi * cv;
    const T v2 = v[3 * i + 2] + (f2 + fext[3 * i + 2]) * mi * cv;
    v[3 * i] = v0;
    v[3 * i + 1] = v1;
    v[3 * i + 2] = v2;
    const bool fx = fixed[i] != 0;
    unext[3 * i] = fx ? T(0) : u0 + v0 * cu;
    unext[3 * i + 1] = fx ? T(0) : u1 + v1 * cu;
    unext[3 * i + 2] = fx ? T(0) : u2 + v2 * cu;
}

// Before the loop: v_1/2 = v0 + (f_ext/m) (0 + dt)/2 (the internal force of
// the undeformed state is 0), u_1 = v_1/2 dt (0 at clamped nodes), frame 0 = 0.
template <class T>
__global__ void k_explicit_init_tc(int nd, T *__restrict__ v, T *__restrict__ u1, const T *__restrict__ minv,
                                   const T *__restrict__ fext, const int *__restrict__ fixed,
                                   const TimeCtl *__restrict__ tc, T *__restrict__ frames)
{
    const int i = blockIdx.x * EX_WG + threadIdx.x;
    if (i >= nd) return;
    const T cv = (T)tc->cv, cu = (T)tc->cu;
    const bool f0 = tc->max_frames > 0;
    for (int c = 0; c < 3; ++c) {
        T vv = v[3 * i + c] + fext[3 * i + c] * minv[i] * cv;
        v[3 * i + c] = vv;
        u1[3 * i + c] = fixed[i] ? T(0) : vv * cu;
        if (f0) frames[3 * i + c] = T(0);
    }
}

template <class T>
static void launch_dt(ofea_ctx *c, ExplicitWS *w, double lam, double mu, const T *ucur)
{
    const int nb8 = (w->ne[1] + DT_HEX - 1) / DT_HEX;
    if (w->ne[1])
        hipLaunchKernelGGL(k_dt_hex<T>, dim3(nb8), dim3(EX_WG), 0, c->stream, w->ne[1], w->conn[1], c->d_xyz, ucur,
                           lam, mu, (const TimeCtl *)w->tc, w->dtpart);
    if (w->ne[0])
        hipLaunchKernelGGL(k_dt_tet<T>, dim3((w->ne[0] + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream, w->ne[0],
                           w->conn[0], c->d_xyz, ucur, lam, mu, (const TimeCtl *)w->tc, w->dtpart + nb8);
}

static int dt_parts(const ExplicitWS *w)
{
    return (w->ne[1] + DT_HEX - 1) / DT_HEX + (w->ne[0] + EX_WG - 1) / EX_WG;
}

template <class T>
static void launch_step_tc(ofea_ctx *c, ExplicitWS *w, T lam, T mu, double lam_d, double mu_d, bool auto_dt,
                           const T *ucur, T *unext)
{
    launch_forces<T>(c, w, lam, mu, ucur, &w->tc->done);
    if (auto_dt) launch_dt<T>(c, w, lam_d, mu_d, ucur);
    hipLaunchKernelGGL(k_time_ctl, dim3(1), dim3(EX_WG), 0, c->stream, w->tc, (const double *)w->dtpart,
                       dt_parts(w), 0, w->fidx, w->ft);
    hipLaunchKernelGGL(k_node_update_tc<T>, dim3(8 * xcd_q(w->nd)), dim3(EX_WG), 0, c->stream, w->nd, xcd_q(w->nd),
                       w->inc_ptr, w->inc, w->ne[0], w->ne[1], (const T *)w->fe[0], (const T *)w->fe[1], ucur, unext,
                       (T *)w->v, (const T *)w->minv, (const T *)w->fext, w->fixed, (const TimeCtl *)w->tc,
                       (T *)w->frames);
}

template <class T>
static int run_timed(ofea_ctx *c, double E, double nu, const double *mass, const double *f_ext, const double *v0,
                     const int *fixed_mask, double time_total, double time_step, int output_gap, int max_steps,
                     int max_frames, double *frames, int *frame_index, double *frame_time, int *n_frames,
                     int *n_steps)
{
    ExplicitWS *w = ews(c);
    const int nd = c->nd;
    const size_t n3 = 3 * (size_t)nd;
    int rc = upload_inputs<T>(c, w, mass, f_ext, v0, fixed_mask);
    if (rc) return rc;
    if (!w->tc) OFEA_HIP(hipMalloc((void **)&w->tc, sizeof(TimeCtl)));
    if (!w->dtpart) OFEA_HIP(hipMalloc((void **)&w->dtpart, sizeof(double) * std::max(1, dt_parts(w))));
    const int cap = std::max(1, max_frames);
    if (w->frames_cap < cap) {
        drop_graphs(w);  // the captured node kernels hold the frame buffer's address
        for (void **p : {&w->frames, (void **)&w->fidx, (void **)&w->ft})
            if (*p) {
                hipFree(*p);
                *p = nullptr;
            }
        OFEA_HIP(hipMalloc(&w->frames, sizeof(T) * n3 * cap));
        OFEA_HIP(hipMalloc((void **)&w->fidx, sizeof(int) * cap));
        OFEA_HIP(hipMalloc((void **)&w->ft, sizeof(double) * cap));
        w->frames_cap = cap;
    }
    const bool auto_dt = time_step == 0.0;
    TimeCtl S{};
    S.dt = auto_dt ? 1.7976931348623157e308 : time_step;
    S.total = time_total;
    S.gap = output_gap;
    S.auto_dt = auto_dt ? 1 : 0;
    S.max_steps = max_steps;
    S.max_frames = max_frames;
    OFEA_HIP(hipMemcpy(w->tc, &S, sizeof(TimeCtl), hipMemcpyHostToDevice));
    double lam_d, mu_d;
    lame<double>(E, nu, lam_d, mu_d);
    const T lam = (T)lam_d, mu = (T)mu_d;
    // set-up before the loop: the automatic step of the undeformed mesh (u = 0)
    OFEA_HIP(hipMemsetAsync(w->u[1], 0, sizeof(T) * n3, c->stream));
    if (auto_dt) launch_dt<T>(c, w, lam_d, mu_d, (const T *)w->u[1]);
    hipLaunchKernelGGL(k_time_ctl, dim3(1), dim3(EX_WG), 0, c->stream, w->tc, (const double *)w->dtpart,
                       dt_parts(w), 1, w->fidx, w->ft);
    hipLaunchKernelGGL(k_explicit_init_tc<T>, dim3((nd + EX_WG - 1) / EX_WG), dim3(EX_WG), 0, c->stream, nd,
                       (T *)w->v, (T *)w->u[0], (const T *)w->minv, (const T *)w->fext, w->fixed,
                       (const TimeCtl *)w->tc, (T *)w->frames);
    OFEA_HIP(hipGetLastError());
    if (w->texec && (w->tg_E != E || w->tg_nu != nu || w->tg_auto != (int)auto_dt)) {
        hipGraphExecDestroy(w->texec);
        hipGraphDestroy(w->tgraph);
        w->texec = nullptr;
        w->tgraph = nullptr;
    }
    if (!w->texec) {
        OFEA_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < EX_GRAPH_STEPS; ++s)
            launch_step_tc<T>(c, w, lam, mu, lam_d, mu_d, auto_dt, (const T *)w->u[s & 1], (T *)w->u[(s + 1) & 1]);
        OFEA_HIP(hipStreamEndCapture(c->stream, &w->tgraph));
        OFEA_HIP(hipGraphInstantiate(&w->texec, w->tgraph, nullptr, nullptr, 0));
        w->tg_E = E;
        w->tg_nu = nu;
        w->tg_auto = auto_dt ? 1 : 0;
    }
    // Graphs run until the device reports the end; steps after it return at
    // once.  The host reads the state after each batch and sizes the next
    // batch from the remaining time and the last step.
    int batch = 1;
    for (;;) {
        for (int g = 0; g < batch; ++g) OFEA_HIP(hipGraphLaunch(w->texec, c->stream));
        OFEA_HIP(hipMemcpyAsync(&S, w->tc, sizeof(TimeCtl), hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
        if (S.done) break;
        const double rem = S.dt_old > 0.0 ? (S.total - S.t) / S.dt_old : 0.0;
        const double graphs = rem / EX_GRAPH_STEPS + output_gap / (double)EX_GRAPH_STEPS + 1.0;
        batch = (int)std::min(64.0, std::max(1.0, std::ceil(graphs)));
        const int left = (max_steps - S.steps + EX_GRAPH_STEPS - 1) / EX_GRAPH_STEPS + 1;
        batch = std::min(batch, std::max(1, left));
    }
    if (n_steps) *n_steps = S.steps;
    if (n_frames) *n_frames = S.nsaved;
    const int keep = std::min(S.nsaved, max_frames);
    if (keep > 0) {
        std::vector<T> h(n3 * keep);
        OFEA_HIP(hipMemcpyAsync(h.data(), w->frames, sizeof(T) * n3 * keep, hipMemcpyDeviceToHost, c->stream));
        if (frame_index)
            OFEA_HIP(hipMemcpyAsync(frame_index, w->fidx, sizeof(int) * keep, hipMemcpyDeviceToHost, c->stream));
        if (frame_time)
            OFEA_HIP(hipMemcpyAsync(frame_time, w->ft, sizeof(double) * keep, hipMemcpyDeviceToHost, c->stream));
        OFEA_HIP(hipStreamSynchronize(c->stream));
        if (frames)
            for (size_t i = 0; i < n3 * keep; ++i) frames[i] = (double)h[i];
    }
    if (S.overflow) return fail(OFEA_ERR_LIMIT, "explicit time loop reached max_steps before time_total");
    if (S.nsaved > max_frames) return fail(OFEA_ERR_LIMIT, "more outputs than max_frames (the first ones are kept)");
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

int explicit_analysis(ofea_ctx *c, double E, double nu, const double *mass, const double *f_ext, const double *v0,
                      const int *fixed_mask, double time_total, double time_step, int output_gap, int precision,
                      int max_steps, int max_frames, double *frames, int *frame_index, double *frame_time,
                      int *n_frames, int *n_steps)
{
    int rc = ensure_incidence(c);
    if (rc) return rc;
    if (precision == OFEA_FP32)
        return run_timed<float>(c, E, nu, mass, f_ext, v0, fixed_mask, time_total, time_step, output_gap, max_steps,
                                max_frames, frames, frame_index, frame_time, n_frames, n_steps);
    return run_timed<double>(c, E, nu, mass, f_ext, v0, fixed_mask, time_total, time_step, output_gap, max_steps,
                             max_frames, frames, frame_index, frame_time, n_frames, n_steps);
}

}  // namespace ofea
