// post.hip -- element strain / stress and their nodal averages on the device.
//
// Replaces simulation_post::get_cauchy_stress_3d (src/post_process.cpp:79-168)
// with the element routine get_stress_node_hammer
// (src/Elements/C3D4_hammer.cpp:185-226).  Two kernels:
//   k_post_elem: one thread per element gathers its node coordinates and
//     displacements (a host vector, or the context's resident free-DOF
//     solution read through the nodes' free ranks, 0 at clamped nodes) and
//     writes its 7
//     values as one 64-byte row:
//       OFEA_POST_STRAIN  the reference's output, B u_e: engineering strain
//                         [xx, yy, zz, xy, yz, xz] (the reference leaves
//                         D * B u commented out), and its "Mises" formula
//                         sqrt(0.5 ((e0-e1)^2 + (e1-e2)^2 + (e2-e0)^2
//                         + 6 (e3^2 + e4^2 + e5^2))) on those values;
//       OFEA_POST_STRESS  Cauchy stress D B u_e (small strain) and its von
//                         Mises value, same formula;
//     C3D4 strain is constant; the reference has no C3D8 routine (ele_base's
//     no-op), here C3D8 takes the volume average of the 2x2x2 Gauss-point
//     strains, sum_g detJ_g B_g u_e / sum_g detJ_g;
//   k_post_node: eight lanes per node (lane = component) sum the rows of the
//     node's elements in ascending element order (the reference's loop
//     order) and divide by the number of element incidences; a node in no
//     element keeps 0.
// On all-tet meshes k_post_chunks computes the nodal values in one pass
// instead, over chunks of consecutive nodes whose distinct tets it evaluates
// once per chunk (no element rows in memory); meshes beyond its chunk
// capacity use k_post_pack + k_post_tet_fused (each tet evaluated at its 4
// nodes); k_post_elem then runs only when the element values are asked for.
#include <algorithm>
#include <vector>

#include "elem_device.h"
#include "ofea_internal.h"

namespace ofea {
namespace {

constexpr int POST_WG = 256;

struct PostWs {
    int nd = 0, ne = 0;
    int *conn = nullptr;     // ne x 8, 0-based; C3D4 rows keep the nodes' free ranks in slots 4-7
    int *rank8 = nullptr;    // ne x 8 free ranks of C3D8 nodes (meshes with C3D8 only)
    int *etype = nullptr;    // ne
    int *inc_ptr = nullptr;  // nd + 1
    int *inc = nullptr;      // element of each incidence, ascending per node
    double *E = nullptr;     // ne per-element moduli or null
    double *u = nullptr;     // 3 nd
    double *ev = nullptr;    // ne x 8: the 7 element values of element e in row e
    double *ev7 = nullptr;   // 7 x ne, component-major copy for the caller (allocated on first request)
    double *nv = nullptr;    // 7 x nd
    double *xu = nullptr;    // nd x 6 packed coordinates and displacement (all-tet meshes beyond the chunk plan)
    // chunk plan of all-tet meshes (k_post_chunks): fixed strides per chunk
    int nchunks = 0;
    int *ch_hdr = nullptr;       // 4 per chunk: first owned node, owned nodes, distinct elements, incidence offset
    int *ch_node = nullptr;      // PC_NTAB per chunk: node id (owned nodes first, in order), -1 = none
    uint32_t *ch_eloc = nullptr; // PC_ECAP per chunk: 4 x 8-bit node slots per element, ascending element
    int *ch_eid = nullptr;       // PC_ECAP per chunk: element index (per-element moduli)
    uint16_t *ch_ist = nullptr;  // PC_NOWN + 1 per chunk: first incidence of each owned node, chunk-relative
    uint8_t *ch_inc = nullptr;   // the owned nodes' elements as chunk slots, ascending, chunk after chunk
};

// Chunk plan of the fused nodal pass on all-tet meshes: a chunk owns a run of
// consecutive nodes (all nodes, clamped ones too) and lists the distinct
// elements of those nodes (computed once per chunk) and every node they touch.
constexpr int PC_ECAP = 248;  // distinct elements per chunk (one per thread)
constexpr int PC_NTAB = 160;  // nodes per chunk, owned and touched
constexpr int PC_NOWN = 32;   // owned nodes per chunk (eight lanes each)
constexpr int PC_ICAP = 1024; // incidences of the owned nodes per chunk

__device__ inline double mises_form(const double s[6])
{
    const double a = s[0] - s[1], b = s[1] - s[2], c = s[2] - s[0];
    return sqrt(0.5 * (a * a + b * b + c * c + 6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
}

// strain rows of B (src/Elements/C3D4_hammer.cpp:205-219) applied to u
template <int NN>
__device__ inline void strain_of(const double g[3][NN], const double *U, double w, double eps[6])
{
#pragma unroll
    for (int a = 0; a < NN; ++a) {
        const double ux = U[3 * a], uy = U[3 * a + 1], uz = U[3 * a + 2];
        eps[0] += w * g[0][a] * ux;
        eps[1] += w * g[1][a] * uy;
        eps[2] += w * g[2][a] * uz;
        eps[3] += w * (g[1][a] * ux + g[0][a] * uy);
        eps[4] += w * (g[2][a] * uy + g[1][a] * uz);
        eps[5] += w * (g[2][a] * ux + g[0][a] * uz);
    }
}

// One thread per element.  RESIDENT: the displacement is the free-DOF solution
// x read through the nodes' free ranks (0 at clamped nodes); otherwise the
// full 3 nd vector u.
// Output row e of ev holds the 7 values in doubles 0..6 (64-byte rows).
// Consecutive workgroups go to one XCD (xcd_remap), so elements that share
// nodes find them in the same L2.
template <bool RESIDENT>
__global__ __launch_bounds__(POST_WG) void k_post_elem(int ne, const int *__restrict__ conn,
                                                       const int *__restrict__ rank8, const int *__restrict__ etype,
                                                       const double *__restrict__ xyz, const double *__restrict__ u,
                                                       const double *__restrict__ Eel, double E, double nu,
                                                       int quantity, double *__restrict__ ev)
{
    const int e = xcd_remap(blockIdx.x, gridDim.x) * POST_WG + threadIdx.x;
    if (e >= ne) return;
    const int *cn = conn + 8 * (size_t)e;
    // free rank of each element node (-1 clamped): C3D4 keeps them in slots 4-7
    // of its connectivity row, C3D8 in rank8
    const int *rk = !RESIDENT ? nullptr : (etype[e] == OFEA_C3D4 ? cn + 4 : rank8 + 8 * (size_t)e);
    auto load_u = [&](int a, int node, double *U) {
        if (RESIDENT) {
            const int r = rk[a];
            const double *x = u + 3 * (size_t)max(r, 0);
            U[0] = r >= 0 ? x[0] : 0.0;
            U[1] = r >= 0 ? x[1] : 0.0;
            U[2] = r >= 0 ? x[2] : 0.0;
        } else {
            U[0] = u[3 * (size_t)node];
            U[1] = u[3 * (size_t)node + 1];
            U[2] = u[3 * (size_t)node + 2];
        }
    };
    double eps[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (etype[e] == OFEA_C3D4) {
        double X[12], U[12];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int node = cn[a];
#pragma unroll
            for (int c = 0; c < 3; ++c) X[3 * a + c] = xyz[3 * node + c];
            load_u(a, node, U + 3 * a);
        }
        double g[3][4];
        tet_grads(X, g);
        strain_of<4>(g, U, 1.0, eps);
    } else {
        double X[24], U[24];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const int node = cn[a];
#pragma unroll
            for (int c = 0; c < 3; ++c) X[3 * a + c] = xyz[3 * node + c];
            load_u(a, node, U + 3 * a);
        }
        double vol = 0.0;
        for (int gp = 0; gp < 8; ++gp) {
            double g[3][8];
            const double detJ = hex_point(X, gp, g);
            strain_of<8>(g, U, detJ, eps);
            vol += detJ;
        }
        const double iv = 1.0 / vol;
#pragma unroll
        for (int i = 0; i < 6; ++i) eps[i] *= iv;
    }
    double out[8];
    if (quantity == OFEA_POST_STRESS) {
        double lam, mu;
        lame<double>(Eel ? Eel[e] : E, nu, lam, mu);
        const double tr = eps[0] + eps[1] + eps[2];
#pragma unroll
        for (int i = 0; i < 3; ++i) out[i] = lam * tr + 2.0 * mu * eps[i];
#pragma unroll
        for (int i = 3; i < 6; ++i) out[i] = mu * eps[i];
    } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) out[i] = eps[i];
    }
    out[6] = mises_form(out);
    out[7] = 0.0;
    double2 *row = reinterpret_cast<double2 *>(ev + 8 * (size_t)e);
#pragma unroll
    for (int i = 0; i < 4; ++i) row[i] = make_double2(out[2 * i], out[2 * i + 1]);
}

// One 48-byte record per node for the fused pass: x, y, z, ux, uy, uz (the
// displacement of a clamped node is 0 in the resident form).
template <bool RESIDENT>
__global__ __launch_bounds__(POST_WG) void k_post_pack(int nd, const double *__restrict__ xyz,
                                                       const double *__restrict__ u, const int *__restrict__ resort,
                                                       double *__restrict__ xu)
{
    const int i = blockIdx.x * POST_WG + threadIdx.x;
    if (i >= nd) return;
    double r[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = xyz[3 * (size_t)i + c];
    if (RESIDENT) {
        const int f = resort[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) r[3 + c] = f >= 0 ? u[3 * (size_t)f + c] : 0.0;
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) r[3 + c] = u[3 * (size_t)i + c];
    }
    double2 *dst = reinterpret_cast<double2 *>(xu + 6 * (size_t)i);
    dst[0] = make_double2(r[0], r[1]);
    dst[1] = make_double2(r[2], r[3]);
    dst[2] = make_double2(r[4], r[5]);
}

// The 7 values of tet e (strain or stress and the Mises form) from the
// packed node records and the 16-byte tet connectivity rows (conn4, ne x 4,
// 0-based): 1 + 4 x 3 sixteen-byte loads per tet.
__device__ __forceinline__ void tet_values(int e, const int *__restrict__ conn4, const double *__restrict__ xu,
                                           const double *__restrict__ Eel, double E, double nu, int quantity,
                                           double out[7])
{
    const int4 nodes = *reinterpret_cast<const int4 *>(conn4 + 4 * (size_t)e);
    const int nn[4] = {nodes.x, nodes.y, nodes.z, nodes.w};
    double X[12], U[12];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double2 *r = reinterpret_cast<const double2 *>(xu + 6 * (size_t)nn[a]);
        const double2 p0 = r[0], p1 = r[1], p2 = r[2];
        X[3 * a] = p0.x;
        X[3 * a + 1] = p0.y;
        X[3 * a + 2] = p1.x;
        U[3 * a] = p1.y;
        U[3 * a + 1] = p2.x;
        U[3 * a + 2] = p2.y;
    }
    double g[3][4];
    tet_grads(X, g);
    double eps[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    strain_of<4>(g, U, 1.0, eps);
    if (quantity == OFEA_POST_STRESS) {
        double lam, mu;
        lame<double>(Eel ? Eel[e] : E, nu, lam, mu);
        const double tr = eps[0] + eps[1] + eps[2];
#pragma unroll
        for (int i = 0; i < 3; ++i) out[i] = lam * tr + 2.0 * mu * eps[i];
#pragma unroll
        for (int i = 3; i < 6; ++i) out[i] = mu * eps[i];
    } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) out[i] = eps[i];
    }
    out[6] = mises_form(out);
}

// Nodal averages of an all-tet mesh in one pass, without element rows in
// memory: eight lanes per node take its incidences k0 + j, k0 + j + 8, ...
// and recompute those tets' values from the packed node records (each tet is
// met by its 4 nodes); the lanes' sums are added in a fixed tree (lanes 4
// apart, then 2, then 1) and lane 0 writes the 7 components [c][node].  Same
// sums as k_post_elem + k_post_node up to the order of the additions.
__global__ __launch_bounds__(POST_WG) void k_post_tet_fused(int nd, const int *__restrict__ inc_ptr,
                                                            const int *__restrict__ inc,
                                                            const int *__restrict__ conn4,
                                                            const double *__restrict__ xu,
                                                            const double *__restrict__ Eel, double E, double nu,
                                                            int quantity, double *__restrict__ nv)
{
    const int t = xcd_remap(blockIdx.x, gridDim.x) * POST_WG + threadIdx.x;
    const int i = t >> 3, j = t & 7;
    if (i >= nd) return;  // whole 8-lane groups leave together
    const int k0 = inc_ptr[i], k1 = inc_ptr[i + 1];
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = k0 + j; k < k1; k += 8) {
        double v[7];
        tet_values(inc[k], conn4, xu, Eel, E, nu, quantity, v);
#pragma unroll
        for (int c = 0; c < 7; ++c) acc[c] += v[c];
    }
#pragma unroll
    for (int o = 4; o > 0; o >>= 1)
#pragma unroll
        for (int c = 0; c < 7; ++c) acc[c] += __shfl_down(acc[c], o, 8);
    if (j == 0) {
#pragma unroll
        for (int c = 0; c < 7; ++c) nv[(size_t)c * nd + i] = k1 > k0 ? acc[c] / (double)(k1 - k0) : 0.0;
    }
}

// Nodal averages of an all-tet mesh over the chunk plan.  One workgroup per
// chunk:
//   staging: lane i < PC_NTAB gathers node i's coordinates and displacement
//            (the resident free-DOF solution through the plan's free rank, 0
//            at clamped nodes, or the caller's full vector) into LDS;
//   elements: lane e computes the 7 values of chunk element e into LDS rows;
//   nodes:   eight lanes per owned node (owned nodes are the first slots of
//            the node table) take the node's incident elements (chunk slots,
//            ascending, staged in LDS) j, j + 8, ..., and add the lanes' sums
//            in a fixed tree; the average divides by the incidence count.
// Each tet is computed once per chunk it touches; no per-node records in
// memory.
__global__ __launch_bounds__(POST_WG) void k_post_chunks(int nchunks, const int *__restrict__ hdr,
                                                         const int *__restrict__ cnode, const int *__restrict__ resort,
                                                         const uint32_t *__restrict__ celoc,
                                                         const int *__restrict__ ceid,
                                                         const uint16_t *__restrict__ cist,
                                                         const uint8_t *__restrict__ cinc, const double *__restrict__ xyz,
                                                         const double *__restrict__ u, int resident,
                                                         const double *__restrict__ Eel, double E, double nu,
                                                         int quantity, int nd, double *__restrict__ nv)
{
    __shared__ double s_n[6][PC_NTAB];        // x, y, z, ux, uy, uz of the chunk's nodes
    __shared__ double s_v[7][PC_ECAP + 1];    // element values
    __shared__ uint8_t s_inc[PC_ICAP];        // the owned nodes' incident elements (chunk slots)
    __shared__ uint16_t s_ist[PC_NOWN + 1];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int t = threadIdx.x;
    const int4 h = reinterpret_cast<const int4 *>(hdr)[chunk];
    const int n0 = h.x, nown = h.y, nel = h.z;
    const int node = t < PC_NTAB ? cnode[(size_t)chunk * PC_NTAB + t] : -1;
    const uint32_t el4 = t < PC_ECAP ? celoc[(size_t)chunk * PC_ECAP + t] : ~0u;
    const uint16_t ist = t <= PC_NOWN ? cist[(size_t)chunk * (PC_NOWN + 1) + t] : (uint16_t)0;
    const int rank = (resident && node >= 0) ? resort[node] : -1;
    const int ei = (Eel && t < nel) ? ceid[(size_t)chunk * PC_ECAP + t] : -1;
    double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (node >= 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] = xyz[3 * (size_t)node + c];
        if (resident) {
            if (rank >= 0)
#pragma unroll
                for (int c = 0; c < 3; ++c) r[3 + c] = u[3 * (size_t)rank + c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) r[3 + c] = u[3 * (size_t)node + c];
        }
    }
    double Ee = E;
    if (ei >= 0) Ee = Eel[ei];
    if (t < PC_NTAB)
#pragma unroll
        for (int c = 0; c < 6; ++c) s_n[c][t] = r[c];
    if (t <= PC_NOWN) s_ist[t] = ist;
    __syncthreads();
    const int ninc = s_ist[nown];
    for (int k = t; k < ninc; k += POST_WG) s_inc[k] = cinc[(size_t)h.w + k];
    if (t < nel) {
        double X[12], U[12];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int sl = (el4 >> (8 * a)) & 255;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                X[3 * a + c] = s_n[c][sl];
                U[3 * a + c] = s_n[3 + c][sl];
            }
        }
        double g[3][4];
        tet_grads(X, g);
        double eps[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        strain_of<4>(g, U, 1.0, eps);
        double out[7];
        if (quantity == OFEA_POST_STRESS) {
            double lam, mu;
            lame<double>(Ee, nu, lam, mu);
            const double tr = eps[0] + eps[1] + eps[2];
#pragma unroll
            for (int i = 0; i < 3; ++i) out[i] = lam * tr + 2.0 * mu * eps[i];
#pragma unroll
            for (int i = 3; i < 6; ++i) out[i] = mu * eps[i];
        } else {
#pragma unroll
            for (int i = 0; i < 6; ++i) out[i] = eps[i];
        }
        out[6] = mises_form(out);
#pragma unroll
        for (int c = 0; c < 7; ++c) s_v[c][t] = out[c];
    }
    __syncthreads();
    const int s = t >> 3, l = t & 7;  // owned node slot s, lane l
    if (s >= nown) return;  // whole 8-lane groups leave together
    const int k0 = s_ist[s], k1 = s_ist[s + 1], cnt = k1 - k0;
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = k0 + l; k < k1; k += 8) {
        const int j = s_inc[k];
#pragma unroll
        for (int c = 0; c < 7; ++c) acc[c] += s_v[c][j];
    }
#pragma unroll
    for (int o = 4; o > 0; o >>= 1)
#pragma unroll
        for (int c = 0; c < 7; ++c) acc[c] += __shfl_down(acc[c], o, 8);
    if (l == 0) {
        const int i = n0 + s;
#pragma unroll
        for (int c = 0; c < 7; ++c) nv[(size_t)c * nd + i] = cnt > 0 ? acc[c] / (double)cnt : 0.0;
    }
}

// Eight lanes per node, lane c < 7 owns component c: the lanes of a node read
// one 64-byte element row per incidence, in ascending element order.  Output
// component-major [c][node] (the layout the caller receives).
__global__ __launch_bounds__(POST_WG) void k_post_node(int nd, const int *__restrict__ inc_ptr,
                                                       const int *__restrict__ inc, const double *__restrict__ ev,
                                                       double *__restrict__ nv)
{
    const int t = xcd_remap(blockIdx.x, gridDim.x) * POST_WG + threadIdx.x;
    const int i = t >> 3, c = t & 7;
    if (i >= nd) return;
    const int k0 = inc_ptr[i], k1 = inc_ptr[i + 1];
    double acc = 0.0;
#pragma unroll 4
    for (int k = k0; k < k1; ++k) acc += ev[8 * (size_t)inc[k] + c];
    if (c < 7) nv[(size_t)c * nd + i] = k1 > k0 ? acc / (double)(k1 - k0) : 0.0;
}

// element rows [e][8] -> component-major [c][e] for the caller
__global__ __launch_bounds__(POST_WG) void k_post_unpack(int ne, const double *__restrict__ ev, double *__restrict__ out)
{
    const int t = blockIdx.x * POST_WG + threadIdx.x;
    if (t >= 7 * ne) return;
    const int c = t / ne, e = t - c * ne;
    out[t] = ev[8 * (size_t)e + c];
}

void free_ws(PostWs *w)
{
    if (!w) return;
    dev_free(w->conn);
    dev_free(w->rank8);
    dev_free(w->etype);
    dev_free(w->inc_ptr);
    dev_free(w->inc);
    dev_free(w->E);
    dev_free(w->u);
    dev_free(w->ev);
    dev_free(w->ev7);
    dev_free(w->nv);
    dev_free(w->xu);
    dev_free(w->ch_hdr);
    dev_free(w->ch_node);
    dev_free(w->ch_ist);
    dev_free(w->ch_inc);
    dev_free(w->ch_eloc);
    dev_free(w->ch_eid);
    delete w;
}

bool build_chunks(ofea_ctx *c, PostWs *w, const std::vector<int> &conn0, const std::vector<int> &ptr,
                  const std::vector<int> &inc);

int build(ofea_ctx *c, PostWs *w)
{
    w->nd = c->nd;
    w->ne = c->ne;
    const int nd = c->nd, ne = c->ne;
    std::vector<int> conn0(8 * (size_t)ne, 0), ptr(nd + 1, 0);
    for (int e = 0; e < ne; ++e) {
        const int nn = c->etype[e] == OFEA_C3D8 ? 8 : 4;
        for (int a = 0; a < nn; ++a) {
            const int node = c->conn[8 * (size_t)e + a] - 1;
            if (node < 0 || node >= nd) return fail(OFEA_ERR_ARG, "element node out of range");
            conn0[8 * (size_t)e + a] = node;
            ptr[node + 1]++;
        }
    }
    std::vector<int> rank8;
    for (int e = 0; e < ne; ++e) {
        int *row = conn0.data() + 8 * (size_t)e;
        if (c->etype[e] == OFEA_C3D4) {
            for (int a = 0; a < 4; ++a) row[4 + a] = c->resort[row[a]];
        } else {
            if (rank8.empty()) rank8.assign(8 * (size_t)ne, -1);
            for (int a = 0; a < 8; ++a) rank8[8 * (size_t)e + a] = c->resort[row[a]];
        }
    }
    for (int i = 0; i < nd; ++i) ptr[i + 1] += ptr[i];
    std::vector<int> inc(ptr[nd]), fill(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < ne; ++e) {  // ascending e: the reference's accumulation order per node
        const int nn = c->etype[e] == OFEA_C3D8 ? 8 : 4;
        for (int a = 0; a < nn; ++a) inc[fill[conn0[8 * (size_t)e + a]]++] = e;
    }
    int rc;
    if ((rc = dev_upload(w->conn, conn0.data(), conn0.size()))) return rc;
    if (!rank8.empty() && (rc = dev_upload(w->rank8, rank8.data(), rank8.size()))) return rc;
    if ((rc = dev_upload(w->etype, c->etype.data(), c->etype.size()))) return rc;
    if ((rc = dev_upload(w->inc_ptr, ptr.data(), ptr.size()))) return rc;
    if ((rc = dev_upload(w->inc, inc.data(), std::max<size_t>(inc.size(), 1)))) return rc;
    if ((rc = dev_upload<double>(w->u, nullptr, 3 * (size_t)nd))) return rc;
    if ((rc = dev_upload<double>(w->ev, nullptr, 8 * (size_t)ne))) return rc;
    if ((rc = dev_upload<double>(w->nv, nullptr, 7 * (size_t)nd))) return rc;
    w->nchunks = 0;
    if (!c->has_type[OFEA_C3D8] && nd > 0 && !build_chunks(c, w, conn0, ptr, inc)) w->nchunks = 0;
    return OFEA_OK;
}

// Chunk plan of k_post_chunks (all-tet meshes): chunks of consecutive nodes
// grown while their distinct elements, touched nodes and owned nodes fit.
// Returns false (the fused per-node pass is kept) when one node alone does
// not fit.
bool build_chunks(ofea_ctx *c, PostWs *w, const std::vector<int> &conn0, const std::vector<int> &ptr,
                  const std::vector<int> &inc)
{
    const int nd = c->nd;
    std::vector<int> hdr, cnode, ceid;
    std::vector<uint32_t> celoc;
    std::vector<uint16_t> cist;
    std::vector<uint8_t> cinc;
    std::vector<int> eslot(c->ne, -1), nslot(nd, -1);
    std::vector<int> els, others;
    int i = 0;
    while (i < nd) {
        const int n0 = i;
        els.clear();
        others.clear();
        int ntouched = 0;
        int nincid = 0;
        while (i < nd && i - n0 < PC_NOWN) {
            // what node i adds: its elements not yet in the chunk and their nodes
            int add_e = 0, add_n = nslot[i] < 0 ? 1 : 0;
            std::vector<int> new_nodes;
            for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
                const int e = inc[k];
                if (eslot[e] >= 0) continue;
                eslot[e] = -2;
                ++add_e;
                for (int a = 0; a < 4; ++a) {
                    const int v = conn0[8 * (size_t)e + a];
                    if (v != i && nslot[v] == -1) nslot[v] = -3, new_nodes.push_back(v);
                }
            }
            const bool fits = (int)els.size() + add_e <= PC_ECAP &&
                              ntouched + add_n + (int)new_nodes.size() <= PC_NTAB &&
                              nincid + (ptr[i + 1] - ptr[i]) <= PC_ICAP;
            for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
                const int e = inc[k];
                if (eslot[e] != -2) continue;
                eslot[e] = fits ? 0 : -1;
                if (fits) els.push_back(e);
            }
            for (int v : new_nodes) nslot[v] = fits ? 0 : -1;
            if (!fits) {
                if (i == n0) {
                    for (int e : els) eslot[e] = -1;
                    for (int v = 0; v < nd; ++v) nslot[v] = -1;
                    return false;
                }
                break;
            }
            if (nslot[i] < 0) ++ntouched;  // an owned node that no earlier element touched
            nslot[i] = 0;
            for (int v : new_nodes) others.push_back(v);
            ntouched += (int)new_nodes.size();
            nincid += ptr[i + 1] - ptr[i];
            ++i;
        }
        // node table: owned nodes n0..i-1 first, then the other touched nodes ascending
        std::sort(els.begin(), els.end());
        std::vector<int> tab;
        for (int v = n0; v < i; ++v) tab.push_back(v);
        std::sort(others.begin(), others.end());
        for (int v : others)
            if (v < n0 || v >= i) tab.push_back(v);
        if ((int)tab.size() > PC_NTAB) {  // cannot happen: counted above
            return false;
        }
        for (int k = 0; k < (int)tab.size(); ++k) nslot[tab[k]] = k;
        hdr.insert(hdr.end(), {n0, i - n0, (int)els.size(), (int)cinc.size()});
        const size_t nb = cnode.size(), eb = celoc.size();
        cnode.resize(nb + PC_NTAB, -1);
        celoc.resize(eb + PC_ECAP, ~0u);
        ceid.resize(eb + PC_ECAP, -1);
        for (int k = 0; k < (int)tab.size(); ++k) cnode[nb + k] = tab[k];
        for (int k = 0; k < (int)els.size(); ++k) eslot[els[k]] = k;
        // owned nodes' incidences as chunk element slots (inc is ascending per node)
        const size_t ib = cist.size();
        cist.resize(ib + PC_NOWN + 1, 0);
        int acc = 0;
        for (int v = n0; v < i; ++v) {
            cist[ib + (v - n0)] = (uint16_t)acc;
            for (int k = ptr[v]; k < ptr[v + 1]; ++k) cinc.push_back((uint8_t)eslot[inc[k]]);
            acc += ptr[v + 1] - ptr[v];
        }
        for (int v = i - n0; v <= PC_NOWN; ++v) cist[ib + v] = (uint16_t)acc;
        for (int k = 0; k < (int)els.size(); ++k) {
            uint32_t wv = 0;
            for (int a = 0; a < 4; ++a) wv |= (uint32_t)nslot[conn0[8 * (size_t)els[k] + a]] << (8 * a);
            celoc[eb + k] = wv;
            ceid[eb + k] = els[k];
        }
        for (int e : els) eslot[e] = -1;
        for (int v : tab) nslot[v] = -1;
        w->nchunks++;
    }
    int rc;
    if ((rc = dev_upload(w->ch_hdr, hdr.data(), hdr.size()))) return false;
    if ((rc = dev_upload(w->ch_node, cnode.data(), cnode.size()))) return false;
    if ((rc = dev_upload(w->ch_ist, cist.data(), cist.size()))) return false;
    if ((rc = dev_upload(w->ch_inc, cinc.data(), std::max<size_t>(cinc.size(), 1)))) return false;
    if ((rc = dev_upload(w->ch_eloc, celoc.data(), celoc.size()))) return false;
    if ((rc = dev_upload(w->ch_eid, ceid.data(), ceid.size()))) return false;
    return true;
}

int ensure(ofea_ctx *c, PostWs *&w)
{
    w = (PostWs *)c->post_ws;
    if (w) return OFEA_OK;
    w = new PostWs();
    int rc = build(c, w);
    if (rc) {
        free_ws(w);
        w = nullptr;
        return rc;
    }
    c->post_ws = w;
    return OFEA_OK;
}

}  // namespace

void post_free(ofea_ctx *c)
{
    free_ws((PostWs *)c->post_ws);
    c->post_ws = nullptr;
}

int post_nodal(ofea_ctx *c, double E, double nu, const double *u_full, int quantity, double *elem_out,
               double *node_out)
{
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    if (!u_full && !c->d_x) return fail(OFEA_ERR_STATE, "no displacement given and no resident solution");
    if (quantity != OFEA_POST_STRAIN && quantity != OFEA_POST_STRESS)
        return fail(OFEA_ERR_ARG, "quantity must be OFEA_POST_STRAIN or OFEA_POST_STRESS");
    if (quantity == OFEA_POST_STRESS && !(nu > -1.0 && nu < 0.5))
        return fail(OFEA_ERR_ARG, "Poisson ratio must lie in (-1, 0.5)");
    PostWs *w = nullptr;
    int rc = ensure(c, w);
    if (rc) return rc;
    // per-element moduli follow ofea_set_elem_E
    dev_free(w->E);
    if (!c->E_elem.empty() && (rc = dev_upload(w->E, c->E_elem.data(), c->E_elem.size()))) return rc;
    const int nd = c->nd, ne = c->ne;
    const dim3 ge((ne + POST_WG - 1) / POST_WG), gn((8 * nd + POST_WG - 1) / POST_WG), blk(POST_WG);
    const bool resident = u_full == nullptr;  // else the context's free-DOF solution, read through the free ranks
    const double *uu = resident ? c->d_x : w->u;
    if (!resident) OFEA_HIP(hipMemcpyAsync(w->u, u_full, sizeof(double) * 3 * nd, hipMemcpyHostToDevice, c->stream));
    // all-tet meshes: nodal values in one pass (element rows only when asked for);
    // with hexes (8 Gauss points each) the element pass runs once and its rows are summed
    const bool fused = !c->has_type[OFEA_C3D8];
    if (!fused || elem_out) {
        if (resident)
            hipLaunchKernelGGL(k_post_elem<true>, ge, blk, 0, c->stream, ne, w->conn, w->rank8, w->etype, c->d_xyz,
                               uu, w->E, E, nu, quantity, w->ev);
        else
            hipLaunchKernelGGL(k_post_elem<false>, ge, blk, 0, c->stream, ne, w->conn, w->rank8, w->etype, c->d_xyz,
                               uu, w->E, E, nu, quantity, w->ev);
        OFEA_HIP(hipGetLastError());
    }
    if (fused && w->nchunks > 0) {
        hipLaunchKernelGGL(k_post_chunks, dim3(w->nchunks), blk, 0, c->stream, w->nchunks, w->ch_hdr, w->ch_node,
                           c->d_resort, w->ch_eloc, w->ch_eid, w->ch_ist, w->ch_inc, c->d_xyz, uu, resident ? 1 : 0,
                           w->E, E, nu, quantity, nd, w->nv);
    } else if (fused) {
        if (!w->xu && (rc = dev_upload<double>(w->xu, nullptr, 6 * (size_t)nd))) return rc;
        const dim3 gp((nd + POST_WG - 1) / POST_WG);
        if (resident)
            hipLaunchKernelGGL(k_post_pack<true>, gp, blk, 0, c->stream, nd, c->d_xyz, uu, c->d_resort, w->xu);
        else
            hipLaunchKernelGGL(k_post_pack<false>, gp, blk, 0, c->stream, nd, c->d_xyz, uu, c->d_resort, w->xu);
        // all tets: the type-local connectivity of the assembly plans is in global element order
        hipLaunchKernelGGL(k_post_tet_fused, gn, blk, 0, c->stream, nd, w->inc_ptr, w->inc, c->dp[OFEA_C3D4].conn,
                           w->xu, w->E, E, nu, quantity, w->nv);
    } else {
        hipLaunchKernelGGL(k_post_node, gn, blk, 0, c->stream, nd, w->inc_ptr, w->inc, w->ev, w->nv);
    }
    OFEA_HIP(hipGetLastError());
    if (elem_out) {
        if (!w->ev7 && (rc = dev_upload<double>(w->ev7, nullptr, 7 * (size_t)ne))) return rc;
        hipLaunchKernelGGL(k_post_unpack, dim3((7 * ne + POST_WG - 1) / POST_WG), blk, 0, c->stream, ne, w->ev,
                           w->ev7);
        OFEA_HIP(hipGetLastError());
        OFEA_HIP(hipMemcpyAsync(elem_out, w->ev7, sizeof(double) * 7 * ne, hipMemcpyDeviceToHost, c->stream));
    }
    if (node_out)
        OFEA_HIP(hipMemcpyAsync(node_out, w->nv, sizeof(double) * 7 * nd, hipMemcpyDeviceToHost, c->stream));
    if (elem_out || node_out) OFEA_HIP(hipStreamSynchronize(c->stream));
    return OFEA_OK;
}

}  // namespace ofea
