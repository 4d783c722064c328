// sfem.hip -- node-based smoothed FEM (C3D4) on the device.
//
// Replaces SFEM3D::build / SFEM_C3D4::build_ele_stiff_mat / shapeGradient
// (src/SFEM3D.cpp:8-163, src/Elements/SFEM_C3D4.cpp:53-427) and
// assamble_stiffness::SFEM_build_CSR / SFEM_fill_CSR_sparse_mat
// (src/assemble.cpp:291-414), driven as CAE_process::implict_SFEManalysis
// (src/cae.cpp:379-416).
//
// Smoothing domain k = the cell of node k: in each tet around k, the part
// bounded by three interior facets (edge midpoint, two face centroids, the
// tet centroid) plus, where the tet's face through k lies on the surface,
// the corner quad of that face.  Per member i of ring(k) (the nodes of the
// tets around k):
//   G_i  = (1 / V_k) sum over the cell boundary of  int N_i n dA
//   GG_i = l (1 / V_k) sum of  int grad N_i (x) n dA,  l = (3 V_k / (4 pi))^(1/3)
// with V_k = sum |V_e| / 4 and n pointing out of the cell.  The reference's
// six-term loop, sum over m of (V_k/6) B(G + s GG[:, in])^T D B(G + s GG[:, in])
// (s = -1 for m < 3, +1 otherwise, in = m mod 3), expands to
//   K_k = V_k K(G, G) + (V_k / 3) sum_in K(a_in, a_in),  a_in = GG[:, in],
//   K(x, y)_ij[c][d] = lambda x_i[c] y_j[d] + mu x_i[d] y_j[c] + mu delta_cd x_i . y_j,
// the form used here (with M the sum of the outer products of the scaled
// rows, K = lambda M + mu M^T + mu tr(M) I).  Facet integrals: each facet is
// a planar convex quad; N_i is linear on it, so two triangles' vector areas
// times mean corner values integrate it exactly.
//
// Kernels:
//   k_sfem_grad   one workgroup per chunk of domains: lane = (domain,
//                 element) integrates the element's facets of the cell; lane
//                 = (domain, member) then sums them in ascending element
//                 order and writes a 12-double row {G sqrt(V), a_in sqrt(V/3)}
//   k_sfem_gather one lane per 3x3 block (r, s) of the pattern sums its
//                 contributors (domains containing both r and s, ascending
//                 domain: the reference's summation order) and writes the 9
//                 CSC values once; no atomics.
//   k_sfem_gather_chunks  the same sums over a chunk plan: the rows of the
//                 domains around a chunk of spatially compact columns are
//                 staged in LDS once (streaming loads, so they do not evict
//                 the partly written value lines from L2), then one lane per
//                 block reads its contributor pairs from LDS; the plain
//                 gather serves meshes whose rings exceed the LDS.
#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <vector>

#include "elem_device.h"
#include "ofea_internal.h"

namespace ofea {
namespace {

constexpr int SF_PAIRS = 256;  // (domain, element) pairs and members per workgroup
constexpr int SF_MAXE = 96;    // tets around one node
constexpr int SF_MAXM = 128;   // nodes in one node's ring

struct SfemWs {
    int nd = 0, ndom = 0;
    long long nrows = 0;            // total (domain, member) rows
    int *conn = nullptr;            // ne x 4, 0-based
    unsigned char *fflag = nullptr; // ne: bit f = face f on the surface (faces 123, 124, 234, 134)
    int *dom_mptr = nullptr;        // ndom + 1 into members
    int *dom_member = nullptr;      // member node ids (ascending per domain)
    int *dom_eptr = nullptr;        // ndom + 1 into elements
    int *dom_elem = nullptr;        // incident elements (ascending)
    double *vol = nullptr;          // ndom: V_k
    double *rows = nullptr;         // nrows x 12 (scaled G and a_0..a_2)
    int *chunk_dom = nullptr;       // nchunks + 1
    int nchunks = 0;
    int *mdom = nullptr;            // domain of each member row
    int *pdom = nullptr;            // domain of each (domain, element) pair
    int *cptr = nullptr;            // blocks + 1
    int2 *contrib = nullptr;        // (row of r, row of s) per contributor
    long long ncontrib = 0;
    int *blk_r = nullptr;           // column node of each block
    // chunk plan of k_sfem_gather_chunks (gnchunks = 0: the plain gather)
    int gnchunks = 0;
    int *g_col_ptr = nullptr, *g_cols = nullptr, *g_dom_ptr = nullptr, *g_blk_ptr = nullptr, *g_cptr = nullptr;
    int2 *g_doms = nullptr;
    uint32_t *g_con = nullptr;
};

__device__ inline void d_cross(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Adds int_facet N_j n dA (gsum) and int_facet n dA (nsum) of one planar quad
// facet with corners P[4] (cyclic order); N_j at the corners is phi[c][j]; n
// is oriented away from the point ref.  N_j is linear, so its integral over
// each triangle of the split (0, 1, 2) + (0, 2, 3) is the triangle's vector
// area times the mean corner value: exact, as the reference's 2x2 Gauss rule
// on the bilinear map is for a planar quad.
__device__ inline void facet(const double P[4][3], const double phi[4][4], const double ref[3], double gsum[4][3],
                             double nsum[3])
{
    double e1[3], e2[3], e3[3], S1[3], S2[3];
    for (int c = 0; c < 3; ++c) {
        e1[c] = P[1][c] - P[0][c];
        e2[c] = P[2][c] - P[0][c];
        e3[c] = P[3][c] - P[0][c];
    }
    d_cross(e1, e2, S1);  // twice the vector areas
    d_cross(e2, e3, S2);
    double dir = 0.0;
    for (int c = 0; c < 3; ++c) dir += (S1[c] + S2[c]) * (0.25 * (P[0][c] + P[1][c] + P[2][c] + P[3][c]) - ref[c]);
    const double h = dir < 0.0 ? -0.5 : 0.5;
    for (int c = 0; c < 3; ++c) {
        S1[c] *= h;
        S2[c] *= h;
        nsum[c] += S1[c] + S2[c];
    }
    const double t = 1.0 / 3.0;
    for (int j = 0; j < 4; ++j) {
        const double w1 = t * (phi[0][j] + phi[1][j] + phi[2][j]), w2 = t * (phi[0][j] + phi[2][j] + phi[3][j]);
        for (int c = 0; c < 3; ++c) gsum[j][c] += w1 * S1[c] + w2 * S2[c];
    }
}

// Contribution of one tet (corners x[4], rotated so that corner 0 is the
// domain node) to the 4 corners: gsum[j] = int N_j n over the cell's facets
// in this tet, nsum = int n (for GG = grad N_j (x) nsum).  bflag bit b: the
// face opposite rotated corner b+1 (corners {0, .., } without b+1) is a
// surface face.
__device__ inline void tet_cell(const double x[4][3], int bflag, double gsum[4][3], double nsum[3])
{
    double m1[3], m2[3], m3[3], c1[3], c2[3], c3[3], cc[3];
    for (int c = 0; c < 3; ++c) {
        m1[c] = 0.5 * (x[0][c] + x[1][c]);
        m2[c] = 0.5 * (x[0][c] + x[2][c]);
        m3[c] = 0.5 * (x[0][c] + x[3][c]);
        c1[c] = (x[0][c] + x[1][c] + x[2][c]) / 3.0;  // face 0-1-2
        c2[c] = (x[0][c] + x[2][c] + x[3][c]) / 3.0;  // face 0-2-3
        c3[c] = (x[0][c] + x[1][c] + x[3][c]) / 3.0;  // face 0-1-3
        cc[c] = 0.25 * (x[0][c] + x[1][c] + x[2][c] + x[3][c]);
    }
    const double t = 1.0 / 3.0, h = 0.5, f = 0.25;
    // interior facets: (c3, m1, c1, cc), (c1, m2, c2, cc), (c2, m3, c3, cc), outward from corner 0
    {
        double P[4][3];
        for (int c = 0; c < 3; ++c) { P[0][c] = c3[c]; P[1][c] = m1[c]; P[2][c] = c1[c]; P[3][c] = cc[c]; }
        const double phi[4][4] = {{t, t, 0, t}, {h, h, 0, 0}, {t, t, t, 0}, {f, f, f, f}};
        facet(P, phi, x[0], gsum, nsum);
    }
    {
        double P[4][3];
        for (int c = 0; c < 3; ++c) { P[0][c] = c1[c]; P[1][c] = m2[c]; P[2][c] = c2[c]; P[3][c] = cc[c]; }
        const double phi[4][4] = {{t, t, t, 0}, {h, 0, h, 0}, {t, 0, t, t}, {f, f, f, f}};
        facet(P, phi, x[0], gsum, nsum);
    }
    {
        double P[4][3];
        for (int c = 0; c < 3; ++c) { P[0][c] = c2[c]; P[1][c] = m3[c]; P[2][c] = c3[c]; P[3][c] = cc[c]; }
        const double phi[4][4] = {{t, 0, t, t}, {h, 0, 0, h}, {t, t, 0, t}, {f, f, f, f}};
        facet(P, phi, x[0], gsum, nsum);
    }
    // surface corner quads of corner 0, outward from the tet centroid
    if (bflag & 1) {  // face 0-2-3 (opposite corner 1): corner 0, mid 0-2, centroid, mid 0-3
        double P[4][3];
        for (int c = 0; c < 3; ++c) { P[0][c] = m2[c]; P[1][c] = c2[c]; P[2][c] = m3[c]; P[3][c] = x[0][c]; }
        const double phi[4][4] = {{h, 0, h, 0}, {t, 0, t, t}, {h, 0, 0, h}, {1, 0, 0, 0}};
        facet(P, phi, cc, gsum, nsum);
    }
    if (bflag & 4) {  // face 0-1-2 (opposite corner 3)
        double P[4][3];
        for (int c = 0; c < 3; ++c) { P[0][c] = m1[c]; P[1][c] = c1[c]; P[2][c] = m2[c]; P[3][c] = x[0][c]; }
        const double phi[4][4] = {{h, h, 0, 0}, {t, t, t, 0}, {h, 0, h, 0}, {1, 0, 0, 0}};
        facet(P, phi, cc, gsum, nsum);
    }
    if (bflag & 2) {  // face 0-1-3 (opposite corner 2)
        double P[4][3];
        for (int c = 0; c < 3; ++c) { P[0][c] = m3[c]; P[1][c] = c3[c]; P[2][c] = m1[c]; P[3][c] = x[0][c]; }
        const double phi[4][4] = {{h, 0, 0, h}, {t, t, 0, t}, {h, h, 0, 0}, {1, 0, 0, 0}};
        facet(P, phi, cc, gsum, nsum);
    }
}

// face index (faces 123, 124, 234, 134 of the element's own node order) of
// the face opposite local node p
__device__ inline int face_opposite(int p) { return p == 0 ? 2 : (p == 1 ? 3 : (p == 2 ? 1 : 0)); }

// One workgroup per chunk of consecutive domains (<= SF_PAIRS (domain,
// element) pairs and <= SF_PAIRS members).  Phase 1, lane = pair: the
// element's facets of the domain's cell, per rotated corner j the integral
// of N_j n, the integral of n and grad N_j, into LDS.  Phase 2, lane =
// (domain, member): the member's entries summed in ascending element order,
// scaled, written as a 12-double row {G sqrt(V), a_0 sqrt(V/3), a_1 .., a_2 ..}
// so that the block gather needs no volume.
__global__ __launch_bounds__(SF_PAIRS) void k_sfem_grad(int nchunks, const int *__restrict__ chunk_dom,
                                                        const double *__restrict__ xyz, const int *__restrict__ conn,
                                                        const unsigned char *__restrict__ fflag,
                                                        const int *__restrict__ mptr, const int *__restrict__ member,
                                                        const int *__restrict__ mdom, const int *__restrict__ eptr,
                                                        const int *__restrict__ elem, const int *__restrict__ pdom,
                                                        const double *__restrict__ vol, double *__restrict__ rows)
{
    __shared__ double s_g[SF_PAIRS][4][3];
    __shared__ double s_t[SF_PAIRS][4][3];  // grad N of the element's rotated corners
    __shared__ double s_n[SF_PAIRS][3];
    __shared__ int s_node[SF_PAIRS][4];
    const int chunk = xcd_remap(blockIdx.x, nchunks);
    const int d0 = chunk_dom[chunk], d1 = chunk_dom[chunk + 1];
    const int p0 = eptr[d0], p1 = eptr[d1];
    const int p = p0 + (int)threadIdx.x;
    if (p < p1) {
        const int k = pdom[p], e = elem[p], q = p - p0;
        const int *cn = conn + 4 * (size_t)e;
        int pos = 0;
        for (int a = 1; a < 4; ++a)
            if (cn[a] == k) pos = a;
        double x[4][3], X[12];
        int nodes[4];
        for (int j = 0; j < 4; ++j) {
            nodes[j] = cn[(pos + j) & 3];
            for (int c = 0; c < 3; ++c) x[j][c] = X[3 * j + c] = xyz[3 * nodes[j] + c];
        }
        // surface flags of the faces opposite rotated corners 1, 3, 2
        const int ff = fflag[e];
        int b = 0;
        if (ff >> face_opposite((pos + 1) & 3) & 1) b |= 1;
        if (ff >> face_opposite((pos + 2) & 3) & 1) b |= 2;
        if (ff >> face_opposite((pos + 3) & 3) & 1) b |= 4;
        double gsum[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, nsum[3] = {0, 0, 0};
        tet_cell(x, b, gsum, nsum);
        double g[3][4];
        tet_grads(X, g);
        for (int j = 0; j < 4; ++j) {
            s_node[q][j] = nodes[j];
            for (int c = 0; c < 3; ++c) {
                s_g[q][j][c] = gsum[j][c];
                s_t[q][j][c] = g[c][j];
            }
        }
        for (int c = 0; c < 3; ++c) s_n[q][c] = nsum[c];
    }
    __syncthreads();
    const int m = mptr[d0] + (int)threadIdx.x;
    if (m >= mptr[d1]) return;
    const int k = mdom[m], node = member[m];
    const double V = vol[k];
    double G[3] = {0, 0, 0}, GG[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int q = eptr[k] - p0; q < eptr[k + 1] - p0; ++q)
        for (int j = 0; j < 4; ++j)
            if (s_node[q][j] == node) {
                for (int c = 0; c < 3; ++c) G[c] += s_g[q][j][c];
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) GG[a][c] += s_t[q][j][a] * s_n[q][c];
            }
    const double l = cbrt(3.0 * V / (4.0 * 3.14159265358979323846));
    const double iv = 1.0 / V, sg = sqrt(V), sa = sqrt(V / 3.0);
    double *row = rows + 12 * (size_t)m;
    for (int c = 0; c < 3; ++c) row[c] = G[c] * iv * sg;
    for (int in = 0; in < 3; ++in)
        for (int a = 0; a < 3; ++a) row[3 + 3 * in + a] = GG[a][in] * iv * l * sa;
}

// K_rs = lambda M + mu M^T + mu tr(M) I with M = sum over the block's
// domains of the outer products of the scaled rows (G and the three a_in)
__global__ __launch_bounds__(256) void k_sfem_gather(int nblk, const int *__restrict__ blk_r, const int *__restrict__ node_blk,
                                                     const int *__restrict__ col_ptr, const int *__restrict__ cptr,
                                                     const int2 *__restrict__ contrib, const double *__restrict__ rows,
                                                     double lam, double mu, double *__restrict__ val)
{
    const int b = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (b >= nblk) return;
    const int r = blk_r[b];
    const int k = b - node_blk[r];
    double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // contributors two at a time: both pairs and then all four rows are
    // loaded before the sums, which keep the ascending order
    auto load_rows = [&](int2 pr, double x[12], double y[12]) {
        const double2 *x2 = reinterpret_cast<const double2 *>(rows + 12 * (size_t)pr.x);
        const double2 *y2 = reinterpret_cast<const double2 *>(rows + 12 * (size_t)pr.y);
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double2 u = x2[i], v = y2[i];
            x[2 * i] = u.x;
            x[2 * i + 1] = u.y;
            y[2 * i] = v.x;
            y[2 * i + 1] = v.y;
        }
    };
    auto add = [&](const double x[12], const double y[12]) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d) M[3 * c + d] += x[3 * v + c] * y[3 * v + d];
    };
    int t = cptr[b];
    const int t1 = cptr[b + 1];
    for (; t + 1 < t1; t += 2) {
        const int2 p0 = contrib[t], p1 = contrib[t + 1];
        double x0[12], y0[12], x1[12], y1[12];
        load_rows(p0, x0, y0);
        load_rows(p1, x1, y1);
        add(x0, y0);
        add(x1, y1);
    }
    if (t < t1) {
        double x0[12], y0[12];
        load_rows(contrib[t], x0, y0);
        add(x0, y0);
    }
    const double tr = M[0] + M[4] + M[8];
    const int base = col_ptr[3 * r], stride = col_ptr[3 * r + 1] - base;
    double *dst = val + base + 3 * k;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) dst[c * stride + d] = lam * M[3 * c + d] + mu * M[3 * d + c] + (c == d ? mu * tr : 0.0);
}

// The same block sums over a chunk plan (built by sfem_setup): one workgroup
// per chunk of spatially compact columns.  The rows of every domain that
// touches a chunk column are staged in LDS first (each domain's rows are
// contiguous in memory: one wave copies a domain at a time), then one lane
// per block of the chunk's columns sums its contributors from LDS -- the
// same pairs in the same ascending-domain order as k_sfem_gather, so the
// values are the same bits.
constexpr int SG_WG = 1024;
constexpr int SG_ROWS = 1600;  // staged rows per chunk (12 doubles each)
constexpr int SG_COLS = 64;   // columns per chunk
__global__ __launch_bounds__(SG_WG) void k_sfem_gather_chunks(int nchunks, const int *__restrict__ ch_col_ptr,
                                                              const int *__restrict__ ch_cols,
                                                              const int *__restrict__ ch_dom_ptr,
                                                              const int2 *__restrict__ ch_doms,
                                                              const int *__restrict__ ch_blk_ptr,
                                                              const int *__restrict__ gcptr,
                                                              const uint32_t *__restrict__ gcon,
                                                              const int *__restrict__ dom_mptr,
                                                              const int *__restrict__ node_blk,
                                                              const int *__restrict__ col_ptr,
                                                              const double *__restrict__ rows, double lam, double mu,
                                                              double *__restrict__ val)
{
    __shared__ __align__(16) double s_rows[SG_ROWS * 12];
    __shared__ int s_grow[SG_ROWS];  // global row of each staged row
    __shared__ int s_cb[SG_COLS + 1], s_col[SG_COLS], s_nrow;
    const int ch = xcd_remap(blockIdx.x, nchunks), t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int c0 = ch_col_ptr[ch], ncol = ch_col_ptr[ch + 1] - c0;
    if (t == 0) {
        s_cb[0] = 0;
        for (int i = 0; i < ncol; ++i) {
            const int r = ch_cols[c0 + i];
            s_col[i] = r;
            s_cb[i + 1] = s_cb[i] + node_blk[r + 1] - node_blk[r];
        }
    }
    // the global row of every staged row (one thread per domain), then one
    // flat copy of 16-byte pieces, eight loads issued before their stores
    const int d0 = ch_dom_ptr[ch], ndom = ch_dom_ptr[ch + 1] - d0;
    for (int i = t; i < ndom; i += SG_WG) {
        const int2 e = ch_doms[d0 + i];  // domain, first LDS row
        const int g0 = dom_mptr[e.x], cnt = dom_mptr[e.x + 1] - g0;
        for (int j = 0; j < cnt; ++j) s_grow[e.y + j] = g0 + j;
        if (i == ndom - 1) s_nrow = e.y + cnt;
    }
    __syncthreads();
    {
        const int npc = 6 * s_nrow;  // 16-byte pieces
        double2 *dst = reinterpret_cast<double2 *>(s_rows);
        constexpr int U = 8;
        for (int i0 = 0; i0 < npc; i0 += U * SG_WG) {
            double2 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * SG_WG + t;
                if (i < npc) {
                    const double *q = rows + 12 * (size_t)s_grow[i / 6] + 2 * (i % 6);
                    v[u].x = __builtin_nontemporal_load(q);
                    v[u].y = __builtin_nontemporal_load(q + 1);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * SG_WG + t;
                if (i < npc) dst[i] = v[u];
            }
        }
    }
    __syncthreads();
    const int nb = s_cb[ncol], gb0 = ch_blk_ptr[ch];
    for (int bl = t; bl < nb; bl += SG_WG) {
        int ci = 0;
        while (s_cb[ci + 1] <= bl) ++ci;
        const int r = s_col[ci], k = bl - s_cb[ci];
        double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = gcptr[gb0 + bl]; q < gcptr[gb0 + bl + 1]; ++q) {
            const uint32_t u = gcon[q];
            const double *x = s_rows + 12 * (u & 0xFFFFu), *y = s_rows + 12 * (u >> 16);
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d) M[3 * c + d] += x[3 * v + c] * y[3 * v + d];
        }
        const double tr = M[0] + M[4] + M[8];
        const int base = col_ptr[3 * r], stride = col_ptr[3 * r + 1] - base;
        double *dst = val + base + 3 * k;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                dst[c * stride + d] = lam * M[3 * c + d] + mu * M[3 * d + c] + (c == d ? mu * tr : 0.0);
    }
}

void free_ws(SfemWs *w)
{
    if (!w) return;
    dev_free(w->conn);
    dev_free(w->fflag);
    dev_free(w->dom_mptr);
    dev_free(w->dom_member);
    dev_free(w->dom_eptr);
    dev_free(w->dom_elem);
    dev_free(w->vol);
    dev_free(w->rows);
    dev_free(w->cptr);
    dev_free(w->contrib);
    dev_free(w->blk_r);
    dev_free(w->chunk_dom);
    dev_free(w->mdom);
    dev_free(w->pdom);
    dev_free(w->g_col_ptr);
    dev_free(w->g_cols);
    dev_free(w->g_dom_ptr);
    dev_free(w->g_blk_ptr);
    dev_free(w->g_cptr);
    dev_free(w->g_doms);
    dev_free(w->g_con);
    delete w;
}

}  // namespace

struct SfemHost {
    std::vector<int> conn;                // ne x 4, 0-based
    std::vector<unsigned char> fflag;     // ne
    std::vector<int> mptr, member, eptr, elem;
    std::vector<double> vol;
};

// Smoothing domains of a C3D4 mesh (SFEM3D::eleData / nodeData,
// src/SFEM3D.cpp:31-134): surface faces are the faces met once; per node its
// tets (ascending), its ring (sorted node ids of those tets) and V_k = sum |V_e| / 4.
int sfem_domains(int nd, int ne, const int *conn8, const double *xyz, SfemHost &H)
{
    H.conn.resize(4 * (size_t)ne);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 4; ++a) {
            const int v = conn8[8 * (size_t)e + a] - 1;
            if (v < 0 || v >= nd) return fail(OFEA_ERR_ARG, "element node out of range");
            H.conn[4 * (size_t)e + a] = v;
        }
    static const int FACES[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};
    std::unordered_map<unsigned long long, int> count;
    count.reserve(4 * (size_t)ne);
    auto key = [&](int e, int f) {
        int v[3] = {H.conn[4 * (size_t)e + FACES[f][0]], H.conn[4 * (size_t)e + FACES[f][1]],
                    H.conn[4 * (size_t)e + FACES[f][2]]};
        std::sort(v, v + 3);
        return ((unsigned long long)v[0] << 42) | ((unsigned long long)v[1] << 21) | (unsigned long long)v[2];
    };
    if (nd >= (1 << 21)) return fail(OFEA_ERR_LIMIT, "too many nodes for the face key");
    for (int e = 0; e < ne; ++e)
        for (int f = 0; f < 4; ++f) count[key(e, f)] ^= 1;  // a face met twice is interior
    H.fflag.assign(ne, 0);
    for (int e = 0; e < ne; ++e)
        for (int f = 0; f < 4; ++f)
            if (count[key(e, f)]) H.fflag[e] |= (unsigned char)(1 << f);
    H.eptr.assign(nd + 1, 0);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 4; ++a) H.eptr[H.conn[4 * (size_t)e + a] + 1]++;
    for (int i = 0; i < nd; ++i) H.eptr[i + 1] += H.eptr[i];
    H.elem.resize(H.eptr[nd]);
    std::vector<int> fill(H.eptr.begin(), H.eptr.end() - 1);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 4; ++a) H.elem[fill[H.conn[4 * (size_t)e + a]]++] = e;
    H.vol.assign(nd, 0.0);
    std::vector<double> ve(ne);
    for (int e = 0; e < ne; ++e) {
        const double *p[4];
        for (int a = 0; a < 4; ++a) p[a] = xyz + 3 * (size_t)H.conn[4 * (size_t)e + a];
        double u[3], v[3], w[3];
        for (int c = 0; c < 3; ++c) {
            u[c] = p[0][c] - p[3][c];
            v[c] = p[1][c] - p[3][c];
            w[c] = p[2][c] - p[3][c];
        }
        ve[e] = std::fabs(u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                          u[2] * (v[0] * w[1] - v[1] * w[0])) / 6.0;
    }
    H.mptr.assign(nd + 1, 0);
    H.member.clear();
    std::vector<int> ring;
    for (int k = 0; k < nd; ++k) {
        ring.clear();
        for (int t = H.eptr[k]; t < H.eptr[k + 1]; ++t) {
            const int e = H.elem[t];
            H.vol[k] += ve[e] / 4.0;
            for (int a = 0; a < 4; ++a) ring.push_back(H.conn[4 * (size_t)e + a]);
        }
        std::sort(ring.begin(), ring.end());
        ring.erase(std::unique(ring.begin(), ring.end()), ring.end());
        if (H.eptr[k + 1] - H.eptr[k] > SF_MAXE) return fail(OFEA_ERR_LIMIT, "a node has more than 96 tets");
        if ((int)ring.size() > SF_MAXM) return fail(OFEA_ERR_LIMIT, "a node ring has more than 128 nodes");
        H.member.insert(H.member.end(), ring.begin(), ring.end());
        H.mptr[k + 1] = (int)H.member.size();
    }
    return OFEA_OK;
}

// SFEM_build_CSR (src/assemble.cpp:291-350): column j holds row i when some
// domain's ring contains both (rows ascending, clamped DOFs dropped), built
// node-block-wise like build_pattern.
int sfem_pattern(int nd, const SfemHost &H, const int *resort, Pattern &P)
{
    P = Pattern();
    int nfree = 0;
    for (int i = 0; i < nd; ++i)
        if (resort[i] >= 0) nfree = std::max(nfree, resort[i] + 1);
    P.nfree = nfree;
    P.n = 3 * nfree;
    std::vector<std::vector<int>> nb(nfree);
    for (int k = 0; k < nd; ++k)
        for (int a = H.mptr[k]; a < H.mptr[k + 1]; ++a) {
            const int r = resort[H.member[a]];
            if (r < 0) continue;
            for (int b = H.mptr[k]; b < H.mptr[k + 1]; ++b) {
                const int s = resort[H.member[b]];
                if (s >= 0) nb[r].push_back(s);
            }
        }
    P.node_blk.assign(nfree + 1, 0);
    for (int r = 0; r < nfree; ++r) {
        auto &v = nb[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        if (v.size() > 1024) return fail(OFEA_ERR_LIMIT, "an SFEM node couples more than 1024 nodes");
        P.node_blk[r + 1] = P.node_blk[r] + (int)v.size();
    }
    const long long nblk = P.node_blk[nfree];
    if (9 * nblk >= (1LL << 31)) return fail(OFEA_ERR_LIMIT, "pattern exceeds int32 indices");
    P.nnz = 9 * nblk;
    P.bcol.resize(nblk);
    P.diag_k.assign(nfree, -1);
    P.col_ptr.assign(P.n + 1, 0);
    P.row_idx.resize(P.nnz);
    for (int r = 0; r < nfree; ++r) {
        const auto &v = nb[r];
        const int deg = (int)v.size();
        std::copy(v.begin(), v.end(), P.bcol.begin() + P.node_blk[r]);
        for (int k = 0; k < deg; ++k)
            if (v[k] == r) P.diag_k[r] = k;
        for (int c = 0; c < 3; ++c) {
            const long long base = 9LL * P.node_blk[r] + (long long)c * 3 * deg;
            P.col_ptr[3 * r + c] = (int)base;
            for (int k = 0; k < deg; ++k)
                for (int d = 0; d < 3; ++d) P.row_idx[base + 3 * k + d] = 3 * v[k] + d;
        }
    }
    P.col_ptr[P.n] = (int)P.nnz;
    return OFEA_OK;
}

void sfem_free(ofea_ctx *c)
{
    free_ws((SfemWs *)c->sfem_ws);
    c->sfem_ws = nullptr;
}

// Chunk plan of k_sfem_gather_chunks: the free columns in Morton order of
// their coordinates, cut greedily into chunks of at most SG_COLS columns whose
// domains (the ring of each column's node) hold at most SG_ROWS rows; per
// chunk its columns, its domains (ascending, with their first LDS row) and,
// per block of its columns, the contributors as (LDS row of r, LDS row of s)
// in the order of the plain gather's list.  Returns false when one column's
// domains alone exceed SG_ROWS (the plain gather then serves the mesh).
struct GatherPlan2 {
    std::vector<int> col_ptr{0}, cols, dom_ptr{0}, blk_ptr{0}, cptr{0};
    std::vector<int2> doms;
    std::vector<uint32_t> con;
};

static bool sfem_chunk_plan(int nd, const double *xyz, const SfemHost &H, const int *resort, const Pattern &P,
                            const std::vector<int> &cnt, const std::vector<int2> &contrib, GatherPlan2 &G)
{
    const int nf = P.nfree;
    std::vector<int> node_of(nf, -1);
    for (int i = 0; i < nd; ++i)
        if (resort[i] >= 0) node_of[resort[i]] = i;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int r = 0; r < nf; ++r)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], xyz[3 * (size_t)node_of[r] + c]);
            hi[c] = std::max(hi[c], xyz[3 * (size_t)node_of[r] + c]);
        }
    auto spread = [](unsigned v) {  // 10 bits -> every third bit
        v &= 1023u;
        v = (v | (v << 16)) & 0x030000FFu;
        v = (v | (v << 8)) & 0x0300F00Fu;
        v = (v | (v << 4)) & 0x030C30C3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    };
    std::vector<std::pair<unsigned, int>> key(nf);
    for (int r = 0; r < nf; ++r) {
        unsigned q[3];
        for (int c = 0; c < 3; ++c) {
            const double ext = hi[c] - lo[c];
            const double u = ext > 0.0 ? (xyz[3 * (size_t)node_of[r] + c] - lo[c]) / ext : 0.0;
            q[c] = (unsigned)std::min(1023.0, std::max(0.0, u * 1023.0));
        }
        key[r] = {spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2), r};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> dom_mark(nd, -1), chunk_doms, chunk_cols, row_dom(H.member.size());
    for (int k = 0; k < nd; ++k)
        for (int a = H.mptr[k]; a < H.mptr[k + 1]; ++a) row_dom[a] = k;
    auto rows_of = [&](int k) { return H.mptr[k + 1] - H.mptr[k]; };
    auto flush = [&]() {
        if (chunk_cols.empty()) return;
        std::sort(chunk_doms.begin(), chunk_doms.end());
        std::vector<int> lrow0(chunk_doms.size());
        int lr = 0;
        for (size_t i = 0; i < chunk_doms.size(); ++i) {
            lrow0[i] = lr;
            G.doms.push_back(make_int2(chunk_doms[i], lr));
            dom_mark[chunk_doms[i]] = (int)i;  // position in the chunk's sorted list
            lr += rows_of(chunk_doms[i]);
        }
        for (int r : chunk_cols) {
            G.cols.push_back(r);
            for (int b = P.node_blk[r]; b < P.node_blk[r + 1]; ++b) {
                for (int q = cnt[b]; q < cnt[b + 1]; ++q) {
                    const int2 pr = contrib[q];  // rows (k, r), (k, s): the same domain k
                    const int kdom = row_dom[pr.x];
                    const int base = lrow0[dom_mark[kdom]] - H.mptr[kdom];
                    G.con.push_back((uint32_t)(base + pr.x) | ((uint32_t)(base + pr.y) << 16));
                }
                G.cptr.push_back((int)G.con.size());
            }
        }
        for (int k : chunk_doms) dom_mark[k] = -1;
        G.col_ptr.push_back((int)G.cols.size());
        G.dom_ptr.push_back((int)G.doms.size());
        G.blk_ptr.push_back((int)G.cptr.size() - 1);
        chunk_doms.clear();
        chunk_cols.clear();
    };
    int rows = 0;
    std::vector<int> add;
    for (const auto &kr : key) {
        const int r = kr.second, n = node_of[r];
        add.clear();
        int extra = 0;
        for (int a = H.mptr[n]; a < H.mptr[n + 1]; ++a) {  // the domains holding n: its ring
            const int k = H.member[a];
            if (dom_mark[k] == -2) continue;
            dom_mark[k] = -2;
            add.push_back(k);
            extra += rows_of(k);
        }
        for (int k : add) dom_mark[k] = -1;
        int fresh = 0;  // domains already in the chunk do not count again
        for (int k : add)
            if (std::find(chunk_doms.begin(), chunk_doms.end(), k) == chunk_doms.end()) fresh += rows_of(k);
        if (extra > SG_ROWS) return false;
        if (!chunk_cols.empty() && ((int)chunk_cols.size() >= SG_COLS || rows + fresh > SG_ROWS)) {
            flush();
            rows = 0;
            fresh = extra;
        }
        for (int k : add)
            if (std::find(chunk_doms.begin(), chunk_doms.end(), k) == chunk_doms.end()) chunk_doms.push_back(k);
        chunk_cols.push_back(r);
        rows += fresh;
    }
    flush();
    return G.con.size() < (1ull << 31);
}

int sfem_setup(ofea_ctx *c)
{
    if (c->nd == 0) return fail(OFEA_ERR_STATE, "ofea_set_mesh first");
    if (c->has_type[OFEA_C3D8]) return fail(OFEA_ERR_ARG, "the smoothed method takes C3D4 meshes only");
    SfemHost H;
    int rc = sfem_domains(c->nd, c->ne, c->conn.data(), c->xyz.data(), H);
    if (rc) return rc;
    Pattern P;
    if ((rc = sfem_pattern(c->nd, H, c->resort.data(), P))) return rc;
    // contributors of every block (r, k), ascending domain
    const int nblk = P.node_blk[P.nfree];
    std::vector<int> cnt(nblk + 1, 0);
    auto block_of = [&](int r, int s) {
        const int *b = P.bcol.data() + P.node_blk[r], *e = P.bcol.data() + P.node_blk[r + 1];
        return P.node_blk[r] + (int)(std::lower_bound(b, e, s) - b);
    };
    for (int k = 0; k < c->nd; ++k)
        for (int a = H.mptr[k]; a < H.mptr[k + 1]; ++a) {
            const int r = c->resort[H.member[a]];
            if (r < 0) continue;
            for (int b = H.mptr[k]; b < H.mptr[k + 1]; ++b) {
                const int s = c->resort[H.member[b]];
                if (s >= 0) cnt[block_of(r, s) + 1]++;
            }
        }
    for (int b = 0; b < nblk; ++b) cnt[b + 1] += cnt[b];
    if (cnt[nblk] < 0 || (long long)cnt[nblk] >= (1LL << 31)) return fail(OFEA_ERR_LIMIT, "too many contributors");
    std::vector<int2> contrib(cnt[nblk]);
    std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
    for (int k = 0; k < c->nd; ++k)  // ascending domain: the reference's summation order
        for (int a = H.mptr[k]; a < H.mptr[k + 1]; ++a) {
            const int r = c->resort[H.member[a]];
            if (r < 0) continue;
            for (int b = H.mptr[k]; b < H.mptr[k + 1]; ++b) {
                const int s = c->resort[H.member[b]];
                if (s >= 0) contrib[fillp[block_of(r, s)]++] = make_int2(a, b);
            }
        }
    std::vector<int> blk_r(nblk);
    for (int r = 0; r < P.nfree; ++r)
        for (int b = P.node_blk[r]; b < P.node_blk[r + 1]; ++b) blk_r[b] = r;
    // device: replace the context's pattern (the FEM plans stay but no longer match)
    hipStreamSynchronize(c->stream);
    solver_free(c);
    batch_free(c);
    learned_free(c);
    ca_free(c);
    sfem_free(c);
    c->P = std::move(P);
    c->sfem = true;
    c->assembled = false;
    if ((rc = dev_upload(c->d_col_ptr, c->P.col_ptr.data(), c->P.col_ptr.size()))) return rc;
    if ((rc = dev_upload(c->d_node_blk, c->P.node_blk.data(), c->P.node_blk.size()))) return rc;
    if ((rc = dev_upload(c->d_bcol, c->P.bcol.data(), std::max<size_t>(c->P.bcol.size(), 1)))) return rc;
    if ((rc = dev_upload(c->d_diag_k, c->P.diag_k.data(), std::max<size_t>(c->P.diag_k.size(), 1)))) return rc;
    if ((rc = dev_upload<double>(c->d_val, nullptr, (size_t)std::max(c->P.nnz, 1LL)))) return rc;
    if ((rc = dev_upload<double>(c->d_rhs, nullptr, (size_t)std::max(c->P.n, 1)))) return rc;
    if ((rc = dev_upload<double>(c->d_x, nullptr, (size_t)std::max(c->P.n, 1)))) return rc;
    c->rhs_set = false;
    auto *w = new SfemWs();
    c->sfem_ws = w;
    w->nd = c->nd;
    w->ndom = c->nd;
    w->nrows = (long long)H.member.size();
    w->ncontrib = (long long)contrib.size();
    if ((rc = dev_upload(w->conn, H.conn.data(), H.conn.size()))) return rc;
    if ((rc = dev_upload(w->fflag, H.fflag.data(), H.fflag.size()))) return rc;
    if ((rc = dev_upload(w->dom_mptr, H.mptr.data(), H.mptr.size()))) return rc;
    if ((rc = dev_upload(w->dom_member, H.member.data(), std::max<size_t>(H.member.size(), 1)))) return rc;
    if ((rc = dev_upload(w->dom_eptr, H.eptr.data(), H.eptr.size()))) return rc;
    if ((rc = dev_upload(w->dom_elem, H.elem.data(), std::max<size_t>(H.elem.size(), 1)))) return rc;
    if ((rc = dev_upload(w->vol, H.vol.data(), H.vol.size()))) return rc;
    if ((rc = dev_upload<double>(w->rows, nullptr, 12 * (size_t)std::max(w->nrows, 1LL)))) return rc;
    {  // chunks of consecutive domains: <= SF_PAIRS pairs and <= SF_PAIRS members each
        std::vector<int> chunk{0}, mdom(H.member.size()), pdom(std::max<size_t>(H.elem.size(), 1));
        int np = 0, nm = 0;
        for (int k = 0; k < c->nd; ++k) {
            const int dp = H.eptr[k + 1] - H.eptr[k], dm = H.mptr[k + 1] - H.mptr[k];
            if (np + dp > SF_PAIRS || nm + dm > SF_PAIRS) {
                chunk.push_back(k);
                np = nm = 0;
            }
            np += dp;
            nm += dm;
            for (int a = H.mptr[k]; a < H.mptr[k + 1]; ++a) mdom[a] = k;
            for (int a = H.eptr[k]; a < H.eptr[k + 1]; ++a) pdom[a] = k;
        }
        chunk.push_back(c->nd);
        w->nchunks = (int)chunk.size() - 1;
        if ((rc = dev_upload(w->chunk_dom, chunk.data(), chunk.size()))) return rc;
        if ((rc = dev_upload(w->mdom, mdom.data(), std::max<size_t>(mdom.size(), 1)))) return rc;
        if ((rc = dev_upload(w->pdom, pdom.data(), pdom.size()))) return rc;
    }
    if ((rc = dev_upload(w->cptr, cnt.data(), cnt.size()))) return rc;
    if ((rc = dev_upload(w->contrib, contrib.data(), std::max<size_t>(contrib.size(), 1)))) return rc;
    int *d_blk_r = nullptr;
    if ((rc = dev_upload(d_blk_r, blk_r.data(), std::max<size_t>(blk_r.size(), 1)))) return rc;
    w->blk_r = d_blk_r;
    GatherPlan2 G;
    if (sfem_chunk_plan(c->nd, c->xyz.data(), H, c->resort.data(), c->P, cnt, contrib, G)) {
        w->gnchunks = (int)G.col_ptr.size() - 1;
        if ((rc = dev_upload(w->g_col_ptr, G.col_ptr.data(), G.col_ptr.size()))) return rc;
        if ((rc = dev_upload(w->g_cols, G.cols.data(), std::max<size_t>(G.cols.size(), 1)))) return rc;
        if ((rc = dev_upload(w->g_dom_ptr, G.dom_ptr.data(), G.dom_ptr.size()))) return rc;
        if ((rc = dev_upload(w->g_doms, G.doms.data(), std::max<size_t>(G.doms.size(), 1)))) return rc;
        if ((rc = dev_upload(w->g_blk_ptr, G.blk_ptr.data(), G.blk_ptr.size()))) return rc;
        if ((rc = dev_upload(w->g_cptr, G.cptr.data(), G.cptr.size()))) return rc;
        if ((rc = dev_upload(w->g_con, G.con.data(), std::max<size_t>(G.con.size(), 1)))) return rc;
    }
    return OFEA_OK;
}

int sfem_assemble(ofea_ctx *c, double E, double nu, double *nz_val)
{
    SfemWs *w = (SfemWs *)c->sfem_ws;
    if (!c->sfem || !w) return fail(OFEA_ERR_STATE, "ofea_sfem_setup first");
    if (!(nu > -1.0 && nu < 0.5)) return fail(OFEA_ERR_ARG, "Poisson ratio must lie in (-1, 0.5)");
    double lam, mu;
    lame<double>(E, nu, lam, mu);
    if (w->nchunks > 0)
        hipLaunchKernelGGL(k_sfem_grad, dim3(w->nchunks), dim3(SF_PAIRS), 0, c->stream, w->nchunks, w->chunk_dom,
                           c->d_xyz, w->conn, w->fflag, w->dom_mptr, w->dom_member, w->mdom, w->dom_eptr, w->dom_elem,
                           w->pdom, w->vol, w->rows);
    OFEA_HIP(hipGetLastError());
    const int nblk = c->P.node_blk[c->P.nfree];
    if (w->gnchunks > 0)
        hipLaunchKernelGGL(k_sfem_gather_chunks, dim3(w->gnchunks), dim3(SG_WG), 0, c->stream, w->gnchunks,
                           w->g_col_ptr, w->g_cols, w->g_dom_ptr, w->g_doms, w->g_blk_ptr, w->g_cptr, w->g_con,
                           w->dom_mptr, c->d_node_blk, c->d_col_ptr, w->rows, lam, mu, c->d_val);
    else if (nblk > 0)
        hipLaunchKernelGGL(k_sfem_gather, dim3((nblk + 255) / 256), dim3(256), 0, c->stream, nblk, w->blk_r,
                           c->d_node_blk, c->d_col_ptr, w->cptr, w->contrib, w->rows, lam, mu, c->d_val);
    OFEA_HIP(hipGetLastError());
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

}  // namespace ofea
