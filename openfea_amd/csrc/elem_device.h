// elem_device.h -- element geometry shared by the assembly, explicit and
// total-Lagrangian kernels (device code, templated on the arithmetic type).
#pragma once

#include <hip/hip_runtime.h>

namespace ofea {

// D entries of build_cons_mat (src/Elements/C3D8.cpp:21-35): lambda = fac_a,
// mu = fac_b with fac = E(1-nu)/((1+nu)(1-2nu)).
template <class T>
__host__ __device__ inline void lame(T E, T nu, T &lam, T &mu)
{
    T fac = (E * (T(1) - nu)) / ((T(1) + nu) * (T(1) - T(2) * nu));
    lam = fac * nu / (T(1) - nu);
    mu = fac * (T(1) - T(2) * nu) / (T(2) * (T(1) - nu));
}

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X; this
// bijection gives consecutive logical tiles to the same XCD (shared L2).
__device__ inline int xcd_remap(int i, int n)
{
    int q = n >> 3, rem = n & 7, x = i & 7, j = i >> 3;
    return x < rem ? x * (q + 1) + j : rem * (q + 1) + (x - rem) * q + j;
}

template <class T>
__device__ inline T det3(const T J[9])
{
    // six-term expansion of src/Elements/C3D8_Gauss.cpp:58-63
    return J[0] * J[4] * J[8] + J[1] * J[5] * J[6] + J[2] * J[3] * J[7] - J[2] * J[4] * J[6] - J[5] * J[7] * J[0] -
           J[8] * J[1] * J[3];
}

template <class T>
__device__ inline void inv3(const T J[9], T Ji[9])
{
    T c00 = J[4] * J[8] - J[5] * J[7];
    T c01 = J[5] * J[6] - J[3] * J[8];
    T c02 = J[3] * J[7] - J[4] * J[6];
    T id = T(1) / (J[0] * c00 + J[1] * c01 + J[2] * c02);
    Ji[0] = c00 * id;
    Ji[1] = (J[2] * J[7] - J[1] * J[8]) * id;
    Ji[2] = (J[1] * J[5] - J[2] * J[4]) * id;
    Ji[3] = c01 * id;
    Ji[4] = (J[0] * J[8] - J[2] * J[6]) * id;
    Ji[5] = (J[2] * J[3] - J[0] * J[5]) * id;
    Ji[6] = c02 * id;
    Ji[7] = (J[1] * J[6] - J[0] * J[7]) * id;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) * id;
}

// Corner signs of the Abaqus C3D8 node order, which is also the Gauss-point
// order of src/Elements/C3D8_Gauss.cpp:85-92.
//   x: - + + - - + + -   y: - - + + - - + +   z: - - - - + + + +
__host__ __device__ inline int hsgn(int a, int axis)
{
    int bit = axis == 0 ? (((a + 1) >> 1) & 1) : axis == 1 ? ((a >> 1) & 1) : ((a >> 2) & 1);
    return bit ? 1 : -1;
}

// Trilinear shape-function derivatives (src/Elements/C3D8_Gauss.cpp:26-51).
template <class T>
__device__ inline void hex_dN(int a, T r, T s, T t, T &dr, T &ds, T &dt)
{
    T sr = (T)hsgn(a, 0), ss = (T)hsgn(a, 1), st = (T)hsgn(a, 2);
    dr = T(0.125) * sr * (T(1) + ss * s) * (T(1) + st * t);
    ds = T(0.125) * ss * (T(1) + sr * r) * (T(1) + st * t);
    dt = T(0.125) * st * (T(1) + sr * r) * (T(1) + ss * s);
}

// Physical gradients g[c][a] of the 8 shape functions at Gauss point gp of the
// element with coordinates X (node-major), and the signed detJ (weight 1).
template <class T>
__device__ inline T hex_point(const T X[24], int gp, T g[3][8])
{
    const T q = T(0.57735026918962576451);  // 1/sqrt(3)
    T r = T(hsgn(gp, 0)) * q, s = T(hsgn(gp, 1)) * q, t = T(hsgn(gp, 2)) * q;
    T dN[3][8];
#pragma unroll
    for (int a = 0; a < 8; ++a) hex_dN<T>(a, r, s, t, dN[0][a], dN[1][a], dN[2][a]);
    T J[9], Ji[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T acc = T(0);
#pragma unroll
            for (int a = 0; a < 8; ++a) acc += dN[i][a] * X[3 * a + j];
            J[3 * i + j] = acc;
        }
    T detJ = det3(J);
    inv3(J, Ji);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 8; ++a) g[c][a] = Ji[3 * c] * dN[0][a] + Ji[3 * c + 1] * dN[1][a] + Ji[3 * c + 2] * dN[2][a];
    return detJ;
}

// Constant gradients of the linear tet (dN = [-1 1 0 0; -1 0 1 0; -1 0 0 1],
// src/Elements/C3D4_hammer.cpp:26-29); returns the signed detJ = 6V.
template <class T>
__device__ inline T tet_grads(const T X[12], T g[3][4])
{
    T J[9], Ji[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) J[3 * i + j] = X[3 * (i + 1) + j] - X[j];
    T detJ = det3(J);
    inv3(J, Ji);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        g[c][1] = Ji[3 * c + 0];
        g[c][2] = Ji[3 * c + 1];
        g[c][3] = Ji[3 * c + 2];
        g[c][0] = -(Ji[3 * c + 0] + Ji[3 * c + 1] + Ji[3 * c + 2]);
    }
    return detJ;
}

// Select g[.][a] for a lane-dependent a without dynamic register indexing.
template <class T, int NN>
__device__ inline void pick_col(const T g[3][NN], int a, T out[3])
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T v = T(0);
#pragma unroll
        for (int b = 0; b < NN; ++b) v = (b == a) ? g[c][b] : v;
        out[c] = v;
    }
}

// 8-lane reduce-scatter over aligned 8-lane groups: every lane produces 8
// blocks of NV values (blk(b, out)); afterwards lane j of the group holds the
// group sum of block j.  Three xor-shuffle halving steps (4, 2, 1).
template <int NV, class T, class BlockFn>
__device__ inline void reduce_scatter8(int lane8, const BlockFn &blk, T out[NV])
{
    const bool h2 = (lane8 >> 2) & 1, h1 = (lane8 >> 1) & 1, h0 = lane8 & 1;
    T acc1[4][NV];
#pragma unroll
    for (int bl = 0; bl < 4; ++bl) {
        T lo[NV], hi[NV];
        blk(bl, lo);
        blk(4 + bl, hi);
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            T mine = h2 ? hi[i] : lo[i], other = h2 ? lo[i] : hi[i];
            acc1[bl][i] = mine + __shfl_xor(other, 4);
        }
    }
    T acc2[2][NV];
#pragma unroll
    for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            T lo = acc1[b2][i], hi = acc1[2 + b2][i];
            T mine = h1 ? hi : lo, other = h1 ? lo : hi;
            acc2[b2][i] = mine + __shfl_xor(other, 2);
        }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        T mine = h0 ? acc2[1][i] : acc2[0][i], other = h0 ? acc2[0][i] : acc2[1][i];
        out[i] = mine + __shfl_xor(other, 1);
    }
}

// Full sum over an aligned 8-lane group (every lane gets the sum).
template <class T>
__device__ inline T group8_sum(T v)
{
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 1);
    return v;
}

}  // namespace ofea
