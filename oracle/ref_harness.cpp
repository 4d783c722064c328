// ref_harness.cpp -- C entry points over the reference's own element code.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's element
// translation units, taken where they lie under the reference checkout
// (src/Elements/C3D8.cpp, C3D8_Gauss.cpp, C3D4.cpp, C3D4_hammer.cpp), into
// oracle/_ref/libofea_ref.so by oracle/Makefile.  The library is git-ignored
// (no reference code enters the history) but, like every library built in
// the tree, it is copied with the working tree to the machines that run the
// tests and the benchmark.  It is loaded only as a checker or a CPU baseline:
// by the golden-vector script under tests/golden/, by tests that compare the
// restatement with it where it exists, and by bench.py's CPU-baseline leg.
//
// Each entry point instantiates the registered element class
// (src/Elements/C3D8.cpp:19, src/Elements/C3D4.cpp:19), sets the material
// (src/include/elastic_mat.h:17-22), builds D (build_cons_mat) and calls the
// reference member named in its comment.
#include "include/C3D8.h"
#include "include/C3D4.h"

#include <algorithm>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Eigen::MatrixXd;

template <class Elem>
void setup(Elem &el, double E, double nu, double rho)
{
    CAE::elastic_mat m{E, nu, rho};
    el.set_matrial(m);
    el.build_cons_mat();
}

MatrixXd coords_mat(const double *X, int nn)
{
    MatrixXd M(nn, 3);
    for (int a = 0; a < nn; ++a)
        for (int c = 0; c < 3; ++c) M(a, c) = X[3 * a + c];
    return M;
}

void copy_out(const MatrixXd &K, double *out)
{
    for (int i = 0; i < K.rows(); ++i)
        for (int j = 0; j < K.cols(); ++j) out[i * K.cols() + j] = K(i, j);
}

template <class Elem>
void ke(const double *X, double E, double nu, int nn, double *Ke)
{
    Elem el;
    setup(el, E, nu, 1.0);
    MatrixXd Xm = coords_mat(X, nn);
    MatrixXd K(3 * nn, 3 * nn);
    el.build_ele_stiff_mat(Xm, K);  // virtual ele_base::build_ele_stiff_mat
    copy_out(K, Ke);
}

template <class Elem>
void nl(const double *X, const double *u, double E, double nu, int nn, double *K_out, double *f_out)
{
    Elem el;
    setup(el, E, nu, 1.0);
    MatrixXd Xm = coords_mat(X, nn);
    MatrixXd U(3, nn);  // node_dis is 3 x nn (src/nl_assemble.cpp:116-131)
    for (int a = 0; a < nn; ++a)
        for (int c = 0; c < 3; ++c) U(c, a) = u[3 * a + c];
    MatrixXd K(3 * nn, 3 * nn), f(3 * nn, 1);
    el.build_ele_nl_stiff_mat(Xm, U, K, f);
    copy_out(K, K_out);
    for (int i = 0; i < 3 * nn; ++i) f_out[i] = f(i, 0);
}

template <class Elem>
void mass(const double *X, double rho, int nn, double *m_out)
{
    Elem el;
    setup(el, 1.0, 0.3, rho);
    std::vector<int> topo(nn);
    std::vector<std::vector<double>> xyz(nn, std::vector<double>(3));
    for (int a = 0; a < nn; ++a) {
        topo[a] = a + 1;
        for (int c = 0; c < 3; ++c) xyz[a][c] = X[3 * a + c];
    }
    std::vector<double> M(nn, 0.0);
    el.build_ele_mass(topo, xyz, M);
    for (int a = 0; a < nn; ++a) m_out[a] = M[a];
}

template <class Elem>
void in_force(const double *X, const double *du, double E, double nu, int nn, double *f_inout)
{
    Elem el;
    setup(el, E, nu, 1.0);
    std::vector<int> topo(nn);
    std::vector<std::vector<double>> xyz(nn, std::vector<double>(3));
    std::vector<double> d(3 * nn), F(3 * nn);
    for (int a = 0; a < nn; ++a) {
        topo[a] = a + 1;
        for (int c = 0; c < 3; ++c) {
            xyz[a][c] = X[3 * a + c];
            d[3 * a + c] = du[3 * a + c];
            F[3 * a + c] = f_inout[3 * a + c];
        }
    }
    std::vector<double> stress, strain;
    el.cal_in_force(topo, xyz, d, stress, strain, F);
    for (int i = 0; i < 3 * nn; ++i) f_inout[i] = F[i];
}

template <class Elem>
double timestep(const double *X, double E, double nu, int nn, double dt)
{
    Elem el;
    setup(el, E, nu, 1.0);
    MatrixXd Xm = coords_mat(X, nn);
    el.update_timestep(Xm, dt);
    return dt;
}

// column-major lookup of row r in column j (rows sorted), as
// fill_CSR_sparse_mat finds each entry (src/assemble.cpp:95-151)
inline int csc_find(const int *col_ptr, const int *row_idx, int j, int r)
{
    const int *b = row_idx + col_ptr[j], *e = row_idx + col_ptr[j + 1];
    const int *p = std::lower_bound(b, e, r);
    return (p != e && *p == r) ? (int)(p - row_idx) : -1;
}

}  // namespace

extern "C" {

// K on a prebuilt CSC pattern from the reference's own element stiffness
// (build_ele_stiff_mat of the registered element classes) and the scatter of
// fill_CSR_sparse_mat (src/assemble.cpp:95-151): every element's Ke added, in
// element order, at the free DOFs (resort >= 0, 1-based conn, 8 slots per
// element, etype 0 = C3D4, 1 = C3D8; E_elem per element or NULL for E).
// With threads > 1 (OpenMP) the element
// matrices of a chunk are computed in parallel and the scatter is split by
// column ranges, each thread adding the chunk's elements in order into its
// own columns, so the sums do not depend on the thread count.  Returns the
// threads used.
int ref_fill_csc(int ne, const int *conn, const int *etype, const double *xyz, const int *resort, double E,
                 const double *E_elem, double nu, const int *col_ptr, const int *row_idx, double *val, int ncol,
                 int threads)
{
    const int T = threads > 0 ? threads : 1;
    const int CH = 4096;
    std::vector<double> buf((size_t)CH * 576);
    std::vector<int> dofs((size_t)CH * 24);
    for (int e0 = 0; e0 < ne; e0 += CH) {
        const int e1 = std::min(ne, e0 + CH);
#pragma omp parallel for schedule(static) num_threads(T)
        for (int e = e0; e < e1; ++e) {
            const int nn = etype[e] == 1 ? 8 : 4;
            double X[24];
            int *dof = &dofs[(size_t)(e - e0) * 24];
            for (int a = 0; a < nn; ++a) {
                const int node = conn[8 * e + a] - 1, f = resort[node];
                for (int c = 0; c < 3; ++c) {
                    X[3 * a + c] = xyz[3 * node + c];
                    dof[3 * a + c] = f >= 0 ? 3 * f + c : -1;
                }
            }
            double *Ke = &buf[(size_t)(e - e0) * 576];
            const double Ee = E_elem ? E_elem[e] : E;
            if (nn == 8) ke<CAE::hex_ele_elastic>(X, Ee, nu, 8, Ke);
            else ke<CAE::tetra_ele_elastic>(X, Ee, nu, 4, Ke);
        }
#pragma omp parallel num_threads(T)
        {
            int t = 0, nt = 1;
#ifdef _OPENMP
            t = omp_get_thread_num();
            nt = omp_get_num_threads();
#endif
            const int c0 = (int)((long long)ncol * t / nt), c1 = (int)((long long)ncol * (t + 1) / nt);
            for (int e = e0; e < e1; ++e) {
                const int nn = etype[e] == 1 ? 8 : 4;
                const double *Ke = &buf[(size_t)(e - e0) * 576];
                const int *dof = &dofs[(size_t)(e - e0) * 24];
                for (int j = 0; j < 3 * nn; ++j) {
                    if (dof[j] < c0 || dof[j] >= c1) continue;
                    for (int i = 0; i < 3 * nn; ++i) {
                        if (dof[i] < 0) continue;
                        const int k = csc_find(col_ptr, row_idx, dof[j], dof[i]);
                        if (k >= 0) val[k] += Ke[i * 3 * nn + j];
                    }
                }
            }
        }
    }
    return T;
}

// hex_ele_elastic::build_ele_stiff_mat (src/Elements/C3D8.cpp:45-49)
void ref_c3d8_ke(const double *X, double E, double nu, double *Ke) { ke<CAE::hex_ele_elastic>(X, E, nu, 8, Ke); }
// tetra_ele_elastic::build_ele_stiff_mat (src/Elements/C3D4.cpp:45-54)
void ref_c3d4_ke(const double *X, double E, double nu, double *Ke) { ke<CAE::tetra_ele_elastic>(X, E, nu, 4, Ke); }
// hex_ele_elastic::build_ele_nl_stiff_mat (src/Elements/C3D8.cpp:51-56)
void ref_c3d8_nl(const double *X, const double *u, double E, double nu, double *K, double *f)
{
    nl<CAE::hex_ele_elastic>(X, u, E, nu, 8, K, f);
}
// tetra_ele_elastic::build_ele_nl_stiff_mat (src/Elements/C3D4.cpp:56-60)
void ref_c3d4_nl(const double *X, const double *u, double E, double nu, double *K, double *f)
{
    nl<CAE::tetra_ele_elastic>(X, u, E, nu, 4, K, f);
}
// build_ele_mass (src/Elements/C3D8.cpp:58-108, src/Elements/C3D4.cpp:75-100)
void ref_c3d8_mass(const double *X, double rho, double *m) { mass<CAE::hex_ele_elastic>(X, rho, 8, m); }
void ref_c3d4_mass(const double *X, double rho, double *m) { mass<CAE::tetra_ele_elastic>(X, rho, 4, m); }
// cal_in_force (src/Elements/C3D8.cpp:158-193, src/Elements/C3D4.cpp:102-138)
void ref_c3d8_in_force(const double *X, const double *du, double E, double nu, double *f)
{
    in_force<CAE::hex_ele_elastic>(X, du, E, nu, 8, f);
}
void ref_c3d4_in_force(const double *X, const double *du, double E, double nu, double *f)
{
    in_force<CAE::tetra_ele_elastic>(X, du, E, nu, 4, f);
}
// update_timestep (src/Elements/C3D8.cpp:195-219, src/Elements/C3D4.cpp:140-164)
double ref_c3d8_timestep(const double *X, double E, double nu, double dt)
{
    return timestep<CAE::hex_ele_elastic>(X, E, nu, 8, dt);
}
double ref_c3d4_timestep(const double *X, double E, double nu, double dt)
{
    return timestep<CAE::tetra_ele_elastic>(X, E, nu, 4, dt);
}

}  // extern "C"
