/* openfea_amd_cae.hpp -- C++17 host layer with the reference's call surface.
 *
 * Header-only classes over the C ABI of openfea_amd.h that keep the names,
 * argument meaning and error behaviour of the reference's implicit pipeline
 * (CAE_process::pre_info + implict_analysis, src/cae.cpp:168-196 and
 * 198-301), so that code written against the reference's classes runs on the
 * GPU by changing its includes and namespace:
 *
 *   ReadInfo            src/include/data2cae.h:29-77, src/data2cae.cpp:18-492
 *   data_management     src/include/data_management.h:45-104 (the fields of
 *                       the linear static path)
 *   elastic_mat         src/include/elastic_mat.h:17-22
 *   set_BCs             src/include/set_bcs.h:13-25, src/set_bc.cpp:20-77
 *   assamble_stiffness  src/include/assemble.h:35-72, src/assemble.cpp:19-151
 *   solution_api        src/include/linear_solution.h:27, src/linear_solution.cpp:17-64
 *   assamble_nl_stiffness src/include/nl_assemble.h:32-55, src/nl_assemble.cpp:19-166
 *   solution_nl_api     src/include/linear_solution.h:28, src/linear_solution.cpp:66-107
 *   CAE_process         src/include/cae.h:47-91: pre_info, implict_analysis
 *                       (linear and the NLFEA Newton branch, src/cae.cpp:168-301)
 *                       and explicit_analysis (src/cae.cpp:420-574)
 *   simulation_post     src/include/post_process.h:25-45, src/post_process.cpp:18-168
 *   data_process        src/include/data2vtk.h:27-40, src/data2vtk.cpp:47-427
 *
 * Errors follow the reference: a message on std::cerr, then std::exit(1)
 * (the reference prints and exits in its reader and in PARDISO);
 * solution_api returns false for a solver string it does not know.  The
 * direct solvers ("Pardiso_*", "SuperLU_*") are not part of this library:
 * solution_api returns false for them.  The "GPU_PCG*" strings select the
 * device PCG methods; the learned ones need ofea_learned_* on the context
 * (assamble_stiffness::gpu_ctx()) first.
 *
 * The reader matches the reference's keywords "Set-load", "*Cload" and
 * "Set-fix" only (the values src/main.cpp passes); other keywords are
 * rejected.  Keyword-level decks (*Boundary, *Dsload, sections) are read by
 * the C ABI's ofea_inp_* extension getters. */
#ifndef OPENFEA_AMD_CAE_HPP
#define OPENFEA_AMD_CAE_HPP

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "openfea_amd.h"

namespace ofea_cae {

inline void check(int rc, const char *what)
{
    if (rc < 0) {
        std::cerr << "openfea_amd: " << what << ": " << ofea_last_error() << std::endl;
        std::exit(1);
    }
}

struct elastic_mat {
    double young_modulus = 0.0;
    double poisson_ratio = 0.0;
    double density = 0.0;
};

class data_management {
public:
    int ne_ = 0, nd_ = 0;
    std::vector<std::vector<double>> coords_;    // nd x 3
    std::vector<std::vector<int>> node_topos_;   // ne x 8, 1-based; C3D4 uses the first 4
    std::vector<int> load_set_;                  // 1-based node ids
    int load_dof_ = 0;                           // 1 x, 2 y, 3 z
    double load_value_ = 0.0;
    std::vector<int> dis_bc_set_;                // 1-based node ids
    std::vector<int> resort_free_nodes_;         // nd: free-node rank or -1
    std::vector<double> single_load_vec_;        // 3 x free nodes
    std::vector<double> single_dis_vec_;         // 3 x free nodes
    std::vector<double> single_full_dis_vec_;    // 3 nd
    std::vector<std::vector<double>> stress_mat_;       // ne x 7: xx yy zz xy yz xz Mises
    std::vector<std::vector<double>> stress_node_mat_;  // nd x 7
    std::vector<int> ele_list_idx_;              // per element: ELE_TYPES value
    double time_total_ = 0.0, time_step_ = 0.0;  // explicit: total time; step (0 = automatic)
    bool NLFEA = false;                           // geometric nonlinear analysis
    double res_lmit = 1.0e-6;                     // its residual limit (data_management.h:78)
    std::map<std::string, int> ELE_TYPES = {{"C3D4", OFEA_C3D4}, {"C3D8", OFEA_C3D8}};

    // flat arrays in the layout of the C ABI
    std::vector<double> flat_coords() const
    {
        std::vector<double> xyz;
        xyz.reserve(3 * coords_.size());
        for (const auto &p : coords_) xyz.insert(xyz.end(), p.begin(), p.begin() + 3);
        return xyz;
    }
    std::vector<int> flat_topos() const
    {
        std::vector<int> conn(8 * node_topos_.size(), 0);
        for (size_t e = 0; e < node_topos_.size(); ++e)
            for (size_t a = 0; a < node_topos_[e].size() && a < 8; ++a) conn[8 * e + a] = node_topos_[e][a];
        return conn;
    }
};

// Reads the deck once (ofea_inp_open) and hands out its parts in the
// reference's method order.
class ReadInfo {
public:
    std::string path_;

    ReadInfo() = default;
    explicit ReadInfo(std::string path1) : path_(std::move(path1)) {}
    ReadInfo(const ReadInfo &) = delete;
    ReadInfo &operator=(const ReadInfo &) = delete;
    ~ReadInfo()
    {
        if (deck_) ofea_inp_close(deck_);
    }

    void read_ele_node_num(data_management &d)
    {
        open();
        d.nd_ = nd_;
        d.ne_ = ne_;
    }

    void read_geo_mesh(data_management &d)
    {
        open();
        std::vector<double> xyz(3 * (size_t)nd_);
        std::vector<int> conn(8 * (size_t)ne_), etype(ne_);
        check(ofea_inp_get(deck_, xyz.data(), conn.data(), etype.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr),
              "read_geo_mesh");
        d.nd_ = nd_;
        d.ne_ = ne_;
        d.coords_.assign(nd_, std::vector<double>(3));
        for (int i = 0; i < nd_; ++i)
            for (int c = 0; c < 3; ++c) d.coords_[i][c] = xyz[3 * (size_t)i + c];
        d.node_topos_.assign(ne_, std::vector<int>(8));
        for (int e = 0; e < ne_; ++e)
            for (int a = 0; a < 8; ++a) d.node_topos_[e][a] = conn[8 * (size_t)e + a];
        d.ele_list_idx_ = etype;
    }

    void read_load_bcs(const std::string &load_set_keyword, const std::string &load_value_keyword,
                       data_management &d)
    {
        if (load_set_keyword != "Set-load" || load_value_keyword != "*Cload") {
            std::cerr << "openfea_amd: read_load_bcs reads the keywords Set-load and *Cload only" << std::endl;
            std::exit(1);
        }
        open();
        d.load_set_.assign(nload_, 0);
        check(ofea_inp_get(deck_, nullptr, nullptr, nullptr, nullptr, d.load_set_.data(), &d.load_dof_,
                           &d.load_value_, nullptr, nullptr, nullptr, nullptr, nullptr),
              "read_load_bcs");
    }

    void read_dis_bcs(const std::string &dis_set_keyword, data_management &d)
    {
        if (dis_set_keyword != "Set-fix") {
            std::cerr << "openfea_amd: read_dis_bcs reads the keyword Set-fix only" << std::endl;
            std::exit(1);
        }
        open();
        d.dis_bc_set_.assign(nfix_, 0);
        check(ofea_inp_get(deck_, nullptr, nullptr, nullptr, d.dis_bc_set_.data(), nullptr, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr),
              "read_dis_bcs");
    }

    // a deck without *Elastic / *Density keeps the material passed in, as
    // the reference keeps its constructor's values
    void read_mat(elastic_mat &mat)
    {
        open();
        double E, nu, rho;
        check(ofea_inp_get(deck_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &E, &nu, &rho,
                           nullptr, nullptr),
              "read_mat");
        if (!std::isnan(E)) mat.young_modulus = E;
        if (!std::isnan(nu)) mat.poisson_ratio = nu;
        if (!std::isnan(rho)) mat.density = rho;
    }

    void read_time(data_management &d)
    {
        open();
        check(ofea_inp_get(deck_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr, &d.time_step_, &d.time_total_),
              "read_time");
    }

    const ofea_deck *deck() { return open(); }

private:
    ofea_deck *deck_ = nullptr;
    int nd_ = 0, ne_ = 0, nfix_ = 0, nload_ = 0;

    ofea_deck *open()
    {
        if (!deck_) {
            check(ofea_inp_open(path_.c_str(), &deck_), "ReadInfo");
            check(ofea_inp_sizes(deck_, &nd_, &ne_, &nfix_, &nload_), "ReadInfo");
        }
        return deck_;
    }
};

class set_BCs {
public:
    // sorts dis_bc_set_ ascending and numbers the free nodes in node order
    void build_free_index(data_management &d)
    {
        d.resort_free_nodes_.assign(d.nd_, -1);
        const int nfree = ofea_build_free_index(d.nd_, d.dis_bc_set_.data(), (int)d.dis_bc_set_.size(),
                                                d.resort_free_nodes_.data());
        check(nfree, "build_free_index");
    }

    // F[3 r + dof - 1] = load_value_ for every load-set node (assignment)
    void build_single_load(data_management &d)
    {
        int nfree = 0;
        for (int r : d.resort_free_nodes_)
            if (r >= 0) ++nfree;
        d.single_load_vec_.assign(3 * (size_t)nfree, 0.0);
        check(ofea_build_single_load(3 * nfree, d.resort_free_nodes_.data(), d.load_set_.data(),
                                     (int)d.load_set_.size(), d.load_dof_, d.load_value_, d.single_load_vec_.data()),
              "build_single_load");
    }
};

// The CSC arrays of the reference (col_idx_ n+1, row_idx_ nnz, nz_val_ nnz),
// built and filled on the GPU; K stays resident in HBM for solution_api.
class assamble_stiffness {
public:
    int num_row_ = 0, num_col_ = 0, num_nz_val_ = 0;
    std::vector<double> nz_val_;
    std::vector<int> row_idx_;
    std::vector<int> col_idx_;

    assamble_stiffness() = default;
    explicit assamble_stiffness(int device) : device_(device) {}
    assamble_stiffness(const assamble_stiffness &) = delete;
    assamble_stiffness &operator=(const assamble_stiffness &) = delete;
    ~assamble_stiffness()
    {
        if (ctx_) ofea_destroy(ctx_);
    }

    void build_CSR(data_management &d)
    {
        if (!ctx_) check(ofea_create(&ctx_, device_), "build_CSR");
        const std::vector<double> xyz = d.flat_coords();
        const std::vector<int> conn = d.flat_topos();
        check(ofea_set_mesh(ctx_, d.nd_, xyz.data(), d.ne_, conn.data(), d.ele_list_idx_.data(),
                            d.resort_free_nodes_.data()),
              "build_CSR");
        int n = 0;
        long long nnz = 0;
        check(ofea_pattern(ctx_, &n, &nnz, nullptr, nullptr), "build_CSR");
        if (nnz > 0x7fffffffLL) {
            std::cerr << "openfea_amd: build_CSR: nnz exceeds the int32 index range" << std::endl;
            std::exit(1);
        }
        num_row_ = num_col_ = n;
        num_nz_val_ = (int)nnz;
        col_idx_.assign(n + 1, 0);
        row_idx_.assign(nnz, 0);
        check(ofea_pattern(ctx_, nullptr, nullptr, col_idx_.data(), row_idx_.data()), "build_CSR");
    }

    // zero-initialised sum over the elements in ascending order (FP64); the
    // values stay resident on the device and are copied into nz_val_
    void fill_CSR_sparse_mat(data_management &, elastic_mat &mat)
    {
        if (!ctx_) {
            std::cerr << "openfea_amd: fill_CSR_sparse_mat before build_CSR" << std::endl;
            std::exit(1);
        }
        nz_val_.assign(num_nz_val_, 0.0);
        check(ofea_assemble(ctx_, mat.young_modulus, mat.poisson_ratio, nz_val_.data()), "fill_CSR_sparse_mat");
    }

    ofea_ctx *gpu_ctx() const { return ctx_; }

private:
    int device_ = 0;
    ofea_ctx *ctx_ = nullptr;
};

// Relative residual and iteration limit of the GPU solves of solution_api.
struct gpu_solve_options {
    double rtol = 1e-10;
    int maxit = 200000;
    int iters = 0;        // filled by the last solve
    double relres = 0.0;  // true FP64 residual of the last solve
};

inline gpu_solve_options &solution_options()
{
    static gpu_solve_options o;
    return o;
}

// K u = F with the resident K; F = single_load_vec_, u -> single_dis_vec_.
// The load vector is left unchanged.  A solve that misses rtol returns true
// with solution_options().relres above rtol (the reference's solvers return
// true once they ran).
inline bool solution_api(assamble_stiffness &item_ass, data_management &d, const std::string &solver_type)
{
    static const std::map<std::string, int> methods = {
        {"GPU_PCG", OFEA_PCG_JACOBI},
        {"GPU_PCG_BLOCK", OFEA_PCG_BLOCK_JACOBI},
        {"GPU_PCG_MIXED", OFEA_PCG_MIXED},
        {"GPU_PCG_MIXED32", OFEA_PCG_MIXED32},
        {"GPU_PCG_LEARNED", OFEA_PCG_LEARNED},
        {"GPU_PCG_LEARNED_MIXED", OFEA_PCG_LEARNED_MIXED},
        {"GPU_PCG_TWO_LEVEL", OFEA_PCG_TWO_LEVEL},
        {"GPU_PCG_TWO_LEVEL_MIXED", OFEA_PCG_TWO_LEVEL_MIXED},
        {"GPU_PCG_LEARNED_TWO_LEVEL", OFEA_PCG_LEARNED_TWO_LEVEL},
        {"GPU_PCG_LEARNED_TWO_LEVEL_MIXED", OFEA_PCG_LEARNED_TWO_LEVEL_MIXED}};
    const auto m = methods.find(solver_type);
    if (m == methods.end()) {
        std::cerr << "solution_api: no solver named " << solver_type << std::endl;
        return false;
    }
    if (!item_ass.gpu_ctx()) {
        std::cerr << "openfea_amd: solution_api before build_CSR/fill_CSR_sparse_mat" << std::endl;
        std::exit(1);
    }
    gpu_solve_options &o = solution_options();
    d.single_dis_vec_.assign(d.single_load_vec_.size(), 0.0);
    check(ofea_solve(item_ass.gpu_ctx(), m->second, d.single_load_vec_.data(), d.single_dis_vec_.data(), o.rtol,
                     o.maxit, &o.iters, &o.relres),
          "solution_api");
    return true;
}

// The tangent of the total-Lagrangian St.Venant-Kirchhoff element and the
// internal force on the reference's CSC arrays (col_idx n+1, row_idx nnz,
// nz_val nnz; the same pattern as assamble_stiffness), built and filled on
// the GPU; K_T stays resident for solution_nl_api.
class assamble_nl_stiffness {
public:
    int num_row = 0, num_col = 0, num_nz_val = 0;
    std::vector<double> nz_val;
    std::vector<int> row_idx;
    std::vector<int> col_idx;

    assamble_nl_stiffness() = default;
    explicit assamble_nl_stiffness(int device) : device_(device) {}
    assamble_nl_stiffness(const assamble_nl_stiffness &) = delete;
    assamble_nl_stiffness &operator=(const assamble_nl_stiffness &) = delete;
    ~assamble_nl_stiffness()
    {
        if (ctx_) ofea_destroy(ctx_);
    }

    void build_CSR(data_management &d)
    {
        if (!ctx_) check(ofea_create(&ctx_, device_), "build_CSR");
        const std::vector<double> xyz = d.flat_coords();
        const std::vector<int> conn = d.flat_topos();
        check(ofea_set_mesh(ctx_, d.nd_, xyz.data(), d.ne_, conn.data(), d.ele_list_idx_.data(),
                            d.resort_free_nodes_.data()),
              "build_CSR");
        int n = 0;
        long long nnz = 0;
        check(ofea_pattern(ctx_, &n, &nnz, nullptr, nullptr), "build_CSR");
        if (nnz > 0x7fffffffLL) {
            std::cerr << "openfea_amd: build_CSR: nnz exceeds the int32 index range" << std::endl;
            std::exit(1);
        }
        num_row = num_col = n;
        num_nz_val = (int)nnz;
        col_idx.assign(n + 1, 0);
        row_idx.assign(nnz, 0);
        check(ofea_pattern(ctx_, nullptr, nullptr, col_idx.data(), row_idx.data()), "build_CSR");
    }

    // the free DOFs of element ele_id (3 per node; 3 r + c for free rank r,
    // negative for a clamped node), as src/nl_assemble.cpp:78-91
    void build_ele_dofs(std::vector<int> &item_ele_dofs, data_management &d, int ele_id, int num_nodes)
    {
        item_ele_dofs.assign(3 * (size_t)num_nodes, 0);
        for (int i = 0; i < num_nodes; ++i) {
            const int r = d.resort_free_nodes_[d.node_topos_[ele_id][i] - 1];
            for (int c = 0; c < 3; ++c) item_ele_dofs[3 * i + c] = 3 * r + c;
        }
    }

    // K_T into nz_val (zeroed first) and f_int added into inter_force_vec at
    // the free-DOF displacement current_dis_vec
    void fill_CSR_sparse_mat(data_management &, elastic_mat &mat, std::vector<double> &current_dis_vec,
                             std::vector<double> &inter_force_vec)
    {
        if (!ctx_) {
            std::cerr << "openfea_amd: fill_CSR_sparse_mat before build_CSR" << std::endl;
            std::exit(1);
        }
        if (current_dis_vec.size() != (size_t)num_row || inter_force_vec.size() != (size_t)num_row) {
            std::cerr << "openfea_amd: fill_CSR_sparse_mat: vectors must hold 3 x free nodes" << std::endl;
            std::exit(1);
        }
        nz_val.assign(num_nz_val, 0.0);
        std::vector<double> f(num_row, 0.0);
        check(ofea_nl_assemble(ctx_, mat.young_modulus, mat.poisson_ratio, current_dis_vec.data(), nz_val.data(),
                               f.data()),
              "fill_CSR_sparse_mat");
        for (int i = 0; i < num_row; ++i) inter_force_vec[i] = inter_force_vec[i] + f[i];
    }

    ofea_ctx *gpu_ctx() const { return ctx_; }

private:
    int device_ = 0;
    ofea_ctx *ctx_ = nullptr;
};

// K_T x = b with the resident tangent of the last fill_CSR_sparse_mat; x is
// resized to b.  The "GPU_PCG*" strings as in solution_api; other strings
// return false.
inline bool solution_nl_api(assamble_nl_stiffness &item_ass, data_management &, std::vector<double> &b,
                            std::vector<double> &x, const std::string &solver_type)
{
    static const std::map<std::string, int> methods = {
        {"GPU_PCG", OFEA_PCG_JACOBI},
        {"GPU_PCG_BLOCK", OFEA_PCG_BLOCK_JACOBI},
        {"GPU_PCG_MIXED", OFEA_PCG_MIXED},
        {"GPU_PCG_MIXED32", OFEA_PCG_MIXED32},
        {"GPU_PCG_LEARNED", OFEA_PCG_LEARNED},
        {"GPU_PCG_LEARNED_MIXED", OFEA_PCG_LEARNED_MIXED},
        {"GPU_PCG_TWO_LEVEL", OFEA_PCG_TWO_LEVEL},
        {"GPU_PCG_TWO_LEVEL_MIXED", OFEA_PCG_TWO_LEVEL_MIXED},
        {"GPU_PCG_LEARNED_TWO_LEVEL", OFEA_PCG_LEARNED_TWO_LEVEL},
        {"GPU_PCG_LEARNED_TWO_LEVEL_MIXED", OFEA_PCG_LEARNED_TWO_LEVEL_MIXED}};
    const auto m = methods.find(solver_type);
    if (m == methods.end()) {
        std::cout << "Please check your flag of solver !!!";
        return false;
    }
    if (!item_ass.gpu_ctx()) {
        std::cerr << "openfea_amd: solution_nl_api before build_CSR/fill_CSR_sparse_mat" << std::endl;
        std::exit(1);
    }
    gpu_solve_options &o = solution_options();
    x.assign(b.size(), 0.0);
    check(ofea_solve(item_ass.gpu_ctx(), m->second, b.data(), x.data(), o.rtol, o.maxit, &o.iters, &o.relres),
          "solution_nl_api");
    return true;
}

// Strain/stress recovery needs the mesh on a device context: pass the
// assembly's context (assamble_stiffness::gpu_ctx()) or let the object make
// its own on first use.
class simulation_post {
public:
    simulation_post() = default;
    explicit simulation_post(ofea_ctx *shared) : ctx_(shared) {}
    simulation_post(const simulation_post &) = delete;
    simulation_post &operator=(const simulation_post &) = delete;
    ~simulation_post()
    {
        if (own_) ofea_destroy(ctx_);
    }

    bool reset_displacement(data_management &d)
    {
        d.single_full_dis_vec_.assign(3 * (size_t)d.nd_, 0.0);
        check(ofea_reset_displacement(d.nd_, d.resort_free_nodes_.data(), d.single_dis_vec_.data(),
                                      d.single_full_dis_vec_.data()),
              "reset_displacement");
        return true;
    }

    // What the reference computes and writes as S_*: the engineering strain
    // B u_e per element and its Mises-form value, averaged at the nodes
    // (OFEA_POST_STRAIN).  OFEA_POST_STRESS gives the Cauchy stress D B u_e
    // of the material instead.
    bool get_cauchy_stress_3d(data_management &d, const elastic_mat &mat = elastic_mat{},
                              int quantity = OFEA_POST_STRAIN)
    {
        if (!ctx_) {
            check(ofea_create(&ctx_, 0), "get_cauchy_stress_3d");
            own_ = true;
            const std::vector<double> xyz = d.flat_coords();
            const std::vector<int> conn = d.flat_topos();
            check(ofea_set_mesh(ctx_, d.nd_, xyz.data(), d.ne_, conn.data(), d.ele_list_idx_.data(),
                                d.resort_free_nodes_.data()),
                  "get_cauchy_stress_3d");
        }
        if (d.single_full_dis_vec_.size() != 3 * (size_t)d.nd_) reset_displacement(d);
        std::vector<double> elem(7 * (size_t)d.ne_), node(7 * (size_t)d.nd_);
        check(ofea_post_nodal(ctx_, mat.young_modulus, mat.poisson_ratio, d.single_full_dis_vec_.data(), quantity,
                              elem.data(), node.data()),
              "get_cauchy_stress_3d");
        d.stress_mat_.assign(d.ne_, std::vector<double>(7));
        for (int e = 0; e < d.ne_; ++e)
            for (int k = 0; k < 7; ++k) d.stress_mat_[e][k] = elem[(size_t)k * d.ne_ + e];
        d.stress_node_mat_.assign(d.nd_, std::vector<double>(7));
        for (int i = 0; i < d.nd_; ++i)
            for (int k = 0; k < 7; ++k) d.stress_node_mat_[i][k] = node[(size_t)k * d.nd_ + i];
        return true;
    }

private:
    ofea_ctx *ctx_ = nullptr;
    bool own_ = false;
};

class data_process {
public:
    // legacy ASCII VTK in the reference's dialect; the S_* fields are written
    // when get_cauchy_stress_3d has filled stress_node_mat_
    void export_dis_2_vtk(data_management &d, const std::string &path, double scale_dis)
    {
        const std::vector<double> xyz = d.flat_coords();
        const std::vector<int> conn = d.flat_topos();
        std::vector<double> fields;
        if (d.stress_node_mat_.size() == (size_t)d.nd_) {
            fields.assign(7 * (size_t)d.nd_, 0.0);
            for (int i = 0; i < d.nd_; ++i)
                for (int k = 0; k < 7; ++k) fields[(size_t)k * d.nd_ + i] = d.stress_node_mat_[i][k];
        }
        check(ofea_write_vtk(path.c_str(), d.nd_, xyz.data(), d.ne_, conn.data(), d.ele_list_idx_.data(),
                             d.single_full_dis_vec_.data(), scale_dis, fields.empty() ? nullptr : fields.data()),
              "export_dis_2_vtk");
    }
};

// Run options (src/include/cae.h:33-45).
struct controlPara {
    std::string work_path;
    std::string ifile_name;
    std::string ofile_name;
    int file_type = 1;
    int analysis_type = 1;  // 1 implicit, 2 explicit
    bool sfem_flag = false;
    bool reanalysis_flag = false;
    bool is_save_stiffness = true;
};

// The reference's driver class for the implicit (linear or NLFEA Newton)
// and explicit analyses.  Where the reference names a direct solver
// ("Pardiso_class"), gpu_solver_ names the GPU PCG method used instead.
// The deck is option_.ifile_name, or the constructor's path when that is
// empty.
class CAE_process {
public:
    controlPara option_;
    std::string path_;
    elastic_mat mat_;
    data_management data_cae_;
    std::string gpu_solver_ = "GPU_PCG_TWO_LEVEL";
    int explicit_precision_ = OFEA_FP64;  // arithmetic of the explicit element and node phases
    int explicit_max_steps_ = 10000000;   // guard: the reference's loop has none
    int device_ = 0;

    CAE_process() = default;
    explicit CAE_process(std::string path) : path_(std::move(path)) {}
    CAE_process(std::string path, elastic_mat mat) : path_(std::move(path)), mat_(mat) {}

    // src/cae.cpp:168-196
    void pre_info(const std::string &load_set_keyword, const std::string &load_value_keyword,
                  const std::string &dis_set_keyword)
    {
        ReadInfo item_info(option_.ifile_name.empty() ? path_ : option_.ifile_name);
        item_info.read_ele_node_num(data_cae_);
        item_info.read_geo_mesh(data_cae_);
        item_info.read_load_bcs(load_set_keyword, load_value_keyword, data_cae_);
        item_info.read_dis_bcs(dis_set_keyword, data_cae_);
        item_info.read_mat(mat_);
        if (option_.analysis_type == 2) item_info.read_time(data_cae_);
    }

    // src/cae.cpp:198-301.  The linear branch assembles K and solves with
    // gpu_solver_; the NLFEA branch runs the reference's Newton loop
    // (res_norm = ||F - f_int||_2 / (3 nfree) > res_lmit, at most 100 loops)
    // on assamble_nl_stiffness and solution_nl_api.  Then the displacement of
    // every node, the nodal S_* fields and the VTK file.  is_save_stiffness
    // is accepted for the signature: the CA re-analysis of this library
    // keeps K on the device itself (ofea_ca_*).
    void implict_analysis(const std::string &result_path, bool is_save_stiffness = false)
    {
        (void)is_save_stiffness;
        set_BCs item_bcs;
        item_bcs.build_free_index(data_cae_);
        item_bcs.build_single_load(data_cae_);
        ofea_ctx *ctx = nullptr;
        assamble_nl_stiffness item_nl_assam(device_);
        assamble_stiffness item_assam(device_);
        if (data_cae_.NLFEA) {
            double res_norm = 1.;
            const size_t n = data_cae_.single_load_vec_.size();
            item_nl_assam.build_CSR(data_cae_);
            std::vector<double> current_dis_vec(n, 0.), inter_force_vec(n, 0.), res_vec(n, 0.), incre_dis_vec(n);
            int loop = 0;
            while (res_norm > data_cae_.res_lmit && loop < 100) {
                std::fill(inter_force_vec.begin(), inter_force_vec.end(), 0.);
                item_nl_assam.fill_CSR_sparse_mat(data_cae_, mat_, current_dis_vec, inter_force_vec);
                for (size_t i = 0; i < n; ++i) res_vec[i] = data_cae_.single_load_vec_[i] - inter_force_vec[i];
                if (!solution_nl_api(item_nl_assam, data_cae_, res_vec, incre_dis_vec, gpu_solver_)) std::exit(1);
                res_norm = 0.;
                for (size_t i = 0; i < n; ++i) {
                    current_dis_vec[i] = current_dis_vec[i] + incre_dis_vec[i];
                    res_norm = res_norm + res_vec[i] * res_vec[i];
                }
                res_norm = std::sqrt(res_norm) / (double)n;
                loop++;
                std::cout << "this is loop " << loop << ", and res is " << res_norm << std::endl;
            }
            std::cout << "the loop is over " << std::endl;
            data_cae_.single_dis_vec_ = current_dis_vec;
            ctx = item_nl_assam.gpu_ctx();
        } else {
            item_assam.build_CSR(data_cae_);
            item_assam.fill_CSR_sparse_mat(data_cae_, mat_);
            if (!solution_api(item_assam, data_cae_, gpu_solver_)) std::exit(1);
            ctx = item_assam.gpu_ctx();
        }
        simulation_post post_item(ctx);
        post_item.reset_displacement(data_cae_);
        post_item.get_cauchy_stress_3d(data_cae_, mat_);
        data_process item_output;
        item_output.export_dis_2_vtk(data_cae_, result_path, 1.0);
    }

    // src/cae.cpp:420-574 on the GPU (ofea_explicit_analysis): lumped mass at
    // mat_.density, the load value of load_set_ stored as an int, dis_bc_set_
    // clamped, time_step_ (0 = automatic) and time_total_ of data_cae_, 10
    // outputs; every output k is written as result_path + "_" + k + ".vtk"
    // with single_full_dis_vec_ holding that displacement.  path_abaqus is
    // unused, as in the reference.
    void explicit_analysis(const std::string &result_path, const std::string &path_abaqus)
    {
        (void)path_abaqus;
        std::cout << "Explicit solving ......\n";
        data_management &d = data_cae_;
        std::vector<int> resort(d.nd_, -1);
        check(ofea_build_free_index(d.nd_, d.dis_bc_set_.data(), (int)d.dis_bc_set_.size(), resort.data()),
              "explicit_analysis");
        std::vector<int> fixed(d.nd_);
        for (int i = 0; i < d.nd_; ++i) fixed[i] = resort[i] < 0 ? 1 : 0;
        ofea_ctx *ctx = nullptr;
        check(ofea_create(&ctx, device_), "explicit_analysis");
        struct Guard {
            ofea_ctx *c;
            ~Guard() { ofea_destroy(c); }
        } guard{ctx};
        const std::vector<double> xyz = d.flat_coords();
        const std::vector<int> conn = d.flat_topos();
        check(ofea_set_mesh(ctx, d.nd_, xyz.data(), d.ne_, conn.data(), d.ele_list_idx_.data(), resort.data()),
              "explicit_analysis");
        std::vector<double> mass(d.nd_), f_ext(3 * (size_t)d.nd_);
        check(ofea_lumped_mass(ctx, mat_.density, mass.data()), "explicit_analysis");
        check(ofea_build_explicit_load(d.nd_, d.load_set_.data(), (int)d.load_set_.size(), d.load_dof_,
                                       d.load_value_, f_ext.data()),
              "explicit_analysis");
        const int output_gap = 10, max_frames = output_gap + 2;
        std::vector<double> frames(3 * (size_t)d.nd_ * max_frames), times(max_frames);
        std::vector<int> index(max_frames);
        int n_frames = 0, n_steps = 0;
        check(ofea_explicit_analysis(ctx, mat_.young_modulus, mat_.poisson_ratio, mass.data(), f_ext.data(), nullptr,
                                     fixed.data(), d.time_total_, d.time_step_, output_gap, explicit_precision_,
                                     explicit_max_steps_, max_frames, frames.data(), index.data(), times.data(),
                                     &n_frames, &n_steps),
              "explicit_analysis");
        data_process item_output;
        for (int k = 0; k < n_frames; ++k) {
            d.single_full_dis_vec_.assign(frames.begin() + 3 * (size_t)d.nd_ * k,
                                          frames.begin() + 3 * (size_t)d.nd_ * (k + 1));
            item_output.export_dis_2_vtk(d, result_path + "_" + std::to_string(index[k]) + ".vtk", 1.0);
        }
        explicit_steps_ = n_steps;
        std::cout << "Explicit finished\n";
    }

    int explicit_steps() const { return explicit_steps_; }

private:
    int explicit_steps_ = 0;
};

}  // namespace ofea_cae

#endif
