// The reference's linear static pipeline -- CAE_process::pre_info and
// implict_analysis (src/cae.cpp:168-301) -- written against the C++ host
// layer include/openfea_amd_cae.hpp, in the reference's call order.
//
//   implicit_pipeline DECK OUT_PREFIX [SOLVER]   full pipeline on device 0
//   implicit_pipeline DECK OUT_PREFIX --host-only reader, set_BCs and the
//                                                  host pattern (no device)
//
// Raw little-endian arrays go to OUT_PREFIX.<name>.<i32|f64>; the full run
// also writes OUT_PREFIX.vtk.  The last line printed is a summary.
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "openfea_amd_cae.hpp"

using namespace ofea_cae;

template <class T>
static void dump(const std::string &path, const std::vector<T> &v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(1);
    }
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s DECK OUT_PREFIX [SOLVER | --host-only]\n", argv[0]);
        return 2;
    }
    const std::string deck = argv[1], out = argv[2];
    const std::string mode = argc > 3 ? argv[3] : "GPU_PCG";

    // pre_info (src/cae.cpp:168-196)
    data_management d;
    elastic_mat mat{2.1e5, 0.3, 7800};  // the material src/main.cpp:45 hands to CAE_process
    ReadInfo info(deck);
    info.read_ele_node_num(d);
    info.read_geo_mesh(d);
    info.read_load_bcs("Set-load", "*Cload", d);
    info.read_dis_bcs("Set-fix", d);
    info.read_mat(mat);

    // implict_analysis (src/cae.cpp:198-207)
    set_BCs item_bcs;
    item_bcs.build_free_index(d);
    item_bcs.build_single_load(d);
    dump(out + ".resort.i32", d.resort_free_nodes_);
    dump(out + ".load.f64", d.single_load_vec_);
    dump(out + ".fixsorted.i32", d.dis_bc_set_);

    if (mode == "--host-only") {
        int n = 0;
        long long nnz = 0;
        const std::vector<int> conn = d.flat_topos();
        check(ofea_build_csr(d.nd_, d.ne_, conn.data(), d.ele_list_idx_.data(), d.resort_free_nodes_.data(), &n,
                             &nnz, nullptr, nullptr),
              "ofea_build_csr");
        std::vector<int> col(n + 1), row(nnz);
        check(ofea_build_csr(d.nd_, d.ne_, conn.data(), d.ele_list_idx_.data(), d.resort_free_nodes_.data(),
                             nullptr, nullptr, col.data(), row.data()),
              "ofea_build_csr");
        dump(out + ".col.i32", col);
        dump(out + ".row.i32", row);
        std::printf("summary nd=%d ne=%d n=%d nnz=%lld E=%.17g nu=%.17g load_dof=%d load_value=%.17g\n", d.nd_,
                    d.ne_, n, nnz, mat.young_modulus, mat.poisson_ratio, d.load_dof_, d.load_value_);
        return 0;
    }

    // assembly and solve (src/cae.cpp:255-276)
    assamble_stiffness item_assam;
    item_assam.build_CSR(d);
    item_assam.fill_CSR_sparse_mat(d, mat);
    dump(out + ".col.i32", item_assam.col_idx_);
    dump(out + ".row.i32", item_assam.row_idx_);
    dump(out + ".val.f64", item_assam.nz_val_);
    if (!solution_api(item_assam, d, mode)) return 3;
    const bool unknown_rejected = !solution_api(item_assam, d, "no_such_solver");

    // post-processing and output (src/cae.cpp:284-292)
    simulation_post post_item(item_assam.gpu_ctx());
    post_item.reset_displacement(d);
    post_item.get_cauchy_stress_3d(d, mat);
    data_process item_output;
    item_output.export_dis_2_vtk(d, out + ".vtk", 1.0);
    dump(out + ".u.f64", d.single_full_dis_vec_);
    std::vector<double> S;
    for (const auto &row : d.stress_node_mat_) S.insert(S.end(), row.begin(), row.end());
    dump(out + ".S.f64", S);
    std::printf("summary nd=%d ne=%d n=%d nnz=%d iters=%d relres=%.3e unknown_rejected=%d\n", d.nd_, d.ne_,
                item_assam.num_row_, item_assam.num_nz_val_, solution_options().iters, solution_options().relres,
                unknown_rejected ? 1 : 0);
    return 0;
}
