// The reference's driver class CAE_process (src/include/cae.h:47-91) on the
// C++ host layer include/openfea_amd_cae.hpp: pre_info, then the NLFEA
// branch of implict_analysis (src/cae.cpp:209-253) or explicit_analysis
// (src/cae.cpp:420-574), called as src/main.cpp calls them.
//
//   cae_process DECK OUT_PREFIX nl [SOLVER]
//       NLFEA Newton loop; writes OUT_PREFIX.vtk and OUT_PREFIX.u.f64
//   cae_process DECK OUT_PREFIX explicit TIME_TOTAL [TIME_STEP [fp32]]
//       explicit loop; writes OUT_PREFIX_<k>.vtk and the last output as
//       OUT_PREFIX.u.f64
//
// The last line printed is a summary.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "openfea_amd_cae.hpp"

using namespace ofea_cae;

static void dump(const std::string &path, const std::vector<double> &v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(1);
    }
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s DECK OUT_PREFIX nl [SOLVER] | explicit TIME_TOTAL [TIME_STEP [fp32]]\n",
                     argv[0]);
        return 2;
    }
    const std::string deck = argv[1], out = argv[2], mode = argv[3];
    CAE_process cae_item(deck, elastic_mat{2.1e5, 0.3, 7800});  // src/main.cpp:155-163
    if (mode == "nl") {
        cae_item.data_cae_.NLFEA = true;
        if (argc > 4) cae_item.gpu_solver_ = argv[4];
        cae_item.pre_info("Set-load", "*Cload", "Set-fix");
        cae_item.implict_analysis(out + ".vtk");
        dump(out + ".u.f64", cae_item.data_cae_.single_full_dis_vec_);
        std::printf("summary mode=nl nd=%d ne=%d pcg_iters_last=%d\n", cae_item.data_cae_.nd_,
                    cae_item.data_cae_.ne_, solution_options().iters);
        return 0;
    }
    if (mode == "explicit") {
        if (argc < 5) return 2;
        cae_item.option_.analysis_type = 2;
        cae_item.pre_info("Set-load", "*Cload", "Set-fix");
        // the time can be set by hand after pre_info (src/main.cpp:168-171)
        cae_item.data_cae_.time_total_ = std::atof(argv[4]);
        cae_item.data_cae_.time_step_ = argc > 5 ? std::atof(argv[5]) : 0.0;
        if (argc > 6 && std::string(argv[6]) == "fp32") cae_item.explicit_precision_ = OFEA_FP32;
        cae_item.explicit_analysis(out, "");
        dump(out + ".u.f64", cae_item.data_cae_.single_full_dis_vec_);
        std::printf("summary mode=explicit nd=%d ne=%d steps=%d\n", cae_item.data_cae_.nd_, cae_item.data_cae_.ne_,
                    cae_item.explicit_steps());
        return 0;
    }
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
