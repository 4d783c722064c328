"""openfea_amd -- MI355X-native (gfx950) implementation of OpenFEA's element hot path.

The numerical work lives in the C-ABI library libofea_hip.so (hand-written HIP
kernels, include/openfea_amd.h); this package is its host-side mirror of the
reference's call surface (fea.py) plus synthetic workload generators (meshgen.py).
"""
from . import _capi, meshgen  # noqa: F401
from .fea import (AssambleStiffness, CA_ReAnalysis, CAInfo, Context, Deck, build_csr, build_free_index,  # noqa: F401
                  build_single_load, export_dis_2_vtk, get_cauchy_stress_3d, implict_analysis, implict_SFEManalysis,
                  read_ca_info, read_inp,
                  reset_ca_displacement, reset_displacement, solution_api)

__all__ = ["AssambleStiffness", "CA_ReAnalysis", "CAInfo", "Context", "Deck", "build_csr", "build_free_index",
           "build_single_load", "export_dis_2_vtk", "get_cauchy_stress_3d", "implict_analysis", "implict_SFEManalysis",
           "read_ca_info", "read_inp",
           "reset_ca_displacement", "reset_displacement", "solution_api", "meshgen"]
