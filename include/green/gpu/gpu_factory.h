/*
 * Drop-in for green-gpu's <green/gpu/gpu_factory.h> (gpu_factory.h:29-102):
 * the three functions green-mbpt's custom-kernel registration calls, with
 * the same signatures, backed by the gfx950 library through the C ABI
 * (include/green_gpu_amd.h).  Build: add include/ to the include path and
 * link libgreen_gpu_amd.so and HDF5 (green-mbpt links HDF5 already).
 */
#ifndef GREEN_GPU_FACTORY_H
#define GREEN_GPU_FACTORY_H

#include <functional>
#include <memory>
#include <tuple>

#include "amd_kernels.h"

namespace green::gpu {
  using bz_utils_t = symmetry::brillouin_zone_utils;
  using x_type     = ztensor<4>;
  using G_type     = utils::shared_object<ztensor<5>>;

  /**
   * HF kernel extension (reference: gpu_factory.h:47-62).  nso != nao selects
   * the two-component kernel, as in the reference; X2C must agree with it.
   */
  inline std::tuple<std::shared_ptr<void>, std::function<x_type(const x_type&)>> custom_hf_kernel(
      bool X2C, const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ, double madelung,
      const bz_utils_t& bz_utils, const ztensor<4>& S_k) {
    if (X2C != (nso != nao)) throw std::runtime_error("custom_hf_kernel: X2C requires nso = 2 nao");
    auto kernel = std::make_shared<amd::hf_kernel>(p, nao, nso, ns, NQ, madelung, bz_utils, S_k);
    std::function<x_type(const x_type&)> callback = [kernel](const x_type& dm) -> x_type { return kernel->solve(dm); };
    return std::tuple{std::static_pointer_cast<void>(kernel), callback};
  }

  /**
   * GW kernel extension (reference: gpu_factory.h:78-89).  The linear solver
   * comes from the "cuda_linear_solver" parameter and the precision from
   * P_sp && Sigma_sp (gw_gpu_kernel.h:64), as in the reference, unless
   * "gpu_precision" names one; "gpu_autotune" and "gpu_graph" switch on the
   * launch-variant tuning and graph replay.
   */
  inline std::tuple<std::shared_ptr<void>, std::function<void(G_type&, G_type&)>> custom_gw_kernel(
      bool X2C, const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ, const grids::transformer_t& ft,
      const bz_utils_t& bz_utils, const ztensor<4>& S_k) {
    (void)S_k;
    if (X2C != (nso != nao)) throw std::runtime_error("custom_gw_kernel: X2C requires nso = 2 nao");
    auto kernel = std::make_shared<amd::gw_kernel>(p, nao, nso, ns, NQ, ft, bz_utils,
                                                   p["cuda_linear_solver"].as<LinearSolverType>());
    std::function<void(G_type&, G_type&)> callback = [kernel](G_type& g, G_type& s) { kernel->solve(g, s); };
    return std::tuple{std::static_pointer_cast<void>(kernel), callback};
  }

  /**
   * Kernel parameters (reference: gpu_factory.h:95-102).  The memory-mode and
   * tau-batch switches are accepted for compatibility; this implementation
   * sizes its device batches itself.
   */
  inline void custom_kernel_parameters(params::params& p) {
    p.define<int>("verbose", "verbosity level of the GPU kernels", 0);
    p.define<LinearSolverType>("cuda_linear_solver", "solver of the screened-interaction equation: LU or Cholesky",
                               LinearSolverType::LU);
    p.define<bool>("cuda_low_gpu_memory", "accepted for compatibility (device batches are sized automatically)");
    p.define<bool>("cuda_low_cpu_memory",
                   "host has little memory: read the integral chunks of each pair batch on demand (one chunk "
                   "cached) instead of holding every stored pair");
    p.define<size_t>("nt_batch", "accepted for compatibility (tau batches are sized automatically)", 0);
    // no counterpart in the reference: arithmetic, launch-variant tuning and
    // graph replay of this implementation
    p.define<std::string>("gpu_precision",
                          "contraction arithmetic: fp64, fp32, bf16x3 (split-bf16 x3 on the matrix cores), i8x5 or "
                          "i8x4 (FP64 operands sliced onto the int8 matrix cores, five or four slices); "
                          "empty: fp32 for GW when P_sp and Sigma_sp, else fp64",
                          "");
    p.define<bool>("gpu_autotune", "time every GEMM launch variant on the first GW iteration, keep the fastest", false);
    p.define<bool>("gpu_graph", "replay every GW iteration as one captured HIP graph", false);
  }

}  // namespace green::gpu

#endif  // GREEN_GPU_FACTORY_H
