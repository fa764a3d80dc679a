// GEMM call-site benchmark: launches the six GEMM call sites of one GW
// iteration with the shapes, strides and batching the engine uses (pairs per
// batch kb and P0 partial slices as the plan chooses them; defaults: the
// nk = 216 system, nao = 104, NQ = 600, ns = 2, nts = 110, one pair per
// batch, 8 P0 slices), through the library's own dispatch, and prints the
// per-launch median time of each (HIP events on the launch stream).
// Profilers attribute the same kernels (one kernel symbol per call site), so
// their per-kernel durations can be compared with these.
//
//   gemm_site_bench <fp64|fp32|bf16x3|i8x5|i8x4> [--nao N] [--nq N] [--nts N] [--kb N]
//                   [--p0-split S] [--reps R] [--sites a,b,..] [--calibrate-copy]
//
// i8x5 / i8x4 force the int8 emulation at every site (the FP64 kernel is fp64).
// --calibrate-copy ends with one plain 1 GiB device-to-device copy, the
// reference point for scaling HBM byte counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "zgemm.hpp"

using namespace ggpu;

namespace {

struct CallSite {
  const char* name;
  int site, M, N, K, KB, nb0, nb1;
  int64_t ars, acs, akbs, abs0, abs1;
  int64_t brs, bcs, bkbs, bbs0, bbs1;
  int64_t crs, ccs, cbs0, cbs1;
  int conj;      // 1: A conjugated, 2: B conjugated
  int out_low;   // C in the reduced type in the reduced modes
  double alpha, beta;
};

// Operand values: uniform in [-1, 1) from a counter hash of the element index
// (constant data would lower the power draw, and with it the clocks, below
// what real operands see).
__global__ void fill_kernel(float* x, int64_t n, uint32_t salt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t h = (uint32_t)i * 0x9E3779B1u ^ (uint32_t)(i >> 32) ^ salt;
    h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
    x[i] = (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
  }
}
__global__ void fill_kernel64(double* x, int64_t n, uint32_t salt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t h = (uint32_t)i * 0x9E3779B1u ^ (uint32_t)(i >> 32) ^ salt;
    h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
    x[i] = (double)h * (2.0 / 4294967296.0) - 1.0;
  }
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "HIP error %s at %s\n", hipGetErrorString(e), what);
    std::exit(1);
  }
}

// Largest element offset + 1 a strided view reaches.
int64_t extent(int64_t r, int64_t nr, int64_t c, int64_t nc, int64_t kb, int64_t nkb, int64_t b0, int64_t nb0,
               int64_t b1, int64_t nb1) {
  return (nr - 1) * r + (nc - 1) * c + (nkb - 1) * kb + (nb0 - 1) * b0 + (nb1 - 1) * b1 + 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr,
                 "usage: %s <fp64|fp32|bf16x3|i8x5|i8x4> [--nao N] [--nq N] [--nts N] [--kb N] [--p0-split S] "
                 "[--reps R] [--sites a,b] [--variant V] [--calibrate-copy]\n",
                 argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  if (mode != "fp64" && mode != "fp32" && mode != "bf16x3" && mode != "i8x5" && mode != "i8x4") {
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
  }
  const Compute comp = mode == "fp64" ? Compute::F64 : mode == "fp32" ? Compute::F32
                       : mode == "i8x5" ? Compute::I8X5
                       : mode == "i8x4" ? Compute::I8X4 : Compute::BF16X3;
  int reps = 10, nao = 104, NQ = 600, nts = 110, kb = 1, S = 8, variant = -2;
  std::string only;
  bool calibrate = false;
  for (int i = 2; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() { return i + 1 < argc ? std::atoi(argv[++i]) : 0; };
    if (a == "--calibrate-copy") calibrate = true;
    else if (a == "--nao") nao = next();
    else if (a == "--nq") NQ = next();
    else if (a == "--nts") nts = next();
    else if (a == "--kb") kb = next();
    else if (a == "--p0-split") S = next();
    else if (a == "--reps") reps = std::max(1, next());
    else if (a == "--variant") variant = next();
    else if (a == "--sites" && i + 1 < argc) only = argv[++i];
    else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  if (nao < 1 || NQ < 1 || nts < 4 || nts % 2 || kb < 1 || S < 1) {
    std::fprintf(stderr, "bad sizes\n");
    return 2;
  }
  const int half = nts / 2, ns = 2;
  const int64_t n2 = (int64_t)nao * nao, stride = (int64_t)NQ * n2, Q2 = (int64_t)NQ * NQ;
  const int64_t xs = (int64_t)half * stride, ys = (int64_t)nts * stride;
  // P0 partial slices as GwPlan::p0_for_q forms them: E = kb ns (pair, spin)
  // entries, slice j sums entries j, j + S, ...; with fewer entries than
  // slices every entry is its own slice
  const int E = kb * ns, full = E / S;
  const int p0_kb = full > 0 ? full : 1, p0_slices = full > 0 ? S : E;
  const std::vector<CallSite> sites = {
      // X1 = G(k, beta - tau) V (one pair, one spin)
      {"p0_gv", kSiteP0X1, half * nao, nao * NQ, nao, 1, 1, 1, nao, 1, 0, 0, 0, (int64_t)nao * NQ, 1, 0, 0, 0,
       (int64_t)nao * NQ, 1, 0, 0, 0, 1, 1.0, 0.0},
      // X2 = G(k+q, tau) V^dagger per tau
      {"p0_gvh", kSiteP0X2, NQ * nao, nao, nao, 1, half, 1, nao, 1, 0, 0, 0, 1, nao, 0, n2, 0, nao, 1, stride, 0, 1,
       1, 1.0, 0.0},
      // P0 partial slices += X1 . X2 over the (pair, spin) K-batch
      {"p0_contract", kSiteP0, NQ, NQ, (int)n2, p0_kb, half, p0_slices, 1, NQ, (int64_t)S * xs, stride, xs, 1, n2,
       (int64_t)S * xs, stride, xs, NQ, 1, Q2, half * Q2, 0, 0, -1.0, 1.0},
      // Y1 = V G(k-q, tau) per tau
      {"sig_vg", kSiteSigY1, NQ * nao, nao, nao, 1, nts, 1, nao, 1, 0, 0, 0, nao, 1, 0, n2, 0, nao, 1, stride, 0, 0,
       1, 1.0, 0.0},
      // Y2 = P~(tau)^T Y1 per tau
      {"sig_py", kSiteSigY2, NQ, (int)n2, NQ, 1, nts, 1, 1, NQ, 0, Q2, 0, n2, 1, 0, stride, 0, (int64_t)nts * n2, 1,
       n2, 0, 0, 1, 1.0, 0.0},
      // Sigma -= Y2 . V^dagger over the Q K-batch, for every pair and spin of a batch
      {"sig_contract", kSiteSigma, nts * nao, nao, nao, NQ, kb, ns, nao, 1, (int64_t)nts * n2, ns * ys, ys, 1, nao,
       n2, stride, 0, nao, 1, (int64_t)ns * nts * n2, (int64_t)nts * n2, 2, 0, -1.0, 1.0},
  };
  // a workspace as the plans have one (int8 slices, pre-split bf16 planes)
  // and the launch variant at every site: --variant, else the int8
  // emulation in the int8 modes and the built-in choice otherwise
  void* work = nullptr;
  size_t work_bytes = 0;
  int variants[kNumSites];
  if (variant != -2 && !gemm_variant_valid(comp, variant)) {
    std::fprintf(stderr, "variant %d not valid in mode %s\n", variant, mode.c_str());
    return 2;
  }
  std::fill(variants, variants + kNumSites, variant != -2 ? variant : int8_slices(comp) ? (int)kVariantInt8 : -1);
  if (int8_slices(comp) || comp == Compute::BF16X3) {
    size_t fr = 0, tot = 0;
    check(hipMemGetInfo(&fr, &tot), "meminfo");
    work_bytes = std::min<size_t>((size_t)24 << 30, fr / 4);
    check(hipMalloc(&work, work_bytes), "malloc workspace");
    set_gemm_workspace(work, work_bytes);
  }
  set_site_variants(variants);

  std::printf("GEMM call sites, nao=%d NQ=%d ns=%d nts=%d, %d pairs per batch, %d P0 slices, mode %s, %d repetitions\n",
              nao, NQ, ns, nts, kb, S, mode.c_str(), reps);
  std::printf("%-13s %8s %8s %8s %5s %6s %12s %10s\n", "site", "M", "N", "K", "KB", "batch", "ms/launch", "TFLOP/s");
  for (const CallSite& s : sites) {
    if (!only.empty() && ("," + only + ",").find("," + std::string(s.name) + ",") == std::string::npos) continue;
    const bool low = complex64_operands(comp);
    const size_t ein = low ? 8 : 16, eout = (low && s.out_low) ? 8 : 16;
    const int64_t ea = extent(s.ars, s.M, s.acs, s.K, s.akbs, s.KB, s.abs0, s.nb0, s.abs1, s.nb1);
    const int64_t eb = extent(s.brs, s.K, s.bcs, s.N, s.bkbs, s.KB, s.bbs0, s.nb0, s.bbs1, s.nb1);
    const int64_t ec = extent(s.crs, s.M, s.ccs, s.N, 0, 1, s.cbs0, s.nb0, s.cbs1, s.nb1);
    void *A = nullptr, *B = nullptr, *C = nullptr;
    check(hipMalloc(&A, ea * ein), "malloc A");
    check(hipMalloc(&B, eb * ein), "malloc B");
    check(hipMalloc(&C, ec * eout), "malloc C");
    // real parts and imaginary parts are separate scalars: 2 per element
    if (low) {
      hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, nullptr, (float*)A, 2 * ea, 1u);
      hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, nullptr, (float*)B, 2 * eb, 2u);
    } else {
      hipLaunchKernelGGL(fill_kernel64, dim3(4096), dim3(256), 0, nullptr, (double*)A, 2 * ea, 1u);
      hipLaunchKernelGGL(fill_kernel64, dim3(4096), dim3(256), 0, nullptr, (double*)B, 2 * eb, 2u);
    }
    check(hipGetLastError(), "fill");
    check(hipMemset(C, 0, ec * eout), "clear C");
    GemmDesc d;
    d.site = s.site;
    d.compute = comp;
    d.out_f32 = (low && s.out_low) ? 1 : 0;
    d.M = s.M; d.N = s.N; d.K = s.K; d.KB = s.KB; d.nb0 = s.nb0; d.nb1 = s.nb1;
    d.A.ptr = A; d.A.buf = A; d.A.buf_elems = ea;
    d.A.rs = s.ars; d.A.cs = s.acs; d.A.kbs = s.akbs; d.A.bs0 = s.abs0; d.A.bs1 = s.abs1; d.A.conj = s.conj & 1;
    d.B.ptr = B; d.B.buf = B; d.B.buf_elems = eb;
    d.B.rs = s.brs; d.B.cs = s.bcs; d.B.kbs = s.bkbs; d.B.bs0 = s.bbs0; d.B.bs1 = s.bbs1; d.B.conj = (s.conj >> 1) & 1;
    d.C.ptr = C; d.C.buf = C; d.C.buf_elems = ec;
    d.C.rs = s.crs; d.C.cs = s.ccs; d.C.bs0 = s.cbs0; d.C.bs1 = s.cbs1;
    d.alpha_re = s.alpha; d.beta_re = s.beta;
    try {
      gemm(d, nullptr);   // warm-up (and the bounds check of every operand)
    } catch (const std::exception& e) {
      std::fprintf(stderr, "%s: %s\n", s.name, e.what());
      return 1;
    }
    check(hipDeviceSynchronize(), "warm-up");
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    std::vector<float> t;
    for (int r = 0; r < reps; ++r) {
      check(hipEventRecord(e0, nullptr), "record");
      gemm(d, nullptr);
      check(hipEventRecord(e1, nullptr), "record");
      check(hipEventSynchronize(e1), "sync");
      float ms = 0;
      check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    const double med = t[t.size() / 2];
    const double flop = 8.0 * s.M * s.N * (double)s.K * s.KB * s.nb0 * s.nb1;
    std::printf("%-13s %8d %8d %8d %5d %6d %12.4f %10.2f\n", s.name, s.M, s.N, s.K, s.KB, s.nb0 * s.nb1, med,
                flop / (med * 1e-3) / 1e12);
    std::fflush(stdout);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(A);
    hipFree(B);
    hipFree(C);
  }
  if (work) {
    set_gemm_workspace(nullptr, 0);
    set_site_variants(nullptr);
    hipFree(work);
  }
  if (calibrate) {
    const size_t bytes = (size_t)1 << 30;
    void *src = nullptr, *dst = nullptr;
    check(hipMalloc(&src, bytes), "malloc src");
    check(hipMalloc(&dst, bytes), "malloc dst");
    check(hipMemset(src, 1, bytes), "fill src");
    check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "copy");
    check(hipDeviceSynchronize(), "copy sync");
    std::printf("calibration copy: %zu bytes read, %zu bytes written\n", bytes, bytes);
    hipFree(src);
    hipFree(dst);
  }
  return 0;
}
