// Batched complex GEMM engine for gfx950 (MFMA).
//
//   C[b0,b1] (M x N) = alpha * sum_{kb < KB} sum_{k < K} op(A)[m][k] op(B)[k][n]  +  beta * C
//
// Every contraction of the HF/GW path is expressed through this one
// descriptor: each operand is a 2-D strided view (row stride, column stride)
// with an extra "K-batch" stride that lets one launch reduce over an outer
// index (spins, k-pairs, auxiliary index Q), and two batch indices with their
// own strides (or a per-b1 offset table for gathers over k-pairs).  Operand
// conjugation is applied while staging tiles; there is no transposed or
// conjugated copy of any operand.
#pragma once

#include "common.hpp"

namespace ggpu {

enum class Compute : int {
  F64 = 0,     // complex128 operands, v_mfma_f64_16x16x4_f64, FP64 accumulation
  F32 = 1,     // complex64 operands, v_mfma_f32_16x16x4_f32, FP32 accumulation
  BF16X3 = 2,  // complex64 operands split hi+lo into bf16, 3 products on
               // v_mfma_f32_16x16x32_bf16, FP32 accumulation
  I8X5 = 3,    // complex128 operands split into five signed 7-bit slices,
               // exact int32 slice products (s + t <= 5) on
               // v_mfma_i32_16x16x64_i8, FP64 recombination: ~2^-34 of the
               // operand scale
  I8X4 = 4,    // as I8X5 with four slices (s + t <= 4): ~2^-27 of the operand scale
};

// Operands of this compute mode are complex64 (else complex128).
inline bool complex64_operands(Compute c) { return c == Compute::F32 || c == Compute::BF16X3; }
// Slices of the int8 emulation of this compute mode (0: not emulated).
inline int int8_slices(Compute c) { return c == Compute::I8X5 ? 5 : c == Compute::I8X4 ? 4 : 0; }

struct Operand {
  const void* ptr = nullptr;   // element 0 of batch (0,0)
  int64_t rs = 0, cs = 0;      // stride of the first / second logical index (elements)
  int64_t kbs = 0;             // stride of the K-batch index
  int64_t bs0 = 0, bs1 = 0;    // batch strides
  const int64_t* off1 = nullptr;  // optional device table: offset of batch b1 (replaces bs1)
  int conj = 0;
  // Host-side bounds check: the allocation [buf, buf + buf_elems) that every
  // touched element must lie in (buf == nullptr: unchecked).
  const void* buf = nullptr;
  int64_t buf_elems = 0;
  const int64_t* off1_host = nullptr;  // host copy of off1 for the bounds check
};

struct OutOperand {
  void* ptr = nullptr;
  int64_t rs = 0, cs = 0;
  int64_t bs0 = 0, bs1 = 0;
  const int64_t* off1 = nullptr;
  const void* buf = nullptr;
  int64_t buf_elems = 0;
  const int64_t* off1_host = nullptr;
};

// Call sites of the HF/GW path.  Each heavy site launches its own kernel
// symbol (so profilers attribute time per site) and is timed separately when
// a GemmProfiler is installed.
enum Site : int {
  kSiteOther = 0,
  kSiteP0X1 = 1,     // X1 = G(k, beta - tau) V          (P0, first product)
  kSiteP0X2 = 2,     // X2 = G(k+q, tau) V^dagger        (P0, second product)
  kSiteP0 = 3,       // P0 += X1 . X2                    (P0 contraction)
  kSiteSigY1 = 4,    // Y1 = V G(k-q, tau)               (Sigma, first product)
  kSiteSigY2 = 5,    // Y2 = P~(tau) . Y1                (Sigma, screened-interaction product)
  kSiteSigma = 6,    // Sigma -= Y2 . V^dagger           (Sigma contraction)
  kSiteIR = 7,       // tau <-> Omega transforms
  kSiteDyson = 8,    // Cholesky updates and triangular-solve updates
  kSiteHF = 9,       // HF exchange / Hartree / Ewald products
  kSiteRotate = 10,  // k- and q-symmetry rotations
  kNumSites = 11,
};

struct GemmDesc;
// Hook called around every launch (installed per thread).
struct GemmProfiler {
  virtual ~GemmProfiler() = default;
  virtual void before(const GemmDesc& d, hipStream_t s) = 0;
  virtual void after(const GemmDesc& d, hipStream_t s) = 0;
};
void set_gemm_profiler(GemmProfiler* p);
// Workgroup tile of the split-bf16 x3 GEMMs: chosen per shape (-1, the
// default) or fixed to one of these for every launch (all give the same sums).
// (the W3 tiles allocate registers for three waves per SIMD; their LDS
// allows three workgroups per CU).
enum {
  kBf16Tile64x64 = 0, kBf16Tile128x64 = 1, kBf16Tile64x128 = 4, kBf16Tile160x64 = 5,
  kBf16Tile96x64W3 = 6, kBf16Tile64x96W3 = 7, kBf16Tile64x64W3 = 8,
  // 64 x 64, A split once and kept in registers while the workgroup walks
  // several batches (one A for every batch, K <= 64, no K-batch; other
  // launches take the 64 x 64 three-wave tile)
  kBf16TileAstat64x64 = 9,
  // 128 x 128 from operands split once into fragment-native bf16 planes in
  // the plan's workspace (no workspace: the 160 x 64 tile)
  kBf16TileFrag128x128 = 10,
  kBf16TileFrag128x128W8 = 11,   // the same with eight waves (two per SIMD)
  kBf16TileFrag128x128W16 = 12,  // and with sixteen (four per SIMD)
  kBf16TileFrag128x192W8 = 13    // 128 x 192, eight waves
};
// Variant of a Compute::I8X5 / I8X4 launch that runs the int8 emulation with
// the mode's slice count; kVariantInt8x5 runs it with five slices in either
// mode (FP64-accurate products, so any site of a four-slice plan may take
// it); the FP64 kernel variants (8, 16, 83, 84) run the launch on the FP64
// kernel instead.  Built-in choice: the emulation at the Sigma P.Y1 product
// and the P0 contraction (the sites where it is faster on the benchmark
// shapes), the FP64 kernel everywhere else; per-site variants override it.
enum { kVariantInt8 = 200, kVariantInt8x5 = 201 };
// Launch variant per call site for the launches of the calling thread
// (table of kNumSites entries, -1: the built-in choice; nullptr: none):
// for FP64 / FP32 8 or 16 forces the LDS k-depth, 83 or 84 k-depth 8 with
// registers allocated for 3 or 4 waves per SIMD (all give the same sums);
// for split-bf16 a kBf16Tile* value; in the int8 modes kVariantInt8 or an
// FP64 kernel variant.
void set_site_variants(const int* table);
// Device workspace of the Compute::I8X5 / I8X4 launches of the calling thread
// (nullptr: none; such launches then fail with an error).
void set_gemm_workspace(void* ptr, size_t bytes);
// Variants worth timing for a compute mode (the autotuner's candidates);
// returns their count, writes at most cap of them.
int gemm_variant_candidates(Compute c, int* out, int cap);
bool gemm_variant_valid(Compute c, int variant);
// Whether the autotuner may pick the int8 emulation variant `variant`
// (kVariantInt8 or kVariantInt8x5) at this call site: five slices anywhere;
// four slices only at the P0 contraction and the Sigma P.Y1 product
// (four-slice emulation at all six contraction sites leaves ~1.1e-6 Ha in
// Sigma of the nk = 216, nao = 104, NQ = 600 benchmark q, two sites
// ~3.5e-7).  An explicit per-site variant may still pick it.
inline bool int8_site_tunable(Compute c, int site, int variant = kVariantInt8) {
  if (c == Compute::I8X5 || variant == kVariantInt8x5) return true;
  return c == Compute::I8X4 && (site == kSiteP0 || site == kSiteSigY2);
}

struct GemmDesc {
  int site = kSiteOther;
  int batch_major = 0;   // set by gemm(): batch index fastest in the tile order
  int M = 0, N = 0, K = 0, KB = 1;
  int nb0 = 1, nb1 = 1;
  Operand A, B;
  OutOperand C;
  double alpha_re = 1.0, alpha_im = 0.0;
  double beta_re = 0.0, beta_im = 0.0;
  Compute compute = Compute::F64;
  int out_f32 = 0;             // C holds complex64 (only with F32 / BF16X3 compute)
};

// Throws Error(kErrArgument) if any address the launch would touch falls
// outside an operand's declared allocation.
void gemm_check_bounds(const GemmDesc& d);

// Enqueue the GEMM on `stream`.
void gemm(const GemmDesc& d, hipStream_t stream);

// Algorithmic FLOPs of the launch: 8 per complex multiply-add.
inline double gemm_flops(const GemmDesc& d) {
  return 8.0 * d.M * (double)d.N * d.K * d.KB * d.nb0 * d.nb1;
}

}  // namespace ggpu
