// FP64-accurate complex GEMM on the gfx950 int8 matrix cores
// (v_mfma_i32_16x16x64_i8): every FP64 operand is split into five signed
// 7-bit slices with one power-of-two scale per row of A and per column of B
// (Ozaki-style splitting), the slice products with s + t <= 5 are summed
// exactly in int32 (one accumulator per slice diagonal s + t), and the
// diagonals are recombined in FP64:
//
//   x 2^sigma = sum_s x_s 128^-s + r,  |x_s| <= 127,  |r| < 128^-4 (relative to the row maximum)
//   (A B)[m][n] = 2^-(sigma_m + sigma_n) sum_d 128^-d sum_{s+t=d} (A_s B_t)[m][n]
//
// The dropped terms (s + t > 5, and the slice remainders) leave a relative
// error of about 2^-36 of max|A_m| max|B_n| K per element (a numpy model on
// Gaussian operands: ~2e-10 of max|C|), far below complex64 arithmetic.
// Complex products use four real products per slice pair with the
// imaginary part of A also stored negated, so re and im accumulate exactly:
//   re += Ar_s Br_t + (-Ai_s) Bi_t,   im += Ar_s Bi_t + Ai_s Br_t.
//
// A launch runs three kernels: the per-row / per-column maximum binary
// exponent (over K and the K-batch), the slicing of each operand into a
// canonical int8 layout, and the int8 GEMM with the FP64 epilogue (alpha,
// beta, strided or complex64 output).
#include "ozaki.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>

namespace ggpu {
namespace {

constexpr int NSL = 5;              // slices per operand
constexpr int DMAX = 5;             // kept slice pairs: s + t <= DMAX (19 pairs)
constexpr int PA = 3 * NSL;         // A planes per k block: re, im, -im of every slice
constexpr int PB = 2 * NSL;         // B planes: re, im of every slice
constexpr int KBLK = 64;            // k per block: one v_mfma_i32_16x16x64_i8
constexpr int TILE = 64;            // rows of a workgroup tile (both operands)
constexpr int NT = 256;
constexpr int LDS_PAD = 16;         // bytes of padding per LDS row (bank spread)
constexpr int ROWA = PA * KBLK;     // bytes of one A row of a k block in global memory
constexpr int ROWB = PB * KBLK;
// Largest K * KB one launch may sum: every int32 diagonal accumulator holds at
// most NSL pairs * 2 products * K * 127^2 < 2^31 (no diagonal has more than
// NSL slice pairs).
constexpr int64_t KMAX = 12800;

typedef int v4i __attribute__((ext_vector_type(4)));

struct SliceView {
  const double2* src = nullptr;   // element (row 0, k 0, kb 0, batch 0)
  int64_t row_s = 0, k_s = 0, kb_s = 0, bs0 = 0, bs1 = 0;
  const int64_t* off1 = nullptr;  // device: offset of batch b1 (replaces bs1)
  int rows = 0, K = 0, KB = 1, nb1 = 1;
  int conj = 0;
  int nplanes = 0;                // PA (A side, re / im / -im) or PB (B side)
  int Rp = 0, nkb = 0;            // padded rows, k blocks
  int* E = nullptr;               // [batch][Rp]: max binary exponent (INT_MIN: all zero)
  int8_t* out = nullptr;          // [batch][kblock][Rp][nplanes][64]
};

// Load the 64 x 64 tile (rows r0.., flattened k k0..) of batch b into LDS,
// zero outside the operand; lanes run along whichever index is contiguous.
// The flattened k of one launch stays below 2^31 (K * KB <= 12800).
__device__ inline void load_tile(const SliceView& v, int64_t b, int r0, int k0, double2 (*T)[KBLK + 1]) {
  const int t = threadIdx.x;
  const int kt = v.K * v.KB;
  const bool k_fast = v.k_s == 1 || v.row_s != 1;
  const int64_t b0 = b / v.nb1, b1 = b - b0 * v.nb1;
  const double2* base = v.src + b0 * v.bs0 + (v.off1 ? v.off1[b1] : b1 * v.bs1);
  for (int it = 0; it < TILE * KBLK / NT; ++it) {
    int r, c;
    if (k_fast) { r = it * (NT / KBLK) + t / KBLK; c = t % KBLK; }
    else { c = it * (NT / TILE) + t / TILE; r = t % TILE; }
    const int row = r0 + r, kk = k0 + c;
    double2 x = make_double2(0.0, 0.0);
    if (row < v.rows && kk < kt) {
      const int kb = kk / v.K, k = kk - kb * v.K;
      x = base[(int64_t)row * v.row_s + (int64_t)k * v.k_s + (int64_t)kb * v.kb_s];
    }
    T[r][c] = x;
  }
}

__device__ inline int max_exponent(double a) {
  if (a == 0.0) return INT_MIN;
  int e;
  (void)frexp(a, &e);
  return e;
}

// E[b][row] = max over (k, kb) of the binary exponent of max(|re|, |im|).
// Grid: (k blocks, row tiles, batches); E starts at INT_MIN.
__global__ __launch_bounds__(NT) void oz_exponent_kernel(const SliceView v) {
  __shared__ double2 T[TILE][KBLK + 1];
  __shared__ int red[NT];
  const int64_t b = blockIdx.z;
  const int r0 = blockIdx.y * TILE;
  const int k0 = blockIdx.x * KBLK;
  load_tile(v, b, r0, k0, T);
  __syncthreads();
  const int t = threadIdx.x, r = t / 4, q = t % 4;
  int e = INT_MIN;
  for (int j = 0; j < 16; ++j) {
    const double2 x = T[r][16 * q + j];
    e = max(e, max_exponent(fmax(fabs(x.x), fabs(x.y))));
  }
  red[t] = e;
  __syncthreads();
  if (q == 0) {
    e = max(max(red[t], red[t + 1]), max(red[t + 2], red[t + 3]));
    if (e != INT_MIN && r0 + r < v.rows) atomicMax(&v.E[b * v.Rp + r0 + r], e);
  }
}

__device__ inline int sigma_of(int e) { return e == INT_MIN ? 0 : 7 - e; }

// Canonical slices: thread (row r, quarter q) turns 16 consecutive k of its
// row into NSL digits of re and im, packed 16 bytes per plane.
__global__ __launch_bounds__(NT) void oz_slice_kernel(const SliceView v) {
  __shared__ double2 T[TILE][KBLK + 1];
  const int64_t b = blockIdx.z;
  const int r0 = blockIdx.y * TILE;
  const int kblk = blockIdx.x;
  load_tile(v, b, r0, kblk * KBLK, T);
  __syncthreads();
  const int t = threadIdx.x, r = t / 4, q = t % 4;
  const int row = r0 + r;
  const int sg = row < v.rows ? sigma_of(v.E[b * v.Rp + row]) : 0;
  unsigned int wre[NSL][4], wim[NSL][4], wneg[NSL][4];
  for (int s = 0; s < NSL; ++s)
    for (int w = 0; w < 4; ++w) { wre[s][w] = 0u; wim[s][w] = 0u; wneg[s][w] = 0u; }
  for (int j = 0; j < 16; ++j) {
    const double2 x = T[r][16 * q + j];
    double xr = ldexp(x.x, sg), xi = ldexp(v.conj ? -x.y : x.y, sg);
    for (int s = 0; s < NSL; ++s) {
      const double dr = trunc(xr), di = trunc(xi);
      xr = (xr - dr) * 128.0;
      xi = (xi - di) * 128.0;
      const int ir = (int)dr, ii = (int)di;
      const int sh = 8 * (j & 3);
      wre[s][j >> 2] |= (unsigned int)(ir & 0xff) << sh;
      wim[s][j >> 2] |= (unsigned int)(ii & 0xff) << sh;
      wneg[s][j >> 2] |= (unsigned int)((-ii) & 0xff) << sh;
    }
  }
  int8_t* dst = v.out + (((int64_t)b * v.nkb + kblk) * v.Rp + row) * ((int64_t)v.nplanes * KBLK) + 16 * q;
  const int per = v.nplanes / NSL;   // 3 (re, im, -im) or 2 (re, im)
  for (int s = 0; s < NSL; ++s) {
    *reinterpret_cast<uint4*>(dst + (per * s + 0) * KBLK) = make_uint4(wre[s][0], wre[s][1], wre[s][2], wre[s][3]);
    *reinterpret_cast<uint4*>(dst + (per * s + 1) * KBLK) = make_uint4(wim[s][0], wim[s][1], wim[s][2], wim[s][3]);
    if (per == 3)
      *reinterpret_cast<uint4*>(dst + (per * s + 2) * KBLK) =
          make_uint4(wneg[s][0], wneg[s][1], wneg[s][2], wneg[s][3]);
  }
}

struct OzGemm {
  const int8_t* Aq = nullptr;
  const int8_t* Bq = nullptr;
  const int* EA = nullptr;
  const int* EB = nullptr;
  int M = 0, N = 0, Mp = 0, Np = 0, nkb = 0;
  int nb0 = 1, nb1 = 1, a_shared = 0, b_shared = 0, batch_major = 0;
  void* C = nullptr;
  int64_t c_rs = 0, c_cs = 0, c_bs0 = 0, c_bs1 = 0;
  const int64_t* c_off1 = nullptr;
  double alpha_re = 1.0, alpha_im = 0.0, beta_re = 0.0, beta_im = 0.0;
};

// Workgroups are dispatched round-robin over the 8 XCDs; the linear id is
// remapped so every XCD works on a contiguous range of tiles (tiles sharing
// an operand panel share an L2).
__device__ inline int64_t oz_xcd_logical(int64_t id, int64_t total) {
  const int64_t xcd = id % 8, slot = id / 8;
  const int64_t per = total / 8, rem = total % 8;
  return (xcd < rem) ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

template <class TO>
__device__ inline void oz_store(const OzGemm& g, int64_t b0, int64_t b1, int row, int col, double vr, double vi) {
  TO* C = static_cast<TO*>(g.C) + b0 * g.c_bs0 + (g.c_off1 ? g.c_off1[b1] : b1 * g.c_bs1) + row * g.c_rs +
          col * g.c_cs;
  double cr = g.alpha_re * vr - g.alpha_im * vi, ci = g.alpha_re * vi + g.alpha_im * vr;
  if (g.beta_re != 0.0 || g.beta_im != 0.0) {
    const TO o = *C;
    cr += g.beta_re * (double)o.x - g.beta_im * (double)o.y;
    ci += g.beta_re * (double)o.y + g.beta_im * (double)o.x;
  }
  TO out;
  out.x = cr;
  out.y = ci;
  *C = out;
}

// 64 x 64 output tile per workgroup, four waves of 32 x 32 (2 x 2 blocks of
// the 16 x 16 MFMA output).  Every k block stages all A and B planes in LDS;
// the next k block is loaded from global memory into registers while the
// current one is multiplied (one wave per SIMD, so the MFMAs must cover the
// load latency).
template <class TO>
__global__ __launch_bounds__(NT) void oz_gemm_kernel(const OzGemm g) {
  __shared__ __attribute__((aligned(16))) int8_t sA[TILE * (ROWA + LDS_PAD)];
  __shared__ __attribute__((aligned(16))) int8_t sB[TILE * (ROWB + LDS_PAD)];
  const int tiles_m = (g.M + TILE - 1) / TILE, tiles_n = (g.N + TILE - 1) / TILE;
  int64_t lg = oz_xcd_logical(blockIdx.x, gridDim.x);
  int mt, nt;
  int64_t b0, b1;
  if (g.batch_major) {
    const int64_t nb = (int64_t)g.nb0 * g.nb1, bb = lg % nb;
    lg /= nb;
    mt = (int)(lg % tiles_m);
    nt = (int)(lg / tiles_m);
    b1 = bb % g.nb1;
    b0 = bb / g.nb1;
  } else {
    mt = (int)(lg % tiles_m);
    lg /= tiles_m;
    nt = (int)(lg % tiles_n);
    lg /= tiles_n;
    b1 = lg % g.nb1;
    b0 = lg / g.nb1;
  }
  const int m0 = mt * TILE, n0 = nt * TILE;
  const int64_t bA = g.a_shared ? 0 : b0 * g.nb1 + b1, bB = g.b_shared ? 0 : b0 * g.nb1 + b1;
  const int8_t* gA = g.Aq + ((bA * g.nkb) * g.Mp + m0) * (int64_t)ROWA;
  const int8_t* gB = g.Bq + ((bB * g.nkb) * g.Np + n0) * (int64_t)ROWB;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;

  v4i re[2][2][DMAX + 1], im[2][2][DMAX + 1];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int d = 0; d <= DMAX; ++d) { re[i][j][d] = v4i{0, 0, 0, 0}; im[i][j][d] = v4i{0, 0, 0, 0}; }

  // The next k block's A and B tiles (15 and 10 16-byte chunks per thread)
  // are held in named registers between the loads and the LDS stores (an
  // array of them is not kept in registers by the compiler).
  static_assert(TILE * ROWA / 16 / NT == 15 && TILE * ROWB / 16 / NT == 10, "chunk counts");
#define OZ_A15(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14)
#define OZ_B10(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9)
#define OZ_DECL_A(c) uint4 ra##c;
#define OZ_DECL_B(c) uint4 rb##c;
#define OZ_LOAD_A(c) ra##c = srcA[(c) * NT + t];
#define OZ_LOAD_B(c) rb##c = srcB[(c) * NT + t];
#define OZ_STORE_A(c) *reinterpret_cast<uint4*>(sA + (((c) * NT + t) / (ROWA / 16)) * (ROWA + LDS_PAD) + \
                                                 16 * (((c) * NT + t) % (ROWA / 16))) = ra##c;
#define OZ_STORE_B(c) *reinterpret_cast<uint4*>(sB + (((c) * NT + t) / (ROWB / 16)) * (ROWB + LDS_PAD) + \
                                                 16 * (((c) * NT + t) % (ROWB / 16))) = rb##c;
  OZ_A15(OZ_DECL_A)
  OZ_B10(OZ_DECL_B)
  {
    const uint4* srcA = reinterpret_cast<const uint4*>(gA);
    const uint4* srcB = reinterpret_cast<const uint4*>(gB);
    OZ_A15(OZ_LOAD_A)
    OZ_B10(OZ_LOAD_B)
  }
  const int8_t* aRow[2];
  aRow[0] = sA + (wm + lr) * (ROWA + LDS_PAD) + 16 * lk;
  aRow[1] = aRow[0] + 16 * (ROWA + LDS_PAD);
  for (int kb = 0; kb < g.nkb; ++kb) {
    OZ_A15(OZ_STORE_A)
    OZ_B10(OZ_STORE_B)
    __syncthreads();
    if (kb + 1 < g.nkb) {
      const uint4* srcA = reinterpret_cast<const uint4*>(gA + (int64_t)(kb + 1) * g.Mp * ROWA);
      const uint4* srcB = reinterpret_cast<const uint4*>(gB + (int64_t)(kb + 1) * g.Np * ROWB);
      OZ_A15(OZ_LOAD_A)
      OZ_B10(OZ_LOAD_B)
    }
    v4i bf[2][PB];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int p = 0; p < PB; ++p)
        bf[j][p] = *reinterpret_cast<const v4i*>(sB + (wn + 16 * j + lr) * (ROWB + LDS_PAD) + p * KBLK + 16 * lk);
    // A fragments (re, im, -im) of slice s for block row i, fetched one
    // (i, s) step ahead of their MFMAs; within a step the products of one
    // accumulator are four MFMAs apart.
    v4i cur[3], nxt[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) cur[q] = *reinterpret_cast<const v4i*>(aRow[0] + q * KBLK);
#pragma unroll
    for (int step = 0; step < 2 * NSL; ++step) {
      const int i = step / NSL, s = step % NSL;
      if (step + 1 < 2 * NSL) {
        const int i2 = (step + 1) / NSL, s2 = (step + 1) % NSL;
#pragma unroll
        for (int q = 0; q < 3; ++q) nxt[q] = *reinterpret_cast<const v4i*>(aRow[i2] + (3 * s2 + q) * KBLK);
      }
#pragma unroll
      for (int u = 0; u < NSL && u + s <= DMAX; ++u) {
        const int d = s + u;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          re[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cur[0], bf[j][2 * u], re[i][j][d], 0, 0, 0);
          im[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cur[0], bf[j][2 * u + 1], im[i][j][d], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          re[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cur[2], bf[j][2 * u + 1], re[i][j][d], 0, 0, 0);
          im[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cur[1], bf[j][2 * u], im[i][j][d], 0, 0, 0);
        }
      }
      if (step + 1 < 2 * NSL) {
#pragma unroll
        for (int q = 0; q < 3; ++q) cur[q] = nxt[q];
      }
    }
    __syncthreads();
  }
  // epilogue: lane (lr, lk), register r holds row 4 lk + r, column lr of each 16 x 16 block
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + 4 * lk + r, col = n0 + wn + 16 * j + lr;
        if (row >= g.M || col >= g.N) continue;
        double vr = (double)re[i][j][DMAX][r], vi = (double)im[i][j][DMAX][r];
#pragma unroll
        for (int d = DMAX - 1; d >= 0; --d) {
          vr = vr * 0.0078125 + (double)re[i][j][d][r];
          vi = vi * 0.0078125 + (double)im[i][j][d][r];
        }
        const int ea = g.EA[bA * g.Mp + row], eb = g.EB[bB * g.Np + col];
        const int sc = -(sigma_of(ea) + sigma_of(eb));
        oz_store<TO>(g, b0, b1, row, col, ldexp(vr, sc), ldexp(vi, sc));
      }
#undef OZ_A15
#undef OZ_B10
#undef OZ_DECL_A
#undef OZ_DECL_B
#undef OZ_LOAD_A
#undef OZ_LOAD_B
#undef OZ_STORE_A
#undef OZ_STORE_B
}

__global__ void oz_fill_kernel(int* p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

struct Plan {
  int64_t Mp, Np, nkb, nbA, nbB;
  size_t bytesEA, bytesEB, bytesA, bytesB;
  size_t total() const { return bytesEA + bytesEB + bytesA + bytesB + 4 * 256; }
};

Plan plan_for(int M, int N, int64_t kt, int nb0, int nb1, bool a_shared, bool b_shared) {
  Plan p;
  p.Mp = round_up(M, TILE);
  p.Np = round_up(N, TILE);
  p.nkb = (kt + KBLK - 1) / KBLK;
  p.nbA = a_shared ? 1 : (int64_t)nb0 * nb1;
  p.nbB = b_shared ? 1 : (int64_t)nb0 * nb1;
  p.bytesEA = round_up((size_t)p.nbA * p.Mp * sizeof(int), 256);
  p.bytesEB = round_up((size_t)p.nbB * p.Np * sizeof(int), 256);
  p.bytesA = round_up((size_t)p.nbA * p.nkb * p.Mp * ROWA, 256);
  p.bytesB = round_up((size_t)p.nbB * p.nkb * p.Np * ROWB, 256);
  return p;
}

void fill_int(int* p, int64_t n, int v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(oz_fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  GGPU_HIP(hipGetLastError());
}

void slice_operand(SliceView v, int64_t nb, hipStream_t s) {
  require(nb >= 1 && nb <= 65535 && v.Rp / TILE <= 65535 && v.nkb >= 1, "int8 emulation: operand too large");
  fill_int(v.E, nb * v.Rp, INT_MIN, s);
  const dim3 grid((unsigned)v.nkb, (unsigned)(v.Rp / TILE), (unsigned)nb);
  hipLaunchKernelGGL(oz_exponent_kernel, grid, dim3(NT), 0, s, v);
  GGPU_HIP(hipGetLastError());
  hipLaunchKernelGGL(oz_slice_kernel, grid, dim3(NT), 0, s, v);
  GGPU_HIP(hipGetLastError());
}

// One launch over batches [b0s, b0s + nb0) and K-batch [kb0, kb0 + KBc).
void run_part(const GemmDesc& d, int b0s, int nb0, int kb0, int KBc, bool first_k, char* work, size_t work_bytes,
              hipStream_t s) {
  const bool a_shared = d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr;
  const bool b_shared = d.B.bs0 == 0 && d.B.bs1 == 0 && d.B.off1 == nullptr;
  const int64_t kt = (int64_t)d.K * KBc;
  const Plan p = plan_for(d.M, d.N, kt, nb0, d.nb1, a_shared, b_shared);
  require(p.total() <= work_bytes, "int8 emulation: workspace too small for one batch");
  int* EA = reinterpret_cast<int*>(work);
  int* EB = reinterpret_cast<int*>(work + p.bytesEA);
  int8_t* Aq = reinterpret_cast<int8_t*>(work + p.bytesEA + p.bytesEB);
  int8_t* Bq = Aq + p.bytesA;

  SliceView a;
  a.src = static_cast<const double2*>(d.A.ptr) + (int64_t)b0s * d.A.bs0 + (int64_t)kb0 * d.A.kbs;
  a.row_s = d.A.rs; a.k_s = d.A.cs; a.kb_s = d.A.kbs; a.bs0 = d.A.bs0; a.bs1 = d.A.bs1; a.off1 = d.A.off1;
  a.rows = d.M; a.K = d.K; a.KB = KBc; a.nb1 = a_shared ? 1 : d.nb1;
  a.conj = d.A.conj; a.nplanes = PA; a.Rp = (int)p.Mp; a.nkb = (int)p.nkb; a.E = EA; a.out = Aq;
  slice_operand(a, p.nbA, s);
  SliceView b;
  b.src = static_cast<const double2*>(d.B.ptr) + (int64_t)b0s * d.B.bs0 + (int64_t)kb0 * d.B.kbs;
  b.row_s = d.B.cs; b.k_s = d.B.rs; b.kb_s = d.B.kbs; b.bs0 = d.B.bs0; b.bs1 = d.B.bs1; b.off1 = d.B.off1;
  b.rows = d.N; b.K = d.K; b.KB = KBc; b.nb1 = b_shared ? 1 : d.nb1;
  b.conj = d.B.conj; b.nplanes = PB; b.Rp = (int)p.Np; b.nkb = (int)p.nkb; b.E = EB; b.out = Bq;
  slice_operand(b, p.nbB, s);

  OzGemm g;
  g.Aq = Aq; g.Bq = Bq; g.EA = EA; g.EB = EB;
  g.M = d.M; g.N = d.N; g.Mp = (int)p.Mp; g.Np = (int)p.Np; g.nkb = (int)p.nkb;
  g.nb0 = nb0; g.nb1 = d.nb1; g.a_shared = a_shared; g.b_shared = b_shared;
  g.batch_major = (nb0 * (int64_t)d.nb1 > 1) && a_shared;
  g.C = static_cast<char*>(d.C.ptr) + (int64_t)b0s * d.C.bs0 * (d.out_f32 ? sizeof(float2) : sizeof(double2));
  g.c_rs = d.C.rs; g.c_cs = d.C.cs; g.c_bs0 = d.C.bs0; g.c_bs1 = d.C.bs1; g.c_off1 = d.C.off1;
  g.alpha_re = d.alpha_re; g.alpha_im = d.alpha_im;
  g.beta_re = first_k ? d.beta_re : 1.0;
  g.beta_im = first_k ? d.beta_im : 0.0;
  const int64_t tiles = (int64_t)(p.Mp / TILE) * (p.Np / TILE) * nb0 * d.nb1;
  require(tiles < ((int64_t)1 << 31), "int8 emulation: grid too large");
  if (d.out_f32) hipLaunchKernelGGL((oz_gemm_kernel<float2>), dim3((unsigned)tiles), dim3(NT), 0, s, g);
  else hipLaunchKernelGGL((oz_gemm_kernel<double2>), dim3((unsigned)tiles), dim3(NT), 0, s, g);
  GGPU_HIP(hipGetLastError());
}

}  // namespace

size_t gemm_int8_emulated_workspace(const GemmDesc& d) {
  const bool a_shared = d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr;
  const bool b_shared = d.B.bs0 == 0 && d.B.bs1 == 0 && d.B.off1 == nullptr;
  const int KBc = (int)std::max<int64_t>(1, std::min<int64_t>(d.KB, KMAX / std::max(1, d.K)));
  return plan_for(d.M, d.N, (int64_t)d.K * KBc, 1, d.nb1, a_shared, b_shared).total();
}

void gemm_int8_emulated(const GemmDesc& d, void* work, size_t work_bytes, hipStream_t s) {
  require(work != nullptr && work_bytes > 0, "int8 emulation: no workspace installed");
  require(d.K <= KMAX, "int8 emulation: K above " + std::to_string(KMAX));
  const bool a_shared = d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr;
  const bool b_shared = d.B.bs0 == 0 && d.B.bs1 == 0 && d.B.off1 == nullptr;
  // K-batch chunks small enough that no int32 diagonal sum can overflow
  const int KBc = (int)std::max<int64_t>(1, std::min<int64_t>(d.KB, KMAX / std::max(1, d.K)));
  // batch ranges (along b0) whose slices fit the workspace
  int nb0c = d.nb0;
  while (nb0c > 1 && plan_for(d.M, d.N, (int64_t)d.K * KBc, nb0c, d.nb1, a_shared, b_shared).total() > work_bytes)
    nb0c = (nb0c + 1) / 2;
  for (int kb0 = 0; kb0 < d.KB; kb0 += KBc) {
    const int kbn = std::min(KBc, d.KB - kb0);
    for (int b0s = 0; b0s < d.nb0; b0s += nb0c)
      run_part(d, b0s, std::min(nb0c, d.nb0 - b0s), kb0, kbn, kb0 == 0, static_cast<char*>(work), work_bytes, s);
  }
}

}  // namespace ggpu
