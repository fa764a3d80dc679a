// Complex GEMM with FP64 operands on the gfx950 int8 matrix cores
// (v_mfma_i32_16x16x64_i8): every FP64 operand is split into NSL signed
// 7-bit slices (NSL = 5 or 4) with one power-of-two scale per row of A and per
// column of B (Ozaki-style splitting), the slice products with s + t <= NSL
// are summed exactly in int32 (one accumulator per slice diagonal s + t), and
// the diagonals are recombined in FP64:
//
//   x 2^sigma = sum_s x_s 128^-s + r,  |x_s| <= 127,  |r| < 128^-(NSL-1) (in units of 2^-7 of the row maximum)
//   (A B)[m][n] = 2^-(sigma_m + sigma_n) sum_d 128^-d sum_{s+t=d} (A_s B_t)[m][n]
//
// The slice remainders (and the dropped pairs s + t > NSL, smaller still)
// leave an error of about 2^-(7 NSL - 1) of max|A_m| max|B_n| per term of the
// k sum: 2^-34 with five slices (FP64-accurate products for the contractions
// here), 2^-27 with four.
// Complex products use four real products per slice pair; the negated
// imaginary slice of A is formed in registers (bytewise), so re and im
// accumulate exactly:
//   re += Ar_s Br_t + (-Ai_s) Bi_t,   im += Ar_s Bi_t + Ai_s Br_t.
//
// Slices are stored fragment-native: for every (batch, k block of 64, block
// of 16 rows, plane) one 1 KiB fragment whose lane l holds row l % 16 and the
// 16 k bytes 16 (l / 16) .. + 15 -- the operand layout of the MFMA, so a
// fragment moves from HBM to LDS and from LDS to registers as one contiguous
// 1 KiB wave access (LDS reads conflict-free, no padding).  Planes: re and im
// of every slice (2 NSL).
//
// A launch runs the slicing of both operands (per-row / per-column maximum
// binary exponent, then the digits: from an LDS-resident copy of each 16-row
// block when its k extent fits (one HBM read), in one fused pass reading the
// rows twice when it is short, else an exponent pass and a digit pass) and
// the int8 GEMM with the FP64 epilogue (alpha, beta, strided or complex64
// output): one tile per workgroup, or, for short k extents, one workgroup per
// CU walking tiles.
#include "ozaki.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>

namespace ggpu {
namespace {

// NSL (a template argument): slices per operand, 5 or 4; the slice pairs
// s + t <= NSL are kept (19 or 13 pairs), one int32 accumulator per diagonal.
constexpr int KBLK = 64;            // k per block: one v_mfma_i32_16x16x64_i8
constexpr int TILE = 64;            // rows of a workgroup tile (both operands)
constexpr int RB = 16;              // rows of a fragment
constexpr int FRAG = 1024;          // bytes of a fragment (64 lanes x 16 bytes)
constexpr int NT = 256;
// one operand of a tile for one k block: 4 row blocks x 2 NSL planes (re and
// im of every slice); a stage holds both operands, two stages the LDS
template <int NSL>
constexpr int stage_bytes() { return 2 * (TILE / RB) * 2 * NSL * FRAG; }
// Largest K * KB one launch may sum: every int32 diagonal accumulator holds at
// most 5 pairs * 2 products * K * 127^2 < 2^31 (no diagonal has more than 5
// slice pairs).
constexpr int64_t KMAX = 12800;
// Largest summed k extent sliced in one fused pass (the rows' second read
// then comes from cache).
constexpr int FUSE_KT = 2048;

typedef int v4i __attribute__((ext_vector_type(4)));

struct SliceView {
  const double2* src = nullptr;   // element (row 0, k 0, kb 0, batch 0)
  int64_t row_s = 0, k_s = 0, kb_s = 0, bs0 = 0, bs1 = 0;
  const int64_t* off1 = nullptr;  // device: offset of batch b1 (replaces bs1)
  // rows in groups of rgrp (0: one group), group g at g * rgrp_s: batches
  // merged into the row index
  int rgrp = 0;
  int64_t rgrp_s = 0;
  int rows = 0, K = 0, KB = 1, nb1 = 1;
  int conj = 0;
  int Rp = 0, nkb = 0;            // padded rows, k blocks
  int* E = nullptr;               // [batch][Rp]: max binary exponent (INT_MIN: all zero)
  int8_t* out = nullptr;          // [batch][kblock][Rp / 16][2 NSL planes][64 lanes][16 bytes]
};

__device__ inline const double2* batch_base(const SliceView& v, int64_t b) {
  const int64_t b0 = b / v.nb1, b1 = b - b0 * v.nb1;
  return v.src + b0 * v.bs0 + (v.off1 ? v.off1[b1] : b1 * v.bs1);
}

// Offset of a row within its batch.
__device__ inline int64_t row_off(const SliceView& v, int row) {
  return v.rgrp ? (int64_t)(row / v.rgrp) * v.rgrp_s + (int64_t)(row % v.rgrp) * v.row_s : (int64_t)row * v.row_s;
}

// Element (row, flattened k) of one batch; zero outside the operand.  The
// flattened k of one launch stays below 2^31 (K * KB <= 12800).
__device__ inline double2 element(const SliceView& v, const double2* base, int row, int kk) {
  if (row >= v.rows || kk >= v.K * v.KB) return make_double2(0.0, 0.0);
  const int kb = kk / v.K, k = kk - kb * v.K;
  return base[row_off(v, row) + (int64_t)k * v.k_s + (int64_t)kb * v.kb_s];
}

__device__ inline int max_exponent(double2 x) {
  const double a = fmax(fabs(x.x), fabs(x.y));
  if (a == 0.0) return INT_MIN;
  int e;
  (void)frexp(a, &e);
  return e;
}

__device__ inline int sigma_of(int e) { return e == INT_MIN ? 0 : 7 - e; }

// Lanes run along whichever index is contiguous in memory.
__device__ inline bool k_fast(const SliceView& v) { return v.k_s == 1 || v.row_s != 1; }

// E[b][row] = max over (k, kb) of the binary exponent of max(|re|, |im|),
// for operands with a long k extent.  Grid: (k blocks, row tiles of 64,
// batches); E starts at INT_MIN.
__global__ __launch_bounds__(NT) void oz_exponent_kernel(const SliceView v) {
  __shared__ int red[TILE][KBLK + 1];
  const int64_t b = blockIdx.z;
  const double2* base = batch_base(v, b);
  const int r0 = blockIdx.y * TILE, k0 = blockIdx.x * KBLK, t = threadIdx.x;
  // a wave reads 64 consecutive elements of one row (k contiguous) or of
  // one k (rows contiguous)
  const bool kf = k_fast(v);
  for (int j = 0; j < TILE * KBLK / NT; ++j) {
    const int r = kf ? (t >> 6) + 4 * j : t & 63, c = kf ? t & 63 : (t >> 6) + 4 * j;
    red[r][c] = max_exponent(element(v, base, r0 + r, k0 + c));
  }
  __syncthreads();
  if (t < TILE) {
    int e = INT_MIN;
    for (int c = 0; c < KBLK; ++c) e = max(e, red[t][c]);
    if (e != INT_MIN && r0 + t < v.rows) atomicMax(&v.E[b * v.Rp + r0 + t], e);
  }
}

// Digits of one block of 16 rows: wave w of the workgroup takes k blocks
// w, w + 4, .. (FUSED) or the one k block 4 blockIdx.x + w; lane (lr, lk)
// loads the 16 k of its fragment lane (row lr, k 16 lk ..) straight into
// registers, turns them into NSL digits of re and im and stores one 16-byte
// lane of each of the NP fragments (every store a contiguous 1 KiB per wave).
// FUSED: the workgroup first finds its rows' exponents itself (grid: (1,
// Rp / 16, batches)); otherwise they come from E (grid: (k groups of 4,
// Rp / 16, batches)).
template <bool FUSED, int NSL>
__global__ __launch_bounds__(NT) void oz_slice_kernel(const SliceView v) {
  constexpr int NP = 2 * NSL;
  __shared__ int red[NT];
  __shared__ int rowexp[RB];
  const int64_t b = blockIdx.z;
  const int rblk = blockIdx.y, r0 = rblk * RB, t = threadIdx.x;
  const double2* base = batch_base(v, b);
  const int kt = v.K * v.KB;
  if (FUSED) {
    // thread t: row t / 16 (k contiguous) or t % 16 (rows contiguous), every 16th k
    const bool kf = k_fast(v);
    const int rr = kf ? t / RB : t % RB, kc = kf ? t % RB : t / RB;
    int e = INT_MIN;
    for (int kk = kc; kk < kt; kk += RB) e = max(e, max_exponent(element(v, base, r0 + rr, kk)));
    red[rr * RB + kc] = e;
    __syncthreads();
    if (t < RB) {
      e = INT_MIN;
      for (int j = 0; j < RB; ++j) e = max(e, red[t * RB + j]);
      if (r0 + t < v.rows) v.E[b * v.Rp + r0 + t] = e;
      rowexp[t] = e;
    }
    __syncthreads();
  }
  const int lane = t & 63, w = t >> 6, lr = lane & 15, lk = lane >> 4;
  const int row = r0 + lr;
  const bool live = row < v.rows;
  const int sg = live ? sigma_of(FUSED ? rowexp[lr] : v.E[b * v.Rp + row]) : 0;
  const int rblocks = v.Rp / RB;
  const double2* rowp = base + row_off(v, live ? row : 0);
  const int kb_first = FUSED ? w : (int)blockIdx.x * 4 + w;
  const int kb_end = FUSED ? v.nkb : min(v.nkb, kb_first + 1);
  for (int kb = kb_first; kb < kb_end; kb += 4) {
    // flattened k of element j: kk0 + j = q K + k
    const int kk0 = kb * KBLK + 16 * lk;
    int q = kk0 / v.K, k = kk0 - q * v.K;
    double2 x[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      x[j] = (live && kk0 + j < kt) ? rowp[(int64_t)k * v.k_s + (int64_t)q * v.kb_s] : make_double2(0.0, 0.0);
      if (++k == v.K) { k = 0; ++q; }
    }
    unsigned int wre[NSL][4], wim[NSL][4];
#pragma unroll
    for (int s = 0; s < NSL; ++s)
#pragma unroll
      for (int c = 0; c < 4; ++c) { wre[s][c] = 0u; wim[s][c] = 0u; }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      double xr = ldexp(x[j].x, sg), xi = ldexp(v.conj ? -x[j].y : x[j].y, sg);
      const int sh = 8 * (j & 3);
#pragma unroll
      for (int s = 0; s < NSL; ++s) {
        const double dr = trunc(xr), di = trunc(xi);
        xr = (xr - dr) * 128.0;
        xi = (xi - di) * 128.0;
        wre[s][j >> 2] |= (unsigned int)((int)dr & 0xff) << sh;
        wim[s][j >> 2] |= (unsigned int)((int)di & 0xff) << sh;
      }
    }
    int8_t* dst = v.out + (((int64_t)b * v.nkb + kb) * rblocks + rblk) * ((int64_t)NP * FRAG) + 16 * lane;
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      *reinterpret_cast<uint4*>(dst + (2 * s) * FRAG) = make_uint4(wre[s][0], wre[s][1], wre[s][2], wre[s][3]);
      *reinterpret_cast<uint4*>(dst + (2 * s + 1) * FRAG) = make_uint4(wim[s][0], wim[s][1], wim[s][2], wim[s][3]);
    }
  }
}

// Exponents and digits of one block of 16 rows of a row-contiguous operand
// whose whole k extent fits in LDS (K * KB <= LDS_KT): the workgroup stages
// the block once (16 rows x K * KB complex doubles, k-major: the 16 rows of
// one k are one 256-byte run), finds the rows' exponents from LDS and writes
// the digits from LDS -- one HBM read of the operand.  Grid: (1, Rp / 16,
// batches); one workgroup per CU (the staging fills the LDS).
constexpr int LDS_KT = 632;

// Digit pass for k-contiguous operands with a long k extent (after the
// exponent pass): one wave per (16-row block, k block), which first copies
// its 16 rows x 64 k into its own LDS area with load-to-LDS copies -- one
// instruction per row, 64 lanes reading 1 KiB contiguous -- and then reads
// each lane's 16 k values from there (rows 1040 bytes apart: conflict-free).
// Reading the 16 k values straight from memory instead touches 64 lines per
// load instruction.  Lanes past the operand read a valid element and are
// masked below; same digits as oz_slice_kernel.
constexpr int KF_ROW = KBLK * 16 + 16;   // bytes per staged row (64 complex + pad)
template <int NSL>
__global__ __launch_bounds__(64) void oz_digits_kf_kernel(const SliceView v) {
  constexpr int NP = 2 * NSL;
  __shared__ __attribute__((aligned(16))) int8_t T[RB * KF_ROW];
  const int64_t b = blockIdx.z;
  const int rblk = blockIdx.y, r0 = rblk * RB, kb = blockIdx.x;
  const int lane = threadIdx.x, lr = lane & 15, lk = lane >> 4;
  const double2* base = batch_base(v, b);
  const int kt = v.K * v.KB;
  {
    const int kk = min(kb * KBLK + lane, kt - 1);
    const int q = kk / v.K, k = kk - q * v.K;
    const int64_t koff = (int64_t)k * v.k_s + (int64_t)q * v.kb_s;
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int row = (r0 + i < v.rows) ? r0 + i : 0;
      __builtin_amdgcn_global_load_lds((const void*)(base + row_off(v, row) + koff),
                                       (__attribute__((address_space(3))) void*)(T + i * KF_ROW), 16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0x3f70);   // vmcnt(0): this wave's copies have landed (the area is its own)
  }
  const int row = r0 + lr;
  const bool live = row < v.rows;
  const int sg = live ? sigma_of(v.E[b * v.Rp + row]) : 0;
  const int rblocks = v.Rp / RB;
  unsigned int wre[NSL][4], wim[NSL][4];
#pragma unroll
  for (int s = 0; s < NSL; ++s)
#pragma unroll
    for (int c = 0; c < 4; ++c) { wre[s][c] = 0u; wim[s][c] = 0u; }
  const double2* Tl = reinterpret_cast<const double2*>(T + lr * KF_ROW) + 16 * lk;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int k = kb * KBLK + 16 * lk + j;
    const double2 x = (live && k < kt) ? Tl[j] : make_double2(0.0, 0.0);
    double xr = ldexp(x.x, sg), xi = ldexp(v.conj ? -x.y : x.y, sg);
    const int sh = 8 * (j & 3);
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      const double dr = trunc(xr), di = trunc(xi);
      xr = (xr - dr) * 128.0;
      xi = (xi - di) * 128.0;
      wre[s][j >> 2] |= (unsigned int)((int)dr & 0xff) << sh;
      wim[s][j >> 2] |= (unsigned int)((int)di & 0xff) << sh;
    }
  }
  int8_t* dst = v.out + (((int64_t)b * v.nkb + kb) * rblocks + rblk) * ((int64_t)NP * FRAG) + 16 * lane;
#pragma unroll
  for (int s = 0; s < NSL; ++s) {
    *reinterpret_cast<uint4*>(dst + (2 * s) * FRAG) = make_uint4(wre[s][0], wre[s][1], wre[s][2], wre[s][3]);
    *reinterpret_cast<uint4*>(dst + (2 * s + 1) * FRAG) = make_uint4(wim[s][0], wim[s][1], wim[s][2], wim[s][3]);
  }
}

template <int NSL, int NTH>
__global__ __launch_bounds__(NTH) void oz_slice_lds_kernel(const SliceView v) {
  constexpr int NWV = NTH / 64, KC = NTH / RB;   // waves; k columns of the exponent pass
  constexpr int NP = 2 * NSL;
  __shared__ double2 T[LDS_KT][RB];
  __shared__ int red[RB * NWV];
  __shared__ int rowexp[RB];
  const int64_t b = blockIdx.z;
  const int rblk = blockIdx.y, r0 = rblk * RB, t = threadIdx.x;
  const double2* base = batch_base(v, b);
  const int kt = v.K * v.KB;
  // stage (rows contiguous only): one load-to-LDS copy moves 4 k x 16 rows
  // (64 lanes x 16 bytes, lane l -> T[k0 + l / 16][l % 16]); every copy of the
  // block is in flight at once.  Lanes past the operand read a valid element
  // (row 0 / the last k) and are masked where T is used.
  {
    const int lane = t & 63, w = t >> 6, r = lane % RB;
    const int row = (r0 + r < v.rows) ? r0 + r : 0;
    for (int k0 = 4 * w; k0 < kt; k0 += 4 * NWV) {
      const int kk = min(k0 + lane / RB, kt - 1);
      const int q = kk / v.K, k = kk - q * v.K;
      const double2* src = base + row_off(v, row) + (int64_t)k * v.k_s + (int64_t)q * v.kb_s;
      __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)&T[k0][0], 16, 0,
                                       0);
    }
    __builtin_amdgcn_s_waitcnt(0x3f70);   // vmcnt(0): this wave's copies have landed
  }
  __syncthreads();
  {
    const int rr = t % RB, kc = t / RB;
    int e = INT_MIN;
    if (r0 + rr < v.rows)
      for (int k = kc; k < kt; k += KC) e = max(e, max_exponent(T[k][rr]));
    // the four k columns of a row within a wave (lanes rr, rr + 16, ..), then across waves
    e = max(e, __shfl_xor(e, 16));
    e = max(e, __shfl_xor(e, 32));
    if ((t & 63) < RB) red[rr * NWV + (t >> 6)] = e;
  }
  __syncthreads();
  if (t < RB) {
    int e = INT_MIN;
    for (int j = 0; j < NWV; ++j) e = max(e, red[t * NWV + j]);
    if (r0 + t < v.rows) v.E[b * v.Rp + r0 + t] = e;
    rowexp[t] = e;
  }
  __syncthreads();
  const int lane = t & 63, w = t >> 6, lr = lane & 15, lk = lane >> 4;
  const int row = r0 + lr;
  const int sg = row < v.rows ? sigma_of(rowexp[lr]) : 0;
  const int rblocks = v.Rp / RB;
  for (int kb = w; kb < v.nkb; kb += NWV) {
    unsigned int wre[NSL][4], wim[NSL][4];
#pragma unroll
    for (int s = 0; s < NSL; ++s)
#pragma unroll
      for (int c = 0; c < 4; ++c) { wre[s][c] = 0u; wim[s][c] = 0u; }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int k = kb * KBLK + 16 * lk + j;
      const double2 x = (k < kt && row < v.rows) ? T[k][lr] : make_double2(0.0, 0.0);
      double xr = ldexp(x.x, sg), xi = ldexp(v.conj ? -x.y : x.y, sg);
      const int sh = 8 * (j & 3);
#pragma unroll
      for (int s = 0; s < NSL; ++s) {
        const double dr = trunc(xr), di = trunc(xi);
        xr = (xr - dr) * 128.0;
        xi = (xi - di) * 128.0;
        wre[s][j >> 2] |= (unsigned int)((int)dr & 0xff) << sh;
        wim[s][j >> 2] |= (unsigned int)((int)di & 0xff) << sh;
      }
    }
    int8_t* dst = v.out + (((int64_t)b * v.nkb + kb) * rblocks + rblk) * ((int64_t)NP * FRAG) + 16 * lane;
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      *reinterpret_cast<uint4*>(dst + (2 * s) * FRAG) = make_uint4(wre[s][0], wre[s][1], wre[s][2], wre[s][3]);
      *reinterpret_cast<uint4*>(dst + (2 * s + 1) * FRAG) = make_uint4(wim[s][0], wim[s][1], wim[s][2], wim[s][3]);
    }
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
  int cgrp = 0;                   // columns in groups of cgrp at cgrp_s (merged batches)
  int64_t cgrp_s = 0;
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
  const int64_t coff = g.cgrp ? (int64_t)(col / g.cgrp) * g.cgrp_s + (int64_t)(col % g.cgrp) * g.c_cs
                             : (int64_t)col * g.c_cs;
  TO* C = static_cast<TO*>(g.C) + b0 * g.c_bs0 + (g.c_off1 ? g.c_off1[b1] : b1 * g.c_bs1) + row * g.c_rs + coff;
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

// Bytewise two's-complement negation of 16 packed int8 digits (|digit| <= 127).
__device__ inline v4i neg_bytes(v4i x) {
  v4i r;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const unsigned int n = ~(unsigned int)x[c];
    r[c] = (int)(((n & 0x7f7f7f7fu) + 0x01010101u) ^ (n & 0x80808080u));
  }
  return r;
}

__device__ inline v4i lds_frag(const int8_t* p) { return *reinterpret_cast<const v4i*>(p); }

// One 64 x 64 output tile per workgroup, 2 x WNW waves of 32 x (64 / WNW)
// (2 x 4 / WNW blocks of the 16 x 16 MFMA output): four waves of 32 x 32,
// one per SIMD, or eight of 32 x 16, two per SIMD.  Both operands'
// fragments of a k block (80 KiB) are copied HBM -> LDS by the load-to-LDS
// path into one of two stages while the other stage is multiplied, one
// barrier per k block.
// Within a k block every wave walks its A slices from the last to the first,
// so the B fragments it needs grow by one slice per step and are read from
// LDS one step ahead of their first use.
//
// Tile order: workgroups are dispatched round-robin over the 8 XCDs; the
// linear id is remapped so every XCD works on a contiguous range of tiles,
// and within a batch the tiles run in groups of GM row tiles, column after
// column, so the ~32 tiles an XCD runs at once cover GM row panels of A (kept
// in its L2 while the group sweeps the columns) and ~32 / GM column panels of B.
constexpr int GM = 4;

template <class TO, int NSL, int WNW>
__global__ __launch_bounds__(128 * WNW) void oz_gemm_kernel(const OzGemm g) {
  constexpr int DMAX = NSL, NP = 2 * NSL, STAGE = stage_bytes<NSL>(), STAGE_A = STAGE / 2;
  constexpr int NW = 2 * WNW, RJ = 4 / WNW;   // waves; column blocks per wave (2 row blocks each)
  __shared__ __attribute__((aligned(1024))) int8_t lds[2 * STAGE];
  const int tiles_m = (g.M + TILE - 1) / TILE, tiles_n = (g.N + TILE - 1) / TILE;
  int64_t L = oz_xcd_logical(blockIdx.x, gridDim.x);
  int mt, nt;
  int64_t b0, b1;
  if (g.batch_major) {
    const int64_t nb = (int64_t)g.nb0 * g.nb1, bb = L % nb;
    L /= nb;
    mt = (int)(L % tiles_m);
    nt = (int)(L / tiles_m);
    b1 = bb % g.nb1;
    b0 = bb / g.nb1;
  } else {
    const int per_batch = tiles_m * tiles_n;
    const int64_t bb = L / per_batch;
    const int r = (int)(L - bb * per_batch);
    const int mg = r / (GM * tiles_n), rem = r - mg * (GM * tiles_n);
    const int gm = min(GM, tiles_m - mg * GM);   // row tiles in this group
    nt = rem / gm;
    mt = mg * GM + (rem - nt * gm);
    b1 = bb % g.nb1;
    b0 = bb / g.nb1;
  }
  const int m0 = mt * TILE, n0 = nt * TILE;
  const int64_t bA = g.a_shared ? 0 : b0 * g.nb1 + b1, bB = g.b_shared ? 0 : b0 * g.nb1 + b1;
  // the tile's A fragments (4 row blocks x NP planes) of one k block are
  // contiguous, and so are its B fragments; consecutive k blocks are
  // kstrideA / kstrideB bytes apart
  const int64_t kstrideA = (int64_t)(g.Mp / RB) * NP * FRAG, kstrideB = (int64_t)(g.Np / RB) * NP * FRAG;
  const int8_t* gA = g.Aq + bA * g.nkb * kstrideA + (int64_t)(m0 / RB) * NP * FRAG;
  const int8_t* gB = g.Bq + bB * g.nkb * kstrideB + (int64_t)(n0 / RB) * NP * FRAG;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int wm = (w / WNW) * 32, wn = (w % WNW) * 16 * RJ;
  const int lr = lane & 15, lk = lane >> 4;

  // wave w copies fragments w, w + 4, .. of the stage: NSL of A, then NSL of B
  auto copy_stage = [&](int kb, int st) {
    const int8_t* srcA = gA + (int64_t)kb * kstrideA + 16 * lane;
    const int8_t* srcB = gB + (int64_t)kb * kstrideB + 16 * lane;
    int8_t* dst = lds + st * STAGE;
#pragma unroll
    for (int it = 0; it < STAGE_A / FRAG / NW; ++it) {
      const int f = w + NW * it;
      __builtin_amdgcn_global_load_lds((const void*)(srcA + f * FRAG),
                                       (__attribute__((address_space(3))) void*)(dst + f * FRAG), 16, 0, 0);
    }
#pragma unroll
    for (int it = 0; it < STAGE_A / FRAG / NW; ++it) {
      const int f = w + NW * it;
      __builtin_amdgcn_global_load_lds((const void*)(srcB + f * FRAG),
                                       (__attribute__((address_space(3))) void*)(dst + STAGE_A + f * FRAG), 16, 0,
                                       0);
    }
  };

  // exponents of the rows / columns this lane stores (fetched with the first
  // k block)
  int ea[2][4], eb[RJ];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = m0 + wm + 16 * i + 4 * lk + r;
      ea[i][r] = row < g.M ? g.EA[bA * g.Mp + row] : INT_MIN;
    }
#pragma unroll
  for (int j = 0; j < RJ; ++j) {
    const int col = n0 + wn + 16 * j + lr;
    eb[j] = col < g.N ? g.EB[bB * g.Np + col] : INT_MIN;
  }

  v4i re[2][RJ][DMAX + 1], im[2][RJ][DMAX + 1];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < RJ; ++j)
#pragma unroll
      for (int d = 0; d <= DMAX; ++d) { re[i][j][d] = v4i{0, 0, 0, 0}; im[i][j][d] = v4i{0, 0, 0, 0}; }

  copy_stage(0, 0);
  for (int kb = 0; kb < g.nkb; ++kb) {
    const int st = kb & 1;
    // stage st has landed for every wave, and every wave is done with the
    // stage the next copy overwrites
    __syncthreads();
    if (kb + 1 < g.nkb) copy_stage(kb + 1, st ^ 1);
    const int8_t* sA = lds + st * STAGE + (wm / RB) * NP * FRAG + 16 * lane;
    const int8_t* sB = lds + st * STAGE + STAGE_A + (wn / RB) * NP * FRAG + 16 * lane;
    // B fragments of slice u: bf[j][2u] (re), bf[j][2u + 1] (im), block column j
    v4i bf[RJ][NP];
#pragma unroll
    for (int u = 0; u < NSL && u <= DMAX - (NSL - 1); ++u)
#pragma unroll
      for (int j = 0; j < RJ; ++j) {
        bf[j][2 * u] = lds_frag(sB + (j * NP + 2 * u) * FRAG);
        bf[j][2 * u + 1] = lds_frag(sB + (j * NP + 2 * u + 1) * FRAG);
      }
    v4i ar = lds_frag(sA + (2 * (NSL - 1)) * FRAG), ai = lds_frag(sA + (2 * (NSL - 1) + 1) * FRAG);
#pragma unroll
    for (int step = 0; step < 2 * NSL; ++step) {
      const int i = step / NSL, s = NSL - 1 - step % NSL;
      v4i ar2 = ar, ai2 = ai;
      if (step + 1 < 2 * NSL) {
        const int i2 = (step + 1) / NSL, s2 = NSL - 1 - (step + 1) % NSL;
        ar2 = lds_frag(sA + (i2 * NP + 2 * s2) * FRAG);
        ai2 = lds_frag(sA + (i2 * NP + 2 * s2 + 1) * FRAG);
      }
      // the slice of B the next step uses first
      if (i == 0 && s >= 1 && DMAX - s + 1 < NSL) {
        const int u = DMAX - s + 1;
#pragma unroll
        for (int j = 0; j < RJ; ++j) {
          bf[j][2 * u] = lds_frag(sB + (j * NP + 2 * u) * FRAG);
          bf[j][2 * u + 1] = lds_frag(sB + (j * NP + 2 * u + 1) * FRAG);
        }
      }
      const v4i nai = neg_bytes(ai);
      // the products with Ar for every slice u of the step, then those with Ai
#pragma unroll
      for (int u = 0; u < NSL && u + s <= DMAX; ++u) {
        const int d = s + u;
#pragma unroll
        for (int j = 0; j < RJ; ++j) {
          re[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ar, bf[j][2 * u], re[i][j][d], 0, 0, 0);
          im[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ar, bf[j][2 * u + 1], im[i][j][d], 0, 0, 0);
        }
      }
#pragma unroll
      for (int u = 0; u < NSL && u + s <= DMAX; ++u) {
        const int d = s + u;
#pragma unroll
        for (int j = 0; j < RJ; ++j) {
          re[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(nai, bf[j][2 * u + 1], re[i][j][d], 0, 0, 0);
          im[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ai, bf[j][2 * u], im[i][j][d], 0, 0, 0);
        }
      }
      ar = ar2;
      ai = ai2;
      // keep each step's MFMAs together, behind the reads issued for the next step
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // epilogue: lane (lr, lk), register r holds row 4 lk + r, column lr of each 16 x 16 block
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < RJ; ++j)
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
        const int sc = -(sigma_of(ea[i][r]) + sigma_of(eb[j]));
        oz_store<TO>(g, b0, b1, row, col, ldexp(vr, sc), ldexp(vi, sc));
      }
}

// Coordinates of logical tile L: within a batch, groups of GM row tiles,
// column after column (or the batch index fastest when every batch shares A).
struct OzTile {
  int m0, n0;
  int64_t b0, b1, bA, bB;
};
__device__ inline OzTile oz_tile(const OzGemm& g, int64_t L, int tiles_m, int tiles_n) {
  int mt, nt;
  OzTile c;
  if (g.batch_major) {
    const int64_t nb = (int64_t)g.nb0 * g.nb1, bb = L % nb;
    L /= nb;
    mt = (int)(L % tiles_m);
    nt = (int)(L / tiles_m);
    c.b1 = bb % g.nb1;
    c.b0 = bb / g.nb1;
  } else {
    const int per_batch = tiles_m * tiles_n;
    const int64_t bb = L / per_batch;
    const int r = (int)(L - bb * per_batch);
    const int mg = r / (GM * tiles_n), rem = r - mg * (GM * tiles_n);
    const int gm = min(GM, tiles_m - mg * GM);   // row tiles in this group
    nt = rem / gm;
    mt = mg * GM + (rem - nt * gm);
    c.b1 = bb % g.nb1;
    c.b0 = bb / g.nb1;
  }
  c.m0 = mt * TILE;
  c.n0 = nt * TILE;
  c.bA = g.a_shared ? 0 : c.b0 * g.nb1 + c.b1;
  c.bB = g.b_shared ? 0 : c.b0 * g.nb1 + c.b1;
  return c;
}

// Variant for short k extents (a few k blocks per tile), persistent (waves
// laid out as in oz_gemm_kernel): each workgroup walks several tiles
// (workgroup w, on XCD w % 8,
// takes every nx-th tile of that XCD's contiguous range, nx = the XCD's
// workgroups), so the tiles an XCD runs at once are neighbours in the tile
// order.  The copy of a tile's first k block is issued with the previous
// tile's last k block, so the epilogue of one tile runs while the next
// tile's first stage lands; the exponents of the rows and columns a lane
// stores are fetched with the tile's last k block.
template <class TO, int NSL, int WNW>
__global__ __launch_bounds__(128 * WNW) void oz_gemm_walk_kernel(const OzGemm g, int64_t ntiles) {
  constexpr int DMAX = NSL, NP = 2 * NSL, STAGE = stage_bytes<NSL>(), STAGE_A = STAGE / 2;
  constexpr int NW = 2 * WNW, RJ = 4 / WNW;   // waves; column blocks per wave (2 row blocks each)
  __shared__ __attribute__((aligned(1024))) int8_t lds[2 * STAGE];
  const int tiles_m = (g.M + TILE - 1) / TILE, tiles_n = (g.N + TILE - 1) / TILE;
  const int64_t P = gridDim.x, x = blockIdx.x % 8, slot = blockIdx.x / 8;
  const int64_t nx = (P - x + 7) / 8;
  const int64_t per = ntiles / 8, rem = ntiles % 8;
  const int64_t end = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + per + (x < rem ? 1 : 0);
  int64_t L = end - per - (x < rem ? 1 : 0) + slot;
  if (L >= end) return;   // uniform over the workgroup

  // the tile's A fragments (4 row blocks x NP planes) of one k block are
  // contiguous, and so are its B fragments; consecutive k blocks are
  // kstrideA / kstrideB bytes apart
  const int64_t kstrideA = (int64_t)(g.Mp / RB) * NP * FRAG, kstrideB = (int64_t)(g.Np / RB) * NP * FRAG;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int wm = (w / WNW) * 32, wn = (w % WNW) * 16 * RJ;
  const int lr = lane & 15, lk = lane >> 4;

  // wave w copies fragments w, w + 4, .. of the stage: NSL of A, then NSL of B
  auto copy_stage = [&](const OzTile& c, int kb, int st) {
    const int8_t* srcA = g.Aq + c.bA * g.nkb * kstrideA + (int64_t)(c.m0 / RB) * NP * FRAG +
                         (int64_t)kb * kstrideA + 16 * lane;
    const int8_t* srcB = g.Bq + c.bB * g.nkb * kstrideB + (int64_t)(c.n0 / RB) * NP * FRAG +
                         (int64_t)kb * kstrideB + 16 * lane;
    int8_t* dst = lds + st * STAGE;
#pragma unroll
    for (int it = 0; it < STAGE_A / FRAG / NW; ++it) {
      const int f = w + NW * it;
      __builtin_amdgcn_global_load_lds((const void*)(srcA + f * FRAG),
                                       (__attribute__((address_space(3))) void*)(dst + f * FRAG), 16, 0, 0);
    }
#pragma unroll
    for (int it = 0; it < STAGE_A / FRAG / NW; ++it) {
      const int f = w + NW * it;
      __builtin_amdgcn_global_load_lds((const void*)(srcB + f * FRAG),
                                       (__attribute__((address_space(3))) void*)(dst + STAGE_A + f * FRAG), 16, 0,
                                       0);
    }
  };
  // exponents of the rows / columns this lane stores
  auto load_exps = [&](const OzTile& c, int (&ea)[2][4], int (&eb)[RJ]) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = c.m0 + wm + 16 * i + 4 * lk + r;
        ea[i][r] = row < g.M ? g.EA[c.bA * g.Mp + row] : INT_MIN;
      }
#pragma unroll
    for (int j = 0; j < RJ; ++j) {
      const int col = c.n0 + wn + 16 * j + lr;
      eb[j] = col < g.N ? g.EB[c.bB * g.Np + col] : INT_MIN;
    }
  };

  OzTile cur = oz_tile(g, L, tiles_m, tiles_n);
  v4i re[2][RJ][DMAX + 1], im[2][RJ][DMAX + 1];
  // the products of one k block from stage st
  auto compute = [&](int st) {
    const int8_t* sA = lds + st * STAGE + (wm / RB) * NP * FRAG + 16 * lane;
    const int8_t* sB = lds + st * STAGE + STAGE_A + (wn / RB) * NP * FRAG + 16 * lane;
    // B fragments of slice u: bf[j][2u] (re), bf[j][2u + 1] (im), block column j
    v4i bf[RJ][NP];
#pragma unroll
    for (int u = 0; u < NSL && u <= DMAX - (NSL - 1); ++u)
#pragma unroll
      for (int j = 0; j < RJ; ++j) {
        bf[j][2 * u] = lds_frag(sB + (j * NP + 2 * u) * FRAG);
        bf[j][2 * u + 1] = lds_frag(sB + (j * NP + 2 * u + 1) * FRAG);
      }
    v4i ar = lds_frag(sA + (2 * (NSL - 1)) * FRAG), ai = lds_frag(sA + (2 * (NSL - 1) + 1) * FRAG);
#pragma unroll
    for (int step = 0; step < 2 * NSL; ++step) {
      const int i = step / NSL, s = NSL - 1 - step % NSL;
      v4i ar2 = ar, ai2 = ai;
      if (step + 1 < 2 * NSL) {
        const int i2 = (step + 1) / NSL, s2 = NSL - 1 - (step + 1) % NSL;
        ar2 = lds_frag(sA + (i2 * NP + 2 * s2) * FRAG);
        ai2 = lds_frag(sA + (i2 * NP + 2 * s2 + 1) * FRAG);
      }
      // the slice of B the next step uses first
      if (i == 0 && s >= 1 && DMAX - s + 1 < NSL) {
        const int u = DMAX - s + 1;
#pragma unroll
        for (int j = 0; j < RJ; ++j) {
          bf[j][2 * u] = lds_frag(sB + (j * NP + 2 * u) * FRAG);
          bf[j][2 * u + 1] = lds_frag(sB + (j * NP + 2 * u + 1) * FRAG);
        }
      }
      const v4i nai = neg_bytes(ai);
      // the products with Ar for every slice u of the step, then those with Ai
#pragma unroll
      for (int u = 0; u < NSL && u + s <= DMAX; ++u) {
        const int d = s + u;
#pragma unroll
        for (int j = 0; j < RJ; ++j) {
          re[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ar, bf[j][2 * u], re[i][j][d], 0, 0, 0);
          im[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ar, bf[j][2 * u + 1], im[i][j][d], 0, 0, 0);
        }
      }
#pragma unroll
      for (int u = 0; u < NSL && u + s <= DMAX; ++u) {
        const int d = s + u;
#pragma unroll
        for (int j = 0; j < RJ; ++j) {
          re[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(nai, bf[j][2 * u + 1], re[i][j][d], 0, 0, 0);
          im[i][j][d] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ai, bf[j][2 * u], im[i][j][d], 0, 0, 0);
        }
      }
      ar = ar2;
      ai = ai2;
      // keep each step's MFMAs together, behind the reads issued for the next step
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  int st = 0;
  copy_stage(cur, 0, 0);
  for (;;) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < RJ; ++j)
#pragma unroll
        for (int d = 0; d <= DMAX; ++d) { re[i][j][d] = v4i{0, 0, 0, 0}; im[i][j][d] = v4i{0, 0, 0, 0}; }
    for (int kb = 0; kb + 1 < g.nkb; ++kb, st ^= 1) {
      // stage st has landed for every wave, and every wave is done with the
      // stage the next copy overwrites
      __syncthreads();
      copy_stage(cur, kb + 1, st ^ 1);
      compute(st);
    }
    // the last k block: the exponents the epilogue needs (issued before the
    // next copy, so waiting for them leaves that copy in flight), then the
    // next tile's first k block
    __syncthreads();
    int ea[2][4], eb[RJ];
    load_exps(cur, ea, eb);
    const bool more = L + nx < end;
    const OzTile nxt = oz_tile(g, more ? L + nx : L, tiles_m, tiles_n);
    if (more) copy_stage(nxt, 0, st ^ 1);
    compute(st);
    st ^= 1;
    // epilogue: lane (lr, lk), register r holds row 4 lk + r, column lr of each 16 x 16 block
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < RJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = cur.m0 + wm + 16 * i + 4 * lk + r, col = cur.n0 + wn + 16 * j + lr;
          if (row >= g.M || col >= g.N) continue;
          double vr = (double)re[i][j][DMAX][r], vi = (double)im[i][j][DMAX][r];
#pragma unroll
          for (int d = DMAX - 1; d >= 0; --d) {
            vr = vr * 0.0078125 + (double)re[i][j][d][r];
            vi = vi * 0.0078125 + (double)im[i][j][d][r];
          }
          const int sc = -(sigma_of(ea[i][r]) + sigma_of(eb[j]));
          oz_store<TO>(g, cur.b0, cur.b1, row, col, ldexp(vr, sc), ldexp(vi, sc));
        }
    if (!more) break;
    L += nx;
    cur = nxt;
  }
}

__global__ void oz_fill_kernel(int* p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Compute units of the current device (cached per device).
int compute_units() {
  static int cus[64] = {0};
  int dev = 0;
  GGPU_HIP(hipGetDevice(&dev));
  require(dev >= 0 && dev < 64, "int8 emulation: device index out of range");
  if (cus[dev] == 0) {
    int n = 0;
    GGPU_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    cus[dev] = std::max(1, n);
  }
  return cus[dev];
}

struct Plan {
  int64_t Mp, Np, nkb, nbA, nbB;
  size_t bytesEA, bytesEB, bytesA, bytesB;
  size_t total() const { return bytesEA + bytesEB + bytesA + bytesB + 4 * 256; }
};

Plan plan_for(int M, int N, int64_t kt, int nb0, int nb1, bool a_shared, bool b_shared, int nsl) {
  const int np = 2 * nsl;
  Plan p;
  p.Mp = round_up(M, TILE);
  p.Np = round_up(N, TILE);
  p.nkb = (kt + KBLK - 1) / KBLK;
  p.nbA = a_shared ? 1 : (int64_t)nb0 * nb1;
  p.nbB = b_shared ? 1 : (int64_t)nb0 * nb1;
  p.bytesEA = round_up((size_t)p.nbA * p.Mp * sizeof(int), 256);
  p.bytesEB = round_up((size_t)p.nbB * p.Np * sizeof(int), 256);
  p.bytesA = round_up((size_t)p.nbA * p.nkb * p.Mp * np * KBLK, 256);
  p.bytesB = round_up((size_t)p.nbB * p.nkb * p.Np * np * KBLK, 256);
  return p;
}

void fill_int(int* p, int64_t n, int v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(oz_fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  GGPU_HIP(hipGetLastError());
}

template <int NSL>
void slice_operand(SliceView v, int64_t nb, hipStream_t s) {
  require(nb >= 1 && nb <= 65535 && v.Rp / RB <= 65535 && v.nkb >= 1, "int8 emulation: operand too large");
  const int64_t kt = (int64_t)v.K * v.KB;
  if (kt <= LDS_KT && v.row_s == 1 && v.k_s != 1) {
    // eight waves: the copies and the digit work of the block spread over two
    // waves per SIMD (12 % faster than four at sig_py's Y1)
    hipLaunchKernelGGL((oz_slice_lds_kernel<NSL, 512>), dim3(1, (unsigned)(v.Rp / RB), (unsigned)nb), dim3(512), 0, s,
                       v);
    GGPU_HIP(hipGetLastError());
    return;
  }
  if (kt <= FUSE_KT) {
    hipLaunchKernelGGL((oz_slice_kernel<true, NSL>), dim3(1, (unsigned)(v.Rp / RB), (unsigned)nb), dim3(NT), 0, s, v);
    GGPU_HIP(hipGetLastError());
    return;
  }
  fill_int(v.E, nb * v.Rp, INT_MIN, s);
  hipLaunchKernelGGL(oz_exponent_kernel, dim3((unsigned)v.nkb, (unsigned)(v.Rp / TILE), (unsigned)nb), dim3(NT),
                     0, s, v);
  GGPU_HIP(hipGetLastError());
  if (v.k_s == 1) {   // k-contiguous rows: stage each wave's block through LDS
    hipLaunchKernelGGL((oz_digits_kf_kernel<NSL>), dim3((unsigned)v.nkb, (unsigned)(v.Rp / RB), (unsigned)nb),
                       dim3(64), 0, s, v);
    GGPU_HIP(hipGetLastError());
    return;
  }
  const int ngroups = (v.nkb + 3) / 4;
  hipLaunchKernelGGL((oz_slice_kernel<false, NSL>), dim3((unsigned)ngroups, (unsigned)(v.Rp / RB), (unsigned)nb),
                     dim3(NT), 0, s, v);
  GGPU_HIP(hipGetLastError());
}

// The int8 GEMM of one launch part: `tiles` 64 x 64 tiles of nkb k blocks.
template <int NSL>
void launch_gemm(const OzGemm& g, int64_t tiles, int64_t nkb, bool out_f32, hipStream_t s) {
  require(tiles < ((int64_t)1 << 31), "int8 emulation: grid too large");
  if (nkb <= 4) {
    // short k extents: one workgroup per CU (the two LDS stages fill it),
    // each walking tiles, so a tile's first copy hides behind the previous
    // tile's last k block (K = nao = 104 launches: 5-11 % faster; on the
    // K = 600 and K = 10816 launches walking tiles cost 1-2 %)
    const unsigned grid = (unsigned)std::min<int64_t>(tiles, compute_units());
    // eight waves (two per SIMD): 4-14 % faster than four at the K = nao sites
    if (out_f32) hipLaunchKernelGGL((oz_gemm_walk_kernel<float2, NSL, 4>), dim3(grid), dim3(512), 0, s, g, tiles);
    else hipLaunchKernelGGL((oz_gemm_walk_kernel<double2, NSL, 4>), dim3(grid), dim3(512), 0, s, g, tiles);
  } else if (out_f32) {
    // eight waves (two per SIMD): 1-3 % faster than four on every long-K site
    hipLaunchKernelGGL((oz_gemm_kernel<float2, NSL, 4>), dim3((unsigned)tiles), dim3(512), 0, s, g);
  } else {
    hipLaunchKernelGGL((oz_gemm_kernel<double2, NSL, 4>), dim3((unsigned)tiles), dim3(512), 0, s, g);
  }
  GGPU_HIP(hipGetLastError());
}

// One launch over batches [b0s, b0s + nb0) and K-batch [kb0, kb0 + KBc).
template <int NSL>
void run_part(const GemmDesc& d, int b0s, int nb0, int kb0, int KBc, bool first_k, char* work, size_t work_bytes,
              hipStream_t s) {
  const bool a_shared = d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr;
  const bool b_shared = d.B.bs0 == 0 && d.B.bs1 == 0 && d.B.off1 == nullptr;
  const int64_t kt = (int64_t)d.K * KBc;
  // One A for batches along b0 only (the G.V products of the P0 and Sigma
  // chains): the batches become column groups of one product, C[:, (b, n)]
  // = A B_b, so N = nao is not padded to the tile once per batch (p0_gvh and
  // sig_vg at nao = 104: 11 % faster).  Same integer sums per output.
  const bool merge = a_shared && d.nb1 == 1 && nb0 > 1 && d.KB == 1 && !b_shared && d.B.off1 == nullptr &&
                     d.C.off1 == nullptr && (int64_t)d.N * nb0 < ((int64_t)1 << 30);
  if (merge) {
    const int Nm = d.N * nb0;
    const Plan p = plan_for(d.M, Nm, kt, 1, 1, true, false, NSL);
    require(p.total() <= work_bytes, "int8 emulation: workspace too small for one batch");
    int* EA = reinterpret_cast<int*>(work);
    int* EB = reinterpret_cast<int*>(work + p.bytesEA);
    int8_t* Aq = reinterpret_cast<int8_t*>(work + p.bytesEA + p.bytesEB);
    int8_t* Bq = Aq + p.bytesA;
    SliceView a;
    a.src = static_cast<const double2*>(d.A.ptr);
    a.row_s = d.A.rs; a.k_s = d.A.cs; a.kb_s = d.A.kbs;
    a.rows = d.M; a.K = d.K; a.KB = 1; a.nb1 = 1;
    a.conj = d.A.conj; a.Rp = (int)p.Mp; a.nkb = (int)p.nkb; a.E = EA; a.out = Aq;
    slice_operand<NSL>(a, 1, s);
    SliceView b;
    b.src = static_cast<const double2*>(d.B.ptr) + (int64_t)b0s * d.B.bs0;
    b.row_s = d.B.cs; b.k_s = d.B.rs; b.kb_s = d.B.kbs; b.rgrp = d.N; b.rgrp_s = d.B.bs0;
    b.rows = Nm; b.K = d.K; b.KB = 1; b.nb1 = 1;
    b.conj = d.B.conj; b.Rp = (int)p.Np; b.nkb = (int)p.nkb; b.E = EB; b.out = Bq;
    slice_operand<NSL>(b, 1, s);
    OzGemm g;
    g.Aq = Aq; g.Bq = Bq; g.EA = EA; g.EB = EB;
    g.M = d.M; g.N = Nm; g.Mp = (int)p.Mp; g.Np = (int)p.Np; g.nkb = (int)p.nkb;
    g.nb0 = 1; g.nb1 = 1; g.a_shared = 1; g.b_shared = 0; g.batch_major = 0;
    g.C = static_cast<char*>(d.C.ptr) + (int64_t)b0s * d.C.bs0 * (d.out_f32 ? sizeof(float2) : sizeof(double2));
    g.c_rs = d.C.rs; g.c_cs = d.C.cs; g.cgrp = d.N; g.cgrp_s = d.C.bs0;
    g.alpha_re = d.alpha_re; g.alpha_im = d.alpha_im;
    g.beta_re = first_k ? d.beta_re : 1.0;
    g.beta_im = first_k ? d.beta_im : 0.0;
    launch_gemm<NSL>(g, (int64_t)(p.Mp / TILE) * (p.Np / TILE), p.nkb, d.out_f32, s);
    return;
  }
  const Plan p = plan_for(d.M, d.N, kt, nb0, d.nb1, a_shared, b_shared, NSL);
  require(p.total() <= work_bytes, "int8 emulation: workspace too small for one batch");
  int* EA = reinterpret_cast<int*>(work);
  int* EB = reinterpret_cast<int*>(work + p.bytesEA);
  int8_t* Aq = reinterpret_cast<int8_t*>(work + p.bytesEA + p.bytesEB);
  int8_t* Bq = Aq + p.bytesA;

  SliceView a;
  a.src = static_cast<const double2*>(d.A.ptr) + (int64_t)b0s * d.A.bs0 + (int64_t)kb0 * d.A.kbs;
  a.row_s = d.A.rs; a.k_s = d.A.cs; a.kb_s = d.A.kbs; a.bs0 = d.A.bs0; a.bs1 = d.A.bs1; a.off1 = d.A.off1;
  a.rows = d.M; a.K = d.K; a.KB = KBc; a.nb1 = a_shared ? 1 : d.nb1;
  a.conj = d.A.conj; a.Rp = (int)p.Mp; a.nkb = (int)p.nkb; a.E = EA; a.out = Aq;
  slice_operand<NSL>(a, p.nbA, s);
  SliceView b;
  b.src = static_cast<const double2*>(d.B.ptr) + (int64_t)b0s * d.B.bs0 + (int64_t)kb0 * d.B.kbs;
  b.row_s = d.B.cs; b.k_s = d.B.rs; b.kb_s = d.B.kbs; b.bs0 = d.B.bs0; b.bs1 = d.B.bs1; b.off1 = d.B.off1;
  b.rows = d.N; b.K = d.K; b.KB = KBc; b.nb1 = b_shared ? 1 : d.nb1;
  b.conj = d.B.conj; b.Rp = (int)p.Np; b.nkb = (int)p.nkb; b.E = EB; b.out = Bq;
  slice_operand<NSL>(b, p.nbB, s);

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
  launch_gemm<NSL>(g, (int64_t)(p.Mp / TILE) * (p.Np / TILE) * nb0 * d.nb1, p.nkb, d.out_f32, s);
}

}  // namespace

size_t gemm_int8_emulated_workspace(const GemmDesc& d, int slices) {
  const bool a_shared = d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr;
  const bool b_shared = d.B.bs0 == 0 && d.B.bs1 == 0 && d.B.off1 == nullptr;
  const int KBc = (int)std::max<int64_t>(1, std::min<int64_t>(d.KB, KMAX / std::max(1, d.K)));
  return plan_for(d.M, d.N, (int64_t)d.K * KBc, 1, d.nb1, a_shared, b_shared, slices).total();
}

void gemm_int8_emulated(const GemmDesc& d, int slices, void* work, size_t work_bytes, hipStream_t s) {
  require(slices == 4 || slices == 5, "int8 emulation: 4 or 5 slices");
  require(work != nullptr && work_bytes > 0, "int8 emulation: no workspace installed");
  require(d.K <= KMAX, "int8 emulation: K above " + std::to_string(KMAX));
  const bool a_shared = d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr;
  const bool b_shared = d.B.bs0 == 0 && d.B.bs1 == 0 && d.B.off1 == nullptr;
  // K-batch chunks small enough that no int32 diagonal sum can overflow
  const int KBc = (int)std::max<int64_t>(1, std::min<int64_t>(d.KB, KMAX / std::max(1, d.K)));
  // batch ranges (along b0) whose slices fit the workspace
  int nb0c = d.nb0;
  while (nb0c > 1 &&
         plan_for(d.M, d.N, (int64_t)d.K * KBc, nb0c, d.nb1, a_shared, b_shared, slices).total() > work_bytes)
    nb0c = (nb0c + 1) / 2;
  for (int kb0 = 0; kb0 < d.KB; kb0 += KBc) {
    const int kbn = std::min(KBc, d.KB - kb0);
    for (int b0s = 0; b0s < d.nb0; b0s += nb0c) {
      const int nb = std::min(nb0c, d.nb0 - b0s);
      char* w = static_cast<char*>(work);
      if (slices == 5) run_part<5>(d, b0s, nb, kb0, kbn, kb0 == 0, w, work_bytes, s);
      else run_part<4>(d, b0s, nb, kb0, kbn, kb0 == 0, w, work_bytes, s);
    }
  }
}

}  // namespace ggpu
