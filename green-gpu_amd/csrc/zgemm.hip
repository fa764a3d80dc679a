// Batched complex GEMM on the gfx950 matrix cores.
//
// Workgroup = 4 wave64s computing a 64 x 64 complex tile of C; each wave owns
// a 32 x 32 quadrant = 2 x 2 blocks of the 16 x 16 MFMA output, kept as
// separate real / imaginary accumulators (4 real MFMAs per complex block and
// k-step: re += ar*br - ai*bi, im += ar*bi + ai*br).
//
// Tiles of A (64 x 16) and B (16 x 64) are staged through LDS as split
// real / imaginary planes, k-major, so the MFMA operand reads (lane -> row
// lane%16, k lane/16) are unit-stride across 16 lanes.  The next tile is
// prefetched from global memory into registers while the current one is
// multiplied.  Global loads follow whichever operand dimension is contiguous
// (template flags), so every operand layout of the HF/GW chains streams with
// full 64-byte segments.
//
// Split-bf16 x3 has its own kernels further down: tiles from 64 x 64 to
// 160 x 64 that split complex64 operands into bf16 hi / lo planes while
// staging them, and an A-stationary 64 x 64 tile that splits one shared A
// once and walks several batches.  Each call site is its own kernel symbol;
// the launch variant per site comes from the built-in choice, a per-thread
// table (the plan's autotuned choice) or a global override.
#include "zgemm.hpp"
#include "ozaki.hpp"

#include <algorithm>
#include <cstdio>
#include <type_traits>

namespace ggpu {
namespace {

constexpr int BM = 64;
constexpr int BN = 64;
constexpr int BKT = 16;
constexpr int NTHREADS = 256;
constexpr int PAD = 2;

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <class T> struct Cplx;
template <> struct Cplx<double> { using type = double2; };
template <> struct Cplx<float> { using type = float2; };

template <class T, class V>
__device__ inline V mfma(T a, T b, V c);
template <>
__device__ inline v4d mfma<double, v4d>(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <>
__device__ inline v4f mfma<float, v4f>(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class T> struct Acc;
template <> struct Acc<double> { using type = v4d; };
template <> struct Acc<float> { using type = v4f; };

// Row (within the 16 x 16 block) of accumulator register r of lane group
// lk = lane / 16.  Measured on gfx950: v_mfma_f64_16x16x4_f64 holds rows
// lk + 4 r, v_mfma_f32_16x16x4_f32 holds rows 4 lk + r; the column is lane % 16
// for both.
template <class T> __device__ inline int acc_row(int lk, int r);
template <> __device__ inline int acc_row<double>(int lk, int r) { return lk + 4 * r; }
template <> __device__ inline int acc_row<float>(int lk, int r) { return 4 * lk + r; }

// Workgroup -> tile mapping.  Workgroups are dispatched round-robin over the
// 8 XCDs (each with its own L2); the linear id is remapped so that every XCD
// receives a contiguous range of logical tiles, ordered m-tile fastest, then
// n-tile, then batch: the workgroups that share a B panel (same n-tile) and an
// A panel (same batch) run on the same XCD at the same time.
constexpr int NXCD = 8;
__device__ inline int64_t xcd_logical(int64_t id, int64_t total) {
  const int64_t xcd = id % NXCD, slot = id / NXCD;
  const int64_t per = total / NXCD, rem = total % NXCD;
  return (xcd < rem) ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

// Logical tile -> (m-tile, n-tile, batch b0, batch b1).  Default order:
// m-tile fastest, then n-tile, then batch.  With d.batch_major the batch
// index runs fastest instead: chosen when every batch shares A (A's batch
// strides are zero), so the workgroups reusing one A panel run back to back
// and find it in L2 instead of re-reading A from memory once per batch.
struct TileCoord { int mt, nt; int64_t b0, b1; };
__device__ inline TileCoord tile_coord(const GemmDesc& d, int tiles_m, int tiles_n) {
  int64_t lg = xcd_logical(blockIdx.x, gridDim.x);
  TileCoord c;
  if (d.batch_major) {
    const int64_t nb = (int64_t)d.nb0 * d.nb1;
    const int64_t b = lg % nb;
    lg /= nb;
    c.mt = (int)(lg % tiles_m);
    c.nt = (int)(lg / tiles_m);
    c.b1 = b % d.nb1;
    c.b0 = b / d.nb1;
  } else {
    c.mt = (int)(lg % tiles_m);
    lg /= tiles_m;
    c.nt = (int)(lg % tiles_n);
    lg /= tiles_n;
    c.b1 = lg % d.nb1;
    c.b0 = lg / d.nb1;
  }
  return c;
}

// T: arithmetic type of operands and MFMA; TO: complex element type of C.
// AKC: A's k index is the contiguous one (else its m index).
// BKC: B's k index is the contiguous one (else its n index).
// GAUSS: 3 real products per complex product (t1 = ar br, t2 = ai bi,
//   t3 = (ar + ai)(br + bi); re = t1 - t2, im = t3 - t1 - t2) instead of 4.
// BK: k per LDS stage (16, or 8 for half the LDS and twice the workgroups
// per CU).
// WPE: waves per SIMD the register allocation must allow (1: no constraint).
template <class T, class TO, bool AKC, bool BKC, bool GAUSS, int SITE, int BK = BKT, int WPE = 1>
__global__ __launch_bounds__(NTHREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void gemm_mfma_kernel(
    const GemmDesc d) {
  using CT = typename Cplx<T>::type;
  using AV = typename Acc<T>::type;
  constexpr int EPT = BM * BK / NTHREADS;   // complex elements per thread and operand
  __shared__ T sA[2][2][BK][BM + PAD];   // [buffer][re/im][k][m]
  __shared__ T sB[2][2][BK][BN + PAD];   // [buffer][re/im][k][n]

  const int tid = threadIdx.x;
  const int tiles_m = (d.M + BM - 1) / BM;
  const int tiles_n = (d.N + BN - 1) / BN;
  const TileCoord tc = tile_coord(d, tiles_m, tiles_n);
  const int m0 = tc.mt * BM, n0 = tc.nt * BN;
  const int64_t b1 = tc.b1, b0 = tc.b0;

  const CT* A = static_cast<const CT*>(d.A.ptr) + b0 * d.A.bs0 + (d.A.off1 ? d.A.off1[b1] : b1 * d.A.bs1);
  const CT* B = static_cast<const CT*>(d.B.ptr) + b0 * d.B.bs0 + (d.B.off1 ? d.B.off1[b1] : b1 * d.B.bs1);
  TO* C = static_cast<TO*>(d.C.ptr) + b0 * d.C.bs0 + (d.C.off1 ? d.C.off1[b1] : b1 * d.C.bs1);

  // Loader coordinates inside the tile: EPT consecutive complex elements per
  // thread along the operand's contiguous index.
  constexpr int TK = BK / EPT, TMN = BM / EPT;   // threads per k-row / per m-row
  int a_m[EPT], a_k[EPT], b_k[EPT], b_n[EPT];
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    if (AKC) { a_m[j] = tid / TK; a_k[j] = (tid % TK) * EPT + j; }
    else     { a_k[j] = tid / TMN; a_m[j] = (tid % TMN) * EPT + j; }
    if (BKC) { b_n[j] = tid / TK; b_k[j] = (tid % TK) * EPT + j; }
    else     { b_k[j] = tid / TMN; b_n[j] = (tid % TMN) * EPT + j; }
  }

  const int ktiles = (d.K + BK - 1) / BK;
  const int niter = d.KB * ktiles;
  CT ra[EPT], rb[EPT];

  auto gload = [&](int it, CT (&xa)[EPT], CT (&xb)[EPT]) {
    const int kb = it / ktiles;
    const int k0 = (it - kb * ktiles) * BK;
    const CT* Ab = A + kb * d.A.kbs;
    const CT* Bb = B + kb * d.B.kbs;
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int gm = m0 + a_m[j], gk = k0 + a_k[j];
      xa[j] = (gm < d.M && gk < d.K) ? Ab[gm * d.A.rs + gk * d.A.cs] : CT{0, 0};
      const int gk2 = k0 + b_k[j], gn = n0 + b_n[j];
      xb[j] = (gk2 < d.K && gn < d.N) ? Bb[gk2 * d.B.rs + gn * d.B.cs] : CT{0, 0};
    }
  };
  auto sstore = [&](int buf, const CT (&xa)[EPT], const CT (&xb)[EPT]) {
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      sA[buf][0][a_k[j]][a_m[j]] = xa[j].x;
      sA[buf][1][a_k[j]][a_m[j]] = d.A.conj ? -xa[j].y : xa[j].y;
      sB[buf][0][b_k[j]][b_n[j]] = xb[j].x;
      sB[buf][1][b_k[j]][b_n[j]] = d.B.conj ? -xb[j].y : xb[j].y;
    }
  };

  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  AV acc[GAUSS ? 3 : 2][2][2];
#pragma unroll
  for (int a = 0; a < (GAUSS ? 3 : 2); ++a)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[a][i][j] = AV{0, 0, 0, 0};

  // One k-step of 4: fragments from LDS, then the block products.
  auto kstep = [&](int buf, int kk) {
    T ar[2], ai[2], br[2], bi[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      ar[i] = sA[buf][0][kk + lk][wm + 16 * i + lr];
      ai[i] = sA[buf][1][kk + lk][wm + 16 * i + lr];
      br[i] = sB[buf][0][kk + lk][wn + 16 * i + lr];
      bi[i] = sB[buf][1][kk + lk][wn + 16 * i + lr];
    }
    if (GAUSS) {
      T as[2], bs[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) { as[i] = ar[i] + ai[i]; bs[i] = br[i] + bi[i]; }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[0][i][j] = mfma<T, AV>(ar[i], br[j], acc[0][i][j]);
          acc[1][i][j] = mfma<T, AV>(ai[i], bi[j], acc[1][i][j]);
          acc[GAUSS ? 2 : 0][i][j] = mfma<T, AV>(as[i], bs[j], acc[GAUSS ? 2 : 0][i][j]);
        }
    } else {
      T nai[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) nai[i] = -ai[i];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[0][i][j] = mfma<T, AV>(ar[i], br[j], acc[0][i][j]);
          acc[1][i][j] = mfma<T, AV>(ar[i], bi[j], acc[1][i][j]);
          acc[0][i][j] = mfma<T, AV>(nai[i], bi[j], acc[0][i][j]);
          acc[1][i][j] = mfma<T, AV>(ai[i], br[j], acc[1][i][j]);
        }
    }
  };

  // All k-steps of k-tile it (LDS buffer it & 1).
  auto ksteps = [&](int it) {
    const int buf = it & 1;
    const int k0 = (it % ktiles) * BK;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4)
      if (k0 + kk < d.K) kstep(buf, kk);   // uniform: skip steps past the end of K
  };

  if (niter > 0) {
    gload(0, ra, rb);
    sstore(0, ra, rb);
  }
  __syncthreads();
  for (int it = 0; it < niter; ++it) {
    if (it + 1 < niter) gload(it + 1, ra, rb);   // next k-tile, into registers
    ksteps(it);
    if (it + 1 < niter) sstore((it & 1) ^ 1, ra, rb);
    __syncthreads();
  }

  const bool use_beta = (d.beta_re != 0.0 || d.beta_im != 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + acc_row<T>(lk, r);
        const int col = n0 + wn + 16 * j + lr;
        if (row < d.M && col < d.N) {
          double xr, xi;
          if (GAUSS) {
            const double t1 = acc[0][i][j][r], t2 = acc[1][i][j][r], t3 = acc[GAUSS ? 2 : 0][i][j][r];
            xr = t1 - t2;
            xi = t3 - t1 - t2;
          } else {
            xr = acc[0][i][j][r];
            xi = acc[1][i][j][r];
          }
          double vr = d.alpha_re * xr - d.alpha_im * xi;
          double vi = d.alpha_re * xi + d.alpha_im * xr;
          TO* p = C + row * d.C.rs + col * d.C.cs;
          if (use_beta) {
            const TO o = *p;
            vr += d.beta_re * o.x - d.beta_im * o.y;
            vi += d.beta_re * o.y + d.beta_im * o.x;
          }
          using TS = decltype(TO{}.x);
          *p = TO{static_cast<TS>(vr), static_cast<TS>(vi)};
        }
      }
}

// ---------------------------------------------------------------------------
// Split-bf16 x3 kernel.  Operands are complex64 in memory; every element x is
// split while it is staged into LDS as hi = bf16(x), lo = bf16(x - hi), and
// each real product is formed as hi*hi + hi*lo + lo*hi on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation (the lo*lo term, ~2^-16
// relative, is dropped).  A complex 16x16x32 block costs 12 MFMAs:
//   re += ar br - ai bi,  im += ar bi + ai br  (3 products each).
// LDS holds 4 planes per operand (re_hi, im_hi, re_lo, im_lo) with k
// contiguous per row, so every MFMA fragment is one 16-byte LDS read.
// ---------------------------------------------------------------------------
constexpr int XBK = 32;          // k per MFMA step
constexpr int XPAD = 8;          // bf16 elements of row padding
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

__device__ inline unsigned short bf16_rn(float x) {
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

template <class TO, bool AKC, bool BKC, int SITE>
__global__ __launch_bounds__(NTHREADS) void gemm_bf16x3_kernel(const GemmDesc d) {
  __shared__ __attribute__((aligned(16))) unsigned short sA[4][BM][XBK + XPAD];
  __shared__ __attribute__((aligned(16))) unsigned short sB[4][BN][XBK + XPAD];

  const int tid = threadIdx.x;
  const int tiles_m = (d.M + BM - 1) / BM;
  const int tiles_n = (d.N + BN - 1) / BN;
  const TileCoord tc = tile_coord(d, tiles_m, tiles_n);
  const int m0 = tc.mt * BM, n0 = tc.nt * BN;
  const int64_t b1 = tc.b1, b0 = tc.b0;
  const float2* A = static_cast<const float2*>(d.A.ptr) + b0 * d.A.bs0 + (d.A.off1 ? d.A.off1[b1] : b1 * d.A.bs1);
  const float2* B = static_cast<const float2*>(d.B.ptr) + b0 * d.B.bs0 + (d.B.off1 ? d.B.off1[b1] : b1 * d.B.bs1);
  TO* C = static_cast<TO*>(d.C.ptr) + b0 * d.C.bs0 + (d.C.off1 ? d.C.off1[b1] : b1 * d.C.bs1);

  // loader: 8 complex elements of A and of B per thread per k-tile
  int am[8], ak[8], bn[8], bk[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (AKC) { am[j] = tid >> 2; ak[j] = ((tid & 3) << 3) + j; }
    else     { ak[j] = tid >> 3; am[j] = ((tid & 7) << 3) + j; }
    if (BKC) { bn[j] = tid >> 2; bk[j] = ((tid & 3) << 3) + j; }
    else     { bk[j] = tid >> 3; bn[j] = ((tid & 7) << 3) + j; }
  }
  const int ktiles = (d.K + XBK - 1) / XBK;
  const int niter = d.KB * ktiles;
  float2 ra[8], rb[8];
  auto gload = [&](int it) {
    const int kb = it / ktiles;
    const int k0 = (it - kb * ktiles) * XBK;
    const float2* Ab = A + kb * d.A.kbs;
    const float2* Bb = B + kb * d.B.kbs;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int gm = m0 + am[j], gk = k0 + ak[j];
      ra[j] = (gm < d.M && gk < d.K) ? Ab[gm * d.A.rs + gk * d.A.cs] : float2{0.f, 0.f};
      const int gk2 = k0 + bk[j], gn = n0 + bn[j];
      rb[j] = (gk2 < d.K && gn < d.N) ? Bb[gk2 * d.B.rs + gn * d.B.cs] : float2{0.f, 0.f};
    }
  };
  auto split_store = [&](unsigned short (*S)[BM][XBK + XPAD], int row, int k, float re, float im) {
    const unsigned short rh = bf16_rn(re), ih = bf16_rn(im);
    S[0][row][k] = rh;
    S[1][row][k] = ih;
    S[2][row][k] = bf16_rn(re - bf16_f(rh));
    S[3][row][k] = bf16_rn(im - bf16_f(ih));
  };
  auto sstore = [&]() {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      split_store(sA, am[j], ak[j], ra[j].x, d.A.conj ? -ra[j].y : ra[j].y);
      split_store(sB, bn[j], bk[j], rb[j].x, d.B.conj ? -rb[j].y : rb[j].y);
    }
  };

  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  v4f cr[2][2], ci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) { cr[i][j] = v4f{0, 0, 0, 0}; ci[i][j] = v4f{0, 0, 0, 0}; }

  auto frag = [&](unsigned short (*S)[BM][XBK + XPAD], int p, int row) -> v8bf {
    const v4u u = *reinterpret_cast<const v4u*>(&S[p][row][8 * lk]);
    return __builtin_bit_cast(v8bf, u);
  };
  auto negf = [&](unsigned short (*S)[BM][XBK + XPAD], int p, int row) -> v8bf {
    v4u u = *reinterpret_cast<const v4u*>(&S[p][row][8 * lk]);
    u ^= v4u{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    return __builtin_bit_cast(v8bf, u);
  };

  if (niter > 0) gload(0);
  for (int it = 0; it < niter; ++it) {
    sstore();
    __syncthreads();
    if (it + 1 < niter) gload(it + 1);
    v8bf arh[2], aih[2], arl[2], ail[2], naih[2], nail[2], brh[2], bih[2], brl[2], bil[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int ra_ = wm + 16 * i + lr, rb_ = wn + 16 * i + lr;
      arh[i] = frag(sA, 0, ra_); aih[i] = frag(sA, 1, ra_); arl[i] = frag(sA, 2, ra_); ail[i] = frag(sA, 3, ra_);
      naih[i] = negf(sA, 1, ra_); nail[i] = negf(sA, 3, ra_);
      brh[i] = frag(sB, 0, rb_); bih[i] = frag(sB, 1, rb_); brl[i] = frag(sB, 2, rb_); bil[i] = frag(sB, 3, rb_);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        v4f r = cr[i][j], m = ci[i][j];
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh[i], brh[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh[i], bih[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh[i], brl[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh[i], bil[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arl[i], brh[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arl[i], bih[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih[i], bih[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih[i], brh[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih[i], bil[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih[i], brl[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nail[i], bih[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ail[i], brh[j], m, 0, 0, 0);
        cr[i][j] = r;
        ci[i][j] = m;
      }
    __syncthreads();
  }

  const bool use_beta = (d.beta_re != 0.0 || d.beta_im != 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + 4 * lk + r;
        const int col = n0 + wn + 16 * j + lr;
        if (row < d.M && col < d.N) {
          const double xr = cr[i][j][r], xi = ci[i][j][r];
          double vr = d.alpha_re * xr - d.alpha_im * xi;
          double vi = d.alpha_re * xi + d.alpha_im * xr;
          TO* p = C + row * d.C.rs + col * d.C.cs;
          if (use_beta) {
            const TO o = *p;
            vr += d.beta_re * o.x - d.beta_im * o.y;
            vi += d.beta_re * o.y + d.beta_im * o.x;
          }
          using TS = decltype(TO{}.x);
          *p = TO{static_cast<TS>(vr), static_cast<TS>(vi)};
        }
      }
}

// ---------------------------------------------------------------------------
// Large-tile split-bf16 x3 kernel.  Workgroup tile (32 WMB) x (32 WNB) of C,
// 4 waves in 2 x 2, each wave WMB x WNB blocks of 16 x 16.  The bigger tile
// halves the operand traffic per MFMA against the 64 x 64 kernel above, and
// the loader is organised in groups of 8 consecutive k of one row:
//   k contiguous in memory:  group g -> row g / 4, k0 = 8 (g % 4)
//                            (4 lanes stream one 256-byte row segment)
//   row contiguous:          group g -> row g % rows, k0 = 8 (g / rows)
//                            (lanes stream consecutive rows of one k)
// so every global load instruction is coalesced in both orientations and
// each group is split into the four bf16 planes with one 16-byte LDS store
// per plane.
// ---------------------------------------------------------------------------
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
// (a, b) -> packed bf16 hi = rne(a), rne(b) and lo = rne(a - hi_a), rne(b - hi_b),
// element a in the low half (v_cvt_pk_bf16_f32 does the rounding)
__device__ inline void split2(float a, float b, unsigned& hi, unsigned& lo) {
  const unsigned h = __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{a, b}, v2bf));
  const float ha = __uint_as_float(h << 16), hb = __uint_as_float(h & 0xffff0000u);
  lo = __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{a - ha, b - hb}, v2bf));
  hi = h;
}

template <class TO, bool AKC, bool BKC, int WMB, int WNB, int SITE, int WPE = 1>
__global__ __launch_bounds__(NTHREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void gemm_bf16x3_big_kernel(
    const GemmDesc d) {
  constexpr int TM = 32 * WMB, TN = 32 * WNB, LDK = XBK + XPAD;
  constexpr int GA = (TM * 4 + NTHREADS - 1) / NTHREADS;   // 8-element groups per thread
  constexpr int GB = (TN * 4 + NTHREADS - 1) / NTHREADS;
  __shared__ __attribute__((aligned(16))) unsigned short sA[4][TM][LDK];
  __shared__ __attribute__((aligned(16))) unsigned short sB[4][TN][LDK];

  const int tid = threadIdx.x;
  const int tiles_m = (d.M + TM - 1) / TM;
  const int tiles_n = (d.N + TN - 1) / TN;
  const TileCoord tc = tile_coord(d, tiles_m, tiles_n);
  const int m0 = tc.mt * TM, n0 = tc.nt * TN;
  const int64_t b1 = tc.b1, b0 = tc.b0;
  const float2* A = static_cast<const float2*>(d.A.ptr) + b0 * d.A.bs0 + (d.A.off1 ? d.A.off1[b1] : b1 * d.A.bs1);
  const float2* B = static_cast<const float2*>(d.B.ptr) + b0 * d.B.bs0 + (d.B.off1 ? d.B.off1[b1] : b1 * d.B.bs1);
  TO* C = static_cast<TO*>(d.C.ptr) + b0 * d.C.bs0 + (d.C.off1 ? d.C.off1[b1] : b1 * d.C.bs1);

  int arow[GA], ak0[GA], brow[GB], bk0[GB];
#pragma unroll
  for (int u = 0; u < GA; ++u) {
    const int g = tid + u * NTHREADS;
    arow[u] = AKC ? (g >> 2) : (g % TM);
    ak0[u] = AKC ? ((g & 3) << 3) : ((g / TM) << 3);
    if (g >= TM * 4) arow[u] = -1;
  }
#pragma unroll
  for (int u = 0; u < GB; ++u) {
    const int g = tid + u * NTHREADS;
    brow[u] = BKC ? (g >> 2) : (g % TN);
    bk0[u] = BKC ? ((g & 3) << 3) : ((g / TN) << 3);
    if (g >= TN * 4) brow[u] = -1;
  }
  const int ktiles = (d.K + XBK - 1) / XBK;
  const int niter = d.KB * ktiles;
  float2 ra[GA][8], rb[GB][8];
  auto gload = [&](int it) {
    const int kb = it / ktiles;
    const int kt = (it - kb * ktiles) * XBK;
    const float2* Ab = A + kb * d.A.kbs;
    const float2* Bb = B + kb * d.B.kbs;
#pragma unroll
    for (int u = 0; u < GA; ++u) {
      const int gm = m0 + arow[u];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int gk = kt + ak0[u] + j;
        ra[u][j] = (arow[u] >= 0 && gm < d.M && gk < d.K) ? Ab[gm * d.A.rs + gk * d.A.cs] : float2{0.f, 0.f};
      }
    }
#pragma unroll
    for (int u = 0; u < GB; ++u) {
      const int gn = n0 + brow[u];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int gk = kt + bk0[u] + j;
        rb[u][j] = (brow[u] >= 0 && gn < d.N && gk < d.K) ? Bb[gk * d.B.rs + gn * d.B.cs] : float2{0.f, 0.f};
      }
    }
  };
  // split 8 consecutive-k elements into the re_hi, im_hi, re_lo, im_lo planes
  auto split8 = [&](unsigned short (*S)[LDK], int rows, int row, int k0, const float2* x, bool conj) {
    unsigned hr[4], hi[4], lr_[4], li[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float s = conj ? -1.f : 1.f;
      split2(x[2 * q].x, x[2 * q + 1].x, hr[q], lr_[q]);
      split2(s * x[2 * q].y, s * x[2 * q + 1].y, hi[q], li[q]);
    }
    *reinterpret_cast<v4u*>(&S[0 * rows + row][k0]) = v4u{hr[0], hr[1], hr[2], hr[3]};
    *reinterpret_cast<v4u*>(&S[1 * rows + row][k0]) = v4u{hi[0], hi[1], hi[2], hi[3]};
    *reinterpret_cast<v4u*>(&S[2 * rows + row][k0]) = v4u{lr_[0], lr_[1], lr_[2], lr_[3]};
    *reinterpret_cast<v4u*>(&S[3 * rows + row][k0]) = v4u{li[0], li[1], li[2], li[3]};
  };
  auto sstore = [&]() {
#pragma unroll
    for (int u = 0; u < GA; ++u)
      if (arow[u] >= 0) split8(&sA[0][0], TM, arow[u], ak0[u], ra[u], d.A.conj != 0);
#pragma unroll
    for (int u = 0; u < GB; ++u)
      if (brow[u] >= 0) split8(&sB[0][0], TN, brow[u], bk0[u], rb[u], d.B.conj != 0);
  };

  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * (16 * WMB), wn = (w & 1) * (16 * WNB);
  const int lr = lane & 15, lk = lane >> 4;
  v4f cr[WMB][WNB], ci[WMB][WNB];
#pragma unroll
  for (int i = 0; i < WMB; ++i)
#pragma unroll
    for (int j = 0; j < WNB; ++j) { cr[i][j] = v4f{0, 0, 0, 0}; ci[i][j] = v4f{0, 0, 0, 0}; }
  const v4u sign = v4u{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};

  if (niter > 0) gload(0);
  for (int it = 0; it < niter; ++it) {
    sstore();
    __syncthreads();
    if (it + 1 < niter) gload(it + 1);
    v8bf brh[WNB], bih[WNB], brl[WNB], bil[WNB];
#pragma unroll
    for (int j = 0; j < WNB; ++j) {
      const int rb_ = wn + 16 * j + lr;
      brh[j] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sB[0][rb_][8 * lk]));
      bih[j] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sB[1][rb_][8 * lk]));
      brl[j] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sB[2][rb_][8 * lk]));
      bil[j] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sB[3][rb_][8 * lk]));
    }
#pragma unroll
    for (int i = 0; i < WMB; ++i) {
      const int ra_ = wm + 16 * i + lr;
      const v4u uih = *reinterpret_cast<const v4u*>(&sA[1][ra_][8 * lk]);
      const v4u uil = *reinterpret_cast<const v4u*>(&sA[3][ra_][8 * lk]);
      const v8bf arh = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sA[0][ra_][8 * lk]));
      const v8bf arl = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sA[2][ra_][8 * lk]));
      const v8bf aih = __builtin_bit_cast(v8bf, uih), ail = __builtin_bit_cast(v8bf, uil);
      const v8bf naih = __builtin_bit_cast(v8bf, uih ^ sign), nail = __builtin_bit_cast(v8bf, uil ^ sign);
      // product-major order: consecutive updates of one accumulator are
      // 2 WNB - 1 MFMAs apart
#define GGPU_X3_STEP(acc, a, b)                                                       \
  _Pragma("unroll") for (int j = 0; j < WNB; ++j) acc[i][j] =                         \
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
      GGPU_X3_STEP(cr, arh, brh) GGPU_X3_STEP(ci, arh, bih)
      GGPU_X3_STEP(cr, arh, brl) GGPU_X3_STEP(ci, arh, bil)
      GGPU_X3_STEP(cr, arl, brh) GGPU_X3_STEP(ci, arl, bih)
      GGPU_X3_STEP(cr, naih, bih) GGPU_X3_STEP(ci, aih, brh)
      GGPU_X3_STEP(cr, naih, bil) GGPU_X3_STEP(ci, aih, brl)
      GGPU_X3_STEP(cr, nail, bih) GGPU_X3_STEP(ci, ail, brh)
#undef GGPU_X3_STEP
    }
    __syncthreads();
  }

  const bool use_beta = (d.beta_re != 0.0 || d.beta_im != 0.0);
#pragma unroll
  for (int i = 0; i < WMB; ++i)
#pragma unroll
    for (int j = 0; j < WNB; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + 4 * lk + r;
        const int col = n0 + wn + 16 * j + lr;
        if (row < d.M && col < d.N) {
          const double xr = cr[i][j][r], xi = ci[i][j][r];
          double vr = d.alpha_re * xr - d.alpha_im * xi;
          double vi = d.alpha_re * xi + d.alpha_im * xr;
          TO* p = C + row * d.C.rs + col * d.C.cs;
          if (use_beta) {
            const TO o = *p;
            vr += d.beta_re * o.x - d.beta_im * o.y;
            vi += d.beta_re * o.y + d.beta_im * o.x;
          }
          using TS = decltype(TO{}.x);
          *p = TO{static_cast<TS>(vr), static_cast<TS>(vi)};
        }
      }
}

// A-stationary split-bf16 x3 kernel, for K <= 2 XBK, no K-batch and one A
// shared by every batch (the G.V products of the P0 and Sigma chains multiply
// one V by every tau).  A workgroup splits its A tile once, keeps its waves'
// A fragments (four bf16 planes, both k-steps) in registers and walks `bpw`
// consecutive batches, staging only B per batch (the next B is loaded while
// the current one is multiplied).  Per output element the products are
// accumulated in the order of gemm_bf16x3_big_kernel, so the sums are equal.
template <class TO, bool AKC, bool BKC, int SITE>
__global__ __launch_bounds__(NTHREADS) __attribute__((amdgpu_waves_per_eu(2))) void gemm_bf16x3_astat_kernel(
    const GemmDesc d, int bpw) {
  constexpr int WMB = 2, WNB = 2;
  constexpr int TM = 32 * WMB, TN = 32 * WNB, KT = 2 * XBK, LDK2 = KT + XPAD;
  constexpr int GA = (TM * (KT / 8) + NTHREADS - 1) / NTHREADS;   // 8-element k-groups per thread
  constexpr int GB = (TN * (KT / 8) + NTHREADS - 1) / NTHREADS;
  __shared__ __attribute__((aligned(16))) unsigned short sA[4][TM][LDK2];
  __shared__ __attribute__((aligned(16))) unsigned short sB[4][TN][LDK2];

  const int tid = threadIdx.x;
  const int tiles_m = (d.M + TM - 1) / TM;
  const int tiles_n = (d.N + TN - 1) / TN;
  const int64_t nbatch = (int64_t)d.nb0 * d.nb1;
  // m-tile fastest: the workgroups sharing a run of B tiles sit side by side
  int64_t lg = xcd_logical(blockIdx.x, gridDim.x);
  const int mt = (int)(lg % tiles_m);
  lg /= tiles_m;
  const int nt = (int)(lg % tiles_n);
  const int64_t chunk = lg / tiles_n;
  const int m0 = mt * TM, n0 = nt * TN;
  const float2* A = static_cast<const float2*>(d.A.ptr);

  constexpr int GPR = KT / 8;   // groups per row
  int arow[GA], ak0[GA], brow[GB], bk0[GB];
#pragma unroll
  for (int u = 0; u < GA; ++u) {
    const int g = tid + u * NTHREADS;
    arow[u] = AKC ? (g / GPR) : (g % TM);
    ak0[u] = AKC ? ((g % GPR) << 3) : ((g / TM) << 3);
    if (g >= TM * GPR) arow[u] = -1;
  }
#pragma unroll
  for (int u = 0; u < GB; ++u) {
    const int g = tid + u * NTHREADS;
    brow[u] = BKC ? (g / GPR) : (g % TN);
    bk0[u] = BKC ? ((g % GPR) << 3) : ((g / TN) << 3);
    if (g >= TN * GPR) brow[u] = -1;
  }
  auto split8 = [&](unsigned short (*S)[LDK2], int rows, int row, int k0, const float2* x, bool conj) {
    unsigned hr[4], hi[4], lr_[4], li[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float sg = conj ? -1.f : 1.f;
      split2(x[2 * q].x, x[2 * q + 1].x, hr[q], lr_[q]);
      split2(sg * x[2 * q].y, sg * x[2 * q + 1].y, hi[q], li[q]);
    }
    *reinterpret_cast<v4u*>(&S[0 * rows + row][k0]) = v4u{hr[0], hr[1], hr[2], hr[3]};
    *reinterpret_cast<v4u*>(&S[1 * rows + row][k0]) = v4u{hi[0], hi[1], hi[2], hi[3]};
    *reinterpret_cast<v4u*>(&S[2 * rows + row][k0]) = v4u{lr_[0], lr_[1], lr_[2], lr_[3]};
    *reinterpret_cast<v4u*>(&S[3 * rows + row][k0]) = v4u{li[0], li[1], li[2], li[3]};
  };

  // A tile: load, split, keep the fragments in registers
  {
    float2 ra[GA][8];
#pragma unroll
    for (int u = 0; u < GA; ++u) {
      const int gm = m0 + arow[u];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int gk = ak0[u] + j;
        ra[u][j] = (arow[u] >= 0 && gm < d.M && gk < d.K) ? A[gm * d.A.rs + gk * d.A.cs] : float2{0.f, 0.f};
      }
    }
#pragma unroll
    for (int u = 0; u < GA; ++u)
      if (arow[u] >= 0) split8(&sA[0][0], TM, arow[u], ak0[u], ra[u], d.A.conj != 0);
  }
  __syncthreads();
  const int lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * (16 * WMB), wn = (w & 1) * (16 * WNB);
  const int lr = lane & 15, lk = lane >> 4;
  const int ksteps = (d.K + XBK - 1) / XBK;   // 1 or 2
  v4u a_rh[WMB][2], a_ih[WMB][2], a_rl[WMB][2], a_il[WMB][2];
#pragma unroll
  for (int i = 0; i < WMB; ++i)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int ra_ = wm + 16 * i + lr, kc = ks * XBK + 8 * lk;
      a_rh[i][ks] = *reinterpret_cast<const v4u*>(&sA[0][ra_][kc]);
      a_ih[i][ks] = *reinterpret_cast<const v4u*>(&sA[1][ra_][kc]);
      a_rl[i][ks] = *reinterpret_cast<const v4u*>(&sA[2][ra_][kc]);
      a_il[i][ks] = *reinterpret_cast<const v4u*>(&sA[3][ra_][kc]);
    }
  const v4u sign = v4u{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};

  const int64_t b_first = chunk * bpw;
  const int nb = (int)(nbatch - b_first < bpw ? nbatch - b_first : bpw);
  float2 rb[GB][8];
  auto gload_b = [&](int j) {
    const int64_t b = b_first + j, b0 = b / d.nb1, b1 = b % d.nb1;
    const float2* B = static_cast<const float2*>(d.B.ptr) + b0 * d.B.bs0 + (d.B.off1 ? d.B.off1[b1] : b1 * d.B.bs1);
#pragma unroll
    for (int u = 0; u < GB; ++u) {
      const int gn = n0 + brow[u];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int gk = bk0[u] + q;
        rb[u][q] = (brow[u] >= 0 && gn < d.N && gk < d.K) ? B[gk * d.B.rs + gn * d.B.cs] : float2{0.f, 0.f};
      }
    }
  };
  const bool use_beta = (d.beta_re != 0.0 || d.beta_im != 0.0);
  if (nb > 0) gload_b(0);
  for (int j = 0; j < nb; ++j) {
    __syncthreads();   // every wave is done with the previous B (and, first time, with sA)
#pragma unroll
    for (int u = 0; u < GB; ++u)
      if (brow[u] >= 0) split8(&sB[0][0], TN, brow[u], bk0[u], rb[u], d.B.conj != 0);
    __syncthreads();
    if (j + 1 < nb) gload_b(j + 1);
    v4f cr[WMB][WNB], ci[WMB][WNB];
#pragma unroll
    for (int i = 0; i < WMB; ++i)
#pragma unroll
      for (int q = 0; q < WNB; ++q) { cr[i][q] = v4f{0, 0, 0, 0}; ci[i][q] = v4f{0, 0, 0, 0}; }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      if (ks >= ksteps) break;   // uniform
      v8bf brh[WNB], bih[WNB], brl[WNB], bil[WNB];
#pragma unroll
      for (int q = 0; q < WNB; ++q) {
        const int rb_ = wn + 16 * q + lr, kc = ks * XBK + 8 * lk;
        brh[q] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sB[0][rb_][kc]));
        bih[q] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sB[1][rb_][kc]));
        brl[q] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sB[2][rb_][kc]));
        bil[q] = __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(&sB[3][rb_][kc]));
      }
#pragma unroll
      for (int i = 0; i < WMB; ++i) {
        const v8bf arh = __builtin_bit_cast(v8bf, a_rh[i][ks]), arl = __builtin_bit_cast(v8bf, a_rl[i][ks]);
        const v8bf aih = __builtin_bit_cast(v8bf, a_ih[i][ks]), ail = __builtin_bit_cast(v8bf, a_il[i][ks]);
        const v8bf naih = __builtin_bit_cast(v8bf, a_ih[i][ks] ^ sign), nail = __builtin_bit_cast(v8bf, a_il[i][ks] ^ sign);
#define GGPU_X3_STEP(acc, a, b)                                                       \
  _Pragma("unroll") for (int q = 0; q < WNB; ++q) acc[i][q] =                         \
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[q], acc[i][q], 0, 0, 0);
        GGPU_X3_STEP(cr, arh, brh) GGPU_X3_STEP(ci, arh, bih)
        GGPU_X3_STEP(cr, arh, brl) GGPU_X3_STEP(ci, arh, bil)
        GGPU_X3_STEP(cr, arl, brh) GGPU_X3_STEP(ci, arl, bih)
        GGPU_X3_STEP(cr, naih, bih) GGPU_X3_STEP(ci, aih, brh)
        GGPU_X3_STEP(cr, naih, bil) GGPU_X3_STEP(ci, aih, brl)
        GGPU_X3_STEP(cr, nail, bih) GGPU_X3_STEP(ci, ail, brh)
#undef GGPU_X3_STEP
      }
    }
    const int64_t b = b_first + j, b0 = b / d.nb1, b1 = b % d.nb1;
    TO* C = static_cast<TO*>(d.C.ptr) + b0 * d.C.bs0 + (d.C.off1 ? d.C.off1[b1] : b1 * d.C.bs1);
#pragma unroll
    for (int i = 0; i < WMB; ++i)
#pragma unroll
      for (int q = 0; q < WNB; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + 16 * i + 4 * lk + r;
          const int col = n0 + wn + 16 * q + lr;
          if (row < d.M && col < d.N) {
            const double xr = cr[i][q][r], xi = ci[i][q][r];
            double vr = d.alpha_re * xr - d.alpha_im * xi;
            double vi = d.alpha_re * xi + d.alpha_im * xr;
            TO* p = C + row * d.C.rs + col * d.C.cs;
            if (use_beta) {
              const TO o = *p;
              vr += d.beta_re * o.x - d.beta_im * o.y;
              vi += d.beta_re * o.y + d.beta_im * o.x;
            }
            using TS = decltype(TO{}.x);
            *p = TO{static_cast<TS>(vr), static_cast<TS>(vi)};
          }
        }
  }
}

// Whether the A-stationary kernel applies: one A for every batch, no
// K-batch, K within its two k-steps, and several batches to walk.
inline bool astat_fits(const GemmDesc& d) {
  return d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr && d.KB == 1 && d.K <= 2 * XBK &&
         (int64_t)d.nb0 * d.nb1 >= 2;
}

thread_local GemmProfiler* g_profiler = nullptr;
thread_local void* g_workspace = nullptr;
thread_local size_t g_workspace_bytes = 0;
thread_local const int* g_site_variants = nullptr;   // [kNumSites], -1: built-in choice

// ---------------------------------------------------------------------------
// Split-bf16 x3 from pre-split, fragment-native operands.  A split kernel
// turns each complex64 operand into bf16 planes (re_hi, im_hi, re_lo, im_lo:
// hi = bf16(x), lo = bf16(x - hi), as split2) stored as 1 KiB fragments, one
// per (batch, k block of 32, 16-row block, plane): lane l holds row l % 16
// and the 8 k values 8 (l / 16) .. + 7 -- the operand layout of
// v_mfma_f32_16x16x32_bf16 -- so the GEMM moves every fragment HBM -> LDS
// with one load-to-LDS wave instruction and LDS -> registers with one
// conflict-free 16-byte read per lane, and splits nothing.  Each K-batch
// entry's K is padded to a multiple of 32 on its own (the k blocks of the
// other split-bf16 tiles), and every output sums its products in their
// order, so all split-bf16 tiles give the same sums.
// ---------------------------------------------------------------------------
constexpr int FB_KB = 32;        // k per fragment / MFMA step
constexpr int FB_TILE = 128;     // workgroup tile rows (columns: 128 or 192)
constexpr int FB_FRAG = 1024;
constexpr int FB_NP = 4;         // planes per element
// both operands' fragments of one k block: 64 KiB (128 columns) or 80 KiB (192)
template <int TN>
constexpr int fb_stage() { return (FB_TILE / 16 + TN / 16) * FB_NP * FB_FRAG; }

struct FbView {
  const float2* src = nullptr;   // element (row 0, k 0, kb 0, batch 0)
  int64_t row_s = 0, k_s = 0, kb_s = 0, bs0 = 0, bs1 = 0;
  const int64_t* off1 = nullptr;
  int rows = 0, K = 0, KB = 1, nb1 = 1, conj = 0;
  int Rp = 0, kpb = 0;           // padded rows; k blocks per K-batch entry
  int8_t* out = nullptr;
};

// One wave per (16-row block, k block): lane l splits row l % 16, k values
// 8 (l / 16) .. + 7 and writes its 16 bytes of each plane's fragment.
__global__ __launch_bounds__(256) void fb_split_kernel(const FbView v, int64_t nkb) {
  const int lane = threadIdx.x & 63;
  const int64_t kblk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (kblk >= nkb) return;
  const int rb = blockIdx.y;
  const int64_t b = blockIdx.z;
  const int64_t b0 = b / v.nb1, b1 = b - b0 * v.nb1;
  const float2* base = v.src + b0 * v.bs0 + (v.off1 ? v.off1[b1] : b1 * v.bs1);
  const int row = rb * 16 + (lane & 15);
  const int kb = (int)(kblk / v.kpb);
  const int k0 = (int)(kblk - (int64_t)kb * v.kpb) * FB_KB + 8 * (lane >> 4);
  const float s = v.conj ? -1.f : 1.f;
  float2 x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = k0 + j;
    x[j] = (row < v.rows && k < v.K) ? base[kb * v.kb_s + row * v.row_s + k * v.k_s] : float2{0.f, 0.f};
  }
  unsigned hr[4], hi[4], lr_[4], li[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    split2(x[2 * q].x, x[2 * q + 1].x, hr[q], lr_[q]);
    split2(s * x[2 * q].y, s * x[2 * q + 1].y, hi[q], li[q]);
  }
  int8_t* o = v.out + ((b * nkb + kblk) * (v.Rp / 16) + rb) * (int64_t)FB_NP * FB_FRAG + 16 * lane;
  *reinterpret_cast<v4u*>(o + 0 * FB_FRAG) = v4u{hr[0], hr[1], hr[2], hr[3]};
  *reinterpret_cast<v4u*>(o + 1 * FB_FRAG) = v4u{hi[0], hi[1], hi[2], hi[3]};
  *reinterpret_cast<v4u*>(o + 2 * FB_FRAG) = v4u{lr_[0], lr_[1], lr_[2], lr_[3]};
  *reinterpret_cast<v4u*>(o + 3 * FB_FRAG) = v4u{li[0], li[1], li[2], li[3]};
}

struct FbGemm {
  const int8_t* Aq = nullptr;
  const int8_t* Bq = nullptr;
  int M = 0, N = 0, Mp = 0, Np = 0;
  int64_t nkb = 0;
  int nb0 = 1, nb1 = 1, a_shared = 0, b_shared = 0, batch_major = 0;
  void* C = nullptr;
  int64_t c_rs = 0, c_cs = 0, c_bs0 = 0, c_bs1 = 0;
  const int64_t* c_off1 = nullptr;
  double alpha_re = 1.0, alpha_im = 0.0, beta_re = 0.0, beta_im = 0.0;
};

// 128 x TN tile per workgroup of WMW x WNW waves, each wave RI x RJ blocks
// of the 16 x 16 MFMA output: for TN = 128, 2 x 2 waves of 64 x 64 (one wave
// per SIMD), 4 x 2 of 32 x 64 (two), 4 x 4 of 32 x 32 (four); for TN = 192,
// 2 x 4 waves of 64 x 48 (two).  A stage holds both operands' fragments of
// one k block; two stages, one barrier per k block; a wave keeps its B
// fragments of the k block in registers and walks its row blocks, reading
// the next row block's A fragments ahead.
template <class TO, int WMW, int WNW, int TN>
__global__ __launch_bounds__(64 * WMW * WNW) void gemm_bf16x3_frag_kernel(const FbGemm g) {
  constexpr int NW = WMW * WNW, RI = 8 / WMW, RJ = TN / 16 / WNW;
  constexpr int STAGE = fb_stage<TN>(), STAGE_A = (FB_TILE / 16) * FB_NP * FB_FRAG;
  constexpr int STAGE_B = STAGE - STAGE_A;
  __shared__ __attribute__((aligned(1024))) int8_t lds[2 * STAGE];
  const int tiles_m = g.Mp / FB_TILE, tiles_n = g.Np / TN;
  int64_t L = xcd_logical(blockIdx.x, gridDim.x);
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
    mt = (int)(L % tiles_m);
    L /= tiles_m;
    nt = (int)(L % tiles_n);
    L /= tiles_n;
    b1 = L % g.nb1;
    b0 = L / g.nb1;
  }
  const int m0 = mt * FB_TILE, n0 = nt * TN;
  const int64_t bA = g.a_shared ? 0 : b0 * g.nb1 + b1, bB = g.b_shared ? 0 : b0 * g.nb1 + b1;
  const int64_t kstrideA = (int64_t)(g.Mp / 16) * FB_NP * FB_FRAG, kstrideB = (int64_t)(g.Np / 16) * FB_NP * FB_FRAG;
  const int8_t* gA = g.Aq + bA * g.nkb * kstrideA + (int64_t)(m0 / 16) * FB_NP * FB_FRAG;
  const int8_t* gB = g.Bq + bB * g.nkb * kstrideB + (int64_t)(n0 / 16) * FB_NP * FB_FRAG;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int wm = (w / WNW) * 16 * RI, wn = (w % WNW) * 16 * RJ;
  const int lr = lane & 15, lk = lane >> 4;

  // wave w copies fragments w, w + NW, .. of each operand's half of the stage
  auto copy_stage = [&](int64_t kb, int st) {
    const int8_t* srcA = gA + kb * kstrideA + 16 * lane;
    const int8_t* srcB = gB + kb * kstrideB + 16 * lane;
    int8_t* dst = lds + st * STAGE;
#pragma unroll
    for (int it = 0; it < STAGE_A / FB_FRAG / NW; ++it) {
      const int f = w + NW * it;
      __builtin_amdgcn_global_load_lds((const void*)(srcA + f * FB_FRAG),
                                       (__attribute__((address_space(3))) void*)(dst + f * FB_FRAG), 16, 0, 0);
    }
#pragma unroll
    for (int it = 0; it < STAGE_B / FB_FRAG / NW; ++it) {
      const int f = w + NW * it;
      __builtin_amdgcn_global_load_lds((const void*)(srcB + f * FB_FRAG),
                                       (__attribute__((address_space(3))) void*)(dst + STAGE_A + f * FB_FRAG), 16,
                                       0, 0);
    }
  };
  auto frag = [](const int8_t* p) { return *reinterpret_cast<const v4u*>(p); };

  v4f cr[RI][RJ], ci[RI][RJ];
#pragma unroll
  for (int i = 0; i < RI; ++i)
#pragma unroll
    for (int j = 0; j < RJ; ++j) { cr[i][j] = v4f{0, 0, 0, 0}; ci[i][j] = v4f{0, 0, 0, 0}; }
  const v4u sign = v4u{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};

  copy_stage(0, 0);
  for (int64_t kb = 0; kb < g.nkb; ++kb) {
    const int st = (int)(kb & 1);
    // stage st has landed for every wave, and every wave is done with the
    // stage the next copy overwrites
    __syncthreads();
    if (kb + 1 < g.nkb) copy_stage(kb + 1, st ^ 1);
    const int8_t* sA = lds + st * STAGE + (wm / 16) * FB_NP * FB_FRAG + 16 * lane;
    const int8_t* sB = lds + st * STAGE + STAGE_A + (wn / 16) * FB_NP * FB_FRAG + 16 * lane;
    v8bf brh[RJ], bih[RJ], brl[RJ], bil[RJ];
#pragma unroll
    for (int j = 0; j < RJ; ++j) {
      brh[j] = __builtin_bit_cast(v8bf, frag(sB + (j * FB_NP + 0) * FB_FRAG));
      bih[j] = __builtin_bit_cast(v8bf, frag(sB + (j * FB_NP + 1) * FB_FRAG));
      brl[j] = __builtin_bit_cast(v8bf, frag(sB + (j * FB_NP + 2) * FB_FRAG));
      bil[j] = __builtin_bit_cast(v8bf, frag(sB + (j * FB_NP + 3) * FB_FRAG));
    }
    v4u a0 = frag(sA + 0 * FB_FRAG), a1 = frag(sA + 1 * FB_FRAG), a2 = frag(sA + 2 * FB_FRAG),
        a3 = frag(sA + 3 * FB_FRAG);
#pragma unroll
    for (int i = 0; i < RI; ++i) {
      v4u n0_ = a0, n1_ = a1, n2_ = a2, n3_ = a3;
      if (i + 1 < RI) {
        n0_ = frag(sA + ((i + 1) * FB_NP + 0) * FB_FRAG);
        n1_ = frag(sA + ((i + 1) * FB_NP + 1) * FB_FRAG);
        n2_ = frag(sA + ((i + 1) * FB_NP + 2) * FB_FRAG);
        n3_ = frag(sA + ((i + 1) * FB_NP + 3) * FB_FRAG);
      }
      const v8bf arh = __builtin_bit_cast(v8bf, a0), aih = __builtin_bit_cast(v8bf, a1);
      const v8bf arl = __builtin_bit_cast(v8bf, a2), ail = __builtin_bit_cast(v8bf, a3);
      const v8bf naih = __builtin_bit_cast(v8bf, a1 ^ sign), nail = __builtin_bit_cast(v8bf, a3 ^ sign);
      // per accumulator the order of gemm_bf16x3_big_kernel
#define GGPU_FB_STEP(acc, a, b)                                                        \
  _Pragma("unroll") for (int j = 0; j < RJ; ++j) acc[i][j] =                           \
      __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
      GGPU_FB_STEP(cr, arh, brh) GGPU_FB_STEP(ci, arh, bih)
      GGPU_FB_STEP(cr, arh, brl) GGPU_FB_STEP(ci, arh, bil)
      GGPU_FB_STEP(cr, arl, brh) GGPU_FB_STEP(ci, arl, bih)
      GGPU_FB_STEP(cr, naih, bih) GGPU_FB_STEP(ci, aih, brh)
      GGPU_FB_STEP(cr, naih, bil) GGPU_FB_STEP(ci, aih, brl)
      GGPU_FB_STEP(cr, nail, bih) GGPU_FB_STEP(ci, ail, brh)
#undef GGPU_FB_STEP
      a0 = n0_; a1 = n1_; a2 = n2_; a3 = n3_;
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  const bool use_beta = (g.beta_re != 0.0 || g.beta_im != 0.0);
  TO* C = static_cast<TO*>(g.C) + b0 * g.c_bs0 + (g.c_off1 ? g.c_off1[b1] : b1 * g.c_bs1);
#pragma unroll
  for (int i = 0; i < RI; ++i)
#pragma unroll
    for (int j = 0; j < RJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + 4 * lk + r;
        const int col = n0 + wn + 16 * j + lr;
        if (row < g.M && col < g.N) {
          const double xr = cr[i][j][r], xi = ci[i][j][r];
          double vr = g.alpha_re * xr - g.alpha_im * xi;
          double vi = g.alpha_re * xi + g.alpha_im * xr;
          TO* p = C + row * g.c_rs + col * g.c_cs;
          if (use_beta) {
            const TO o = *p;
            vr += g.beta_re * o.x - g.beta_im * o.y;
            vi += g.beta_re * o.y + g.beta_im * o.x;
          }
          using TS = decltype(TO{}.x);
          *p = TO{static_cast<TS>(vr), static_cast<TS>(vi)};
        }
      }
}

inline int64_t fb_round(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Bytes of split planes one launch over nb0 batches needs (B padded to TN).
inline size_t fb_bytes(const GemmDesc& d, int nb0, bool a_shared, bool b_shared, int TN) {
  const int64_t kpb = (d.K + FB_KB - 1) / FB_KB, nkb = kpb * d.KB;
  const int64_t nbA = a_shared ? 1 : (int64_t)nb0 * d.nb1, nbB = b_shared ? 1 : (int64_t)nb0 * d.nb1;
  const int64_t per_row_block = nkb * FB_NP * FB_FRAG;
  return (size_t)(nbA * (fb_round(d.M, FB_TILE) / 16) + nbB * (fb_round(d.N, TN) / 16)) * per_row_block;
}

// Whether the pre-split tile can run this launch with the workspace it has
// (one batch of its planes must fit).
inline bool fb_fits(const GemmDesc& d, int TN) {
  const bool a_shared = d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr;
  const bool b_shared = d.B.bs0 == 0 && d.B.bs1 == 0 && d.B.off1 == nullptr;
  return g_workspace != nullptr && fb_bytes(d, 1, a_shared, b_shared, TN) <= g_workspace_bytes;
}

template <class TO, int WMW, int WNW, int TN = 128>
void launch_bf16x3_frag(const GemmDesc& d, hipStream_t stream) {
  const bool a_shared = d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr;
  const bool b_shared = d.B.bs0 == 0 && d.B.bs1 == 0 && d.B.off1 == nullptr;
  const int kpb = (d.K + FB_KB - 1) / FB_KB;
  const int64_t nkb = (int64_t)kpb * d.KB;
  const int Mp = (int)fb_round(d.M, FB_TILE), Np = (int)fb_round(d.N, TN);
  int nb0c = d.nb0;   // batches per part: their planes fit the workspace
  while (nb0c > 1 && fb_bytes(d, nb0c, a_shared, b_shared, TN) > g_workspace_bytes) nb0c = (nb0c + 1) / 2;
  require(fb_bytes(d, nb0c, a_shared, b_shared, TN) <= g_workspace_bytes, "gemm: split-bf16 workspace too small");
  require(Mp / 16 <= 65535 && Np / 16 <= 65535, "gemm: split-bf16 operand too large");
  const size_t elem = d.out_f32 ? sizeof(float2) : sizeof(double2);
  for (int b0s = 0; b0s < d.nb0; b0s += nb0c) {
    const int nb = std::min(nb0c, d.nb0 - b0s);
    const int64_t nbA = a_shared ? 1 : (int64_t)nb * d.nb1, nbB = b_shared ? 1 : (int64_t)nb * d.nb1;
    int8_t* Aq = static_cast<int8_t*>(g_workspace);
    int8_t* Bq = Aq + (size_t)nbA * (Mp / 16) * nkb * FB_NP * FB_FRAG;
    for (int op = 0; op < 2; ++op) {
      const Operand& o = op == 0 ? d.A : d.B;
      FbView v;
      v.src = static_cast<const float2*>(o.ptr) + (int64_t)b0s * o.bs0;
      // A(m, k) = A[m rs + k cs]; B(k, n) = B[k rs + n cs]: rows are m / n
      v.row_s = op == 0 ? o.rs : o.cs;
      v.k_s = op == 0 ? o.cs : o.rs;
      v.kb_s = o.kbs; v.bs0 = o.bs0; v.bs1 = o.bs1; v.off1 = o.off1;
      v.rows = op == 0 ? d.M : d.N; v.K = d.K; v.KB = d.KB;
      v.nb1 = (op == 0 ? a_shared : b_shared) ? 1 : d.nb1;
      v.conj = o.conj; v.Rp = op == 0 ? Mp : Np; v.kpb = kpb;
      v.out = op == 0 ? Aq : Bq;
      const int64_t nbat = op == 0 ? nbA : nbB;
      require(nbat <= 65535 && (nkb + 3) / 4 < ((int64_t)1 << 31), "gemm: split-bf16 operand too large");
      hipLaunchKernelGGL(fb_split_kernel, dim3((unsigned)((nkb + 3) / 4), (unsigned)(v.Rp / 16), (unsigned)nbat),
                         dim3(256), 0, stream, v, nkb);
      GGPU_HIP(hipGetLastError());
    }
    FbGemm g;
    g.Aq = Aq; g.Bq = Bq; g.M = d.M; g.N = d.N; g.Mp = Mp; g.Np = Np; g.nkb = nkb;
    g.nb0 = nb; g.nb1 = d.nb1; g.a_shared = a_shared; g.b_shared = b_shared;
    g.batch_major = (nb * (int64_t)d.nb1 > 1) && a_shared;
    g.C = static_cast<char*>(d.C.ptr) + (int64_t)b0s * d.C.bs0 * (int64_t)elem;
    g.c_rs = d.C.rs; g.c_cs = d.C.cs; g.c_bs0 = d.C.bs0; g.c_bs1 = d.C.bs1; g.c_off1 = d.C.off1;
    g.alpha_re = d.alpha_re; g.alpha_im = d.alpha_im; g.beta_re = d.beta_re; g.beta_im = d.beta_im;
    const int64_t tiles = (int64_t)(Mp / FB_TILE) * (Np / TN) * nb * d.nb1;
    require(tiles < ((int64_t)1 << 31), "gemm: grid too large");
    hipLaunchKernelGGL((gemm_bf16x3_frag_kernel<TO, WMW, WNW, TN>), dim3((unsigned)tiles), dim3(64 * WMW * WNW), 0,
                       stream, g);
    GGPU_HIP(hipGetLastError());
  }
}

template <class TO, int SITE>
void launch_bf16x3_astat(const GemmDesc& d, hipStream_t stream) {
  constexpr int TM = 64, TN = 64;
  const int64_t tiles = (int64_t)((d.M + TM - 1) / TM) * ((d.N + TN - 1) / TN);
  const int64_t nbatch = (int64_t)d.nb0 * d.nb1;
  // batches per workgroup: as many as keep about 2048 workgroups in the grid
  const int bpw = (int)std::max<int64_t>(1, std::min<int64_t>(16, tiles * nbatch / 2048));
  const int64_t total = tiles * ((nbatch + bpw - 1) / bpw);
  require(total < (int64_t)1 << 31, "gemm: grid too large");
  dim3 grid((unsigned)total), block(NTHREADS);
  const bool akc = (d.A.cs == 1) || (d.A.rs != 1);
  const bool bkc = (d.B.rs == 1) || (d.B.cs != 1);
  if (akc && bkc)
    hipLaunchKernelGGL((gemm_bf16x3_astat_kernel<TO, true, true, SITE>), grid, block, 0, stream, d, bpw);
  else if (akc)
    hipLaunchKernelGGL((gemm_bf16x3_astat_kernel<TO, true, false, SITE>), grid, block, 0, stream, d, bpw);
  else if (bkc)
    hipLaunchKernelGGL((gemm_bf16x3_astat_kernel<TO, false, true, SITE>), grid, block, 0, stream, d, bpw);
  else
    hipLaunchKernelGGL((gemm_bf16x3_astat_kernel<TO, false, false, SITE>), grid, block, 0, stream, d, bpw);
  GGPU_HIP(hipGetLastError());
}

template <class TO, int WMB, int WNB, int SITE, int WPE = 1>
void launch_bf16x3_big(const GemmDesc& d, hipStream_t stream) {
  constexpr int TM = 32 * WMB, TN = 32 * WNB;
  const int64_t tiles = (int64_t)((d.M + TM - 1) / TM) * ((d.N + TN - 1) / TN);
  const int64_t total = tiles * d.nb0 * d.nb1;
  require(total < (int64_t)1 << 31, "gemm: grid too large");
  dim3 grid((unsigned)total), block(NTHREADS);
  const bool akc = (d.A.cs == 1) || (d.A.rs != 1);
  const bool bkc = (d.B.rs == 1) || (d.B.cs != 1);
  if (akc && bkc)
    hipLaunchKernelGGL((gemm_bf16x3_big_kernel<TO, true, true, WMB, WNB, SITE, WPE>), grid, block, 0, stream, d);
  else if (akc)
    hipLaunchKernelGGL((gemm_bf16x3_big_kernel<TO, true, false, WMB, WNB, SITE, WPE>), grid, block, 0, stream, d);
  else if (bkc)
    hipLaunchKernelGGL((gemm_bf16x3_big_kernel<TO, false, true, WMB, WNB, SITE, WPE>), grid, block, 0, stream, d);
  else
    hipLaunchKernelGGL((gemm_bf16x3_big_kernel<TO, false, false, WMB, WNB, SITE, WPE>), grid, block, 0, stream, d);
  GGPU_HIP(hipGetLastError());
}

// Fraction of a tiled extent that is real work.
inline double tile_use(int n, int t) { return (double)n / (double)(((n + t - 1) / t) * t); }

template <class TO, int SITE>
void launch_bf16x3_site(const GemmDesc& d, hipStream_t stream, int variant) {
  int mode = variant;
  if (mode < 0) {
    // Two or three waves per SIMD (6- to 10-block wave tiles) beat bigger
    // tiles.  Large shapes take the tile measured fastest for their call site
    // inside a GW iteration on MI355X (nao = 52, NQ = 300, nts = 110; every
    // tile forced in turn, per-launch times from the site profiler) unless
    // its tile waste is more than 10 % above the best fit of the two-wave
    // tiles; other shapes take that best fit.  (Repeated launches of one GEMM
    // on the same operands rank the tiles differently: warm caches.)
    mode = kBf16Tile64x64;
    if (d.M >= 64 && d.N >= 48 && (d.M >= 128 || d.N >= 128)) {
      const int cand[5] = {kBf16Tile128x64, kBf16Tile64x128, kBf16Tile160x64, kBf16Tile96x64W3, kBf16Tile64x64W3};
      const double use[5] = {tile_use(d.M, 128) * tile_use(d.N, 64), tile_use(d.M, 64) * tile_use(d.N, 128),
                             tile_use(d.M, 160) * tile_use(d.N, 64), tile_use(d.M, 96) * tile_use(d.N, 64),
                             tile_use(d.M, 64) * tile_use(d.N, 64)};
      int best = 0;
      for (int c = 1; c < 3; ++c)
        if (use[c] > use[best] + 0.03) best = c;
      int pref = -1;
      switch (SITE) {
        case kSiteSigY1: case kSiteSigma: pref = 0; break;   // 128 x 64
        case kSiteSigY2: pref = 1; break;                    // 64 x 128
        case kSiteP0X2: case kSiteP0: pref = 2; break;       // 160 x 64
        case kSiteP0X1: pref = 3; break;                     // 96 x 64, three waves
        default: break;
      }
      mode = cand[(pref >= 0 && use[pref] >= 0.9 * use[best]) ? pref : best];
    }
  }
  switch (mode) {
    case kBf16Tile128x64: return launch_bf16x3_big<TO, 4, 2, SITE>(d, stream);
    case kBf16Tile64x128: return launch_bf16x3_big<TO, 2, 4, SITE>(d, stream);
    case kBf16Tile160x64: return launch_bf16x3_big<TO, 5, 2, SITE>(d, stream);
    case kBf16Tile96x64W3: return launch_bf16x3_big<TO, 3, 2, SITE, 3>(d, stream);
    case kBf16Tile64x96W3: return launch_bf16x3_big<TO, 2, 3, SITE, 3>(d, stream);
    case kBf16Tile64x64W3: return launch_bf16x3_big<TO, 2, 2, SITE, 3>(d, stream);
    case kBf16TileAstat64x64:
      if (astat_fits(d)) return launch_bf16x3_astat<TO, SITE>(d, stream);
      return launch_bf16x3_big<TO, 2, 2, SITE, 3>(d, stream);   // a 64 x 64 tile, A staged per batch
    case kBf16TileFrag128x128:
      if (fb_fits(d, 128)) return launch_bf16x3_frag<TO, 2, 2>(d, stream);
      return launch_bf16x3_big<TO, 5, 2, SITE>(d, stream);   // no workspace: 160 x 64
    case kBf16TileFrag128x128W8:
      if (fb_fits(d, 128)) return launch_bf16x3_frag<TO, 4, 2>(d, stream);
      return launch_bf16x3_big<TO, 5, 2, SITE>(d, stream);
    case kBf16TileFrag128x128W16:
      if (fb_fits(d, 128)) return launch_bf16x3_frag<TO, 4, 4>(d, stream);
      return launch_bf16x3_big<TO, 5, 2, SITE>(d, stream);
    case kBf16TileFrag128x192W8:
      if (fb_fits(d, 192)) return launch_bf16x3_frag<TO, 2, 4, 192>(d, stream);
      return launch_bf16x3_big<TO, 5, 2, SITE>(d, stream);
    default: break;
  }
  const int64_t tiles = (int64_t)((d.M + BM - 1) / BM) * ((d.N + BN - 1) / BN);
  const int64_t total = tiles * d.nb0 * d.nb1;
  require(total < (int64_t)1 << 31, "gemm: grid too large");
  dim3 grid((unsigned)total), block(NTHREADS);
  const bool akc = (d.A.cs == 1) || (d.A.rs != 1);
  const bool bkc = (d.B.rs == 1) || (d.B.cs != 1);
  if (akc && bkc) hipLaunchKernelGGL((gemm_bf16x3_kernel<TO, true, true, SITE>), grid, block, 0, stream, d);
  else if (akc) hipLaunchKernelGGL((gemm_bf16x3_kernel<TO, true, false, SITE>), grid, block, 0, stream, d);
  else if (bkc) hipLaunchKernelGGL((gemm_bf16x3_kernel<TO, false, true, SITE>), grid, block, 0, stream, d);
  else hipLaunchKernelGGL((gemm_bf16x3_kernel<TO, false, false, SITE>), grid, block, 0, stream, d);
  GGPU_HIP(hipGetLastError());
}

template <class TO>
void launch_bf16x3(const GemmDesc& d, hipStream_t stream, int v) {
  switch (d.site) {
    case kSiteP0X1: launch_bf16x3_site<TO, kSiteP0X1>(d, stream, v); break;
    case kSiteP0X2: launch_bf16x3_site<TO, kSiteP0X2>(d, stream, v); break;
    case kSiteP0: launch_bf16x3_site<TO, kSiteP0>(d, stream, v); break;
    case kSiteSigY1: launch_bf16x3_site<TO, kSiteSigY1>(d, stream, v); break;
    case kSiteSigY2: launch_bf16x3_site<TO, kSiteSigY2>(d, stream, v); break;
    case kSiteSigma: launch_bf16x3_site<TO, kSiteSigma>(d, stream, v); break;
    default: launch_bf16x3_site<TO, kSiteOther>(d, stream, v); break;
  }
}

template <class T, class TO, bool GAUSS, int SITE, int BK, int WPE>
void launch_site_bk(const GemmDesc& d, hipStream_t stream) {
  const int64_t tiles = (int64_t)((d.M + BM - 1) / BM) * ((d.N + BN - 1) / BN);
  const int64_t total = tiles * d.nb0 * d.nb1;
  require(total < (int64_t)1 << 31, "gemm: grid too large");
  dim3 grid((unsigned)total), block(NTHREADS);
  const bool akc = (d.A.cs == 1) || (d.A.rs != 1);
  const bool bkc = (d.B.rs == 1) || (d.B.cs != 1);
  if (akc && bkc) hipLaunchKernelGGL((gemm_mfma_kernel<T, TO, true, true, GAUSS, SITE, BK, WPE>), grid, block, 0, stream, d);
  else if (akc) hipLaunchKernelGGL((gemm_mfma_kernel<T, TO, true, false, GAUSS, SITE, BK, WPE>), grid, block, 0, stream, d);
  else if (bkc) hipLaunchKernelGGL((gemm_mfma_kernel<T, TO, false, true, GAUSS, SITE, BK, WPE>), grid, block, 0, stream, d);
  else hipLaunchKernelGGL((gemm_mfma_kernel<T, TO, false, false, GAUSS, SITE, BK, WPE>), grid, block, 0, stream, d);
  GGPU_HIP(hipGetLastError());
}

template <class T, class TO, bool GAUSS, int SITE>
void launch_site(const GemmDesc& d, hipStream_t stream, int variant) {
  // Variants measured on MI355X (nao = 52, NQ = 300, nts = 110; medians of
  // interleaved rounds): staging 8 k per LDS step is faster than 16 or equal
  // at every call site, and FP64 register allocation for three waves per
  // SIMD (152-156 VGPRs, accumulators out of the AGPRs, no spills) instead
  // of two gains 7-17 % at every site.  FP32 runs five waves unconstrained.
  // All variants give the same sums bit for bit.
  constexpr bool f64 = std::is_same<T, double>::value;
  constexpr int pref = f64 ? 83 : 8;
  switch (variant > 0 ? variant : pref) {
    case 8: return launch_site_bk<T, TO, GAUSS, SITE, 8, 1>(d, stream);
    case 83: return launch_site_bk<T, TO, GAUSS, SITE, 8, 3>(d, stream);
    case 84: return launch_site_bk<T, TO, GAUSS, SITE, 8, 4>(d, stream);
    default: return launch_site_bk<T, TO, GAUSS, SITE, 16, 1>(d, stream);
  }
}

// FP64 uses the 3-product (Gauss) complex scheme; FP32 keeps 4 products
// (the Gauss recombination t3 - t1 - t2 loses relative accuracy when the
// result is small against |A||B|, which FP32 cannot afford).
template <class T, class TO>
void launch_typed(const GemmDesc& d, hipStream_t stream, int v) {
  constexpr bool G = std::is_same<T, double>::value;
  switch (d.site) {
    case kSiteP0X1: launch_site<T, TO, G, kSiteP0X1>(d, stream, v); break;
    case kSiteP0X2: launch_site<T, TO, G, kSiteP0X2>(d, stream, v); break;
    case kSiteP0: launch_site<T, TO, G, kSiteP0>(d, stream, v); break;
    case kSiteSigY1: launch_site<T, TO, G, kSiteSigY1>(d, stream, v); break;
    case kSiteSigY2: launch_site<T, TO, G, kSiteSigY2>(d, stream, v); break;
    case kSiteSigma: launch_site<T, TO, G, kSiteSigma>(d, stream, v); break;
    default: launch_site<T, TO, G, kSiteOther>(d, stream, v); break;
  }
}


// Range of element offsets (relative to ptr) touched by a 2-D view with the
// given extents and strides, over all batch indices.
struct Range { int64_t lo = 0, hi = 0; };
inline void extend(Range& r, int64_t n, int64_t stride) {
  if (n <= 1 || stride == 0) return;
  const int64_t span = (n - 1) * stride;
  if (span > 0) r.hi += span; else r.lo += span;
}

void check_operand(const char* name, const void* ptr, const void* buf, int64_t buf_elems, size_t esize,
                   const int64_t* off_host, int nb1, Range r) {
  if (!buf) return;
  const int64_t base = (int64_t)(((const char*)ptr - (const char*)buf) / (int64_t)esize);
  int64_t lo = base + r.lo, hi = base + r.hi;
  if (off_host) {
    int64_t omin = off_host[0], omax = off_host[0];
    for (int i = 1; i < nb1; ++i) { omin = std::min(omin, off_host[i]); omax = std::max(omax, off_host[i]); }
    lo += omin; hi += omax;
  }
  if (lo < 0 || hi >= buf_elems) {
    char msg[256];
    snprintf(msg, sizeof(msg), "gemm operand %s addresses [%lld, %lld] outside its allocation of %lld elements",
             name, (long long)lo, (long long)hi, (long long)buf_elems);
    throw Error(kErrArgument, msg);
  }
}

}  // namespace

void gemm_check_bounds(const GemmDesc& d) {
  require(d.M >= 0 && d.N >= 0 && d.K >= 0 && d.KB >= 1 && d.nb0 >= 1 && d.nb1 >= 1,
          "gemm: bad dimensions");
  require(d.nb0 <= 65535 && d.nb1 <= 65535, "gemm: batch count above 65535");
  if (d.M == 0 || d.N == 0) return;
  Range ra, rb, rc;
  extend(ra, d.M, d.A.rs); extend(ra, d.K, d.A.cs); extend(ra, d.KB, d.A.kbs);
  extend(ra, d.nb0, d.A.bs0); if (!d.A.off1) extend(ra, d.nb1, d.A.bs1);
  extend(rb, d.K, d.B.rs); extend(rb, d.N, d.B.cs); extend(rb, d.KB, d.B.kbs);
  extend(rb, d.nb0, d.B.bs0); if (!d.B.off1) extend(rb, d.nb1, d.B.bs1);
  extend(rc, d.M, d.C.rs); extend(rc, d.N, d.C.cs);
  extend(rc, d.nb0, d.C.bs0); if (!d.C.off1) extend(rc, d.nb1, d.C.bs1);
  const size_t ein = complex64_operands(d.compute) ? sizeof(float2) : sizeof(double2);
  const size_t eout = d.out_f32 ? sizeof(float2) : sizeof(double2);
  require(!d.A.off1 || d.A.off1_host || !d.A.buf, "gemm: A offset table without host copy");
  require(!d.B.off1 || d.B.off1_host || !d.B.buf, "gemm: B offset table without host copy");
  require(!d.C.off1 || d.C.off1_host || !d.C.buf, "gemm: C offset table without host copy");
  if (d.K > 0) {
    check_operand("A", d.A.ptr, d.A.buf, d.A.buf_elems, ein, d.A.off1_host, d.nb1, ra);
    check_operand("B", d.B.ptr, d.B.buf, d.B.buf_elems, ein, d.B.off1_host, d.nb1, rb);
  }
  check_operand("C", d.C.ptr, d.C.buf, d.C.buf_elems, eout, d.C.off1_host, d.nb1, rc);
}

void set_gemm_profiler(GemmProfiler* p) { g_profiler = p; }
void set_site_variants(const int* table) { g_site_variants = table; }
void set_gemm_workspace(void* ptr, size_t bytes) {
  g_workspace = ptr;
  g_workspace_bytes = ptr ? bytes : 0;
}

int gemm_variant_candidates(Compute c, int* out, int cap) {
  static const int f64[] = {83, 8, 16};
  static const int f32[] = {8, 16, 83};
  static const int bf[] = {kBf16Tile128x64, kBf16Tile64x128, kBf16Tile160x64, kBf16Tile96x64W3,
                           kBf16Tile64x96W3, kBf16Tile64x64W3, kBf16Tile64x64, kBf16TileAstat64x64,
                           kBf16TileFrag128x128, kBf16TileFrag128x128W8, kBf16TileFrag128x128W16,
                           kBf16TileFrag128x192W8};
  if (int8_slices(c)) {                  // int8 emulation or the FP64 kernel
    // four-slice plans may also take five slices where four are over budget
    static const int i8x5[] = {kVariantInt8, 83}, i8x4[] = {kVariantInt8, kVariantInt8x5, 83};
    const int* v = c == Compute::I8X4 ? i8x4 : i8x5;
    const int n = c == Compute::I8X4 ? 3 : 2;
    for (int i = 0; i < n && i < cap; ++i) out[i] = v[i];
    return n;
  }
  const int* v = c == Compute::F64 ? f64 : c == Compute::F32 ? f32 : bf;
  const int n = c == Compute::F64 ? 3 : c == Compute::F32 ? 3 : 12;
  for (int i = 0; i < n && i < cap; ++i) out[i] = v[i];
  return n;
}

bool gemm_variant_valid(Compute c, int v) {
  if (v == -1) return true;
  int cand[16];
  const int n = gemm_variant_candidates(c, cand, 16);
  return std::find(cand, cand + n, v) != cand + n;
}

void gemm(const GemmDesc& dd, hipStream_t stream) {
  gemm_check_bounds(dd);
  if (dd.M == 0 || dd.N == 0) return;
  GemmDesc d = dd;
  d.batch_major = (d.nb0 * (int64_t)d.nb1 > 1) && d.A.bs0 == 0 && d.A.bs1 == 0 && d.A.off1 == nullptr;
  const int v = (g_site_variants && d.site >= 0 && d.site < kNumSites) ? g_site_variants[d.site] : -1;
  if (g_profiler) g_profiler->before(d, stream);
  switch (d.compute) {
    case Compute::F64:
      require(!d.out_f32, "gemm: FP64 compute writes complex128");
      launch_typed<double, double2>(d, stream, v);
      break;
    case Compute::F32:
      if (d.out_f32) launch_typed<float, float2>(d, stream, v);
      else launch_typed<float, double2>(d, stream, v);
      break;
    case Compute::BF16X3:
      if (d.out_f32) launch_bf16x3<float2>(d, stream, v);
      else launch_bf16x3<double2>(d, stream, v);
      break;
    case Compute::I8X5:
    case Compute::I8X4: {
      const bool emulate =
          v == kVariantInt8 || v == kVariantInt8x5 || (v == -1 && (d.site == kSiteSigY2 || d.site == kSiteP0));
      if (emulate) {
        gemm_int8_emulated(d, v == kVariantInt8x5 ? 5 : int8_slices(d.compute), g_workspace, g_workspace_bytes,
                           stream);
      } else {
        require(!d.out_f32, "gemm: FP64 kernel writes complex128");
        GemmDesc f = d;
        f.compute = Compute::F64;
        launch_typed<double, double2>(f, stream, v);
      }
      break;
    }
    default:
      throw Error(kErrArgument, "gemm: compute mode not available");
  }
  if (g_profiler) g_profiler->after(d, stream);
}

}  // namespace ggpu
