// Bandwidth-bound and small-matrix kernels of the HF/GW path (gfx950).
#include "kernels.hpp"

#include <algorithm>
#include <cmath>

namespace ggpu {
namespace {

constexpr int TILE = 32;

__device__ inline z64 cj(z64 a) { return z64{a.x, -a.y}; }
__device__ inline z64 cmul(z64 a, z64 b) { return z64{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline z64 csub(z64 a, z64 b) { return z64{a.x - b.x, a.y - b.y}; }

template <class T> __device__ inline c32 narrow(T v) { return c32{(float)v.x, (float)v.y}; }
template <class T> __device__ inline T cjt(T a) { return T{a.x, -a.y}; }
template <class T> __device__ inline T scl(T a, double f) {
  using R = decltype(a.x);
  return T{(R)(f * a.x), (R)(f * a.y)};
}

// 32 x 32 tiles through LDS; block = 32 x 8 threads.
template <class TI, class TO>
__global__ __launch_bounds__(256) void block_copy_kernel(const BlockCopy a) {
  __shared__ TO tile[TILE][TILE + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int c0 = blockIdx.x * TILE, r0 = blockIdx.y * TILE;
  const int outer = blockIdx.z / a.n_inner, inner = blockIdx.z - outer * a.n_inner;
  const TI* in = static_cast<const TI*>(a.in) + (a.in_off ? a.in_off[outer] : 0) + inner * a.in_bs;
  TO* out = static_cast<TO*>(a.out) + (a.out_off ? a.out_off[outer] : 0) + inner * a.out_bs;
  const bool tr = a.trans ? a.trans[outer] != 0 : a.trans_all != 0;
  const bool cg = a.conj ? a.conj[outer] != 0 : a.conj_all != 0;
  if (tr) {
    // in is cols x rows: element (r, c) of the output is in[c][r].
#pragma unroll
    for (int j = 0; j < TILE; j += 8) {
      const int c = c0 + ty + j, r = r0 + tx;
      if (c < a.cols && r < a.rows) {
        if constexpr (sizeof(TI) == sizeof(TO)) tile[ty + j][tx] = in[c * a.in_ld + r];
        else tile[ty + j][tx] = narrow(in[c * a.in_ld + r]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < TILE; j += 8) {
      const int r = r0 + ty + j, c = c0 + tx;
      if (r < a.rows && c < a.cols) {
        TO v = tile[tx][ty + j];
        if (a.scale != 1.0) v = scl(v, a.scale);
        out[r * a.out_ld + c] = cg ? cjt(v) : v;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < TILE; j += 8) {
      const int r = r0 + ty + j, c = c0 + tx;
      if (r < a.rows && c < a.cols) {
        TO v;
        if constexpr (sizeof(TI) == sizeof(TO)) v = in[r * a.in_ld + c];
        else v = narrow(in[r * a.in_ld + c]);
        if (a.scale != 1.0) v = scl(v, a.scale);
        out[r * a.out_ld + c] = cg ? cjt(v) : v;
      }
    }
  }
}

__global__ void to_c32_kernel(const z64* in, c32* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = c32{(float)in[i].x, (float)in[i].y};
}

__global__ void slab_gather_kernel(const z64* in, const int64_t* src_off, z64* out,
                                   const int64_t* dst_off, int n_slabs, int64_t len, double scale) {
  const int64_t total = (int64_t)n_slabs * len;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = e / len, i = e - o * len;
    const z64 v = in[src_off[o] + i];
    out[dst_off[o] + i] = z64{scale * v.x, scale * v.y};
  }
}

__global__ __launch_bounds__(256) void herm_one_minus_kernel(z64* M, z64* A, int n) {
  __shared__ z64 t1[TILE][TILE + 1];
  __shared__ z64 t2[TILE][TILE + 1];
  // lower-triangular tile pair index -> (bi, bj), bi >= bj
  const int x = blockIdx.x;
  int bi = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= x) ++bi;
  while (bi * (bi + 1) / 2 > x) --bi;
  const int bj = x - bi * (bi + 1) / 2;
  const int64_t w = blockIdx.y;
  z64* Mw = M + w * (int64_t)n * n;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int ri = bi * TILE, rj = bj * TILE;
#pragma unroll
  for (int j = 0; j < TILE; j += 8) {
    const int r = ty + j;
    if (ri + r < n && rj + tx < n) t1[r][tx] = Mw[(int64_t)(ri + r) * n + rj + tx];   // tile (bi, bj)
    if (rj + r < n && ri + tx < n) t2[r][tx] = Mw[(int64_t)(rj + r) * n + ri + tx];   // tile (bj, bi)
  }
  __syncthreads();
  z64* Aw = A ? A + w * (int64_t)n * n : nullptr;
#pragma unroll
  for (int j = 0; j < TILE; j += 8) {
    const int r = ty + j;
    // element (ri + r, rj + tx) of H = (M + M^dagger)/2
    if (ri + r < n && rj + tx < n) {
      const z64 a = t1[r][tx], b = t2[tx][r];
      z64 h = z64{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
      const int64_t idx = (int64_t)(ri + r) * n + rj + tx;
      Mw[idx] = h;
      if (Aw) Aw[idx] = z64{(ri + r == rj + tx ? 1.0 : 0.0) - h.x, -h.y};
    }
    // element (rj + r, ri + tx)
    if (bi != bj && rj + r < n && ri + tx < n) {
      const z64 a = t2[r][tx], b = t1[tx][r];
      z64 h = z64{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
      const int64_t idx = (int64_t)(rj + r) * n + ri + tx;
      Mw[idx] = h;
      if (Aw) Aw[idx] = z64{-h.x, -h.y};
    }
  }
}

__global__ void zero_kernel(z64* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z64{0.0, 0.0};
}
__global__ void zero_int_kernel(int* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
__global__ void sum_partials_kernel(const z64* in, int64_t stride, int np, z64* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int p = 0; p < np; ++p) { const z64 v = in[p * stride + i]; re += v.x; im += v.y; }
    out[i] = z64{re, im};
  }
}

constexpr int NBMAX = 32;

// One workgroup per matrix; the nb x nb diagonal block lives in LDS.
__global__ __launch_bounds__(256) void chol_diag_kernel(z64* A, int n, int j0, int nb, int* info) {
  __shared__ z64 L[NBMAX][NBMAX + 1];
  const int64_t w = blockIdx.x;
  z64* Aw = A + w * (int64_t)n * n;
  const int tid = threadIdx.x;
  for (int e = tid; e < nb * nb; e += blockDim.x) {
    const int r = e / nb, c = e - r * nb;
    L[r][c] = (c <= r) ? Aw[(int64_t)(j0 + r) * n + j0 + c] : z64{0.0, 0.0};
  }
  __syncthreads();
  for (int c = 0; c < nb; ++c) {
    if (tid == 0) {
      const double d = L[c][c].x;
      if (!(d > 0.0)) {                    // not positive definite (or NaN)
        if (info[w] == 0) info[w] = j0 + c + 1;
        L[c][c] = z64{1.0, 0.0};
      } else {
        L[c][c] = z64{sqrt(d), 0.0};
      }
    }
    __syncthreads();
    const double inv = 1.0 / L[c][c].x;
    for (int r = c + 1 + tid; r < nb; r += blockDim.x) L[r][c] = z64{L[r][c].x * inv, L[r][c].y * inv};
    __syncthreads();
    const int m = nb - c - 1;
    for (int e = tid; e < m * m; e += blockDim.x) {
      const int r = c + 1 + e / m, s = c + 1 + e % m;
      if (s <= r) L[r][s] = csub(L[r][s], cmul(L[r][c], cj(L[s][c])));
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += blockDim.x) {
    const int r = e / nb, c = e - r * nb;
    if (c <= r) Aw[(int64_t)(j0 + r) * n + j0 + c] = L[r][c];
  }
}

// Rows below the diagonal block: x L11^dagger = a.  64 rows per workgroup.
__global__ __launch_bounds__(64) void chol_panel_kernel(z64* A, int n, int j0, int nb) {
  __shared__ z64 L[NBMAX][NBMAX + 1];
  __shared__ z64 X[64][NBMAX + 1];
  const int64_t w = blockIdx.y;
  z64* Aw = A + w * (int64_t)n * n;
  const int tid = threadIdx.x;
  const int r0 = j0 + nb + blockIdx.x * 64;
  for (int e = tid; e < nb * nb; e += 64) {
    const int r = e / nb, c = e - r * nb;
    L[r][c] = Aw[(int64_t)(j0 + r) * n + j0 + c];
  }
  for (int e = tid; e < 64 * nb; e += 64) {
    const int r = e / nb, c = e - r * nb;
    if (r0 + r < n) X[r][c] = Aw[(int64_t)(r0 + r) * n + j0 + c];
  }
  __syncthreads();
  if (r0 + tid < n) {
    for (int c = 0; c < nb; ++c) {
      z64 acc = X[tid][c];
      for (int s = 0; s < c; ++s) acc = csub(acc, cmul(X[tid][s], cj(L[c][s])));
      const double inv = 1.0 / L[c][c].x;
      X[tid][c] = z64{acc.x * inv, acc.y * inv};
    }
  }
  __syncthreads();
  for (int e = tid; e < 64 * nb; e += 64) {
    const int r = e / nb, c = e - r * nb;
    if (r0 + r < n) Aw[(int64_t)(r0 + r) * n + j0 + c] = X[r][c];
  }
}

// Triangular solve with the nb x nb diagonal block T at (i0, i0) of the
// matrix A (row stride lda) on rows i0 .. i0+nb of B, columns bcol0 ..
// bcol0+ncols; 64 right-hand-side columns per workgroup (one per thread).
// KIND: kTriLower  T lower, non-unit (Cholesky L, forward)
//       kTriUpperH T^dagger upper from the lower L (Cholesky, backward)
//       kTriUnitLower  T unit lower (LU L, forward)
//       kTriUpper  T upper, non-unit (LU U, backward)
enum { kTriLower = 0, kTriUpperH = 1, kTriUnitLower = 2, kTriUpper = 3 };

__device__ inline z64 cdiv(z64 a, z64 d) {
  const double m = 1.0 / (d.x * d.x + d.y * d.y);
  return z64{(a.x * d.x + a.y * d.y) * m, (a.y * d.x - a.x * d.y) * m};
}

template <int KIND>
__global__ __launch_bounds__(64) void trsm_block_kernel(const z64* A, int lda, int64_t a_ms, z64* B, int ldb,
                                                        int64_t b_ms, int bcol0, int ncols, int i0, int nb) {
  __shared__ z64 T[NBMAX][NBMAX + 1];
  __shared__ z64 Y[NBMAX][64 + 1];
  const int64_t w = blockIdx.y;
  const z64* Aw = A + w * a_ms;
  z64* Bw = B + w * b_ms + bcol0;
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * 64;
  for (int e = tid; e < nb * nb; e += 64) {
    const int r = e / nb, c = e - r * nb;
    T[r][c] = Aw[(int64_t)(i0 + r) * lda + i0 + c];
  }
  for (int e = tid; e < nb * 64; e += 64) {
    const int r = e / 64, c = e - r * 64;
    if (c0 + c < ncols) Y[r][c] = Bw[(int64_t)(i0 + r) * ldb + c0 + c];
  }
  __syncthreads();
  if (c0 + tid < ncols) {
    if (KIND == kTriLower || KIND == kTriUnitLower) {
      for (int r = 0; r < nb; ++r) {
        z64 acc = Y[r][tid];
        for (int s = 0; s < r; ++s) acc = csub(acc, cmul(T[r][s], Y[s][tid]));
        if (KIND == kTriLower) {
          const double inv = 1.0 / T[r][r].x;
          acc = z64{acc.x * inv, acc.y * inv};
        }
        Y[r][tid] = acc;
      }
    } else {
      for (int r = nb - 1; r >= 0; --r) {
        z64 acc = Y[r][tid];
        if (KIND == kTriUpperH) {
          for (int s = r + 1; s < nb; ++s) acc = csub(acc, cmul(cj(T[s][r]), Y[s][tid]));
          const double inv = 1.0 / T[r][r].x;
          Y[r][tid] = z64{acc.x * inv, acc.y * inv};
        } else {
          for (int s = r + 1; s < nb; ++s) acc = csub(acc, cmul(T[r][s], Y[s][tid]));
          Y[r][tid] = cdiv(acc, T[r][r]);
        }
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < nb * 64; e += 64) {
    const int r = e / 64, c = e - r * 64;
    if (c0 + c < ncols) Bw[(int64_t)(i0 + r) * ldb + c0 + c] = Y[r][c];
  }
}

// LU with partial pivoting of the panel of columns j0 .. j0+nb, rows j0 .. n
// (LAPACK zgetf2 semantics on the panel): one workgroup per matrix, the
// panel in global memory.  Pivot = first row of largest |re| + |im| (izamax);
// the row swaps are applied inside the panel only (lu_row_swaps applies them
// to the other columns).  ipiv[w * n + c] = pivot row of column c (0-based);
// info[w] = column + 1 of the first exactly zero pivot (0 = none).
constexpr int LU_THREADS = 256;

__global__ __launch_bounds__(LU_THREADS) void lu_panel_kernel(z64* A, int n, int j0, int nb, int* ipiv, int* info) {
  __shared__ double sval[LU_THREADS];
  __shared__ int sidx[LU_THREADS];
  __shared__ z64 srow[NBMAX];
  const int64_t w = blockIdx.x;
  z64* Aw = A + w * (int64_t)n * n;
  const int tid = threadIdx.x;
  for (int c = j0; c < j0 + nb; ++c) {
    double best = -1.0;
    int bi = c;                                // also the pivot if every entry is NaN
    for (int r = c + tid; r < n; r += LU_THREADS) {
      const z64 v = Aw[(int64_t)r * n + c];
      const double a = fabs(v.x) + fabs(v.y);
      if (a > best) { best = a; bi = r; }      // rows visited in increasing order: ties keep the first
    }
    sval[tid] = best;
    sidx[tid] = bi;
    __syncthreads();
    for (int h = LU_THREADS / 2; h > 0; h >>= 1) {
      if (tid < h) {
        const double o = sval[tid + h];
        const int oi = sidx[tid + h];
        if (o > sval[tid] || (o == sval[tid] && oi < sidx[tid])) { sval[tid] = o; sidx[tid] = oi; }
      }
      __syncthreads();
    }
    const int p = sidx[0];
    if (tid == 0) ipiv[w * n + c] = p;
    // swap rows c and p inside the panel
    if (p != c && tid < nb) {
      const int64_t a = (int64_t)c * n + j0 + tid, b = (int64_t)p * n + j0 + tid;
      const z64 t = Aw[a];
      Aw[a] = Aw[b];
      Aw[b] = t;
    }
    __syncthreads();
    const z64 piv = Aw[(int64_t)c * n + c];
    const bool zero_piv = piv.x == 0.0 && piv.y == 0.0;
    if (zero_piv) {
      if (tid == 0 && info[w] == 0) info[w] = c + 1;
    } else {
      const double m = 1.0 / (piv.x * piv.x + piv.y * piv.y);
      const z64 inv = z64{piv.x * m, -piv.y * m};
      for (int r = c + 1 + tid; r < n; r += LU_THREADS) {
        const int64_t a = (int64_t)r * n + c;
        Aw[a] = cmul(Aw[a], inv);
      }
    }
    if (tid < j0 + nb - c - 1) srow[tid] = Aw[(int64_t)c * n + c + 1 + tid];
    __syncthreads();
    // rank-1 update of the panel's remaining columns
    const int m = j0 + nb - c - 1;
    if (m > 0 && !zero_piv) {
      const int rows = n - c - 1;
      for (int e = tid; e < rows * m; e += LU_THREADS) {
        const int r = c + 1 + e / m, cc = e - (e / m) * m;
        const int64_t a = (int64_t)r * n + c + 1 + cc;
        Aw[a] = csub(Aw[a], cmul(Aw[(int64_t)r * n + c], srow[cc]));
      }
    }
    __syncthreads();
  }
}

// Apply the row interchanges ipiv[w * ldp + i], i = i_begin .. i_end (in
// order), to the columns of B outside [skip_lo, skip_hi): one thread per
// column.  B: nw matrices of ldb columns, row stride ldb, matrix stride b_ms.
__global__ __launch_bounds__(256) void lu_row_swaps_kernel(z64* B, int ldb, int64_t b_ms, int ncols,
                                                           const int* ipiv, int ldp, int i_begin, int i_end,
                                                           int skip_lo, int skip_hi) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncols || (c >= skip_lo && c < skip_hi)) return;
  const int64_t w = blockIdx.y;
  z64* Bw = B + w * b_ms;
  const int* pw = ipiv + w * ldp;
  for (int i = i_begin; i < i_end; ++i) {
    const int p = pw[i];
    if (p != i) {
      const z64 t = Bw[(int64_t)i * ldb + c];
      Bw[(int64_t)i * ldb + c] = Bw[(int64_t)p * ldb + c];
      Bw[(int64_t)p * ldb + c] = t;
    }
  }
}

__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Complex Gaussian sample for counter e (two uniforms from one hash chain,
// Box-Muller).  The host generator of the test suite uses the same formulas.
__device__ inline z64 gauss_c(uint64_t seed, uint64_t e, double sigma) {
  const uint64_t h1 = splitmix64(seed * 0xD1B54A32D192ED03ull + e);
  const uint64_t h2 = splitmix64(h1);
  const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
  const double rad = sigma * sqrt(-2.0 * log(u1));
  const double ang = 6.283185307179586 * u2;
  double sn, cs;
  sincos(ang, &sn, &cs);
  return z64{rad * cs, rad * sn};
}

// One pair per grid row (blockIdx.y), the pair's NQ * nao^2 elements along
// x: no 64-bit index division per element (they cost more than the Gaussian
// itself).  Counter of element (Q, i, j) of stored pair p:
// (p * NQ + Q) * nao^2 + i * nao + j.
__global__ void synth_integrals_kernel(z64* V, int64_t p0, const int* diag, int NQ, int nao, uint64_t seed,
                                       double sigma) {
  const int nn = nao * nao;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;   // element of the pair
  if (e >= NQ * nn) return;
  const int64_t lp = blockIdx.y;
  const int Q = e / nn, ij = e - Q * nn;
  const uint64_t base = ((uint64_t)(p0 + lp) * NQ + Q) * nn;
  z64 v = gauss_c(seed, base + ij, sigma);
  if (diag[lp]) {
    const int i = ij / nao, j = ij - i * nao;
    const z64 t = gauss_c(seed, base + j * nao + i, sigma);
    v = z64{0.5 * (v.x + t.x), 0.5 * (v.y - t.y)};
  }
  V[lp * NQ * nn + e] = v;
}

inline int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

void block_copy(const BlockCopy& a, hipStream_t s) {
  if (a.rows <= 0 || a.cols <= 0 || a.n_outer <= 0 || a.n_inner <= 0) return;
  require(a.n_inner <= 65535, "block_copy: inner batch above 65535");
  require(!(a.in_f32 && !a.out_f32), "block_copy: complex64 -> complex128 not supported");
  // grid.z is limited to 65535: launch the outer batches in chunks
  const int per = 65535 / a.n_inner;
  for (int o0 = 0; o0 < a.n_outer; o0 += per) {
    BlockCopy c = a;
    c.n_outer = std::min(per, a.n_outer - o0);
    if (c.in_off) c.in_off += o0;
    if (c.out_off) c.out_off += o0;
    if (c.trans) c.trans += o0;
    if (c.conj) c.conj += o0;
    dim3 grid((c.cols + TILE - 1) / TILE, (c.rows + TILE - 1) / TILE, c.n_outer * c.n_inner);
    if (!c.in_f32 && !c.out_f32) hipLaunchKernelGGL((block_copy_kernel<z64, z64>), grid, dim3(256), 0, s, c);
    else if (!c.in_f32) hipLaunchKernelGGL((block_copy_kernel<z64, c32>), grid, dim3(256), 0, s, c);
    else hipLaunchKernelGGL((block_copy_kernel<c32, c32>), grid, dim3(256), 0, s, c);
    GGPU_HIP(hipGetLastError());
  }
}

void slab_gather(const z64* in, const int64_t* src_off, z64* out, const int64_t* dst_off, int n_slabs,
                 int64_t len, double scale, hipStream_t s) {
  if (n_slabs <= 0 || len <= 0) return;
  hipLaunchKernelGGL(slab_gather_kernel, dim3(grid_for(n_slabs * len, 256)), dim3(256), 0, s, in, src_off,
                     out, dst_off, n_slabs, len, scale);
  GGPU_HIP(hipGetLastError());
}

void herm_one_minus(z64* M, z64* A, int n, int nw, hipStream_t s) {
  if (n <= 0 || nw <= 0) return;
  const int T = (n + TILE - 1) / TILE;
  require(nw <= 65535, "herm_one_minus: too many matrices");
  hipLaunchKernelGGL(herm_one_minus_kernel, dim3(T * (T + 1) / 2, nw), dim3(256), 0, s, M, A, n);
  GGPU_HIP(hipGetLastError());
}

void to_c32(const z64* in, c32* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(to_c32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, in, out, n);
  GGPU_HIP(hipGetLastError());
}

void zero(z64* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
  GGPU_HIP(hipGetLastError());
}

void zero_int(int* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(zero_int_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
  GGPU_HIP(hipGetLastError());
}

void sum_partials(const z64* in, int64_t stride, int np, z64* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(sum_partials_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, in, stride, np, out, n);
  GGPU_HIP(hipGetLastError());
}

void chol_diag_block(z64* A, int n, int nw, int j0, int nb, int* info, hipStream_t s) {
  require(nb >= 1 && nb <= NBMAX && j0 >= 0 && j0 + nb <= n, "chol_diag_block: bad block");
  require(nw <= 2147483647, "chol_diag_block: nw");
  hipLaunchKernelGGL(chol_diag_kernel, dim3(nw), dim3(256), 0, s, A, n, j0, nb, info);
  GGPU_HIP(hipGetLastError());
}

void chol_panel(z64* A, int n, int nw, int j0, int nb, hipStream_t s) {
  require(nb >= 1 && nb <= NBMAX && j0 >= 0 && j0 + nb <= n, "chol_panel: bad block");
  const int rows = n - j0 - nb;
  if (rows <= 0) return;
  require(nw <= 65535, "chol_panel: nw");
  hipLaunchKernelGGL(chol_panel_kernel, dim3((rows + 63) / 64, nw), dim3(64), 0, s, A, n, j0, nb);
  GGPU_HIP(hipGetLastError());
}

template <int KIND>
static void trsm_launch(const z64* A, int lda, int64_t a_ms, z64* B, int ldb, int64_t b_ms, int bcol0, int ncols,
                        int nw, int i0, int nb, hipStream_t s) {
  if (ncols <= 0) return;
  hipLaunchKernelGGL((trsm_block_kernel<KIND>), dim3((ncols + 63) / 64, nw), dim3(64), 0, s, A, lda, a_ms, B, ldb,
                     b_ms, bcol0, ncols, i0, nb);
  GGPU_HIP(hipGetLastError());
}

void trsm_lower_block(const z64* L, z64* B, int n, int nrhs, int nw, int i0, int nb, hipStream_t s) {
  require(nb >= 1 && nb <= NBMAX && i0 >= 0 && i0 + nb <= n && nw <= 65535, "trsm_lower_block: bad block");
  trsm_launch<kTriLower>(L, n, (int64_t)n * n, B, nrhs, (int64_t)n * nrhs, 0, nrhs, nw, i0, nb, s);
}

void trsm_upper_h_block(const z64* L, z64* B, int n, int nrhs, int nw, int i0, int nb, hipStream_t s) {
  require(nb >= 1 && nb <= NBMAX && i0 >= 0 && i0 + nb <= n && nw <= 65535, "trsm_upper_h_block: bad block");
  trsm_launch<kTriUpperH>(L, n, (int64_t)n * n, B, nrhs, (int64_t)n * nrhs, 0, nrhs, nw, i0, nb, s);
}

void lu_panel(z64* A, int n, int nw, int j0, int nb, int* ipiv, int* info, hipStream_t s) {
  require(nb >= 1 && nb <= NBMAX && j0 >= 0 && j0 + nb <= n, "lu_panel: bad block");
  hipLaunchKernelGGL(lu_panel_kernel, dim3(nw), dim3(LU_THREADS), 0, s, A, n, j0, nb, ipiv, info);
  GGPU_HIP(hipGetLastError());
}

void lu_row_swaps(z64* B, int ldb, int ncols, int nw, const int* ipiv, int ldp, int i_begin, int i_end, int skip_lo,
                  int skip_hi, hipStream_t s) {
  require(i_begin >= 0 && i_end <= ldp && nw <= 65535 && ncols <= ldb, "lu_row_swaps: bad range");
  if (ncols <= 0 || i_end <= i_begin) return;
  hipLaunchKernelGGL(lu_row_swaps_kernel, dim3((ncols + 255) / 256, nw), dim3(256), 0, s, B, ldb,
                     (int64_t)ldp * ldb, ncols, ipiv, ldp, i_begin, i_end, skip_lo, skip_hi);
  GGPU_HIP(hipGetLastError());
}

void trsm_unit_lower_cols(const z64* A, z64* B, int n, int nw, int i0, int nb, int bcol0, int ncols, hipStream_t s) {
  require(nb >= 1 && nb <= NBMAX && i0 >= 0 && i0 + nb <= n && nw <= 65535 && bcol0 >= 0 && bcol0 + ncols <= n,
          "trsm_unit_lower_cols: bad block");
  trsm_launch<kTriUnitLower>(A, n, (int64_t)n * n, B, n, (int64_t)n * n, bcol0, ncols, nw, i0, nb, s);
}

void trsm_upper_cols(const z64* A, z64* B, int n, int nw, int i0, int nb, int bcol0, int ncols, hipStream_t s) {
  require(nb >= 1 && nb <= NBMAX && i0 >= 0 && i0 + nb <= n && nw <= 65535 && bcol0 >= 0 && bcol0 + ncols <= n,
          "trsm_upper_cols: bad block");
  trsm_launch<kTriUpper>(A, n, (int64_t)n * n, B, n, (int64_t)n * n, bcol0, ncols, nw, i0, nb, s);
}

void synth_integrals(z64* V, int64_t p0, int64_t np, const int* pair_is_diag, int NQ, int nao, uint64_t seed,
                     double sigma, hipStream_t s) {
  if (np <= 0) return;
  const int64_t per_pair = (int64_t)NQ * nao * nao;
  require(per_pair < ((int64_t)1 << 31), "synth_integrals: NQ * nao^2 above 2^31");
  const unsigned bx = (unsigned)((per_pair + 255) / 256);
  // grid.y is limited to 65535: launch the pairs in chunks
  for (int64_t c = 0; c < np; c += 65535) {
    const int64_t n = std::min<int64_t>(65535, np - c);
    hipLaunchKernelGGL(synth_integrals_kernel, dim3(bx, (unsigned)n), dim3(256), 0, s, V + c * per_pair, p0 + c,
                       pair_is_diag + c, NQ, nao, seed, sigma);
    GGPU_HIP(hipGetLastError());
  }
}

}  // namespace ggpu
