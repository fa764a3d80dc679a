// Bandwidth-bound and small-matrix kernels of the HF/GW path (gfx950).
#pragma once

#include "common.hpp"

namespace ggpu {

// out[b][r][c] = op(in[b]) where op is a copy (trans = 0) or a transpose
// (trans = 1, reads in[b][c][r]) of an rows x cols block, optionally
// complex-conjugated.  Batch b = (outer, inner): element offsets are
// in_off[outer] + inner * in_bs (same for out); flags are per outer batch.
// Elements are multiplied by `scale`; a null in_off/out_off means offset 0.
struct BlockCopy {
  const void* in = nullptr;          // complex128, or complex64 if in_f32
  void* out = nullptr;               // complex128, or complex64 if out_f32
  int in_f32 = 0, out_f32 = 0;
  int rows = 0, cols = 0;          // shape of the OUTPUT block
  int64_t in_ld = 0, out_ld = 0;   // leading dimensions (row strides)
  int n_inner = 1;
  int64_t in_bs = 0, out_bs = 0;
  int n_outer = 1;
  const int64_t* in_off = nullptr;   // device, n_outer entries
  const int64_t* out_off = nullptr;  // device, n_outer entries
  const int* trans = nullptr;        // device, n_outer entries (nullptr: trans_all)
  const int* conj = nullptr;         // device, n_outer entries (nullptr: conj_all)
  int trans_all = 0, conj_all = 0;
  double scale = 1.0;                // applied to every element
};
void block_copy(const BlockCopy& a, hipStream_t s);

// Contiguous-slab gather: out[o][i] = scale * in[src[o]][i] for i < len
// (complex elements; src offsets in elements).  Used for the host <-> device
// layout changes of G and Sigma.
void slab_gather(const z64* in, const int64_t* src_off, z64* out, const int64_t* dst_off,
                 int n_slabs, int64_t len, double scale, hipStream_t s);

// For every w: H = (M + M^dagger)/2 (real diagonal), written back to M, and
// A = I - H written to A (if A != nullptr).  M, A: nw x n x n.
void herm_one_minus(z64* M, z64* A, int n, int nw, hipStream_t s);

// out[i] = (complex64) in[i]
void to_c32(const z64* in, c32* out, int64_t n, hipStream_t s);

// Fill n complex elements with zero.
void zero(z64* p, int64_t n, hipStream_t s);
void zero_int(int* p, int64_t n, hipStream_t s);

// out[i] = sum_p in[p * stride + i], i < n.
void sum_partials(const z64* in, int64_t stride, int np, z64* out, int64_t n, hipStream_t s);

// ---- batched blocked Cholesky (lower, A = L L^dagger), nw matrices n x n ----
// Factor the nb x nb diagonal block at (j0, j0) in place; info[w] = column+1
// of the first non-positive pivot (0 = fine).
void chol_diag_block(z64* A, int n, int nw, int j0, int nb, int* info, hipStream_t s);
// A[j0+nb:n, j0:j0+nb] <- A[...] * L11^-dagger
void chol_panel(z64* A, int n, int nw, int j0, int nb, hipStream_t s);
// B[i0:i0+nb, :] <- L[i0:i0+nb, i0:i0+nb]^-1 B[i0:i0+nb, :]    (lower, forward)
void trsm_lower_block(const z64* L, z64* B, int n, int nrhs, int nw, int i0, int nb, hipStream_t s);
// B[i0:i0+nb, :] <- (L[i0:i0+nb, i0:i0+nb]^dagger)^-1 B[i0:i0+nb, :] (backward)
void trsm_upper_h_block(const z64* L, z64* B, int n, int nrhs, int nw, int i0, int nb, hipStream_t s);

// ---- batched blocked LU with partial pivoting (getrf / getrs), nw matrices n x n ----
// Factor the panel of columns j0 .. j0+nb (rows j0 .. n) in place with row
// interchanges inside the panel; ipiv[w * n + c] = pivot row of column c;
// info[w] = column + 1 of the first exactly zero pivot (0 = none).
void lu_panel(z64* A, int n, int nw, int j0, int nb, int* ipiv, int* info, hipStream_t s);
// Apply the interchanges ipiv[w * ldp + i], i = i_begin .. i_end in order, to
// the columns [0, ncols) of B (row stride ldb, matrix stride ldp * ldb)
// outside [skip_lo, skip_hi).
void lu_row_swaps(z64* B, int ldb, int ncols, int nw, const int* ipiv, int ldp, int i_begin, int i_end, int skip_lo,
                  int skip_hi, hipStream_t s);
// B[i0:i0+nb, bcol0:bcol0+ncols] <- T^-1 B[...], T the nb x nb diagonal block
// of A at (i0, i0), unit lower (LU's L) or upper non-unit (LU's U); A and B
// are n x n (B may be A itself when the column ranges do not overlap).
void trsm_unit_lower_cols(const z64* A, z64* B, int n, int nw, int i0, int nb, int bcol0, int ncols, hipStream_t s);
void trsm_upper_cols(const z64* A, z64* B, int n, int nw, int i0, int nb, int bcol0, int ncols, hipStream_t s);

// ---- synthetic density-fitted integrals (counter-based, reproducible) ----
// V[p][Q][i][j] for stored pairs p in [p0, p0+np) of the pair list k1 >= k2,
// Re/Im ~ N(0, sigma^2); diagonal pairs (k1 == k2) are Hermitised per Q.
void synth_integrals(z64* V, int64_t p0, int64_t np, const int* pair_is_diag, int NQ, int nao,
                     uint64_t seed, double sigma, hipStream_t s);

}  // namespace ggpu
