// Host-side engines of the HF / GW path: symmetry tables, the device
// integral store, and the plans behind the C ABI.
#pragma once

#include "../../include/green_gpu_amd.h"
#include "common.hpp"
#include "kernels.hpp"
#include "zgemm.hpp"

#include <array>

namespace ggpu {

// How stored pair data is turned into V^Q(k1,k2): per-Q transpose and/or
// complex conjugation of the stored block at position `pos`
// (df_integral_t.h:146-235: conj-transpose if k1 < k2, then conj for a
// conjugated pair, transpose for a transposed pair).
struct PairView {
  int64_t pos;
  int trans;
  int conj;
};

struct SymTables {
  int nk = 0, ink = 0, nq = 0, inq = 0;
  std::vector<int64_t> k_f2r, k_r2f, k_tr, q_f2r, q_r2f, q_tr, k1_from_k2q, k2_from_k1q;
  std::vector<std::vector<int64_t>> k_stars, q_stars;
  std::vector<z64> U_k, U_q;  // empty = identity
  std::vector<int64_t> conj_kpair, trans_kpair, irre_pos;
  int64_t num_stored = 0;

  void load(const ggpu_symmetry& s, int nao, int NQ);
  PairView view(int64_t k1, int64_t k2) const;
  bool k_identity() const { return U_k.empty(); }
  bool q_identity() const { return U_q.empty(); }
};

struct PairTable;

// The stored k-pair integrals, (num_stored, NQ, nao, nao): either resident in
// HBM, or in pinned host memory and streamed pair batch by pair batch into a
// double-buffered device staging area on a copy stream of their own, each
// batch's copy overlapping the previous batch's compute.
class IntegralStore {
 public:
  ~IntegralStore();
  // Where the stored pairs live (before loading): kResDevice (HBM),
  // kResHost (pinned host memory, copied per pair batch into staging),
  // kResGenerated (synthetic pairs generated per batch into staging on the
  // device; only synth() fills such a store) or kResReader (each batch's
  // pairs fetched through a caller's reader into pinned host staging, then
  // copied like kResHost; only set_reader() fills such a store).
  enum { kResDevice = 0, kResHost = 1, kResGenerated = 2, kResReader = 3 };
  typedef int (*Reader)(void* user, int64_t pos, double* dst);
  void set_residence(int mode, int slots, int64_t pair_elems);
  void set_reader(Reader fn, void* user, int64_t num_stored);
  void load_host(const double* V_stored, int64_t n_elems, hipStream_t s);
  void synth(const SymTables& t, int NQ, int nao, uint64_t seed, double sigma, hipStream_t s);
  bool loaded() const { return loaded_; }
  bool staged() const { return mode_ != kResDevice; }
  bool reader() const { return mode_ == kResReader; }
  int64_t elems() const { return n_elems_; }
  double bytes_streamed() const { return streamed_; }
  double bytes_generated() const { return generated_; }
  // Device view of pairs [i0, i0 + n) of `tab` and the element offset of each
  // (n entries, device): the store itself when resident in HBM, else the
  // staging buffer the batch was copied to.  `s` waits for the copy; call
  // done(s) after the last read of the view to release the buffer.
  const z64* view(const PairTable& tab, int i0, int n, hipStream_t s, const int64_t** offsets);
  void done(hipStream_t s);
  // Device time (ms) the copy stream spent copying (or generating) staged
  // batches since the last call, from an event pair around every batch;
  // call after the streams have drained.
  double collect_staging_ms();
  // Graph capture of an iteration on stream s: begin_capture() makes the
  // copy stream join s's capture (and forgets buffer state recorded outside
  // it), end_capture() joins it back; the staging copies or generation
  // kernels then become nodes of the graph.  Not for the reader residence
  // (its host-side reads cannot be captured).
  void begin_capture(hipStream_t s);
  void end_capture(hipStream_t s);
  void abort_capture() { capturing_ = false; pre_ = Pending{}; cur_ = -1; }

 private:
  void alloc_host(int64_t n_elems);
  std::vector<hipEvent_t> tpool_;          // (start, end) event pairs of staged batches
  size_t tused_ = 0;
  void issue(const PairTable& tab, int i0, int n, int buf);
  DeviceBuffer<z64> V_;
  z64* host_ = nullptr;
  // reader residence: the callback and two pinned batch buffers (one per
  // device staging buffer; each is refilled after its last copy completed)
  Reader reader_ = nullptr;
  void* reader_user_ = nullptr;
  z64* hstage_[2] = {nullptr, nullptr};
  int mode_ = kResDevice;
  bool loaded_ = false;
  // generated residence: the synth() parameters and per-position diagonal flags
  uint64_t seed_ = 0;
  double sigma_ = 0;
  int NQ_ = 0, nao_ = 0;
  DeviceBuffer<int> diag_;
  int64_t n_elems_ = 0, pair_elems_ = 0;
  int slots_ = 0;
  hipStream_t copy_ = nullptr;
  hipEvent_t ev_copy_[2] = {nullptr, nullptr}, ev_used_[2] = {nullptr, nullptr};
  DeviceBuffer<z64> stage_;
  DeviceBuffer<int64_t> stage_off_;        // [2][slots]: (buf * slots + b) * pair_elems
  struct Pending { const PairTable* tab = nullptr; int i0 = -1, n = 0, buf = -1; } pre_;
  int cur_ = -1, next_ = 0;
  bool capturing_ = false;
  bool used_in_capture_[2] = {false, false};   // ev_used_[b] recorded inside the capture
  hipEvent_t ev_join_[2] = {nullptr, nullptr};
  double streamed_ = 0, generated_ = 0;
};

// Per-outer-batch tables for BlockCopy (device arrays uploaded once).
struct PairTable {
  DeviceBuffer<int64_t> in_off;   // pos * NQ * nao^2
  DeviceBuffer<int> trans, conj;
  std::vector<PairView> host;
  void build(const std::vector<PairView>& views, int64_t stride, hipStream_t s);
};

// Unfold pairs [i0, i0+n) of a PairTable into dst[b] (NQ x nao x nao each).
void unfold_pairs(IntegralStore& st, const PairTable& tab, int i0, int n, const int64_t* out_off_dev,
                  void* dst, int out_f32, int NQ, int nao, hipStream_t s);

// Rotate IBZ matrices into the full BZ: out[k][s][t] = U_k X[t][s][ibz(k)] U_k^dagger
// (conjugated for time-reversal images).  X has host layout (nt, ns, ink, n, n),
// out has (nk, ns, nt, n, n).
class FullBZRotator {
 public:
  void init(const SymTables& t, int n, int ns, int nt, hipStream_t s);
  void run(const z64* X, z64* out, hipStream_t s);

 private:
  const SymTables* t_ = nullptr;
  int n_ = 0, ns_ = 0, nt_ = 0;
  DeviceBuffer<int64_t> src_, dst_;
  DeviceBuffer<z64> Uk_, tmp_;
};

// Times every GEMM launch with a pair of events on its stream and sums
// time, algorithmic FLOPs and launch counts per call site.
class SiteProfiler : public GemmProfiler {
 public:
  ~SiteProfiler() override;
  void before(const GemmDesc& d, hipStream_t s) override;
  void after(const GemmDesc& d, hipStream_t s) override;
  void reset();
  void collect();   // after the stream has been synchronised
  double ms[kNumSites] = {}, flops[kNumSites] = {};
  int64_t launches[kNumSites] = {};

 private:
  struct Rec { int site; double flops; size_t ev; };
  std::vector<hipEvent_t> pool_;
  std::vector<Rec> recs_;
  size_t used_ = 0;
};

// Tables of a BlockCopy between spinor matrices (nso = 2 nao) and their
// nao x nao spin blocks, one entry per outer batch.
struct SpinorTables {
  DeviceBuffer<int64_t> in_off, out_off;
  DeviceBuffer<int> trans, conj;
  std::vector<int64_t> h_in, h_out;   // host copies for the bounds check
  int n = 0;
  void build(const std::vector<int64_t>& in, const std::vector<int64_t>& out, const std::vector<int>& tr,
             const std::vector<int>& cj);
};

struct GwPlan {
  // ns: spin blocks the contractions run over (the two-component plan runs
  // over the four pseudo-spins aa, bb, ab, ba); nso: size of the caller's
  // matrices (nao, or 2 nao for X2C)
  int nao, nso, ns, NQ, nts, nw_b, half;
  bool x2c = false;
  Compute compute;
  int solver;
  SymTables sym;
  IntegralStore store;
  hipStream_t stream = nullptr;
  hipEvent_t ev[7];

  DeviceBuffer<z64> Tf_wt;   // folded tau -> Omega transform, (nw_b, half-1)
  DeviceBuffer<z64> T_tw;    // (nts, nw_b)
  DeviceBuffer<z64> G_host_dev, Gf, Grev, Sig_acc, Sig_host_dev;
  DeviceBuffer<int64_t> rev_src, rev_dst;   // Gf -> Grev (tau reversed) slab offsets
  DeviceBuffer<c32> Gf32, Grev32, Pt32;     // complex64 copies for the reduced-precision modes
  // P0 partial sums (split of the pair/spin reduction): more slices, more
  // workgroups per contraction launch (8: 2.6 % faster than 4 at nao = 52,
  // NQ = 300); the slices are a K-batch of the tau -> Omega transform.
  int p0_split = 8;
  DeviceBuffer<z64> P0, P0w, Amat, Pt, Prot, Ptmp, Uq_dev;
  DeviceBuffer<int> info;
  DeviceBuffer<int> ipiv;                // LU pivots, (nw_b, NQ)
  DeviceBuffer<z64> work;
  DeviceBuffer<char> ozwork;            // slices and scales of the int8-emulated FP64 products
  DeviceBuffer<int64_t> bidx_off;       // b * NQ * nao^2
  DeviceBuffer<int64_t> sig_src, sig_dst;
  std::vector<PairTable> p0_tables;     // per IBZ q
  std::vector<std::vector<PairTable>> sig_tables;  // per IBZ q, per star member
  FullBZRotator rot;
  int kb = 1;                           // pairs per batch
  double last_ms[6] = {0, 0, 0, 0, 0, 0};
  double streamed_last = 0;             // integral bytes copied host -> device by the last iterate
  double generated_last = 0;            // integral bytes generated on the device by the last iterate
  double staging_ms_last = 0;           // copy-stream time of those copies / generation (ms)
  // This device's share of the IBZ q-points: iq = share_rank, share_rank +
  // share_size, ... (the partial Sigma of all shares sums to the full one).
  int share_rank = 0, share_size = 1;
  bool profiling = false;
  SiteProfiler prof;
  std::vector<hipEvent_t> qev;          // per-q phase events
  // Launch variant per GEMM call site (-1: the built-in choice), picked by
  // autotune() or set by the caller; installed around the plan's launches.
  int variants[kNumSites];
  // Autotuner record: candidates and the per-site time (ms) of each, the best
  // over the rounds (negative: the site did not run); wall time of the tuning.
  std::vector<int> tune_cand;
  std::vector<std::vector<double>> tune_ms;   // [candidate][site]
  double tune_wall_ms = 0;
  int batch_limit = 0;                  // pair batches per q (and star member); 0: all
  int last_q = -1;                      // IBZ q whose intermediates P0 / P0w / Pt hold
  // Graph replay: the iteration's launches captured once into a HIP graph
  // (for the G / Sigma addresses, variants and q share of the capture) and
  // replayed by every later iterate; needs the integrals resident in HBM.
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  const void* graph_G = nullptr;
  const void* graph_S = nullptr;
  int graph_captures = 0;               // captures made (a new one after every change)
  int config_gen = 0;                   // bumped by every change that alters the launch sequence
  void invalidate_graph();              // after such a change: drop the graph, bump config_gen

  SpinorTables x2c_split, x2c_merge;   // G spinor -> pseudo-spin blocks, Sigma back

  GwPlan(int nao, int nso, int ns, int NQ, int nts, int nw_b, const ggpu_symmetry& s, const double* T_wt,
         const double* T_tw, int precision, int solver, size_t work_bytes);
  ~GwPlan();

  bool lowp() const { return complex64_operands(compute); }
  void iterate_device(const z64* G_dev_hostlayout, z64* Sigma_dev_hostlayout);
  void iterate(const double* G, double* Sigma);
  void polarization(const double* G, int iq, int which, double* out);
  // The intermediates the last iterate (or polarization call) left for its
  // last q: which = 0 P0(tau), 1 P(Omega), 2 P(tau); *iq = that IBZ q.
  void last_polarization(int which, int* iq, double* out);
  // Time every launch variant of the contraction call sites (P0 and Sigma
  // chains) on the first q of this plan's share, `batches` pair batches per
  // chain, best of `rounds`, and keep the fastest per site.
  void autotune(const z64* G_dev_hostlayout, int rounds, int batches);
  void set_variants(const int* v);
  // The launches of one iteration on the plan's stream; with timed, the
  // phase events around them.
  void enqueue_iteration(const z64* G_dev_hostlayout, z64* Sigma_dev_hostlayout, bool timed);
  void check_info();

 private:
  void stage_G(const z64* G_dev_hostlayout);
  void capture(const z64* G_dev_hostlayout, z64* Sigma_dev_hostlayout);
  void drop_graph();
  void p0_for_q(int iq);
  void dyson_for_q();
  void cholesky_solve_dyson();
  void lu_solve_dyson();
  void sigma_for_q(int iq);
};

// Several independent GW plans replayed as one HIP graph: each plan's
// iteration is a branch forked from and joined to the batch's own stream, so
// one launch runs all systems concurrently.  Plans must keep their integrals
// resident in HBM; a change to any plan (q share, variants, integrals) makes
// the next run capture again.  Plans must outlive the batch and must not be
// used from another thread while it runs.
struct GwBatch {
  std::vector<GwPlan*> plans;
  std::vector<const z64*> G;
  std::vector<z64*> S;
  std::vector<int> gens;                // config_gen of each plan at the capture
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t exec = nullptr;
  int captures = 0;
  double last_ms = 0;

  GwBatch(GwPlan* const* plans, int n, const void* const* G_dev, void* const* S_dev);
  ~GwBatch();
  void run();

 private:
  void capture();
};

struct HfPlan {
  // ns: the caller's spins; nps: density blocks the exchange runs over
  // (3 for X2C: aa, bb, ab); nso: the caller's matrix size
  int nao, nso, ns, nps, NQ;
  bool x2c = false;
  SymTables sym;
  IntegralStore store;
  hipStream_t stream = nullptr;
  FullBZRotator rot;
  DeviceBuffer<z64> dm_host_dev, dm_fbz, Kpart, Kout, work;
  DeviceBuffer<int64_t> bidx_off;
  std::vector<PairTable> tables;   // per IBZ k: pairs (k, k2) for all k2
  int kb = 1, nsplit = 1;
  // Arithmetic of the two exchange products (X = V D, K += X V^dagger); the
  // partial sums of K, J and the Ewald term are always FP64.
  Compute compute = Compute::F64;
  DeviceBuffer<c32> dm32;               // complex64 copy of dm_fbz (reduced-precision modes)
  DeviceBuffer<char> ozwork;            // slices and scales of the int8-emulated FP64 products
  void set_precision(int precision);
  // Launch variant per call site of the exchange products (-1: built-in).
  int variants[kNumSites];
  // This device's share of the IBZ k-points: ik = share_rank, share_rank +
  // share_size, ...; the other rows of K and F are zero, so the results of
  // all shares sum to the full ones.
  int share_rank = 0, share_size = 1;

  // Two-electron HF energy of the last solve (FP64 traces on the device,
  // star-weighted sum on the host); valid after solve().
  double energy = 0.0;
  bool energy_valid = false;

  HfPlan(int nao, int nso, int ns, int NQ, const ggpu_symmetry& s);
  ~HfPlan();
  void exchange(const double* dm, double* K);
  // The exchange on caller device buffers (dm, K: (ns, ink, nso, nso)) with
  // the K partial sums in the library's static per-GPU scratch array instead
  // of the plan's partial buffer: no device allocation.  Needs the integrals
  // resident in HBM.  Calls that overlap on one GPU race on the scratch.
  void exchange_device_noalloc(const z64* dm_dev, z64* K_dev);
  z64* scratch = nullptr;            // this device's copy of the static scratch array
  int64_t scratch_elems = 0;
  void solve(const double* dm, const double* S_k, double madelung, double* F);

 private:
  void stage_dm(const z64* dm_dev);   // caller-layout dm on the device -> dm_fbz (full BZ, density blocks)
  // Kout[ik][s] from dm_fbz; the K partial sums of the split (k2, Q)
  // reduction go to `part` (nsp slices of nps nao^2)
  void exchange_device(z64* part, int nsp);
  void fetch(double* out);  // Kout -> caller layout (ns, ink, nso, nso)
  std::vector<PairTable> diag_tables;  // (k, k) pairs: [0] all full k, [1] IBZ representatives
  DeviceBuffer<z64> jvec, S_dev, Ftmp, dm2c_fbz, F2c, traces;
  DeviceBuffer<int64_t> kout_src, kout_dst;   // Kout (ik, s) -> caller layout (s, ik) slab offsets
  void compute_energy();    // from dm_host_dev and the Fock matrices of solve()
  SpinorTables x2c_split, x2c_merge;
};

}  // namespace ggpu
