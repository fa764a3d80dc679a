// Host orchestration of the HF exchange and GW iteration on one gfx950 GPU.
//
// Device data layout (HBM-resident for the whole plan / iteration):
//   V store   (num_stored, NQ, nao, nao)   stored k-pairs, as the reference's integral reader holds them
//   Gf        (nk, ns, nts, nao, nao)      G on the full BZ (rotated once per iteration)
//   Sig_acc   (ink, ns, nts, nao, nao)     Sigma accumulator for the IBZ k
//   P0        (nts/2, NQ, NQ)              bare bubble of the current q, t < nts/2
//   P0w/Amat  (nw_b, NQ, NQ)               P0(Omega) -> right-hand side / (1 - P0) factor
//   Pt        (nts, NQ, NQ)                dressed polarization of the current q
//   work      pair-batch intermediates (X1, X2 for P0; Y1, Y2 for Sigma)
#include "engine.hpp"
#include "ozaki.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace ggpu {

// Launch a GEMM with every operand tied to its allocation, so that the
// host-side bounds check covers the whole launch.
template <class BA, class BB, class BC>
static void gemm_on(GemmDesc g, const BA& a, const BB& b, const BC& c, hipStream_t s, int site = -1) {
  if (site >= 0) g.site = site;
  g.A.buf = a.get(); g.A.buf_elems = (int64_t)a.size();
  g.B.buf = b.get(); g.B.buf_elems = (int64_t)b.size();
  g.C.buf = c.get(); g.C.buf_elems = (int64_t)c.size();
  gemm(g, s);
}

// View of caller memory with a known element count (device pointers handed in).
struct Span {
  const void* p;
  size_t n;
  const void* get() const { return p; }
  size_t size() const { return n; }
};

// ---------------------------------------------------------------------------
// Symmetry tables
// ---------------------------------------------------------------------------
static bool is_identity(const double* U, int64_t count, int n) {
  for (int64_t m = 0; m < count; ++m)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double* z = U + 2 * ((m * n + i) * (int64_t)n + j);
        if (z[0] != (i == j ? 1.0 : 0.0) || z[1] != 0.0) return false;
      }
  return true;
}

static std::vector<int64_t> copy_idx(const int64_t* p, int64_t n, const char* name) {
  require(p != nullptr, std::string("symmetry: ") + name + " is NULL");
  return std::vector<int64_t>(p, p + n);
}

static void check_range(const std::vector<int64_t>& v, int64_t lo, int64_t hi, const char* name) {
  for (auto x : v) require(x >= lo && x < hi, std::string("symmetry: ") + name + " entry out of range");
}

void SymTables::load(const ggpu_symmetry& s, int nao, int NQ) {
  require(s.nk > 0 && s.ink > 0 && s.nq > 0 && s.inq > 0 && s.ink <= s.nk && s.inq <= s.nq,
          "symmetry: bad k/q counts");
  require(s.nq == s.nk, "symmetry: the q mesh must equal the k mesh");
  nk = s.nk; ink = s.ink; nq = s.nq; inq = s.inq;
  k_f2r = copy_idx(s.k_full_to_reduced, nk, "k_full_to_reduced");
  k_r2f = copy_idx(s.k_reduced_to_full, ink, "k_reduced_to_full");
  k_tr = copy_idx(s.k_tr_conj, nk, "k_tr_conj");
  q_f2r = copy_idx(s.q_full_to_reduced, nq, "q_full_to_reduced");
  q_r2f = copy_idx(s.q_reduced_to_full, inq, "q_reduced_to_full");
  q_tr = copy_idx(s.q_tr_conj, nq, "q_tr_conj");
  k1_from_k2q = copy_idx(s.k1_from_k2q, (int64_t)nk * nq, "k1_from_k2q");
  k2_from_k1q = copy_idx(s.k2_from_k1q, (int64_t)nk * nq, "k2_from_k1q");
  check_range(k_f2r, 0, ink, "k_full_to_reduced");
  check_range(k_r2f, 0, nk, "k_reduced_to_full");
  check_range(q_f2r, 0, inq, "q_full_to_reduced");
  check_range(q_r2f, 0, nq, "q_reduced_to_full");
  check_range(k1_from_k2q, 0, nk, "k1_from_k2q");
  check_range(k2_from_k1q, 0, nk, "k2_from_k1q");
  auto stars = [&](const int64_t* off, const int64_t* mem, int n, int64_t full, const char* nm) {
    require(off && mem, std::string("symmetry: ") + nm + " star arrays are NULL");
    std::vector<std::vector<int64_t>> out(n);
    require(off[0] == 0 && off[n] == full, std::string("symmetry: ") + nm + " stars must cover the mesh");
    for (int i = 0; i < n; ++i) {
      require(off[i + 1] > off[i], std::string("symmetry: empty ") + nm + " star");
      out[i].assign(mem + off[i], mem + off[i + 1]);
      check_range(out[i], 0, full, nm);
    }
    return out;
  };
  k_stars = stars(s.k_star_offsets, s.k_star_members, ink, nk, "k");
  q_stars = stars(s.q_star_offsets, s.q_star_members, inq, nq, "q");
  U_k.clear();
  U_q.clear();
  if (s.U_k && !is_identity(s.U_k, nk, nao)) {
    U_k.resize((size_t)nk * nao * nao);
    std::memcpy(U_k.data(), s.U_k, U_k.size() * sizeof(z64));
  }
  if (s.U_q && !is_identity(s.U_q, nq, NQ)) {
    U_q.resize((size_t)nq * NQ * NQ);
    std::memcpy(U_q.data(), s.U_q, U_q.size() * sizeof(z64));
  }
  const int64_t npair = (int64_t)nk * (nk + 1) / 2;
  if (s.conj_kpair || s.trans_kpair || s.irre_pos) {
    conj_kpair = copy_idx(s.conj_kpair, npair, "conj_kpair");
    trans_kpair = copy_idx(s.trans_kpair, npair, "trans_kpair");
    irre_pos = copy_idx(s.irre_pos, npair, "irre_pos");
    num_stored = s.num_kpair_stored;
    check_range(conj_kpair, 0, npair, "conj_kpair");
    check_range(trans_kpair, 0, npair, "trans_kpair");
    check_range(irre_pos, -1, num_stored, "irre_pos");
  } else {
    conj_kpair.resize(npair);
    trans_kpair.resize(npair);
    irre_pos.resize(npair);
    for (int64_t i = 0; i < npair; ++i) conj_kpair[i] = trans_kpair[i] = irre_pos[i] = i;
    num_stored = npair;
    require(s.num_kpair_stored == 0 || s.num_kpair_stored == npair,
            "symmetry: num_kpair_stored must be nk(nk+1)/2 without pair tables");
  }
  require(num_stored > 0, "symmetry: no stored pairs");
}

PairView SymTables::view(int64_t k1, int64_t k2) const {
  int64_t idx = (k1 >= k2) ? k1 * (k1 + 1) / 2 + k2 : k2 * (k2 + 1) / 2 + k1;
  const bool neg = k1 < k2;
  int kind = 0;  // 0 direct, 1 conjugated, 2 transposed
  if (conj_kpair[idx] != idx) { kind = 1; idx = conj_kpair[idx]; }
  else if (trans_kpair[idx] != idx) { kind = 2; idx = trans_kpair[idx]; }
  const int64_t pos = irre_pos[idx];
  require(pos >= 0 && pos < num_stored, "integral pair is not stored");
  PairView v;
  v.pos = pos;
  v.trans = (neg ? 1 : 0) ^ (kind == 2 ? 1 : 0);
  v.conj = (neg ? 1 : 0) ^ (kind == 1 ? 1 : 0);
  return v;
}

// ---------------------------------------------------------------------------
// Integral store and pair unfolding
// ---------------------------------------------------------------------------
IntegralStore::~IntegralStore() {
  if (copy_) (void)hipStreamSynchronize(copy_);
  for (auto e : tpool_) (void)hipEventDestroy(e);
  for (int b = 0; b < 2; ++b) {
    if (ev_copy_[b]) (void)hipEventDestroy(ev_copy_[b]);
    if (ev_used_[b]) (void)hipEventDestroy(ev_used_[b]);
    if (ev_join_[b]) (void)hipEventDestroy(ev_join_[b]);
  }
  if (copy_) (void)hipStreamDestroy(copy_);
  if (host_) (void)hipHostFree(host_);
  for (auto h : hstage_)
    if (h) (void)hipHostFree(h);
}

void IntegralStore::set_residence(int mode, int slots, int64_t pair_elems) {
  require(!loaded_, "integrals: choose the residence before loading integrals", kErrState);
  require(mode == kResDevice || mode == kResHost || mode == kResGenerated || mode == kResReader,
          "integrals: unknown residence");
  mode_ = mode;
  if (mode == kResDevice) return;
  require(slots >= 1 && pair_elems > 0, "integrals: bad staging size");
  slots_ = slots;
  pair_elems_ = pair_elems;
  if (!copy_) {
    GGPU_HIP(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b) {
      GGPU_HIP(hipEventCreateWithFlags(&ev_copy_[b], hipEventDisableTiming));
      GGPU_HIP(hipEventCreateWithFlags(&ev_used_[b], hipEventDisableTiming));
      GGPU_HIP(hipEventCreateWithFlags(&ev_join_[b], hipEventDisableTiming));
    }
  }
  stage_.resize((size_t)2 * slots * pair_elems);
  std::vector<int64_t> off((size_t)2 * slots);
  for (int i = 0; i < 2 * slots; ++i) off[i] = i * pair_elems;
  stage_off_.resize(off.size());
  GGPU_HIP(hipMemcpy(stage_off_.get(), off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
}

void IntegralStore::alloc_host(int64_t n_elems) {
  if (host_ && n_elems_ == n_elems) return;
  if (host_) GGPU_HIP(hipHostFree(host_));
  host_ = nullptr;
  GGPU_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_), (size_t)n_elems * sizeof(z64), hipHostMallocDefault));
}

void IntegralStore::set_reader(Reader fn, void* user, int64_t num_stored) {
  require(fn != nullptr, "integrals: NULL pair reader");
  require(mode_ == kResReader, "integrals: a pair reader needs the READER residence", kErrState);
  for (int b = 0; b < 2; ++b)
    if (!hstage_[b])
      GGPU_HIP(hipHostMalloc(reinterpret_cast<void**>(&hstage_[b]), (size_t)slots_ * pair_elems_ * sizeof(z64),
                             hipHostMallocDefault));
  reader_ = fn;
  reader_user_ = user;
  n_elems_ = num_stored * pair_elems_;
  loaded_ = true;
}

void IntegralStore::load_host(const double* V_stored, int64_t n_elems, hipStream_t s) {
  require(V_stored != nullptr, "integrals: NULL pointer");
  require(mode_ != kResGenerated, "integrals: a generated store takes no loaded integrals", kErrState);
  require(mode_ != kResReader, "integrals: a reader store takes no loaded integrals", kErrState);
  if (mode_ == kResHost) {
    alloc_host(n_elems);
    std::memcpy(host_, V_stored, (size_t)n_elems * sizeof(z64));
  } else {
    V_.resize(n_elems);
    GGPU_HIP(hipMemcpyAsync(V_.get(), V_stored, n_elems * sizeof(z64), hipMemcpyHostToDevice, s));
    GGPU_HIP(hipStreamSynchronize(s));
  }
  n_elems_ = n_elems;
  loaded_ = true;
}

void IntegralStore::synth(const SymTables& t, int NQ, int nao, uint64_t seed, double sigma, hipStream_t s) {
  require(mode_ != kResReader, "integrals: a reader store takes no synthetic integrals", kErrState);
  // diagonal flag per stored position: find the pair index that maps to it
  std::vector<int> diag(t.num_stored, 0);
  for (int64_t k1 = 0; k1 < t.nk; ++k1) {
    const int64_t idx = k1 * (k1 + 1) / 2 + k1;
    if (t.conj_kpair[idx] == idx && t.trans_kpair[idx] == idx && t.irre_pos[idx] >= 0) diag[t.irre_pos[idx]] = 1;
  }
  DeviceBuffer<int> d(diag.size());
  GGPU_HIP(hipMemcpy(d.get(), diag.data(), diag.size() * sizeof(int), hipMemcpyHostToDevice));
  const int64_t per = (int64_t)NQ * nao * nao, n = t.num_stored * per;
  if (mode_ == kResGenerated) {
    // nothing is stored: issue() generates each batch of pairs on demand
    require(per == pair_elems_, "integrals: pair size differs from the staging size");
    diag_ = std::move(d);
    seed_ = seed;
    sigma_ = sigma;
    NQ_ = NQ;
    nao_ = nao;
  } else if (mode_ == kResHost) {
    // generate on the device in chunks of pairs, copy each to pinned memory
    alloc_host(n);
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(t.num_stored, ((int64_t)1 << 30) / (per * 16)));
    DeviceBuffer<z64> tmp((size_t)chunk * per);
    for (int64_t p0 = 0; p0 < t.num_stored; p0 += chunk) {
      const int64_t np = std::min(chunk, t.num_stored - p0);
      synth_integrals(tmp.get(), p0, np, d.get() + p0, NQ, nao, seed, sigma, s);
      GGPU_HIP(hipMemcpyAsync(host_ + p0 * per, tmp.get(), (size_t)np * per * sizeof(z64), hipMemcpyDeviceToHost, s));
      GGPU_HIP(hipStreamSynchronize(s));
    }
  } else {
    V_.resize((size_t)n);
    synth_integrals(V_.get(), 0, t.num_stored, d.get(), NQ, nao, seed, sigma, s);
    GGPU_HIP(hipStreamSynchronize(s));
  }
  n_elems_ = n;
  loaded_ = true;
}

// Copy pairs [i0, i0 + n) of tab into staging buffer buf, after the last
// reader of that buffer is done.
void IntegralStore::issue(const PairTable& tab, int i0, int n, int buf) {
  require(n <= slots_, "integrals: batch larger than the staging area");
  // inside a capture only events recorded in it may be waited on; the copy
  // stream joined the capture behind everything captured before
  if (!capturing_ || used_in_capture_[buf]) GGPU_HIP(hipStreamWaitEvent(copy_, ev_used_[buf], 0));
  while (!capturing_ && tpool_.size() < tused_ + 2) {
    hipEvent_t e;
    GGPU_HIP(hipEventCreate(&e));
    tpool_.push_back(e);
  }
  const size_t bytes = (size_t)pair_elems_ * sizeof(z64);
  if (mode_ == kResReader) {
    // the previous copies out of this host buffer have completed before it
    // is refilled (an event never recorded completes at once)
    GGPU_HIP(hipEventSynchronize(ev_copy_[buf]));
    z64* h = hstage_[buf];
    for (int i = 0; i < n; ++i) {
      const int64_t pos = tab.host[i0 + i].pos;
      require(pos >= 0 && (pos + 1) * pair_elems_ <= n_elems_, "integrals: pair outside store");
      const int rc = reader_(reader_user_, pos, reinterpret_cast<double*>(h + (int64_t)i * pair_elems_));
      require(rc == 0, "integrals: the pair reader failed for stored pair " + std::to_string(pos) + " (status " +
                           std::to_string(rc) + ")");
    }
  }
  if (!capturing_) GGPU_HIP(hipEventRecord(tpool_[tused_], copy_));
  if (mode_ == kResReader)
    GGPU_HIP(hipMemcpyAsync(stage_.get() + (int64_t)buf * slots_ * pair_elems_, hstage_[buf], bytes * n,
                            hipMemcpyHostToDevice, copy_));
  for (int i = 0; i < n && mode_ != kResReader; ++i) {
    const int64_t pos = tab.host[i0 + i].pos;
    require(pos >= 0 && (pos + 1) * pair_elems_ <= n_elems_, "integrals: pair outside store");
    z64* dst = stage_.get() + ((int64_t)buf * slots_ + i) * pair_elems_;
    if (mode_ == kResGenerated)
      synth_integrals(dst, pos, 1, diag_.get() + pos, NQ_, nao_, seed_, sigma_, copy_);
    else
      GGPU_HIP(hipMemcpyAsync(dst, host_ + pos * pair_elems_, bytes, hipMemcpyHostToDevice, copy_));
  }
  GGPU_HIP(hipEventRecord(ev_copy_[buf], copy_));
  if (!capturing_) {
    GGPU_HIP(hipEventRecord(tpool_[tused_ + 1], copy_));
    tused_ += 2;
  }
  (mode_ == kResGenerated ? generated_ : streamed_) += (double)bytes * n;
}

const z64* IntegralStore::view(const PairTable& tab, int i0, int n, hipStream_t s, const int64_t** offsets) {
  require(i0 >= 0 && i0 + n <= (int)tab.host.size(), "integrals: pair range");
  if (mode_ == kResDevice) {
    *offsets = tab.in_off.get() + i0;
    return V_.get();
  }
  require(cur_ < 0, "integrals: previous staged batch still in use");
  int buf;
  if (pre_.tab == &tab && pre_.i0 == i0 && pre_.n == n) {
    buf = pre_.buf;
  } else {
    buf = next_;
    issue(tab, i0, n, buf);
  }
  GGPU_HIP(hipStreamWaitEvent(s, ev_copy_[buf], 0));
  cur_ = buf;
  // prefetch the batch that usually follows into the other buffer
  const int j0 = i0 + n, total = (int)tab.host.size();
  if (j0 < total) {
    const int m = std::min(n, total - j0);
    issue(tab, j0, m, buf ^ 1);
    pre_ = Pending{&tab, j0, m, buf ^ 1};
  } else {
    pre_ = Pending{};
  }
  next_ = buf ^ 1;
  *offsets = stage_off_.get() + (int64_t)buf * slots_;
  return stage_.get();
}

double IntegralStore::collect_staging_ms() {
  double ms = 0;
  if (copy_) GGPU_HIP(hipStreamSynchronize(copy_));
  for (size_t i = 0; i + 1 < tused_; i += 2) {
    float t = 0;
    GGPU_HIP(hipEventElapsedTime(&t, tpool_[i], tpool_[i + 1]));
    ms += t;
  }
  tused_ = 0;
  return ms;
}

void IntegralStore::done(hipStream_t s) {
  if (mode_ == kResDevice || cur_ < 0) return;
  GGPU_HIP(hipEventRecord(ev_used_[cur_], s));
  if (capturing_) used_in_capture_[cur_] = true;
  cur_ = -1;
}

void IntegralStore::begin_capture(hipStream_t s) {
  if (mode_ == kResDevice) return;
  require(mode_ != kResReader, "integrals: graph capture cannot replay a pair reader", kErrState);
  require(cur_ < 0, "integrals: a staged batch is still in use", kErrState);
  GGPU_HIP(hipEventRecord(ev_join_[0], s));
  GGPU_HIP(hipStreamWaitEvent(copy_, ev_join_[0], 0));
  capturing_ = true;
  used_in_capture_[0] = used_in_capture_[1] = false;
  pre_ = Pending{};
  next_ = 0;
}

void IntegralStore::end_capture(hipStream_t s) {
  if (mode_ == kResDevice || !capturing_) return;
  capturing_ = false;
  GGPU_HIP(hipEventRecord(ev_join_[1], copy_));
  GGPU_HIP(hipStreamWaitEvent(s, ev_join_[1], 0));
  // a prefetched batch of the capture is not valid outside it
  pre_ = Pending{};
}

void PairTable::build(const std::vector<PairView>& views, int64_t stride, hipStream_t s) {
  (void)s;
  host = views;
  std::vector<int64_t> off(views.size());
  std::vector<int> tr(views.size()), cj(views.size());
  for (size_t i = 0; i < views.size(); ++i) {
    off[i] = views[i].pos * stride;
    tr[i] = views[i].trans;
    cj[i] = views[i].conj;
  }
  in_off.resize(off.size());
  trans.resize(tr.size());
  conj.resize(cj.size());
  GGPU_HIP(hipMemcpy(in_off.get(), off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  GGPU_HIP(hipMemcpy(trans.get(), tr.data(), tr.size() * sizeof(int), hipMemcpyHostToDevice));
  GGPU_HIP(hipMemcpy(conj.get(), cj.data(), cj.size() * sizeof(int), hipMemcpyHostToDevice));
}

void unfold_pairs(IntegralStore& st, const PairTable& tab, int i0, int n, const int64_t* out_off_dev,
                  void* dst, int out_f32, int NQ, int nao, hipStream_t s) {
  require(i0 >= 0 && i0 + n <= (int)tab.host.size(), "unfold_pairs: range");
  for (int i = i0; i < i0 + n; ++i)
    require((tab.host[i].pos + 1) * (int64_t)NQ * nao * nao <= st.elems(), "unfold_pairs: pair outside store");
  const int64_t* offs = nullptr;
  const z64* src = st.view(tab, i0, n, s, &offs);
  BlockCopy c;
  c.in = src;
  c.out = dst;
  c.out_f32 = out_f32;
  c.rows = nao;
  c.cols = nao;
  c.in_ld = nao;
  c.out_ld = nao;
  c.n_inner = NQ;
  c.in_bs = (int64_t)nao * nao;
  c.out_bs = (int64_t)nao * nao;
  c.n_outer = n;
  c.in_off = offs;
  c.out_off = out_off_dev;
  c.trans = tab.trans.get() + i0;
  c.conj = tab.conj.get() + i0;
  block_copy(c, s);
  st.done(s);
}

// ---------------------------------------------------------------------------
// IBZ -> full BZ rotation
// ---------------------------------------------------------------------------
void FullBZRotator::init(const SymTables& t, int n, int ns, int nt, hipStream_t s) {
  (void)s;
  t_ = &t;
  n_ = n; ns_ = ns; nt_ = nt;
  bool any_tr = false;
  for (auto x : t.k_tr) any_tr |= (x != 0);
  const int64_t n2 = (int64_t)n * n;
  if (t.k_identity() && !any_tr) {
    std::vector<int64_t> src((size_t)t.nk * ns * nt), dst(src.size());
    size_t o = 0;
    for (int k = 0; k < t.nk; ++k)
      for (int sp = 0; sp < ns; ++sp)
        for (int tt = 0; tt < nt; ++tt, ++o) {
          src[o] = (((int64_t)tt * ns + sp) * t.ink + t.k_f2r[k]) * n2;
          dst[o] = (((int64_t)k * ns + sp) * nt + tt) * n2;
        }
    src_.resize(src.size());
    dst_.resize(dst.size());
    GGPU_HIP(hipMemcpy(src_.get(), src.data(), src.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    GGPU_HIP(hipMemcpy(dst_.get(), dst.data(), dst.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  } else {
    std::vector<z64> U((size_t)t.nk * n2);
    if (t.k_identity()) {
      for (int k = 0; k < t.nk; ++k)
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) U[k * n2 + i * n + j] = z64{i == j ? 1.0 : 0.0, 0.0};
    } else {
      U = t.U_k;
    }
    Uk_.resize(U.size());
    GGPU_HIP(hipMemcpy(Uk_.get(), U.data(), U.size() * sizeof(z64), hipMemcpyHostToDevice));
    tmp_.resize((size_t)ns * nt * n2);
  }
}

void FullBZRotator::run(const z64* X, z64* out, hipStream_t s) {
  const SymTables& t = *t_;
  const int64_t n2 = (int64_t)n_ * n_;
  const size_t in_elems = (size_t)nt_ * ns_ * t.ink * n2, out_elems = (size_t)t.nk * ns_ * nt_ * n2;
  if (src_.size()) {
    slab_gather(X, src_.get(), out, dst_.get(), (int)src_.size(), n2, 1.0, s);
    return;
  }
  for (int k = 0; k < t.nk; ++k) {
    const int tr = t.k_tr[k] != 0;
    GemmDesc g;
    g.M = n_; g.N = n_; g.K = n_; g.nb0 = nt_; g.nb1 = ns_;
    g.A.ptr = Uk_.get() + k * n2; g.A.rs = n_; g.A.cs = 1; g.A.conj = tr;
    g.B.ptr = X + t.k_f2r[k] * n2; g.B.rs = n_; g.B.cs = 1;
    g.B.bs0 = (int64_t)ns_ * t.ink * n2; g.B.bs1 = (int64_t)t.ink * n2; g.B.conj = tr;
    g.C.ptr = tmp_.get(); g.C.rs = n_; g.C.cs = 1; g.C.bs0 = n2; g.C.bs1 = (int64_t)nt_ * n2;
    gemm_on(g, Uk_, Span{X, in_elems}, tmp_, s, kSiteRotate);
    GemmDesc h;
    h.M = n_; h.N = n_; h.K = n_; h.nb0 = nt_; h.nb1 = ns_;
    h.A.ptr = tmp_.get(); h.A.rs = n_; h.A.cs = 1; h.A.bs0 = n2; h.A.bs1 = (int64_t)nt_ * n2;
    h.B.ptr = Uk_.get() + k * n2; h.B.rs = 1; h.B.cs = n_; h.B.conj = !tr;
    h.C.ptr = out + (int64_t)k * ns_ * nt_ * n2; h.C.rs = n_; h.C.cs = 1; h.C.bs0 = n2; h.C.bs1 = (int64_t)nt_ * n2;
    gemm_on(h, tmp_, Uk_, Span{out, out_elems}, s, kSiteRotate);
  }
}

// ---------------------------------------------------------------------------
// Per-site GEMM timing
// ---------------------------------------------------------------------------
SiteProfiler::~SiteProfiler() {
  for (auto e : pool_) (void)hipEventDestroy(e);
}

void SiteProfiler::before(const GemmDesc& d, hipStream_t s) {
  while (pool_.size() < used_ + 2) {
    hipEvent_t e;
    GGPU_HIP(hipEventCreate(&e));
    pool_.push_back(e);
  }
  GGPU_HIP(hipEventRecord(pool_[used_], s));
  recs_.push_back(Rec{d.site, gemm_flops(d), used_});
  used_ += 2;
}

void SiteProfiler::after(const GemmDesc&, hipStream_t s) {
  GGPU_HIP(hipEventRecord(pool_[recs_.back().ev + 1], s));
}

void SiteProfiler::reset() {
  used_ = 0;
  recs_.clear();
  for (int i = 0; i < kNumSites; ++i) { ms[i] = 0; flops[i] = 0; launches[i] = 0; }
}

void SiteProfiler::collect() {
  for (const Rec& r : recs_) {
    float t = 0;
    GGPU_HIP(hipEventElapsedTime(&t, pool_[r.ev], pool_[r.ev + 1]));
    ms[r.site] += t;
    flops[r.site] += r.flops;
    launches[r.site] += 1;
  }
  recs_.clear();
  used_ = 0;
}

// ---------------------------------------------------------------------------
// Spinor <-> spin-block tables (two-component plans)
// ---------------------------------------------------------------------------
// Spin blocks in pseudo-spin order: (row block, column block) of aa, bb, ab, ba
// (x2c_gw_gpu_kernel::copy_Gk_2c, gw_gpu_kernel.cpp:513-523).
static const int kX2cBlock[4][2] = {{0, 0}, {1, 1}, {0, 1}, {1, 0}};

void SpinorTables::build(const std::vector<int64_t>& in, const std::vector<int64_t>& out, const std::vector<int>& tr,
                         const std::vector<int>& cj) {
  n = (int)in.size();
  h_in = in;
  h_out = out;
  in_off.resize(in.size());
  out_off.resize(out.size());
  trans.resize(tr.size());
  conj.resize(cj.size());
  GGPU_HIP(hipMemcpy(in_off.get(), in.data(), in.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  GGPU_HIP(hipMemcpy(out_off.get(), out.data(), out.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  GGPU_HIP(hipMemcpy(trans.get(), tr.data(), tr.size() * sizeof(int), hipMemcpyHostToDevice));
  GGPU_HIP(hipMemcpy(conj.get(), cj.data(), cj.size() * sizeof(int), hipMemcpyHostToDevice));
}

// Split: for every matrix m < count, block b < nblk of the spinor matrix
// in[m] (nso x nso, slab stride in_ms) goes to out[m * nblk + b] (nao x nao,
// slab stride out_ms; inner batch of n_inner slabs handled by the copy).
static void build_split(SpinorTables& t, int count, int nblk, int nao, int64_t in_ms, int64_t out_ms) {
  const int nso = 2 * nao;
  std::vector<int64_t> in, out;
  std::vector<int> zero;
  for (int m = 0; m < count; ++m)
    for (int b = 0; b < nblk; ++b) {
      in.push_back(m * in_ms + (int64_t)kX2cBlock[b][0] * nao * nso + (int64_t)kX2cBlock[b][1] * nao);
      out.push_back(((int64_t)m * nblk + b) * out_ms);
      zero.push_back(0);
    }
  t.build(in, out, zero, zero);
}

// Copies nao x nao blocks; in_elems / out_elems are the buffer sizes the
// offsets are checked against on the host before the launch.
static void spinor_copy(const SpinorTables& t, const z64* in, size_t in_elems, int64_t in_ld, int64_t in_bs,
                        z64* out, size_t out_elems, int64_t out_ld, int64_t out_bs, int nao, int n_inner,
                        double scale, hipStream_t s) {
  const int64_t span_in = (int64_t)(n_inner - 1) * in_bs + (int64_t)(nao - 1) * in_ld + nao;
  const int64_t span_out = (int64_t)(n_inner - 1) * out_bs + (int64_t)(nao - 1) * out_ld + nao;
  for (int i = 0; i < t.n; ++i)
    require(t.h_in[i] >= 0 && t.h_in[i] + span_in <= (int64_t)in_elems && t.h_out[i] >= 0 &&
                t.h_out[i] + span_out <= (int64_t)out_elems,
            "spinor block copy out of bounds");
  BlockCopy c;
  c.in = in; c.out = out;
  c.rows = nao; c.cols = nao;
  c.in_ld = in_ld; c.out_ld = out_ld;
  c.n_inner = n_inner; c.in_bs = in_bs; c.out_bs = out_bs;
  c.n_outer = t.n;
  c.in_off = t.in_off.get(); c.out_off = t.out_off.get();
  c.trans = t.trans.get(); c.conj = t.conj.get();
  c.scale = scale;
  block_copy(c, s);
}

// ---------------------------------------------------------------------------
// GW plan
// ---------------------------------------------------------------------------
GwPlan::GwPlan(int nao_, int nso_, int ns_, int NQ_, int nts_, int nw_b_, const ggpu_symmetry& s, const double* T_wt,
               const double* T_tw_h, int precision, int solver_, size_t work_bytes)
    : nao(nao_), nso(nso_), ns(ns_), NQ(NQ_), nts(nts_), nw_b(nw_b_), half(nts_ / 2), solver(solver_) {
  require(nao > 0 && NQ > 0, "gw: nao and NQ must be positive");
  require(nso == nao || nso == 2 * nao, "gw: nso must be nao or 2 nao");
  require(ns == 1 || ns == 2, "gw: ns must be 1 or 2");
  std::fill(variants, variants + kNumSites, -1);
  x2c = nso == 2 * nao;
  require(!x2c || ns == 1, "gw: a two-component (nso = 2 nao) plan needs ns = 1");
  if (x2c) ns = 4;   // pseudo-spins aa, bb, ab, ba
  require(nts % 2 == 0, "gw: nts must be even");
  require(nts >= 4, "gw: nts must be at least 4");
  require(nts >= nw_b, "gw: Nt is not greater than or equal to Nw_b");
  require(nw_b > 0, "gw: nw_b must be positive");
  require(solver == GGPU_SOLVER_CHOLESKY || solver == GGPU_SOLVER_LU, "gw: unknown linear solver");
  require(precision == GGPU_PRECISION_FP64 || precision == GGPU_PRECISION_FP32 ||
              precision == GGPU_PRECISION_BF16X3 || precision == GGPU_PRECISION_INT8X5 ||
              precision == GGPU_PRECISION_INT8X4,
          "gw: unknown precision");
  require(T_wt && T_tw_h, "gw: NULL transform matrices");
  compute = precision == GGPU_PRECISION_FP64     ? Compute::F64
            : precision == GGPU_PRECISION_FP32   ? Compute::F32
            : precision == GGPU_PRECISION_INT8X5 ? Compute::I8X5
            : precision == GGPU_PRECISION_INT8X4 ? Compute::I8X4
                                                 : Compute::BF16X3;
  sym.load(s, nso, NQ);
  if (x2c)
    for (auto t : sym.k_tr)
      require(t == 0, "gw: two-component plans do not support time-reversal k images");
  GGPU_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  for (auto& e : ev) GGPU_HIP(hipEventCreate(&e));

  // tau -> Omega with the mirror P0(beta - tau) = P0(tau) folded into the
  // matrix (the reference copies the mirrored slices, cugw_qpt.cu:184-191)
  const int ni = nts - 2;
  std::vector<z64> tf((size_t)nw_b * (half - 1));
  const z64* Tw = reinterpret_cast<const z64*>(T_wt);
  for (int w = 0; w < nw_b; ++w)
    for (int t = 1; t < half; ++t) {
      const z64 a = Tw[(int64_t)w * ni + (t - 1)], b = Tw[(int64_t)w * ni + (nts - 2 - t)];
      tf[(size_t)w * (half - 1) + (t - 1)] = z64{a.x + b.x, a.y + b.y};
    }
  Tf_wt.resize(tf.size());
  GGPU_HIP(hipMemcpy(Tf_wt.get(), tf.data(), tf.size() * sizeof(z64), hipMemcpyHostToDevice));
  T_tw.resize((size_t)nts * nw_b);
  GGPU_HIP(hipMemcpy(T_tw.get(), T_tw_h, T_tw.bytes(), hipMemcpyHostToDevice));

  const int64_t n2 = (int64_t)nao * nao, Q2 = (int64_t)NQ * NQ;
  G_host_dev.resize((size_t)nts * ns * sym.ink * n2);
  Gf.resize((size_t)sym.nk * ns * nts * n2);
  Grev.resize(Gf.size());
  if (lowp()) {
    Gf32.resize(Gf.size());
    Grev32.resize(Gf.size());
  }
  {
    std::vector<int64_t> src((size_t)sym.nk * ns * nts), dst(src.size());
    size_t o = 0;
    for (int k = 0; k < sym.nk; ++k)
      for (int sp = 0; sp < ns; ++sp)
        for (int t = 0; t < nts; ++t, ++o) {
          // X2C: the -tau factor of pseudo-spin ab is the ba block and vice
          // versa (copy_Gk_2c with minus_t, gw_gpu_kernel.cpp:517-521)
          const int sp_src = (x2c && sp >= 2) ? 5 - sp : sp;
          src[o] = (((int64_t)k * ns + sp_src) * nts + (nts - 1 - t)) * n2;
          dst[o] = (((int64_t)k * ns + sp) * nts + t) * n2;
        }
    rev_src.resize(src.size());
    rev_dst.resize(dst.size());
    GGPU_HIP(hipMemcpy(rev_src.get(), src.data(), src.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    GGPU_HIP(hipMemcpy(rev_dst.get(), dst.data(), dst.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  }
  Sig_acc.resize((size_t)sym.ink * ns * nts * n2);
  Sig_host_dev.resize(Sig_acc.size());
  P0.resize((size_t)p0_split * half * Q2);
  P0w.resize((size_t)nw_b * Q2);
  Amat.resize((size_t)nw_b * Q2);
  Pt.resize((size_t)nts * Q2);
  if (lowp()) Pt32.resize((size_t)nts * Q2);
  if (!sym.q_identity()) {
    Prot.resize((size_t)nts * Q2);
    Ptmp.resize((size_t)nts * Q2);
    Uq_dev.resize(sym.U_q.size());
    GGPU_HIP(hipMemcpy(Uq_dev.get(), sym.U_q.data(), Uq_dev.bytes(), hipMemcpyHostToDevice));
  }
  info.resize(nw_b);
  if (solver == GGPU_SOLVER_LU) ipiv.resize((size_t)nw_b * NQ);

  // pair tables
  const int64_t stride = (int64_t)NQ * n2;
  p0_tables.resize(sym.inq);
  sig_tables.resize(sym.inq);
  for (int iq = 0; iq < sym.inq; ++iq) {
    const int64_t q = sym.q_r2f[iq];
    std::vector<PairView> v;
    for (int ik = 0; ik < sym.ink; ++ik)
      for (auto k : sym.k_stars[ik]) v.push_back(sym.view(k, sym.k1_from_k2q[k * sym.nq + q]));
    p0_tables[iq].build(v, stride, stream);
    sig_tables[iq].resize(sym.q_stars[iq].size());
    for (size_t j = 0; j < sym.q_stars[iq].size(); ++j) {
      const int64_t qd = sym.q_stars[iq][j];
      std::vector<PairView> u;
      for (int ik = 0; ik < sym.ink; ++ik) {
        const int64_t k = sym.k_r2f[ik];
        u.push_back(sym.view(k, sym.k2_from_k1q[k * sym.nq + qd]));
      }
      sig_tables[iq][j].build(u, stride, stream);
    }
  }

  // workspace for pair batches
  const int64_t per_pair_p0 = (2 + 2LL * ns * half) * stride;
  const int64_t per_pair_sig = (1 + 2LL * ns * nts) * stride;
  const int64_t per_pair = std::max(per_pair_p0, per_pair_sig) * (int64_t)(lowp() ? sizeof(c32) : sizeof(z64));
  size_t budget = work_bytes;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    GGPU_HIP(hipMemGetInfo(&fr, &tot));
    budget = std::min<size_t>((size_t)(0.6 * fr), (size_t)48 << 30);
  }
  kb = (int)std::max<int64_t>(1, std::min<int64_t>(budget / per_pair, std::max(sym.nk, sym.ink)));
  kb = std::min(kb, 4096);
  work.resize(((size_t)kb * per_pair + sizeof(z64) - 1) / sizeof(z64));
  if (int8_slices(compute) || compute == Compute::BF16X3) {
    // int8 slices (or pre-split bf16 planes) of the operands of one
    // contraction launch (launches whose slices do not fit are split over
    // their batches)
    size_t fr = 0, tot = 0;
    GGPU_HIP(hipMemGetInfo(&fr, &tot));
    ozwork.resize(std::min<size_t>((size_t)24 << 30, fr / 4));
  }
  std::vector<int64_t> bo(kb);
  for (int b = 0; b < kb; ++b) bo[b] = b * stride;
  bidx_off.resize(kb);
  GGPU_HIP(hipMemcpy(bidx_off.get(), bo.data(), bo.size() * sizeof(int64_t), hipMemcpyHostToDevice));

  if (x2c) {
    // G: rotated spinor slabs (k, t) -> pseudo-spin blocks (k, ss, t);
    // Sigma: pseudo-spin blocks (ik, ss, t) -> spinor layout (t, ik)
    const int64_t so2 = (int64_t)nso * nso;
    build_split(x2c_split, sym.nk, 4, nao, (int64_t)nts * so2, (int64_t)nts * n2);
    std::vector<int64_t> in, out;
    std::vector<int> z;
    for (int ik = 0; ik < sym.ink; ++ik)
      for (int b = 0; b < 4; ++b) {
        in.push_back(((int64_t)ik * 4 + b) * nts * n2);
        out.push_back(ik * so2 + (int64_t)kX2cBlock[b][0] * nao * nso + (int64_t)kX2cBlock[b][1] * nao);
        z.push_back(0);
      }
    x2c_merge.build(in, out, z, z);
    rot.init(sym, nso, 1, nts, stream);
    return;
  }
  // Sigma (ik, s, t) -> host layout (t, s, ik)
  std::vector<int64_t> src((size_t)sym.ink * ns * nts), dst(src.size());
  size_t o = 0;
  for (int ik = 0; ik < sym.ink; ++ik)
    for (int sp = 0; sp < ns; ++sp)
      for (int t = 0; t < nts; ++t, ++o) {
        src[o] = (((int64_t)ik * ns + sp) * nts + t) * n2;
        dst[o] = (((int64_t)t * ns + sp) * sym.ink + ik) * n2;
      }
  sig_src.resize(src.size());
  sig_dst.resize(dst.size());
  GGPU_HIP(hipMemcpy(sig_src.get(), src.data(), src.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  GGPU_HIP(hipMemcpy(sig_dst.get(), dst.data(), dst.size() * sizeof(int64_t), hipMemcpyHostToDevice));

  rot.init(sym, nao, ns, nts, stream);
}

GwPlan::~GwPlan() {
  if (stream) (void)hipStreamSynchronize(stream);
  drop_graph();
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (auto& e : qev) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

void GwPlan::stage_G(const z64* G_dev) {
  if (x2c) {
    // spinor G on the full BZ into Grev (used as scratch), then its blocks
    rot.run(G_dev, Grev.get(), stream);
    spinor_copy(x2c_split, Grev.get(), Grev.size(), nso, (int64_t)nso * nso, Gf.get(), Gf.size(), nao,
                (int64_t)nao * nao, nao, nts, 1.0, stream);
  } else {
    rot.run(G_dev, Gf.get(), stream);
  }
  slab_gather(Gf.get(), rev_src.get(), Grev.get(), rev_dst.get(), (int)rev_src.size(), (int64_t)nao * nao, 1.0,
              stream);
  if (lowp()) {
    to_c32(Gf.get(), Gf32.get(), (int64_t)Gf.size(), stream);
    to_c32(Grev.get(), Grev32.get(), (int64_t)Grev.size(), stream);
  }
}

// P0(q, tau_t)[Q'][Q] += pref * Tr[G(k, beta - tau_t) V^{Q'} G(k+q, tau_t) V^{Q dagger}]
// for t < nts/2 (cugw_qkpt.cu:202-246), as three GEMMs per pair batch:
//   X1[t][n][m][Q'] = sum_p G(k, nts-1-t)[n][p] V^{Q'}[p][m]
//   X2[t][Q][p][i]  = sum_m G(k+q, t)[i][m] conj(V^Q[p][m])
//   P0[t][Q'][Q]   += pref * sum_{b, s, (n, m)} X1[t][(n,m)][Q'] X2[t][Q][(n,m)]
void GwPlan::p0_for_q(int iq) {
  const int64_t n2 = (int64_t)nao * nao, stride = (int64_t)NQ * n2;
  const PairTable& tab = p0_tables[iq];
  const int64_t q = sym.q_r2f[iq];
  const int npairs = (int)tab.host.size();
  const double pref = (ns == 1) ? -2.0 : -1.0;
  const bool lp = lowp();
  const int64_t es = lp ? (int64_t)sizeof(c32) : (int64_t)sizeof(z64);
  zero(P0.get(), (int64_t)P0.size(), stream);
  // full-BZ k of every pair, in table order
  std::vector<int64_t> ks;
  for (int ik = 0; ik < sym.ink; ++ik)
    for (auto k : sym.k_stars[ik]) ks.push_back(k);
  char* Vdir = reinterpret_cast<char*>(work.get());
  char* VpmQ = Vdir + kb * stride * es;
  char* X1 = VpmQ + kb * stride * es;
  char* X2 = X1 + (int64_t)kb * ns * half * stride * es;
  const int64_t xs = (int64_t)half * stride;  // one (pair, spin) slab of X1 / X2 (elements)
  const Span wk{work.get(), work.bytes() / (size_t)es};
  const void* Gr = lp ? (const void*)Grev32.get() : (const void*)Grev.get();
  const void* Gp = lp ? (const void*)Gf32.get() : (const void*)Gf.get();
  const Span gsp{Gp, Gf.size()}, grsp{Gr, Grev.size()};
  const int p0_end = batch_limit > 0 ? std::min(npairs, batch_limit * kb) : npairs;
  for (int i0 = 0; i0 < p0_end; i0 += kb) {
    const int n = std::min(kb, npairs - i0);
    unfold_pairs(store, tab, i0, n, bidx_off.get(), Vdir, lp, NQ, nao, stream);
    BlockCopy tr;
    tr.in = Vdir; tr.out = VpmQ; tr.in_f32 = lp; tr.out_f32 = lp;
    tr.rows = (int)n2; tr.cols = NQ; tr.in_ld = n2; tr.out_ld = NQ;
    tr.n_inner = n; tr.in_bs = stride; tr.out_bs = stride; tr.trans_all = 1;
    block_copy(tr, stream);
    for (int b = 0; b < n; ++b) {
      const int64_t k = ks[i0 + b], k1 = sym.k1_from_k2q[k * sym.nq + q];
      for (int s = 0; s < ns; ++s) {
        // rows (t, n) of the tau-reversed G: one tall GEMM for all t < nts/2
        GemmDesc g;
        g.compute = compute; g.out_f32 = lp;
        g.M = half * nao; g.N = nao * NQ; g.K = nao;
        g.A.ptr = (const char*)Gr + (k * ns + s) * nts * n2 * es; g.A.rs = nao; g.A.cs = 1;
        g.B.ptr = VpmQ + b * stride * es; g.B.rs = (int64_t)nao * NQ; g.B.cs = 1;
        g.C.ptr = X1 + (b * ns + s) * xs * es; g.C.rs = (int64_t)nao * NQ; g.C.cs = 1;
        g.site = kSiteP0X1;
        gemm_on(g, grsp, wk, wk, stream);
        GemmDesc h;
        h.compute = compute; h.out_f32 = lp;
        h.M = NQ * nao; h.N = nao; h.K = nao; h.nb0 = half;
        h.A.ptr = Vdir + b * stride * es; h.A.rs = nao; h.A.cs = 1; h.A.conj = 1;
        h.B.ptr = (const char*)Gp + (k1 * ns + s) * nts * n2 * es; h.B.rs = 1; h.B.cs = nao; h.B.bs0 = n2;
        h.C.ptr = X2 + (b * ns + s) * xs * es; h.C.rs = nao; h.C.cs = 1; h.C.bs0 = stride;
        h.site = kSiteP0X2;
        gemm_on(h, wk, gsp, wk, stream);
      }
    }
    // P0 partial slice j accumulates (pair, spin) entries j, j + S, j + 2S, ...
    // (FP64 partials also in the reduced-precision modes)
    const int E = n * ns, S = p0_split, full = E / S, rem = E - full * S;
    const int64_t Q2 = (int64_t)NQ * NQ;
    for (int part = 0; part < 2; ++part) {
      const int slices = part == 0 ? (full > 0 ? S : 0) : rem;
      if (slices == 0) continue;
      GemmDesc p;
      p.compute = compute; p.out_f32 = 0;
      p.M = NQ; p.N = NQ; p.K = (int)n2; p.KB = part == 0 ? full : 1; p.nb0 = half; p.nb1 = slices;
      const int64_t first = part == 0 ? 0 : (int64_t)full * S * xs;
      p.A.ptr = X1 + first * es; p.A.rs = 1; p.A.cs = NQ; p.A.kbs = S * xs; p.A.bs0 = stride; p.A.bs1 = xs;
      p.B.ptr = X2 + first * es; p.B.rs = 1; p.B.cs = n2; p.B.kbs = S * xs; p.B.bs0 = stride; p.B.bs1 = xs;
      p.C.ptr = P0.get(); p.C.rs = NQ; p.C.cs = 1; p.C.bs0 = Q2; p.C.bs1 = (int64_t)half * Q2;
      p.alpha_re = pref; p.beta_re = 1.0;
      p.site = kSiteP0;
      gemm_on(p, wk, wk, P0, stream);
    }
  }
}

// tau -> Omega, Dyson solve, Omega -> tau for the current q (cugw_qpt.cu:174-358)
void GwPlan::dyson_for_q() {
  const int64_t Q2 = (int64_t)NQ * NQ;
  const int n = NQ;
  {
    GemmDesc g;
    g.M = nw_b; g.N = (int)Q2; g.K = half - 1; g.KB = p0_split;
    g.A.ptr = Tf_wt.get(); g.A.rs = half - 1; g.A.cs = 1; g.A.kbs = 0;
    g.B.ptr = P0.get() + Q2; g.B.rs = Q2; g.B.cs = 1; g.B.kbs = (int64_t)half * Q2;
    g.C.ptr = P0w.get(); g.C.rs = Q2; g.C.cs = 1;
    g.alpha_re = 1.0 / sym.nk;
    g.site = kSiteIR;
    gemm_on(g, Tf_wt, P0, P0w, stream);
  }
  herm_one_minus(P0w.get(), Amat.get(), n, nw_b, stream);
  if (solver == GGPU_SOLVER_LU) {
    lu_solve_dyson();
  } else {
    cholesky_solve_dyson();
  }
  {
    GemmDesc g;
    g.M = nts; g.N = (int)Q2; g.K = nw_b;
    g.A.ptr = T_tw.get(); g.A.rs = nw_b; g.A.cs = 1;
    g.B.ptr = P0w.get(); g.B.rs = Q2; g.B.cs = 1;
    g.C.ptr = Pt.get(); g.C.rs = Q2; g.C.cs = 1;
    g.site = kSiteIR;
    gemm_on(g, T_tw, P0w, Pt, stream);
  }
  if (lowp()) to_c32(Pt.get(), Pt32.get(), (int64_t)Pt.size(), stream);
}

// (1 - P0) X = P0 by a batched blocked Cholesky factorisation (potrf /
// potrs semantics, cugw_qpt.cu:223-264): right-looking, 32-column blocks
// factored in LDS, trailing updates on the matrix cores; X overwrites P0w.
void GwPlan::cholesky_solve_dyson() {
  const int64_t Q2 = (int64_t)NQ * NQ;
  const int n = NQ;
  const int nb = 32;
  for (int j0 = 0; j0 < n; j0 += nb) {
    const int jb = std::min(nb, n - j0);
    chol_diag_block(Amat.get(), n, nw_b, j0, jb, info.get(), stream);
    chol_panel(Amat.get(), n, nw_b, j0, jb, stream);
    const int rem = n - j0 - jb;
    if (rem > 0) {
      GemmDesc h;
      h.M = rem; h.N = rem; h.K = jb; h.nb0 = nw_b;
      h.A.ptr = Amat.get() + (int64_t)(j0 + jb) * n + j0; h.A.rs = n; h.A.cs = 1; h.A.bs0 = Q2;
      h.B.ptr = Amat.get() + (int64_t)(j0 + jb) * n + j0; h.B.rs = 1; h.B.cs = n; h.B.bs0 = Q2; h.B.conj = 1;
      h.C.ptr = Amat.get() + (int64_t)(j0 + jb) * n + (j0 + jb); h.C.rs = n; h.C.cs = 1; h.C.bs0 = Q2;
      h.alpha_re = -1.0; h.beta_re = 1.0;
      h.site = kSiteDyson;
      gemm_on(h, Amat, Amat, Amat, stream);
    }
  }
  // solve L Y = P0 (forward), L^dagger X = Y (backward); X overwrites P0w
  for (int i0 = 0; i0 < n; i0 += nb) {
    const int ib = std::min(nb, n - i0);
    if (i0 > 0) {
      GemmDesc g;
      g.M = ib; g.N = n; g.K = i0; g.nb0 = nw_b;
      g.A.ptr = Amat.get() + (int64_t)i0 * n; g.A.rs = n; g.A.cs = 1; g.A.bs0 = Q2;
      g.B.ptr = P0w.get(); g.B.rs = n; g.B.cs = 1; g.B.bs0 = Q2;
      g.C.ptr = P0w.get() + (int64_t)i0 * n; g.C.rs = n; g.C.cs = 1; g.C.bs0 = Q2;
      g.alpha_re = -1.0; g.beta_re = 1.0;
      g.site = kSiteDyson;
      gemm_on(g, Amat, P0w, P0w, stream);
    }
    trsm_lower_block(Amat.get(), P0w.get(), n, n, nw_b, i0, ib, stream);
  }
  const int nblk = (n + nb - 1) / nb;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int i0 = bi * nb, ib = std::min(nb, n - i0);
    const int rest = n - i0 - ib;
    if (rest > 0) {
      GemmDesc g;
      g.M = ib; g.N = n; g.K = rest; g.nb0 = nw_b;
      g.A.ptr = Amat.get() + (int64_t)(i0 + ib) * n + i0; g.A.rs = 1; g.A.cs = n; g.A.bs0 = Q2; g.A.conj = 1;
      g.B.ptr = P0w.get() + (int64_t)(i0 + ib) * n; g.B.rs = n; g.B.cs = 1; g.B.bs0 = Q2;
      g.C.ptr = P0w.get() + (int64_t)i0 * n; g.C.rs = n; g.C.cs = 1; g.C.bs0 = Q2;
      g.alpha_re = -1.0; g.beta_re = 1.0;
      g.site = kSiteDyson;
      gemm_on(g, Amat, P0w, P0w, stream);
    }
    trsm_upper_h_block(Amat.get(), P0w.get(), n, n, nw_b, i0, ib, stream);
  }
}

// (1 - P0) X = P0 by a batched blocked LU factorisation with partial
// pivoting (getrf / getrs semantics, cugw_qpt.cu:266-322): right-looking,
// 32-column panels factored with row interchanges, the interchanges applied
// to the columns left and right of the panel, U12 = L11^-1 A12, and the
// trailing update on the matrix cores; then the pivots are applied to the
// right-hand side, L (unit lower) and U solved blockwise, and X is
// Hermitised again (cu_compute_Pq.cu:44-53).  X overwrites P0w.
void GwPlan::lu_solve_dyson() {
  const int64_t Q2 = (int64_t)NQ * NQ;
  const int n = NQ;
  const int nb = 32;
  for (int j0 = 0; j0 < n; j0 += nb) {
    const int jb = std::min(nb, n - j0);
    lu_panel(Amat.get(), n, nw_b, j0, jb, ipiv.get(), info.get(), stream);
    lu_row_swaps(Amat.get(), n, n, nw_b, ipiv.get(), n, j0, j0 + jb, j0, j0 + jb, stream);
    const int rem = n - j0 - jb;
    if (rem > 0) {
      trsm_unit_lower_cols(Amat.get(), Amat.get(), n, nw_b, j0, jb, j0 + jb, rem, stream);
      GemmDesc h;   // A22 -= L21 U12
      h.M = rem; h.N = rem; h.K = jb; h.nb0 = nw_b;
      h.A.ptr = Amat.get() + (int64_t)(j0 + jb) * n + j0; h.A.rs = n; h.A.cs = 1; h.A.bs0 = Q2;
      h.B.ptr = Amat.get() + (int64_t)j0 * n + (j0 + jb); h.B.rs = n; h.B.cs = 1; h.B.bs0 = Q2;
      h.C.ptr = Amat.get() + (int64_t)(j0 + jb) * n + (j0 + jb); h.C.rs = n; h.C.cs = 1; h.C.bs0 = Q2;
      h.alpha_re = -1.0; h.beta_re = 1.0;
      h.site = kSiteDyson;
      gemm_on(h, Amat, Amat, Amat, stream);
    }
  }
  // P X = L U X = B: interchange the rows of B, then L Y = P B, U X = Y
  lu_row_swaps(P0w.get(), n, n, nw_b, ipiv.get(), n, 0, n, 0, 0, stream);
  for (int i0 = 0; i0 < n; i0 += nb) {
    const int ib = std::min(nb, n - i0);
    if (i0 > 0) {
      GemmDesc g;
      g.M = ib; g.N = n; g.K = i0; g.nb0 = nw_b;
      g.A.ptr = Amat.get() + (int64_t)i0 * n; g.A.rs = n; g.A.cs = 1; g.A.bs0 = Q2;
      g.B.ptr = P0w.get(); g.B.rs = n; g.B.cs = 1; g.B.bs0 = Q2;
      g.C.ptr = P0w.get() + (int64_t)i0 * n; g.C.rs = n; g.C.cs = 1; g.C.bs0 = Q2;
      g.alpha_re = -1.0; g.beta_re = 1.0;
      g.site = kSiteDyson;
      gemm_on(g, Amat, P0w, P0w, stream);
    }
    trsm_unit_lower_cols(Amat.get(), P0w.get(), n, nw_b, i0, ib, 0, n, stream);
  }
  const int nblk = (n + nb - 1) / nb;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int i0 = bi * nb, ib = std::min(nb, n - i0);
    const int rest = n - i0 - ib;
    if (rest > 0) {
      GemmDesc g;
      g.M = ib; g.N = n; g.K = rest; g.nb0 = nw_b;
      g.A.ptr = Amat.get() + (int64_t)i0 * n + (i0 + ib); g.A.rs = n; g.A.cs = 1; g.A.bs0 = Q2;
      g.B.ptr = P0w.get() + (int64_t)(i0 + ib) * n; g.B.rs = n; g.B.cs = 1; g.B.bs0 = Q2;
      g.C.ptr = P0w.get() + (int64_t)i0 * n; g.C.rs = n; g.C.cs = 1; g.C.bs0 = Q2;
      g.alpha_re = -1.0; g.beta_re = 1.0;
      g.site = kSiteDyson;
      gemm_on(g, Amat, P0w, P0w, stream);
    }
    trsm_upper_cols(Amat.get(), P0w.get(), n, nw_b, i0, ib, 0, n, stream);
  }
  herm_one_minus(P0w.get(), nullptr, n, nw_b, stream);
}

// Sigma(k, tau) -= sum_{Q,Q'} P~(q, tau)[Q'][Q] V^{Q'}(k,k-q) G(k-q, tau) V^{Q dagger}(k,k-q)
// for every q in the star of the IBZ q and every IBZ k (cugw_qkpt.cu:248-338):
//   Y1[t][Q'][i][n] = sum_m V^{Q'}[i][m] G(k-q, t)[m][n]
//   Y2[Q][t][i][n]  = sum_Q' P~[t][Q'][Q] Y1[t][Q'][i][n]
//   Sigma[t][i][j] -= sum_{Q,m} Y2[Q][t][i][m] conj(V^Q[j][m])
void GwPlan::sigma_for_q(int iq) {
  const int64_t n2 = (int64_t)nao * nao, stride = (int64_t)NQ * n2, Q2 = (int64_t)NQ * NQ;
  const bool lp = lowp();
  const int64_t es = lp ? (int64_t)sizeof(c32) : (int64_t)sizeof(z64);
  char* Vdir = reinterpret_cast<char*>(work.get());
  char* Y1 = Vdir + kb * stride * es;
  char* Y2 = Y1 + (int64_t)kb * ns * nts * stride * es;
  const int64_t ys = (int64_t)nts * stride;   // one (pair, spin) slab of Y1 / Y2 (elements)
  const Span wk{work.get(), work.bytes() / (size_t)es};
  const void* Gp = lp ? (const void*)Gf32.get() : (const void*)Gf.get();
  const Span gsp{Gp, Gf.size()};
  for (size_t j = 0; j < sym.q_stars[iq].size(); ++j) {
    const int64_t qd = sym.q_stars[iq][j];
    const int tr = sym.q_tr[qd] != 0;
    const void* Pq = lp ? (const void*)Pt32.get() : (const void*)Pt.get();
    if (!sym.q_identity()) {
      // P~ = U_q P U_q^dagger in FP64 (conjugated at use for a time-reversal image)
      GemmDesc a;
      a.M = NQ; a.N = NQ; a.K = NQ; a.nb0 = nts;
      a.A.ptr = Uq_dev.get() + qd * Q2; a.A.rs = NQ; a.A.cs = 1;
      a.B.ptr = Pt.get(); a.B.rs = NQ; a.B.cs = 1; a.B.bs0 = Q2;
      a.C.ptr = Ptmp.get(); a.C.rs = NQ; a.C.cs = 1; a.C.bs0 = Q2;
      a.site = kSiteRotate;
      gemm_on(a, Uq_dev, Pt, Ptmp, stream);
      GemmDesc c;
      c.M = NQ; c.N = NQ; c.K = NQ; c.nb0 = nts;
      c.A.ptr = Ptmp.get(); c.A.rs = NQ; c.A.cs = 1; c.A.bs0 = Q2;
      c.B.ptr = Uq_dev.get() + qd * Q2; c.B.rs = 1; c.B.cs = NQ; c.B.conj = 1;
      c.C.ptr = Prot.get(); c.C.rs = NQ; c.C.cs = 1; c.C.bs0 = Q2;
      c.site = kSiteRotate;
      gemm_on(c, Ptmp, Uq_dev, Prot, stream);
      if (lp) {
        to_c32(Prot.get(), Pt32.get(), (int64_t)Prot.size(), stream);
        Pq = Pt32.get();
      } else {
        Pq = Prot.get();
      }
    }
    const Span psp{Pq, Pt.size()};
    const PairTable& tab = sig_tables[iq][j];
    const int sig_end = batch_limit > 0 ? std::min(sym.ink, batch_limit * kb) : sym.ink;
    for (int i0 = 0; i0 < sig_end; i0 += kb) {
      const int n = std::min(kb, sym.ink - i0);
      unfold_pairs(store, tab, i0, n, bidx_off.get(), Vdir, lp, NQ, nao, stream);
      for (int b = 0; b < n; ++b) {
        const int64_t k = sym.k_r2f[i0 + b];
        const int64_t k1 = sym.k2_from_k1q[k * sym.nq + qd];
        for (int s = 0; s < ns; ++s) {
          GemmDesc g;
          g.compute = compute; g.out_f32 = lp;
          g.M = NQ * nao; g.N = nao; g.K = nao; g.nb0 = nts;
          g.A.ptr = Vdir + b * stride * es; g.A.rs = nao; g.A.cs = 1;
          g.B.ptr = (const char*)Gp + (k1 * ns + s) * nts * n2 * es; g.B.rs = nao; g.B.cs = 1; g.B.bs0 = n2;
          g.C.ptr = Y1 + (b * ns + s) * ys * es; g.C.rs = nao; g.C.cs = 1; g.C.bs0 = stride;
          g.site = kSiteSigY1;
          gemm_on(g, wk, gsp, wk, stream);
          GemmDesc h;
          h.compute = compute; h.out_f32 = lp;
          h.M = NQ; h.N = (int)n2; h.K = NQ; h.nb0 = nts;
          h.A.ptr = Pq; h.A.rs = 1; h.A.cs = NQ; h.A.bs0 = Q2; h.A.conj = tr;
          h.B.ptr = Y1 + (b * ns + s) * ys * es; h.B.rs = n2; h.B.cs = 1; h.B.bs0 = stride;
          h.C.ptr = Y2 + (b * ns + s) * ys * es; h.C.rs = (int64_t)nts * n2; h.C.cs = 1; h.C.bs0 = n2;
          h.site = kSiteSigY2;
          gemm_on(h, psp, wk, wk, stream);
        }
      }
      GemmDesc f;
      f.compute = compute; f.out_f32 = 0;       // Sigma accumulates in FP64
      f.M = nts * nao; f.N = nao; f.K = nao; f.KB = NQ; f.nb0 = n; f.nb1 = ns;
      f.A.ptr = Y2; f.A.rs = nao; f.A.cs = 1; f.A.kbs = (int64_t)nts * n2; f.A.bs0 = ns * ys; f.A.bs1 = ys;
      f.B.ptr = Vdir; f.B.rs = 1; f.B.cs = nao; f.B.kbs = n2; f.B.bs0 = stride; f.B.bs1 = 0; f.B.conj = 1;
      f.C.ptr = Sig_acc.get() + (int64_t)i0 * ns * nts * n2; f.C.rs = nao; f.C.cs = 1;
      f.C.bs0 = (int64_t)ns * nts * n2; f.C.bs1 = (int64_t)nts * n2;
      f.alpha_re = -1.0; f.beta_re = 1.0;
      f.site = kSiteSigma;
      gemm_on(f, wk, wk, Sig_acc, stream);
    }
  }
}

void GwPlan::check_info() {
  std::vector<int> h(nw_b);
  GGPU_HIP(hipMemcpy(h.data(), info.get(), nw_b * sizeof(int), hipMemcpyDeviceToHost));
  for (int w = 0; w < nw_b; ++w)
    if (h[w] != 0) {
      if (solver == GGPU_SOLVER_LU)
        throw Error(kErrSingular, "gw: 1 - P0(Omega) is singular (Omega index " + std::to_string(w) +
                                      ", zero pivot in column " + std::to_string(h[w]) + ")");
      throw Error(kErrNotPositiveDefinite, "gw: 1 - P0(Omega) is not positive definite (Omega index " +
                                               std::to_string(w) + ", column " + std::to_string(h[w]) + ")");
    }
}

// Install a GEMM profiler / per-site launch variants for the calling thread
// while a plan runs.
struct ProfilerGuard {
  explicit ProfilerGuard(GemmProfiler* p) { set_gemm_profiler(p); }
  ~ProfilerGuard() { set_gemm_profiler(nullptr); }
};
struct VariantGuard {
  explicit VariantGuard(const int* v, DeviceBuffer<char>* ws = nullptr) {
    set_site_variants(v);
    if (ws) set_gemm_workspace(ws->get(), ws->size());
  }
  ~VariantGuard() {
    set_site_variants(nullptr);
    set_gemm_workspace(nullptr, 0);
  }
};

// Sites whose launch variant the autotuner picks: the six contraction sites,
// all in the plan's arithmetic (IR, Dyson and rotations stay FP64).
static bool tunable_site(int s) { return s >= kSiteP0X1 && s <= kSiteSigma; }

void GwPlan::set_variants(const int* v) {
  require(v != nullptr, "gw: NULL variant table");
  for (int s = 0; s < kNumSites; ++s) {
    require(v[s] == -1 || (tunable_site(s) && gemm_variant_valid(compute, v[s])),
            "gw: launch variant " + std::to_string(v[s]) + " not available for call site " + std::to_string(s) +
                " in this plan's precision");
  }
  std::copy(v, v + kNumSites, variants);
  invalidate_graph();
}

void GwPlan::autotune(const z64* G_dev, int rounds, int batches) {
  require(store.loaded(), "gw: integrals not loaded", kErrState);
  require(G_dev != nullptr, "gw: NULL G");
  require(rounds >= 1 && rounds <= 8, "gw: autotune rounds must be 1 .. 8");
  require(batches >= 1, "gw: autotune needs at least one pair batch");
  int cand[16];
  const int nc = gemm_variant_candidates(compute, cand, 16);
  tune_cand.assign(cand, cand + nc);
  tune_ms.assign(nc, std::vector<double>(kNumSites, -1.0));
  tune_wall_ms = 0;
  invalidate_graph();
  const int iq = share_rank;       // first q of this device's share
  if (iq >= sym.inq) return;
  hipEvent_t e0 = ev[0], e1 = ev[6];
  GGPU_HIP(hipEventRecord(e0, stream));
  SiteProfiler tp;
  ProfilerGuard guard(&tp);
  struct Limit {
    int& r;
    int saved;
    Limit(int& x, int v) : r(x), saved(x) { r = v; }
    ~Limit() { r = saved; }
  } limit(batch_limit, batches);
  stage_G(G_dev);
  int trial[kNumSites];
  // the int8 emulation is a candidate only where int8_site_tunable allows it
  auto tried = [&](int c, int s) {
    const bool i8 = cand[c] == kVariantInt8 || cand[c] == kVariantInt8x5;
    return tunable_site(s) && (!i8 || int8_site_tunable(compute, s, cand[c]));
  };
  for (int r = 0; r < rounds; ++r)
    for (int c = 0; c < nc; ++c) {
      for (int s = 0; s < kNumSites; ++s) trial[s] = tried(c, s) ? cand[c] : -1;
      VariantGuard vg(trial, &ozwork);
      tp.reset();
      p0_for_q(iq);
      sigma_for_q(iq);
      GGPU_HIP(hipStreamSynchronize(stream));
      tp.collect();
      for (int s = 0; s < kNumSites; ++s) {
        if (!tried(c, s) || tp.launches[s] == 0) continue;
        double& best = tune_ms[c][s];
        best = best < 0 ? tp.ms[s] : std::min(best, tp.ms[s]);
      }
    }
  GGPU_HIP(hipEventRecord(e1, stream));
  GGPU_HIP(hipStreamSynchronize(stream));
  float w = 0;
  GGPU_HIP(hipEventElapsedTime(&w, e0, e1));
  tune_wall_ms = w;
  (void)store.collect_staging_ms();
  for (int s = 0; s < kNumSites; ++s) {
    variants[s] = -1;
    if (!tunable_site(s)) continue;
    double best = -1;
    for (int c = 0; c < nc; ++c)
      if (tune_ms[c][s] >= 0 && (best < 0 || tune_ms[c][s] < best)) { best = tune_ms[c][s]; variants[s] = cand[c]; }
  }
}

void GwPlan::enqueue_iteration(const z64* G_dev, z64* S_dev, bool timed) {
  const int64_t n2 = (int64_t)nao * nao;
  if (timed) GGPU_HIP(hipEventRecord(ev[0], stream));
  stage_G(G_dev);
  if (timed) GGPU_HIP(hipEventRecord(ev[1], stream));
  zero(Sig_acc.get(), Sig_acc.size(), stream);
  zero_int(info.get(), nw_b, stream);
  for (int iq = share_rank; iq < sym.inq; iq += share_size) {
    if (timed) GGPU_HIP(hipEventRecord(qev[4 * iq + 0], stream));
    last_q = iq;
    p0_for_q(iq);
    if (timed) GGPU_HIP(hipEventRecord(qev[4 * iq + 1], stream));
    dyson_for_q();
    if (timed) GGPU_HIP(hipEventRecord(qev[4 * iq + 2], stream));
    sigma_for_q(iq);
    if (timed) GGPU_HIP(hipEventRecord(qev[4 * iq + 3], stream));
  }
  if (timed) GGPU_HIP(hipEventRecord(ev[5], stream));
  if (x2c)
    spinor_copy(x2c_merge, Sig_acc.get(), Sig_acc.size(), nao, n2, S_dev, Sig_host_dev.size(), nso,
                (int64_t)sym.ink * nso * nso, nao, nts, 1.0 / sym.nk, stream);
  else
    slab_gather(Sig_acc.get(), sig_src.get(), S_dev, sig_dst.get(), (int)sig_src.size(), n2, 1.0 / sym.nk, stream);
  if (timed) GGPU_HIP(hipEventRecord(ev[6], stream));
}

void GwPlan::drop_graph() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  gexec = nullptr;
  graph_G = graph_S = nullptr;
}

void GwPlan::invalidate_graph() {
  drop_graph();
  ++config_gen;
}

// Record the launches of one iteration into a graph (nothing runs) and
// instantiate it.  Every buffer is zeroed by a kernel, not a memset node.
void GwPlan::capture(const z64* G_dev, z64* S_dev) {
  drop_graph();
  GGPU_HIP(hipStreamSynchronize(stream));
  GGPU_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    store.begin_capture(stream);
    enqueue_iteration(G_dev, S_dev, false);
    store.end_capture(stream);
  } catch (...) {
    store.abort_capture();
    (void)hipStreamEndCapture(stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  GGPU_HIP(hipStreamEndCapture(stream, &graph));
  const hipError_t e = hipGraphInstantiateWithFlags(&gexec, graph, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    gexec = nullptr;
    GGPU_HIP(e);
  }
  graph_G = G_dev;
  graph_S = S_dev;
  ++graph_captures;
}

void GwPlan::iterate_device(const z64* G_dev, z64* S_dev) {
  require(store.loaded(), "gw: integrals not loaded", kErrState);
  while (qev.size() < (size_t)4 * sym.inq) {
    hipEvent_t e;
    GGPU_HIP(hipEventCreate(&e));
    qev.push_back(e);
  }
  if (use_graph) {
    require(!store.reader(), "gw: graph replay cannot replay a pair reader", kErrState);
    require(!profiling, "gw: per-site profiling does not apply to graph replay", kErrState);
    if (!gexec || graph_G != G_dev || graph_S != S_dev) {
      VariantGuard vguard(variants, &ozwork);
      capture(G_dev, S_dev);
    }
    GGPU_HIP(hipEventRecord(ev[0], stream));
    GGPU_HIP(hipGraphLaunch(gexec, stream));
    GGPU_HIP(hipEventRecord(ev[6], stream));
    GGPU_HIP(hipStreamSynchronize(stream));
    float tot = 0;
    GGPU_HIP(hipEventElapsedTime(&tot, ev[0], ev[6]));
    last_ms[0] = tot;
    for (int i = 1; i < 6; ++i) last_ms[i] = 0;
    streamed_last = generated_last = staging_ms_last = 0;
    check_info();
    return;
  }
  ProfilerGuard guard(profiling ? &prof : nullptr);
  VariantGuard vguard(variants, &ozwork);
  if (profiling) prof.reset();
  const double streamed0 = store.bytes_streamed(), generated0 = store.bytes_generated();
  (void)store.collect_staging_ms();
  enqueue_iteration(G_dev, S_dev, true);
  GGPU_HIP(hipStreamSynchronize(stream));
  double t_p0 = 0, t_dy = 0, t_sig = 0;
  for (int iq = share_rank; iq < sym.inq; iq += share_size) {
    float a = 0, b = 0, c = 0;
    GGPU_HIP(hipEventElapsedTime(&a, qev[4 * iq + 0], qev[4 * iq + 1]));
    GGPU_HIP(hipEventElapsedTime(&b, qev[4 * iq + 1], qev[4 * iq + 2]));
    GGPU_HIP(hipEventElapsedTime(&c, qev[4 * iq + 2], qev[4 * iq + 3]));
    t_p0 += a; t_dy += b; t_sig += c;
  }
  float tot = 0, g = 0, sg = 0;
  GGPU_HIP(hipEventElapsedTime(&tot, ev[0], ev[6]));
  GGPU_HIP(hipEventElapsedTime(&g, ev[0], ev[1]));
  GGPU_HIP(hipEventElapsedTime(&sg, ev[5], ev[6]));
  last_ms[0] = tot; last_ms[1] = g; last_ms[2] = t_p0; last_ms[3] = t_dy; last_ms[4] = t_sig; last_ms[5] = sg;
  streamed_last = store.bytes_streamed() - streamed0;
  generated_last = store.bytes_generated() - generated0;
  staging_ms_last = store.collect_staging_ms();
  if (profiling) prof.collect();
  check_info();
}

GwBatch::GwBatch(GwPlan* const* p, int n, const void* const* G_dev, void* const* S_dev) {
  require(n >= 1 && p && G_dev && S_dev, "gw batch: need at least one plan and its G / Sigma");
  for (int i = 0; i < n; ++i) {
    require(p[i] != nullptr && G_dev[i] && S_dev[i], "gw batch: NULL plan, G or Sigma");
    require(std::find(plans.begin(), plans.end(), p[i]) == plans.end(), "gw batch: a plan may appear only once");
    require(p[i]->store.loaded(), "gw batch: system " + std::to_string(i) + ": integrals not loaded", kErrState);
    require(!p[i]->store.staged(),
            "gw batch: system " + std::to_string(i) + ": graph replay needs the integrals resident in HBM", kErrState);
    plans.push_back(p[i]);
    G.push_back(static_cast<const z64*>(G_dev[i]));
    S.push_back(static_cast<z64*>(S_dev[i]));
  }
  GGPU_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  GGPU_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  GGPU_HIP(hipEventCreate(&t0));
  GGPU_HIP(hipEventCreate(&t1));
  join.resize(n, nullptr);
  for (auto& e : join) GGPU_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

GwBatch::~GwBatch() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  for (auto e : join)
    if (e) (void)hipEventDestroy(e);
  for (auto e : {fork, t0, t1})
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

// Fork every plan's stream from the batch stream inside one capture, record
// each plan's iteration on its own stream and join them back.
void GwBatch::capture() {
  if (exec) (void)hipGraphExecDestroy(exec);
  exec = nullptr;
  for (auto* p : plans) GGPU_HIP(hipStreamSynchronize(p->stream));
  GGPU_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    GGPU_HIP(hipEventRecord(fork, stream));
    for (size_t i = 0; i < plans.size(); ++i) {
      GwPlan& p = *plans[i];
      GGPU_HIP(hipStreamWaitEvent(p.stream, fork, 0));
      VariantGuard vg(p.variants, &p.ozwork);
      p.enqueue_iteration(G[i], S[i], false);
      GGPU_HIP(hipEventRecord(join[i], p.stream));
      GGPU_HIP(hipStreamWaitEvent(stream, join[i], 0));
    }
  } catch (...) {
    (void)hipStreamEndCapture(stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  GGPU_HIP(hipStreamEndCapture(stream, &graph));
  const hipError_t e = hipGraphInstantiateWithFlags(&exec, graph, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    exec = nullptr;
    GGPU_HIP(e);
  }
  gens.clear();
  for (auto* p : plans) gens.push_back(p->config_gen);
  ++captures;
}

void GwBatch::run() {
  bool stale = exec == nullptr;
  for (size_t i = 0; i < plans.size() && !stale; ++i) stale = gens[i] != plans[i]->config_gen;
  for (size_t i = 0; i < plans.size(); ++i) {
    require(!plans[i]->store.staged(),
            "gw batch: system " + std::to_string(i) + ": graph replay needs the integrals resident in HBM", kErrState);
    require(!plans[i]->profiling, "gw batch: per-site profiling does not apply to graph replay", kErrState);
  }
  if (stale) capture();
  GGPU_HIP(hipEventRecord(t0, stream));
  GGPU_HIP(hipGraphLaunch(exec, stream));
  GGPU_HIP(hipEventRecord(t1, stream));
  GGPU_HIP(hipStreamSynchronize(stream));
  float ms = 0;
  GGPU_HIP(hipEventElapsedTime(&ms, t0, t1));
  last_ms = ms;
  for (size_t i = 0; i < plans.size(); ++i) {
    try {
      plans[i]->check_info();
    } catch (const Error& e) {
      throw Error(e.code, "system " + std::to_string(i) + ": " + e.what());
    }
  }
}

void GwPlan::iterate(const double* G, double* Sigma) {
  require(G && Sigma, "gw: NULL G or Sigma");
  GGPU_HIP(hipMemcpyAsync(G_host_dev.get(), G, G_host_dev.bytes(), hipMemcpyHostToDevice, stream));
  iterate_device(G_host_dev.get(), Sig_host_dev.get());
  GGPU_HIP(hipMemcpyAsync(Sigma, Sig_host_dev.get(), Sig_host_dev.bytes(), hipMemcpyDeviceToHost, stream));
  GGPU_HIP(hipStreamSynchronize(stream));
}

void GwPlan::polarization(const double* G, int iq, int which, double* out) {
  require(store.loaded(), "gw: integrals not loaded", kErrState);
  require(iq >= 0 && iq < sym.inq, "gw: iq out of range");
  require(which >= 0 && which <= 2, "gw: which must be 0, 1 or 2");
  require(G && out, "gw: NULL pointer");
  VariantGuard vguard(variants, &ozwork);
  GGPU_HIP(hipMemcpyAsync(G_host_dev.get(), G, G_host_dev.bytes(), hipMemcpyHostToDevice, stream));
  stage_G(G_host_dev.get());
  zero_int(info.get(), nw_b, stream);
  last_q = iq;
  p0_for_q(iq);
  if (which != 0) dyson_for_q();
  GGPU_HIP(hipStreamSynchronize(stream));
  (void)store.collect_staging_ms();
  if (which != 0) check_info();
  int q_done = -1;
  last_polarization(which, &q_done, out);
}

void GwPlan::last_polarization(int which, int* iq, double* out) {
  require(which >= 0 && which <= 2, "gw: which must be 0, 1 or 2");
  require(out != nullptr && iq != nullptr, "gw: NULL pointer");
  require(last_q >= 0, "gw: no q has been computed yet", kErrState);
  const int64_t Q2 = (int64_t)NQ * NQ;
  GGPU_HIP(hipStreamSynchronize(stream));
  *iq = last_q;
  if (which == 0) {
    // sum the split-K partial slices, scale by 1/nk and mirror tau
    std::vector<z64> h((size_t)p0_split * half * Q2);
    GGPU_HIP(hipMemcpy(h.data(), P0.get(), h.size() * sizeof(z64), hipMemcpyDeviceToHost));
    z64* o = reinterpret_cast<z64*>(out);
    const double sc = 1.0 / sym.nk;
    for (int t = 0; t < half; ++t)
      for (int64_t e = 0; e < Q2; ++e) {
        z64 v{0.0, 0.0};
        for (int j = 0; j < p0_split; ++j) {
          const z64 u = h[((size_t)j * half + t) * Q2 + e];
          v.x += u.x;
          v.y += u.y;
        }
        o[(size_t)t * Q2 + e] = z64{v.x * sc, v.y * sc};
        o[(size_t)(nts - 1 - t) * Q2 + e] = z64{v.x * sc, v.y * sc};
      }
    return;
  }
  const z64* src = (which == 1) ? P0w.get() : Pt.get();
  const size_t cnt = (which == 1 ? (size_t)nw_b : (size_t)nts) * Q2;
  GGPU_HIP(hipMemcpy(out, src, cnt * sizeof(z64), hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------------------
// Static scratch of the no-allocation HF exchange
// ---------------------------------------------------------------------------
// K partial sums of ggpu_hf_exchange_device_noalloc: 4 Mi complex doubles
// (64 MiB) per GPU, held until the process ends.  Kernels never name the
// array: host code passes the device address hipGetSymbolAddress returns for
// the current device (each GPU has its own copy at its own address).
constexpr int64_t kHfScratchElems = (int64_t)1 << 22;
static __device__ __align__(256) double scratch[2 * kHfScratchElems];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_addr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
static z64* scratch_on_current_device() {
  int dev = 0;
  GGPU_HIP(hipGetDevice(&dev));
  require(dev >= 0 && dev < kMaxScratchDevices, "scratch: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    GGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "scratch: no device address");
    g_scratch_addr[dev] = p;
  }
  return static_cast<z64*>(g_scratch_addr[dev]);
}

// ---------------------------------------------------------------------------
// HF plan
// ---------------------------------------------------------------------------
HfPlan::HfPlan(int nao_, int nso_, int ns_, int NQ_, const ggpu_symmetry& s)
    : nao(nao_), nso(nso_), ns(ns_), nps(ns_), NQ(NQ_) {
  require(nao > 0 && NQ > 0, "hf: nao and NQ must be positive");
  require(nso == nao || nso == 2 * nao, "hf: nso must be nao or 2 nao");
  require(ns == 1 || ns == 2, "hf: Invalid value of \"ns\"");
  x2c = nso == 2 * nao;
  require(!x2c || ns == 1, "hf: a two-component (nso = 2 nao) plan needs ns = 1");
  if (x2c) nps = 3;   // density blocks aa, bb, ab (cuhf_utils with pseudo spin 3)
  std::fill(variants, variants + kNumSites, -1);
  sym.load(s, nso, NQ);
  if (x2c)
    for (auto t : sym.k_tr)
      require(t == 0, "hf: two-component plans do not support time-reversal k images");
  GGPU_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  const int64_t n2 = (int64_t)nao * nao, stride = (int64_t)NQ * n2, so2 = (int64_t)nso * nso;
  tables.resize(sym.ink);
  for (int ik = 0; ik < sym.ink; ++ik) {
    const int64_t k = sym.k_r2f[ik];
    std::vector<PairView> v;
    for (int k2 = 0; k2 < sym.nk; ++k2) v.push_back(sym.view(k, k2));
    tables[ik].build(v, stride, stream);
  }
  dm_host_dev.resize((size_t)ns * sym.ink * so2);
  dm_fbz.resize((size_t)sym.nk * nps * n2);
  const int64_t per_pair = (1 + nps) * stride * (int64_t)sizeof(z64);
  size_t fr = 0, tot = 0;
  GGPU_HIP(hipMemGetInfo(&fr, &tot));
  const size_t budget = std::min<size_t>((size_t)(0.4 * fr), (size_t)16 << 30);
  kb = (int)std::max<int64_t>(1, std::min<int64_t>(budget / per_pair, sym.nk));
  work.resize((size_t)kb * (per_pair / sizeof(z64)));
  // enough (b, Q) slices that a contraction launch (one nao x nao tile per
  // density block and slice) has about 1024 workgroups on the 256 CUs
  const int tiles = ((nao + 63) / 64) * ((nao + 63) / 64);
  nsplit = std::max(64, std::min(512, 1024 / (tiles * nps)));
  Kpart.resize((size_t)nsplit * nps * n2);
  Kout.resize((size_t)sym.ink * nps * n2);
  jvec.resize(NQ);
  S_dev.resize((size_t)ns * sym.ink * so2);
  Ftmp.resize((size_t)nps * sym.ink * n2);
  diag_tables.resize(2);
  {
    std::vector<PairView> all, ibz;
    for (int k = 0; k < sym.nk; ++k) all.push_back(sym.view(k, k));
    for (int ik = 0; ik < sym.ink; ++ik) ibz.push_back(sym.view(sym.k_r2f[ik], sym.k_r2f[ik]));
    diag_tables[0].build(all, stride, stream);
    diag_tables[1].build(ibz, stride, stream);
  }
  std::vector<int64_t> bo(kb);
  for (int b = 0; b < kb; ++b) bo[b] = b * stride;
  bidx_off.resize(kb);
  GGPU_HIP(hipMemcpy(bidx_off.get(), bo.data(), bo.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  // set-up of the no-allocation exchange: this device's scratch address and
  // the Kout -> caller-layout tables
  scratch = scratch_on_current_device();
  scratch_elems = kHfScratchElems;
  if (!x2c) {
    std::vector<int64_t> src((size_t)ns * sym.ink), dst(src.size());
    for (int ik = 0; ik < sym.ink; ++ik)
      for (int sp = 0; sp < ns; ++sp) {
        src[(size_t)ik * ns + sp] = ((int64_t)ik * ns + sp) * n2;
        dst[(size_t)ik * ns + sp] = ((int64_t)sp * sym.ink + ik) * n2;
      }
    kout_src.resize(src.size());
    kout_dst.resize(dst.size());
    GGPU_HIP(hipMemcpy(kout_src.get(), src.data(), src.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    GGPU_HIP(hipMemcpy(kout_dst.get(), dst.data(), dst.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  }
  if (x2c) {
    dm2c_fbz.resize((size_t)sym.nk * so2);
    F2c.resize((size_t)sym.ink * so2);
    build_split(x2c_split, sym.nk, 3, nao, so2, n2);
    // Fock blocks aa, bb, ab from Kout[ik][0..2]; ba = ab^dagger
    // (copy_2c_Fock_from_device_to_host, cuda_common.cu:176-205)
    std::vector<int64_t> in, out;
    std::vector<int> tr, cj;
    for (int ik = 0; ik < sym.ink; ++ik)
      for (int b = 0; b < 4; ++b) {
        in.push_back(((int64_t)ik * 3 + (b == 3 ? 2 : b)) * n2);
        out.push_back(ik * so2 + (int64_t)kX2cBlock[b][0] * nao * nso + (int64_t)kX2cBlock[b][1] * nao);
        tr.push_back(b == 3);
        cj.push_back(b == 3);
      }
    x2c_merge.build(in, out, tr, cj);
    rot.init(sym, nso, 1, 1, stream);
  } else {
    rot.init(sym, nao, ns, 1, stream);
  }
}

void HfPlan::set_precision(int precision) {
  require(precision >= GGPU_PRECISION_FP64 && precision <= GGPU_PRECISION_INT8X4, "hf: unknown precision");
  compute = precision == GGPU_PRECISION_FP64     ? Compute::F64
            : precision == GGPU_PRECISION_FP32   ? Compute::F32
            : precision == GGPU_PRECISION_BF16X3 ? Compute::BF16X3
            : precision == GGPU_PRECISION_INT8X5 ? Compute::I8X5
                                                 : Compute::I8X4;
  dm32.resize(complex64_operands(compute) ? dm_fbz.size() : 0);
  if (int8_slices(compute) && ozwork.size() == 0) {
    size_t fr = 0, tot = 0;
    GGPU_HIP(hipMemGetInfo(&fr, &tot));
    ozwork.resize(std::min<size_t>((size_t)8 << 30, fr / 8));
  }
}

HfPlan::~HfPlan() {
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
}

// K_s(k) = (px/nk) sum_{k2} sum_Q V^Q(k,k2) D_s(k2) V^Q(k,k2)^dagger  (cuhf_utils.cu:108-159):
//   X[s][b][Q][i][c] = sum_a V^Q(k,k2_b)[i][a] D_s(k2_b)[a][c]
//   K[s][i][j]     += (px/nk) sum_{b,Q,c} X[s][b][Q][i][c] conj(V^Q(k,k2_b)[j][c])
// The (b, Q) reduction is split into `nsplit` slices written to partial
// buffers, so each launch has nsplit * ns workgroups instead of ns.
void HfPlan::exchange_device(z64* kpart, int nsp) {
  const int64_t n2 = (int64_t)nao * nao, stride = (int64_t)NQ * n2;
  const double px = (nps == 1) ? -0.5 : -1.0;   // -1 for ns = 2 and the X2C blocks
  const bool lp = complex64_operands(compute);
  const int64_t es = lp ? (int64_t)sizeof(c32) : (int64_t)sizeof(z64);
  char* Vdir = reinterpret_cast<char*>(work.get());
  char* X = Vdir + kb * stride * es;
  VariantGuard vguard(variants, &ozwork);
  const int64_t xs = (int64_t)kb * stride;          // one spin slab of X (elements)
  const Span wk{work.get(), work.bytes() / (size_t)es};
  if (lp) {
    require(dm32.size() == dm_fbz.size(), "hf: reduced-precision density buffer missing");
    to_c32(dm_fbz.get(), dm32.get(), (int64_t)dm_fbz.size(), stream);
  }
  const void* dm = lp ? (const void*)dm32.get() : (const void*)dm_fbz.get();
  const Span dsp{dm, dm_fbz.size()};
  if (share_size > 1) zero(Kout.get(), Kout.size(), stream);
  const int64_t part_elems = (int64_t)nsp * nps * n2;
  for (int ik = share_rank; ik < sym.ink; ik += share_size) {
    zero(kpart, part_elems, stream);
    for (int i0 = 0; i0 < sym.nk; i0 += kb) {
      const int n = std::min(kb, sym.nk - i0);
      unfold_pairs(store, tables[ik], i0, n, bidx_off.get(), Vdir, lp, NQ, nao, stream);
      {
        // X[s][b] = V(k, k2 = i0 + b) D_s(k2) for every density block s and
        // pair b of the batch, in one launch
        GemmDesc g;
        g.compute = compute; g.out_f32 = lp;
        g.M = NQ * nao; g.N = nao; g.K = nao; g.nb0 = nps; g.nb1 = n;
        g.A.ptr = Vdir; g.A.rs = nao; g.A.cs = 1; g.A.bs0 = 0; g.A.bs1 = stride;
        g.B.ptr = (const char*)dm + (int64_t)i0 * nps * n2 * es; g.B.rs = nao; g.B.cs = 1; g.B.bs0 = n2;
        g.B.bs1 = (int64_t)nps * n2;
        g.C.ptr = X; g.C.rs = nao; g.C.cs = 1; g.C.bs0 = xs; g.C.bs1 = stride;
        gemm_on(g, wk, dsp, wk, stream, kSiteHF);
      }
      const int total = n * NQ;
      const int ch = (total + nsp - 1) / nsp;
      const int full = total / ch, rem = total - full * ch;
      for (int part = 0; part < 2; ++part) {
        const int slices = part == 0 ? full : (rem > 0 ? 1 : 0);
        if (slices == 0) continue;
        const int kbat = part == 0 ? ch : rem;
        const int64_t first = part == 0 ? 0 : (int64_t)full * ch;
        GemmDesc f;
        f.compute = compute; f.out_f32 = 0;       // K partial sums in FP64
        f.M = nao; f.N = nao; f.K = nao; f.KB = kbat; f.nb0 = nps; f.nb1 = slices;
        f.A.ptr = X + first * n2 * es; f.A.rs = nao; f.A.cs = 1; f.A.kbs = n2; f.A.bs0 = xs;
        f.A.bs1 = (int64_t)ch * n2;
        f.B.ptr = Vdir + first * n2 * es; f.B.rs = 1; f.B.cs = nao; f.B.kbs = n2; f.B.bs0 = 0;
        f.B.bs1 = (int64_t)ch * n2;
        f.B.conj = 1;
        f.C.ptr = kpart + (part == 0 ? 0 : (int64_t)full * nps * n2); f.C.rs = nao; f.C.cs = 1;
        f.C.bs0 = n2; f.C.bs1 = (int64_t)nps * n2;
        f.alpha_re = px / sym.nk; f.beta_re = 1.0;
        gemm_on(f, wk, wk, Span{kpart, (size_t)part_elems}, stream, kSiteHF);
      }
    }
    sum_partials(kpart, (int64_t)nps * n2, nsp, Kout.get() + (int64_t)ik * nps * n2, (int64_t)nps * n2, stream);
  }
}

static void ibz_to_host_layout(const std::vector<z64>& h, double* out, int ink, int ns, int64_t n2) {
  z64* o = reinterpret_cast<z64*>(out);
  for (int ik = 0; ik < ink; ++ik)
    for (int s = 0; s < ns; ++s)
      std::memcpy(o + ((int64_t)s * ink + ik) * n2, h.data() + ((int64_t)ik * ns + s) * n2, n2 * sizeof(z64));
}

void HfPlan::stage_dm(const z64* dm_dev) {
  if (x2c) {
    rot.run(dm_dev, dm2c_fbz.get(), stream);              // (nk, 1, 1, nso, nso)
    spinor_copy(x2c_split, dm2c_fbz.get(), dm2c_fbz.size(), nso, 0, dm_fbz.get(), dm_fbz.size(), nao, 0, nao, 1, 1.0,
                stream);
  } else {
    rot.run(dm_dev, dm_fbz.get(), stream);                // (nk, ns, 1, nao, nao)
  }
}

void HfPlan::fetch(double* out) {
  const int64_t n2 = (int64_t)nao * nao;
  (void)store.collect_staging_ms();   // staging timers are not reported for HF
  if (x2c) {
    spinor_copy(x2c_merge, Kout.get(), Kout.size(), nao, 0, F2c.get(), F2c.size(), nso, 0, nao, 1, 1.0, stream);
    GGPU_HIP(hipMemcpyAsync(out, F2c.get(), F2c.bytes(), hipMemcpyDeviceToHost, stream));
    GGPU_HIP(hipStreamSynchronize(stream));
    return;
  }
  std::vector<z64> h(Kout.size());
  GGPU_HIP(hipMemcpyAsync(h.data(), Kout.get(), Kout.bytes(), hipMemcpyDeviceToHost, stream));
  GGPU_HIP(hipStreamSynchronize(stream));
  ibz_to_host_layout(h, out, sym.ink, ns, n2);
}

// E_2 = (1/nk) sum_{s, k in IBZ} w_k (1/2) Re Tr[D_s(k) F_s(k)], w_k = |star(k)|:
// the two-electron HF energy of the full Brillouin zone (Tr[D F] is unchanged
// by the unitary k rotation and only conjugated under time reversal).  The
// reference returns only the Fock matrix (hf_gpu_kernel.cpp:66-104); the
// energy is green-mbpt's.  With ns = 1 the density is the spin-summed one, so
// the same formula holds.  Traces accumulate in FP64 on the device.
void HfPlan::compute_energy() {
  const int n = x2c ? nso : nao;
  const int64_t n2 = (int64_t)n * n;
  const int nsp = x2c ? 1 : ns;
  traces.resize((size_t)nsp * sym.ink);
  GemmDesc g;   // tr[s][ik] = sum_a sum_b D[a][b] F[b][a]: K index b, K-batch a
  g.M = 1; g.N = 1; g.K = n; g.KB = n; g.nb0 = nsp; g.nb1 = sym.ink;
  g.A.ptr = dm_host_dev.get(); g.A.rs = 0; g.A.cs = 1; g.A.kbs = n;
  g.A.bs0 = (int64_t)sym.ink * n2; g.A.bs1 = n2;
  const z64* F = x2c ? F2c.get() : Kout.get();
  g.B.ptr = F; g.B.rs = n; g.B.cs = 0; g.B.kbs = 1;
  g.B.bs0 = x2c ? 0 : n2; g.B.bs1 = x2c ? n2 : (int64_t)ns * n2;
  g.C.ptr = traces.get(); g.C.rs = 0; g.C.cs = 0; g.C.bs0 = sym.ink; g.C.bs1 = 1;
  if (x2c) gemm_on(g, dm_host_dev, F2c, traces, stream, kSiteHF);
  else gemm_on(g, dm_host_dev, Kout, traces, stream, kSiteHF);
  std::vector<z64> h(traces.size());
  GGPU_HIP(hipMemcpyAsync(h.data(), traces.get(), traces.bytes(), hipMemcpyDeviceToHost, stream));
  GGPU_HIP(hipStreamSynchronize(stream));
  double e = 0.0;
  for (int s = 0; s < nsp; ++s)
    for (int ik = 0; ik < sym.ink; ++ik) e += (double)sym.k_stars[ik].size() * 0.5 * h[(size_t)s * sym.ink + ik].x;
  energy = e / sym.nk;
  energy_valid = true;
}

void HfPlan::exchange(const double* dm, double* K) {
  require(store.loaded(), "hf: integrals not loaded", kErrState);
  require(dm && K, "hf: NULL pointer");
  GGPU_HIP(hipMemcpyAsync(dm_host_dev.get(), dm, dm_host_dev.bytes(), hipMemcpyHostToDevice, stream));
  stage_dm(dm_host_dev.get());
  exchange_device(Kpart.get(), nsplit);
  fetch(K);
}

// The exchange on caller device buffers with the K partial sums in the static
// scratch array: the same launches as exchange() (fewer partial slices when
// the plan's would not fit the scratch), and no device allocation -- every
// other buffer it touches was allocated when the plan was created.
void HfPlan::exchange_device_noalloc(const z64* dm_dev, z64* K_dev) {
  require(store.loaded(), "hf: integrals not loaded", kErrState);
  require(dm_dev && K_dev, "hf: NULL pointer");
  require(!store.staged(), "hf no-alloc exchange: needs the integrals resident in HBM", kErrState);
  require(!complex64_operands(compute) || dm32.size() == dm_fbz.size(),
          "hf: reduced-precision density buffer missing");
  const int64_t n2 = (int64_t)nao * nao, so2 = (int64_t)nso * nso;
  const int64_t slice = (int64_t)nps * n2;
  require(slice <= scratch_elems,
          "hf no-alloc exchange: one partial slice of " + std::to_string(slice) +
              " complex elements exceeds the static scratch of " + std::to_string(scratch_elems));
  require(scratch == scratch_on_current_device(), "hf no-alloc exchange: plan created on another device");
  const int nsp = (int)std::max<int64_t>(1, std::min<int64_t>(nsplit, scratch_elems / slice));
  stage_dm(dm_dev);
  exchange_device(scratch, nsp);
  if (x2c)
    spinor_copy(x2c_merge, Kout.get(), Kout.size(), nao, 0, K_dev, (size_t)sym.ink * so2, nso, 0, nao, 1, 1.0, stream);
  else
    slab_gather(Kout.get(), kout_src.get(), K_dev, kout_dst.get(), (int)kout_src.size(), n2, 1.0, stream);
  GGPU_HIP(hipStreamSynchronize(stream));
}

// F = J + K - c madelung S D S  (hf_gpu_kernel.cpp:66-104, :141-201); the
// two-component variant follows hf_gpu_kernel.cpp:253-327.
void HfPlan::solve(const double* dm, const double* S_k, double madelung, double* F) {
  require(store.loaded(), "hf: integrals not loaded", kErrState);
  require(dm && S_k && F, "hf: NULL pointer");
  const int64_t n2 = (int64_t)nao * nao, stride = (int64_t)NQ * n2, so2 = (int64_t)nso * nso;
  GGPU_HIP(hipMemcpyAsync(dm_host_dev.get(), dm, dm_host_dev.bytes(), hipMemcpyHostToDevice, stream));
  GGPU_HIP(hipMemcpyAsync(S_dev.get(), S_k, S_dev.bytes(), hipMemcpyHostToDevice, stream));
  stage_dm(dm_host_dev.get());
  exchange_device(Kpart.get(), nsplit);                // Kout[ik][s]
  // Hartree: jvec[Q] = (1/nk) sum_{k', s} sum_{a,b} V^Q(k',k')[a][b] D_s(k')[b][a]
  // (X2C: the density is D_aa + D_bb, blocks 0 and 1)
  const int nj = x2c ? 2 : ns;
  z64* Vdir = work.get();
  zero(jvec.get(), jvec.size(), stream);
  for (int i0 = 0; i0 < sym.nk; i0 += kb) {
    const int n = std::min(kb, sym.nk - i0);
    unfold_pairs(store, diag_tables[0], i0, n, bidx_off.get(), Vdir, 0, NQ, nao, stream);
    for (int b = 0; b < n; ++b)
      for (int s = 0; s < nj; ++s) {
        GemmDesc g;   // (NQ x 1) = sum_a sum_b V[Q][a][b] D[b][a]: K index b, K-batch a
        g.M = NQ; g.N = 1; g.K = nao; g.KB = nao;
        g.A.ptr = Vdir + b * stride; g.A.rs = n2; g.A.cs = 1; g.A.kbs = nao;
        g.B.ptr = dm_fbz.get() + ((int64_t)(i0 + b) * nps + s) * n2; g.B.rs = nao; g.B.cs = 0; g.B.kbs = 1;
        g.C.ptr = jvec.get(); g.C.rs = 1; g.C.cs = 1;
        g.alpha_re = 1.0 / sym.nk; g.beta_re = 1.0;
        gemm_on(g, work, dm_fbz, jvec, stream, kSiteHF);
      }
  }
  // F[ik][s] = K + sum_Q jvec[Q] V^Q(k,k) - c madelung S D S
  for (int i0 = 0; i0 < sym.ink; i0 += kb) {
    const int n = std::min(kb, sym.ink - i0);
    unfold_pairs(store, diag_tables[1], i0, n, bidx_off.get(), Vdir, 0, NQ, nao, stream);
    for (int b = 0; b < n; ++b) {
      GemmDesc g;   // (1 x nao^2) += jvec^T (1 x NQ) . V (NQ x nao^2), for every spin (X2C: aa, bb)
      g.M = 1; g.N = (int)n2; g.K = NQ; g.nb0 = nj;
      g.A.ptr = jvec.get(); g.A.rs = 0; g.A.cs = 1;
      g.B.ptr = Vdir + b * stride; g.B.rs = n2; g.B.cs = 1;
      g.C.ptr = Kout.get() + (int64_t)(i0 + b) * nps * n2; g.C.rs = 0; g.C.cs = 1; g.C.bs0 = n2;
      g.beta_re = 1.0;
      gemm_on(g, jvec, work, Kout, stream, kSiteHF);
    }
  }
  if (x2c) {
    // Ewald on the aa, bb, ab blocks with the AO overlap S_aa on both sides
    // (x2c_hf_gpu_kernel::add_Ewald, hf_gpu_kernel.cpp:304-327); ba follows
    // as the adjoint of ab when the blocks are merged
    for (int b = 0; b < 3; ++b) {
      GemmDesc g;
      g.M = nao; g.N = nao; g.K = nao; g.nb0 = sym.ink;
      g.A.ptr = S_dev.get(); g.A.rs = nso; g.A.cs = 1; g.A.bs0 = so2;
      g.B.ptr = dm_host_dev.get() + (int64_t)kX2cBlock[b][0] * nao * nso + (int64_t)kX2cBlock[b][1] * nao;
      g.B.rs = nso; g.B.cs = 1; g.B.bs0 = so2;
      g.C.ptr = Ftmp.get() + b * n2; g.C.rs = nao; g.C.cs = 1; g.C.bs0 = 3 * n2;
      gemm_on(g, S_dev, dm_host_dev, Ftmp, stream, kSiteHF);
    }
    GemmDesc h;
    h.M = nao; h.N = nao; h.K = nao; h.nb0 = sym.ink; h.nb1 = 3;
    h.A.ptr = Ftmp.get(); h.A.rs = nao; h.A.cs = 1; h.A.bs0 = 3 * n2; h.A.bs1 = n2;
    h.B.ptr = S_dev.get(); h.B.rs = nso; h.B.cs = 1; h.B.bs0 = so2; h.B.bs1 = 0;
    h.C.ptr = Kout.get(); h.C.rs = nao; h.C.cs = 1; h.C.bs0 = 3 * n2; h.C.bs1 = n2;
    h.alpha_re = -madelung; h.beta_re = 1.0;
    gemm_on(h, Ftmp, S_dev, Kout, stream, kSiteHF);
  } else {
    // Ewald: Ftmp[s][ik] = S D, then Kout[ik][s] -= c madelung (S D) S
    const double c = (ns == 2) ? 1.0 : 0.5;
    GemmDesc g;
    g.M = nao; g.N = nao; g.K = nao; g.nb0 = ns * sym.ink;
    g.A.ptr = S_dev.get(); g.A.rs = nao; g.A.cs = 1; g.A.bs0 = n2;
    g.B.ptr = dm_host_dev.get(); g.B.rs = nao; g.B.cs = 1; g.B.bs0 = n2;
    g.C.ptr = Ftmp.get(); g.C.rs = nao; g.C.cs = 1; g.C.bs0 = n2;
    gemm_on(g, S_dev, dm_host_dev, Ftmp, stream, kSiteHF);
    for (int s = 0; s < ns; ++s) {
      GemmDesc h;
      h.M = nao; h.N = nao; h.K = nao; h.nb0 = sym.ink;
      h.A.ptr = Ftmp.get() + (int64_t)s * sym.ink * n2; h.A.rs = nao; h.A.cs = 1; h.A.bs0 = n2;
      h.B.ptr = S_dev.get() + (int64_t)s * sym.ink * n2; h.B.rs = nao; h.B.cs = 1; h.B.bs0 = n2;
      h.C.ptr = Kout.get() + (int64_t)s * n2; h.C.rs = nao; h.C.cs = 1; h.C.bs0 = (int64_t)ns * n2;
      h.alpha_re = -c * madelung; h.beta_re = 1.0;
      gemm_on(h, Ftmp, S_dev, Kout, stream, kSiteHF);
    }
  }
  // rows of IBZ k outside this plan's share hold only J and Ewald: clear them
  if (share_size > 1)
    for (int ik = 0; ik < sym.ink; ++ik)
      if (ik % share_size != share_rank) zero(Kout.get() + (int64_t)ik * nps * n2, (int64_t)nps * n2, stream);
  energy_valid = false;
  fetch(F);
  compute_energy();
}

}  // namespace ggpu
