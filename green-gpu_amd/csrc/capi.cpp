// extern "C" entry points: every call catches the library's exceptions and
// turns them into a status code plus a thread-local message.
#include "../../include/green_gpu_amd.h"
#include "engine.hpp"

#include <algorithm>
#include <exception>
#include <new>
#include <string>
#include <thread>
#include <vector>

struct ggpu_hf_plan { ggpu::HfPlan* impl; };
struct ggpu_gw_plan { ggpu::GwPlan* impl; };
struct ggpu_gw_batch { ggpu::GwBatch* impl = nullptr; };

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return GGPU_OK;
  } catch (const ggpu::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return GGPU_ERR_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return GGPU_ERR_ARGUMENT;
  }
}
}  // namespace

extern "C" {

const char* ggpu_last_error(void) { return g_last_error.c_str(); }

int ggpu_device_count(int* count) {
  return guarded([&] {
    ggpu::require(count != nullptr, "NULL count");
    int n = 0;
    GGPU_HIP(hipGetDeviceCount(&n));
    *count = n;
  });
}

int ggpu_device_memory(size_t* free_bytes, size_t* total_bytes) {
  return guarded([&] {
    ggpu::require(free_bytes && total_bytes, "NULL pointer");
    GGPU_HIP(hipMemGetInfo(free_bytes, total_bytes));
  });
}

int ggpu_hf_plan_create(ggpu_hf_plan** plan, int nao, int nso, int ns, int NQ, const ggpu_symmetry* sym) {
  return guarded([&] {
    ggpu::require(plan && sym, "NULL plan or symmetry");
    *plan = nullptr;
    auto* p = new ggpu_hf_plan{nullptr};
    try {
      p->impl = new ggpu::HfPlan(nao, nso, ns, NQ, *sym);
    } catch (...) {
      delete p;
      throw;
    }
    *plan = p;
  });
}

int ggpu_hf_set_integral_residence(ggpu_hf_plan* plan, int residence) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    auto* p = plan->impl;
    p->store.set_residence(residence, p->kb, (int64_t)p->NQ * p->nao * p->nao);
  });
}

int ggpu_hf_set_precision(ggpu_hf_plan* plan, int precision) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->set_precision(precision);
  });
}

int ggpu_hf_set_variants(ggpu_hf_plan* plan, const int* variants) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && variants, "NULL argument");
    ggpu::HfPlan& p = *plan->impl;
    for (int s = 0; s < ggpu::kNumSites; ++s)
      ggpu::require(variants[s] == -1 || (s == ggpu::kSiteHF && ggpu::gemm_variant_valid(p.compute, variants[s])),
                    "hf: launch variant " + std::to_string(variants[s]) + " not available for call site " +
                        std::to_string(s) + " in this plan's precision");
    std::copy(variants, variants + ggpu::kNumSites, p.variants);
  });
}

int ggpu_hf_set_k_share(ggpu_hf_plan* plan, int rank, int size) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    ggpu::require(size >= 1 && rank >= 0 && rank < size, "k share: need 0 <= rank < size");
    plan->impl->share_rank = rank;
    plan->impl->share_size = size;
  });
}

int ggpu_hf_load_integrals(ggpu_hf_plan* plan, const double* V_stored) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    auto* p = plan->impl;
    p->store.load_host(V_stored, p->sym.num_stored * (int64_t)p->NQ * p->nao * p->nao, p->stream);
  });
}

int ggpu_hf_set_integral_reader(ggpu_hf_plan* plan, ggpu_pair_reader reader, void* user) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && reader, "NULL argument");
    auto* p = plan->impl;
    p->store.set_residence(ggpu::IntegralStore::kResReader, p->kb, (int64_t)p->NQ * p->nao * p->nao);
    p->store.set_reader(reader, user, p->sym.num_stored);
  });
}

int ggpu_hf_synth_integrals(ggpu_hf_plan* plan, uint64_t seed, double sigma) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    auto* p = plan->impl;
    p->store.synth(p->sym, p->NQ, p->nao, seed, sigma, p->stream);
  });
}

int ggpu_hf_exchange(ggpu_hf_plan* plan, const double* dm, double* K_out) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->exchange(dm, K_out);
  });
}

int ggpu_hf_solve(ggpu_hf_plan* plan, const double* dm, const double* S_k, double madelung, double* F_out) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->solve(dm, S_k, madelung, F_out);
  });
}

int ggpu_hf_exchange_device_noalloc(ggpu_hf_plan* plan, const void* dm_device, void* K_device) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->exchange_device_noalloc(static_cast<const ggpu::z64*>(dm_device), static_cast<ggpu::z64*>(K_device));
  });
}

int ggpu_hf_energy(ggpu_hf_plan* plan, double* energy) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && energy, "NULL argument");
    ggpu::require(plan->impl->energy_valid, "hf: no solve has run yet", ggpu::kErrState);
    *energy = plan->impl->energy;
  });
}

void ggpu_hf_plan_destroy(ggpu_hf_plan* plan) {
  if (!plan) return;
  delete plan->impl;
  delete plan;
}

int ggpu_gw_plan_create(ggpu_gw_plan** plan, int nao, int nso, int ns, int NQ, int nts, int nw_b, const ggpu_symmetry* sym,
                        const double* T_wt, const double* T_tw, int precision, int solver, size_t work_bytes) {
  return guarded([&] {
    ggpu::require(plan && sym, "NULL plan or symmetry");
    *plan = nullptr;
    auto* p = new ggpu_gw_plan{nullptr};
    try {
      p->impl = new ggpu::GwPlan(nao, nso, ns, NQ, nts, nw_b, *sym, T_wt, T_tw, precision, solver, work_bytes);
    } catch (...) {
      delete p;
      throw;
    }
    *plan = p;
  });
}

int ggpu_gw_set_integral_residence(ggpu_gw_plan* plan, int residence) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    auto* p = plan->impl;
    p->invalidate_graph();
    p->store.set_residence(residence, p->kb, (int64_t)p->NQ * p->nao * p->nao);
  });
}

int ggpu_gw_streamed_bytes(ggpu_gw_plan* plan, double* bytes) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && bytes, "NULL argument");
    *bytes = plan->impl->streamed_last;
  });
}

int ggpu_gw_generated_bytes(ggpu_gw_plan* plan, double* bytes) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && bytes, "NULL argument");
    *bytes = plan->impl->generated_last;
  });
}

int ggpu_gw_staging_ms(ggpu_gw_plan* plan, double* ms) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && ms, "NULL argument");
    *ms = plan->impl->staging_ms_last;
  });
}

int ggpu_gw_set_q_share(ggpu_gw_plan* plan, int rank, int size) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    ggpu::require(size >= 1 && rank >= 0 && rank < size, "q share: need 0 <= rank < size");
    plan->impl->share_rank = rank;
    plan->impl->share_size = size;
    plan->impl->invalidate_graph();
  });
}

int ggpu_gw_load_integrals(ggpu_gw_plan* plan, const double* V_stored) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    auto* p = plan->impl;
    p->invalidate_graph();
    p->store.load_host(V_stored, p->sym.num_stored * (int64_t)p->NQ * p->nao * p->nao, p->stream);
  });
}

int ggpu_gw_set_integral_reader(ggpu_gw_plan* plan, ggpu_pair_reader reader, void* user) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && reader, "NULL argument");
    auto* p = plan->impl;
    p->invalidate_graph();
    p->store.set_residence(ggpu::IntegralStore::kResReader, p->kb, (int64_t)p->NQ * p->nao * p->nao);
    p->store.set_reader(reader, user, p->sym.num_stored);
  });
}

int ggpu_gw_synth_integrals(ggpu_gw_plan* plan, uint64_t seed, double sigma) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    auto* p = plan->impl;
    p->invalidate_graph();
    p->store.synth(p->sym, p->NQ, p->nao, seed, sigma, p->stream);
  });
}

int ggpu_gw_iterate(ggpu_gw_plan* plan, const double* G, double* Sigma_out) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->iterate(G, Sigma_out);
  });
}

int ggpu_gw_iterate_device(ggpu_gw_plan* plan, const void* G_device, void* Sigma_device) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && G_device && Sigma_device, "NULL argument");
    plan->impl->iterate_device(static_cast<const ggpu::z64*>(G_device), static_cast<ggpu::z64*>(Sigma_device));
  });
}

int ggpu_gw_iterate_many(ggpu_gw_plan* const* plans, int n, const void* const* G_device, void* const* Sigma_device) {
  return guarded([&] {
    ggpu::require(plans && G_device && Sigma_device && n >= 1, "NULL argument or no systems");
    for (int i = 0; i < n; ++i) {
      ggpu::require(plans[i] && plans[i]->impl && G_device[i] && Sigma_device[i], "NULL argument");
      for (int j = 0; j < i; ++j) ggpu::require(plans[i] != plans[j], "the same plan appears twice");
    }
    int dev = 0;
    GGPU_HIP(hipGetDevice(&dev));
    // one host thread per system: each plan enqueues on its own stream, so
    // the systems' kernels overlap on the device
    std::vector<std::exception_ptr> err(n);
    std::vector<std::thread> th;
    th.reserve(n);
    for (int i = 0; i < n; ++i)
      th.emplace_back([&, i] {
        try {
          GGPU_HIP(hipSetDevice(dev));
          plans[i]->impl->iterate_device(static_cast<const ggpu::z64*>(G_device[i]),
                                         static_cast<ggpu::z64*>(Sigma_device[i]));
        } catch (...) {
          err[i] = std::current_exception();
        }
      });
    for (auto& t : th) t.join();
    for (int i = 0; i < n; ++i) {
      if (!err[i]) continue;
      try {
        std::rethrow_exception(err[i]);
      } catch (const ggpu::Error& e) {
        throw ggpu::Error(e.code, "system " + std::to_string(i) + ": " + e.what());
      }
    }
  });
}

int ggpu_gw_batch_create(ggpu_gw_batch** batch, ggpu_gw_plan* const* plans, int n, const void* const* G_device,
                         void* const* Sigma_device) {
  return guarded([&] {
    ggpu::require(batch != nullptr && plans != nullptr && n >= 1, "NULL argument");
    *batch = nullptr;
    std::vector<ggpu::GwPlan*> impl(n);
    for (int i = 0; i < n; ++i) {
      ggpu::require(plans[i] && plans[i]->impl, "gw batch: NULL plan");
      impl[i] = plans[i]->impl;
    }
    auto* b = new ggpu_gw_batch;
    try {
      b->impl = new ggpu::GwBatch(impl.data(), n, G_device, Sigma_device);
    } catch (...) {
      delete b;
      throw;
    }
    *batch = b;
  });
}

int ggpu_gw_batch_run(ggpu_gw_batch* batch, double* ms) {
  return guarded([&] {
    ggpu::require(batch && batch->impl, "NULL batch");
    batch->impl->run();
    if (ms) *ms = batch->impl->last_ms;
  });
}

int ggpu_gw_batch_captures(ggpu_gw_batch* batch, int* captures) {
  return guarded([&] {
    ggpu::require(batch && batch->impl && captures, "NULL argument");
    *captures = batch->impl->captures;
  });
}

void ggpu_gw_batch_destroy(ggpu_gw_batch* batch) {
  if (!batch) return;
  delete batch->impl;
  delete batch;
}

int ggpu_gw_polarization(ggpu_gw_plan* plan, const double* G, int iq, int which, double* out) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->polarization(G, iq, which, out);
  });
}

int ggpu_gw_last_polarization(ggpu_gw_plan* plan, int which, int* iq, double* out) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->last_polarization(which, iq, out);
  });
}

int ggpu_gw_set_pair_batch_limit(ggpu_gw_plan* plan, int batches) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    ggpu::require(batches >= 0, "gw: pair batch limit must be >= 0");
    plan->impl->invalidate_graph();
    plan->impl->batch_limit = batches;
  });
}

int ggpu_gw_last_timings(ggpu_gw_plan* plan, double* ms6) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && ms6, "NULL argument");
    for (int i = 0; i < 6; ++i) ms6[i] = plan->impl->last_ms[i];
  });
}

int ggpu_gw_set_profiling(ggpu_gw_plan* plan, int enable) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->profiling = enable != 0;
  });
}

int ggpu_gw_site_stats(ggpu_gw_plan* plan, int site, double* ms, double* flops, int64_t* launches) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && ms && flops && launches, "NULL argument");
    ggpu::require(site >= 0 && site < ggpu::kNumSites, "site out of range");
    *ms = plan->impl->prof.ms[site];
    *flops = plan->impl->prof.flops[site];
    *launches = plan->impl->prof.launches[site];
  });
}

int ggpu_gw_set_graph(ggpu_gw_plan* plan, int enable) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->invalidate_graph();
    plan->impl->use_graph = enable != 0;
  });
}

int ggpu_gw_graph_captures(ggpu_gw_plan* plan, int* captures) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && captures, "NULL argument");
    *captures = plan->impl->graph_captures;
  });
}

int ggpu_gw_autotune(ggpu_gw_plan* plan, const double* G, int rounds, int batches) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && G, "NULL argument");
    ggpu::GwPlan& p = *plan->impl;
    GGPU_HIP(hipMemcpyAsync(p.G_host_dev.get(), G, p.G_host_dev.bytes(), hipMemcpyHostToDevice, p.stream));
    p.autotune(p.G_host_dev.get(), rounds, batches);
  });
}

int ggpu_gw_autotune_device(ggpu_gw_plan* plan, const void* G_device, int rounds, int batches) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && G_device, "NULL argument");
    plan->impl->autotune(static_cast<const ggpu::z64*>(G_device), rounds, batches);
  });
}

int ggpu_gw_autotune_report(ggpu_gw_plan* plan, int cap, int* n, int* variants, double* ms, double* wall_ms) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && n, "NULL argument");
    const ggpu::GwPlan& p = *plan->impl;
    *n = (int)p.tune_cand.size();
    const int m = std::min(cap, *n);
    for (int c = 0; c < m; ++c) {
      if (variants) variants[c] = p.tune_cand[c];
      if (ms)
        for (int s = 0; s < ggpu::kNumSites; ++s) ms[(size_t)c * ggpu::kNumSites + s] = p.tune_ms[c][s];
    }
    if (wall_ms) *wall_ms = p.tune_wall_ms;
  });
}

int ggpu_gw_get_variants(ggpu_gw_plan* plan, int* variants) {
  return guarded([&] {
    ggpu::require(plan && plan->impl && variants, "NULL argument");
    std::copy(plan->impl->variants, plan->impl->variants + ggpu::kNumSites, variants);
  });
}

int ggpu_gw_set_variants(ggpu_gw_plan* plan, const int* variants) {
  return guarded([&] {
    ggpu::require(plan && plan->impl, "NULL plan");
    plan->impl->set_variants(variants);
  });
}

void ggpu_gw_plan_destroy(ggpu_gw_plan* plan) {
  if (!plan) return;
  delete plan->impl;
  delete plan;
}

}  // extern "C"
