/*
 * C++ adapter between green-mbpt's custom-kernel plugin API and the gfx950
 * C ABI (green_gpu_amd.h).  These classes take the places of
 * scalar_hf_gpu_kernel / x2c_hf_gpu_kernel (hf_gpu_kernel.h:68-100) and
 * scalar_gw_gpu_kernel / x2c_gw_gpu_kernel (gw_gpu_kernel.h:131-206):
 * they translate green-symmetry / green-grids objects into the plain tables
 * of ggpu_symmetry, read the density-fitted integrals from disk once, keep
 * one device plan per kernel object, and rethrow ABI status codes as
 * std::runtime_error (std::domain_error for a Dyson matrix that is not
 * positive definite under Cholesky or singular under LU).
 *
 * Threading and MPI: like the reference, every rank of a node calls solve();
 * the node's rank 0 does the device work on shared objects, the others wait
 * at the shared object's fences.  Spreading q-points over several devices
 * is not done by this adapter.
 */
#ifndef GREEN_GPU_AMD_KERNELS_H
#define GREEN_GPU_AMD_KERNELS_H

#include <Eigen/Dense>
#include <green/grids/transformer_t.h>
#include <green/ndarray/ndarray.h>
#include <green/params/params.h>
#include <green/symmetry/symmetry.h>
#include <green/utils/mpi_shared.h>
#include <green/utils/mpi_utils.h>

#include <algorithm>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../green_gpu_amd.h"
#include "amd_df_reader.h"

namespace green::gpu {

  // Same enumerators as the reference's common_defs.h:137-139, so parameter
  // files and command lines stay valid.
  enum class LinearSolverType { Cholesky, LU };

  template <size_t Dim>
  using ztensor = green::ndarray::ndarray<std::complex<double>, Dim>;

  namespace amd {

    inline void check(int rc) {
      if (rc == GGPU_OK) return;
      const std::string msg = ggpu_last_error();
      if (rc == GGPU_ERR_NOT_POSITIVE_DEFINITE || rc == GGPU_ERR_SINGULAR) throw std::domain_error(msg);
      throw std::runtime_error(msg);
    }

    // Arithmetic of the contractions: the "gpu_precision" parameter (fp64,
    // fp32, bf16x3, i8x5 or i8x4) when set, else `fallback`.
    inline int precision_param(const params::params& p, int fallback) {
      const std::string s = p["gpu_precision"].as<std::string>();
      if (s.empty()) return fallback;
      if (s == "fp64") return GGPU_PRECISION_FP64;
      if (s == "fp32") return GGPU_PRECISION_FP32;
      if (s == "bf16x3") return GGPU_PRECISION_BF16X3;
      if (s == "i8x5") return GGPU_PRECISION_INT8X5;
      if (s == "i8x4") return GGPU_PRECISION_INT8X4;
      throw std::runtime_error("gpu_precision must be fp64, fp32, bf16x3, i8x5 or i8x4, not \"" + s + "\"");
    }

    // Owned copies of every table ggpu_symmetry points to
    // (the data make_cu_symmetry_data gathers, gw_gpu_kernel.cpp:34-91,
    // plus the k-pair lists df_integral_t uses, df_integral_t.h:68-86).
    class symmetry_tables {
    public:
      template <typename BZ>
      symmetry_tables(const BZ& bz, size_t nso, size_t NQ, bool with_U_k) {
        const auto& ksym  = bz.k_symmetry();
        const auto& qsym  = bz.q_symmetry();
        const auto& kqmap = bz.k_q_map();
        for (auto x : ksym.full_to_reduced()) k_f2r_.push_back(static_cast<int64_t>(x));
        for (auto x : ksym.reduced_to_full()) k_r2f_.push_back(static_cast<int64_t>(x));
        for (auto x : ksym.tr_conj_list()) k_tr_.push_back(static_cast<int64_t>(x));
        for (auto x : qsym.full_to_reduced()) q_f2r_.push_back(static_cast<int64_t>(x));
        for (auto x : qsym.reduced_to_full()) q_r2f_.push_back(static_cast<int64_t>(x));
        for (auto x : qsym.tr_conj_list()) q_tr_.push_back(static_cast<int64_t>(x));
        const size_t nk = k_f2r_.size(), ink = k_r2f_.size(), nq = q_f2r_.size(), inq = q_r2f_.size();
        k_off_.push_back(0);
        for (size_t ik = 0; ik < ink; ++ik) {
          for (auto k : ksym.star(ik)) k_mem_.push_back(static_cast<int64_t>(k));
          k_off_.push_back(static_cast<int64_t>(k_mem_.size()));
        }
        q_off_.push_back(0);
        for (size_t iq = 0; iq < inq; ++iq) {
          for (auto q : qsym.star(iq)) q_mem_.push_back(static_cast<int64_t>(q));
          q_off_.push_back(static_cast<int64_t>(q_mem_.size()));
        }
        k1_.resize(nk * nq);
        k2_.resize(nk * nq);
        for (size_t k = 0; k < nk; ++k)
          for (size_t q = 0; q < nq; ++q) {
            k1_[k * nq + q] = static_cast<int64_t>(kqmap.k1_from_k2q(k, q));
            k2_[k * nq + q] = static_cast<int64_t>(kqmap.k2_from_k1q(k, q));
          }
        using RowMat = Eigen::Matrix<std::complex<double>, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor>;
        if (with_U_k) {
          RowMat U(nso, nso);
          U_k_.resize(nk * nso * nso);
          for (size_t k = 0; k < nk; ++k) {
            ksym.k_sym_transform_ao(U, k);
            std::memcpy(U_k_.data() + k * nso * nso, U.data(), nso * nso * sizeof(std::complex<double>));
          }
        }
        RowMat V(NQ, NQ);
        U_q_.resize(nq * NQ * NQ);
        for (size_t q = 0; q < nq; ++q) {
          qsym.q_sym_transform_p0(V, q);
          std::memcpy(U_q_.data() + q * NQ * NQ, V.data(), NQ * NQ * sizeof(std::complex<double>));
        }
        const size_t npair = nk * (nk + 1) / 2;
        for (size_t i = 0; i < npair; ++i) {
          conj_.push_back(static_cast<int64_t>(ksym.conj_kpair_list()[i]));
          trans_.push_back(static_cast<int64_t>(ksym.trans_kpair_list()[i]));
          irre_.push_back(static_cast<int64_t>(ksym.irre_pos_kpair(i)));
        }
        num_stored_ = static_cast<int64_t>(ksym.num_kpair_stored());

        s_.nk  = static_cast<int32_t>(nk);
        s_.ink = static_cast<int32_t>(ink);
        s_.nq  = static_cast<int32_t>(nq);
        s_.inq = static_cast<int32_t>(inq);
        s_.k_full_to_reduced = k_f2r_.data();
        s_.k_reduced_to_full = k_r2f_.data();
        s_.k_tr_conj         = k_tr_.data();
        s_.k_star_offsets    = k_off_.data();
        s_.k_star_members    = k_mem_.data();
        s_.q_full_to_reduced = q_f2r_.data();
        s_.q_reduced_to_full = q_r2f_.data();
        s_.q_tr_conj         = q_tr_.data();
        s_.q_star_offsets    = q_off_.data();
        s_.q_star_members    = q_mem_.data();
        s_.k1_from_k2q       = k1_.data();
        s_.k2_from_k1q       = k2_.data();
        s_.U_k               = U_k_.empty() ? nullptr : reinterpret_cast<const double*>(U_k_.data());
        s_.U_q               = reinterpret_cast<const double*>(U_q_.data());
        s_.conj_kpair        = conj_.data();
        s_.trans_kpair       = trans_.data();
        s_.irre_pos          = irre_.data();
        s_.num_kpair_stored  = num_stored_;
      }
      const ggpu_symmetry* get() const { return &s_; }
      size_t               num_stored() const { return static_cast<size_t>(num_stored_); }

    private:
      std::vector<int64_t>              k_f2r_, k_r2f_, k_tr_, q_f2r_, q_r2f_, q_tr_, k_off_, k_mem_, q_off_, q_mem_;
      std::vector<int64_t>              k1_, k2_, conj_, trans_, irre_;
      std::vector<std::complex<double>> U_k_, U_q_;
      int64_t                           num_stored_ = 0;
      ggpu_symmetry                     s_{};
    };

    // cuda_low_cpu_memory: the reference reads the chunk holding each k-pair
    // when a (k1, k2) loop needs it, caching one chunk
    // (df_integral_t::read_integrals, df_integral_t.h:61-84); here the plan
    // asks for the stored pairs of each pair batch through
    // ggpu_*_set_integral_reader, and this cache reads their chunk files.
    class chunk_cache {
    public:
      chunk_cache(const std::string& path, size_t nao, size_t NQ, size_t num_stored) :
          reader_(path, nao, NQ), num_stored_(num_stored) {}

      // ggpu_pair_reader
      static int read(void* user, int64_t pos, double* dst) noexcept {
        try {
          auto& c = *static_cast<chunk_cache*>(user);
          if (pos < 0 || static_cast<size_t>(pos) >= c.num_stored_) return 2;
          const size_t cs = c.reader_.chunk_size(), first = static_cast<size_t>(pos) / cs * cs;
          if (static_cast<long>(first) != c.current_) {
            const size_t count = std::min(cs, c.num_stored_ - first);
            c.chunk_.resize(count * c.reader_.pair_elements());
            c.reader_.read_chunk(first, count, c.chunk_.data());
            c.current_ = static_cast<long>(first);
          }
          std::memcpy(dst, c.chunk_.data() + (static_cast<size_t>(pos) - first) * c.reader_.pair_elements(),
                      c.reader_.pair_elements() * sizeof(std::complex<double>));
          return 0;
        } catch (...) {
          return 1;
        }
      }

    private:
      df_reader                         reader_;
      size_t                            num_stored_;
      long                              current_ = -1;
      std::vector<std::complex<double>> chunk_;
    };

    // HF kernel: hf_gpu_kernel::solve (hf_gpu_kernel.cpp:66-104) returns the
    // Fock contribution J + K - Ewald on the IBZ; X2C when nso == 2 nao.
    class hf_kernel {
    public:
      template <typename BZ>
      hf_kernel(const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ, double madelung, const BZ& bz,
                const ztensor<4>& S_k) :
          nso_(nso), ns_(ns), madelung_(madelung), S_k_(S_k), tables_(bz, nso, NQ, true) {
        ggpu_hf_plan* h = nullptr;
        check(ggpu_hf_plan_create(&h, static_cast<int>(nao), static_cast<int>(nso), static_cast<int>(ns),
                                  static_cast<int>(NQ), tables_.get()));
        plan_.reset(h);
        // the reference's exchange is FP64 (hf_gpu_kernel.h:67-84)
        check(ggpu_hf_set_precision(plan_.get(), precision_param(p, GGPU_PRECISION_FP64)));
        const std::string path = p["dfintegral_hf_file"].as<std::string>();
        if (p["cuda_low_cpu_memory"].as<bool>()) {
          chunks_ = std::make_unique<chunk_cache>(path, nao, NQ, tables_.num_stored());
          check(ggpu_hf_set_integral_reader(plan_.get(), &chunk_cache::read, chunks_.get()));
        } else {
          df_reader  reader(path, nao, NQ);
          const auto V = reader.read_entire(tables_.num_stored());
          check(ggpu_hf_load_integrals(plan_.get(), reinterpret_cast<const double*>(V.data())));
        }
      }

      ztensor<4> solve(const ztensor<4>& dm) {
        const auto& sh = dm.shape();
        ztensor<4>  F(sh[0], sh[1], sh[2], sh[3]);
        check(ggpu_hf_solve(plan_.get(), reinterpret_cast<const double*>(dm.data()),
                            reinterpret_cast<const double*>(S_k_.data()), madelung_,
                            reinterpret_cast<double*>(F.data())));
        return F;
      }

    private:
      std::unique_ptr<chunk_cache> chunks_;   // outlives the plan that reads through it
      std::unique_ptr<ggpu_hf_plan, void (*)(ggpu_hf_plan*)> plan_{nullptr, ggpu_hf_plan_destroy};
      size_t            nso_, ns_;
      double            madelung_;
      const ztensor<4>& S_k_;   // borrowed, as the reference stores it (hf_gpu_kernel.h:61)
      symmetry_tables   tables_;
    };

    // GW kernel: gw_gpu_kernel::solve (gw_gpu_kernel.cpp:163-205) overwrites
    // sigma with Sigma(tau) / nk; X2C when nso == 2 nao.
    class gw_kernel {
    public:
      template <typename BZ>
      gw_kernel(const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ, const grids::transformer_t& ft,
                const BZ& bz, LinearSolverType solver) : tables_(bz, nso, NQ, true) {
        const auto&  T_tw = ft.Ttn_FB();    // (nts, nw_b), row-major
        const auto&  T_wt = ft.Tnt_BF();    // (nw_b, nts - 2), row-major
        const int    nts  = static_cast<int>(ft.sd().repn_fermi().nts());
        const int    nw_b = static_cast<int>(ft.sd().repn_bose().nw());
        const bool   sp   = p["P_sp"].as<bool>() && p["Sigma_sp"].as<bool>();
        const int    prec = precision_param(p, sp ? GGPU_PRECISION_FP32 : GGPU_PRECISION_FP64);
        const int    lin  = solver == LinearSolverType::Cholesky ? GGPU_SOLVER_CHOLESKY : GGPU_SOLVER_LU;
        ggpu_gw_plan* h = nullptr;
        check(ggpu_gw_plan_create(&h, static_cast<int>(nao), static_cast<int>(nso), static_cast<int>(ns),
                                  static_cast<int>(NQ), nts, nw_b, tables_.get(),
                                  reinterpret_cast<const double*>(T_wt.data()),
                                  reinterpret_cast<const double*>(T_tw.data()), prec, lin, 0));
        plan_.reset(h);
        const std::string path = p["dfintegral_file"].as<std::string>();
        if (p["cuda_low_cpu_memory"].as<bool>()) {
          chunks_ = std::make_unique<chunk_cache>(path, nao, NQ, tables_.num_stored());
          check(ggpu_gw_set_integral_reader(plan_.get(), &chunk_cache::read, chunks_.get()));
        } else {
          df_reader  reader(path, nao, NQ);
          const auto V = reader.read_entire(tables_.num_stored());
          check(ggpu_gw_load_integrals(plan_.get(), reinterpret_cast<const double*>(V.data())));
        }
        if (p["gpu_graph"].as<bool>()) check(ggpu_gw_set_graph(plan_.get(), 1));
        autotune_ = p["gpu_autotune"].as<bool>();
      }

      template <typename G_type, typename S_type>
      void solve(G_type& g, S_type& sigma) {
        sigma.fence();
        if (!utils::context().node_rank) {
          const double* G = reinterpret_cast<const double*>(g.object().data());
          if (autotune_) {   // once, on the first iteration's own G
            check(ggpu_gw_autotune(plan_.get(), G, 2, 1));
            autotune_ = false;
          }
          check(ggpu_gw_iterate(plan_.get(), G, reinterpret_cast<double*>(sigma.object().data())));
        }
        sigma.fence();
      }

    private:
      std::unique_ptr<chunk_cache> chunks_;   // outlives the plan that reads through it
      std::unique_ptr<ggpu_gw_plan, void (*)(ggpu_gw_plan*)> plan_{nullptr, ggpu_gw_plan_destroy};
      symmetry_tables tables_;
      bool            autotune_ = false;
    };

  }  // namespace amd
}  // namespace green::gpu

#endif  // GREEN_GPU_AMD_KERNELS_H
