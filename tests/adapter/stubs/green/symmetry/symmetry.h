// Test stand-in for green::symmetry::brillouin_zone_utils: the accessors the
// adapter reads, backed by plain tables filled by the test.
#pragma once
#include <complex>
#include <cstring>
#include <vector>
namespace green::symmetry {
  struct table_set {
    std::vector<size_t>              f2r, r2f;
    std::vector<long>                tr;
    std::vector<std::vector<long>>   stars;
    std::vector<std::complex<double>> U;   // [n][m][m]
    size_t                           m = 0;
    std::vector<size_t>              conj, trans;
    std::vector<long>                irre;
    size_t                           nstored = 0;
    const std::vector<size_t>&       full_to_reduced() const { return f2r; }
    const std::vector<size_t>&       reduced_to_full() const { return r2f; }
    const std::vector<long>&         tr_conj_list() const { return tr; }
    const std::vector<long>&         star(size_t i) const { return stars[i]; }
    template <class Mat> void        k_sym_transform_ao(Mat& out, size_t k) const { copy(out, k); }
    template <class Mat> void        q_sym_transform_p0(Mat& out, size_t q) const { copy(out, q); }
    const std::vector<size_t>&       conj_kpair_list() const { return conj; }
    const std::vector<size_t>&       trans_kpair_list() const { return trans; }
    long                             irre_pos_kpair(size_t i) const { return irre[i]; }
    size_t                           num_kpair_stored() const { return nstored; }
    template <class Mat> void copy(Mat& out, size_t i) const {
      std::memcpy(out.data(), U.data() + i * m * m, m * m * sizeof(std::complex<double>));
    }
  };
  struct kq_map {
    size_t              nq = 0;
    std::vector<size_t> k1, k2;
    size_t              k1_from_k2q(size_t k, size_t q) const { return k1[k * nq + q]; }
    size_t              k2_from_k1q(size_t k, size_t q) const { return k2[k * nq + q]; }
  };
  class brillouin_zone_utils {
  public:
    table_set        k, q;
    kq_map           map;
    const table_set& k_symmetry() const { return k; }
    const table_set& q_symmetry() const { return q; }
    const kq_map&    k_q_map() const { return map; }
  };
}  // namespace green::symmetry
