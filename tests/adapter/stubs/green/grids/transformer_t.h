// Test stand-in for green::grids::transformer_t: grid sizes and the two
// row-major bosonic transforms.
#pragma once
#include <Eigen/Dense>
#include <complex>
namespace green::grids {
  using cmat = Eigen::Matrix<std::complex<double>, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor>;
  struct repn {
    size_t nts_ = 0, nw_ = 0;
    size_t nts() const { return nts_; }
    size_t nw() const { return nw_; }
    size_t ni() const { return nts_ - 2; }
  };
  struct sparse_data {
    repn        fermi, bose;
    const repn& repn_fermi() const { return fermi; }
    const repn& repn_bose() const { return bose; }
  };
  class transformer_t {
  public:
    transformer_t(size_t nts, size_t nw_b) : tw_(nts, nw_b), wt_(nw_b, nts - 2) {
      sd_.fermi.nts_ = nts;
      sd_.bose.nw_   = nw_b;
      sd_.bose.nts_  = nts;
    }
    const sparse_data& sd() const { return sd_; }
    const cmat&        Ttn_FB() const { return tw_; }
    const cmat&        Tnt_BF() const { return wt_; }
    cmat&              Ttn_FB() { return tw_; }
    cmat&              Tnt_BF() { return wt_; }
  private:
    sparse_data sd_;
    cmat        tw_, wt_;
  };
}  // namespace green::grids
