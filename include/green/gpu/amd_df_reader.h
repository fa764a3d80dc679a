/*
 * Reader of green-mbpt's density-fitted Coulomb integral files, feeding the
 * stored k-pair array that ggpu_hf_load_integrals / ggpu_gw_load_integrals
 * take.  Replaces df_integral_t::read_entire / read_a_chunk
 * (src/green/gpu/df_integral_t.h:53-144) on the integrating side, which
 * already links HDF5.
 *
 * File format (df_integral_t.h:53-60, :136-144):
 *   <path>/meta.h5        dataset "chunk_size" (integer): pairs per chunk
 *   <path>/VQ_<c>.h5      dataset "/<c>" of doubles, c = first pair of the
 *                         chunk (a multiple of chunk_size), shape
 *                         (chunk_size, NQ, nao, 2 nao): complex numbers as
 *                         interleaved (re, im) doubles, pairs in stored order
 * The last chunk may be shorter than chunk_size pairs; its file still holds
 * chunk_size pairs' worth of data when written by green-mbpt, and only the
 * stored pairs are copied.
 */
#ifndef GREEN_GPU_AMD_DF_READER_H
#define GREEN_GPU_AMD_DF_READER_H

#include <hdf5.h>
#include <hdf5_hl.h>

#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace green::gpu::amd {

  class df_reader {
  public:
    df_reader(std::string path, size_t nao, size_t NQ) : path_(std::move(path)), nao_(nao), NQ_(NQ) {
      hid_t file = H5Fopen((path_ + "/meta.h5").c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
      if (file < 0) throw std::runtime_error("df_reader: cannot open " + path_ + "/meta.h5");
      long cs = 0;
      const herr_t rc = H5LTread_dataset_long(file, "chunk_size", &cs);
      H5Fclose(file);
      if (rc < 0 || cs <= 0) throw std::runtime_error("df_reader: cannot read chunk_size from " + path_ + "/meta.h5");
      chunk_size_ = static_cast<size_t>(cs);
    }

    size_t chunk_size() const { return chunk_size_; }
    size_t pair_elements() const { return NQ_ * nao_ * nao_; }

    // Read pairs [c_id, c_id + count) of chunk c_id into out (count * NQ * nao^2
    // complex); c_id must be a multiple of chunk_size, count <= chunk_size.
    void read_chunk(size_t c_id, size_t count, std::complex<double>* out) {
      if (c_id % chunk_size_ != 0 || count > chunk_size_)
        throw std::runtime_error("df_reader: bad chunk request " + std::to_string(c_id));
      const std::string fname = path_ + "/VQ_" + std::to_string(c_id) + ".h5";
      const std::string dset  = "/" + std::to_string(c_id);
      hid_t file = H5Fopen(fname.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
      if (file < 0) throw std::runtime_error("df_reader: cannot open " + fname);
      int rank = 0;
      if (H5LTget_dataset_ndims(file, dset.c_str(), &rank) < 0 || rank != 4) {
        H5Fclose(file);
        throw std::runtime_error("df_reader: " + fname + dset + " is not a 4-d dataset");
      }
      hsize_t dims[4];
      H5LTget_dataset_info(file, dset.c_str(), dims, nullptr, nullptr);
      if (dims[1] != NQ_ || dims[2] != nao_ || dims[3] != 2 * nao_ || dims[0] < count) {
        H5Fclose(file);
        throw std::runtime_error("df_reader: " + fname + " has the wrong shape for NQ=" + std::to_string(NQ_) +
                                 ", nao=" + std::to_string(nao_));
      }
      buf_.resize(static_cast<size_t>(dims[0]) * pair_elements());
      const herr_t rc = H5LTread_dataset_double(file, dset.c_str(), reinterpret_cast<double*>(buf_.data()));
      H5Fclose(file);
      if (rc < 0) throw std::runtime_error("failure reading VQij with chunk id = " + std::to_string(c_id));
      std::memcpy(out, buf_.data(), count * pair_elements() * sizeof(std::complex<double>));
    }

    // All num_stored pairs, in stored order (what read_entire assembles).
    std::vector<std::complex<double>> read_entire(size_t num_stored) {
      std::vector<std::complex<double>> V(num_stored * pair_elements());
      for (size_t c = 0; c < num_stored; c += chunk_size_) {
        const size_t count = std::min(chunk_size_, num_stored - c);
        read_chunk(c, count, V.data() + c * pair_elements());
      }
      return V;
    }

  private:
    std::string                       path_;
    size_t                            nao_, NQ_, chunk_size_ = 0;
    std::vector<std::complex<double>> buf_;
  };

}  // namespace green::gpu::amd

#endif  // GREEN_GPU_AMD_DF_READER_H
