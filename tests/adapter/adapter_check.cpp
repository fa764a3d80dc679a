// Exercises the C++ adapter (include/green/gpu/gpu_factory.h) against the
// green-* stand-ins of tests/adapter/stubs:
//   1. writes the stored integrals of a system in green-mbpt's chunked HDF5
//      format and reads them back with amd::df_reader;
//   2. converts brillouin_zone_utils tables into ggpu_symmetry and checks
//      every array;
//   3. registers the kernel parameters;
//   4. creates the HF kernel through custom_hf_kernel: without a GPU this
//      must fail with the library's device error, with one it solves and
//      writes the Fock matrices to <dir>/F.bin;
//   5. when GW inputs are present (gw_sizes.txt), writes the GW integrals in
//      the chunked format, builds a transformer_t from the given bosonic
//      transforms and runs custom_gw_kernel's callback on G (LU solver, as
//      green-mbpt's default) twice, with the integrals read whole and read
//      chunk by chunk per pair batch (cuda_low_cpu_memory): without a GPU it
//      must fail with the device error, with one it writes Sigma(tau) to
//      <dir>/Sigma.bin and <dir>/Sigma_chunks.bin.
// Input (written by tests/test_adapter.py into <dir>): sizes.txt, tables.txt,
// V.bin (stored pairs, complex128), dm.bin and S.bin (ns, ink, nso, nso);
// for GW: gw_sizes.txt (nts, nw_b, chunk), T_tw.bin (nts, nw_b), T_wt.bin
// (nw_b, nts - 2), gw_V.bin (stored pairs), gw_G.bin (nts, ns, ink, nso, nso).
// An optional second argument names an integral directory written by
// green-mbpt; its pairs are read and written to <dir>/V_ref.bin.
#include <green/gpu/gpu_factory.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>

using cd = std::complex<double>;

static int fails = 0;
#define EXPECT(c, msg)                                      \
  do {                                                      \
    if (!(c)) {                                             \
      std::cout << "FAIL: " << msg << std::endl;            \
      ++fails;                                              \
    }                                                       \
  } while (0)

static std::vector<cd> read_bin(const std::string& f, size_t n) {
  std::vector<cd> v(n);
  std::ifstream in(f, std::ios::binary);
  in.read(reinterpret_cast<char*>(v.data()), n * sizeof(cd));
  if (!in) throw std::runtime_error("short read of " + f);
  return v;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  std::map<std::string, double> sz;
  {
    std::ifstream in(dir + "/sizes.txt");
    std::string k;
    double v;
    while (in >> k >> v) sz[k] = v;
  }
  const size_t nao = sz["nao"], nso = sz["nso"], ns = sz["ns"], NQ = sz["NQ"], chunk = sz["chunk"];
  std::map<std::string, std::vector<long>> t;
  {
    std::ifstream in(dir + "/tables.txt");
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream s(line);
      std::string name;
      s >> name;
      long x;
      while (s >> x) t[name].push_back(x);
    }
  }
  auto as_size = [](const std::vector<long>& v) { return std::vector<size_t>(v.begin(), v.end()); };
  green::symmetry::brillouin_zone_utils bz;
  bz.k.f2r = as_size(t["k_f2r"]);
  bz.k.r2f = as_size(t["k_r2f"]);
  bz.k.tr  = t["k_tr"];
  bz.q.f2r = as_size(t["q_f2r"]);
  bz.q.r2f = as_size(t["q_r2f"]);
  bz.q.tr  = t["q_tr"];
  for (size_t i = 0; i < bz.k.r2f.size(); ++i) bz.k.stars.push_back(t["k_star_" + std::to_string(i)]);
  for (size_t i = 0; i < bz.q.r2f.size(); ++i) bz.q.stars.push_back(t["q_star_" + std::to_string(i)]);
  const size_t nk = bz.k.f2r.size(), nq = bz.q.f2r.size();
  bz.map.nq = nq;
  bz.map.k1 = as_size(t["k1"]);
  bz.map.k2 = as_size(t["k2"]);
  bz.k.conj    = as_size(t["conj"]);
  bz.k.trans   = as_size(t["trans"]);
  bz.k.irre    = t["irre"];
  bz.k.nstored = t["nstored"][0];
  bz.k.m       = nso;
  bz.k.U       = read_bin(dir + "/U_k.bin", nk * nso * nso);
  bz.q.m       = NQ;
  bz.q.U       = read_bin(dir + "/U_q.bin", nq * NQ * NQ);
  const size_t nstored = bz.k.nstored, pe = NQ * nao * nao;
  const auto   V = read_bin(dir + "/V.bin", nstored * pe);

  // 1. chunked HDF5 integrals, last chunk short
  const std::string df = dir + "/df";
  {
    hid_t f = H5Fcreate((df + "/meta.h5").c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    long  cs = static_cast<long>(chunk);
    H5LTmake_dataset_long(f, "chunk_size", 0, nullptr, &cs);
    H5Fclose(f);
    for (size_t c = 0; c < nstored; c += chunk) {
      const size_t  cnt     = std::min(chunk, nstored - c);
      hsize_t       dims[4] = {cnt, NQ, nao, 2 * nao};
      const std::string name = df + "/VQ_" + std::to_string(c) + ".h5";
      hid_t g = H5Fcreate(name.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
      H5LTmake_dataset_double(g, ("/" + std::to_string(c)).c_str(), 4, dims,
                              reinterpret_cast<const double*>(V.data() + c * pe));
      H5Fclose(g);
    }
  }
  green::gpu::amd::df_reader reader(df, nao, NQ);
  EXPECT(reader.chunk_size() == chunk, "chunk_size");
  const auto Vr = reader.read_entire(nstored);
  EXPECT(Vr.size() == V.size() && std::memcmp(Vr.data(), V.data(), V.size() * sizeof(cd)) == 0,
         "df_reader round trip");
  std::cout << "df_reader: " << (nstored + chunk - 1) / chunk << " chunks read" << std::endl;

  // 2. symmetry tables
  green::gpu::amd::symmetry_tables tabs(bz, nso, NQ, true);
  const ggpu_symmetry*             s = tabs.get();
  EXPECT(s->nk == (int)nk && s->ink == (int)bz.k.r2f.size() && s->nq == (int)nq && s->inq == (int)bz.q.r2f.size(),
         "counts");
  for (size_t k = 0; k < nk; ++k) {
    EXPECT(s->k_full_to_reduced[k] == (int64_t)bz.k.f2r[k], "k_full_to_reduced");
    EXPECT(s->k_tr_conj[k] == bz.k.tr[k], "k_tr_conj");
    for (size_t q = 0; q < nq; ++q) {
      EXPECT(s->k1_from_k2q[k * nq + q] == (int64_t)bz.map.k1[k * nq + q], "k1_from_k2q");
      EXPECT(s->k2_from_k1q[k * nq + q] == (int64_t)bz.map.k2[k * nq + q], "k2_from_k1q");
    }
  }
  for (size_t i = 0; i < bz.k.r2f.size(); ++i) {
    EXPECT(s->k_reduced_to_full[i] == (int64_t)bz.k.r2f[i], "k_reduced_to_full");
    const auto& st = bz.k.stars[i];
    EXPECT(s->k_star_offsets[i + 1] - s->k_star_offsets[i] == (int64_t)st.size(), "k star size");
    for (size_t j = 0; j < st.size(); ++j) EXPECT(s->k_star_members[s->k_star_offsets[i] + j] == st[j], "k star");
  }
  for (size_t i = 0; i < bz.q.r2f.size(); ++i) {
    const auto& st = bz.q.stars[i];
    for (size_t j = 0; j < st.size(); ++j) EXPECT(s->q_star_members[s->q_star_offsets[i] + j] == st[j], "q star");
  }
  for (size_t i = 0; i < nk * (nk + 1) / 2; ++i) {
    EXPECT(s->conj_kpair[i] == (int64_t)bz.k.conj[i] && s->trans_kpair[i] == (int64_t)bz.k.trans[i] &&
               s->irre_pos[i] == bz.k.irre[i],
           "pair tables");
  }
  EXPECT(s->num_kpair_stored == (int64_t)nstored, "num_kpair_stored");
  EXPECT(std::memcmp(s->U_k, bz.k.U.data(), bz.k.U.size() * sizeof(cd)) == 0, "U_k");
  EXPECT(std::memcmp(s->U_q, bz.q.U.data(), bz.q.U.size() * sizeof(cd)) == 0, "U_q");
  std::cout << "symmetry tables: checked" << std::endl;

  // 3. parameters
  green::params::params p;
  green::gpu::custom_kernel_parameters(p);
  for (const char* k : {"verbose", "cuda_linear_solver", "cuda_low_gpu_memory", "cuda_low_cpu_memory", "nt_batch",
                        "gpu_precision", "gpu_autotune", "gpu_graph"})
    EXPECT(p.defined(k), std::string("parameter ") + k);
  p.set("dfintegral_hf_file", df);

  // 4. HF kernel through the factory
  const size_t             ink = bz.k.r2f.size();
  green::gpu::ztensor<4>   S_k(ns, ink, nso, nso), dm(ns, ink, nso, nso);
  {
    auto a = read_bin(dir + "/S.bin", S_k.size()), b = read_bin(dir + "/dm.bin", dm.size());
    std::memcpy(S_k.data(), a.data(), a.size() * sizeof(cd));
    std::memcpy(dm.data(), b.data(), b.size() * sizeof(cd));
  }
  try {
    (void)green::gpu::custom_hf_kernel(nso == nao, p, nao, nso, ns, NQ, sz["madelung"], bz, S_k);
    EXPECT(false, "X2C flag that contradicts nso must be rejected");
  } catch (const std::runtime_error& e) {
    EXPECT(std::string(e.what()).find("X2C") != std::string::npos, "X2C mismatch message");
  }
  int ndev = 0;
  if (ggpu_device_count(&ndev) != GGPU_OK) ndev = 0;
  try {
    auto [kernel, solve] = green::gpu::custom_hf_kernel(nso != nao, p, nao, nso, ns, NQ, sz["madelung"], bz, S_k);
    auto F = solve(dm);
    std::ofstream out(dir + "/F.bin", std::ios::binary);
    out.write(reinterpret_cast<const char*>(F.data()), F.size() * sizeof(cd));
    std::cout << "hf kernel: solved on the device" << std::endl;
    EXPECT(ndev > 0, "kernel ran without a device");
  } catch (const std::runtime_error& e) {
    std::cout << "hf kernel: " << e.what() << std::endl;
    EXPECT(ndev == 0, std::string("kernel failed with a device present: ") + e.what());
  }
  // 5. GW kernel through the factory
  std::ifstream gws(dir + "/gw_sizes.txt");
  if (gws) {
    std::map<std::string, double> gz;
    {
      std::string k;
      double      v;
      while (gws >> k >> v) gz[k] = v;
    }
    const size_t nts = gz["nts"], nw_b = gz["nw_b"], gchunk = gz["chunk"];
    green::grids::transformer_t ft(nts, nw_b);
    {
      const auto tw = read_bin(dir + "/T_tw.bin", nts * nw_b), wt = read_bin(dir + "/T_wt.bin", nw_b * (nts - 2));
      std::memcpy(ft.Ttn_FB().data(), tw.data(), tw.size() * sizeof(cd));
      std::memcpy(ft.Tnt_BF().data(), wt.data(), wt.size() * sizeof(cd));
    }
    const auto        Vg  = read_bin(dir + "/gw_V.bin", nstored * pe);
    const std::string dfg = dir + "/df_gw";
    {
      hid_t f = H5Fcreate((dfg + "/meta.h5").c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
      long  cs = static_cast<long>(gchunk);
      H5LTmake_dataset_long(f, "chunk_size", 0, nullptr, &cs);
      H5Fclose(f);
      for (size_t c = 0; c < nstored; c += gchunk) {
        const size_t      cnt     = std::min(gchunk, nstored - c);
        hsize_t           dims[4] = {cnt, NQ, nao, 2 * nao};
        const std::string name    = dfg + "/VQ_" + std::to_string(c) + ".h5";
        hid_t             g       = H5Fcreate(name.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
        H5LTmake_dataset_double(g, ("/" + std::to_string(c)).c_str(), 4, dims,
                                reinterpret_cast<const double*>(Vg.data() + c * pe));
        H5Fclose(g);
      }
    }
    p.set("dfintegral_file", dfg);
    p.set("cuda_linear_solver", "LU");
    using G_type = green::gpu::G_type;
    green::gpu::ztensor<5> G0(nts, ns, ink, nso, nso);
    {
      const auto g = read_bin(dir + "/gw_G.bin", G0.size());
      std::memcpy(G0.data(), g.data(), g.size() * sizeof(cd));
    }
    G_type g_sh(std::move(G0));
    // the integrals read whole, then chunk by chunk per pair batch (cuda_low_cpu_memory)
    for (const char* low : {"false", "true"}) {
      p.set("cuda_low_cpu_memory", low);
      const std::string tag = std::string(low) == "true" ? " (chunked reads)" : "";
      G_type            s_sh(green::gpu::ztensor<5>(nts, ns, ink, nso, nso));
      try {
        auto [kernel, solve] = green::gpu::custom_gw_kernel(nso != nao, p, nao, nso, ns, NQ, ft, bz, S_k);
        solve(g_sh, s_sh);
        const auto&   S = s_sh.object();
        std::ofstream out(dir + (tag.empty() ? "/Sigma.bin" : "/Sigma_chunks.bin"), std::ios::binary);
        out.write(reinterpret_cast<const char*>(S.data()), S.size() * sizeof(cd));
        std::cout << "gw kernel" << tag << ": solved on the device" << std::endl;
        EXPECT(ndev > 0, "gw kernel ran without a device");
      } catch (const std::runtime_error& e) {
        std::cout << "gw kernel" << tag << ": " << e.what() << std::endl;
        EXPECT(ndev == 0, std::string("gw kernel failed with a device present: ") + e.what());
      }
    }
    p.set("cuda_low_cpu_memory", "false");
  }
  // optional: a directory written by green-mbpt itself
  if (argc > 2) {
    green::gpu::amd::df_reader ref(argv[2], nao, NQ);
    const auto                 Vref = ref.read_entire(nstored);
    std::ofstream              out(dir + "/V_ref.bin", std::ios::binary);
    out.write(reinterpret_cast<const char*>(Vref.data()), Vref.size() * sizeof(cd));
    std::cout << "df_reader: read " << argv[2] << " (chunk_size " << ref.chunk_size() << ")" << std::endl;
  }
  std::cout << (fails ? "FAILED" : "ALL OK") << std::endl;
  return fails ? 1 : 0;
}
