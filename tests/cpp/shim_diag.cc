// Drives the drop-in C++ API (include/sbd_amd/sbd.h) the way the reference's sample driver
// drives sbd (samples/selected_basis_diagonalization/main.cc:17-226): integrals handed over as
// an in-memory FCIDump, strings as bit_length-20 words, then sbd::tpb::diag, sbd::mult and
// sbd::Correlation.  Test program: tests/test_gpu_shim.py runs it and checks the outputs
// against the CPU oracle.
//
//   shim_diag <instance file> <out prefix> [tolerance]
// Instance file: char[4] "SBDI", u32 version, i32 L, nel_a, nel_b, words (version 1: one),
//                i64 na, nb, nent, u64 alpha[na words], u64 beta[nb words], f64 val[nent],
//                i32 idx[4 nent]; a string is `words` uint64 words, least significant first
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "sbd_amd/sbd.h"

int main(int argc, char** argv) {
  int provided;
  MPI_Init_thread(&argc, &argv, MPI_THREAD_FUNNELED, &provided);
  if (argc < 3) {
    std::fprintf(stderr, "usage: shim_diag <instance> <out prefix> [tolerance]\n");
    return 2;
  }
  std::ifstream f(argv[1], std::ios::binary);
  char magic[4];
  uint32_t ver;
  int32_t hdr[4];
  int64_t n3[3];
  f.read(magic, 4);
  f.read(reinterpret_cast<char*>(&ver), 4);
  f.read(reinterpret_cast<char*>(hdr), 16);
  f.read(reinterpret_cast<char*>(n3), 24);
  const int W = ver == 2 ? hdr[3] : 1;
  std::vector<uint64_t> a(n3[0] * W), b(n3[1] * W);
  std::vector<double> val(n3[2]);
  std::vector<int32_t> idx(4 * n3[2]);
  f.read(reinterpret_cast<char*>(a.data()), 8 * n3[0] * W);
  f.read(reinterpret_cast<char*>(b.data()), 8 * n3[1] * W);
  f.read(reinterpret_cast<char*>(val.data()), 8 * n3[2]);
  f.read(reinterpret_cast<char*>(idx.data()), 16 * n3[2]);
  if (!f) return 3;
  const int L = hdr[0];
  const std::string out = argv[2];

  sbd::FCIDump fcidump;
  fcidump.header["NORB"] = std::to_string(L);
  fcidump.header["NELEC"] = std::to_string(hdr[1] + hdr[2]);
  for (size_t e = 0; e < val.size(); ++e)
    fcidump.integrals.emplace_back(val[e], idx[4 * e], idx[4 * e + 1], idx[4 * e + 2], idx[4 * e + 3]);
  auto words = [&](const std::vector<uint64_t>& s) {
    std::vector<std::vector<size_t>> r;
    for (size_t k = 0; k < s.size() / W; ++k) {
      std::string str;
      for (int p = L - 1; p >= 0; --p) str += ((s[k * W + p / 64] >> (p % 64)) & 1u) ? '1' : '0';
      r.push_back(sbd::from_string(str, 20, L));
    }
    return r;
  };
  auto adet = words(a), bdet = words(b);

  // extra arguments after the tolerance are options in the sample driver's form
  // (e.g. --amd_solver block --amd_precision fp32)
  sbd::tpb::SBD d = sbd::tpb::generate_sbd_data(argc, argv);
  d.max_it = 200;
  d.max_nb = 10;
  d.eps = argc > 3 ? std::atof(argv[3]) : 1e-8;
  d.do_rdm = 1;
  double energy = 0.0;
  std::vector<double> density;
  std::vector<std::vector<size_t>> co;
  std::vector<std::vector<double>> one, two;
  // --load <file>: start from a saved wavefunction (LoadWavefunction) instead of HF
  std::string load;
  for (int i = 4; i + 1 < argc; ++i)
    if (std::string(argv[i]) == "--load") load = argv[i + 1];
  sbd::tpb::diag(MPI_COMM_WORLD, d, fcidump, adet, bdet, load, out + ".wf", energy, density, co, one, two);

  // a second use of the API pieces: mult on a fixed vector, accumulated onto ones
  int LL, N;
  double I0;
  sbd::oneInt<double> I1;
  sbd::twoInt<double> I2;
  sbd::SetupIntegrals(fcidump, LL, N, I0, I1, I2);
  MPI_Comm h, bc, t;
  sbd::TaskCommunicator(MPI_COMM_WORLD, 1, 1, 1, 1, h, bc, t);
  std::vector<sbd::TaskHelpers> helper;
  std::vector<std::vector<size_t>> sm;
  sbd::MakeHelpers(adet, bdet, 20, L, helper, sm, h, bc, t, 1, 1);
  std::vector<double> hii;
  sbd::makeQChamDiagTerms(adet, bdet, 20, L, helper, I0, I1, I2, hii, h, bc, t);
  std::vector<double> x(hii.size()), y(hii.size(), 1.0);
  for (size_t i = 0; i < x.size(); ++i) x[i] = std::cos(0.37 * static_cast<double>(i));
  sbd::mult(hii, x, y, adet, bdet, 20, L, 1, 1, helper, I0, I1, I2, h, bc, t);
  sbd::FreeHelpers(helper);

  auto dump = [&](const std::string& tag, const std::vector<double>& v) {
    std::ofstream o(out + "." + tag, std::ios::binary);
    o.write(reinterpret_cast<const char*>(v.data()), 8 * v.size());
  };
  std::vector<double> o1, o2;
  for (auto& v : one) o1.insert(o1.end(), v.begin(), v.end());
  for (auto& v : two) o2.insert(o2.end(), v.begin(), v.end());
  dump("one", o1);
  dump("two", o2);
  dump("density", density);
  dump("hii", hii);
  dump("multy", y);
  std::printf("{\"energy\": %.17g, \"carryover\": %zu, \"ndet\": %zu}\n", energy, co.size(), hii.size());
  MPI_Finalize();
  return 0;
}
