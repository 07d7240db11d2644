// ref_driver.cc -- runs the compiled r-ccs-cms/sbd tpb path on an instance file.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ from the reference
// headers where they lie (include path to the reference tree), against the image's own
// MPICH and MKL.  It writes fixture files (tests/golden/) and times the reference's CPU
// sigma build for the benchmark's CPU baseline.  The product never links or runs it.
//
// The pipeline mirrors tpb::diag (chemistry/tpb/sbdiag.h:102-456) at MPI world size 1, with
// the integrals handed over in memory (no 6-digit FCIDUMP text broadcast, fcidump.h:166).
//
// Instance file (little endian):
//   char[4] "SBDI", u32 version, i32 L, i32 nel_a, i32 nel_b, i32 words (version 1: 0, one word),
//   i64 na, i64 nb, i64 nent, u64 alpha[na * words], u64 beta[nb * words], f64 val[nent],
//   i32 idx[4*nent]; a string is `words` uint64 words, least significant first
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <iostream>
#include <map>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>
#include <unistd.h>
#include <omp.h>

#include "sbd/sbd.h"
#include "mpi.h"

namespace {

struct Inst {
  int L = 0, nel_a = 0, nel_b = 0, W = 1;
  std::vector<uint64_t> a, b;
  std::vector<double> val;
  std::vector<int32_t> idx;
};

Inst read_instance(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  char magic[4];
  uint32_t ver;
  int32_t hdr[4];
  int64_t n3[3];
  f.read(magic, 4);
  f.read(reinterpret_cast<char*>(&ver), 4);
  f.read(reinterpret_cast<char*>(hdr), 16);
  f.read(reinterpret_cast<char*>(n3), 24);
  if (std::memcmp(magic, "SBDI", 4) != 0 || (ver != 1 && ver != 2)) throw std::runtime_error("bad instance");
  Inst in;
  in.L = hdr[0];
  in.nel_a = hdr[1];
  in.nel_b = hdr[2];
  in.W = ver == 2 ? hdr[3] : 1;
  in.a.resize(n3[0] * in.W);
  in.b.resize(n3[1] * in.W);
  in.val.resize(n3[2]);
  in.idx.resize(4 * n3[2]);
  f.read(reinterpret_cast<char*>(in.a.data()), 8 * n3[0] * in.W);
  f.read(reinterpret_cast<char*>(in.b.data()), 8 * n3[1] * in.W);
  f.read(reinterpret_cast<char*>(in.val.data()), 8 * n3[2]);
  f.read(reinterpret_cast<char*>(in.idx.data()), 16 * n3[2]);
  if (!f) throw std::runtime_error("short instance file");
  return in;
}

// W-word uint64 masks (bit p = orbital p) -> reference words of bit_length bits
std::vector<std::vector<size_t>> to_words(const std::vector<uint64_t>& s, int L, size_t bl, int W) {
  size_t nw = (L + bl - 1) / bl;
  std::vector<std::vector<size_t>> out(s.size() / W, std::vector<size_t>(nw, 0));
  for (size_t k = 0; k < out.size(); ++k)
    for (int p = 0; p < L; ++p)
      if ((s[k * W + p / 64] >> (p % 64)) & 1u) out[k][p / bl] |= size_t(1) << (p % bl);
  return out;
}

template <class T>
void write_vec(const std::string& path, const std::vector<T>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), sizeof(T) * v.size());
}

std::vector<double> read_vec(const std::string& path, size_t n) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<double> v(n);
  f.read(reinterpret_cast<char*>(v.data()), 8 * n);
  if (!f) throw std::runtime_error("short vector file " + path);
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  int provided;
  MPI_Init_thread(&argc, &argv, MPI_THREAD_FUNNELED, &provided);
  MPI_Comm comm = MPI_COMM_WORLD;

  std::string inpath, outp = "ref", op = "all", vecpath, apath, bpath;
  size_t norb = 0;
  int max_it = 100, max_nb = 10, reps = 1;
  double eps = 1e-8;
  size_t bit_length = 20;
  std::vector<double> thresholds;
  std::vector<size_t> kepts;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() { return std::string(argv[++i]); };
    if (a == "--in") inpath = next();
    else if (a == "--out") outp = next();
    else if (a == "--op") op = next();
    else if (a == "--vec") vecpath = next();
    else if (a == "--iteration") max_it = std::stoi(next());
    else if (a == "--block") max_nb = std::stoi(next());
    else if (a == "--tolerance") eps = std::stod(next());
    else if (a == "--bit_length") bit_length = std::stoul(next());
    else if (a == "--reps") reps = std::stoi(next());
    else if (a == "--adets") apath = next();
    else if (a == "--bdets") bpath = next();
    else if (a == "--norb") norb = std::stoul(next());
    else if (a == "--threshold") thresholds.push_back(std::stod(next()));
    else if (a == "--kept") kepts.push_back(std::stoul(next()));
  }
  if (op == "files") {
    // determinant files through LoadAlphaDets (words written flat), and with --vec the
    // wavefunction through SaveMatrixFormWF as .bin and .txt
    std::vector<std::vector<size_t>> adet, bdet;
    sbd::LoadAlphaDets(apath, adet, bit_length, norb);
    sbd::LoadAlphaDets(bpath, bdet, bit_length, norb);
    std::vector<size_t> fa, fb;
    for (auto& v : adet) fa.insert(fa.end(), v.begin(), v.end());
    for (auto& v : bdet) fb.insert(fb.end(), v.begin(), v.end());
    write_vec(outp + ".adet.u64", fa);
    write_vec(outp + ".bdet.u64", fb);
    if (!vecpath.empty()) {
      std::vector<double> W = read_vec(vecpath, adet.size() * bdet.size());
      MPI_Comm h_comm, b_comm, t_comm;
      sbd::TaskCommunicator(comm, 1, 1, 1, 1, h_comm, b_comm, t_comm);
      sbd::SaveMatrixFormWF(outp + ".wf.bin", adet, bdet, 1, 1, h_comm, b_comm, t_comm, W);
      sbd::SaveMatrixFormWF(outp + ".wf.txt", adet, bdet, 1, 1, h_comm, b_comm, t_comm, W);
    }
    std::printf("{\"op\":\"files\",\"na\":%zu,\"nb\":%zu}\n", adet.size(), bdet.size());
    MPI_Finalize();
    return 0;
  }
  if (op == "fcidump") {
    // LoadFCIDump (framework/fcidump.h:27-95) on a text file, then SetupIntegrals
    // (chemistry/basic/makeintegrals.h:10-55): the header as "key\tvalue" lines, the integral
    // records, I0, I1 and I2
    sbd::FCIDump d = sbd::LoadFCIDump(inpath);
    {
      std::ofstream h(outp + ".header.txt", std::ios::binary);
      for (const auto& kv : d.header) h << kv.first << '\t' << kv.second << '\n';
    }
    std::vector<double> vals;
    std::vector<int32_t> idx;
    for (const auto& t : d.integrals) {
      vals.push_back(std::get<0>(t));
      idx.push_back(std::get<1>(t));
      idx.push_back(std::get<2>(t));
      idx.push_back(std::get<3>(t));
      idx.push_back(std::get<4>(t));
    }
    write_vec(outp + ".ints.f64", vals);
    write_vec(outp + ".ints.i32", idx);
    int L = 0, N = 0;
    double I0 = 0.0;
    sbd::oneInt<double> I1;
    sbd::twoInt<double> I2;
    sbd::SetupIntegrals(d, L, N, I0, I1, I2);
    std::vector<double> i0{I0};
    write_vec(outp + ".I0.f64", i0);
    write_vec(outp + ".I1.f64", I1.store);
    write_vec(outp + ".I2.f64", I2.store);
    std::printf("{\"op\":\"fcidump\",\"norb\":%d,\"nelec\":%d,\"nint\":%zu}\n", L, N,
                d.integrals.size());
    MPI_Finalize();
    return 0;
  }
  Inst inst = read_instance(inpath);

  sbd::FCIDump fcidump;
  fcidump.header["NORB"] = std::to_string(inst.L);
  fcidump.header["NELEC"] = std::to_string(inst.nel_a + inst.nel_b);
  for (size_t e = 0; e < inst.val.size(); ++e)
    fcidump.integrals.emplace_back(inst.val[e], inst.idx[4 * e], inst.idx[4 * e + 1],
                                   inst.idx[4 * e + 2], inst.idx[4 * e + 3]);
  auto adet = to_words(inst.a, inst.L, bit_length, inst.W);
  auto bdet = to_words(inst.b, inst.L, bit_length, inst.W);

  int L, N;
  double I0;
  sbd::oneInt<double> I1;
  sbd::twoInt<double> I2;
  sbd::SetupIntegrals(fcidump, L, N, I0, I1, I2);

  MPI_Comm h_comm, b_comm, t_comm;
  sbd::TaskCommunicator(comm, 1, 1, 1, 1, h_comm, b_comm, t_comm);
  std::vector<sbd::TaskHelpers> helper;
  std::vector<std::vector<size_t>> sharedMemory;
  auto t_help0 = std::chrono::high_resolution_clock::now();
  sbd::MakeHelpers(adet, bdet, bit_length, L, helper, sharedMemory, h_comm, b_comm, t_comm, 1, 1);
  sbd::RemakeHelpers(adet, bdet, bit_length, L, helper, sharedMemory, h_comm, b_comm, t_comm, 1, 1);
  double t_help = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t_help0).count();

  std::vector<double> hii;
  sbd::makeQChamDiagTerms(adet, bdet, bit_length, L, helper, I0, I1, I2, hii, h_comm, b_comm, t_comm);
  const size_t ndet = adet.size() * bdet.size();

  std::vector<std::string> ops;
  {
    std::stringstream ss(op);
    std::string tok;
    while (std::getline(ss, tok, ',')) ops.push_back(tok);
  }
  auto has = [&](const char* k) { return std::find(ops.begin(), ops.end(), k) != ops.end(); };
  auto want = [&](const char* k) { return has("all") || has(k); };

  if (want("ints")) {
    write_vec(outp + ".I1.f64", I1.store);
    write_vec(outp + ".I2.f64", I2.store);
    std::vector<double> i0{I0};
    write_vec(outp + ".I0.f64", i0);
  }
  if (want("conn")) {
    // task order at one rank: [type 2, type 1, type 0]; every task carries the full lists
    const auto& h = helper[0];
    std::vector<int64_t> out;
    auto dump = [&](size_t n, size_t* len, const std::vector<size_t*>& sm, const std::string& tag) {
      std::vector<int64_t> cnt(n), idx;
      for (size_t i = 0; i < n; ++i) {
        cnt[i] = static_cast<int64_t>(len[i]);
        for (size_t j = 0; j < len[i]; ++j) idx.push_back(static_cast<int64_t>(sm[i][j]));
      }
      write_vec(outp + "." + tag + ".cnt.i64", cnt);
      write_vec(outp + "." + tag + ".idx.i64", idx);
    };
    size_t na = h.braAlphaEnd - h.braAlphaStart, nb = h.braBetaEnd - h.braBetaStart;
    dump(na, h.SinglesFromAlphaLen, h.SinglesFromAlphaSM, "sa");
    dump(na, h.DoublesFromAlphaLen, h.DoublesFromAlphaSM, "da");
    dump(nb, h.SinglesFromBetaLen, h.SinglesFromBetaSM, "sb");
    dump(nb, h.DoublesFromBetaLen, h.DoublesFromBetaSM, "db");
    std::vector<int64_t> types;
    for (auto& t : helper) types.push_back(static_cast<int64_t>(t.taskType));
    write_vec(outp + ".tasks.i64", types);
  }
  if (want("hdiag")) write_vec(outp + ".hii.f64", hii);
  if (want("mult")) {
    std::vector<double> x = read_vec(vecpath, ndet);
    std::vector<double> y(ndet, 0.0);
    sbd::mult(hii, x, y, adet, bdet, bit_length, static_cast<size_t>(L), 1, 1, helper, I0, I1, I2,
              h_comm, b_comm, t_comm);
    write_vec(outp + ".sigma.f64", y);
  }
  if (has("timemult")) {
    std::vector<double> x(ndet), y(ndet);
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (auto& v : x) v = u(g);
    double best = 1e300, tot = 0.0;
    for (int r = 0; r < reps; ++r) {
      std::fill(y.begin(), y.end(), 0.0);
      auto t0 = std::chrono::high_resolution_clock::now();
      sbd::mult(hii, x, y, adet, bdet, bit_length, static_cast<size_t>(L), 1, 1, helper, I0, I1,
                I2, h_comm, b_comm, t_comm);
      double dt = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
      best = std::min(best, dt);
      tot += dt;
    }
    std::printf("{\"op\":\"timemult\",\"ndet\":%zu,\"threads\":%d,\"reps\":%d,\"best_s\":%.6f,"
                "\"mean_s\":%.6f,\"helper_s\":%.6f}\n",
                ndet, omp_get_max_threads(), reps, best, tot / reps, t_help);
  }
  if (has("rdmvec")) {
    // Correlation (chemistry/tpb/correlation.h:194-443) and OccupationDensity
    // (chemistry/tpb/occupation.h:10-95) of a given vector
    std::vector<double> W = read_vec(vecpath, ndet);
    std::vector<std::vector<double>> one, two;
    sbd::Correlation(W, adet, bdet, bit_length, static_cast<size_t>(L), 1, 1, helper, h_comm,
                     b_comm, t_comm, one, two);
    std::vector<double> o, t;
    for (auto& v : one) o.insert(o.end(), v.begin(), v.end());
    for (auto& v : two) t.insert(t.end(), v.begin(), v.end());
    write_vec(outp + ".one.f64", o);
    write_vec(outp + ".two.f64", t);
    std::vector<int> oIdx(L);
    std::iota(oIdx.begin(), oIdx.end(), 0);
    std::vector<double> dens;
    sbd::OccupationDensity(oIdx, W, adet, bdet, bit_length, 1, 1, b_comm, dens);
    write_vec(outp + ".density.f64", dens);
  }
  if (has("carryover")) {
    // CarryOverAlphaDet (chemistry/tpb/rdmat.h:60-132) of a given vector: the threshold form
    // for every --threshold and the kept form for every --kept (selected alpha indices, in the
    // reference's order, matched back to the sorted string table)
    std::vector<double> W = read_vec(vecpath, ndet);
    std::vector<int64_t> out;  // per case: count, then indices
    auto index_of = [&](const std::vector<size_t>& w) {
      for (size_t k = 0; k < adet.size(); ++k)
        if (adet[k] == w) return static_cast<int64_t>(k);
      return int64_t(-1);
    };
    std::vector<double> discarded;
    for (double th : thresholds) {
      std::vector<std::vector<size_t>> rdet;
      sbd::CarryOverAlphaDet(W, adet, bdet, 1, 1, b_comm, rdet, th);
      out.push_back(static_cast<int64_t>(rdet.size()));
      for (auto& w : rdet) out.push_back(index_of(w));
    }
    for (size_t kept : kepts) {
      std::vector<std::vector<size_t>> rdet;
      double dw = 0.0;
      sbd::CarryOverAlphaDet(W, adet, bdet, 1, 1, b_comm, kept, rdet, dw);
      out.push_back(static_cast<int64_t>(rdet.size()));
      for (auto& w : rdet) out.push_back(index_of(w));
      discarded.push_back(dw);
    }
    write_vec(outp + ".carry.i64", out);
    write_vec(outp + ".discarded.f64", discarded);
  }
  if (want("davidson")) {
    std::vector<double> W;
    sbd::BasisInitVector(W, adet, bdet, 1, 1, h_comm, b_comm, t_comm, 0);
    std::stringstream cap;
    auto* old = std::cout.rdbuf(cap.rdbuf());
    std::cout.precision(17);
    auto t0 = std::chrono::high_resolution_clock::now();
    sbd::Davidson(hii, W, adet, bdet, bit_length, static_cast<size_t>(L), 1, 1, helper, I0, I1, I2,
                  h_comm, b_comm, t_comm, max_it, max_nb, eps, 86400.0);
    double t_dav = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    std::cout.rdbuf(old);
    // " Davidson iteration it.ib (tol=R): E0 E1 ..."
    std::vector<double> trace, resid;
    std::string line;
    while (std::getline(cap, line)) {
      auto p = line.find("(tol=");
      if (p == std::string::npos) continue;
      auto q = line.find(")", p);
      resid.push_back(std::stod(line.substr(p + 5, q - p - 5)));
      std::istringstream es(line.substr(line.find(":", q) + 1));
      double e0;
      es >> e0;
      trace.push_back(e0);
    }
    std::vector<double> C(W.size(), 0.0);
    sbd::mult(hii, W, C, adet, bdet, bit_length, static_cast<size_t>(L), 1, 1, helper, I0, I1, I2,
              h_comm, b_comm, t_comm);
    double E = 0.0;
    sbd::InnerProduct(W, C, E, b_comm);
    write_vec(outp + ".trace.f64", trace);
    write_vec(outp + ".resid.f64", resid);
    write_vec(outp + ".W.f64", W);
    std::vector<double> e{E, t_dav};
    write_vec(outp + ".energy.f64", e);
    if (want("rdm")) {
      std::vector<std::vector<double>> one, two;
      sbd::Correlation(W, adet, bdet, bit_length, static_cast<size_t>(L), 1, 1, helper, h_comm,
                       b_comm, t_comm, one, two);
      std::vector<double> o, t;
      for (auto& v : one) o.insert(o.end(), v.begin(), v.end());
      for (auto& v : two) t.insert(t.end(), v.begin(), v.end());
      write_vec(outp + ".one.f64", o);
      write_vec(outp + ".two.f64", t);
      std::vector<int> oIdx(L);
      std::iota(oIdx.begin(), oIdx.end(), 0);
      std::vector<double> dens;
      sbd::OccupationDensity(oIdx, W, adet, bdet, bit_length, 1, 1, b_comm, dens);
      write_vec(outp + ".density.f64", dens);
    }
    std::printf("{\"op\":\"davidson\",\"energy\":%.17g,\"steps\":%zu,\"seconds\":%.6f,\"threads\":%d}\n",
                E, trace.size(), t_dav, omp_get_max_threads());
  }
  sbd::FreeHelpers(helper);
  MPI_Finalize();
  return 0;
}
