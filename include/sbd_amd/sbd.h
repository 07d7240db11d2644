// sbd_amd/sbd.h -- drop-in C++ header API of r-ccs-cms/sbd's tensor-product-basis path, served
// by the MI355X engine (libsbdx.so, C ABI in sbdx.h).
//
// A caller of the reference replaces `#include "sbd/sbd.h"` by `#include "sbd_amd/sbd.h"` and
// links libsbdx.so; the functions below keep the reference's names, argument lists, argument
// meaning, output resizing and exception behaviour (ElemT = double).  The whole wavefunction
// lives on one GPU, so every communicator must have size 1 (std::invalid_argument otherwise,
// as the reference does for a communicator-size mismatch, helper.h:200-202).
//
// Reference interfaces (include/sbd/ in r-ccs-cms/sbd):
//   oneInt, twoInt                 chemistry/basic/integrals.h:10-67
//   FCIDump, LoadFCIDump, printFCIDump  framework/fcidump.h:18-152
//   SetupIntegrals                 chemistry/basic/makeintegrals.h:10-55
//   from_string, makestring,       framework/bit_manipulation.h:135-172, :324-330
//   sort_bitarray, LoadAlphaDets   chemistry/basic/makedeterminants.h:160-185 (.txt and .bin)
//   DecodeAlphaDets                chemistry/basic/makedeterminants.h:63-108
//   LoadFromAlphaDets              chemistry/basic/makedeterminants.h:110-149 (one rank)
//   change_bitlength, get_extension framework/bit_manipulation.h:863-934, :1017-1038
//   TaskHelpers, TaskCommunicator, chemistry/tpb/helper.h:12-36, :184-218, :432-574, :650-953
//   MakeHelpers, RemakeHelpers, FreeHelpers
//   makeQChamDiagTerms             chemistry/tpb/qcham.h:314-366
//   mult (on the fly)              chemistry/tpb/mult.h:387-648 (accumulates into Wb)
//   BasisInitVector                chemistry/tpb/davidson.h:101-140
//   Davidson (with/without time)   chemistry/tpb/davidson.h:412-430, :661-923
//   InnerProduct, Normalize        framework/dm_vector.h:18-50
//   Correlation                    chemistry/tpb/correlation.h:194-443
//   OccupationDensity              chemistry/tpb/occupation.h:10-95
//   CarryOverAlphaDet              chemistry/tpb/rdmat.h:60-132
//   SaveWavefunction               chemistry/tpb/restart.h:20-67 (one rank's file)
//   tpb::SBD, generate_sbd_data,   chemistry/tpb/sbdiag.h:12-85, :102-456
//   tpb::diag
//   LoadWavefunction               chemistry/tpb/restart.h:414-657 (one rank's file)
//   SaveMatrixFormWF               chemistry/tpb/restart.h:662-780 (.bin, .dat, .txt)
//   tpb::diag method 1             makeQCham + stored-H mult (qcham.h:10-312, mult.h:13-123)
//                                  as sbdx_store_hamiltonian
// Not provided here: the generic operator layer (include/sbd/basic/).
#ifndef SBD_AMD_SBD_H
#define SBD_AMD_SBD_H

#include <mpi.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "sbdx.h"

namespace sbd {

// ------------------------------------------------------------------------------ integrals

template <typename ElemT>
class oneInt {
 public:
  std::vector<ElemT> store;
  int norbs;
  inline ElemT& operator()(int i, int j) { return store.at(i * norbs + j); }
  inline ElemT Value(int i, int j) const { return store.at(i * norbs + j); }
};

template <typename ElemT>
class twoInt {
 public:
  std::vector<ElemT> store;
  ElemT maxEntry;
  ElemT zero;
  int norbs;
  std::vector<ElemT> DirectMat;
  std::vector<ElemT> ExchangeMat;
  twoInt() : maxEntry(100.0), zero(0.0) {}
  static int pair(int i, int j) { return std::max(i, j) * (std::max(i, j) + 1) / 2 + std::min(i, j); }
  inline ElemT& operator()(int i, int j, int k, int l) {
    zero = ElemT(0.0);
    if (!((i % 2 == j % 2) && (k % 2 == l % 2))) return zero;
    int a = pair(i / 2, j / 2), b = pair(k / 2, l / 2);
    return store[std::max(a, b) * (std::max(a, b) + 1) / 2 + std::min(a, b)];
  }
  inline ElemT Value(int i, int j, int k, int l) const {
    if (!((i % 2 == j % 2) && (k % 2 == l % 2))) return zero;
    int a = pair(i / 2, j / 2), b = pair(k / 2, l / 2);
    return store[std::max(a, b) * (std::max(a, b) + 1) / 2 + std::min(a, b)];
  }
  inline ElemT& Direct(int i, int j) { return DirectMat[i + norbs * j]; }
  inline ElemT& Exchange(int i, int j) { return ExchangeMat[i + norbs * j]; }
  inline ElemT DirectValue(int i, int j) const { return DirectMat[i + norbs * j]; }
  inline ElemT ExchangeValue(int i, int j) const { return ExchangeMat[i + norbs * j]; }
};

struct FCIDump {
  std::map<std::string, std::string> header;
  std::vector<std::tuple<double, int, int, int, int>> integrals;
};

// FCIDUMP text (framework/fcidump.h:27-95, read the same way so that a file gives the same
// header and integral records):
//  * blank lines and lines that start with '!' are skipped everywhere;
//  * header lines run up to the first line containing "&END" (case as written), which is
//    dropped whole; "&FCI" is cut out of a header line and the lines are joined with nothing
//    between them;
//  * the joined header splits at ','; an item holding '=' gives key (before the first '=')
//    and value (after it), both trimmed of blanks and tabs; later items replace earlier ones;
//    items without '=' (the further values of ORBSYM=1,1,...) are dropped;
//  * every later line is one "value i j k l" record.  A line that does not start with five
//    numbers is skipped (the reference stores it with indices it never set).
inline FCIDump LoadFCIDump(const std::string& filename) {
  std::ifstream in(filename);
  if (!in.is_open()) throw std::runtime_error("Failed to open FCIDUMP file.");
  FCIDump d;
  std::string line, head;
  bool in_header = true;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '!') continue;
    if (in_header) {
      if (line.find("&END") != std::string::npos) {
        in_header = false;
        continue;
      }
      const size_t fci = line.find("&FCI");
      if (fci != std::string::npos) line.erase(fci, 4);
      head += line;
      continue;
    }
    std::istringstream ls(line);
    double v = 0.0;
    int i = 0, j = 0, k = 0, l = 0;
    if (ls >> v >> i >> j >> k >> l) d.integrals.emplace_back(v, i, j, k, l);
  }
  auto trim = [](std::string t) {
    const size_t b = t.find_first_not_of(" \t");
    if (b == std::string::npos) return std::string();
    return t.substr(b, t.find_last_not_of(" \t") - b + 1);
  };
  std::stringstream hs(head);
  std::string item;
  while (std::getline(hs, item, ',')) {
    const size_t eq = item.find('=');
    if (eq == std::string::npos) continue;
    d.header[trim(item.substr(0, eq))] = trim(item.substr(eq + 1));
  }
  return d;
}

// The header ("key = value" per line, in key order) and the integral records ("value i j k l"
// per line, values at the stream's default 6 significant digits) on standard output
// (framework/fcidump.h:141-152).
inline void printFCIDump(const FCIDump& fcidump) {
  std::cout << "Header:" << std::endl;
  for (const auto& kv : fcidump.header) std::cout << kv.first << " = " << kv.second << std::endl;
  std::cout << "\nIntegrals:" << std::endl;
  for (const auto& r : fcidump.integrals)
    std::cout << std::get<0>(r) << " " << std::get<1>(r) << " " << std::get<2>(r) << " " << std::get<3>(r) << " "
              << std::get<4>(r) << std::endl;
}

template <typename ElemT>
void SetupIntegrals(const FCIDump& fcidump, int& L, int& N, ElemT& I0, oneInt<ElemT>& I1,
                    twoInt<ElemT>& I2) {
  bool have_norb = false;
  for (const auto& [key, value] : fcidump.header) {
    if (key == "NORB") {
      L = std::atoi(value.c_str());
      have_norb = true;
    }
    if (key == "NELEC") N = std::atoi(value.c_str());
  }
  // the reference goes on with whatever L held and writes out of bounds
  if (!have_norb || L <= 0) throw std::invalid_argument("SetupIntegrals: FCIDUMP header has no positive NORB");
  for (const auto& rec : fcidump.integrals) {
    const int i = std::get<1>(rec), j = std::get<2>(rec), k = std::get<3>(rec), l = std::get<4>(rec);
    auto in = [L](int p) { return p >= 1 && p <= L; };
    const bool ok = (i == 0 && j == 0 && k == 0 && l == 0) || (k == 0 && l == 0 && in(i) && in(j)) ||
                    (in(i) && in(j) && in(k) && in(l));
    if (!ok) throw std::out_of_range("SetupIntegrals: integral record with an index outside 1..NORB");
  }
  // Storage (chemistry/basic/integrals.h:10-67): I1 over spin orbitals 2p + s, row-major, with
  // the same h_pq in both spin blocks and both orders; I2 once per spatial quadruple, at the
  // packed slot of the pair of packed pairs (8-fold symmetry); the all-zero record is the core
  // energy.  Records are applied in file order, so a repeated record keeps its last value.
  I1.norbs = 2 * L;
  I1.store.assign(static_cast<size_t>(4) * L * L, ElemT(0.0));
  const size_t npair = static_cast<size_t>(L) * (L + 1) / 2;
  I2.norbs = L;
  I2.store.assign(npair * (npair + 1) / 2, ElemT(0.0));
  auto packed = [](size_t a, size_t b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; };
  auto eri_slot = [&](int p, int q, int r, int t) { return packed(packed(p, q), packed(r, t)); };
  for (const auto& rec : fcidump.integrals) {
    const ElemT v = ElemT(std::get<0>(rec));
    const int p = std::get<1>(rec) - 1, q = std::get<2>(rec) - 1;
    const int r = std::get<3>(rec) - 1, t = std::get<4>(rec) - 1;
    if (p < 0 && q < 0 && r < 0 && t < 0) {
      I0 = v;
    } else if (r < 0 && t < 0) {
      for (int spin = 0; spin < 2; ++spin) {
        I1.store[static_cast<size_t>(2 * p + spin) * I1.norbs + 2 * q + spin] = v;
        I1.store[static_cast<size_t>(2 * q + spin) * I1.norbs + 2 * p + spin] = v;
      }
    } else {
      I2.store[eri_slot(p, q, r, t)] = v;
    }
  }
  // Coulomb J_pq = (pp|qq) and exchange K_pq = (pq|qp), column-major L x L
  I2.DirectMat.assign(static_cast<size_t>(L) * L, ElemT(0.0));
  I2.ExchangeMat.assign(static_cast<size_t>(L) * L, ElemT(0.0));
  for (int q = 0; q < L; ++q)
    for (int pp = 0; pp < L; ++pp) {
      I2.DirectMat[pp + static_cast<size_t>(L) * q] = I2.store[eri_slot(pp, pp, q, q)];
      I2.ExchangeMat[pp + static_cast<size_t>(L) * q] = I2.store[eri_slot(pp, q, q, pp)];
    }
}

// ------------------------------------------------------------------------------ strings

inline std::vector<size_t> from_string(const std::string& s, size_t bit_length, size_t L) {
  std::vector<size_t> r((L + bit_length - 1) / bit_length, 0);
  for (size_t i = 0; i < L; ++i)
    if (s[L - 1 - i] == '1') r[i / bit_length] |= size_t(1) << (i % bit_length);
  return r;
}

inline std::string makestring(const std::vector<size_t>& c, size_t bit_length, size_t L) {
  std::string s;
  for (size_t i = L; i > 0; i--) s += (c[(i - 1) / bit_length] >> ((i - 1) % bit_length)) & 1 ? '1' : '0';
  return s;
}

inline bool string_less(const std::vector<size_t>& a, const std::vector<size_t>& b) {
  for (size_t n = a.size(); n > 0; n--) {
    if (a[n - 1] < b[n - 1]) return true;
    if (a[n - 1] > b[n - 1]) return false;
  }
  return false;
}

inline void sort_bitarray(std::vector<std::vector<size_t>>& a) {
  std::sort(a.begin(), a.end(), string_less);
  a.erase(std::unique(a.begin(), a.end()), a.end());
}

// lower-cased text after the last '.', empty without one (bit_manipulation.h:1017-1038)
inline std::string get_extension(const std::string& path) {
  const size_t dot = path.rfind('.');
  if (dot == std::string::npos) return std::string();
  std::string ext = path.substr(dot + 1);
  for (auto& ch : ext) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
  return ext;
}

// Repack a bit string held bit_length_a bits per word into bit_length_b bits per word; bit q of
// the string is bit q % w of word q / w, and the result has ceil(size * a / b) words
// (bit_manipulation.h:863-893, and the overload over a set of strings at :901-934).
inline void change_bitlength(size_t bit_length_a, std::vector<size_t>& b, size_t bit_length_b) {
  const size_t nbits = b.size() * bit_length_a;
  std::vector<size_t> out((nbits + bit_length_b - 1) / bit_length_b, 0);
  for (size_t q = 0; q < nbits; ++q)
    if ((b[q / bit_length_a] >> (q % bit_length_a)) & 1) out[q / bit_length_b] |= size_t(1) << (q % bit_length_b);
  b.swap(out);
}
inline void change_bitlength(size_t bit_length_a, std::vector<std::vector<size_t>>& b, size_t bit_length_b) {
  for (auto& v : b) change_bitlength(bit_length_a, v, bit_length_b);
}

// Binary determinant file: back-to-back records of ceil(norb / 8) bytes, each the string as a
// big-endian unsigned integer whose bit p is orbital p (the top 8 ceil(norb / 8) - norb bits of
// the first byte are padding and dropped); a trailing partial record is ignored.  Each string
// comes back one bit per word, entry p = orbital p (makedeterminants.h:63-108).
inline std::vector<std::vector<size_t>> DecodeAlphaDets(const std::string& filename, size_t norb) {
  std::ifstream in(filename, std::ios::binary);
  if (!in) throw std::runtime_error("Failed to open file");
  const size_t nbyte = (norb + 7) / 8;
  std::vector<std::vector<size_t>> all;
  std::vector<unsigned char> rec(nbyte);
  while (in.read(reinterpret_cast<char*>(rec.data()), static_cast<std::streamsize>(nbyte))) {
    std::vector<size_t> bits(norb);
    for (size_t p = 0; p < norb; ++p) bits[p] = (rec[nbyte - 1 - p / 8] >> (p % 8)) & 1;
    all.push_back(std::move(bits));
  }
  return all;
}

// .txt: one string per line (from_string); .bin: DecodeAlphaDets then bit_length words.  Any
// other extension loads nothing and leaves adet as given, as the reference does; either way
// the result is sorted and de-duplicated (makedeterminants.h:160-185).
inline void LoadAlphaDets(const std::string& file, std::vector<std::vector<size_t>>& adet,
                          size_t bit_length, size_t total_bit_length) {
  const std::string ext = get_extension(file);
  if (ext == "txt") {
    std::ifstream in(file);
    if (!in.is_open()) throw std::runtime_error("Failed to open alpha det file.");
    std::string line;
    adet.clear();
    while (std::getline(in, line)) adet.push_back(from_string(line, bit_length, total_bit_length));
  } else if (ext == "bin") {
    adet = DecodeAlphaDets(file, total_bit_length);
    change_bitlength(1, adet, bit_length);
  }
  sort_bitarray(adet);
}

// ------------------------------------------------------------------------------ engine glue

namespace amd_detail {

inline void check(int rc) {
  if (rc == SBDX_OK) return;
  std::string msg = std::string("libsbdx: ") + sbdx_last_error();
  if (rc == SBDX_ERR_ARG) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

inline void require_single(MPI_Comm c) {
  int s = 1;
  MPI_Comm_size(c, &s);
  if (s != 1) throw std::invalid_argument("sbd_amd: the wavefunction lives on one GPU; communicators must have size 1");
}

// 64-bit words per packed string: one up to 64 orbitals, two up to 128
inline int words_for(size_t L) {
  if (L > 128) throw std::invalid_argument("sbd_amd: at most 128 orbitals");
  return L > 64 ? 2 : 1;
}

// bit_length-bit words -> `words_for(L)` uint64 words per string, least significant first
// (the semantics of change_bitlength, framework/bit_manipulation.h:863-893)
inline std::vector<uint64_t> pack(const std::vector<std::vector<size_t>>& d, size_t bit_length, size_t L) {
  const int W = words_for(L);
  std::vector<uint64_t> out(d.size() * W, 0);
  for (size_t k = 0; k < d.size(); ++k)
    for (size_t p = 0; p < L; ++p)
      if ((d[k][p / bit_length] >> (p % bit_length)) & 1u) out[k * W + p / 64] |= uint64_t(1) << (p % 64);
  return out;
}

// One GPU context per helper set; integrals are (re)loaded when the caller's change.
struct Engine {
  sbdx_ctx* ctx = nullptr;
  size_t L = 0, na = 0, nb = 0;
  double I0 = 0.0;
  std::vector<double> I1, I2;
  bool have = false;
  // a, b: packed strings (pack), words_for(norb) words each
  Engine(size_t norb, const std::vector<uint64_t>& a, const std::vector<uint64_t>& b)
      : L(norb), na(a.size() / words_for(norb)), nb(b.size() / words_for(norb)) {
    check(sbdx_create_words(static_cast<int>(norb), words_for(norb), static_cast<int64_t>(na), a.data(),
                            static_cast<int64_t>(nb), b.data(), &ctx));
  }
  ~Engine() { sbdx_destroy(ctx); }
  void ensure(double i0, const oneInt<double>& i1, const twoInt<double>& i2) {
    if (have && i0 == I0 && i1.store == I1 && i2.store == I2) return;
    I0 = i0;
    I1 = i1.store;
    I2 = i2.store;
    check(sbdx_set_integrals(ctx, I0, I1.data(), I2.data(), static_cast<int64_t>(I2.size())));
    have = true;
  }
};

}  // namespace amd_detail

// Configurations of the generic operator layer from a binary determinant file
// (makedeterminants.h:110-149): the n strings of the file (in file order, DecodeAlphaDets)
// serve as both spins, and configuration Na * ib + ia holds the 2 norb bits of string ia
// followed by those of string ib, bit_length bits per word.  One rank: every pair.
inline void LoadFromAlphaDets(const std::string& filename, std::vector<std::vector<size_t>>& config,
                              size_t norb, size_t bit_length, MPI_Comm comm) {
  amd_detail::require_single(comm);
  const std::vector<std::vector<size_t>> dets = DecodeAlphaDets(filename, norb);
  const size_t n = dets.size();
  config.assign(n * n, std::vector<size_t>());
  std::vector<size_t> bits(2 * norb);
  for (size_t ib = 0; ib < n; ++ib)
    for (size_t ia = 0; ia < n; ++ia) {
      std::copy(dets[ia].begin(), dets[ia].end(), bits.begin());
      std::copy(dets[ib].begin(), dets[ib].end(), bits.begin() + norb);
      std::vector<size_t> w = bits;
      change_bitlength(1, w, bit_length);
      config[n * ib + ia] = std::move(w);
    }
}

// Solver used by the drop-in Davidson below; no counterpart in the reference, whose algorithm
// the defaults keep.  generate_sbd_data also reads --amd_solver block|reference,
// --amd_precision fp32|fp64, --amd_block k and --amd_num_block m into it.
struct AmdSolver {
  int method = 0;      // 0: the reference's algorithm; 1: block thick-restart Davidson
  int precision = 0;   // SBDX_FP64, or SBDX_FP32 (mixed precision, method 1)
  int block = 0;       // vectors per block (method 1)
  int num_block = 0;   // blocks per subspace (0: the caller's num_block)
};
inline AmdSolver& amd_solver() {
  static AmdSolver s;
  return s;
}

struct TaskHelpers {
  size_t braAlphaStart, braAlphaEnd, ketAlphaStart, ketAlphaEnd;
  size_t braBetaStart, braBetaEnd, ketBetaStart, ketBetaEnd;
  size_t taskType, adetShift, bdetShift;
  std::shared_ptr<amd_detail::Engine> engine;  // the GPU context holding the lists
};

inline void TaskCommunicator(MPI_Comm comm, int h_comm_size, int adet_comm_size, int bdet_comm_size,
                             int task_comm_size, MPI_Comm& h_comm, MPI_Comm& b_comm, MPI_Comm& t_comm) {
  int mpi_size = 1;
  MPI_Comm_size(comm, &mpi_size);
  if (h_comm_size * adet_comm_size * bdet_comm_size * task_comm_size != mpi_size)
    throw std::invalid_argument("MPI Size of twister is not a square of a integer");
  amd_detail::require_single(comm);
  MPI_Comm_dup(comm, &h_comm);
  MPI_Comm_dup(comm, &b_comm);
  MPI_Comm_dup(comm, &t_comm);
}

// The GPU builds the single/double lists of both spins (helper.h:132-182 semantics); at one
// rank the reference's three tasks [2, 1, 0] all span the full ranges.
inline void MakeHelpers(const std::vector<std::vector<size_t>>& adets, const std::vector<std::vector<size_t>>& bdets,
                        size_t bit_length, size_t norb, std::vector<TaskHelpers>& helper,
                        std::vector<std::vector<size_t>>& sharedMemory, MPI_Comm h_comm, MPI_Comm b_comm,
                        MPI_Comm t_comm, size_t adet_comm_size, size_t bdet_comm_size) {
  for (MPI_Comm c : {h_comm, b_comm, t_comm}) amd_detail::require_single(c);
  if (adet_comm_size != 1 || bdet_comm_size != 1) throw std::invalid_argument("sbd_amd: comm sizes must be 1");
  auto eng = std::make_shared<amd_detail::Engine>(norb, amd_detail::pack(adets, bit_length, norb),
                                                  amd_detail::pack(bdets, bit_length, norb));
  helper.clear();
  for (size_t type : {size_t(2), size_t(1), size_t(0)}) {
    TaskHelpers t{0, adets.size(), 0, adets.size(), 0, bdets.size(), 0, bdets.size(), type, 0, 0, eng};
    helper.push_back(t);
  }
  sharedMemory.assign(helper.size(), {});
}

inline void RemakeHelpers(const std::vector<std::vector<size_t>>&, const std::vector<std::vector<size_t>>&, size_t,
                          size_t, std::vector<TaskHelpers>& helper, std::vector<std::vector<size_t>>&, MPI_Comm,
                          MPI_Comm, MPI_Comm, size_t, size_t) {
  if (helper.empty() || !helper[0].engine) throw std::invalid_argument("RemakeHelpers before MakeHelpers");
}

inline void FreeHelpers(std::vector<TaskHelpers>& helper) {
  for (auto& h : helper) h.engine.reset();
}

inline amd_detail::Engine& engine_of(const std::vector<TaskHelpers>& helper) {
  if (helper.empty() || !helper[0].engine) throw std::invalid_argument("sbd_amd: helpers not built");
  return *helper[0].engine;
}

// ------------------------------------------------------------------------------ hot path

template <typename ElemT>
void makeQChamDiagTerms(const std::vector<std::vector<size_t>>& adets, const std::vector<std::vector<size_t>>& bdets,
                        const size_t bit_length, const size_t norbs, const std::vector<TaskHelpers>& helper,
                        ElemT& I0, oneInt<ElemT>& I1, twoInt<ElemT>& I2, std::vector<ElemT>& hii,
                        MPI_Comm h_comm, MPI_Comm b_comm, MPI_Comm t_comm) {
  static_assert(std::is_same<ElemT, double>::value, "sbd_amd serves ElemT = double");
  (void)adets; (void)bdets; (void)bit_length; (void)norbs; (void)h_comm; (void)b_comm; (void)t_comm;
  auto& e = engine_of(helper);
  e.ensure(I0, I1, I2);
  hii.resize(e.na * e.nb);
  amd_detail::check(sbdx_hdiag(e.ctx, hii.data()));
}

// Wb += H Wk.  H_ii comes from the engine (the same values makeQChamDiagTerms returns).
template <typename ElemT>
void mult(const std::vector<ElemT>& hii, const std::vector<ElemT>& Wk, std::vector<ElemT>& Wb,
          const std::vector<std::vector<size_t>>& adets, const std::vector<std::vector<size_t>>& bdets,
          const size_t bit_length, const size_t norbs, const size_t adet_comm_size, const size_t bdet_comm_size,
          const std::vector<TaskHelpers>& helper, const ElemT& I0, const oneInt<ElemT>& I1,
          const twoInt<ElemT>& I2, MPI_Comm h_comm, MPI_Comm b_comm, MPI_Comm t_comm) {
  static_assert(std::is_same<ElemT, double>::value, "sbd_amd serves ElemT = double");
  (void)hii; (void)adets; (void)bdets; (void)bit_length; (void)norbs; (void)adet_comm_size; (void)bdet_comm_size;
  (void)h_comm; (void)b_comm; (void)t_comm;
  auto& e = engine_of(helper);
  e.ensure(I0, I1, I2);
  std::vector<double> s(Wk.size());
  amd_detail::check(sbdx_sigma(e.ctx, 1, Wk.data(), s.data()));
  for (size_t i = 0; i < s.size(); ++i) Wb[i] += s[i];
}

template <typename ElemT>
void InnerProduct(const std::vector<ElemT>& X, const std::vector<ElemT>& Y, ElemT& res, MPI_Comm comm) {
  ElemT sum = ElemT(0.0);
  for (size_t i = 0; i < X.size(); ++i) sum += X[i] * Y[i];
  MPI_Allreduce(&sum, &res, 1, MPI_DOUBLE, MPI_SUM, comm);
}

template <typename ElemT, typename RealT>
void Normalize(std::vector<ElemT>& X, RealT& res, MPI_Comm comm) {
  RealT sum = 0.0;
  for (auto v : X) sum += v * v;
  MPI_Allreduce(&sum, &res, 1, MPI_DOUBLE, MPI_SUM, comm);
  res = std::sqrt(res);
  ElemT f = ElemT(1.0 / res);
  for (auto& v : X) v *= f;
}

template <typename ElemT>
void BasisInitVector(std::vector<ElemT>& W, const std::vector<std::vector<size_t>>& adet,
                     const std::vector<std::vector<size_t>>& bdet, const size_t adet_comm_size,
                     const size_t bdet_comm_size, MPI_Comm h_comm, MPI_Comm b_comm, MPI_Comm t_comm, int init) {
  (void)adet_comm_size; (void)bdet_comm_size; (void)h_comm; (void)b_comm; (void)t_comm;
  W.assign(adet.size() * bdet.size(), ElemT(0.0));
  if (init == 0) {
    W[0] = ElemT(1.0);
  } else if (init == 1) {
    std::mt19937 gen(static_cast<unsigned>(std::chrono::high_resolution_clock::now().time_since_epoch().count()));
    std::uniform_real_distribution<double> dist(-1, 1);
    double s = 0.0;
    for (auto& v : W) { v = dist(gen); s += v * v; }
    for (auto& v : W) v /= std::sqrt(s);
  }
}

template <typename ElemT, typename RealT>
void Davidson(const std::vector<ElemT>& hii, std::vector<ElemT>& W, const std::vector<std::vector<size_t>>& adets,
              const std::vector<std::vector<size_t>>& bdets, const size_t bit_length, const size_t norbs,
              const size_t adet_comm_size, const size_t bdet_comm_size, const std::vector<TaskHelpers>& helper,
              const ElemT& I0, const oneInt<ElemT>& I1, const twoInt<ElemT>& I2, MPI_Comm h_comm, MPI_Comm b_comm,
              MPI_Comm t_comm, int max_iteration, int num_block, RealT eps, RealT max_time) {
  static_assert(std::is_same<ElemT, double>::value, "sbd_amd serves ElemT = double");
  (void)hii; (void)adets; (void)bdets; (void)bit_length; (void)norbs; (void)adet_comm_size; (void)bdet_comm_size;
  (void)h_comm; (void)b_comm; (void)t_comm;
  auto& e = engine_of(helper);
  e.ensure(I0, I1, I2);
  sbdx_davidson_opts o;
  sbdx_davidson_defaults(&o);
  o.max_iteration = max_iteration;
  o.num_block = num_block;
  o.eps = eps;
  o.max_time = max_time;
  o.init = 1;
  o.verbose = 1;
  const AmdSolver& as = amd_solver();
  if (as.method == 1) {
    o.method = 1;
    o.precision = as.precision;
    o.block = as.block;
    if (as.num_block > 0) o.num_block = as.num_block;
  }
  sbdx_davidson_result r;
  amd_detail::check(sbdx_davidson(e.ctx, &o, W.data(), &r, nullptr, nullptr, 0));
}

template <typename ElemT, typename RealT>
void Davidson(const std::vector<ElemT>& hii, std::vector<ElemT>& W, const std::vector<std::vector<size_t>>& adets,
              const std::vector<std::vector<size_t>>& bdets, const size_t bit_length, const size_t norbs,
              const size_t adet_comm_size, const size_t bdet_comm_size, const std::vector<TaskHelpers>& helper,
              const ElemT& I0, const oneInt<ElemT>& I1, const twoInt<ElemT>& I2, MPI_Comm h_comm, MPI_Comm b_comm,
              MPI_Comm t_comm, int max_iteration, int num_block, RealT eps) {
  Davidson(hii, W, adets, bdets, bit_length, norbs, adet_comm_size, bdet_comm_size, helper, I0, I1, I2, h_comm,
           b_comm, t_comm, max_iteration, num_block, eps, RealT(86400.0));
}

template <typename ElemT>
void Correlation(const std::vector<ElemT>& W, const std::vector<std::vector<size_t>>& adet,
                 const std::vector<std::vector<size_t>>& bdet, const size_t bit_length, const size_t norb,
                 const size_t adet_comm_size, const size_t bdet_comm_size, const std::vector<TaskHelpers>& helper,
                 MPI_Comm h_comm, MPI_Comm b_comm, MPI_Comm t_comm, std::vector<std::vector<ElemT>>& onebody,
                 std::vector<std::vector<ElemT>>& twobody) {
  (void)adet; (void)bdet; (void)bit_length; (void)adet_comm_size; (void)bdet_comm_size;
  (void)h_comm; (void)b_comm; (void)t_comm;
  auto& e = engine_of(helper);
  const size_t L2 = norb * norb, L4 = L2 * L2;
  std::vector<double> one(2 * L2), two(4 * L4);
  amd_detail::check(sbdx_rdm(e.ctx, W.data(), one.data(), two.data()));
  onebody.assign(2, std::vector<ElemT>());
  twobody.assign(4, std::vector<ElemT>());
  for (int s = 0; s < 2; ++s) onebody[s].assign(one.begin() + s * L2, one.begin() + (s + 1) * L2);
  for (int s = 0; s < 4; ++s) twobody[s].assign(two.begin() + s * L4, two.begin() + (s + 1) * L4);
}

// res[2 io + s] for the orbitals in oIdx
template <typename ElemT>
void OccupationDensity(const std::vector<int>& oIdx, const std::vector<ElemT>& W,
                       const std::vector<std::vector<size_t>>& adet, const std::vector<std::vector<size_t>>& bdet,
                       const size_t bit_length, size_t adet_comm_size, size_t bdet_comm_size, MPI_Comm b_comm,
                       std::vector<ElemT>& res) {
  (void)adet_comm_size; (void)bdet_comm_size;
  amd_detail::require_single(b_comm);
  res.assign(2 * oIdx.size(), ElemT(0.0));
  const size_t nb = bdet.size();
  for (size_t ia = 0; ia < adet.size(); ++ia)
    for (size_t ib = 0; ib < nb; ++ib) {
      const double w2 = W[ia * nb + ib] * W[ia * nb + ib];
      for (size_t io = 0; io < oIdx.size(); ++io) {
        const int o = oIdx[io];
        if ((adet[ia][o / bit_length] >> (o % bit_length)) & 1u) res[2 * io] += w2;
        if ((bdet[ib][o / bit_length] >> (o % bit_length)) & 1u) res[2 * io + 1] += w2;
      }
    }
}

// ------------------------------------------------------------------------------ post-processing

template <typename ElemT, typename RealT>
void CarryOverAlphaDet(const std::vector<ElemT>& W, const std::vector<std::vector<size_t>>& adet,
                       const std::vector<std::vector<size_t>>& bdet, const size_t, const size_t, MPI_Comm,
                       std::vector<std::vector<size_t>>& rdet, RealT threshold) {
  const size_t na = adet.size(), nb = bdet.size();
  std::vector<RealT> D(na, 0.0);
  for (size_t ia = 0; ia < na; ++ia)
    for (size_t ib = 0; ib < nb; ++ib) D[ia] += W[ia * nb + ib] * W[ia * nb + ib];
  std::vector<size_t> idx(na);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](size_t i, size_t j) { return D[i] > D[j]; });
  std::vector<RealT> S(na);
  RealT sum = 0.0;
  for (size_t i = 0; i < na; ++i) S[i] = (sum += D[idx[i]]);
  // the smallest prefix whose weight exceeds 1 - threshold, plus one string; capped at every
  // string (the reference reads past its table when the total weight never exceeds the target)
  const size_t imax = std::min(
      na, static_cast<size_t>(std::upper_bound(S.begin(), S.end(), RealT(1.0) - threshold) - S.begin()) + 1);
  rdet.resize(imax, std::vector<size_t>(adet[0].size()));
  for (size_t k = 0; k < imax; ++k) rdet[k] = adet[idx[k]];
}

template <typename ElemT, typename RealT>
void CarryOverAlphaDet(const std::vector<ElemT>& W, const std::vector<std::vector<size_t>>& adet,
                       const std::vector<std::vector<size_t>>& bdet, const size_t, const size_t, MPI_Comm,
                       size_t kept, std::vector<std::vector<size_t>>& rdet, RealT& discarded_weight) {
  const size_t na = adet.size(), nb = bdet.size();
  std::vector<RealT> D(na, 0.0);
  for (size_t ia = 0; ia < na; ++ia)
    for (size_t ib = 0; ib < nb; ++ib) D[ia] += W[ia * nb + ib] * W[ia * nb + ib];
  std::vector<size_t> idx(na);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](size_t i, size_t j) { return D[i] > D[j]; });
  kept = std::min(kept, na);  // at most every string (the reference reads past its table)
  rdet.resize(kept, std::vector<size_t>(adet[0].size()));
  RealT sum = 0.0;
  for (size_t k = 0; k < kept; ++k) {
    rdet[k] = adet[idx[k]];
    sum += D[idx[k]];
  }
  discarded_weight = 1.0 - sum;
}

template <typename ElemT>
void SaveWavefunction(const std::string file, const std::vector<std::vector<size_t>>& adet,
                      const std::vector<std::vector<size_t>>& bdet, size_t, size_t, MPI_Comm, MPI_Comm,
                      MPI_Comm, std::vector<ElemT>& W) {
  size_t ar = adet.size(), br = bdet.size(), len = adet[0].size();
  std::ostringstream name;
  name << file << std::setw(6) << std::setfill('0') << 0;
  std::ofstream f(name.str(), std::ios::binary);
  f.write(reinterpret_cast<const char*>(&ar), sizeof(size_t));
  f.write(reinterpret_cast<const char*>(&br), sizeof(size_t));
  f.write(reinterpret_cast<const char*>(&len), sizeof(size_t));
  for (const auto& a : adet) f.write(reinterpret_cast<const char*>(a.data()), sizeof(size_t) * len);
  for (const auto& b : bdet) f.write(reinterpret_cast<const char*>(b.data()), sizeof(size_t) * len);
  f.write(reinterpret_cast<const char*>(W.data()), sizeof(ElemT) * W.size());
}

// restart.h:414-657 at one rank: read the rank-0 file (file + "000000") written by
// SaveWavefunction, find every saved alpha and beta string in the current basis (exact match of
// the bit_length words), and copy the saved amplitude of each pair present in both; the rest
// of W is zero.  The reference leaves W zero when the file is missing; here a W with no
// overlap at all raises std::runtime_error, since a zero start vector cannot start Davidson.
template <typename ElemT>
void LoadWavefunction(const std::string& file, const std::vector<std::vector<size_t>>& adet,
                      const std::vector<std::vector<size_t>>& bdet, const size_t adet_comm_size,
                      const size_t bdet_comm_size, const MPI_Comm h_comm, const MPI_Comm b_comm,
                      const MPI_Comm t_comm, std::vector<ElemT>& W) {
  amd_detail::require_single(h_comm);
  amd_detail::require_single(b_comm);
  amd_detail::require_single(t_comm);
  if (adet_comm_size != 1 || bdet_comm_size != 1) throw std::invalid_argument("sbd_amd: comm sizes must be 1");
  W.assign(adet.size() * bdet.size(), ElemT(0));
  std::ostringstream name;
  name << file << std::setw(6) << std::setfill('0') << 0;
  std::ifstream f(name.str(), std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("LoadWavefunction: cannot open " + name.str());
  size_t na = 0, nb = 0, len = 0;
  f.read(reinterpret_cast<char*>(&na), sizeof(size_t));
  f.read(reinterpret_cast<char*>(&nb), sizeof(size_t));
  f.read(reinterpret_cast<char*>(&len), sizeof(size_t));
  if (!f || len == 0) throw std::runtime_error("LoadWavefunction: bad header in " + name.str());
  std::vector<std::vector<size_t>> la(na, std::vector<size_t>(len)), lb(nb, std::vector<size_t>(len));
  for (auto& a : la) f.read(reinterpret_cast<char*>(a.data()), sizeof(size_t) * len);
  for (auto& b : lb) f.read(reinterpret_cast<char*>(b.data()), sizeof(size_t) * len);
  std::vector<ElemT> lw(na * nb);
  f.read(reinterpret_cast<char*>(lw.data()), sizeof(ElemT) * lw.size());
  if (!f) throw std::runtime_error("LoadWavefunction: truncated file " + name.str());
  std::map<std::vector<size_t>, size_t> ia_of, ib_of;
  for (size_t i = 0; i < adet.size(); ++i) ia_of.emplace(adet[i], i);
  for (size_t i = 0; i < bdet.size(); ++i) ib_of.emplace(bdet[i], i);
  std::vector<long long> ma(na, -1), mb(nb, -1);
  for (size_t i = 0; i < na; ++i) {
    auto it = ia_of.find(la[i]);
    if (it != ia_of.end()) ma[i] = static_cast<long long>(it->second);
  }
  for (size_t i = 0; i < nb; ++i) {
    auto it = ib_of.find(lb[i]);
    if (it != ib_of.end()) mb[i] = static_cast<long long>(it->second);
  }
  bool any = false;
  for (size_t i = 0; i < na; ++i) {
    if (ma[i] < 0) continue;
    for (size_t j = 0; j < nb; ++j) {
      if (mb[j] < 0) continue;
      const ElemT v = lw[i * nb + j];
      W[static_cast<size_t>(ma[i]) * bdet.size() + static_cast<size_t>(mb[j])] = v;
      any = any || v != ElemT(0);
    }
  }
  if (!any) throw std::runtime_error("LoadWavefunction: the saved wavefunction has no overlap with this basis");
}

// restart.h:662-780 at one rank: the wavefunction as an |adet| x |bdet| matrix, alpha index
// major.  .bin: the raw ElemT values in that order; .dat / .txt: one line "value # ia ib" per
// entry, written with the stream's default 6 significant digits (the reference raises the
// precision of std::cout, not of the file stream); other extensions write nothing.  A size
// mismatch is reported on std::cout after writing, as in the reference.
template <typename ElemT>
void SaveMatrixFormWF(const std::string filename, std::vector<std::vector<size_t>>& adet,
                      std::vector<std::vector<size_t>>& bdet, size_t adet_comm_size, size_t bdet_comm_size,
                      MPI_Comm h_comm, MPI_Comm b_comm, MPI_Comm t_comm, std::vector<ElemT>& W) {
  amd_detail::require_single(h_comm);
  amd_detail::require_single(b_comm);
  amd_detail::require_single(t_comm);
  if (adet_comm_size != 1 || bdet_comm_size != 1) throw std::invalid_argument("sbd_amd: comm sizes must be 1");
  std::cout.precision(16);
  const std::string ext = get_extension(filename);
  const size_t na = adet.size(), nb = bdet.size();
  const size_t nw = std::min(W.size(), na * nb);
  if (ext == "bin") {
    std::ofstream f(filename, std::ios::binary);
    f.write(reinterpret_cast<const char*>(W.data()), static_cast<std::streamsize>(sizeof(ElemT) * nw));
  } else if (ext == "dat" || ext == "txt") {
    std::ofstream f(filename);
    for (size_t q = 0; q < nw; ++q) f << W[q] << " # " << q / nb << " " << q % nb << std::endl;
  } else {
    return;
  }
  if (W.size() != na * nb) std::cout << " SaveMatrixFormWF: total size is not compatible " << std::endl;
}

// ------------------------------------------------------------------------------ driver

namespace tpb {

struct SBD {
  int task_comm_size = 1;
  int adet_comm_size = 1;
  int bdet_comm_size = 1;
  int h_comm_size = 1;
  int method = 0;
  int max_it = 1;
  int max_nb = 10;
  double eps = 1.0e-4;
  double max_time = 86400.0;
  int init = 0;
  int do_shuffle = 0;
  int do_rdm = 0;
  double ratio = 0.0;
  double threshold = 0.01;
  size_t bit_length = 20;
};

inline SBD generate_sbd_data(int argc, char* argv[]) {
  SBD d;
  for (int i = 0; i + 1 < argc; i++) {
    const std::string a = argv[i];
    const char* v = argv[i + 1];
    bool used = true;
    if (a == "--adet_comm_size") d.adet_comm_size = std::atoi(v);
    else if (a == "--bdet_comm_size") d.bdet_comm_size = std::atoi(v);
    else if (a == "--task_comm_size") d.task_comm_size = std::atoi(v);
    else if (a == "--method") d.method = std::atoi(v);
    else if (a == "--iteration") d.max_it = std::atoi(v);
    else if (a == "--block") d.max_nb = std::atoi(v);
    else if (a == "--tolerance") d.eps = std::atof(v);
    else if (a == "--carryover_ratio") d.ratio = std::atof(v);
    else if (a == "--carryover_threshold") d.threshold = std::atof(v);
    else if (a == "--shuffle") d.do_shuffle = std::atoi(v);
    else if (a == "--rdm") d.do_rdm = std::atoi(v);
    else if (a == "--bit_length") d.bit_length = std::atoi(v);
    else if (a == "--amd_solver") amd_solver().method = std::string(v) == "block" ? 1 : 0;
    else if (a == "--amd_precision") amd_solver().precision = std::string(v) == "fp32" ? SBDX_FP32 : SBDX_FP64;
    else if (a == "--amd_block") amd_solver().block = std::atoi(v);
    else if (a == "--amd_num_block") amd_solver().num_block = std::atoi(v);
    else used = false;
    if (used) i++;
  }
  return d;
}

// sbdiag.h:102-456 at MPI world size 1, method 0.
inline void diag(const MPI_Comm& comm, const SBD& sbd_data, const FCIDump& fcidump,
                 const std::vector<std::vector<size_t>>& adet, const std::vector<std::vector<size_t>>& bdet,
                 const std::string& loadname, const std::string& savename, double& energy,
                 std::vector<double>& density, std::vector<std::vector<size_t>>& carryover_bitstrings,
                 std::vector<std::vector<double>>& one_p_rdm, std::vector<std::vector<double>>& two_p_rdm) {
  int mpi_rank = 0;
  MPI_Comm_rank(comm, &mpi_rank);
  if (sbd_data.method != 0 && sbd_data.method != 1)
    throw std::invalid_argument("sbd_amd: method must be 0 (on the fly) or 1 (stored Hamiltonian)");
  int L = 0, N = 0;
  double I0 = 0.0;
  oneInt<double> I1;
  twoInt<double> I2;
  SetupIntegrals(fcidump, L, N, I0, I1, I2);
  MPI_Comm h_comm, b_comm, t_comm;
  TaskCommunicator(comm, sbd_data.h_comm_size, sbd_data.adet_comm_size, sbd_data.bdet_comm_size,
                   sbd_data.task_comm_size, h_comm, b_comm, t_comm);
  const size_t bl = sbd_data.bit_length;
  std::vector<TaskHelpers> helper;
  std::vector<std::vector<size_t>> shared;
  auto t0 = std::chrono::high_resolution_clock::now();
  MakeHelpers(adet, bdet, bl, L, helper, shared, h_comm, b_comm, t_comm, 1, 1);
  RemakeHelpers(adet, bdet, bl, L, helper, shared, h_comm, b_comm, t_comm, 1, 1);
  auto t1 = std::chrono::high_resolution_clock::now();
  if (mpi_rank == 0)
    std::cout << " Elapsed time for helper construction " << std::chrono::duration<double>(t1 - t0).count()
              << " (sec) " << std::endl;
  std::vector<double> W;
  if (loadname.empty()) BasisInitVector(W, adet, bdet, 1, 1, h_comm, b_comm, t_comm, sbd_data.init);
  else LoadWavefunction(loadname, adet, bdet, 1, 1, h_comm, b_comm, t_comm, W);
  std::vector<double> hii;
  makeQChamDiagTerms(adet, bdet, bl, L, helper, I0, I1, I2, hii, h_comm, b_comm, t_comm);
  if (sbd_data.method == 1) {  // makeQCham (qcham.h:10-312): H stored once, mult reads it
    int64_t nnz = 0;
    amd_detail::check(sbdx_store_hamiltonian(engine_of(helper).ctx, 1, &nnz));
    if (mpi_rank == 0) std::cout << " Stored Hamiltonian elements " << nnz << std::endl;
  }
  Davidson(hii, W, adet, bdet, bl, static_cast<size_t>(L), 1, 1, helper, I0, I1, I2, h_comm, b_comm, t_comm,
           sbd_data.max_it, sbd_data.max_nb, sbd_data.eps, sbd_data.max_time);
  std::vector<double> C(W.size(), 0.0);
  mult(hii, W, C, adet, bdet, bl, static_cast<size_t>(L), 1, 1, helper, I0, I1, I2, h_comm, b_comm, t_comm);
  double E = 0.0;
  InnerProduct(W, C, E, b_comm);
  if (mpi_rank == 0) {
    std::cout.precision(16);
    std::cout << " Energy = " << E << std::endl;
  }
  energy = E;
  if (sbd_data.do_rdm == 0) {
    std::vector<int> oIdx(L);
    std::iota(oIdx.begin(), oIdx.end(), 0);
    OccupationDensity(oIdx, W, adet, bdet, bl, 1, 1, b_comm, density);
  } else {
    Correlation(W, adet, bdet, bl, static_cast<size_t>(L), 1, 1, helper, h_comm, b_comm, t_comm, one_p_rdm, two_p_rdm);
    density.resize(2 * L);
    for (int io = 0; io < L; io++) {
      density[2 * io + 0] = one_p_rdm[0][io + L * io];
      density[2 * io + 1] = one_p_rdm[1][io + L * io];
    }
  }
  if (sbd_data.ratio == 0.0) {
    CarryOverAlphaDet(W, adet, bdet, 1, 1, b_comm, carryover_bitstrings, sbd_data.threshold);
  } else {
    size_t kept = static_cast<size_t>(sbd_data.ratio * adet.size());
    double tw = 0.0;
    CarryOverAlphaDet(W, adet, bdet, 1, 1, b_comm, kept, carryover_bitstrings, tw);
    if (mpi_rank == 0) std::cout << " truncated weight in carry-over = " << tw << std::endl;
  }
  if (!savename.empty()) SaveWavefunction(savename, adet, bdet, 1, 1, h_comm, b_comm, t_comm, W);
  FreeHelpers(helper);
  MPI_Comm_free(&h_comm);
  MPI_Comm_free(&b_comm);
  MPI_Comm_free(&t_comm);
}

}  // namespace tpb
}  // namespace sbd

#endif  // SBD_AMD_SBD_H
