// common.h -- shared definitions of libsbdx (MI355X / gfx950 selected-basis engine).
//
// Working basis.  The reference orders spin-orbitals interleaved (alpha p -> 2p, beta p ->
// 2p+1; chemistry/basic/determinants.h:98-119), so an alpha excitation's fermionic sign also
// counts beta electrons.  The engine works in the spin-blocked order instead (all alpha
// creators, then all beta): there every excitation sign depends on one spin string only, and
//   H_interleaved = S H_blocked S,   S(ia,ib) = (-1)^{#{(p in alpha, q in beta) : q < p}},
// a diagonal +-1.  Vectors are converted with S at the API boundary only; inside the Davidson
// iteration every vector stays in the blocked basis (S is orthogonal and diagonal, so dot
// products, norms and the diagonal preconditioner are unchanged).
#pragma once

#include <hip/hip_runtime.h>

#include "sbdx.h"

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace sbdx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SBDX_HIP(call)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      throw ::sbdx::Error(-3, std::string(#call) + ": " + hipGetErrorString(e_) + " at " +   \
                                  __FILE__ + ":" + std::to_string(__LINE__));                \
  } while (0)

#define SBDX_CHECK_LAUNCH() SBDX_HIP(hipGetLastError())

// Owning device allocation.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    SBDX_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  void ensure(size_t count) { if (n < count || p == nullptr) alloc(count); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  T* get() const { return p; }
  size_t bytes() const { return n * sizeof(T); }
};

// Excitation tables of one spin, CSR over bra strings x, kets in ascending index order.
// exc_s packs a single excitation: bits 0-6 annihilated orbital i (occupied in the ket),
// bits 7-13 created orbital a (occupied in the bra), bit 14 set if the blocked sign is -1
// (exc_pack / exc_i / exc_a / exc_neg).
struct SpinTab {
  int64_t n = 0;
  DevBuf<uint64_t> str;          // n strings of sbdx_ctx::W words each, least significant first
  DevBuf<int64_t> off_s, off_d;
  DevBuf<uint32_t> ket_s, ket_d, exc_s;
  DevBuf<double> val_s, val_d;   // singles: sign * (h_ai + same-spin sum); doubles: element
  DevBuf<float> val_s32, val_d32;
  DevBuf<double> e1;             // same-spin diagonal energy of each string
  int64_t nnz_s = 0, nnz_d = 0;
  int max_s = 0, max_d = 0;      // longest row of each list
  // the lists cut into nblk ket blocks of ceil(n / nblk) kets: offsets in (row, block) order,
  // entry row * nblk + b starts block b of the row (n * nblk + 1 entries), built on first use
  struct KetBlocks {
    int nblk = 0;
    DevBuf<int64_t> s, d;
  };
  KetBlocks kblk[6];
};

// Per-phase timing of sigma builds (HIP events on the launch stream), for measurement only.
enum Phase { kPhaseAlpha = 0, kPhaseMixed = 1, kPhaseTranspose = 2, kPhaseBeta = 3, kPhaseAddT = 4,
             kPhases = 5 };
struct Prof {
  bool on = false;
  hipEvent_t ev[kPhases + 1] = {};
  double ms[kPhases] = {};
  int64_t calls = 0;
};

// Beta-single entries regrouped for the mixed (alpha single x beta single) pass: columns are
// cut into tiles of consecutive beta strings holding at most cap entries; inside a tile
// entries are sorted by ket column jb.  jm = (jb, meta), meta = slot (bits 0-15: the entry's
// index in the beta list minus the tile's first index, so each column's entries have
// consecutive slots in the list's own order) | pb = b*L + j (16-29) | sign (bit 31).
struct MixTiles {
  int ntile = 0;
  int max_cols = 0;
  int64_t cap = 0;           // entries per tile (threads x entries per thread); 0: empty
  DevBuf<int64_t> ent_off;   // ntile + 1
  DevBuf<int32_t> col0, ncol;
  DevBuf<uint2> jm;
};

struct Workspace {
  DevBuf<double> partial;        // block partial sums of reductions
  DevBuf<double> scal;           // device scalars
  DevBuf<unsigned char> xT, yT;  // transposed copies for the beta part (any precision)
  DevBuf<unsigned char> tmp_in, tmp_out;
  DevBuf<unsigned char> sub_x, sub_y;  // one sub-block of a block built in sub-blocks
};

}  // namespace sbdx

struct sbdx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int L = 0;
  int W = 1;                       // 64-bit words per string: 1 (L <= 64) or 2 (L <= 128)
  int64_t na = 0, nb = 0;
  sbdx::SpinTab A, B;
  sbdx::MixTiles mix[6];           // mixed-pass tiles, one per tile size in use
  bool have_ints = false;
  double I0 = 0.0;
  std::vector<double> h_I1, h_I2;
  sbdx::DevBuf<double> h1a, h1b;   // L x L one-electron blocks (alpha, beta)
  sbdx::DevBuf<double> eri4;       // dense (pq|rs), index ((p L + q) L + r) L + s
  sbdx::DevBuf<float> eri4_32;
  sbdx::DevBuf<double> hii;        // H_ii, blocked and interleaved bases agree
  sbdx::DevBuf<float> hii32;
  // Coulomb tables of the same-spin singles' spectator term: coul_a[p * nb + ib] =
  // sum_{j in beta(ib)} (p|jj) for every orbital pair p = a L + i (alpha-row pass), coul_b
  // [p * na + ia] = sum_{j in alpha(ia)} (p|jj) (beta-row pass); j ascending
  sbdx::DevBuf<double> coul_a, coul_b;
  sbdx::DevBuf<float> coul_a32, coul_b32;
  sbdx::Workspace ws;
  sbdx::Prof prof;
  int mix_variant_tuned[3][4] = {}; // autotuned mixed-kernel variant [FP64|FP32|FP32 two-sum][K slot]
  int fp32_acc = SBDX_ACC_FP64;     // accumulation of FP32 sigma builds (sbdx_set_fp32_accumulation)
  bool stored = false;             // H stored as CSR (stored.hip): sigma builds use it
  // determinant 0's block of H, for the block solver's start vectors: per spin, 1 for the
  // strings in the connected component of string 0 (strings joined by a single or double
  // excitation); state 0 = not built, 1 = both spins connected (no masks), 2 = masks built
  int block0_state = 0;
  sbdx::DevBuf<uint8_t> block0_a, block0_b;
  // set while sbdx_sigma_device_noalloc runs: the sigma path then takes its transposed pair
  // from noalloc_xT/yT and treats any allocation it would make as an error
  bool no_alloc = false;
  void* noalloc_xT = nullptr;
  void* noalloc_yT = nullptr;
  sbdx::DevBuf<int64_t> h_rowptr;
  sbdx::DevBuf<uint32_t> h_col;
  sbdx::DevBuf<double> h_val;
  sbdx::DevBuf<float> h_val32;
  int64_t h_nnz = 0;
  bool have_mlp = false;           // learned preconditioner set (sbdx_mlp_set)
  sbdx::DevBuf<float> mlp_w;       // its parameters, rounded to bfloat16
  sbdx::DevBuf<float> mlp_pa, mlp_pb;  // per-string first-layer sums (alpha includes b1)
  int64_t ndet() const { return na * nb; }
};

// a group of contexts on one device solved together (batch.hip)
struct sbdx_batch {
  std::vector<sbdx_ctx*> ctx;
  int device = 0;
  hipStream_t stream = nullptr;
};

namespace sbdx {

// ----------------------------------------------------------------------------- bit helpers

__host__ __device__ inline int popc64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __popcll(x);
#else
  return __builtin_popcountll(x);
#endif
}

__host__ __device__ inline int ctz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __ffsll(static_cast<unsigned long long>(x)) - 1;
#else
  return __builtin_ctzll(x);
#endif
}

// bits in [lo, hi), 0 <= lo <= hi <= 64
__host__ __device__ inline uint64_t mask_range(int lo, int hi) {
  uint64_t up = hi >= 64 ? ~0ull : ((1ull << hi) - 1ull);
  uint64_t dn = lo >= 64 ? ~0ull : ((1ull << lo) - 1ull);
  return up & ~dn;
}

// Sign of a single excitation i -> a applied to ket string y (one spin, blocked order):
// (-1)^{occupied orbitals strictly between i and a}.  Equals the reference's parity
// (determinants.h:205-256, :408) evaluated on a one-spin determinant.
__host__ __device__ inline int sign_single(uint64_t y, int i, int a) {
  int lo = i < a ? i : a, hi = i < a ? a : i;
  return (popc64(y & mask_range(lo + 1, hi)) & 1) ? -1 : 1;
}

// determinants.h:205-256 parity on a one-spin string: [st, en) popcount, flip if st set
__host__ __device__ inline int parity_ref(uint64_t y, int st, int en) {
  int s = (popc64(y & mask_range(st, en)) & 1) ? -1 : 1;
  return ((y >> st) & 1ull) ? -s : s;
}

// determinants.h:424-442 on a one-spin string: I<J annihilated (ket), A<B created (bra)
__host__ __device__ inline int sign_double(uint64_t y, int I, int J, int A, int B) {
  int s = parity_ref(y, I < A ? I : A, I < A ? A : I);
  s *= parity_ref(y, J < B ? J : B, J < B ? B : J);
  if (A > J || B < I) s = -s;
  return s;
}

// Interleaved <-> blocked sign of determinant (a, b).
__host__ __device__ inline int det_sign(uint64_t a, uint64_t b) {
  int c = 0;
  while (a) {
    int p = ctz64(a);
    c += popc64(b & mask_range(0, p));
    a &= a - 1;
  }
  return (c & 1) ? -1 : 1;
}

// ----------------------------------------------------------------------------- multi-word strings
// A string of W 64-bit words, orbital p at bit p % 64 of word p / 64 (W = 1: L <= 64, W = 2:
// L <= 128).  The kernels that read strings take W as a template argument; with W = 1 every
// helper below is the one-word expression above.
template <int W>
struct Str {
  uint64_t w[W];
};

template <int W>
__host__ __device__ inline Str<W> ld_str(const uint64_t* s, int64_t i) {
  Str<W> r;
#pragma unroll
  for (int k = 0; k < W; ++k) r.w[k] = s[i * W + k];
  return r;
}
template <int W>
__host__ __device__ inline int popc(const Str<W>& a) {
  int c = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) c += popc64(a.w[k]);
  return c;
}
template <int W>
__host__ __device__ inline Str<W> andn(const Str<W>& a, const Str<W>& b) {  // a & ~b
  Str<W> r;
#pragma unroll
  for (int k = 0; k < W; ++k) r.w[k] = a.w[k] & ~b.w[k];
  return r;
}
template <int W>
__host__ __device__ inline bool test(const Str<W>& a, int p) {
  return (a.w[p >> 6] >> (p & 63)) & 1ull;
}
// lowest set orbital (the string is not empty)
template <int W>
__host__ __device__ inline int first(const Str<W>& a) {
#pragma unroll
  for (int k = 0; k < W - 1; ++k)
    if (a.w[k]) return 64 * k + ctz64(a.w[k]);
  return 64 * (W - 1) + ctz64(a.w[W - 1]);
}
// the string without its lowest set orbital
template <int W>
__host__ __device__ inline Str<W> drop_first(Str<W> a) {
#pragma unroll
  for (int k = 0; k < W; ++k)
    if (a.w[k]) {
      a.w[k] &= a.w[k] - 1;
      break;
    }
  return a;
}
// f(p) for every set orbital p, ascending
template <int W, class F>
__host__ __device__ inline void for_each_bit(const Str<W>& a, F&& f) {
#pragma unroll
  for (int k = 0; k < W; ++k)
    for (uint64_t o = a.w[k]; o; o &= o - 1) f(64 * k + ctz64(o));
}
// set orbitals in [lo, hi), 0 <= lo <= hi <= 64 W
template <int W>
__host__ __device__ inline int count_range(const Str<W>& a, int lo, int hi) {
  int c = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int l = lo - 64 * k, h = hi - 64 * k;
    c += popc64(a.w[k] & mask_range(l < 0 ? 0 : (l > 64 ? 64 : l), h < 0 ? 0 : (h > 64 ? 64 : h)));
  }
  return c;
}
template <int W>
__host__ __device__ inline int sign_single(const Str<W>& y, int i, int a) {
  const int lo = i < a ? i : a, hi = i < a ? a : i;
  return (count_range(y, lo + 1, hi) & 1) ? -1 : 1;
}
template <int W>
__host__ __device__ inline int parity_ref(const Str<W>& y, int st, int en) {
  const int s = (count_range(y, st, en) & 1) ? -1 : 1;
  return test(y, st) ? -s : s;
}
template <int W>
__host__ __device__ inline int sign_double(const Str<W>& y, int I, int J, int A, int B) {
  int s = parity_ref(y, I < A ? I : A, I < A ? A : I);
  s *= parity_ref(y, J < B ? J : B, J < B ? B : J);
  if (A > J || B < I) s = -s;
  return s;
}
template <int W>
__host__ __device__ inline int det_sign(const Str<W>& a, const Str<W>& b) {
  int c = 0;
  for_each_bit(a, [&](int p) { c += count_range(b, 0, p); });
  return (c & 1) ? -1 : 1;
}

// f(std::integral_constant<int, W>) for a context's string width
template <class F>
inline void with_words(int W, F&& f) {
  if (W == 1) f(std::integral_constant<int, 1>{});
  else f(std::integral_constant<int, 2>{});
}

// packed single excitation (SpinTab::exc_s)
__host__ __device__ inline uint32_t exc_pack(int i, int a, bool neg) {
  return static_cast<uint32_t>(i) | (static_cast<uint32_t>(a) << 7) | (neg ? (1u << 14) : 0u);
}
__host__ __device__ inline int exc_i(uint32_t e) { return static_cast<int>(e & 127u); }
__host__ __device__ inline int exc_a(uint32_t e) { return static_cast<int>((e >> 7) & 127u); }
__host__ __device__ inline bool exc_neg(uint32_t e) { return (e >> 14) & 1u; }
// orbital pair a L + i of the excitation: the row of its integrals (a i | . .)
__host__ __device__ inline uint32_t exc_pair(uint32_t e, int L) {
  return static_cast<uint32_t>(exc_a(e)) * static_cast<uint32_t>(L) + static_cast<uint32_t>(exc_i(e));
}

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// index of a block width 1, 2, 4, 8
inline int k_slot(int k) { return k == 1 ? 0 : k == 2 ? 1 : k == 4 ? 2 : 3; }

// the tuned mixed-kernel variant of a precision, block width and the context's current FP32
// accumulation (each accumulation keeps its own choice); 0: none
inline int& tuned_variant(sbdx_ctx& c, int prec, int k) {
  return c.mix_variant_tuned[prec == SBDX_FP64 ? 0 : (c.fp32_acc == SBDX_ACC_TWOSUM ? 2 : 1)][k_slot(k)];
}
inline int tuned_variant(const sbdx_ctx& c, int prec, int k) {
  return c.mix_variant_tuned[prec == SBDX_FP64 ? 0 : (c.fp32_acc == SBDX_ACC_TWOSUM ? 2 : 1)][k_slot(k)];
}

// ----------------------------------------------------------------------------- launchers

void build_tables(sbdx_ctx& c, const uint64_t* alpha, const uint64_t* beta);  // tables.hip
void build_integral_tables(sbdx_ctx& c);        // tables.hip: values, H_ii, S
void ensure_fp32(sbdx_ctx& c);                  // tables.hip: float copies
void reset_mix_tiles(sbdx_ctx& c);              // sigma.hip: drop cached mixed-pass tiles
void prepare_mix_tiles(sbdx_ctx& c, int prec, int k);  // sigma.hip: build the default variant's tiles
void* noalloc_scratch(int device);              // noalloc.hip: device address of the scratch array
std::vector<int> mix_variant_candidates(int k); // sigma.hip: autotuner search space
bool mix_variant_fits(const sbdx_ctx& c, int prec, int k, int variant);  // its LDS fits a CU
// sigma of nvec vectors stored one after another (vector-major)
void sigma_blocked(sbdx_ctx& c, int prec, int nvec, const void* x, void* y, hipStream_t s);
// fused sigma of a block of k vectors stored interleaved, element [det * k + v], k in {1,2,4,8}
void sigma_block(sbdx_ctx& c, int prec, int k, const void* x, void* y, hipStream_t s);
void apply_sign(sbdx_ctx& c, int prec, int nvec, const void* x, void* y, hipStream_t s);
void apply_sign_block(sbdx_ctx& c, int prec, int k, const void* x, void* y, hipStream_t s);
// stored Hamiltonian (stored.hip)
int64_t store_hamiltonian(sbdx_ctx& c);
void drop_stored(sbdx_ctx& c);
void sigma_stored(sbdx_ctx& c, int prec, int k, const void* x, void* y, hipStream_t s);
// batched lockstep Davidson of many small problems (batch.hip)
void batch_davidson(sbdx_batch& b, const sbdx_davidson_opts& o, double* const* W, sbdx_davidson_result* res);
// learned preconditioner (mlp.hip): T[det*k+v] = R/(theta_v - H_ii) (1 + m), P = float|double
void mlp_precondition(sbdx_ctx& c, int prec, int k, const double* R, const double* theta,
                      const double* rms, void* T, hipStream_t s);

}  // namespace sbdx
