// sigma.hip -- the on-the-fly H C product (chemistry/tpb/mult.h:387-648) in the blocked basis.
//
// sigma = D C + H_aa C + H_bb C + H_ab C, with C an n_alpha x n_beta row-major matrix, for a
// block of K vectors stored interleaved (element [det * K + v]); K = 1 is the single-vector
// case.  One gather then loads K contiguous values, and list decoding, integral lookups and
// matrix-element arithmetic are shared by the K vectors.
//
//  * Row pass (alpha rows, task type 2 + diagonal): one workgroup = one wavefront per (alpha
//    row ia, 64 C consecutive beta columns, C = 2 where a K-value unit is 8 bytes), so the
//    row's excitation lists are wavefront-uniform (scalar loads) and every gathered operand
//    C[ja, ib..ib+64C-1] is one coalesced segment; the pass runs in ket blocks whose segments
//    fit an XCD's L2.
//  * Mixed pass (task type 0): k_sigma_mixed, described at the kernel.
//  * Beta part (task type 1): C is transposed, the same row kernel runs with the beta tables,
//    and the result is added back transposed.
// Every pass updates sigma by its owner; there are no global atomics.
//
// Matrix elements (blocked order, signs from one spin string each):
//   alpha single ja -> ia (i -> a):  s_a (F_a + sum_{j in beta(ib)} (ai|jj))
//   alpha double:                     stored value (sign and exchange included)
//   alpha single x beta single:       s_a s_b (ai|bj)
// Arithmetic: FP64 (FMA), or FP32 data whose products (exact in FP64) are accumulated in FP64
// and rounded to FP32 once per pass (alpha rows, mixed, beta rows) when the pass adds its sum
// to the output element, or (SBDX_ACC_TWOSUM) FP32 data accumulated in FP32 with compensated
// partials (runs of up to 16 FMA terms, each run added to the sum by TwoSum).  Every output
// element is summed in a fixed order, so a build is bitwise reproducible.
#include "common.h"

#include <vector>

#include <algorithm>
#include <cstdlib>
#include <utility>

namespace sbdx {
namespace {

// Lanes per workgroup of the row kernels: one wavefront.  Every row gathers a segment of
// kRowTile x C columns from each of its ~750 ket rows, and the rows in flight on an XCD reach
// nearly all ket rows, so the segment size sets the L2 footprint (c2, FP64: 20,000 ket rows x
// 512 B = 10 MB at 64 columns against 40 MB at 256); the ket blocks bound it further.
constexpr int kRowTile = 64;

struct AccD {
  using Wide = double;
  double s = 0.0;
  __device__ inline void add(double p) { s += p; }
  __device__ inline void fma(double a, double b) { s = __fma_rn(a, b, s); }
  __device__ inline void flush() {}
  __device__ inline double get() const { return s; }
  __device__ inline Wide wide() const { return s; }
  __device__ inline double apply(double y) const { return y + s; }
};

// FP32 data, FP64 accumulation: the product of two FP32 numbers is exact in FP64 (24 + 24
// significand bits), so every term enters the sum unrounded and only the FP64 additions round;
// one conversion per operand and one FP64 FMA per term, against ten FP32 operations for an
// error-free FP32 product plus TwoSum compensation of similar accuracy.  Partial sums that are
// combined (the mixed pass's per-entry sums) stay in FP64; each pass rounds to FP32 once, when
// it adds its sum to the output element.
struct AccX {
  using Wide = double;
  double s = 0.0;
  __device__ inline void fma(float a, float b) {
    s = __fma_rn(static_cast<double>(a), static_cast<double>(b), s);
  }
  __device__ inline void add(double p) { s += p; }
  __device__ inline void flush() {}
  __device__ inline float get() const { return static_cast<float>(s); }
  __device__ inline Wide wide() const { return s; }
  __device__ inline float apply(float y) const { return static_cast<float>(static_cast<double>(y) + s); }
};

// compensated FP32 (SBDX_ACC_TWOSUM): the terms of a short run (one batch of gathers, at most
// 16 terms) are summed with FMA into the FP32 partial p, and flush() adds p to the running sum
// s exactly, as s' + err (TwoSum), the rounding errors gathered in c.  Partials from elsewhere
// (the previous ket block's sum, the mixed pass's per-entry sums) enter the same way.  One FMA
// per term and one TwoSum per run.
struct AccF {
  using Wide = float;
  float s = 0.0f, c = 0.0f, p = 0.0f;
  __device__ inline void add(float q) {
    const float t = __fadd_rn(s, q);
    const float bq = __fsub_rn(t, s);
    c += __fadd_rn(__fsub_rn(s, __fsub_rn(t, bq)), __fsub_rn(q, bq));
    s = t;
  }
  __device__ inline void fma(float a, float b) { p = __fmaf_rn(a, b, p); }
  __device__ inline void flush() {
    add(p);
    p = 0.0f;
  }
  __device__ inline float get() const { return s + (c + p); }
  __device__ inline Wide wide() const { return s + (c + p); }
  __device__ inline float apply(float y) {
    flush();
    add(y);
    return get();
  }
};

// accumulator of a precision and FP32 accumulation mode
template <typename T, int kAcc> struct AccOf;
template <int kAcc> struct AccOf<double, kAcc> { using type = AccD; };
template <> struct AccOf<float, SBDX_ACC_FP64> { using type = AccX; };
template <> struct AccOf<float, SBDX_ACC_TWOSUM> { using type = AccF; };

// K contiguous values, loaded and stored as whole 16-byte words where possible
template <typename T, int K>
struct alignas(K * sizeof(T) >= 16 ? 16 : K * sizeof(T)) VecK {
  T v[K];
};

template <typename T, int K>
__device__ inline VecK<T, K> ldk(const T* p) {
  return *reinterpret_cast<const VecK<T, K>*>(p);
}

template <typename T, int K>
__device__ inline void stk(T* p, const VecK<T, K>& a) {
  *reinterpret_cast<VecK<T, K>*>(p) = a;
}

// Raw buffer resource over `bytes` bytes at p (stride 0; dword 3 = 0x00020000, the GFX9-family
// raw-buffer format): a load addresses it with a 32-bit per-lane byte offset and no 64-bit
// address arithmetic, and a load past `bytes` returns zeros.
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), static_cast<short>(0), static_cast<int>(bytes),
                                           0x00020000);
}

// K contiguous values at byte offset off of a buffer resource
template <typename T, int K>
__device__ inline VecK<T, K> ldk_buf(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  constexpr int B = K * static_cast<int>(sizeof(T));
  VecK<T, K> a;
  if constexpr (B == 4) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    __builtin_memcpy(&a, &v, 4);
  } else if constexpr (B == 8) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    __builtin_memcpy(&a, &v, 8);
  } else {
#pragma unroll
    for (int c = 0; c < B / 16; ++c) {
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * c, 0, 0);
      __builtin_memcpy(reinterpret_cast<char*>(&a) + 16 * c, &v, 16);
    }
  }
  return a;
}

struct RowTabs {
  const int64_t* off_s;
  const uint32_t* ket_s;
  const uint32_t* exc_s;
  const int64_t* off_d;
  const uint32_t* ket_d;
};

// Same-spin part (and, for the alpha rows, the diagonal): y[row, col] = sum over the row's
// singles and doubles of value * x[ket, col], for K interleaved vectors.  One workgroup = one
// wavefront = one row x 64 C columns, C consecutive columns per lane: the row's list entries
// are the same for every lane (scalar loads), and the gathers of one entry are one contiguous
// segment across the lanes.  A single's value is its stored same-spin part plus the spectator string's Coulomb
// term, read from the Coulomb table row of its orbital pair (coalesced, in flight with the
// gathers).
template <typename T, bool kDiag, int K, int kAcc, int C>
__global__ __launch_bounds__(kRowTile) void k_sigma_rows(
    int64_t nrow, int64_t ncol, int L, RowTabs rt, int nblk, int blk, const T* __restrict__ val_s,
    const T* __restrict__ val_d, const T* __restrict__ coul, const T* __restrict__ x, T* __restrict__ y,
    const T* __restrict__ hii) {
  constexpr int KC = K * C;  // values per lane per ket row
  // gathers issued back to back per wavefront before their FMAs
  constexpr int U = KC * sizeof(T) <= 8 ? 16 : (KC * sizeof(T) == 16 ? (C > 1 ? 8 : 4) : 2);
  const int64_t g = blockIdx.x;
  const int64_t tile = g / nrow;
  const int64_t row = g - tile * nrow;
  const int64_t col = (tile * kRowTile + threadIdx.x) * C;  // ncol is a multiple of C
  if (col >= ncol) return;
  typename AccOf<T, kAcc>::type acc[C][K];
  const int64_t me = row * ncol + col;
  if (blk > 0) {  // a later ket block adds to the sums of the earlier ones
    const VecK<T, KC> ym = ldk<T, KC>(y + me * K);
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int v = 0; v < K; ++v) acc[c][v].add(ym.v[c * K + v]);
  } else if (kDiag) {
    const VecK<T, KC> xm = ldk<T, KC>(x + me * K);
    const VecK<T, C> h = ldk<T, C>(hii + me);
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int v = 0; v < K; ++v) acc[c][v].fma(h.v[c], xm.v[c * K + v]);
  }
  auto fma_all = [&](T w, const VecK<T, KC>& xv) {
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int v = 0; v < K; ++v) acc[c][v].fma(w, xv.v[c * K + v]);
  };
  auto flush_all = [&] {  // the compensated FP32 accumulation closes a run of terms
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int v = 0; v < K; ++v) acc[c][v].flush();
  };
  flush_all();

  // same-spin doubles
  const int64_t rb = row * nblk + blk;
  const int64_t d0 = rt.off_d[rb], d1 = rt.off_d[rb + 1];
  const T* __restrict__ xc = x + col * K;
  const int64_t stride = ncol * K;
  int64_t k = d0;
  for (; k + U <= d1; k += U) {
    VecK<T, KC> xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = ldk<T, KC>(xc + static_cast<int64_t>(rt.ket_d[k + u]) * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) fma_all(val_d[k + u], xv[u]);
    flush_all();
  }
  for (; k < d1; ++k) fma_all(val_d[k], ldk<T, KC>(xc + static_cast<int64_t>(rt.ket_d[k]) * stride));
  flush_all();

  // same-spin singles: value = stored part + sign * Coulomb table [pair, col]
  const int64_t s0 = rt.off_s[rb], s1 = rt.off_s[rb + 1];
  const T* __restrict__ cc = coul + col;
  auto pair_of = [&](int64_t q) {  // a L + i, in 64 bits (the form that keeps the loads interleaved)
    const uint32_t e = rt.exc_s[q];
    return static_cast<int64_t>(exc_a(e)) * L + exc_i(e);
  };
  // the C columns' values of single q are distinct (their Coulomb terms differ)
  auto fma_single = [&](int64_t q, const VecK<T, C>& d, const VecK<T, KC>& xv) {
    const bool neg = exc_neg(rt.exc_s[q]);
    const T base = val_s[q];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const T w = base + (neg ? -d.v[c] : d.v[c]);
#pragma unroll
      for (int v = 0; v < K; ++v) acc[c][v].fma(w, xv.v[c * K + v]);
    }
  };
  k = s0;
  for (; k + U <= s1; k += U) {
    VecK<T, KC> xv[U];
    VecK<T, C> dv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      xv[u] = ldk<T, KC>(xc + static_cast<int64_t>(rt.ket_s[k + u]) * stride);
      dv[u] = ldk<T, C>(cc + pair_of(k + u) * ncol);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) fma_single(k + u, dv[u], xv[u]);
    flush_all();
  }
  for (; k < s1; ++k)
    fma_single(k, ldk<T, C>(cc + pair_of(k) * ncol), ldk<T, KC>(xc + static_cast<int64_t>(rt.ket_s[k]) * stride));
  flush_all();
  VecK<T, KC> out;
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int v = 0; v < K; ++v) out.v[c * K + v] = acc[c][v].get();
  stk<T, KC>(y + me * K, out);
}

// Alpha single x beta single (task type 0, mult.h:569-596):
//   y[ia, ib] += sum_{k in S_a(ia)} sum_{m in S_b(ib)} s_a s_b (a_k i_k | b_m j_m) x[ja_k, jb_m].
// One workgroup = (alpha row ia, beta column tile).  The tile's beta-single entries (sorted by
// jb) are held in registers for the whole loop over the row's alpha singles, so each entry is
// loaded once per workgroup; for every alpha single the integral row V[(a_k i_k), :] (sign
// folded in) is staged in double-buffered LDS, and the ket row x[ja_k, :] is gathered along
// sorted jb.  Workgroups are numbered so that all tiles of one row run on the same XCD and
// share that row's ket rows in L2.  At the end every entry writes its sum to LDS at its slot in
// column order (the beta list's own order: column by column, kets ascending), and every column
// sums its contiguous run of slots in that order (a fixed order: sigma is bitwise
// reproducible); sigma is updated by its owner, with no atomics.
constexpr int kMixMaxCols = 1024;
constexpr size_t kMixMaxLds = 160 * 1024;  // LDS of one CU
// LDS bytes of the row's alpha-single list (ket row and packed word per single), 16-aligned
__host__ __device__ constexpr size_t mix_list_bytes(int max_s) {
  return (static_cast<size_t>(max_s) * 8 + 15) / 16 * 16;
}
// LDS bytes of a variant: the row's alpha-single list and the double-buffered integral rows of
// its stages, reused at the end for one vector's entry sums (tile entries x accumulator width).
__host__ __device__ constexpr size_t mix_lds_bytes(int L, int max_s, int stages, int elem, int tile, int wide) {
  const size_t main = mix_list_bytes(max_s) + 2 * static_cast<size_t>(stages) * L * L * elem;
  const size_t fold = static_cast<size_t>(tile) * wide;
  return main > fold ? main : fold;
}
// Variants of the kernel: entries per thread (kEPT), threads per workgroup (kMixThreads) and
// alpha singles staged per barrier (kStage), coded as 1000000 (kStage - 1) + 100 * threads +
// entries per thread.  Every workgroup of a row gathers from all of
// the row's ket rows, touching most of their cache lines whatever its tile size, so larger
// tiles (threads x entries) mean fewer line fetches per row; more entries per thread mean
// more registers and fewer waves.  The autotuner picks among them per (precision, K).
// V row elements fetched per thread: kVPerThread * kMixThreads covers L^2 (2048: L <= 45,
// 4096: L <= 64).  kStage alpha singles are staged per barrier (their V rows in LDS, their
// gathers in flight together); the sums still take the alpha singles in ascending order, so
// the result does not depend on kStage.  More than 64 orbitals: 16,384 / kMixThreads covers
// L^2 <= 16,384 (the 1024-thread one-stage variants, where two such rows fit the LDS: FP64 up
// to ~99 orbitals, FP32 up to 128), or kVPerThread = 0: every entry reads its V value from the
// row in global memory (L2-resident while the workgroup runs) instead of LDS.

template <typename T, int kEPT, int kVPerThread, int K, int kMixThreads, int kStage, int kAcc, int kVW>
__global__ __launch_bounds__(kMixThreads) void k_sigma_mixed(
    int64_t na, int64_t nb, int L, int max_s, const int64_t* __restrict__ a_off,
    const uint32_t* __restrict__ a_ket, const uint32_t* __restrict__ a_exc, int ntile,
    const int64_t* __restrict__ t_off, const int32_t* __restrict__ t_col0, const int32_t* __restrict__ t_ncol,
    const uint2* __restrict__ e_jm, const int64_t* __restrict__ b_off, const T* __restrict__ V,
    const T* __restrict__ x, T* __restrict__ y) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L2 = L * L;
  // the row's alpha singles (ket row, packed excitation word), then the double-buffered V rows
  // [2][kStage][L2]; the singles are read from LDS, so their loads never wait on the gathers
  uint32_t* s_ket = reinterpret_cast<uint32_t*>(smem_raw);
  uint32_t* s_exc = s_ket + max_s;
  T* vs = reinterpret_cast<T*>(smem_raw + mix_list_bytes(max_s));

  // XCD-aware numbering: workgroup g runs on XCD g % 8; XCD x takes rows ia = x (mod 8), all
  // tiles of a row back to back
  const int64_t g = blockIdx.x;
  const int64_t xcd = g & 7, slot = g >> 3;
  const int64_t tile = slot % ntile;
  const int64_t ia = xcd + 8 * (slot / ntile);
  if (ia >= na) return;
  const int tid = threadIdx.x;
  const int64_t k0 = a_off[ia];
  const int nk = static_cast<int>(a_off[ia + 1] - k0);
  if (nk == 0) return;
  for (int t = tid; t < nk; t += kMixThreads) {
    s_ket[t] = a_ket[k0 + t];
    s_exc[t] = a_exc[k0 + t];
  }

  const int64_t e0 = t_off[tile], e1 = t_off[tile + 1];
  const int col0 = t_col0[tile], ncol = t_ncol[tile];
  uint32_t joff[kEPT], meta[kEPT];
  typename AccOf<T, kAcc>::type acc[kEPT][K];
  // padding entries (past the tile's end) read column 0 and pair 0, which are always valid, so
  // the entry loop below has no branches and all its gathers can be in flight at once; their
  // sums are dropped at the end (slot 0xffff).  joff: byte offset of the entry's ket column in
  // a ket row
#pragma unroll
  for (int j = 0; j < kEPT; ++j) {
    const int64_t e = e0 + tid + static_cast<int64_t>(j) * kMixThreads;
    const uint2 jm = e < e1 ? e_jm[e] : make_uint2(0u, 0x0000ffffu);
    joff[j] = jm.x * static_cast<uint32_t>(K * sizeof(T));
    meta[j] = jm.y;
  }
  __syncthreads();  // the row's singles are staged

  // V rows of the next stage's alpha singles: fetched into registers before the current
  // stage's gathers (their latency hides under them), written to LDS after them; kVW values
  // per load (16 bytes where L^2 values of a row are a whole number of 16-byte words)
  constexpr int kVV = kVPerThread / kVW > 0 ? kVPerThread / kVW : 1;  // loads per thread
  VecK<T, kVW> pre[kStage][kVV];
  bool pre_neg[kStage];
  auto fetch = [&](int kk) {
    if constexpr (kVPerThread == 0) return;
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      if (kk + q >= nk) break;
      const uint32_t e = __builtin_amdgcn_readfirstlane(s_exc[kk + q]);
      const T* src = V + static_cast<int64_t>(exc_pair(e, L)) * L2;
      pre_neg[q] = exc_neg(e);
#pragma unroll
      for (int u = 0; u < kVV; ++u) {
        const int p = (tid + u * kMixThreads) * kVW;
        if (p < L2) pre[q][u] = ldk<T, kVW>(src + p);  // (commit skips the rest)
      }
    }
  };
  auto commit = [&](T* dst, int kk) {
    if constexpr (kVPerThread == 0) return;
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
      if (kk + q >= nk) break;
#pragma unroll
      for (int u = 0; u < kVV; ++u) {
        const int p = (tid + u * kMixThreads) * kVW;
        if (p < L2) {
          VecK<T, kVW> w = pre[q][u];
#pragma unroll
          for (int i = 0; i < kVW; ++i) w.v[i] = pre_neg[q] ? -w.v[i] : w.v[i];
          stk<T, kVW>(dst + q * L2 + p, w);
        }
      }
    }
  };
  fetch(0);
  commit(vs, 0);
  const int64_t row_bytes = nb * K * static_cast<int64_t>(sizeof(T));
  int buf = 0;
  // runs of 16 alpha singles (16 terms per entry and vector); the compensated FP32 accumulation
  // closes each run (a no-op for the other accumulators), outside the inner loop so that it is
  // not if-converted into every stage
  static_assert(16 % kStage == 0, "a run holds whole stages");
  for (int kk = 0; kk < nk;) {
    const int krun = kk + 16 < nk ? kk + 16 : nk;
    for (; kk < krun; kk += kStage, buf ^= 1) {
      T* cur = vs + buf * kStage * L2;
      T* nxt = vs + (buf ^ 1) * kStage * L2;
      __syncthreads();  // cur is staged; nobody reads nxt any more
      const bool more = kk + kStage < nk;
      if (more) fetch(kk + kStage);
      VecK<T, K> xv[kStage][kEPT];
#pragma unroll
      for (int q = 0; q < kStage; ++q) {
        if (kk + q >= nk) break;
        const uint32_t ja = __builtin_amdgcn_readfirstlane(s_ket[kk + q]);
        const __amdgpu_buffer_rsrc_t xr =
            buf_rsrc(reinterpret_cast<const unsigned char*>(x) + ja * row_bytes, static_cast<uint32_t>(row_bytes));
#pragma unroll
        for (int j = 0; j < kEPT; ++j) xv[q][j] = ldk_buf<T, K>(xr, joff[j]);
      }
#pragma unroll
      for (int q = 0; q < kStage; ++q) {
        if (kk + q >= nk) break;
        const T* vrow;
        bool vneg = false;
        if constexpr (kVPerThread == 0) {
          const uint32_t e = __builtin_amdgcn_readfirstlane(s_exc[kk + q]);
          vrow = V + static_cast<int64_t>(exc_pair(e, L)) * L2;
          vneg = exc_neg(e);
        } else {
          vrow = cur + q * L2;
        }
#pragma unroll
        for (int j = 0; j < kEPT; ++j) {
          T w = vrow[(meta[j] >> 16) & 0x3fff];
          if constexpr (kVPerThread == 0) w = vneg ? -w : w;
#pragma unroll
          for (int v = 0; v < K; ++v) acc[j][v].fma(w, xv[q][j].v[v]);
        }
      }
      if (more) commit(nxt, kk + kStage);
    }
#pragma unroll
    for (int j = 0; j < kEPT; ++j)
#pragma unroll
      for (int v = 0; v < K; ++v) acc[j][v].flush();
  }

  // Fold, one vector at a time: every entry writes its sum (sign of the beta single applied)
  // to LDS at its slot = its index in the beta list minus the tile's first index, so column c
  // of the tile owns the contiguous slots b_off[col0 + c] - b_off[col0] ..
  // b_off[col0 + c + 1] - b_off[col0] - 1 and sums them in ascending order.
  using Acc = typename AccOf<T, kAcc>::type;
  using Wide = typename Acc::Wide;
  constexpr int kCols = kMixMaxCols / kMixThreads;
  Wide* stage = reinterpret_cast<Wide*>(smem_raw);
  const int64_t base = b_off[col0];
  int cs[kCols], ce[kCols];
#pragma unroll
  for (int u = 0; u < kCols; ++u) {
    const int c = tid + u * kMixThreads;
    cs[u] = c < ncol ? static_cast<int>(b_off[col0 + c] - base) : 0;
    ce[u] = c < ncol ? static_cast<int>(b_off[col0 + c + 1] - base) : 0;
  }
  T* yr = y + (ia * nb + col0) * K;
#pragma unroll
  for (int v = 0; v < K; ++v) {
    __syncthreads();  // the LDS is free: integral rows (first vector) or the previous vector
#pragma unroll
    for (int j = 0; j < kEPT; ++j) {
      const uint32_t slot = meta[j] & 0xffffu;
      const Wide w = acc[j][v].wide();
      if (slot != 0xffffu) stage[slot] = meta[j] >> 31 ? -w : w;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kCols; ++u) {
      const int c = tid + u * kMixThreads;
      if (c < ncol) {
        Acc col;
        for (int q = cs[u]; q < ce[u]; ++q) col.add(stage[q]);
        yr[c * K + v] = col.apply(yr[c * K + v]);
      }
    }
  }
}

// out[c * R + r] = in[r * C + c] on K-value units; S x S tiles staged in LDS
template <typename T, int K>
__global__ __launch_bounds__(256) void k_transpose(const T* __restrict__ in, T* __restrict__ out,
                                                   int64_t R, int64_t C) {
  constexpr int S = K * sizeof(T) <= 8 ? 64 : (K * sizeof(T) <= 32 ? 32 : 16);
  __shared__ VecK<T, K> tile[S][S + 1];
  const int64_t tc = cdiv(C, S);
  const int64_t br = blockIdx.x / tc, bc = blockIdx.x % tc;
  const int tx = threadIdx.x % S, ty = threadIdx.x / S;
  constexpr int STEP = 256 / S;
  for (int k = ty; k < S; k += STEP) {
    const int64_t r = br * S + k, c = bc * S + tx;
    if (r < R && c < C) tile[k][tx] = ldk<T, K>(in + (r * C + c) * K);
  }
  __syncthreads();
  for (int k = ty; k < S; k += STEP) {
    const int64_t c = bc * S + k, r = br * S + tx;
    if (r < R && c < C) stk<T, K>(out + (c * R + r) * K, tile[tx][k]);
  }
}

// y[r * C + c] += inT[c * R + r] on K-value units
template <typename T, int K>
__global__ __launch_bounds__(256) void k_add_transposed(const T* __restrict__ inT, T* __restrict__ y,
                                                        int64_t R, int64_t C) {
  constexpr int S = K * sizeof(T) <= 8 ? 64 : (K * sizeof(T) <= 32 ? 32 : 16);
  __shared__ VecK<T, K> tile[S][S + 1];
  const int64_t tr = cdiv(R, S);
  const int64_t bc = blockIdx.x / tr, br = blockIdx.x % tr;
  const int tx = threadIdx.x % S, ty = threadIdx.x / S;
  constexpr int STEP = 256 / S;
  for (int k = ty; k < S; k += STEP) {
    const int64_t c = bc * S + k, r = br * S + tx;
    if (r < R && c < C) tile[k][tx] = ldk<T, K>(inT + (c * R + r) * K);
  }
  __syncthreads();
  for (int k = ty; k < S; k += STEP) {
    const int64_t r = br * S + k, c = bc * S + tx;
    if (r < R && c < C) {
      VecK<T, K> a = ldk<T, K>(y + (r * C + c) * K);
#pragma unroll
      for (int v = 0; v < K; ++v) a.v[v] += tile[tx][k].v[v];
      stk<T, K>(y + (r * C + c) * K, a);
    }
  }
}

template <typename T, int K, int W>
__global__ __launch_bounds__(256) void k_apply_sign(const uint64_t* __restrict__ sa,
                                                    const uint64_t* __restrict__ sb, int64_t na,
                                                    int64_t nb, const T* __restrict__ x,
                                                    T* __restrict__ y) {
  const int64_t nt = cdiv(nb, 256);
  const int64_t ia = blockIdx.x / nt;
  const int64_t ib = (blockIdx.x % nt) * 256 + threadIdx.x;
  if (ib >= nb) return;
  const bool neg = det_sign(ld_str<W>(sa, ia), ld_str<W>(sb, ib)) < 0;
  const int64_t k = (ia * nb + ib) * K;
  VecK<T, K> a = ldk<T, K>(x + k);
#pragma unroll
  for (int v = 0; v < K; ++v) a.v[v] = neg ? -a.v[v] : a.v[v];
  stk<T, K>(y + k, a);
}

// A block of K = 4 or 8 vectors can also be built as sub-blocks of S = 2 (or 4) consecutive
// vectors: the variant code S (kSplit2, kSplit4) of width K means "gather S of the K
// interleaved values of every determinant into an S-wide block, build its sigma, scatter it
// back", K / S times.  The fused K-wide kernels keep one K-value unit per lane and entry, so
// their registers and L2 segments grow with K; at 4e8 determinants K = 2 sub-blocks cost less
// per vector than the fused K = 4 and K = 8 builds.
constexpr int kSplit2 = 2, kSplit4 = 4;

// The mixed-kernel variant a fused K-wide build uses: the context's autotuned choice when there
// is one (sbdx_autotune, sbdx_set_mix_variant), else the autotuner's choice on the
// 4e8-determinant 36-orbital workload, or for more than 45 orbitals (integral rows of more than
// 2048 values: the large variants drop to one wave per SIMD) its choice on the 1e8-determinant
// 48-orbital workload.
int builtin_variant(const sbdx_ctx& c, int prec, int K) {
  // compensated FP32 (three words of state per entry and vector): the autotuner's choices at
  // 4e8 determinants, 512 x 16 for one vector and 512 x 8 x 2 stages for two
  if (prec != 0 && c.fp32_acc == SBDX_ACC_TWOSUM && c.L * c.L <= 2048) {
    if (K == 1) return 51216;
    if (K == 2) return 1051208;
  }
  if (c.L > 64) {
    // V rows staged in LDS by the 1024-thread variants where two rows fit, else read from
    // global memory by the 256-thread ones
    const int es = prec == 0 ? 8 : 4;
    if (K == 1) return mix_lds_bytes(c.L, c.A.max_s, 1, es, 1024 * 8, 8) <= kMixMaxLds ? 102408 : 25616;
    if (K == 2) return mix_lds_bytes(c.L, c.A.max_s, 1, es, 1024 * 4, 8) <= kMixMaxLds ? 102404 : 25608;
    return 25604;
  }
  if (c.L * c.L > 2048) {
    if (K == 1) return 25616;
    if (K == 2) return 25608;
  }
  if (K == 1) return prec == 0 ? 1102408 : 102408;
  if (K == 2) return prec == 0 ? 102408 : 25616;
  if (K == 4) return prec == 0 ? 1102404 : 25608;
  return 25604;
}
int fused_variant(const sbdx_ctx& c, int prec, int K) {
  const int tuned = tuned_variant(c, prec, K);
  if (tuned >= 100) return tuned;  // a kernel variant, not a sub-block width
  return builtin_variant(c, prec, K);
}

// What a K-wide build runs: the tuned variant or sub-block width, else K = 2 sub-blocks for
// K = 4 and 8 up to 45 orbitals (measured faster per vector at 4e8 determinants), else the
// fused default.
int mix_variant(const sbdx_ctx& c, int prec, int K) {
  const int tuned = tuned_variant(c, prec, K);
  if (tuned) return tuned;
  if (K >= 4 && c.L * c.L <= 2048) return kSplit2;
  if (K >= 4 && c.L > 64 && builtin_variant(c, prec, 2) == 102404) return kSplit2;  // LDS-staged pairs
  return fused_variant(c, prec, K);
}

// Regroup the beta singles into mixed-pass tiles of at most cap entries (host side, once per
// integral set and tile size: the packed excitation words carry the orbital pair and sign).
const MixTiles& mix_tiles(sbdx_ctx& c, int64_t cap) {
  MixTiles* slot = nullptr;
  for (auto& m : c.mix) {
    if (m.cap == cap && m.ent_off.get()) return m;
    if (!slot && m.cap == 0) slot = &m;
  }
  if (c.no_alloc)
    throw Error(SBDX_ERR_STATE, "the mixed-pass tiles of this precision's kernel variant are not built: run "
                                "sbdx_sigma_device once with this precision first");
  // tile sets are never evicted: captured graphs (batches) hold their addresses.  The
  // variants use four tile sizes (1024 to 8192 entries), fewer than the slots.
  if (!slot) throw Error(SBDX_ERR_STATE, "no free slot for another mixed-pass tile size");
  MixTiles& mt = *slot;
  mt = MixTiles{};
  mt.cap = cap;
  const SpinTab& B = c.B;
  const int L = c.L;
  const int64_t nb = B.n;
  if (B.nnz_s == 0 || c.A.nnz_s == 0) {
    mt.ent_off.alloc(1);
    return mt;
  }
  std::vector<int64_t> off(nb + 1);
  std::vector<uint32_t> ket(B.nnz_s), exc(B.nnz_s);
  SBDX_HIP(hipMemcpy(off.data(), B.off_s.get(), (nb + 1) * 8, hipMemcpyDeviceToHost));
  SBDX_HIP(hipMemcpy(ket.data(), B.ket_s.get(), B.nnz_s * 4, hipMemcpyDeviceToHost));
  SBDX_HIP(hipMemcpy(exc.data(), B.exc_s.get(), B.nnz_s * 4, hipMemcpyDeviceToHost));
  std::vector<int64_t> t_off{0};
  std::vector<int32_t> t_col0, t_ncol;
  std::vector<uint2> e_jm;
  e_jm.reserve(B.nnz_s);
  struct Ent {
    uint32_t jb, meta;
  };
  std::vector<Ent> buf;
  int64_t ib = 0;
  while (ib < nb) {
    const int64_t start = ib;
    int64_t cnt = 0;
    while (ib < nb && ib - start < kMixMaxCols && cnt + (off[ib + 1] - off[ib]) <= cap) {
      cnt += off[ib + 1] - off[ib];
      ++ib;
    }
    if (ib == start) throw Error(-1, "beta string with more singles than a mixed tile holds");
    if (cap > 65535) throw Error(-1, "mixed tile larger than its 16-bit entry slots");
    buf.clear();
    for (int64_t b = start; b < ib; ++b)
      for (int64_t k = off[b]; k < off[b + 1]; ++k) {
        const uint32_t e = exc[k];
        const uint32_t pb = exc_pair(e, L);
        const uint32_t slot = static_cast<uint32_t>(k - off[start]);
        const uint32_t meta = slot | (pb << 16) | (exc_neg(e) ? 1u << 31 : 0u);
        buf.push_back(Ent{ket[k], meta});
      }
    std::stable_sort(buf.begin(), buf.end(), [](const Ent& u, const Ent& v) { return u.jb < v.jb; });
    for (size_t r = 0; r < buf.size(); ++r) e_jm.push_back(make_uint2(buf[r].jb, buf[r].meta));
    t_col0.push_back(static_cast<int32_t>(start));
    t_ncol.push_back(static_cast<int32_t>(ib - start));
    t_off.push_back(static_cast<int64_t>(e_jm.size()));
    mt.max_cols = std::max<int>(mt.max_cols, static_cast<int>(ib - start));
  }
  mt.ntile = static_cast<int>(t_col0.size());
  mt.ent_off.alloc(t_off.size());
  mt.col0.alloc(t_col0.size());
  mt.ncol.alloc(t_ncol.size());
  mt.jm.alloc(e_jm.size());
  SBDX_HIP(hipMemcpy(mt.ent_off.get(), t_off.data(), t_off.size() * 8, hipMemcpyHostToDevice));
  SBDX_HIP(hipMemcpy(mt.col0.get(), t_col0.data(), t_col0.size() * 4, hipMemcpyHostToDevice));
  SBDX_HIP(hipMemcpy(mt.ncol.get(), t_ncol.data(), t_ncol.size() * 4, hipMemcpyHostToDevice));
  SBDX_HIP(hipMemcpy(mt.jm.get(), e_jm.data(), e_jm.size() * sizeof(uint2), hipMemcpyHostToDevice));
  return mt;
}

template <typename T, int K, int kAcc>
void launch_mixed(sbdx_ctx& c, const T* x, T* y, const T* V, hipStream_t st) {
  const int var = fused_variant(c, std::is_same<T, double>::value ? 0 : 1, K);
  const int stages = var / 1000000 + 1, threads = (var % 1000000) / 100;
  const MixTiles& mt = mix_tiles(c, static_cast<int64_t>(var % 100) * threads);
  if (mt.ntile == 0) return;
  const int64_t blocks = 8 * cdiv(c.na, 8) * mt.ntile;
  using Wide = typename AccOf<T, kAcc>::type::Wide;
  // more than 64 orbitals: the 1024-thread one-stage variants stage V rows of up to 16,384
  // values in LDS (where two fit), the 256-thread ones read them from global memory
  const bool global_v = c.L > 64 && threads == 256;
  auto launch = [&](auto kernel, int threads, int ept) {
    const size_t lds = mix_lds_bytes(c.L, c.A.max_s, global_v ? 0 : stages, static_cast<int>(sizeof(T)),
                                     threads * ept, static_cast<int>(sizeof(Wide)));
    if (c.nb * K * static_cast<int64_t>(sizeof(T)) >= (int64_t(1) << 32))
      throw Error(SBDX_ERR_ARG, "beta string count too large for 32-bit ket-row offsets of the mixed kernel");
    if (lds > kMixMaxLds) throw Error(SBDX_ERR_ARG, "mixed-kernel variant needs more LDS than a CU has for this problem");
    if (lds > 65536)
      SBDX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    kernel<<<static_cast<unsigned>(blocks), threads, lds, st>>>(
        c.na, c.nb, c.L, c.A.max_s, c.A.off_s.get(), c.A.ket_s.get(), c.A.exc_s.get(), mt.ntile, mt.ent_off.get(),
        mt.col0.get(), mt.ncol.get(), mt.jm.get(), c.B.off_s.get(), V, x, y);
    SBDX_CHECK_LAUNCH();
  };
  const bool small_v = static_cast<int>(c.L) * c.L <= 2048;
  // integral rows loaded 16 bytes at a time when a row is a whole number of 16-byte words
  constexpr int VW = 16 / static_cast<int>(sizeof(T));
  const bool wide_v = (static_cast<int64_t>(c.L) * c.L * sizeof(T)) % 16 == 0;
  auto pick = [&](auto e, auto t, auto sg) {
    constexpr int E = decltype(e)::value, NT = decltype(t)::value, SG = decltype(sg)::value;
    if (c.L > 64) {
      if constexpr (SG == 1 && NT == 256) {
        launch(&k_sigma_mixed<T, E, 0, K, NT, 1, kAcc, 1>, NT, E);
      } else if constexpr (SG == 1 && NT == 1024) {
        if (wide_v) launch(&k_sigma_mixed<T, E, 16384 / NT, K, NT, 1, kAcc, VW>, NT, E);
        else launch(&k_sigma_mixed<T, E, 16384 / NT, K, NT, 1, kAcc, 1>, NT, E);
      } else {
        throw Error(SBDX_ERR_ARG, "above 64 orbitals the mixed kernel runs its one-stage 256- and 1024-thread variants only");
      }
    } else if (wide_v)
      small_v ? launch(&k_sigma_mixed<T, E, 2048 / NT, K, NT, SG, kAcc, VW>, NT, E)
              : launch(&k_sigma_mixed<T, E, 4096 / NT, K, NT, SG, kAcc, VW>, NT, E);
    else
      small_v ? launch(&k_sigma_mixed<T, E, 2048 / NT, K, NT, SG, kAcc, 1>, NT, E)
              : launch(&k_sigma_mixed<T, E, 4096 / NT, K, NT, SG, kAcc, 1>, NT, E);
  };
  using std::integral_constant;
  using I4 = integral_constant<int, 4>;
  using I8 = integral_constant<int, 8>;
  using I16 = integral_constant<int, 16>;
  using I32 = integral_constant<int, 32>;
  using T256 = integral_constant<int, 256>;
  using T512 = integral_constant<int, 512>;
  using T1024 = integral_constant<int, 1024>;
  using S1 = integral_constant<int, 1>;
  using S2 = integral_constant<int, 2>;
  using S4 = integral_constant<int, 4>;
  // the candidates of mix_variant_candidates(K); anything else is refused
  if constexpr (K == 1) {
    if (var == 25632) pick(I32{}, T256{}, S1{});
    else if (var == 25616) pick(I16{}, T256{}, S1{});
    else if (var == 51216) pick(I16{}, T512{}, S1{});
    else if (var == 102408) pick(I8{}, T1024{}, S1{});
    else if (var == 1102408) pick(I8{}, T1024{}, S2{});
    else if (var == 3102408) pick(I8{}, T1024{}, S4{});
    else if (var == 1051216) pick(I16{}, T512{}, S2{});
    else if (var == 1025616) pick(I16{}, T256{}, S2{});
    else throw Error(SBDX_ERR_ARG, "unknown mixed-kernel variant");
  } else if constexpr (K == 2) {
    if (var == 102404) pick(I4{}, T1024{}, S1{});
    else if (var == 25608) pick(I8{}, T256{}, S1{});
    else if (var == 102408) pick(I8{}, T1024{}, S1{});
    else if (var == 25616) pick(I16{}, T256{}, S1{});
    else if (var == 51216) pick(I16{}, T512{}, S1{});
    else if (var == 1102404) pick(I4{}, T1024{}, S2{});
    else if (var == 3102404) pick(I4{}, T1024{}, S4{});
    else if (var == 1051208) pick(I8{}, T512{}, S2{});
    else throw Error(SBDX_ERR_ARG, "unknown mixed-kernel variant");
  } else if constexpr (K == 4) {
    if (var == 25608) pick(I8{}, T256{}, S1{});
    else if (var == 25604) pick(I4{}, T256{}, S1{});
    else if (var == 102404) pick(I4{}, T1024{}, S1{});
    else if (var == 1102404) pick(I4{}, T1024{}, S2{});
    else throw Error(SBDX_ERR_ARG, "unknown mixed-kernel variant");
  } else {
    if (var == 25604) pick(I4{}, T256{}, S1{});
    else if (var == 51204) pick(I4{}, T512{}, S1{});
    else throw Error(SBDX_ERR_ARG, "unknown mixed-kernel variant");
  }
}

// The rows in flight on an XCD reach nearly every ket row, so a row pass's L2 working set is
// (ket rows) x (its 64-column segment of K words): 10 MB for 20,000 FP64 ket rows against an
// XCD's 4 MB L2.  The pass is therefore cut into ket blocks whose segments fit in kRowL2Budget:
// launch b sums the entries whose ket lies in block b and adds them to the earlier blocks'
// sums (the diagonal and the first block write the output).
constexpr size_t kRowL2Budget = size_t(3) << 20;
constexpr int kMaxKetBlocks = 16;

int ket_blocks_for(int64_t nket, int K, size_t word) {
  const size_t ws = static_cast<size_t>(nket) * kRowTile * K * word;
  return static_cast<int>(std::min<int64_t>(kMaxKetBlocks, std::max<int64_t>(1, cdiv(ws, kRowL2Budget))));
}

// first index of each (row, block) in a ket-sorted CSR list, blocks of bw kets
__global__ __launch_bounds__(256) void k_ket_blocks(const int64_t* __restrict__ off, const uint32_t* __restrict__ ket,
                                                    int64_t nrow, int64_t bw, int nblk, int64_t* __restrict__ out) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (q > nrow * nblk) return;
  if (q == nrow * nblk) {
    out[q] = off[nrow];
    return;
  }
  const int64_t row = q / nblk, b = q - row * nblk;
  int64_t lo = off[row], hi = off[row + 1];
  const int64_t bound = b * bw;
  while (lo < hi) {  // first entry with ket >= bound
    const int64_t mid = (lo + hi) / 2;
    if (static_cast<int64_t>(ket[mid]) < bound) lo = mid + 1;
    else hi = mid;
  }
  out[q] = lo;
}

// the lists of one spin cut into nblk ket blocks (built once per block count)
RowTabs ket_block_tabs(sbdx_ctx& c, SpinTab& t, int64_t nket, int nblk, hipStream_t st) {
  RowTabs rt{t.off_s.get(), t.ket_s.get(), t.exc_s.get(), t.off_d.get(), t.ket_d.get()};
  if (nblk == 1) return rt;
  SpinTab::KetBlocks* kb = nullptr;
  for (auto& k : t.kblk)
    if (k.nblk == nblk) kb = &k;
  if (!kb) {
    if (c.no_alloc)
      throw Error(SBDX_ERR_STATE, "the ket blocks of this precision's row passes are not built: run "
                                  "sbdx_sigma_device once with this precision first");
    for (auto& k : t.kblk)
      if (k.nblk == 0) {
        kb = &k;
        break;
      }
    if (!kb) throw Error(SBDX_ERR_STATE, "no free slot for another ket-block count");
    const int64_t n = t.n, bw = cdiv(nket, nblk);
    kb->s.alloc(n * nblk + 1);
    kb->d.alloc(n * nblk + 1);
    const unsigned grid = static_cast<unsigned>(cdiv(n * nblk + 1, 256));
    k_ket_blocks<<<grid, 256, 0, st>>>(t.off_s.get(), t.ket_s.get(), n, bw, nblk, kb->s.get());
    k_ket_blocks<<<grid, 256, 0, st>>>(t.off_d.get(), t.ket_d.get(), n, bw, nblk, kb->d.get());
    SBDX_CHECK_LAUNCH();
    kb->nblk = nblk;
  }
  rt.off_s = kb->s.get();
  rt.off_d = kb->d.get();
  return rt;
}

// The row kernel, one launch per ket block.  Lane width: the vector memory path takes a lane
// quad per clock for 8- and 16-byte loads and 16 lanes for 4-byte ones (a 64-lane load of
// 8-byte values costs as many L1 accesses as one of 16-byte values), so K words of 8 bytes are
// loaded two columns per lane (C = 2, 16-byte lanes, 128 columns per wavefront); 4- and 16-byte
// K words keep one column per lane.  The ket blocks are sized for the C-column segments.
template <typename T, bool kDiag, int K, int kAcc>
void launch_rows(sbdx_ctx& c, SpinTab& t, int64_t nrow, int64_t ncol, const T* val_s, const T* val_d,
                 const T* coul, const T* x, T* y, const T* hii, hipStream_t st) {
  constexpr int CW = K * sizeof(T) == 8 ? 2 : 1;
  auto run = [&](auto cc) {
    constexpr int C = decltype(cc)::value;
    const int nblk = ket_blocks_for(nrow, C * K, sizeof(T));  // kets of a row pass are its rows
    const RowTabs rt = ket_block_tabs(c, t, nrow, nblk, st);
    for (int b = 0; b < nblk; ++b) {
      k_sigma_rows<T, kDiag, K, kAcc, C><<<static_cast<unsigned>(nrow * cdiv(ncol, kRowTile * C)), kRowTile, 0,
                                           st>>>(nrow, ncol, c.L, rt, nblk, b, val_s, val_d, coul, x, y, hii);
      SBDX_CHECK_LAUNCH();
    }
  };
  // C columns per lane need whole 16-byte words per row
  if (CW > 1 && ncol % CW == 0) run(std::integral_constant<int, CW>{});
  else run(std::integral_constant<int, 1>{});
}

template <typename T, int K, int kAcc>
void sigma_k(sbdx_ctx& c, const T* x, T* y, hipStream_t st, const T* vs_a, const T* vd_a,
             const T* vs_b, const T* vd_b, const T* V, const T* hii, const T* coul_a, const T* coul_b) {
  const int64_t na = c.na, nb = c.nb;
  const int64_t n = na * nb;
  Prof& pf = c.prof;
  auto mark = [&](int k) {
    if (pf.on) SBDX_HIP(hipEventRecord(pf.ev[k], st));
  };
  mark(0);
  // alpha rows: diagonal + alpha singles and doubles
  launch_rows<T, true, K, kAcc>(c, c.A, na, nb, vs_a, vd_a, coul_a, x, y, hii, st);
  SBDX_CHECK_LAUNCH();
  mark(1);
  // alpha single x beta single
  launch_mixed<T, K, kAcc>(c, x, y, V, st);
  mark(2);
  // beta rows on the transposed vector
  if (!c.no_alloc) {
    c.ws.xT.ensure(n * K * sizeof(T));
    c.ws.yT.ensure(n * K * sizeof(T));
  }
  T* xT = reinterpret_cast<T*>(c.no_alloc ? c.noalloc_xT : c.ws.xT.get());
  T* yT = reinterpret_cast<T*>(c.no_alloc ? c.noalloc_yT : c.ws.yT.get());
  constexpr int S = K * sizeof(T) <= 8 ? 64 : (K * sizeof(T) <= 32 ? 32 : 16);
  const unsigned tblocks = static_cast<unsigned>(cdiv(na, S) * cdiv(nb, S));
  k_transpose<T, K><<<tblocks, 256, 0, st>>>(x, xT, na, nb);
  SBDX_CHECK_LAUNCH();
  mark(3);
  launch_rows<T, false, K, kAcc>(c, c.B, nb, na, vs_b, vd_b, coul_b, xT, yT, nullptr, st);
  SBDX_CHECK_LAUNCH();
  mark(4);
  k_add_transposed<T, K><<<tblocks, 256, 0, st>>>(yT, y, na, nb);
  SBDX_CHECK_LAUNCH();
  mark(5);
  if (pf.on) {
    SBDX_HIP(hipEventSynchronize(pf.ev[kPhases]));
    for (int k = 0; k < kPhases; ++k) {
      float ms = 0.0f;
      SBDX_HIP(hipEventElapsedTime(&ms, pf.ev[k], pf.ev[k + 1]));
      pf.ms[k] += ms;
    }
    pf.calls++;
  }
}

// sub[d * S + v] = x[d * K + s0 + v] and back, v < S (one determinant per thread)
template <typename T, int K, int S>
__global__ __launch_bounds__(256) void k_sub_gather(const T* __restrict__ x, T* __restrict__ sub, int64_t n, int s0) {
  const int64_t d = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (d < n) stk<T, S>(sub + d * S, ldk<T, S>(x + d * K + s0));
}
template <typename T, int K, int S>
__global__ __launch_bounds__(256) void k_sub_scatter(const T* __restrict__ sub, T* __restrict__ y, int64_t n, int s0) {
  const int64_t d = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (d < n) stk<T, S>(y + d * K + s0, ldk<T, S>(sub + d * S));
}

// a K-wide block built as K / S sub-blocks of width S (the kSplit variants)
template <typename T, int K, int S, int kAcc, typename... Args>
void sigma_split(sbdx_ctx& c, const T* x, T* y, hipStream_t st, Args... args) {
  static_assert(S < K && K % S == 0, "sub-blocks divide the block");
  if (c.no_alloc) throw Error(SBDX_ERR_STATE, "sub-block builds allocate their sub-block buffers");
  const int64_t n = c.ndet();
  c.ws.sub_x.ensure(n * S * sizeof(T));
  c.ws.sub_y.ensure(n * S * sizeof(T));
  T* sx = reinterpret_cast<T*>(c.ws.sub_x.get());
  T* sy = reinterpret_cast<T*>(c.ws.sub_y.get());
  const unsigned g = static_cast<unsigned>(cdiv(n, 256));
  for (int s0 = 0; s0 < K; s0 += S) {
    k_sub_gather<T, K, S><<<g, 256, 0, st>>>(x, sx, n, s0);
    SBDX_CHECK_LAUNCH();
    sigma_k<T, S, kAcc>(c, sx, sy, st, args...);
    k_sub_scatter<T, K, S><<<g, 256, 0, st>>>(sy, y, n, s0);
    SBDX_CHECK_LAUNCH();
  }
  if (c.prof.on) c.prof.calls -= K / S - 1;  // the phases of one block build, over its sub-blocks
}

template <typename T>
void sigma_dispatch(sbdx_ctx& c, int k, const T* x, T* y, hipStream_t st) {
  const bool f = std::is_same<T, float>::value;
  if (f) ensure_fp32(c);
  const T* vs_a = f ? reinterpret_cast<const T*>(c.A.val_s32.get()) : reinterpret_cast<const T*>(c.A.val_s.get());
  const T* vd_a = f ? reinterpret_cast<const T*>(c.A.val_d32.get()) : reinterpret_cast<const T*>(c.A.val_d.get());
  const T* vs_b = f ? reinterpret_cast<const T*>(c.B.val_s32.get()) : reinterpret_cast<const T*>(c.B.val_s.get());
  const T* vd_b = f ? reinterpret_cast<const T*>(c.B.val_d32.get()) : reinterpret_cast<const T*>(c.B.val_d.get());
  const T* V = f ? reinterpret_cast<const T*>(c.eri4_32.get()) : reinterpret_cast<const T*>(c.eri4.get());
  const T* hii = f ? reinterpret_cast<const T*>(c.hii32.get()) : reinterpret_cast<const T*>(c.hii.get());
  const T* ca = f ? reinterpret_cast<const T*>(c.coul_a32.get()) : reinterpret_cast<const T*>(c.coul_a.get());
  const T* cb = f ? reinterpret_cast<const T*>(c.coul_b32.get()) : reinterpret_cast<const T*>(c.coul_b.get());
  const int split = k >= 4 ? mix_variant(c, f ? 1 : 0, k) : 0;
  auto go = [&](auto acc) {
    constexpr int A = decltype(acc)::value;
    switch (k) {
      case 1: sigma_k<T, 1, A>(c, x, y, st, vs_a, vd_a, vs_b, vd_b, V, hii, ca, cb); break;
      case 2: sigma_k<T, 2, A>(c, x, y, st, vs_a, vd_a, vs_b, vd_b, V, hii, ca, cb); break;
      case 4:
        if (split == kSplit2) sigma_split<T, 4, 2, A>(c, x, y, st, vs_a, vd_a, vs_b, vd_b, V, hii, ca, cb);
        else sigma_k<T, 4, A>(c, x, y, st, vs_a, vd_a, vs_b, vd_b, V, hii, ca, cb);
        break;
      case 8:
        if (split == kSplit2) sigma_split<T, 8, 2, A>(c, x, y, st, vs_a, vd_a, vs_b, vd_b, V, hii, ca, cb);
        else if (split == kSplit4) sigma_split<T, 8, 4, A>(c, x, y, st, vs_a, vd_a, vs_b, vd_b, V, hii, ca, cb);
        else sigma_k<T, 8, A>(c, x, y, st, vs_a, vd_a, vs_b, vd_b, V, hii, ca, cb);
        break;
      default: throw Error(-1, "block width must be 1, 2, 4 or 8");
    }
  };
  // FP64 has one accumulation; FP32 follows the context's mode
  if (f && c.fp32_acc == SBDX_ACC_TWOSUM) go(std::integral_constant<int, SBDX_ACC_TWOSUM>{});
  else go(std::integral_constant<int, SBDX_ACC_FP64>{});
}

template <typename T>
void sign_dispatch(sbdx_ctx& c, int k, const T* x, T* y, hipStream_t st) {
  const unsigned blocks = static_cast<unsigned>(c.na * cdiv(c.nb, 256));
  with_words(c.W, [&](auto words) {
    constexpr int W = decltype(words)::value;
    const uint64_t *sa = c.A.str.get(), *sb = c.B.str.get();
    switch (k) {
      case 1: k_apply_sign<T, 1, W><<<blocks, 256, 0, st>>>(sa, sb, c.na, c.nb, x, y); break;
      case 2: k_apply_sign<T, 2, W><<<blocks, 256, 0, st>>>(sa, sb, c.na, c.nb, x, y); break;
      case 4: k_apply_sign<T, 4, W><<<blocks, 256, 0, st>>>(sa, sb, c.na, c.nb, x, y); break;
      case 8: k_apply_sign<T, 8, W><<<blocks, 256, 0, st>>>(sa, sb, c.na, c.nb, x, y); break;
      default: throw Error(-1, "block width must be 1, 2, 4 or 8");
    }
  });
  SBDX_CHECK_LAUNCH();
}

}  // namespace

void reset_mix_tiles(sbdx_ctx& c) {
  for (auto& m : c.mix) m = MixTiles{};
}

// build (once) the mixed-pass tiles of the kernel variant a sigma build of precision prec and
// block width k uses, so a later build does not allocate them
void prepare_mix_tiles(sbdx_ctx& c, int prec, int k) {
  int var = mix_variant(c, prec, k);
  if (var < k) var = fused_variant(c, prec, var);  // sub-blocks of width var, fused
  const int threads = (var % 1000000) / 100;
  (void)mix_tiles(c, static_cast<int64_t>(var % 100) * threads);
}

// variants (1000000 (stages - 1) + 100 threads + entries per thread) of the mixed kernel for
// block width K, and for K = 4 and 8 the sub-block widths (kSplit2, kSplit4)
std::vector<int> mix_variant_candidates(int K) {
  if (K == 1) return {25632, 25616, 51216, 102408, 1102408, 3102408, 1051216, 1025616};
  if (K == 2) return {102404, 25608, 102408, 25616, 51216, 1102404, 3102404, 1051208};
  if (K == 4) return {25608, 25604, 102404, 1102404, kSplit2};
  return {25604, 51204, kSplit2, kSplit4};
}

// whether a variant's LDS (double-buffered V rows of its stages, reused by the fold for one
// vector's entry sums) fits a CU for this problem
bool mix_variant_fits(const sbdx_ctx& c, int prec, int K, int var) {
  if (var < K) return mix_variant_fits(c, prec, var, fused_variant(c, prec, var));  // fused sub-blocks
  const int stages = var / 1000000 + 1, es = prec == 0 ? 8 : 4;
  const int threads = (var % 1000000) / 100, ept = var % 100;
  if (c.L > 64) {  // one stage; 256 threads: V from global memory, 1024: V rows in LDS
    if (stages != 1 || (threads != 256 && threads != 1024)) return false;
    if (threads == 1024 && ept * K > 8) return false;  // would spill: 128 registers at 1024 threads
    return mix_lds_bytes(c.L, c.A.max_s, threads == 256 ? 0 : 1, es, threads * ept, 8) <= kMixMaxLds;
  }
  return mix_lds_bytes(c.L, c.A.max_s, stages, es, threads * ept, 8) <= kMixMaxLds;
}

void sigma_block(sbdx_ctx& c, int prec, int k, const void* x, void* y, hipStream_t st) {
  if (!c.have_ints) throw Error(-2, "integrals not set");
  if (c.stored) {
    sigma_stored(c, prec, k, x, y, st);
    return;
  }
  if (prec == 0) sigma_dispatch<double>(c, k, static_cast<const double*>(x), static_cast<double*>(y), st);
  else sigma_dispatch<float>(c, k, static_cast<const float*>(x), static_cast<float*>(y), st);
}

void sigma_blocked(sbdx_ctx& c, int prec, int nvec, const void* x, void* y, hipStream_t st) {
  const int64_t n = c.ndet();
  const size_t es = prec == 0 ? 8 : 4;
  for (int v = 0; v < nvec; ++v)
    sigma_block(c, prec, 1, static_cast<const char*>(x) + v * n * es, static_cast<char*>(y) + v * n * es, st);
}

void apply_sign_block(sbdx_ctx& c, int prec, int k, const void* x, void* y, hipStream_t st) {
  if (prec == 0) sign_dispatch<double>(c, k, static_cast<const double*>(x), static_cast<double*>(y), st);
  else sign_dispatch<float>(c, k, static_cast<const float*>(x), static_cast<float*>(y), st);
}

void apply_sign(sbdx_ctx& c, int prec, int nvec, const void* x, void* y, hipStream_t st) {
  const int64_t n = c.ndet();
  const size_t es = prec == 0 ? 8 : 4;
  for (int v = 0; v < nvec; ++v)
    apply_sign_block(c, prec, 1, static_cast<const char*>(x) + v * n * es, static_cast<char*>(y) + v * n * es, st);
}

}  // namespace sbdx
