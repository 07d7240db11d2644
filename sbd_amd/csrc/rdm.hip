// rdm.hip -- 1- and 2-RDMs (chemistry/tpb/correlation.h:194-443) and occupations
// (chemistry/tpb/occupation.h:10-95) from a blocked-basis vector W'.
//
// The reference visits every (bra, ket) pair of the sigma loops and scatters W_bra W_ket sgn
// into gamma and Gamma.  Here the same sums are grouped by excitation class so the GPU only
// produces compact reductions over the spectator spin:
//   diagonal:        rowsq[ia] = sum_ib W'^2, colsq[ib], t[ia][q] = sum_ib W'^2 n_q(ib)
//   alpha single k:  rho_k = sum_ib W'[ia,ib] W'[ja,ib], rho_kq = same weighted by n_q(ib)
//   alpha double k:  rho_k
//   (beta: the same on the transposed vector)
//   alpha x beta:    G[pa][pb] = sum s_a s_b W'[ia,ib] W'[ja,jb] over single pairs with
//                    orbital pairs pa = (a, i), pb = (b, j)
// and the host expands them into the reference's index patterns (ZeroDiffCorrelation
// :15-44, OneDiffCorrelation :49-86, TwoDiffCorrelation :91-130).  In the blocked basis the
// interleaved signs reduce to one-spin signs (see common.h).  No global atomics: row and pair
// sums are wavefront reductions in a fixed order, and the mixed part keeps per-workgroup
// partials in LDS, written once per alpha single and summed in ascending order.
#include "common.h"

#include <algorithm>
#include <cstring>

namespace sbdx {
namespace {

constexpr int kB = 256;

// Sum of v over the 64 lanes of a wavefront in a fixed order (xor butterfly: every lane ends
// with the same value, the same on every run).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Per-lane sums p and, for the first LQ orbitals, p n_q(col): the bins of a column are its
// string's occupied orbitals, taken without branches (select) so they stay in registers.
template <int LQ>
struct Bins {
  double b[LQ];
  __device__ inline void zero() {
#pragma unroll
    for (int q = 0; q < LQ; ++q) b[q] = 0.0;
  }
  __device__ inline void add(uint64_t str, double p) {
#pragma unroll
    for (int q = 0; q < LQ; ++q) b[q] += ((str >> q) & 1ull) ? p : 0.0;
  }
};

// One workgroup per row (4 wavefronts over the columns): rowsq[row] = sum_col W^2 and, when t
// is given, t[row * L + q] = sum_col W^2 n_q(col).  Lane sums -> wavefront sums -> the four
// wavefronts added in order: no atomics, the same bits on every run.  Strings of SW words: the
// bins cover word `word` (orbitals 64 word + q), one launch per word.
template <int LQ, int SW = 1>
__global__ __launch_bounds__(kB) void k_rowstats(const double* __restrict__ W, int64_t ncol,
                                                 const uint64_t* __restrict__ col_str, int L,
                                                 double* __restrict__ rowsq, double* __restrict__ t,
                                                 int word = 0) {
  __shared__ double part[kB / 64][LQ + 1];
  const int64_t row = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double s = 0.0;
  Bins<LQ> bins;
  bins.zero();
  for (int64_t col = threadIdx.x; col < ncol; col += kB) {
    const double w = W[row * ncol + col];
    const double w2 = w * w;
    s += w2;
    if (t) bins.add(col_str[col * SW + word], w2);
  }
  s = wave_sum(s);
  if (lane == 0) part[wv][0] = s;
  if (t) {
#pragma unroll
    for (int q = 0; q < LQ; ++q) {
      const double v = wave_sum(bins.b[q]);
      if (lane == 0) part[wv][1 + q] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x <= LQ) {
    double r = 0.0;
    for (int k = 0; k < kB / 64; ++k) r += part[k][threadIdx.x];
    if (threadIdx.x == 0) rowsq[row] = r;
    else if (t && 64 * word + threadIdx.x - 1 < L) t[row * L + 64 * word + threadIdx.x - 1] = r;
  }
}

// One workgroup per row, one wavefront per G listed kets (the row's kets dealt round-robin over
// the four wavefronts, G at a time): rho_k = sum_col W[row, col] W[ket_k, col] and, with kBins,
// rho_kq = the same weighted by n_q(col) for q < L.  out[k * (1 + L)] = rho_k,
// out[k * (1 + L) + 1 + q] = rho_kq (stride 1 without bins).  Each lane takes C consecutive
// columns (16-byte loads for C = 2) and the G kets share every load of the bra row.  No
// barriers, no atomics.
// C consecutive doubles (C = 2: one 16-byte load; p 16-byte aligned)
template <int C>
struct Cols {
  double v[C];
};
template <int C>
__device__ inline Cols<C> ld_cols(const double* p) {
  Cols<C> r;
  if constexpr (C == 2) {
    const double2 t = *reinterpret_cast<const double2*>(p);
    r.v[0] = t.x;
    r.v[1] = t.y;
  } else {
    r.v[0] = *p;
  }
  return r;
}

template <int LQ, bool kBins, int G, int C, int SW = 1>
__global__ __launch_bounds__(kB) void k_pairdots(const double* __restrict__ W, int64_t ncol,
                                                 const int64_t* __restrict__ off,
                                                 const uint32_t* __restrict__ ket,
                                                 const uint64_t* __restrict__ col_str, int L,
                                                 double* __restrict__ out, int word = 0) {
  static_assert(!kBins || G == 1, "the weighted sums keep one ket's bins per lane");
  const int64_t row = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int stride = kBins ? 1 + L : 1;
  const double* __restrict__ wr = W + row * ncol;
  const int64_t kend = off[row + 1];
  for (int64_t k0 = off[row] + wv * G; k0 < kend; k0 += (kB / 64) * G) {
    const double* __restrict__ wk[G];
#pragma unroll
    for (int g = 0; g < G; ++g)  // past the row's kets: reread the first one, drop its sum
      wk[g] = W + static_cast<int64_t>(ket[k0 + g < kend ? k0 + g : k0]) * ncol;
    double s[G];
#pragma unroll
    for (int g = 0; g < G; ++g) s[g] = 0.0;
    Bins<kBins ? LQ : 1> bins;
    bins.zero();
#pragma unroll 2
    for (int64_t col = static_cast<int64_t>(lane) * C; col < ncol; col += 64 * C) {
      const Cols<C> a = ld_cols<C>(wr + col);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const Cols<C> b = ld_cols<C>(wk[g] + col);
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const double p = a.v[c] * b.v[c];
          s[g] += p;
          if constexpr (kBins) bins.add(col_str[(col + c) * SW + word], p);
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const double v = wave_sum(s[g]);
      if (lane == 0 && k0 + g < kend) out[(k0 + g) * stride] = v;
    }
    if (kBins) {
#pragma unroll
      for (int q = 0; q < (kBins ? LQ : 1); ++q) {
        const double v = wave_sum(bins.b[q]);
        if (lane == 0 && 64 * word + q < L) out[k0 * stride + 1 + 64 * word + q] = v;
      }
    }
  }
}

// Alpha single x beta single part of the 2-RDM:
//   Z[k][pb] = sum over beta singles (ib -> jb, pair pb, sign s_b) of s_b W[ia, ib] W[ja, jb]
// for every alpha single k = (ia <- ja) of a chunk of alpha rows, and then
//   G[pa][pb] = sum over alpha singles k with pair pa of s_a Z[k][pb]   (k_rdm_fold).
// Beta singles are cut by pair range (kPbr pairs each) and, inside a range, sorted by jb and
// cut into tiles of kThreads x kEpt entries, so the gathers of W[ja, jb] along a tile touch few
// lines.  One workgroup = (alpha row ia, pair range r): its tile entries and their bra values
// W[ia, ib] sit in registers; the row's alpha singles are taken kmax at a time, each with an
// LDS bin row of the range's pairs (LDS atomic adds: per-workgroup partials; the bins of one
// alpha single are written once to Z, with no global atomics, and G is summed from Z in
// ascending k).
constexpr int kRdmThreads = 512, kRdmEpt = 8;
constexpr int kRdmTile = kRdmThreads * kRdmEpt;
constexpr int kRdmBinBytes = 64 * 1024;

__global__ __launch_bounds__(kRdmThreads) void k_rdm_mixed(
    const double* __restrict__ W, int64_t nb, int64_t row0, int64_t nrow, int L2, int pbr, int kmax,
    const int64_t* __restrict__ a_off, const uint32_t* __restrict__ a_ket, const int64_t* __restrict__ r_tile,
    const int64_t* __restrict__ t_off, const uint32_t* __restrict__ e_jb, const uint32_t* __restrict__ e_ib,
    const uint32_t* __restrict__ e_meta, int64_t k_base, double* __restrict__ Z) {
  extern __shared__ __align__(16) double bins[];  // [kmax][pbr]
  const int nrange = static_cast<int>(gridDim.x / nrow);
  const int64_t ia = row0 + blockIdx.x / nrange;
  const int r = static_cast<int>(blockIdx.x % nrange);
  const int tid = threadIdx.x;
  const int64_t k0 = a_off[ia], k1 = a_off[ia + 1];
  const int pb0 = r * pbr;
  const int npb = L2 - pb0 < pbr ? L2 - pb0 : pbr;
  const double* __restrict__ wbra = W + ia * nb;
  for (int64_t kc = k0; kc < k1; kc += kmax) {
    const int nk = static_cast<int>(k1 - kc < kmax ? k1 - kc : kmax);
    for (int q = tid; q < nk * pbr; q += kRdmThreads) bins[q] = 0.0;
    __syncthreads();
    for (int64_t t = r_tile[r]; t < r_tile[r + 1]; ++t) {
      const int64_t e0 = t_off[t], e1 = t_off[t + 1];
      uint32_t jb[kRdmEpt], meta[kRdmEpt];
      double wb[kRdmEpt];
#pragma unroll
      for (int j = 0; j < kRdmEpt; ++j) {
        const int64_t e = e0 + tid + static_cast<int64_t>(j) * kRdmThreads;
        const bool ok = e < e1;
        jb[j] = ok ? e_jb[e] : 0u;
        meta[j] = ok ? e_meta[e] : 0u;
        const double w = ok ? wbra[e_ib[e]] : 0.0;
        wb[j] = meta[j] >> 31 ? -w : w;
      }
      for (int q = 0; q < nk; ++q) {
        const double* __restrict__ xr = W + static_cast<int64_t>(a_ket[kc + q]) * nb;
        double xv[kRdmEpt];
#pragma unroll
        for (int j = 0; j < kRdmEpt; ++j) xv[j] = xr[jb[j]];
        double* __restrict__ row = bins + q * pbr;
#pragma unroll
        for (int j = 0; j < kRdmEpt; ++j) {
          const double p = wb[j] * xv[j];
          if (p != 0.0) atomicAdd(row + (meta[j] & 0xffffu), p);
        }
      }
    }
    __syncthreads();
    for (int q = tid; q < nk * pbr; q += kRdmThreads) {
      const int k = q / pbr, pb = q - k * pbr;
      if (pb < npb) Z[(kc + k - k_base) * L2 + pb0 + pb] = bins[q];
    }
    __syncthreads();
  }
}

// G[pa][pb] += sum over the alpha singles of the chunk with pair pa (ascending k) of
// s_a Z[k][pb]; one workgroup per pa.
__global__ __launch_bounds__(kB) void k_rdm_fold(const double* __restrict__ Z, int L2,
                                                 const int64_t* __restrict__ pa_off,
                                                 const int64_t* __restrict__ pa_k, int64_t k_base,
                                                 double* __restrict__ G) {
  const int64_t pa = blockIdx.x;
  const int64_t q0 = pa_off[pa], q1 = pa_off[pa + 1];
  if (q0 == q1) return;
  for (int pb = threadIdx.x; pb < L2; pb += kB) {
    double s = 0.0;
    for (int64_t q = q0; q < q1; ++q) {
      const int64_t kk = pa_k[q];
      const int64_t k = kk < 0 ? ~kk : kk;  // negative: s_a = -1
      const double z = Z[(k - k_base) * L2 + pb];
      s += kk < 0 ? -z : z;
    }
    G[pa * L2 + pb] += s;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_transpose_d(const T* __restrict__ in, T* __restrict__ out,
                                                     int64_t R, int64_t C) {
  __shared__ T tile[64][65];
  const int64_t tc = cdiv(C, 64);
  const int64_t br = blockIdx.x / tc, bc = blockIdx.x % tc;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int k = ty; k < 64; k += 4) {
    const int64_t r = br * 64 + k, c = bc * 64 + tx;
    if (r < R && c < C) tile[k][tx] = in[r * C + c];
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const int64_t c = bc * 64 + k, r = br * 64 + tx;
    if (r < R && c < C) out[c * R + r] = tile[tx][k];
  }
}

template <class T>
std::vector<T> to_host(const DevBuf<T>& d, int64_t n) {
  std::vector<T> h(std::max<int64_t>(n, 1));
  if (n) SBDX_HIP(hipMemcpy(h.data(), d.get(), n * sizeof(T), hipMemcpyDeviceToHost));
  return h;
}

struct HostSpin {
  int64_t n = 0;
  std::vector<uint64_t> str;  // n strings of the context's W words
  std::vector<int64_t> off_s, off_d;
  std::vector<uint32_t> ket_s, ket_d, exc_s;
};

HostSpin host_spin(const SpinTab& t, int W) {
  HostSpin h;
  h.n = t.n;
  h.str = to_host(t.str, t.n * W);
  h.off_s = to_host(t.off_s, t.n + 1);
  h.off_d = to_host(t.off_d, t.n + 1);
  h.ket_s = to_host(t.ket_s, t.nnz_s);
  h.ket_d = to_host(t.ket_d, t.nnz_d);
  h.exc_s = to_host(t.exc_s, t.nnz_s);
  return h;
}

struct Rdm {
  int L;
  double* one;
  double* two;
  int64_t L2, L3, L4;
  Rdm(int l, double* o, double* t) : L(l), one(o), two(t) {
    L2 = static_cast<int64_t>(L) * L;
    L3 = L2 * L;
    L4 = L3 * L;
  }
  double& g(int s, int i, int j) { return one[s * L2 + i + L * j]; }
  double& G(int st, int i, int j, int k, int l) { return two[st * L4 + i + L * j + L2 * k + L3 * l]; }
};

// Same-spin singles of spin s, summed per orbital pair (bra-only p, ket-only q) before they are
// expanded: w0[p, q] = sum of sign * rho, w1[p, q, j] = the same over the singles whose bra
// string holds j (same-spin spectators), w2[p, q, j] = sum of sign * rho_j (other-spin
// spectators j).  Adding into these L^3 arrays keeps the per-single work in cache; expand()
// then writes OneDiffCorrelation's index patterns (correlation.h:49-86) once per (p, q, j).
struct SinglesByPair {
  int L;
  std::vector<double> w0, w1, w2;
  explicit SinglesByPair(int l)
      : L(l), w0(static_cast<size_t>(l) * l, 0.0), w1(static_cast<size_t>(l) * l * l, 0.0),
        w2(static_cast<size_t>(l) * l * l, 0.0) {}
  template <int SW>
  void add(int p, int q, const Str<SW>& spec_same, double w, const double* wq) {
    const size_t pq = static_cast<size_t>(p) * L + q;
    w0[pq] += w;
    double* a1 = &w1[pq * L];
    for_each_bit(spec_same, [&](int j) { a1[j] += w; });
    double* a2 = &w2[pq * L];
    for (int j = 0; j < L; ++j) a2[j] += wq[j];
  }
  void expand(Rdm& r, int s) const {
    const int t = 1 - s;
    for (int p = 0; p < L; ++p)
      for (int q = 0; q < L; ++q) {
        const size_t pq = static_cast<size_t>(p) * L + q;
        r.g(s, p, q) += w0[pq];
        for (int j = 0; j < L; ++j) {
          const double w = w1[pq * L + j];
          if (w != 0.0) {
            r.G(s + 2 * s, q, j, p, j) += w;
            r.G(s + 2 * s, j, q, j, p) += w;
            r.G(s + 2 * s, q, j, j, p) -= w;
            r.G(s + 2 * s, j, q, p, j) -= w;
          }
          const double v = w2[pq * L + j];
          if (v != 0.0) {
            r.G(s + 2 * t, q, j, p, j) += v;
            r.G(t + 2 * s, j, q, j, p) += v;
          }
        }
      }
  }
};

// TwoDiffCorrelation (correlation.h:91-130) with interleaved spin-orbitals c0<c1 (bra-only)
// and d0<d1 (ket-only); w already carries the blocked-basis sign.
void two_diff(Rdm& r, int c0, int c1, int d0, int d1, double w) {
  const int I = c0, J = c1, A = d0, B = d1;
  const int oi = I / 2, si = I % 2, oa = A / 2, sa = A % 2;
  const int oj = J / 2, sj = J % 2, ob = B / 2, sb = B % 2;
  if (si == sa) {
    r.G(si + 2 * sj, oa, ob, oi, oj) += w;
    r.G(sj + 2 * si, ob, oa, oj, oi) += w;
  }
  if (si == sb) {
    r.G(si + 2 * sj, oa, ob, oj, oi) -= w;
    r.G(sj + 2 * si, ob, oa, oi, oj) -= w;
  }
}

// Pair-range tiles of the beta singles for the mixed part (k_rdm_mixed), built on the host
// from the beta lists: entries cut by pair range, sorted by (jb, ib) inside a range, and cut
// into tiles of kRdmTile entries.
struct RdmTiles {
  int pbr = 0, nrange = 0, kmax = 0;
  std::vector<int64_t> r_tile, t_off;
  std::vector<uint32_t> jb, ib, meta;
};

RdmTiles rdm_tiles(const HostSpin& B, int L, int max_alpha_singles) {
  RdmTiles T;
  const int L2 = L * L;
  T.kmax = std::max(1, std::min(max_alpha_singles, 32));
  int pbr = std::min<int>(L2, kRdmBinBytes / (8 * T.kmax));
  T.nrange = static_cast<int>(cdiv(L2, pbr));
  T.pbr = static_cast<int>(cdiv(L2, T.nrange));
  struct Ent {
    uint32_t r, jb, ib, meta;
  };
  std::vector<Ent> ent;
  const int64_t nb = B.n;
  ent.reserve(B.ket_s.size());
  for (int64_t ib = 0; ib < nb; ++ib)
    for (int64_t k = B.off_s[ib]; k < B.off_s[ib + 1]; ++k) {
      const uint32_t e = B.exc_s[k];
      const uint32_t pb = exc_pair(e, L);
      const uint32_t r = pb / static_cast<uint32_t>(T.pbr);
      ent.push_back(Ent{r, B.ket_s[k], static_cast<uint32_t>(ib),
                        (pb - r * static_cast<uint32_t>(T.pbr)) | (exc_neg(e) ? 1u << 31 : 0u)});
    }
  std::sort(ent.begin(), ent.end(), [](const Ent& u, const Ent& v) {
    return u.r != v.r ? u.r < v.r : (u.jb != v.jb ? u.jb < v.jb : u.ib < v.ib);
  });
  T.jb.resize(ent.size());
  T.ib.resize(ent.size());
  T.meta.resize(ent.size());
  for (size_t q = 0; q < ent.size(); ++q) {
    T.jb[q] = ent[q].jb;
    T.ib[q] = ent[q].ib;
    T.meta[q] = ent[q].meta;
  }
  T.t_off.push_back(0);
  size_t q = 0;
  for (int r = 0; r < T.nrange; ++r) {
    T.r_tile.push_back(static_cast<int64_t>(T.t_off.size()) - 1);
    size_t end = q;
    while (end < ent.size() && static_cast<int>(ent[end].r) == r) ++end;
    while (q < end) {
      q = std::min(end, q + static_cast<size_t>(kRdmTile));
      T.t_off.push_back(static_cast<int64_t>(q));
    }
  }
  T.r_tile.push_back(static_cast<int64_t>(T.t_off.size()) - 1);
  return T;
}

template <class T>
DevBuf<T> to_dev(const std::vector<T>& h, hipStream_t st) {
  DevBuf<T> d(std::max<size_t>(h.size(), 1));
  if (!h.empty()) SBDX_HIP(hipMemcpy(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// G[pa][pb] (pa = a L + i, pb = b L + j) of the alpha single x beta single part, on the device.
void rdm_mixed(sbdx_ctx& c, const double* dW, const HostSpin& A, const HostSpin& B, double* G) {
  const int L = c.L, L2 = L * L;
  const int64_t na = c.na, nb = c.nb;
  hipStream_t st = c.stream;
  if (c.A.nnz_s == 0 || c.B.nnz_s == 0) return;
  const RdmTiles T = rdm_tiles(B, L, c.A.max_s);
  auto d_rt = to_dev(T.r_tile, st);
  auto d_to = to_dev(T.t_off, st);
  auto d_jb = to_dev(T.jb, st);
  auto d_ib = to_dev(T.ib, st);
  auto d_meta = to_dev(T.meta, st);
  // alpha rows in chunks whose Z (alpha singles x L^2 FP64) stays under zcap bytes
  const int64_t zcap = int64_t(6) << 30;
  const int64_t per_single = static_cast<int64_t>(L2) * 8;
  int64_t max_rows_singles = 0;
  for (int64_t ia = 0; ia < na; ++ia) max_rows_singles = std::max(max_rows_singles, A.off_s[ia + 1] - A.off_s[ia]);
  const int64_t zk = std::max<int64_t>(max_rows_singles, std::min<int64_t>(c.A.nnz_s, zcap / per_single));
  DevBuf<double> Z(zk * L2);
  const size_t lds = static_cast<size_t>(T.kmax) * T.pbr * 8;
  int64_t row0 = 0;
  while (row0 < na) {
    int64_t row1 = row0;
    while (row1 < na && A.off_s[row1 + 1] - A.off_s[row0] <= zk) ++row1;
    const int64_t k_base = A.off_s[row0], nk = A.off_s[row1] - k_base;
    if (nk > 0) {
      // the chunk's alpha singles grouped by pair pa, ascending k; ~k marks s_a = -1
      std::vector<int64_t> pa_off(L2 + 1, 0), pa_k(nk);
      for (int64_t k = k_base; k < k_base + nk; ++k) {
        ++pa_off[exc_pair(A.exc_s[k], L) + 1];
      }
      for (int p = 0; p < L2; ++p) pa_off[p + 1] += pa_off[p];
      std::vector<int64_t> fill(pa_off.begin(), pa_off.end() - 1);
      for (int64_t k = k_base; k < k_base + nk; ++k) {
        const uint32_t e = A.exc_s[k];
        pa_k[fill[exc_pair(e, L)]++] = exc_neg(e) ? ~k : k;
      }
      auto d_pa_off = to_dev(pa_off, st);
      auto d_pa_k = to_dev(pa_k, st);
      const int64_t nrow = row1 - row0;
      k_rdm_mixed<<<static_cast<unsigned>(nrow * T.nrange), kRdmThreads, lds, st>>>(
          dW, nb, row0, nrow, L2, T.pbr, T.kmax, c.A.off_s.get(), c.A.ket_s.get(), d_rt.get(), d_to.get(),
          d_jb.get(), d_ib.get(), d_meta.get(), k_base, Z.get());
      SBDX_CHECK_LAUNCH();
      k_rdm_fold<<<static_cast<unsigned>(L2), kB, 0, st>>>(Z.get(), L2, d_pa_off.get(), d_pa_k.get(), k_base, G);
      SBDX_CHECK_LAUNCH();
      SBDX_HIP(hipStreamSynchronize(st));  // host vectors of this chunk go out of scope
    }
    row0 = row1;
  }
}

}  // namespace

void occupation_blocked(sbdx_ctx& c, const double* dW, double* density) {
  const int L = c.L;
  const int64_t na = c.na, nb = c.nb, n = c.ndet();
  hipStream_t st = c.stream;
  DevBuf<double> rowsq(na), colsq(nb), wT(n);
  k_rowstats<1><<<static_cast<unsigned>(na), kB, 0, st>>>(dW, nb, c.B.str.get(), L, rowsq.get(), nullptr);
  SBDX_CHECK_LAUNCH();
  k_transpose_d<double><<<static_cast<unsigned>(cdiv(na, 64) * cdiv(nb, 64)), 256, 0, st>>>(dW, wT.get(), na, nb);
  SBDX_CHECK_LAUNCH();
  k_rowstats<1><<<static_cast<unsigned>(nb), kB, 0, st>>>(wT.get(), na, c.A.str.get(), L, colsq.get(), nullptr);
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipStreamSynchronize(st));
  auto ra = to_host(rowsq, na), cb = to_host(colsq, nb);
  const int SW = c.W;
  auto sa = to_host(c.A.str, na * SW), sb = to_host(c.B.str, nb * SW);
  auto bit = [&](const std::vector<uint64_t>& s, int64_t x, int p) { return (s[x * SW + (p >> 6)] >> (p & 63)) & 1u; };
  for (int i = 0; i < 2 * L; ++i) density[i] = 0.0;
  for (int64_t ia = 0; ia < na; ++ia)
    for (int p = 0; p < L; ++p)
      if (bit(sa, ia, p)) density[2 * p] += ra[ia];
  for (int64_t ib = 0; ib < nb; ++ib)
    for (int p = 0; p < L; ++p)
      if (bit(sb, ib, p)) density[2 * p + 1] += cb[ib];
}

void rdm_blocked(sbdx_ctx& c, const double* dW, double* onebody, double* twobody) {
  const int L = c.L;
  const int64_t na = c.na, nb = c.nb, n = c.ndet();
  const int64_t L2 = static_cast<int64_t>(L) * L, L4 = L2 * L2;
  hipStream_t st = c.stream;
  DevBuf<double> wT(n), rowsq(na), colsq(nb), ta(na * L);
  DevBuf<double> rs_a((c.A.nnz_s + 1) * (L + 1)), rd_a(c.A.nnz_d + 1);
  DevBuf<double> rs_b((c.B.nnz_s + 1) * (L + 1)), rd_b(c.B.nnz_d + 1);
  DevBuf<double> G(L4);
  SBDX_HIP(hipMemsetAsync(G.get(), 0, L4 * 8, st));
  k_transpose_d<double><<<static_cast<unsigned>(cdiv(na, 64) * cdiv(nb, 64)), 256, 0, st>>>(dW, wT.get(), na, nb);
  SBDX_CHECK_LAUNCH();
  auto stats = [&](auto lq) {
    constexpr int LQ = decltype(lq)::value;
    const unsigned ga = static_cast<unsigned>(na), gb = static_cast<unsigned>(nb);
    k_rowstats<1><<<gb, kB, 0, st>>>(wT.get(), na, c.A.str.get(), L, colsq.get(), nullptr);
    // the spectator-weighted sums, one launch per string word (the plain sums of the second
    // launch rewrite the same values)
    with_words(c.W, [&](auto words) {
      constexpr int SW = decltype(words)::value;
      for (int word = 0; word < SW; ++word) {
        k_rowstats<LQ, SW><<<ga, kB, 0, st>>>(dW, nb, c.B.str.get(), L, rowsq.get(), ta.get(), word);
        SBDX_CHECK_LAUNCH();
        k_pairdots<LQ, true, 1, 1, SW><<<ga, kB, 0, st>>>(dW, nb, c.A.off_s.get(), c.A.ket_s.get(), c.B.str.get(), L,
                                                         rs_a.get(), word);
        k_pairdots<LQ, true, 1, 1, SW><<<gb, kB, 0, st>>>(wT.get(), na, c.B.off_s.get(), c.B.ket_s.get(),
                                                         c.A.str.get(), L, rs_b.get(), word);
      }
    });
    // the same-spin doubles: eight kets per wavefront, two columns per lane where the rows hold
    // whole 16-byte words
    auto doubles = [&](unsigned g, const double* w, int64_t ncol, const SpinTab& t, double* out) {
      if (ncol % 2 == 0) k_pairdots<1, false, 8, 2><<<g, kB, 0, st>>>(w, ncol, t.off_d.get(), t.ket_d.get(), nullptr, L, out);
      else k_pairdots<1, false, 8, 1><<<g, kB, 0, st>>>(w, ncol, t.off_d.get(), t.ket_d.get(), nullptr, L, out);
    };
    doubles(ga, dW, nb, c.A, rd_a.get());
    doubles(gb, wT.get(), na, c.B, rd_b.get());
    SBDX_CHECK_LAUNCH();
  };
  if (L <= 32) stats(std::integral_constant<int, 32>{});
  else if (L <= 48) stats(std::integral_constant<int, 48>{});
  else stats(std::integral_constant<int, 64>{});
  const HostSpin A = host_spin(c.A, c.W), B = host_spin(c.B, c.W);
  rdm_mixed(c, dW, A, B, G.get());
  SBDX_HIP(hipStreamSynchronize(st));
  auto hr = to_host(rowsq, na), hc = to_host(colsq, nb), ht = to_host(ta, na * L);
  auto hsa = to_host(rs_a, c.A.nnz_s * (L + 1)), hda = to_host(rd_a, c.A.nnz_d);
  auto hsb = to_host(rs_b, c.B.nnz_s * (L + 1)), hdb = to_host(rd_b, c.B.nnz_d);
  auto hG = to_host(G, L4);

  std::memset(onebody, 0, sizeof(double) * 2 * L2);
  std::memset(twobody, 0, sizeof(double) * 4 * L4);
  Rdm r(L, onebody, twobody);

  with_words(c.W, [&](auto words) {
    constexpr int SW = decltype(words)::value;
    // diagonal (ZeroDiffCorrelation), grouped by spin pair
    for (int64_t ia = 0; ia < na; ++ia) {
      const Str<SW> s = ld_str<SW>(A.str.data(), ia);
      for (int p = 0; p < L; ++p) {
        if (!test(s, p)) continue;
        r.g(0, p, p) += hr[ia];
        for (int q = 0; q < L; ++q) {
          if (q == p || !test(s, q)) continue;
          r.G(0, p, q, p, q) += hr[ia];
          r.G(0, p, q, q, p) -= hr[ia];
        }
        for (int q = 0; q < L; ++q) {
          const double v = ht[ia * L + q];
          r.G(0 + 2 * 1, p, q, p, q) += v;
          r.G(1 + 2 * 0, q, p, q, p) += v;
        }
      }
    }
    for (int64_t ib = 0; ib < nb; ++ib) {
      const Str<SW> s = ld_str<SW>(B.str.data(), ib);
      for (int p = 0; p < L; ++p) {
        if (!test(s, p)) continue;
        r.g(1, p, p) += hc[ib];
        for (int q = 0; q < L; ++q) {
          if (q == p || !test(s, q)) continue;
          r.G(3, p, q, p, q) += hc[ib];
          r.G(3, p, q, q, p) -= hc[ib];
        }
      }
    }
    // same-spin singles and doubles
    for (int spin = 0; spin < 2; ++spin) {
      const HostSpin& H = spin == 0 ? A : B;
      const std::vector<double>& rs = spin == 0 ? hsa : hsb;
      const std::vector<double>& rd = spin == 0 ? hda : hdb;
      const int64_t nrow = spin == 0 ? na : nb;
      SinglesByPair sp(L);
      // doubles summed per orbital quadruple (bra-only c0 < c1, ket-only d0 < d1) first, then
      // expanded once per quadruple
      std::vector<double> dq(static_cast<size_t>(L4), 0.0);
      for (int64_t x = 0; x < nrow; ++x) {
        const Str<SW> sx = ld_str<SW>(H.str.data(), x);
        for (int64_t k = H.off_s[x]; k < H.off_s[x + 1]; ++k) {
          const uint32_t e = H.exc_s[k];
          const int q = exc_i(e), p = exc_a(e);  // q: ket-only (annihilated), p: bra-only
          const double sg = exc_neg(e) ? -1.0 : 1.0;
          double wq[128];
          for (int j = 0; j < L; ++j) wq[j] = sg * rs[k * (L + 1) + 1 + j];
          sp.add(p, q, sx, sg * rs[k * (L + 1)], wq);
        }
        for (int64_t k = H.off_d[x]; k < H.off_d[x + 1]; ++k) {
          const Str<SW> sy = ld_str<SW>(H.str.data(), H.ket_d[k]);
          const Str<SW> cbits = andn(sx, sy), dbits = andn(sy, sx);
          const int c0 = first(cbits), c1 = first(drop_first(cbits));
          const int d0 = first(dbits), d1 = first(drop_first(dbits));
          const int sg = sign_double(sx, c0, c1, d0, d1);
          dq[((static_cast<size_t>(c0) * L + c1) * L + d0) * L + d1] += sg * rd[k];
        }
      }
      sp.expand(r, spin);
      for (int c0 = 0; c0 < L; ++c0)
        for (int c1 = c0 + 1; c1 < L; ++c1)
          for (int d0 = 0; d0 < L; ++d0)
            for (int d1 = d0 + 1; d1 < L; ++d1) {
              const double w = dq[((static_cast<size_t>(c0) * L + c1) * L + d0) * L + d1];
              if (w != 0.0) two_diff(r, 2 * c0 + spin, 2 * c1 + spin, 2 * d0 + spin, 2 * d1 + spin, w);
            }
    }
  });
  // alpha single x beta single
  for (int a = 0; a < L; ++a)
    for (int i = 0; i < L; ++i)
      for (int b = 0; b < L; ++b)
        for (int j = 0; j < L; ++j) {
          const double v = hG[(static_cast<int64_t>(a) * L + i) * L2 + b * L + j];
          if (v == 0.0) continue;
          // bra-only {alpha a, beta b}, ket-only {alpha i, beta j} in interleaved indices
          const int ca = 2 * a, cb = 2 * b + 1, da = 2 * i, db = 2 * j + 1;
          const int c0 = std::min(ca, cb), c1 = std::max(ca, cb);
          const int d0 = std::min(da, db), d1 = std::max(da, db);
          const int si = c0 % 2, sa_ = d0 % 2;
          two_diff(r, c0, c1, d0, d1, si == sa_ ? v : -v);
        }
}

}  // namespace sbdx
