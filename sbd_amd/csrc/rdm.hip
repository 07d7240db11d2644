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
// interleaved signs reduce to one-spin signs (see common.h).
#include "common.h"

#include <algorithm>
#include <cstring>

namespace sbdx {
namespace {

constexpr int kB = 256;

__device__ inline double block_reduce(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kB / 64; ++k) r += sh[k];
  __syncthreads();
  return r;
}

// one block per row: rowsq[row], t[row * L + q] = sum_col W^2 n_q(col)
__global__ __launch_bounds__(kB) void k_rowstats(const double* __restrict__ W, int64_t nrow, int64_t ncol,
                                                 const uint64_t* __restrict__ col_str, int L,
                                                 double* __restrict__ rowsq, double* __restrict__ t) {
  __shared__ double bins[64];
  __shared__ double sh[kB / 64];
  const int64_t row = blockIdx.x;
  if (threadIdx.x < 64) bins[threadIdx.x] = 0.0;
  __syncthreads();
  double s = 0.0;
  for (int64_t col = threadIdx.x; col < ncol; col += kB) {
    const double w = W[row * ncol + col];
    const double w2 = w * w;
    s += w2;
    if (t) {
      uint64_t o = col_str[col];
      while (o) {
        atomicAdd(&bins[ctz64(o)], w2);
        o &= o - 1;
      }
    }
  }
  s = block_reduce(s, sh);
  if (threadIdx.x == 0) rowsq[row] = s;
  if (t && threadIdx.x < L) t[row * L + threadIdx.x] = bins[threadIdx.x];
}

// one block per row: for every listed ket k of the row, rho (and rho_q if nq > 0)
__global__ __launch_bounds__(kB) void k_pairdots(const double* __restrict__ W, int64_t nrow, int64_t ncol,
                                                 const int64_t* __restrict__ off,
                                                 const uint32_t* __restrict__ ket,
                                                 const uint64_t* __restrict__ col_str, int L, int nq,
                                                 double* __restrict__ out) {
  __shared__ double bins[64];
  __shared__ double sh[kB / 64];
  const int64_t row = blockIdx.x;
  const int stride = 1 + nq;
  for (int64_t k = off[row]; k < off[row + 1]; ++k) {
    if (nq && threadIdx.x < 64) bins[threadIdx.x] = 0.0;
    __syncthreads();
    const double* wk = W + static_cast<int64_t>(ket[k]) * ncol;
    const double* wr = W + row * ncol;
    double s = 0.0;
    for (int64_t col = threadIdx.x; col < ncol; col += kB) {
      const double p = wr[col] * wk[col];
      s += p;
      if (nq) {
        uint64_t o = col_str[col];
        while (o) {
          atomicAdd(&bins[ctz64(o)], p);
          o &= o - 1;
        }
      }
    }
    s = block_reduce(s, sh);
    if (threadIdx.x == 0) out[k * stride] = s;
    if (nq && threadIdx.x < L) out[k * stride + 1 + threadIdx.x] = bins[threadIdx.x];
    __syncthreads();
  }
}

// alpha single x beta single products binned by (pa, pb); one thread per determinant
__global__ __launch_bounds__(kB) void k_mixed(const double* __restrict__ W, int64_t na, int64_t nb, int L,
                                              const int64_t* __restrict__ aoff, const uint32_t* __restrict__ aket,
                                              const uint32_t* __restrict__ aexc,
                                              const int64_t* __restrict__ boff, const uint32_t* __restrict__ bket,
                                              const uint32_t* __restrict__ bexc, double* __restrict__ G) {
  const int64_t nt = cdiv(nb, kB);
  const int64_t ia = blockIdx.x / nt;
  const int64_t ib = (blockIdx.x % nt) * kB + threadIdx.x;
  if (ib >= nb) return;
  const double wb = W[ia * nb + ib];
  if (wb == 0.0) return;
  const int64_t L2 = static_cast<int64_t>(L) * L;
  for (int64_t k = aoff[ia]; k < aoff[ia + 1]; ++k) {
    const uint32_t e = aexc[k];
    const int64_t pa = ((e >> 6) & 63) * L + (e & 63);
    const double sa = (e >> 12) & 1u ? -wb : wb;
    const double* wr = W + static_cast<int64_t>(aket[k]) * nb;
    for (int64_t m = boff[ib]; m < boff[ib + 1]; ++m) {
      const uint32_t f = bexc[m];
      const int64_t pb = ((f >> 6) & 63) * L + (f & 63);
      const double v = sa * wr[bket[m]];
      atomicAdd(&G[pa * L2 + pb], (f >> 12) & 1u ? -v : v);
    }
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
  std::vector<uint64_t> str;
  std::vector<int64_t> off_s, off_d;
  std::vector<uint32_t> ket_s, ket_d, exc_s;
};

HostSpin host_spin(const SpinTab& t) {
  HostSpin h;
  h.str = to_host(t.str, t.n);
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

// Same-spin single pair of spin s: bra-only orbital p, ket-only q, bra string (same spin)
// spec_same; weight w = sign * rho; wq[o] = sign * rho_q for other-spin spectators.
// Restates OneDiffCorrelation (correlation.h:49-86) summed over the spectator spin.
void one_diff(Rdm& r, int s, int p, int q, uint64_t spec_same, double w, const double* wq) {
  const int t = 1 - s;
  r.g(s, p, q) += w;
  uint64_t o = spec_same;
  while (o) {
    const int j = ctz64(o);
    o &= o - 1;
    r.G(s + 2 * s, q, j, p, j) += w;
    r.G(s + 2 * s, j, q, j, p) += w;
    r.G(s + 2 * s, q, j, j, p) -= w;
    r.G(s + 2 * s, j, q, p, j) -= w;
  }
  for (int j = 0; j < r.L; ++j) {
    const double v = wq[j];
    if (v == 0.0) continue;
    r.G(s + 2 * t, q, j, p, j) += v;
    r.G(t + 2 * s, j, q, j, p) += v;
  }
}

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

}  // namespace

void occupation_blocked(sbdx_ctx& c, const double* dW, double* density) {
  const int L = c.L;
  const int64_t na = c.na, nb = c.nb, n = c.ndet();
  hipStream_t st = c.stream;
  DevBuf<double> rowsq(na), colsq(nb), wT(n);
  k_rowstats<<<static_cast<unsigned>(na), kB, 0, st>>>(dW, na, nb, c.B.str.get(), L, rowsq.get(), nullptr);
  SBDX_CHECK_LAUNCH();
  k_transpose_d<double><<<static_cast<unsigned>(cdiv(na, 64) * cdiv(nb, 64)), 256, 0, st>>>(dW, wT.get(), na, nb);
  SBDX_CHECK_LAUNCH();
  k_rowstats<<<static_cast<unsigned>(nb), kB, 0, st>>>(wT.get(), nb, na, c.A.str.get(), L, colsq.get(), nullptr);
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipStreamSynchronize(st));
  auto ra = to_host(rowsq, na), cb = to_host(colsq, nb);
  auto sa = to_host(c.A.str, na), sb = to_host(c.B.str, nb);
  for (int i = 0; i < 2 * L; ++i) density[i] = 0.0;
  for (int64_t ia = 0; ia < na; ++ia)
    for (int p = 0; p < L; ++p)
      if ((sa[ia] >> p) & 1u) density[2 * p] += ra[ia];
  for (int64_t ib = 0; ib < nb; ++ib)
    for (int p = 0; p < L; ++p)
      if ((sb[ib] >> p) & 1u) density[2 * p + 1] += cb[ib];
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
  k_rowstats<<<static_cast<unsigned>(na), kB, 0, st>>>(dW, na, nb, c.B.str.get(), L, rowsq.get(), ta.get());
  k_rowstats<<<static_cast<unsigned>(nb), kB, 0, st>>>(wT.get(), nb, na, c.A.str.get(), L, colsq.get(), nullptr);
  SBDX_CHECK_LAUNCH();
  k_pairdots<<<static_cast<unsigned>(na), kB, 0, st>>>(dW, na, nb, c.A.off_s.get(), c.A.ket_s.get(), c.B.str.get(), L, L, rs_a.get());
  k_pairdots<<<static_cast<unsigned>(na), kB, 0, st>>>(dW, na, nb, c.A.off_d.get(), c.A.ket_d.get(), c.B.str.get(), L, 0, rd_a.get());
  k_pairdots<<<static_cast<unsigned>(nb), kB, 0, st>>>(wT.get(), nb, na, c.B.off_s.get(), c.B.ket_s.get(), c.A.str.get(), L, L, rs_b.get());
  k_pairdots<<<static_cast<unsigned>(nb), kB, 0, st>>>(wT.get(), nb, na, c.B.off_d.get(), c.B.ket_d.get(), c.A.str.get(), L, 0, rd_b.get());
  SBDX_CHECK_LAUNCH();
  k_mixed<<<static_cast<unsigned>(na * cdiv(nb, kB)), kB, 0, st>>>(
      dW, na, nb, L, c.A.off_s.get(), c.A.ket_s.get(), c.A.exc_s.get(), c.B.off_s.get(), c.B.ket_s.get(),
      c.B.exc_s.get(), G.get());
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipStreamSynchronize(st));

  const HostSpin A = host_spin(c.A), B = host_spin(c.B);
  auto hr = to_host(rowsq, na), hc = to_host(colsq, nb), ht = to_host(ta, na * L);
  auto hsa = to_host(rs_a, c.A.nnz_s * (L + 1)), hda = to_host(rd_a, c.A.nnz_d);
  auto hsb = to_host(rs_b, c.B.nnz_s * (L + 1)), hdb = to_host(rd_b, c.B.nnz_d);
  auto hG = to_host(G, L4);

  std::memset(onebody, 0, sizeof(double) * 2 * L2);
  std::memset(twobody, 0, sizeof(double) * 4 * L4);
  Rdm r(L, onebody, twobody);

  // diagonal (ZeroDiffCorrelation), grouped by spin pair
  for (int64_t ia = 0; ia < na; ++ia) {
    const uint64_t s = A.str[ia];
    for (int p = 0; p < L; ++p) {
      if (!((s >> p) & 1u)) continue;
      r.g(0, p, p) += hr[ia];
      for (int q = 0; q < L; ++q) {
        if (q == p || !((s >> q) & 1u)) continue;
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
    const uint64_t s = B.str[ib];
    for (int p = 0; p < L; ++p) {
      if (!((s >> p) & 1u)) continue;
      r.g(1, p, p) += hc[ib];
      for (int q = 0; q < L; ++q) {
        if (q == p || !((s >> q) & 1u)) continue;
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
    for (int64_t x = 0; x < nrow; ++x) {
      const uint64_t sx = H.str[x];
      for (int64_t k = H.off_s[x]; k < H.off_s[x + 1]; ++k) {
        const uint32_t e = H.exc_s[k];
        const int q = e & 63, p = (e >> 6) & 63;  // q: ket-only (annihilated), p: bra-only
        const double sg = (e >> 12) & 1u ? -1.0 : 1.0;
        double wq[64];
        for (int j = 0; j < L; ++j) wq[j] = sg * rs[k * (L + 1) + 1 + j];
        one_diff(r, spin, p, q, sx, sg * rs[k * (L + 1)], wq);
      }
      for (int64_t k = H.off_d[x]; k < H.off_d[x + 1]; ++k) {
        const uint64_t sy = H.str[H.ket_d[k]];
        uint64_t cbits = sx & ~sy, dbits = sy & ~sx;
        const int c0 = ctz64(cbits);
        cbits &= cbits - 1;
        const int c1 = ctz64(cbits);
        const int d0 = ctz64(dbits);
        dbits &= dbits - 1;
        const int d1 = ctz64(dbits);
        const int sg = sign_double(sx, c0, c1, d0, d1);
        two_diff(r, 2 * c0 + spin, 2 * c1 + spin, 2 * d0 + spin, 2 * d1 + spin, sg * rd[k]);
      }
    }
  }
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
