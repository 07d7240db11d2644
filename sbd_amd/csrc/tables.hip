// tables.hip -- string tables, excitation lists, matrix-element tables and H_ii on the GPU.
//
// Excitation lists (chemistry/tpb/helper.h:132-182): for every bra string x, the kets y with
// popcount(x ^ y) == 2 (singles) and == 4 (doubles), in ascending y.  One wavefront owns one
// bra string and scans the kets 64 at a time: XOR + popcount per lane, a ballot per class,
// and the lane's slot is the popcount of the ballot below it, so the lists come out in
// ascending ket order with no sort and no atomics.  A count pass sizes the CSR, a fill pass
// writes it.
#include "common.h"

#include <algorithm>
#include <numeric>

namespace sbdx {
namespace {

constexpr int kWave = 64;

// orbitals in which strings x and y differ (bits of the mask only)
template <int W>
__device__ inline int str_diff(const Str<W>& x, const uint64_t* __restrict__ s, int64_t y, const Str<W>& mask) {
  int d = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) d += popc64((x.w[k] ^ s[y * W + k]) & mask.w[k]);
  return d;
}

template <int W>
__global__ void k_count_exc(const uint64_t* __restrict__ s, int64_t n, Str<W> mask,
                            int64_t* __restrict__ ns, int64_t* __restrict__ nd) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t x = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (x >= n) return;  // whole wavefront leaves together
  const Str<W> sx = ld_str<W>(s, x);
  int64_t cs = 0, cd = 0;
  for (int64_t y0 = 0; y0 < n; y0 += kWave) {
    const int64_t y = y0 + lane;
    const int d = (y < n) ? str_diff(sx, s, y, mask) : 0;
    cs += popc64(__ballot(d == 2));
    cd += popc64(__ballot(d == 4));
  }
  if (lane == 0) {
    ns[x] = cs;
    nd[x] = cd;
  }
}

template <int W>
__global__ void k_fill_exc(const uint64_t* __restrict__ s, int64_t n, Str<W> mask,
                           const int64_t* __restrict__ off_s, uint32_t* __restrict__ ket_s,
                           const int64_t* __restrict__ off_d, uint32_t* __restrict__ ket_d) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t x = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (x >= n) return;
  const Str<W> sx = ld_str<W>(s, x);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  int64_t ps = off_s[x], pd = off_d[x];
  for (int64_t y0 = 0; y0 < n; y0 += kWave) {
    const int64_t y = y0 + lane;
    const int d = (y < n) ? str_diff(sx, s, y, mask) : 0;
    const uint64_t b2 = __ballot(d == 2), b4 = __ballot(d == 4);
    if (d == 2) ket_s[ps + popc64(b2 & below)] = static_cast<uint32_t>(y);
    if (d == 4) ket_d[pd + popc64(b4 & below)] = static_cast<uint32_t>(y);
    ps += popc64(b2);
    pd += popc64(b4);
  }
}

// Matrix-element tables of one spin (blocked order).  One thread per list entry; the bra of
// entry k is found by binary search in the offsets.
__device__ inline int64_t row_of(const int64_t* off, int64_t n, int64_t k) {
  int64_t lo = 0, hi = n;  // off[lo] <= k < off[hi]
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= k) lo = mid; else hi = mid;
  }
  return lo;
}

// singles: value = sign * (h_ai + sum_{j in ket} ((ai|jj) - (aj|ji)))  (determinants.h:400-422
// restricted to the excited spin; the other spin's Coulomb part is added per column)
template <int W>
__global__ void k_single_values(const uint64_t* __restrict__ s, int64_t n, const int64_t* __restrict__ off,
                                const uint32_t* __restrict__ ket, int64_t nnz, int L,
                                const double* __restrict__ h1, const double* __restrict__ eri4,
                                uint32_t* __restrict__ exc, double* __restrict__ val) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int64_t x = row_of(off, n, k);
  const Str<W> sx = ld_str<W>(s, x), sy = ld_str<W>(s, ket[k]);
  const int a = first(andn(sx, sy)), i = first(andn(sy, sx));
  const int sg = sign_single(sy, i, a);
  const int64_t L2 = static_cast<int64_t>(L) * L, L3 = L2 * L;
  double F = h1[a * L + i];
  for_each_bit(sy, [&](int j) { F += eri4[a * L3 + i * L2 + j * L + j] - eri4[a * L3 + j * L2 + j * L + i]; });
  exc[k] = exc_pack(i, a, sg < 0);
  val[k] = sg * F;
}

// doubles: value = sign * ((AI|BJ) - (AJ|BI))  (determinants.h:424-442 on one spin)
template <int W>
__global__ void k_double_values(const uint64_t* __restrict__ s, int64_t n, const int64_t* __restrict__ off,
                                const uint32_t* __restrict__ ket, int64_t nnz, int L,
                                const double* __restrict__ eri4, double* __restrict__ val) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int64_t x = row_of(off, n, k);
  const Str<W> sx = ld_str<W>(s, x), sy = ld_str<W>(s, ket[k]);
  const Str<W> d = andn(sy, sx), c = andn(sx, sy);
  const int I = first(d), J = first(drop_first(d));
  const int A = first(c), B = first(drop_first(c));
  const int sg = sign_double(sy, I, J, A, B);
  const int64_t L2 = static_cast<int64_t>(L) * L, L3 = L2 * L;
  val[k] = sg * (eri4[A * L3 + I * L2 + B * L + J] - eri4[A * L3 + J * L2 + B * L + I]);
}

// same-spin diagonal energy of each string: sum_p h_pp + sum_{p<q} ((pp|qq) - (pq|qp))
template <int W>
__global__ void k_string_energy(const uint64_t* __restrict__ s, int64_t n, int L,
                                const double* __restrict__ h1, const double* __restrict__ eri4,
                                double* __restrict__ e1) {
  const int64_t x = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int64_t L2 = static_cast<int64_t>(L) * L, L3 = L2 * L;
  double e = 0.0;
  Str<W> o = ld_str<W>(s, x);
  for (int m = popc(o); m > 0; --m) {
    const int p = first(o);
    o = drop_first(o);
    e += h1[p * L + p];
    for_each_bit(o, [&](int q) { e += eri4[p * L3 + p * L2 + q * L + q] - eri4[p * L3 + q * L2 + q * L + p]; });
  }
  e1[x] = e;
}

// H_ii(ia, ib) = I0 + E_a(ia) + E_b(ib) + sum_{p in a, q in b} (pp|qq)  (qcham.h:314-366;
// the same terms as ZeroExcite, determinants.h:374-398, grouped by spin)
template <int W>
__global__ void k_hdiag(const uint64_t* __restrict__ sa, const uint64_t* __restrict__ sb,
                        int64_t na, int64_t nb, int L, double I0, const double* __restrict__ ea,
                        const double* __restrict__ eb, const double* __restrict__ eri4,
                        double* __restrict__ hii, float* __restrict__ hii32) {
  __shared__ double u[64 * W];
  const int64_t ntile = cdiv(nb, blockDim.x);
  const int64_t ia = blockIdx.x / ntile;
  const int64_t ib = (blockIdx.x % ntile) * blockDim.x + threadIdx.x;
  const int64_t L2 = static_cast<int64_t>(L) * L, L3 = L2 * L;
  for (int q = threadIdx.x; q < L; q += blockDim.x) {
    double acc = 0.0;
    for_each_bit(ld_str<W>(sa, ia), [&](int p) { acc += eri4[p * L3 + p * L2 + q * L + q]; });
    u[q] = acc;
  }
  __syncthreads();
  if (ib >= nb) return;
  double h = ea[ia] + eb[ib];
  for_each_bit(ld_str<W>(sb, ib), [&](int q) { h += u[q]; });
  h += I0;
  hii[ia * nb + ib] = h;
  if (hii32) hii32[ia * nb + ib] = static_cast<float>(h);
}

__global__ void k_to_float(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = static_cast<float>(x[i]);
}

void build_spin(SpinTab& t, const uint64_t* host_strings, int64_t n, int L, int W, hipStream_t st) {
  t.n = n;
  t.str.alloc(n * W);
  SBDX_HIP(hipMemcpyAsync(t.str.get(), host_strings, n * W * sizeof(uint64_t), hipMemcpyHostToDevice, st));
  DevBuf<int64_t> ns(n), nd(n);
  const int waves_per_block = 4;
  const dim3 grid(static_cast<unsigned>(cdiv(n, waves_per_block)));
  with_words(W, [&](auto words) {
    constexpr int WW = decltype(words)::value;
    Str<WW> mask;
    for (int k = 0; k < WW; ++k) mask.w[k] = mask_range(0, std::min(64, std::max(0, L - 64 * k)));
    k_count_exc<WW><<<grid, waves_per_block * kWave, 0, st>>>(t.str.get(), n, mask, ns.get(), nd.get());
  });
  SBDX_CHECK_LAUNCH();
  std::vector<int64_t> hs(n), hd(n);
  SBDX_HIP(hipMemcpyAsync(hs.data(), ns.get(), n * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SBDX_HIP(hipMemcpyAsync(hd.data(), nd.get(), n * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SBDX_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> os(n + 1, 0), od(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    os[i + 1] = os[i] + hs[i];
    od[i + 1] = od[i] + hd[i];
  }
  t.nnz_s = os[n];
  t.nnz_d = od[n];
  t.max_s = static_cast<int>(n ? *std::max_element(hs.begin(), hs.end()) : 0);
  t.max_d = static_cast<int>(n ? *std::max_element(hd.begin(), hd.end()) : 0);
  t.off_s.alloc(n + 1);
  t.off_d.alloc(n + 1);
  t.ket_s.alloc(t.nnz_s);
  t.ket_d.alloc(t.nnz_d);
  SBDX_HIP(hipMemcpyAsync(t.off_s.get(), os.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  SBDX_HIP(hipMemcpyAsync(t.off_d.get(), od.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  with_words(W, [&](auto words) {
    constexpr int WW = decltype(words)::value;
    Str<WW> mask;
    for (int k = 0; k < WW; ++k) mask.w[k] = mask_range(0, std::min(64, std::max(0, L - 64 * k)));
    k_fill_exc<WW><<<grid, waves_per_block * kWave, 0, st>>>(t.str.get(), n, mask, t.off_s.get(), t.ket_s.get(),
                                                             t.off_d.get(), t.ket_d.get());
  });
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipStreamSynchronize(st));
}

void spin_values(SpinTab& t, int L, int W, const double* h1, const double* eri4, hipStream_t st) {
  t.exc_s.alloc(t.nnz_s);
  t.val_s.alloc(t.nnz_s);
  t.val_d.alloc(t.nnz_d);
  t.e1.alloc(t.n);
  const int bs = 256;
  with_words(W, [&](auto words) {
    constexpr int WW = decltype(words)::value;
    if (t.nnz_s)
      k_single_values<WW><<<static_cast<unsigned>(cdiv(t.nnz_s, bs)), bs, 0, st>>>(
          t.str.get(), t.n, t.off_s.get(), t.ket_s.get(), t.nnz_s, L, h1, eri4, t.exc_s.get(), t.val_s.get());
    SBDX_CHECK_LAUNCH();
    if (t.nnz_d)
      k_double_values<WW><<<static_cast<unsigned>(cdiv(t.nnz_d, bs)), bs, 0, st>>>(
          t.str.get(), t.n, t.off_d.get(), t.ket_d.get(), t.nnz_d, L, eri4, t.val_d.get());
    SBDX_CHECK_LAUNCH();
    k_string_energy<WW><<<static_cast<unsigned>(cdiv(t.n, bs)), bs, 0, st>>>(t.str.get(), t.n, L, h1, eri4,
                                                                             t.e1.get());
  });
  SBDX_CHECK_LAUNCH();
  t.val_s32.release();
  t.val_d32.release();
}

}  // namespace

void build_tables(sbdx_ctx& c, const uint64_t* alpha, const uint64_t* beta) {
  build_spin(c.A, alpha, c.na, c.L, c.W, c.stream);
  build_spin(c.B, beta, c.nb, c.L, c.W, c.stream);
}

// out[p * ncol + col] = sum over the occupied orbitals j of col_str[col], ascending, of
// eri4[p L^2 + j L + j]
template <int W>
__global__ __launch_bounds__(256) void k_coul_table(const uint64_t* __restrict__ col_str, int64_t ncol, int L,
                                                    const double* __restrict__ eri4, double* __restrict__ out) {
  const int64_t nt = cdiv(ncol, 256);
  const int64_t p = blockIdx.x / nt;
  const int64_t col = (blockIdx.x % nt) * 256 + threadIdx.x;
  if (col >= ncol) return;
  const int64_t L2 = static_cast<int64_t>(L) * L;
  const double* __restrict__ vp = eri4 + p * L2;
  double s = 0.0;
  for_each_bit(ld_str<W>(col_str, col), [&](int j) { s += vp[j * L + j]; });
  out[p * ncol + col] = s;
}

void build_integral_tables(sbdx_ctx& c) {
  drop_stored(c);  // a stored H belongs to the previous integrals
  const int L = c.L;
  const int n2 = 2 * L;
  const int64_t L2 = static_cast<int64_t>(L) * L, L4 = L2 * L2;
  std::vector<double> h1a(L2), h1b(L2), eri(L4);
  for (int p = 0; p < L; ++p)
    for (int q = 0; q < L; ++q) {
      h1a[p * L + q] = c.h_I1[(2 * p) * n2 + 2 * q];
      h1b[p * L + q] = c.h_I1[(2 * p + 1) * n2 + 2 * q + 1];
    }
  auto pair = [](int64_t i, int64_t j) { return i > j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; };
  for (int p = 0; p < L; ++p)
    for (int q = 0; q < L; ++q)
      for (int r = 0; r < L; ++r)
        for (int s = 0; s < L; ++s) {
          int64_t a = pair(p, q), b = pair(r, s);
          int64_t hi = a > b ? a : b, lo = a > b ? b : a;
          eri[((static_cast<int64_t>(p) * L + q) * L + r) * L + s] = c.h_I2[hi * (hi + 1) / 2 + lo];
        }
  c.h1a.alloc(L2);
  c.h1b.alloc(L2);
  c.eri4.alloc(L4);
  hipStream_t st = c.stream;
  SBDX_HIP(hipMemcpyAsync(c.h1a.get(), h1a.data(), L2 * 8, hipMemcpyHostToDevice, st));
  SBDX_HIP(hipMemcpyAsync(c.h1b.get(), h1b.data(), L2 * 8, hipMemcpyHostToDevice, st));
  SBDX_HIP(hipMemcpyAsync(c.eri4.get(), eri.data(), L4 * 8, hipMemcpyHostToDevice, st));
  spin_values(c.A, L, c.W, c.h1a.get(), c.eri4.get(), st);
  spin_values(c.B, L, c.W, c.h1b.get(), c.eri4.get(), st);
  const int64_t n = c.ndet();
  c.hii.alloc(n);
  c.hii32.release();
  c.eri4_32.release();
  const int bs = 256;
  const int64_t ntile = cdiv(c.nb, bs);
  c.coul_a.alloc(L2 * c.nb);
  c.coul_b.alloc(L2 * c.na);
  c.coul_a32.release();
  c.coul_b32.release();
  with_words(c.W, [&](auto words) {
    constexpr int WW = decltype(words)::value;
    k_hdiag<WW><<<static_cast<unsigned>(c.na * ntile), bs, 0, st>>>(
        c.A.str.get(), c.B.str.get(), c.na, c.nb, L, c.I0, c.A.e1.get(), c.B.e1.get(), c.eri4.get(),
        c.hii.get(), nullptr);
    SBDX_CHECK_LAUNCH();
    k_coul_table<WW><<<static_cast<unsigned>(L2 * cdiv(c.nb, 256)), 256, 0, st>>>(c.B.str.get(), c.nb, L,
                                                                                    c.eri4.get(), c.coul_a.get());
    k_coul_table<WW><<<static_cast<unsigned>(L2 * cdiv(c.na, 256)), 256, 0, st>>>(c.A.str.get(), c.na, L,
                                                                                    c.eri4.get(), c.coul_b.get());
  });
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipStreamSynchronize(st));
  reset_mix_tiles(c);
  c.have_ints = true;
  prepare_mix_tiles(c, SBDX_FP64, 1);  // FP64 single-vector builds allocate nothing more
}

void ensure_fp32(sbdx_ctx& c) {
  if (c.hii32.get()) return;
  if (c.no_alloc)
    throw Error(SBDX_ERR_STATE, "the FP32 tables are not built: run sbdx_sigma_device once with SBDX_FP32 first");
  hipStream_t st = c.stream;
  const int bs = 256;
  auto conv = [&](const DevBuf<double>& x, DevBuf<float>& y, int64_t n) {
    y.alloc(n);
    if (n) k_to_float<<<static_cast<unsigned>(cdiv(n, bs)), bs, 0, st>>>(x.get(), y.get(), n);
    SBDX_CHECK_LAUNCH();
  };
  for (SpinTab* t : {&c.A, &c.B}) {
    conv(t->val_s, t->val_s32, t->nnz_s);
    conv(t->val_d, t->val_d32, t->nnz_d);
  }
  const int64_t L4 = static_cast<int64_t>(c.L) * c.L * c.L * c.L;
  conv(c.eri4, c.eri4_32, L4);
  conv(c.hii, c.hii32, c.ndet());
  const int64_t L2 = static_cast<int64_t>(c.L) * c.L;
  conv(c.coul_a, c.coul_a32, L2 * c.nb);
  conv(c.coul_b, c.coul_b32, L2 * c.na);
  SBDX_HIP(hipStreamSynchronize(st));
}

}  // namespace sbdx
