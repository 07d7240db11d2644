// stored.hip -- the stored-Hamiltonian path (the reference's method 1: makeQCham,
// chemistry/tpb/qcham.h:10-312, and its mult, mult.h:13-123).
//
// The whole H (blocked basis) is built once as CSR on the GPU from the excitation lists: row
// r = ia * nb + ib holds, in this order, the diagonal, the alpha doubles and singles, the beta
// doubles and singles, and the alpha-single x beta-single pairs.  Row lengths come from the list
// offsets, the row pointer from a device prefix sum, and one wavefront fills a row (its lanes
// write consecutive entries).  Element values follow the on-the-fly kernels exactly (same
// tables, same per-column Coulomb term).  A context with a stored H uses it for every sigma
// build (any solver, any block width): one wavefront per row reads the row's (column, value)
// pairs coalesced and gathers x.
#include "common.h"
#include "sbdx.h"


#include <algorithm>

namespace sbdx {
namespace {

constexpr int kWaves = 4;  // wavefronts per workgroup

struct Lists {
  const int64_t* off_s;
  const uint32_t* ket_s;
  const uint32_t* exc_s;
  const double* val_s;
  const int64_t* off_d;
  const uint32_t* ket_d;
  const double* val_d;
  const uint64_t* str;
};

__global__ void k_csr_len(Lists A, Lists B, int64_t na, int64_t nb, int64_t* __restrict__ len) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = na * nb;
  if (r > n) return;
  if (r == n) {
    len[n] = 0;
    return;
  }
  const int64_t ia = r / nb, ib = r - ia * nb;
  const int64_t sa = A.off_s[ia + 1] - A.off_s[ia], da = A.off_d[ia + 1] - A.off_d[ia];
  const int64_t sb = B.off_s[ib + 1] - B.off_s[ib], db = B.off_d[ib + 1] - B.off_d[ib];
  len[r] = 1 + sa + da + sb + db + sa * sb;
}

// value of a single excitation entry q of one spin, with the spectator string's Coulomb term
template <int W>
__device__ inline double single_value(const Lists& T, int64_t q, const Str<W>& spectator, const double* V, int L) {
  const uint32_t e = T.exc_s[q];
  const int64_t L2 = static_cast<int64_t>(L) * L;
  const double* vp = V + static_cast<int64_t>(exc_pair(e, L)) * L2;
  double coul = 0.0;
  for_each_bit(spectator, [&](int j) { coul += vp[j * L + j]; });
  return T.val_s[q] + (exc_neg(e) ? -coul : coul);
}

template <int W>
__global__ __launch_bounds__(64 * kWaves) void k_csr_fill(Lists A, Lists B, int64_t na, int64_t nb, int L,
                                                          const double* __restrict__ V,
                                                          const double* __restrict__ hii,
                                                          const int64_t* __restrict__ rowptr,
                                                          uint32_t* __restrict__ col, double* __restrict__ val) {
  const int lane = threadIdx.x & 63;
  const int64_t n = na * nb;
  const int64_t L2 = static_cast<int64_t>(L) * L;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6); r < n;
       r += static_cast<int64_t>(gridDim.x) * kWaves) {
    const int64_t ia = r / nb, ib = r - ia * nb;
    int64_t p = rowptr[r];
    if (lane == 0) {
      col[p] = static_cast<uint32_t>(r);
      val[p] = hii[r];
    }
    ++p;
    // alpha doubles and singles: (ja, ib)
    const int64_t ad0 = A.off_d[ia], ad1 = A.off_d[ia + 1];
    for (int64_t q = ad0 + lane; q < ad1; q += 64) {
      col[p + q - ad0] = static_cast<uint32_t>(static_cast<int64_t>(A.ket_d[q]) * nb + ib);
      val[p + q - ad0] = A.val_d[q];
    }
    p += ad1 - ad0;
    const int64_t as0 = A.off_s[ia], as1 = A.off_s[ia + 1];
    const Str<W> bstr = ld_str<W>(B.str, ib), astr = ld_str<W>(A.str, ia);
    for (int64_t q = as0 + lane; q < as1; q += 64) {
      col[p + q - as0] = static_cast<uint32_t>(static_cast<int64_t>(A.ket_s[q]) * nb + ib);
      val[p + q - as0] = single_value(A, q, bstr, V, L);
    }
    p += as1 - as0;
    // beta doubles and singles: (ia, jb)
    const int64_t bd0 = B.off_d[ib], bd1 = B.off_d[ib + 1];
    for (int64_t q = bd0 + lane; q < bd1; q += 64) {
      col[p + q - bd0] = static_cast<uint32_t>(ia * nb + B.ket_d[q]);
      val[p + q - bd0] = B.val_d[q];
    }
    p += bd1 - bd0;
    const int64_t bs0 = B.off_s[ib], bs1 = B.off_s[ib + 1];
    for (int64_t q = bs0 + lane; q < bs1; q += 64) {
      col[p + q - bs0] = static_cast<uint32_t>(ia * nb + B.ket_s[q]);
      val[p + q - bs0] = single_value(B, q, astr, V, L);
    }
    p += bs1 - bs0;
    // alpha single x beta single: (ja, jb), s_a s_b (a i | b j)
    const int64_t ns = bs1 - bs0, npair = (as1 - as0) * ns;
    for (int64_t t = lane; t < npair; t += 64) {
      const int64_t qa = as0 + t / ns, qb = bs0 + t % ns;
      const uint32_t ea = A.exc_s[qa], eb = B.exc_s[qb];
      const double v = V[static_cast<int64_t>(exc_pair(ea, L)) * L2 + exc_pair(eb, L)];
      const bool neg = exc_neg(ea) != exc_neg(eb);
      col[p + t] = static_cast<uint32_t>(static_cast<int64_t>(A.ket_s[qa]) * nb + B.ket_s[qb]);
      val[p + t] = neg ? -v : v;
    }
  }
}

__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[i] = static_cast<float>(in[i]);
}

// y = H x for K interleaved vectors; one wavefront per row
template <typename T, int K>
__global__ __launch_bounds__(64 * kWaves) void k_spmv(const int64_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                                      const T* __restrict__ val, const T* __restrict__ x,
                                                      T* __restrict__ y, int64_t n) {
  const int lane = threadIdx.x & 63;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6); r < n;
       r += static_cast<int64_t>(gridDim.x) * kWaves) {
    T acc[K];
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] = T(0);
    const int64_t p0 = rowptr[r], p1 = rowptr[r + 1];
    for (int64_t p = p0 + lane; p < p1; p += 64) {
      const T w = val[p];
      const int64_t c = col[p];
#pragma unroll
      for (int v = 0; v < K; ++v) acc[v] = fma(w, x[c * K + v], acc[v]);
    }
#pragma unroll
    for (int v = 0; v < K; ++v) {
      T s = acc[v];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
      if (lane == 0) y[r * K + v] = s;
    }
  }
}

// Exclusive prefix sum of n int64 values: each workgroup scans 1024 of them (wavefront scans
// of per-thread sums, then the four wavefronts' totals), the workgroups' totals are scanned on
// the host, and a second pass adds each workgroup's offset.  Integer sums: exact.
constexpr int kScanItems = 4;
constexpr int kScanBlock = 256 * kScanItems;

__global__ __launch_bounds__(256) void k_scan_local(const int64_t* __restrict__ in, int64_t* __restrict__ out,
                                                    int64_t n, int64_t* __restrict__ totals) {
  __shared__ int64_t wave_total[4];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanBlock + static_cast<int64_t>(threadIdx.x) * kScanItems;
  int64_t v[kScanItems];
  int64_t t = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < n ? in[base + i] : 0;
    t += v[i];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int64_t x = t;  // inclusive scan of the per-thread sums across the wavefront
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wave_total[wv] = x;
  __syncthreads();
  int64_t woff = 0;
  for (int w = 0; w < wv; ++w) woff += wave_total[w];
  int64_t run = woff + x - t;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 255) totals[blockIdx.x] = woff + x;
}

__global__ __launch_bounds__(256) void k_scan_add(int64_t* __restrict__ out, int64_t n,
                                                  const int64_t* __restrict__ offsets) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) out[i] += offsets[i / kScanBlock];
}

void exclusive_scan(const int64_t* in, int64_t* out, int64_t n, hipStream_t st) {
  const int64_t nblk = cdiv(n, kScanBlock);
  DevBuf<int64_t> totals(nblk);
  k_scan_local<<<static_cast<unsigned>(nblk), 256, 0, st>>>(in, out, n, totals.get());
  SBDX_CHECK_LAUNCH();
  std::vector<int64_t> h(nblk);
  SBDX_HIP(hipMemcpyAsync(h.data(), totals.get(), nblk * 8, hipMemcpyDeviceToHost, st));
  SBDX_HIP(hipStreamSynchronize(st));
  int64_t run = 0;
  for (int64_t b = 0; b < nblk; ++b) {
    const int64_t t = h[b];
    h[b] = run;
    run += t;
  }
  SBDX_HIP(hipMemcpyAsync(totals.get(), h.data(), nblk * 8, hipMemcpyHostToDevice, st));
  k_scan_add<<<static_cast<unsigned>(cdiv(n, 256)), 256, 0, st>>>(out, n, totals.get());
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipStreamSynchronize(st));  // the host buffer h is read by the copy above
}

Lists lists(const SpinTab& t) {
  return Lists{t.off_s.get(), t.ket_s.get(), t.exc_s.get(), t.val_s.get(), t.off_d.get(), t.ket_d.get(), t.val_d.get(),
               t.str.get()};
}

}  // namespace

int64_t store_hamiltonian(sbdx_ctx& c) {
  if (!c.have_ints) throw Error(SBDX_ERR_STATE, "integrals not set");
  const int64_t n = c.ndet();
  if (n >= (int64_t(1) << 31)) throw Error(SBDX_ERR_ARG, "stored Hamiltonian: too many determinants");
  hipStream_t st = c.stream;
  DevBuf<int64_t> rowptr(n + 1), len(n + 1);
  const Lists A = lists(c.A), B = lists(c.B);
  k_csr_len<<<static_cast<unsigned>(cdiv(n + 1, 256)), 256, 0, st>>>(A, B, c.na, c.nb, len.get());
  SBDX_CHECK_LAUNCH();
  exclusive_scan(len.get(), rowptr.get(), n + 1, st);
  int64_t nnz = 0;
  SBDX_HIP(hipMemcpyAsync(&nnz, rowptr.get() + n, 8, hipMemcpyDeviceToHost, st));
  SBDX_HIP(hipStreamSynchronize(st));
  size_t free_b = 0, total_b = 0;
  SBDX_HIP(hipMemGetInfo(&free_b, &total_b));
  const double need = static_cast<double>(nnz) * 16.0;
  if (need > 0.8 * static_cast<double>(free_b))
    throw Error(SBDX_ERR_DEVICE, "stored Hamiltonian needs " + std::to_string(need / 1e9) + " GB (" +
                                     std::to_string(nnz) + " elements); use the on-the-fly sigma");
  c.h_col.alloc(nnz);
  c.h_val.alloc(nnz);
  c.h_val32.release();
  const unsigned g = static_cast<unsigned>(std::min<int64_t>(cdiv(n, kWaves), 65536));
  with_words(c.W, [&](auto words) {
    k_csr_fill<decltype(words)::value><<<g, 64 * kWaves, 0, st>>>(A, B, c.na, c.nb, c.L, c.eri4.get(), c.hii.get(),
                                                              rowptr.get(), c.h_col.get(), c.h_val.get());
  });
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipStreamSynchronize(st));
  c.h_rowptr = std::move(rowptr);
  c.h_nnz = nnz;
  c.stored = true;
  return nnz;
}

void drop_stored(sbdx_ctx& c) {
  c.stored = false;
  c.h_rowptr.release();
  c.h_col.release();
  c.h_val.release();
  c.h_val32.release();
  c.h_nnz = 0;
}

void sigma_stored(sbdx_ctx& c, int prec, int k, const void* x, void* y, hipStream_t st) {
  const int64_t n = c.ndet();
  const unsigned g = static_cast<unsigned>(std::min<int64_t>(cdiv(n, kWaves), 65536));
  auto go = [&](auto tag) {
    using T = decltype(tag);
    const T* val;
    if constexpr (std::is_same<T, float>::value) {
      if (!c.h_val32.get()) {
        c.h_val32.alloc(c.h_nnz);
        k_to_float<<<4096, 256, 0, st>>>(c.h_val.get(), c.h_val32.get(), c.h_nnz);
        SBDX_CHECK_LAUNCH();
      }
      val = c.h_val32.get();
    } else {
      val = c.h_val.get();
    }
    const T* xx = static_cast<const T*>(x);
    T* yy = static_cast<T*>(y);
    switch (k) {
      case 1: k_spmv<T, 1><<<g, 64 * kWaves, 0, st>>>(c.h_rowptr.get(), c.h_col.get(), val, xx, yy, n); break;
      case 2: k_spmv<T, 2><<<g, 64 * kWaves, 0, st>>>(c.h_rowptr.get(), c.h_col.get(), val, xx, yy, n); break;
      case 4: k_spmv<T, 4><<<g, 64 * kWaves, 0, st>>>(c.h_rowptr.get(), c.h_col.get(), val, xx, yy, n); break;
      case 8: k_spmv<T, 8><<<g, 64 * kWaves, 0, st>>>(c.h_rowptr.get(), c.h_col.get(), val, xx, yy, n); break;
      default: throw Error(SBDX_ERR_ARG, "block width must be 1, 2, 4 or 8");
    }
    SBDX_CHECK_LAUNCH();
  };
  if (prec == SBDX_FP32) go(float{});
  else go(double{});
}

}  // namespace sbdx
