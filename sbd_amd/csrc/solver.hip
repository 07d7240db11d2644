// solver.hip -- block thick-restart Davidson with mixed precision (method 1).
//
// Subspace: the anchor block X (K vectors, FP64) plus up to m-1 correction blocks stored in
// precision P (FP64, or FP32 for the mixed-precision solver).  Every restart recomputes the
// anchor's sigma in FP64, so the residual that decides convergence is exact FP64; the correction
// blocks and their sigmas carry P rounding, which enters the Ritz vectors only through their
// coefficients, which vanish as the iteration converges (an outer FP64 refinement loop).
// Rayleigh-Ritz solves H_sub u = e S_sub u with both matrices accumulated in FP64 from the stored
// blocks, by canonical orthogonalization, so FP32 blocks need not be exactly orthonormal.  New
// blocks: preconditioned residuals, two passes of classical Gram-Schmidt against the subspace in
// FP64 arithmetic, then Loewdin orthonormalization of the block itself.  Thick restart keeps the
// K Ritz vectors of the whole subspace as the next anchor, and the correction part of those
// Ritz vectors (their component outside the old anchor, with its sigma, a linear combination of
// the correction blocks' sigmas) as the first block of the next subspace: the restarted method
// then keeps the previous search direction, as LOBPCG does, instead of losing it.
//
// Every block product is tall-skinny: the kernels stream their blocks once ([N][K] interleaved,
// K contiguous values per determinant) and reduce K x K partial sums per workgroup.
//
// The single-vector solver that follows the reference step for step (chemistry/tpb/davidson.h:
// 661-923) is davidson_fp64 (davidson.hip, method 0).
#include "common.h"
#include "sbdx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>

namespace sbdx {

void jacobi_syev_host(int n, double* a, int lda, double* w);  // davidson.hip

namespace {

constexpr int kB = 256;
constexpr int kMaxBlocks = 32;

inline unsigned nblk(int64_t n) {
  return static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>(cdiv(n, kB * 4), 1), 2048));
}

__device__ inline double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// partial[b][a*K + c] = sum over this workgroup's rows of A[n][a] * B[n][c]
template <typename TA, typename TB, int K>
__global__ __launch_bounds__(kB) void k_gram(const TA* __restrict__ A, const TB* __restrict__ B, int64_t n,
                                             double* __restrict__ partial) {
  __shared__ double sh[kB / 64][K * K];
  double acc[K][K];
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int c = 0; c < K; ++c) acc[a][c] = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    double av[K], bv[K];
#pragma unroll
    for (int a = 0; a < K; ++a) av[a] = static_cast<double>(A[i * K + a]);
#pragma unroll
    for (int c = 0; c < K; ++c) bv[c] = static_cast<double>(B[i * K + c]);
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int c = 0; c < K; ++c) acc[a][c] = __fma_rn(av[a], bv[c], acc[a][c]);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const double s = wsum(acc[a][c]);
      if (lane == 0) sh[w][a * K + c] = s;
    }
  __syncthreads();
  for (int q = threadIdx.x; q < K * K; q += kB) {
    double s = 0.0;
#pragma unroll
    for (int ww = 0; ww < kB / 64; ++ww) s += sh[ww][q];
    partial[static_cast<int64_t>(blockIdx.x) * K * K + q] = s;
  }
}

// Several blocks against one: partial[g][(q K + a) K + c] = sum over this workgroup's
// determinants of L_q[det K + a] R[det K + c], for NL left blocks (FP64 or FP32 each) and one
// right block.  Consecutive lanes take consecutive determinants, so every block is read in
// whole lines and the right block once for all NL left ones.  The products are HBM-bound
// (NL K^2 flops per NL K + K values read), so they run on the vector ALUs: a matrix-core
// version (v_mfma_f64_16x16x4f64 over LDS-staged chunks) measured 1.4-3.6x slower, since the
// FP64 shape contracts 4 determinants per instruction and a K-wide right block fills K of its
// 16 output columns.
struct LeftSet {
  const void* p[32];
  int dbl[32];  // 1: FP64 block, 0: FP32 block
};

template <typename TR, int K, int NL>
__global__ __launch_bounds__(kB) void k_multigram(LeftSet L, const TR* __restrict__ R, int64_t n,
                                                  double* __restrict__ partial) {
  __shared__ double sh[kB / 64][NL * K * K];
  double acc[NL][K][K];
#pragma unroll
  for (int q = 0; q < NL; ++q)
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int c = 0; c < K; ++c) acc[q][a][c] = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    double r[K];
#pragma unroll
    for (int c = 0; c < K; ++c) r[c] = static_cast<double>(R[i * K + c]);
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      double l[K];
      if (L.dbl[q]) {
#pragma unroll
        for (int a = 0; a < K; ++a) l[a] = static_cast<const double*>(L.p[q])[i * K + a];
      } else {
#pragma unroll
        for (int a = 0; a < K; ++a) l[a] = static_cast<double>(static_cast<const float*>(L.p[q])[i * K + a]);
      }
#pragma unroll
      for (int a = 0; a < K; ++a)
#pragma unroll
        for (int c = 0; c < K; ++c) acc[q][a][c] = __fma_rn(l[a], r[c], acc[q][a][c]);
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NL; ++q)
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const double s = wsum(acc[q][a][c]);
        if (lane == 0) sh[w][(q * K + a) * K + c] = s;
      }
  __syncthreads();
  for (int t = threadIdx.x; t < NL * K * K; t += kB) {
    double s = 0.0;
#pragma unroll
    for (int ww = 0; ww < kB / 64; ++ww) s += sh[ww][t];
    partial[static_cast<int64_t>(blockIdx.x) * NL * K * K + t] = s;
  }
}

// left blocks per k_multigram pass: NL K^2 accumulators per thread, at most 32 (64 for K = 8)
template <int K>
constexpr int multigram_max_nl() {
  return 32 / (K * K) > 1 ? 32 / (K * K) : 1;
}

// out[q] = sum_b partial[b][q], one workgroup per q
__global__ __launch_bounds__(kB) void k_sum_partials(const double* __restrict__ partial, int nblock, int m,
                                                    double* __restrict__ out) {
  __shared__ double sh[kB / 64];
  const int q = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblock; b += kB) s += partial[static_cast<int64_t>(b) * m + q];
  s = wsum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < kB / 64; ++w) r += sh[w];
    out[q] = r;
  }
}

struct BlockPtrs {
  const void* b[kMaxBlocks];
  const void* hb[kMaxBlocks];
};

// Subspace combination per determinant: y = X u_0 + sum_{i>=1} B_i u_i, hy likewise from the
// sigma blocks (u_i: K x K blocks of coef, row-major [a][v]).  mode 0 writes the residual
// hy - y diag(e), mode 1 the Ritz vectors y.  Workgroup partial sums of |out_v|^2.
template <typename P, int K>
__global__ __launch_bounds__(kB) void k_combine(const double* __restrict__ X, const double* __restrict__ HX,
                                                BlockPtrs bp, int nblocks, const double* __restrict__ coef,
                                                const double* __restrict__ e, int mode, double* __restrict__ out,
                                                int64_t n, double* __restrict__ partial) {
  __shared__ double cs[kMaxBlocks * K * K];
  __shared__ double es[K];
  __shared__ double sh[kB / 64][K];
  for (int q = threadIdx.x; q < nblocks * K * K; q += kB) cs[q] = coef[q];
  if (threadIdx.x < K) es[threadIdx.x] = e[threadIdx.x];
  __syncthreads();
  double nrm[K];
#pragma unroll
  for (int v = 0; v < K; ++v) nrm[v] = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    double y[K], hy[K];
#pragma unroll
    for (int v = 0; v < K; ++v) y[v] = hy[v] = 0.0;
    {
      double xa[K], ha[K];
#pragma unroll
      for (int a = 0; a < K; ++a) {
        xa[a] = X[i * K + a];
        ha[a] = mode == 0 ? HX[i * K + a] : 0.0;
      }
#pragma unroll
      for (int a = 0; a < K; ++a)
#pragma unroll
        for (int v = 0; v < K; ++v) {
          y[v] = __fma_rn(xa[a], cs[a * K + v], y[v]);
          hy[v] = __fma_rn(ha[a], cs[a * K + v], hy[v]);
        }
    }
    for (int blk = 1; blk < nblocks; ++blk) {
      const P* B = static_cast<const P*>(bp.b[blk]);
      const P* HB = static_cast<const P*>(bp.hb[blk]);
      const double* cb = cs + blk * K * K;
      double xa[K], ha[K];
#pragma unroll
      for (int a = 0; a < K; ++a) {
        xa[a] = static_cast<double>(B[i * K + a]);
        ha[a] = mode == 0 ? static_cast<double>(HB[i * K + a]) : 0.0;
      }
#pragma unroll
      for (int a = 0; a < K; ++a)
#pragma unroll
        for (int v = 0; v < K; ++v) {
          y[v] = __fma_rn(xa[a], cb[a * K + v], y[v]);
          hy[v] = __fma_rn(ha[a], cb[a * K + v], hy[v]);
        }
    }
#pragma unroll
    for (int v = 0; v < K; ++v) {
      const double o = mode == 0 ? hy[v] - es[v] * y[v] : y[v];
      out[i * K + v] = o;
      nrm[v] = __fma_rn(o, o, nrm[v]);
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < K; ++v) {
    const double a = wsum(nrm[v]);
    if (lane == 0) sh[w][v] = a;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int ww = 0; ww < kB / 64; ++ww) s += sh[ww][threadIdx.x];
    partial[static_cast<int64_t>(blockIdx.x) * K + threadIdx.x] = s;
  }
}

// T = R / (e_v - H_ii) with the reference's regulariser (davidson.h:855-862), stored in P
template <typename P, int K>
__global__ __launch_bounds__(kB) void k_precond_diag(const double* __restrict__ R, const double* __restrict__ e,
                                                     const double* __restrict__ hii, P* __restrict__ T, int64_t n) {
  double es[K];
#pragma unroll
  for (int v = 0; v < K; ++v) es[v] = e[v];
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    const double h = hii[i];
#pragma unroll
    for (int v = 0; v < K; ++v) {
      const double den = es[v] - h;
      const double d = fabs(den) > 1.0e-12 ? den : den - 1.0e-12;
      T[i * K + v] = static_cast<P>(R[i * K + v] / d);
    }
  }
}

// T -= X c_0 + sum_{i>=1} B_i c_i (the projection step of classical Gram-Schmidt), FP64 math
template <typename P, int K>
__global__ __launch_bounds__(kB) void k_project_out(P* __restrict__ T, const double* __restrict__ X, BlockPtrs bp,
                                                    int nblocks, const double* __restrict__ coef, int64_t n) {
  __shared__ double cs[kMaxBlocks * K * K];
  for (int q = threadIdx.x; q < nblocks * K * K; q += kB) cs[q] = coef[q];
  __syncthreads();
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    double t[K];
#pragma unroll
    for (int v = 0; v < K; ++v) t[v] = static_cast<double>(T[i * K + v]);
#pragma unroll
    for (int a = 0; a < K; ++a) {
      const double xa = X[i * K + a];
#pragma unroll
      for (int v = 0; v < K; ++v) t[v] = __fma_rn(-xa, cs[a * K + v], t[v]);
    }
    for (int blk = 1; blk < nblocks; ++blk) {
      const P* B = static_cast<const P*>(bp.b[blk]);
      const double* cb = cs + blk * K * K;
#pragma unroll
      for (int a = 0; a < K; ++a) {
        const double ba = static_cast<double>(B[i * K + a]);
#pragma unroll
        for (int v = 0; v < K; ++v) t[v] = __fma_rn(-ba, cb[a * K + v], t[v]);
      }
    }
#pragma unroll
    for (int v = 0; v < K; ++v) T[i * K + v] = static_cast<P>(t[v]);
  }
}

// The correction part of the Ritz block and its sigma: P = sum_{i>=1} B_i u_i,
// HP = sum_{i>=1} HB_i u_i (FP64 sums, stored in P); the search direction kept across a restart
template <typename P, int K>
__global__ __launch_bounds__(kB) void k_corr(BlockPtrs bp, int nblocks, const double* __restrict__ coef,
                                             P* __restrict__ outP, P* __restrict__ outHP, int64_t n) {
  __shared__ double cs[kMaxBlocks * K * K];
  for (int q = threadIdx.x; q < nblocks * K * K; q += kB) cs[q] = coef[q];
  __syncthreads();
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    double y[K], hy[K];
#pragma unroll
    for (int v = 0; v < K; ++v) y[v] = hy[v] = 0.0;
    for (int blk = 1; blk < nblocks; ++blk) {
      const P* B = static_cast<const P*>(bp.b[blk]);
      const P* HB = static_cast<const P*>(bp.hb[blk]);
      const double* cb = cs + blk * K * K;
#pragma unroll
      for (int a = 0; a < K; ++a) {
        const double xa = static_cast<double>(B[i * K + a]), ha = static_cast<double>(HB[i * K + a]);
#pragma unroll
        for (int v = 0; v < K; ++v) {
          y[v] = __fma_rn(xa, cb[a * K + v], y[v]);
          hy[v] = __fma_rn(ha, cb[a * K + v], hy[v]);
        }
      }
    }
#pragma unroll
    for (int v = 0; v < K; ++v) {
      outP[i * K + v] = static_cast<P>(y[v]);
      outHP[i * K + v] = static_cast<P>(hy[v]);
    }
  }
}

// P -= X c, HP -= HX c (c: K x K row-major [a][v]); FP64 arithmetic
template <typename P, int K>
__global__ __launch_bounds__(kB) void k_project_pair(P* __restrict__ Pv, P* __restrict__ HPv, const double* __restrict__ X,
                                                     const double* __restrict__ HX, const double* __restrict__ coef,
                                                     int64_t n) {
  __shared__ double cs[K * K];
  for (int q = threadIdx.x; q < K * K; q += kB) cs[q] = coef[q];
  __syncthreads();
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    double p[K], hp[K];
#pragma unroll
    for (int v = 0; v < K; ++v) {
      p[v] = static_cast<double>(Pv[i * K + v]);
      hp[v] = static_cast<double>(HPv[i * K + v]);
    }
#pragma unroll
    for (int a = 0; a < K; ++a) {
      const double xa = X[i * K + a], ha = HX[i * K + a];
#pragma unroll
      for (int v = 0; v < K; ++v) {
        p[v] = __fma_rn(-xa, cs[a * K + v], p[v]);
        hp[v] = __fma_rn(-ha, cs[a * K + v], hp[v]);
      }
    }
#pragma unroll
    for (int v = 0; v < K; ++v) {
      Pv[i * K + v] = static_cast<P>(p[v]);
      HPv[i * K + v] = static_cast<P>(hp[v]);
    }
  }
}

// Y <- Y M (K x K, row-major [a][v]) in place
template <typename TY, int K>
__global__ __launch_bounds__(kB) void k_rotate(TY* __restrict__ Y, const double* __restrict__ M, int64_t n) {
  __shared__ double ms[K * K];
  for (int q = threadIdx.x; q < K * K; q += kB) ms[q] = M[q];
  __syncthreads();
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    double y[K], o[K];
#pragma unroll
    for (int a = 0; a < K; ++a) {
      y[a] = static_cast<double>(Y[i * K + a]);
      o[a] = 0.0;
    }
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int v = 0; v < K; ++v) o[v] = __fma_rn(y[a], ms[a * K + v], o[v]);
#pragma unroll
    for (int v = 0; v < K; ++v) Y[i * K + v] = static_cast<TY>(o[v]);
  }
}

// dst[i * ds + od] = src[i * ss + os] for i < n (one column of an interleaved block)
__global__ __launch_bounds__(kB) void k_copy_col(double* __restrict__ dst, int ds, int od,
                                                 const double* __restrict__ src, int ss, int os, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB)
    dst[i * ds + od] = src[i * ss + os];
}

// the kq smallest H_ii over i >= 1 per workgroup (kq <= 8): values and indices, ascending
// the kq lowest H_ii of each workgroup's share, determinant 0 excluded; with masks (okA, okB)
// only determinants whose alpha and beta strings are both marked
__global__ __launch_bounds__(kB) void k_lowest(const double* __restrict__ hii, int64_t n, int kq,
                                               const uint8_t* __restrict__ okA, const uint8_t* __restrict__ okB,
                                               int64_t nb, double* __restrict__ val, int64_t* __restrict__ idx) {
  __shared__ double sv[kB * 8];
  __shared__ int64_t si[kB * 8];
  double v[8];
  int64_t ix[8];
  for (int a = 0; a < 8; ++a) {
    v[a] = 1e300;
    ix[a] = -1;
  }
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x + 1; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    if (okA && !(okA[i / nb] && okB[i % nb])) continue;
    const double h = hii[i];
    if (h < v[kq - 1]) {
      int p = kq - 1;
      while (p > 0 && v[p - 1] > h) {
        v[p] = v[p - 1];
        ix[p] = ix[p - 1];
        --p;
      }
      v[p] = h;
      ix[p] = i;
    }
  }
  for (int a = 0; a < kq; ++a) {
    sv[threadIdx.x * 8 + a] = v[a];
    si[threadIdx.x * 8 + a] = ix[a];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bv[8];
    int64_t bi[8];
    for (int a = 0; a < 8; ++a) {
      bv[a] = 1e300;
      bi[a] = -1;
    }
    for (int t = 0; t < kB; ++t)
      for (int a = 0; a < kq; ++a) {
        const double h = sv[t * 8 + a];
        const int64_t hi = si[t * 8 + a];
        if (hi < 0 || !(h < bv[kq - 1])) continue;
        int p = kq - 1;
        while (p > 0 && bv[p - 1] > h) {
          bv[p] = bv[p - 1];
          bi[p] = bi[p - 1];
          --p;
        }
        bv[p] = h;
        bi[p] = hi;
      }
    for (int a = 0; a < kq; ++a) {
      val[blockIdx.x * kq + a] = bv[a];
      idx[blockIdx.x * kq + a] = bi[a];
    }
  }
}

// Union-find over one spin's strings joined by their single and double excitations: 1 for the
// strings in string 0's component.  Returns whether that component holds every string.
bool component_of_string0(const SpinTab& t, std::vector<uint8_t>& mark) {
  const int64_t n = t.n;
  std::vector<int64_t> off_s(n + 1), off_d(n + 1);
  std::vector<uint32_t> ket_s(std::max<int64_t>(t.nnz_s, 1)), ket_d(std::max<int64_t>(t.nnz_d, 1));
  SBDX_HIP(hipMemcpy(off_s.data(), t.off_s.get(), (n + 1) * 8, hipMemcpyDeviceToHost));
  SBDX_HIP(hipMemcpy(off_d.data(), t.off_d.get(), (n + 1) * 8, hipMemcpyDeviceToHost));
  if (t.nnz_s) SBDX_HIP(hipMemcpy(ket_s.data(), t.ket_s.get(), t.nnz_s * 4, hipMemcpyDeviceToHost));
  if (t.nnz_d) SBDX_HIP(hipMemcpy(ket_d.data(), t.ket_d.get(), t.nnz_d * 4, hipMemcpyDeviceToHost));
  std::vector<int64_t> parent(n);
  for (int64_t x = 0; x < n; ++x) parent[x] = x;
  auto find = [&](int64_t x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  for (int64_t x = 0; x < n; ++x) {
    for (int64_t k = off_s[x]; k < off_s[x + 1]; ++k) parent[find(ket_s[k])] = find(x);
    for (int64_t k = off_d[x]; k < off_d[x + 1]; ++k) parent[find(ket_d[k])] = find(x);
  }
  const int64_t r0 = find(0);
  mark.assign(n, 0);
  bool all = true;
  for (int64_t x = 0; x < n; ++x) {
    mark[x] = find(x) == r0;
    all = all && mark[x];
  }
  return all;
}

// Determinant 0's block of H (built once per context): H couples no determinant inside it to
// one outside, so a Davidson run started from determinant 0 alone -- the reference's, init 0 --
// stays in it.  State 1: every string connected (the whole basis is one block).
void build_block0(sbdx_ctx& c) {
  if (c.block0_state) return;
  std::vector<uint8_t> ma, mb;
  const bool all_a = component_of_string0(c.A, ma), all_b = component_of_string0(c.B, mb);
  if (all_a && all_b) {
    c.block0_state = 1;
    return;
  }
  c.block0_a.alloc(c.na);
  c.block0_b.alloc(c.nb);
  SBDX_HIP(hipMemcpy(c.block0_a.get(), ma.data(), c.na, hipMemcpyHostToDevice));
  SBDX_HIP(hipMemcpy(c.block0_b.get(), mb.data(), c.nb, hipMemcpyHostToDevice));
  c.block0_state = 2;
}

__global__ void k_set_entries(double* __restrict__ X, const int64_t* __restrict__ pos, int m) {
  const int q = threadIdx.x;
  if (q < m) X[pos[q]] = 1.0;
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename P, int K>
struct BlockSolver {
  sbdx_ctx& c;
  const sbdx_davidson_opts& o;
  hipStream_t st;
  const int64_t n;
  const int m;                      // blocks in a full subspace: anchor + m-1 corrections
  const int nroots;                 // roots that must converge (<= K)
  DevBuf<double> X, HX, R;          // anchor, its FP64 sigma, residual / next anchor
  std::vector<DevBuf<P>> B, HB;     // corrections 1..m-1 and their sigmas (B[0] unused)
  DevBuf<P> T;                      // block under construction
  DevBuf<P> Pd, HPd;                // search direction kept across a restart, and its sigma
  bool have_dir = false;
  DevBuf<double> partial, small;    // reduction partials; staged coefficients
  std::vector<double> Hs, Ss;       // (m K) x (m K) subspace matrices, row-major
  bool use_mlp;                     // learned preconditioner in use
  int cycle = 0;                    // restart cycle of the current step
  int fallback_step = 0;            // step at which it was dropped (0: never)
  static constexpr int prec = std::is_same<P, float>::value ? SBDX_FP32 : SBDX_FP64;

  BlockSolver(sbdx_ctx& cc, const sbdx_davidson_opts& oo)
      : c(cc), o(oo), st(cc.stream), n(cc.ndet()),
        m(std::max(2, std::min(oo.num_block, kMaxBlocks))), nroots(std::max(1, oo.nroots)),
        use_mlp(oo.preconditioner == 1) {
    X.alloc(n * K);
    HX.alloc(n * K);
    R.alloc(n * K);
    T.alloc(n * K);
    Pd.alloc(n * K);
    HPd.alloc(n * K);
    B.resize(m);
    HB.resize(m);
    for (int i = 1; i < m; ++i) {
      B[i].alloc(n * K);
      HB[i].alloc(n * K);
    }
    partial.alloc(static_cast<size_t>(2048) * std::max(K * K, 32));  // k_gram, k_multigram
    small.alloc(kMaxBlocks * K * K + K);
    Hs.assign(static_cast<size_t>(m) * K * m * K, 0.0);
    Ss = Hs;
  }

  void upload(double* dst, const double* src, size_t count) {
    SBDX_HIP(hipMemcpyAsync(dst, src, count * 8, hipMemcpyHostToDevice, st));
    SBDX_HIP(hipStreamSynchronize(st));
  }
  const void* bptr(int i) const { return i == 0 ? static_cast<const void*>(X.get()) : B[i].get(); }
  const void* hbptr(int i) const { return i == 0 ? static_cast<const void*>(HX.get()) : HB[i].get(); }
  BlockPtrs ptrs(int nblocks) const {
    BlockPtrs p{};
    for (int i = 0; i < nblocks; ++i) {
      p.b[i] = bptr(i);
      p.hb[i] = hbptr(i);
    }
    return p;
  }

  // out (host, K x K row-major) = A^T B
  template <typename TA, typename TB>
  void gram(const TA* a, const TB* b, double* out) {
    const unsigned g = nblk(n);
    k_gram<TA, TB, K><<<g, kB, 0, st>>>(a, b, n, partial.get());
    SBDX_CHECK_LAUNCH();
    k_sum_partials<<<K * K, kB, 0, st>>>(partial.get(), static_cast<int>(g), K * K, small.get());
    SBDX_CHECK_LAUNCH();
    SBDX_HIP(hipMemcpyAsync(out, small.get(), K * K * 8, hipMemcpyDeviceToHost, st));
    SBDX_HIP(hipStreamSynchronize(st));
  }
  template <typename TB>
  void gram_with(int i, const TB* b, double* out) {  // B_i^T b
    if (i == 0) gram(X.get(), b, out);
    else gram(static_cast<const P*>(bptr(i)), b, out);
  }
  // out[i][a][c] (host, K x K per left block) = L_i^T R for every left block, in passes of up
  // to multigram_max_nl<K>() left blocks (power-of-two groups) over one read of R each;
  // lefts: (pointer, is FP64)
  void multigram(const std::vector<std::pair<const void*, bool>>& lefts, const void* R, double* out) {
    const unsigned g = nblk(n);
    size_t b0 = 0;
    while (b0 < lefts.size()) {
      int nl = multigram_max_nl<K>();
      while (static_cast<size_t>(nl) > lefts.size() - b0) nl /= 2;
      LeftSet ls{};
      for (int q = 0; q < nl; ++q) {
        ls.p[q] = lefts[b0 + q].first;
        ls.dbl[q] = lefts[b0 + q].second ? 1 : 0;
      }
      const P* r = static_cast<const P*>(R);
      switch (nl) {
        case 1: k_multigram<P, K, 1><<<g, kB, 0, st>>>(ls, r, n, partial.get()); break;
        case 2: k_multigram<P, K, (multigram_max_nl<K>() >= 2 ? 2 : 1)><<<g, kB, 0, st>>>(ls, r, n, partial.get()); break;
        case 4: k_multigram<P, K, (multigram_max_nl<K>() >= 4 ? 4 : 1)><<<g, kB, 0, st>>>(ls, r, n, partial.get()); break;
        case 8: k_multigram<P, K, (multigram_max_nl<K>() >= 8 ? 8 : 1)><<<g, kB, 0, st>>>(ls, r, n, partial.get()); break;
        case 16: k_multigram<P, K, (multigram_max_nl<K>() >= 16 ? 16 : 1)><<<g, kB, 0, st>>>(ls, r, n, partial.get()); break;
        default: k_multigram<P, K, (multigram_max_nl<K>() >= 32 ? 32 : 1)><<<g, kB, 0, st>>>(ls, r, n, partial.get()); break;
      }
      SBDX_CHECK_LAUNCH();
      k_sum_partials<<<nl * K * K, kB, 0, st>>>(partial.get(), static_cast<int>(g), nl * K * K, small.get());
      SBDX_CHECK_LAUNCH();
      SBDX_HIP(hipMemcpyAsync(out + b0 * K * K, small.get(), nl * K * K * 8, hipMemcpyDeviceToHost, st));
      SBDX_HIP(hipStreamSynchronize(st));
      b0 += nl;
    }
  }

  // the new row j >= 1 of the subspace matrices in two passes: [X, HX, B_1..B_j]^T B_j (overlaps
  // and the anchor's exact coupling (H X)^T B_j) and [B_1..B_j]^T HB_j
  void fill_row(int j) {
    std::vector<std::pair<const void*, bool>> ls{{X.get(), true}, {HX.get(), true}};
    for (int i = 1; i <= j; ++i) ls.push_back({bptr(i), prec == SBDX_FP64});
    std::vector<double> o1(ls.size() * K * K);
    multigram(ls, bptr(j), o1.data());
    std::vector<std::pair<const void*, bool>> lh;
    for (int i = 1; i <= j; ++i) lh.push_back({bptr(i), prec == SBDX_FP64});
    std::vector<double> o2(lh.size() * K * K);
    multigram(lh, hbptr(j), o2.data());
    const int ld = m * K;
    auto put = [&](int i, const double* h, const double* sv) {
      for (int a = 0; a < K; ++a)
        for (int b = 0; b < K; ++b) {
          Hs[(i * K + a) * ld + j * K + b] = Hs[(j * K + b) * ld + i * K + a] = h[a * K + b];
          Ss[(i * K + a) * ld + j * K + b] = Ss[(j * K + b) * ld + i * K + a] = sv[a * K + b];
        }
    };
    put(0, o1.data() + K * K, o1.data());
    for (int i = 1; i <= j; ++i) put(i, o2.data() + (i - 1) * K * K, o1.data() + (i + 1) * K * K);
  }

  // subspace matrices: blocks (i, j) and (j, i) from B_i^T HB_j and B_i^T B_j.  The anchor's
  // coupling to a correction block is taken as (H X)^T B_j with the exact FP64 sigma of X, not
  // as X^T (H B_j) with the correction's precision-P sigma: it is the first-order term that
  // decides the correction's coefficient, and its P rounding (|H| eps_P) would otherwise set
  // a floor under the residual once |r| falls to that size.
  void fill_block(int i, int j) {
    std::vector<double> h(K * K), s(K * K);
    if (j == 0) {
      gram_with(i, HX.get(), h.data());
      gram_with(i, X.get(), s.data());
    } else if (i == 0) {
      gram(HX.get(), static_cast<const P*>(bptr(j)), h.data());
      gram(X.get(), static_cast<const P*>(bptr(j)), s.data());
    } else {
      gram_with(i, static_cast<const P*>(hbptr(j)), h.data());
      gram_with(i, static_cast<const P*>(bptr(j)), s.data());
    }
    const int ld = m * K;
    for (int a = 0; a < K; ++a)
      for (int b = 0; b < K; ++b) {
        Hs[(i * K + a) * ld + j * K + b] = Hs[(j * K + b) * ld + i * K + a] = h[a * K + b];
        Ss[(i * K + a) * ld + j * K + b] = Ss[(j * K + b) * ld + i * K + a] = s[a * K + b];
      }
  }

  // lowest K Ritz pairs of the first nblocks blocks; U: (nblocks K) x K row-major
  bool rayleigh_ritz(int nblocks, std::vector<double>& U, std::vector<double>& E) {
    const int d = nblocks * K, ld = m * K;
    std::vector<double> S(d * d), Hm(d * d), sw(d);
    for (int a = 0; a < d; ++a)
      for (int b = 0; b < d; ++b) {
        S[a + d * b] = Ss[a * ld + b];
        Hm[a * d + b] = 0.5 * (Hs[a * ld + b] + Hs[b * ld + a]);
      }
    jacobi_syev_host(d, S.data(), d, sw.data());  // S = Q diag(sw) Q^T, Q in the columns of S
    const double smax = sw[d - 1];
    std::vector<int> keep;
    for (int a = 0; a < d; ++a)
      if (sw[a] > 1e-12 * smax) keep.push_back(a);
    const int r = static_cast<int>(keep.size());
    if (r < K) return false;
    std::vector<double> Z(d * r);  // Z = Q_keep diag(sw)^-1/2, d x r row-major
    for (int a = 0; a < d; ++a)
      for (int q = 0; q < r; ++q) Z[a * r + q] = S[a + d * keep[q]] / std::sqrt(sw[keep[q]]);
    std::vector<double> HZ(d * r, 0.0), Hr(r * r, 0.0), ew(r);
    for (int a = 0; a < d; ++a)
      for (int q = 0; q < r; ++q) {
        double s = 0.0;
        for (int b = 0; b < d; ++b) s += Hm[a * d + b] * Z[b * r + q];
        HZ[a * r + q] = s;
      }
    for (int p = 0; p < r; ++p)
      for (int q = p; q < r; ++q) {
        double s = 0.0;
        for (int a = 0; a < d; ++a) s += Z[a * r + p] * HZ[a * r + q];
        Hr[p + r * q] = Hr[q + r * p] = s;
      }
    jacobi_syev_host(r, Hr.data(), r, ew.data());
    U.assign(d * K, 0.0);
    E.assign(K, 0.0);
    for (int v = 0; v < K; ++v) {
      E[v] = ew[v];
      for (int a = 0; a < d; ++a) {
        double s = 0.0;
        for (int q = 0; q < r; ++q) s += Z[a * r + q] * Hr[q + r * v];
        U[a * K + v] = s;
      }
    }
    return true;
  }

  // R <- residuals (mode 0) or Ritz vectors (mode 1) of coefficients U; returns column norms
  void combine(int nblocks, const std::vector<double>& U, const std::vector<double>& E, int mode,
               std::vector<double>& nrm) {
    std::vector<double> stage(nblocks * K * K + K);
    for (int i = 0; i < nblocks; ++i)
      for (int a = 0; a < K; ++a)
        for (int v = 0; v < K; ++v) stage[i * K * K + a * K + v] = U[(i * K + a) * K + v];
    for (int v = 0; v < K; ++v) stage[nblocks * K * K + v] = E[v];
    upload(small.get(), stage.data(), stage.size());
    const unsigned g = nblk(n);
    k_combine<P, K><<<g, kB, 0, st>>>(X.get(), HX.get(), ptrs(nblocks), nblocks, small.get(),
                                      small.get() + nblocks * K * K, mode, R.get(), n, partial.get());
    SBDX_CHECK_LAUNCH();
    std::vector<double> h(static_cast<size_t>(g) * K);
    SBDX_HIP(hipMemcpyAsync(h.data(), partial.get(), h.size() * 8, hipMemcpyDeviceToHost, st));
    SBDX_HIP(hipStreamSynchronize(st));
    nrm.assign(K, 0.0);
    for (unsigned b = 0; b < g; ++b)
      for (int v = 0; v < K; ++v) nrm[v] += h[b * K + v];
    for (int v = 0; v < K; ++v) nrm[v] = std::sqrt(nrm[v]);
  }

  // Y <- Y M with M = Q diag(l)^-1/2 from Y^T Y = Q diag(l) Q^T (directions with l below
  // 1e-14 max(l) become zero columns)
  template <typename TY>
  void loewdin(TY* Y) {
    std::vector<double> G(K * K), Gc(K * K), w(K), M(K * K, 0.0);
    gram(Y, Y, G.data());
    for (int a = 0; a < K; ++a)
      for (int b = 0; b < K; ++b) Gc[a + K * b] = G[a * K + b];
    jacobi_syev_host(K, Gc.data(), K, w.data());
    const double lmax = std::max(w[K - 1], 1e-300);
    for (int a = 0; a < K; ++a)
      for (int v = 0; v < K; ++v) M[a * K + v] = w[v] > 1e-14 * lmax ? Gc[a + K * v] / std::sqrt(w[v]) : 0.0;
    upload(small.get(), M.data(), M.size());
    k_rotate<TY, K><<<nblk(n), kB, 0, st>>>(Y, small.get(), n);
    SBDX_CHECK_LAUNCH();
  }

  // preconditioned residuals in T: the learned preconditioner while it is in use (falling
  // back for good if it yields a non-finite block), else the diagonal one
  void precondition(const std::vector<double>& E, const std::vector<double>& nr, int nstep) {
    if (use_mlp && cycle >= o.precond_warmup) {
      std::vector<double> rms(K);
      for (int v = 0; v < K; ++v) rms[v] = nr[v] / std::sqrt(static_cast<double>(n));
      mlp_precondition(c, prec, K, R.get(), E.data(), rms.data(), T.get(), st);
      std::vector<double> G(K * K);
      gram(T.get(), T.get(), G.data());
      bool finite = true;
      for (double g : G) finite = finite && std::isfinite(g);
      if (finite) return;
      use_mlp = false;
      fallback_step = nstep;
    }
    upload(small.get(), E.data(), K);
    k_precond_diag<P, K><<<nblk(n), kB, 0, st>>>(R.get(), small.get(), c.hii.get(), T.get(), n);
    SBDX_CHECK_LAUNCH();
  }

  // new correction block in T from the residual block in R
  void expand(int nblocks, const std::vector<double>& E, const std::vector<double>& nr, int nstep) {
    precondition(E, nr, nstep);
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double> coef(nblocks * K * K);
      std::vector<std::pair<const void*, bool>> ls{{X.get(), true}};
      for (int i = 1; i < nblocks; ++i) ls.push_back({bptr(i), prec == SBDX_FP64});
      multigram(ls, T.get(), coef.data());
      upload(small.get(), coef.data(), coef.size());
      k_project_out<P, K><<<nblk(n), kB, 0, st>>>(T.get(), X.get(), ptrs(nblocks), nblocks, small.get(), n);
      SBDX_CHECK_LAUNCH();
      if (pass == 0) loewdin(T.get());  // rescale before the second pass
    }
    loewdin(T.get());
  }

  // the kept direction, orthogonalised against the new anchor (its sigma updated with the exact
  // FP64 sigma of the anchor) and orthonormalised, becomes block 1 of the new subspace
  void adopt_direction() {
    std::vector<double> cxp(K * K), G(K * K), Gc(K * K), w(K), M(K * K, 0.0);
    gram(X.get(), Pd.get(), cxp.data());
    upload(small.get(), cxp.data(), cxp.size());
    k_project_pair<P, K><<<nblk(n), kB, 0, st>>>(Pd.get(), HPd.get(), X.get(), HX.get(), small.get(), n);
    SBDX_CHECK_LAUNCH();
    gram(Pd.get(), Pd.get(), G.data());
    for (int a = 0; a < K; ++a)
      for (int b = 0; b < K; ++b) Gc[a + K * b] = G[a * K + b];
    jacobi_syev_host(K, Gc.data(), K, w.data());
    const double lmax = std::max(w[K - 1], 1e-300);
    for (int a = 0; a < K; ++a)
      for (int v = 0; v < K; ++v) M[a * K + v] = w[v] > 1e-14 * lmax ? Gc[a + K * v] / std::sqrt(w[v]) : 0.0;
    upload(small.get(), M.data(), M.size());
    k_rotate<P, K><<<nblk(n), kB, 0, st>>>(Pd.get(), small.get(), n);
    SBDX_CHECK_LAUNCH();
    k_rotate<P, K><<<nblk(n), kB, 0, st>>>(HPd.get(), small.get(), n);
    SBDX_CHECK_LAUNCH();
    std::swap(B[1], Pd);
    std::swap(HB[1], HPd);
    fill_row(1);
    have_dir = false;
  }

  void initial_block(const double* dW0) {
    SBDX_HIP(hipMemsetAsync(X.get(), 0, n * K * 8, st));
    if (K > 1) {
      // columns 1..K-1: unit vectors on the K-1 lowest H_ii other than determinant 0; from the
      // HF start (init 0), inside determinant 0's block of H where the basis splits into
      // blocks (else a unit vector in another block can lead root 1 to that block's lowest
      // eigenvalue, not the one the reference's Davidson finds from determinant 0), unless
      // that block holds fewer than K determinants
      const int kq = K - 1;
      const unsigned g = nblk(n);
      DevBuf<double> val(g * kq);
      DevBuf<int64_t> idx(g * kq);
      auto lowest = [&](const uint8_t* okA, const uint8_t* okB) {
        k_lowest<<<g, kB, 0, st>>>(c.hii.get(), n, kq, okA, okB, c.nb, val.get(), idx.get());
        SBDX_CHECK_LAUNCH();
        std::vector<double> hv(g * kq);
        std::vector<int64_t> hi(g * kq);
        SBDX_HIP(hipMemcpyAsync(hv.data(), val.get(), hv.size() * 8, hipMemcpyDeviceToHost, st));
        SBDX_HIP(hipMemcpyAsync(hi.data(), idx.get(), hi.size() * 8, hipMemcpyDeviceToHost, st));
        SBDX_HIP(hipStreamSynchronize(st));
        std::vector<std::pair<double, int64_t>> cand;
        for (size_t q = 0; q < hv.size(); ++q)
          if (hi[q] >= 0) cand.emplace_back(hv[q], hi[q]);
        std::sort(cand.begin(), cand.end());
        return cand;
      };
      if (o.init == 0) build_block0(c);
      std::vector<std::pair<double, int64_t>> cand;
      if (o.init == 0 && c.block0_state == 2) cand = lowest(c.block0_a.get(), c.block0_b.get());
      if (static_cast<int>(cand.size()) < kq) cand = lowest(nullptr, nullptr);
      if (static_cast<int>(cand.size()) < kq) throw Error(SBDX_ERR_ARG, "fewer determinants than the block width");
      std::vector<int64_t> pos(kq);
      for (int v = 1; v < K; ++v) pos[v - 1] = cand[v - 1].second * K + v;
      DevBuf<int64_t> dp(kq);
      SBDX_HIP(hipMemcpyAsync(dp.get(), pos.data(), kq * 8, hipMemcpyHostToDevice, st));
      k_set_entries<<<1, 64, 0, st>>>(X.get(), dp.get(), kq);
      SBDX_CHECK_LAUNCH();
      SBDX_HIP(hipStreamSynchronize(st));
    }
    k_copy_col<<<nblk(n), kB, 0, st>>>(X.get(), K, 0, dW0, 1, 0, n);
    SBDX_CHECK_LAUNCH();
    loewdin(X.get());
  }

  void run(double* dW, sbdx_davidson_result& res, double* trace, double* resid, int trace_cap) {
    const double t0 = now();
    res.seconds_to_tol = -1.0;
    res.steps_to_tol = -1;
    initial_block(dW);
    int nstep = 0;
    bool conv = false, stop = false;
    double prev_anchor = 0.0;
    std::vector<double> U, E(K, 0.0), nr(K, 0.0), ny;
    for (int it = 0; it < o.max_iteration && !stop; ++it) {
      cycle = it;
      // anchor: exact FP64 sigma of X
      sigma_block(c, SBDX_FP64, K, X.get(), HX.get(), st);
      std::fill(Hs.begin(), Hs.end(), 0.0);
      std::fill(Ss.begin(), Ss.end(), 0.0);
      fill_block(0, 0);
      int nblocks = 1;
      for (;;) {
        if (!rayleigh_ritz(nblocks, U, E)) throw Error(SBDX_ERR_DEVICE, "Davidson subspace lost its rank");
        const bool exact = nblocks == 1;
        combine(nblocks, U, E, 0, nr);
        if (nstep < trace_cap) {
          if (trace) trace[nstep] = E[0];
          if (resid) resid[nstep] = nr[0];
        }
        ++nstep;
        double worst = 0.0;
        for (int v = 0; v < nroots; ++v) worst = std::max(worst, nr[v]);
        if (o.verbose) {
          std::printf(" Davidson iteration %d.%d (tol=%.16g)%s:", it, nblocks - 1, nr[0], exact ? "" : " (mixed)");
          for (int v = 0; v < std::min(K, std::max(4, nroots)); ++v) std::printf(" %.16g", E[v]);
          std::printf("\n");
          if (exact && nroots > 1) {  // every requested root's residual on the exact steps
            std::printf("   residuals:");
            for (int v = 0; v < nroots; ++v) std::printf(" %.3e", nr[v]);
            std::printf("\n");
          }
          std::fflush(stdout);
        }
        const bool near_ref = o.energy_ref != 0.0 && std::fabs(E[0] - o.energy_ref) < o.energy_tol;
        if (exact) {
          // stall test of the learned preconditioner on the exact residual of every restart:
          // a cycle that does not halve it switches to the diagonal preconditioner for good
          if (use_mlp && it > o.precond_warmup && worst > 0.5 * prev_anchor) {
            use_mlp = false;
            fallback_step = nstep;
          }
          prev_anchor = worst;
          if (near_ref && res.seconds_to_tol < 0) {
            res.seconds_to_tol = now() - t0;
            res.steps_to_tol = nstep;
            if (o.stop_at_tol) stop = true;
          }
          if (worst < o.eps) conv = stop = true;
        }
        if (stop || now() - t0 > o.max_time) {
          stop = true;
          break;
        }
        // a mixed-precision estimate that meets a criterion is confirmed by a restart
        if (!exact && (worst < o.eps || (near_ref && res.seconds_to_tol < 0))) break;
        if (nblocks == m) break;
        if (exact && have_dir && m > 2) {
          adopt_direction();
          nblocks = 2;
          continue;
        }
        expand(nblocks, E, nr, nstep);
        std::swap(B[nblocks], T);
        sigma_block(c, prec, K, B[nblocks].get(), HB[nblocks].get(), st);
        fill_row(nblocks);
        ++nblocks;
      }
      // thick restart (or the final vectors): X <- normalised Ritz vectors of the subspace; the
      // correction part of the Ritz block is kept as the next cycle's first search direction
      combine(nblocks, U, E, 1, ny);
      if (nblocks > 1 && !stop) {
        std::vector<double> stage(nblocks * K * K);
        for (int i = 0; i < nblocks; ++i)
          for (int a = 0; a < K; ++a)
            for (int v = 0; v < K; ++v) stage[i * K * K + a * K + v] = U[(i * K + a) * K + v];
        upload(small.get(), stage.data(), stage.size());
        k_corr<P, K><<<nblk(n), kB, 0, st>>>(ptrs(nblocks), nblocks, small.get(), Pd.get(), HPd.get(), n);
        SBDX_CHECK_LAUNCH();
        have_dir = true;
      }
      std::vector<double> M(K * K, 0.0);
      for (int v = 0; v < K; ++v) M[v * K + v] = ny[v] > 0 ? 1.0 / ny[v] : 0.0;
      upload(small.get(), M.data(), M.size());
      k_rotate<double, K><<<nblk(n), kB, 0, st>>>(R.get(), small.get(), n);
      SBDX_CHECK_LAUNCH();
      std::swap(X, R);
    }
    // return the Ritz vectors vector-major in dW
    for (int v = 0; v < K; ++v) {
      k_copy_col<<<nblk(n), kB, 0, st>>>(dW + v * n, 1, 0, X.get(), K, v, n);
      SBDX_CHECK_LAUNCH();
    }
    SBDX_HIP(hipStreamSynchronize(st));
    res.seconds = now() - t0;
    res.steps = nstep;
    double worst = 0.0;
    for (int v = 0; v < nroots; ++v) worst = std::max(worst, nr[v]);
    res.residual = worst;
    res.converged = conv ? 1 : 0;
    res.energy = E[0];
    res.precond_fallback = fallback_step;
  }
};

template <typename P>
void run_width(sbdx_ctx& c, const sbdx_davidson_opts& o, int k, double* dW, sbdx_davidson_result& res,
               double* trace, double* resid, int cap) {
  switch (k) {
    case 1: { BlockSolver<P, 1> s(c, o); s.run(dW, res, trace, resid, cap); break; }
    case 2: { BlockSolver<P, 2> s(c, o); s.run(dW, res, trace, resid, cap); break; }
    case 4: { BlockSolver<P, 4> s(c, o); s.run(dW, res, trace, resid, cap); break; }
    case 8: { BlockSolver<P, 8> s(c, o); s.run(dW, res, trace, resid, cap); break; }
    default: throw Error(SBDX_ERR_ARG, "block width must be 1, 2, 4 or 8");
  }
}

}  // namespace

// vectors per block: the next power of two >= max(nroots, block)
int davidson_block_width(const sbdx_davidson_opts& o) {
  const int k = std::max(std::max(1, o.nroots), o.block);
  int w = 1;
  while (w < k) w <<= 1;
  return w;
}

// dW: block-width * ndet doubles (blocked basis, vector-major); vector 0 holds the initial guess
// on entry, the Ritz vectors (lowest first) on return.
void davidson_block(sbdx_ctx& c, const sbdx_davidson_opts& o, double* dW, sbdx_davidson_result& res,
                    double* trace, double* resid, int trace_cap) {
  const int k = davidson_block_width(o);
  if (o.precision == SBDX_FP32) {
    ensure_fp32(c);
    run_width<float>(c, o, k, dW, res, trace, resid, trace_cap);
  } else {
    run_width<double>(c, o, k, dW, res, trace, resid, trace_cap);
  }
}

}  // namespace sbdx
