// davidson.hip -- lowest-root Davidson on the GPU (chemistry/tpb/davidson.h:661-923).
//
// Per step, as the reference: sigma of the newest basis vector; the new column of the
// subspace matrix <C_j|HC_i>; the (ib+1)-dimensional eigenproblem on the host; the Ritz vector
// W and residual R = E0 W - sum_k u_k HC_k; both normalised (the residual norm before
// normalisation is the convergence test); the correction R/(E0 - H_ii) (with the reference's
// 1e-12 regulariser); one modified Gram-Schmidt pass; normalisation; a wall-clock deadline
// from the mean step time.  After num_block steps the basis restarts from W.
//
// All basis vectors stay resident in HBM in the blocked basis; the streaming passes are fused
// so each step reads every vector once where the algorithm allows (multi-dot reads HC_ib once
// against all C_j; the Ritz kernel forms W, R and both squared norms in one pass).
#include "common.h"
#include "sbdx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace sbdx {

constexpr int kMaxVec = 64;

struct PtrSet {
  const double* p[kMaxVec];
};
struct CoefSet {
  double c[kMaxVec];
};

namespace {

constexpr int kBlock = 256;

inline unsigned reduce_blocks(int64_t n) {
  int64_t b = cdiv(n, kBlock * 4);
  return static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>(b, 1), 2048));
}

__device__ inline double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) r += sh[k];
  __syncthreads();
  return r;
}

// partial[b * m + k] = sum over this block's share of X_k . y
__global__ __launch_bounds__(kBlock) void k_multidot(PtrSet X, int m, const double* __restrict__ y,
                                                    int64_t n, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  for (int k = 0; k < m; ++k) {
    double s = 0.0;
    const double* __restrict__ xk = X.p[k];
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock)
      s += xk[i] * y[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x * m + k] = s;
  }
}

// out[k] = sum_b partial[b * m + k]; one block per k
__global__ __launch_bounds__(kBlock) void k_finish(const double* __restrict__ partial, int nblk, int m,
                                                  double* __restrict__ out) {
  __shared__ double sh[kBlock / 64];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kBlock) s += partial[b * m + k];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[k] = s;
}

// W = sum_k u_k C_k ; R = -sum_k u_k HC_k + e0 W ; partial sums of |W|^2 and |R|^2
__global__ __launch_bounds__(kBlock) void k_ritz(PtrSet C, PtrSet HC, CoefSet u, int m, double e0,
                                                double* __restrict__ W, double* __restrict__ R,
                                                int64_t n, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  double sw = 0.0, sr = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kBlock) {
    double w = 0.0, r = 0.0;
    for (int k = 0; k < m; ++k) {
      w += C.p[k][i] * u.c[k];
      r += HC.p[k][i] * (-u.c[k]);
    }
    r += e0 * w;
    W[i] = w;
    R[i] = r;
    sw += w * w;
    sr += r * r;
  }
  sw = block_sum(sw, sh);
  sr = block_sum(sr, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x * 2] = sw;
    partial[blockIdx.x * 2 + 1] = sr;
  }
}

// W *= fw ; c = (R * fr) / (e0 - d) with the reference's regulariser (davidson.h:855-862);
// R is normalised in place too (it is not read again, but keeps the reference's state)
__global__ __launch_bounds__(kBlock) void k_precond(double* __restrict__ W, double fw,
                                                   double* __restrict__ R, double fr, double e0,
                                                   const double* __restrict__ d,
                                                   double* __restrict__ c, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kBlock) {
    W[i] *= fw;
    const double r = R[i] * fr;
    R[i] = r;
    if (c) {
      const double den = e0 - d[i];
      c[i] = fabs(den) > 1.0e-12 ? r / den : r / (den - 1.0e-12);
    }
  }
}

// y += -(*olap) x  (olap on the device: no host round trip inside Gram-Schmidt)
__global__ __launch_bounds__(kBlock) void k_axpy_dev(double* __restrict__ y, const double* __restrict__ x,
                                                    const double* __restrict__ olap, int64_t n) {
  const double a = -olap[0];
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kBlock)
    y[i] += x[i] * a;
}

// x *= 1/sqrt(*nrm2)
__global__ __launch_bounds__(kBlock) void k_scale_dev(double* __restrict__ x, const double* __restrict__ nrm2,
                                                     int64_t n) {
  const double f = 1.0 / sqrt(nrm2[0]);
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kBlock)
    x[i] *= f;
}

__global__ void k_set_unit(double* __restrict__ x, int64_t n, double v) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kBlock)
    x[i] = (i == 0) ? v : 0.0;
}

// Cyclic Jacobi for the small symmetric subspace problem (the reference calls LAPACK dsyev
// 'V','U' here, framework/hp_numeric.h:500-516).  a: n x n column-major with leading
// dimension lda, upper triangle read; out: eigenvectors in columns, w ascending.
void jacobi_syev(int n, double* a, int lda, double* w) {
  std::vector<double> m(static_cast<size_t>(n) * n), v(static_cast<size_t>(n) * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) m[i + n * j] = a[std::min(i, j) + lda * std::max(i, j)];
  for (int i = 0; i < n; ++i) v[i + n * i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += m[i + n * j] * m[i + n * j];
        if (i != j) off += m[i + n * j] * m[i + n * j];
      }
    if (off <= 1e-34 * (tot > 0 ? tot : 1.0)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = m[p + n * q];
        if (apq == 0.0) continue;
        const double theta = (m[q + n * q] - m[p + n * p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {
          const double kp = m[k + n * p], kq = m[k + n * q];
          m[k + n * p] = cs * kp - sn * kq;
          m[k + n * q] = sn * kp + cs * kq;
        }
        for (int k = 0; k < n; ++k) {
          const double pk = m[p + n * k], qk = m[q + n * k];
          m[p + n * k] = cs * pk - sn * qk;
          m[q + n * k] = sn * pk + cs * qk;
        }
        for (int k = 0; k < n; ++k) {
          const double kp = v[k + n * p], kq = v[k + n * q];
          v[k + n * p] = cs * kp - sn * kq;
          v[k + n * q] = sn * kp + cs * kq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return m[x * (n + 1)] < m[y * (n + 1)]; });
  for (int j = 0; j < n; ++j) {
    const int s = ord[j];
    w[j] = m[s * (n + 1)];
    int imax = 0;
    for (int i = 1; i < n; ++i)
      if (std::fabs(v[i + n * s]) > std::fabs(v[imax + n * s])) imax = i;
    const double sg = v[imax + n * s] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) a[i + lda * j] = sg * v[i + n * s];
  }
}

struct Reducer {
  sbdx_ctx& c;
  hipStream_t st;
  explicit Reducer(sbdx_ctx& cc, hipStream_t s) : c(cc), st(s) {
    c.ws.partial.ensure(2048 * kMaxVec);
    c.ws.scal.ensure(kMaxVec * 2);
  }
  // out_dev[k] = X_k . y for k < m (device result)
  void multidot_dev(const double* const* X, int m, const double* y, int64_t n, double* out_dev) {
    PtrSet ps{};
    for (int k = 0; k < m; ++k) ps.p[k] = X[k];
    const unsigned nb = reduce_blocks(n);
    k_multidot<<<nb, kBlock, 0, st>>>(ps, m, y, n, c.ws.partial.get());
    SBDX_CHECK_LAUNCH();
    k_finish<<<m, kBlock, 0, st>>>(c.ws.partial.get(), nb, m, out_dev);
    SBDX_CHECK_LAUNCH();
  }
  void multidot(const double* const* X, int m, const double* y, int64_t n, double* out_host) {
    multidot_dev(X, m, y, n, c.ws.scal.get());
    SBDX_HIP(hipMemcpyAsync(out_host, c.ws.scal.get(), m * sizeof(double), hipMemcpyDeviceToHost, st));
    SBDX_HIP(hipStreamSynchronize(st));
  }
};

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// Davidson on device vector dW (blocked basis, FP64 arithmetic, lowest root).
void davidson_fp64(sbdx_ctx& c, const sbdx_davidson_opts& o, double* dW, sbdx_davidson_result& res,
                   double* trace, double* resid, int trace_cap) {
  const int64_t n = c.ndet();
  const int nb = std::max(1, std::min(o.num_block, kMaxVec));
  hipStream_t st = c.stream;
  std::vector<DevBuf<double>> C(nb), HC(nb);
  for (int k = 0; k < nb; ++k) {
    C[k].alloc(n);
    HC[k].alloc(n);
  }
  DevBuf<double> R(n);
  Reducer red(c, st);
  const unsigned sb = reduce_blocks(n);
  std::vector<double> H(static_cast<size_t>(nb) * nb, 0.0), U(static_cast<size_t>(nb) * nb, 0.0), E(nb);
  std::vector<double> onestep(static_cast<size_t>(nb) * std::max(1, o.max_iteration), 0.0);
  std::vector<double> hpart(2 * sb);
  int nstep = 0;
  bool cont = true, conv = false;
  double norm_R = 0.0, e_last = 0.0;
  res.seconds_to_tol = -1.0;
  res.steps_to_tol = -1;
  const double t0 = now();

  for (int it = 0; it < o.max_iteration && cont; ++it) {
    SBDX_HIP(hipMemcpyAsync(C[0].get(), dW, n * 8, hipMemcpyDeviceToDevice, st));
    for (int ib = 0; ib < nb; ++ib) {
      const double ts = now();
      sigma_blocked(c, SBDX_FP64, 1, C[ib].get(), HC[ib].get(), st);
      std::vector<const double*> cp(ib + 1);
      for (int k = 0; k <= ib; ++k) cp[k] = C[k].get();
      std::vector<double> col(ib + 1);
      red.multidot(cp.data(), ib + 1, HC[ib].get(), n, col.data());
      for (int jb = 0; jb <= ib; ++jb) {
        H[jb + nb * ib] = col[jb];
        H[ib + nb * jb] = col[jb];
      }
      for (int jb = 0; jb <= ib; ++jb)
        for (int kb = 0; kb <= ib; ++kb) U[jb + nb * kb] = H[jb + nb * kb];
      jacobi_syev(ib + 1, U.data(), nb, E.data());

      PtrSet pc{}, ph{};
      CoefSet u{};
      for (int k = 0; k <= ib; ++k) {
        pc.p[k] = C[k].get();
        ph.p[k] = HC[k].get();
        u.c[k] = U[k];
      }
      k_ritz<<<sb, kBlock, 0, st>>>(pc, ph, u, ib + 1, E[0], dW, R.get(), n, c.ws.partial.get());
      SBDX_CHECK_LAUNCH();
      SBDX_HIP(hipMemcpyAsync(hpart.data(), c.ws.partial.get(), 2 * sb * 8, hipMemcpyDeviceToHost, st));
      SBDX_HIP(hipStreamSynchronize(st));
      double nw2 = 0.0, nr2 = 0.0;
      for (unsigned b = 0; b < sb; ++b) {
        nw2 += hpart[2 * b];
        nr2 += hpart[2 * b + 1];
      }
      const double norm_W = std::sqrt(nw2);
      norm_R = std::sqrt(nr2);
      e_last = E[0];
      if (nstep < trace_cap) {
        if (trace) trace[nstep] = E[0];
        if (resid) resid[nstep] = norm_R;
      }
      ++nstep;
      if (o.verbose) {
        std::printf(" Davidson iteration %d.%d (tol=%.16g):", it, ib, norm_R);
        for (int p = 0; p < std::min(ib + 1, 4); ++p) std::printf(" %.16g", E[p]);
        std::printf("\n");
        std::fflush(stdout);
      }
      if (o.energy_ref != 0.0 && res.seconds_to_tol < 0 && std::fabs(E[0] - o.energy_ref) < o.energy_tol) {
        res.seconds_to_tol = now() - t0;
        res.steps_to_tol = nstep;
      }
      const bool at_tol = o.stop_at_tol && res.steps_to_tol == nstep;
      const bool last = norm_R < o.eps || at_tol;
      double* cnew = (!last && ib < nb - 1) ? C[ib + 1].get() : nullptr;
      k_precond<<<sb, kBlock, 0, st>>>(dW, 1.0 / norm_W, R.get(), 1.0 / norm_R, E[0], c.hii.get(), cnew, n);
      SBDX_CHECK_LAUNCH();
      if (last) {
        conv = norm_R < o.eps;
        cont = false;
        break;
      }
      if (cnew) {
        double* olap = c.ws.scal.get() + kMaxVec;
        for (int kb = 0; kb <= ib; ++kb) {
          const double* xk = C[kb].get();
          red.multidot_dev(&xk, 1, cnew, n, olap);
          k_axpy_dev<<<sb, kBlock, 0, st>>>(cnew, xk, olap, n);
          SBDX_CHECK_LAUNCH();
        }
        red.multidot_dev(&cnew, 1, cnew, n, olap);
        k_scale_dev<<<sb, kBlock, 0, st>>>(cnew, olap, n);
        SBDX_CHECK_LAUNCH();
      }
      SBDX_HIP(hipStreamSynchronize(st));
      const double te = now();
      onestep[static_cast<size_t>(it) * nb + ib] = te - ts;
      double ave = 0.0;
      for (int ks = 0; ks <= it * nb + ib; ++ks) ave += onestep[ks];
      ave /= (it * nb + ib + 1);
      if ((te - t0) + ave > o.max_time) {
        cont = false;
        break;
      }
    }
  }
  SBDX_HIP(hipStreamSynchronize(st));
  res.seconds = now() - t0;
  res.steps = nstep;
  res.residual = norm_R;
  res.converged = conv ? 1 : 0;
  // reported energy: <W|H|W> from one more sigma (chemistry/tpb/sbdiag.h:239-262)
  sigma_blocked(c, SBDX_FP64, 1, dW, R.get(), st);
  double e = 0.0;
  const double* wp = dW;
  red.multidot(&wp, 1, R.get(), n, &e);
  res.energy = e;
  (void)e_last;
}

void jacobi_syev_host(int n, double* a, int lda, double* w) { jacobi_syev(n, a, lda, w); }

void init_hf(sbdx_ctx& c, double* dW, hipStream_t st) {
  // HF vector W[0] = 1 (davidson.h:101-140, init 0) expressed in the blocked basis
  Str<2> a0{}, b0{};  // the first strings, high word zero when W = 1
  SBDX_HIP(hipMemcpy(a0.w, c.A.str.get(), 8 * c.W, hipMemcpyDeviceToHost));
  SBDX_HIP(hipMemcpy(b0.w, c.B.str.get(), 8 * c.W, hipMemcpyDeviceToHost));
  const double v = det_sign(a0, b0) < 0 ? -1.0 : 1.0;
  k_set_unit<<<reduce_blocks(c.ndet()), kBlock, 0, st>>>(dW, c.ndet(), v);
  SBDX_CHECK_LAUNCH();
}

}  // namespace sbdx
