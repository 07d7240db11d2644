// batch.hip -- many independent small selected-basis problems solved together.
//
// A batch advances the thick-restart Davidson of every instance (one vector per block, FP64 or
// mixed precision, diagonal preconditioner; the algorithm of solver.hip) in lockstep.  Each
// lockstep step is one hipGraph launch that holds, for every active instance, the residual of
// the current Ritz vector, the preconditioned correction, two classical Gram-Schmidt passes and
// its normalisation (all with device-resident coefficients), the new block's sigma build, and
// the new row of the subspace matrices.  The host then reads every instance's small results in
// one copy, solves the subspace eigenproblems, and writes the coefficients back in one copy.
// Graphs are instantiated once per (block index, active set) and replayed: for a step the
// launch count is one whatever the number of instances, and instances of very different sizes
// share the GPU inside one launch: each problem's sigma build is its own branch of the graph
// (captured on side streams), so small problems run beside large ones.
//
// Batched kernels take grid (workgroups per instance, instances); instance i grid-strides over
// its own n_i, and an instance that has converged is skipped by its workgroups at once (a device
// flag) and left out of the sigma builds of the graphs captured after it.
#include "common.h"
#include "sbdx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <type_traits>
#include <vector>

namespace sbdx {

void jacobi_syev_host(int n, double* a, int lda, double* w);  // davidson.hip
void init_hf(sbdx_ctx& c, double* dW, hipStream_t st);       // davidson.hip

namespace {

constexpr int kB = 256;
constexpr int kMaxBlocks = 24;
constexpr int kSlots = 2 * kMaxBlocks + 2;  // per-instance reduction slots

struct Inst {
  int64_t n;
  const double* hii;
  double* X;         // anchor (FP64)
  double* HX;        // its sigma (FP64)
  double* R;         // residual / next anchor (FP64)
  void* B[kMaxBlocks];   // corrections (precision P), B[0] unused
  void* HB[kMaxBlocks];
  const double* coef;    // u[0..nb) then E at [kMaxBlocks]
  double* res;           // reduction results, kSlots
  double* part;          // workgroup partials [G][kSlots]
  const int* active;
};

__device__ inline double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// reduce `m` per-thread values into part[blockIdx.x][slot0 + q]
template <int M>
__device__ inline void block_store(double (&acc)[M], int m, double* part, int slot0) {
  __shared__ double sh[kB / 64][M];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = 0; q < m; ++q) {
    const double s = wsum(acc[q]);
    if (lane == 0) sh[w][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < m) {
    double s = 0.0;
    for (int ww = 0; ww < kB / 64; ++ww) s += sh[ww][threadIdx.x];
    part[static_cast<int64_t>(blockIdx.x) * kSlots + slot0 + threadIdx.x] = s;
  }
  __syncthreads();
}

// the instance's block pointers, staged in LDS once per workgroup (the loops below index them
// per element; from global memory they would be reloaded every iteration)
__device__ inline void stage_ptrs(const Inst& I, int nb, const void** sB, const void** sHB) {
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    sB[k] = I.B[k];
    sHB[k] = I.HB[k];
  }
  __syncthreads();
}

// res[slot0 + q] = sum over workgroups of part[g][slot0 + q], one thread per q per instance
__global__ void k_breduce(const Inst* __restrict__ in, int G, int slot0, int m) {
  const Inst& I = in[blockIdx.y];
  if (!*I.active) return;
  const int q = threadIdx.x;
  if (q >= m) return;
  double s = 0.0;
  for (int g = 0; g < G; ++g) s += I.part[static_cast<int64_t>(g) * kSlots + slot0 + q];
  I.res[slot0 + q] = s;
}

// mode 0: R = sum_k u_k HB_k - E sum_k u_k B_k (k = 0 is X, HX), |R|^2 -> slot 2 kMaxBlocks
// mode 1: R = sum_k u_k B_k (the Ritz vector), |R|^2 -> same slot
template <typename P>
__global__ __launch_bounds__(kB) void k_bcombine(const Inst* __restrict__ in, int nb, int mode) {
  const Inst& I = in[blockIdx.y];
  if (!*I.active) return;
  __shared__ const void* sB[kMaxBlocks];
  __shared__ const void* sHB[kMaxBlocks];
  __shared__ double u[kMaxBlocks];
  stage_ptrs(I, nb, sB, sHB);
  for (int k = threadIdx.x; k < nb; k += kB) u[k] = I.coef[k];
  __syncthreads();
  const double e = I.coef[kMaxBlocks];
  double acc[1] = {0.0};
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < I.n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    double y = u[0] * I.X[i], hy = mode == 0 ? u[0] * I.HX[i] : 0.0;
    for (int k = 1; k < nb; ++k) {
      y = fma(u[k], static_cast<double>(static_cast<const P*>(sB[k])[i]), y);
      if (mode == 0) hy = fma(u[k], static_cast<double>(static_cast<const P*>(sHB[k])[i]), hy);
    }
    const double o = mode == 0 ? hy - e * y : y;
    I.R[i] = o;
    acc[0] = fma(o, o, acc[0]);
  }
  block_store<1>(acc, 1, I.part, 2 * kMaxBlocks);
}

// B[j] = R / (E - H_ii) (the reference's regulariser, davidson.h:855-862)
template <typename P>
__global__ __launch_bounds__(kB) void k_bprecond(const Inst* __restrict__ in, int j) {
  const Inst& I = in[blockIdx.y];
  if (!*I.active) return;
  const double e = I.coef[kMaxBlocks];
  P* T = static_cast<P*>(I.B[j]);
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < I.n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    const double den = e - I.hii[i];
    T[i] = static_cast<P>(I.R[i] / (fabs(den) > 1.0e-12 ? den : den - 1.0e-12));
  }
}

// slots [0, j): B_k . B_j (k = 0: X); slot j: B_j . B_j
template <typename P>
__global__ __launch_bounds__(kB) void k_bdots_t(const Inst* __restrict__ in, int j) {
  const Inst& I = in[blockIdx.y];
  if (!*I.active) return;
  __shared__ const void* sB[kMaxBlocks];
  __shared__ const void* sHB[kMaxBlocks];
  stage_ptrs(I, j + 1, sB, sHB);
  const P* T = static_cast<const P*>(sB[j]);
  double acc[kMaxBlocks + 1];
  for (int k = 0; k <= j; ++k) acc[k] = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < I.n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    const double t = static_cast<double>(T[i]);
    acc[0] = fma(I.X[i], t, acc[0]);
    for (int k = 1; k < j; ++k) acc[k] = fma(static_cast<double>(static_cast<const P*>(sB[k])[i]), t, acc[k]);
    acc[j] = fma(t, t, acc[j]);
  }
  block_store<kMaxBlocks + 1>(acc, j + 1, I.part, 0);
}

// B_j <- (B_j - sum_{k<j} c_k B_k) / sqrt(|B_j|^2 - sum c_k^2), coefficients from res
template <typename P>
__global__ __launch_bounds__(kB) void k_bproject(const Inst* __restrict__ in, int j) {
  const Inst& I = in[blockIdx.y];
  if (!*I.active) return;
  __shared__ double c[kMaxBlocks];
  __shared__ double scale;
  if (threadIdx.x == 0) {
    double nn = I.res[j];
    for (int k = 0; k < j; ++k) {
      c[k] = I.res[k];
      nn -= c[k] * c[k];
    }
    scale = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
  }
  __syncthreads();
  __shared__ const void* sB[kMaxBlocks];
  __shared__ const void* sHB[kMaxBlocks];
  stage_ptrs(I, j + 1, sB, sHB);
  P* T = static_cast<P*>(const_cast<void*>(sB[j]));
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < I.n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    double t = static_cast<double>(T[i]) - c[0] * I.X[i];
    for (int k = 1; k < j; ++k) t = fma(-c[k], static_cast<double>(static_cast<const P*>(sB[k])[i]), t);
    T[i] = static_cast<P>(t * scale);
  }
}

// new row of the subspace matrices: slot k: H(k, j), slot kMaxBlocks + k: S(k, j), k <= j;
// the anchor's coupling is (H X)^T B_j with the exact FP64 sigma of X
template <typename P>
__global__ __launch_bounds__(kB) void k_brow(const Inst* __restrict__ in, int j) {
  const Inst& I = in[blockIdx.y];
  if (!*I.active) return;
  __shared__ const void* sB[kMaxBlocks];
  __shared__ const void* sHB[kMaxBlocks];
  stage_ptrs(I, j + 1, sB, sHB);
  const P* Bj = static_cast<const P*>(sB[j]);
  const P* Hj = static_cast<const P*>(sHB[j]);
  double h[kMaxBlocks], s[kMaxBlocks];
  for (int k = 0; k <= j; ++k) h[k] = s[k] = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < I.n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    const double b = static_cast<double>(Bj[i]), hb = static_cast<double>(Hj[i]);
    h[0] = fma(I.HX[i], b, h[0]);
    s[0] = fma(I.X[i], b, s[0]);
    for (int k = 1; k <= j; ++k) {
      const double bk = static_cast<double>(static_cast<const P*>(sB[k])[i]);
      h[k] = fma(bk, hb, h[k]);
      s[k] = fma(bk, b, s[k]);
    }
  }
  block_store<kMaxBlocks>(h, j + 1, I.part, 0);
  block_store<kMaxBlocks>(s, j + 1, I.part, kMaxBlocks);
}

// anchor: slot 0 = X . HX, slot kMaxBlocks = X . X
__global__ __launch_bounds__(kB) void k_banchor(const Inst* __restrict__ in) {
  const Inst& I = in[blockIdx.y];
  if (!*I.active) return;
  double h[1] = {0.0}, s[1] = {0.0};
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < I.n;
       i += static_cast<int64_t>(gridDim.x) * kB) {
    h[0] = fma(I.X[i], I.HX[i], h[0]);
    s[0] = fma(I.X[i], I.X[i], s[0]);
  }
  block_store<1>(h, 1, I.part, 0);
  block_store<1>(s, 1, I.part, kMaxBlocks);
}

// X <- R * coef[0] (normalised next anchor)
__global__ __launch_bounds__(kB) void k_bnext(const Inst* __restrict__ in) {
  const Inst& I = in[blockIdx.y];
  if (!*I.active) return;
  const double s = I.coef[0];
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < I.n;
       i += static_cast<int64_t>(gridDim.x) * kB)
    I.X[i] = I.R[i] * s;
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// lowest eigenpair of H u = e S u (d x d, row-major, S positive definite up to rounding) by
// canonical orthogonalization; false if S has lost its rank
bool ritz1(int d, const double* Hs, const double* Ss, double* u, double* e) {
  std::vector<double> S(d * d), sw(d);
  for (int a = 0; a < d; ++a)
    for (int b = 0; b < d; ++b) S[a + d * b] = Ss[a * d + b];
  jacobi_syev_host(d, S.data(), d, sw.data());
  const double smax = sw[d - 1];
  std::vector<int> keep;
  for (int a = 0; a < d; ++a)
    if (sw[a] > 1e-12 * smax) keep.push_back(a);
  const int r = static_cast<int>(keep.size());
  if (r < 1) return false;
  std::vector<double> Z(d * r), Hr(r * r), ew(r);
  for (int a = 0; a < d; ++a)
    for (int q = 0; q < r; ++q) Z[a * r + q] = S[a + d * keep[q]] / std::sqrt(sw[keep[q]]);
  for (int p = 0; p < r; ++p)
    for (int q = p; q < r; ++q) {
      double s = 0.0;
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) s += Z[a * r + p] * 0.5 * (Hs[a * d + b] + Hs[b * d + a]) * Z[b * r + q];
      Hr[p + r * q] = Hr[q + r * p] = s;
    }
  jacobi_syev_host(r, Hr.data(), r, ew.data());
  *e = ew[0];
  for (int a = 0; a < d; ++a) {
    double s = 0.0;
    for (int q = 0; q < r; ++q) s += Z[a * r + q] * Hr[q];
    u[a] = s;
  }
  return true;
}

}  // namespace

}  // namespace sbdx

namespace sbdx {
namespace {

template <typename P>
struct BatchSolver {
  sbdx_batch& b;
  const sbdx_davidson_opts& o;
  const int cnt;
  const int m;
  hipStream_t st;
  unsigned G = 1;                       // workgroups per instance
  static constexpr int prec = std::is_same<P, float>::value ? SBDX_FP32 : SBDX_FP64;
  std::vector<DevBuf<double>> X, HX, R, coef, res, part;
  std::vector<std::vector<DevBuf<P>>> B, HB;
  DevBuf<Inst> dinst;
  DevBuf<int> dactive;
  std::vector<int> active;
  std::vector<double> h_coef, h_res;    // pinned-size host mirrors, cnt x slots
  std::vector<std::vector<double>> Hs, Ss;  // per instance m x m
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (phase, epoch) -> executable
  int epoch = 0;

  BatchSolver(sbdx_batch& bb, const sbdx_davidson_opts& oo)
      : b(bb), o(oo), cnt(static_cast<int>(bb.ctx.size())), m(std::max(2, std::min(oo.num_block, kMaxBlocks))),
        st(bb.stream) {
    int64_t nmax = 1;
    X.resize(cnt); HX.resize(cnt); R.resize(cnt); coef.resize(cnt); res.resize(cnt); part.resize(cnt);
    B.resize(cnt); HB.resize(cnt);
    for (int i = 0; i < cnt; ++i) nmax = std::max(nmax, b.ctx[i]->ndet());
    G = static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>(cdiv(nmax, kB * 8), 1), 512));
    std::vector<Inst> hi(cnt);
    dactive.alloc(cnt);
    for (int i = 0; i < cnt; ++i) {
      sbdx_ctx& c = *b.ctx[i];
      const int64_t n = c.ndet();
      X[i].alloc(n); HX[i].alloc(n); R[i].alloc(n);
      coef[i].alloc(kMaxBlocks + 1); res[i].alloc(kSlots); part[i].alloc(static_cast<size_t>(G) * kSlots);
      B[i].resize(m); HB[i].resize(m);
      Inst in{};
      in.n = n;
      in.hii = c.hii.get();
      in.X = X[i].get(); in.HX = HX[i].get(); in.R = R[i].get();
      for (int j = 1; j < m; ++j) {
        B[i][j].alloc(n); HB[i][j].alloc(n);
        in.B[j] = B[i][j].get(); in.HB[j] = HB[i][j].get();
      }
      in.coef = coef[i].get(); in.res = res[i].get(); in.part = part[i].get();
      in.active = dactive.get() + i;
      hi[i] = in;
    }
    dinst.alloc(cnt);
    SBDX_HIP(hipMemcpy(dinst.get(), hi.data(), cnt * sizeof(Inst), hipMemcpyHostToDevice));
    active.assign(cnt, 1);
    SBDX_HIP(hipMemcpy(dactive.get(), active.data(), cnt * sizeof(int), hipMemcpyHostToDevice));
    h_coef.assign(static_cast<size_t>(cnt) * (kMaxBlocks + 1), 0.0);
    h_res.assign(static_cast<size_t>(cnt) * kSlots, 0.0);
    Hs.assign(cnt, std::vector<double>(m * m, 0.0));
    Ss = Hs;
    make_side_streams();
    // warm every instance's sigma path eagerly: tables, FP32 copies and workspaces are
    // allocated outside graph capture
    for (int i = 0; i < cnt; ++i) {
      sigma_block(*b.ctx[i], SBDX_FP64, 1, X[i].get(), HX[i].get(), st);
      if (prec == SBDX_FP32) sigma_block(*b.ctx[i], prec, 1, B[i][1].get(), HB[i][1].get(), st);
    }
    SBDX_HIP(hipStreamSynchronize(st));
  }
  ~BatchSolver() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    for (auto e : join_ev) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    for (auto s : side) (void)hipStreamDestroy(s);
  }

  // The problems' sigma builds are independent: inside a capture they go to kSide side streams
  // (fork from and join back to the main stream through events), so each problem's chain is its
  // own branch of the graph and small problems run beside large ones.
  static constexpr int kSide = 8;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join_ev;
  hipEvent_t fork_ev = nullptr;
  void make_side_streams() {
    side.resize(kSide);
    join_ev.resize(kSide);
    for (int q = 0; q < kSide; ++q) {
      SBDX_HIP(hipStreamCreate(&side[q]));  // default priority
      SBDX_HIP(hipEventCreateWithFlags(&join_ev[q], hipEventDisableTiming));
    }
    SBDX_HIP(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
  }
  // sigma of every active problem, in P or FP64, from x_of(i) into y_of(i)
  template <class FX, class FY>
  void sigma_all(int sprec, FX x_of, FY y_of) {
    SBDX_HIP(hipEventRecord(fork_ev, st));
    for (int q = 0; q < kSide; ++q) SBDX_HIP(hipStreamWaitEvent(side[q], fork_ev, 0));
    int slot = 0;
    for (int i = 0; i < cnt; ++i)
      if (active[i]) sigma_block(*b.ctx[i], sprec, 1, x_of(i), y_of(i), side[slot++ % kSide]);
    for (int q = 0; q < kSide; ++q) {
      SBDX_HIP(hipEventRecord(join_ev[q], side[q]));
      SBDX_HIP(hipStreamWaitEvent(st, join_ev[q], 0));
    }
  }

  dim3 grid() const { return dim3(G, static_cast<unsigned>(cnt)); }

  // run `body` (launches on st) as a graph cached under (phase, current active set)
  template <class F>
  void graph(int phase, F&& body) {
    auto key = std::make_pair(phase, epoch);
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      hipGraph_t g;
      SBDX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
      try {
        body();
      } catch (...) {
        hipGraph_t dummy;
        (void)hipStreamEndCapture(st, &dummy);
        throw;
      }
      SBDX_HIP(hipStreamEndCapture(st, &g));
      hipGraphExec_t ex;
      SBDX_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
      it = graphs.emplace(key, ex).first;
    }
    SBDX_HIP(hipGraphLaunch(it->second, st));
  }

  void reduce(int slot0, int mcount) {
    k_breduce<<<dim3(1, cnt), 64, 0, st>>>(dinst.get(), static_cast<int>(G), slot0, mcount);
    SBDX_CHECK_LAUNCH();
  }
  void upload_coef() {
    for (int i = 0; i < cnt; ++i)
      SBDX_HIP(hipMemcpyAsync(coef[i].get(), h_coef.data() + i * (kMaxBlocks + 1), (kMaxBlocks + 1) * 8,
                              hipMemcpyHostToDevice, st));
  }
  void download_res() {
    for (int i = 0; i < cnt; ++i)
      if (active[i])
        SBDX_HIP(hipMemcpyAsync(h_res.data() + i * kSlots, res[i].get(), kSlots * 8, hipMemcpyDeviceToHost, st));
    SBDX_HIP(hipStreamSynchronize(st));
  }
  // anchor phase: sigma of every active X in FP64 and its Rayleigh data
  void anchor_phase() {
    graph(1000, [&] {
      sigma_all(SBDX_FP64, [&](int i) { return X[i].get(); }, [&](int i) { return HX[i].get(); });
      k_banchor<<<grid(), kB, 0, st>>>(dinst.get());
      SBDX_CHECK_LAUNCH();
      reduce(0, 1);
      reduce(kMaxBlocks, 1);
    });
    download_res();
    for (int i = 0; i < cnt; ++i)
      if (active[i]) {
        Hs[i][0] = h_res[i * kSlots];
        Ss[i][0] = h_res[i * kSlots + kMaxBlocks];
      }
  }

  // expansion step to block j: residual of the current Ritz vector (coefficients in h_coef),
  // correction, orthogonalisation, sigma, new subspace row
  void expand_phase(int j) {
    upload_coef();
    graph(j, [&] {
      k_bcombine<P><<<grid(), kB, 0, st>>>(dinst.get(), j, 0);
      SBDX_CHECK_LAUNCH();
      reduce(2 * kMaxBlocks, 1);
      k_bprecond<P><<<grid(), kB, 0, st>>>(dinst.get(), j);
      SBDX_CHECK_LAUNCH();
      for (int pass = 0; pass < 2; ++pass) {
        k_bdots_t<P><<<grid(), kB, 0, st>>>(dinst.get(), j);
        SBDX_CHECK_LAUNCH();
        reduce(0, j + 1);
        k_bproject<P><<<grid(), kB, 0, st>>>(dinst.get(), j);
        SBDX_CHECK_LAUNCH();
      }
      sigma_all(prec, [&](int i) { return B[i][j].get(); }, [&](int i) { return HB[i][j].get(); });
      k_brow<P><<<grid(), kB, 0, st>>>(dinst.get(), j);
      SBDX_CHECK_LAUNCH();
      reduce(0, j + 1);
      reduce(kMaxBlocks, j + 1);
    });
    download_res();
  }

  // residual only (no expansion): for the last step of a cycle
  void residual_phase(int nb) {
    upload_coef();
    graph(2000 + nb, [&] {
      k_bcombine<P><<<grid(), kB, 0, st>>>(dinst.get(), nb, 0);
      SBDX_CHECK_LAUNCH();
      reduce(2 * kMaxBlocks, 1);
    });
    download_res();
  }

  // next anchor X = normalised Ritz vector (coefficients in h_coef)
  void restart_phase(int nb) {
    upload_coef();
    graph(3000 + nb, [&] {
      k_bcombine<P><<<grid(), kB, 0, st>>>(dinst.get(), nb, 1);
      SBDX_CHECK_LAUNCH();
      reduce(2 * kMaxBlocks, 1);
    });
    download_res();
    for (int i = 0; i < cnt; ++i) {
      const double nn = h_res[i * kSlots + 2 * kMaxBlocks];
      h_coef[i * (kMaxBlocks + 1)] = nn > 0 ? 1.0 / std::sqrt(nn) : 0.0;
    }
    upload_coef();
    graph(4000, [&] {
      k_bnext<<<grid(), kB, 0, st>>>(dinst.get());
      SBDX_CHECK_LAUNCH();
    });
  }

  void run(double* const* W, sbdx_davidson_result* out) {
    const double t0 = now();
    for (int i = 0; i < cnt; ++i) {
      std::memset(&out[i], 0, sizeof(out[i]));
      out[i].seconds_to_tol = -1.0;
      out[i].steps_to_tol = -1;
      init_hf(*b.ctx[i], X[i].get(), st);
    }
    std::vector<double> E(cnt, 0.0), nrm(cnt, 0.0);
    std::vector<int> steps(cnt, 0);
    int left = cnt, captured = cnt;
    for (int it = 0; it < o.max_iteration && left > 0; ++it) {
      anchor_phase();
      int nb = 1;
      for (;;) {
        // Rayleigh-Ritz for every active instance on nb blocks
        for (int i = 0; i < cnt; ++i) {
          if (!active[i]) continue;
          std::vector<double> H(nb * nb), S(nb * nb), u(nb);
          for (int a = 0; a < nb; ++a)
            for (int c = 0; c < nb; ++c) {
              H[a * nb + c] = Hs[i][a * m + c];
              S[a * nb + c] = Ss[i][a * m + c];
            }
          double e = 0.0;
          if (!ritz1(nb, H.data(), S.data(), u.data(), &e)) throw Error(SBDX_ERR_DEVICE, "batched Davidson subspace lost its rank");
          E[i] = e;
          for (int a = 0; a < kMaxBlocks; ++a) h_coef[i * (kMaxBlocks + 1) + a] = a < nb ? u[a] : 0.0;
          h_coef[i * (kMaxBlocks + 1) + kMaxBlocks] = e;
        }
        const bool last = nb == m;
        if (last) residual_phase(nb);
        else expand_phase(nb);
        bool changed = false;
        for (int i = 0; i < cnt; ++i) {
          if (!active[i]) continue;
          nrm[i] = std::sqrt(std::max(0.0, h_res[i * kSlots + 2 * kMaxBlocks]));
          ++steps[i];
          const bool exact = nb == 1;
          if (exact && o.energy_ref != 0.0 && out[i].seconds_to_tol < 0 && std::fabs(E[i] - o.energy_ref) < o.energy_tol) {
            out[i].seconds_to_tol = now() - t0;
            out[i].steps_to_tol = steps[i];
          }
          if (exact && nrm[i] < o.eps) {
            out[i].converged = 1;
            out[i].energy = E[i];
            out[i].residual = nrm[i];
            out[i].steps = steps[i];
            out[i].seconds = now() - t0;
            active[i] = 0;  // X is its solution (normalised at the restart)
            --left;
            changed = true;
          }
        }
        if (changed) {
          SBDX_HIP(hipMemcpyAsync(dactive.get(), active.data(), cnt * sizeof(int), hipMemcpyHostToDevice, st));
          SBDX_HIP(hipStreamSynchronize(st));
          // converged problems drop out of the batched kernels at once (device flags); the
          // graphs, whose sigma builds still include them, are recaptured only when a quarter
          // of their problems have left, which bounds both the wasted sigma work and the
          // number of captures
          if (4 * left <= 3 * captured) {
            ++epoch;
            captured = left;
          }
        }
        if (left == 0 || now() - t0 > o.max_time) break;
        if (last) break;
        for (int i = 0; i < cnt; ++i)
          if (active[i])
            for (int a = 0; a <= nb; ++a) {
              Hs[i][a * m + nb] = Hs[i][nb * m + a] = h_res[i * kSlots + a];
              Ss[i][a * m + nb] = Ss[i][nb * m + a] = h_res[i * kSlots + kMaxBlocks + a];
            }
        ++nb;
      }
      if (left == 0 || now() - t0 > o.max_time) break;
      restart_phase(nb);
    }
    for (int i = 0; i < cnt; ++i) {
      if (!out[i].converged) {
        out[i].energy = E[i];
        out[i].residual = nrm[i];
        out[i].steps = steps[i];
        out[i].seconds = now() - t0;
      }
      if (W && W[i]) {
        const int64_t n = b.ctx[i]->ndet();
        DevBuf<double> tmp(n);
        apply_sign(*b.ctx[i], SBDX_FP64, 1, X[i].get(), tmp.get(), st);
        SBDX_HIP(hipMemcpyAsync(W[i], tmp.get(), n * 8, hipMemcpyDeviceToHost, st));
        SBDX_HIP(hipStreamSynchronize(st));
      }
    }
  }
};

}  // namespace

void batch_davidson(sbdx_batch& b, const sbdx_davidson_opts& o, double* const* W, sbdx_davidson_result* res) {
  if (o.precision == SBDX_FP32) {
    for (auto* c : b.ctx) ensure_fp32(*c);
    BatchSolver<float> s(b, o);
    s.run(W, res);
  } else {
    BatchSolver<double> s(b, o);
    s.run(W, res);
  }
}

}  // namespace sbdx
