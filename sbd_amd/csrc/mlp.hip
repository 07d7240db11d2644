// mlp.hip -- the learned Davidson preconditioner: recording of training pairs, training
// (forward, backward, Adam) and inference, all on the GPU.
//
// Network: inputs x (130: alpha occupation bits, beta occupation bits, xh, xr), hidden width 64,
//   z1 = b1 + x W1, a1 = relu(z1), z2 = b2 + a1 W2, a2 = relu(z2), m = b3 + a2 . w3,
// parameters rounded to bfloat16, every product and sum in FP32.  The GEMM-shaped parts run on
// the matrix cores (v_mfma_f32_16x16x4_f32): one wavefront owns a tile of 64 samples and keeps
// its activations in the MFMA accumulator layout; transposes go through LDS.
//
// MFMA 16x16x4 lane layout (l = lane, lr = l % 16, lg = l / 16): lane l supplies A[lr][lg] and
// B[lg][lr]; accumulator register r of lane l holds D[4 lg + r][lr].  In every tile below,
// sample index s = 16 it + 4 lg + r and hidden unit u = 16 jt + lr ("C layout").
//
// Inference inside the solver replaces the bit part of layer 1 by per-string sums
// pa[ia] = b1 + sum_{p in alpha} W1[p], pb[ib] = sum_{p in beta} W1[64 + p] (computed once per
// network and context), so a determinant costs two row gathers plus the 64 x 64 layer.
#include "common.h"
#include "sbdx.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <type_traits>
#include <vector>

namespace sbdx {
namespace {

constexpr int H = 64;
constexpr int NIN = 130;
constexpr int KPAD = 132;  // layer-1 reduction length padded to the MFMA K of 4
constexpr int OFF_W1 = 0, OFF_B1 = NIN * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H, NPARAM = OFF_B3 + 1;
static_assert(NPARAM == SBDX_MLP_NPARAM, "parameter layout");
constexpr int LD = H + 1;  // LDS row stride of 64 x 64 tiles (bank-conflict padding)

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline v4f mfma(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

__host__ __device__ inline float bf16_round(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  b = (b + 0x7fffu + ((b >> 16) & 1u)) & 0xffff0000u;
  float r;
  memcpy(&r, &b, 4);
  return r;
}

// input k of a sample (k >= 130: zero padding)
__device__ inline float xin(uint64_t a, uint64_t b, float xh, float xr, int k) {
  if (k < 64) return static_cast<float>((a >> k) & 1ull);
  if (k < 128) return static_cast<float>((b >> (k - 64)) & 1ull);
  if (k == 128) return xh;
  if (k == 129) return xr;
  return 0.f;
}

__device__ inline float sum16(float v) {  // over the 16 lanes of a lane group
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 8, 64);
  return v;
}
__device__ inline float sum_groups(float v) {  // over the 4 lane groups (same lr)
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

struct TileShared {
  float w2[H][LD];  // W2[k][j]
  float b1[H], b2[H], w3[H];
  float a1[64][LD];  // a1[s][k]
  float dz[64][LD];  // dz2[s][j], then dz1[s][k]
  uint64_t sa[64], sb[64];
  float xh[64], xr[64], u[64], v[64];
};

__device__ inline void load_small_weights(TileShared& sh, const float* __restrict__ W) {
  for (int q = threadIdx.x; q < H * H; q += 64) sh.w2[q / H][q % H] = W[OFF_W2 + q];
  for (int q = threadIdx.x; q < H; q += 64) {
    sh.b1[q] = W[OFF_B1 + q];
    sh.b2[q] = W[OFF_B2 + q];
    sh.w3[q] = W[OFF_W3 + q];
  }
}

// layers 1-3 for the tile staged in sh; leaves a1 in sh.a1, returns a2 (C layout) and m
__device__ inline void forward_tile(TileShared& sh, const float* __restrict__ W, v4f (&y)[4][4], float (&m)[4][4]) {
  const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
  v4f z[4][4];
#pragma unroll
  for (int it = 0; it < 4; ++it)
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) z[it][jt] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < KPAD / 4; ++ks) {
    const int k = 4 * ks + lg;
    float bw[4];
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) bw[jt] = k < NIN ? W[OFF_W1 + k * H + 16 * jt + lr] : 0.f;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int s = 16 * it + lr;
      const float a = xin(sh.sa[s], sh.sb[s], sh.xh[s], sh.xr[s], k);
#pragma unroll
      for (int jt = 0; jt < 4; ++jt) z[it][jt] = mfma(a, bw[jt], z[it][jt]);
    }
  }
#pragma unroll
  for (int it = 0; it < 4; ++it)
#pragma unroll
    for (int jt = 0; jt < 4; ++jt)
#pragma unroll
      for (int r = 0; r < 4; ++r) sh.a1[16 * it + 4 * lg + r][16 * jt + lr] = fmaxf(z[it][jt][r] + sh.b1[16 * jt + lr], 0.f);
  __syncthreads();
#pragma unroll
  for (int it = 0; it < 4; ++it)
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) y[it][jt] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < H / 4; ++ks) {
    const int k = 4 * ks + lg;
    float bw[4];
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) bw[jt] = sh.w2[k][16 * jt + lr];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const float a = sh.a1[16 * it + lr][k];
#pragma unroll
      for (int jt = 0; jt < 4; ++jt) y[it][jt] = mfma(a, bw[jt], y[it][jt]);
    }
  }
  const float b3 = W[OFF_B3];
#pragma unroll
  for (int it = 0; it < 4; ++it)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float p = 0.f;
#pragma unroll
      for (int jt = 0; jt < 4; ++jt) {
        const float a2 = fmaxf(y[it][jt][r] + sh.b2[16 * jt + lr], 0.f);
        y[it][jt][r] = a2;
        p = fmaf(a2, sh.w3[16 * jt + lr], p);
      }
      m[it][r] = sum16(p) + b3;
    }
}

// stage samples [base, end) (through perm when given) of S into sh
__device__ inline void stage_samples(TileShared& sh, const sbdx_mlp_sample* __restrict__ S,
                                     const int64_t* __restrict__ perm, int64_t base, int64_t end) {
  const int l = threadIdx.x;
  const int64_t j = base + l;
  if (j < end) {
    const sbdx_mlp_sample s = S[perm ? perm[j] : j];
    sh.sa[l] = s.alpha;
    sh.sb[l] = s.beta;
    sh.xh[l] = bf16_round(s.xh);
    sh.xr[l] = bf16_round(s.xr);
    sh.u[l] = s.u;
    sh.v[l] = s.v;
  } else {
    sh.sa[l] = sh.sb[l] = 0;
    sh.xh[l] = sh.xr[l] = sh.u[l] = sh.v[l] = 0.f;
  }
}

// One tile of 64 samples per workgroup (one wavefront): forward, loss and backward; the tile's
// gradient sum goes to gpart[blockIdx.x][NPARAM], its loss sum to lpart[blockIdx.x].
__global__ __launch_bounds__(64) void k_mlp_grad(const sbdx_mlp_sample* __restrict__ S, const int64_t* __restrict__ perm,
                                                 int64_t off, int64_t end, float inv_n, const float* __restrict__ W,
                                                 float* __restrict__ gpart, float* __restrict__ lpart) {
  __shared__ TileShared sh;
  const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
  stage_samples(sh, S, perm, off + static_cast<int64_t>(blockIdx.x) * 64, end);
  load_small_weights(sh, W);
  __syncthreads();
  v4f y[4][4];
  float m[4][4];
  forward_tile(sh, W, y, m);
  float* g = gpart + static_cast<int64_t>(blockIdx.x) * NPARAM;

  // loss, dL/dm, dz2 = dm w3 [a2 > 0]; dw3, db2 partials per lane (unit 16 jt + lr)
  float dw3[4] = {0.f, 0.f, 0.f, 0.f}, db2[4] = {0.f, 0.f, 0.f, 0.f}, db3 = 0.f, loss = 0.f;
#pragma unroll
  for (int it = 0; it < 4; ++it)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int s = 16 * it + 4 * lg + r;
      const float u = sh.u[s], v = sh.v[s];
      const float e = fmaf(u, 1.f + m[it][r], -v);
      const float dm = 2.f * e * u * inv_n;
      if (lr == 0) {
        loss = fmaf(e * e, inv_n, loss);
        db3 += dm;
      }
#pragma unroll
      for (int jt = 0; jt < 4; ++jt) {
        const float a2 = y[it][jt][r];
        dw3[jt] = fmaf(dm, a2, dw3[jt]);
        const float d = a2 > 0.f ? dm * sh.w3[16 * jt + lr] : 0.f;
        db2[jt] += d;
        sh.dz[s][16 * jt + lr] = d;
      }
    }
  __syncthreads();
  // dW2 = a1^T dz2 (reduction over the tile's samples)
  for (int kt = 0; kt < 4; ++kt) {
    v4f acc[4];
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) acc[jt] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 16; ++ks) {
      const int s = 4 * ks + lg;
      const float a = sh.a1[s][16 * kt + lr];
#pragma unroll
      for (int jt = 0; jt < 4; ++jt) acc[jt] = mfma(a, sh.dz[s][16 * jt + lr], acc[jt]);
    }
#pragma unroll
    for (int jt = 0; jt < 4; ++jt)
#pragma unroll
      for (int r = 0; r < 4; ++r) g[OFF_W2 + (16 * kt + 4 * lg + r) * H + 16 * jt + lr] = acc[jt][r];
  }
  // da1 = dz2 W2^T, dz1 = da1 [a1 > 0]
  v4f d[4][4];
#pragma unroll
  for (int it = 0; it < 4; ++it)
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) d[it][kt] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < 16; ++ks) {
    const int j = 4 * ks + lg;
    float bw[4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) bw[kt] = sh.w2[16 * kt + lr][j];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const float a = sh.dz[16 * it + lr][j];
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) d[it][kt] = mfma(a, bw[kt], d[it][kt]);
    }
  }
  __syncthreads();  // every read of dz2 is done
  float db1[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int it = 0; it < 4; ++it)
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int s = 16 * it + 4 * lg + r, k = 16 * kt + lr;
        const float dz1 = sh.a1[s][k] > 0.f ? d[it][kt][r] : 0.f;
        db1[kt] += dz1;
        sh.dz[s][k] = dz1;
      }
  __syncthreads();
  // dW1 = X^T dz1
  for (int pt = 0; pt < (NIN + 15) / 16; ++pt) {
    v4f acc[4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) acc[kt] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < 16; ++ks) {
      const int s = 4 * ks + lg;
      const float a = xin(sh.sa[s], sh.sb[s], sh.xh[s], sh.xr[s], 16 * pt + lr);
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) acc[kt] = mfma(a, sh.dz[s][16 * kt + lr], acc[kt]);
    }
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int p = 16 * pt + 4 * lg + r;
        if (p < NIN) g[OFF_W1 + p * H + 16 * kt + lr] = acc[kt][r];
      }
  }
  // vector parameters: reduce over the lane groups
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const float s1 = sum_groups(db1[t]), s2 = sum_groups(db2[t]), s3 = sum_groups(dw3[t]);
    if (lg == 0) {
      g[OFF_B1 + 16 * t + lr] = s1;
      g[OFF_B2 + 16 * t + lr] = s2;
      g[OFF_W3 + 16 * t + lr] = s3;
    }
  }
  const float tb3 = sum_groups(db3), tl = sum_groups(loss);
  if (l == 0) {
    g[OFF_B3] = tb3;
    lpart[blockIdx.x] = tl;
  }
}

__global__ __launch_bounds__(64) void k_mlp_forward(const sbdx_mlp_sample* __restrict__ S, int64_t n,
                                                    const float* __restrict__ W, float* __restrict__ out) {
  __shared__ TileShared sh;
  const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * 64;
  stage_samples(sh, S, nullptr, base, n);
  load_small_weights(sh, W);
  __syncthreads();
  v4f y[4][4];
  float m[4][4];
  forward_tile(sh, W, y, m);
  if (lr == 0)
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t j = base + 16 * it + 4 * lg + r;
        if (j < n) out[j] = m[it][r];
      }
}

// Sum of tile partials (fixed order, deterministic), then either the gradient out (adam == 0)
// or one Adam step on the FP32 master parameters and their bfloat16-rounded working copy.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ gpart, const float* __restrict__ lpart,
                                                    int ntile, int adam, float lr, float b1, float b2, float eps,
                                                    float wd, float c1, float c2, float* __restrict__ master,
                                                    float* __restrict__ mom1, float* __restrict__ mom2,
                                                    float* __restrict__ work, float* __restrict__ grad_out,
                                                    double* __restrict__ loss_out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < NPARAM) {
    float g = 0.f;
    for (int t = 0; t < ntile; ++t) g += gpart[static_cast<int64_t>(t) * NPARAM + q];
    if (adam) {
      const float m1 = b1 * mom1[q] + (1.f - b1) * g;
      const float m2 = b2 * mom2[q] + (1.f - b2) * g * g;
      mom1[q] = m1;
      mom2[q] = m2;
      const float p = master[q] - lr * ((m1 * c1) / (sqrtf(m2 * c2) + eps) + wd * master[q]);
      master[q] = p;
      work[q] = bf16_round(p);
    } else {
      grad_out[q] = g;
    }
  }
  if (q == 0) {
    double s = 0.0;
    for (int t = 0; t < ntile; ++t) s += lpart[t];
    *loss_out = s;
  }
}

__global__ void k_round_params(const float* __restrict__ in, float* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < NPARAM) out[q] = bf16_round(in[q]);
}

// per-string first-layer sums: out[s][j] = (add_bias ? b1[j] : 0) + sum_{p in str[s]} W1[row0 + p][j]
__global__ __launch_bounds__(256) void k_mlp_project(const uint64_t* __restrict__ str, int64_t n, int row0, int add_bias,
                                                     const float* __restrict__ W, float* __restrict__ out) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (q >= n * H) return;
  const int64_t s = q / H;
  const int j = static_cast<int>(q % H);
  uint64_t x = str[s];
  float acc = add_bias ? W[OFF_B1 + j] : 0.f;
  while (x) {
    const int p = ctz64(x);
    acc += W[OFF_W1 + (row0 + p) * H + j];
    x &= x - 1;
  }
  out[q] = acc;
}

template <int K>
struct ColParams {
  double theta[K];
  double inv_rms[K];
};

struct PrecShared {
  float w2[H][LD];
  float b2[H], w3[H], w1h[H], w1r[H];
  float a1[64][LD];
  float xh[64], xr[64];
  double theta[8], inv_rms[8];
};

// T[i*K + v] = R[i*K + v] / den(theta_v - H_ii) * (1 + m_iv); one wavefront per tile of 64
// determinants, grid-stride over tiles
template <typename P, int K>
__global__ __launch_bounds__(64) void k_mlp_precond(const double* __restrict__ R, const double* __restrict__ hii,
                                                    const float* __restrict__ W, const float* __restrict__ pa,
                                                    const float* __restrict__ pb, int64_t nbeta, ColParams<K> cp,
                                                    P* __restrict__ T, int64_t n) {
  __shared__ PrecShared sh;
  const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
  for (int q = l; q < H * H; q += 64) sh.w2[q / H][q % H] = W[OFF_W2 + q];
  for (int q = l; q < H; q += 64) {
    sh.b2[q] = W[OFF_B2 + q];
    sh.w3[q] = W[OFF_W3 + q];
    sh.w1h[q] = W[OFF_W1 + 128 * H + q];
    sh.w1r[q] = W[OFF_W1 + 129 * H + q];
  }
  if (l < K) {
    sh.theta[l] = cp.theta[l];
    sh.inv_rms[l] = cp.inv_rms[l];
  }
  const float b3 = W[OFF_B3];
  __syncthreads();
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * 64; base < n; base += static_cast<int64_t>(gridDim.x) * 64) {
#pragma unroll 1
    for (int v = 0; v < K; ++v) {
      {
        const int64_t i = base + l;
        float fh = 0.f, fr = 0.f;
        if (i < n) {
          fh = static_cast<float>(asinh(hii[i] - sh.theta[v]));
          fr = static_cast<float>(asinh(fabs(R[i * K + v]) * sh.inv_rms[v]));
        }
        sh.xh[l] = bf16_round(fh);
        sh.xr[l] = bf16_round(fr);
      }
      __syncthreads();
      // layer 1 from the per-string sums
#pragma unroll
      for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int s = 16 * it + 4 * lg + r;
          int64_t i = base + s;
          if (i >= n) i = n - 1;  // padding rows: any valid determinant, output discarded
          const int64_t ia = i / nbeta, ib = i - ia * nbeta;
#pragma unroll
          for (int jt = 0; jt < 4; ++jt) {
            const int u = 16 * jt + lr;
            const float z = pa[ia * H + u] + pb[ib * H + u] + sh.xh[s] * sh.w1h[u] + sh.xr[s] * sh.w1r[u];
            sh.a1[s][u] = fmaxf(z, 0.f);
          }
        }
      __syncthreads();
      v4f y[4][4];
#pragma unroll
      for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int jt = 0; jt < 4; ++jt) y[it][jt] = v4f{0.f, 0.f, 0.f, 0.f};
      for (int ks = 0; ks < H / 4; ++ks) {
        const int k = 4 * ks + lg;
        float bw[4];
#pragma unroll
        for (int jt = 0; jt < 4; ++jt) bw[jt] = sh.w2[k][16 * jt + lr];
#pragma unroll
        for (int it = 0; it < 4; ++it) {
          const float a = sh.a1[16 * it + lr][k];
#pragma unroll
          for (int jt = 0; jt < 4; ++jt) y[it][jt] = mfma(a, bw[jt], y[it][jt]);
        }
      }
#pragma unroll
      for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float p = 0.f;
#pragma unroll
          for (int jt = 0; jt < 4; ++jt) p = fmaf(fmaxf(y[it][jt][r] + sh.b2[16 * jt + lr], 0.f), sh.w3[16 * jt + lr], p);
          const float m = sum16(p) + b3;
          const int64_t i = base + 16 * it + 4 * lg + r;
          if (lr == 0 && i < n) {
            const double den = sh.theta[v] - hii[i];
            const double dd = fabs(den) > 1.0e-12 ? den : den - 1.0e-12;
            T[i * K + v] = static_cast<P>(R[i * K + v] / dd * (1.0 + static_cast<double>(m)));
          }
        }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------- training-pair recording

__device__ inline double hash01(uint64_t i, uint64_t seed) {
  // splitmix64 finaliser (Steele, Lea, Flood 2014) of seed ^ ((i + 1) * golden ratio)
  uint64_t z = seed ^ ((i + 1) * 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
}

struct RecScalars {
  double sx, theta, sxs, proj;  // x scale, Rayleigh quotient, signed x* scale, x^.x*^
  double vnorm, c, q_total, target, inv_rms;
  uint64_t seed;
};

// per determinant (blocked basis): r, d, t
__device__ inline void rec_terms(const double* __restrict__ x, const double* __restrict__ hx,
                                 const double* __restrict__ xs, const double* __restrict__ hii, int64_t i,
                                 const RecScalars& k, double& r, double& d, double& t) {
  const double xi = x[i] * k.sx;
  r = hx[i] * k.sx - k.theta * xi;
  d = xs[i] * k.sxs - k.proj * xi;
  const double den = k.theta - hii[i];
  t = r / (fabs(den) > 1.0e-12 ? den : den - 1.0e-12);
}

// mode 0: sums of x.x, x.hx, x.xs, xs.xs; mode 1: r.r, t.d, d.d; mode 2: sum of u^2 + v^2
__global__ __launch_bounds__(256) void k_rec_reduce(int mode, const double* __restrict__ x, const double* __restrict__ hx,
                                                    const double* __restrict__ xs, const double* __restrict__ hii,
                                                    RecScalars k, int64_t n, double* __restrict__ partial) {
  __shared__ double sh[4][4];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
    if (mode == 0) {
      const double xi = x[i], si = xs[i];
      a0 = fma(xi, xi, a0);
      a1 = fma(xi, hx[i], a1);
      a2 = fma(xi, si, a2);
      a3 = fma(si, si, a3);
    } else {
      double r, d, t;
      rec_terms(x, hx, xs, hii, i, k, r, d, t);
      if (mode == 1) {
        a0 = fma(r, r, a0);
        a1 = fma(t, d, a1);
        a2 = fma(d, d, a2);
      } else {
        const double u = t / k.vnorm, v = k.c * d / k.vnorm;
        a0 = fma(u, u, fma(v, v, a0));
      }
    }
  }
  double acc[4] = {a0, a1, a2, a3};
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = 0; q < 4; ++q) {
    double v = acc[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) sh[w][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) partial[blockIdx.x * 4 + threadIdx.x] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

__global__ __launch_bounds__(256) void k_rec_sample(const double* __restrict__ x, const double* __restrict__ hx,
                                                    const double* __restrict__ xs, const double* __restrict__ hii,
                                                    const uint64_t* __restrict__ astr, const uint64_t* __restrict__ bstr,
                                                    int64_t nbeta, RecScalars k, int64_t n,
                                                    sbdx_mlp_sample* __restrict__ out, int64_t cap,
                                                    unsigned long long* __restrict__ counter) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
    double r, d, t;
    rec_terms(x, hx, xs, hii, i, k, r, d, t);
    const double u = t / k.vnorm, v = k.c * d / k.vnorm;
    const double p = fmin(1.0, k.target * (u * u + v * v) / k.q_total);
    if (!(hash01(static_cast<uint64_t>(i), k.seed) < p)) continue;
    const unsigned long long slot = atomicAdd(counter, 1ull);
    if (static_cast<int64_t>(slot) >= cap) continue;
    const int64_t ia = i / nbeta, ib = i - ia * nbeta;
    const uint64_t a = astr[ia], b = bstr[ib];
    const double w = det_sign(a, b) / sqrt(p);  // back to the caller's basis, importance weight
    sbdx_mlp_sample s;
    s.alpha = a;
    s.beta = b;
    s.xh = static_cast<float>(asinh(hii[i] - k.theta));
    s.xr = static_cast<float>(asinh(fabs(r) * k.inv_rms));
    s.u = static_cast<float>(u * w);
    s.v = static_cast<float>(v * w);
    out[slot] = s;
  }
}

inline unsigned grid_for(int64_t n, int per) {
  return static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>(cdiv(n, per), 1), 4096));
}

struct Stream {
  hipStream_t s = nullptr;
  Stream() { SBDX_HIP(hipStreamCreate(&s)); }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

// one full-batch gradient (and loss) of rounded parameters Wr over samples dS[0, n)
void loss_grad_dev(const sbdx_mlp_sample* dS, int64_t n, const float* Wr, DevBuf<float>& gpart, DevBuf<float>& lpart,
                   float* dgrad, double* dloss, hipStream_t st) {
  const int64_t ntile = std::max<int64_t>(1, cdiv(n, 64));
  gpart.ensure(ntile * NPARAM);
  lpart.ensure(ntile);
  k_mlp_grad<<<static_cast<unsigned>(ntile), 64, 0, st>>>(dS, nullptr, 0, n, 1.f / static_cast<float>(std::max<int64_t>(n, 1)), Wr,
                                                          gpart.get(), lpart.get());
  SBDX_CHECK_LAUNCH();
  k_mlp_reduce<<<cdiv(NPARAM, 256), 256, 0, st>>>(gpart.get(), lpart.get(), static_cast<int>(ntile), 0, 0.f, 0.f, 0.f,
                                                  0.f, 0.f, 0.f, 0.f, nullptr, nullptr, nullptr, nullptr, dgrad, dloss);
  SBDX_CHECK_LAUNCH();
}

}  // namespace

void mlp_precondition(sbdx_ctx& c, int prec, int k, const double* R, const double* theta, const double* rms, void* T,
                      hipStream_t st) {
  if (!c.have_mlp) throw Error(SBDX_ERR_STATE, "preconditioner 1 needs a network (sbdx_mlp_set)");
  const int64_t n = c.ndet();
  const unsigned g = static_cast<unsigned>(std::min<int64_t>(cdiv(n, 64), 8192));
  auto go = [&](auto tag, auto kk) {
    using P = decltype(tag);
    constexpr int K = decltype(kk)::value;
    ColParams<K> cp;
    for (int v = 0; v < K; ++v) {
      cp.theta[v] = theta[v];
      cp.inv_rms[v] = rms[v] > 0 ? 1.0 / rms[v] : 0.0;
    }
    k_mlp_precond<P, K><<<g, 64, 0, st>>>(R, c.hii.get(), c.mlp_w.get(), c.mlp_pa.get(), c.mlp_pb.get(), c.nb, cp,
                                          static_cast<P*>(T), n);
    SBDX_CHECK_LAUNCH();
  };
  auto by_k = [&](auto tag) {
    switch (k) {
      case 1: go(tag, std::integral_constant<int, 1>{}); break;
      case 2: go(tag, std::integral_constant<int, 2>{}); break;
      case 4: go(tag, std::integral_constant<int, 4>{}); break;
      case 8: go(tag, std::integral_constant<int, 8>{}); break;
      default: throw Error(SBDX_ERR_ARG, "block width must be 1, 2, 4 or 8");
    }
  };
  if (prec == SBDX_FP32) by_k(float{});
  else by_k(double{});
}

void mlp_set(sbdx_ctx& c, const float* params) {
  if (!params) {
    c.have_mlp = false;
    c.mlp_w.release();
    c.mlp_pa.release();
    c.mlp_pb.release();
    return;
  }
  if (c.L > 64) throw Error(SBDX_ERR_ARG, "the network's inputs hold at most 64 orbitals per spin");
  std::vector<float> w(NPARAM);
  for (int q = 0; q < NPARAM; ++q) {
    if (!std::isfinite(params[q])) throw Error(SBDX_ERR_ARG, "network parameters must be finite");
    w[q] = bf16_round(params[q]);
  }
  c.mlp_w.ensure(NPARAM);
  c.mlp_pa.ensure(c.na * H);
  c.mlp_pb.ensure(c.nb * H);
  SBDX_HIP(hipMemcpyAsync(c.mlp_w.get(), w.data(), NPARAM * 4, hipMemcpyHostToDevice, c.stream));
  k_mlp_project<<<grid_for(c.na * H, 256), 256, 0, c.stream>>>(c.A.str.get(), c.na, 0, 1, c.mlp_w.get(), c.mlp_pa.get());
  SBDX_CHECK_LAUNCH();
  k_mlp_project<<<grid_for(c.nb * H, 256), 256, 0, c.stream>>>(c.B.str.get(), c.nb, 64, 0, c.mlp_w.get(), c.mlp_pb.get());
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipStreamSynchronize(c.stream));
  c.have_mlp = true;
}

// dX, dXs: device vectors in the blocked basis
int64_t mlp_record(sbdx_ctx& c, const double* dX, const double* dXs, double target, uint64_t seed,
                   sbdx_mlp_sample* out, int64_t cap, double* theta_out) {
  if (c.L > 64) throw Error(SBDX_ERR_ARG, "the network's inputs hold at most 64 orbitals per spin");
  const int64_t n = c.ndet();
  hipStream_t st = c.stream;
  DevBuf<double> hx(n), part(4096 * 4);
  sigma_blocked(c, SBDX_FP64, 1, dX, hx.get(), st);
  const unsigned g = grid_for(n, 1024);
  RecScalars k{};
  auto reduce = [&](int mode, double* s) {
    k_rec_reduce<<<g, 256, 0, st>>>(mode, dX, hx.get(), dXs, c.hii.get(), k, n, part.get());
    SBDX_CHECK_LAUNCH();
    std::vector<double> h(static_cast<size_t>(g) * 4);
    SBDX_HIP(hipMemcpyAsync(h.data(), part.get(), h.size() * 8, hipMemcpyDeviceToHost, st));
    SBDX_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < 4; ++q) s[q] = 0.0;
    for (unsigned b = 0; b < g; ++b)
      for (int q = 0; q < 4; ++q) s[q] += h[b * 4 + q];
  };
  double s0[4], s1[4], s2[4];
  reduce(0, s0);
  if (!(s0[0] > 0.0) || !(s0[3] > 0.0)) throw Error(SBDX_ERR_ARG, "record needs nonzero vectors");
  k.sx = 1.0 / std::sqrt(s0[0]);
  k.theta = s0[1] / s0[0];
  k.sxs = (s0[2] < 0 ? -1.0 : 1.0) / std::sqrt(s0[3]);
  k.proj = std::fabs(s0[2]) / std::sqrt(s0[0] * s0[3]);
  reduce(1, s1);
  if (!(s1[2] > 0.0) || s1[1] == 0.0) throw Error(SBDX_ERR_ARG, "record needs an iterate different from the solution");
  k.c = s1[1] / s1[2];
  k.vnorm = std::fabs(k.c) * std::sqrt(s1[2]);
  k.inv_rms = s1[0] > 0 ? 1.0 / std::sqrt(s1[0] / static_cast<double>(n)) : 0.0;
  reduce(2, s2);
  k.q_total = s2[0];
  k.target = target;
  k.seed = seed;
  DevBuf<unsigned long long> counter(1);
  DevBuf<sbdx_mlp_sample> dout(std::max<int64_t>(cap, 1));
  SBDX_HIP(hipMemsetAsync(counter.get(), 0, 8, st));
  k_rec_sample<<<g, 256, 0, st>>>(dX, hx.get(), dXs, c.hii.get(), c.A.str.get(), c.B.str.get(), c.nb, k, n, dout.get(), cap,
                                  counter.get());
  SBDX_CHECK_LAUNCH();
  unsigned long long cnt = 0;
  SBDX_HIP(hipMemcpyAsync(&cnt, counter.get(), 8, hipMemcpyDeviceToHost, st));
  SBDX_HIP(hipStreamSynchronize(st));
  const int64_t keep = std::min<int64_t>(static_cast<int64_t>(cnt), cap);
  if (keep > 0) SBDX_HIP(hipMemcpy(out, dout.get(), keep * sizeof(sbdx_mlp_sample), hipMemcpyDeviceToHost));
  if (theta_out) *theta_out = k.theta;
  return static_cast<int64_t>(cnt);
}

void mlp_loss_grad(const sbdx_mlp_sample* S, int64_t n, const float* params, double* loss, float* grad) {
  Stream st;
  std::vector<float> w(NPARAM);
  for (int q = 0; q < NPARAM; ++q) w[q] = bf16_round(params[q]);
  DevBuf<sbdx_mlp_sample> dS(std::max<int64_t>(n, 1));
  DevBuf<float> dW(NPARAM), dG(NPARAM), gpart, lpart;
  DevBuf<double> dL(1);
  SBDX_HIP(hipMemcpyAsync(dS.get(), S, n * sizeof(sbdx_mlp_sample), hipMemcpyHostToDevice, st.s));
  SBDX_HIP(hipMemcpyAsync(dW.get(), w.data(), NPARAM * 4, hipMemcpyHostToDevice, st.s));
  loss_grad_dev(dS.get(), n, dW.get(), gpart, lpart, dG.get(), dL.get(), st.s);
  SBDX_HIP(hipMemcpyAsync(grad, dG.get(), NPARAM * 4, hipMemcpyDeviceToHost, st.s));
  SBDX_HIP(hipMemcpyAsync(loss, dL.get(), 8, hipMemcpyDeviceToHost, st.s));
  SBDX_HIP(hipStreamSynchronize(st.s));
}

void mlp_forward(const sbdx_mlp_sample* S, int64_t n, const float* params, float* m) {
  Stream st;
  std::vector<float> w(NPARAM);
  for (int q = 0; q < NPARAM; ++q) w[q] = bf16_round(params[q]);
  DevBuf<sbdx_mlp_sample> dS(std::max<int64_t>(n, 1));
  DevBuf<float> dW(NPARAM), dm(std::max<int64_t>(n, 1));
  SBDX_HIP(hipMemcpyAsync(dS.get(), S, n * sizeof(sbdx_mlp_sample), hipMemcpyHostToDevice, st.s));
  SBDX_HIP(hipMemcpyAsync(dW.get(), w.data(), NPARAM * 4, hipMemcpyHostToDevice, st.s));
  k_mlp_forward<<<static_cast<unsigned>(std::max<int64_t>(1, cdiv(n, 64))), 64, 0, st.s>>>(dS.get(), n, dW.get(), dm.get());
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipMemcpyAsync(m, dm.get(), n * 4, hipMemcpyDeviceToHost, st.s));
  SBDX_HIP(hipStreamSynchronize(st.s));
}

void mlp_train(const sbdx_mlp_sample* S, int64_t n, const sbdx_mlp_train_opts& o, float* params, double* trace,
               int trace_cap) {
  Stream st;
  const int64_t batch = std::max<int64_t>(64, std::min<int64_t>(o.batch, n));
  const int64_t ntile = cdiv(batch, 64);
  DevBuf<sbdx_mlp_sample> dS(std::max<int64_t>(n, 1));
  DevBuf<int64_t> perm(std::max<int64_t>(n, 1));
  DevBuf<float> master(NPARAM), m1(NPARAM), m2(NPARAM), work(NPARAM), gpart(ntile * NPARAM), lpart(ntile);
  DevBuf<double> dl(1);
  SBDX_HIP(hipMemcpyAsync(dS.get(), S, n * sizeof(sbdx_mlp_sample), hipMemcpyHostToDevice, st.s));
  SBDX_HIP(hipMemcpyAsync(master.get(), params, NPARAM * 4, hipMemcpyHostToDevice, st.s));
  SBDX_HIP(hipMemsetAsync(m1.get(), 0, NPARAM * 4, st.s));
  SBDX_HIP(hipMemsetAsync(m2.get(), 0, NPARAM * 4, st.s));
  k_round_params<<<cdiv(NPARAM, 256), 256, 0, st.s>>>(master.get(), work.get());
  SBDX_CHECK_LAUNCH();
  std::vector<int64_t> hp(n);
  std::iota(hp.begin(), hp.end(), 0);
  std::mt19937_64 rng(o.seed);
  int64_t step = 0;
  for (int ep = 0; ep < o.epochs; ++ep) {
    std::shuffle(hp.begin(), hp.end(), rng);
    SBDX_HIP(hipMemcpyAsync(perm.get(), hp.data(), n * 8, hipMemcpyHostToDevice, st.s));
    double ep_loss = 0.0;
    int64_t nb = 0;
    for (int64_t off = 0; off < n; off += batch) {
      const int64_t end = std::min(n, off + batch);
      const int64_t tiles = cdiv(end - off, 64);
      ++step;
      const double c1 = 1.0 / (1.0 - std::pow(o.beta1, static_cast<double>(step)));
      const double c2 = 1.0 / (1.0 - std::pow(o.beta2, static_cast<double>(step)));
      k_mlp_grad<<<static_cast<unsigned>(tiles), 64, 0, st.s>>>(dS.get(), perm.get(), off, end,
                                                                 1.f / static_cast<float>(end - off), work.get(),
                                                                 gpart.get(), lpart.get());
      SBDX_CHECK_LAUNCH();
      k_mlp_reduce<<<cdiv(NPARAM, 256), 256, 0, st.s>>>(gpart.get(), lpart.get(), static_cast<int>(tiles), 1,
                                                        static_cast<float>(o.lr), static_cast<float>(o.beta1),
                                                        static_cast<float>(o.beta2), static_cast<float>(o.eps),
                                                        static_cast<float>(o.weight_decay), static_cast<float>(c1),
                                                        static_cast<float>(c2), master.get(), m1.get(), m2.get(),
                                                        work.get(), nullptr, dl.get());
      SBDX_CHECK_LAUNCH();
      double bl = 0.0;
      SBDX_HIP(hipMemcpyAsync(&bl, dl.get(), 8, hipMemcpyDeviceToHost, st.s));
      SBDX_HIP(hipStreamSynchronize(st.s));
      ep_loss += bl;
      ++nb;
    }
    ep_loss /= static_cast<double>(std::max<int64_t>(nb, 1));
    if (trace && ep < trace_cap) trace[ep] = ep_loss;
    if (o.verbose) {
      std::printf(" mlp epoch %d: loss %.6e\n", ep, ep_loss);
      std::fflush(stdout);
    }
  }
  SBDX_HIP(hipMemcpyAsync(params, master.get(), NPARAM * 4, hipMemcpyDeviceToHost, st.s));
  SBDX_HIP(hipStreamSynchronize(st.s));
}

}  // namespace sbdx
