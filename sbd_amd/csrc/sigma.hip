// sigma.hip -- the on-the-fly H C product (chemistry/tpb/mult.h:387-648) in the blocked basis.
//
// sigma = D C + H_aa C + H_bb C + H_ab C, with C an n_alpha x n_beta row-major matrix.
//
//  * Row pass (alpha rows): one workgroup per (alpha row ia, tile of 256 beta columns); every
//    wavefront covers 64 consecutive columns of the same row, so the row's excitation lists are
//    wavefront-uniform (scalar loads) and every gathered operand C[ja, ib..ib+63] is one
//    coalesced 512-byte (FP64) segment.  The diagonal, the alpha singles and doubles (type 2)
//    and the alpha-single x beta-single part (type 0) are fused here, so sigma is written once,
//    by its owner, with no atomics.
//  * Beta part (type 1): C is transposed, the same row kernel runs with the beta tables, and
//    the result is added back transposed.
//
// Matrix elements (blocked order, signs from one spin string each):
//   alpha single ja -> ia (i -> a):  s_a (F_a + sum_{j in beta(ib)} (ai|jj))
//   alpha double:                     stored value (sign and exchange included)
//   alpha single x beta single:       s_a s_b (ai|bj)
// Arithmetic: FP64 (plain) or FP32 with compensated accumulation (error-free product and
// sum transformations, i.e. Ogita-Rump-Oishi Dot2 per output element).
#include "common.h"

namespace sbdx {
namespace {

constexpr int kTile = 256;

struct AccD {
  double s = 0.0;
  __device__ inline void add(double p) { s += p; }
  __device__ inline void fma(double a, double b) { s = __fma_rn(a, b, s); }
  __device__ inline double get() const { return s; }
};

// compensated FP32: p = a*b exactly as p + e (FMA), s + p exactly as t + err (TwoSum)
struct AccF {
  float s = 0.0f, c = 0.0f;
  __device__ inline void add(float p) {
    const float t = __fadd_rn(s, p);
    const float bp = __fsub_rn(t, s);
    c += __fadd_rn(__fsub_rn(s, __fsub_rn(t, bp)), __fsub_rn(p, bp));
    s = t;
  }
  __device__ inline void fma(float a, float b) {
    const float p = __fmul_rn(a, b);
    c += __fmaf_rn(a, b, -p);
    add(p);
  }
  __device__ inline float get() const { return s + c; }
};

template <typename T> struct AccOf;
template <> struct AccOf<double> { using type = AccD; };
template <> struct AccOf<float> { using type = AccF; };

struct RowTabs {
  const int64_t* off_s;
  const uint32_t* ket_s;
  const uint32_t* exc_s;
  const int64_t* off_d;
  const uint32_t* ket_d;
};

template <typename T, bool kFuse>
__global__ __launch_bounds__(kTile) void k_sigma_rows(
    int64_t nrow, int64_t ncol, int L, RowTabs rt, const T* __restrict__ val_s,
    const T* __restrict__ val_d, const uint64_t* __restrict__ col_str, const T* __restrict__ V,
    const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ hii, RowTabs ct) {
  const int64_t g = blockIdx.x;
  const int64_t tile = g / nrow;
  const int64_t row = g - tile * nrow;
  const int64_t col = tile * kTile + threadIdx.x;
  if (col >= ncol) return;
  const int64_t L2 = static_cast<int64_t>(L) * L;
  typename AccOf<T>::type acc;
  const int64_t me = row * ncol + col;
  if (kFuse) acc.fma(hii[me], x[me]);

  // same-spin doubles
  const int64_t d0 = rt.off_d[row], d1 = rt.off_d[row + 1];
  for (int64_t k = d0; k < d1; ++k) acc.fma(val_d[k], x[static_cast<int64_t>(rt.ket_d[k]) * ncol + col]);

  // same-spin singles, with the spectator string's Coulomb term
  const uint64_t ycol = col_str[col];
  const int64_t s0 = rt.off_s[row], s1 = rt.off_s[row + 1];
  for (int64_t k = s0; k < s1; ++k) {
    const uint32_t e = rt.exc_s[k];
    const int i = e & 63, a = (e >> 6) & 63;
    const T* vp = V + (static_cast<int64_t>(a) * L + i) * L2;
    T coul = T(0);
    uint64_t o = ycol;
    while (o) {
      const int j = ctz64(o);
      o &= o - 1;
      coul += vp[j * L + j];
    }
    const T v = val_s[k] + ((e >> 12) & 1u ? -coul : coul);
    acc.fma(v, x[static_cast<int64_t>(rt.ket_s[k]) * ncol + col]);
  }

  if (kFuse) {
    // alpha single (row, wavefront-uniform) x beta single (column, per lane)
    const int64_t m0 = ct.off_s[col], m1 = ct.off_s[col + 1];
    for (int64_t k = s0; k < s1; ++k) {
      const uint32_t e = rt.exc_s[k];
      const int i = e & 63, a = (e >> 6) & 63;
      const T* vp = V + (static_cast<int64_t>(a) * L + i) * L2;
      const T* xr = x + static_cast<int64_t>(rt.ket_s[k]) * ncol;
      typename AccOf<T>::type part;
      for (int64_t m = m0; m < m1; ++m) {
        const uint32_t eb = ct.exc_s[m];
        const T v = vp[((eb >> 6) & 63) * L + (eb & 63)];
        part.fma((eb >> 12) & 1u ? -v : v, xr[ct.ket_s[m]]);
      }
      const T p = part.get();
      acc.add((e >> 12) & 1u ? -p : p);
    }
  }
  y[me] = acc.get();
}

// out[c * R + r] = in[r * C + c]
template <typename T>
__global__ __launch_bounds__(256) void k_transpose(const T* __restrict__ in, T* __restrict__ out,
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

// y[r * C + c] += inT[c * R + r]
template <typename T>
__global__ __launch_bounds__(256) void k_add_transposed(const T* __restrict__ inT, T* __restrict__ y,
                                                        int64_t R, int64_t C) {
  __shared__ T tile[64][65];
  const int64_t tr = cdiv(R, 64);
  const int64_t bc = blockIdx.x / tr, br = blockIdx.x % tr;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int k = ty; k < 64; k += 4) {
    const int64_t c = bc * 64 + k, r = br * 64 + tx;
    if (r < R && c < C) tile[k][tx] = inT[c * R + r];
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const int64_t r = br * 64 + k, c = bc * 64 + tx;
    if (r < R && c < C) y[r * C + c] += tile[tx][k];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_apply_sign(const uint64_t* __restrict__ sa,
                                                    const uint64_t* __restrict__ sb, int64_t na,
                                                    int64_t nb, const T* __restrict__ x,
                                                    T* __restrict__ y) {
  const int64_t nt = cdiv(nb, 256);
  const int64_t ia = blockIdx.x / nt;
  const int64_t ib = (blockIdx.x % nt) * 256 + threadIdx.x;
  if (ib >= nb) return;
  const int s = det_sign(sa[ia], sb[ib]);
  const int64_t k = ia * nb + ib;
  y[k] = s < 0 ? -x[k] : x[k];
}

template <typename T>
void sigma_one(sbdx_ctx& c, const T* x, T* y, hipStream_t st, const T* vs_a, const T* vd_a,
               const T* vs_b, const T* vd_b, const T* V, const T* hii) {
  const int64_t na = c.na, nb = c.nb;
  const int64_t n = na * nb;
  RowTabs ta{c.A.off_s.get(), c.A.ket_s.get(), c.A.exc_s.get(), c.A.off_d.get(), c.A.ket_d.get()};
  RowTabs tb{c.B.off_s.get(), c.B.ket_s.get(), c.B.exc_s.get(), c.B.off_d.get(), c.B.ket_d.get()};
  Prof& pf = c.prof;
  auto mark = [&](int k) {
    if (pf.on) SBDX_HIP(hipEventRecord(pf.ev[k], st));
  };
  mark(0);
  // alpha rows: diagonal + alpha S/D + alpha-S x beta-S
  {
    const int64_t blocks = na * cdiv(nb, kTile);
    k_sigma_rows<T, true><<<static_cast<unsigned>(blocks), kTile, 0, st>>>(
        na, nb, c.L, ta, vs_a, vd_a, c.B.str.get(), V, x, y, hii, tb);
    SBDX_CHECK_LAUNCH();
  }
  mark(1);
  // beta rows on the transposed vector
  c.ws.xT.ensure(n * sizeof(T));
  c.ws.yT.ensure(n * sizeof(T));
  T* xT = reinterpret_cast<T*>(c.ws.xT.get());
  T* yT = reinterpret_cast<T*>(c.ws.yT.get());
  const unsigned tblocks = static_cast<unsigned>(cdiv(na, 64) * cdiv(nb, 64));
  k_transpose<T><<<tblocks, 256, 0, st>>>(x, xT, na, nb);
  SBDX_CHECK_LAUNCH();
  mark(2);
  {
    const int64_t blocks = nb * cdiv(na, kTile);
    k_sigma_rows<T, false><<<static_cast<unsigned>(blocks), kTile, 0, st>>>(
        nb, na, c.L, tb, vs_b, vd_b, c.A.str.get(), V, xT, yT, nullptr, ta);
    SBDX_CHECK_LAUNCH();
  }
  mark(3);
  k_add_transposed<T><<<tblocks, 256, 0, st>>>(yT, y, na, nb);
  SBDX_CHECK_LAUNCH();
  mark(4);
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

}  // namespace

void sigma_blocked(sbdx_ctx& c, int prec, int nvec, const void* x, void* y, hipStream_t st) {
  if (!c.have_ints) throw Error(-2, "integrals not set");
  const int64_t n = c.ndet();
  for (int v = 0; v < nvec; ++v) {
    if (prec == 0) {
      sigma_one<double>(c, static_cast<const double*>(x) + v * n, static_cast<double*>(y) + v * n, st,
                        c.A.val_s.get(), c.A.val_d.get(), c.B.val_s.get(), c.B.val_d.get(),
                        c.eri4.get(), c.hii.get());
    } else {
      ensure_fp32(c);
      sigma_one<float>(c, static_cast<const float*>(x) + v * n, static_cast<float*>(y) + v * n, st,
                       c.A.val_s32.get(), c.A.val_d32.get(), c.B.val_s32.get(), c.B.val_d32.get(),
                       c.eri4_32.get(), c.hii32.get());
    }
  }
}

void apply_sign(sbdx_ctx& c, int prec, int nvec, const void* x, void* y, hipStream_t st) {
  const int64_t n = c.ndet();
  const unsigned blocks = static_cast<unsigned>(c.na * cdiv(c.nb, 256));
  for (int v = 0; v < nvec; ++v) {
    if (prec == 0)
      k_apply_sign<double><<<blocks, 256, 0, st>>>(c.A.str.get(), c.B.str.get(), c.na, c.nb,
                                                   static_cast<const double*>(x) + v * n,
                                                   static_cast<double*>(y) + v * n);
    else
      k_apply_sign<float><<<blocks, 256, 0, st>>>(c.A.str.get(), c.B.str.get(), c.na, c.nb,
                                                  static_cast<const float*>(x) + v * n,
                                                  static_cast<float*>(y) + v * n);
    SBDX_CHECK_LAUNCH();
  }
}

}  // namespace sbdx
