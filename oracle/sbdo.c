/*
 * sbdo.c -- CPU oracle (test infrastructure only; see sbdo.h for scope and pinning).
 *
 * Restates r-ccs-cms/sbd, include/sbd/chemistry/{basic,tpb}/ and framework/dm_vector.h, at
 * one MPI rank.  Per determinant, sigma accumulates in the reference's order (diagonal,
 * alpha singles, alpha doubles, beta singles, beta doubles, alpha-single x beta-single;
 * mult.h:448-597), and every matrix element sums its terms in the reference's order, so with
 * the same compiler flags (no FMA contraction) sigma matches the reference bit for bit.
 */
#include "sbdo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- integrals */

static inline int64_t pair_index(int i, int j) {
  int a = i > j ? i : j, b = i > j ? j : i;
  return (int64_t)a * (a + 1) / 2 + b;
}

static inline int64_t packed_index(int I, int J, int K, int L) {
  int64_t ij = pair_index(I, J), kl = pair_index(K, L);
  int64_t a = ij > kl ? ij : kl, b = ij > kl ? kl : ij;
  return a * (a + 1) / 2 + b;
}

/* integrals.h:51-59: spin-orbital indices, zero unless spin(i)=spin(j), spin(k)=spin(l) */
double sbdo_value(const sbdo_ints *in, int i, int j, int k, int l) {
  if (!(((i & 1) == (j & 1)) && ((k & 1) == (l & 1)))) return 0.0;
  return in->I2[packed_index(i >> 1, j >> 1, k >> 1, l >> 1)];
}

/* makeintegrals.h:10-55 */
int sbdo_setup_integrals(int L, int64_t nent, const double *val, const int32_t *idx4,
                         double *I0, double *I1, double *I2, double *Dm, double *Xm) {
  const int n2 = 2 * L;
  const int64_t B = (int64_t)L * (L + 1) / 2;
  memset(I1, 0, sizeof(double) * (size_t)n2 * n2);
  memset(I2, 0, sizeof(double) * (size_t)(B * (B + 1) / 2));
  *I0 = 0.0;
  for (int64_t e = 0; e < nent; ++e) {
    int i = idx4[4 * e], j = idx4[4 * e + 1], k = idx4[4 * e + 2], l = idx4[4 * e + 3];
    double v = val[e];
    if (i == 0 && j == 0 && k == 0 && l == 0) {
      *I0 = v;
    } else if (k == l && k == 0) {
      int p = i - 1, q = j - 1;
      I1[(2 * p) * n2 + 2 * q] = v;
      I1[(2 * p + 1) * n2 + 2 * q + 1] = v;
      I1[(2 * q) * n2 + 2 * p] = v;
      I1[(2 * q + 1) * n2 + 2 * p + 1] = v;
    } else {
      if (i < 1 || j < 1 || k < 1 || l < 1 || i > L || j > L || k > L || l > L) return -1;
      I2[packed_index(i - 1, j - 1, k - 1, l - 1)] = v;
    }
  }
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) {
      Dm[i + L * j] = I2[packed_index(i, i, j, j)];
      Xm[i + L * j] = I2[packed_index(i, j, j, i)];
    }
  return 0;
}

/* ---------------------------------------------------------------- bits */

static inline int getbit(const uint64_t d[SBDO_DW], int x) { return (int)((d[x >> 6] >> (x & 63)) & 1u); }

/* determinants.h:98-119: alpha orbital p -> bit 2p, beta orbital p -> bit 2p+1 */
void sbdo_det_from_ab(const uint64_t *a, const uint64_t *b, int L, uint64_t det[SBDO_DW]) {
  for (int w = 0; w < SBDO_DW; ++w) det[w] = 0;
  for (int p = 0; p < L; ++p) {
    if ((a[p >> 6] >> (p & 63)) & 1u) det[(2 * p) >> 6] |= (uint64_t)1 << ((2 * p) & 63);
    if ((b[p >> 6] >> (p & 63)) & 1u) det[(2 * p + 1) >> 6] |= (uint64_t)1 << ((2 * p + 1) & 63);
  }
}

/* popcount of bits in [start, end) */
static inline int popcount_range(const uint64_t d[SBDO_DW], int start, int end) {
  int n = 0;
  for (int w = 0; w < SBDO_DW; ++w) {
    int lo = w * 64, hi = lo + 64;
    int s = start > lo ? start : lo, e = end < hi ? end : hi;
    if (s >= e) continue;
    uint64_t m = (e - lo == 64) ? ~(uint64_t)0 : (((uint64_t)1 << (e - lo)) - 1);
    m &= ~((((uint64_t)1) << (s - lo)) - 1);
    n += __builtin_popcountll(d[w] & m);
  }
  return n;
}

/* determinants.h:205-256: (-1)^{#set bits in [start,end)}, flipped again if bit start is set */
int sbdo_parity(const uint64_t det[SBDO_DW], int start, int end) {
  int s = (popcount_range(det, start, end) & 1) ? -1 : 1;
  if (getbit(det, start)) s = -s;
  return s;
}

/* ---------------------------------------------------------------- Slater-Condon */

/* determinants.h:374-398 */
double sbdo_zero_excite(const uint64_t det[SBDO_DW], const sbdo_ints *in) {
  const int n2 = 2 * in->L, L = in->L;
  int closed[256];
  int nc = 0;
  for (int w = 0; w < SBDO_DW; ++w) {
    uint64_t bits = det[w];
    while (bits) {
      int pos = __builtin_ctzll(bits);
      int x = w * 64 + pos;
      if (x < n2) closed[nc++] = x;
      bits &= bits - 1;
    }
  }
  double energy = 0.0;
  for (int i = 0; i < nc; ++i) {
    int I = closed[i];
    energy += in->I1[I * n2 + I];
    for (int j = i + 1; j < nc; ++j) {
      int J = closed[j];
      energy += in->Dm[(I / 2) + L * (J / 2)];
      if ((I % 2) == (J % 2)) energy -= in->Xm[(I / 2) + L * (J / 2)];
    }
  }
  return energy + in->I0;
}

/* determinants.h:400-422: ket = det, i occupied in ket only, a occupied in bra only */
static double one_excite(const uint64_t det[SBDO_DW], int i, int a, const sbdo_ints *in) {
  const int n2 = 2 * in->L;
  double sgn = (double)sbdo_parity(det, i < a ? i : a, i < a ? a : i);
  double energy = in->I1[a * n2 + i];
  for (int w = 0; w < SBDO_DW; ++w) {
    uint64_t bits = det[w];
    while (bits) {
      int pos = __builtin_ctzll(bits);
      int j = w * 64 + pos;
      energy += (sbdo_value(in, a, i, j, j) - sbdo_value(in, a, j, j, i));
      bits &= bits - 1;
    }
  }
  energy *= sgn;
  return energy;
}

/* determinants.h:424-442 */
static double two_excite(const uint64_t det[SBDO_DW], int i, int j, int a, int b, const sbdo_ints *in) {
  double sgn = 1.0;
  int I = i < j ? i : j, J = i < j ? j : i;
  int A = a < b ? a : b, B = a < b ? b : a;
  sgn *= (double)sbdo_parity(det, I < A ? I : A, I < A ? A : I);
  sgn *= (double)sbdo_parity(det, J < B ? J : B, J < B ? B : J);
  if (A > J || B < I) sgn *= -1.0;
  return sgn * (sbdo_value(in, A, I, B, J) - sbdo_value(in, A, J, B, I));
}

/* determinants.h:505-564: Hij(bra, ket) */
double sbdo_hij(const uint64_t bra[SBDO_DW], const uint64_t ket[SBDO_DW], const sbdo_ints *in) {
  int c[4], d[4];
  int nc = 0, nd = 0;
  for (int w = 0; w < SBDO_DW; ++w) {
    uint64_t dc = bra[w] & ~ket[w], dd = ket[w] & ~bra[w];
    while (dc) {
      int pos = __builtin_ctzll(dc);
      if (nc < 4) c[nc] = w * 64 + pos;
      nc++;
      dc &= dc - 1;
    }
    while (dd) {
      int pos = __builtin_ctzll(dd);
      if (nd < 4) d[nd] = w * 64 + pos;
      nd++;
      dd &= dd - 1;
    }
  }
  if (nc == 0) return sbdo_zero_excite(ket, in);
  if (nc == 1) return one_excite(ket, d[0], c[0], in);
  if (nc == 2) return two_excite(ket, d[0], d[1], c[0], c[1], in);
  return 0.0;
}

/* ---------------------------------------------------------------- connectivity */

/* orbitals (below L) in which two strings of W words differ */
static inline int string_difference(const uint64_t *x, const uint64_t *y, int W, int L) {
  int d = 0;
  for (int w = 0; w < W; ++w) {
    const int bits = L - 64 * w;
    const uint64_t m = bits >= 64 ? ~(uint64_t)0 : (bits <= 0 ? 0 : (((uint64_t)1 << bits) - 1));
    d += __builtin_popcountll((x[w] ^ y[w]) & m);
  }
  return d;
}

/* helper.h:132-182 */
void sbdo_count_excitations(const uint64_t *s, int64_t n, int W, int L, int64_t *nsingle,
                            int64_t *ndouble) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    int64_t ns = 0, nd = 0;
    for (int64_t j = 0; j < n; ++j) {
      int d = string_difference(s + i * W, s + j * W, W, L);
      if (d == 2) ns++;
      else if (d == 4) nd++;
    }
    nsingle[i] = ns;
    ndouble[i] = nd;
  }
}

void sbdo_fill_excitations(const uint64_t *s, int64_t n, int W, int L, const int64_t *off_s,
                           int64_t *idx_s, const int64_t *off_d, int64_t *idx_d) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    int64_t ps = off_s[i], pd = off_d[i];
    for (int64_t j = 0; j < n; ++j) {
      int d = string_difference(s + i * W, s + j * W, W, L);
      if (d == 2) idx_s[ps++] = j;
      else if (d == 4) idx_d[pd++] = j;
    }
  }
}

/* ---------------------------------------------------------------- H_ii, sigma */

/* qcham.h:314-366 */
void sbdo_hdiag(const sbdo_basis *bs, const sbdo_ints *in, double *hii) {
  const int64_t na = bs->na, nb = bs->nb;
#pragma omp parallel for schedule(static)
  for (int64_t ia = 0; ia < na; ++ia) {
    uint64_t det[SBDO_DW];
    for (int64_t ib = 0; ib < nb; ++ib) {
      sbdo_det_from_ab(bs->a + ia * bs->W, bs->b + ib * bs->W, bs->L, det);
      hii[ia * nb + ib] = sbdo_zero_excite(det, in);
    }
  }
}

/* mult.h:387-648 at one rank: tasks type 2 (alpha S+D), 1 (beta S+D), 0 (alpha S x beta S),
 * preceded by the diagonal; each bra element is owned by one thread. */
void sbdo_mult(const sbdo_basis *bs, const sbdo_ints *in, const double *hii,
               const double *Wk, double *Wb) {
  const int64_t na = bs->na, nb = bs->nb;
  const int L = bs->L;
  const sbdo_conn *A = &bs->ca, *Bc = &bs->cb;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t ia = 0; ia < na; ++ia) {
    uint64_t deti[SBDO_DW], detj[SBDO_DW];
    for (int64_t ib = 0; ib < nb; ++ib) {
      const int64_t bra = ia * nb + ib;
      double acc = Wb[bra];
      acc += hii[bra] * Wk[bra];
      sbdo_det_from_ab(bs->a + ia * bs->W, bs->b + ib * bs->W, L, deti);
      /* type 2: alpha singles then alpha doubles (mult.h:498-532) */
      for (int64_t k = A->off_s[ia]; k < A->off_s[ia + 1]; ++k) {
        int64_t ja = A->idx_s[k];
        sbdo_det_from_ab(bs->a + ja * bs->W, bs->b + ib * bs->W, L, detj);
        acc += sbdo_hij(deti, detj, in) * Wk[ja * nb + ib];
      }
      for (int64_t k = A->off_d[ia]; k < A->off_d[ia + 1]; ++k) {
        int64_t ja = A->idx_d[k];
        sbdo_det_from_ab(bs->a + ja * bs->W, bs->b + ib * bs->W, L, detj);
        acc += sbdo_hij(deti, detj, in) * Wk[ja * nb + ib];
      }
      /* type 1: beta singles then beta doubles (mult.h:534-566) */
      for (int64_t k = Bc->off_s[ib]; k < Bc->off_s[ib + 1]; ++k) {
        int64_t jb = Bc->idx_s[k];
        sbdo_det_from_ab(bs->a + ia * bs->W, bs->b + jb * bs->W, L, detj);
        acc += sbdo_hij(deti, detj, in) * Wk[ia * nb + jb];
      }
      for (int64_t k = Bc->off_d[ib]; k < Bc->off_d[ib + 1]; ++k) {
        int64_t jb = Bc->idx_d[k];
        sbdo_det_from_ab(bs->a + ia * bs->W, bs->b + jb * bs->W, L, detj);
        acc += sbdo_hij(deti, detj, in) * Wk[ia * nb + jb];
      }
      /* type 0: alpha single x beta single (mult.h:569-596) */
      for (int64_t k = A->off_s[ia]; k < A->off_s[ia + 1]; ++k) {
        int64_t ja = A->idx_s[k];
        for (int64_t m = Bc->off_s[ib]; m < Bc->off_s[ib + 1]; ++m) {
          int64_t jb = Bc->idx_s[m];
          sbdo_det_from_ab(bs->a + ja * bs->W, bs->b + jb * bs->W, L, detj);
          acc += sbdo_hij(deti, detj, in) * Wk[ja * nb + jb];
        }
      }
      Wb[bra] = acc;
    }
  }
}

/* ---------------------------------------------------------------- vectors */

double sbdo_inner(const double *x, const double *y, int64_t n) {
  double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
  for (int64_t i = 0; i < n; ++i) sum += x[i] * y[i];
  return sum;
}

/* dm_vector.h:32-50: returns the norm before scaling; scales by the reciprocal */
double sbdo_normalize(double *x, int64_t n) {
  double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
  for (int64_t i = 0; i < n; ++i) sum += x[i] * x[i];
  double res = sqrt(sum);
  double factor = 1.0 / res;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) x[i] *= factor;
  return res;
}

/* Cyclic Jacobi rotations to machine precision, then ascending sort. */
void sbdo_syev(int n, double *a, int lda, double *w) {
  double *v = (double *)calloc((size_t)n * n, sizeof(double));
  double *m = (double *)malloc(sizeof(double) * (size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) m[i + n * j] = a[(i <= j ? i : j) + lda * (i <= j ? j : i)];
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
        double apq = m[p + n * q];
        if (apq == 0.0) continue;
        double app = m[p + n * p], aqq = m[q + n * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double mkp = m[k + n * p], mkq = m[k + n * q];
          m[k + n * p] = c * mkp - s * mkq;
          m[k + n * q] = s * mkp + c * mkq;
        }
        for (int k = 0; k < n; ++k) {
          double mpk = m[p + n * k], mqk = m[q + n * k];
          m[p + n * k] = c * mpk - s * mqk;
          m[q + n * k] = s * mpk + c * mqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = v[k + n * p], vkq = v[k + n * q];
          v[k + n * p] = c * vkp - s * vkq;
          v[k + n * q] = s * vkp + c * vkq;
        }
      }
  }
  int *ord = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  for (int i = 1; i < n; ++i) {
    int t = ord[i], k = i;
    while (k > 0 && m[ord[k - 1] * (n + 1)] > m[t * (n + 1)]) { ord[k] = ord[k - 1]; --k; }
    ord[k] = t;
  }
  for (int j = 0; j < n; ++j) {
    int src = ord[j];
    w[j] = m[src * (n + 1)];
    /* sign convention: largest-magnitude component positive */
    int imax = 0;
    for (int i = 1; i < n; ++i)
      if (fabs(v[i + n * src]) > fabs(v[imax + n * src])) imax = i;
    double sg = v[imax + n * src] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) a[i + lda * j] = sg * v[i + n * src];
  }
  free(ord);
  free(m);
  free(v);
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* davidson.h:661-923 with h/b/t communicators of size 1 */
int sbdo_davidson(const sbdo_basis *bs, const sbdo_ints *in, const double *hii, double *W,
                  int max_it, int nb, double eps, double max_time, double *trace,
                  double *resid, int trace_cap) {
  const int64_t n = bs->na * bs->nb;
  const double eps_reg = 1.0e-12;
  double **C = (double **)malloc(sizeof(double *) * nb);
  double **HC = (double **)malloc(sizeof(double *) * nb);
  for (int k = 0; k < nb; ++k) {
    C[k] = (double *)malloc(sizeof(double) * n);
    HC[k] = (double *)malloc(sizeof(double) * n);
    memcpy(C[k], W, sizeof(double) * n);
    memcpy(HC[k], W, sizeof(double) * n);
  }
  double *R = (double *)malloc(sizeof(double) * n);
  const double *dii = hii;
  double *H = (double *)calloc((size_t)nb * nb, sizeof(double));
  double *U = (double *)calloc((size_t)nb * nb, sizeof(double));
  double *E = (double *)malloc(sizeof(double) * nb);
  double *steps = (double *)calloc((size_t)nb * (max_it > 0 ? max_it : 1), sizeof(double));
  int do_continue = 1, nstep = 0;
  double t0 = now_seconds();

  for (int it = 0; it < max_it; ++it) {
    memcpy(C[0], W, sizeof(double) * n);
    for (int ib = 0; ib < nb; ++ib) {
      double ts = now_seconds();
      memset(HC[ib], 0, sizeof(double) * n);
      sbdo_mult(bs, in, hii, C[ib], HC[ib]);
      for (int jb = 0; jb <= ib; ++jb) {
        H[jb + nb * ib] = sbdo_inner(C[jb], HC[ib], n);
        H[ib + nb * jb] = H[jb + nb * ib];
      }
      for (int jb = 0; jb <= ib; ++jb)
        for (int kb = 0; kb <= ib; ++kb) U[jb + nb * kb] = H[jb + nb * kb];
      sbdo_syev(ib + 1, U, nb, E);

      double x = U[0];
#pragma omp parallel for schedule(static)
      for (int64_t is = 0; is < n; ++is) W[is] = C[0][is] * x;
      x = -1.0 * U[0];
#pragma omp parallel for schedule(static)
      for (int64_t is = 0; is < n; ++is) R[is] = HC[0][is] * x;
      for (int kb = 1; kb <= ib; ++kb) {
        x = U[kb];
#pragma omp parallel for schedule(static)
        for (int64_t is = 0; is < n; ++is) W[is] += C[kb][is] * x;
        x = -1.0 * U[kb];
#pragma omp parallel for schedule(static)
        for (int64_t is = 0; is < n; ++is) R[is] += HC[kb][is] * x;
      }
#pragma omp parallel for schedule(static)
      for (int64_t is = 0; is < n; ++is) R[is] += E[0] * W[is];

      sbdo_normalize(W, n);
      double norm_R = sbdo_normalize(R, n);
      if (nstep < trace_cap) {
        trace[nstep] = E[0];
        resid[nstep] = norm_R;
      }
      nstep++;
      if (norm_R < eps) {
        do_continue = 0;
        break;
      }
      if (ib < nb - 1) {
        double *c = C[ib + 1];
        const double e0 = E[0];
#pragma omp parallel for schedule(static)
        for (int64_t is = 0; is < n; ++is) {
          if (fabs(e0 - dii[is]) > eps_reg) c[is] = R[is] / (e0 - dii[is]);
          else c[is] = R[is] / (e0 - dii[is] - eps_reg);
        }
        for (int kb = 0; kb < ib + 1; ++kb) {
          double olap = sbdo_inner(C[kb], c, n);
          olap *= -1.0;
#pragma omp parallel for schedule(static)
          for (int64_t is = 0; is < n; ++is) c[is] += C[kb][is] * olap;
        }
        sbdo_normalize(c, n);
      }
      double te = now_seconds();
      steps[it * nb + ib] = te - ts;
      double ave = 0.0;
      for (int ks = 0; ks <= it * nb + ib; ++ks) ave += steps[ks];
      ave /= (it * nb + ib + 1);
      if ((te - t0) + ave > max_time) {
        do_continue = 0;
        break;
      }
    }
    if (!do_continue) break;
    memcpy(C[0], W, sizeof(double) * n);
  }
  for (int k = 0; k < nb; ++k) {
    free(C[k]);
    free(HC[k]);
  }
  free(C);
  free(HC);
  free(R);
  free(H);
  free(U);
  free(E);
  free(steps);
  return nstep;
}

/* ---------------------------------------------------------------- RDMs */

typedef struct {
  int L;
  double *one; /* 2 L^2 */
  double *two; /* 4 L^4 */
} rdm_acc;

/* correlation.h:15-44 */
static void zero_diff_corr(const uint64_t det[SBDO_DW], double wi, rdm_acc *r) {
  const int L = r->L, n2 = 2 * L;
  const int64_t L2 = (int64_t)L * L, L3 = L2 * L, L4 = L3 * L;
  int closed[256];
  int nc = 0;
  for (int w = 0; w < SBDO_DW; ++w) {
    uint64_t bits = det[w];
    while (bits) {
      int pos = __builtin_ctzll(bits);
      int x = w * 64 + pos;
      if (x < n2) closed[nc++] = x;
      bits &= bits - 1;
    }
  }
  double ww = wi * wi;
  for (int i = 0; i < nc; ++i) {
    int oi = closed[i] / 2, si = closed[i] % 2;
    r->one[si * L2 + oi + L * oi] += ww;
    for (int j = i + 1; j < nc; ++j) {
      int oj = closed[j] / 2, sj = closed[j] % 2;
      r->two[(si + 2 * sj) * L4 + oi + L * oj + L2 * oi + L3 * oj] += ww;
      r->two[(sj + 2 * si) * L4 + oj + L * oi + L2 * oj + L3 * oi] += ww;
      if (si == sj) {
        r->two[(si + 2 * sj) * L4 + oi + L * oj + L2 * oj + L3 * oi] += -ww;
        r->two[(sj + 2 * si) * L4 + oj + L * oi + L2 * oi + L3 * oj] += -ww;
      }
    }
  }
}

/* correlation.h:49-86: i = bra-only spin-orbital, a = ket-only */
static void one_diff_corr(const uint64_t det[SBDO_DW], double wi, double wj, int i, int a,
                          rdm_acc *r) {
  const int L = r->L;
  const int64_t L2 = (int64_t)L * L, L3 = L2 * L, L4 = L3 * L;
  double sgn = (double)sbdo_parity(det, i < a ? i : a, i < a ? a : i);
  int oi = i / 2, si = i % 2, oa = a / 2;
  r->one[si * L2 + oi + L * oa] += wi * wj * sgn;
  for (int w = 0; w < SBDO_DW; ++w) {
    uint64_t bits = det[w];
    while (bits) {
      int pos = __builtin_ctzll(bits);
      int soj = w * 64 + pos;
      int oj = soj / 2, sj = soj % 2;
      r->two[(si + 2 * sj) * L4 + oa + oj * L + oi * L2 + oj * L3] += wi * wj * sgn;
      r->two[(sj + 2 * si) * L4 + oj + oa * L + oj * L2 + oi * L3] += wi * wj * sgn;
      if (si == sj) {
        r->two[(si + 2 * sj) * L4 + oa + oj * L + oj * L2 + oi * L3] += wi * wj * (-sgn);
        r->two[(sj + 2 * si) * L4 + oj + oa * L + oi * L2 + oj * L3] += wi * wj * (-sgn);
      }
      bits &= bits - 1;
    }
  }
}

/* correlation.h:91-130 */
static void two_diff_corr(const uint64_t det[SBDO_DW], double wi, double wj, int i, int j, int a,
                          int b, rdm_acc *r) {
  const int L = r->L;
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2;
  double sgn = 1.0;
  int I = i < j ? i : j, J = i < j ? j : i;
  int A = a < b ? a : b, B = a < b ? b : a;
  sgn *= (double)sbdo_parity(det, I < A ? I : A, I < A ? A : I);
  sgn *= (double)sbdo_parity(det, J < B ? J : B, J < B ? B : J);
  if (A > J || B < I) sgn *= -1.0;
  int oi = I / 2, si = I % 2, oa = A / 2, sa = A % 2;
  int oj = J / 2, sj = J % 2, ob = B / 2, sb = B % 2;
  (void)sa;
  if (si == sa) {
    r->two[(si + 2 * sj) * L4 + oa + L * ob + L2 * (oi + L * oj)] += sgn * wi * wj;
    r->two[(sj + 2 * si) * L4 + ob + L * oa + L2 * (oj + L * oi)] += sgn * wi * wj;
  }
  if (si == sb) {
    r->two[(si + 2 * sj) * L4 + oa + L * ob + L2 * (oj + L * oi)] += (-sgn) * wi * wj;
    r->two[(sj + 2 * si) * L4 + ob + L * oa + L2 * (oi + L * oj)] += (-sgn) * wi * wj;
  }
}

/* correlation.h:135-189: DetI = bra, DetJ = ket */
static void corr_term(const uint64_t deti[SBDO_DW], const uint64_t detj[SBDO_DW], double wi, double wj,
                      rdm_acc *r) {
  int c[4], d[4];
  int nc = 0, nd = 0;
  for (int w = 0; w < SBDO_DW; ++w) {
    uint64_t dc = deti[w] & ~detj[w], dd = detj[w] & ~deti[w];
    while (dc) {
      if (nc < 4) c[nc] = w * 64 + __builtin_ctzll(dc);
      nc++;
      dc &= dc - 1;
    }
    while (dd) {
      if (nd < 4) d[nd] = w * 64 + __builtin_ctzll(dd);
      nd++;
      dd &= dd - 1;
    }
  }
  if (nc == 0) zero_diff_corr(deti, wi, r);
  else if (nc == 1) one_diff_corr(deti, wi, wj, c[0], d[0], r);
  else if (nc == 2) two_diff_corr(deti, wi, wj, c[0], c[1], d[0], d[1], r);
}

/* correlation.h:194-443 */
void sbdo_correlation(const sbdo_basis *bs, const double *W, double *onebody,
                      double *twobody) {
  const int L = bs->L;
  const int64_t na = bs->na, nb = bs->nb;
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2;
  memset(onebody, 0, sizeof(double) * 2 * L2);
  memset(twobody, 0, sizeof(double) * 4 * L4);
  rdm_acc main_acc = {L, onebody, twobody};
  for (int64_t ia = 0; ia < na; ++ia)
    for (int64_t ib = 0; ib < nb; ++ib) {
      uint64_t det[SBDO_DW];
      sbdo_det_from_ab(bs->a + ia * bs->W, bs->b + ib * bs->W, L, det);
      zero_diff_corr(det, W[ia * nb + ib], &main_acc);
    }
  const sbdo_conn *A = &bs->ca, *Bc = &bs->cb;
#pragma omp parallel
  {
    rdm_acc loc;
    loc.L = L;
    loc.one = (double *)calloc((size_t)(2 * L2), sizeof(double));
    loc.two = (double *)calloc((size_t)(4 * L4), sizeof(double));
#pragma omp for schedule(dynamic, 1)
    for (int64_t ia = 0; ia < na; ++ia) {
      uint64_t deti[SBDO_DW], detj[SBDO_DW];
      for (int64_t ib = 0; ib < nb; ++ib) {
        double wi = W[ia * nb + ib];
        sbdo_det_from_ab(bs->a + ia * bs->W, bs->b + ib * bs->W, L, deti);
        for (int64_t k = A->off_s[ia]; k < A->off_s[ia + 1]; ++k) {
          int64_t ja = A->idx_s[k];
          sbdo_det_from_ab(bs->a + ja * bs->W, bs->b + ib * bs->W, L, detj);
          corr_term(deti, detj, wi, W[ja * nb + ib], &loc);
        }
        for (int64_t k = A->off_d[ia]; k < A->off_d[ia + 1]; ++k) {
          int64_t ja = A->idx_d[k];
          sbdo_det_from_ab(bs->a + ja * bs->W, bs->b + ib * bs->W, L, detj);
          corr_term(deti, detj, wi, W[ja * nb + ib], &loc);
        }
        for (int64_t k = Bc->off_s[ib]; k < Bc->off_s[ib + 1]; ++k) {
          int64_t jb = Bc->idx_s[k];
          sbdo_det_from_ab(bs->a + ia * bs->W, bs->b + jb * bs->W, L, detj);
          corr_term(deti, detj, wi, W[ia * nb + jb], &loc);
        }
        for (int64_t k = Bc->off_d[ib]; k < Bc->off_d[ib + 1]; ++k) {
          int64_t jb = Bc->idx_d[k];
          sbdo_det_from_ab(bs->a + ia * bs->W, bs->b + jb * bs->W, L, detj);
          corr_term(deti, detj, wi, W[ia * nb + jb], &loc);
        }
        for (int64_t k = A->off_s[ia]; k < A->off_s[ia + 1]; ++k) {
          int64_t ja = A->idx_s[k];
          for (int64_t m = Bc->off_s[ib]; m < Bc->off_s[ib + 1]; ++m) {
            int64_t jb = Bc->idx_s[m];
            sbdo_det_from_ab(bs->a + ja * bs->W, bs->b + jb * bs->W, L, detj);
            corr_term(deti, detj, wi, W[ja * nb + jb], &loc);
          }
        }
      }
    }
#pragma omp critical
    {
      for (int64_t i = 0; i < 2 * L2; ++i) onebody[i] += loc.one[i];
      for (int64_t i = 0; i < 4 * L4; ++i) twobody[i] += loc.two[i];
    }
    free(loc.one);
    free(loc.two);
  }
}

/* occupation.h:10-95 with the full orbital range on one rank */
void sbdo_occupation(const sbdo_basis *bs, const double *W, double *density) {
  const int L = bs->L;
  const int64_t na = bs->na, nb = bs->nb;
  for (int i = 0; i < 2 * L; ++i) density[i] = 0.0;
  for (int64_t ia = 0; ia < na; ++ia)
    for (int64_t ib = 0; ib < nb; ++ib) {
      double w2 = W[ia * nb + ib] * W[ia * nb + ib];
      for (int io = 0; io < L; ++io) {
        density[2 * io] += w2 * (double)((bs->a[ia * bs->W + (io >> 6)] >> (io & 63)) & 1u);
        density[2 * io + 1] += w2 * (double)((bs->b[ib * bs->W + (io >> 6)] >> (io & 63)) & 1u);
      }
    }
}

void sbdo_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int sbdo_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
