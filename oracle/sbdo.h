/*
 * sbdo -- CPU oracle for the selected-basis diagonalization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C (OpenMP) restatement of r-ccs-cms/sbd's
 * tensor-product-basis path, used to check the HIP engine (tests, smoke check) and as the
 * "port" CPU baseline in the benchmark.  Nothing in the product links, loads or calls it.
 *
 * Storage width: the reference packs `bit_length` (default 20) bits per size_t word.  The
 * oracle packs 64 bits per word (strings: W = 1 uint64 for L <= 64, W = 2 for L <= 128, least
 * significant word first; interleaved determinants: SBDO_DW = 4 uint64).  Every function below
 * depends only on bit positions, never on the word width, so the semantics are the
 * reference's.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   setup_integrals  -- compiled reference fixtures; 8-fold symmetry identities
 *   parity / hij     -- brute-force second quantization on small systems (exact signs)
 *   excitations      -- closed form nA(L-nA), C(nA,2)C(L-nA,2) on full spaces; brute force
 *   hdiag            -- closed form; compiled reference fixtures
 *   mult             -- dense-H brute force on tiny systems; compiled reference (bit-exact)
 *   davidson         -- scipy eigvalsh on the dense H; compiled reference energy trace
 *   correlation      -- RDM sum rules, energy identity; compiled reference fixtures
 *   occupation       -- equals diag of the 1-RDM; trace = electron count
 */
#ifndef SBDO_H
#define SBDO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBDO_DW 4      /* uint64 words of an interleaved determinant (2L <= 256 bits) */

typedef struct {
  int L;               /* spatial orbitals (<= 128) */
  double I0;           /* core energy */
  const double *I1;    /* (2L)x(2L) spin-orbital one-electron matrix, I1[i*2L+j] */
  const double *I2;    /* packed spatial (pq|rs), reference twoInt::store layout */
  const double *Dm;    /* Direct(i,j) = (ii|jj), Dm[i+L*j] */
  const double *Xm;    /* Exchange(i,j) = (ij|ji), Xm[i+L*j] */
} sbdo_ints;

/* CSR excitation lists for one spin: row x -> kets in ascending order */
typedef struct {
  const int64_t *off_s, *idx_s;   /* singles: popcount(x ^ x') == 2 */
  const int64_t *off_d, *idx_d;   /* doubles: popcount(x ^ x') == 4 */
} sbdo_conn;

typedef struct {
  int L;
  int64_t na, nb;
  const uint64_t *a, *b;   /* na (nb) strings of W words each */
  sbdo_conn ca, cb;
  int W;
} sbdo_basis;

/* chemistry/basic/makeintegrals.h:10-55 (caller sizes: I1 4L^2, I2 B(B+1)/2 with
 * B=L(L+1)/2, Dm/Xm L^2).  idx4 holds (i,j,k,l) 1-based per entry. */
int sbdo_setup_integrals(int L, int64_t nent, const double *val, const int32_t *idx4,
                         double *I0, double *I1, double *I2, double *Dm, double *Xm);

/* chemistry/basic/integrals.h:51-59 */
double sbdo_value(const sbdo_ints *in, int i, int j, int k, int l);

/* chemistry/basic/determinants.h:98-119; a, b: one string each (ceil(L/64) words) */
void sbdo_det_from_ab(const uint64_t *a, const uint64_t *b, int L, uint64_t det[SBDO_DW]);

/* chemistry/basic/determinants.h:205-256: returns the +-1 factor */
int sbdo_parity(const uint64_t det[SBDO_DW], int start, int end);

/* chemistry/basic/determinants.h:374-398, :505-564 */
double sbdo_zero_excite(const uint64_t det[SBDO_DW], const sbdo_ints *in);
double sbdo_hij(const uint64_t bra[SBDO_DW], const uint64_t ket[SBDO_DW], const sbdo_ints *in);

/* chemistry/tpb/helper.h:132-182 (GenerateExcitation).  Pass 1 counts, pass 2 fills.
 * s: n strings of W words each. */
void sbdo_count_excitations(const uint64_t *s, int64_t n, int W, int L, int64_t *nsingle,
                            int64_t *ndouble);
void sbdo_fill_excitations(const uint64_t *s, int64_t n, int W, int L, const int64_t *off_s,
                           int64_t *idx_s, const int64_t *off_d, int64_t *idx_d);

/* chemistry/tpb/qcham.h:314-366 */
void sbdo_hdiag(const sbdo_basis *bs, const sbdo_ints *in, double *hii);

/* chemistry/tpb/mult.h:387-648: Wb += H Wk (accumulates, like the reference) */
void sbdo_mult(const sbdo_basis *bs, const sbdo_ints *in, const double *hii,
               const double *Wk, double *Wb);

/* framework/dm_vector.h:18-50 */
double sbdo_inner(const double *x, const double *y, int64_t n);
double sbdo_normalize(double *x, int64_t n);

/* Symmetric eigensolver for the Davidson subspace (stands in for LAPACK dsyev 'V','U',
 * framework/hp_numeric.h:500-516): n x n column-major, leading dimension lda; on exit
 * a holds eigenvectors in columns, w ascending eigenvalues. */
void sbdo_syev(int n, double *a, int lda, double *w);

/* chemistry/tpb/davidson.h:661-923.  W: initial vector in, Ritz vector out.
 * trace[k] = lowest Ritz value at step k (k < trace_cap), resid[k] = ||r|| at step k.
 * Returns the number of Davidson steps taken (== sigma builds). */
int sbdo_davidson(const sbdo_basis *bs, const sbdo_ints *in, const double *hii, double *W,
                  int max_it, int num_block, double eps, double max_time, double *trace,
                  double *resid, int trace_cap);

/* chemistry/tpb/correlation.h:194-443: onebody 2*L^2, twobody 4*L^4 (overwritten) */
void sbdo_correlation(const sbdo_basis *bs, const double *W, double *onebody,
                      double *twobody);

/* chemistry/tpb/occupation.h:10-95 at one rank: density[2p+s] */
void sbdo_occupation(const sbdo_basis *bs, const double *W, double *density);

void sbdo_set_threads(int n);
int sbdo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
