/*
 * sbdx.h -- C ABI of the MI355X selected-basis diagonalization engine (libsbdx.so).
 *
 * One context holds one selected-basis problem on one GPU: the alpha and beta string tables,
 * their single/double excitation lists, and the integrals.  Every entry point returns 0 on
 * success and a negative sbdx_status on failure (message: sbdx_last_error()); no C++
 * exception crosses this boundary.  Host buffers are owned by the caller, device buffers by
 * the context.  Calls on one context must not overlap (one context per host thread/stream).
 *
 * Vectors are determinant-indexed, alpha-major: W[ia * n_beta + ib] (the reference's layout,
 * chemistry/tpb/mult.h:503-504, qcham.h:359).  Strings are uint64 bit masks, bit p = orbital
 * p, i.e. the reference's `bit_length`-bit words concatenated (framework/bit_manipulation.h:
 * 863-893 change_bitlength to 64): one word each for norb <= 64 (sbdx_create), `words` words
 * each, least significant first, for norb <= 128 (sbdx_create_words).
 *
 * Entry points and the reference interface each one stands in for (file:line in r-ccs-cms/sbd
 * include/sbd/):
 *   sbdx_create            MakeHelpers / RemakeHelpers    chemistry/tpb/helper.h:432, :650
 *   sbdx_create_words      the same, strings of any bit_length up to 128 orbitals
 *   sbdx_set_integrals     SetupIntegrals (consumer side) chemistry/basic/makeintegrals.h:10
 *   sbdx_get_excitations   TaskHelpers lists              chemistry/tpb/helper.h:12-36, :132-182
 *   sbdx_hdiag             makeQChamDiagTerms             chemistry/tpb/qcham.h:314-366
 *   sbdx_sigma             mult (on-the-fly)              chemistry/tpb/mult.h:387-648
 *   sbdx_davidson          Davidson (+ max_time)          chemistry/tpb/davidson.h:661-923
 *   sbdx_rdm               Correlation                    chemistry/tpb/correlation.h:194-443
 *   sbdx_occupation        OccupationDensity              chemistry/tpb/occupation.h:10-95
 *   sbdx_destroy           FreeHelpers                    chemistry/tpb/helper.h:339-350
 */
#ifndef SBDX_H
#define SBDX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SBDX_OK = 0,
  SBDX_ERR_ARG = -1,       /* invalid argument (sizes, null pointers, norb > 128, ...) */
  SBDX_ERR_STATE = -2,     /* call out of order (e.g. sigma before set_integrals) */
  SBDX_ERR_DEVICE = -3,    /* HIP runtime error (allocation, launch) */
  SBDX_ERR_NOCONV = -4     /* not an error of the call: Davidson stopped before tolerance */
} sbdx_status;

/* Arithmetic of a sigma build. */
enum { SBDX_FP64 = 0, SBDX_FP32 = 1 };

/* Accumulation of FP32 sigma builds (FP32 integrals and vectors either way).
 * SBDX_ACC_FP64: each product of two FP32 numbers is exact in FP64 and is added unrounded into
 *   an FP64 accumulator, rounded to FP32 once per output (the default; fewer instructions).
 * SBDX_ACC_TWOSUM: FP32 accumulation with compensated partials: runs of up to 16 terms are
 *   summed by FP32 FMA, and each run is added to the running sum by TwoSum, its rounding
 *   errors gathered in a compensation term. */
enum { SBDX_ACC_FP64 = 0, SBDX_ACC_TWOSUM = 1 };

typedef struct sbdx_ctx sbdx_ctx;

/* Create a context on the current HIP device.  alpha/beta: n_alpha/n_beta uint64 strings (any
 * order; each string's index is its row/column in W).  Builds the excitation lists on the GPU. */
int sbdx_create(int norb, int64_t n_alpha, const uint64_t *alpha, int64_t n_beta,
                const uint64_t *beta, sbdx_ctx **out);

/* sbdx_create for strings of `words` (1 or 2) uint64 words each, least significant word first
 * (orbital p = bit p % 64 of word p / 64), norb <= 64 words <= 128; bits at or above norb must
 * be clear.  Above 64 orbitals the mixed pass runs its one-stage variants: 1024 threads stage
 * each integral row of L^2 values in LDS where two rows fit (FP64 up to ~99 orbitals, FP32 up
 * to 128), 256 threads read it from global memory; the learned preconditioner is not available
 * (its inputs hold 64 orbitals per spin). */
int sbdx_create_words(int norb, int words, int64_t n_alpha, const uint64_t *alpha, int64_t n_beta,
                      const uint64_t *beta, sbdx_ctx **out);

/* Integrals in the reference's storage: I1 = oneInt::store, (2 norb)^2 spin-orbital entries,
 * I1[i*2L+j] (even index = alpha, odd = beta); I2 = twoInt::store, packed spatial (pq|rs) of
 * n_I2 = B(B+1)/2 entries with B = norb(norb+1)/2 (chemistry/basic/integrals.h:10-67). */
int sbdx_set_integrals(sbdx_ctx *ctx, double I0, const double *I1, const double *I2,
                       int64_t n_I2);

/* Sizes of the excitation lists of one spin (0 = alpha, 1 = beta). */
int sbdx_excitation_counts(sbdx_ctx *ctx, int spin, int64_t *n_single, int64_t *n_double);

/* Copy one spin's lists to host CSR: off_* have n+1 entries; idx_* hold ket string indices in
 * ascending order per bra string (the reference's SinglesFrom* / DoublesFrom* lists). */
int sbdx_get_excitations(sbdx_ctx *ctx, int spin, int64_t *off_s, int64_t *idx_s,
                         int64_t *off_d, int64_t *idx_d);

int64_t sbdx_ndet(const sbdx_ctx *ctx);

/* H_ii for every determinant (host buffer of sbdx_ndet entries). */
int sbdx_hdiag(sbdx_ctx *ctx, double *hii);

/* S = H C for nvec vectors (host buffers, nvec * ndet doubles each, vector-major).  Unlike
 * the reference's mult the output is overwritten, not accumulated. */
int sbdx_sigma(sbdx_ctx *ctx, int nvec, const double *C, double *S);

/* Device-resident variant: dC, dS are device pointers of nvec*ndet elements of the type given
 * by prec (SBDX_FP64: double, SBDX_FP32: float); stream may be NULL (context stream). */
int sbdx_sigma_device(sbdx_ctx *ctx, int prec, int nvec, const void *dC, void *dS,
                      void *stream);

/* No counterpart in the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): sbdx_sigma_device for one
 * vector (nvec = 1), enqueued on `stream` (the context's stream if NULL) without allocating or
 * synchronising.  Its scratch is a fixed per-device array of 64 MiB: ndet may be at most
 * 2,097,152 in FP64 (4,194,304 in FP32), larger problems get SBDX_ERR_ARG.  Needs the tables of
 * its precision built beforehand: FP64 after sbdx_set_integrals; FP32 after one sbdx_sigma_device
 * call with SBDX_FP32 (SBDX_ERR_STATE otherwise).  Not with a stored Hamiltonian.  Calls that
 * overlap on one device share the scratch and race. */
int sbdx_sigma_device_noalloc(sbdx_ctx *ctx, int prec, const void *dC, void *dS, void *stream);

/* Fused sigma of a block of k vectors (k = 1, 2, 4 or 8) stored interleaved: element
 * [det * k + v] of dC / dS (device pointers, type by prec).  One pass over the excitation lists
 * serves all k vectors (mult, chemistry/tpb/mult.h:387-648, applied to k vectors at once). */
int sbdx_sigma_block_device(sbdx_ctx *ctx, int prec, int k, const void *dC, void *dS,
                            void *stream);

/* Stored Hamiltonian (the reference's method 1: makeQCham, chemistry/tpb/qcham.h:10-312, with
 * its stored-H mult, mult.h:13-123): enable = 1 builds H as CSR in HBM (diagonal, same-spin
 * singles and doubles, alpha-single x beta-single pairs; 12 bytes per element, refused with
 * SBDX_ERR_DEVICE beyond 80% of free memory) and every later sigma build of this context, in
 * every solver, is a sparse matrix-vector product with it; enable = 0 frees it.  *nnz
 * (optional): stored elements.  Setting new integrals drops it. */
int sbdx_store_hamiltonian(sbdx_ctx *ctx, int enable, int64_t *nnz);

/* Davidson options (defaults by sbdx_davidson_defaults = the reference's SBD struct,
 * chemistry/tpb/sbdiag.h:12-30, with the iteration cap raised to converge). */
typedef struct {
  int max_iteration;      /* restarts (reference: max_it) */
  int num_block;          /* subspace size per restart, in blocks (reference: max_nb) */
  double eps;             /* residual tolerance ||E0 W - H W||, every requested root */
  double max_time;        /* wall-clock budget in seconds */
  int init;               /* 0: W[0] = 1 (HF, reference init 0); 1: vector 0 of W as given */
  int precision;          /* SBDX_FP64, or SBDX_FP32 = mixed precision (method 1 only): FP32
                             correction blocks and sigmas (accumulated as set by
                             sbdx_set_fp32_accumulation), FP64 anchor sigma, Rayleigh-Ritz,
                             orthogonalization and residual */
  int nroots;             /* lowest roots returned and converged, 1..8 (> 1: method 1) */
  int preconditioner;     /* 0: diagonal (Jacobi); 1: learned (method 1; falls back to 0 for
                             good when a restart cycle fails to halve the exact residual) */
  int verbose;            /* print the reference's per-step line */
  double energy_ref;      /* if nonzero: record time to |E0 - energy_ref| < energy_tol, judged
                             on FP64 Ritz values only */
  double energy_tol;
  int method;             /* 0: the reference's algorithm (one vector, restart to the Ritz
                             vector every num_block steps); 1: block thick-restart Davidson */
  int block;              /* method 1: vectors per block (1, 2, 4, 8; rounded up to a power of
                             two >= nroots); 0 = nroots.  The first block: the start vector
                             and unit vectors on the block - 1 lowest H_ii; with init 0 these
                             are taken inside determinant 0's block of H (the determinants
                             whose strings are linked to determinant 0's by chains of single
                             and double excitations), where the reference's Davidson stays */
  int stop_at_tol;        /* 1: stop once energy_ref is reached (time-to-solution runs) */
  int precond_warmup;     /* preconditioner 1: restart cycles run with the diagonal
                             preconditioner before the learned one takes over (default 1:
                             the first cycle, from the start vector, far from the training
                             distribution of Ritz values) */
} sbdx_davidson_opts;

typedef struct {
  double energy;          /* <W|H|W> of the returned vector (reference: sbdiag.h:239-262) */
  double residual;        /* last ||r||, the largest over the requested roots */
  int steps;              /* sigma builds inside the iteration */
  int converged;
  double seconds;         /* wall time of the iteration */
  double seconds_to_tol;  /* time at which |E0 - energy_ref| < energy_tol, or -1 */
  int steps_to_tol;
  int precond_fallback;   /* preconditioner 1 stalled and the diagonal one took over at this
                             step (0: never) */
} sbdx_davidson_result;

void sbdx_davidson_defaults(sbdx_davidson_opts *opts);

/* W: host buffer of nroots*ndet doubles, vector-major (in: initial guess in vector 0 if
 * init=1; out: the nroots lowest Ritz vectors).  trace/resid (optional, may be NULL): lowest
 * Ritz value and residual norm per step.  Returns SBDX_OK also when the iteration stops before
 * eps (res->converged = 0). */
int sbdx_davidson(sbdx_ctx *ctx, const sbdx_davidson_opts *opts, double *W,
                  sbdx_davidson_result *res, double *trace, double *resid, int trace_cap);

/* Device-resident variant: dW is a device buffer of nroots*ndet doubles. */
int sbdx_davidson_device(sbdx_ctx *ctx, const sbdx_davidson_opts *opts, void *dW,
                         sbdx_davidson_result *res, double *trace, double *resid,
                         int trace_cap);

/* 1- and 2-RDM of W (host): onebody[s][i + L*j] (2 L^2), twobody[s + 2t][i + L j + L^2 k + L^3 l]
 * (4 L^4) = <c+_{is} c+_{jt} c_{lt} c_{ks}> (the reference's layout, user-guide.md:136-139). */
int sbdx_rdm(sbdx_ctx *ctx, const double *W, double *onebody, double *twobody);

/* Device-resident variant: dW is a device buffer of ndet doubles (the interleaved basis, as
 * sbdx_davidson_device leaves it); the RDMs are written to host memory. */
int sbdx_rdm_device(sbdx_ctx *ctx, const void *dW, double *onebody, double *twobody);

/* Occupations density[2p + s] = <n_{p,s}> of W (host). */
int sbdx_occupation(sbdx_ctx *ctx, const double *W, double *density);

/* ---- Learned preconditioner (no counterpart in the reference; its fallback, the diagonal
 * preconditioner, is the reference's, chemistry/tpb/davidson.h:855-862).
 *
 * A 3-layer MLP of width 64 maps the features of determinant i -- occupation bits of its alpha
 * and beta strings, xh = asinh(H_ii - theta), xr = asinh(|r_i| / rms(r)) -- to m_i, and the
 * correction becomes t_i = r_i / (theta - H_ii) * (1 + m_i).  Parameters: one float vector of
 * SBDX_MLP_NPARAM entries, used rounded to bfloat16 (weights BF16, accumulation FP32):
 *   W1[130][64] (rows: alpha orbital p, 64 + beta orbital p, 128 = xh, 129 = xr), b1[64],
 *   W2[64][64] ([in][out]), b2[64], w3[64], b3.
 *   z1 = b1 + x W1; a1 = relu(z1); z2 = b2 + a1 W2; a2 = relu(z2); m = b3 + a2 . w3.
 * Training pairs come from the solver's own iterates: for an iterate x and the converged
 * vector x*, d = x* - (x . x*) x is the exact correction, t = r / (theta - H_ii) the diagonal
 * one; u = t / |c d|, v = c d / |c d| with c = (t . d) / (d . d); loss = mean (u (1 + m) - v)^2.
 */
#define SBDX_MLP_NPARAM 12609

typedef struct {
  uint64_t alpha, beta;   /* occupation bit masks of the determinant's strings */
  float xh, xr;           /* features (see above) */
  float u, v;             /* diagonal correction and target, importance-weighted */
} sbdx_mlp_sample;

typedef struct {
  int epochs;
  int batch;              /* samples per optimizer step */
  double lr, beta1, beta2, eps, weight_decay;   /* Adam */
  uint64_t seed;          /* shuffling */
  int verbose;
} sbdx_mlp_train_opts;

void sbdx_mlp_train_defaults(sbdx_mlp_train_opts *opts);

/* Record training pairs at iterate dX against the converged dXstar (device vectors of ndet
 * doubles, the caller's basis; need not be normalised).  Determinant i is kept with probability
 * p_i = min(1, target * (u_i^2 + v_i^2) / sum_j (u_j^2 + v_j^2)) (a counter-based hash of i and
 * seed decides), and its u, v are scaled by 1 / sqrt(p_i).  Writes min(count, cap) samples to
 * out (host) in no particular order; *count = number kept; *theta = Rayleigh quotient of dX. */
int sbdx_mlp_record(sbdx_ctx *ctx, const void *dX, const void *dXstar, double target,
                    uint64_t seed, sbdx_mlp_sample *out, int64_t cap, int64_t *count,
                    double *theta);

/* Mean loss over n samples (host) and its gradient with respect to the bfloat16-rounded
 * parameters (current HIP device). */
int sbdx_mlp_loss_grad(const sbdx_mlp_sample *samples, int64_t n, const float *params,
                       double *loss, float *grad);

/* Network output m for n samples. */
int sbdx_mlp_forward(const sbdx_mlp_sample *samples, int64_t n, const float *params, float *m);

/* Adam training on the GPU; params: in = initial, out = trained (FP32 master values).
 * loss_trace (optional): mean loss of each epoch. */
int sbdx_mlp_train(const sbdx_mlp_sample *samples, int64_t n, const sbdx_mlp_train_opts *opts,
                   float *params, double *loss_trace, int trace_cap);

/* Network used by preconditioner 1 on this context (NULL removes it). */
int sbdx_mlp_set(sbdx_ctx *ctx, const float *params);

/* Apply preconditioner 1 to one residual (host vectors of ndet doubles, caller's basis). */
int sbdx_mlp_apply(sbdx_ctx *ctx, double theta, const double *r, double *t);

/* ---- Batches of independent problems, no counterpart in the reference (sbd solves one problem
 * per MPI job): many small contexts solved together by a lockstep block Davidson (one vector
 * per block, thick restart, FP64 or mixed precision, diagonal preconditioner), each lockstep
 * step replayed as one HIP graph that holds the kernels of every unconverged problem. */
typedef struct sbdx_batch sbdx_batch;

/* ctxs: count contexts on the current device, integrals set; the batch does not own them. */
int sbdx_batch_create(int count, sbdx_ctx *const *ctxs, sbdx_batch **out);

/* Solve every problem from its HF vector (opts: nroots 1, block 0/1, preconditioner 0,
 * num_block 2..24; method is ignored).  W (optional): count host buffers of ndet_i doubles
 * (NULL entries skipped) for the returned vectors; res: count results. */
int sbdx_batch_davidson(sbdx_batch *batch, const sbdx_davidson_opts *opts, double *const *W,
                        sbdx_davidson_result *res);

void sbdx_batch_destroy(sbdx_batch *batch);

/* Autotuner, no counterpart in the reference: times every variant of the sigma build's
 * alpha-single x beta-single kernel (coded 1000000 (stages - 1) + 100 * threads per workgroup +
 * entries per thread: threads x entries set its tile shape, stages the alpha singles taken per
 * barrier) that fits this problem, and for k = 4 and 8 the sub-block widths (codes 2 and 4: the
 * block built as k / code blocks of that width, each gathered out of the interleaved block and
 * scattered back), for block width k and precision prec on this context's problem, `reps`
 * builds each, and keeps the fastest for every later build of that (prec, k).
 * candidates/ms (optional, cap entries): the variants and their mean ms per build; *chosen:
 * the variant kept; *ncand: variants tried. */
int sbdx_autotune(sbdx_ctx *ctx, int prec, int k, int reps, int *chosen, int *candidates,
                  double *ms, int cap, int *ncand);

/* No counterpart in the reference: the accumulation of this context's FP32 sigma builds
 * (SBDX_ACC_FP64 or SBDX_ACC_TWOSUM; SBDX_ERR_ARG otherwise).  Each accumulation keeps its own
 * kernel-variant choice: sbdx_autotune and sbdx_set_mix_variant act on the one in effect. */
int sbdx_set_fp32_accumulation(sbdx_ctx *ctx, int mode);

/* No counterpart in the reference: pin the alpha-single x beta-single kernel variant for
 * (prec, k) to one of the autotuner's candidates (the values sbdx_autotune reports), or back to
 * the built-in default with variant 0; SBDX_ERR_ARG for any other value.  For tests and for
 * reproducible timing. */
int sbdx_set_mix_variant(sbdx_ctx *ctx, int prec, int k, int variant);

/* Measurement hook, no counterpart in the reference: when enabled, every sigma build records
 * HIP events on its launch stream around its phases (0 alpha rows, 1 alpha-single x
 * beta-single, 2 transpose, 3 beta rows, 4 transposed add) and accumulates their durations.
 * Enabling resets the sums. */
int sbdx_profile(sbdx_ctx *ctx, int enable);
int sbdx_profile_read(sbdx_ctx *ctx, double *ms, int cap, int64_t *calls);

/* Measurement hook, no counterpart in the reference: bandwidth ceiling of a memory level of the
 * current device, in GB/s, over `reps` timed launches.  SBDX_PROBE_HBM: STREAM-style copy of
 * `bytes` (2 x bytes moved per launch; use a buffer far above the L2s and the MALL).
 * SBDX_PROBE_L2: every workgroup reads the whole of an L2-resident buffer of `bytes` (at most
 * 2 MiB) in 16-byte loads (workgroups x bytes moved per launch, L2 to L1).  Allocates. */
#define SBDX_PROBE_HBM 0
#define SBDX_PROBE_L2 1
int sbdx_bandwidth_probe(int kind, int64_t bytes, int reps, double *gbs);

void sbdx_destroy(sbdx_ctx *ctx);

/* Message of the last failed call on this host thread ("" if none). */
const char *sbdx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
