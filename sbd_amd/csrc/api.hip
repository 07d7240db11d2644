// api.hip -- the extern "C" boundary of libsbdx (declared in include/sbdx.h).
//
// Every entry point converts C++ exceptions into a negative status plus a thread-local
// message; host vectors are converted between the reference's interleaved basis and the
// engine's blocked basis here and nowhere else.
#include "common.h"
#include "sbdx.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

namespace sbdx {
void davidson_fp64(sbdx_ctx& c, const sbdx_davidson_opts& o, double* dW, sbdx_davidson_result& res,
                   double* trace, double* resid, int trace_cap);
void init_hf(sbdx_ctx& c, double* dW, hipStream_t st);
void davidson_block(sbdx_ctx& c, const sbdx_davidson_opts& o, double* dW, sbdx_davidson_result& res,
                    double* trace, double* resid, int trace_cap);
int davidson_block_width(const sbdx_davidson_opts& o);
void rdm_blocked(sbdx_ctx& c, const double* dW, double* onebody, double* twobody);
void occupation_blocked(sbdx_ctx& c, const double* dW, double* density);
void mlp_set(sbdx_ctx& c, const float* params);
int64_t mlp_record(sbdx_ctx& c, const double* dX, const double* dXs, double target, uint64_t seed,
                   sbdx_mlp_sample* out, int64_t cap, double* theta_out);
void mlp_loss_grad(const sbdx_mlp_sample* S, int64_t n, const float* params, double* loss, float* grad);
void mlp_forward(const sbdx_mlp_sample* S, int64_t n, const float* params, float* m);
void mlp_train(const sbdx_mlp_sample* S, int64_t n, const sbdx_mlp_train_opts& o, float* params, double* trace,
               int trace_cap);
double bandwidth_probe(int kind, int64_t bytes, int reps);
void sigma_noalloc(sbdx_ctx& c, int prec, const void* dC, void* dS, hipStream_t st);
}  // namespace sbdx

namespace {

thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_err.clear();
    return SBDX_OK;
  } catch (const sbdx::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return SBDX_ERR_DEVICE;
  } catch (const std::exception& e) {
    g_err = e.what();
    return SBDX_ERR_ARG;
  }
}

void need(bool ok, const char* msg) {
  if (!ok) throw sbdx::Error(SBDX_ERR_ARG, msg);
}

void need_ints(const sbdx_ctx* c) {
  if (!c) throw sbdx::Error(SBDX_ERR_ARG, "null context");
  if (!c->have_ints) throw sbdx::Error(SBDX_ERR_STATE, "integrals not set (call sbdx_set_integrals)");
}

// host vector (interleaved basis) -> device (blocked basis) and back
void upload_blocked(sbdx_ctx& c, const double* host, double* dev, sbdx::DevBuf<double>& tmp) {
  const int64_t n = c.ndet();
  tmp.ensure(n);
  SBDX_HIP(hipMemcpyAsync(tmp.get(), host, n * 8, hipMemcpyHostToDevice, c.stream));
  sbdx::apply_sign(c, SBDX_FP64, 1, tmp.get(), dev, c.stream);
}

void download_user(sbdx_ctx& c, const double* dev, double* host, sbdx::DevBuf<double>& tmp) {
  const int64_t n = c.ndet();
  tmp.ensure(n);
  sbdx::apply_sign(c, SBDX_FP64, 1, dev, tmp.get(), c.stream);
  SBDX_HIP(hipMemcpyAsync(host, tmp.get(), n * 8, hipMemcpyDeviceToHost, c.stream));
  SBDX_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace

extern "C" {

int sbdx_create_words(int norb, int words, int64_t n_alpha, const uint64_t* alpha, int64_t n_beta,
                      const uint64_t* beta, sbdx_ctx** out) {
  return guarded([&] {
    need(out != nullptr, "out is null");
    *out = nullptr;
    need(words == 1 || words == 2, "words must be 1 or 2");
    need(norb >= 1 && norb <= 64 * words, "norb must be in [1, 64 words]");
    need(n_alpha >= 1 && n_beta >= 1, "empty string set");
    need(n_alpha < (int64_t(1) << 32) && n_beta < (int64_t(1) << 32), "too many strings");
    need(alpha && beta, "null string table");
    // the context keeps W = 1 word per string up to 64 orbitals and 2 above
    const int W = norb > 64 ? 2 : 1;
    std::vector<uint64_t> sa(n_alpha * W), sb(n_beta * W);
    auto take = [&](const uint64_t* in, int64_t n, std::vector<uint64_t>& outv, const char* bits_msg,
                    const char* count_msg) {
      int pop0 = 0;
      for (int64_t i = 0; i < n; ++i) {
        int pop = 0;
        for (int k = 0; k < words; ++k) {
          const int bits = std::min(64, std::max(0, norb - 64 * k));
          const uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
          const uint64_t v = in[i * words + k];
          need((v & ~mask) == 0, bits_msg);
          if (k < W) outv[i * W + k] = v;
          pop += sbdx::popc64(v);
        }
        // every string of one spin must hold the same electron count: the excitation tables
        // decode a popcount-2 (-4) difference as exactly one (two) electron(s) moved
        if (i == 0) pop0 = pop;
        need(pop == pop0, count_msg);
      }
    };
    take(alpha, n_alpha, sa, "alpha string has bits >= norb", "alpha strings differ in electron count");
    take(beta, n_beta, sb, "beta string has bits >= norb", "beta strings differ in electron count");
    auto* c = new sbdx_ctx();
    try {
      SBDX_HIP(hipGetDevice(&c->device));
      (void)sbdx::noalloc_scratch(c->device);  // once per device: the no-allocation sigma's scratch
      SBDX_HIP(hipStreamCreate(&c->stream));
      c->L = norb;
      c->W = W;
      c->na = n_alpha;
      c->nb = n_beta;
      sbdx::build_tables(*c, sa.data(), sb.data());
    } catch (...) {
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
      throw;
    }
    *out = c;
  });
}

int sbdx_create(int norb, int64_t n_alpha, const uint64_t* alpha, int64_t n_beta,
                const uint64_t* beta, sbdx_ctx** out) {
  if (norb > 64) return guarded([&] { need(false, "norb must be in [1, 64] (sbdx_create_words takes up to 128)"); });
  return sbdx_create_words(norb, 1, n_alpha, alpha, n_beta, beta, out);
}

int sbdx_set_integrals(sbdx_ctx* c, double I0, const double* I1, const double* I2, int64_t n_I2) {
  return guarded([&] {
    need(c && I1 && I2, "null argument");
    const int64_t B = static_cast<int64_t>(c->L) * (c->L + 1) / 2;
    need(n_I2 == B * (B + 1) / 2, "n_I2 must be B(B+1)/2 with B = norb(norb+1)/2");
    c->I0 = I0;
    c->h_I1.assign(I1, I1 + 4 * c->L * c->L);
    c->h_I2.assign(I2, I2 + n_I2);
    sbdx::build_integral_tables(*c);
  });
}

int sbdx_excitation_counts(sbdx_ctx* c, int spin, int64_t* n_single, int64_t* n_double) {
  return guarded([&] {
    need(c && (spin == 0 || spin == 1), "bad argument");
    const sbdx::SpinTab& t = spin == 0 ? c->A : c->B;
    if (n_single) *n_single = t.nnz_s;
    if (n_double) *n_double = t.nnz_d;
  });
}

int sbdx_get_excitations(sbdx_ctx* c, int spin, int64_t* off_s, int64_t* idx_s, int64_t* off_d,
                         int64_t* idx_d) {
  return guarded([&] {
    need(c && (spin == 0 || spin == 1), "bad argument");
    const sbdx::SpinTab& t = spin == 0 ? c->A : c->B;
    std::vector<uint32_t> tmp;
    auto get = [&](const sbdx::DevBuf<int64_t>& off, const sbdx::DevBuf<uint32_t>& ket, int64_t nnz,
                   int64_t* ho, int64_t* hi) {
      if (ho) SBDX_HIP(hipMemcpy(ho, off.get(), (t.n + 1) * 8, hipMemcpyDeviceToHost));
      if (hi && nnz) {
        tmp.resize(nnz);
        SBDX_HIP(hipMemcpy(tmp.data(), ket.get(), nnz * 4, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < nnz; ++k) hi[k] = tmp[k];
      }
    };
    get(t.off_s, t.ket_s, t.nnz_s, off_s, idx_s);
    get(t.off_d, t.ket_d, t.nnz_d, off_d, idx_d);
  });
}

int64_t sbdx_ndet(const sbdx_ctx* c) { return c ? c->ndet() : 0; }

int sbdx_hdiag(sbdx_ctx* c, double* hii) {
  return guarded([&] {
    need_ints(c);
    need(hii != nullptr, "null output");
    SBDX_HIP(hipMemcpy(hii, c->hii.get(), c->ndet() * 8, hipMemcpyDeviceToHost));
  });
}

int sbdx_sigma(sbdx_ctx* c, int nvec, const double* C, double* S) {
  return guarded([&] {
    need_ints(c);
    need(C && S && nvec >= 1, "bad argument");
    const int64_t n = c->ndet();
    sbdx::DevBuf<double> x(n), y(n), tmp;
    for (int v = 0; v < nvec; ++v) {
      upload_blocked(*c, C + v * n, x.get(), tmp);
      sbdx::sigma_blocked(*c, SBDX_FP64, 1, x.get(), y.get(), c->stream);
      download_user(*c, y.get(), S + v * n, tmp);
    }
  });
}

int sbdx_sigma_device(sbdx_ctx* c, int prec, int nvec, const void* dC, void* dS, void* stream) {
  return guarded([&] {
    need_ints(c);
    need(dC && dS && nvec >= 1 && (prec == SBDX_FP64 || prec == SBDX_FP32), "bad argument");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const int64_t n = c->ndet();
    const size_t es = prec == SBDX_FP64 ? 8 : 4;
    c->ws.tmp_in.ensure(n * es);
    c->ws.tmp_out.ensure(n * es);
    for (int v = 0; v < nvec; ++v) {
      const char* xin = static_cast<const char*>(dC) + v * n * es;
      char* yout = static_cast<char*>(dS) + v * n * es;
      sbdx::apply_sign(*c, prec, 1, xin, c->ws.tmp_in.get(), st);
      sbdx::sigma_blocked(*c, prec, 1, c->ws.tmp_in.get(), c->ws.tmp_out.get(), st);
      sbdx::apply_sign(*c, prec, 1, c->ws.tmp_out.get(), yout, st);
    }
    SBDX_HIP(hipStreamSynchronize(st));
  });
}

int sbdx_sigma_device_noalloc(sbdx_ctx* c, int prec, const void* dC, void* dS, void* stream) {
  return guarded([&] {
    need_ints(c);
    need(dC && dS && (prec == SBDX_FP64 || prec == SBDX_FP32), "bad argument");
    sbdx::sigma_noalloc(*c, prec, dC, dS, stream ? static_cast<hipStream_t>(stream) : c->stream);
  });
}

int sbdx_sigma_block_device(sbdx_ctx* c, int prec, int k, const void* dC, void* dS, void* stream) {
  return guarded([&] {
    need_ints(c);
    need(dC && dS && (k == 1 || k == 2 || k == 4 || k == 8) && (prec == SBDX_FP64 || prec == SBDX_FP32),
         "bad argument (k must be 1, 2, 4 or 8)");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const int64_t n = c->ndet();
    const size_t es = prec == SBDX_FP64 ? 8 : 4;
    c->ws.tmp_in.ensure(n * k * es);
    c->ws.tmp_out.ensure(n * k * es);
    sbdx::apply_sign_block(*c, prec, k, dC, c->ws.tmp_in.get(), st);
    sbdx::sigma_block(*c, prec, k, c->ws.tmp_in.get(), c->ws.tmp_out.get(), st);
    sbdx::apply_sign_block(*c, prec, k, c->ws.tmp_out.get(), dS, st);
    SBDX_HIP(hipStreamSynchronize(st));
  });
}

void sbdx_davidson_defaults(sbdx_davidson_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_iteration = 100;
  o->num_block = 10;
  o->eps = 1.0e-4;
  o->max_time = 86400.0;
  o->init = 0;
  o->precision = SBDX_FP64;
  o->nroots = 1;
  o->preconditioner = 0;
  o->verbose = 0;
  o->energy_ref = 0.0;
  o->energy_tol = 1.0e-6;
  o->method = 0;
  o->block = 0;
  o->stop_at_tol = 0;
  o->precond_warmup = 1;
}

// validates the options; returns the vectors per block of the device work buffer
static int check_opts(const sbdx_ctx* c, const sbdx_davidson_opts* o) {
  need(o != nullptr, "null options");
  need(o->max_iteration >= 1 && o->num_block >= 1 && o->num_block <= 64, "bad iteration/block");
  need(o->nroots >= 1 && o->nroots <= 8, "nroots must be 1..8");
  need(o->precision == SBDX_FP64 || o->precision == SBDX_FP32, "precision must be SBDX_FP64 or SBDX_FP32");
  need(o->preconditioner == 0 || o->preconditioner == 1, "preconditioner must be 0 (diagonal) or 1 (learned)");
  need(o->preconditioner == 0 || c->have_mlp, "preconditioner 1 needs a network (sbdx_mlp_set)");
  need(o->method == 0 || o->method == 1, "method must be 0 or 1");
  need(o->block >= 0 && o->block <= 8, "block must be 0..8");
  need(o->precond_warmup >= 0, "precond_warmup must be >= 0");
  if (o->method == 0) {
    need(o->nroots == 1 && o->block <= 1 && o->precision == SBDX_FP64 && o->preconditioner == 0,
         "method 0 (the reference algorithm) is FP64, one root, diagonal preconditioner; use method 1");
    return 1;
  }
  const int k = sbdx::davidson_block_width(*o);
  need(c->ndet() >= k, "fewer determinants than vectors per block");
  return k;
}

static void run_davidson(sbdx_ctx& c, const sbdx_davidson_opts& o, double* w, sbdx_davidson_result& res,
                         double* trace, double* resid, int trace_cap) {
  if (o.method == 0) sbdx::davidson_fp64(c, o, w, res, trace, resid, trace_cap);
  else sbdx::davidson_block(c, o, w, res, trace, resid, trace_cap);
}

int sbdx_davidson_device(sbdx_ctx* c, const sbdx_davidson_opts* o, void* dW, sbdx_davidson_result* res,
                         double* trace, double* resid, int trace_cap) {
  return guarded([&] {
    need_ints(c);
    const int k = check_opts(c, o);
    need(dW && res, "null argument");
    std::memset(res, 0, sizeof(*res));
    const int64_t n = c->ndet();
    double* user = static_cast<double*>(dW);
    sbdx::DevBuf<double> w(n * k);
    if (o->init == 0) sbdx::init_hf(*c, w.get(), c->stream);
    else sbdx::apply_sign(*c, SBDX_FP64, 1, user, w.get(), c->stream);
    run_davidson(*c, *o, w.get(), *res, trace, resid, trace_cap);
    for (int v = 0; v < o->nroots; ++v) sbdx::apply_sign(*c, SBDX_FP64, 1, w.get() + v * n, user + v * n, c->stream);
    SBDX_HIP(hipStreamSynchronize(c->stream));
  });
}

int sbdx_davidson(sbdx_ctx* c, const sbdx_davidson_opts* o, double* W, sbdx_davidson_result* res,
                  double* trace, double* resid, int trace_cap) {
  return guarded([&] {
    need_ints(c);
    const int k = check_opts(c, o);
    need(W && res, "null argument");
    std::memset(res, 0, sizeof(*res));
    const int64_t n = c->ndet();
    sbdx::DevBuf<double> w(n * k), tmp;
    if (o->init == 0) sbdx::init_hf(*c, w.get(), c->stream);
    else upload_blocked(*c, W, w.get(), tmp);
    run_davidson(*c, *o, w.get(), *res, trace, resid, trace_cap);
    for (int v = 0; v < o->nroots; ++v) download_user(*c, w.get() + v * n, W + v * n, tmp);
  });
}

int sbdx_rdm(sbdx_ctx* c, const double* W, double* onebody, double* twobody) {
  return guarded([&] {
    need_ints(c);
    need(W && onebody && twobody, "null argument");
    sbdx::DevBuf<double> w(c->ndet()), tmp;
    upload_blocked(*c, W, w.get(), tmp);
    sbdx::rdm_blocked(*c, w.get(), onebody, twobody);
  });
}

int sbdx_rdm_device(sbdx_ctx* c, const void* dW, double* onebody, double* twobody) {
  return guarded([&] {
    need_ints(c);
    need(dW && onebody && twobody, "null argument");
    sbdx::DevBuf<double> w(c->ndet());
    sbdx::apply_sign(*c, SBDX_FP64, 1, dW, w.get(), c->stream);
    sbdx::rdm_blocked(*c, w.get(), onebody, twobody);
  });
}

int sbdx_occupation(sbdx_ctx* c, const double* W, double* density) {
  return guarded([&] {
    need(c && W && density, "null argument");
    sbdx::DevBuf<double> w(c->ndet()), tmp;
    upload_blocked(*c, W, w.get(), tmp);
    sbdx::occupation_blocked(*c, w.get(), density);
  });
}

void sbdx_mlp_train_defaults(sbdx_mlp_train_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->epochs = 20;
  o->batch = 16384;
  o->lr = 1.0e-3;
  o->beta1 = 0.9;
  o->beta2 = 0.999;
  o->eps = 1.0e-8;
  o->weight_decay = 0.0;
  o->seed = 0;
  o->verbose = 0;
}

int sbdx_mlp_record(sbdx_ctx* c, const void* dX, const void* dXstar, double target, uint64_t seed,
                    sbdx_mlp_sample* out, int64_t cap, int64_t* count, double* theta) {
  return guarded([&] {
    need_ints(c);
    need(dX && dXstar && count && cap >= 0 && (out || cap == 0), "null argument");
    need(target > 0.0, "target must be positive");
    const int64_t n = c->ndet();
    sbdx::DevBuf<double> x(n), xs(n);
    sbdx::apply_sign(*c, SBDX_FP64, 1, dX, x.get(), c->stream);
    sbdx::apply_sign(*c, SBDX_FP64, 1, dXstar, xs.get(), c->stream);
    *count = sbdx::mlp_record(*c, x.get(), xs.get(), target, seed, out, cap, theta);
  });
}

int sbdx_mlp_loss_grad(const sbdx_mlp_sample* samples, int64_t n, const float* params, double* loss, float* grad) {
  return guarded([&] {
    need(samples && params && loss && grad && n > 0, "null argument or empty sample set");
    sbdx::mlp_loss_grad(samples, n, params, loss, grad);
  });
}

int sbdx_mlp_forward(const sbdx_mlp_sample* samples, int64_t n, const float* params, float* m) {
  return guarded([&] {
    need(samples && params && m && n > 0, "null argument or empty sample set");
    sbdx::mlp_forward(samples, n, params, m);
  });
}

int sbdx_mlp_train(const sbdx_mlp_sample* samples, int64_t n, const sbdx_mlp_train_opts* o, float* params,
                   double* loss_trace, int trace_cap) {
  return guarded([&] {
    need(samples && o && params && n > 0, "null argument or empty sample set");
    need(o->epochs >= 0 && o->batch >= 1 && o->lr > 0.0, "bad training options");
    sbdx::mlp_train(samples, n, *o, params, loss_trace, trace_cap);
  });
}

int sbdx_mlp_set(sbdx_ctx* c, const float* params) {
  return guarded([&] {
    need(c != nullptr, "null context");
    sbdx::mlp_set(*c, params);
  });
}

int sbdx_mlp_apply(sbdx_ctx* c, double theta, const double* r, double* t) {
  return guarded([&] {
    need_ints(c);
    need(r && t, "null argument");
    const int64_t n = c->ndet();
    sbdx::DevBuf<double> rb(n), tb(n), tmp;
    upload_blocked(*c, r, rb.get(), tmp);
    double nr = 0.0;
    for (int64_t i = 0; i < n; ++i) nr += r[i] * r[i];
    const double rms = std::sqrt(nr / static_cast<double>(n));
    sbdx::mlp_precondition(*c, SBDX_FP64, 1, rb.get(), &theta, &rms, tb.get(), c->stream);
    download_user(*c, tb.get(), t, tmp);
  });
}

int sbdx_autotune(sbdx_ctx* c, int prec, int k, int reps, int* chosen, int* candidates, double* ms, int cap,
                  int* ncand) {
  return guarded([&] {
    need_ints(c);
    need((k == 1 || k == 2 || k == 4 || k == 8) && (prec == SBDX_FP64 || prec == SBDX_FP32) && reps >= 1,
         "bad argument (k must be 1, 2, 4 or 8)");
    const int64_t n = c->ndet();
    const size_t es = prec == SBDX_FP64 ? 8 : 4;
    sbdx::DevBuf<unsigned char> x(n * k * es), y(n * k * es);
    SBDX_HIP(hipMemsetAsync(x.get(), 0, n * k * es, c->stream));
    hipEvent_t e0, e1;
    SBDX_HIP(hipEventCreate(&e0));
    SBDX_HIP(hipEventCreate(&e1));
    std::vector<int> cand;
    for (int v : sbdx::mix_variant_candidates(k))
      if (sbdx::mix_variant_fits(*c, prec, k, v)) cand.push_back(v);
    int& slot = sbdx::tuned_variant(*c, prec, k);  // the context's current FP32 accumulation
    int best = 0;
    double best_ms = 0.0;
    try {
      for (size_t q = 0; q < cand.size(); ++q) {
        slot = cand[q];
        sbdx::sigma_block(*c, prec, k, x.get(), y.get(), c->stream);  // builds this variant's tiles
        SBDX_HIP(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps; ++r) sbdx::sigma_block(*c, prec, k, x.get(), y.get(), c->stream);
        SBDX_HIP(hipEventRecord(e1, c->stream));
        SBDX_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        SBDX_HIP(hipEventElapsedTime(&t, e0, e1));
        const double mean = static_cast<double>(t) / reps;
        if (ms && static_cast<int>(q) < cap) ms[q] = mean;
        if (candidates && static_cast<int>(q) < cap) candidates[q] = cand[q];
        if (best == 0 || mean < best_ms) {
          best = cand[q];
          best_ms = mean;
        }
      }
    } catch (...) {
      slot = 0;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    slot = best;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (chosen) *chosen = best;
    if (ncand) *ncand = static_cast<int>(cand.size());
  });
}

int sbdx_set_fp32_accumulation(sbdx_ctx* c, int mode) {
  return guarded([&] {
    need(c != nullptr, "null context");
    need(mode == SBDX_ACC_FP64 || mode == SBDX_ACC_TWOSUM, "mode must be SBDX_ACC_FP64 or SBDX_ACC_TWOSUM");
    c->fp32_acc = mode;
  });
}

int sbdx_set_mix_variant(sbdx_ctx* c, int prec, int k, int variant) {
  return guarded([&] {
    need(c != nullptr, "null context");
    need((k == 1 || k == 2 || k == 4 || k == 8) && (prec == SBDX_FP64 || prec == SBDX_FP32),
         "bad argument (k must be 1, 2, 4 or 8)");
    const std::vector<int> cand = sbdx::mix_variant_candidates(k);
    need(variant == 0 || std::find(cand.begin(), cand.end(), variant) != cand.end(),
         "not a mixed-kernel variant for this block width");
    need(variant == 0 || sbdx::mix_variant_fits(*c, prec, k, variant),
         "this mixed-kernel variant needs more LDS than a CU has for this problem");
    sbdx::tuned_variant(*c, prec, k) = variant;
  });
}

int sbdx_store_hamiltonian(sbdx_ctx* c, int enable, int64_t* nnz) {
  return guarded([&] {
    need_ints(c);
    if (enable) {
      const int64_t m = sbdx::store_hamiltonian(*c);
      if (nnz) *nnz = m;
    } else {
      sbdx::drop_stored(*c);
      if (nnz) *nnz = 0;
    }
  });
}

int sbdx_batch_create(int count, sbdx_ctx* const* ctxs, sbdx_batch** out) {
  return guarded([&] {
    need(out != nullptr && ctxs != nullptr && count >= 1, "bad argument");
    *out = nullptr;
    int dev = 0;
    SBDX_HIP(hipGetDevice(&dev));
    for (int i = 0; i < count; ++i) {
      need_ints(ctxs[i]);
      need(ctxs[i]->device == dev, "every context of a batch must live on the current device");
      for (int j = 0; j < i; ++j) need(ctxs[j] != ctxs[i], "a context appears twice in the batch");
    }
    auto* b = new sbdx_batch();
    b->ctx.assign(ctxs, ctxs + count);
    b->device = dev;
    if (hipStreamCreate(&b->stream) != hipSuccess) {
      delete b;
      throw sbdx::Error(SBDX_ERR_DEVICE, "stream creation failed");
    }
    *out = b;
  });
}

int sbdx_batch_davidson(sbdx_batch* b, const sbdx_davidson_opts* o, double* const* W, sbdx_davidson_result* res) {
  return guarded([&] {
    need(b != nullptr && o != nullptr && res != nullptr, "null argument");
    need(o->max_iteration >= 1 && o->num_block >= 2 && o->num_block <= 24, "num_block must be 2..24");
    need(o->nroots == 1 && o->block <= 1 && o->preconditioner == 0,
         "a batch solves the lowest root with one vector per block and the diagonal preconditioner");
    need(o->precision == SBDX_FP64 || o->precision == SBDX_FP32, "bad precision");
    sbdx::batch_davidson(*b, *o, W, res);
  });
}

void sbdx_batch_destroy(sbdx_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int sbdx_profile(sbdx_ctx* c, int enable) {
  return guarded([&] {
    need(c != nullptr, "null context");
    sbdx::Prof& p = c->prof;
    if (enable && !p.ev[0])
      for (auto& e : p.ev) SBDX_HIP(hipEventCreate(&e));
    p.on = enable != 0;
    for (double& m : p.ms) m = 0.0;
    p.calls = 0;
  });
}

int sbdx_profile_read(sbdx_ctx* c, double* ms, int cap, int64_t* calls) {
  return guarded([&] {
    need(c && ms, "null argument");
    for (int k = 0; k < cap && k < sbdx::kPhases; ++k) ms[k] = c->prof.ms[k];
    if (calls) *calls = c->prof.calls;
  });
}

int sbdx_bandwidth_probe(int kind, int64_t bytes, int reps, double* gbs) {
  return guarded([&] {
    need(gbs != nullptr, "null argument");
    *gbs = sbdx::bandwidth_probe(kind, bytes, reps);
  });
}

void sbdx_destroy(sbdx_ctx* c) {
  if (!c) return;
  for (auto& e : c->prof.ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  hipStream_t s = c->stream;
  delete c;
  if (s) (void)hipStreamDestroy(s);
}

const char* sbdx_last_error(void) { return g_err.c_str(); }

}  // extern "C"
