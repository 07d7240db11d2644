// noalloc.hip -- the sigma build of sbdx_sigma_device for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// Its scratch -- the input in the blocked basis, the blocked output, and the transposed pair of
// the beta pass, four vectors of ndet words -- lives in the file-scope device array `scratch`
// below (kScratchBytes on every device the library runs on, held until the process ends).  The
// array's host-side name is not a device address: its address on a device is taken once per
// device with hipGetSymbolAddress, with that device current, when the first context is created
// there (sbdx_create), and only that address is ever handed to kernels.  Every call on a device
// shares that device's array, which keeps the last call's contents (each call overwrites all
// it reads before reading it): calls that overlap on one device race.
#include "common.h"

#include <mutex>
#include <vector>

namespace sbdx {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 4 vectors of 2,097,152 FP64 words

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch;  // per device: the device address of `scratch`
}  // namespace

// the device address of `scratch` on `device`, which must be the current device; resolved once
void* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0) throw Error(SBDX_ERR_ARG, "bad device");
  if (g_scratch.size() <= static_cast<size_t>(device)) g_scratch.resize(device + 1, nullptr);
  if (!g_scratch[device]) {
    void* p = nullptr;
    SBDX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(SBDX_ERR_DEVICE, "no device address for the scratch array");
    g_scratch[device] = p;
  }
  return g_scratch[device];
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// sigma of one vector (interleaved basis, device buffers), enqueued on st: the signs S at the
// boundary, the build in the blocked basis, all scratch from the device's array; no device
// allocation, no synchronisation
void sigma_noalloc(sbdx_ctx& c, int prec, const void* dC, void* dS, hipStream_t st) {
  if (c.stored) throw Error(SBDX_ERR_STATE, "not available with a stored Hamiltonian");
  const int64_t n = c.ndet();
  const size_t es = prec == SBDX_FP64 ? 8 : 4;
  const size_t need = 4 * static_cast<size_t>(n) * es;
  if (need > kScratchBytes)
    throw Error(SBDX_ERR_ARG, "problem too large for the no-allocation sigma: 4 x ndet x word size = " +
                                  std::to_string(need) + " bytes of scratch, the maximum is " +
                                  std::to_string(kScratchBytes));
  int dev = -1;
  SBDX_HIP(hipGetDevice(&dev));
  if (dev != c.device) throw Error(SBDX_ERR_STATE, "the context's device is not the current device");
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (static_cast<size_t>(dev) < g_scratch.size()) base = static_cast<unsigned char*>(g_scratch[dev]);
  }
  if (!base) throw Error(SBDX_ERR_STATE, "scratch address not resolved on this device");
  void* xb = base;
  void* yb = base + n * es;
  c.noalloc_xT = base + 2 * n * es;
  c.noalloc_yT = base + 3 * n * es;
  c.no_alloc = true;
  try {
    apply_sign_block(c, prec, 1, dC, xb, st);
    sigma_block(c, prec, 1, xb, yb, st);
    apply_sign_block(c, prec, 1, yb, dS, st);
  } catch (...) {
    c.no_alloc = false;
    c.noalloc_xT = c.noalloc_yT = nullptr;
    throw;
  }
  c.no_alloc = false;
  c.noalloc_xT = c.noalloc_yT = nullptr;
}

}  // namespace sbdx
