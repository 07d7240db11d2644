// probe.hip -- bandwidth ceilings of the memory levels the sigma build runs against, measured on
// the device it runs on (measurement only; no counterpart in the reference).
//
//  * SBDX_PROBE_HBM: a STREAM-style copy of `bytes` (read once, written once, far larger than
//    the L2s and the MALL): bytes moved = 2 x bytes per repetition.
//  * SBDX_PROBE_L2: every workgroup streams the whole of an L2-resident buffer of `bytes`
//    (<= 2 MiB, so every XCD's 4 MiB L2 holds its own copy; each CU's 32 KiB L1 cannot), in
//    whole 16-byte loads: bytes moved = workgroups x bytes per repetition, all L2 -> L1.
// Both time `reps` launches with HIP events on the null stream and return GB/s.
#include "common.h"
#include "sbdx.h"

namespace sbdx {
namespace {

// four 16-byte words per thread in flight: workgroup g copies chunks g, g + grid, ... of
// 4 x 256 words
__global__ __launch_bounds__(256) void k_probe_copy(const uint4* __restrict__ in, uint4* __restrict__ out, int64_t n) {
  constexpr int U = 4;
  const int64_t chunk = U * 256;
  for (int64_t c = blockIdx.x; c * chunk < n; c += gridDim.x) {
    const int64_t base = c * chunk + threadIdx.x;
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * 256 < n) v[u] = in[base + u * 256];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * 256 < n) out[base + u * 256] = v[u];
  }
}

// each workgroup reads all n words of buf, starting at a workgroup-dependent offset so the
// workgroups of one XCD do not move in lockstep over the same lines; the xor of what a thread
// read is stored so the loads are not removed
__global__ __launch_bounds__(256) void k_probe_l2(const uint4* __restrict__ buf, int64_t n, uint32_t* __restrict__ sink) {
  uint32_t acc = 0;
  const int64_t start = (static_cast<int64_t>(blockIdx.x) * 4099) % n;
  for (int64_t j = threadIdx.x; j < n; j += 256) {
    int64_t i = start + j;
    if (i >= n) i -= n;
    const uint4 v = buf[i];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  sink[static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x] = acc;
}

}  // namespace

double bandwidth_probe(int kind, int64_t bytes, int reps) {
  if (reps < 1 || bytes < 4096 || bytes % 16 != 0) throw Error(SBDX_ERR_ARG, "bad probe arguments");
  if (kind != SBDX_PROBE_HBM && kind != SBDX_PROBE_L2) throw Error(SBDX_ERR_ARG, "unknown probe kind");
  if (kind == SBDX_PROBE_L2 && bytes > (2 << 20)) throw Error(SBDX_ERR_ARG, "L2 probe buffer above 2 MiB");
  int dev = 0, cus = 0;
  SBDX_HIP(hipGetDevice(&dev));
  SBDX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned grid = static_cast<unsigned>(cus) * 8;
  const int64_t n = bytes / 16;
  DevBuf<uint4> a(n);
  DevBuf<uint4> b(kind == SBDX_PROBE_HBM ? n : 1);
  DevBuf<uint32_t> sink(static_cast<size_t>(grid) * 256);
  SBDX_HIP(hipMemset(a.get(), 1, n * 16));
  hipEvent_t e0, e1;
  SBDX_HIP(hipEventCreate(&e0));
  SBDX_HIP(hipEventCreate(&e1));
  hipStream_t st = nullptr;  // the null stream: nothing else of this process runs meanwhile
  auto once = [&] {
    if (kind == SBDX_PROBE_HBM) k_probe_copy<<<grid, 256, 0, st>>>(a.get(), b.get(), n);
    else k_probe_l2<<<grid, 256, 0, st>>>(a.get(), n, sink.get());
  };
  once();  // first touch / warm every L2
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) once();
  SBDX_CHECK_LAUNCH();
  SBDX_HIP(hipEventRecord(e1, st));
  SBDX_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  SBDX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double moved = kind == SBDX_PROBE_HBM ? 2.0 * static_cast<double>(bytes) * reps
                                              : static_cast<double>(grid) * static_cast<double>(bytes) * reps;
  return moved / (static_cast<double>(ms) * 1e-3) / 1e9;
}

}  // namespace sbdx
