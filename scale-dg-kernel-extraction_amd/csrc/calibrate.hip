// Streaming calibration: the HBM rate a kernel with the stage kernels' access
// mix reaches on this GPU, as the practical ceiling beside the datasheet peak.
//
// A stage launch reads q, u, v, w (and q0 on two stages of three) and writes
// one field: about five bytes read per byte written.  stream_probe<NR, ASYNC>
// reads NR arrays and writes one (out = sum of the inputs), with
// non-temporal stores as the stage kernels' outputs, in two load forms:
//   ASYNC = true:  global_load_lds_dwordx4 into LDS (no VGPRs held), then LDS
//                  reads, as the staged fields of stage_f64, 16 bytes per lane
//                  per array;
//   ASYNC = false: one 8-byte load per lane and array into a register.  This
//                  is the fastest streaming form measured on MI355X: 6.1-6.3
//                  TB/s, against 5.4-5.9 for 16-byte loads (one to four per
//                  lane), two to eight 8-byte loads per lane, 512- and
//                  1024-lane workgroups, or cached stores.
// Workgroups take consecutive spans in launch order, so the 8 XCDs (dealt
// workgroups round-robin) work on neighbouring addresses at any time; giving
// each XCD a contiguous eighth of the arrays instead costs 5-10% (5.5-5.7
// TB/s in every load form).
#include <algorithm>

#include "dgk_internal.hpp"

namespace dgk {
namespace {

constexpr int kProbeThreads = 256;
// doubles per array per workgroup: 16 bytes per lane (LDS form), 8 (register form)
template <bool ASYNC>
constexpr int kProbeSpan = ASYNC ? 2 * kProbeThreads : kProbeThreads;
// n is a multiple of this (of both spans)
constexpr int kProbeDoubles = 2048;

template <int NR, bool ASYNC>
__global__ void __launch_bounds__(kProbeThreads)
stream_probe(const double *__restrict__ in, int64_t n, double *__restrict__ out) {
  // n is a multiple of kProbeDoubles (checked on the host); array r at in + r n
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  constexpr int PER = kProbeSpan<ASYNC>;
  const int64_t b0 = (int64_t)blockIdx.x * PER;
  if constexpr (ASYNC) {
    double2 acc = {0.0, 0.0};
    __shared__ double s[NR * PER];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const double *gp = in + (int64_t)r * n + b0 + 2 * (wave * 64 + lane);
      __attribute__((address_space(3))) void *lp =
          (__attribute__((address_space(3))) void *)(s + r * PER + 2 * wave * 64);
#if defined(__HIP_DEVICE_COMPILE__)
      __builtin_amdgcn_global_load_lds((const void *)gp, lp, 16, 0, 0);
#else
      (void)gp;
      (void)lp;
#endif
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const double2 x = *(const double2 *)(s + r * PER + 2 * tid);
      acc.x += x.x;
      acc.y += x.y;
    }
    double *o = out + b0 + 2 * tid;
    __builtin_nontemporal_store(acc.x, o);
    __builtin_nontemporal_store(acc.y, o + 1);
  } else {
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < NR; ++r) acc += in[(int64_t)r * n + b0 + tid];
    __builtin_nontemporal_store(acc, out + b0 + tid);
    (void)wave;
    (void)lane;
  }
}

template <int NR, bool ASYNC>
int time_probe(const double *in, int64_t n, double *out, int reps, hipStream_t st,
               double *gbs) {
  const int nblk = (int)(n / kProbeSpan<ASYNC>);
  hipEvent_t e0, e1;
  DGK_HIP(hipEventCreate(&e0));
  DGK_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL((stream_probe<NR, ASYNC>), dim3(nblk), dim3(kProbeThreads), 0, st, in,
                     n, out);  // warm-up
  DGK_HIP(hipGetLastError());
  DGK_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL((stream_probe<NR, ASYNC>), dim3(nblk), dim3(kProbeThreads), 0, st, in,
                       n, out);
  DGK_HIP(hipGetLastError());
  DGK_HIP(hipEventRecord(e1, st));
  DGK_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  DGK_HIP(hipEventElapsedTime(&ms, e0, e1));
  DGK_HIP(hipEventDestroy(e0));
  DGK_HIP(hipEventDestroy(e1));
  *gbs = (double)(NR + 1) * 8.0 * (double)n * reps / (ms * 1e-3) / 1e9;
  return DGK_OK;
}

}  // namespace
}  // namespace dgk

using namespace dgk;

extern "C" int dgk_calibrate_stream(int64_t n, int reps, double *gbs_async,
                                    double *gbs_regs) {
  clear_error();
  if (n < kProbeDoubles || n % kProbeDoubles || reps < 1 || !gbs_async || !gbs_regs)
    return set_error(DGK_ERR_ARG, "n must be a positive multiple of " +
                                      std::to_string(kProbeDoubles) + ", reps >= 1");
  if (n / kProbeSpan<false> > (int64_t)INT32_MAX)
    return set_error(DGK_ERR_ARG, "n too large");
  constexpr int NR = 5;
  double *in = nullptr, *out = nullptr;
  hipStream_t st = nullptr;
  int rc = DGK_OK;
  auto fail = [&](int code, const std::string &m) {
    rc = set_error(code, m);
  };
  if (hipMalloc(&in, sizeof(double) * (size_t)n * NR) != hipSuccess ||
      hipMalloc(&out, sizeof(double) * (size_t)n) != hipSuccess) {
    (void)hipGetLastError();
    fail(DGK_ERR_NOMEM, "hipMalloc of the calibration arrays failed");
  } else if (hipStreamCreate(&st) != hipSuccess ||
             hipMemsetAsync(in, 0, sizeof(double) * (size_t)n * NR, st) != hipSuccess) {
    (void)hipGetLastError();
    fail(DGK_ERR_DEVICE, "stream or memset failed");
  } else {
    rc = time_probe<NR, true>(in, n, out, reps, st, gbs_async);
    if (rc == DGK_OK) rc = time_probe<NR, false>(in, n, out, reps, st, gbs_regs);
  }
  if (st) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  if (in) (void)hipFree(in);
  if (out) (void)hipFree(out);
  return rc;
}
