// fs_peak.hip -- issue-rate microbenchmark of the VALU instructions the
// residue kernels are built from, for the roofline the benchmark prices them
// against (no reference counterpart).  Each lane runs 8 independent chains of
// one instruction kind; a launch of 8 workgroups of 256 threads per CU keeps
// 8 waves on every SIMD.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fs_batch_internal.hpp"
#include "fs_internal.hpp"
#include "../../include/factsieve.h"

namespace {

constexpr int kChains = 8;

template <int K>
__global__ __launch_bounds__(256) void k_peak(uint64_t* out, uint32_t iters, uint32_t seed) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t a = 0x9E3779B1u | seed;  // odd multiplier
  if (K == 3) {  // v_fma_f64: d = d * A + B
    double d[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) d[c] = (double)(t + c) * 0x1p-40;
    const double A = 0.999999, B = (double)seed * 0x1p-60;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int c = 0; c < kChains; ++c) d[c] = __builtin_fma(d[c], A, B);
    }
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += d[c];
    out[t] = __builtin_bit_cast(uint64_t, s);
  } else if (K == 0) {  // v_mad_u64_u32: x = lo32(x) * A + x (64-bit accumulate)
    uint64_t x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) x[c] = (uint64_t)t * 0x100000001ull + c;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int c = 0; c < kChains; ++c) x[c] = (uint64_t)(uint32_t)x[c] * a + x[c];
    }
    uint64_t s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s ^= x[c];
    out[t] = s;
  } else {  // K == 1: v_mul_lo_u32, K == 2: v_mul_hi_u32
    uint32_t y[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) y[c] = t * 2654435761u + c + 1;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int c = 0; c < kChains; ++c) y[c] = K == 1 ? y[c] * a : __umulhi(y[c], a);
    }
    uint32_t s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s ^= y[c];
    out[t] = s;
  }
}

}  // namespace

// Lane-operations per second of instruction `kind` (0 v_mad_u64_u32,
// 1 v_mul_lo_u32, 2 v_mul_hi_u32, 3 v_fma_f64) over one timed launch of
// `iters` iterations of 8 chains per lane, after one warm-up launch.
extern "C" int fs_valu_peak(int kind, uint32_t iters, double* lane_ops_per_s, double* ms) {
  if (kind < 0 || kind > 3 || iters == 0 || !lane_ops_per_s) return FS_E_ARG;
  int rc = fs_init(0);
  if (rc) return rc;
  const uint32_t blocks = fs_device_cus() * 8, threads = blocks * 256;
  uint64_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&out, (size_t)threads * 8);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float t_ms = 0;
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
    e = hipEventRecord(e0, 0);
    switch (kind) {
      case 0: k_peak<0><<<blocks, 256>>>(out, iters, 1u); break;
      case 1: k_peak<1><<<blocks, 256>>>(out, iters, 1u); break;
      case 2: k_peak<2><<<blocks, 256>>>(out, iters, 1u); break;
      default: k_peak<3><<<blocks, 256>>>(out, iters, 1u); break;
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_ms, e0, e1);
  }
  if (out) (void)hipFree(out);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) {
    fs_set_hip_error("valu peak", (int)e);
    return FS_E_HIP;
  }
  *lane_ops_per_s = (double)threads * iters * kChains / (t_ms * 1e-3);
  if (ms) *ms = t_ms;
  return FS_OK;
}
