// Measurement probes: a plain copy kernel (calibrates rocprof's FETCH_SIZE /
// WRITE_SIZE against a known byte count) and an LDS read-bandwidth kernel
// (measures the LDS peak the hop-loop kernel is compared with).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "kmc_types.h"

namespace kmc {

__global__ void __launch_bounds__(256) kmc_probe_copy_kernel(const float4* __restrict__ src,
                                                            float4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ds_read_b128 stream: each lane reads 16 B at a fixed base plus 16 immediate
// offsets 4 KiB apart (a wave reads 1 KiB contiguous per load, no bank
// conflicts, no address arithmetic); the empty asm makes the base opaque so the
// loads are issued in every iteration.
__global__ void __launch_bounds__(256) kmc_probe_lds_b128_kernel(int iters, uint32_t* out) {
  __shared__ __align__(16) uint4 buf[4096];   // 64 KiB
  for (int i = threadIdx.x; i < 4096; i += 256) buf[i] = make_uint4(i, 3 * i, 5 * i, 7 * i);
  __syncthreads();
  uint4 acc = make_uint4(0, 0, 0, 0);
  int base = threadIdx.x;
  for (int it = 0; it < iters; ++it) {
    asm volatile("" : "+v"(base));   // an index, so the loads stay LDS loads
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint4 v = buf[base + k * 256];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
  }
  if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x12345678u) out[blockIdx.x] = acc.x;   // keeps the loads live
}

// Each thread streams WIDTH-byte LDS reads over a 64 KiB buffer `iters` times.
template <int WIDTH>
__global__ void __launch_bounds__(256) kmc_probe_lds_kernel(int iters, uint32_t* out) {
  __shared__ __align__(16) uint32_t buf[16384];
  for (int i = threadIdx.x; i < 16384; i += 256) buf[i] = i * 2654435761u;
  __syncthreads();
  constexpr int WORDS = WIDTH / 4;
  uint32_t acc = 0;
  int base = threadIdx.x * WORDS;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int idx = (base + k * 256 * WORDS) & (16384 - WORDS);
      if constexpr (WORDS == 4) {
        const uint4 v = *reinterpret_cast<const uint4*>(&buf[idx]);
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
      } else if constexpr (WORDS == 2) {
        const uint2 v = *reinterpret_cast<const uint2*>(&buf[idx]);
        acc ^= v.x ^ v.y;
      } else {
        acc ^= buf[idx];
      }
    }
    base += 8 * 256 * WORDS + WORDS;
  }
  if (acc == 0x12345678u) out[blockIdx.x] = acc;   // keeps the loads live
}

// Dependent-instruction latency: one wave runs `iters` x 32 instructions that
// each depend on the previous one, timed with the shader clock.  kind 0:
// v_add_f64 on all lanes; 1: v_add_f64 on lane 0 alone; 2: v_add_f32 on all
// lanes; 3: v_fma_f64 on all lanes; 4: four independent v_add_f64 chains
// interleaved (issue rate, per instruction).
#define KMC_X4(s) s s s s
#define KMC_X32(s) KMC_X4(KMC_X4(s)) KMC_X4(KMC_X4(s))
__global__ void __launch_bounds__(64) kmc_probe_chain_kernel(int kind, int iters, double* out) {
  double a = 1.0 + threadIdx.x * 1e-3, b = 1e-9, c = 2.0, d = 3.0, e = 4.0;
  float f = 1.0f + threadIdx.x, g = 1e-7f;
  asm volatile("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d), "+v"(e), "+v"(f), "+v"(g));
  uint64_t t0 = 0, t1 = 0;
  if (kind == 1 && threadIdx.x != 0) return;
  t0 = __builtin_amdgcn_s_memtime();
  for (int it = 0; it < iters; ++it) {
    if (kind <= 1)
      asm volatile(KMC_X32("v_add_f64 %0, %0, %1\n\t") : "+v"(a) : "v"(b));
    else if (kind == 2)
      asm volatile(KMC_X32("v_add_f32 %0, %0, %1\n\t") : "+v"(f) : "v"(g));
    else if (kind == 3)
      asm volatile(KMC_X32("v_fma_f64 %0, %0, %1, %1\n\t") : "+v"(a) : "v"(b));
    else
      asm volatile(KMC_X4(KMC_X4("v_add_f64 %0, %0, %4\n\tv_add_f64 %1, %1, %4\n\t"
                                 "v_add_f64 %2, %2, %4\n\tv_add_f64 %3, %3, %4\n\t"))
                   : "+v"(a), "+v"(c), "+v"(d), "+v"(e) : "v"(b));
  }
  t1 = __builtin_amdgcn_s_memtime();
  if (threadIdx.x == 0) {
    out[0] = (double)(t1 - t0) / ((double)iters * (kind == 4 ? 64 : 32));
    out[1] = a + c + d + e + (double)f;
  }
}
#undef KMC_X32
#undef KMC_X4

}  // namespace kmc

using namespace kmc;

extern "C" {

int kmc_probe_copy(int64_t bytes, int repeats, double* ms_per_copy) {
  if (bytes <= 0 || repeats <= 0) return KMC_ERR_ARGS;
  const size_t n = (size_t)bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess) return KMC_ERR_NOMEM;
  if (hipMalloc(&b, n * sizeof(float4)) != hipSuccess) {
    (void)hipFree(a);
    return KMC_ERR_NOMEM;
  }
  (void)hipMemset(a, 1, n * sizeof(float4));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(kmc_probe_copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < repeats; ++r)
    hipLaunchKernelGGL(kmc_probe_copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
  (void)hipEventRecord(e1, 0);
  hipError_t st = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (ms_per_copy) *ms_per_copy = ms / repeats;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return st == hipSuccess ? KMC_OK : KMC_ERR_DEVICE;
}

int kmc_probe_lds(int width_bytes, int blocks, int iters, double* bytes_per_s) {
  if (blocks <= 0 || iters <= 0) return KMC_ERR_ARGS;
  uint32_t* out = nullptr;
  if (hipMalloc(&out, sizeof(uint32_t) * blocks) != hipSuccess) return KMC_ERR_NOMEM;
  // width 0: the b128 stream (16 loads of 16 B per lane and iteration)
  void (*k)(int, uint32_t*) = width_bytes == 0    ? kmc_probe_lds_b128_kernel
                              : width_bytes == 16 ? kmc_probe_lds_kernel<16>
                              : width_bytes == 8  ? kmc_probe_lds_kernel<8>
                                                  : kmc_probe_lds_kernel<4>;
  const double per_iter = width_bytes == 0 ? 16.0 * 16 : 8.0 * (width_bytes == 16 ? 16 : (width_bytes == 8 ? 8 : 4));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, 1, out);
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, iters, out);
  (void)hipEventRecord(e1, 0);
  hipError_t st = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (bytes_per_s) *bytes_per_s = (double)blocks * 256 * per_iter * (double)iters / (ms * 1e-3);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return st == hipSuccess ? KMC_OK : KMC_ERR_DEVICE;
}

int kmc_probe_chain(int kind, int iters, double* cycles_per_instruction) {
  if (kind < 0 || kind > 4 || iters <= 0 || !cycles_per_instruction) return KMC_ERR_ARGS;
  double* out = nullptr;
  if (hipMalloc(&out, 2 * sizeof(double)) != hipSuccess) return KMC_ERR_NOMEM;
  hipLaunchKernelGGL(kmc_probe_chain_kernel, dim3(1), dim3(64), 0, 0, kind, 16, out);
  hipLaunchKernelGGL(kmc_probe_chain_kernel, dim3(1), dim3(64), 0, 0, kind, iters, out);
  double h[2] = {0, 0};
  hipError_t st = hipMemcpy(h, out, sizeof h, hipMemcpyDeviceToHost);
  (void)hipFree(out);
  *cycles_per_instruction = h[0];
  return st == hipSuccess ? KMC_OK : KMC_ERR_DEVICE;
}

}  // extern "C"
