// Measurement probes for the roofline figures bench.py reports. They are not
// part of the drop-in boundary.
//
//  * odba_probe_cell_mix: the VALU issue rate of the DTW cell update's own
//    instruction mix (v_sub_f32, v_mul_f32, v_min3_f32, v_add_f32 on
//    independent chains) at a given number of 256-thread workgroups per CU.
//    On gfx950 this mix issues above the 64 lane-ops/CU/clock of a single
//    16-lane pass per SIMD, so the ceiling for the swath kernel is taken from
//    this measurement and from the published FP32 vector rate, not from the
//    16-lanes-per-SIMD formula.
//  * odba_probe_path_mix: the same for the path cell (v_sub_f32, v_mul_f32,
//    three v_add_f32, v_min3_f32, two compares into lane masks and two
//    v_addc_co_u32 that shift the step bits into the lane's plane words:
//    10 lane-ops per cell), with the path kernels' own dependencies: 16
//    columns per lane chained through the left value, the left edge from
//    the previous lane by a DPP wave shift, the diagonal from the previous
//    row. It is the ceiling the path kernels are measured against.
//  * odba_probe_copy: a float4 grid-stride copy of a known byte count, the
//    calibration kernel for FETCH_SIZE / WRITE_SIZE counter readings.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kChains = 16;

__global__ void __launch_bounds__(256) k_cell_mix(float* __restrict__ out, int iters, float c1) {
  float x[kChains], e[kChains];
#pragma unroll
  for (int k = 0; k < kChains; k++) {
    x[k] = (float)(threadIdx.x + k) * 1e-3f;
    e[k] = (float)k * 0.25f + 1.0f;
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < kChains; k++) {
      const float d = x[k] - c1;
      const float dd = d * d;
      x[k] = fminf(fminf(x[k], e[k]), dd) + dd;
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < kChains; k++) s += x[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__device__ __forceinline__ float wave_shr1(float v, float old) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xF, 0xF, false));
}

// One step per iteration: every lane updates its 16 columns of one row.
// Lane 0's left edge and its row value come from `seed`, the others' from
// the previous lane.
__global__ void __launch_bounds__(256) k_path_mix(uint32_t* __restrict__ out, int iters, float seed) {
  const int lane = threadIdx.x & 63;
  float x[16], D[16];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    x[q] = (float)(lane + q) * 1e-3f;
    D[q] = (float)q * 0.5f;
  }
  float y = (float)lane * 0.01f, dgin = 0.0f;
  uint32_t acc = 0;
  for (int it = 0; it < iters; ++it) {
    const float lin = wave_shr1(D[15], seed);
    y = wave_shr1(y, y + seed);
    float dg = dgin, lf = lin;
    uint32_t wa = 0, wb = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const float d = y - x[q];
      const float dd = d * d;
      const float U = D[q] + dd, R = lf + dd, G = dg + dd;
      const float v = fminf(fminf(U, R), G);
      const unsigned long long na = __builtin_amdgcn_fcmpf(G, v, 14);  // unordered or not equal
      const unsigned long long nb = __builtin_amdgcn_fcmpf(U, v, 1);   // ordered and equal
      unsigned long long co;
      asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(wa), "=s"(co) : "v"(wa), "s"(na));
      asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(wb), "=s"(co) : "v"(wb), "s"(nb));
      dg = D[q];
      D[q] = v;
      lf = v;
    }
    dgin = lin;
    acc ^= (wb << 16) | wa;
  }
  uint32_t s = acc;
#pragma unroll
  for (int q = 0; q < 16; q++) s ^= __float_as_uint(D[q]);
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n4) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

}  // namespace

extern "C" {

// Runs the cell-mix probe once untimed and once timed. *lane_ops_per_s gets
// the timed launch's rate (4 lane-ops per chain step), *ms its duration.
int odba_probe_cell_mix(int device, int blocks_per_cu, int iters, double* lane_ops_per_s, double* ms_out) {
  if (blocks_per_cu < 1 || blocks_per_cu > 32 || iters < 1) return 1;
  if (hipSetDevice(device) != hipSuccess) return 2;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 2;
  const int blocks = p.multiProcessorCount * blocks_per_cu;
  float* out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return 3;
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipLaunchKernelGGL(k_cell_mix, dim3(blocks), dim3(256), 0, 0, out, 16, 0.5f);
  hipEventRecord(a, 0);
  hipLaunchKernelGGL(k_cell_mix, dim3(blocks), dim3(256), 0, 0, out, iters, 0.5f);
  hipEventRecord(b, 0);
  const hipError_t e = hipEventSynchronize(b);
  float ms = 0.0f;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipFree(out);
  if (e != hipSuccess || hipGetLastError() != hipSuccess) return 4;
  const double ops = (double)blocks * 256.0 * iters * kChains * 4.0;
  if (lane_ops_per_s) *lane_ops_per_s = ops / (ms * 1e-3);
  if (ms_out) *ms_out = ms;
  return 0;
}

// Path cell mix, as odba_probe_cell_mix: *lane_ops_per_s counts 10 lane-ops
// per cell (16 cells per lane and iteration).
int odba_probe_path_mix(int device, int blocks_per_cu, int iters, double* lane_ops_per_s, double* ms_out) {
  if (blocks_per_cu < 1 || blocks_per_cu > 32 || iters < 1) return 1;
  if (hipSetDevice(device) != hipSuccess) return 2;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 2;
  const int blocks = p.multiProcessorCount * blocks_per_cu;
  uint32_t* out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)) != hipSuccess) return 3;
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipLaunchKernelGGL(k_path_mix, dim3(blocks), dim3(256), 0, 0, out, 16, 1e-4f);
  hipEventRecord(a, 0);
  hipLaunchKernelGGL(k_path_mix, dim3(blocks), dim3(256), 0, 0, out, iters, 1e-4f);
  hipEventRecord(b, 0);
  const hipError_t e = hipEventSynchronize(b);
  float ms = 0.0f;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipFree(out);
  if (e != hipSuccess || hipGetLastError() != hipSuccess) return 4;
  const double ops = (double)blocks * 256.0 * iters * 16.0 * 10.0;
  if (lane_ops_per_s) *lane_ops_per_s = ops / (ms * 1e-3);
  if (ms_out) *ms_out = ms;
  return 0;
}

// Copies `bytes` (a multiple of 16) between two fresh buffers: one untimed
// launch, then one timed launch. *ms gets the timed launch's duration.
int odba_probe_copy(int device, uint64_t bytes, double* ms_out) {
  if (bytes == 0 || bytes % 16) return 1;
  if (hipSetDevice(device) != hipSuccess) return 2;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 2;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return 3;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    hipFree(a);
    return 3;
  }
  hipMemset(a, 1, bytes);
  hipMemset(b, 0, bytes);
  const long long n4 = (long long)(bytes / 16);
  const int blocks = p.multiProcessorCount * 16;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n4);
  hipEventRecord(e0, 0);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n4);
  hipEventRecord(e1, 0);
  const hipError_t e = hipEventSynchronize(e1);
  float ms = 0.0f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess || hipGetLastError() != hipSuccess) return 4;
  if (ms_out) *ms_out = ms;
  return 0;
}

}  // extern "C"
