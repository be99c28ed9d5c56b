// Adaptive k-segmentation for gfx950 (adaptive_device_segmentation<T>,
// segmentation.hpp:48-337): one 256-thread workgroup per (sequence, chunk of
// 256*ds samples). Same arithmetic as the reference, in the same integer
// types; the Bellman DP is reorganised: for a fixed segment count p every
// right end n is independent, so each thread owns one n and scans its
// candidate left ends itself (first strict minimum, ascending), instead of
// one thread scanning after a barrier per n. The binary search over k, the
// noise test and the per-segment medians follow the reference step by step.

#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include <stdint.h>

#include "odba_dtw.hpp"

namespace odba {

template <typename T> struct KL;
template <> struct KL<float> {
  static constexpr bool kFloat = true;
  __device__ static float mn() { return FLT_MIN; }
  __device__ static float mx() { return FLT_MAX; }
};
template <> struct KL<double> {
  static constexpr bool kFloat = true;
  __device__ static double mn() { return DBL_MIN; }
  __device__ static double mx() { return DBL_MAX; }
};
template <> struct KL<int32_t> {
  static constexpr bool kFloat = false;
  __device__ static int32_t mn() { return INT_MIN; }
  __device__ static int32_t mx() { return INT_MAX; }
};
template <> struct KL<uint32_t> {
  static constexpr bool kFloat = false;
  __device__ static uint32_t mn() { return 0u; }
  __device__ static uint32_t mx() { return UINT_MAX; }
};
template <> struct KL<uint64_t> {
  static constexpr bool kFloat = false;
  __device__ static uint64_t mn() { return 0ull; }
  __device__ static uint64_t mx() { return ULLONG_MAX; }
};

// 32-lane shuffle-down reduction with the reference's comparison
// (warpReduceMax / warpReduceMin, cuda_utils.hpp:20-40): lanes past the
// group keep their own value, NaN never replaces an accumulator.
template <typename T, bool MAX> __device__ __forceinline__ T red32(T v) {
  for (int off = 16; off > 0; off >>= 1) {
    T t = __shfl_down(v, off, 32);
    if (MAX) { if (t > v) v = t; }
    else { if (t < v) v = t; }
  }
  return v;
}

__device__ __forceinline__ unsigned char d2u8(double d) {
  if (d != d || d <= 0.0) return 0;
  if (d >= 255.0) return 255;
  return (unsigned char)d;
}

__host__ __device__ inline long long kseg_align(long long x, long long a) { return (x + a - 1) / a * a; }

// Extent e of the DP search: bytes left after the reference's shared-memory
// layout up to `squares`, divided by 6 bytes per downsampled point.
__host__ __device__ inline int kseg_extent(int mk, int N_ds, int tsize, int lds_bytes) {
  long long off = (long long)(mk + 1) * 2;
  off = kseg_align(off, 4) + 4LL * (N_ds + 1);
  off += 4LL * 2 * (N_ds + 1);
  off = kseg_align(off, tsize) + (long long)tsize * (N_ds + 2);
  off = kseg_align(off, 4);
  return N_ds > 0 ? (int)(((long long)lds_bytes - off) / (6LL * N_ds)) : 0;
}

template <typename T>
__global__ void __launch_bounds__(256) k_kseg(const T* __restrict__ data, const long long* __restrict__ offs,
                                              const int* __restrict__ lens, const long long* __restrict__ out_base,
                                              const long long* __restrict__ path_base, int spb, int mk, int minseg,
                                              int lds_bytes, unsigned short* __restrict__ kpath_all,
                                              T* __restrict__ out) {
  extern __shared__ unsigned char smem[];
  __shared__ T dq[256];
  __shared__ T wmx[8], wmn[8];
  __shared__ T s_max, s_min;
  __shared__ uint32_t kdist[2 * 257];
  __shared__ unsigned short bp[258];
  __shared__ int s_flag;

  const int seq = blockIdx.x, chunk = blockIdx.y, t = threadIdx.x;
  const int orig_N = lens[seq];
  if (chunk * spb >= orig_N) return;
  const int ds = (minseg + 2) / 3;
  const T* ser = data + offs[seq] + (long long)chunk * spb;
  const int N = ((chunk + 1) * spb > orig_N) ? orig_N % spb : spb;
  int expected_k = mk;
  if ((chunk + 1) * spb > orig_N) {
    const int per = (spb + mk - 1) / mk;
    expected_k = ((orig_N % spb) + per - 1) / per;
  }
  const int N_ds = N / ds;
  const int e = kseg_extent(mk, N_ds, (int)sizeof(T), lds_bytes);
  uint32_t* squares = (uint32_t*)smem;
  unsigned short* sums = (unsigned short*)(squares + (size_t)N_ds * e);
  unsigned short* kpath = kpath_all + path_base[seq] + (long long)chunk * (256LL * mk);
  T* outp = out + out_base[seq] + (long long)chunk * mk;

  // downsample, with the reference's loop bound ds*t + t + i < N (segmentation.hpp:112-121)
  if (t * ds <= N - ds) {
    T acc = T(0);
    int cnt = 0;
    for (int i = 0; i < ds && ds * t + t + i < N; i++) { acc = (T)(acc + ser[ds * t + i]); cnt++; }
    // 0/0 gives NaN for floating types; integer 0/0 has no defined value in
    // the reference (parity unpinned) and is taken as 0 here.
    if (cnt) dq[t] = (T)(acc / (T)cnt);
    else dq[t] = KL<T>::kFloat ? (T)(acc / (T)cnt) : T(0);
  }
  __syncthreads();
  // min / max over full warps of 32 (segmentation.hpp:126-150)
  {
    T vmx = t < N_ds ? dq[t] : KL<T>::mn();
    T vmn = t < N_ds ? dq[t] : KL<T>::mx();
    vmn = red32<T, false>(vmn);
    vmx = red32<T, true>(vmx);
    const int lane = t & 31, wid = t >> 5;
    if (!lane && wid < N_ds / 32) { wmx[wid] = vmx; wmn[wid] = vmn; }
    __syncthreads();
    if (!wid) {
      T a = lane < N_ds / 32 ? wmx[lane] : KL<T>::mn();
      a = red32<T, true>(a);
      T b = lane < N_ds / 32 ? wmn[lane] : KL<T>::mx();
      b = red32<T, false>(b);
      if (!lane) { s_max = a; s_min = b; }
    }
    __syncthreads();
  }
  const T Tmax = s_max, Tmin = s_min;
  auto SUMS = [&](int l, int r) -> unsigned short& { return sums[(r - l) + e * l]; };
  auto SQ = [&](int l, int r) -> uint32_t& { return squares[(r - l) + e * l]; };
  if (t < N_ds) {
    const T q = (T)((T)(dq[t] - Tmin) / (T)(Tmax - Tmin + 1));
    const unsigned short v = d2u8(256.0 * (double)q);
    SUMS(t, t) = v;
    SQ(t, t) = (uint32_t)v * (uint32_t)v;
  }
  __syncthreads();
  if (t < N_ds) {
    for (int sp = 1; sp < e; sp++) {
      if (t + sp < N_ds) {
        SUMS(t, t + sp) = (unsigned short)(SUMS(t, t + sp - 1) + SUMS(t + sp, t + sp));
        SQ(t, t + sp) = SQ(t, t + sp - 1) + SQ(t + sp, t + sp);
      }
    }
  }
  __syncthreads();
  auto DIST = [&](int l, int r) -> uint32_t {  // 1-based (segmentation.hpp:31-35)
    const int a = l - 1, b = r - 1;
    const int s = SUMS(a, b);
    const int ms = s * s / (b - a + 1);
    return SQ(a, b) - (uint32_t)ms;
  };

  short smallest_noisy = (short)(expected_k + 1);
  short test_k = (short)((expected_k + 1) / 2);
  short delta = (short)(test_k / 2);
  bool ex = false;
  while (!ex) {
    if (t < N_ds) {
      kdist[N_ds * 1 + t + 1] = (t + 1 <= e) ? DIST(1, t + 1) : (uint32_t)(INT_MAX / 2);
      if (t < test_k) kpath[(long long)N_ds * t + t] = (unsigned short)(test_k - 1);
    }
    __syncthreads();
    for (int p = 2; p <= test_k; p++) {
      const int n = t + 1;
      uint32_t best = 0;
      int bidx = -1;
      if (n >= p && n <= N_ds) {
        int minval = INT_MAX, minidx = -1;
        const int start = (n - e + 1 > p) ? n - e + 1 : p;
        for (int idx = start; idx - start < e && idx <= n; idx++) {
          const uint32_t o = kdist[N_ds * ((p - 1) & 1) + idx - 1] + DIST(idx, n);
          if (o < (uint32_t)minval) { minval = (int)o; minidx = idx; }
        }
        best = (uint32_t)minval;
        bidx = minidx;
      }
      __syncthreads();  // all reads of row p-1 done before row p (same slot as p-2) is written
      if (n >= p && n <= N_ds) {
        kdist[N_ds * (p & 1) + n] = best;
        kpath[(long long)N_ds * (p - 1) + n - 1] = (unsigned short)((unsigned short)bidx - 1);
      }
      __syncthreads();
    }
    if (t == 0) {
      bp[test_k] = (unsigned short)N_ds;
      for (int p = test_k - 1; p >= 1; p--) {
        const int nn = bp[p + 1] - 1;
        bp[p] = (nn >= 0 && nn < N_ds) ? kpath[(long long)N_ds * p + nn] : (unsigned short)0;
      }
      bp[0] = 0;
      bool noisy = false;
      for (int i = 1; i <= test_k; i++)
        if ((int)bp[i] - (int)bp[i - 1] < minseg / ds) noisy = true;
      s_flag = noisy ? 1 : 0;
    }
    __threadfence_block();
    __syncthreads();
    // identical in every thread; made wave-uniform so the loop below stays
    // one uniform loop around its barriers
    const bool noisy = __builtin_amdgcn_readfirstlane(s_flag) != 0;
    if (delta == 0 || test_k + delta > mk) ex = true;
    else if (test_k != 1 && noisy) { smallest_noisy = test_k; test_k = (short)(test_k - delta); }
    else {
      if (test_k == 1 || smallest_noisy == test_k + 1) ex = true;
      else test_k = (short)(test_k + delta);
    }
    delta = (short)((delta + 1) / 2);
    __syncthreads();
  }
  // medians (segmentation.hpp:280-333); test_k here is the last one evaluated
  // in the loop, whose breakpoints are in bp[] ... the loop body above has
  // already moved test_k on exit only through the `ex` branches, which leave
  // it unchanged.
  T* orig = (T*)smem;  // the DP tables are no longer needed (the reference reuses the same space)
  __syncthreads();
  for (int i = 0; i < ds; i++) {
    const int pos = t * ds + i;
    if (pos < spb && chunk * spb + pos < orig_N) orig[pos] = ser[pos];
  }
  // Past the chunk's samples the buffer reads as zero, up to spb + 2 values:
  // an even-length median reads one or two values past its segment's end
  // (segmentation.hpp:325-330), and a chunk of fewer samples than ds has an
  // empty segment whose median reads orig[0] and orig[1].
  for (int pos = N + t; pos < spb + 2; pos += 256) orig[pos] = T(0);
  __syncthreads();
  // A median that reads past its segment sees the next segment's values as
  // loaded, before that segment is sorted (the segments are taken one after
  // the other): every thread reads its two values past the end now.
  const bool seg = t > 0 && t <= test_k;
  int left = 0, right = 0;
  T past0 = T(0), past1 = T(0);
  if (seg) {
    right = bp[t] * ds;
    if (right > N) right = N;
    left = bp[t - 1] * ds;
    past0 = orig[right];
    past1 = orig[right + 1];
  }
  __syncthreads();
  if (t > test_k && t <= mk) {
    outp[t - 1] = KL<T>::mx();
  } else if (seg) {
    for (int i = left; i < right - 1; ++i)
      for (int j = i + 1; j < right; ++j)
        if (orig[i] > orig[j]) { T tmp = orig[i]; orig[i] = orig[j]; orig[j] = tmp; }
    auto at = [&](int i) -> T { return i < right ? orig[i] : (i == right ? past0 : past1); };
    const int mid = (left + right) / 2;
    if ((left - right) % 2) outp[t - 1] = at(mid);
    else outp[t - 1] = (T)((at(mid) + at(mid + 1)) / T(2));
  }
}

template <typename T>
static hipError_t kseg_t(hipStream_t st, const void* data, const long long* offs, const int* lens, int n, int maxch,
                         const long long* out_base, const long long* path_base, int spb, int mk, int minseg,
                         int lds_bytes, unsigned short* kpath, void* out, size_t dyn) {
  if (dyn > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)k_kseg<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_kseg<T>, dim3(n, maxch), dim3(256), dyn, st, (const T*)data, offs, lens, out_base, path_base,
                     spb, mk, minseg, lds_bytes, kpath, (T*)out);
  return hipGetLastError();
}

hipError_t launch_kseg(int dtype, hipStream_t st, const void* data, const long long* offs, const int* lens, int n,
                       int maxch, const long long* out_base, const long long* path_base, int spb, int mk, int minseg,
                       int lds_bytes, unsigned short* kpath, void* out, size_t dyn) {
  switch (dtype) {
    case ODBA_F32: return kseg_t<float>(st, data, offs, lens, n, maxch, out_base, path_base, spb, mk, minseg, lds_bytes, kpath, out, dyn);
    case ODBA_F64: return kseg_t<double>(st, data, offs, lens, n, maxch, out_base, path_base, spb, mk, minseg, lds_bytes, kpath, out, dyn);
    case ODBA_I32: return kseg_t<int32_t>(st, data, offs, lens, n, maxch, out_base, path_base, spb, mk, minseg, lds_bytes, kpath, out, dyn);
    case ODBA_U32: return kseg_t<uint32_t>(st, data, offs, lens, n, maxch, out_base, path_base, spb, mk, minseg, lds_bytes, kpath, out, dyn);
    case ODBA_U64: return kseg_t<uint64_t>(st, data, offs, lens, n, maxch, out_base, path_base, spb, mk, minseg, lds_bytes, kpath, out, dyn);
  }
  return hipErrorInvalidValue;
}

int kseg_extent_host(int mk, int N_ds, int tsize, int lds_bytes) { return kseg_extent(mk, N_ds, tsize, lds_bytes); }

}  // namespace odba
