// Measurement kernels, never called on the product path: the FP64 oracle on the GPU
// (a plain FP64 GEMM and relative Frobenius error norms, used to measure the error of
// the FP32 / BF16 / split-BF16 paths at full size) and the L2 -> CU stream-rate
// kernel behind the fused chains' bandwidth bound.
#include "fg_internal.h"

namespace fg {
namespace {

constexpr int T64 = 64;  // output tile edge
constexpr int K64 = 16;  // k per stage

// 256 threads, each computes a 4 x 4 block of a 64 x 64 tile, fma in fp64.
__global__ __launch_bounds__(256) void f64_gemm_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                                       double *__restrict__ C, long M, long N, long K) {
  __shared__ double sA[K64][T64 + 1];
  __shared__ double sB[K64][T64 + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const long m0 = (long)blockIdx.y * T64, n0 = (long)blockIdx.x * T64;
  double acc[4][4] = {};
  for (long k0 = 0; k0 < K; k0 += K64) {
    for (int e = threadIdx.x; e < K64 * T64; e += 256) {
      int kk = e % K64, mm = e / K64;
      long m = m0 + mm, k = k0 + kk;
      sA[kk][mm] = (m < M && k < K) ? A[(size_t)m * K + k] : 0.0;
      int nn = e % T64, kb = e / T64;
      long n = n0 + nn, k2 = k0 + kb;
      sB[kb][nn] = (n < N && k2 < K) ? B[(size_t)k2 * N + n] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < K64; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sA[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sB[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      long m = m0 + ty * 4 + i, n = n0 + tx * 4 + j;
      if (m < M && n < N) C[(size_t)m * N + n] = acc[i][j];
    }
}

__device__ inline double load_as_f64(int kind, const void *p, size_t e) {
  switch (kind) {
    case 0: return (double)((const float *)p)[e];
    case 1: return (double)bf16_to_f32(((const uint16_t *)p)[e]);
    case 2: return (double)((const int *)p)[e];
    default: return ((const double *)p)[e];
  }
}

__global__ void to_f64_kernel(int kind, const void *src, double *dst, long n) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (size_t)n) dst[e] = load_as_f64(kind, src, e);
}

constexpr int NBLK = 1024;

// Pass 1: each block reduces a fixed stride of elements into (sum (x-ref)^2, sum ref^2).
__global__ __launch_bounds__(256) void err_partial_kernel(int kind, const void *x, const double *ref, long n,
                                                          double *part) {
  __shared__ double s0[256], s1[256];
  double d2 = 0.0, r2 = 0.0;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < (size_t)n; e += (size_t)NBLK * 256) {
    double r = ref[e], d = load_as_f64(kind, x, e) - r;
    d2 += d * d;
    r2 += r * r;
  }
  s0[threadIdx.x] = d2;
  s1[threadIdx.x] = r2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

// Pass 2: one thread adds the block partials in block order.
__global__ void err_final_kernel(const double *part, double *out2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int i = 0; i < NBLK; ++i) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  out2[0] = a;
  out2[1] = b;
}

// The fused chains' weight stream on its own: wave w of every workgroup reads its
// own contiguous slice of a region all workgroups share (8 slices), 16 B per lane,
// 16 loads of 1 KB in flight per wave, `rounds` x 16 loads per wave. The loaded
// values are folded into one word per lane so no load is dead.
__global__ __launch_bounds__(512) void stream_kernel(const uint4 *__restrict__ region, unsigned slice_frags,
                                                     int rounds, unsigned *__restrict__ out) {
  constexpr int D = 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint4 *base = region + (size_t)wave * slice_frags * 64 + lane;
  unsigned f = 0;
  uint4 v[D], acc = make_uint4(0u, 0u, 0u, 0u);
  // every component feeds its own accumulator, so each load stays one 16-byte load
  auto fold = [&](const uint4 &x) {
    acc.x ^= x.x;
    acc.y += x.y;
    acc.z ^= x.z;
    acc.w += x.w;
  };
#pragma unroll
  for (int d = 0; d < D; ++d) {
    v[d] = base[(size_t)f * 64];
    f = f + 1 == slice_frags ? 0 : f + 1;
  }
  for (int it = 0; it < rounds; ++it) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      fold(v[d]);
      v[d] = base[(size_t)f * 64];
      f = f + 1 == slice_frags ? 0 : f + 1;
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) fold(v[d]);
  out[(size_t)blockIdx.x * 512 + threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

// Seeded operands for timing runs (the autotuner): element i of a buffer from
// splitmix64(seed + i): fp32 uniform in [-1, 1), bf16 the same rounded down to bf16,
// int8 uniform in [-127, 127]. Timing on zeros would not load the MFMA pipes (nor
// the clock) the way real data does.
__global__ void fill_random_kernel(int esz, void *p, long n, unsigned long long seed) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)n; i += (size_t)gridDim.x * blockDim.x) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float u = (float)(z >> 40) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
    if (esz == 4) ((float *)p)[i] = u;
    else if (esz == 2) ((uint16_t *)p)[i] = (uint16_t)(__builtin_bit_cast(uint32_t, u) >> 16);
    else ((int8_t *)p)[i] = (int8_t)((int)(z % 255u) - 127);
  }
}

}  // namespace

hipError_t launch_fill_random(int esz, void *p, long n, unsigned long long seed, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
  hipLaunchKernelGGL(fill_random_kernel, dim3((unsigned)blocks), dim3(256), 0, s, esz, p, n, seed);
  return hipGetLastError();
}

// region: 8 * slice_frags KB, out: workgroups * 512 words.
hipError_t launch_stream_probe(const void *region, unsigned slice_frags, int rounds, int workgroups, unsigned *out,
                               hipStream_t s) {
  hipLaunchKernelGGL(stream_kernel, dim3((unsigned)workgroups), dim3(512), 0, s, (const uint4 *)region, slice_frags,
                     rounds, out);
  return hipGetLastError();
}

hipError_t launch_f64_gemm(const double *A, const double *B, double *C, long M, long N, long K, hipStream_t s) {
  dim3 grid((unsigned)((N + T64 - 1) / T64), (unsigned)((M + T64 - 1) / T64));
  hipLaunchKernelGGL(f64_gemm_kernel, grid, dim3(256), 0, s, A, B, C, M, N, K);
  return hipGetLastError();
}

hipError_t launch_to_f64(int src_kind, const void *src, double *dst, long n, hipStream_t s) {
  hipLaunchKernelGGL(to_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src_kind, src, dst, n);
  return hipGetLastError();
}

// out2 must hold 2 doubles followed by 2 * 1024 doubles of scratch.
hipError_t launch_err_norms(int x_kind, const void *x, const double *ref, long n, double *out2, hipStream_t s) {
  double *part = out2 + 2;
  hipLaunchKernelGGL(err_partial_kernel, dim3(NBLK), dim3(256), 0, s, x_kind, x, ref, n, part);
  hipLaunchKernelGGL(err_final_kernel, dim3(1), dim3(64), 0, s, part, out2);
  return hipGetLastError();
}

}  // namespace fg
