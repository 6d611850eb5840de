// Patch matrix of an NCHW image (im2col) for chains whose B comes from an image.
// The fused chain kernel gathers the same elements itself while it stages its B tile
// (chain_kernel.hip, stage_b); this kernel forms B in memory for the unfused paths.
#include "fg_internal.h"

namespace fg {
namespace {

// A thread forms V = 16 / sizeof(E) consecutive columns of B (one 16-byte store per
// row when N % V == 0) for RB consecutive rows: the columns' (image, oy, ox) and the
// rows' (c, ky, kx) are decomposed once and then stepped. Element offsets are 32-bit
// (the padded image has fewer than 2^31 elements, checked by the caller).
constexpr int RB = 8;
template <typename E>
__global__ __launch_bounds__(256) void im2col_kernel(const E *__restrict__ X, E *__restrict__ B, ImgGeom g, long K1,
                                                     long N) {
  constexpr int V = 16 / sizeof(E);
  const long n0 = ((long)blockIdx.x * 256 + threadIdx.x) * V;
  if (n0 >= N) return;
  const int ohw = g.OH * g.OW, HW = g.H * g.W;
  int y0[V], x0[V], base[V];
  {
    int b = (int)(n0 / ohw), r = (int)(n0 - (long)b * ohw);
    int oy = r / g.OW, ox = r - oy * g.OW;
#pragma unroll
    for (int j = 0; j < V; ++j) {
      y0[j] = oy * g.sh - g.ph;
      x0[j] = ox * g.sw - g.pw;
      base[j] = b * g.C * HW;
      if (++ox == g.OW) {
        ox = 0;
        if (++oy == g.OH) {
          oy = 0;
          ++b;
        }
      }
    }
  }
  const int kk = g.kh * g.kw;
  const bool full = n0 + V <= N && N % V == 0 && ((uintptr_t)B & 15) == 0;
  for (long k0 = (long)blockIdx.y * RB; k0 < K1; k0 += (long)gridDim.y * RB) {
  int c = (int)(k0 / kk), rr = (int)(k0 - (long)c * kk);
  int ky = rr / g.kw, kx = rr - ky * g.kw;
  for (int q = 0; q < RB && k0 + q < K1; ++q) {
    const int dy = ky * g.dh, dx = kx * g.dw;
    E v[V];
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const int iy = y0[j] + dy, ix = x0[j] + dx;
      // columns past N (the tail thread) belong to no image: never loaded
      v[j] = (n0 + j < N && (unsigned)iy < (unsigned)g.H && (unsigned)ix < (unsigned)g.W)
                 ? X[base[j] + c * HW + iy * g.W + ix]
                 : E(0);
    }
    E *row = B + (k0 + q) * N + n0;
    if (full) {
      uint4 w;
      __builtin_memcpy(&w, v, 16);
      *(uint4 *)row = w;
    } else {
      for (int j = 0; j < V && n0 + j < N; ++j) row[j] = v[j];
    }
    if (++kx == g.kw) {
      kx = 0;
      if (++ky == g.kh) {
        ky = 0;
        ++c;
      }
    }
  }
  }
}

template <typename E>
hipError_t launch_t(const void *X, void *B, const ImgGeom &g, long K1, long N, hipStream_t s) {
  constexpr long V = 16 / sizeof(E);
  const long ygroups = (K1 + RB - 1) / RB;  // row groups, grid-strided past 65535
  dim3 grid((unsigned)((N + 256 * V - 1) / (256 * V)), (unsigned)(ygroups < 65535 ? ygroups : 65535));
  hipLaunchKernelGGL(im2col_kernel<E>, grid, dim3(256), 0, s, (const E *)X, (E *)B, g, K1, N);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_im2col(int esz, const void *X, void *B, const ImgGeom &g, long K1, long N, hipStream_t s) {
  if (K1 <= 0 || N <= 0 || (N + 255) / 256 > 0x7fffffffL) return hipErrorInvalidValue;
  switch (esz) {
    case 4: return launch_t<uint32_t>(X, B, g, K1, N, s);
    case 2: return launch_t<uint16_t>(X, B, g, K1, N, s);
    case 1: return launch_t<uint8_t>(X, B, g, K1, N, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace fg
