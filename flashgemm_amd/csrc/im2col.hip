// Patch matrix of an NCHW image (im2col) for chains whose B comes from an image.
// The fused chain kernel gathers the same elements itself while it stages its B tile
// (chain_kernel.hip, stage_b); this kernel forms B in memory for the unfused paths.
#include "fg_internal.h"

namespace fg {
namespace {

// One thread per column n of B (coalesced stores along n), a grid-stride loop over
// the rows k; element offsets are 32-bit (the padded image has fewer than 2^31
// elements, checked by the caller).
template <typename E>
__global__ __launch_bounds__(256) void im2col_kernel(const E *__restrict__ X, E *__restrict__ B, ImgGeom g, long K1,
                                                     long N) {
  const long n = (long)blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const int ohw = g.OH * g.OW;
  const int b = (int)(n / ohw), r = (int)(n - (long)b * ohw);
  const int oy = r / g.OW, ox = r - oy * g.OW;
  const int y0 = oy * g.sh - g.ph, x0 = ox * g.sw - g.pw;
  const int kk = g.kh * g.kw;
  for (long k = blockIdx.y; k < K1; k += gridDim.y) {
    const int c = (int)(k / kk), rr = (int)(k - (long)c * kk);
    const int ky = rr / g.kw, kx = rr - ky * g.kw;
    const int iy = y0 + ky * g.dh, ix = x0 + kx * g.dw;
    E v = E(0);
    if ((unsigned)iy < (unsigned)g.H && (unsigned)ix < (unsigned)g.W) v = X[((b * g.C + c) * g.H + iy) * g.W + ix];
    B[k * N + n] = v;
  }
}

template <typename E>
hipError_t launch_t(const void *X, void *B, const ImgGeom &g, long K1, long N, hipStream_t s) {
  dim3 grid((unsigned)((N + 255) / 256), (unsigned)(K1 < 65535 ? K1 : 65535));
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
