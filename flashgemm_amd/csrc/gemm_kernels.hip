// Tiled MFMA GEMM for gfx950: C(M x N) = A(M x K) . B(K x N), row-major.
//
// Replaces the reference's packing micro-kernels and register-blocked AVX-512
// kernels (singlegemm/src/kernel.h, kernelMGN.h, PACK_x86.h, PACK_B.h;
// multigemm/src/kernel_*.h) with one workgroup-level design:
//   * two tile shapes: 256 threads = 4 waves in a 2 x 2 grid on a 128 x 128 output
//     tile (each wave 64 x 64 = 2 x 2 blocks of 32 x 32, one MFMA tile each, or four
//     16 x 16 ones for FP32), and for large problems
//     512 threads = 8 waves in a 2 x 4 grid on a 256 x 256 tile (each wave
//     128 x 64 = 4 x 2 blocks), which halves the operand bytes per MFMA; a single
//     FP32 GEMM with at most one 128 x 128 tile per CU (2048^3) runs each tile with
//     two groups of 4 waves splitting its k stages (512 threads, two waves per SIMD),
//     the halves added in a fixed order through LDS;
//   * A and B staged global -> registers -> LDS, double buffered, in the MFMA
//     operand layout: 16-byte k-groups per row (A) or per column (B), placed by a
//     row and group permutation (slot()) under which the staging writes and the
//     fragment reads hit distinct banks;
//   * kinds: FP32 (v_mfma_f32_16x16x4_f32), BF16 (v_mfma_f32_32x32x16_bf16),
//     INT8 (v_mfma_i32_32x32x32_i8 with the u8 bias identity
//     u8(a).b = s8(a ^ 0x80).b + 128.b, the 128.sum_k b term added in the
//     epilogue), and FP32 split into bf16 hi/lo planes with three products.
//   * epilogues: store C, requantize into the next GEMM's operand (fp32 / bf16
//     RNE / int8 saturate), or write a split-K partial slab.
#include <type_traits>

#include "fg_internal.h"
#include "fg_runtime.h"

namespace fg {
namespace {

// Tile configurations: BT x BT output tile, WM x WN waves, each wave a
// (BT/WM) x (BT/WN) block of 32 x 32 MFMA tiles (IM x JN of them).
template <int BT_, int WM_, int WN_> struct Tile {
  static constexpr int BT = BT_, WM = WM_, WN = WN_, NW = WM_ * WN_, NT = NW * 64;
  static constexpr int IM = BT_ / WM_ / 32, JN = BT_ / WN_ / 32;
  // one operand tile per stage: 8 k-groups of 16 bytes x BT rows (both planes for F32X3)
  static constexpr int TU4 = 8 * BT_;
};
using TileS = Tile<128, 2, 2>;  // 256 threads, 64 x 64 per wave, 64 KB LDS
using TileL = Tile<256, 2, 4>;  // 512 threads, 128 x 64 per wave, 128 KB LDS

template <int KIND> struct KT;
// BK: k per pipeline stage. KG: k per 16-byte LDS group (per plane for F32X3).
// KC: k consumed per MFMA step (both lane halves). RB/CW: rows x columns of B one
// thread stages per stage.
template <> struct KT<K_F32> {
  static constexpr int BK = 32, KG = 4, ESZ = 4, KC = 16, RB = 4, CW = 4;
  using Acc = f32x16;
};
template <> struct KT<K_BF16> {
  static constexpr int BK = 64, KG = 8, ESZ = 2, KC = 16, RB = 8, CW = 4;
  using Acc = f32x16;
};
template <> struct KT<K_I8> {
  static constexpr int BK = 128, KG = 16, ESZ = 1, KC = 32, RB = 16, CW = 4;
  using Acc = i32x16;
};
template <> struct KT<K_F32X3> {
  static constexpr int BK = 32, KG = 8, ESZ = 4, KC = 16, RB = 8, CW = 2;
  using Acc = f32x16;
};

// Operand tile per stage: BK/KG groups x BT rows x 16 B, or for F32X3 two planes
// (hi, lo) of BK/8 groups x BT rows x 16 B: 8 x BT 16-byte slots either way.
// Row r of group g sits at g*BT + (f(r) ^ p(g)) with
// f(r) = r ^ 4(r & 1) ^ ((r >> 3) & 1) ^ 2((r >> 4) & 1), a permutation within every
// aligned 32 rows, and p(g) = 2(g & 3) + ((g >> 2) & 1), a permutation of the low three
// bits by group. Measured bank-conflict free (SQ_LDS_BANK_CONFLICT, a probe of each
// pattern) for every access here: 32 consecutive rows of two groups (32x32 fragment
// reads), 16 consecutive rows of four groups (FP32's 16x16 fragment reads; with
// p(g) = g & 7 these took 4 conflict cycles per read), one row in 8 groups (A staging),
// rows r and r+1 in 4 groups (split-bf16 A staging, 8-byte halves), every second and
// every fourth column of one group (B staging with 2 or 4 columns per thread).
template <int BT>
__device__ __forceinline__ int slot(int g, int r) {
  return g * BT + ((r ^ ((r & 1) << 2) ^ ((r >> 3) & 1) ^ (((r >> 4) & 1) << 1)) ^ (((g & 3) << 1) | ((g >> 2) & 1)));
}

// ---------------------------------------------------------------- staging --
// Word 3 of a raw buffer resource on gfx9 (32-bit data format, no swizzle, no index
// stride): what the interior tiles' buffer loads use.
constexpr int BUF_DW3 = 0x00020000;

template <int KIND, class TL>
struct Stager {
  using T = KT<KIND>;
  static constexpr int NT = TL::NT, BT = TL::BT, TILE_U4 = TL::TU4;
  uint32_t a[16];  // 4 x 16 B of A
  uint32_t b[16];  // RB rows x CW columns of B

  // Operands are staged in one of two ways, by their memory layout:
  //   rows: an operand whose k runs along memory rows (A as stored M x K, or a
  //     transposed B stored N x K): 16-byte k-units per row, load_rows / store_rows;
  //   cols: an operand whose k runs down the columns (B as stored K x N, or a
  //     transposed A stored K x M): RB rows x CW columns per thread, transposed into
  //     k-groups on the way to LDS, load_cols / store_cols.
  // Transposed operands are supported for the FP32 kinds (no INT8 u8 bias or column
  // sums on those paths).
  //
  // Rows [m0, m0+BT) x k [k0, k0+BK) of a row-major R x K operand: 8 BT 16-byte units,
  // 4 per thread (INT8: the u8 -> s8 bias of A is applied here).
  __device__ __forceinline__ void load_rows(const char *A, long M, long K, long kend, long m0, long k0,
                                            bool vec, int t, uint32_t (&a)[16]) {
    if (vec && m0 + BT <= M && k0 + T::BK <= kend && (size_t)BT * K * T::ESZ < (1u << 31)) {
      // interior tile (uniform branch), as buffer loads: the tile's corner is the
      // resource base, the thread's row and 16-byte unit one vector offset and the
      // q-th row step a scalar one, so no per-load 64-bit address arithmetic (unit
      // u = t + NT q lies in row (t >> 3) + q NT / 8, same column unit t & 7)
      const char *base = A + ((size_t)m0 * K + k0) * T::ESZ;
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, 0x7fffffff, BUF_DW3);
      const unsigned rowb = (unsigned)K * T::ESZ;
      const unsigned vo = (unsigned)(t >> 3) * rowb + (unsigned)(t & 7) * 16u;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint4 val = __builtin_bit_cast(
            uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)vo, (int)((unsigned)(q * (NT / 8)) * rowb), 0));
        const uint32_t x = KIND == K_I8 ? 0x80808080u : 0u;  // u8 -> s8 bias
        a[4 * q + 0] = val.x ^ x; a[4 * q + 1] = val.y ^ x; a[4 * q + 2] = val.z ^ x; a[4 * q + 3] = val.w ^ x;
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int u = t + NT * q;
      int row = u >> 3, v = u & 7;
      long m = m0 + row;
      long k = k0 + (long)v * (16 / T::ESZ);
      uint4 val = {0u, 0u, 0u, 0u};
      if (m < M) {
        const char *p = A + ((size_t)m * K + k) * T::ESZ;
        if (vec && k + 16 / T::ESZ <= kend) {
          val = *(const uint4 *)p;
        } else {
          uint32_t w[4] = {0u, 0u, 0u, 0u};
          for (int e = 0; e < 16 / T::ESZ; ++e)
            if (k + e < kend) {
              uint32_t x;
              if (T::ESZ == 4) x = *(const uint32_t *)(p + 4 * e);
              else if (T::ESZ == 2) x = *(const uint16_t *)(p + 2 * e);
              else x = *(const uint8_t *)(p + e);
              w[(e * T::ESZ) >> 2] |= x << (((e * T::ESZ) & 3) * 8);
            }
          val = {w[0], w[1], w[2], w[3]};
        }
      }
      if (KIND == K_I8) {
        // u8 -> s8 bias (x ^ 0x80 = x - 128 as signed). Zero padding must stay
        // zero, so only flip bytes that lie inside the matrix.
        uint32_t mask[4] = {0u, 0u, 0u, 0u};
        if (m < M)
          for (int e = 0; e < 16; ++e)
            if (k + e < kend) mask[e >> 2] |= 0x80u << ((e & 3) * 8);
        val.x ^= mask[0]; val.y ^= mask[1]; val.z ^= mask[2]; val.w ^= mask[3];
      }
      a[4 * q + 0] = val.x; a[4 * q + 1] = val.y; a[4 * q + 2] = val.z; a[4 * q + 3] = val.w;
    }
  }

  __device__ __forceinline__ void store_rows(uint4 *sA, int t, const uint32_t (&a)[16]) const {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int u = t + NT * q;
      int row = u >> 3, v = u & 7;
      if (KIND == K_F32X3) {
        // 4 floats = k 4v..4v+3 -> group v/2, half v%2 of the hi and lo planes.
        uint32_t h01, l01, h23, l23;
        split_pk(__builtin_bit_cast(float, a[4 * q]), __builtin_bit_cast(float, a[4 * q + 1]), h01, l01);
        split_pk(__builtin_bit_cast(float, a[4 * q + 2]), __builtin_bit_cast(float, a[4 * q + 3]), h23, l23);
        int g = v >> 1, half = v & 1;
        uint2 *hp = (uint2 *)(sA + slot<BT>(g, row)) + half;
        uint2 *lp = (uint2 *)(sA + TILE_U4 / 2 + slot<BT>(g, row)) + half;
        *hp = make_uint2(h01, h23);
        *lp = make_uint2(l01, l23);
      } else {
        sA[slot<BT>(v, row)] = make_uint4(a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]);
      }
    }
  }

  // k [k0, k0+BK) x columns [n0, n0+BT) of a row-major K x N operand: thread = RB
  // rows x CW columns.
  __device__ __forceinline__ void load_cols(const char *B, long N, long kend, long n0, long k0, bool vec,
                                            int t, uint32_t (&b)[16]) {
    constexpr int NCB = BT / T::CW;  // column blocks
    int cb = t % NCB, kg = t / NCB;
    long n = n0 + (long)cb * T::CW;
    long kb = k0 + (long)kg * T::RB;
    constexpr int RBYTES = T::CW * T::ESZ;  // bytes per row piece: 16, 8, 4 or 8
    if (vec && n0 + BT <= N && k0 + T::BK <= kend && (size_t)T::BK * N * T::ESZ < (1u << 31)) {
      // interior tile (uniform branch), as buffer loads: the tile's corner is the
      // resource base, the thread's first row and column block one vector offset, the
      // r-th row a scalar one
      const char *base = B + ((size_t)k0 * N + n0) * T::ESZ;
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, 0x7fffffff, BUF_DW3);
      const unsigned rowb = (unsigned)N * T::ESZ;
      const unsigned vo = (unsigned)(kg * T::RB) * rowb + (unsigned)(cb * T::CW * T::ESZ);
#pragma unroll
      for (int r = 0; r < T::RB; ++r) {
        const int so = (int)((unsigned)r * rowb);
        if (RBYTES == 16) {
          const uint4 v = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)vo, so, 0));
          b[4 * r + 0] = v.x; b[4 * r + 1] = v.y; b[4 * r + 2] = v.z; b[4 * r + 3] = v.w;
        } else if (RBYTES == 8) {
          const uint2 v = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)vo, so, 0));
          b[2 * r + 0] = v.x; b[2 * r + 1] = v.y;
        } else {
          b[r] = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)vo, so, 0);
        }
      }
      return;
    }
#pragma unroll
    for (int r = 0; r < T::RB; ++r) {
      long k = kb + r;
      const char *p = B + ((size_t)k * N + n) * T::ESZ;
      bool rowok = k < kend;
      if (RBYTES == 16) {
        uint4 v = {0u, 0u, 0u, 0u};
        if (rowok && vec && n + T::CW <= N) v = *(const uint4 *)p;
        else if (rowok)
          for (int c = 0; c < T::CW; ++c)
            if (n + c < N) (&v.x)[c] = *(const uint32_t *)(p + 4 * c);
        b[4 * r + 0] = v.x; b[4 * r + 1] = v.y; b[4 * r + 2] = v.z; b[4 * r + 3] = v.w;
      } else if (RBYTES == 8) {
        uint2 v = {0u, 0u};
        if (rowok && vec && n + T::CW <= N) v = *(const uint2 *)p;
        else if (rowok) {
          uint32_t w[2] = {0u, 0u};
          for (int c = 0; c < T::CW; ++c)
            if (n + c < N) {
              uint32_t x = (T::ESZ == 4) ? *(const uint32_t *)(p + 4 * c) : *(const uint16_t *)(p + 2 * c);
              w[(c * T::ESZ) >> 2] |= x << (((c * T::ESZ) & 3) * 8);
            }
          v = make_uint2(w[0], w[1]);
        }
        b[2 * r + 0] = v.x; b[2 * r + 1] = v.y;
      } else {
        uint32_t v = 0u;
        if (rowok && vec && n + T::CW <= N) v = *(const uint32_t *)p;
        else if (rowok)
          for (int c = 0; c < T::CW; ++c)
            if (n + c < N) v |= (uint32_t)(*(const uint8_t *)(p + c)) << (8 * c);
        b[r] = v;
      }
    }
  }

  // Rows of B become k-groups of a column: a 4 x 4 byte transpose per 4 rows (8
  // v_perm_b32) for INT8, one v_perm_b32 per pair of rows for BF16.
  // INT8 also adds the s8 sums of the 4 columns it stages into cs (the column sums of
  // the workgroup's B panel, for the u8 bias).
  __device__ __forceinline__ void store_cols(uint4 *sB, int t, int (&cs)[4], const uint32_t (&b)[16]) const {
    constexpr int NCB = BT / T::CW;
    int cb = t % NCB, g = t / NCB;
    uint32_t wi[4][4];  // INT8: wi[c][d] = rows 4d..4d+3 of column c
    if (KIND == K_I8) {
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const uint32_t t0 = __builtin_amdgcn_perm(b[4 * d + 1], b[4 * d], 0x05010400u);
        const uint32_t t1 = __builtin_amdgcn_perm(b[4 * d + 3], b[4 * d + 2], 0x05010400u);
        const uint32_t t2 = __builtin_amdgcn_perm(b[4 * d + 1], b[4 * d], 0x07030602u);
        const uint32_t t3 = __builtin_amdgcn_perm(b[4 * d + 3], b[4 * d + 2], 0x07030602u);
        wi[0][d] = __builtin_amdgcn_perm(t1, t0, 0x05040100u);
        wi[1][d] = __builtin_amdgcn_perm(t1, t0, 0x07060302u);
        wi[2][d] = __builtin_amdgcn_perm(t3, t2, 0x05040100u);
        wi[3][d] = __builtin_amdgcn_perm(t3, t2, 0x07060302u);
      }
    }
#pragma unroll
    for (int c = 0; c < T::CW; ++c) {
      int n = cb * T::CW + c;
      if (KIND == K_F32) {
        sB[slot<BT>(g, n)] = make_uint4(b[c], b[4 + c], b[8 + c], b[12 + c]);
      } else if (KIND == K_BF16) {
        const uint32_t sel = (c & 1) ? 0x07060302u : 0x05040100u;  // high or low halves of both words
        uint32_t w[4];
#pragma unroll
        for (int d = 0; d < 4; ++d)
          w[d] = __builtin_amdgcn_perm(b[2 * (2 * d + 1) + (c >> 1)], b[2 * (2 * d) + (c >> 1)], sel);
        sB[slot<BT>(g, n)] = make_uint4(w[0], w[1], w[2], w[3]);
      } else if (KIND == K_I8) {
        sB[slot<BT>(g, n)] = make_uint4(wi[c][0], wi[c][1], wi[c][2], wi[c][3]);
#pragma unroll
        for (int d = 0; d < 4; ++d) cs[c] = __builtin_amdgcn_sdot4((int)wi[c][d], 0x01010101, cs[c], false);
      } else {  // F32X3: 8 floats of column c -> hi and lo planes
        uint32_t h[4], l[4];
        for (int r = 0; r < 4; ++r)
          split_pk(__builtin_bit_cast(float, b[4 * r + c]), __builtin_bit_cast(float, b[4 * r + 2 + c]), h[r], l[r]);
        sB[slot<BT>(g, n)] = make_uint4(h[0], h[1], h[2], h[3]);
        sB[TILE_U4 / 2 + slot<BT>(g, n)] = make_uint4(l[0], l[1], l[2], l[3]);
      }
    }
  }
};

// ------------------------------------------------------------------- MFMA --
template <int KIND, class TL>
__device__ __forceinline__ void mma_step(const uint4 *sA, const uint4 *sB, int kc, int wm, int wn, int lane,
                                         typename KT<KIND>::Acc (&acc)[TL::IM][TL::JN]) {
  using T = KT<KIND>;
  constexpr int BT = TL::BT, IM = TL::IM, JN = TL::JN, HP = TL::TU4 / 2;
  const int h = lane >> 5, l32 = lane & 31;
  const int r0 = wm * (BT / TL::WM) + l32, c0 = wn * (BT / TL::WN) + l32;
  if (KIND == K_F32X3) {
    const int g = kc / 8 + h;
    uint4 ah[IM], al[IM], bh[JN], bl[JN];
#pragma unroll
    for (int i = 0; i < IM; ++i) {
      ah[i] = sA[slot<BT>(g, r0 + i * 32)];
      al[i] = sA[HP + slot<BT>(g, r0 + i * 32)];
    }
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      bh[j] = sB[slot<BT>(g, c0 + j * 32)];
      bl[j] = sB[HP + slot<BT>(g, c0 + j * 32)];
    }
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j) {
        // small terms first, then the dominant hi.hi product
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, al[i]),
                                                            __builtin_bit_cast(bf16x8, bh[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ah[i]),
                                                            __builtin_bit_cast(bf16x8, bl[j]), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ah[i]),
                                                            __builtin_bit_cast(bf16x8, bh[j]), acc[i][j], 0, 0, 0);
      }
    return;
  }
  if constexpr (KIND == K_F32) {
    // 16x16x4 MFMAs (the f32 shape with the higher issue rate on gfx950: 156 against 148
    // TFLOP/s from registers): lane quad q holds k = kc + 4q + s (s = 0..3, one 16-byte
    // group) of row (column) lane % 16 of each 16-row (16-column) sub-block; sub-block
    // (ii, jj) of 32 x 32 block (i, j) accumulates in elements 4 (2 ii + jj) .. + 3 of
    // acc[i][j]. Every accumulator gets its MFMAs in k order, 4 JN others between them.
    const int q = lane >> 4, l16 = lane & 15;
    const int g4 = kc / T::KG + q;
    const int rr = wm * (BT / TL::WM) + l16, cc = wn * (BT / TL::WN) + l16;
    // (row blocks one at a time, so the 256 x 256 tile's fragments fit its registers)
    f32x4 bv[2 * JN];
#pragma unroll
    for (int j = 0; j < 2 * JN; ++j) bv[j] = __builtin_bit_cast(f32x4, sB[slot<BT>(g4, cc + j * 16)]);
#pragma unroll
    for (int i = 0; i < IM; ++i) {
      const f32x4 av[2] = {__builtin_bit_cast(f32x4, sA[slot<BT>(g4, rr + i * 32)]),
                           __builtin_bit_cast(f32x4, sA[slot<BT>(g4, rr + i * 32 + 16)])};
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int j = 0; j < JN; ++j)
#pragma unroll
          for (int sb = 0; sb < 4; ++sb) {
            f32x4 c = {acc[i][j][4 * sb], acc[i][j][4 * sb + 1], acc[i][j][4 * sb + 2], acc[i][j][4 * sb + 3]};
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(av[sb >> 1][s], bv[2 * j + (sb & 1)][s], c, 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[i][j][4 * sb + t] = c[t];
          }
    }
    return;
  }
  const int g = kc / T::KG + h;
  uint4 a[IM], b[JN];
#pragma unroll
  for (int i = 0; i < IM; ++i) a[i] = sA[slot<BT>(g, r0 + i * 32)];
#pragma unroll
  for (int j = 0; j < JN; ++j) b[j] = sB[slot<BT>(g, c0 + j * 32)];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      if (KIND == K_BF16) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a[i]),
                                                            __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0, 0, 0);
      } else {
        acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(i32x4, a[i]),
                                                          __builtin_bit_cast(i32x4, b[j]), acc[i][j], 0, 0, 0);
      }
    }
}

// Map a linear workgroup id to (tile_m, tile_n). Consecutive ids are dispatched
// round-robin over the 8 XCDs; regroup them so that each XCD receives a contiguous
// run of logical tiles, and walk the tiles in bands of 4 tile rows (all columns of a
// band, 4 rows at a time), so the tiles an XCD runs together share few A row panels
// and few B column panels in its L2 (2048^3: 4 + 8 panels instead of 16 + 2).
__device__ __forceinline__ void tile_of(int bid, int tiles_m, int tiles_n, int &tm, int &tn) {
  constexpr int BAND = 4;
  const int total = tiles_m * tiles_n;
  int L = bid;
  if ((total & 7) == 0) L = (bid & 7) * (total >> 3) + (bid >> 3);
  const int band = BAND * tiles_n, st = L / band, l = L - st * band;
  const int rows = tiles_m - st * BAND < BAND ? tiles_m - st * BAND : BAND;  // the last band may be short
  tm = st * BAND + l % rows;
  tn = l / rows;
}

// One BT x BT output tile at (m0, n0) over k in [kbeg, kend), by one workgroup of
// KG x TL::NT threads; smem holds KG x 4 x TL::TU4 slots ([group][stage][A|B]).
// EPI_PARTIAL writes slab `slab` of the split-K partials. TRANS bit 0: A is stored
// transposed (K x M, op(A) = A^T); bit 1: B is stored transposed (N x K). FP32 kinds
// only.
// KG = 2 (FP32 kinds, C stored): two groups of TL::NW waves share the tile, group g
// taking the g-th half of its k stages, so a CU holding one tile still has two waves
// per SIMD; group 1 hands its accumulators to group 0 through LDS, which adds them
// (first half + second half, a fixed order) and stores C.
template <int KIND, int EPI, class TL, int TRANS, int KG = 1>
__device__ __forceinline__ void gemm_tile(const void *__restrict__ Av, const void *__restrict__ Bv,
                                          void *__restrict__ Cv, const int *__restrict__ colsum, long M, long N,
                                          long K, long kbeg, long kend, long m0, long n0, int slab, int vecA,
                                          int vecB, uint4 *smem_all) {
  using T = KT<KIND>;
  using Acc = typename T::Acc;
  constexpr int BT = TL::BT, IM = TL::IM, JN = TL::JN, TILE_U4 = TL::TU4;
  static_assert(KG == 1 || ((KIND == K_F32 || KIND == K_F32X3) && EPI == EPI_STORE), "k groups: FP32 kinds, C");
  // the group index is wave-uniform (NT is a multiple of 64); readfirstlane tells the
  // compiler so, which keeps the staging's buffer resources in scalar registers (derived
  // from a per-thread value they were rebuilt per lane in a loop around every load)
  const int grp = KG == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)(threadIdx.x / TL::NT));
  const int t = KG == 1 ? (int)threadIdx.x : (int)(threadIdx.x - grp * TL::NT), lane = t & 63, wave = t >> 6;
  const int wm = wave / TL::WN, wn = wave % TL::WN;
  uint4 *const smem = smem_all + grp * 4 * TILE_U4;

  Acc acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0;

  const char *A = (const char *)Av;
  const char *B = (const char *)Bv;
  static_assert(TRANS == 0 || KIND == K_F32 || KIND == K_F32X3, "transposed operands: FP32 kinds only");
  constexpr bool TA = TRANS & 1, TB = (TRANS >> 1) & 1;
  Stager<KIND, TL> st;
  int cs[4] = {0, 0, 0, 0}, cs_unused[4];
  auto load = [&](long k0) {
    if (TA) st.load_cols(A, M, kend, m0, k0, vecA, t, st.a);
    else st.load_rows(A, M, K, kend, m0, k0, vecA, t, st.a);
    if (TB) st.load_rows(B, N, K, kend, n0, k0, vecB, t, st.b);
    else st.load_cols(B, N, kend, n0, k0, vecB, t, st.b);
  };
  auto store = [&](uint4 *dst) {
    if (TA) st.store_cols(dst, t, cs_unused, st.a);
    else st.store_rows(dst, t, st.a);
    if (TB) st.store_rows(dst + TILE_U4, t, st.b);
    else st.store_cols(dst + TILE_U4, t, cs, st.b);
  };
  const int total = (int)((kend - kbeg + T::BK - 1) / T::BK);
  // stages of this group: [first, first + nstages); every group runs `iters` barriers
  const int iters = KG == 1 ? total : (total + 1) / 2;
  const int first = grp * iters;
  const int nstages = KG == 1 ? total : (grp == 0 ? iters : total - iters);
  const long gk = kbeg + (long)first * T::BK;
  if (nstages > 0) {
    load(gk);
    store(smem);
  }
  __syncthreads();
  for (int s = 0; s < iters; ++s) {
    uint4 *cur = smem + (s & 1) * 2 * TILE_U4;
    uint4 *nxt = smem + ((s + 1) & 1) * 2 * TILE_U4;
    const bool more = s + 1 < nstages;
    if (more) load(gk + (long)(s + 1) * T::BK);
    if (KG == 1 || s < nstages) {
#pragma unroll
      for (int kc = 0; kc < T::BK; kc += T::KC) mma_step<KIND, TL>(cur, cur + TILE_U4, kc, wm, wn, lane, acc);
    }
    if (more) store(nxt);
    __syncthreads();
  }
  if (KG == 2) {
    // group 1's accumulators -> LDS (lane-contiguous), group 0 adds them
    float *xs = (float *)smem_all;
    constexpr int PER_WAVE = IM * JN * 16 * 64;
    if (grp == 1) {
#pragma unroll
      for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) xs[wave * PER_WAVE + ((i * JN + j) * 16 + r) * 64 + lane] = (float)acc[i][j][r];
    }
    __syncthreads();
    if (grp == 1) return;
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][j][r] += xs[wave * PER_WAVE + ((i * JN + j) * 16 + r) * 64 + lane];
  }
  // INT8 without a column-sum array: the workgroup staged every row of its B panel
  // (of its k-slice, for a split-K partial), so it sums the panel's columns itself
  // (threads of one column block added in LDS, integers: exact in any order). A partial
  // slab then carries its slice's share of the bias, and the slabs add up to the
  // unsplit result exactly.
  const int *csum_tile = nullptr;
  if (KIND == K_I8 && colsum == nullptr) {
    constexpr int NCB = BT / KT<K_I8>::CW, NG = TL::NT / NCB;
    int *red = (int *)smem;  // NG x BT partial sums, then BT totals
    const int cb = t % NCB, g = t / NCB;
#pragma unroll
    for (int c = 0; c < 4; ++c) red[g * BT + cb * 4 + c] = cs[c];
    __syncthreads();
    for (int n = t; n < BT; n += TL::NT) {
      int sum = 0;
#pragma unroll
      for (int q = 0; q < NG; ++q) sum += red[q * BT + n];
      red[NG * BT + n] = sum;
    }
    __syncthreads();
    csum_tile = red + NG * BT;
  }

  // ---- epilogue: lane holds rows 8(r/4) + 4h + (r%4) of column lane%32 per block
  const int h = lane >> 5, l32 = lane & 31;
  // bytes per element of C: partial slabs and the INT8 STORE are 32-bit, REQUANT
  // writes the next operand (int8, bf16, fp32)
  constexpr int CE = (EPI == EPI_REQUANT && KIND == K_I8) ? 1 : ((EPI == EPI_REQUANT && KIND == K_BF16) ? 2 : 4);
  auto put = [&](char *p, const typename KT<KIND>::Acc &a, int r, int bias) {
    if (EPI == EPI_PARTIAL) {
      if (KIND == K_I8) *(int *)p = (int)((unsigned)a[r] + (unsigned)bias);
      else *(float *)p = (float)a[r];
    } else if (KIND == K_I8) {
      const int v = (int)((unsigned)a[r] + (unsigned)bias);
      if (EPI == EPI_STORE) *(int *)p = v;
      else *(int8_t *)p = sat_i8(v);
    } else if (KIND == K_BF16 && EPI == EPI_REQUANT) {
      *(uint16_t *)p = bf16_dev((float)a[r]);
    } else {
      *(float *)p = (float)a[r];
    }
  };
  char *const Cb = (char *)Cv + (EPI == EPI_PARTIAL ? (size_t)slab * M * N * CE : 0);
  // element r of a 32 x 32 block's accumulator, for this lane: row lrow + row_of(r),
  // column lcol + col_of(r) of the block. 32x32 MFMAs (BF16, INT8): rows 8(r/4) + 4h +
  // (r%4) of column l32; FP32's 16x16 sub-blocks: rows 16 (r/8) + 4q + (r%4) of column
  // 16 ((r/4) % 2) + lane % 16.
  constexpr bool S16 = KIND == K_F32;
  const int lrow = S16 ? 4 * (lane >> 4) : 4 * h, lcol = S16 ? (lane & 15) : l32;
  auto row_of = [](int r) { return S16 ? 16 * (r >> 3) + (r & 3) : 8 * (r >> 2) + (r & 3); };
  auto col_of = [](int r) { return S16 ? 16 * ((r >> 2) & 1) : 0; };
  if (m0 + BT <= M && n0 + BT <= N) {
    // whole tile inside C: one lane pointer per block, rows at uniform strides
    const size_t rs = (size_t)N * CE;
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      const long n = n0 + wn * (BT / TL::WN) + j * 32 + lcol;
      int bias = 0;
      if (KIND == K_I8 && (EPI != EPI_PARTIAL || !colsum)) bias = 128 * (colsum ? colsum[n] : csum_tile[n - n0]);
#pragma unroll
      for (int i = 0; i < IM; ++i) {
        char *const p = Cb + ((size_t)(m0 + wm * (BT / TL::WM) + i * 32 + lrow) * N + n) * CE;
#pragma unroll
        for (int r = 0; r < 16; ++r)
          put(p + (size_t)row_of(r) * rs + (size_t)col_of(r) * CE, acc[i][j], r, bias);
      }
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < JN; ++j) {
    const long nb = n0 + wn * (BT / TL::WN) + j * 32 + lcol;
    int bias = 0;
    if (KIND == K_I8 && (EPI != EPI_PARTIAL || !colsum) && nb < N)
      bias = 128 * (colsum ? colsum[nb] : csum_tile[nb - n0]);
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long m = m0 + wm * (BT / TL::WM) + i * 32 + lrow + row_of(r), n = nb + col_of(r);
        if (m >= M || n >= N) continue;
        put(Cb + ((size_t)m * N + n) * CE, acc[i][j], r, bias);
      }
  }
}

template <int KIND, int EPI, class TL, int TRANS, int KG = 1>
__global__ __launch_bounds__(TL::NT * KG) void gemm_kernel(const void *__restrict__ Av, const void *__restrict__ Bv,
                                                           void *__restrict__ Cv, const int *__restrict__ colsum,
                                                           long M, long N, long K, long kper, int tiles_m,
                                                           int tiles_n, int vecA, int vecB) {
  extern __shared__ uint4 smem[];
  int tm, tn;
  tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const long kbeg = (long)blockIdx.y * kper;
  const long kend = (kbeg + kper < K) ? kbeg + kper : K;
  gemm_tile<KIND, EPI, TL, TRANS, KG>(Av, Bv, Cv, colsum, M, N, K, kbeg, kend, (long)tm * TL::BT,
                                      (long)tn * TL::BT, (int)blockIdx.y, vecA, vecB, smem);
}

// Grouped launch over a LevelTable: workgroup b runs item b, one 128 x 128 tile of one
// GEMM (or one k-slice of it, for a split descriptor) with the 256-thread tile of
// gemm_kernel, with the GEMM's kind and epilogue. One workgroup per item: a persistent
// loop around the tile costs registers (INT8 to one wave per SIMD).
__global__ __launch_bounds__(TileS::NT) void grouped_gemm_kernel(const char *__restrict__ table) {
  extern __shared__ uint4 smem[];
  const LevelTable *hdr = (const LevelTable *)table;
  if ((int)blockIdx.x >= hdr->num_items) return;
  const GemmDesc *descs = (const GemmDesc *)(table + hdr->descs_off);
  const int2 w = ((const int2 *)(table + hdr->items_off))[blockIdx.x];
  const GemmDesc &d = descs[w.x];
  const int tile = w.y / d.splits, slice = w.y - tile * d.splits;
  const int tm = tile / d.tiles_n, tn = tile - tm * d.tiles_n;
  const long m0 = (long)tm * TileS::BT, n0 = (long)tn * TileS::BT;
  const long kb = (long)slice * d.kper, ke = kb + d.kper < d.K ? kb + d.kper : d.K;
  const int va = d.vec & 1, vb = (d.vec >> 1) & 1;
#define FG_LEVEL_TILE(KIND, EPI) \
  gemm_tile<KIND, EPI, TileS, 0>(d.A, d.B, d.C, nullptr, d.M, d.N, d.K, kb, ke, m0, n0, slice, va, vb, smem)
#define FG_LEVEL_KIND(KIND)                                            \
  do {                                                                 \
    if (d.epi == EPI_STORE) FG_LEVEL_TILE(KIND, EPI_STORE);            \
    else if (d.epi == EPI_REQUANT) FG_LEVEL_TILE(KIND, EPI_REQUANT);   \
    else FG_LEVEL_TILE(KIND, EPI_PARTIAL);                             \
  } while (0)
  switch (d.kind) {
    case K_F32: FG_LEVEL_KIND(K_F32); break;
    case K_BF16: FG_LEVEL_KIND(K_BF16); break;
    case K_I8: FG_LEVEL_KIND(K_I8); break;
    case K_F32X3: FG_LEVEL_KIND(K_F32X3); break;
  }
#undef FG_LEVEL_KIND
#undef FG_LEVEL_TILE
}

// The split GEMMs of a level: workgroup b sums 256 elements of one descriptor over its
// slabs, slab 0 first (a fixed order: the bits depend on the split, not on the launch),
// and writes C in the descriptor's epilogue. INT8 slabs already hold their slices'
// share of the u8 bias, so the integer sum is the unsplit result.
__global__ __launch_bounds__(256) void grouped_reduce_kernel(const char *__restrict__ table) {
  const ReduceTable *hdr = (const ReduceTable *)table;
  if ((int)blockIdx.x >= hdr->num_items) return;
  const ReduceDesc *descs = (const ReduceDesc *)(table + hdr->descs_off);
  const int2 w = ((const int2 *)(table + hdr->items_off))[blockIdx.x];
  const ReduceDesc &d = descs[w.x];
  const size_t e = (size_t)w.y * 256 + threadIdx.x;
  if (e >= (size_t)d.MN) return;
  const size_t MN = (size_t)d.MN;
  if (d.kind == K_I8) {
    const unsigned *p = (const unsigned *)d.P + e;
    unsigned s = 0;
    for (int z = 0; z < d.splits; ++z) s += p[(size_t)z * MN];
    if (d.epi == EPI_STORE) ((int *)d.C)[e] = (int)s;
    else ((int8_t *)d.C)[e] = sat_i8((int)s);
  } else {
    const float *p = (const float *)d.P + e;
    float s = p[0];
    for (int z = 1; z < d.splits; ++z) s += p[(size_t)z * MN];
    if (d.kind == K_BF16 && d.epi == EPI_REQUANT) ((uint16_t *)d.C)[e] = bf16_dev(s);
    else ((float *)d.C)[e] = s;
  }
}

// INT8 column sums sum_k s8(B[k][n]) for the u8 bias term (exact in int32 for
// K < 2^24). A workgroup covers 256 columns x CS_ROWS rows: thread (ty, tx) sums
// CS_RPT rows of 16 adjacent columns (one 16-byte load per row, 16 threads cover
// 256 contiguous bytes of a row), the 16 row slices are added in LDS and each
// column meets the other row blocks in one integer atomic (exact and order
// independent, so the result is deterministic). out[] must be zero first
// (colsum_zero_kernel, a fill kernel rather than a memset so the pair can sit inside
// a captured graph).
constexpr int CS_RPT = 8;
constexpr int CS_ROWS = 16 * CS_RPT;

__global__ void colsum_zero_kernel(int *__restrict__ out, long N) {
  long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) out[n] = 0;
}

__global__ __launch_bounds__(256) void colsum_i8_kernel(const int8_t *__restrict__ B, long K, long N,
                                                        int *__restrict__ out) {
  __shared__ int part[16][256];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const long n = (long)blockIdx.x * 256 + tx * 16;
  const long k0 = (long)blockIdx.y * CS_ROWS + ty * CS_RPT;
  int s[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) s[c] = 0;
  if (n + 16 <= N && (N & 15) == 0 && ((uintptr_t)B & 15) == 0 && k0 + CS_RPT <= K) {
#pragma unroll
    for (int r = 0; r < CS_RPT; ++r) {
      const uint4 w = *(const uint4 *)(B + (size_t)(k0 + r) * N + n);
      const uint32_t q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int c = 0; c < 16; ++c) s[c] += (int)(int8_t)(q[c >> 2] >> (8 * (c & 3)));
    }
  } else {
    for (long k = k0; k < k0 + CS_RPT && k < K; ++k)
      for (int c = 0; c < 16; ++c)
        if (n + c < N) s[c] += B[(size_t)k * N + n + c];
  }
#pragma unroll
  for (int c = 0; c < 16; ++c) part[ty][tx * 16 + c] = s[c];
  __syncthreads();
  const long col = (long)blockIdx.x * 256 + threadIdx.x;
  if (col >= N) return;
  int t = 0;
  for (int y = 0; y < 16; ++y) t += part[y][threadIdx.x];
  atomicAdd(&out[col], t);
}

// Split-K reduction: a workgroup owns RE consecutive elements; RL lanes per element
// each sum the slabs z = l, l + RL, ... in order, and lane sums are combined in
// lane order through LDS, so the result does not depend on scheduling (nor on RE,
// which only sets how many elements a workgroup takes: 64, or 16 for small outputs
// with many slabs, so those spread over more CUs). A lane's loads are issued four
// slabs at a time and added in slab order. With has_ex it also applies the epilogue
// extensions (epi_i8 / epi_f, row m = e / N).
constexpr int RE = 64;
template <int KIND, int EPI, int RL, int RE_ = RE>
__global__ __launch_bounds__(RE_ * RL) void splitk_reduce_kernel(const void *__restrict__ P, int splits, long M,
                                                                  long N, const int *__restrict__ colsum,
                                                                  void *__restrict__ C, EpiOne ex, int has_ex) {
  using Sum = typename std::conditional<KIND == K_I8, unsigned, float>::type;
  __shared__ Sum part[RL][RE_];
  const int x = threadIdx.x % RE_, l = threadIdx.x / RE_;
  const size_t MN = (size_t)M * N;
  const size_t e = (size_t)blockIdx.x * RE_ + x;
  Sum acc = 0;
  if (e < MN) {
    const Sum *src = (const Sum *)P + e;
    int z = l;
    for (; z + 3 * RL < splits; z += 4 * RL) {
      const Sum v0 = src[(size_t)z * MN], v1 = src[(size_t)(z + RL) * MN];
      const Sum v2 = src[(size_t)(z + 2 * RL) * MN], v3 = src[(size_t)(z + 3 * RL) * MN];
      acc += v0;
      acc += v1;
      acc += v2;
      acc += v3;
    }
    for (; z < splits; z += RL) acc += src[(size_t)z * MN];
  }
  part[l][x] = acc;
  __syncthreads();
  if (l != 0 || e >= MN) return;
  Sum s = part[0][x];
  for (int q = 1; q < RL; ++q) s += part[q][x];
  const long m = (long)(e / N);
  if (KIND == K_I8) {
    s += 128u * (unsigned)colsum[e % N];
    if (has_ex) {
      const int bias = ex.bias ? ((const int *)ex.bias)[m] : 0;
      if (EPI == EPI_STORE) ((int *)C)[e] = epi_i32((int)s, bias, ex.relu);
      else ((int8_t *)C)[e] = (int8_t)epi_i8((int)s, bias, ex.scale != nullptr, ex.scale ? ex.scale[m] : 1.0f, ex.zp,
                                            ex.relu);
    } else if (EPI == EPI_STORE) {
      ((int *)C)[e] = (int)s;
    } else {
      ((int8_t *)C)[e] = sat_i8((int)s);
    }
  } else {
    if (has_ex)
      s = epi_f(s, ex.scale != nullptr, ex.scale ? ex.scale[m] : 1.0f, ex.bias != nullptr,
                ex.bias ? ((const float *)ex.bias)[m] : 0.0f, ex.relu);
    if (KIND == K_BF16 && EPI == EPI_REQUANT) ((uint16_t *)C)[e] = bf16_dev(s);
    else ((float *)C)[e] = s;
  }
}

// Reference-layout re-interpretation of a requantized intermediate Z -> X (gather).
template <typename E>
__global__ void ref_permute_kernel(const E *__restrict__ Z, E *__restrict__ X, long rows, RefPanels p,
                                   bool is_int8) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)rows * p.N) return;
  long k = (long)(e / p.N), n = (long)(e % p.N);
  long p0, w, r, c;
  ref_panel_of(p, n, p0, w);
  if (!ref_src(is_int8, w, k, n - p0, r, c) || r >= rows) { X[e] = (E)0; return; }
  X[e] = Z[(size_t)r * p.N + p0 + c];
}

template <int KIND, int EPI, class TL, int TRANS, int KG = 1>
hipError_t launch_tile(const GemmArgs &g, hipStream_t s) {
  using T = KT<KIND>;
  int tiles_m = (int)((g.M + TL::BT - 1) / TL::BT), tiles_n = (int)((g.N + TL::BT - 1) / TL::BT);
  int splits = g.splits > 0 ? g.splits : 1;
  long kper = ((g.K + splits - 1) / splits + T::BK - 1) / T::BK * T::BK;
  splits = (int)((g.K + kper - 1) / kper);
  int esz = T::ESZ;
  // vector loads: 16-byte k-units along rows (rows layout) or CW-element pieces of a row (cols layout)
  auto vec_rows = [&](const void *p, long ld) { return (ld * esz) % 16 == 0 && ((uintptr_t)p & 15) == 0; };
  auto vec_cols = [&](const void *p, long ld) {
    return ld % T::CW == 0 && ((uintptr_t)p & (T::CW * esz - 1)) == 0;
  };
  bool vecA = (TRANS & 1) ? vec_cols(g.A, g.M) : vec_rows(g.A, g.K);
  bool vecB = (TRANS & 2) ? vec_rows(g.B, g.K) : vec_cols(g.B, g.N);
  const size_t lds = KG * 4 * (size_t)TL::TU4 * sizeof(uint4);
  auto kfn = gemm_kernel<KIND, EPI, TL, TRANS, KG>;
  hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid(tiles_m * tiles_n, splits);
  hipLaunchKernelGGL(kfn, grid, dim3(TL::NT * KG), lds, s, g.A, g.B, g.C, g.colsum, g.M, g.N, g.K, kper, tiles_m,
                     tiles_n, (int)vecA, (int)vecB);
  return hipGetLastError();
}

// The 256 x 256 tile reads half the operand bytes per MFMA of the 128 x 128 tile;
// it is used when it still gives about one tile per CU.
// A single FP32 GEMM (g.kgroups) with at most one 128 x 128 tile per CU and a long K
// runs two k groups per tile (8 waves), for two waves per SIMD.
template <int KIND, int EPI, int TRANS>
hipError_t launch_trans(const GemmArgs &g, hipStream_t s) {
  const long big = ((g.M + 255) / 256) * ((g.N + 255) / 256) * (g.splits > 1 ? g.splits : 1);
  if (big >= 192) return launch_tile<KIND, EPI, TileL, TRANS>(g, s);
  if constexpr ((KIND == K_F32 || KIND == K_F32X3) && EPI == EPI_STORE) {
    const long small = ((g.M + 127) / 128) * ((g.N + 127) / 128);
    if (g.kgroups && g.splits <= 1 && small <= num_cus() && g.K >= 16 * KT<KIND>::BK)
      return launch_tile<KIND, EPI, TileS, TRANS, 2>(g, s);
  }
  return launch_tile<KIND, EPI, TileS, TRANS>(g, s);
}

template <int KIND, int EPI>
hipError_t launch_one(const GemmArgs &g, hipStream_t s) {
  if constexpr (KIND == K_F32 || KIND == K_F32X3) {
    if constexpr (EPI != EPI_REQUANT) {  // transposed operands: single FP32 GEMMs (C or split-K slabs)
      switch (g.trans) {
        case 1: return launch_trans<KIND, EPI, 1>(g, s);
        case 2: return launch_trans<KIND, EPI, 2>(g, s);
        case 3: return launch_trans<KIND, EPI, 3>(g, s);
      }
    }
  }
  if (g.trans != 0) return hipErrorInvalidValue;
  return launch_trans<KIND, EPI, 0>(g, s);
}

template <int KIND>
hipError_t launch_kind(const GemmArgs &g, hipStream_t s) {
  switch (g.epi) {
    case EPI_STORE: return launch_one<KIND, EPI_STORE>(g, s);
    case EPI_REQUANT: return launch_one<KIND, EPI_REQUANT>(g, s);
    case EPI_PARTIAL: return launch_one<KIND, EPI_PARTIAL>(g, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipErrorInvalidValue;
  switch (g.kind) {
    case K_F32: return launch_kind<K_F32>(g, s);
    case K_BF16: return launch_kind<K_BF16>(g, s);
    case K_I8: return launch_kind<K_I8>(g, s);
    case K_F32X3: return launch_kind<K_F32X3>(g, s);
  }
  return hipErrorInvalidValue;
}

template <int KIND>
void gemm_desc_kind(GemmDesc *d, const void *A, const void *B, long N, long K) {
  using T = KT<KIND>;
  const bool va = (K * T::ESZ) % 16 == 0 && ((uintptr_t)A & 15) == 0;
  const bool vb = N % T::CW == 0 && ((uintptr_t)B & (T::CW * T::ESZ - 1)) == 0;
  d->vec = (va ? 1 : 0) | (vb ? 2 : 0);
}

void gemm_desc(GemmDesc *d, int kind, int epi, const void *A, const void *B, void *C, long M, long N, long K) {
  d->A = A;
  d->B = B;
  d->C = C;
  d->M = M;
  d->N = N;
  d->K = K;
  d->kind = kind;
  d->epi = epi;
  d->tiles_n = (int)((N + TileS::BT - 1) / TileS::BT);
  d->kper = K;
  d->splits = 1;
  d->pad = 0;
  switch (kind) {
    case K_F32: gemm_desc_kind<K_F32>(d, A, B, N, K); break;
    case K_BF16: gemm_desc_kind<K_BF16>(d, A, B, N, K); break;
    case K_I8: gemm_desc_kind<K_I8>(d, A, B, N, K); break;
    default: gemm_desc_kind<K_F32X3>(d, A, B, N, K); break;
  }
}

int gemm_desc_split(GemmDesc *d, int splits, void *slabs) {
  const long bk = d->kind == K_I8 ? KT<K_I8>::BK : (d->kind == K_BF16 ? KT<K_BF16>::BK : KT<K_F32>::BK);
  if (splits <= 1) return 1;
  const long kper = ((d->K + splits - 1) / splits + bk - 1) / bk * bk;
  const int used = (int)((d->K + kper - 1) / kper);
  if (used <= 1) return 1;
  d->kper = kper;
  d->splits = used;
  d->epi = EPI_PARTIAL;
  d->C = slabs;
  return used;
}

hipError_t launch_grouped_reduce(const void *table, int grid, hipStream_t s) {
  if (grid <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(grouped_reduce_kernel, dim3(grid), dim3(256), 0, s, (const char *)table);
  return hipGetLastError();
}

hipError_t launch_grouped_gemm(const void *table, int grid, hipStream_t s) {
  if (grid <= 0) return hipErrorInvalidValue;
  const size_t lds = 4 * (size_t)TileS::TU4 * sizeof(uint4);
  hipError_t e = hipFuncSetAttribute((const void *)grouped_gemm_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(grouped_gemm_kernel, dim3(grid), dim3(TileS::NT), lds, s, (const char *)table);
  return hipGetLastError();
}

__device__ __forceinline__ void split3(float x, uint16_t &h, uint16_t &m, uint16_t &l) {
  h = bf16_dev(x);
  const float r1 = x - __builtin_bit_cast(float, (uint32_t)h << 16);
  m = bf16_dev(r1);
  l = bf16_dev(r1 - __builtin_bit_cast(float, (uint32_t)m << 16));
}

// A (M x K) -> A6 (M x 6K), column blocks [hi hi mid hi lo mid]; one element per thread.
__global__ __launch_bounds__(256) void split3_a_kernel(const float *__restrict__ A, long M, long K,
                                                       uint16_t *__restrict__ A6) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= M * K) return;
  const long m = e / K, k = e - m * K;
  uint16_t h, md, l;
  split3(A[e], h, md, l);
  uint16_t *row = A6 + m * 6 * K + k;
  row[0] = h; row[K] = h; row[2 * K] = md; row[3 * K] = h; row[4 * K] = l; row[5 * K] = md;
}

// B (K x N) -> B6 (6K x N), row blocks [hi mid hi lo hi mid].
__global__ __launch_bounds__(256) void split3_b_kernel(const float *__restrict__ B, long K, long N,
                                                       uint16_t *__restrict__ B6) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= K * N) return;
  uint16_t h, md, l;
  split3(B[e], h, md, l);
  const long blk = K * N;
  B6[e] = h; B6[blk + e] = md; B6[2 * blk + e] = h; B6[3 * blk + e] = l; B6[4 * blk + e] = h; B6[5 * blk + e] = md;
}

hipError_t launch_split3(const float *A, long M, long K, uint16_t *A6, const float *B, long N, uint16_t *B6,
                         hipStream_t s) {
  hipLaunchKernelGGL(split3_a_kernel, dim3((unsigned)((M * K + 255) / 256)), dim3(256), 0, s, A, M, K, A6);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(split3_b_kernel, dim3((unsigned)((K * N + 255) / 256)), dim3(256), 0, s, B, K, N, B6);
  return hipGetLastError();
}

hipError_t launch_colsum_i8(const int8_t *B, long K, long N, int *out, hipStream_t s) {
  hipLaunchKernelGGL(colsum_zero_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, out, N);
  dim3 grid((unsigned)((N + 255) / 256), (unsigned)((K + CS_ROWS - 1) / CS_ROWS));
  hipLaunchKernelGGL(colsum_i8_kernel, grid, dim3(256), 0, s, B, K, N, out);
  return hipGetLastError();
}

hipError_t launch_splitk_reduce(int kind, int epi, const void *P, int splits, long M, long N,
                                const int *colsum, void *C, const EpiOne *ex, hipStream_t s) {
  const EpiOne exv = ex ? *ex : EpiOne{};
  const int has_ex = ex != nullptr;
  size_t MN = (size_t)M * N;
  dim3 grid((unsigned)((MN + RE - 1) / RE));
  // lanes per element: enough to keep every slab read in flight, at most 16
  const int rl = splits >= 16 ? 16 : splits >= 8 ? 8 : splits >= 4 ? 4 : splits >= 2 ? 2 : 1;
  // many slabs of a small output: 16 elements per workgroup, so the reads spread over
  // about one workgroup per CU instead of a quarter of them
  const bool narrow = rl == 16 && MN <= (size_t)16 * 1024 * 4;
  dim3 grid16((unsigned)((MN + 15) / 16));
#define FG_RED1(K_, E_, L_) \
  hipLaunchKernelGGL((splitk_reduce_kernel<K_, E_, L_>), grid, dim3(RE * L_), 0, s, P, splits, M, N, colsum, C, \
                     exv, has_ex)
#define FG_RED16(K_, E_)                                                                                          \
  hipLaunchKernelGGL((splitk_reduce_kernel<K_, E_, 16, 16>), grid16, dim3(256), 0, s, P, splits, M, N, colsum, C, \
                     exv, has_ex)
#define FG_RED(K_, E_)                          \
  do {                                          \
    if (narrow) {                               \
      FG_RED16(K_, E_);                         \
      break;                                    \
    }                                           \
    switch (rl) {                               \
      case 16: FG_RED1(K_, E_, 16); break;      \
      case 8: FG_RED1(K_, E_, 8); break;        \
      case 4: FG_RED1(K_, E_, 4); break;        \
      case 2: FG_RED1(K_, E_, 2); break;        \
      default: FG_RED1(K_, E_, 1); break;       \
    }                                           \
  } while (0)
  if (kind == K_I8) {
    if (epi == EPI_STORE) FG_RED(K_I8, EPI_STORE); else FG_RED(K_I8, EPI_REQUANT);
  } else if (kind == K_BF16) {
    if (epi == EPI_STORE) FG_RED(K_BF16, EPI_STORE); else FG_RED(K_BF16, EPI_REQUANT);
  } else {
    FG_RED(K_F32, EPI_STORE);
  }
#undef FG_RED
#undef FG_RED1
#undef FG_RED16
  return hipGetLastError();
}

hipError_t launch_ref_permute(bool is_int8, const void *Z, void *X, long rows, const RefPanels &p,
                              hipStream_t s) {
  size_t n = (size_t)rows * p.N;
  dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (is_int8)
    hipLaunchKernelGGL((ref_permute_kernel<int8_t>), grid, block, 0, s, (const int8_t *)Z, (int8_t *)X, rows, p, true);
  else
    hipLaunchKernelGGL((ref_permute_kernel<uint16_t>), grid, block, 0, s, (const uint16_t *)Z, (uint16_t *)X, rows, p,
                       false);
  return hipGetLastError();
}

// ------------------------------------------------------------ skinny FP32 --
// FP32 GEMMs with a small M or N (at most one 64-wide tile across it) are mostly
// k-traffic: a 128 x 128 tile would waste half or three quarters of its MFMAs, so
// they get a 64 x 64 tile per workgroup, split-K over the grid, no LDS. Each of the
// 4 waves owns one 32 x 32 quadrant of the tile over the workgroup's whole k-slice,
// so no wave waits for another and none exchanges data: it writes its quadrant of the
// partial slab (or of C) itself. A wave streams its operands straight into MFMA
// fragment registers, SK_PD groups of 32 k ahead of its MFMAs: A[m][k..k+3] as 16
// bytes per lane, B[k + s][n] as coalesced dwords. Within a group the k order is
// fixed and the even and odd MFMAs of a k-step go to two accumulators added at the
// end, so results do not depend on timing. That is the kernel for short slices
// (split-K-heavy shapes); for long slices skinny_tile_kernel below puts every wave on
// the whole tile instead (twice the operand reuse), and its four wave tiles are summed
// through LDS, an exchange that cost ~1.7 us of the 64 x 64 x 65536 GEMM.
namespace {

constexpr int SK_TILE = 64;
constexpr int SK_GROUP = 32;  // k per group (4 MFMA k-steps of 8)
constexpr int SK_QGROUPS = 8; // groups per slice of the quadrant kernel (256 k)

struct SkFrag {
  f32x4 a[4];     // [k-step t]: A[row][k + 8t + 4h .. + 3]
  float b[4][4];  // [k-step t][s]: B[k + 8t + 4h + s][col]
};

// One group of a quadrant for lane (h, l32): A row `row`, B column `col`, k [k, k + 32)
// clipped to ke. fast: the quadrant lies inside A and B, the whole group inside the
// slice and the A rows are 16-byte aligned (K % 4 == 0).
__device__ __forceinline__ void sk_load(const float *__restrict__ A, const float *__restrict__ B, long M, long N,
                                        long K, long ke, long row, long col, long k, bool fast, int h, SkFrag &f) {
  if (fast) {
#pragma unroll
    for (int t = 0; t < 4; ++t) f.a[t] = *(const f32x4 *)(A + row * K + k + 8 * t + 4 * h);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) f.b[t][s] = B[(k + 8 * t + 4 * h + s) * N + col];
    return;
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const long kk = k + 8 * t + 4 * h + s;
      f.a[t][s] = (row < M && kk < ke) ? A[row * K + kk] : 0.0f;
      f.b[t][s] = (col < N && kk < ke) ? B[kk * N + col] : 0.0f;
    }
}

// FULL: every quadrant inside C (M, N multiples of 64), every slice a whole SK_QGROUPS
// groups inside K, K % 4 == 0 and A 16-byte aligned: no guarded loads and no data-dependent
// branch in the unrolled slice, so each group waits only for its own loads. With the
// guarded path compiled in, the merged control flow made every group wait for all
// loads in flight (vmcnt(0)), which drained the prefetch ring once per group.
template <int SK_PD, bool FULL>
__global__ __launch_bounds__(256) void skinny_f32_kernel(const float *__restrict__ A, const float *__restrict__ B,
                                                         float *__restrict__ C, long M, long N, long K, long kper,
                                                         int tiles_n, int partial) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, l32 = lane & 31;
  // quadrant (wave >> 1, wave & 1) of the workgroup's 64 x 64 tile
  const long m0 = (long)(blockIdx.x / tiles_n) * SK_TILE + 32 * (wave >> 1);
  const long n0 = (long)(blockIdx.x % tiles_n) * SK_TILE + 32 * (wave & 1);
  if (!FULL && (m0 >= M || n0 >= N)) return;  // the quadrant lies outside C (no barrier follows)
  const long kb = (long)blockIdx.y * kper;
  const long ke = FULL ? kb + kper : (kb + kper < K ? kb + kper : K);
  const long row = m0 + l32, col = n0 + l32;
  const bool inside = FULL || (m0 + 32 <= M && n0 + 32 <= N && (K & 3) == 0 && ((uintptr_t)A & 15) == 0);
  f32x16 acc, acc2;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    acc[r] = 0.0f;
    acc2[r] = 0.0f;
  }
  // the slice holds at most SK_QGROUPS groups (kper == SK_QGROUPS * SK_GROUP); groups
  // past the end of K (the last slice) are skipped with a uniform branch
  const int ng = FULL ? SK_QGROUPS : (int)((ke - kb + SK_GROUP - 1) / SK_GROUP);
  auto load = [&](int g, SkFrag &f) {
    const long k = kb + (long)g * SK_GROUP;
    sk_load(A, B, M, N, K, ke, row, col, k, FULL || (inside && k + SK_GROUP <= ke), h, f);
  };
  auto mma = [&](const SkFrag &f) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int s = 0; s < 4; s += 2) {
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(f.a[t][s], f.b[t][s], acc, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_32x32x2f32(f.a[t][s + 1], f.b[t][s + 1], acc2, 0, 0, 0);
      }
  };
  // ring of SK_PD + 1 fragment buffers, the whole slice unrolled at compile time so
  // every buffer index is a constant (registers, not scratch) and the compiler sees
  // the full load / MFMA schedule (a runtime-length loop measured 13.2 against 9.5 us)
  // (the scheduling barriers keep each group's loads SK_PD groups ahead of its MFMAs:
  // left alone, the scheduler reuses a few registers and issues every load just before
  // the MFMA that reads it)
  SkFrag f[SK_PD + 1];
#pragma unroll
  for (int g = 0; g < SK_PD; ++g)
    if (g < ng) load(g, f[g]);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int g = 0; g < SK_QGROUPS; ++g) {
    if (g >= ng) break;
    if (g + SK_PD < ng) load(g + SK_PD, f[(g + SK_PD) % (SK_PD + 1)]);
    __builtin_amdgcn_sched_barrier(0);
    mma(f[g % (SK_PD + 1)]);
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] += acc2[r];
  // lane holds rows 8(r/4) + 4h + (r%4) of column l32 of the quadrant
  float *out = partial ? C + (size_t)blockIdx.y * M * N : C;
  if (!FULL && col >= N) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long m = m0 + 8 * (r >> 2) + 4 * h + (r & 3);
    if (FULL || m < M) out[m * N + col] = acc[r];
  }
}

// Long slices: each of the 4 waves computes the whole 64 x 64 tile for the groups
// g = wave, wave + 4, ... of the slice (so each loaded A fragment meets two B column
// blocks and each B fragment two A row blocks), one group ahead; the four wave tiles
// are summed through LDS in wave order.
struct SkTileFrag {
  f32x4 a[4][2];     // [k-step][row block]: k = 8t + 4h .. 8t + 4h + 3
  float b[4][2][4];  // [k-step][column block][s]: B[8t + 4h + s][32j + l32]
};

__device__ __forceinline__ void sk_tile_load(const float *__restrict__ A, const float *__restrict__ B, long M,
                                             long N, long K, long ke, long m0, long n0, long k, bool fast, int h,
                                             int l32, SkTileFrag &f) {
  if (fast) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
        f.a[t][i] = *(const f32x4 *)(A + (m0 + 32 * i + l32) * K + k + 8 * t + 4 * h);
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int s = 0; s < 4; ++s) f.b[t][j][s] = B[(k + 8 * t + 4 * h + s) * N + n0 + 32 * j + l32];
    }
    return;
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const long m = m0 + 32 * i + l32;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const long kk = k + 8 * t + 4 * h + s;
        f.a[t][i][s] = (m < M && kk < ke) ? A[m * K + kk] : 0.0f;
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const long n = n0 + 32 * j + l32;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const long kk = k + 8 * t + 4 * h + s;
        f.b[t][j][s] = (n < N && kk < ke) ? B[kk * N + n] : 0.0f;
      }
    }
  }
}

// FULL: the tile inside C (M, N multiples of 64), every slice whole and a multiple of
// 8 groups (each wave an even number of groups), K % 4 == 0, A 16-byte aligned: every
// load is unconditional (the prefetch past the slice's end re-reads the wave's current
// group), so each group waits only for its own loads instead of all in flight.
template <bool FULL>
__global__ __launch_bounds__(256) void skinny_tile_kernel(const float *__restrict__ A, const float *__restrict__ B,
                                                          float *__restrict__ C, long M, long N, long K, long kper,
                                                          int tiles_n, int partial) {
  __shared__ float red[SK_TILE * SK_TILE];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, l32 = lane & 31;
  const long m0 = (long)(blockIdx.x / tiles_n) * SK_TILE, n0 = (long)(blockIdx.x % tiles_n) * SK_TILE;
  const long kb = (long)blockIdx.y * kper;
  const long ke = FULL ? kb + kper : (kb + kper < K ? kb + kper : K);
  const bool inside = FULL || (m0 + SK_TILE <= M && n0 + SK_TILE <= N && (K & 3) == 0 && ((uintptr_t)A & 15) == 0);
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  SkTileFrag f[2];
  long k = kb + (long)wave * SK_GROUP;
  if (FULL || k < ke) sk_tile_load(A, B, M, N, K, ke, m0, n0, k, FULL || (inside && k + SK_GROUP <= ke), h, l32, f[0]);
  auto mma = [&](const SkTileFrag &g) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int i = 0; i < 2; ++i)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(g.a[t][i][s], g.b[t][j][s], acc[i][j], 0, 0, 0);
  };
  if (FULL) {
    for (; k < ke; k += 8 * SK_GROUP) {  // two groups per iteration, buffers alternating
      const long k1 = k + 4 * SK_GROUP, k2 = k + 8 * SK_GROUP;
      sk_tile_load(A, B, M, N, K, ke, m0, n0, k1, true, h, l32, f[1]);
      __builtin_amdgcn_sched_barrier(0);
      mma(f[0]);
      __builtin_amdgcn_sched_barrier(0);
      sk_tile_load(A, B, M, N, K, ke, m0, n0, k2 < ke ? k2 : k, true, h, l32, f[0]);
      __builtin_amdgcn_sched_barrier(0);
      mma(f[1]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  for (; !FULL && k < ke; k += 8 * SK_GROUP) {  // two groups per iteration, buffers alternating
    const long k1 = k + 4 * SK_GROUP, k2 = k + 8 * SK_GROUP;
    if (k1 < ke) sk_tile_load(A, B, M, N, K, ke, m0, n0, k1, inside && k1 + SK_GROUP <= ke, h, l32, f[1]);
    __builtin_amdgcn_sched_barrier(0);
    mma(f[0]);
    __builtin_amdgcn_sched_barrier(0);
    if (k1 >= ke) break;
    if (k2 < ke) sk_tile_load(A, B, M, N, K, ke, m0, n0, k2, inside && k2 + SK_GROUP <= ke, h, l32, f[0]);
    __builtin_amdgcn_sched_barrier(0);
    mma(f[1]);
    __builtin_amdgcn_sched_barrier(0);
  }
  // lane holds rows 8(r/4) + 4h + (r%4) of column 32j + l32 in row block i
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int row = 32 * i + 8 * (r >> 2) + 4 * h + (r & 3), col = 32 * j + l32;
            if (w == 0) red[row * SK_TILE + col] = acc[i][j][r];
            else red[row * SK_TILE + col] += acc[i][j][r];
          }
    }
    __syncthreads();
  }
  float *out = partial ? C + (size_t)blockIdx.y * M * N : C;
  for (int e = threadIdx.x; e < SK_TILE * SK_TILE; e += 256) {
    const long m = m0 + e / SK_TILE, n = n0 + e % SK_TILE;
    if (m < M && n < N) out[m * N + n] = red[e];
  }
}

}  // namespace

bool skinny_f32_shape(long M, long N) { return M <= SK_TILE || N <= SK_TILE; }

int skinny_f32_splits(long M, long N, long K, long *kper_out) {
  const long tiles = ((M + SK_TILE - 1) / SK_TILE) * ((N + SK_TILE - 1) / SK_TILE);
  long want = (256 + tiles - 1) / tiles;  // about one workgroup per CU (measured faster than two)
  long kper = (K + want - 1) / want;
  kper = (kper + 4 * SK_GROUP - 1) / (4 * SK_GROUP) * (4 * SK_GROUP);
  if (kper < 8 * SK_GROUP) kper = 8 * SK_GROUP;  // at least two groups per wave
  // short slices: the quadrant kernel's fixed 256-k slice (more slabs, same work)
  if (kper <= 16 * SK_GROUP) kper = SK_QGROUPS * SK_GROUP;
  *kper_out = kper;
  return (int)((K + kper - 1) / kper);
}

hipError_t launch_skinny_f32(const float *A, const float *B, float *C, long M, long N, long K, long kper,
                             int splits, hipStream_t s) {
  const int tiles_m = (int)((M + SK_TILE - 1) / SK_TILE), tiles_n = (int)((N + SK_TILE - 1) / SK_TILE);
  const dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)splits);
  const int partial = splits > 1;
  // Short slices (the split-K-heavy shapes, e.g. 64 x 64 x 65536 with 256 slices of
  // 256 k): one quadrant per wave, nothing exchanged (kernel 13.0 against 16.8 us for
  // that shape, rocprofv3). Long slices (tall shapes such as 8192 x 64 x 8192, 4096 k
  // per slice): every wave on the whole tile, which feeds each loaded fragment to two
  // MFMAs instead of one (~100 against 140 us for that shape).
  // every quadrant of every 64 x 64 tile inside C, every slice whole
  const bool full = M % SK_TILE == 0 && N % SK_TILE == 0 && K % kper == 0 && K % 4 == 0 && ((uintptr_t)A & 15) == 0;
  const bool full_tile = full && kper % (8 * SK_GROUP) == 0;
  if (kper == SK_QGROUPS * SK_GROUP && full)
    hipLaunchKernelGGL((skinny_f32_kernel<2, true>), grid, dim3(256), 0, s, A, B, C, M, N, K, kper, tiles_n, partial);
  else if (kper == SK_QGROUPS * SK_GROUP)
    hipLaunchKernelGGL((skinny_f32_kernel<2, false>), grid, dim3(256), 0, s, A, B, C, M, N, K, kper, tiles_n,
                       partial);
  else if (full_tile)
    hipLaunchKernelGGL(skinny_tile_kernel<true>, grid, dim3(256), 0, s, A, B, C, M, N, K, kper, tiles_n, partial);
  else
    hipLaunchKernelGGL(skinny_tile_kernel<false>, grid, dim3(256), 0, s, A, B, C, M, N, K, kper, tiles_n, partial);
  return hipGetLastError();
}

}  // namespace fg
