// Fused GEMM chain for gfx950: C = A_n ... A_1 B with every intermediate kept in LDS.
//
// This is the MI355X counterpart of the reference's per-panel chain loop
// (multigemm/src/flashgemm.c:123-148), where each thread walks 32-column panels of
// B and runs every GEMM of the chain, the epilogue of GEMM i writing straight into
// GEMM i+1's packed B (flashgemm.c:132). Here a workgroup owns a tile of `TN`
// columns of B and runs every GEMM of the chain on it:
//   * the weights are first re-laid out by pack_a_kernel into MFMA fragment order
//     (the counterpart of the reference's per-call A pack, PACK_x86.h:616-647):
//     one wave's 16-byte-per-lane fragment of 32 rows is 1 KB contiguous, so the
//     hot loop's weight loads are fully coalesced, rows and K are zero padded (no
//     bounds logic), INT8 bytes carry the u8 -> s8 bias (x ^ 0x80) and FP32 for the
//     split-bf16 form is stored as its bf16 hi/lo halves;
//   * the B tile (K1 x TN) is staged once into an LDS arena in the MFMA B-operand
//     layout (16-byte k-groups per column, slot g*TN + n);
//   * GEMM i reads X_i and writes its requantized output Y_i (fp32 / bf16 RNE /
//     int8 saturate) right after it in a circular arena, so X_{i+1} = Y_i never
//     leaves the CU; each wave holds its final row pass in registers until every
//     wave is done with X_i, which lets Y_i reuse X_i's rows: the arena needs
//     max_i (K_i + M_i - one pass of rows) rather than max_i (K_i + M_i);
//   * each wave owns row blocks of the weights and streams their packed fragments
//     as one sequence, PF k-steps per chunk, the next chunk in flight across row
//     passes and GEMM boundaries; B fragments are read one step ahead;
//   * INT8 uses u8(a).b = s8(a ^ 0x80).b + 128 sum_k b: the column sums of X_0 are
//     taken while the tile is staged and those of each Y_i by the epilogue that
//     writes it, and they seed the next GEMM's accumulators;
//   * only the last GEMM writes HBM (C, fp32 or int32, row-major).
// Variants (template parameters of chain_kernel): wide tiles whose waves split over
// the columns (G, for chains with few rows per GEMM), B gathered from an NCHW image
// while it is staged (IMG, the convolution front end), and epilogues with per-row
// bias / scale, INT8 zero point and ReLU (EPX).
#include "fg_internal.h"

namespace fg {
namespace {

constexpr int NW = 8;  // waves per workgroup
// Word 3 of a raw buffer resource on gfx9 (32-bit data format, no swizzle, no index
// stride): what the buffer loads of the weight stream use.
constexpr int BUF_DW3 = 0x00020000;
constexpr int NTH = NW * 64;
constexpr int MAXG = CHAIN_MAXG;
constexpr size_t LDS_LIMIT = 160 * 1024;

template <int KIND> struct CT;
// KG: k per 16-byte LDS group (per plane for F32X3); KC: k per MFMA step;
// KA: A elements per lane-half per step; ESZ: global element bytes; RP: row blocks
// per wave pass;
// PF: k-steps per prefetch chunk; FS: packed bytes per lane per fragment.
// FP32 runs 16x16x4 MFMAs (measured 99% of their issue rate from registers, against
// 94% for 32x32x2): a k-step is 16 k, its fragment two 16-row halves.
template <> struct CT<K_F32> {
  static constexpr int KG = 4, KC = 16, KA = 4, ESZ = 4, RP = 1, PF = 4, FS = 32;
  using Acc = f32x16;
};
template <> struct CT<K_BF16> {
  static constexpr int KG = 8, KC = 16, KA = 8, ESZ = 2, RP = 2, PF = 8, FS = 16;
  using Acc = f32x16;
};
template <> struct CT<K_I8> {
  static constexpr int KG = 16, KC = 32, KA = 16, ESZ = 1, RP = 2, PF = 8, FS = 16;
  using Acc = i32x16;
};
template <> struct CT<K_F32X3> {
  static constexpr int KG = 8, KC = 16, KA = 8, ESZ = 4, RP = 1, PF = 4, FS = 32;
  using Acc = f32x16;
};

template <int TN>
__device__ __forceinline__ int cslot(int g, int n) { return g * TN + n; }

// threadIdx.x passed through an empty asm: addresses derived from it are recomputed
// where they are used (two VALU ops) instead of being hoisted out of the GEMM loop and
// kept live, or spilled to scratch, across the whole chain. A spill reload would also
// wait for every weight load in flight.
__device__ __forceinline__ int fresh_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// --------------------------------------------------------- packed weights --
// Packed layout of A (M x K) for kind KIND: fragments f = (rg * kspad + ks) * RP + i
// for row group rg (RP * 32 rows), k-step ks < kspad and row block i; fragment f is
// 64 lanes x FS bytes at f * 64 * FS, lane l = h*32 + r holding row
// (rg*RP + i)*32 + r, elements k = ks*KC + h*KA .. + KA-1 (zero outside A). For
// F32X3 the fragment is the 1 KB hi half then the 1 KB lo half. For F32 it is two
// 1 KB halves of 16 rows: lane l = q*16 + r of half a holds row rg*32 + 16a + r,
// elements k = ks*16 + 4q .. + 3.
template <int KIND>
__host__ __device__ inline long kspad_of(long K) {
  using T = CT<KIND>;
  const long ks = (K + T::KC - 1) / T::KC;
  return (ks + T::PF - 1) / T::PF * T::PF;
}

template <int KIND>
__host__ __device__ inline long packed_frags(long M, long K) {
  using T = CT<KIND>;
  const long groups = (M + 32 * T::RP - 1) / (32 * T::RP);
  return groups * kspad_of<KIND>(K) * T::RP;
}

// One lane of packed fragment f of A.
template <int KIND>
__device__ __forceinline__ void pack_frag(const char *__restrict__ A, long M, long K, char *__restrict__ out, long f,
                                          int lane) {
  using T = CT<KIND>;
  if constexpr (KIND == K_F32) {
    const int q = lane >> 4, r16 = lane & 15;
    const unsigned kspad = (unsigned)kspad_of<KIND>(K);
    const unsigned fu = (unsigned)f, rg = fu / kspad, ks = fu - rg * kspad;
    const long k0 = (long)ks * 16 + 4 * q;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const long row = (long)rg * 32 + 16 * a + r16;
      uint32_t w[4] = {0u, 0u, 0u, 0u};
      const char *src = A + ((size_t)row * K + k0) * 4;
      if (row < M && k0 + 4 <= K && ((uintptr_t)src & 15) == 0) {
        const uint4 v = *(const uint4 *)src;
        w[0] = v.x;
        w[1] = v.y;
        w[2] = v.z;
        w[3] = v.w;
      } else if (row < M) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (k0 + e < K) w[e] = *(const uint32_t *)(src + 4 * e);
      }
      *(uint4 *)(out + f * 64 * T::FS + a * 1024 + lane * 16) = make_uint4(w[0], w[1], w[2], w[3]);
    }
    return;
  }
  const int h = lane >> 5, r = lane & 31;
  // 32-bit index math: a packed matrix has far fewer than 2^31 fragments, and a
  // 64-bit division is a long software sequence on the GPU
  const unsigned kspad = (unsigned)kspad_of<KIND>(K);
  const unsigned fu = (unsigned)f, per_rg = kspad * T::RP;
  const unsigned rg = fu / per_rg, rem = fu - rg * per_rg;
  const unsigned ks = rem / T::RP, i = rem % T::RP;
  const long row = (long)(rg * T::RP + i) * 32 + r;
  const long k0 = (long)ks * T::KC + h * T::KA;
  uint32_t w[T::KA * T::ESZ / 4];
#pragma unroll
  for (int d = 0; d < T::KA * T::ESZ / 4; ++d) w[d] = 0u;
  const char *src = A + ((size_t)row * K + k0) * T::ESZ;
  if (row < M && k0 + T::KA <= K && ((uintptr_t)src & 15) == 0) {
    // interior, aligned: the lane's KA elements are 16-byte loads
#pragma unroll
    for (int d = 0; d < T::KA * T::ESZ / 16; ++d) {
      const uint4 v = *(const uint4 *)(src + 16 * d);
      w[4 * d] = v.x;
      w[4 * d + 1] = v.y;
      w[4 * d + 2] = v.z;
      w[4 * d + 3] = v.w;
    }
    if (KIND == K_I8)
#pragma unroll
      for (int d = 0; d < 4; ++d) w[d] ^= 0x80808080u;  // u8 -> s8 bias
  } else if (row < M) {
#pragma unroll
    for (int e = 0; e < T::KA; ++e)
      if (k0 + e < K) {
        uint32_t x;
        if (T::ESZ == 4) x = *(const uint32_t *)(src + 4 * e);
        else if (T::ESZ == 2) x = *(const uint16_t *)(src + 2 * e);
        else x = (uint32_t)(*(const uint8_t *)(src + e) ^ 0x80u);  // u8 -> s8 bias
        w[(e * T::ESZ) >> 2] |= x << (((e * T::ESZ) & 3) * 8);
      }
  }
  char *dst = out + f * 64 * T::FS + lane * 16;
  if (KIND == K_F32X3) {
    uint32_t hi[4], lo[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      split_pk(__builtin_bit_cast(float, w[2 * e]), __builtin_bit_cast(float, w[2 * e + 1]), hi[e], lo[e]);
    *(uint4 *)dst = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    *(uint4 *)(dst + 1024) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
  } else {
    *(uint4 *)dst = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// Up to PACK_MAX matrices packed by one launch: fragment f belongs to matrix m when
// fstart[m] <= f < fstart[m+1]; four fragments (one per wave) per workgroup.
template <int KIND>
__global__ __launch_bounds__(256) void pack_a_kernel(PackJob J) {
  const long f = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (f >= J.fstart[J.n]) return;
  int m = 0;
  while (m + 1 < J.n && f >= J.fstart[m + 1]) ++m;
  pack_frag<KIND>((const char *)J.A[m], J.M[m], J.K[m], (char *)J.out[m], f - J.fstart[m], threadIdx.x & 63);
}

// One lane's packed A fragment: 16 bytes, or hi and lo halves for F32X3, or the two
// 16-row halves for F32.
template <int KIND> struct AFrag { uint4 v[KIND == K_F32X3 || KIND == K_F32 ? 2 : 1]; };

// One lane's B fragments for one k-step: TNB column blocks (hi and lo planes for
// F32X3; for F32 two 16-column halves per block, b[2j + half]).
template <int KIND, int TNB> struct BFrag {
  uint4 b[KIND == K_F32 ? 2 * TNB : TNB];
  uint4 lo[KIND == K_F32X3 ? TNB : 1];
};

// One k-step of MFMAs. FP32: 16x16x4 MFMAs, k sub-step s outermost, so the
// 4 TNB products between two MFMAs on one accumulator hide the MFMA latency; lane
// q*16 + l of half a of the A fragment holds row 16a + l, of b[2j + b] column
// 16b + l of block j, both at k = 4q + s, and sub-block (a, b) of block j is acc
// elements 4(2a + b) .. + 3.
template <int KIND, int TNB>
__device__ __forceinline__ void mma_k(const BFrag<KIND, TNB> &bf, const AFrag<KIND> (&af)[CT<KIND>::RP],
                                      typename CT<KIND>::Acc (&acc)[CT<KIND>::RP][TNB]) {
  constexpr int RP = CT<KIND>::RP;
  if constexpr (KIND == K_F32) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < TNB; ++j)
#pragma unroll
        for (int i = 0; i < RP; ++i)
#pragma unroll
          for (int sb = 0; sb < 4; ++sb) {
            const f32x4 av = __builtin_bit_cast(f32x4, af[i].v[sb >> 1]);
            const f32x4 bv = __builtin_bit_cast(f32x4, bf.b[2 * j + (sb & 1)]);
            f32x4 c = {acc[i][j][4 * sb], acc[i][j][4 * sb + 1], acc[i][j][4 * sb + 2], acc[i][j][4 * sb + 3]};
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], bv[s], c, 0, 0, 0);
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[i][j][4 * sb + e] = c[e];
          }
    return;
  }
#pragma unroll
  for (int j = 0; j < TNB; ++j) {
#pragma unroll
    for (int i = 0; i < RP; ++i) {
      if (KIND == K_F32X3) {
        // small terms first, then the dominant hi.hi product
        const bf16x8 ah = __builtin_bit_cast(bf16x8, af[i].v[0]), al = __builtin_bit_cast(bf16x8, af[i].v[1]);
        const bf16x8 bh = __builtin_bit_cast(bf16x8, bf.b[j]), bl = __builtin_bit_cast(bf16x8, bf.lo[j]);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[i][j], 0, 0, 0);
      } else if (KIND == K_BF16) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, af[i].v[0]),
                                                            __builtin_bit_cast(bf16x8, bf.b[j]), acc[i][j], 0, 0, 0);
      } else {
        acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(i32x4, af[i].v[0]),
                                                          __builtin_bit_cast(i32x4, bf.b[j]), acc[i][j], 0, 0, 0);
      }
    }
  }
}

// Stage the B tile (rows [0, K1p) x columns [n0, n0+TN)) into the arena at row 0.
// The fused path requires N % 4 == 0 and a 4-element aligned B, so each 4-column
// group lies entirely inside or outside the matrix; out-of-range rows and groups
// read a clamped address and are zeroed with a select. `issue` runs after the first
// round of tile loads and before their LDS writes: the caller's weight prefetch goes
// there, so the tile's loads are not queued behind it. INT8 adds each column's sum of
// s8 X_0 into colsum (zeroed beforehand). IMG: B is the patch matrix of the image at
// B (geometry img), gathered element by element; N % 4 and alignment do not matter.
template <int KIND, int TN, bool IMG, bool FASTB, typename F>
__device__ __forceinline__ void stage_b(uint4 *arena, int lo_off, const char *B, long K1, long K1p, long N, long n0,
                                        int *colsum, const ImgGeom &img, F &&issue) {
  using T = CT<KIND>;
  // one task = KG rows x 4 columns of one k-group
  const int ngroups = (int)(K1p / T::KG);
  const int tasks = ngroups * (TN / 4);
  // Raw loads first, then the transposition into slots, so all of a thread's
  // first-round loads are in flight together. Rows past K1 and groups past N read a
  // clamped (valid) address and are zeroed afterwards. When B spans less than 4 GiB
  // the addresses are 32-bit offsets from B, and a group inside B (every task but
  // those at the K1 edge) adds a uniform row stride per row.
  using Raw = typename std::conditional<T::ESZ == 4, uint4,
                                        typename std::conditional<T::ESZ == 2, uint2, uint32_t>::type>::type;
  const bool off32 = (size_t)K1 * (size_t)N * T::ESZ < (1ull << 32);
  // tasks per thread loaded before `issue` (BF16 tiles of 64 columns have two)
  constexpr int FIRST = KIND == K_BF16 ? 2 : 1;
  Raw x[FIRST][T::KG];
  // Patch gather: column n = (b*OH + oy)*OW + ox, row k = (c*kh + ky)*kw + kx reads
  // pixel (oy*sh - ph + ky*dh, ox*sw - pw + kx*dw) of channel c of image b, zero
  // outside the image. One division per task for its first column and first row;
  // the other columns and rows step their coordinates. Lanes outside the matrix or
  // the image load element 0 and select zero.
  auto load_img = [&](int task, Raw(&x)[T::KG]) {
    const int g = task / (TN / 4), cb = task % (TN / 4);
    const long nb = n0 + cb * 4;
    const int ohw = img.OH * img.OW, HW = img.H * img.W;
    int cy[4], cx[4], cbase[4];
    bool cok[4];
    {
      const int n = (int)(nb < N ? nb : 0);
      int b = n / ohw, r = n - b * ohw;
      int oy = r / img.OW, ox = r - oy * img.OW;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        cok[j] = nb + j < N;
        cy[j] = oy * img.sh - img.ph;
        cx[j] = ox * img.sw - img.pw;
        cbase[j] = (b * img.C * img.H + cy[j]) * img.W + cx[j];
        if (++ox == img.OW) {
          ox = 0;
          if (++oy == img.OH) {
            oy = 0;
            ++b;
          }
        }
      }
    }
    const int k0 = g * T::KG, kk = img.kh * img.kw;
    if (k0 >= K1) {  // a group of the chunk padding (wave-uniform: a wave covers one group)
#pragma unroll
      for (int r = 0; r < T::KG; ++r) x[r] = Raw{};
      return;
    }
    int c = k0 / kk, rr = k0 - c * kk;
    int ky = rr / img.kw, kx = rr - ky * img.kw;
#pragma unroll
    for (int r = 0; r < T::KG; ++r) {
      const bool rok = k0 + r < K1;
      const int dy = ky * img.dh, dx = kx * img.dw;
      const int roff = c * HW + dy * img.W + dx;
      if constexpr (T::ESZ == 1) {
        if (img.vec4) {
          // INT8 with stride 1 along x and OW % 4 == 0: the 4 columns are 4 adjacent
          // pixels of one image row, bytes base .. base + 3. Two aligned dword loads
          // (indices clamped into the image; a clamped dword only feeds bytes that
          // lie outside it, which the mask clears) and a byte funnel shift.
          const int base = cbase[0] + roff, ix0 = cx[0] + dx;
          const bool rowok = rok && (unsigned)(cy[0] + dy) < (unsigned)img.H;
          const int q = (base - (base & 3)) >> 2;
          const int ql = q < 0 ? 0 : (q >= img.nq ? img.nq - 1 : q);
          const int qh = q + 1 < 0 ? 0 : (q + 1 >= img.nq ? img.nq - 1 : q + 1);
          const uint32_t lo = ((const uint32_t *)B)[ql], hi = ((const uint32_t *)B)[qh];
          const uint32_t v = __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)(base & 3));
          uint32_t mask = 0u;
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (rowok && cok[j] && (unsigned)(ix0 + j) < (unsigned)img.W) mask |= 0xffu << (8 * j);
          x[r] = v & mask;
          if (++kx == img.kw) {
            kx = 0;
            if (++ky == img.kh) {
              ky = 0;
              ++c;
            }
          }
          continue;
        }
      }
      uint32_t e[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool ok = rok && cok[j] && (unsigned)(cy[j] + dy) < (unsigned)img.H &&
                        (unsigned)(cx[j] + dx) < (unsigned)img.W;
        const unsigned off = ok ? (unsigned)(cbase[j] + roff) : 0u;
        uint32_t v;
        if constexpr (T::ESZ == 4) v = *(const uint32_t *)(B + (size_t)off * 4);
        else if constexpr (T::ESZ == 2) v = *(const uint16_t *)(B + (size_t)off * 2);
        else v = *(const uint8_t *)(B + off);
        e[j] = ok ? v : 0u;
      }
      if constexpr (T::ESZ == 4) x[r] = make_uint4(e[0], e[1], e[2], e[3]);
      else if constexpr (T::ESZ == 2) x[r] = make_uint2(e[0] | (e[1] << 16), e[2] | (e[3] << 16));
      else x[r] = e[0] | (e[1] << 8) | (e[2] << 16) | (e[3] << 24);
      if (++kx == img.kw) {
        kx = 0;
        if (++ky == img.kh) {
          ky = 0;
          ++c;
        }
      }
    }
  };
  // INNER: the whole tile lies inside B (columns n0 .. n0 + TN - 1 < N, K1 a whole
  // number of chunks, 32-bit offsets): a task's rows are one 32-bit base plus a uniform
  // row stride, with no clamps and no edge zeroing (the generic path keeps 64-bit clamps
  // and masks live, some hundred vector instructions per wave before its first load)
  const unsigned rbu = (unsigned)N * T::ESZ, cbase = (unsigned)n0 * T::ESZ;
  auto load = [&](auto inner_c, int task, Raw(&x)[T::KG]) {
    if constexpr (IMG) {
      load_img(task, x);
      return;
    }
    if constexpr (decltype(inner_c)::value) {
      const unsigned g = (unsigned)task / (TN / 4), cb = (unsigned)task % (TN / 4);
      const unsigned o0 = g * T::KG * rbu + cbase + cb * 4 * T::ESZ;
#pragma unroll
      for (int r = 0; r < T::KG; ++r) x[r] = *(const Raw *)(B + (o0 + (unsigned)r * rbu));
      return;
    }
    const int g = task / (TN / 4), cb = task % (TN / 4);
    const long nb = n0 + cb * 4, k0 = (long)g * T::KG;
    const long nbc = nb < N ? nb : 0;
    if (off32) {
      const unsigned rb = (unsigned)N * T::ESZ, c32 = (unsigned)nbc * T::ESZ;
      if (k0 + T::KG <= K1) {
        const unsigned o0 = (unsigned)k0 * rb + c32;
#pragma unroll
        for (int r = 0; r < T::KG; ++r) x[r] = *(const Raw *)(B + (o0 + (unsigned)r * rb));
      } else {
#pragma unroll
        for (int r = 0; r < T::KG; ++r) {
          const unsigned kr = (unsigned)(k0 + r < K1 ? k0 + r : K1 - 1);
          x[r] = *(const Raw *)(B + (kr * rb + c32));
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < T::KG; ++r) {
        const long kr = k0 + r < K1 ? k0 + r : K1 - 1;
        x[r] = *(const Raw *)(B + ((size_t)kr * N + nbc) * T::ESZ);
      }
    }
  };
  // element (row r, column c) of a task as raw bits: component c of x[r] (FP32)
  auto comp = [](const uint4 &v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w)); };
  // Transpose a task (KG rows x 4 columns) into its 4 slots: byte permutes for INT8
  // (a 4 x 4 byte transpose is 8 v_perm_b32) and BF16 (one v_perm_b32 per pair of rows).
  auto put = [&](auto inner_c, int task, Raw(&x)[T::KG]) {
    const int g = task / (TN / 4), cb = task % (TN / 4);
    const long k0 = (long)g * T::KG;
    const bool nok = n0 + cb * 4 < N;
    if (!decltype(inner_c)::value && !(nok && k0 + T::KG <= K1)) {  // edge task
#pragma unroll
      for (int r = 0; r < T::KG; ++r)
        if (!(nok && k0 + r < K1)) x[r] = Raw{};
    }
    uint4 *const sl = arena + cslot<TN>(g, cb * 4);
    if constexpr (KIND == K_F32) {
#pragma unroll
      for (int c = 0; c < 4; ++c) sl[c] = make_uint4(comp(x[0], c), comp(x[1], c), comp(x[2], c), comp(x[3], c));
    } else if constexpr (KIND == K_BF16) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const uint32_t sel = (c & 1) ? 0x07060302u : 0x05040100u;  // high or low halves of both words
        uint32_t w[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
          w[e] = __builtin_amdgcn_perm(c < 2 ? x[2 * e + 1].x : x[2 * e + 1].y, c < 2 ? x[2 * e].x : x[2 * e].y, sel);
        sl[c] = make_uint4(w[0], w[1], w[2], w[3]);
      }
    } else if constexpr (KIND == K_I8) {
      uint32_t w[4][4];  // w[c][d]: rows 4d..4d+3 of column c
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const uint32_t t0 = __builtin_amdgcn_perm(x[4 * d + 1], x[4 * d], 0x05010400u);      // r0.b0 r1.b0 r0.b1 r1.b1
        const uint32_t t1 = __builtin_amdgcn_perm(x[4 * d + 3], x[4 * d + 2], 0x05010400u);  // r2.b0 r3.b0 r2.b1 r3.b1
        const uint32_t t2 = __builtin_amdgcn_perm(x[4 * d + 1], x[4 * d], 0x07030602u);      // r0.b2 r1.b2 r0.b3 r1.b3
        const uint32_t t3 = __builtin_amdgcn_perm(x[4 * d + 3], x[4 * d + 2], 0x07030602u);
        w[0][d] = __builtin_amdgcn_perm(t1, t0, 0x05040100u);
        w[1][d] = __builtin_amdgcn_perm(t1, t0, 0x07060302u);
        w[2][d] = __builtin_amdgcn_perm(t3, t2, 0x05040100u);
        w[3][d] = __builtin_amdgcn_perm(t3, t2, 0x07060302u);
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        int cs = 0;
#pragma unroll
        for (int d = 0; d < 4; ++d) cs = __builtin_amdgcn_sdot4((int)w[c][d], 0x01010101, cs, false);
        sl[c] = make_uint4(w[c][0], w[c][1], w[c][2], w[c][3]);
        atomicAdd(&colsum[cb * 4 + c], cs);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        uint32_t h[4], l[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          split_pk(__builtin_bit_cast(float, comp(x[2 * r], c)), __builtin_bit_cast(float, comp(x[2 * r + 1], c)), h[r],
                   l[r]);
        sl[c] = make_uint4(h[0], h[1], h[2], h[3]);
        sl[lo_off + c] = make_uint4(l[0], l[1], l[2], l[3]);
      }
    }
  };
  const int tid = fresh_tid();
  auto drive = [&](auto inner_c) {
#pragma unroll
    for (int f = 0; f < FIRST; ++f)
      if (tid + f * NTH < tasks) load(inner_c, tid + f * NTH, x[f]);
    __builtin_amdgcn_sched_barrier(0);
    issue();
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int f = 0; f < FIRST; ++f)
      if (tid + f * NTH < tasks) put(inner_c, tid + f * NTH, x[f]);
    for (int task = tid + FIRST * NTH; task < tasks; task += NTH) {
      load(inner_c, task, x[0]);
      put(inner_c, task, x[0]);
    }
  };
  // wave-uniform (n0, N and K1 are the workgroup's), so every thread takes one branch
  // and the barrier inside issue() is reached by all of them. The grouped kernel
  // (FASTB false) keeps the generic path only: its code, four kinds of chain tile,
  // already exceeds the instruction cache, and the second copy of the staging code
  // cost its mixed batches more than the fast path saved.
  if constexpr (FASTB) {
    if (!IMG && off32 && n0 + TN <= N && K1p == K1) {
      drive(std::true_type{});
      return;
    }
  }
  drive(std::false_type{});
}

// Intermediate epilogue in the reference's layout (ChainParams::refl). The reference's
// GEMM i+1 reads, inside each 32-column panel, X[k][c] = Z[src_row][src_col]
// (ref_src: INT8 (k&~3) + (c>>3), 4(c&7) + (k&3); BF16 (k&~1) + (c>>4), 2(c&15) + (k&1)),
// a bijection within 4 (INT8) or 2 (BF16) rows of the panel. When every panel is a
// 32-aligned block of the tile, GEMM i's epilogue applies the inverse while it writes
// the arena: Z[m][j] goes to
//   INT8: X[4(m>>2) + (j&3)][8(m&3) + (j>>2)]     BF16: X[2(m>>1) + (j&1)][16(m&1) + (j>>1)]
// of its panel. Lane (h, l32) holds rows m = 8q + 4h + e (e = 0..3) of panel column
// j = l32, so its four values land in one X row (INT8) or two (BF16), in four panel
// columns: byte / halfword LDS stores. Every M_i is a multiple of 32 on this path (no
// padded rows inside a block). INT8 also adds each value to the column sum of the X
// column it lands in (the next GEMM's u8 bias).
template <int KIND, int TN, int TNW>
__device__ __forceinline__ void store_rows_refl(const ChainParams &P, int gi, long rb0_l, uint4 *arena,
                                                const typename CT<KIND>::Acc (&acc)[CT<KIND>::RP][TNW], int *csum,
                                                int cb0) {
  using T = CT<KIND>;
  constexpr int RP = T::RP;
  const int S = P.S;
  const int lane = fresh_tid() & 63;
  const int h = lane >> 5, l32 = lane & 31;
  const int Mp = (int)P.M[gi];
  const int rb0 = (int)rb0_l;
  int cs[TNW][4];
#pragma unroll
  for (int j = 0; j < TNW; ++j)
#pragma unroll
    for (int e = 0; e < 4; ++e) cs[j][e] = 0;
#pragma unroll
  for (int i = 0; i < RP; ++i) {
    const int rbase = rb0 + i * 32;
    if (rbase >= Mp) continue;
    int rs = P.ybase[gi] + rbase;  // arena row of the block (a multiple of 32; no wrap inside)
    if (rs >= S) rs -= S;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int j = 0; j < TNW; ++j) {
        const int col0 = (cb0 + j) * 32;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (KIND == K_I8) {
            const int k = rs + 8 * q + 4 * h + (l32 & 3);
            const int c = col0 + 8 * e + (l32 >> 2);
            const int8_t v = sat_i8((int)acc[i][j][4 * q + e]);
            ((int8_t *)(arena + (k / T::KG) * TN + c))[k % T::KG] = v;
            cs[j][e] += v;
          } else {
            const int k = rs + 8 * q + 4 * h + 2 * (e >> 1) + (l32 & 1);
            const int c = col0 + 16 * (e & 1) + (l32 >> 1);
            ((uint16_t *)(arena + (k / T::KG) * TN + c))[k % T::KG] = bf16_dev((float)acc[i][j][4 * q + e]);
          }
        }
      }
    }
  }
  if (KIND == K_I8)
#pragma unroll
    for (int j = 0; j < TNW; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) atomicAdd(&csum[(cb0 + j) * 32 + 8 * e + (l32 >> 2)], cs[j][e]);
}

// Epilogue of one pass: element r of acc[i][j] is row lrow + crow(r) and column
// lcol + ccol(r) of the 32 x 32 block (rb0 + 32 i, column block j): for the 32x32
// MFMAs rows 4h + 8(r/4) + (r%4) of column l32, for FP32's 16x16 MFMAs (sub-block
// r/4 = 2a + b) rows 4q + 16a + (r%4) of column l16 + 16b, q = lane / 16. The last GEMM writes C, the others Y_gi into the arena (INT8
// also adds the s8 column sums of what it writes into csum: the next GEMM's bias).
// Index math is 32-bit (arena rows and M fit easily) and the row-bound checks are
// only evaluated in the block that straddles M. EPX applies GEMM gi's epilogue
// extensions (P.epi.g[gi]: per-row bias and scale, zero point, ReLU; epi_i8, epi_i32,
// epi_f); rows past M read no bias or scale.
template <int KIND, int TN, int TNW, bool EPX>
__device__ __forceinline__ void store_rows_impl(const ChainParams &P, int gi, long rb0_l, long n0, uint4 *arena,
                                                const typename CT<KIND>::Acc (&acc)[CT<KIND>::RP][TNW], int *csum,
                                                int cb0) {
  using T = CT<KIND>;
  constexpr int RP = T::RP;
  const int S = P.S;
  const int SG = S / T::KG;
  const int lo_off = (KIND == K_F32X3) ? SG * TN : 0;
  const int lane = fresh_tid() & 63;
  const int h = lane >> 5, l32 = lane & 31;
  const int lrow = KIND == K_F32 ? 4 * (lane >> 4) : 4 * h, lcol = KIND == K_F32 ? (lane & 15) : l32;
  auto crow = [](int r) { return KIND == K_F32 ? 16 * (r >> 3) + (r & 3) : 8 * (r >> 2) + (r & 3); };
  auto ccol = [](int r) { return KIND == K_F32 ? 16 * ((r >> 2) & 1) : 0; };
  const int M = (int)P.M[gi], Mp = (M + 31) & ~31;
  const int rb0 = (int)rb0_l;
  const EpiOne ex = EPX ? P.epi.g[gi] : EpiOne{};
  // per-row bias bits and scales of GEMM gi in LDS (filled at kernel start)
  const uint32_t *lb = nullptr;
  const float *ls = nullptr;
  if (EPX) {
    int rows = 0, off = 0;
    for (int g = 0; g < P.num_gemm; ++g) {
      off += g < gi ? (int)P.M[g] : 0;
      rows += (int)P.M[g];
    }
    lb = (const uint32_t *)((const char *)arena + P.epi_off) + off;
    ls = (const float *)((const uint32_t *)((const char *)arena + P.epi_off) + rows) + off;
  }
  const bool has_bias = ex.bias != nullptr, has_scale = ex.scale != nullptr;
  if (gi == P.num_gemm - 1 && (KIND == K_I8 || KIND == K_BF16) && P.out_requant) {
    // a fused segment that ends inside a longer chain: its last GEMM writes the next
    // GEMM's operand to HBM, requantized exactly as the unfused GEMM's epilogue does
    // (int8 saturate, bf16 RNE), row-major M x N
    const long tile_w = (P.N - n0 < TN) ? P.N - n0 : TN;
#pragma unroll
    for (int i = 0; i < RP; ++i) {
      const int rbase = rb0 + i * 32;
      if (rbase >= Mp) continue;
#pragma unroll
      for (int j = 0; j < TNW; ++j) {
        const int n = (cb0 + j) * 32 + l32;
        if (n >= tile_w) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int m = rbase + 4 * h + 8 * (r >> 2) + (r & 3);
          if (m >= M) continue;
          const size_t e = (size_t)m * P.N + n0 + n;
          if (KIND == K_I8) ((int8_t *)P.C)[e] = sat_i8((int)acc[i][j][r]);
          else ((uint16_t *)P.C)[e] = bf16_dev((float)acc[i][j][r]);
        }
      }
    }
    return;
  }
  if (gi == P.num_gemm - 1) {
    const long tile_w = (P.N - n0 < TN) ? P.N - n0 : TN;
    const size_t rs = (size_t)P.N * 4;  // bytes per row of C
#pragma unroll
    for (int i = 0; i < RP; ++i) {
      const int rbase = rb0 + i * 32;
      if (rbase >= Mp) continue;
      const bool full = rbase + 32 <= M;
#pragma unroll
      for (int j = 0; j < TNW; ++j) {
        const int n = (cb0 + j) * 32 + lcol;
        char *base = (char *)P.C + ((size_t)(rbase + lrow) * P.N + n0 + n) * 4;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int m = rbase + lrow + crow(r);
          if ((full || m < M) && n + ccol(r) < tile_w) {
            char *const dst = base + crow(r) * rs + ccol(r) * 4;
            if (KIND == K_I8) {
              int v = (int)acc[i][j][r];
              if (EPX) v = epi_i32(v, (int)lb[m], ex.relu);
              __builtin_nontemporal_store(v, (int *)dst);
            } else {
              float v = (float)acc[i][j][r];
              if (EPX) v = epi_f(v, has_scale, ls[m], has_bias, __builtin_bit_cast(float, lb[m]), ex.relu);
              __builtin_nontemporal_store(v, (float *)dst);
            }
          }
        }
      }
    }
    return;
  }
  if constexpr (KIND == K_I8 || KIND == K_BF16) {
    if (P.refl) {
      store_rows_refl<KIND, TN, TNW>(P, gi, rb0, arena, acc, csum, cb0);
      return;
    }
  }
  // rows rbase + lrow + crow(4q) + (0..3) of column lcol + ccol(4q), q = 0..3: four
  // consecutive rows per q. Padded rows (m >= M) are stored as zero: they are the
  // next GEMM's k-padding. A 32-row block starts on a multiple of 32 and S is a
  // multiple of a chunk, so no block wraps: its slots are a wave-uniform block base,
  // a per-lane offset (lrow, lcol) and a compile-time offset per (q, j), which the
  // LDS stores take as immediates. Row r of the block lies in group r / KG at byte
  // (r % KG) * EB of its slot (lrow and crow(4q) are multiples of 4, and of KG when
  // KG is 4).
  constexpr int EB = (KIND == K_F32) ? 4 : (KIND == K_I8 ? 1 : 2);  // arena bytes per element (per plane)
  const int lane_off = lcol * 16 + (lrow / T::KG) * TN * 16 + (lrow % T::KG) * EB;
  int cs[TNW];
#pragma unroll
  for (int j = 0; j < TNW; ++j) cs[j] = 0;
  // A block of 32 real rows (every block but a ragged last one) takes a path with no
  // row masks: a wave-uniform branch, so the mask arithmetic is not computed (and
  // selected away) for every value of every full block.
  auto block = [&](const int i, auto full_c) {
    constexpr bool full = decltype(full_c)::value;
    const int rbase = rb0 + i * 32;
    int rs = P.ybase[gi] + rbase;  // arena row of the block, a multiple of 32
    if (rs >= S) rs -= S;
    char *const blk = (char *)arena + (rs / T::KG) * TN * 16 + lane_off;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int m0r = rbase + lrow + crow(4 * q);
      const int valid = full ? 4 : (M - m0r < 0 ? 0 : (M - m0r > 4 ? 4 : M - m0r));
      // per-row extension operands of the 4 rows (rows past M use row M - 1)
      int eb[4];
      float es[4], ef[4];
      if (EPX) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int m = m0r + e < M ? m0r + e : M - 1;
          eb[e] = (int)lb[m];
          ef[e] = __builtin_bit_cast(float, lb[m]);
          es[e] = ls[m];
        }
      }
#pragma unroll
      for (int j = 0; j < TNW; ++j) {
        char *const dst =
            blk + ((crow(4 * q) / T::KG) * TN + (cb0 + j) * 32 + ccol(4 * q)) * 16 + (crow(4 * q) % T::KG) * EB;
        if (KIND == K_I8) {
          int a4[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            a4[e] = (int)acc[i][j][4 * q + e];
            if (EPX) a4[e] = epi_i8(a4[e], eb[e], has_scale, es[e], ex.zp, ex.relu);
          }
          uint32_t w = sat_pk_i8(a4[0], a4[1], a4[2], a4[3]);
          if (!full) w &= valid == 0 ? 0u : (0xffffffffu >> (8 * (4 - valid)));
          *(uint32_t *)dst = w;
          cs[j] = __builtin_amdgcn_sdot4((int)w, 0x01010101, cs[j], false);
        } else {
          float v[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            float a = (float)acc[i][j][4 * q + e];
            if (EPX) a = epi_f(a, has_scale, es[e], has_bias, ef[e], ex.relu);
            v[e] = (full || e < valid) ? a : 0.0f;
          }
          if (KIND == K_F32) {
            *(uint4 *)dst = make_uint4(__builtin_bit_cast(uint32_t, v[0]), __builtin_bit_cast(uint32_t, v[1]),
                                       __builtin_bit_cast(uint32_t, v[2]), __builtin_bit_cast(uint32_t, v[3]));
          } else if (KIND == K_BF16) {
            *(uint2 *)dst = make_uint2(cvt_pk_bf16(v[0], v[1]), cvt_pk_bf16(v[2], v[3]));
          } else {
            uint32_t h01, l01, h23, l23;
            split_pk(v[0], v[1], h01, l01);
            split_pk(v[2], v[3], h23, l23);
            *(uint2 *)dst = make_uint2(h01, h23);
            *(uint2 *)(dst + (size_t)lo_off * 16) = make_uint2(l01, l23);
          }
        }
      }
    }
  };
#pragma unroll
  for (int i = 0; i < RP; ++i) {
    const int rbase = rb0 + i * 32;
    if (rbase >= Mp) continue;
    if (rbase + 32 <= M) block(i, std::true_type{});
    else block(i, std::false_type{});
  }
  if (KIND == K_I8)
#pragma unroll
    for (int j = 0; j < TNW; ++j) atomicAdd(&csum[(cb0 + j) * 32 + l32], cs[j]);  // integer: order independent
}

// One workgroup runs the whole chain on columns [n0, n0 + TN) of B. P.A[i] are
// packed weights (pack_a_kernel layout for KIND).
//
// The arena is circular (P.S rows). X_i starts at xbase[i] and spans K_i rounded up
// to a chunk (PF k-steps, the k-padding of the packed weights), Y_i follows it at
// ybase[i] (mod S); regions start on chunk boundaries, so no chunk of X wraps and
// every B read of a chunk is the chunk's base plus an immediate offset. Rows of X
// past K are zero. Each wave keeps its final row pass of every GEMM in registers
// until all waves are done with X_i, then writes it, possibly over X_i; earlier
// passes go to rows that never meet X_i. Per wave, weights and B are double
// buffered a whole chunk ahead: while chunk c's MFMAs run, chunk c+1's weight
// loads (global) and B reads (LDS) are in flight. The weight stream continues
// across row passes and GEMM boundaries.
//
// G > 1 splits the waves into G column groups: wave w owns the TNB / G column blocks
// starting at block (w % G) * TNB / G and walks the rows of every GEMM with the
// NW / G waves of its group. Wide tiles with G = TNB (one 32-column block per wave)
// keep every wave busy in chains whose M_i are a few row blocks, where G = 1 would
// leave most waves without rows.
template <int KIND, int TNB, bool IMG = false, bool EPX = false, int G = 1, bool RAWA = false, bool FASTB = true>
__device__ __forceinline__ void chain_tile(const ChainParams &P, const long n0, uint4 *arena) {
  using T = CT<KIND>;
  using Acc = typename T::Acc;
  static_assert(TNB % G == 0 && NW % G == 0, "column groups must split the tile and the waves");
  constexpr int TN = TNB * 32;
  constexpr int TNW = TNB / G;               // column blocks per wave
  constexpr int NWR = NW / G;                // waves per column group
  constexpr int RP = T::RP;
  constexpr int PF = T::PF;
  constexpr int FRAG = 64 * T::FS;           // bytes per packed fragment
  constexpr long PASS = (long)NWR * RP * 32;  // rows of one pass of a column group
  constexpr long CS = (long)PF * T::KC;      // rows of one chunk
  const int SG = P.S / T::KG;                // arena k-groups
  const int lo_off = (KIND == K_F32X3) ? SG * TN : 0;  // lo plane (F32X3), in slots
  // INT8: three buffers of TN column sums; GEMM i reads its bias from buffer i % 3,
  // its epilogues sum Y_i into buffer (i + 1) % 3, and it clears buffer (i + 2) % 3
  // for GEMM i + 1 (the barrier that ends GEMM i orders the clear before any use)
  int *colsum = (int *)(arena + (size_t)SG * TN * (KIND == K_F32X3 ? 2 : 1));

  const int lane = threadIdx.x & 63;
  const int wave_id = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
  const int wave = wave_id / G;           // row position inside the column group
  const int cb0 = (wave_id % G) * TNW;    // first column block of this wave
  const int h = lane >> 5, l32 = lane & 31;

  // first packed-weight chunk of this wave in GEMM gi (nullptr: no rows there);
  // fragment (ks, i) of a pass at base + (ks * RP + i) * FRAG. The chunk pointers are
  // wave-uniform (scalar registers); each load adds the lane's 16-byte offset, so
  // the weight stream costs one vector register for addressing.
  //
  // RAW (FP32 with every M_i % 32 == 0 and K_i a whole number of chunks): P.A[gi] are
  // the row-major weights themselves, read in the packed fragments' element order
  // (lane q * 16 + r of half a: row 16a + r of the block, k 4q..4q+3 of the step) at
  // row stride K_gi, so the per-call pack is skipped; the lane offset and the byte
  // offset of the second half (16 rows) then depend on the GEMM and travel with the
  // chunk pointer (a chunk of the next GEMM is fetched early).
  constexpr bool RAW = RAWA && KIND == K_F32;
  static_assert(!RAW || RP == 1, "row-major weights: one row block per wave pass");
  auto first_ptr = [&](int gi) -> const char * {
    if ((long)wave * RP * 32 >= (P.M[gi] + 31) / 32 * 32) return nullptr;
    if (RAW) return (const char *)P.A[gi] + (size_t)wave * RP * 32 * P.K[gi] * 4;
    return (const char *)P.A[gi] + (size_t)wave * kspad_of<KIND>(P.K[gi]) * RP * FRAG;
  };
  const unsigned lane16 = (unsigned)lane * 16u;
  // lane offset and second-half offset (wave-uniform) of GEMM gi's weight stream
  struct Lane {
    unsigned v;
    int h1;
  };
  auto lane_of = [&](int gi) -> Lane {
    if (RAW) return Lane{(unsigned)(((long)(lane & 15) * P.K[gi] + 4 * (lane >> 4)) * 4), (int)(16 * 4 * P.K[gi])};
    return Lane{lane16, 1024};
  };
  constexpr size_t STEP = RAW ? (size_t)T::KC * 4 : (size_t)RP * FRAG;  // bytes per k-step
  AFrag<KIND> fa[2][PF][RP];
  // the RP fragments of k-step s of the chunk at base
  // (packed: the carried lane offset is not read, so it costs no register; a carried
  // copy of lane16 the compiler cannot prove constant made the INT8 64-column kernel
  // spill)
  //
  // Buffer loads: the chunk pointer (wave-uniform) becomes the resource base, the step
  // a scalar offset, the fragment an immediate and the lane offset the one vector
  // operand, so a load costs no vector address arithmetic. With flat 64-bit addresses
  // each step spent vector adds with carries on them, and a microbenchmark of the INT8
  // 64-column loop (256 workgroups, weights from L2) ran at 209 ns per k-step that way
  // against 158 ns with buffer loads. The record count only has to cover the chunk's
  // reads (offsets < 2 GiB), which are inside the weights.
  auto load_step = [&](const char *base, Lane ln, int s, AFrag<KIND>(&dst)[RP]) {
    const unsigned lo = RAW ? ln.v : lane16;
    const int h1 = RAW ? ln.h1 : 1024;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, 0x7fffffff, BUF_DW3);
#pragma unroll
    for (int i = 0; i < RP; ++i) {
      const int so = (int)((size_t)s * STEP + (size_t)i * FRAG);
      dst[i].v[0] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)lo, so, 0));
      if (KIND == K_F32X3 || KIND == K_F32)
        dst[i].v[1] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)lo, so + h1, 0));
    }
  };
  // Issued strictly in step order: the loop waits for step s of a chunk with a count
  // of the loads issued after it, and a reordered burst (step 0 issued last) made the
  // wait analysis assume so on every later chunk too, a near-full drain of the stream
  // once per two chunks.
  auto load_chunk = [&](const char *base, Lane lo, AFrag<KIND>(&dst)[PF][RP]) {
#pragma unroll
    for (int s = 0; s < PF; ++s) {
      load_step(base, lo, s, dst[s]);
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  constexpr size_t CHUNK = (size_t)PF * STEP;

  // B fragments of one k-step: column block j of this lane at slot
  // (g0 + 2 s + h) * TN + 32 j + l32 for step s of the chunk whose first group is g0;
  // FP32 (four groups per step): 16-column half u of block j at slot
  // (g0 + 4 s + q) * TN + 32 j + 16 u + l16, q = lane / 16, l16 = lane % 16
  constexpr int GPS = T::KC / T::KG;  // arena groups per k-step
  const uint4 *lane_b = KIND == K_F32 ? arena + (lane >> 4) * TN + cb0 * 32 + (lane & 15)
                                      : arena + h * TN + cb0 * 32 + l32;
  auto read_b = [&](int g0, int s, BFrag<KIND, TNW> &dst) {
    const uint4 *p = lane_b + (g0 + GPS * s) * TN;
#pragma unroll
    for (int j = 0; j < (KIND == K_F32 ? 2 * TNW : TNW); ++j) {
      dst.b[j] = p[(KIND == K_F32 ? 16 : 32) * j];
      if (KIND == K_F32X3) dst.lo[j] = p[lo_off + 32 * j];
    }
  };

  // invariant: at the start of each pass fa[0] holds (or has in flight) chunk 0
  const char *apk = first_ptr(0);
  Lane alane = lane_of(0);  // lane offsets of the GEMM apk belongs to
  const long K1c = (P.K[0] + CS - 1) / CS * CS;
  // A wave with no rows in GEMM 0 loads (and ignores) wave 0's first chunk, which
  // always exists: the loads stay unconditional and the wait counts after them exact.
  // Between the tile's first loads and their LDS writes: the weight prefetch, and
  // (INT8) the clearing of the column sums the tile's writes add into, with the
  // barrier that orders the two (so it overlaps the tile's load latency).
  stage_b<KIND, TN, IMG, FASTB>(arena, lo_off, (const char *)P.B, P.K[0], K1c, P.N, n0, colsum, P.img, [&] {
    load_chunk(apk ? apk : (const char *)P.A[0], alane, fa[0]);
    if (KIND == K_I8) {
      for (int t = fresh_tid(); t < 2 * TN; t += NTH) colsum[t] = 0;
      __syncthreads();
    }
  });
  if (EPX) {
    // every GEMM's per-row bias and scale into LDS (bias bits, then scales, rows of
    // GEMM g after those of GEMMs < g): the epilogues read them with LDS loads, so
    // they never wait behind the weight stream's global loads
    uint32_t *const eb = (uint32_t *)((char *)arena + P.epi_off);
    int rows = 0;
    for (int g = 0; g < P.num_gemm; ++g) rows += (int)P.M[g];
    float *const es = (float *)(eb + rows);
    int off = 0;
    for (int g = 0; g < P.num_gemm; ++g) {
      const uint32_t *bsrc = (const uint32_t *)P.epi.g[g].bias;
      const float *ssrc = P.epi.g[g].scale;
      for (int t = fresh_tid(); t < (int)P.M[g]; t += NTH) {
        eb[off + t] = bsrc ? bsrc[t] : 0u;
        es[off + t] = ssrc ? ssrc[t] : 1.0f;
      }
      off += (int)P.M[g];
    }
  }
  __syncthreads();

  // rows [Mp, Mc) of Y_gi (the k-padding of the next GEMM) must read as zero
  auto zero_pad = [&](int gi) {
    const long Mp = (P.M[gi] + 31) / 32 * 32, Mc = kspad_of<KIND>(P.M[gi]) * T::KC;
    const int zg0 = (int)((P.ybase[gi] + Mp) / T::KG), nzg = (int)((Mc - Mp) / T::KG);
    for (int t = fresh_tid(); t < nzg * TN; t += NTH) {
      const int g = zg0 + t / TN < SG ? zg0 + t / TN : zg0 + t / TN - SG;
      arena[cslot<TN>(g, t % TN)] = make_uint4(0u, 0u, 0u, 0u);
      if (KIND == K_F32X3) arena[lo_off + cslot<TN>(g, t % TN)] = make_uint4(0u, 0u, 0u, 0u);
    }
  };

  for (int gi = 0; gi < P.num_gemm; ++gi) {
    const long M = P.M[gi], K = P.K[gi];
    const long Mp = (M + 31) / 32 * 32;
    const int xg0 = P.xbase[gi] / T::KG;
    const bool last = gi == P.num_gemm - 1;
    // k-steps padded to whole chunks: the packed weights are zero there and so are
    // the rows of X
    const int kspad = (int)kspad_of<KIND>(K);
    const int nchunks = kspad / PF;
    const int ksteps = (int)((K + T::KC - 1) / T::KC);  // k-steps with real data
    const size_t pass_stride = RAW ? (size_t)NWR * RP * 32 * K * 4 : (size_t)NWR * kspad * RP * FRAG;
    const char *next_gemm = last ? nullptr : first_ptr(gi + 1);
    const Lane next_lane = last ? alane : lane_of(gi + 1);
    auto chunk_g0 = [&](int c) {  // first arena group of chunk c of X_gi
      const int g = xg0 + GPS * PF * c;
      return g < SG ? g : g - SG;
    };
    // INT8: the accumulators start at the bias 128 * sum_k X[k][n] of their column,
    // so the MFMAs accumulate u8(a).b directly.
    int bias0[TNW];
#pragma unroll
    for (int j = 0; j < TNW; ++j)
      bias0[j] = (KIND == K_I8) ? 128 * colsum[(gi % 3) * TN + (cb0 + j) * 32 + (fresh_tid() & 31)] : 0;
    int *const csum_next = colsum + ((gi + 1) % 3) * TN;
    if (KIND == K_I8)
      for (int t = fresh_tid(); t < TN; t += NTH) colsum[((gi + 2) % 3) * TN + t] = 0;

    Acc acc[RP][TNW];
    long held_rb0 = -1;  // final pass kept in acc until X_gi is dead
    for (long rb0 = (long)wave * RP * 32; rb0 < Mp; rb0 += PASS) {
#pragma unroll
      for (int i = 0; i < RP; ++i)
#pragma unroll
        for (int j = 0; j < TNW; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) acc[i][j][r] = bias0[j];
      const bool final_pass = rb0 + PASS >= Mp;
      const char *nxt = final_pass ? next_gemm : apk + pass_stride;
      const Lane after_lane = (final_pass && next_gemm) ? next_lane : alane;
      // With no next pass the prefetch re-reads this pass's first chunk: the loads
      // stay unconditional, so their wait counts are exact.
      const char *after = nxt ? nxt : apk;

      // B fragments one step ahead, double buffered (PF is even, so the buffer of
      // step ks is ks & 1 across chunks); steps past the data (the last chunk's
      // padding) are skipped with a uniform branch.
      BFrag<KIND, TNW> bq[2];
      auto run_chunk = [&](int c, const AFrag<KIND>(&a)[PF][RP]) {
        const int g0 = chunk_g0(c), gn = chunk_g0(c + 1 < nchunks ? c + 1 : 0);
#pragma unroll
        for (int s = 0; s < PF; ++s) {
          if (c * PF + s >= ksteps) break;
          if (s + 1 < PF) read_b(g0, s + 1, bq[(s + 1) & 1]);
          else read_b(gn, 0, bq[0]);
          __builtin_amdgcn_sched_barrier(0);
          mma_k<KIND, TNW>(bq[s & 1], a[s], acc);
          __builtin_amdgcn_sched_barrier(0);
        }
      };
      // The same, with the weight loads of the chunk at nbase issued step by step
      // between the MFMAs (k-step s of the next chunk during k-step s of this one):
      // loads issued in one burst per chunk stall the waves behind the vector-memory
      // queue, spread out they overlap the MFMAs. The loads are unconditional (also
      // in the skipped padding steps), so every fragment of the next chunk arrives.
      auto run_chunk_il = [&](int c, const AFrag<KIND>(&a)[PF][RP], const char *nbase, Lane nlane,
                              AFrag<KIND>(&na)[PF][RP]) {
        const int g0 = chunk_g0(c), gn = chunk_g0(c + 1 < nchunks ? c + 1 : 0);
#pragma unroll
        for (int s = 0; s < PF; ++s) {
          load_step(nbase, nlane, s, na[s]);
          __builtin_amdgcn_sched_barrier(0);
          if (c * PF + s < ksteps) {
            if (s + 1 < PF) read_b(g0, s + 1, bq[(s + 1) & 1]);
            else read_b(gn, 0, bq[0]);
            __builtin_amdgcn_sched_barrier(0);
            mma_k<KIND, TNW>(bq[s & 1], a[s], acc);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      };
      read_b(chunk_g0(0), 0, bq[0]);
      int c = 0;
      for (; c + 1 < nchunks; c += 2) {
        if constexpr (KIND == K_I8 || KIND == K_BF16) {  // measured faster interleaved
          run_chunk_il(c, fa[0], apk + (c + 1) * CHUNK, alane, fa[1]);
          run_chunk_il(c + 1, fa[1], c + 2 < nchunks ? apk + (c + 2) * CHUNK : after,
                       c + 2 < nchunks ? alane : after_lane, fa[0]);
        } else {  // FP32 (16 MFMAs per fragment): one burst per chunk measured faster
          load_chunk(apk + (c + 1) * CHUNK, alane, fa[1]);
          __builtin_amdgcn_sched_barrier(0);
          run_chunk(c, fa[0]);
          load_chunk(c + 2 < nchunks ? apk + (c + 2) * CHUNK : after, c + 2 < nchunks ? alane : after_lane, fa[0]);
          __builtin_amdgcn_sched_barrier(0);
          run_chunk(c + 1, fa[1]);
        }
      }
      if (c < nchunks) {  // odd chunk count: the next loads wait for this chunk
        run_chunk(c, fa[0]);
        __builtin_amdgcn_sched_barrier(0);
        load_chunk(after, after_lane, fa[0]);
      }
      apk = after;
      alane = after_lane;
      if (final_pass && !last && ((P.hold >> gi) & 1)) held_rb0 = rb0;
      else store_rows_impl<KIND, TN, TNW, EPX>(P, gi, rb0, n0, arena, acc, csum_next, cb0);
    }
    if ((long)wave * RP * 32 >= Mp && next_gemm) {  // no rows here: fetch the next GEMM's first chunk
      apk = next_gemm;
      alane = next_lane;
      load_chunk(apk, alane, fa[0]);
    }
    if (last) break;  // C is written; nothing in LDS is read again
    if (!((P.hold >> gi) & 1)) {
      // Y_gi does not meet X_gi: its padding rows are cleared now and one barrier
      // both retires X_gi and publishes Y_gi
      zero_pad(gi);
      __syncthreads();
      continue;
    }
    __syncthreads();  // every wave is done reading X_gi
    if (held_rb0 >= 0) store_rows_impl<KIND, TN, TNW, EPX>(P, gi, held_rb0, n0, arena, acc, csum_next, cb0);
    zero_pad(gi);
    __syncthreads();  // Y_gi complete
  }
}

// Workgroup b -> tile. Workgroups are dealt to the 8 XCDs round-robin (b % 8); this
// gives each XCD a contiguous range of tiles, so neighbouring tiles that share cache
// lines of B (64-column INT8 tiles use half a line per row) meet in one L2. A
// bijection on [0, G) for any G.
__device__ __forceinline__ int xcd_tile(int b, int G) {
  const int q = G / 8, r = G % 8, x = b % 8, i = b / 8;
  return x * q + (x < r ? x : r) + i;
}

template <int KIND, int TNB, bool IMG, bool EPX, int G, bool RAWA = false>
__global__ __launch_bounds__(NTH) void chain_kernel(ChainParams P) {
  extern __shared__ uint4 arena[];
  {
    // Touch every 64-byte line of the arguments with independent scalar loads: one
    // round trip, instead of one serialized miss per line where a field is first used
    // (about 0.5 us of the kernel's start).
    const unsigned *kp = reinterpret_cast<const unsigned *>(&P);
    unsigned kw = 0;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(ChainParams) / 4); i += 16) kw ^= kp[i];
    asm volatile("" ::"s"(kw));
  }
  chain_tile<KIND, TNB, IMG, EPX, G, RAWA>(P, (long)xcd_tile(blockIdx.x, gridDim.x) * (TNB * 32), arena);
}

constexpr int RP_of(int kind) {
  return kind == K_F32 ? CT<K_F32>::RP : kind == K_BF16 ? CT<K_BF16>::RP : kind == K_I8 ? CT<K_I8>::RP
                                                                                        : CT<K_F32X3>::RP;
}
constexpr int KG_of(int kind) {
  return kind == K_F32 ? CT<K_F32>::KG : kind == K_BF16 ? CT<K_BF16>::KG : kind == K_I8 ? CT<K_I8>::KG
                                                                                        : CT<K_F32X3>::KG;
}
// rows of one chunk (PF k-steps): the granule of X regions and of the weights' k-padding
constexpr int CS_of(int kind) {
  return kind == K_F32 ? CT<K_F32>::PF * CT<K_F32>::KC
         : kind == K_BF16 ? CT<K_BF16>::PF * CT<K_BF16>::KC
         : kind == K_I8   ? CT<K_I8>::PF * CT<K_I8>::KC
                          : CT<K_F32X3>::PF * CT<K_F32X3>::KC;
}

// LDS bytes: S rows (both planes for F32X3) and the three INT8 column-sum buffers.
size_t arena_bytes(int kind, int S, int tile_n) {
  const size_t groups = (size_t)S / KG_of(kind);
  return groups * tile_n * 16 * (kind == K_F32X3 ? 2 : 1) + (kind == K_I8 ? 3 * tile_n * sizeof(int) : 0);
}

// Column groups of a fused tile (chain_tile's G): 256-column tiles, and INT8/BF16
// tiles of 128 columns, give each wave one 32-column block and all rows; narrower
// tiles give each wave all columns of a row band.
constexpr int wave_groups(int kind, int tile_n) {
  return tile_n == 256 ? 8 : ((tile_n == 128 && (kind == K_I8 || kind == K_BF16)) ? 4 : 1);
}

// Circular arena of S rows: X_0 (the B tile) at row 0, Y_i right after X_i (mod S)
// and X_{i+1} = Y_i; every region spans its k-extent rounded up to a chunk and starts
// on a chunk boundary (S is a multiple of a chunk). Per GEMM: when Kc_i + Mc_i rows
// fit the LDS at this tile width, Y_i never meets X_i and every pass is stored as soon
// as it is done; otherwise GEMM i is held (bit i of *hold): all rows of Y_i but each
// wave's final pass are written while X_i is still read, so they must not meet it,
// Kc_i + (passes_i - 1) * PASS <= S, and the final passes are written over X_i after a
// barrier. Only the GEMMs that need it are held (the BF16 3-chain at 64 columns holds
// its second GEMM, 512 + 1024 rows, not its first, 512 + 512).
int plan_arena(int kind, const long *M, const long *K, int n, int tile_n, int *xbase, int *ybase, int *hold,
               size_t budget = LDS_LIMIT) {
  const long cs = CS_of(kind);
  auto rc = [cs](long v) { return (v + cs - 1) / cs * cs; };
  auto r32 = [](long v) { return (v + 31) / 32 * 32; };
  const long pass = (long)(NW / wave_groups(kind, tile_n)) * RP_of(kind) * 32;  // rows per pass
  auto fits = [&](long rows) { return tile_n > 0 && arena_bytes(kind, (int)rc(rows), tile_n) <= budget; };
  long S = rc(K[0]);
  *hold = 0;
  for (int i = 0; i < n; ++i) {
    const long kc = rc(K[i]);
    S = std::max(S, kc);
    if (i < n - 1) {
      const long wide = kc + rc(M[i]);
      if (fits(wide)) {
        S = std::max(S, wide);
      } else {
        const long passes = (r32(M[i]) + pass - 1) / pass;
        S = std::max(S, std::max(rc(M[i]), kc + (passes - 1) * pass));
        *hold |= 1 << i;
      }
    }
  }
  // A pass (PASS rows) can be shorter than a chunk (waves split over column groups),
  // so the held bound Kc_i + (passes_i - 1) * PASS need not be a whole number of
  // chunks; S must be, or a region starting on a chunk boundary could run past S
  // inside a chunk (its reads are one base plus immediate offsets, never wrapped).
  S = rc(S);
  long x = 0;
  for (int i = 0; i < n; ++i) {
    xbase[i] = (int)x;
    ybase[i] = (int)((x + rc(K[i])) % S);
    x = ybase[i];
  }
  return (int)S;
}


template <int KIND, int TNB, int G>
hipError_t launch_chain_t(const ChainArgs &c, hipStream_t s) {
  ChainParams P{};
  make_chain_params(c, &P);
  if (!c.img && !chain_fused_aligned(P)) return hipErrorInvalidValue;
  // reference-layout intermediates: INT8 / BF16 matrix chains with the plain epilogues
  if (c.refl && (c.img || c.epi || (KIND != K_I8 && KIND != K_BF16))) return hipErrorInvalidValue;
  // a segment's requantized output takes the plain epilogues
  if (c.out_requant && c.epi) return hipErrorInvalidValue;
  constexpr int TN = TNB * 32;
  const size_t lds = c.epi ? (size_t)P.epi_off + chain_epi_lds_bytes(c.M, c.num_gemm) : arena_bytes(KIND, P.S, TN);
  if (lds > LDS_LIMIT) return hipErrorInvalidValue;
  static_assert(wave_groups(KIND, TNB * 32) == G, "tile width and column groups disagree");
  auto kfn = c.img ? (c.epi ? chain_kernel<KIND, TNB, true, true, G> : chain_kernel<KIND, TNB, true, false, G>)
                   : (c.epi ? chain_kernel<KIND, TNB, false, true, G> : chain_kernel<KIND, TNB, false, false, G>);
  if (c.raw_a) {  // row-major FP32 weights (chain_raw_weights_ok), matrix B, plain epilogues
    if constexpr (KIND == K_F32) {
      if (c.img || c.epi) return hipErrorInvalidValue;
      kfn = chain_kernel<K_F32, TNB, false, false, G, true>;
    } else {
      return hipErrorInvalidValue;
    }
  }
  hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid((unsigned)((c.N + TN - 1) / TN));
  hipLaunchKernelGGL(kfn, grid, dim3(NTH), lds, s, P);
  return hipGetLastError();
}

// Tile widths of the grouped kernel, one per kind.
constexpr int GT_F32 = 1, GT_BF16 = 1, GT_I8 = 1, GT_F32X3 = 1;

// Persistent grouped launch over a packed table (GroupTable header, the request
// descriptors, the tile list): workgroup b runs tiles b, b + grid, ...; each tile is
// a whole chain on TN columns of its request. Counts and offsets are read from the
// table, so one captured graph replays batches of any size. WIDE adds the 64-column
// tiles (gtnb = 2) the autotune table can choose; it is a separate instantiation
// because its registers would cost the 32-column path occupancy and spills.
template <bool WIDE>
__global__ __launch_bounds__(NTH) void grouped_chain_kernel(const char *__restrict__ table) {
  extern __shared__ uint4 arena[];
  const GroupTable *hdr = (const GroupTable *)table;
  const int n = hdr->num_items;
  const ChainParams *reqs = (const ChainParams *)(table + hdr->reqs_off);
  const int2 *items = (const int2 *)(table + hdr->items_off);
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = items[it];
    const ChainParams &P = reqs[w.x];
    const long n0 = (long)w.y * P.gtnb * 32;
    if (WIDE && P.gtnb == 2) {
      switch (P.kind) {
        case K_F32: chain_tile<K_F32, 2, false, false, 1, false, false>(P, n0, arena); break;
        case K_BF16: chain_tile<K_BF16, 2, false, false, 1, false, false>(P, n0, arena); break;
        case K_I8: chain_tile<K_I8, 2, false, false, 1, false, false>(P, n0, arena); break;
        case K_F32X3: chain_tile<K_F32X3, 2, false, false, 1, false, false>(P, n0, arena); break;
      }
    } else {
      switch (P.kind) {
        case K_F32: chain_tile<K_F32, GT_F32, false, false, 1, false, false>(P, n0, arena); break;
        case K_BF16: chain_tile<K_BF16, GT_BF16, false, false, 1, false, false>(P, n0, arena); break;
        case K_I8: chain_tile<K_I8, GT_I8, false, false, 1, false, false>(P, n0, arena); break;
        case K_F32X3: chain_tile<K_F32X3, GT_F32X3, false, false, 1, false, false>(P, n0, arena); break;
      }
    }
    __syncthreads();  // the next tile restages the arena
  }
}

// Tile widths per kind: 32, 64, 96 and 128 columns with every wave on all columns
// (BF16 and INT8 stop at 64: two row blocks per wave leave no registers for more),
// 128 (INT8, BF16) and 256 with the waves split over the columns (wave_groups).
constexpr bool tile_ok_of(int kind, int tile_n) {
  return tile_n == 32 || tile_n == 64 || tile_n == 128 || tile_n == 256 ||
         (tile_n == 96 && kind != K_BF16 && kind != K_I8);
}

template <int KIND>
hipError_t launch_chain_k(const ChainArgs &c, hipStream_t s) {
  constexpr bool two_rows = KIND == K_BF16 || KIND == K_I8;
  switch (c.tile_n) {
    case 32: return launch_chain_t<KIND, 1, 1>(c, s);
    case 64: return launch_chain_t<KIND, 2, 1>(c, s);
    case 128: return launch_chain_t<KIND, 4, two_rows ? 4 : 1>(c, s);
    case 256: return launch_chain_t<KIND, 8, 8>(c, s);
  }
  if constexpr (!two_rows) {
    if (c.tile_n == 96) return launch_chain_t<KIND, 3, 1>(c, s);
  }
  return hipErrorInvalidValue;
}

template <int KIND>
hipError_t launch_pack_t(PackJob J, hipStream_t s) {
  J.fstart[0] = 0;
  for (int m = 0; m < J.n; ++m) J.fstart[m + 1] = J.fstart[m] + packed_frags<KIND>(J.M[m], J.K[m]);
  const long nf = J.fstart[J.n];
  if (nf == 0) return hipSuccess;
  hipLaunchKernelGGL((pack_a_kernel<KIND>), dim3((unsigned)((nf + 3) / 4)), dim3(256), 0, s, J);
  return hipGetLastError();
}

}  // namespace

bool chain_raw_weights_ok(const ChainArgs &c) {
  if (c.kind != K_F32 || c.img || c.epi) return false;
  for (int i = 0; i < c.num_gemm; ++i)
    if (c.M[i] % 32 != 0 || c.K[i] % CS_of(K_F32) != 0 || ((uintptr_t)c.A[i] & 15) != 0) return false;
  return true;
}

size_t chain_epi_lds_bytes(const long *M, int num_gemm) {
  long rows = 0;
  for (int i = 0; i < num_gemm; ++i) rows += M[i];
  return (size_t)rows * 8;
}

size_t chain_fused_lds_bytes(int kind, const long *M, const long *K, int num_gemm, int tile_n, bool epi) {
  int xb[MAXG], yb[MAXG];
  if (num_gemm < 1 || num_gemm > MAXG) return (size_t)-1;
  int hold;
  const size_t extra = epi ? chain_epi_lds_bytes(M, num_gemm) : 0;
  if (extra >= LDS_LIMIT) return (size_t)-1;
  const int S = plan_arena(kind, M, K, num_gemm, tile_n, xb, yb, &hold, LDS_LIMIT - extra);
  return (arena_bytes(kind, S, tile_n) + 15) / 16 * 16 + extra;
}

int chain_fused_max_gemms() { return MAXG; }

bool chain_fused_tile_ok(int kind, int tile_n) { return tile_ok_of(kind, tile_n); }

int chain_fused_rows_per_wave(int kind) { return RP_of(kind) * 32; }

int chain_fused_wave_groups(int kind, int tile_n) { return wave_groups(kind, tile_n); }

size_t packed_a_bytes(int kind, long M, long K) {
  switch (kind) {
    case K_F32: return (size_t)packed_frags<K_F32>(M, K) * 64 * CT<K_F32>::FS;
    case K_BF16: return (size_t)packed_frags<K_BF16>(M, K) * 64 * CT<K_BF16>::FS;
    case K_I8: return (size_t)packed_frags<K_I8>(M, K) * 64 * CT<K_I8>::FS;
    case K_F32X3: return (size_t)packed_frags<K_F32X3>(M, K) * 64 * CT<K_F32X3>::FS;
  }
  return 0;
}

hipError_t launch_pack_a(int kind, const PackJob &J, hipStream_t s) {
  if (J.n < 1 || J.n > PACK_MAX) return hipErrorInvalidValue;
  for (int m = 0; m < J.n; ++m)  // positive shapes; fragment indices fit 32 bits
    if (J.M[m] <= 0 || J.K[m] <= 0 || packed_a_bytes(kind, J.M[m], J.K[m]) / 1024 >= (1ull << 31))
      return hipErrorInvalidValue;
  switch (kind) {
    case K_F32: return launch_pack_t<K_F32>(J, s);
    case K_BF16: return launch_pack_t<K_BF16>(J, s);
    case K_I8: return launch_pack_t<K_I8>(J, s);
    case K_F32X3: return launch_pack_t<K_F32X3>(J, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_chain_fused(const ChainArgs &c, hipStream_t s) {
  if (c.num_gemm < 1 || c.num_gemm > MAXG || c.N <= 0) return hipErrorInvalidValue;
  if (chain_fused_lds_bytes(c.kind, c.M, c.K, c.num_gemm, c.tile_n, c.epi != nullptr) > LDS_LIMIT)
    return hipErrorInvalidValue;
  switch (c.kind) {
    case K_F32: return launch_chain_k<K_F32>(c, s);
    case K_BF16: return launch_chain_k<K_BF16>(c, s);
    case K_I8: return launch_chain_k<K_I8>(c, s);
    case K_F32X3: return launch_chain_k<K_F32X3>(c, s);
  }
  return hipErrorInvalidValue;
}

void make_chain_params(const ChainArgs &c, ChainParams *P) {
  static const int esz[4] = {4, 2, 1, 4};
  const int e = esz[c.kind];
  *P = ChainParams{};
  P->kind = c.kind;
  P->num_gemm = c.num_gemm;
  for (int i = 0; i < c.num_gemm; ++i) {
    P->A[i] = c.A[i];
    P->M[i] = c.M[i];
    P->K[i] = c.K[i];
  }
  const size_t extra = c.epi ? chain_epi_lds_bytes(c.M, c.num_gemm) : 0;
  P->S = plan_arena(c.kind, c.M, c.K, c.num_gemm, c.tile_n, P->xbase, P->ybase, &P->hold,
                    extra < LDS_LIMIT ? LDS_LIMIT - extra : 0);
  P->epi_off = (int)((arena_bytes(c.kind, P->S, c.tile_n) + 15) / 16 * 16);
  P->B = c.B;
  P->C = c.C;
  P->N = c.N;
  P->vecB = (c.N % 4 == 0) && (((uintptr_t)c.B & (4 * e - 1)) == 0);
  if (c.img) P->img = *c.img;
  if (c.epi) P->epi = *c.epi;
  P->refl = c.refl;
  P->out_requant = c.out_requant;
}

bool chain_refl_ok(int kind, const RefPanels &p, int num_gemm, const long *M) {
  if (kind != K_I8 && kind != K_BF16) return false;
  if (p.nb % 32 != 0 || p.ne % 32 != 0) return false;
  for (int i = 0; i + 1 < num_gemm; ++i)
    if (M[i] % 32 != 0) return false;
  return true;
}

bool chain_fused_shape_ok(int kind, long N, int num_gemm, const long *K) {
  (void)kind;
  (void)K;
  return num_gemm >= 1 && num_gemm <= MAXG && N % 4 == 0;
}

bool chain_fused_aligned(const ChainParams &P) { return P.vecB != 0; }

int grouped_tile_n(int kind) {
  switch (kind) {
    case K_F32: return GT_F32 * 32;
    case K_BF16: return GT_BF16 * 32;
    case K_I8: return GT_I8 * 32;
    case K_F32X3: return GT_F32X3 * 32;
  }
  return 0;
}

hipError_t launch_grouped(const void *table, int grid, size_t lds, bool wide, hipStream_t s) {
  if (lds > LDS_LIMIT || grid <= 0) return hipErrorInvalidValue;
  auto kfn = wide ? grouped_chain_kernel<true> : grouped_chain_kernel<false>;
  hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kfn, dim3(grid), dim3(NTH), lds, s, (const char *)table);
  return hipGetLastError();
}

}  // namespace fg
