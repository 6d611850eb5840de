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
//     layout (16-byte k-groups per column, slot g*TN + (n ^ (g & 7)));
//   * GEMM i reads X_i from one end of the arena and writes its requantized output
//     Y_i (fp32 / bf16 RNE / int8 saturate) at the other end, so X_{i+1} = Y_i never
//     leaves the CU; the arena holds max_i (K_i + M_i) rows;
//   * each wave owns row blocks of the weights, streams their packed fragments PF
//     k-steps ahead (double buffered) and reads B fragments one step ahead;
//   * INT8 uses u8(a).b = s8(a ^ 0x80).b + 128 sum_k b: the column sums of each X_i
//     are taken from LDS and seed the accumulators;
//   * only the last GEMM writes HBM (C, fp32 or int32, row-major).
#include "fg_internal.h"

namespace fg {
namespace {

constexpr int NW = 8;  // waves per workgroup
constexpr int NTH = NW * 64;
constexpr int MAXG = CHAIN_MAXG;
constexpr size_t LDS_LIMIT = 160 * 1024;

template <int KIND> struct CT;
// KG: k per 16-byte LDS group (per plane for F32X3); KC: k per MFMA step;
// KA: A elements per lane-half per step; ESZ: global element bytes; LB: LDS bytes
// per stored element (both planes for F32X3); RP: row blocks per wave item;
// PF: k-steps per prefetch chunk; FS: packed bytes per lane per fragment.
template <> struct CT<K_F32> {
  static constexpr int KG = 4, KC = 8, KA = 4, ESZ = 4, LB = 4, RP = 1, PF = 4, FS = 16;
  using Acc = f32x16;
};
template <> struct CT<K_BF16> {
  static constexpr int KG = 8, KC = 16, KA = 8, ESZ = 2, LB = 2, RP = 2, PF = 4, FS = 16;
  using Acc = f32x16;
};
template <> struct CT<K_I8> {
  static constexpr int KG = 16, KC = 32, KA = 16, ESZ = 1, LB = 1, RP = 2, PF = 4, FS = 16;
  using Acc = i32x16;
};
template <> struct CT<K_F32X3> {
  static constexpr int KG = 8, KC = 16, KA = 8, ESZ = 4, LB = 4, RP = 1, PF = 2, FS = 32;
  using Acc = f32x16;
};

template <int TN>
__device__ __forceinline__ int cslot(int g, int n) { return g * TN + (n ^ (g & 7)); }

// --------------------------------------------------------- packed weights --
// Packed layout of A (M x K) for kind KIND: fragments f = (rg * kspad + ks) * RP + i
// for row group rg (RP * 32 rows), k-step ks < kspad and row block i; fragment f is
// 64 lanes x FS bytes at f * 64 * FS, lane l = h*32 + r holding row
// (rg*RP + i)*32 + r, elements k = ks*KC + h*KA .. + KA-1 (zero outside A). For
// F32X3 the fragment is the 1 KB hi half then the 1 KB lo half.
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
  const int h = lane >> 5, r = lane & 31;
  const long kspad = kspad_of<KIND>(K);
  const long i = f % T::RP, ks = (f / T::RP) % kspad, rg = f / (T::RP * kspad);
  const long row = (rg * T::RP + i) * 32 + r;
  const long k0 = ks * T::KC + h * T::KA;
  uint32_t w[T::KA * T::ESZ / 4];
#pragma unroll
  for (int d = 0; d < T::KA * T::ESZ / 4; ++d) w[d] = 0u;
  if (row < M) {
    const char *p = A + ((size_t)row * K + k0) * T::ESZ;
#pragma unroll
    for (int e = 0; e < T::KA; ++e)
      if (k0 + e < K) {
        uint32_t x;
        if (T::ESZ == 4) x = *(const uint32_t *)(p + 4 * e);
        else if (T::ESZ == 2) x = *(const uint16_t *)(p + 2 * e);
        else x = (uint32_t)(*(const uint8_t *)(p + e) ^ 0x80u);  // u8 -> s8 bias
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

// One lane's packed A fragment: 16 bytes, or hi and lo halves for F32X3.
template <int KIND> struct AFrag { uint4 v[KIND == K_F32X3 ? 2 : 1]; };

// One lane's B fragments for one k-step: TNB column blocks (hi and lo planes for
// F32X3), read from the arena one step ahead of their MFMAs.
template <int KIND, int TNB> struct BFrag {
  uint4 b[TNB];
  uint4 lo[KIND == K_F32X3 ? TNB : 1];
};

template <int KIND, int TNB>
__device__ __forceinline__ void read_bfrag(const uint4 *arena, int lo_off, int g, int l32, BFrag<KIND, TNB> &f) {
  constexpr int TN = TNB * 32;
#pragma unroll
  for (int j = 0; j < TNB; ++j) {
    f.b[j] = arena[cslot<TN>(g, j * 32 + l32)];
    if (KIND == K_F32X3) f.lo[j] = arena[lo_off + cslot<TN>(g, j * 32 + l32)];
  }
}

template <int KIND, int TNB>
__device__ __forceinline__ void mma_k(const BFrag<KIND, TNB> &bf, const AFrag<KIND> (&af)[CT<KIND>::RP],
                                      typename CT<KIND>::Acc (&acc)[CT<KIND>::RP][TNB]) {
  constexpr int RP = CT<KIND>::RP;
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
      } else if (KIND == K_F32) {
        // lane half h holds k = kc + 4h + s for s = 0..3 in both operands
        const f32x4 av = __builtin_bit_cast(f32x4, af[i].v[0]);
        const f32x4 bv = __builtin_bit_cast(f32x4, bf.b[j]);
#pragma unroll
        for (int s = 0; s < 4; ++s)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[s], bv[s], acc[i][j], 0, 0, 0);
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
// read a clamped address and are zeroed with a select.
template <int KIND, int TN>
__device__ void stage_b(uint4 *arena, int lo_off, const char *B, long K1, long K1p, long N, long n0) {
  using T = CT<KIND>;
  // one task = KG rows x 4 columns of one k-group
  const int ngroups = (int)(K1p / T::KG);
  const int tasks = ngroups * (TN / 4);
  for (int task = threadIdx.x; task < tasks; task += NTH) {
    int g = task / (TN / 4), cb = task % (TN / 4);
    const long nb = n0 + cb * 4;
    const bool nok = nb < N;
    // v[r][c] raw element bits
    uint32_t v[T::KG][4];
#pragma unroll
    for (int r = 0; r < T::KG; ++r) {
      const long k = (long)g * T::KG + r;
      const bool ok = nok && k < K1;
      const char *p = B + ((size_t)(ok ? k : 0) * N + (ok ? nb : 0)) * T::ESZ;
      if (T::ESZ == 4) {
        uint4 x = *(const uint4 *)p;
        v[r][0] = x.x; v[r][1] = x.y; v[r][2] = x.z; v[r][3] = x.w;
      } else if (T::ESZ == 2) {
        uint2 x = *(const uint2 *)p;
        v[r][0] = x.x & 0xffffu; v[r][1] = x.x >> 16; v[r][2] = x.y & 0xffffu; v[r][3] = x.y >> 16;
      } else {
        uint32_t x = *(const uint32_t *)p;
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] = (x >> (8 * c)) & 0xffu;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) v[r][c] = ok ? v[r][c] : 0u;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int n = cb * 4 + c;
      int sl = cslot<TN>(g, n);
      if (KIND == K_F32) {
        arena[sl] = make_uint4(v[0][c], v[1][c], v[2][c], v[3][c]);
      } else if (KIND == K_BF16) {
        arena[sl] = make_uint4(v[0][c] | (v[1][c] << 16), v[2][c] | (v[3][c] << 16), v[4][c] | (v[5][c] << 16),
                               v[6][c] | (v[7][c] << 16));
      } else if (KIND == K_I8) {
        uint32_t w[4];
#pragma unroll
        for (int d = 0; d < 4; ++d)
          w[d] = v[4 * d][c] | (v[4 * d + 1][c] << 8) | (v[4 * d + 2][c] << 16) | (v[4 * d + 3][c] << 24);
        arena[sl] = make_uint4(w[0], w[1], w[2], w[3]);
      } else {
        uint32_t h[4], l[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          split_pk(__builtin_bit_cast(float, v[2 * r][c]), __builtin_bit_cast(float, v[2 * r + 1][c]), h[r], l[r]);
        arena[sl] = make_uint4(h[0], h[1], h[2], h[3]);
        arena[lo_off + sl] = make_uint4(l[0], l[1], l[2], l[3]);
      }
    }
  }
}

// INT8: colsum[n] = sum over rows [xbase, xbase + Kp) of s8 X[k][n].
template <int TN>
__device__ void colsum_lds(const uint4 *arena, int g0, int ngroups, int *colsum) {
  // thread -> (column, slice of groups); reduce slices with LDS atomics (integer adds
  // are order independent, so the result is deterministic).
  for (int n = threadIdx.x; n < TN; n += NTH) colsum[n] = 0;
  __syncthreads();
  constexpr int SL = NTH / TN;  // slices per column
  int n = threadIdx.x % TN, sl = threadIdx.x / TN;
  if (sl < SL) {
    int s = 0;
    for (int g = sl; g < ngroups; g += SL) {
      uint4 x = arena[cslot<TN>(g0 + g, n)];
      s = __builtin_amdgcn_sdot4((int)x.x, 0x01010101, s, false);
      s = __builtin_amdgcn_sdot4((int)x.y, 0x01010101, s, false);
      s = __builtin_amdgcn_sdot4((int)x.z, 0x01010101, s, false);
      s = __builtin_amdgcn_sdot4((int)x.w, 0x01010101, s, false);
    }
    atomicAdd(&colsum[n], s);
  }
}

// One workgroup runs the whole chain on columns [n0, n0 + TN) of B. P.A[i] are
// packed weights (pack_a_kernel layout for KIND).
template <int KIND, int TNB>
__device__ void chain_tile(const ChainParams &P, const long n0, uint4 *arena) {
  using T = CT<KIND>;
  using Acc = typename T::Acc;
  constexpr int TN = TNB * 32;
  constexpr int RP = T::RP;
  constexpr int PF = T::PF;
  constexpr int FRAG = 64 * T::FS;  // bytes per packed fragment
  const int lo_off = (KIND == K_F32X3) ? (P.S / 8) * TN : 0;  // lo plane (F32X3)
  int *colsum = (int *)(arena + (size_t)(P.S / T::KG) * TN * (KIND == K_F32X3 ? 2 : 1));

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, l32 = lane & 31;
  const long tile_w = (P.N - n0 < TN) ? P.N - n0 : TN;

  const long K1p = (P.K[0] + 31) / 32 * 32;
  stage_b<KIND, TN>(arena, lo_off, (const char *)P.B, P.K[0], K1p, P.N, n0);
  __syncthreads();

  for (int gi = 0; gi < P.num_gemm; ++gi) {
    const long M = P.M[gi], K = P.K[gi];
    const long Mp = (M + 31) / 32 * 32;
    const int xg0 = P.xbase[gi] / T::KG;
    const bool last = gi == P.num_gemm - 1;
    if (KIND == K_I8) {
      colsum_lds<TN>(arena, xg0, (int)((K + 31) / 32 * 32 / T::KG), colsum);
      __syncthreads();
    }
    // k-steps padded to whole chunks (the packed A is zero there); B reads past the
    // X region clamp to its last group, which then meets zero A.
    const int kspad = (int)kspad_of<KIND>(K);
    const int nchunks = kspad / PF;
    const int glast = xg0 + (int)((K + 31) / 32 * 32 / T::KG) - 1;
    // INT8: the accumulators start at the bias 128 * sum_k X[k][n] of their column,
    // so the MFMAs accumulate u8(a).b directly.
    int bias0[TNB];
#pragma unroll
    for (int j = 0; j < TNB; ++j) bias0[j] = (KIND == K_I8) ? 128 * colsum[j * 32 + l32] : 0;
    for (long rb0 = (long)wave * RP * 32; rb0 < Mp; rb0 += (long)NW * RP * 32) {
      Acc acc[RP][TNB];
#pragma unroll
      for (int i = 0; i < RP; ++i)
#pragma unroll
        for (int j = 0; j < TNB; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) acc[i][j][r] = bias0[j];

      // this wave's row group in the packed weights; fragment (ks, i) at
      // apk + (ks * RP + i) * FRAG, this lane's 16 bytes at + lane * 16
      const char *apk = (const char *)P.A[gi] + ((size_t)(rb0 / (32 * RP)) * kspad * RP) * FRAG + lane * 16;
      AFrag<KIND> fa[2][PF][RP];
      auto load_chunk = [&](int c, AFrag<KIND>(&dst)[PF][RP]) {
#pragma unroll
        for (int s = 0; s < PF; ++s)
#pragma unroll
          for (int i = 0; i < RP; ++i) {
            const char *p = apk + ((size_t)(c * PF + s) * RP + i) * FRAG;
            dst[s][i].v[0] = *(const uint4 *)p;
            if (KIND == K_F32X3) dst[s][i].v[1] = *(const uint4 *)(p + 1024);
          }
      };
      // B fragments one step ahead, double buffered (PF is even, so the buffer of
      // step ks is ks & 1 across chunks).
      BFrag<KIND, TNB> bq[2];
      auto run_chunk = [&](int c, const AFrag<KIND>(&src)[PF][RP]) {
#pragma unroll
        for (int s = 0; s < PF; ++s) {
          const int ks = c * PF + s;
          const int gn = xg0 + 2 * (ks + 1) + h;
          read_bfrag<KIND, TNB>(arena, lo_off, gn < glast ? gn : glast, l32, bq[(s + 1) & 1]);
          mma_k<KIND, TNB>(bq[s & 1], src[s], acc);
        }
      };
      read_bfrag<KIND, TNB>(arena, lo_off, (xg0 + h) < glast ? xg0 + h : glast, l32, bq[0]);
      load_chunk(0, fa[0]);
      for (int c = 0; c < nchunks; c += 2) {
        if (c + 1 < nchunks) load_chunk(c + 1, fa[1]);
        run_chunk(c, fa[0]);
        if (c + 1 < nchunks) {
          if (c + 2 < nchunks) load_chunk(c + 2, fa[0]);
          run_chunk(c + 1, fa[1]);
        }
      }

      // epilogue: lane holds rows 8(r/4) + 4h + (r%4) of column j*32 + l32
#pragma unroll
      for (int j = 0; j < TNB; ++j) {
        const int n = j * 32 + l32;
#pragma unroll
        for (int i = 0; i < RP; ++i) {
          const long rbase = rb0 + i * 32;
          if (rbase >= Mp) continue;
          if (last) {
            if (n >= tile_w) continue;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
              long m = rbase + 8 * (r >> 2) + 4 * h + (r & 3);
              if (m >= M) continue;
              size_t off = (size_t)m * P.N + n0 + n;
              if (KIND == K_I8) ((int *)P.C)[off] = (int)acc[i][j][r];
              else ((float *)P.C)[off] = (float)acc[i][j][r];
            }
          } else {
            // rows rbase + 8q + 4h + (0..3), q = 0..3: four consecutive rows per q
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const long row = P.ybase[gi] + rbase + 8 * q + 4 * h;  // arena row, multiple of 4
              const int g = (int)(row / T::KG), pos = (int)(row % T::KG);
              uint4 *dst = arena + cslot<TN>(g, n);
              if (KIND == K_F32) {
                *dst = make_uint4(__builtin_bit_cast(uint32_t, (float)acc[i][j][4 * q]),
                                  __builtin_bit_cast(uint32_t, (float)acc[i][j][4 * q + 1]),
                                  __builtin_bit_cast(uint32_t, (float)acc[i][j][4 * q + 2]),
                                  __builtin_bit_cast(uint32_t, (float)acc[i][j][4 * q + 3]));
              } else if (KIND == K_BF16) {
                *((uint2 *)dst + (pos >> 2)) =
                    make_uint2(cvt_pk_bf16(acc[i][j][4 * q], acc[i][j][4 * q + 1]),
                               cvt_pk_bf16(acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]));
              } else if (KIND == K_I8) {
                // padded rows (m >= M) must stay zero: they enter the next column sum
                const long m0r = rbase + 8 * q + 4 * h;
                uint32_t w = sat_pk_i8((int)acc[i][j][4 * q], (int)acc[i][j][4 * q + 1], (int)acc[i][j][4 * q + 2],
                                       (int)acc[i][j][4 * q + 3]);
                if (m0r + 4 > M) w &= m0r >= M ? 0u : (0xffffffffu >> (8 * (m0r + 4 - M)));
                *((uint32_t *)dst + (pos >> 2)) = w;
              } else {
                uint32_t h01, l01, h23, l23;
                split_pk(acc[i][j][4 * q], acc[i][j][4 * q + 1], h01, l01);
                split_pk(acc[i][j][4 * q + 2], acc[i][j][4 * q + 3], h23, l23);
                *((uint2 *)dst + (pos >> 2)) = make_uint2(h01, h23);
                *((uint2 *)(arena + lo_off + cslot<TN>(g, n)) + (pos >> 2)) = make_uint2(l01, l23);
              }
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

template <int KIND, int TNB>
__global__ __launch_bounds__(NTH) void chain_kernel(ChainParams P) {
  extern __shared__ uint4 arena[];
  chain_tile<KIND, TNB>(P, (long)blockIdx.x * (TNB * 32), arena);
}

constexpr int LB_of(int kind) { return kind == K_F32 ? 4 : kind == K_BF16 ? 2 : kind == K_I8 ? 1 : 4; }

// Arena rows and region placement. X_0 (the B tile) sits at row 0; GEMM i writes
// Y_i at the end opposite to X_i.
int plan_arena(const long *M, const long *K, int n, int *xbase, int *ybase) {
  auto r32 = [](long v) { return (int)((v + 31) / 32 * 32); };
  int S = r32(K[0]);
  for (int i = 0; i < n; ++i) {
    int kx = (i == 0) ? r32(K[0]) : r32(M[i - 1]);
    if (kx > S) S = kx;
    if (i < n - 1 && kx + r32(M[i]) > S) S = kx + r32(M[i]);
  }
  bool at_start = true;  // X_0 at the start
  for (int i = 0; i < n; ++i) {
    int kx = (i == 0) ? r32(K[0]) : r32(M[i - 1]);
    xbase[i] = at_start ? 0 : S - kx;
    if (i < n - 1) ybase[i] = at_start ? S - r32(M[i]) : 0;
    else ybase[i] = 0;
    at_start = !at_start;
  }
  return S;
}

template <int KIND, int TNB>
hipError_t launch_chain_t(const ChainArgs &c, hipStream_t s) {
  using T = CT<KIND>;
  ChainParams P{};
  make_chain_params(c, &P);
  if (!chain_fused_aligned(P)) return hipErrorInvalidValue;
  constexpr int TN = TNB * 32;
  size_t lds = (size_t)P.S * TN * T::LB + (KIND == K_I8 ? TN * sizeof(int) : 0);
  if (lds > LDS_LIMIT) return hipErrorInvalidValue;
  auto kfn = chain_kernel<KIND, TNB>;
  hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid((unsigned)((c.N + TN - 1) / TN));
  hipLaunchKernelGGL(kfn, grid, dim3(NTH), lds, s, P);
  return hipGetLastError();
}

// Tile widths of the grouped kernel, one per kind.
constexpr int GT_F32 = 1, GT_BF16 = 1, GT_I8 = 2, GT_F32X3 = 1;

// Out-of-line tile bodies for the grouped kernel: each kind keeps its own register
// allocation instead of one allocation for the union of all four.
template <int KIND, int TNB>
__device__ __noinline__ void chain_tile_call(const ChainParams &P, long n0, uint4 *arena) {
  chain_tile<KIND, TNB>(P, n0, arena);
}

// Persistent grouped launch over a packed table (GroupTable header, the request
// descriptors, the tile list): workgroup b runs tiles b, b + grid, ...; each tile is
// a whole chain on TN columns of its request. Counts and offsets are read from the
// table, so one captured graph replays batches of any size.
__global__ __launch_bounds__(NTH) void grouped_chain_kernel(const char *__restrict__ table) {
  extern __shared__ uint4 arena[];
  const GroupTable *hdr = (const GroupTable *)table;
  const int n = hdr->num_items;
  const ChainParams *reqs = (const ChainParams *)(table + hdr->reqs_off);
  const int2 *items = (const int2 *)(table + hdr->items_off);
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = items[it];
    const ChainParams &P = reqs[w.x];
    switch (P.kind) {
      case K_F32: chain_tile_call<K_F32, GT_F32>(P, (long)w.y * GT_F32 * 32, arena); break;
      case K_BF16: chain_tile_call<K_BF16, GT_BF16>(P, (long)w.y * GT_BF16 * 32, arena); break;
      case K_I8: chain_tile_call<K_I8, GT_I8>(P, (long)w.y * GT_I8 * 32, arena); break;
      case K_F32X3: chain_tile_call<K_F32X3, GT_F32X3>(P, (long)w.y * GT_F32X3 * 32, arena); break;
    }
    __syncthreads();  // the next tile restages the arena
  }
}

template <int KIND>
hipError_t launch_chain_k(const ChainArgs &c, hipStream_t s) {
  switch (c.tile_n) {
    case 32: return launch_chain_t<KIND, 1>(c, s);
    case 64: return launch_chain_t<KIND, 2>(c, s);
    case 96: return launch_chain_t<KIND, 3>(c, s);
    case 128: return launch_chain_t<KIND, 4>(c, s);
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

size_t chain_fused_lds_bytes(int kind, const long *M, const long *K, int num_gemm, int tile_n) {
  int xb[MAXG], yb[MAXG];
  if (num_gemm < 1 || num_gemm > MAXG) return (size_t)-1;
  int S = plan_arena(M, K, num_gemm, xb, yb);
  return (size_t)S * tile_n * LB_of(kind) + (kind == K_I8 ? tile_n * sizeof(int) : 0);
}

int chain_fused_max_gemms() { return MAXG; }

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
  for (int m = 0; m < J.n; ++m)
    if (J.M[m] <= 0 || J.K[m] <= 0) return hipErrorInvalidValue;
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
  if (chain_fused_lds_bytes(c.kind, c.M, c.K, c.num_gemm, c.tile_n) > LDS_LIMIT) return hipErrorInvalidValue;
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
  P->S = plan_arena(c.M, c.K, c.num_gemm, P->xbase, P->ybase);
  P->B = c.B;
  P->C = c.C;
  P->N = c.N;
  P->vecB = (c.N % 4 == 0) && (((uintptr_t)c.B & (4 * e - 1)) == 0);
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

hipError_t launch_grouped(const void *table, int grid, size_t lds, hipStream_t s) {
  if (lds > LDS_LIMIT || grid <= 0) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void *)grouped_chain_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(grouped_chain_kernel, dim3(grid), dim3(NTH), lds, s, (const char *)table);
  return hipGetLastError();
}

}  // namespace fg
