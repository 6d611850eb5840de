// Fused GEMM chain for gfx950: C = A_n ... A_1 B with every intermediate kept in LDS.
//
// This is the MI355X counterpart of the reference's per-panel chain loop
// (multigemm/src/flashgemm.c:123-148), where each thread walks 32-column panels of
// B and runs every GEMM of the chain, the epilogue of GEMM i writing straight into
// GEMM i+1's packed B (flashgemm.c:132). Here a workgroup owns a tile of `TN`
// columns of B and runs every GEMM of the chain on it:
//   * the B tile (K1 x TN) is staged once into an LDS arena in the MFMA B-operand
//     layout (16-byte k-groups per column, slot g*TN + (n ^ (g & 7)));
//   * GEMM i reads X_i from one end of the arena and writes its requantized output
//     Y_i (fp32 / bf16 RNE / int8 saturate) at the other end, so X_{i+1} = Y_i never
//     leaves the CU; the arena holds max_i (K_i + M_i) rows;
//   * the weights A_i stream from L2 straight into registers (each row block is
//     read by exactly one wave), double buffered PF k-steps ahead;
//   * INT8 uses u8(a).b = s8(a ^ 0x80).b + 128 sum_k b, the column sums of each
//     X_i taken from LDS before its GEMM;
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
// KA: A elements one lane loads per step; ESZ: global element bytes;
// LB: LDS bytes per stored element (both planes for F32X3); RP: row blocks per item.
template <> struct CT<K_F32> {
  static constexpr int KG = 4, KC = 8, KA = 4, ESZ = 4, LB = 4, RP = 1, PF = 4;
  using Acc = f32x16;
};
template <> struct CT<K_BF16> {
  static constexpr int KG = 8, KC = 16, KA = 8, ESZ = 2, LB = 2, RP = 2, PF = 4;
  using Acc = f32x16;
};
template <> struct CT<K_I8> {
  static constexpr int KG = 16, KC = 32, KA = 16, ESZ = 1, LB = 1, RP = 2, PF = 4;
  using Acc = i32x16;
};
template <> struct CT<K_F32X3> {
  static constexpr int KG = 8, KC = 16, KA = 8, ESZ = 4, LB = 4, RP = 1, PF = 2;
  using Acc = f32x16;
};

template <int TN>
__device__ __forceinline__ int cslot(int g, int n) { return g * TN + (n ^ (g & 7)); }

// One lane's A fragment for one k-step: 16 bytes (32 for F32X3, as fp32).
template <int KIND> struct AFrag { uint4 v[KIND == K_F32X3 ? 2 : 1]; };

// One lane's A fragment for k-step `ks`: row `rowp` (a clamped, always valid row
// pointer) at element k = ks*KC + h*KA. The fused path requires K*ESZ % 16 == 0 and
// 16-byte aligned A (the planner sends other shapes to the unfused path), so a
// fragment lies entirely inside or entirely past K. Out-of-range lanes (row >= M,
// k >= K) read a clamped in-bounds address and are zeroed with a select: no
// divergent branches in the k loop.
template <int KIND>
__device__ __forceinline__ AFrag<KIND> load_afrag(const char *rowp, long K, long k, bool row_ok) {
  using T = CT<KIND>;
  constexpr int NV = KIND == K_F32X3 ? 2 : 1;
  const bool ok = row_ok && (k < K);
  const long kc = ok ? k : 0;
  const char *p = rowp + kc * T::ESZ;
  AFrag<KIND> f;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    uint4 v = *(const uint4 *)(p + 16 * q);
    if (KIND == K_I8) {  // u8 -> s8 bias: x ^ 0x80 = x - 128 as signed
      v.x ^= 0x80808080u; v.y ^= 0x80808080u; v.z ^= 0x80808080u; v.w ^= 0x80808080u;
    }
    f.v[q].x = ok ? v.x : 0u;
    f.v[q].y = ok ? v.y : 0u;
    f.v[q].z = ok ? v.z : 0u;
    f.v[q].w = ok ? v.w : 0u;
  }
  return f;
}

__device__ __forceinline__ uint32_t pk(uint16_t lo, uint16_t hi) { return (uint32_t)lo | ((uint32_t)hi << 16); }

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
  if (KIND == K_F32X3) {
    uint4 ah[RP], al[RP];
#pragma unroll
    for (int i = 0; i < RP; ++i) {
      const uint32_t w[8] = {af[i].v[0].x, af[i].v[0].y, af[i].v[0].z, af[i].v[0].w,
                             af[i].v[1].x, af[i].v[1].y, af[i].v[1].z, af[i].v[1].w};
      uint32_t h[4], l[4];
#pragma unroll
      for (int e = 0; e < 4; ++e)
        split_pk(__builtin_bit_cast(float, w[2 * e]), __builtin_bit_cast(float, w[2 * e + 1]), h[e], l[e]);
      ah[i] = make_uint4(h[0], h[1], h[2], h[3]);
      al[i] = make_uint4(l[0], l[1], l[2], l[3]);
    }
#pragma unroll
    for (int j = 0; j < TNB; ++j) {
      const uint4 bh = bf.b[j], bl = bf.lo[j];
#pragma unroll
      for (int i = 0; i < RP; ++i) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, al[i]),
                                                            __builtin_bit_cast(bf16x8, bh), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ah[i]),
                                                            __builtin_bit_cast(bf16x8, bl), acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ah[i]),
                                                            __builtin_bit_cast(bf16x8, bh), acc[i][j], 0, 0, 0);
      }
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < TNB; ++j) {
    const uint4 b = bf.b[j];
#pragma unroll
    for (int i = 0; i < RP; ++i) {
      if (KIND == K_F32) {
        f32x4 av = __builtin_bit_cast(f32x4, af[i].v[0]);
        f32x4 bv = __builtin_bit_cast(f32x4, b);
#pragma unroll
        for (int s = 0; s < 4; ++s)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[s], bv[s], acc[i][j], 0, 0, 0);
      } else if (KIND == K_BF16) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, af[i].v[0]),
                                                            __builtin_bit_cast(bf16x8, b), acc[i][j], 0, 0, 0);
      } else {
        acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(i32x4, af[i].v[0]),
                                                          __builtin_bit_cast(i32x4, b), acc[i][j], 0, 0, 0);
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

// One workgroup runs the whole chain on columns [n0, n0 + TN) of B.
template <int KIND, int TNB>
__device__ void chain_tile(const ChainParams &P, const long n0, uint4 *arena) {
  using T = CT<KIND>;
  using Acc = typename T::Acc;
  constexpr int TN = TNB * 32;
  constexpr int RP = T::RP;
  constexpr int PF = T::PF;
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
    const char *A = (const char *)P.A[gi];
    if (KIND == K_I8) {
      colsum_lds<TN>(arena, xg0, (int)((K + 31) / 32 * 32 / T::KG), colsum);
      __syncthreads();
    }
    // k-steps: `full` ones have every lane's fragment inside K; they run in chunks of
    // PF with no bounds logic. The remaining steps (K tail) use checked loads.
    const int ksteps = (int)((K + T::KC - 1) / T::KC);
    const int kfull = (int)(K / T::KC);
    const int nfc = kfull / PF;  // full chunks
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

      // Per-lane row pointers clamped into the matrix. Rows >= M compute garbage
      // that is never used: the last GEMM does not store them, and an intermediate's
      // padded rows meet zero A columns in the next GEMM (INT8 zeroes them anyway,
      // because they enter its column sums).
      const char *rowp[RP];
      bool row_ok[RP];
#pragma unroll
      for (int i = 0; i < RP; ++i) {
        const long row = rb0 + i * 32 + l32;
        row_ok[i] = row < M;
        rowp[i] = A + ((size_t)(row_ok[i] ? row : M - 1) * K + h * T::KA) * T::ESZ;
      }
      AFrag<KIND> fa[2][PF][RP];
      auto load_chunk = [&](int c, AFrag<KIND>(&dst)[PF][RP]) {
#pragma unroll
        for (int s = 0; s < PF; ++s) {
          const size_t off = (size_t)(c * PF + s) * T::KC * T::ESZ;
#pragma unroll
          for (int i = 0; i < RP; ++i) {
#pragma unroll
            for (int q = 0; q < (KIND == K_F32X3 ? 2 : 1); ++q) {
              uint4 v = *(const uint4 *)(rowp[i] + off + 16 * q);
              if (KIND == K_I8) {  // u8 -> s8 bias
                v.x ^= 0x80808080u; v.y ^= 0x80808080u; v.z ^= 0x80808080u; v.w ^= 0x80808080u;
              }
              dst[s][i].v[q] = v;
            }
          }
        }
      };
      // B fragments one step ahead, double buffered (PF is even, so the buffer of
      // step ks is ks & 1 across chunks). The read for the step after the last
      // clamps to the last step: harmless and branch free.
      BFrag<KIND, TNB> bq[2];
      const int glast = xg0 + 2 * (ksteps - 1) + h;
      auto run_chunk = [&](int c, const AFrag<KIND>(&src)[PF][RP]) {
#pragma unroll
        for (int s = 0; s < PF; ++s) {
          const int ks = c * PF + s;
          const int gn = xg0 + 2 * (ks + 1) + h;
          read_bfrag<KIND, TNB>(arena, lo_off, gn < glast ? gn : glast, l32, bq[(s + 1) & 1]);
          mma_k<KIND, TNB>(bq[s & 1], src[s], acc);
        }
      };
      read_bfrag<KIND, TNB>(arena, lo_off, xg0 + h, l32, bq[0]);
      if (nfc > 0) load_chunk(0, fa[0]);
      for (int c = 0; c < nfc; c += 2) {
        if (c + 1 < nfc) load_chunk(c + 1, fa[1]);
        run_chunk(c, fa[0]);
        if (c + 1 < nfc) {
          if (c + 2 < nfc) load_chunk(c + 2, fa[0]);
          run_chunk(c + 1, fa[1]);
        }
      }
      // K tail: checked loads, one step at a time
      for (int ks = nfc * PF; ks < ksteps; ++ks) {
        AFrag<KIND> af[RP];
        const long k = (long)ks * T::KC + h * T::KA;
#pragma unroll
        for (int i = 0; i < RP; ++i) af[i] = load_afrag<KIND>(rowp[i] - h * T::KA * T::ESZ, K, k, row_ok[i]);
        BFrag<KIND, TNB> bt;
        read_bfrag<KIND, TNB>(arena, lo_off, xg0 + 2 * ks + h, l32, bt);
        mma_k<KIND, TNB>(bt, af, acc);
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

constexpr int KG_of(int kind) { return kind == K_F32 ? 4 : kind == K_BF16 ? 8 : kind == K_I8 ? 16 : 8; }
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

}  // namespace

size_t chain_fused_lds_bytes(int kind, const long *M, const long *K, int num_gemm, int tile_n) {
  int xb[MAXG], yb[MAXG];
  if (num_gemm < 1 || num_gemm > MAXG) return (size_t)-1;
  int S = plan_arena(M, K, num_gemm, xb, yb);
  return (size_t)S * tile_n * LB_of(kind) + (kind == K_I8 ? tile_n * sizeof(int) : 0);
}

int chain_fused_max_gemms() { return MAXG; }

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
    P->vecA[i] = ((c.K[i] * e) % 16 == 0) && (((uintptr_t)c.A[i] & 15) == 0);
  }
  P->S = plan_arena(c.M, c.K, c.num_gemm, P->xbase, P->ybase);
  P->B = c.B;
  P->C = c.C;
  P->N = c.N;
  P->vecB = (c.N % 4 == 0) && (((uintptr_t)c.B & (4 * e - 1)) == 0);
}

bool chain_fused_shape_ok(int kind, long N, int num_gemm, const long *K) {
  static const int esz[4] = {4, 2, 1, 4};
  if (num_gemm < 1 || num_gemm > MAXG || N % 4 != 0) return false;
  for (int i = 0; i < num_gemm; ++i)
    if ((K[i] * esz[kind]) % 16 != 0) return false;
  return true;
}

bool chain_fused_aligned(const ChainParams &P) {
  if (!P.vecB) return false;
  for (int i = 0; i < P.num_gemm; ++i)
    if (!P.vecA[i]) return false;
  return true;
}

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

}  // namespace fg
