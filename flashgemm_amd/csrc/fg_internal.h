// Internal definitions shared by the flashgemm HIP translation units (gfx950 only).
//
// Orientation follows the reference (multigemm/src/flashgemm.h:21-25): a chain
// computes C = A_n ... A_1 B with B K1 x N, each A_i M_i x K_i, C M_n x N, all
// row-major with lda = K_i, ldb = ldc = N. N is the wide dimension; the A_i are the
// weights. Every kernel here produces a (rows of A) x (columns of B) tile.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace fg {

// ----------------------------------------------------------------- dtypes --
enum Dtype : int { DT_F32 = 0, DT_BF16 = 1, DT_INT8 = 2 };

// How an FP32 problem is computed.
enum F32Mode : int {
  F32_MFMA = 0,    // v_mfma_f32_16x16x4_f32, fp32 products and sums
  F32_BF16X3 = 1,  // x = hi + lo (bf16 pair); Ah.Bh + Ah.Bl + Al.Bh on bf16 MFMA
};

// What a kernel does with its fp32/int32 accumulator tile.
enum Epilogue : int {
  EPI_STORE = 0,    // C (fp32 or int32) row-major
  EPI_REQUANT = 1,  // next-GEMM operand: fp32 as is, bf16 RNE, int8 saturate
  EPI_PARTIAL = 2,  // split-K partial (fp32 or int32) into a workspace slab
};

// Compute "kinds" = (operand storage, MFMA) combinations.
// K_F32X6 (three-piece split FP32) is host-side only: it runs as a BF16 GEMM over the
// split operands (launch_split3), never as a kernel kind.
enum Kind : int { K_F32 = 0, K_BF16 = 1, K_I8 = 2, K_F32X3 = 3, K_F32X6 = 4 };

typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(16))) int i32x16;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) int i32x4;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4;
typedef __attribute__((ext_vector_type(2))) unsigned u32x2;

// ------------------------------------------------------------------- bf16 --
// fp32 -> bf16, round to nearest even; a denormal input becomes a signed zero and
// NaN stays a quiet NaN. This is the rounding the reference applies to its BF16
// intermediate (vcvtne2ps2bf16, multigemm/src/kernel_bf16.h:357).
__host__ __device__ inline uint16_t bf16_rne(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  if ((u & 0x7f800000u) == 0) u &= 0x80000000u;
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t b) {
  return __builtin_bit_cast(float, (uint32_t)b << 16);
}

__device__ inline int8_t sat_i8(int v) { return (int8_t)(v > 127 ? 127 : (v < -128 ? -128 : v)); }

// Two fp32 -> two bf16 (low half = first), round to nearest even, one gfx950
// v_cvt_pk_bf16_f32. Unlike bf16_rne above, a denormal fp32 input (|x| < 2^-126)
// rounds to a bf16 denormal instead of flushing to zero.
typedef __attribute__((ext_vector_type(2))) float f32x2_t;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2_t;
__device__ __forceinline__ uint32_t cvt_pk_bf16(float a, float b) {
  f32x2_t v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}
// Device-side fp32 -> bf16 of every kernel here (one rounding everywhere, so the
// fused and unfused paths produce identical intermediates).
__device__ __forceinline__ uint16_t bf16_dev(float x) { return (uint16_t)cvt_pk_bf16(x, 0.0f); }
// hi/lo split of two floats, packed: hi = rne(x), lo = rne(x - hi); x - hi is exact
// in fp32 (Sterbenz), so hi + lo carries 16 significand bits.
__device__ __forceinline__ void split_pk(float a, float b, uint32_t &hi, uint32_t &lo) {
  hi = cvt_pk_bf16(a, b);
  const float ha = __builtin_bit_cast(float, hi << 16), hb = __builtin_bit_cast(float, hi & 0xffff0000u);
  lo = cvt_pk_bf16(a - ha, b - hb);
}
__device__ __forceinline__ void split_bf16(float x, uint16_t &hi, uint16_t &lo) {
  uint32_t h, l;
  split_pk(x, 0.0f, h, l);
  hi = (uint16_t)h;
  lo = (uint16_t)l;
}
// Four int32 saturated to int8 and packed little-endian into one dword: clamps
// become v_med3_i32, the packing two v_perm_b32 (selector byte 0x0c yields zero;
// selectors 0-3 pick bytes of the second operand, 4-7 of the first).
__device__ __forceinline__ uint32_t sat_pk_i8(int a, int b, int c, int d) {
  a = a < -128 ? -128 : (a > 127 ? 127 : a);
  b = b < -128 ? -128 : (b > 127 ? 127 : b);
  c = c < -128 ? -128 : (c > 127 ? 127 : c);
  d = d < -128 ? -128 : (d > 127 ? 127 : d);
  const uint32_t ab = __builtin_amdgcn_perm((uint32_t)b, (uint32_t)a, 0x0c0c0400u);
  const uint32_t cd = __builtin_amdgcn_perm((uint32_t)d, (uint32_t)c, 0x04000c0cu);
  return ab | cd;
}

// ------------------------------------------------------ reference layout --
// The reference's GEMM i epilogue stores the requantized panel so that GEMM i+1
// reads X[k][c] = Z[src_row][panel + src_col] (panel-local column c, panel width w):
//   INT8 w=32: (k&~3)+(c>>3), 4(c&7)+(k&3)      INT8 w=16: (k&~3)+3-(k&3), c
//   BF16 w=32: (k&~1)+(c>>4), 2(c&15)+(k&1)     BF16 w=16: g+(g+1)/2+(c>>3), 2(c&7)+(k&1), g=k>>1
// (kernel_int8.h:329-385,1332-1386 against the packed-B order :6-50;
//  kernel_bf16.h:356-390,1332-1356 against :6-32). Panels start at id*nb + 32j
// inside each thread's range [id*nb, id*nb+NB) (multigemm/src/flashgemm.c:40-51,109-127).
struct RefPanels {
  long nb;     // columns per reference thread
  long ne;     // N % nb
  long num_n;  // threads that own columns
  long N;
};

__host__ __device__ inline void ref_panel_of(const RefPanels &p, long n, long &p0, long &w) {
  long id = n / p.nb;
  long NB = (id < p.num_n - 1) ? p.nb : (p.ne > 0 ? p.ne : p.nb);
  long j = ((n - id * p.nb) / 32) * 32;
  p0 = id * p.nb + j;
  w = (NB - j < 32) ? NB - j : 32;
}

// Returns false when (w) is outside the reference's domain {16, 32}.
__host__ __device__ inline bool ref_src(bool is_int8, long w, long k, long c, long &row, long &col) {
  if (is_int8) {
    if (w == 32) { row = (k & ~3L) + (c >> 3); col = 4 * (c & 7) + (k & 3); return true; }
    if (w == 16) { row = (k & ~3L) + 3 - (k & 3); col = c; return true; }
  } else {
    if (w == 32) { row = (k & ~1L) + (c >> 4); col = 2 * (c & 15) + (k & 1); return true; }
    if (w == 16) { long g = k >> 1; row = g + (g + 1) / 2 + (c >> 3); col = 2 * (c & 7) + (k & 1); return true; }
  }
  return false;
}

// ------------------------------------------------------------- epilogues --
constexpr int CHAIN_MAXG = 8;

// Epilogue extensions of one GEMM (row m = output channel). With every field at its
// default (null pointers, zero point 0, no ReLU) the epilogue is the reference's:
// int8 saturate, bf16 RNE, fp32 as is.
struct EpiOne {
  const void *bias;    // M entries: int32 (INT8) or fp32; nullptr: none
  const float *scale;  // M fp32 entries; nullptr: none (ignored for the INT8 C output)
  int zp;              // INT8 intermediates: zero point, -128..127
  int relu;            // nonzero: clamp below at zero (INT8 intermediates: at zp)
};
struct EpiParams {
  EpiOne g[CHAIN_MAXG];
};

// INT8 intermediate: a = acc + bias (wrapping int32); with a scale, t = fp32(a) *
// scale (RNE conversion and product) rounded half-to-even to an integer, else t = a;
// q = sat8(t + zp), then max(q, zp) under ReLU. t is clamped to [-512, 512] first,
// which cannot change q because zp lies in [-128, 127].
__device__ __forceinline__ int epi_i8(int acc, int bias, bool has_scale, float scale, int zp, int relu) {
  const int a = (int)((unsigned)acc + (unsigned)bias);
  int t;
  if (has_scale) {
    const float f = __fmul_rn((float)a, scale);
    t = (int)__builtin_rintf(__builtin_fminf(__builtin_fmaxf(f, -512.0f), 512.0f));
  } else {
    t = a < -512 ? -512 : (a > 512 ? 512 : a);
  }
  int q = t + zp;
  q = q < -128 ? -128 : (q > 127 ? 127 : q);
  return relu && q < zp ? zp : q;
}
// INT8 C output (int32): acc + bias (wrapping), then max(v, 0) under ReLU.
__device__ __forceinline__ int epi_i32(int acc, int bias, int relu) {
  const int v = (int)((unsigned)acc + (unsigned)bias);
  return relu && v < 0 ? 0 : v;
}
// FP32 / BF16: y = acc * scale, then + bias (two fp32 roundings, never fused), then
// max(y, 0) under ReLU (a NaN becomes 0).
__device__ __forceinline__ float epi_f(float acc, bool has_scale, float scale, bool has_bias, float bias, int relu) {
  float y = acc;
  if (has_scale) y = __fmul_rn(y, scale);
  if (has_bias) y = __fadd_rn(y, bias);
  return relu ? (y > 0.0f ? y : 0.0f) : y;
}

// ------------------------------------------------------------ launchers ---
// Launch helpers implemented in the .hip translation units. All return a
// hipError_t and never synchronize.

struct GemmArgs {
  int kind;         // Kind
  int epi;          // Epilogue
  const void *A;    // M x K row-major (int8 read as u8, bf16 raw bits)
  const void *B;    // K x N row-major (int8 read as s8)
  void *C;          // M x N row-major, element type by (kind, epi)
  const int *colsum;// INT8 split-K: sum_k B[k][n] (int32, N entries); nullptr: the kernel sums its panel
  long M, N, K;
  int splits;       // split-K factor (EPI_PARTIAL writes splits slabs of M x N)
  int trans;        // FP32 kinds, EPI_STORE / EPI_PARTIAL: bit 0 A stored K x M, bit 1 B stored N x K
  int kgroups;      // 1: a single FP32 GEMM may split each tile's k over two wave groups
};

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s);
// FP32 GEMMs with M <= 64 or N <= 64: 64 x 64 tiles, split-K (kper k per slab, slabs
// summed by launch_splitk_reduce when splits > 1).
bool skinny_f32_shape(long M, long N);
int skinny_f32_splits(long M, long N, long K, long *kper);
hipError_t launch_skinny_f32(const float *A, const float *B, float *C, long M, long N, long K, long kper,
                             int splits, hipStream_t s);
hipError_t launch_colsum_i8(const int8_t *B, long K, long N, int *out, hipStream_t s);
// Three-piece split of FP32 operands for C = A.B as one BF16 GEMM over k' = 6K:
// x = hi + mid + lo (each bf16 round-to-nearest-even of the remainder; the remainders
// are exact in fp32); A6 (M x 6K) has column blocks [hi hi mid hi lo mid] of A and
// B6 (6K x N) row blocks [hi mid hi lo hi mid] of B: block j of A6 meets block j of B6, so
// the k' sum is hi.hi + hi.mid + mid.hi + hi.lo + lo.hi + mid.mid (the products whose
// pieces sum to order <= 2; mid.lo, lo.mid, lo.lo are dropped, below 2^-24 relative).
hipError_t launch_split3(const float *A, long M, long K, uint16_t *A6, const float *B, long N, uint16_t *B6,
                         hipStream_t s);
// Sums the slabs and applies the epilogue `epi`, with the extensions of `ex` when it
// is not null (launch_splitk_reduce is also the epilogue pass of unfused GEMMs that
// have extensions).
hipError_t launch_splitk_reduce(int kind, int epi, const void *partials, int splits, long M, long N,
                                const int *colsum, void *C, const EpiOne *ex, hipStream_t s);
hipError_t launch_ref_permute(bool is_int8, const void *Z, void *X, long rows, const RefPanels &p,
                              hipStream_t s);

// Image geometry of a chain whose B is the patch matrix of an NCHW image (im2col):
// B[k][n] = X[b][c][oy*sh - ph + ky*dh][ox*sw - pw + kx*dw] (0 outside the image),
// k = (c*kh + ky)*kw + kx, n = (b*OH + oy)*OW + ox. The image holds fewer than 2^31
// elements, so element offsets are 32-bit.
struct ImgGeom {
  int C, H, W;
  int kh, kw, sh, sw, ph, pw, dh, dw;
  int OH, OW;
  int vec4;  // INT8: sw == 1, OW % 4 == 0, the image 4-byte aligned and a whole number of dwords
  int nq;    // image size in dwords (vec4)
};
// Element size `esz` (1, 2 or 4 bytes): B (K1 x N, K1 = C*kh*kw, N = batch*OH*OW).
hipError_t launch_im2col(int esz, const void *X, void *B, const ImgGeom &g, long K1, long N, hipStream_t s);

struct ChainArgs {
  int kind;                 // Kind
  int num_gemm;
  const void *A[CHAIN_MAXG];  // packed weights for the fused kernels
  long M[CHAIN_MAXG], K[CHAIN_MAXG];
  const void *B;            // K1 x N, or the image when img is set
  void *C;                  // M_n x N (fp32 or int32)
  long N;
  int tile_n;               // columns per workgroup
  const ImgGeom *img;       // nullptr: B is a matrix; else B is gathered from the image
  const EpiParams *epi;     // nullptr: the reference's epilogues; else per-GEMM extensions
  int raw_a;                // A[i] are the row-major FP32 weights, read unpacked (chain_raw_weights_ok)
  int refl;                 // INT8 / BF16: intermediates in the reference's layout (chain_refl_ok)
  int out_requant;          // the chain is a segment of a longer one: C receives the next
                            // GEMM's operand (int8 / bf16 / fp32 M_n x N), not int32 / fp32 C
};
// The fused kernels can write INT8 / BF16 intermediates in the reference's layout
// when every reference panel is a 32-aligned block (every thread's column range a
// multiple of 32, see RefPanels) and every intermediate M_i is a multiple of 32.
bool chain_refl_ok(int kind, const RefPanels &p, int num_gemm, const long *M);
// FP32 chains whose weights need no padding (every M_i % 32 == 0, every K_i a whole
// number of 64-k chunks, 16-byte aligned) on a matrix B with the plain epilogues: the
// fused kernel reads A_i row-major in fragment order and the per-call pack is skipped.
bool chain_raw_weights_ok(const ChainArgs &c);

// Device-side description of one chain for the fused kernels (filled on the host
// by make_chain_params; copied by value or read from a table in device memory).
// A[i] are packed weights (launch_pack_a layout for `kind`).
struct ChainParams {
  int kind;                 // Kind
  const void *A[CHAIN_MAXG];
  long M[CHAIN_MAXG], K[CHAIN_MAXG];
  int xbase[CHAIN_MAXG];    // arena row where X_i starts
  int ybase[CHAIN_MAXG];    // arena row where Y_i is written (i < n-1)
  int num_gemm;
  int S;                    // arena rows
  int hold;                 // bit i: Y_i may overlap X_i, GEMM i's final passes are held (plan_arena)
  const void *B;
  void *C;
  long N;
  int vecB;
  ImgGeom img;              // geometry when B is gathered from an image (chain_kernel IMG)
  EpiParams epi;            // epilogue extensions (chain_kernel EPX)
  int epi_off;              // EPX: LDS byte offset of the per-row bias and scale tables
  int refl;                 // intermediates in the reference's layout (store_rows_refl)
  int out_requant;          // INT8 / BF16: the last GEMM writes the next operand (int8 / bf16), not C
  int gtnb;                 // grouped launch: 32-column blocks per tile of this request (1 or 2)
};

void make_chain_params(const ChainArgs &c, ChainParams *P);
// The fused kernels stage B in groups of 4 columns: N % 4 == 0 (shape) and a B
// base aligned to 4 elements (a B gathered from an image has neither requirement).
bool chain_fused_shape_ok(int kind, long N, int num_gemm, const long *K);
bool chain_fused_aligned(const ChainParams &P);

// Weights in MFMA fragment order for the fused kernels (zero padded; INT8 biased
// to s8; F32X3 split into bf16 hi/lo). packed_a_bytes is the buffer size.
size_t packed_a_bytes(int kind, long M, long K);
constexpr int PACK_MAX = 32;
struct PackJob {  // up to PACK_MAX matrices of one kind in one launch
  const void *A[PACK_MAX];
  void *out[PACK_MAX];
  long M[PACK_MAX], K[PACK_MAX];
  long fstart[PACK_MAX + 1];  // filled by the launcher
  int n;
};
hipError_t launch_pack_a(int kind, const PackJob &J, hipStream_t s);

// Fused chain: every intermediate stays in LDS. Returns hipErrorInvalidValue when the
// largest (K_i + M_i) x tile_n panel does not fit the LDS budget.
hipError_t launch_chain_fused(const ChainArgs &c, hipStream_t s);
// LDS of a fused launch; epi: with the per-row bias and scale tables of the epilogue
// extensions (chain_epi_lds_bytes: 8 bytes per row of every GEMM).
size_t chain_fused_lds_bytes(int kind, const long *M, const long *K, int num_gemm, int tile_n, bool epi = false);
size_t chain_epi_lds_bytes(const long *M, int num_gemm);
int chain_fused_max_gemms();
// Fused tile widths: 32, 64, 128, 256 (96 for the FP32 kinds). Tiles of 256 (and 128
// for INT8/BF16) split the 8 waves over 32-column blocks (chain_fused_wave_groups
// groups), each walking rows in passes of (8 / groups) * chain_fused_rows_per_wave.
bool chain_fused_tile_ok(int kind, int tile_n);
int chain_fused_wave_groups(int kind, int tile_n);
int chain_fused_rows_per_wave(int kind);
constexpr size_t LDS_BYTES_MAX = 160 * 1024;

// Grouped persistent launch. The device table is a GroupTable header followed by
// num_reqs ChainParams at reqs_off and num_items int2 (request index, tile index) at
// items_off. `grid` workgroups walk the tile list; tile widths are fixed per kind
// (grouped_tile_n).
struct GroupTable {
  int num_reqs;
  int num_items;
  int reqs_off;   // bytes from the table start
  int items_off;  // bytes from the table start
};
int grouped_tile_n(int kind);
hipError_t launch_grouped(const void *table, int grid, size_t lds, bool wide, hipStream_t s);

// Grouped unfused GEMMs: one launch runs the 128 x 128 output tiles of many
// independent GEMMs (one level of many chains). Table: LevelTable header, GemmDesc[]
// at descs_off, int2 {desc, tile * splits + slice} items at items_off; tile t of a
// descriptor is (t / tiles_n, t % tiles_n). A descriptor with splits > 1 writes
// EPI_PARTIAL slabs (slice s covers k [s * kper, (s + 1) * kper)) that a grouped
// reduction (ReduceTable) adds in slab order; its split depends on its shape only, so
// a GEMM's bits do not depend on the other GEMMs of the launch.
struct GemmDesc {
  const void *A;  // M x K row-major
  const void *B;  // K x N row-major
  void *C;        // M x N: EPI_STORE fp32 / int32, EPI_REQUANT the next operand, EPI_PARTIAL slabs
  long M, N, K, kper;
  int kind, epi, tiles_n, vec;  // vec: bit 0 A rows 16-byte aligned, bit 1 B columns vectorizable
  int splits, pad;
};
struct LevelTable {
  int num_descs;
  int num_items;
  int descs_off;  // bytes from the table start
  int items_off;  // bytes from the table start
};
constexpr int LEVEL_TILE = 128;
// The reduction of a level's split GEMMs: item {desc, chunk} sums 256 consecutive
// elements over the descriptor's slabs in slab order, then requantizes (EPI_REQUANT)
// or stores (EPI_STORE) into C.
struct ReduceDesc {
  const void *P;  // splits slabs of MN 32-bit partials
  void *C;
  long MN;
  int splits, kind, epi, pad;
};
struct ReduceTable {
  int num_descs;
  int num_items;
  int descs_off;
  int items_off;
};
hipError_t launch_grouped_reduce(const void *table, int grid, hipStream_t s);
// Fills d (vector flags, tile counts) for C = A . B with the given kind and epilogue.
void gemm_desc(GemmDesc *d, int kind, int epi, const void *A, const void *B, void *C, long M, long N, long K);
// Splits d over k into about `splits` slices (whole k-stages; the count used is
// returned, 1 = unsplit and d unchanged): d then writes EPI_PARTIAL slabs at `slabs`.
int gemm_desc_split(GemmDesc *d, int splits, void *slabs);
// One launch per level; each descriptor has its own kind and epilogue (EPI_STORE or
// EPI_REQUANT). grid = the table's item count.
hipError_t launch_grouped_gemm(const void *table, int grid, hipStream_t s);

// FP64 error oracle (never on the product path).
hipError_t launch_f64_gemm(const double *A, const double *B, double *C, long M, long N, long K,
                           hipStream_t s);
hipError_t launch_to_f64(int src_kind, const void *src, double *dst, long n, hipStream_t s);
hipError_t launch_err_norms(int x_kind, const void *x, const double *ref, long n, double *out2,
                            hipStream_t s);
// Seeded random operands of element size esz (4: fp32 in [-1, 1), 2: bf16, 1: int8 in
// [-127, 127]) for timing runs.
hipError_t launch_fill_random(int esz, void *p, long n, unsigned long long seed, hipStream_t s);
// L2 -> CU stream-rate kernel (measurement): region of 8 * slice_frags KB, out of
// workgroups * 512 words.
hipError_t launch_stream_probe(const void *region, unsigned slice_frags, int rounds, int workgroups, unsigned *out,
                               hipStream_t s);

}  // namespace fg
