// GEMM on CDNA4 matrix cores (F15 matmul, F16 Dense affine).
//
// The reference computes every output with one thread looping over K from
// global memory (vulkpy/shader/matmul.comp:23-33, batch_affine.comp:25-39).
// Here a 256-thread block computes a BM x BN tile: operands are staged through
// LDS (K-contiguous rows, padded against bank conflicts), the next K slice is
// prefetched into registers while the current one is multiplied, and four
// waves (2 x 2) each own a (BM/2) x (BN/2) sub-tile built from 32 x 32 MFMA
// blocks:
//   FP32    v_mfma_f32_32x32x2_f32    (operands as stored)
//   BF16    v_mfma_f32_32x32x16_bf16  (operands rounded to nearest-even BF16)
//   BF16X3  the same instruction three times: hi*hi + hi*lo + lo*hi, where
//           x = hi + lo + O(2^-16 |x|) with hi = bf16(x), lo = bf16(x - hi)
// Accumulation is FP32 in every mode.  The epilogue applies alpha, beta*C,
// bias[n] and an activation, and can also emit a BF16 copy of the result.
// Blocks are remapped so that consecutive tiles land on the same XCD (each of
// the 8 XCDs has its own L2) and grouped along M for L2 reuse of B.
#include <type_traits>
#include <utility>

#include "common.h"
#include "fastdiv.h"

namespace vkh {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(16))) float f32x16_t;
// register staging of 8 BF16 values; a native vector (unlike HIP's uint4
// struct) stays in VGPRs when held in arrays across the K loop
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4_t;

// f(integral_constant<int, 0>), ..., f(integral_constant<int, N - 1>): a
// loop unrolled by construction
template <typename F, int... I>
__device__ __forceinline__ void static_for_seq(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}


typedef __attribute__((ext_vector_type(4))) float f32x4_t;

struct GemmArgs {
  const float* A;
  const float* B;
  float* C;
  const float* bias;
  uint16_t* Cbf;
  uint64_t M, N, K, lda, ldb, ldc;
  float alpha, beta;
  int act;
  const float* aux;  // activation output whose derivative multiplies the result
  uint64_t ldaux;
  int aux_act;
  int vec_c;  // C, bias, aux, Cbf rows allow 16-byte (Cbf: 8-byte) accesses
  uint32_t tiles_m, tiles_n;
  uint16_t* Cbft;  // C^T as BF16: Cbft[n * ldt + m] (optional)
  uint64_t ldt;
  // column sums of the final C over each chunk of CS_ROWS rows, in row order
  // from the chunk's first row: cs[(m / CS_ROWS) * N + n] (optional; staged
  // epilogues with 128-row wave slabs only)
  float* cs;
  int no_c;  // 1: the FP32 C is not stored (only Cbf / Cbft / cs)
  // the aux activation's output as BF16 instead of aux (ReLU' only: the BF16
  // copy has the sign, zero-ness and NaN-ness of the value)
  const uint16_t* aux16;
};
constexpr int CS_ROWS = 128;

// SIG = false: the caller has ruled out the sigmoid (whose FP64 exp is most
// of the code of an unrolled epilogue)
template <bool SIG = true>
__device__ __forceinline__ float act_apply(int act, float v) {
  if (act == VKH_ACT_RELU) return (v >= 0.0f || v != v) ? v : 0.0f;  // np.maximum(v, 0)
  if (SIG && act == VKH_ACT_SIGMOID) return 1.0f / (1.0f + (float)exp((double)(-v)));
  return v;
}

// dy * act'(y) given the activation's output y, in the op order of
// nn/layers.py:207-210 (ReLU) and 264-267 (Sigmoid)
__device__ __forceinline__ float act_deriv_mul(int act, float y, float dy) {
  if (act == VKH_ACT_RELU) {
    // max(sign(y), 0) (np.maximum keeps a NaN): 1 for y > 0, y itself for a
    // NaN, +0 otherwise
    const float d = y > 0.f ? 1.f : (y != y ? y : 0.f);
    return d * dy;
  }
  if (act == VKH_ACT_SIGMOID) {
    float d = 1.0f - y;
    d = d * y;
    return d * dy;
  }
  return dy;
}

// Reads 4 consecutive elements of a row (guarded at the row end).
__device__ __forceinline__ float4 load_row4(const float* row, uint64_t c, uint64_t limit, bool vec) {
  if (vec && c + 3 < limit) return *reinterpret_cast<const float4*>(row + c);
  float4 v;
  v.x = c < limit ? row[c] : 0.f;
  v.y = c + 1 < limit ? row[c + 1] : 0.f;
  v.z = c + 2 < limit ? row[c + 2] : 0.f;
  v.w = c + 3 < limit ? row[c + 3] : 0.f;
  return v;
}

// XCD-aware tile order: consecutive remapped ids share an XCD, then tiles are
// grouped by GROUP_M rows of tiles.
__device__ __forceinline__ void tile_coords(const GemmArgs& p, uint32_t& tm, uint32_t& tn) {
  const uint32_t nb = p.tiles_m * p.tiles_n;
  uint32_t b = blockIdx.x;
  if (nb % 8 == 0) b = (b % 8) * (nb / 8) + b / 8;
  constexpr uint32_t GROUP_M = 8;
  const uint32_t width = GROUP_M * p.tiles_n;
  const uint32_t group = b / width;
  const uint32_t first_m = group * GROUP_M;
  const uint32_t gsize = min(p.tiles_m - first_m, GROUP_M);
  tm = first_m + (b % width) % gsize;
  tn = (b % width) / gsize;
}

// ---- packed BF16 path ---------------------------------------------------------------
// BF16 / BF16X3 GEMMs first pack op(A) and op(B)^T into zero-padded, K-major
// BF16 buffers ([Mp][Kp], [Np][Kp]; for BF16X3 also the low parts), so the
// main kernel's inner loop is 16-byte loads, 16-byte LDS stores and MFMAs with
// no conversion and no bounds checks.
constexpr int PK_KALIGN = 64;  // packed K (Kp) is a multiple of every kernel's k slice
constexpr int PK_TR = 128;     // transpose tile rows (r) in k_pack_cols (divides every Rp)
constexpr int PK_TK = 128;     // transpose tile k in k_pack_cols

// dst[r][k] = bf16(op(src)[r][k]) for r < Rp, k < Kp, zero outside [R, K);
// op(src)[r][k] = src[r*ld + k] (TRANS = false) or src[k*ld + r] (TRANS = true).
template <bool SPLIT>
__device__ __forceinline__ void pack8(const float v[8], uint16_t* hi, uint16_t* lo) {
  uint16_t h[8], l[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    h[q] = f2bf(v[q]);
    if (SPLIT) l[q] = f2bf(v[q] - bf2f(h[q]));
  }
  uint4 hv, lv;
  hv.x = h[0] | ((uint32_t)h[1] << 16); hv.y = h[2] | ((uint32_t)h[3] << 16);
  hv.z = h[4] | ((uint32_t)h[5] << 16); hv.w = h[6] | ((uint32_t)h[7] << 16);
  *reinterpret_cast<uint4*>(hi) = hv;
  if (SPLIT) {
    lv.x = l[0] | ((uint32_t)l[1] << 16); lv.y = l[2] | ((uint32_t)l[3] << 16);
    lv.z = l[4] | ((uint32_t)l[5] << 16); lv.w = l[6] | ((uint32_t)l[7] << 16);
    *reinterpret_cast<uint4*>(lo) = lv;
  }
}

// Row-major source: 4 consecutive elements a thread, one thread per 4 of
// the padded [Rp][Kp] (Rp * Kp / 4 < 2^31: the row from one 32-bit
// multiply-high division), so that each load instruction reads one contiguous
// 1 KB run and each store writes 512 B.
__device__ __forceinline__ uint32_t bf2x(float a, float b) { return f2bf(a) | ((uint32_t)f2bf(b) << 16); }

template <bool SPLIT>
__global__ void __launch_bounds__(256)
k_pack_rows(const float* src, uint64_t ld, uint64_t R, uint64_t K, uint16_t* hi, uint16_t* lo,
            uint32_t total4, FastDiv kq, uint64_t Kp, int vec) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= total4) return;
  const uint32_t r = kq.div(t);
  const uint64_t k = (uint64_t)(t - r * kq.d) * 4;
  float v[4];
  if (r < R && vec && k + 3 < K) {
    const float4 a = ld_stream(reinterpret_cast<const float4*>(src + r * ld + k));
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = (r < R && k + q < K) ? src[r * ld + k + q] : 0.f;
  }
  const uint64_t off = (uint64_t)r * Kp + k;
  *reinterpret_cast<uint2*>(hi + off) = make_uint2(bf2x(v[0], v[1]), bf2x(v[2], v[3]));
  if constexpr (SPLIT) {
    float l[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) l[q] = v[q] - bf2f(f2bf(v[q]));
    *reinterpret_cast<uint2*>(lo + off) = make_uint2(bf2x(l[0], l[1]), bf2x(l[2], l[3]));
  }
}

// Transposed source: a PK_TK (k) x PK_TR (r) tile goes through LDS (BF16
// pairs for the plain pack, rows of PK_TR + 1 floats for the split one: the
// column reads of the second phase spread over the banks), so the reads are
// 16-byte loads along r (512-byte row segments) and the writes 16-byte
// stores along k (256-byte row segments of the packed buffer).  Rp is a
// multiple of PK_TR; the last k tile may pass Kp.
template <bool SPLIT>
__global__ void __launch_bounds__(256)
k_pack_cols(const float* src, uint64_t ld, uint64_t R, uint64_t K, uint16_t* hi, uint16_t* lo,
            uint64_t Rp, uint64_t Kp, int vec) {
  const uint64_t r0 = (uint64_t)blockIdx.x * PK_TR, k0 = (uint64_t)blockIdx.y * PK_TK;
  constexpr int R4 = PK_TR / 4;
  constexpr int NL = PK_TK * R4 / 256;  // 16-byte loads per thread, all issued before the LDS writes
  vkh_f4v v[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int f = threadIdx.x + i * 256;
    const int kk = f / R4, rr = (f % R4) * 4;
    const uint64_t r = r0 + rr, k = k0 + kk;
    if (vec && k < K && r + 3 < R) {
      v[i] = *reinterpret_cast<const vkh_f4v*>(src + k * ld + r);
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[i][q] = (k < K && r + q < R) ? src[k * ld + r + q] : 0.f;
    }
  }
  if constexpr (!SPLIT) {
    // BF16 tile: the values are rounded on the way in (the rounding the
    // packed row would apply), so the tile takes 33 KB and four blocks fit
    // a compute unit.  Word w of row kk holds r = 2w (low half) and 2w + 1;
    // a thread of the second phase reads one word of 8 consecutive k and
    // writes the packed rows of both r.
    constexpr int PW = PK_TR / 2 + 1;  // words a row: odd, so the column reads spread over banks
    __shared__ uint32_t tb[PK_TK][PW];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int f = threadIdx.x + i * 256;
      const int kk = f / R4, w = (f % R4) * 2;
      tb[kk][w] = bf2x(v[i][0], v[i][1]);
      tb[kk][w + 1] = bf2x(v[i][2], v[i][3]);
    }
    __syncthreads();
    constexpr int KC = PK_TK / 8;
    for (int f = threadIdx.x; f < (PK_TR / 2) * KC; f += 256) {
      const int rp = f / KC, kc = f % KC;
      const uint64_t k = k0 + kc * 8;
      if (k >= Kp) continue;
      uint32_t wd[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) wd[q] = tb[kc * 8 + q][rp];
      uint4 ev, od;  // rows 2 rp (low halves) and 2 rp + 1 (high halves), k ascending
      ev.x = (wd[0] & 0xffffu) | (wd[1] << 16); od.x = (wd[0] >> 16) | (wd[1] & 0xffff0000u);
      ev.y = (wd[2] & 0xffffu) | (wd[3] << 16); od.y = (wd[2] >> 16) | (wd[3] & 0xffff0000u);
      ev.z = (wd[4] & 0xffffu) | (wd[5] << 16); od.z = (wd[4] >> 16) | (wd[5] & 0xffff0000u);
      ev.w = (wd[6] & 0xffffu) | (wd[7] << 16); od.w = (wd[6] >> 16) | (wd[7] & 0xffff0000u);
      const uint64_t off = (r0 + 2 * rp) * Kp + k;
      *reinterpret_cast<uint4*>(hi + off) = ev;
      *reinterpret_cast<uint4*>(hi + off + Kp) = od;
    }
  } else {  // FP32 tile: the split needs each value's low part too
    __shared__ float t[PK_TK][PK_TR + 1];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int f = threadIdx.x + i * 256;
      const int kk = f / R4, rr = (f % R4) * 4;
#pragma unroll
      for (int q = 0; q < 4; ++q) t[kk][rr + q] = v[i][q];
    }
    __syncthreads();
    constexpr int KC = PK_TK / 8;  // 8-element chunks per packed row in the tile
    for (int f = threadIdx.x; f < PK_TR * KC; f += 256) {
      const int rr = f / KC, kc = f % KC;
      const uint64_t k = k0 + kc * 8;
      if (k >= Kp) continue;
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = t[kc * 8 + q][rr];
      const uint64_t off = (r0 + rr) * Kp + k;
      pack8<SPLIT>(v, hi + off, SPLIT ? lo + off : nullptr);
    }
  }
}

struct PackedArgs {
  const uint16_t* Ah;
  const uint16_t* Al;
  const uint16_t* Bh;
  const uint16_t* Bl;
  uint64_t Kp;
  // split k: block (tile, blockIdx.y = s) accumulates k in [s*kc, (s+1)*kc)
  // and writes its raw accumulators to part[s][M][N]; k_splitk_combine then
  // applies the epilogue.  part == nullptr: the whole k range, epilogue here.
  uint64_t kc;
  float* part;
  GemmArgs e;
};

// Register stage of one packed operand's ROWS x BK slice: 16-byte loads of
// K-contiguous BF16 rows (both parts in split mode), stored to LDS rows of
// LDK = BK + 8 elements (the pad keeps rows 16-byte aligned and spreads the
// rows of a 16-lane LDS access over distinct banks).
template <int ROWS, int BK, bool SPLIT, int THREADS>
struct PackedStager {
  static constexpr int LDK = BK + 8;
  static constexpr int NV = ROWS * BK / 8 / THREADS;  // 16-byte vectors per thread
  static_assert(ROWS * BK / 8 % THREADS == 0, "whole 16-byte vectors per thread");
  static constexpr int PARTS = SPLIT ? 2 : 1;
  u32x4_t r[PARTS][NV];

  __device__ __forceinline__ void load(const uint16_t* hi, const uint16_t* lo, uint64_t row0,
                                       uint64_t Kp, uint64_t k0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = threadIdx.x + i * THREADS;
      const int row = idx / (BK / 8), c = idx % (BK / 8);
      const uint64_t off = (row0 + row) * Kp + k0 + c * 8;
      r[0][i] = *reinterpret_cast<const u32x4_t*>(hi + off);
      if constexpr (SPLIT) r[1][i] = *reinterpret_cast<const u32x4_t*>(lo + off);
    }
  }
  __device__ __forceinline__ void store(uint16_t* hi, uint16_t* lo) const {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = threadIdx.x + i * THREADS;
      const int row = idx / (BK / 8), c = idx % (BK / 8);
      *reinterpret_cast<u32x4_t*>(hi + row * LDK + c * 8) = r[0][i];
      if constexpr (SPLIT) *reinterpret_cast<u32x4_t*>(lo + row * LDK + c * 8) = r[1][i];
    }
  }
};

// 4 consecutive floats p[0 .. cnt) (one 16-byte load when `full`), zeros beyond.
__device__ __forceinline__ f32x4_t load4(const float* p, bool full, int cnt) {
  if (full) return *reinterpret_cast<const f32x4_t*>(p);
  f32x4_t v;
  v[0] = cnt > 0 ? p[0] : 0.f;
  v[1] = cnt > 1 ? p[1] : 0.f;
  v[2] = cnt > 2 ? p[2] : 0.f;
  v[3] = cnt > 3 ? p[3] : 0.f;
  return v;
}

// 4 consecutive BF16 values p[0 .. cnt) as floats (one 8-byte load when
// `full`), zeros beyond.
__device__ __forceinline__ f32x4_t load4_bf(const uint16_t* p, bool full, int cnt) {
  f32x4_t v;
  if (full) {
    const uint2 t = *reinterpret_cast<const uint2*>(p);
    v[0] = __uint_as_float(t.x << 16);
    v[1] = __uint_as_float(t.x & 0xffff0000u);
    v[2] = __uint_as_float(t.y << 16);
    v[3] = __uint_as_float(t.y & 0xffff0000u);
    return v;
  }
  v[0] = cnt > 0 ? bf2f(p[0]) : 0.f;
  v[1] = cnt > 1 ? bf2f(p[1]) : 0.f;
  v[2] = cnt > 2 ? bf2f(p[2]) : 0.f;
  v[3] = cnt > 3 ? bf2f(p[3]) : 0.f;
  return v;
}

// the aux activation's outputs at C[m, n..n+3] (FP32 or BF16 aux)
__device__ __forceinline__ f32x4_t aux4(const GemmArgs& e, uint64_t m, uint64_t n, bool full, int cnt) {
  return e.aux16 ? load4_bf(e.aux16 + m * e.ldaux + n, full, cnt) : load4(e.aux + m * e.ldaux + n, full, cnt);
}

// The final values of 4 consecutive outputs C[m, n..n+3] from their
// accumulators, given C's old values (beta != 0), bias[n..n+3] and the aux
// activation's outputs, in the order alpha*acc (+ beta*C) (+ bias[n]),
// activation, times the aux activation's derivative; stored with 16-byte
// accesses when `full` (cnt outputs of the group in range otherwise).  The
// transposed BF16 copy is written here, one element at a time, unless write_t
// is false (the staged epilogue writes it in whole runs).
// D: x holds the derivative factors themselves (ReLU' decoded from codes)
// bf (full groups): receives the four values as BF16, two a word, when it
// converts them for e.Cbf or when `want_bf` (the caller writes them again)
template <bool SIG = true, bool D = false>
__device__ __forceinline__ float4 finish4(const GemmArgs& e, uint64_t m, uint64_t n, float4 acc, bool full,
                                          int cnt, f32x4_t c, f32x4_t b, f32x4_t x, bool write_t,
                                          uint2* bf = nullptr, bool want_bf = false) {
  float* cp = e.C + m * e.ldc + n;
  float v[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    float t = e.alpha * v[q];
    if (e.beta != 0.f) t += e.beta * c[q];
    if (e.bias) t += b[q];
    t = act_apply<SIG>(e.act, t);
    if (e.aux || e.aux16) t = D ? x[q] * t : act_deriv_mul(e.aux_act, x[q], t);
    v[q] = t;
  }
  if (full) {
    if (!e.no_c) *reinterpret_cast<float4*>(cp) = make_float4(v[0], v[1], v[2], v[3]);
    if (e.Cbf || want_bf) {
      uint2 h;
      h.x = f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
      h.y = f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
      if (e.Cbf) *reinterpret_cast<uint2*>(e.Cbf + m * e.ldc + n) = h;
      if (bf) *bf = h;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < cnt) {
        if (!e.no_c) cp[q] = v[q];
        if (e.Cbf) e.Cbf[m * e.ldc + n + q] = f2bf(v[q]);
      }
    }
  }
  if (e.Cbft && write_t) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < cnt) e.Cbft[(n + q) * e.ldt + m] = f2bf(v[q]);
  }
  return make_float4(v[0], v[1], v[2], v[3]);
}

// finish4 with its inputs read here (C's old values, bias, aux).
__device__ __forceinline__ float4 epilogue4(const GemmArgs& e, uint64_t m, uint64_t n, float4 acc,
                                            bool write_t = true) {
  const bool full = e.vec_c && n + 3 < e.N;
  const int cnt = full ? 4 : (int)min<uint64_t>(4, e.N - n);
  const f32x4_t z = {0.f, 0.f, 0.f, 0.f};
  const f32x4_t c = e.beta != 0.f ? load4(e.C + m * e.ldc + n, full, cnt) : z;
  const f32x4_t b = e.bias ? load4(e.bias + n, full, cnt) : z;
  const f32x4_t x = e.aux || e.aux16 ? aux4(e, m, n, full, cnt) : z;
  return finish4(e, m, n, acc, full, cnt, c, b, x, write_t);
}

// Raw accumulators C[m, n..n+3] of one k chunk into part [M][N] (16-byte
// store when the group is whole and rows are 16-byte aligned).
__device__ __forceinline__ void store_part4(float* part, uint64_t N, uint64_t m, uint64_t n, float4 v) {
  float* pp = part + m * N + n;
  if (N % 4 == 0) {
    *reinterpret_cast<float4*>(pp) = v;
  } else {
    const float a[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (n + q < N) pp[q] = a[q];
  }
}

// Epilogue of a block whose waves each hold TM x TN 32 x 32 accumulator
// blocks of a WM x WN sub-tile at (wm, wn): acc[i][j][r] is C[m, n] with
// m = m0 + wm*WM + i*32 + (r/4)*8 + lhi*4 + r%4 and n = n0 + wn*WN + j*32 +
// l32.  Each wave writes rows i*32 .. i*32+31 of its sub-tile to its own
// 32 x (WN+4) float staging slice of `smem`, then reads them back as 4-float
// row segments (WN/4 lanes per row), so C is accessed in whole row segments.
// With `part` the raw sums go to that k chunk's partial matrix instead.  The
// transposed BF16 copy (e.Cbft): with TST (the caller's LDS holds a second
// area `tbase` of WN x TT_PITCH BF16 a wave after the staging slices) the
// row pass also writes each final value as BF16 into that wave's transposed
// tile, and after every second pass each group of eight lanes writes one
// column's 64 rows as one 128-byte run of C^T (16 bytes a lane).  Without
// it the row pass puts the final values back into the slice and each group
// of four lanes converts one column of 32 rows into a 64-byte run.  The
// caller's operand tiles in `smem` must be dead (a barrier before the call).
constexpr int TT_PITCH = 72;  // transposed tile row: 64 rows of C + 8, 16-byte aligned
template <int WAVES, int WN>
constexpr int tt_bytes() { return WAVES * WN * TT_PITCH * 2; }

template <int TM, int TN, int WM, int WN, bool TST = false>
__device__ __forceinline__ void staged_epilogue(float* smem, f32x16_t (&acc)[TM][TN], const GemmArgs& e,
                                                float* part, uint64_t m0, uint64_t n0, int wm, int wn,
                                                uint16_t* tbase = nullptr) {
  constexpr int SROW = WN + 4, F4R = WN / 4, RPI = 64 / F4R;
  static_assert(64 % F4R == 0, "whole rows per wave pass");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l32 = lane & 31, lhi = lane >> 5;
  float* st = smem + wave * 32 * SROW;
  const uint64_t nb = n0 + wn * WN;
  const bool tr = e.Cbft && !part;
  const bool tst = TST && tr;
  uint16_t* tt = TST ? tbase + wave * WN * TT_PITCH : nullptr;
  // the final values of row segment (rr, cc..cc+3) of pass i into the
  // transposed tile, rows of even passes first (as BF16 pairs already
  // converted, or converted here)
  auto put_t_bf = [&](int i, int rr, int cc, uint2 h) {
    uint16_t* d = tt + cc * TT_PITCH + (i & 1) * 32 + rr;
    d[0] = (uint16_t)h.x;
    d[TT_PITCH] = (uint16_t)(h.x >> 16);
    d[2 * TT_PITCH] = (uint16_t)h.y;
    d[3 * TT_PITCH] = (uint16_t)(h.y >> 16);
  };
  auto put_t = [&](int i, int rr, int cc, float4 f) {
    put_t_bf(i, rr, cc, make_uint2(f2bf(f.x) | ((uint32_t)f2bf(f.y) << 16), f2bf(f.z) | ((uint32_t)f2bf(f.w) << 16)));
  };
  const bool tr_vec = e.ldt % 8 == 0 && ((uintptr_t)e.Cbft & 15u) == 0;
  // column sums: the wave's WM = 128 rows are one chunk; lane c < WN walks
  // column c down the staged slices in row order, 32 rows per pass
  const bool csum = WM == CS_ROWS && e.cs && !part;
  float csacc = 0.f;
  // A lane's row segments: columns n .. n+3 of rows r0 + k*RPI of each pass.
  // In the common case (a whole 4-column group, every row of the pass in
  // range, no sigmoid) no global load waits between the passes' stores (a
  // wait on a load also waits for every store issued before it): bias is
  // read once; a ReLU' aux is read for every pass before the first store and
  // kept as a 2-bit code per value (positive, NaN, else), which is all ReLU'
  // needs (the BF16 aux); other aux values are in registers before their pass,
  // read during the previous pass where they fit (ITER <= 8), else in groups
  // of PF at its start.  Other segments (and the split-k partial sums) take
  // one loop iteration each, with their own loads.
  constexpr int ITER = 32 / RPI, PF = ITER < 8 ? ITER : 8;
  constexpr bool AHEAD = ITER == PF;
  const int cc = (lane % F4R) * 4, r0 = lane / F4R;
  const uint64_t n = nb + cc;
  const bool ncol = n < e.N;
  const bool lane_fast = !part && e.act != VKH_ACT_SIGMOID && e.vec_c && n + 3 < e.N;
  const f32x4_t z = {0.f, 0.f, 0.f, 0.f};
  const f32x4_t bq = lane_fast && e.bias ? *reinterpret_cast<const f32x4_t*>(e.bias + n) : z;
  const bool aux = lane_fast && (e.aux || e.aux16);
  const bool relu_codes = aux && e.aux16;  // (a BF16 aux is ReLU' only)
  const bool xpre = aux && !relu_codes;
  // codes: 8 bits a row segment (2 a value), four segments a word; read as
  // raw bits, a pass a batch
  constexpr int NSEG = TM * ITER;
  uint32_t cw[(NSEG + 3) / 4];
#pragma unroll
  for (int w = 0; w < (NSEG + 3) / 4; ++w) cw[w] = 0u;
  auto put_code = [&](int seg, uint32_t byte) { cw[seg / 4] |= byte << (8 * (seg % 4)); };
  if (relu_codes) {
    constexpr int BP = 1;
#pragma unroll
    for (int i0 = 0; i0 < TM; i0 += BP) {
      uint2 raw[BP * ITER];
#pragma unroll
      for (int g = 0; g < BP * ITER; ++g) {
        const int i = i0 + g / ITER, k = g % ITER;
        const uint64_t m = m0 + wm * WM + i * 32 + r0 + k * RPI;
        raw[g] = (i < TM && m < e.M) ? *reinterpret_cast<const uint2*>(e.aux16 + m * e.ldaux + n)
                                     : make_uint2(0u, 0u);
      }
#pragma unroll
      for (int g = 0; g < BP * ITER; ++g) {
        if (i0 * ITER + g < NSEG) {
          uint32_t byte = 0u;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t hb = ((q < 2 ? raw[g].x : raw[g].y) >> (16 * (q & 1))) & 0xffffu, mag = hb & 0x7fffu;
            byte |= (mag > 0x7f80u ? 2u : ((hb & 0x8000u) == 0u && mag != 0u ? 1u : 0u)) << (2 * q);
          }
          put_code(i0 * ITER + g, byte);
        }
      }
    }
  }
  // ReLU' of segment seg's aux values, as act_deriv_mul finds it: 1
  // (positive), NaN (NaN) or 0
  auto relu_x = [&](int seg) {
    const uint32_t byte = (cw[seg / 4] >> (8 * (seg % 4))) & 0xffu;
    f32x4_t xv;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t c = (byte >> (2 * q)) & 3u;
      xv[q] = c == 1u ? 1.f : (c == 2u ? __builtin_nanf("") : 0.f);
    }
    return xv;
  };
  f32x4_t xq[PF];
#pragma unroll
  for (int k = 0; k < PF; ++k) xq[k] = z;
  auto load_aux = [&](int i, int k0) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const uint64_t m = m0 + wm * WM + i * 32 + r0 + (k0 + k) * RPI;
      xq[k] = m < e.M ? aux4(e, m, n, true, 4) : z;
    }
  };
  if (xpre && AHEAD) load_aux(0, 0);
  // pass i: rows i*32 .. i*32+31 of the sub-tile (a compile-time i, so the
  // accumulators stay in registers however large the pass's code)
  auto pass = [&](auto ic) __attribute__((always_inline)) {
    constexpr int i = decltype(ic)::value;
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r)
        st[((r / 4) * 8 + lhi * 4 + (r % 4)) * SROW + j * 32 + l32] = acc[i][j][r];
    __syncthreads();
    const uint64_t mb = m0 + wm * WM + i * 32;
    if (lane_fast && mb + 31 < e.M) {
#pragma unroll
      for (int k0 = 0; k0 < ITER; k0 += PF) {
        if (xpre && !AHEAD) load_aux(i, k0);
#pragma unroll
        for (int k = 0; k < PF; ++k) {
          const int rr = r0 + (k0 + k) * RPI;
          const float4 v = *reinterpret_cast<const float4*>(st + rr * SROW + cc);
          const f32x4_t c = e.beta != 0.f ? *reinterpret_cast<const f32x4_t*>(e.C + (mb + rr) * e.ldc + n) : z;
          uint2 hb = make_uint2(0u, 0u);
          const float4 f = relu_codes
                               ? finish4<false, true>(e, mb + rr, n, v, true, 4, c, bq, relu_x(i * ITER + k0 + k), false,
                                                      &hb, tst)
                               : finish4<false>(e, mb + rr, n, v, true, 4, c, bq, xq[k], false, &hb, tst);
          if (tst) put_t_bf(i, rr, cc, hb);
          if (csum || (tr && !tst)) *reinterpret_cast<float4*>(st + rr * SROW + cc) = f;
        }
      }
    } else {
      for (int rr = r0; rr < 32; rr += RPI) {
        const uint64_t m = mb + rr;
        if (m < e.M && ncol) {
          const float4 v = *reinterpret_cast<const float4*>(st + rr * SROW + cc);
          if (part) {
            store_part4(part, e.N, m, n, v);
          } else {
            const float4 f = epilogue4(e, m, n, v, false);
            if (tst) put_t(i, rr, cc, f);
            if (csum || (tr && !tst)) *reinterpret_cast<float4*>(st + rr * SROW + cc) = f;
          }
        }
      }
    }
    if (xpre && AHEAD && i + 1 < TM) load_aux(i + 1, 0);
    __syncthreads();
    if (csum && lane < WN) {
      if (mb + 31 < e.M) {
        // the sum stays a chain in row order; its reads go out 8 at a time
#pragma unroll
        for (int r8 = 0; r8 < 32; r8 += 8) {
          float v[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = st[(r8 + q) * SROW + lane];
#pragma unroll
          for (int q = 0; q < 8; ++q) csacc = (i == 0 && r8 + q == 0) ? v[q] : csacc + v[q];
        }
      } else {
        for (int rr = 0; rr < 32; ++rr) {
          if (mb + rr >= e.M) break;
          const float v = st[rr * SROW + lane];
          csacc = (i == 0 && rr == 0) ? v : csacc + v;
        }
      }
    }
    if (tst && ((i & 1) || i == TM - 1)) {
      // rows (i & ~1) * 32 .. of the slab: eight lanes a column, 16 bytes each
      const uint64_t mp = m0 + wm * WM + (i & ~1) * 32;
      const int q = lane & 7, rows = (i & 1) ? 64 : 32;
      if (q * 8 < rows) {
        for (int c = lane >> 3; c < WN; c += 8) {
          const uint64_t nt = nb + c;
          if (nt >= e.N) continue;
          uint16_t* dst = e.Cbft + nt * e.ldt + mp + q * 8;
          const uint16_t* src = tt + c * TT_PITCH + q * 8;
          if (tr_vec && mp + q * 8 + 7 < e.M) {
            *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(src);
          } else {
            for (int r = 0; r < 8; ++r)
              if (mp + q * 8 + r < e.M) dst[r] = src[r];
          }
        }
      }
    }
    if (csum && !(tr && !tst)) __syncthreads();  // st (and tt) are read before the next pass writes them
    if (tr && !tst) {
      // four lanes per column, 8 rows (16 bytes) each: one store covers 16
      // columns in whole 64-byte runs
      const int q = lane & 3;
      for (int c = lane >> 2; c < WN; c += 16) {
        const uint64_t nt = nb + c;
        if (nt >= e.N) continue;
        uint16_t* dst = e.Cbft + nt * e.ldt + mb + q * 8;
        const float* col = st + q * 8 * SROW + c;
        if (tr_vec && mb + 31 < e.M) {
          uint32_t w[4];
#pragma unroll
          for (int t = 0; t < 4; ++t)
            w[t] = f2bf(col[(2 * t) * SROW]) | ((uint32_t)f2bf(col[(2 * t + 1) * SROW]) << 16);
          *reinterpret_cast<uint4*>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
          for (int r = 0; r < 8; ++r)
            if (mb + q * 8 + r < e.M) dst[r] = f2bf(col[r * SROW]);
        }
      }
      __syncthreads();
    }
  };
  static_for<TM>(pass);
  if (csum && lane < WN) {
    const uint64_t m = m0 + wm * WM, nc = nb + lane;
    if (m < e.M && nc < e.N) e.cs[(m / CS_ROWS) * e.N + nc] = csacc;
  }
}

// ---- FP32 path ----------------------------------------------------------------------
// FP32 products on v_mfma_f32_32x32x2_f32 with the operands read as stored
// (no packing pass).  A BM x BN tile per block, WGM x WGN waves of
// (BM/WGM) x (BN/WGN); 32-deep k slices go through registers into
// double-buffered LDS with one barrier per slice, the next slice's global
// loads in flight while the current one is multiplied.  Each operand keeps
// its global layout in LDS:
//   * K-major (A row-major [M][K], B stored [N][K]): rows of 32 floats
//     (128 B) whose 16-byte chunks are XOR-swizzled by row (chunk c of row r
//     sits at position c ^ (r & 7)), so 8 consecutive rows read at one chunk
//     touch 8 distinct bank groups;
//   * otherwise (A stored [K][M], B row-major [K][N]): 32 k rows of the tile
//     width plus 8 floats, so the two half-waves of a fragment read (k rows 4
//     apart) fall on different banks.
// The MFMA takes one float of A and one of B per lane: lane (r, h), r =
// lane % 32, h = lane / 32, supplies row (column) r at its k.  In the 8-deep
// k group g of a slice, MFMA t = 0..3 takes k = 8g + 4h + t, so a K-major
// operand feeds the whole group with one 16-byte read per 32-row block.
// Every output sums its k in this fixed order; slices beyond K are zeros,
// so K needs no padding.
//
// Operand loads never branch per lane: a lane whose element is out of range
// loads from a block of zeros instead, so the compiler issues every load of
// a slice back to back and waits for them only at the LDS store.  (With
// per-lane guarded loads it waits for each load before issuing the next.)
constexpr int F32_BK = 32;

static __device__ __attribute__((aligned(16))) float k_f32_zeros[4];

template <int ROWS, bool KMAJ, int THREADS>
struct F32Stager {
  static constexpr int NV = ROWS * F32_BK / 4 / THREADS;  // float4 per thread
  static_assert(NV * 4 * THREADS == ROWS * F32_BK, "whole float4 per thread");
  static constexpr int LD = KMAJ ? F32_BK : ROWS + 8;     // LDS row (floats)
  static constexpr int FLOATS = KMAJ ? ROWS * F32_BK : F32_BK * LD;
  f32x4_t r[NV];

  // Rows row0.. (K-major: the operand's rows; otherwise its columns) at k0,
  // as runs of 4 elements along the contiguous dimension.  VEC: the operand
  // is 16-byte aligned, its leading dimension and its contiguous extent are
  // multiples of 4, so a run is wholly in or out of range.
  template <bool VEC>
  __device__ __forceinline__ void load(const float* P, uint64_t ld, uint64_t row0, uint64_t nrows,
                                       uint64_t k0, uint64_t K) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = threadIdx.x + i * THREADS;
      uint64_t gr, gc, rlim, clim;  // run row, first column, and their limits
      if constexpr (KMAJ) {
        gr = row0 + idx / 8; gc = k0 + (idx % 8) * 4; rlim = nrows; clim = K;
      } else {
        gr = k0 + idx / (ROWS / 4); gc = row0 + (idx % (ROWS / 4)) * 4; rlim = K; clim = nrows;
      }
      if constexpr (VEC) {
        const bool ok = gr < rlim && gc < clim;
        r[i] = *reinterpret_cast<const f32x4_t*>(ok ? P + (gr * ld + gc) : k_f32_zeros);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const bool ok = gr < rlim && gc + e < clim;
          r[i][e] = *(ok ? P + (gr * ld + gc + e) : k_f32_zeros);
        }
      }
    }
  }

  __device__ __forceinline__ void store(float* s) const {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int idx = threadIdx.x + i * THREADS;
      int off;
      if constexpr (KMAJ) {
        const int row = idx / 8, c = idx % 8;
        off = row * F32_BK + ((c ^ (row & 7)) << 2);
      } else {
        const int kr = idx / (ROWS / 4), c4 = idx % (ROWS / 4);
        off = kr * LD + c4 * 4;
      }
      *reinterpret_cast<f32x4_t*>(s + off) = r[i];
    }
  }

  // values of tile row `row` for the MFMAs t = 0..3 of k group g, half-wave h
  __device__ __forceinline__ static f32x4_t frag(const float* s, int row, int g, int h) {
    if constexpr (KMAJ) {
      const int c = 2 * g + h;
      return *reinterpret_cast<const f32x4_t*>(s + row * F32_BK + ((c ^ (row & 7)) << 2));
    } else {
      const float* q = s + (8 * g + 4 * h) * LD + row;
      return f32x4_t{q[0], q[LD], q[2 * LD], q[3 * LD]};
    }
  }
};

template <bool AK, bool BKM, bool VEC, int BM, int BN, int WGM, int WGN>
__global__ void __launch_bounds__(64 * WGM * WGN, 2)
k_gemm_f32(GemmArgs p, float* part, uint64_t kc) {
  constexpr int THREADS = 64 * WGM * WGN, WAVES = WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  static_assert(WM % 32 == 0 && WN % 32 == 0, "32 x 32 MFMA blocks per wave");
  constexpr int TM = WM / 32, TN = WN / 32;
  using SA = F32Stager<BM, AK, THREADS>;
  using SB = F32Stager<BN, BKM, THREADS>;
  constexpr int MAIN = 2 * (SA::FLOATS + SB::FLOATS), STAGE = WAVES * 32 * (WN + 4);
  constexpr int SMEM_FLOATS = MAIN > STAGE ? MAIN : STAGE;
  constexpr bool TST = 4 * STAGE + tt_bytes<WAVES, WN>() <= 4 * SMEM_FLOATS;
  __shared__ __attribute__((aligned(16))) float smem[SMEM_FLOATS];
  auto As = [&](int b) { return smem + b * SA::FLOATS; };
  auto Bs = [&](int b) { return smem + 2 * SA::FLOATS + b * SB::FLOATS; };

  uint32_t tm, tn;
  tile_coords(p, tm, tn);
  const uint64_t m0 = (uint64_t)tm * BM, n0 = (uint64_t)tn * BN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int l32 = lane & 31, lhi = lane >> 5;
  const int arow = wm * WM + l32, brow = wn * WN + l32;

  SA sa;
  SB sb;
  f32x16_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  // fragments of one k group, double-buffered: group g+1 is read while the
  // MFMAs of group g issue
  constexpr int G = F32_BK / 8;                                // k groups per slice
  constexpr int NM = 4 * TM * TN;                              // MFMAs per group
  constexpr int NDS = (AK ? TM : 2 * TM) + (BKM ? TN : 2 * TN);  // LDS reads per group
  constexpr int WPM = (SA::NV + SB::NV + NM - 1) / NM;         // LDS stores per MFMA
  static_assert(NDS <= NM, "fragment reads fit between the MFMAs of a group");
  f32x4_t fa[2][TM], fb[2][TN];
  auto frags = [&](int buf, int fi, int g) {
#pragma unroll
    for (int i = 0; i < TM; ++i) fa[fi][i] = SA::frag(As(buf), arow + i * 32, g, lhi);
#pragma unroll
    for (int j = 0; j < TN; ++j) fb[fi][j] = SB::frag(Bs(buf), brow + j * 32, g, lhi);
  };

  // k range of this block: all of k, or chunk blockIdx.y of kc (a multiple of
  // F32_BK, so the slices tile the chunk) when k is split
  const uint64_t K = p.K;
  const uint64_t kb = part ? (uint64_t)blockIdx.y * kc : 0;
  const uint64_t ke = part ? min(kb + kc, K) : K;
  const uint64_t nk = (ke - kb + F32_BK - 1) / F32_BK;
  sa.template load<VEC>(p.A, p.lda, m0, p.M, kb, K);
  sb.template load<VEC>(p.B, p.ldb, n0, p.N, kb, K);
  sa.store(As(0));
  sb.store(Bs(0));
  __syncthreads();
  frags(0, 0, 0);
  int cur = 0;
  for (uint64_t kt = 0; kt < nk; ++kt) {
    // the next slice (past K: zeros) into registers
    // (the prefetch after a chunk's last slice is stored but never used)
    sa.template load<VEC>(p.A, p.lda, m0, p.M, kb + (kt + 1) * F32_BK, K);
    sb.template load<VEC>(p.B, p.ldb, n0, p.N, kb + (kt + 1) * F32_BK, K);
    // keep the loads here: left to itself the scheduler sinks them to the
    // LDS store at the end of the slice, where their latency is exposed
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int c = g & 1;
      if (g + 1 < G) frags(cur, c ^ 1, g + 1);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[c][i][t], fb[c][j][t], acc[i][j], 0, 0, 0);
      // every group but the last alternates one MFMA with one fragment read
      // of the next group, so LDS latency hides behind matrix-core work
      if (g + 1 < G) {
#pragma unroll
        for (int q = 0; q < NDS; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
        }
        __builtin_amdgcn_sched_group_barrier(0x008, NM - NDS, 0);
      }
    }
    // the last group's MFMAs alternate with the LDS stores of the next slice
    sa.store(As(cur ^ 1));
    sb.store(Bs(cur ^ 1));
#pragma unroll
    for (int q = 0; q < NM; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);    // MFMA
      __builtin_amdgcn_sched_group_barrier(0x200, WPM, 0);  // DS write
    }
    __syncthreads();
    cur ^= 1;
    frags(cur, 0, 0);
  }
  __syncthreads();  // every wave is done with the operand tiles before staging
  staged_epilogue<TM, TN, WM, WN, TST>(smem, acc, p, part ? part + (uint64_t)blockIdx.y * p.M * p.N : nullptr,
                                       m0, n0, wm, wn, reinterpret_cast<uint16_t*>(smem + STAGE));
}

// Kernel configurations of the FP32 path (tile and wave grid).  They differ
// in how the output is tiled only: every output sums its k in the same order,
// so the choice never changes a result.
typedef void (*F32Launch)(dim3, hipStream_t, const GemmArgs&, bool ta, bool tb, bool vec, float* part,
                          uint64_t kc);
template <int BM, int BN, int WGM, int WGN>
static void launch_f32(dim3 grid, hipStream_t s, const GemmArgs& a, bool ta, bool tb, bool vec, float* part,
                       uint64_t kc) {
  const dim3 block(64 * WGM * WGN);
  // A is K-major unless transposed, B is K-major when transposed
#define VKH_F32_LAUNCH(AK, BKM)                                                                   \
  do {                                                                                            \
    if (vec)                                                                                      \
      hipLaunchKernelGGL((k_gemm_f32<AK, BKM, true, BM, BN, WGM, WGN>), grid, block, 0, s, a, part, kc); \
    else                                                                                          \
      hipLaunchKernelGGL((k_gemm_f32<AK, BKM, false, BM, BN, WGM, WGN>), grid, block, 0, s, a, part, kc); \
  } while (0)
  if (!ta && !tb) VKH_F32_LAUNCH(true, false);
  else if (!ta && tb) VKH_F32_LAUNCH(true, true);
  else if (ta && !tb) VKH_F32_LAUNCH(false, false);
  else VKH_F32_LAUNCH(false, true);
#undef VKH_F32_LAUNCH
}
struct F32Cfg {
  int bm, bn, waves, wm;  // wm: rows of a wave's slab (BM / WGM)
  F32Launch launch;
};
static const F32Cfg F32_CFGS[] = {
    {128, 128, 4, 64, launch_f32<128, 128, 2, 2>},
    {256, 128, 8, 64, launch_f32<256, 128, 4, 2>},
    {128, 256, 8, 64, launch_f32<128, 256, 2, 4>},
    {256, 256, 8, 128, launch_f32<256, 256, 2, 4>},
};
constexpr int F32_NCFG = sizeof(F32_CFGS) / sizeof(F32_CFGS[0]);

// Sums the S k-chunk partials of each output in chunk order and applies the
// GEMM epilogue (alpha, beta * C, bias, activation, aux derivative, BF16 copy).
__global__ void __launch_bounds__(256)
k_splitk_combine(const float* part, uint64_t S, GemmArgs e) {
  const uint64_t gpr = (e.N + 3) / 4, total = e.M * gpr, MN = e.M * e.N;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
    const uint64_t m = t / gpr, n = (t - m * gpr) * 4;
    const float* pp = part + m * e.N + n;
    float4 acc;
    if (e.N % 4 == 0) {
      acc = *reinterpret_cast<const float4*>(pp);
      for (uint64_t s = 1; s < S; ++s) {
        const float4 v = *reinterpret_cast<const float4*>(pp + s * MN);
        acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
      }
    } else {
      float a[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (n + q < e.N) {
          float x = pp[q];
          for (uint64_t s = 1; s < S; ++s) x += pp[s * MN + q];
          a[q] = x;
        }
      }
      acc = make_float4(a[0], a[1], a[2], a[3]);
    }
    epilogue4(e, m, n, acc);
  }
}

// BM x BN tile, WGM x WGN waves of (BM/WGM) x (BN/WGN), double-buffered LDS
// and register prefetch of the next BK-deep K slice; one barrier per slice.
// After the K loop the LDS holds each wave's accumulators, 32 rows at a time,
// so the epilogue reads and writes whole row segments of C.
template <int BM, int BN, int BK, bool SPLIT, int WGM, int WGN>
__global__ void __launch_bounds__(64 * WGM * WGN)
k_gemm_packed(PackedArgs p) {
  static_assert(PK_KALIGN % BK == 0, "Kp must be a multiple of the k slice");
  constexpr int THREADS = 64 * WGM * WGN, WAVES = WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  static_assert(WM % 32 == 0 && WN % 32 == 0 && WN <= 256, "32 x 32 MFMA blocks per wave");
  constexpr int TM = WM / 32, TN = WN / 32;
  constexpr int PARTS = SPLIT ? 2 : 1;
  constexpr int PK_LDK = BK + 8;
  constexpr int A_TILE = BM * PK_LDK, B_TILE = BN * PK_LDK;        // BF16 elements
  constexpr int MAIN_BYTES = 2 * PARTS * (A_TILE + B_TILE) * 2;
  constexpr int SROW = WN + 4;                                      // staging row (floats)
  constexpr int STAGE_BYTES = WAVES * 32 * SROW * 4;
  constexpr int SMEM_BYTES = MAIN_BYTES > STAGE_BYTES ? MAIN_BYTES : STAGE_BYTES;
  constexpr bool TST = STAGE_BYTES + tt_bytes<WAVES, WN>() <= SMEM_BYTES;
  __shared__ __attribute__((aligned(16))) uint16_t smem[SMEM_BYTES / 2];
  // operand tile (buffer, part): As(buf, q) / Bs(buf, q)
  auto As = [&](int buf, int q) { return smem + (buf * PARTS + q) * A_TILE; };
  auto Bs = [&](int buf, int q) { return smem + 2 * PARTS * A_TILE + (buf * PARTS + q) * B_TILE; };

  const GemmArgs& e = p.e;
  uint32_t tm, tn;
  tile_coords(e, tm, tn);
  const uint64_t m0 = (uint64_t)tm * BM, n0 = (uint64_t)tn * BN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int l32 = lane & 31, lhi = lane >> 5;
  const uint64_t Kp = p.Kp;

  PackedStager<BM, BK, SPLIT, THREADS> sa;
  PackedStager<BN, BK, SPLIT, THREADS> sb;

  f32x16_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  // Operand fragments of one 16-deep MFMA k-step, double-buffered in
  // registers: the fragments of step s+1 are read from LDS while the MFMAs
  // of step s issue, so LDS latency hides behind matrix-core work.
  constexpr int KS = BK / 16;
  static_assert(KS % 2 == 0, "the first k-step of every slice uses fragment buffer 0");
  constexpr int NM = TM * TN * (SPLIT ? 3 : 1);  // MFMAs per k-step
  constexpr int NDS = (TM + TN) * PARTS;         // fragment reads per k-step
  constexpr int NVM = (PackedStager<BM, BK, SPLIT, THREADS>::NV + PackedStager<BN, BK, SPLIT, THREADS>::NV) * PARTS;
  constexpr int VPP = (NVM + NDS - 1) / NDS;     // global loads per MFMA of the first step
  constexpr int WPM = (NVM + NM - 1) / NM;       // LDS stores per MFMA of the last step
  bf16x8_t fa[2][TM], fb[2][TN], fal[2][TM], fbl[2][TN];
  const int arow = wm * WM + l32, brow = wn * WN + l32, kcol = lhi * 8;
  auto frags = [&](int buf, int fb_idx, int k) {
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int off = (arow + i * 32) * PK_LDK + k + kcol;
      fa[fb_idx][i] = *reinterpret_cast<const bf16x8_t*>(As(buf, 0) + off);
      if constexpr (SPLIT) fal[fb_idx][i] = *reinterpret_cast<const bf16x8_t*>(As(buf, 1) + off);
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int off = (brow + j * 32) * PK_LDK + k + kcol;
      fb[fb_idx][j] = *reinterpret_cast<const bf16x8_t*>(Bs(buf, 0) + off);
      if constexpr (SPLIT) fbl[fb_idx][j] = *reinterpret_cast<const bf16x8_t*>(Bs(buf, 1) + off);
    }
  };

  const uint64_t kb = p.part ? (uint64_t)blockIdx.y * p.kc : 0;
  const uint64_t nk = ((p.part ? min(kb + p.kc, Kp) : Kp) - kb) / BK;
  sa.load(p.Ah, p.Al, m0, Kp, kb);
  sb.load(p.Bh, p.Bl, n0, Kp, kb);
  sa.store(As(0, 0), As(0, PARTS - 1));
  sb.store(Bs(0, 0), Bs(0, PARTS - 1));
  __syncthreads();
  frags(0, 0, 0);
  int cur = 0;
  for (uint64_t kt = 0; kt < nk; ++kt) {
    // Prefetch of the next slice into registers.  The final iteration loads
    // the last slice again (in bounds, never read): the body has no branch,
    // so the scheduler sees one region.
    const uint64_t kn = kb + (kt + 1 < nk ? kt + 1 : kt) * BK;
    sa.load(p.Ah, p.Al, m0, Kp, kn);
    sb.load(p.Bh, p.Bl, n0, Kp, kn);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int c = s & 1;
      if (s + 1 < KS) frags(cur, c ^ 1, (s + 1) * 16);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          if constexpr (SPLIT) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fal[c][i], fb[c][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[c][i], fbl[c][j], acc[i][j], 0, 0, 0);
          }
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[c][i], fb[c][j], acc[i][j], 0, 0, 0);
        }
      // Issue order within the slice: every k-step but the last alternates
      // one MFMA with one fragment read for the next step (the first step
      // also issues the next slice's global loads, VPP per pair, so that
      // they have the whole slice to arrive); the last step alternates its
      // MFMAs with the LDS stores of the prefetched slice.
      if (s + 1 < KS) {
#pragma unroll
        for (int q = 0; q < NDS; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
          if (s == 0) __builtin_amdgcn_sched_group_barrier(0x020, VPP, 0);  // VMEM read
        }
        if constexpr (NM > NDS) __builtin_amdgcn_sched_group_barrier(0x008, NM - NDS, 0);
      }
    }
    sa.store(As(cur ^ 1, 0), As(cur ^ 1, PARTS - 1));
    sb.store(Bs(cur ^ 1, 0), Bs(cur ^ 1, PARTS - 1));
#pragma unroll
    for (int q = 0; q < NM; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);    // MFMA
      __builtin_amdgcn_sched_group_barrier(0x200, WPM, 0);  // DS write
    }
    __syncthreads();
    cur ^= 1;
    if (kt + 1 < nk) frags(cur, 0, 0);
  }
  __syncthreads();  // every wave is done with the operand tiles before staging
  staged_epilogue<TM, TN, WM, WN, TST>(reinterpret_cast<float*>(smem), acc, e,
                                       p.part ? p.part + (uint64_t)blockIdx.y * e.M * e.N : nullptr,
                                       m0, n0, wm, wn, reinterpret_cast<uint16_t*>(smem) + STAGE_BYTES / 2);
}

// BF16 variant whose operand slices go from global memory straight into LDS
// (global_load_lds_dwordx4: no VGPR staging, no LDS store instructions).  A
// wave instruction fills 8 unpadded rows of 128 B (one BK = 64 slice of 8
// rows): lane l writes row l / 8, 16-byte position l % 8, and fetches the
// global chunk (l % 8) ^ (row & 7) of that row.  The XOR swizzle takes the
// place of the row padding: a fragment read of 8 consecutive rows at one
// chunk touches 8 distinct positions.  The next slice is requested at the
// start of a slice and awaited, with one barrier, at its end.  Every output
// accumulates k in the same order as k_gemm_packed.
constexpr int DIRECT_SMEM_BYTES = 128 * 1024;  // (with 144 KB, room for the transposed staging tile, 1-4% slower)

template <int BM, int BN, int WGM, int WGN>
__global__ void __launch_bounds__(64 * WGM * WGN)
k_gemm_direct(PackedArgs p) {
  constexpr int BK = 64, ROWB = BK * 2;  // a slice row: 64 BF16 = 128 B
  constexpr int WAVES = WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  static_assert(WM % 32 == 0 && WN % 32 == 0, "32 x 32 MFMA blocks per wave");
  static_assert(BM % (8 * WAVES) == 0 && BN % (8 * WAVES) == 0, "whole 8-row loads per wave");
  constexpr int TM = WM / 32, TN = WN / 32;
  constexpr int A_BYTES = BM * ROWB, B_BYTES = BN * ROWB;
  constexpr int MAIN_BYTES = 2 * (A_BYTES + B_BYTES);
  constexpr int SROW = WN + 4;
  constexpr int STAGE_BYTES = WAVES * 32 * SROW * 4;
  constexpr int SMEM_BYTES = MAIN_BYTES > STAGE_BYTES ? MAIN_BYTES : STAGE_BYTES;
  constexpr int A_LOADS = BM / 8 / WAVES, B_LOADS = BN / 8 / WAVES;  // per wave per slice
  // The array's size must not depend on the template parameters: with a
  // dependent size, the host pass of this template emits no launch stub once
  // the LDS-load builtin writes into the array.
  static_assert(SMEM_BYTES <= DIRECT_SMEM_BYTES, "LDS of the direct-load kernel");
  constexpr bool TST = STAGE_BYTES + tt_bytes<WAVES, WN>() <= DIRECT_SMEM_BYTES;
  __shared__ __attribute__((aligned(16))) char smem[DIRECT_SMEM_BYTES];
  auto Abuf = [&](int b) { return smem + b * A_BYTES; };
  auto Bbuf = [&](int b) { return smem + 2 * A_BYTES + b * B_BYTES; };

  const GemmArgs& e = p.e;
  uint32_t tm, tn;
  tile_coords(e, tm, tn);
  const uint64_t m0 = (uint64_t)tm * BM, n0 = (uint64_t)tn * BN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int l32 = lane & 31, lhi = lane >> 5;
  const int lrow = lane >> 3, lpos = lane & 7;
  const uint64_t Kp = p.Kp;

  auto load_slice = [&](int b, uint64_t k0) {
#pragma unroll
    for (int i = 0; i < A_LOADS; ++i) {
      const int r0 = (i * WAVES + wave) * 8, row = r0 + lrow;
      const uint16_t* src = p.Ah + (m0 + row) * Kp + k0 + ((lpos ^ (row & 7)) << 3);
      __builtin_amdgcn_global_load_lds(src, smem + b * A_BYTES + r0 * ROWB, 16, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < B_LOADS; ++i) {
      const int r0 = (i * WAVES + wave) * 8, row = r0 + lrow;
      const uint16_t* src = p.Bh + (n0 + row) * Kp + k0 + ((lpos ^ (row & 7)) << 3);
      __builtin_amdgcn_global_load_lds(src, smem + 2 * A_BYTES + b * B_BYTES + r0 * ROWB, 16, 0, 0);
    }
  };
  // fragment of tile row `row`, 16-byte chunk c (k = 8c .. 8c + 7 of the slice)
  auto frag = [&](const char* buf, int row, int c) {
    return *reinterpret_cast<const bf16x8_t*>(buf + row * ROWB + ((c ^ (row & 7)) << 4));
  };

  f32x16_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  constexpr int KS = BK / 16;
  constexpr int NM = TM * TN, NDS = TM + TN;
  bf16x8_t fa[2][TM], fb[2][TN];
  const int arow = wm * WM + l32, brow = wn * WN + l32;
  auto frags = [&](int buf, int fi, int s) {
#pragma unroll
    for (int i = 0; i < TM; ++i) fa[fi][i] = frag(Abuf(buf), arow + i * 32, 2 * s + lhi);
#pragma unroll
    for (int j = 0; j < TN; ++j) fb[fi][j] = frag(Bbuf(buf), brow + j * 32, 2 * s + lhi);
  };

  const uint64_t kb = p.part ? (uint64_t)blockIdx.y * p.kc : 0;
  const uint64_t nk = ((p.part ? min(kb + p.kc, Kp) : Kp) - kb) / BK;
  load_slice(0, kb);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  frags(0, 0, 0);
  int cur = 0;
  for (uint64_t kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) load_slice(cur ^ 1, kb + (kt + 1) * BK);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int c = s & 1;
      if (s + 1 < KS) frags(cur, c ^ 1, s + 1);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[c][i], fb[c][j], acc[i][j], 0, 0, 0);
      if (s + 1 < KS) {
#pragma unroll
        for (int q = 0; q < NDS; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
        }
        if constexpr (NM > NDS) __builtin_amdgcn_sched_group_barrier(0x008, NM - NDS, 0);
      }
    }
    // the next slice has landed in LDS and every wave is done with this one
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    cur ^= 1;
    if (kt + 1 < nk) frags(cur, 0, 0);
  }
  __syncthreads();  // every wave is done with the operand tiles before staging
  staged_epilogue<TM, TN, WM, WN, TST>(reinterpret_cast<float*>(smem), acc, e,
                                       p.part ? p.part + (uint64_t)blockIdx.y * e.M * e.N : nullptr,
                                       m0, n0, wm, wn, reinterpret_cast<uint16_t*>(smem) + STAGE_BYTES / 2);
}

// Tile choice of the packed kernel: 256 x 256 when both sides fill at least
// one such tile (4 waves of 128 x 128), 128 x 128 otherwise and in split mode
// (whose two BF16 parts double the LDS footprint).
// Kernel configurations of the packed path: tile, k slice and wave grid.
// BF16 keeps one part per operand; split mode (two parts) needs twice the
// LDS per slice, hence its smaller k slice.
typedef void (*PackedLaunch)(dim3, hipStream_t, const PackedArgs&);
template <int BM, int BN, int BK, bool SPLIT, int WGM, int WGN>
static void launch_packed(dim3 grid, hipStream_t s, const PackedArgs& pa) {
  hipLaunchKernelGGL((k_gemm_packed<BM, BN, BK, SPLIT, WGM, WGN>), grid, dim3(64 * WGM * WGN), 0, s, pa);
}
template <int BM, int BN, int WGM, int WGN>
static void launch_direct(dim3 grid, hipStream_t s, const PackedArgs& pa) {
  hipLaunchKernelGGL((k_gemm_direct<BM, BN, WGM, WGN>), grid, dim3(64 * WGM * WGN), 0, s, pa);
}
struct PackedCfg {
  int bm, bn, bk, waves, wm;  // wm: rows of a wave's slab (BM / WGM)
  PackedLaunch launch;
  int direct = 0;  // operand slices loaded from global memory straight into LDS
};
static const PackedCfg BF16_CFGS[] = {
    {256, 256, 64, 4, 128, launch_packed<256, 256, 64, false, 2, 2>},
    {256, 256, 64, 8, 128, launch_packed<256, 256, 64, false, 2, 4>},
    {128, 128, 64, 4, 64, launch_packed<128, 128, 64, false, 2, 2>},
    {256, 128, 64, 4, 128, launch_packed<256, 128, 64, false, 2, 2>},
    {128, 256, 64, 4, 64, launch_packed<128, 256, 64, false, 2, 2>},
    {256, 256, 64, 8, 128, launch_direct<256, 256, 2, 4>, 1},
    // 8 waves of 64 x 32: twice the waves of configuration 2 in the same
    // LDS, for short k, where the operand latency of each tile's few slices
    // is exposed (16384 x 4096 x 256 with its BF16 row and transposed
    // copies: 109 against 143-151 us)
    {128, 128, 64, 8, 64, launch_packed<128, 128, 64, false, 2, 4>},
    // (128 x 256 with 8 waves of 64 x 64: 110 KB of LDS, one block a
    // compute unit, fastest on none of the MLP shapes, so not kept;
    // direct loads with 4 waves of 128 x 128, a third less LDS fragment
    // traffic, one wave per SIMD: 894 against 977 TFLOP/s at 8192^3 and
    // slower on every MLP shape, so not kept)
};
static const PackedCfg SPLIT_CFGS[] = {
    {128, 128, 32, 4, 64, launch_packed<128, 128, 32, true, 2, 2>},
    {256, 128, 32, 8, 64, launch_packed<256, 128, 32, true, 4, 2>},
    {128, 128, 32, 8, 64, launch_packed<128, 128, 32, true, 2, 4>},
};
constexpr int PK_MAXTILE = 256;  // largest bm / bn of any configuration

static const PackedCfg* packed_cfgs(int mode, int& count) {
  if (mode == VKH_GEMM_BF16) { count = (int)(sizeof(BF16_CFGS) / sizeof(BF16_CFGS[0])); return BF16_CFGS; }
  count = (int)(sizeof(SPLIT_CFGS) / sizeof(SPLIT_CFGS[0]));
  return SPLIT_CFGS;
}

// Configuration without tuning, from measurements on MI355X: BF16 uses the
// 8-wave 256 x 256 tile when both sides fill it (8192^3: 1.24 ms against
// 1.43 ms for 4 waves; 65536 x 1024 x 1024: 0.294 against 0.348 ms), split
// mode the 8-wave 256 x 128 tile for long K (8192^3: 3.55 against 4.18 ms)
// and 128 x 128 otherwise (K = 1024: 0.59 against 0.74 ms).
// FP32 takes the 256 x 256 tile when at least 512 such tiles cover C (8192^3:
// 141 against 126-135 TFLOP/s for the other tiles; 65536 x 1024 x 1024: 124
// against 100-113), the 128 x 128 tile otherwise (4096 x 1024 x 784: 92
// against 31 TFLOP/s for 256 x 256, which leaves most CUs idle).
static int default_cfg(int mode, uint64_t M, uint64_t N, uint64_t K) {
  if (mode == VKH_GEMM_FP32) return ceil_div(M, (uint64_t)256) * ceil_div(N, (uint64_t)256) >= 512 ? 3 : 0;
  if (mode == VKH_GEMM_BF16) return (M >= 256 && N >= 256) ? 1 : 2;
  return (M >= 256 && N >= 128 && K >= 4096) ? 1 : 0;
}

// number of kernel configurations of a mode
static int mode_cfgs(int mode) {
  if (mode == VKH_GEMM_FP32) return F32_NCFG;
  int n = 0;
  packed_cfgs(mode, n);
  return n;
}

static std::array<uint64_t, 6> cfg_key(int mode, int ta, int tb, uint64_t M, uint64_t N, uint64_t K) {
  return {(uint64_t)mode, (uint64_t)(ta != 0), (uint64_t)(tb != 0), M, N, K};
}

// Fills p[0..n) with values in [-0.5, 0.5) from a hash of the index (operands
// of the autotuning runs: data-dependent power draw makes zero operands an
// unrepresentative benchmark).
__global__ void __launch_bounds__(256) k_fill_noise(float* p, uint64_t n, uint32_t seed) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    uint32_t h = (uint32_t)i * 0x9e3779b1u ^ (uint32_t)(i >> 32) ^ seed;
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    p[i] = (float)(h >> 8) * (1.0f / 16777216.0f) - 0.5f;
  }
}

// Split k.  A product with few output tiles and a long k leaves most CUs
// idle (4096 x 256 x 16384: 16 tiles of 256 x 256), so k is split into S
// chunks that run as separate blocks; k_splitk_combine sums the chunk
// partials in chunk order.  S and the chunk bounds depend on (M, N, K) only,
// never on the kernel configuration, so every configuration accumulates each
// chunk in the same order and configurations still never change results.
constexpr uint64_t SPLITK_MIN_KP = 4096;    // shorter k: no split
constexpr uint64_t SPLITK_MIN_CHUNK = 1024;  // k per chunk at least
constexpr uint64_t SPLITK_REF_BLOCKS = 256;  // blocks to aim for (one per CU of MI355X)

static uint64_t splitk_chunk(uint64_t M, uint64_t N, uint64_t Kp) {
  const uint64_t tiles = ceil_div(M, 256) * ceil_div(N, 256);
  if (Kp < SPLITK_MIN_KP || tiles * 2 > SPLITK_REF_BLOCKS) return Kp;
  const uint64_t S = std::min(ceil_div(SPLITK_REF_BLOCKS, tiles), Kp / SPLITK_MIN_CHUNK);
  if (S < 2) return Kp;
  return ceil_div(Kp / PK_KALIGN, S) * PK_KALIGN;
}

static size_t operand_bytes(int mode, uint64_t M, uint64_t N, uint64_t Kp) {
  const uint64_t Mp = ceil_div(M, PK_MAXTILE) * PK_MAXTILE, Np = ceil_div(N, PK_MAXTILE) * PK_MAXTILE;
  const uint64_t parts = mode == VKH_GEMM_BF16X3 ? 2 : 1;
  return (size_t)ceil_div((Mp + Np) * Kp * 2 * parts, 256) * 256;
}

// Packing area of an M x N x K product: operands with rows padded to
// PK_MAXTILE, which covers every configuration (so a reservation never
// depends on tuning), then the split-k partials [S][M][N] when k is split.
static size_t packed_bytes(int mode, uint64_t M, uint64_t N, uint64_t K) {
  const uint64_t Kp = ceil_div(std::max<uint64_t>(K, 1), PK_KALIGN) * PK_KALIGN;
  const uint64_t kc = splitk_chunk(M, N, Kp);
  const uint64_t S = ceil_div(Kp, kc);
  return operand_bytes(mode, M, N, Kp) + (S > 1 ? (size_t)(S * M * N * 4) : 0);
}

// Packing area.  Eager calls share a per-context grow-only area (stream order
// keeps uses apart).  A captured GEMM uses its graph's packing region, which
// the graph owns for as long as it can be replayed; it is sized when the
// capture begins from what vkh_gemm_reserve recorded, and a capture that
// needs more fails and records the size for the next capture.
static uint16_t* pack_area(vkh_gpu* g, size_t bytes) {
  if (g->capturing) {
    if (bytes <= g->graph_pack_bytes) return (uint16_t*)g->graph_pack;
    g->graph_pack_hint = std::max(g->graph_pack_hint, bytes);
    fail(VKH_OUT_OF_MEMORY, "graph GEMM packing region too small (" + std::to_string(g->graph_pack_bytes) +
                                " < " + std::to_string(bytes) +
                                " bytes); call vkh_gemm_reserve before the capture, or capture again");
  }
  if (bytes <= g->pack_bytes) return (uint16_t*)g->pack;
  VKH_REQUIRE(g->alloc, "GEMM packing needs a context with its own allocator");
  if (g->pack) g->alloc->release(g->pack);
  g->pack = nullptr;
  g->pack_bytes = 0;
  const size_t want = bytes + bytes / 4;
  g->pack = g->alloc->allocate(want);
  g->pack_bytes = want;
  return (uint16_t*)g->pack;
}

template <bool SPLIT>
static void pack_operand(vkh_gpu* g, const float* src, uint64_t ld, bool trans, uint64_t R,
                         uint64_t K, uint16_t* hi, uint16_t* lo, uint64_t Rp, uint64_t Kp) {
  if (!trans) {
    const int vec = ((uintptr_t)src & 15u) == 0 && ld % 4 == 0;
    const uint64_t total4 = Rp * (Kp / 4);
    VKH_REQUIRE(total4 < (uint64_t(1) << 31), "gemm pack: operand too large");
    hipLaunchKernelGGL((k_pack_rows<SPLIT>), dim3((unsigned)ceil_div(total4, 256)), dim3(256), 0, g->stream,
                       src, ld, R, K, hi, lo, (uint32_t)total4, FastDiv((uint32_t)(Kp / 4)), Kp, vec);
  } else {
    const int vec = ((uintptr_t)src & 15u) == 0 && ld % 4 == 0;
    VKH_REQUIRE(Rp % PK_TR == 0, "gemm pack: padded rows not a multiple of the tile");
    dim3 grid((unsigned)(Rp / PK_TR), (unsigned)ceil_div(Kp, PK_TK));
    hipLaunchKernelGGL((k_pack_cols<SPLIT>), grid, dim3(256), 0, g->stream, src, ld, R, K, hi, lo,
                       Rp, Kp, vec);
  }
  check_launch("gemm pack");
}

// a_pre / b_pre (BF16 mode, K % PK_KALIGN == 0): op(A) / op(B)^T already in
// the packed layout -- BF16 rows of K elements, M / N rounded up to
// PK_MAXTILE rows -- so that operand's packing pass is skipped (the bits are
// the ones it would write)
static void gemm_packed(vkh_gpu* g, int mode, bool ta, bool tb, GemmArgs& a, int cfg,
                        const uint16_t* a_pre = nullptr, const uint16_t* b_pre = nullptr) {
  const bool split = mode == VKH_GEMM_BF16X3;
  int ncfg = 0;
  const PackedCfg* cfgs = packed_cfgs(mode, ncfg);
  VKH_REQUIRE(cfg >= 0 && cfg < ncfg, "gemm: configuration index out of range");
  const PackedCfg& s = cfgs[cfg];
  const uint64_t Mp = ceil_div(a.M, s.bm) * s.bm, Np = ceil_div(a.N, s.bn) * s.bn;
  const uint64_t Kp = ceil_div(std::max<uint64_t>(a.K, 1), PK_KALIGN) * PK_KALIGN;
  uint16_t* base = pack_area(g, packed_bytes(mode, a.M, a.N, a.K));
  const uint16_t* Ah = base;
  uint16_t* Al = split ? base + Mp * Kp : nullptr;
  uint16_t* Bp = split ? Al + Mp * Kp : base + Mp * Kp;
  const uint16_t* Bh = Bp;
  uint16_t* Bl = split ? Bp + Np * Kp : nullptr;
  // op(A)[m][k] and op(B)^T[n][k] = op(B)[k][n]
  if (split) {
    pack_operand<true>(g, a.A, a.lda, ta, a.M, a.K, base, Al, Mp, Kp);
    pack_operand<true>(g, a.B, a.ldb, !tb, a.N, a.K, Bp, Bl, Np, Kp);
  } else {
    if (!a_pre) pack_operand<false>(g, a.A, a.lda, ta, a.M, a.K, base, nullptr, Mp, Kp);
    if (!b_pre) pack_operand<false>(g, a.B, a.ldb, !tb, a.N, a.K, Bp, nullptr, Np, Kp);
  }
  if (a_pre) Ah = a_pre;
  if (b_pre) Bh = b_pre;
  a.tiles_m = (uint32_t)(Mp / s.bm);
  a.tiles_n = (uint32_t)(Np / s.bn);
  const uint64_t nb = (uint64_t)a.tiles_m * a.tiles_n;
  VKH_REQUIRE(nb < (uint64_t(1) << 31), "GEMM too large");
  const uint64_t kc = splitk_chunk(a.M, a.N, Kp);
  const uint64_t S = ceil_div(Kp, kc);
  float* part = nullptr;
  if (S > 1) part = reinterpret_cast<float*>(reinterpret_cast<char*>(base) + operand_bytes(mode, a.M, a.N, Kp));
  PackedArgs pa{Ah, Al, Bh, Bl, Kp, kc, part, a};
  s.launch(dim3((unsigned)nb, (unsigned)S), g->stream, pa);
  check_launch("gemm (packed)");
  if (S > 1) {
    const uint64_t groups = a.M * ceil_div(a.N, 4);
    hipLaunchKernelGGL(k_splitk_combine, dim3(stream_grid(g, groups, 256, 8)), dim3(256), 0, g->stream,
                       (const float*)part, S, a);
    check_launch("gemm split-k combine");
  }
}

// FP32 split k: as above, for fewer than 256 tiles of 128 x 128 (the
// smallest FP32 tile, 2 blocks per CU) and K >= 1024, in chunks of at least
// 256 that are multiples of the 32-deep slice, aiming at 512 blocks.  The
// chunk bounds depend on (M, N, K) only.  4096 x 256 x 16384 (a weight
// gradient over a batch of 16384): 8 chunks, 31 -> 121-126 TFLOP/s.
constexpr uint64_t F32_SPLITK_MIN_K = 1024, F32_SPLITK_MIN_CHUNK = 256, F32_SPLITK_BLOCKS = 512;

static uint64_t f32_splitk_chunk(uint64_t M, uint64_t N, uint64_t K) {
  const uint64_t tiles = ceil_div(M, 128) * ceil_div(N, 128);
  if (K < F32_SPLITK_MIN_K || tiles * 2 > F32_SPLITK_BLOCKS) return K;
  const uint64_t S = std::min(ceil_div(F32_SPLITK_BLOCKS, tiles), K / F32_SPLITK_MIN_CHUNK);
  if (S < 2) return K;
  return ceil_div(ceil_div(K, (uint64_t)F32_BK), S) * F32_BK;
}

// split-k partials [S][M][N] of an FP32 product (0 when k is not split)
static size_t f32_split_bytes(uint64_t M, uint64_t N, uint64_t K) {
  const uint64_t S = ceil_div(K, f32_splitk_chunk(M, N, K));
  return S > 1 ? (size_t)(S * M * N * 4) : 0;
}

static void gemm_f32(vkh_gpu* g, bool ta, bool tb, GemmArgs& a, int cfg) {
  VKH_REQUIRE(cfg >= 0 && cfg < F32_NCFG, "gemm: configuration index out of range");
  const F32Cfg& c = F32_CFGS[cfg];
  a.tiles_m = (uint32_t)ceil_div(a.M, (uint64_t)c.bm);
  a.tiles_n = (uint32_t)ceil_div(a.N, (uint64_t)c.bn);
  const uint64_t nb = (uint64_t)a.tiles_m * a.tiles_n;
  VKH_REQUIRE(nb < (uint64_t(1) << 31), "GEMM too large");
  // 16-byte runs: aligned operands whose leading dimension and contiguous
  // extent (A: K unless transposed, else M; B: N unless transposed, else K)
  // are multiples of 4
  const bool vec = ((uintptr_t)a.A & 15u) == 0 && a.lda % 4 == 0 && (ta ? a.M : a.K) % 4 == 0 &&
                   ((uintptr_t)a.B & 15u) == 0 && a.ldb % 4 == 0 && (tb ? a.K : a.N) % 4 == 0;
  const uint64_t kc = f32_splitk_chunk(a.M, a.N, a.K);
  const uint64_t S = ceil_div(a.K, kc);
  float* part = S > 1 ? reinterpret_cast<float*>(pack_area(g, f32_split_bytes(a.M, a.N, a.K))) : nullptr;
  c.launch(dim3((unsigned)nb, (unsigned)S), g->stream, a, ta, tb, vec, part, kc);
  check_launch("gemm");
  if (S > 1) {
    const uint64_t groups = a.M * ceil_div(a.N, 4);
    hipLaunchKernelGGL(k_splitk_combine, dim3(stream_grid(g, groups, 256, 8)), dim3(256), 0, g->stream,
                       (const float*)part, S, a);
    check_launch("gemm split-k combine");
  }
}

// ---- skinny products: min(M, N) <= 16 ------------------------------------------------
// C = op(A) op(B) with one small output dimension Q is a pass over the large
// operand (the Dense layer with one or ten outputs, its weight gradient):
// memory-bound, so it runs in FP32 on the vector units in every mode.  The
// problem is restated as out[p, q] = sum_k X[p, k] Y[k, q] with P the long
// output dimension; X is read once, Y (Q x K) stays in the caches.
//   row-dot      X K-contiguous: one wave per p, lanes along k, wave sums
//   column-walk  X p-contiguous: one thread per p, k chunks, partial sums
//                combined in chunk order by a second kernel
struct SkinnyArgs {
  const float* X;
  const float* Y;
  uint64_t sxp, sxk, syk, syq, P, K;
  int Q;
  int swap;  // 1: p indexes columns n of C and q rows m
  GemmArgs e;
};

__device__ __forceinline__ void skinny_store(const SkinnyArgs& a, uint64_t p, int q, float acc) {
  const GemmArgs& e = a.e;
  const uint64_t m = a.swap ? (uint64_t)q : p, n = a.swap ? p : (uint64_t)q;
  float* cp = e.C + m * e.ldc + n;
  float v = e.alpha * acc;
  if (e.beta != 0.f) v += e.beta * *cp;
  if (e.bias) v += e.bias[n];
  v = act_apply(e.act, v);
  if (e.aux) v = act_deriv_mul(e.aux_act, e.aux[m * e.ldaux + n], v);
  if (e.aux16) v = act_deriv_mul(e.aux_act, bf2f(e.aux16[m * e.ldaux + n]), v);
  if (!e.no_c) *cp = v;
  if (e.Cbf) e.Cbf[m * e.ldc + n] = f2bf(v);
  if (e.Cbft) e.Cbft[n * e.ldt + m] = f2bf(v);
}

__device__ __forceinline__ float wave_sum64(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int QB>
__global__ void __launch_bounds__(256) k_skinny_rowdot(SkinnyArgs a, int vec) {
  const uint64_t p = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= a.P) return;  // uniform per wave
  float acc[QB];
#pragma unroll
  for (int q = 0; q < QB; ++q) acc[q] = 0.f;
  const float* xr = a.X + p * a.sxp;
  if (vec) {  // X rows and Y columns 16-byte aligned, K % 4 == 0, Y K-contiguous
    for (uint64_t k = (uint64_t)lane * 4; k < a.K; k += 256) {
      const float4 x = *reinterpret_cast<const float4*>(xr + k);
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        if (q < a.Q) {
          const float4 y = *reinterpret_cast<const float4*>(a.Y + q * a.syq + k);
          acc[q] += x.x * y.x;
          acc[q] += x.y * y.y;
          acc[q] += x.z * y.z;
          acc[q] += x.w * y.w;
        }
      }
    }
  } else {
    for (uint64_t k = lane; k < a.K; k += 64) {
      const float x = xr[k];
#pragma unroll
      for (int q = 0; q < QB; ++q)
        if (q < a.Q) acc[q] += x * a.Y[k * a.syk + q * a.syq];
    }
  }
  float mine = 0.f;
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const float s = wave_sum64(acc[q]);
    if (lane == q) mine = s;
  }
  if (lane < a.Q) skinny_store(a, p, lane, mine);
}

// k_skinny_rowdot for Q = 1 with 16-byte operands and K = 256 * NK: the
// column of op(B) (K floats) is loaded into registers once per wave and a
// wave takes RW rows, so that only X is streamed; every row sums its k in the
// order of k_skinny_rowdot (lane-strided float4 steps, then the wave's
// butterfly)
constexpr int ROWDOT_RW = 4;
template <int NK>
__global__ void __launch_bounds__(256) k_skinny_rowdot1(SkinnyArgs a) {
  const int lane = threadIdx.x & 63;
  const uint64_t p0 = ((uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * ROWDOT_RW;
  if (p0 >= a.P) return;  // uniform per wave
  float4 y[NK];
#pragma unroll
  for (int i = 0; i < NK; ++i) y[i] = *reinterpret_cast<const float4*>(a.Y + lane * 4 + i * 256);
  for (int rw = 0; rw < ROWDOT_RW; ++rw) {
    const uint64_t p = p0 + rw;
    if (p >= a.P) break;
    const float* xr = a.X + p * a.sxp;
    float4 x[NK];
#pragma unroll
    for (int i = 0; i < NK; ++i) x[i] = ld_stream(reinterpret_cast<const float4*>(xr + lane * 4 + i * 256));
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < NK; ++i) {
      acc += x[i].x * y[i].x;
      acc += x[i].y * y[i].y;
      acc += x[i].z * y[i].z;
      acc += x[i].w * y[i].w;
    }
    const float s = wave_sum64(acc);
    if (lane == 0) skinny_store(a, p, 0, s);
  }
}

// partial sums part[chunk][q][p] over k in [chunk*kchunk, (chunk+1)*kchunk);
// VW consecutive p per thread (VW = 4: 16-byte loads, P % 4 == 0 and X rows
// 16-byte aligned)
template <int QB, int VW>
__global__ void __launch_bounds__(256)
k_skinny_colwalk(SkinnyArgs a, uint64_t kchunk, float* part) {
  const uint64_t p = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * VW;
  const uint64_t chunk = blockIdx.y;
  if (p >= a.P) return;
  const uint64_t k0 = chunk * kchunk, k1 = min(a.K, k0 + kchunk);
  float acc[QB][VW];
#pragma unroll
  for (int q = 0; q < QB; ++q)
#pragma unroll
    for (int w = 0; w < VW; ++w) acc[q][w] = 0.f;
  for (uint64_t k = k0; k < k1; ++k) {
    float x[VW];
    if constexpr (VW == 4) {
      const float4 t = *reinterpret_cast<const float4*>(a.X + k * a.sxk + p);
      x[0] = t.x; x[1] = t.y; x[2] = t.z; x[3] = t.w;
    } else {
      x[0] = a.X[k * a.sxk + p];
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      if (q < a.Q) {
        const float y = a.Y[k * a.syk + q * a.syq];
#pragma unroll
        for (int w = 0; w < VW; ++w) acc[q][w] += x[w] * y;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < QB; ++q)
    if (q < a.Q)
#pragma unroll
      for (int w = 0; w < VW; ++w) part[(chunk * a.Q + q) * a.P + p + w] = acc[q][w];
}

__global__ void __launch_bounds__(256)
k_skinny_combine(SkinnyArgs a, const float* part, uint64_t nchunk) {
  const uint64_t total = a.P * (uint64_t)a.Q;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
    const int q = (int)(t / a.P);
    const uint64_t p = t - (uint64_t)q * a.P;
    const uint64_t step = (uint64_t)a.Q * a.P;
    const float* pp = part + (uint64_t)q * a.P + p;
    float s = 0.f;
    uint64_t c = 0;
    for (; c + 8 <= nchunk; c += 8) {  // eight loads in flight, summed in chunk order
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = pp[(c + u) * step];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; c < nchunk; ++c) s += pp[c * step];
    skinny_store(a, p, q, s);
  }
}

constexpr uint64_t SKINNY_MAX = 16;  // largest small dimension of the skinny path

// k per chunk (and the chunk count) of the column walk over P outputs and K
// terms: about 1024 threads per compute unit, at least 64 k per chunk; v4:
// four outputs per thread.  The chunks fix the summation order.
static uint64_t colwalk_chunk(vkh_gpu* g, uint64_t P, uint64_t K, bool v4, uint64_t& nchunk_out) {
  const uint64_t pthreads = v4 ? P / 4 : P;
  const uint64_t target = (uint64_t)(g->num_cus > 0 ? g->num_cus : 256) * 1024;
  uint64_t nchunk = std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(target, pthreads), ceil_div(K, 64)));
  nchunk = std::min<uint64_t>(nchunk, 65535);
  const uint64_t kchunk = ceil_div(std::max<uint64_t>(K, 1), nchunk);
  nchunk_out = ceil_div(std::max<uint64_t>(K, 1), kchunk);
  return kchunk;
}

static void gemm_skinny(vkh_gpu* g, bool ta, bool tb, const GemmArgs& e) {
  SkinnyArgs a{};
  a.e = e;
  if (e.N <= e.M) {  // q = n, p = m: X = op(A), Y = op(B)
    a.swap = 0;
    a.P = e.M;
    a.Q = (int)e.N;
    a.X = e.A;
    a.sxp = ta ? 1 : e.lda;
    a.sxk = ta ? e.lda : 1;
    a.Y = e.B;
    a.syk = tb ? 1 : e.ldb;
    a.syq = tb ? e.ldb : 1;
  } else {  // q = m, p = n: X = op(B)^T, Y = op(A)^T
    a.swap = 1;
    a.P = e.N;
    a.Q = (int)e.M;
    a.X = e.B;
    a.sxp = tb ? e.ldb : 1;
    a.sxk = tb ? 1 : e.ldb;
    a.Y = e.A;
    a.syk = ta ? e.lda : 1;
    a.syq = ta ? 1 : e.lda;
  }
  a.K = e.K;
  const int qb = a.Q <= 1 ? 1 : a.Q <= 2 ? 2 : a.Q <= 4 ? 4 : a.Q <= 8 ? 8 : 16;
  if (a.sxk == 1) {
    const int vec = ((uintptr_t)a.X & 15u) == 0 && a.sxp % 4 == 0 && a.K % 4 == 0 && a.syk == 1 &&
                    ((uintptr_t)a.Y & 15u) == 0 && a.syq % 4 == 0;
    if (qb == 1 && vec && a.K % 256 == 0 && a.K >= 1024 && a.K <= 4096) {
      const unsigned blocks1 = (unsigned)ceil_div(a.P, (uint64_t)4 * ROWDOT_RW);
      switch (a.K / 256) {
#define VKH_ROWDOT1(NKV) case NKV: hipLaunchKernelGGL(k_skinny_rowdot1<NKV>, dim3(blocks1), dim3(256), 0, g->stream, a); break;
        VKH_ROWDOT1(4) VKH_ROWDOT1(5) VKH_ROWDOT1(6) VKH_ROWDOT1(7) VKH_ROWDOT1(8) VKH_ROWDOT1(9) VKH_ROWDOT1(10)
        VKH_ROWDOT1(11) VKH_ROWDOT1(12) VKH_ROWDOT1(13) VKH_ROWDOT1(14) VKH_ROWDOT1(15) VKH_ROWDOT1(16)
#undef VKH_ROWDOT1
      }
      check_launch("gemm (skinny row-dot)");
      return;
    }
    const uint64_t blocks = ceil_div(a.P, 4);
    VKH_REQUIRE(blocks < (uint64_t(1) << 31), "GEMM too large");
    dim3 grid((unsigned)blocks), block(256);
    switch (qb) {
      case 1: hipLaunchKernelGGL(k_skinny_rowdot<1>, grid, block, 0, g->stream, a, vec); break;
      case 2: hipLaunchKernelGGL(k_skinny_rowdot<2>, grid, block, 0, g->stream, a, vec); break;
      case 4: hipLaunchKernelGGL(k_skinny_rowdot<4>, grid, block, 0, g->stream, a, vec); break;
      case 8: hipLaunchKernelGGL(k_skinny_rowdot<8>, grid, block, 0, g->stream, a, vec); break;
      default: hipLaunchKernelGGL(k_skinny_rowdot<16>, grid, block, 0, g->stream, a, vec); break;
    }
    check_launch("gemm (skinny row-dot)");
    return;
  }
  const bool v4 = a.P % 4 == 0 && ((uintptr_t)a.X & 15u) == 0 && a.sxk % 4 == 0;
  const uint64_t pthreads = v4 ? a.P / 4 : a.P;
  uint64_t nchunk = 0;
  const uint64_t kchunk = colwalk_chunk(g, a.P, a.K, v4, nchunk);
  Workspace ws(g, sizeof(float) * nchunk * (uint64_t)a.Q * a.P);
  float* part = (float*)ws.get();
  const uint64_t bx = ceil_div(pthreads, 256);
  VKH_REQUIRE(bx < (uint64_t(1) << 31), "GEMM too large");
  dim3 grid((unsigned)bx, (unsigned)nchunk), block(256);
#define VKH_COLWALK(QB)                                                                                 \
  do {                                                                                                  \
    if (v4) hipLaunchKernelGGL((k_skinny_colwalk<QB, 4>), grid, block, 0, g->stream, a, kchunk, part); \
    else hipLaunchKernelGGL((k_skinny_colwalk<QB, 1>), grid, block, 0, g->stream, a, kchunk, part);    \
  } while (0)
  switch (qb) {
    case 1: VKH_COLWALK(1); break;
    case 2: VKH_COLWALK(2); break;
    case 4: VKH_COLWALK(4); break;
    case 8: VKH_COLWALK(8); break;
    default: VKH_COLWALK(16); break;
  }
#undef VKH_COLWALK
  check_launch("gemm (skinny column walk)");
  const uint64_t total = a.P * (uint64_t)a.Q;
  hipLaunchKernelGGL(k_skinny_combine, dim3(full_grid(total, 256)), dim3(256), 0, g->stream, a,
                     (const float*)part, nchunk);
  check_launch("gemm (skinny combine)");
}

// ---- helpers for the nn path -------------------------------------------------------------
__global__ void __launch_bounds__(256) k_to_bf16(const float* src, uint16_t* dst, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = f2bf(src[i]);
}

}  // namespace vkh

using namespace vkh;

extern "C" {

// ---- products with a short k: K <= SMALLK_MAX --------------------------------------
// C = op(A) op(B) with K of a few elements (the input gradient of a Dense
// layer with one or ten outputs) is a pass that writes C (and reads aux or
// beta * C): memory-bound, so like the skinny products it runs in FP32 on the
// vector units in every mode.  Thread (row m, columns n..n+3) sums
// op(A)[m, k] op(B)[k, n + q] over k in order and applies the GEMM epilogue.
constexpr uint64_t SMALLK_MAX = 16;

__global__ void __launch_bounds__(256)
k_gemm_smallk(GemmArgs e, uint64_t sam, uint64_t sak, uint64_t sbk, uint64_t sbn, int vb) {
  const uint64_t n = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 4;
  if (n >= e.N) return;
  const int K = (int)e.K;
  for (uint64_t m = blockIdx.y; m < e.M; m += gridDim.y) {
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < K; ++k) {
      const float av = e.A[m * sam + k * sak];
      float bv[4];
      if (vb && n + 3 < e.N) {
        const float4 t = *reinterpret_cast<const float4*>(e.B + k * sbk + n);
        bv[0] = t.x; bv[1] = t.y; bv[2] = t.z; bv[3] = t.w;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) bv[q] = n + q < e.N ? e.B[k * sbk + (n + q) * sbn] : 0.f;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = k == 0 ? av * bv[q] : acc[q] + av * bv[q];
    }
    epilogue4(e, m, n, make_float4(acc[0], acc[1], acc[2], acc[3]));
  }
}

static void gemm_smallk(vkh_gpu* g, bool ta, bool tb, const GemmArgs& a) {
  const uint64_t sam = ta ? 1 : a.lda, sak = ta ? a.lda : 1;
  const uint64_t sbk = tb ? 1 : a.ldb, sbn = tb ? a.ldb : 1;
  const int vb = !tb && ((uintptr_t)a.B & 15u) == 0 && a.ldb % 4 == 0;
  const uint64_t groups = ceil_div(a.N, 4);
  VKH_REQUIRE(ceil_div(groups, 256) < (uint64_t(1) << 31), "gemm: N too large");
  dim3 grid((unsigned)ceil_div(groups, 256), (unsigned)std::min<uint64_t>(a.M, 65535));
  hipLaunchKernelGGL(k_gemm_smallk, grid, dim3(256), 0, g->stream, a, sam, sak, sbk, sbn, vb);
  check_launch("gemm (small k)");
}

// ---- backward of a one-output Dense layer after an activation ---------------------
// For a last layer y = h w^T + b with w [1, N] and h = act(z) [B, N], one pass
// over h produces everything its backward and the previous layer's bias
// gradient need, with the bits of the separate steps:
//   dx[b, n] = act'(h[b, n]) * (dy[b] * w[n])    (vkh_gemm, K = 1, aux = h)
//   dx_bf, dx_bft: dx as BF16, row-major and transposed (the next products'
//                  packed operands); the FP32 dx is not stored when dx is null
//   dw[n]    = sum_b h[b, n] * dy[b]             (the column walk of vkh_gemm
//                  for dY^T h: 0 + products in k order per chunk of kc_dw rows,
//                  chunks summed in order from 0)
//   db[n]    = sum_b dx[b, n]                    (vkh_bias_grad: per chunk of
//                  kc_db rows in row order from its first element, chunks
//                  summed in order)
// A block stages a 128-row x 32-column tile of dx and of h in LDS (33 KB, four
// blocks a compute unit); the chunk sums walk its columns and the transposed
// copy leaves in 256-byte runs (16 bytes a lane).  kc_dw and kc_db divide 128.
// Each block takes a column of tiles in turn, with the next tile's loads in
// flight during the current tile's column walks and stores.
constexpr int D1_TR = 128, D1_TC = 32, D1_PITCH = D1_TC + 1;  // odd pitch: column walks conflict-free
constexpr int D1_LOADS = D1_TR * D1_TC / 4 / 256;                 // float4 loads of h a thread

__global__ void __launch_bounds__(256)
k_dense1_backward(const float* __restrict__ dy, const float* __restrict__ w, const float* __restrict__ h,
                  int act, uint64_t N, uint64_t row_tiles, float* __restrict__ dx,
                  uint16_t* __restrict__ dx_bf, uint16_t* __restrict__ dx_bft, uint64_t ldt,
                  float* __restrict__ part_dw, int kc_dw, float* __restrict__ part_db, int kc_db) {
  __shared__ float D[D1_TR * D1_PITCH], H[D1_TR * D1_PITCH], Y[D1_TR];
  const uint64_t n0 = (uint64_t)blockIdx.x * D1_TC;
  const int t = threadIdx.x, c4 = t % (D1_TC / 4), r0 = t / (D1_TC / 4);
  const uint64_t n = n0 + c4 * 4;
  const float4 wv = *reinterpret_cast<const float4*>(w + n);
  const float wq[4] = {wv.x, wv.y, wv.z, wv.w};
  const int ndb = part_db ? D1_TR / kc_db : 0, ndw = part_dw ? D1_TR / kc_dw : 0;
  // the block walks row tiles blockIdx.y, + gridDim.y, ...; each tile's
  // loads are issued before the previous tile's column walks and
  // transposed stores, so they are in flight meanwhile
  float4 hv[D1_LOADS];
  float d[D1_LOADS], yv = 0.f;
  auto load_tile = [&](uint64_t tile) {
    const uint64_t b0 = tile * D1_TR;
#pragma unroll
    for (int it = 0; it < D1_LOADS; ++it) {
      const uint64_t b = b0 + r0 + it * (256 / (D1_TC / 4));
      hv[it] = ld_stream(reinterpret_cast<const float4*>(h + b * N + n));
      d[it] = dy[b];
    }
    if (t < D1_TR) yv = dy[b0 + t];
  };
  uint64_t tile = blockIdx.y;
  if (tile < row_tiles) load_tile(tile);
  for (; tile < row_tiles; tile += gridDim.y) {
    const uint64_t b0 = tile * D1_TR;
    if (t < D1_TR) Y[t] = yv;
#pragma unroll
    for (int it = 0; it < D1_LOADS; ++it) {
      const int r = r0 + it * (256 / (D1_TC / 4));
      const uint64_t b = b0 + r;
      const float hq[4] = {hv[it].x, hv[it].y, hv[it].z, hv[it].w};
      float o[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        o[q] = act_deriv_mul(act, hq[q], d[it] * wq[q]);
        D[r * D1_PITCH + c4 * 4 + q] = o[q];
        H[r * D1_PITCH + c4 * 4 + q] = hq[q];
      }
      if (dx) st_stream(reinterpret_cast<float4*>(dx + b * N + n), make_float4(o[0], o[1], o[2], o[3]));
      if (dx_bf) {
        uint2 hb;
        hb.x = f2bf(o[0]) | ((uint32_t)f2bf(o[1]) << 16);
        hb.y = f2bf(o[2]) | ((uint32_t)f2bf(o[3]) << 16);
        *reinterpret_cast<uint2*>(dx_bf + b * N + n) = hb;
      }
    }
    __syncthreads();
    if (tile + gridDim.y < row_tiles) load_tile(tile + gridDim.y);
    // chunk sums: one thread per (chunk, column), rows in order (each sum a
    // chain of dependent adds; its LDS reads go out 8 at a time)
    for (int item = t; item < (ndb + ndw) * D1_TC; item += 256) {
      const int col = item % D1_TC, j = item / D1_TC;
      if (j < ndb) {
        const int rs = j * kc_db;
        float acc = D[rs * D1_PITCH + col];
        int r = 1;
        for (; r + 8 <= kc_db; r += 8) {
          float v[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = D[(rs + r + q) * D1_PITCH + col];
#pragma unroll
          for (int q = 0; q < 8; ++q) acc = acc + v[q];
        }
        for (; r < kc_db; ++r) acc = acc + D[(rs + r) * D1_PITCH + col];
        part_db[(b0 / kc_db + j) * N + n0 + col] = acc;
      } else {
        const int rs = (j - ndb) * kc_dw;
        float acc = 0.f;
        int r = 0;
        for (; r + 8 <= kc_dw; r += 8) {
          float v[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = H[(rs + r + q) * D1_PITCH + col] * Y[rs + r + q];
#pragma unroll
          for (int q = 0; q < 8; ++q) acc += v[q];
        }
        for (; r < kc_dw; ++r) acc += H[(rs + r) * D1_PITCH + col] * Y[rs + r];
        part_dw[(b0 / kc_dw + j - ndb) * N + n0 + col] = acc;
      }
    }
    if (dx_bft) {
      // D1_TC columns x 16 runs of 8 rows; 16 consecutive lanes continue one column
      for (int item = t; item < D1_TC * (D1_TR / 8); item += 256) {
        const int col = item / (D1_TR / 8), seg = item % (D1_TR / 8);
        uint32_t wd[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          wd[q] = f2bf(D[(seg * 8 + 2 * q) * D1_PITCH + col]) |
                  ((uint32_t)f2bf(D[(seg * 8 + 2 * q + 1) * D1_PITCH + col]) << 16);
        *reinterpret_cast<uint4*>(dx_bft + (n0 + col) * ldt + b0 + seg * 8) =
            make_uint4(wd[0], wd[1], wd[2], wd[3]);
      }
    }
    __syncthreads();  // D, H and Y are read before the next tile writes them
  }
}

// chunk lengths of the separate steps (dw: the column walk over N outputs and
// B terms; db: vkh_bias_grad's column sums); false where they do not tile the
// kernel's 128 rows
static bool dense1_chunks(vkh_gpu* g, uint64_t B, uint64_t N, uint64_t& kc_dw, uint64_t& ndw, uint64_t& kc_db) {
  if (B == 0 || B % D1_TR != 0 || N % D1_TC != 0 || B / D1_TR >= 65536 || N / D1_TC >= (uint64_t(1) << 31))
    return false;
  kc_dw = colwalk_chunk(g, N, B, true, ndw);
  kc_db = bias_grad_chunk(g, B, N, true);
  return kc_dw <= (uint64_t)D1_TR && D1_TR % kc_dw == 0 && kc_db <= (uint64_t)D1_TR && D1_TR % kc_db == 0;
}

static void gemm_impl(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N,
                      uint64_t K, float alpha, const float* A, uint64_t lda, const float* B,
                      uint64_t ldb, float beta, float* Cp, uint64_t ldc, const float* bias, int act,
                      uint16_t* c_bf16, const float* aux, uint64_t ldaux, int aux_act,
                      const uint16_t* a_bf16 = nullptr, const uint16_t* b_bf16 = nullptr,
                      uint16_t* c_bf16t = nullptr, uint64_t ldt = 0, float* cs = nullptr,
                      int no_c = 0, const uint16_t* aux16 = nullptr) {
  use(g);
  if (M == 0 || N == 0) return;
  VKH_REQUIRE(Cp || no_c, "gemm: null C");
  VKH_REQUIRE(!(aux && aux16), "gemm: aux given twice");
  VKH_REQUIRE(!aux16 || aux_act == VKH_ACT_RELU, "gemm: a BF16 aux is for ReLU' only");
  VKH_REQUIRE(K == 0 || (A && B), "gemm: null operand");
  VKH_REQUIRE(ldc >= N, "gemm: ldc < N");
  VKH_REQUIRE(trans_a ? lda >= M : lda >= K, "gemm: lda too small");
  VKH_REQUIRE(trans_b ? ldb >= K : ldb >= N, "gemm: ldb too small");
  VKH_REQUIRE(act >= VKH_ACT_NONE && act <= VKH_ACT_SIGMOID, "gemm: unknown activation");
  VKH_REQUIRE(aux_act >= VKH_ACT_NONE && aux_act <= VKH_ACT_SIGMOID, "gemm: unknown aux activation");
  VKH_REQUIRE(!(aux || aux16) || ldaux >= N, "gemm: ldaux < N");
  VKH_REQUIRE(!c_bf16t || ldt >= M, "gemm: ldt < M");
  VKH_REQUIRE(!c_bf16t || ((uintptr_t)c_bf16t & 1u) == 0, "gemm: c_bf16t not 2-byte aligned");
  VKH_REQUIRE(!a_bf16 || ((uintptr_t)a_bf16 & 15u) == 0, "gemm: a_bf16 not 16-byte aligned");
  VKH_REQUIRE(!b_bf16 || ((uintptr_t)b_bf16 & 15u) == 0, "gemm: b_bf16 not 16-byte aligned");
  GemmArgs a{};
  a.A = A; a.B = B; a.C = Cp; a.bias = bias; a.Cbf = c_bf16;
  a.M = M; a.N = N; a.K = K; a.lda = lda; a.ldb = ldb; a.ldc = ldc;
  a.alpha = alpha; a.beta = beta; a.act = act;
  a.aux = aux; a.ldaux = ldaux; a.aux_act = aux_act;
  a.Cbft = c_bf16t; a.ldt = ldt;
  a.cs = cs; a.no_c = no_c; a.aux16 = aux16;
  VKH_REQUIRE(!no_c || beta == 0.0f, "gemm: beta * C needs C");
  a.vec_c = ((uintptr_t)Cp & 15u) == 0 && ldc % 4 == 0 && (!bias || ((uintptr_t)bias & 15u) == 0) &&
            (!aux || (((uintptr_t)aux & 15u) == 0 && ldaux % 4 == 0)) &&
            (!aux16 || (((uintptr_t)aux16 & 7u) == 0 && ldaux % 4 == 0)) &&
            (!c_bf16 || ((uintptr_t)c_bf16 & 7u) == 0);
  VKH_REQUIRE(mode >= VKH_GEMM_FP32 && mode <= VKH_GEMM_BF16X3, "gemm: unknown mode");
  if (std::min(M, N) <= SKINNY_MAX && K > 0) {
    gemm_skinny(g, trans_a, trans_b, a);
    return;
  }
  if (K <= SMALLK_MAX) {
    gemm_smallk(g, trans_a, trans_b, a);
    return;
  }
  switch (mode) {
    case VKH_GEMM_FP32: {
      auto it = g->gemm_cfg.find(cfg_key(mode, trans_a, trans_b, M, N, K));
      gemm_f32(g, trans_a, trans_b, a, it != g->gemm_cfg.end() ? it->second : default_cfg(mode, M, N, K));
      break;
    }
    case VKH_GEMM_BF16:
    case VKH_GEMM_BF16X3: {
      auto it = g->gemm_cfg.find(cfg_key(mode, trans_a, trans_b, M, N, K));
      const bool pre = mode == VKH_GEMM_BF16 && K % PK_KALIGN == 0;
      gemm_packed(g, mode, trans_a, trans_b, a, it != g->gemm_cfg.end() ? it->second : default_cfg(mode, M, N, K),
                  pre ? a_bf16 : nullptr, pre ? b_bf16 : nullptr);
      break;
    }
    default: fail(VKH_INVALID, "gemm: unknown mode");
  }
}

int vkh_gemm_configs(int mode, int* count) {
  return guarded([&] {
    VKH_REQUIRE(count, "null out");
    VKH_REQUIRE(mode >= VKH_GEMM_FP32 && mode <= VKH_GEMM_BF16X3, "unknown mode");
    *count = mode_cfgs(mode);
  });
}

int vkh_gemm_config_info(int mode, int index, int* bm, int* bn, int* bk, int* waves, int* direct) {
  return guarded([&] {
    VKH_REQUIRE(bm && bn && bk && waves, "null out");
    VKH_REQUIRE(mode >= VKH_GEMM_FP32 && mode <= VKH_GEMM_BF16X3, "unknown mode");
    if (mode == VKH_GEMM_FP32) {
      VKH_REQUIRE(index >= 0 && index < F32_NCFG, "configuration index out of range");
      const F32Cfg& c = F32_CFGS[index];
      *bm = c.bm; *bn = c.bn; *bk = F32_BK; *waves = c.waves;
      if (direct) *direct = 0;
      return;
    }
    int n = 0;
    const PackedCfg* c = packed_cfgs(mode, n);
    VKH_REQUIRE(index >= 0 && index < n, "configuration index out of range");
    *bm = c[index].bm; *bn = c[index].bn; *bk = c[index].bk; *waves = c[index].waves;
    if (direct) *direct = c[index].direct;
  });
}

int vkh_gemm_get_config(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N,
                        uint64_t K, int* index, int* tuned) {
  return guarded([&] {
    VKH_REQUIRE(g && index, "null argument");
    VKH_REQUIRE(mode >= VKH_GEMM_FP32 && mode <= VKH_GEMM_BF16X3, "unknown mode");
    auto it = g->gemm_cfg.find(cfg_key(mode, trans_a, trans_b, M, N, K));
    const bool have = it != g->gemm_cfg.end();
    *index = have ? it->second : default_cfg(mode, M, N, K);
    if (tuned) *tuned = have ? 1 : 0;
  });
}

int vkh_gemm_set_config(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N,
                        uint64_t K, int index) {
  return guarded([&] {
    VKH_REQUIRE(g, "null GPU handle");
    VKH_REQUIRE(mode >= VKH_GEMM_FP32 && mode <= VKH_GEMM_BF16X3, "unknown mode");
    const int n = mode_cfgs(mode);
    const auto key = cfg_key(mode, trans_a, trans_b, M, N, K);
    if (index < 0) { g->gemm_cfg.erase(key); return; }
    VKH_REQUIRE(index < n, "configuration index out of range");
    g->gemm_cfg[key] = index;
  });
}

int vkh_gemm_autotune(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N,
                      uint64_t K, int reps, int* best, float* ms_per_config) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(mode >= VKH_GEMM_FP32 && mode <= VKH_GEMM_BF16X3, "unknown mode");
    VKH_REQUIRE(!g->capturing, "gemm_autotune: not allowed during graph capture");
    VKH_REQUIRE(M > 0 && N > 0 && K > 0, "gemm_autotune: empty product");
    VKH_REQUIRE(g->alloc, "gemm_autotune needs a context with its own allocator");
    if (reps < 1) reps = 3;
    const int n = mode_cfgs(mode);
    const bool skinny = std::min(M, N) <= SKINNY_MAX || K <= SMALLK_MAX;
    if (skinny) {  // a single kernel choice
      if (ms_per_config)
        for (int i = 0; i < n; ++i) ms_per_config[i] = -1.0f;
      if (best) *best = default_cfg(mode, M, N, K);
      return;
    }
    // scratch operands in the product's own layout
    const uint64_t na = M * K, nbv = K * N, nc = M * N;
    Workspace wa(g, sizeof(float) * na), wb(g, sizeof(float) * nbv), wc(g, sizeof(float) * nc);
    float* A = (float*)wa.get();
    float* B = (float*)wb.get();
    float* C = (float*)wc.get();
    hipLaunchKernelGGL(k_fill_noise, dim3(full_grid(na, 256)), dim3(256), 0, g->stream, A, na, 1u);
    hipLaunchKernelGGL(k_fill_noise, dim3(full_grid(nbv, 256)), dim3(256), 0, g->stream, B, nbv, 2u);
    check_launch("gemm_autotune fill");
    GemmArgs a{};
    a.A = A; a.B = B; a.C = C;
    a.M = M; a.N = N; a.K = K;
    a.lda = trans_a ? M : K; a.ldb = trans_b ? K : N; a.ldc = N;
    a.alpha = 1.0f; a.beta = 0.0f; a.act = VKH_ACT_NONE; a.aux_act = VKH_ACT_NONE;
    a.vec_c = 1;
    // BF16: the operands packed once, padded for the largest tile (which
    // every configuration accepts), so that only the product is timed (the
    // packing passes are the same for every configuration)
    const bool pre = mode == VKH_GEMM_BF16;
    const uint64_t Kp = ceil_div(K, PK_KALIGN) * PK_KALIGN;
    const uint64_t Mp = ceil_div(M, PK_MAXTILE) * PK_MAXTILE, Np = ceil_div(N, PK_MAXTILE) * PK_MAXTILE;
    Workspace wpa(g, pre ? sizeof(uint16_t) * Mp * Kp : 1), wpb(g, pre ? sizeof(uint16_t) * Np * Kp : 1);
    uint16_t* pa = pre ? (uint16_t*)wpa.get() : nullptr;
    uint16_t* pb = pre ? (uint16_t*)wpb.get() : nullptr;
    if (pre) {
      pack_operand<false>(g, A, a.lda, trans_a, M, K, pa, nullptr, Mp, Kp);
      pack_operand<false>(g, B, a.ldb, !trans_b, N, K, pb, nullptr, Np, Kp);
    }
    hipEvent_t e0, e1;
    VKH_HIP(hipEventCreate(&e0));
    VKH_HIP(hipEventCreate(&e1));
    int bi = -1;
    float bt = 0.0f;
    try {
      auto run = [&](int i) {
        GemmArgs t = a;
        if (mode == VKH_GEMM_FP32) gemm_f32(g, trans_a, trans_b, t, i);
        else gemm_packed(g, mode, trans_a, trans_b, t, i, pa, pb);
      };
      for (int i = 0; i < n; ++i) {
        run(i);  // warm-up (and packing-area growth)
        VKH_HIP(hipEventRecord(e0, g->stream));
        for (int r = 0; r < reps; ++r) run(i);
        VKH_HIP(hipEventRecord(e1, g->stream));
        VKH_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        VKH_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)reps;
        if (ms_per_config) ms_per_config[i] = ms;
        if (bi < 0 || ms < bt) { bi = i; bt = ms; }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g->gemm_cfg[cfg_key(mode, trans_a, trans_b, M, N, K)] = bi;
    if (best) *best = bi;
  });
}

int vkh_gemm_autotune_ext(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N, uint64_t K,
                          float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb, float beta,
                          float* Cp, uint64_t ldc, const float* bias, int act, uint16_t* c_bf16,
                          const float* aux, const uint16_t* aux_bf16, uint64_t ldaux, int aux_act,
                          const uint16_t* a_bf16, const uint16_t* b_bf16, uint16_t* c_bf16t, uint64_t ldt,
                          float* db, int store_c, int reps, int* best, float* ms_per_config) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(mode >= VKH_GEMM_FP32 && mode <= VKH_GEMM_BF16X3, "unknown mode");
    VKH_REQUIRE(!g->capturing, "gemm_autotune: not allowed during graph capture");
    VKH_REQUIRE(M > 0 && N > 0 && K > 0, "gemm_autotune: empty product");
    if (reps < 1) reps = 3;
    (void)db;  // the column sums are not part of the timing
    const int n = mode_cfgs(mode);
    const auto key = cfg_key(mode, trans_a, trans_b, M, N, K);
    // the call's own epilogue; without its column sums an unstored C would
    // leave no output, so then C is stored
    const int no_c = (store_c || !(c_bf16 || c_bf16t)) ? 0 : 1;
    auto run = [&] {
      gemm_impl(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act, c_bf16, aux,
                ldaux, aux_act, a_bf16, b_bf16, c_bf16t, ldt, nullptr, no_c, aux_bf16);
    };
    if (std::min(M, N) <= SKINNY_MAX || K <= SMALLK_MAX) {  // a single kernel choice
      if (ms_per_config)
        for (int i = 0; i < n; ++i) ms_per_config[i] = -1.0f;
      if (best) *best = default_cfg(mode, M, N, K);
      return;
    }
    hipEvent_t e0, e1;
    VKH_HIP(hipEventCreate(&e0));
    VKH_HIP(hipEventCreate(&e1));
    int bi = -1;
    float bt = 0.0f;
    try {
      for (int i = 0; i < n; ++i) {
        g->gemm_cfg[key] = i;
        run();  // warm-up (and packing-area growth)
        VKH_HIP(hipEventRecord(e0, g->stream));
        for (int r = 0; r < reps; ++r) run();
        VKH_HIP(hipEventRecord(e1, g->stream));
        VKH_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        VKH_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)reps;
        if (ms_per_config) ms_per_config[i] = ms;
        if (bi < 0 || ms < bt) { bi = i; bt = ms; }
      }
    } catch (...) {
      g->gemm_cfg.erase(key);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g->gemm_cfg[key] = bi;
    if (best) *best = bi;
  });
}

int vkh_gemm_reserve(vkh_gpu* g, int mode, uint64_t M, uint64_t N, uint64_t K) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(mode >= VKH_GEMM_FP32 && mode <= VKH_GEMM_BF16X3, "gemm_reserve: unknown mode");
    VKH_REQUIRE(!g->capturing, "gemm_reserve: not allowed during graph capture");
    // skinny and short-k products pack nothing; FP32 needs room only for
    // split-k partials
    if (M == 0 || N == 0 || std::min(M, N) <= SKINNY_MAX || K <= SMALLK_MAX) return;
    const size_t bytes = mode == VKH_GEMM_FP32 ? f32_split_bytes(M, N, K) : packed_bytes(mode, M, N, K);
    if (bytes == 0) return;
    g->graph_pack_hint = std::max(g->graph_pack_hint, bytes);
    pack_area(g, bytes);
  });
}

int vkh_gemm(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N, uint64_t K,
             float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb, float beta,
             float* Cp, uint64_t ldc, const float* bias, int act, uint16_t* c_bf16) {
  return guarded([&] {
    gemm_impl(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act,
              c_bf16, nullptr, 0, VKH_ACT_NONE);
  });
}

int vkh_gemm_ex(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N, uint64_t K,
                float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb, float beta,
                float* Cp, uint64_t ldc, const float* bias, int act, uint16_t* c_bf16,
                const float* aux, uint64_t ldaux, int aux_act) {
  return guarded([&] {
    gemm_impl(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act,
              c_bf16, aux, ldaux, aux_act);
  });
}

int vkh_gemm_bf16a(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N, uint64_t K,
                   float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb, float beta,
                   float* Cp, uint64_t ldc, const float* bias, int act, uint16_t* c_bf16,
                   const float* aux, uint64_t ldaux, int aux_act, const uint16_t* a_bf16) {
  return guarded([&] {
    gemm_impl(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act,
              c_bf16, aux, ldaux, aux_act, a_bf16);
  });
}

int vkh_gemm_bf16io(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N, uint64_t K,
                    float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb, float beta,
                    float* Cp, uint64_t ldc, const float* bias, int act, uint16_t* c_bf16,
                    const float* aux, uint64_t ldaux, int aux_act, const uint16_t* a_bf16,
                    const uint16_t* b_bf16, uint16_t* c_bf16t, uint64_t ldt) {
  return guarded([&] {
    gemm_impl(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act,
              c_bf16, aux, ldaux, aux_act, a_bf16, b_bf16, c_bf16t, ldt);
  });
}

// Whether an M x N x K product takes C's column sums in its epilogue with
// vkh_bias_grad's bits: the matrix-core path, a configuration (tuned or
// default) whose waves hold 128-row slabs, no split k, and bias_grad's column
// sums chunked by 128 rows for [M, N].
static bool colsum_ok(vkh_gpu* g, int mode, int ta, int tb, uint64_t M, uint64_t N, uint64_t K) {
  if (std::min(M, N) <= SKINNY_MAX || K <= SMALLK_MAX) return false;
  if (mode < VKH_GEMM_FP32 || mode > VKH_GEMM_BF16X3) return false;
  if (bias_grad_chunk(g, M, N, N % 4 == 0) != (uint64_t)CS_ROWS) return false;
  auto it = g->gemm_cfg.find(cfg_key(mode, ta, tb, M, N, K));
  const int cfg = it != g->gemm_cfg.end() ? it->second : default_cfg(mode, M, N, K);
  if (mode == VKH_GEMM_FP32)
    return F32_CFGS[cfg].wm == CS_ROWS && ceil_div(K, f32_splitk_chunk(M, N, K)) == 1;
  int n = 0;
  const PackedCfg* c = packed_cfgs(mode, n);
  const uint64_t Kp = ceil_div(K, (uint64_t)PK_KALIGN) * PK_KALIGN;
  return c[cfg].wm == CS_ROWS && ceil_div(Kp, splitk_chunk(M, N, Kp)) == 1;
}

int vkh_gemm_colsum_supported(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N,
                              uint64_t K, int* ok) {
  return guarded([&] {
    VKH_REQUIRE(g && ok, "gemm_colsum_supported: null argument");
    *ok = colsum_ok(g, mode, trans_a, trans_b, M, N, K) ? 1 : 0;
  });
}

// vkh_gemm_bf16io with the column sums of C (db, optional), the FP32 C
// optional (store_c) and the aux activation's output as BF16 (aux_bf16)
static void gemm_ext(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N, uint64_t K,
                     float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb, float beta, float* Cp,
                     uint64_t ldc, const float* bias, int act, uint16_t* c_bf16, const float* aux,
                     const uint16_t* aux_bf16, uint64_t ldaux, int aux_act, const uint16_t* a_bf16,
                     const uint16_t* b_bf16, uint16_t* c_bf16t, uint64_t ldt, float* db, int store_c) {
  use(g);
  VKH_REQUIRE(store_c || c_bf16 || c_bf16t || db, "gemm: no output");
  if (!db) {
    gemm_impl(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act, c_bf16, aux,
              ldaux, aux_act, a_bf16, b_bf16, c_bf16t, ldt, nullptr, store_c ? 0 : 1, aux_bf16);
    return;
  }
  VKH_REQUIRE(colsum_ok(g, mode, trans_a, trans_b, M, N, K), "gemm_colsum: shape or configuration not supported");
  const uint64_t nchunk = ceil_div(M, (uint64_t)CS_ROWS);
  Workspace ws(g, sizeof(float) * nchunk * N);
  gemm_impl(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act, c_bf16, aux,
            ldaux, aux_act, a_bf16, b_bf16, c_bf16t, ldt, (float*)ws.get(), store_c ? 0 : 1, aux_bf16);
  sum_chunks(g, (const float*)ws.get(), db, N, nchunk);
}

int vkh_gemm_bf16io_colsum(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N,
                           uint64_t K, float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb,
                           float beta, float* Cp, uint64_t ldc, const float* bias, int act, uint16_t* c_bf16,
                           const float* aux, uint64_t ldaux, int aux_act, const uint16_t* a_bf16,
                           const uint16_t* b_bf16, uint16_t* c_bf16t, uint64_t ldt, float* db, int store_c) {
  return guarded([&] {
    VKH_REQUIRE(db, "gemm_colsum: null db");
    gemm_ext(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act, c_bf16, aux,
             nullptr, ldaux, aux_act, a_bf16, b_bf16, c_bf16t, ldt, db, store_c);
  });
}

int vkh_gemm_bf16io_ext(vkh_gpu* g, int mode, int trans_a, int trans_b, uint64_t M, uint64_t N, uint64_t K,
                        float alpha, const float* A, uint64_t lda, const float* B, uint64_t ldb, float beta,
                        float* Cp, uint64_t ldc, const float* bias, int act, uint16_t* c_bf16, const float* aux,
                        const uint16_t* aux_bf16, uint64_t ldaux, int aux_act, const uint16_t* a_bf16,
                        const uint16_t* b_bf16, uint16_t* c_bf16t, uint64_t ldt, float* db, int store_c) {
  return guarded([&] {
    gemm_ext(g, mode, trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc, bias, act, c_bf16, aux,
             aux_bf16, ldaux, aux_act, a_bf16, b_bf16, c_bf16t, ldt, db, store_c);
  });
}

int vkh_dense1_backward_supported(vkh_gpu* g, uint64_t B, uint64_t N, int* ok) {
  return guarded([&] {
    VKH_REQUIRE(g && ok, "dense1_backward_supported: null argument");
    uint64_t kc_dw = 0, ndw = 0, kc_db = 0;
    *ok = dense1_chunks(g, B, N, kc_dw, ndw, kc_db) ? 1 : 0;
  });
}

int vkh_dense1_backward(vkh_gpu* g, const float* dy, const float* w, const float* h, int act, uint64_t B,
                        uint64_t N, float* dx, uint16_t* dx_bf, uint16_t* dx_bft, uint64_t ldt, float* dw,
                        float* db) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(dy && w && h, "dense1_backward: null pointer");
    VKH_REQUIRE(act >= VKH_ACT_NONE && act <= VKH_ACT_SIGMOID, "dense1_backward: unknown activation");
    auto a16 = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
    VKH_REQUIRE(a16(w) && a16(h) && (!dx || a16(dx)), "dense1_backward: operands not 16-byte aligned");
    VKH_REQUIRE(!dx_bf || ((uintptr_t)dx_bf & 7u) == 0, "dense1_backward: dx_bf not 8-byte aligned");
    VKH_REQUIRE(!dx_bft || (a16(dx_bft) && ldt % 8 == 0 && ldt >= B), "dense1_backward: dx_bft layout");
    uint64_t kc_dw = 0, ndw = 0, kc_db = 0;
    VKH_REQUIRE(dense1_chunks(g, B, N, kc_dw, ndw, kc_db), "dense1_backward: shape not supported");
    const uint64_t ndb = B / kc_db;
    Workspace wdw(g, dw ? sizeof(float) * ndw * N : 1), wdb(g, db ? sizeof(float) * ndb * N : 1);
    float* pdw = dw ? (float*)wdw.get() : nullptr;
    float* pdb = db ? (float*)wdb.get() : nullptr;
    // about four blocks a compute unit, each walking B / 128 / gy row tiles
    const uint64_t col_tiles = N / D1_TC, row_tiles = B / D1_TR;
    const uint64_t gy = std::min<uint64_t>(row_tiles, std::max<uint64_t>(1, ceil_div(4 * (uint64_t)g->num_cus,
                                                                                      col_tiles)));
    dim3 grid((unsigned)col_tiles, (unsigned)gy);
    hipLaunchKernelGGL(k_dense1_backward, grid, dim3(256), 0, g->stream, dy, w, h, act, N, row_tiles, dx, dx_bf,
                       dx_bft, ldt, pdw, (int)kc_dw, pdb, (int)kc_db);
    check_launch("dense1_backward");
    // dw: the column walk's combine (0 + chunks in order; its 1 * sum is the sum)
    if (dw) sum_chunks_from_zero(g, (const float*)pdw, dw, N, ndw);
    if (db) sum_chunks(g, (const float*)pdb, db, N, ndb);
  });
}

int vkh_to_bf16(vkh_gpu* g, const float* src, uint16_t* dst, uint64_t n) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(src && dst, "to_bf16: null pointer");
    unsigned grid = stream_grid(g, n, 256, 8);
    hipLaunchKernelGGL(k_to_bf16, dim3(grid), dim3(256), 0, g->stream, src, dst, n);
    check_launch("to_bf16");
  });
}

}  // extern "C"
