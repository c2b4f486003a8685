/*
 * vk_oracle.c -- CPU restatement of the vulkpy compute shaders and of the
 * Xoshiro128++ host code, used as the correctness oracle and as the timed CPU
 * baseline.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (vulkpy_amd/) may link,
 * load or call this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the vulkpy source tree).  Loops over independent output elements are
 * OpenMP-parallel; results do not depend on the thread count.
 *
 * Decisions shared with the product (documented in DESIGN.md):
 *  - max/min follow numpy.maximum/minimum (NaN propagates) where GLSL leaves
 *    NaN undefined;
 *  - pow is C pow (FP64, rounded once) where GLSL pow is undefined for x < 0;
 *  - transcendentals are FP64 libm results rounded once to FP32.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* OpenMP threads the parallel loops use (the CPU baseline's core count) */
EXPORT int oracle_num_threads(void) { return omp_get_max_threads(); }

/* ---- Xoshiro128++ ------------------------------------------------------- */

static uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

/* _vkarray.cc:627-641 (next_on_cpu) == prng_xoshiro128pp_uint32.comp:32-42 */
static uint32_t next_state(uint32_t s[4]) {
  const uint32_t result = rotl32(s[0] + s[3], 7) + s[0];
  const uint32_t t = s[1] << 9;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl32(s[3], 11);
  return result;
}

/* _vkarray.cc:589-595 */
EXPORT uint64_t oracle_splitmix64(uint64_t x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* _vkarray.cc:597-621: the jump exactly as written (state assigned inside
 * the loop over JUMP words, accumulators never reset) */
EXPORT void oracle_jump_as_written(uint32_t s[4]) {
  static const uint32_t JUMP[4] = {0x8764000b, 0xf542d2d3, 0x6fa035c3, 0x77f2db5b};
  uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int w = 0; w < 4; ++w) {
    for (int b = 0; b < 32; ++b) {
      if (JUMP[w] & ((uint32_t)1 << b)) { s0 ^= s[0]; s1 ^= s[1]; s2 ^= s[2]; s3 ^= s[3]; }
      next_state(s);
    }
    s[0] = s0; s[1] = s1; s[2] = s2; s[3] = s3;
  }
}

/* Vigna's published xoshiro128plusplus.c jump (state assigned after all four
 * words); only used by tests to show the reference differs from it */
EXPORT void oracle_jump_canonical(uint32_t s[4]) {
  static const uint32_t JUMP[4] = {0x8764000b, 0xf542d2d3, 0x6fa035c3, 0x77f2db5b};
  uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int w = 0; w < 4; ++w)
    for (int b = 0; b < 32; ++b) {
      if (JUMP[w] & ((uint32_t)1 << b)) { s0 ^= s[0]; s1 ^= s[1]; s2 ^= s[2]; s3 ^= s[3]; }
      next_state(s);
    }
  s[0] = s0; s[1] = s1; s[2] = s2; s[3] = s3;
}

EXPORT uint32_t oracle_next(uint32_t s[4]) { return next_state(s); }

/* _vkarray.cc:653-675: state 0 from four chained splitmix64 outputs, state i
 * = jump(state i-1), sequentially */
EXPORT void oracle_xoshiro_seed(uint32_t S, uint64_t seed, uint32_t* out) {
  uint32_t s[4];
  for (int i = 0; i < 4; ++i) {
    seed = oracle_splitmix64(seed);
    s[i] = (uint32_t)seed;
  }
  memcpy(out, s, 16);
  for (uint32_t i = 1; i < S; ++i) {
    oracle_jump_as_written(s);
    memcpy(out + 4 * (size_t)i, s, 16);
  }
}

/* _vkarray.cc:697-717 with prng_xoshiro128pp_uint32.comp: chunks of S outputs,
 * thread i of chunk c writes output c*S + i from state i */
EXPORT void oracle_xoshiro_uint32(uint32_t* state, uint32_t S, uint32_t* out, uint64_t n) {
  for (uint64_t shift = 0; shift < n; shift += S) {
    uint64_t size = (n - shift < S) ? n - shift : S;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)size; ++i) out[shift + i] = next_state(state + 4 * i);
  }
}

/* prng_xoshiro128pp_float.comp:33 */
static float to_unit(uint32_t r) {
  uint32_t bits = (r >> 9) | 0x3f800000u;
  float f;
  memcpy(&f, &bits, 4);
  return f - 1.0f;
}

EXPORT void oracle_xoshiro_float(uint32_t* state, uint32_t S, float* out, uint64_t n) {
  for (uint64_t shift = 0; shift < n; shift += S) {
    uint64_t size = (n - shift < S) ? n - shift : S;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)size; ++i) out[shift + i] = to_unit(next_state(state + 4 * i));
  }
}

/* prng_box_muller.comp:19-32 (u holds n + (n & 1) uniforms) */
EXPORT void oracle_box_muller(const float* u, float* out, uint64_t n, float mean, float sd) {
  const uint64_t npairs = (n + 1) / 2;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < (int64_t)npairs; ++p) {
    const uint64_t j = 2 * (uint64_t)p, k = j + 1;
    const float t = 1.0f - u[j];
    const float l = (float)log((double)t);
    const float r = sqrtf(-2.0f * l) * sd;
    const float angle = 6.28318530718f * u[k];
    out[j] = mean + r * (float)sin((double)angle);
    if (k < n) out[k] = mean + r * (float)cos((double)angle);
  }
}

/* prng_randrange.comp:20-27 */
EXPORT void oracle_randrange(const float* u, uint32_t* out, uint64_t n, uint32_t low, uint32_t high) {
  const uint32_t range = high - low + 1u;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = low + (uint32_t)((float)range * u[i]);
}

/* ---- elementwise -------------------------------------------------------- */
static float np_max(float a, float b) { return (a >= b || a != a) ? a : b; }
static float np_min(float a, float b) { return (a <= b || a != a) ? a : b; }

/* op codes as in include/vulkpy_hip.h enum vkh_binop */
static float binop(int op, float a, float b) {
  switch (op) {
    case 0: return a + b;
    case 1: return a - b;
    case 2: return a * b;
    case 3: return a / b;
    case 4: return np_max(a, b);
    case 5: return np_min(a, b);
    case 6: return (float)pow((double)a, (double)b);
    case 7: return b - a;
    case 8: return b / a;
    case 9: return (float)pow((double)b, (double)a);
  }
  return 0.0f;
}

/* add.comp:21-26 and friends (same-shape), *_scalar.comp (b == NULL) */
EXPORT void oracle_binary(int op, const float* a, const float* b, float s, float* c, uint64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) c[i] = binop(op, a[i], b ? b[i] : s);
}

/* add_broadcast.comp:25-44: shapeABC = [a0..an, b0..bn, c0..cn] */
EXPORT void oracle_binary_bcast(int op, const float* a, const float* b, float* c,
                                uint32_t size_a, uint32_t size_b, uint32_t size_c,
                                uint32_t ndim, const uint32_t* shapeABC) {
#pragma omp parallel for schedule(static)
  for (int64_t ci = 0; ci < (int64_t)size_c; ++ci) {
    uint32_t sx = size_a, sy = size_b, sz = size_c;
    uint32_t ai = 0, bi = 0, ci_tmp = (uint32_t)ci;
    for (uint32_t dim = 0; dim < ndim; ++dim) {
      const uint32_t sa = shapeABC[dim], sb = shapeABC[dim + ndim], sc = shapeABC[dim + 2 * ndim];
      sx /= sa; sy /= sb; sz /= sc;
      const uint32_t d = ci_tmp / sz;
      ai += sx * (d < sa - 1 ? d : sa - 1);
      bi += sy * (d < sb - 1 ? d : sb - 1);
      ci_tmp = ci_tmp % sz;
    }
    c[ci] = binop(op, a[ai], b[bi]);
  }
}

/* broadcast.comp:25-44 */
EXPORT void oracle_broadcast_to(const float* a, float* b, uint32_t size_a, uint32_t size_b,
                                uint32_t ndim, const uint32_t* a_shape, const uint32_t* b_shape) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)size_b; ++i) {
    uint32_t i_tmp = (uint32_t)i, j = 0, sa = size_a, sb = size_b;
    for (uint32_t dim = 0; dim < ndim; ++dim) {
      sa /= a_shape[dim];
      sb /= b_shape[dim];
      uint32_t d = i_tmp / sb;
      if (d > a_shape[dim] - 1) d = a_shape[dim] - 1;
      j += d * sa;
      i_tmp = i_tmp % sb;
    }
    b[i] = a[j];
  }
}

/* ---- reductions ---------------------------------------------------------- */
/* sum_axis.comp:20-32, prod_axis.comp, maximum_axis.comp, minimum_axis.comp and
 * their *_rebroadcast.comp: sequential over k for every (i, j) */
EXPORT void oracle_reduce_axis(int op, const float* a, float* b, uint64_t prev, uint64_t axis,
                               uint64_t post, int rebroadcast) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)(prev * post); ++t) {
    const uint64_t i = (uint64_t)t / post, j = (uint64_t)t % post;
    const uint64_t ij = i * axis * post + j;
    float acc = (op == 0) ? 0.0f : (op == 1) ? 1.0f : a[ij];
    for (uint64_t k = 0; k < axis; ++k) {
      const float v = a[k * post + ij];
      if (op == 0) acc += v;
      else if (op == 1) acc *= v;
      else if (op == 2) acc = np_max(acc, v);
      else acc = np_min(acc, v);
    }
    if (rebroadcast) for (uint64_t k = 0; k < axis; ++k) b[k * post + ij] = acc;
    else b[i * post + j] = acc;
  }
}

/* ---- matmul / affine / gather --------------------------------------------- */
/* matmul.comp:23-33 (sequential FP32 accumulation; built with
 * -ffp-contract=off, so every product and sum is rounded separately) */
EXPORT void oracle_matmul(const float* A, const float* B, float* C, uint64_t rowA,
                          uint64_t contract, uint64_t colB) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)(rowA * colB); ++t) {
    const uint64_t row = (uint64_t)t / colB, col = (uint64_t)t % colB;
    float sum = 0.0f;
    for (uint64_t s = 0; s < contract; ++s) {
      sum += A[row * contract + s] * B[s * colB + col];
    }
    C[row * colB + col] = sum;
  }
}

/* batch_affine.comp:25-39: y[b, o] = sum_i w[o, i] x[b, i] + bias[o] */
EXPORT void oracle_batch_affine(const float* W, const float* bias, const float* X, float* Y,
                                uint64_t batch, uint64_t in, uint64_t out) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)(batch * out); ++t) {
    const uint64_t b = (uint64_t)t / out, o = (uint64_t)t % out;
    float sum = 0.0f;
    for (uint64_t i = 0; i < in; ++i) {
      sum += W[o * in + i] * X[b * in + i];
    }
    Y[b * out + o] = sum + bias[o];
  }
}

/* gather.comp:21-26 (index clamped to the last element) */
EXPORT void oracle_gather(const float* a, uint64_t na, const uint32_t* idx, float* c, uint64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t k = idx[i];
    c[i] = a[k < na ? k : na - 1];
  }
}

/* gather_axis.comp:24-43 (index clamped to axis_size - 1) */
EXPORT void oracle_gather_axis(const float* a, const uint32_t* idx, float* c, uint64_t prev,
                               uint64_t axis, uint64_t post, uint64_t nidx) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)(nidx * prev * post); ++t) {
    const uint64_t j = (uint64_t)t % post, q = (uint64_t)t / post;
    const uint64_t i = q % prev, k = q / prev;
    uint64_t bk = idx[k];
    if (bk > axis - 1) bk = axis - 1;
    c[t] = a[(i * axis + bk) * post + j];
  }
}

/* ---- full reductions (the numerically correct semantics) ------------------ */
EXPORT double oracle_sum_f64(const float* a, uint64_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) s += a[i];
  return s;
}
