// Xoshiro128++ PRNG (F18), Box-Muller, randrange, and gather (F17).
//
// Seeding follows vulkpy/_vkarray.cc:643-675 bit for bit: four chained
// splitmix64 outputs (589-595, 658-663) give state 0 and state i is the
// reference's jump() (597-621) applied to state i-1.  That jump assigns the
// state inside the loop over the four JUMP words without resetting its
// accumulators, so it is not Vigna's canonical 2^64 jump, but it is still
// linear over GF(2): the 128x128 bit matrix of one jump lets every host thread
// start at J^lo(state 0) and walk its own range of states sequentially.
//
// Generation replaces the chunked host loop of _vkarray.cc:697-717 (one
// dispatch of `size` threads per chunk, host wait between chunks) by one
// kernel: thread i keeps state i in registers and produces outputs
// i, i + S, i + 2S, ... < n, which is exactly the reference's mapping.
#include "common.h"
#include "bmath.h"

#include <array>
#include <thread>

namespace vkh {

// ---- host seeding ---------------------------------------------------------------
struct St {
  uint32_t w[4];
};

static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

static inline void next_host(St& s) {
  const uint32_t t = s.w[1] << 9;
  s.w[2] ^= s.w[0];
  s.w[3] ^= s.w[1];
  s.w[1] ^= s.w[2];
  s.w[0] ^= s.w[3];
  s.w[2] ^= t;
  s.w[3] = rotl32(s.w[3], 11);
}

static inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// the jump exactly as _vkarray.cc:597-621 writes it
static inline void jump_as_written(St& s) {
  static const uint32_t JUMP[4] = {0x8764000b, 0xf542d2d3, 0x6fa035c3, 0x77f2db5b};
  uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (uint32_t j : JUMP) {
    for (int b = 0; b < 32; ++b) {
      if (j & (1u << b)) { a0 ^= s.w[0]; a1 ^= s.w[1]; a2 ^= s.w[2]; a3 ^= s.w[3]; }
      next_host(s);
    }
    s.w[0] = a0; s.w[1] = a1; s.w[2] = a2; s.w[3] = a3;
  }
}

using Mat = std::array<St, 128>;  // column k = image of basis vector e_k

static St mat_apply(const Mat& M, const St& x) {
  St y{{0, 0, 0, 0}};
  for (int k = 0; k < 128; ++k)
    if (x.w[k >> 5] & (1u << (k & 31)))
      for (int q = 0; q < 4; ++q) y.w[q] ^= M[k].w[q];
  return y;
}

static Mat compose(const Mat& A, const Mat& B) {  // A after B
  Mat C;
  for (int k = 0; k < 128; ++k) C[k] = mat_apply(A, B[k]);
  return C;
}

static Mat identity() {
  Mat I;
  for (int k = 0; k < 128; ++k) {
    I[k] = St{{0, 0, 0, 0}};
    I[k].w[k >> 5] = 1u << (k & 31);
  }
  return I;
}

static Mat jump_matrix() {
  Mat J;
  for (int k = 0; k < 128; ++k) {
    St e{{0, 0, 0, 0}};
    e.w[k >> 5] = 1u << (k & 31);
    jump_as_written(e);
    J[k] = e;
  }
  return J;
}

static Mat mat_pow(Mat base, uint64_t p) {
  Mat r = identity();
  while (p) {
    if (p & 1) r = compose(base, r);
    base = compose(base, base);
    p >>= 1;
  }
  return r;
}

// one step of next() as a matrix (the state update is linear; the ++ output
// scrambler is not part of it)
static Mat step_matrix() {
  Mat M;
  for (int k = 0; k < 128; ++k) {
    St e{{0, 0, 0, 0}};
    e.w[k >> 5] = 1u << (k & 31);
    next_host(e);
    M[k] = e;
  }
  return M;
}

// P_j = M^(Lc * 2^j), j < nbits, as [nbits][128 columns][4 words].  Kept
// per chunk length for the life of the process (generators made by the
// same initializer shapes share chunk lengths).
static void chunk_jumps(uint64_t Lc, int nbits, std::vector<uint32_t>& out) {
  static std::mutex mu;
  static std::unordered_map<uint64_t, std::vector<uint32_t>> cache;
  std::lock_guard<std::mutex> lock(mu);
  std::vector<uint32_t>& v = cache[Lc];
  int have = (int)(v.size() / 512);
  if (have < nbits) {
    Mat P;
    if (have == 0) {
      P = mat_pow(step_matrix(), Lc);
    } else {
      for (int k = 0; k < 128; ++k) std::memcpy(P[k].w, &v[((size_t)(have - 1) * 128 + k) * 4], 16);
      P = compose(P, P);
    }
    v.resize((size_t)nbits * 512);
    for (int j = have; j < nbits; ++j) {
      if (j > have) P = compose(P, P);
      for (int k = 0; k < 128; ++k) std::memcpy(&v[((size_t)j * 128 + k) * 4], P[k].w, 16);
    }
  }
  out.assign(v.begin(), v.begin() + (size_t)nbits * 512);
}

static void seed_states(uint32_t S, uint64_t seed, std::vector<uint32_t>& out) {
  out.resize((size_t)4 * S);
  St s0;
  for (int i = 0; i < 4; ++i) {
    seed = splitmix64(seed);
    s0.w[i] = (uint32_t)seed;
  }
  unsigned T = std::thread::hardware_concurrency();
  if (T < 1) T = 1;
  if (T > 16) T = 16;
  if (S < 4096) T = 1;
  const uint64_t per = ceil_div(S, T);
  Mat J = T > 1 ? jump_matrix() : Mat{};
  auto work = [&](uint64_t lo, uint64_t hi) {
    St s = lo == 0 ? s0 : mat_apply(mat_pow(J, lo), s0);
    for (uint64_t i = lo; i < hi; ++i) {
      if (i > lo) jump_as_written(s);
      std::memcpy(&out[4 * i], s.w, 16);
    }
  };
  if (T == 1) { work(0, S); return; }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; ++t) {
    uint64_t lo = t * per, hi = std::min<uint64_t>(S, lo + per);
    if (lo >= hi) break;
    th.emplace_back(work, lo, hi);
  }
  for (auto& x : th) x.join();
}

// ---- device generation ------------------------------------------------------------
__device__ __forceinline__ uint32_t drotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

// prng_xoshiro128pp_uint32.comp:32-42
__device__ __forceinline__ uint32_t dnext(uint4& s) {
  const uint32_t result = drotl(s.x + s.w, 7) + s.x;
  const uint32_t t = s.y << 9;
  s.z ^= s.x;
  s.w ^= s.y;
  s.y ^= s.z;
  s.x ^= s.w;
  s.z ^= t;
  s.w = drotl(s.w, 11);
  return result;
}

// prng_xoshiro128pp_float.comp:33
__device__ __forceinline__ float to_unit(uint32_t r) {
  return __uint_as_float((r >> 9) | 0x3f800000u) - 1.0f;
}

constexpr int PRNG_BLOCK = 256;

// OUT: 0 = uint32, 1 = float in [0, 1), 2 = randrange
template <int OUT>
__global__ void __launch_bounds__(PRNG_BLOCK)
k_xoshiro(uint4* state, void* out, uint32_t S, uint64_t n, uint32_t low, uint32_t range) {
  const uint64_t i = (uint64_t)blockIdx.x * PRNG_BLOCK + threadIdx.x;
  if (i >= S || i >= n) return;
  uint4 s = state[i];
  for (uint64_t k = i; k < n; k += S) {
    const uint32_t r = dnext(s);
    if (OUT == 0) st_stream1(static_cast<uint32_t*>(out) + k, r);
    else if (OUT == 1) st_stream1(static_cast<float*>(out) + k, to_unit(r));
    else {
      // prng_randrange.comp:24-26: low + uint(float(range) * u)
      const float u = to_unit(r);
      st_stream1(static_cast<uint32_t*>(out) + k, low + __float2uint_rz((float)range * u));
    }
  }
  state[i] = s;
}

// ---- chunked generation for few states ---------------------------------------------
// With S states and n outputs each state takes up to ceil(n / S) sequential
// steps; when S is small (the reference's HeNormal uses 64 states,
// nn/initializers.py:75-89) that serial chain is the whole kernel.  Here
// thread (i, c) produces steps [c*Lc, (c+1)*Lc) of state i: it starts from
// M^(c*Lc) applied to the state (M = one next() step over GF(2)), built
// from the precomputed P_j = M^(Lc * 2^j) for the set bits j of c.  Outputs
// keep the mapping k = t*S + i.  Start states are read from a copy of the
// state buffer; the chunk holding a state's last step writes it back.

// y = P x for a 128x128 bit matrix given by its columns (4 words each)
__device__ __forceinline__ uint4 gf2_apply(const uint32_t* __restrict__ P, uint4 x) {
  uint4 y = make_uint4(0u, 0u, 0u, 0u);
  const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    for (int b = 0; b < 32; ++b) {
      if ((w[q] >> b) & 1u) {
        const uint32_t* col = P + (q * 32 + b) * 4;
        y.x ^= col[0]; y.y ^= col[1]; y.z ^= col[2]; y.w ^= col[3];
      }
    }
  }
  return y;
}

__device__ __forceinline__ uint4 chunk_start(const uint4* init, const uint32_t* jumps, uint64_t i,
                                             uint64_t c) {
  uint4 s = init[i];
  for (int j = 0; (c >> j) != 0; ++j)
    if ((c >> j) & 1u) s = gf2_apply(jumps + (size_t)j * 512, s);
  return s;
}

// steps of state i when n outputs are taken round-robin from S states
__device__ __forceinline__ uint64_t steps_of(uint64_t i, uint64_t S, uint64_t n) {
  return n > i ? (n - 1 - i) / S + 1 : 0;
}

template <int OUT>
__global__ void __launch_bounds__(PRNG_BLOCK)
k_xoshiro_chunked(const uint4* init, uint4* state, const uint32_t* jumps, void* out, uint32_t S,
                  uint64_t n, uint64_t Lc, uint64_t nchunks, uint32_t low, uint32_t range) {
  const uint64_t tid = (uint64_t)blockIdx.x * PRNG_BLOCK + threadIdx.x;
  if (tid >= (uint64_t)S * nchunks) return;
  const uint64_t i = tid % S, c = tid / S;
  const uint64_t L = steps_of(i, S, n);
  const uint64_t t0 = c * Lc;
  if (t0 >= L) return;
  const uint64_t t1 = min(t0 + Lc, L);
  uint4 s = chunk_start(init, jumps, i, c);
  for (uint64_t t = t0; t < t1; ++t) {
    const uint64_t k = t * S + i;
    const uint32_t r = dnext(s);
    if (OUT == 0) st_stream1(static_cast<uint32_t*>(out) + k, r);
    else if (OUT == 1) st_stream1(static_cast<float*>(out) + k, to_unit(r));
    else st_stream1(static_cast<uint32_t*>(out) + k, low + __float2uint_rz((float)range * to_unit(r)));
  }
  if (t1 == L) state[i] = s;
}

// ln t for t in [2^-23, 1] (t = 1 - u of a generated uniform): the hardware
// log2, times ln 2 held as two floats with the rounding error of the leading
// product recovered by an FMA.  On this range (no zero, denormal or infinite
// input) these are the device logf's own steps, so the result is the same
// without its range handling.
constexpr float LN2F_HI = 0x1.62e42ep-1f, LN2F_LO = 0x1.efa39ep-25f;  // hi + lo = ln 2
__device__ __forceinline__ float log_unit(float t) {
  const float l2 = __builtin_amdgcn_logf(t);
  const float p = l2 * LN2F_HI;
  const float e = fmaf(l2, LN2F_LO, fmaf(l2, LN2F_HI, -p));
  return p + e;
}

// Correctly rounded sqrt x for x in [0, 2^96): the hardware square root, then
// whichever neighbour's product with it brackets x (x = 0 stays 0: the lower
// neighbour's residual is NaN, the upper one's is -0).
__device__ __forceinline__ float sqrt_rn(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __int_as_float(__float_as_int(s) - 1), sp = __int_as_float(__float_as_int(s) + 1);
  const float r = fmaf(-sm, s, x) <= 0.0f ? sm : s;
  return fmaf(-sp, s, x) > 0.0f ? sp : r;
}

// Box-Muller of one pair of uniforms (prng_box_muller.comp:25-31) in the
// reference's FP32 operation order; sin and cos come from sincos_small (the
// GLSL.std.450 functions the reference relies on are likewise FP32 within a
// few ulp).  GEN: u1 is a generated uniform in [0, 1), so log and sqrt take
// the range-specific forms above; otherwise (uniforms from a caller's
// buffer) the device logf and sqrtf.
template <bool GEN>
__device__ __forceinline__ void box_muller(float u1, float u2, float mean, float sd, float& o1, float& o2) {
  const float t = 1.0f - u1;
  const float l = GEN ? log_unit(t) : logf(t);
  const float r = (GEN ? sqrt_rn(-2.0f * l) : sqrtf(-2.0f * l)) * sd;
  const float angle = 6.28318530718f * u2;
  float s, c;
  sincos_small(angle, s, c);
  o1 = mean + r * s;
  o2 = mean + r * c;
}

typedef float vkh_f2v __attribute__((ext_vector_type(2)));

// out[j] (j < n) and out[j + 1] (j + 1 < n) of a normal pair; one 8-byte
// nontemporal store when both are in range and out + j is 8-byte aligned
__device__ __forceinline__ void store_pair(float* out, uint64_t j, uint64_t n, float o1, float o2) {
  if (j + 1 < n && ((uintptr_t)(out + j) & 7u) == 0) {
    const vkh_f2v v = {o1, o2};
    __builtin_nontemporal_store(v, reinterpret_cast<vkh_f2v*>(out + j));
  } else {
    if (j < n) st_stream1(out + j, o1);
    if (j + 1 < n) st_stream1(out + j + 1, o2);
  }
}

// Fused uniform + Box-Muller: thread p owns states 2p and 2p+1 (S even) and
// produces normal pairs at uniform indices j = c*S + 2p < nu (nu even).
// PAIRED: n == nu and out 8-byte aligned, so every pair is one 8-byte store.
template <bool PAIRED>
__global__ void __launch_bounds__(PRNG_BLOCK)
k_xoshiro_normal(uint4* state, float* out, uint32_t S, uint64_t nu, uint64_t n, float mean, float sd) {
  const uint64_t p = (uint64_t)blockIdx.x * PRNG_BLOCK + threadIdx.x;
  const uint64_t i = 2 * p;
  if (i >= S || i >= nu) return;
  uint4 s1 = state[i], s2 = state[i + 1];
  for (uint64_t j = i; j < nu; j += S) {
    const float u1 = to_unit(dnext(s1));
    const float u2 = to_unit(dnext(s2));
    float o1, o2;
    box_muller<true>(u1, u2, mean, sd, o1, o2);
    if constexpr (PAIRED) {
      const vkh_f2v v = {o1, o2};
      __builtin_nontemporal_store(v, reinterpret_cast<vkh_f2v*>(out + j));
    } else {
      store_pair(out, j, n, o1, o2);
    }
  }
  state[i] = s1;
  state[i + 1] = s2;
}

// chunked form of k_xoshiro_normal: thread (p, c) advances states 2p and 2p+1
// over steps [c*Lc, (c+1)*Lc); both take the same number of steps because S
// and nu are even
__global__ void __launch_bounds__(PRNG_BLOCK)
k_xoshiro_normal_chunked(const uint4* init, uint4* state, const uint32_t* jumps, float* out,
                         uint32_t S, uint64_t nu, uint64_t n, uint64_t Lc, uint64_t nchunks,
                         float mean, float sd) {
  const uint64_t tid = (uint64_t)blockIdx.x * PRNG_BLOCK + threadIdx.x;
  const uint64_t half = S / 2;
  if (tid >= half * nchunks) return;
  const uint64_t p = tid % half, c = tid / half, i = 2 * p;
  const uint64_t L = steps_of(i, S, nu);
  const uint64_t t0 = c * Lc;
  if (t0 >= L) return;
  const uint64_t t1 = min(t0 + Lc, L);
  uint4 s1 = chunk_start(init, jumps, i, c), s2 = chunk_start(init, jumps, i + 1, c);
  for (uint64_t t = t0; t < t1; ++t) {
    const uint64_t j = t * S + i;
    const float u1 = to_unit(dnext(s1));
    const float u2 = to_unit(dnext(s2));
    float o1, o2;
    box_muller<true>(u1, u2, mean, sd, o1, o2);
    store_pair(out, j, n, o1, o2);
  }
  if (t1 == L) {
    state[i] = s1;
    state[i + 1] = s2;
  }
}

__global__ void __launch_bounds__(PRNG_BLOCK)
k_box_muller(const float* u, float* out, uint64_t npairs, uint64_t n, float mean, float sd) {
  const uint64_t stride = (uint64_t)gridDim.x * PRNG_BLOCK;
  for (uint64_t p = (uint64_t)blockIdx.x * PRNG_BLOCK + threadIdx.x; p < npairs; p += stride) {
    const uint64_t j = 2 * p;
    const float u1 = u[j], u2 = u[j + 1];
    float o1, o2;
    box_muller<false>(u1, u2, mean, sd, o1, o2);
    if (j < n) out[j] = o1;
    if (j + 1 < n) out[j + 1] = o2;
  }
}

// ---- gather --------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_gather(const float* a, uint64_t na, const uint32_t* idx, float* c, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    uint64_t k = idx[i];
    c[i] = a[k < na ? k : na - 1];
  }
}

// c[k, p, j] = a[p, min(idx[k], axis-1), j]
__global__ void __launch_bounds__(256)
k_gather_axis(const float* a, const uint32_t* idx, float* c, uint64_t prev, uint64_t axis,
              uint64_t post, uint64_t nidx) {
  const uint64_t total = nidx * prev * post;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
    const uint64_t j = t % post;
    const uint64_t q = t / post;
    const uint64_t p = q % prev;
    const uint64_t k = q / prev;
    uint64_t b = idx[k];
    if (b >= axis) b = axis - 1;
    c[t] = a[(p * axis + b) * post + j];
  }
}

// out[k, j] = 1 if min(idx[k], classes - 1) == j else 0: the one-hot rows the
// reference gathers from an identity matrix (vkarray.py:239-253), out-of-range
// labels clamped like gather_axis
__global__ void __launch_bounds__(256)
k_onehot(const uint32_t* idx, float* out, uint64_t n, uint32_t classes) {
  const uint64_t total = n * classes;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += stride) {
    const uint64_t k = t / classes;
    const uint32_t j = (uint32_t)(t - k * classes);
    const uint32_t lab = min(idx[k], classes - 1);
    out[t] = lab == j ? 1.0f : 0.0f;
  }
}

}  // namespace vkh

struct vkh_xoshiro {
  vkh_gpu* gpu = nullptr;
  uint32_t size = 0;
  uint32_t* state = nullptr;  // [size][4]
  bool from_alloc = false;
  // jump matrices of the chunked kernels, valid for chunk length jumps_lc
  uint32_t* jumps = nullptr;
  size_t jumps_bytes = 0;
  uint64_t jumps_lc = 0;
  int jumps_nbits = 0;
};

using namespace vkh;

static unsigned prng_grid(uint64_t threads) { return (unsigned)ceil_div(threads, PRNG_BLOCK); }

namespace {

struct ChunkPlan {
  bool use = false;
  uint64_t Lc = 0, C = 1;
  int nbits = 0;
};

// Chunk the per-state chains when `units` parallel chains (states, or state
// pairs) are too few to fill the GPU and each is long: about 1024 threads
// per compute unit, and at least 256 steps per chunk so the start-state
// matrix products stay small next to the generation.  Not during graph
// capture (the jump matrices are uploaded synchronously).
ChunkPlan plan_chunks(const vkh_gpu* g, uint64_t units, uint64_t steps) {
  ChunkPlan p;
  if (g->capturing) return p;
  const uint64_t target = (uint64_t)(g->num_cus > 0 ? g->num_cus : 256) * 1024;
  if (units * 4 >= target) return p;
  const uint64_t C = std::min<uint64_t>(ceil_div(target, units), steps / 256);
  if (C < 2) return p;
  p.Lc = ceil_div(steps, C);
  p.C = ceil_div(steps, p.Lc);
  if (p.C < 2) return p;
  p.use = true;
  while ((uint64_t(1) << p.nbits) < p.C) ++p.nbits;  // bits of c in [0, C)
  return p;
}

const uint32_t* prepare_jumps(vkh_xoshiro* r, const ChunkPlan& p) {
  if (r->jumps && r->jumps_lc == p.Lc && r->jumps_nbits >= p.nbits) return r->jumps;
  std::vector<uint32_t> host;
  chunk_jumps(p.Lc, p.nbits, host);
  const size_t nbytes = host.size() * sizeof(uint32_t);
  vkh_gpu* g = r->gpu;
  if (nbytes > r->jumps_bytes) {
    if (r->jumps) {
      VKH_HIP(hipStreamSynchronize(g->stream));
      VKH_HIP(hipFree(r->jumps));
      r->jumps = nullptr;
      r->jumps_bytes = 0;
    }
    VKH_HIP(hipMalloc(&r->jumps, nbytes));
    r->jumps_bytes = nbytes;
  }
  // synchronous: earlier chunked launches may still read the old matrices,
  // and the host vector dies at the end of this call
  VKH_HIP(hipStreamSynchronize(g->stream));
  VKH_HIP(hipMemcpy(r->jumps, host.data(), nbytes, hipMemcpyHostToDevice));
  r->jumps_lc = p.Lc;
  r->jumps_nbits = p.nbits;
  return r->jumps;
}

template <int OUT>
void launch_uniform(vkh_xoshiro* r, void* out, uint64_t n, uint32_t low, uint32_t range) {
  vkh_gpu* g = r->gpu;
  const ChunkPlan cp = plan_chunks(g, r->size, ceil_div(n, r->size));
  if (cp.use) {
    const uint32_t* jumps = prepare_jumps(r, cp);
    const size_t sb = (size_t)16 * r->size;
    Workspace init(g, sb);
    VKH_HIP(hipMemcpyAsync(init.get(), r->state, sb, hipMemcpyDeviceToDevice, g->stream));
    const uint64_t th = (uint64_t)r->size * cp.C;
    hipLaunchKernelGGL(k_xoshiro_chunked<OUT>, dim3(prng_grid(th)), dim3(PRNG_BLOCK), 0, g->stream,
                       (const uint4*)init.get(), (uint4*)r->state, jumps, out, r->size, n, cp.Lc, cp.C,
                       low, range);
  } else {
    // one state per thread (measured 197 us for 2^28 outputs from 2^20 states;
    // four states per thread with 16-byte stores: 232 us; four steps of a
    // block's 256 states staged through LDS so that each wave stores one
    // 1 KB row segment with 16-byte stores: no faster)
    const uint64_t th = std::min<uint64_t>(r->size, n);
    hipLaunchKernelGGL(k_xoshiro<OUT>, dim3(prng_grid(th)), dim3(PRNG_BLOCK), 0, g->stream,
                       (uint4*)r->state, out, r->size, n, low, range);
  }
  check_launch("xoshiro");
}

}  // namespace

extern "C" {

int vkh_xoshiro_create(vkh_gpu* g, uint32_t size, uint64_t seed, vkh_xoshiro** out) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(out, "null out");
    VKH_REQUIRE(size >= 1, "Xoshiro128pp size must be at least 1");
    std::vector<uint32_t> host;
    seed_states(size, seed, host);
    auto* r = new vkh_xoshiro();
    r->gpu = g;
    r->size = size;
    const size_t nbytes = sizeof(uint32_t) * 4 * (size_t)size;
    try {
      if (g->alloc) {
        r->state = (uint32_t*)g->alloc->allocate(nbytes);
        r->from_alloc = true;
      } else {
        VKH_HIP(hipMalloc(&r->state, nbytes));
      }
      // synchronous copy: the host vector dies at the end of this call
      VKH_HIP(hipMemcpyAsync(r->state, host.data(), nbytes, hipMemcpyHostToDevice, g->stream));
      VKH_HIP(hipStreamSynchronize(g->stream));
    } catch (...) {
      if (r->state) { if (r->from_alloc) g->alloc->release(r->state); else (void)hipFree(r->state); }
      delete r;
      throw;
    }
    *out = r;
  });
}

int vkh_xoshiro_destroy(vkh_xoshiro* r) {
  return guarded([&] {
    if (!r) return;
    use(r->gpu);
    VKH_REQUIRE(!r->gpu->capturing, "xoshiro_destroy: not allowed during graph capture");
    if (r->from_alloc) r->gpu->alloc->release(r->state);
    else { VKH_HIP(hipStreamSynchronize(r->gpu->stream)); VKH_HIP(hipFree(r->state)); }
    if (r->jumps) {
      VKH_HIP(hipStreamSynchronize(r->gpu->stream));
      VKH_HIP(hipFree(r->jumps));
    }
    delete r;
  });
}

int vkh_xoshiro_size(vkh_xoshiro* r, uint32_t* out) {
  return guarded([&] { VKH_REQUIRE(r && out, "null argument"); *out = r->size; });
}

int vkh_xoshiro_state(vkh_xoshiro* r, uint32_t** out) {
  return guarded([&] { VKH_REQUIRE(r && out, "null argument"); *out = r->state; });
}

int vkh_xoshiro_uint32(vkh_xoshiro* r, uint32_t* out, uint64_t n) {
  return guarded([&] {
    VKH_REQUIRE(r, "null generator");
    use(r->gpu);
    if (n == 0) return;
    VKH_REQUIRE(out, "null output");
    launch_uniform<0>(r, out, n, 0u, 0u);
  });
}

int vkh_xoshiro_float(vkh_xoshiro* r, float* out, uint64_t n) {
  return guarded([&] {
    VKH_REQUIRE(r, "null generator");
    use(r->gpu);
    if (n == 0) return;
    VKH_REQUIRE(out, "null output");
    launch_uniform<1>(r, out, n, 0u, 0u);
  });
}

int vkh_xoshiro_randrange(vkh_xoshiro* r, uint32_t* out, uint64_t n, uint32_t low, uint32_t high_incl) {
  return guarded([&] {
    VKH_REQUIRE(r, "null generator");
    use(r->gpu);
    if (n == 0) return;
    VKH_REQUIRE(out, "null output");
    uint32_t range = high_incl - low + 1u;  // uint32 arithmetic as in the shader
    launch_uniform<2>(r, out, n, low, range);
  });
}

int vkh_box_muller(vkh_gpu* g, const float* u, float* out, uint64_t n, float mean, float sd) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(u && out, "null pointer");
    uint64_t npairs = ceil_div(n, 2);
    unsigned grid = stream_grid(g, npairs, PRNG_BLOCK, 8);
    hipLaunchKernelGGL(k_box_muller, dim3(grid), dim3(PRNG_BLOCK), 0, g->stream, u, out, npairs, n, mean, sd);
    check_launch("box_muller");
  });
}

int vkh_xoshiro_normal(vkh_xoshiro* r, float* out, uint64_t n, float mean, float sd) {
  return guarded([&] {
    VKH_REQUIRE(r, "null generator");
    vkh_gpu* g = r->gpu;
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(out, "null output");
    const uint64_t nu = n + (n & 1);  // random.py:105-123: n even, or n + 1 uniforms
    if (r->size % 2 == 0) {
      const ChunkPlan cp = plan_chunks(g, r->size / 2, ceil_div(nu, r->size));
      if (cp.use) {
        const uint32_t* jumps = prepare_jumps(r, cp);
        const size_t sb = (size_t)16 * r->size;
        Workspace init(g, sb);
        VKH_HIP(hipMemcpyAsync(init.get(), r->state, sb, hipMemcpyDeviceToDevice, g->stream));
        const uint64_t th = (uint64_t)(r->size / 2) * cp.C;
        hipLaunchKernelGGL(k_xoshiro_normal_chunked, dim3(prng_grid(th)), dim3(PRNG_BLOCK), 0, g->stream,
                           (const uint4*)init.get(), (uint4*)r->state, jumps, out, r->size, nu, n,
                           cp.Lc, cp.C, mean, sd);
      } else {
        uint64_t th = std::min<uint64_t>(r->size, nu) / 2;
        if (n == nu && ((uintptr_t)out & 7u) == 0)
          hipLaunchKernelGGL(k_xoshiro_normal<true>, dim3(prng_grid(th)), dim3(PRNG_BLOCK), 0, g->stream,
                             (uint4*)r->state, out, r->size, nu, n, mean, sd);
        else
          hipLaunchKernelGGL(k_xoshiro_normal<false>, dim3(prng_grid(th)), dim3(PRNG_BLOCK), 0, g->stream,
                             (uint4*)r->state, out, r->size, nu, n, mean, sd);
      }
      check_launch("xoshiro normal");
      return;
    }
    // odd number of states: uniforms first, then the transform
    Workspace ws(g, sizeof(float) * nu);
    float* u = (float*)ws.get();
    launch_uniform<1>(r, u, nu, 0u, 0u);
    uint64_t npairs = nu / 2;
    unsigned grid = stream_grid(g, npairs, PRNG_BLOCK, 8);
    hipLaunchKernelGGL(k_box_muller, dim3(grid), dim3(PRNG_BLOCK), 0, g->stream, u, out, npairs, n, mean, sd);
    check_launch("box_muller");
  });
}

int vkh_gather(vkh_gpu* g, const float* a, uint64_t na, const uint32_t* idx, float* c, uint64_t n) {
  return guarded([&] {
    use(g);
    if (n == 0) return;
    VKH_REQUIRE(a && idx && c, "gather: null pointer");
    VKH_REQUIRE(na >= 1, "gather from an empty array");
    unsigned grid = stream_grid(g, n, 256, 8);
    hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, g->stream, a, na, idx, c, n);
    check_launch("gather");
  });
}

int vkh_onehot(vkh_gpu* g, const uint32_t* idx, uint64_t n, uint32_t classes, float* out) {
  return guarded([&] {
    use(g);
    VKH_REQUIRE(classes >= 1, "onehot: need at least one class");
    if (n == 0) return;
    VKH_REQUIRE(idx && out, "onehot: null pointer");
    unsigned grid = stream_grid(g, n * classes, 256, 8);
    hipLaunchKernelGGL(k_onehot, dim3(grid), dim3(256), 0, g->stream, idx, out, n, classes);
    check_launch("onehot");
  });
}

int vkh_gather_axis(vkh_gpu* g, const float* a, const uint32_t* idx, float* c, uint64_t prev,
                    uint64_t axis, uint64_t post, uint64_t nidx) {
  return guarded([&] {
    use(g);
    uint64_t total = nidx * prev * post;
    if (total == 0) return;
    VKH_REQUIRE(a && idx && c, "gather_axis: null pointer");
    VKH_REQUIRE(axis >= 1, "gather along an empty axis");
    unsigned grid = stream_grid(g, total, 256, 8);
    hipLaunchKernelGGL(k_gather_axis, dim3(grid), dim3(256), 0, g->stream, a, idx, c, prev, axis, post, nidx);
    check_launch("gather_axis");
  });
}

}  // extern "C"
