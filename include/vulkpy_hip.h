/*
 * vulkpy_hip.h -- C ABI of the MI355X (gfx950) backend for the vulkpy array,
 * PRNG and nn API.
 *
 * Every function returns an int status (VKH_OK on success).  On failure the
 * message is kept in thread-local storage and read with vkh_last_error().
 * Device memory is addressed by plain pointers obtained from vkh_alloc (or
 * from any other allocator on the same device, e.g. torch); sizes are element
 * counts unless a name says "nbytes".  All work is enqueued on the stream
 * owned by (or attached to) a vkh_gpu context and is asynchronous with respect
 * to the host: completion is observed through vkh_event objects or
 * vkh_gpu_wait().
 *
 * The reference exposes one pybind11 module, vulkpy/_vkarray.cc:756-898, whose
 * GPU.submit(spv, ...) takes a shader path and a parameter struct.  Here each
 * shader family is a typed entry point; the comment above each entry names the
 * reference interface (file:line) it replaces.  Reference paths are relative to
 * the vulkpy source tree.
 */
#ifndef VULKPY_HIP_H
#define VULKPY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define VKH_OK            0
#define VKH_ERROR         1  /* generic failure (maps to RuntimeError)        */
#define VKH_TIMEOUT       2  /* event wait timed out ("Timeout at Command Wait") */
#define VKH_INVALID       3  /* bad argument (maps to ValueError)             */
#define VKH_OUT_OF_MEMORY 4
#define VKH_NO_DEVICE     5

/* ---- operation codes ---------------------------------------------------- */
/* binary ops: vulkpy/shader/{add,sub,mul,div,max,min,pow}*.comp,
 * reflected scalar ops rsub/rdiv/rpow_scalar.comp */
enum vkh_binop {
  VKH_ADD = 0, VKH_SUB = 1, VKH_MUL = 2, VKH_DIV = 3,
  VKH_MAX = 4, VKH_MIN = 5, VKH_POW = 6,
  VKH_RSUB = 7, VKH_RDIV = 8, VKH_RPOW = 9
};

/* unary math: vulkpy/shader/{abs,sign,...,invsqrt}.comp (vkarray.py:680-1098) */
enum vkh_unop {
  VKH_ABS = 0, VKH_SIGN = 1, VKH_SIN = 2, VKH_COS = 3, VKH_TAN = 4,
  VKH_ASIN = 5, VKH_ACOS = 6, VKH_ATAN = 7, VKH_SINH = 8, VKH_COSH = 9,
  VKH_TANH = 10, VKH_ASINH = 11, VKH_ACOSH = 12, VKH_ATANH = 13,
  VKH_EXP = 14, VKH_LOG = 15, VKH_EXP2 = 16, VKH_LOG2 = 17,
  VKH_SQRT = 18, VKH_INVSQRT = 19
};

/* reductions: vulkpy/shader/{sum,prod,maximum,minimum}*.comp */
enum vkh_redop { VKH_RSUM = 0, VKH_RPROD = 1, VKH_RMAX = 2, VKH_RMIN = 3 };

/* reduction kernel families that have configurations (vkh_reduce_autotune) */
enum vkh_red_family {
  VKH_RFAM_NONE = -1, /* a single kernel choice (rows shorter than 2048)  */
  VKH_RFAM_ALL = 0,   /* full reduction                                   */
  VKH_RFAM_ROWS = 1,  /* axis with post == 1 and rows of >= 2048 elements */
  VKH_RFAM_COLS = 2   /* strided axis, post > 1                           */
};

/* GEMM arithmetic modes (precision of the products; accumulation is FP32) */
enum vkh_gemm_mode {
  VKH_GEMM_FP32 = 0,   /* FP32 MFMA                                   */
  VKH_GEMM_BF16 = 1,   /* operands rounded to BF16, FP32 accumulate    */
  VKH_GEMM_BF16X3 = 2  /* split-BF16: hi*hi + hi*lo + lo*hi            */
};

/* GEMM epilogue activation */
enum vkh_act { VKH_ACT_NONE = 0, VKH_ACT_RELU = 1, VKH_ACT_SIGMOID = 2 };

typedef struct vkh_gpu vkh_gpu;           /* device + stream + allocator cache */
typedef struct vkh_event vkh_event;       /* completion marker of enqueued work */
typedef struct vkh_xoshiro vkh_xoshiro;   /* Xoshiro128++ state buffer          */

/* ---- errors / library --------------------------------------------------- */
/* thread-local message of the last failed call (empty if none) */
const char* vkh_last_error(void);
/* ABI version of this library */
int vkh_abi_version(int* out);

/* ---- device context ------------------------------------------------------ */
/* number of visible HIP devices */
int vkh_device_count(int* out);
/* replaces createGPU(idx, priority) (_vkarray.cc:759-763) and GPU::GPU
 * (_vkarray.cc:472-501).  One in-order stream at the default priority; the
 * priority argument is accepted and recorded only. */
int vkh_gpu_create(int idx, float priority, vkh_gpu** out);
/* context bound to an existing hipStream_t (e.g. torch's current stream);
 * kernels launched through it go to that stream, and its caching allocator
 * (workspaces, GEMM packing area, vkh_alloc) reuses memory in that stream's
 * order, so the caller must not use the context's memory from other streams
 * without synchronising.  Has no counterpart in the reference (torch interop
 * for the autograd wrappers). */
int vkh_gpu_create_on_stream(int idx, void* stream, vkh_gpu** out);
int vkh_gpu_destroy(vkh_gpu* gpu);
/* replaces GPU::wait (_vkarray.cc:560-562): waits until the stream is idle */
int vkh_gpu_wait(vkh_gpu* gpu);
/* replaces GPU::canSubgroupArithmetic (_vkarray.cc:564-573): always 1 */
int vkh_gpu_can_subgroup_arithmetic(vkh_gpu* gpu, int* out);
int vkh_gpu_get_stream(vkh_gpu* gpu, void** out);
/* device properties: compute units, LDS bytes per block, wavefront size */
int vkh_gpu_info(vkh_gpu* gpu, int* num_cus, int* lds_bytes, int* wave_size);

/* ---- device memory ------------------------------------------------------- */
/* replaces Buffer<T>::Buffer (_vkarray.cc:49-79) and GPU::createBuffer /
 * toBuffer (_vkarray.cc:511-525).  Stream-ordered caching allocator: a freed
 * block may be handed out again to work enqueued later on the same stream. */
int vkh_alloc(vkh_gpu* gpu, size_t nbytes, void** out);
int vkh_free(vkh_gpu* gpu, void* ptr);
/* releases every cached (free) block back to the driver */
int vkh_empty_cache(vkh_gpu* gpu);
/* bytes held by the cache: live allocations and free cached blocks */
int vkh_memory_stats(vkh_gpu* gpu, size_t* live_bytes, size_t* cached_bytes);
/* host -> device copy enqueued on the stream; the host buffer must stay valid
 * and unchanged until an event recorded after this call completes.  VKH_INVALID
 * while the context captures a graph.  Replaces Buffer::set / the host-visible
 * mapping (_vkarray.cc:85, 98-105). */
int vkh_upload(vkh_gpu* gpu, void* dst, const void* src, size_t nbytes);
/* device -> host copy ordered after all work on the stream; returns once the
 * data is in host memory; VKH_INVALID while the context captures a graph
 * (replaces reading the mapped pointer, _vkarray.cc:107) */
int vkh_download(vkh_gpu* gpu, void* dst, const void* src, size_t nbytes);
/* Page-locked host blocks for the host mirrors of arrays, so that vkh_upload /
 * vkh_download move them by DMA without the runtime's staging copies.  The
 * reference's buffers are host-visible mappings instead (_vkarray.cc:61-63),
 * read and written in place.  Freed blocks are cached by size class (up to
 * 4 GiB); live pinned memory is capped at 32 GiB, beyond which vkh_host_alloc
 * fails with VKH_OUT_OF_MEMORY and callers use ordinary host memory.
 * Not tied to a device. */
int vkh_host_alloc(size_t nbytes, void** out);
int vkh_host_free(void* p);
int vkh_host_empty_cache(void);
int vkh_host_stats(size_t* live_bytes, size_t* cached_bytes);
/* device -> device copy on the stream */
int vkh_copy(vkh_gpu* gpu, void* dst, const void* src, size_t nbytes);
/* dst[0..n) = value (vkarray.py:1541 `z[:] = 0.0`, initializers.py:48) */
int vkh_fill_f32(vkh_gpu* gpu, float* dst, float value, size_t n);
int vkh_fill_u32(vkh_gpu* gpu, uint32_t* dst, uint32_t value, size_t n);

/* ---- jobs ---------------------------------------------------------------- */
/* replaces Job (_vkarray.cc:392-457): an event recorded after the work
 * enqueued so far on the stream.  VKH_INVALID during a graph capture (the
 * record would become a graph node that outlives the event); likewise
 * vkh_timer_record. */
int vkh_event_record(vkh_gpu* gpu, vkh_event** out);
/* replaces Job::wait(timeout_ns) (_vkarray.cc:446-456); UINT64_MAX waits
 * forever; returns VKH_TIMEOUT with "Timeout at Command Wait" */
int vkh_event_wait(vkh_event* ev, uint64_t timeout_ns);
int vkh_event_query(vkh_event* ev, int* done);
int vkh_event_destroy(vkh_event* ev);
/* event with timing enabled (for measuring kernels on the context's stream;
 * no counterpart in the reference) and the milliseconds between two of them */
int vkh_timer_record(vkh_gpu* gpu, vkh_event** out);
int vkh_event_elapsed_ms(vkh_event* start, vkh_event* end, float* ms);
/* two timing events that the next elementwise kernel launched on the context
 * outside a graph capture (binary, unary, broadcast and clamp ops: F1-F10)
 * records itself at its own start and end (hipExtLaunchKernelGGL), so
 * vkh_event_elapsed_ms(start, stop) is that kernel's execution time alone,
 * without the event records around it.  Arming again replaces an unused
 * pair.  Destroy the events only after that kernel has run.  Refused during
 * graph capture.  No counterpart in the reference. */
int vkh_timer_next_kernel(vkh_gpu* gpu, vkh_event** start, vkh_event** stop);

/* ---- elementwise (F1-F10) ------------------------------------------------ */
/* F1/F2 same-shape binary and in-place (c may equal a):
 * add.comp ... pow.comp, iadd.comp ... ipow.comp; caller vkarray.py:447-459 */
int vkh_binary(vkh_gpu* gpu, int op, const float* a, const float* b, float* c,
               uint64_t n);
/* F3/F4 scalar and reflected scalar ops (c may equal a):
 * *_scalar.comp, r*_scalar.comp; caller vkarray.py:469-483, 573-583 */
int vkh_binary_scalar(vkh_gpu* gpu, int op, const float* a, float s, float* c,
                      uint64_t n);
/* F5/F6 broadcast binary (NumPy rule).  shapes are left-padded to ndim
 * (ndim <= 16); in-place when c == a and shape_c == shape_a.
 * *_broadcast.comp, i*_broadcast.comp; caller vkarray.py:499-519, 539-557 */
int vkh_binary_bcast(vkh_gpu* gpu, int op, const float* a, const float* b,
                     float* c, int ndim, const uint64_t* shape_a,
                     const uint64_t* shape_b, const uint64_t* shape_c);
/* F7/F8 unary math (b may equal a): abs.comp ... invsqrt.comp */
int vkh_unary(vkh_gpu* gpu, int fn, const float* a, float* b, uint64_t n);
/* F9 clamp(x, lo, hi) with scalar or array bounds, each array broadcast to
 * shape_out in-kernel (no broadcast_to copies).  lo / hi NULL => scalar lo_s /
 * hi_s.  clamp*.comp, iclamp*.comp; caller vkarray.py:1110-1191 */
int vkh_clamp(vkh_gpu* gpu, const float* x, const float* lo, const float* hi,
              float lo_s, float hi_s, float* out, int ndim,
              const uint64_t* shape_x, const uint64_t* shape_lo,
              const uint64_t* shape_hi, const uint64_t* shape_out);
/* F10 materialised broadcast copy: broadcast.comp; caller vkarray.py:1434-1474 */
int vkh_broadcast_to(vkh_gpu* gpu, const float* a, float* out, int ndim,
                     const uint64_t* shape_a, const uint64_t* shape_out);

/* ---- reductions (F11-F14) ----------------------------------------------- */
/* F11/F12 full reduction of n elements into out[0].  Accumulates in FP64;
 * mean != 0 divides the FP64 sum by n before rounding.  Unlike sum.comp /
 * sum_v1.3.comp (vkarray.py:1246-1274) this is correct for every n.
 * Uses workspace from the context's allocator. */
int vkh_reduce_all(vkh_gpu* gpu, int op, const float* a, uint64_t n,
                   float* out, int mean);
/* vkh_reduce_all without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * FP64 block partials go to a file-scope device array of 8192 doubles (64 KiB
 * per device, held until the process ends) instead of a workspace, so the
 * call allocates nothing and its results equal vkh_reduce_all's bit for bit
 * (same kernels, same configuration, same partition).  Returns VKH_INVALID if
 * the launch would need more than 8192 partials (device CUs x blocks per CU of
 * the chosen configuration; at most 4096 on a 256-CU MI355X).  All calls on
 * one GPU share that array: calls that overlap on one GPU (another stream or
 * another context of the same device) race and give undefined results. */
int vkh_reduce_all_noalloc(vkh_gpu* gpu, int op, const float* a, uint64_t n,
                           float* out, int mean);
/* F13/F14 axis reduction over a viewed as [prev, axis, post]:
 * out[p, j] = op_k a[p, k, j]  (rebroadcast == 0, out is [prev, post]) or
 * out[p, k, j] = op_k' a[p, k', j] (rebroadcast != 0, out is [prev, axis, post]).
 * The result is multiplied by scale_mul (if != 1) and then divided by
 * scale_div (if != 1), both in FP32, reproducing mean's `*= n_after/n_before`
 * (vkarray.py:1429-1430) and `/= shape[axis]` (vkarray.py:1427).
 * *_axis.comp, *_axis_rebroadcast.comp; caller vkarray.py:1194-1244 */
int vkh_reduce_axis(vkh_gpu* gpu, int op, const float* a, float* out,
                    uint64_t prev, uint64_t axis, uint64_t post,
                    int rebroadcast, float scale_mul, float scale_div);

/* Reduction kernel configurations.  No counterpart in the reference, whose
 * shaders have fixed 64-thread workgroups (sum_axis.comp:3).  kind 0 is a
 * full reduction of prev * axis * post elements, kind 1 an axis reduction of
 * [prev, axis, post].  A configuration is (unroll, par): unroll = 128-bit
 * loads in flight per thread (4 or 8), par = blocks per CU (families ALL and
 * ROWS) or target threads per CU (COLS: sets how many chunks the axis is
 * split into).  Configurations of the ROWS family and the unroll of every
 * family never change results.  The COLS chunk count fixes the summation
 * order of sums and products, and the ALL blocks per CU the partition of the
 * FP64 partial sums; every configuration keeps the stated error bounds, and
 * the chosen one is used by every later call of the shape on the context,
 * eager or captured.  Index 0 is each family's default. */
int vkh_reduce_family(int kind, uint64_t prev, uint64_t axis, uint64_t post,
                      int* family);
int vkh_reduce_configs(int family, int* count);
int vkh_reduce_config_info(int family, int index, int* unroll, int* par);
/* the configuration a shape uses; *tuned = 1 if chosen, 0 for the default */
int vkh_reduce_get_config(vkh_gpu* gpu, int kind, uint64_t prev,
                          uint64_t axis, uint64_t post, int* index,
                          int* tuned);
/* choose a configuration for a shape (index < 0 restores the default) */
int vkh_reduce_set_config(vkh_gpu* gpu, int kind, uint64_t prev,
                          uint64_t axis, uint64_t post, int index);
/* times every configuration of the shape's family on scratch data (a sum;
 * one warm-up and `reps` timed runs each), keeps the fastest for the shape
 * and returns its index; ms_per_config (optional, vkh_reduce_configs
 * entries) receives each mean time in milliseconds (-1 for a configuration
 * not timed, and for a family with a single kernel choice).  same_order != 0
 * times only the configurations that keep the default summation order, so
 * the choice cannot change a result whatever the timings.  Synchronises the
 * stream; refused during graph capture. */
int vkh_reduce_autotune(vkh_gpu* gpu, int kind, uint64_t prev, uint64_t axis,
                        uint64_t post, int reps, int same_order, int* best,
                        float* ms_per_config);

/* ---- gather (F17) -------------------------------------------------------- */
/* c[i] = a[min(idx[i], n_a - 1)]: gather.comp; caller vkarray.py:1493-1501 */
int vkh_gather(vkh_gpu* gpu, const float* a, uint64_t n_a, const uint32_t* idx,
               float* c, uint64_t n_idx);
/* c[k, p, j] = a[p, min(idx[k], axis-1), j]: gather_axis.comp;
 * caller vkarray.py:1502-1521 */
int vkh_gather_axis(vkh_gpu* gpu, const float* a, const uint32_t* idx,
                    float* c, uint64_t prev, uint64_t axis, uint64_t post,
                    uint64_t n_idx);
/* out[k, j] = (min(idx[k], classes - 1) == j) for k < n, j < classes: the
 * one-hot rows U32Array.to_onehot gathers from np.identity(classes)
 * (vkarray.py:239-253), without the host identity matrix */
int vkh_onehot(vkh_gpu* gpu, const uint32_t* idx, uint64_t n, uint32_t classes,
               float* out);

/* ---- PRNG (F18) ---------------------------------------------------------- */
/* replaces PRNG::Xoshiro128pp ctor (_vkarray.cc:643-675): size states seeded
 * from seed with the splitmix64 chain and the jump exactly as written there */
int vkh_xoshiro_create(vkh_gpu* gpu, uint32_t size, uint64_t seed,
                       vkh_xoshiro** out);
/* waits for the stream: VKH_INVALID during a graph capture */
int vkh_xoshiro_destroy(vkh_xoshiro* rng);
int vkh_xoshiro_size(vkh_xoshiro* rng, uint32_t* out);
/* device pointer of the [size][4] uint32 state (for tests) */
int vkh_xoshiro_state(vkh_xoshiro* rng, uint32_t** out);
/* replaces random_uint32 / random_float (_vkarray.cc:681-717):
 * output k is next() of state k mod size, at that state's (k div size)-th
 * step of this call; states >= n are not advanced.
 * prng_xoshiro128pp_uint32.comp / prng_xoshiro128pp_float.comp */
int vkh_xoshiro_uint32(vkh_xoshiro* rng, uint32_t* out, uint64_t n);
int vkh_xoshiro_float(vkh_xoshiro* rng, float* out, uint64_t n);
/* fused uniform + Box-Muller (random.py:60-124, prng_box_muller.comp,
 * prng_ibox_muller.comp): draws n (n even) or n + 1 (n odd) uniforms and
 * writes n normals */
int vkh_xoshiro_normal(vkh_xoshiro* rng, float* out, uint64_t n, float mean,
                       float stddev);
/* fused uniform + prng_randrange.comp (random.py:126-186); high_incl = high-1 */
int vkh_xoshiro_randrange(vkh_xoshiro* rng, uint32_t* out, uint64_t n,
                          uint32_t low, uint32_t high_incl);
/* Box-Muller of existing uniforms u[0..n_u) (n_u even) into out[0..n):
 * prng_box_muller.comp (out may equal u: prng_ibox_muller.comp) */
int vkh_box_muller(vkh_gpu* gpu, const float* u, float* out, uint64_t n,
                   float mean, float stddev);

/* ---- GEMM (F15/F16) ------------------------------------------------------ */
/* C[M,N] = alpha * op(A) op(B) + beta * C + bias[N] (bias may be NULL), then
 * act.  op(A) is A[M,K] row-major (trans_a=0) or A stored [K,M] (trans_a=1);
 * op(B) is B[K,N] (trans_b=0) or B stored [N,K] (trans_b=1).  Leading
 * dimensions are the row strides in elements.  If c_bf16 != NULL the result is
 * also stored as BF16 bits.  Replaces matmul.comp (vkarray.py:585-605) and
 * batch_affine.comp (nn/layers.py:94-101).
 * Paths: min(M, N) <= 16 (skinny) and K <= 16 (short k) are memory-bound and
 * run in FP32 on the vector units in every mode; otherwise FP32 mode uses
 * FP32 matrix cores on the operands as stored, and the BF16 modes pack their
 * operands for the BF16 matrix cores.  k is split into chunks (partials
 * summed in chunk order) when few output tiles cover C: in BF16 modes fewer
 * than 128 tiles of 256 x 256 with K >= 4096, in FP32 fewer than 256 tiles of
 * 128 x 128 with K >= 1024. */
int vkh_gemm(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
             uint64_t M, uint64_t N, uint64_t K, float alpha,
             const float* A, uint64_t lda, const float* B, uint64_t ldb,
             float beta, float* C, uint64_t ldc, const float* bias, int act,
             uint16_t* c_bf16);

/* BF16 / BF16X3 GEMMs pack their operands into a BF16 area, and products
 * whose k is split (every mode) keep their chunk partials there: a
 * per-context area that grows on demand for eager calls, and a region of the
 * graph's arena for calls captured into a graph.  This grows the former for an
 * M x N x K product ahead of time and records the size so that the next
 * vkh_graph_begin reserves a large enough region (a capture cannot allocate;
 * a captured GEMM that does not fit fails and records its size for the next
 * capture).  Refused during a capture.  No counterpart in the reference. */
int vkh_gemm_reserve(vkh_gpu* gpu, int mode, uint64_t M, uint64_t N,
                     uint64_t K);

/* ---- GEMM autotuning (no counterpart in the reference) -------------------
 * Every mode has several kernel configurations (tile BM x BN, k slice BK,
 * waves per block).  Within a mode every configuration sums each output's k
 * in the same order, so the choice never changes a result.  A context keeps
 * one chosen configuration per (mode, trans_a, trans_b, M, N, K); without
 * one a shape uses the built-in default.  Products with min(M, N) <= 16 or
 * K <= 16 use the vector kernels and no configuration. */
int vkh_gemm_configs(int mode, int* count);
/* direct (optional): 1 when the configuration loads operand slices from
 * global memory straight into LDS, 0 when it stages them through registers */
int vkh_gemm_config_info(int mode, int index, int* bm, int* bn, int* bk,
                         int* waves, int* direct);
/* the configuration a shape uses; *tuned = 1 if chosen, 0 for the default */
int vkh_gemm_get_config(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
                        uint64_t M, uint64_t N, uint64_t K, int* index,
                        int* tuned);
/* choose a configuration for a shape (index < 0 restores the default) */
int vkh_gemm_set_config(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
                        uint64_t M, uint64_t N, uint64_t K, int index);
/* times every configuration on the shape (scratch operands from the
 * context's allocator, one warm-up and `reps` timed runs each; operand
 * packing included in FP32 and BF16X3, the BF16 operands packed once
 * beforehand), keeps the fastest for the shape and returns its index;
 * ms_per_config (optional, vkh_gemm_configs entries) receives each mean time
 * in milliseconds.  Synchronises the stream; refused during graph capture. */
int vkh_gemm_autotune(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
                      uint64_t M, uint64_t N, uint64_t K, int reps, int* best,
                      float* ms_per_config);
/* vkh_gemm_autotune timing the caller's own product: every configuration
 * runs the vkh_gemm_bf16io_ext call given (its operands, packed operands and
 * epilogue: bias, activation, BF16 copies, aux, stored or unstored C; not
 * the column sums db), one warm-up and `reps` timed runs each, and the
 * fastest is kept for the shape.  The outputs' contents are unspecified
 * afterwards.  Synchronises the stream; refused during graph capture.  No
 * counterpart in the reference (a captured training step tunes its products
 * with their epilogues, whose cost differs between tile shapes). */
int vkh_gemm_autotune_ext(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
                          uint64_t M, uint64_t N, uint64_t K, float alpha,
                          const float* A, uint64_t lda, const float* B,
                          uint64_t ldb, float beta, float* C, uint64_t ldc,
                          const float* bias, int act, uint16_t* c_bf16,
                          const float* aux, const uint16_t* aux_bf16,
                          uint64_t ldaux, int aux_act, const uint16_t* a_bf16,
                          const uint16_t* b_bf16, uint16_t* c_bf16t,
                          uint64_t ldt, float* db, int store_c, int reps,
                          int* best, float* ms_per_config);
/* vkh_gemm whose epilogue finally multiplies by act'(y) with y = aux[m, n]
 * (aux_act: VKH_ACT_RELU -> max(sign(y), 0), VKH_ACT_SIGMOID -> (1-y)*y):
 * the backward of an activation fused into the GEMM that produces its input
 * gradient, i.e. Dense.backward's `dy @ W` (nn/layers.py:141) followed by
 * ReLU.backward / Sigmoid.backward (nn/layers.py:207-210, 264-267) */
int vkh_gemm_ex(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
                uint64_t M, uint64_t N, uint64_t K, float alpha,
                const float* A, uint64_t lda, const float* B, uint64_t ldb,
                float beta, float* C, uint64_t ldc, const float* bias, int act,
                uint16_t* c_bf16, const float* aux, uint64_t ldaux,
                int aux_act);
/* vkh_gemm_ex with op(A) also given as BF16 in the layout the BF16 packing
 * pass writes: a_bf16 holds the M rows of op(A), K BF16 values each, with M
 * rounded up to 256 rows allocated (the extra rows are read but reach no
 * element of C), K a multiple of 64 -- for
 * instance the c_bf16 output of the GEMM that produced A.  In BF16 mode,
 * when the matrix-core path runs, A's packing pass is skipped; otherwise
 * a_bf16 is ignored.  The result is identical either way.  No counterpart in
 * the reference: it lets a captured training step hand one layer's
 * activations (or input gradients) to the next product without packing
 * them again. */
int vkh_gemm_bf16a(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
                   uint64_t M, uint64_t N, uint64_t K, float alpha,
                   const float* A, uint64_t lda, const float* B, uint64_t ldb,
                   float beta, float* C, uint64_t ldc, const float* bias,
                   int act, uint16_t* c_bf16, const float* aux,
                   uint64_t ldaux, int aux_act, const uint16_t* a_bf16);
/* vkh_gemm_bf16a that also takes op(B) as BF16 and can write C^T as BF16.
 * b_bf16 (optional) holds the N rows of op(B)^T, K BF16 values each, N
 * rounded up to 256 rows allocated (as for a_bf16), K a multiple of 64;
 * in BF16 mode on the matrix-core path B's packing pass is skipped.
 * c_bf16t (optional) receives C^T rounded to BF16: c_bf16t[n * ldt + m],
 * ldt >= M -- with ldt = M = the next product's K it is that product's
 * packed op(A) or op(B)^T when C is its op(A)^T or op(B).  The FP32 result
 * is the one vkh_gemm_ex computes.  a_bf16 and b_bf16 must be 16-byte
 * aligned, c_bf16t 2-byte aligned (VKH_INVALID otherwise; the same holds
 * for vkh_gemm_bf16a's a_bf16).  No counterpart in the reference: it lets
 * a captured training step hand activations and input gradients to
 * the weight-gradient products, which read them transposed, without a
 * transposing pack. */
int vkh_gemm_bf16io(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
                    uint64_t M, uint64_t N, uint64_t K, float alpha,
                    const float* A, uint64_t lda, const float* B, uint64_t ldb,
                    float beta, float* C, uint64_t ldc, const float* bias,
                    int act, uint16_t* c_bf16, const float* aux,
                    uint64_t ldaux, int aux_act, const uint16_t* a_bf16,
                    const uint16_t* b_bf16, uint16_t* c_bf16t, uint64_t ldt);

/* ---- nn fused kernels (F16, F19 and the nn package compositions) ------- */
/* y = act(x) elementwise: ReLU = x.max(0.0) (nn/layers.py:185),
 * Sigmoid = 1/(1+exp(0-x)) (nn/layers.py:240-244) */
int vkh_activation(vkh_gpu* gpu, int act, const float* x, float* y, uint64_t n);
/* dx = dy * act'(y): ReLU dy*max(sign(y),0) (nn/layers.py:207-210),
 * Sigmoid / Softmax diagonal dy*y*(1-y) (nn/layers.py:264-267, 320-323) */
int vkh_activation_backward(vkh_gpu* gpu, int act, const float* y,
                            const float* dy, float* dx, uint64_t n);
/* exact softmax gradient dx = y * (dy - sum_k dy[k] y[k]) per row of
 * y[rows, cols] = softmax(x), for the autograd wrapper (torch_autograd.py).
 * No counterpart in the reference, whose Softmax.backward is the diagonal
 * form dy * y * (1 - y) (nn/layers.py:320-323; vkh_activation_backward with
 * VKH_ACT_SIGMOID computes that form). */
int vkh_softmax_backward(vkh_gpu* gpu, const float* y, const float* dy,
                         float* dx, uint64_t rows, uint64_t cols);
/* row softmax of x[rows, cols] (nn/layers.py:297-300) */
int vkh_softmax(vkh_gpu* gpu, const float* x, float* y, uint64_t rows,
                uint64_t cols);
/* per-row cross entropy L[r] = sum_c -y*log(x+1e-8) (nn_cross_entropy.comp +
 * sum(axis=1), nn/losses.py:140-147) and dx = -y/(x+1e-8)
 * (nn_cross_entropy_backward.comp) */
int vkh_cross_entropy(vkh_gpu* gpu, const float* x, const float* y, float* L,
                      uint64_t rows, uint64_t cols);
int vkh_cross_entropy_backward(vkh_gpu* gpu, const float* x, const float* y,
                               float* dx, uint64_t n, float scale);
/* fused softmax + cross entropy (nn/losses.py:200-249): sm = softmax(x),
 * L[r] = sum_c -y*log(sm+1e-8) */
int vkh_softmax_cross_entropy(vkh_gpu* gpu, const float* x, const float* y,
                              float* sm, float* L, uint64_t rows,
                              uint64_t cols);
/* per-row losses: MSE L[r] = sum_c (y-x)^2 (nn/losses.py:294-296);
 * Huber L[r] = sum_c 0.5*min(|y-x|^2, |y-x|) (nn/losses.py:365-369) */
int vkh_mse_rows(vkh_gpu* gpu, const float* x, const float* y, float* L,
                 uint64_t rows, uint64_t cols);
int vkh_huber_rows(vkh_gpu* gpu, const float* x, const float* y, float* L,
                   uint64_t rows, uint64_t cols);
/* loss gradients, each multiplied by scale in FP32 afterwards:
 * kind 0 softmax-CE  dx = sm - y          (nn/losses.py:249)
 * kind 1 MSE         dx = (x - y) * 2     (nn/losses.py:318-320)
 * kind 2 Huber       dx = clamp(x-y,-1,1) (nn/losses.py:391-393)
 * kind 3 Huber, exact derivative of the forward 0.5*min(d^2, |d|)
 *        (nn/losses.py:365-369): dx = d for |d| <= 1, 0.5*sign(d) otherwise;
 *        used by the autograd wrapper, no counterpart in the reference */
int vkh_loss_grad(vkh_gpu* gpu, int kind, const float* x, const float* y,
                  float* dx, uint64_t n, float scale);
/* vkh_mse_rows (kind 1) or vkh_huber_rows (kind 2) and vkh_loss_grad of the
 * same kind over the same [rows, cols] x, y, in one launch with the same
 * bits.  No counterpart in the reference (a captured training step's loss
 * is one kernel). */
int vkh_loss_rows_grad(vkh_gpu* gpu, int kind, const float* x, const float* y,
                       float* L, float* dx, uint64_t rows, uint64_t cols,
                       float scale);
/* Adam update in the op order of AdamState.grad2diff + Parameter.update
 * (nn/optimizers.py:235-253, nn/parameters.py:94-95), FP32 arithmetic on
 * scalars formed like the Python ones: beta, 1-beta, 1-beta^t are computed in
 * FP64 and then rounded to FP32; b1t / b2t are the already-updated beta
 * powers */
int vkh_adam(vkh_gpu* gpu, float* w, const float* g, float* m, float* v,
             uint64_t n, float lr, double beta1, double beta2, float eps,
             double b1t, double b2t);
/* bpow[0] *= beta1, bpow[1] *= beta2 in FP64 on the device (the
 * `beta1t *= beta1` of nn/optimizers.py:241-242), so a captured training
 * step needs no host scalars */
int vkh_adam_pow_advance(vkh_gpu* gpu, double* bpow, double beta1,
                         double beta2);
/* Adam update reading the beta powers from device memory bpow[0..1];
 * w_bf16 (may be NULL) also receives the updated weights as BF16.
 * reg_kind != 0 first adds the parameter's regularizer gradient onto g and
 * writes the sum back to g (Parameter.regular_grad, nn/parameters.py:108-116),
 * then updates with it; reg_kind / reg_l1 / reg_s2 as in
 * vkh_regularizer_grad (0 = none).  Used by captured training steps. */
int vkh_adam_device(vkh_gpu* gpu, float* w, float* g, float* m, float* v,
                    uint64_t n, float lr, double beta1, double beta2, float eps,
                    const double* bpow, uint16_t* w_bf16, int reg_kind,
                    float reg_l1, float reg_s2);
/* one parameter of vkh_adam_device_multi: the arguments of vkh_adam_device
 * without the BF16 copy */
#define VKH_ADAM_MULTI_MAX 8
typedef struct vkh_adam_param {
  float* w;
  float* g;
  float* m;
  float* v;
  uint64_t n;
  float lr;
  float eps;
  double beta1;
  double beta2;
  const double* bpow;
  int reg_kind;
  float reg_l1;
  float reg_s2;
} vkh_adam_param;
/* vkh_adam_device on up to VKH_ADAM_MULTI_MAX parameters in one launch,
 * each with the bits vkh_adam_device gives it.  The table is read at the
 * call (captured graphs keep its values).  No counterpart in the reference
 * (a captured training step updates its parameters in one kernel). */
int vkh_adam_device_multi(vkh_gpu* gpu, const vkh_adam_param* params,
                          int count);
/* SGD (nn/optimizers.py:53) and AdaGrad (nn/optimizers.py:134-140), with the
 * optional regularizer term of vkh_adam_device (reg_kind 0 = none) */
int vkh_sgd(vkh_gpu* gpu, float* w, float* g, uint64_t n, float lr,
            int reg_kind, float reg_l1, float reg_s2);
int vkh_adagrad(vkh_gpu* gpu, float* w, float* g, float* h, uint64_t n,
                float lr, float eps, int reg_kind, float reg_l1, float reg_s2);
/* Regularizer gradient of the parameter w in one pass (nn/regularizers.py):
 *   kind 1 Lasso.grad    r = l1 * sign(w)                        (56-69)
 *   kind 2 Ridge.grad    r = s2 * w, s2 = 2 * coeff rounded to FP32 (113-127)
 *   kind 3 Elastic.grad  r = fl(l1 * sign(w)) + fl(s2 * w)         (177-192)
 * accumulate != 0: grad = grad + r (Parameter.regular_grad's add_grad),
 * else grad = r.  Replaces the 1 / 2 / 4 array ops of the composition. */
int vkh_regularizer_grad(vkh_gpu* gpu, int kind, const float* w, float* grad,
                         uint64_t n, float l1, float s2, int accumulate);
/* column sums of dy[rows, cols] accumulated into db (db += sum_r dy[r, :]):
 * Dense bias gradient (nn/layers.py:126-127) */
int vkh_bias_grad(vkh_gpu* gpu, const float* dy, float* db, uint64_t rows,
                  uint64_t cols, float beta);
/* vkh_gemm_bf16io that also writes C's column sums to db[N] with the bits of
 * vkh_bias_grad(C, db, M, N, beta = 0), from its epilogue (the sums of the
 * final values over 128-row chunks in row order, then the chunks in order):
 * the input gradient of a Dense layer and the previous layer's bias gradient
 * (nn/layers.py:126-141) in one product.  store_c = 0 leaves the FP32 C
 * unwritten (its BF16 copies and the sums are still produced; beta must be
 * 0).  No counterpart in the reference.  VKH_INVALID unless
 * vkh_gemm_colsum_supported says the shape is taken: the matrix-core path, a
 * configuration (tuned or default) whose waves hold 128-row slabs, no split
 * k, and vkh_bias_grad chunking [M, N] by 128 rows. */
int vkh_gemm_colsum_supported(vkh_gpu* gpu, int mode, int trans_a,
                              int trans_b, uint64_t M, uint64_t N, uint64_t K,
                              int* ok);
int vkh_gemm_bf16io_colsum(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
                           uint64_t M, uint64_t N, uint64_t K, float alpha,
                           const float* A, uint64_t lda, const float* B,
                           uint64_t ldb, float beta, float* C, uint64_t ldc,
                           const float* bias, int act, uint16_t* c_bf16,
                           const float* aux, uint64_t ldaux, int aux_act,
                           const uint16_t* a_bf16, const uint16_t* b_bf16,
                           uint16_t* c_bf16t, uint64_t ldt, float* db,
                           int store_c);
/* The general form of vkh_gemm_bf16io: db (may be NULL) as in
 * vkh_gemm_bf16io_colsum; store_c = 0 leaves the FP32 C unwritten (C may
 * then be NULL; beta must be 0; c_bf16, c_bf16t or db must be given);
 * aux_bf16 (exclusive with aux, aux_act VKH_ACT_RELU only) gives the aux
 * activation's output as BF16, [M][ldaux] (8-byte rows for the vector path):
 * the BF16 copies this library writes keep the sign, the zero-ness and the
 * NaN-ness of their values, so ReLU' read from one equals ReLU' read from the
 * FP32 value.  A captured training step uses it to drop an FP32 activation
 * whose only other readers take its BF16 copies.  No counterpart in the
 * reference. */
int vkh_gemm_bf16io_ext(vkh_gpu* gpu, int mode, int trans_a, int trans_b,
                        uint64_t M, uint64_t N, uint64_t K, float alpha,
                        const float* A, uint64_t lda, const float* B,
                        uint64_t ldb, float beta, float* C, uint64_t ldc,
                        const float* bias, int act, uint16_t* c_bf16,
                        const float* aux, const uint16_t* aux_bf16,
                        uint64_t ldaux, int aux_act, const uint16_t* a_bf16,
                        const uint16_t* b_bf16, uint16_t* c_bf16t, uint64_t ldt,
                        float* db, int store_c);
/* Backward of a last Dense layer with one output (w: [1, N]) that follows an
 * activation h = act(z) [B, N], in one pass over h, with the results of the
 * separate steps bit for bit (nn/layers.py:126-141 for the layer, 207-210 /
 * 264-267 for the activation, and the previous layer's bias gradient):
 *   dx = act'(h) * (dy w)        as vkh_gemm (K = 1) with aux = h
 *   dx_bf / dx_bft (may be NULL) its BF16 copies, row-major [B][N] and
 *                  transposed [N][ldt] (ldt % 8 == 0, 16-byte aligned)
 *   dw = dy^T h                  as vkh_gemm (1 x N x B), written (beta 0)
 *   db = column sums of dx       as vkh_bias_grad with beta 0
 * dx, dw or db NULL: not stored / not computed (db is still the column sums
 * of the FP32 dx).  No counterpart in the reference.
 * VKH_INVALID unless vkh_dense1_backward_supported(B, N) and the operands
 * are 16-byte aligned (the caller then takes the separate steps). */
int vkh_dense1_backward(vkh_gpu* gpu, const float* dy, const float* w,
                        const float* h, int act, uint64_t B, uint64_t N,
                        float* dx, uint16_t* dx_bf, uint16_t* dx_bft,
                        uint64_t ldt, float* dw, float* db);
/* *ok = 1 when vkh_dense1_backward takes a [B, N] layer on this device:
 * B % 128 == 0, N % 32 == 0 and the column-sum chunk lengths of the
 * separate steps divide 128 rows. */
int vkh_dense1_backward_supported(vkh_gpu* gpu, uint64_t B, uint64_t N, int* ok);
/* dst[i] = BF16(src[i]): round to nearest even, NaN stays NaN, and a nonzero
 * value that would round to zero (|x| < 2^-134) becomes the smallest BF16 of
 * its sign; the rounding of every BF16 operand and copy in this library */
int vkh_to_bf16(vkh_gpu* gpu, const float* src, uint16_t* dst, uint64_t n);

/* ---- graphs ------------------------------------------------------------- */
/* capture everything enqueued on the context's stream between begin and end
 * into one executable hipGraph (no counterpart in the reference) */
int vkh_graph_begin(vkh_gpu* gpu);
int vkh_graph_end(vkh_gpu* gpu, void** graph_exec);
int vkh_graph_launch(vkh_gpu* gpu, void* graph_exec);
/* waits for the stream: VKH_INVALID while the context captures a graph */
int vkh_graph_destroy(void* graph_exec);

#ifdef __cplusplus
}
#endif
#endif /* VULKPY_HIP_H */
