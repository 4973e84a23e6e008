/*
 * libmuscle_hip — MI355X (gfx950) backend for Muscle.jl's tensor-contraction
 * hot path: binary_einsum / binary_einsum! / hadamard / hadamard! /
 * unary_einsum / unary_einsum!.
 *
 * Plain C ABI: every argument is a C scalar, a pointer to plain data or an
 * opaque handle, so a Julia `ccall`, a Python `ctypes` call or any C caller
 * maps onto it one to one.  No exception crosses this boundary; every entry
 * point returns a muscle_hip_status_t and leaves a thread-local message for
 * muscle_hip_last_error().
 *
 * Layout convention (src/Tensor.jl:255-261 of the reference): a tensor is a
 * base pointer plus, per dimension, an integer mode label, an extent and a
 * stride counted in ELEMENTS (Julia's `strides(parent(t))`, column-major for a
 * freshly allocated array, arbitrary for views).  Mode labels are the
 * integers the Julia side gets from
 *     Dict(ind => i for (i, ind) in enumerate(unique(inds_a ∪ inds_b)))
 * exactly as ext/MuscleCUDAExt.jl:24-27 builds them for cuTENSOR.
 *
 * Device pointers are HIP device addresses.  All work is enqueued on the
 * stream set with muscle_hip_set_stream (default: the null stream); calls
 * return once the work is enqueued.
 */
#ifndef MUSCLE_HIP_H
#define MUSCLE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define MUSCLE_HIP_API __attribute__((visibility("default")))
#else
#define MUSCLE_HIP_API
#endif

/* Status codes.  The Julia/Python wrappers map them onto the reference's
 * exceptions: INVALID_ARGUMENT -> ArgumentError (binary_einsum.jl:82-83,
 * hadamard.jl:10), DIM_MISMATCH -> DimensionMismatch (Tensor.jl:16-24),
 * UNSUPPORTED -> ArgumentError("not implemented", binary_einsum.jl:53-55),
 * the rest -> ErrorException. */
typedef enum {
  MUSCLE_HIP_OK = 0,
  MUSCLE_HIP_INVALID_ARGUMENT = 1,
  MUSCLE_HIP_DIM_MISMATCH = 2,
  MUSCLE_HIP_UNSUPPORTED = 3,
  MUSCLE_HIP_HIP_ERROR = 4,
  MUSCLE_HIP_OUT_OF_MEMORY = 5,
  MUSCLE_HIP_NOT_READY = 6,
  MUSCLE_HIP_REJECTED = 7
} muscle_hip_status_t;

/* Element types: Float32, Float64, ComplexF32, ComplexF64 (src/Muscle.jl:46
 * precompiles Float64/ComplexF64 and their mixed pairs). Complex values are
 * stored interleaved (re, im) as in Julia. */
typedef enum {
  MUSCLE_HIP_F32 = 0,
  MUSCLE_HIP_F64 = 1,
  MUSCLE_HIP_C32 = 2,
  MUSCLE_HIP_C64 = 3
} muscle_hip_dtype_t;

/* Arithmetic used for binary_einsum.  NATIVE computes in the promoted
 * element type (FP64 MFMA for Float64/ComplexF64).  The others are the
 * precision ladder for Float64/ComplexF64 operands; they are accepted for
 * Float32/ComplexF32 operands too, where FP32 equals NATIVE.
 *   FP32        operands rounded to FP32, FP32 MFMA, FP32 accumulation.
 *   BF16X3      each operand split into three bfloat16 terms, the six
 *               leading cross products on BF16 MFMA, FP32 accumulation per
 *               K-chunk, chunks summed in FP64.
 *   FP16X2      two float16 terms, three products, FP32 accumulation per
 *               K-chunk, chunks summed in FP64.
 *   OZAKI_INT8  exact int32 MFMA products of int8 data.  Large contractions
 *               (M*N >= 512^2, M*N*K*batch >= 1e9, K <= 2^17 - 128) scale
 *               each row of A and column of B to a 2-norm of 2^T, round to
 *               integers and multiply them modulo up to 13 pairwise coprime
 *               moduli <= 256 (Ozaki scheme II), recombining each output by
 *               the Chinese remainder theorem in FP64; the modulus count is
 *               the smallest whose error model sqrt(K/3) 2^-T meets
 *               target_rel_error.  Complex x complex products run as two
 *               int8 GEMMs per modulus (moduli with a square root of -1)
 *               where those moduli reach the target, else as three (3M).
 *               Other shapes (or targets beyond 13 moduli)
 *               slice power-of-two-scaled rows into balanced base-255 int8
 *               digits (Ozaki scheme I) with FP64 recombination; the slice
 *               count comes from target_rel_error.
 */
typedef enum {
  MUSCLE_HIP_PRECISION_NATIVE = 0,
  MUSCLE_HIP_PRECISION_FP32 = 1,
  MUSCLE_HIP_PRECISION_BF16X3 = 2,
  MUSCLE_HIP_PRECISION_FP16X2 = 3,
  MUSCLE_HIP_PRECISION_OZAKI_INT8 = 4
} muscle_hip_precision_t;

/* One strided tensor operand. */
typedef struct {
  int32_t dtype;          /* muscle_hip_dtype_t */
  int32_t rank;           /* number of dimensions (0 for a scalar) */
  void* data;             /* device address of element (1,1,...,1) */
  const int32_t* modes;   /* [rank] mode labels */
  const int64_t* extents; /* [rank] sizes */
  const int64_t* strides; /* [rank] strides in elements */
} muscle_hip_tensor_t;

/* What the library did for one contraction (optional out-parameter). */
typedef struct {
  int32_t kernel;        /* 0 none, 1 MFMA tile GEMM, 2 reduction, 3 split-precision, 4 elementwise,
                            5 (unused),
                            6 tiled permutation (unary_einsum with no summed modes, transposed through LDS),
                            7 Ozaki int8 in modular form: int8 MFMA products modulo small moduli,
                              recombined by the Chinese remainder theorem (complex x complex: 3M,
                              three products per modulus),
                            8 as 7, complex x complex products split as Z_p[i] = Z_p x Z_p over
                              moduli with a square root of -1 (two products per modulus) */
  int32_t precision;     /* muscle_hip_precision_t actually used */
  int32_t slices;        /* Ozaki slices per operand, or moduli of the modular form (0 otherwise) */
  int32_t swapped;       /* 1 if A and B were exchanged to make C's fastest mode an N mode */
  int64_t m, n, k, batch;/* fused GEMM extents */
  double flops;          /* algorithmic real flops (8*M*N*K*batch for complex) */
  double bytes;          /* algorithmic bytes: every operand element read or written once */
  double err_bound;      /* a-priori relative error estimate of the chosen arithmetic */
} muscle_hip_info_t;

typedef struct muscle_hip_handle_s* muscle_hip_handle_t;

/* --- handle, stream, errors ------------------------------------------- */

/* Creates a handle bound to the current HIP device: plan cache, workspace.
 * Replaces no reference code: BackendCuTENSOR relies on CUDA.jl's implicit
 * task-local handle (ext/MuscleCUDAExt.jl:36-38). */
MUSCLE_HIP_API int muscle_hip_create(muscle_hip_handle_t* handle);
MUSCLE_HIP_API int muscle_hip_destroy(muscle_hip_handle_t handle);
/* Sets the hipStream_t all later calls on this handle enqueue on.  When it
 * differs from the current one, the new stream is first made to wait for
 * everything the handle enqueued on the old stream (one event), because all
 * calls of a handle share its workspace and plan tables; the old stream must
 * still exist at this call.  Calls that must overlap use separate handles. */
MUSCLE_HIP_API int muscle_hip_set_stream(muscle_hip_handle_t handle, void* stream);
MUSCLE_HIP_API int muscle_hip_synchronize(muscle_hip_handle_t handle);
/* Measures the HBM rates this library's kernels can reach on the handle's
 * device: 16-byte stores, then loads, streamed over a fresh buffer of `bytes`
 * (at least 1 MiB; allocated and freed here, synchronizes the handle's
 * stream), five passes each, in TB/s.  The ceilings that write- and
 * read-bound launches are measured against (the nominal 8 TB/s is not
 * reached by either).  Replaces no reference function. */
MUSCLE_HIP_API int muscle_hip_measure_bandwidth(muscle_hip_handle_t handle, int64_t bytes, double* read_tbs,
                                                double* write_tbs);
MUSCLE_HIP_API const char* muscle_hip_status_string(int status);
/* Detail of the last failing call on this thread. */
MUSCLE_HIP_API const char* muscle_hip_last_error(void);
MUSCLE_HIP_API int muscle_hip_version(void);
/* Handle options (muscle_hip_set_option). */
typedef enum {
  /* 3: complex x complex contractions use three real MFMA products per
   *    complex product (Gauss / 3M) -- the default;
   * 4: four real products (4M). */
  MUSCLE_HIP_OPTION_COMPLEX_MULT = 0
} muscle_hip_option_t;
MUSCLE_HIP_API int muscle_hip_set_option(muscle_hip_handle_t handle, int option, int64_t value);

/* Number of cached contraction plans, and drop them all. */
MUSCLE_HIP_API int64_t muscle_hip_plan_cache_size(muscle_hip_handle_t handle);
MUSCLE_HIP_API int muscle_hip_plan_cache_clear(muscle_hip_handle_t handle);

/* --- the hot path ------------------------------------------------------ */

/* C[modes_c] = sum over contracted modes of A[modes_a] * B[modes_b].
 *
 * Replaces binary_einsum(::Backend, inds_c, a, b) (src/Operations/binary_einsum.jl:50,
 * BackendBase at :76-96) and binary_einsum!(::Backend, c, a, b) (:68, BackendBase
 * at :98-121); contraction semantics of cuTENSOR.contract! with alpha=1,
 * beta=0, OP_IDENTITY (ext/MuscleCUDAExt.jl:22-41):
 *   - a mode in A, B and C is a batch (hyper) mode;
 *   - a mode in A and B but not C is contracted;
 *   - a mode of one operand that is absent from C is summed;
 *   - a mode repeated inside one operand is its diagonal.
 * C's dtype must be the promotion of A's and B's (MuscleCUDAExt.jl:16).
 * C is overwritten (beta = 0) and may have any mode order, unlike BackendBase's
 * binary_einsum! (binary_einsum.jl:108).  `precision` is a
 * muscle_hip_precision_t; `target_rel_error` (<= 0: default 1e-12) selects the
 * Ozaki slice count.  `info` may be NULL. */
MUSCLE_HIP_API int muscle_hip_binary_einsum(muscle_hip_handle_t handle,
                                            const muscle_hip_tensor_t* c,
                                            const muscle_hip_tensor_t* a,
                                            const muscle_hip_tensor_t* b,
                                            int precision, double target_rel_error,
                                            muscle_hip_info_t* info);

/* The same contraction as muscle_hip_binary_einsum with NATIVE arithmetic,
 * enqueued without allocating any device memory: no hipMalloc, hipFree or
 * allocating library call on any call, the first included.  It has no
 * counterpart in the reference; it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  The contraction's offset tables and its workspace (split-K
 * or reduction partial sums, casts of mixed-precision operands) live in the
 * library's static device scratch of 32 MiB (33,554,432 bytes) per GPU, held
 * from library load until the process ends; operands are gathered where they
 * lie (no packed copies).  A contraction that needs more scratch is refused
 * with MUSCLE_HIP_INVALID_ARGUMENT and a message giving both sizes.  Every
 * call on one GPU shares that scratch and leaves its contents behind: calls
 * on one handle are ordered by its stream (muscle_hip_set_stream orders a new
 * stream after the old one), but calls from different handles on one GPU
 * that overlap in time race on it.  The offset tables are copied from host
 * memory on each call.  Results are those of muscle_hip_binary_einsum on the
 * same plan. */
MUSCLE_HIP_API int muscle_hip_binary_einsum_noalloc(muscle_hip_handle_t handle,
                                                    const muscle_hip_tensor_t* c,
                                                    const muscle_hip_tensor_t* a,
                                                    const muscle_hip_tensor_t* b,
                                                    muscle_hip_info_t* info);

/* C = A .* broadcast(B), with modes(B) a subset of modes(A) = modes(C), in
 * any relative order; C has A's element type (copy(a), hadamard.jl:38) and
 * may alias A (hadamard!(a, a, b), test/unit/operations/hadamard.jl:14-17).
 * Replaces hadamard(::Backend, a, b) and hadamard!(::Backend, c, a, b)
 * (src/Operations/hadamard.jl:12,27; BackendBase at :34-77).  A real A with a
 * complex B is rejected (Julia would raise InexactError on store). */
MUSCLE_HIP_API int muscle_hip_hadamard(muscle_hip_handle_t handle,
                                       const muscle_hip_tensor_t* c,
                                       const muscle_hip_tensor_t* a,
                                       const muscle_hip_tensor_t* b,
                                       muscle_hip_info_t* info);

/* Y[modes_y] = sum over the modes of X absent from Y, repeated modes of X
 * tied (diagonal).  Replaces unary_einsum(::Backend, inds_y, x) and
 * unary_einsum!(::Backend, y, x) (src/Operations/unary_einsum.jl:35,44), whose
 * semantics are OMEinsum's einsum!((inds(x),), inds(y), ...)
 * (ext/MuscleOMEinsumExt.jl:25-38).  Y has X's element type. */
MUSCLE_HIP_API int muscle_hip_unary_einsum(muscle_hip_handle_t handle,
                                           const muscle_hip_tensor_t* y,
                                           const muscle_hip_tensor_t* x,
                                           muscle_hip_info_t* info);

/* --- factorization (the simple_update path) ----------------------------- */

/* Thin SVD of A viewed as the matrix (U's row modes) x (V's row modes):
 *   A[u..., v...] = sum_s U[u..., s] * S[s] * V[v..., s]
 * U's modes are its row modes (a subset of A's, in the order that makes them
 * the matrix rows, first fastest) followed by the bond mode s; V's likewise;
 * S has the single mode s, of extent min(rows, columns), singular values in
 * decreasing order.  V holds the conjugated right singular vectors, as the
 * reference's Vt.  U and V have A's element type, S its real type.
 * Replaces tensor_svd_thin(::Backend, A; inds_u, inds_v, ind_s)
 * (src/Operations/tensor_svd.jl:100-125; Julia LinearAlgebra.svd).  The
 * matrix SVD comes from the Hermitian eigenproblem of [[0, M], [M^H, 0]]
 * (rocSOLVER syevd/heevd) with two Cholesky-QR passes that make U and V
 * orthonormal; singular vectors of singular values at the solver's noise
 * level are an orthonormal completion, as LAPACK's are for a zero singular
 * value.  The permutations are this library's copies.  Synchronizes the
 * handle's stream twice (the eigenvalues, then the solver flags). */
MUSCLE_HIP_API int muscle_hip_svd_thin(muscle_hip_handle_t handle, const muscle_hip_tensor_t* u,
                                       const muscle_hip_tensor_t* s, const muscle_hip_tensor_t* v,
                                       const muscle_hip_tensor_t* a);

/* Thin QR of A viewed as the matrix (Q's row modes) x (R's column modes):
 *   A[q..., r...] = sum_v Q[q..., v] * R[v, r...]
 * Q's modes are its row modes followed by the bond mode v; R's are v followed
 * by its column modes; v has extent min(rows, columns).  Q has orthonormal
 * columns; R is upper trapezoidal (LAPACK's Householder sign convention, the
 * one Julia's qr uses).  Q and R have A's element type.
 * Replaces tensor_qr_thin(::Backend, A; inds_q, inds_r, ind_virtual)
 * (src/Operations/tensor_qr.jl:45-68; Julia LinearAlgebra.qr); the matrix QR
 * is rocSOLVER's geqrf + orgqr/ungqr.  Enqueues on the handle's stream. */
MUSCLE_HIP_API int muscle_hip_qr_thin(muscle_hip_handle_t handle, const muscle_hip_tensor_t* q,
                                      const muscle_hip_tensor_t* r, const muscle_hip_tensor_t* a);

/* --- batched contractions (one persistent launch) ---------------------- */

typedef struct muscle_hip_batch_s* muscle_hip_batch_t;

/* Tile shapes of the batched kernel (rows of C x columns of C per workgroup);
 * muscle_hip_batch_autotune picks one per size class. */
typedef enum {
  MUSCLE_HIP_TILE_AUTO = -1,
  MUSCLE_HIP_TILE_64x64 = 0,
  MUSCLE_HIP_TILE_32x128 = 1,
  MUSCLE_HIP_TILE_128x32 = 2,
  MUSCLE_HIP_TILE_COUNT = 3
} muscle_hip_tile_t;

typedef struct {
  int64_t pairs;      /* contractions in the batch */
  int64_t items;      /* output tiles of the batched launches */
  int32_t workgroups; /* persistent grid size */
  int32_t graph;      /* 1 once the launch has been captured in a hipGraph */
  double flops;       /* sum of algorithmic flops */
  double bytes;       /* sum of algorithmic bytes (every operand element once) */
  int64_t table_bytes;/* device bytes of offset tables and pair descriptors */
  int64_t shallow_items;/* of `items`, the tiles of pairs with K <= 16 (vector-FMA kernel) */
  int64_t split_items;/* of `items`, K slices of pairs with <= 2 output tiles and K >= 512 */
  int64_t skinny_items;/* of `items`, the items of pairs with min(m, n) <= 16, K <= 1024 and 16+ MFMA tiles (skinny vector-FMA kernel) */
} muscle_hip_batch_info_t;

/* Plans `n` independent binary contractions C_i = A_i * B_i, each with the
 * semantics of muscle_hip_binary_einsum (batch, contracted, open, summed and
 * diagonal modes), to run as ONE persistent kernel launch that works through
 * the output tiles of all of them (largest first).  This replaces a loop of
 * binary_einsum calls over the pairwise contractions of a contraction tree
 * (the callers of src/Operations/binary_einsum.jl:33-51); the reference has
 * no batched form.  All operands share one element type; NATIVE arithmetic.
 * `tiles` (may be NULL) gives a muscle_hip_tile_t per pair; AUTO uses the
 * handle's autotuned choice for the pair's size class, else a shape heuristic.
 * The buffers must stay valid while the batch exists. */
MUSCLE_HIP_API int muscle_hip_batch_create(muscle_hip_handle_t handle, int64_t n,
                                           const muscle_hip_tensor_t* cs,
                                           const muscle_hip_tensor_t* as,
                                           const muscle_hip_tensor_t* bs,
                                           const int32_t* tiles,
                                           muscle_hip_batch_t* batch);
/* mode 0: one persistent launch, captured in a hipGraph on the first run and
 *         replayed afterwards;
 * mode 1: one plain muscle_hip_binary_einsum launch per pair (the comparison
 *         loop, no interpreter in between);
 * mode 2: one persistent launch without graph capture. */
MUSCLE_HIP_API int muscle_hip_batch_run(muscle_hip_handle_t handle, muscle_hip_batch_t batch, int mode);
MUSCLE_HIP_API int muscle_hip_batch_get_info(muscle_hip_batch_t batch, muscle_hip_batch_info_t* info);

/* How one pair of a batch runs: every launch of a batch run belongs to one
 * of these routes. */
typedef enum {
  MUSCLE_HIP_PAIR_EMPTY = 0,   /* C has no elements: nothing runs */
  MUSCLE_HIP_PAIR_TILES = 1,   /* output tiles of the persistent MFMA launch */
  MUSCLE_HIP_PAIR_SPLIT_K = 2, /* few output tiles and a long K: K slices, summed in slice order */
  MUSCLE_HIP_PAIR_SHALLOW = 3, /* K <= 16: the vector-FMA launch */
  MUSCLE_HIP_PAIR_SOLO = 4,    /* 2^29 or more index points: a single-contraction launch of its own */
  MUSCLE_HIP_PAIR_SKINNY = 5   /* min(m, n) <= 16, K <= 1024, 16+ MFMA tiles, not split: the skinny vector-FMA launch (whole short side per item) */
} muscle_hip_pair_route_t;

typedef struct {
  int32_t route;      /* muscle_hip_pair_route_t */
  int32_t tile;       /* muscle_hip_tile_t of the TILES and SPLIT_K routes, -1 otherwise */
  int32_t size_class; /* muscle_hip_tuning_class(m, n) */
  int32_t ksplit;     /* K slices (1 unless SPLIT_K) */
  int64_t m, n, k, batch; /* fused extents of the pair */
} muscle_hip_batch_pair_info_t;

/* The route and fused shape of pair `pair` (0-based, in the order given to
 * muscle_hip_batch_create).  Replaces no reference function: it lets a caller
 * (and the tests) see which launch computes each pair. */
MUSCLE_HIP_API int muscle_hip_batch_get_pair_info(muscle_hip_batch_t batch, int64_t pair,
                                                  muscle_hip_batch_pair_info_t* info);
MUSCLE_HIP_API int muscle_hip_batch_destroy(muscle_hip_batch_t batch);

/* Autotuner: for every size class present in the batch, times each tile shape
 * on up to `samples` of its pairs (persistent launches, `reps` timed runs
 * after a warm-up) and records the fastest in the handle's tuning table,
 * which AUTO then uses.  The table can be saved/loaded as text. */
MUSCLE_HIP_API int muscle_hip_batch_autotune(muscle_hip_handle_t handle, muscle_hip_batch_t batch, int samples,
                                             int reps);
MUSCLE_HIP_API int muscle_hip_tuning_save(muscle_hip_handle_t handle, const char* path);
MUSCLE_HIP_API int muscle_hip_tuning_load(muscle_hip_handle_t handle, const char* path);
/* Size class of a contraction of fused extents m x n (the key of the table). */
MUSCLE_HIP_API int muscle_hip_tuning_class(int64_t m, int64_t n);

/* --- serving: contraction-tree jobs with continuous batching ------------ */

typedef struct muscle_hip_server_s* muscle_hip_server_t;

/* One contraction-tree job.  Operands are numbered 0..n_leaves-1 for the
 * leaves (device tensors owned by the caller, valid until the job is done)
 * and n_leaves + i for the output of node i.  Node i contracts operands lhs[i]
 * and rhs[i] (binary_einsum semantics) into a column-major tensor whose modes
 * are node_modes[node_off[i] .. node_off[i+1]), with extents
 * label_extents[mode].  Nodes are in topological order and the last node is
 * the root.  Every tensor of a job has the leaves' element type.  These are
 * the pairwise contractions a tensor-network caller issues as successive
 * binary_einsum calls (src/Operations/binary_einsum.jl:33-51). */
typedef struct {
  int32_t n_leaves;
  int32_t n_nodes;
  const muscle_hip_tensor_t* leaves;
  const int32_t* lhs;
  const int32_t* rhs;
  const int32_t* node_off;      /* [n_nodes + 1] */
  const int32_t* node_modes;
  int32_t n_labels;
  const int64_t* label_extents; /* [n_labels] */
} muscle_hip_job_t;

typedef struct {
  int32_t status;       /* muscle_hip_status_t of the job (NOT_READY while running) */
  int32_t dtype;
  void* result;         /* device address of the root tensor, column-major in the root's modes */
  int64_t result_elems;
  double submit_s;      /* server clock at submission, admission and completion */
  double admit_s;
  double done_s;
  int64_t ticks;        /* launches the job's contractions took part in */
} muscle_hip_job_result_t;

typedef struct {
  int64_t submitted, completed, rejected;
  int64_t ticks;             /* batched launches (mode 0) or plain calls (mode 1) */
  int64_t launched_pairs;
  int64_t pool_bytes, pool_peak_bytes;
  int64_t active_jobs, waiting_jobs;
} muscle_hip_server_stats_t;

/* Creates a server with its own stream, plan cache and a device pool of
 * pool_bytes for intermediates and results.  mode 0: continuous batching --
 * jobs are admitted in arrival order while their whole intermediate footprint
 * fits in the pool next to the admitted jobs (intermediates are freed, and
 * the reservation shrinks, as soon as their consumer completes), and every
 * scheduler tick launches, as one persistent batched kernel per element
 * type, all contractions of all admitted jobs whose operands have been
 * launched.  mode 1: the plain per-call loop -- one job at a time, one
 * binary_einsum launch per node.  max_active_jobs bounds admission. */
MUSCLE_HIP_API int muscle_hip_server_create(int64_t pool_bytes, int max_active_jobs, int mode,
                                            muscle_hip_server_t* server);
/* Copies the job description and queues it; REJECTED if its intermediates
 * can never fit in the pool.  Its contractions are lowered and planned on the
 * server's planner threads: one the library cannot run (or a job whose
 * tables push it past the pool) ends the job with that status, which wait
 * returns together with the message (muscle_hip_last_error). */
MUSCLE_HIP_API int muscle_hip_server_submit(muscle_hip_server_t server, const muscle_hip_job_t* job,
                                            int64_t* job_id);
/* Waits up to timeout_s (< 0: forever) for a job; NOT_READY on timeout. */
MUSCLE_HIP_API int muscle_hip_server_wait(muscle_hip_server_t server, int64_t job_id, double timeout_s,
                                          muscle_hip_job_result_t* result);
/* Copies a finished job's root tensor (result_elems elements) to dst (device
 * or host memory), synchronously. */
MUSCLE_HIP_API int muscle_hip_server_copy_result(muscle_hip_server_t server, int64_t job_id, void* dst);
/* Returns a finished job's result memory to the pool. */
MUSCLE_HIP_API int muscle_hip_server_release(muscle_hip_server_t server, int64_t job_id);
MUSCLE_HIP_API int muscle_hip_server_get_stats(muscle_hip_server_t server, muscle_hip_server_stats_t* stats);
/* Seconds on the server's clock (the clock of the job timestamps). */
MUSCLE_HIP_API double muscle_hip_server_clock(muscle_hip_server_t server);
/* Drains the admitted jobs and frees the server; jobs still waiting for
 * admission end with MUSCLE_HIP_HIP_ERROR. */
MUSCLE_HIP_API int muscle_hip_server_destroy(muscle_hip_server_t server);

/* Host-only: lower a binary contraction and report the fused GEMM shape and
 * the kernel the library would pick, without touching the GPU.  Replaces no
 * reference function; it exposes the index classification of
 * binary_einsum.jl:77-83 / MuscleReactantExt.jl:94-122 for inspection.
 * class_out (may be NULL) receives, per distinct mode in order of first
 * appearance in (c, a, b): 0 batch, 1 left-open (M), 2 right-open (N),
 * 3 contracted (K), 4 summed in A only, 5 summed in B only. */
MUSCLE_HIP_API int muscle_hip_describe(const muscle_hip_tensor_t* c,
                                       const muscle_hip_tensor_t* a,
                                       const muscle_hip_tensor_t* b,
                                       int precision, double target_rel_error,
                                       muscle_hip_info_t* info,
                                       int32_t* modes_out, int32_t* class_out,
                                       int32_t* nmodes_out);

#ifdef __cplusplus
}
#endif

#endif /* MUSCLE_HIP_H */
