/*
 * force_embed.h -- C ABI of the MI355X force-model graph-embedding hot path.
 *
 * The reference exposes this path only as a header-only C++ class template,
 *   EmbeddingAlgo<INDEX_TYPE, VALUE_TYPE, embedding_dim>
 *     ctor                       cpp/algo/algo.hpp:54-63
 *     algo_force2_vec_ns(...)    cpp/algo/algo.hpp:74-262
 * called once from the driver (cpp/dist_embed.cpp:330-347), which then reads
 * DenseMat::nCoordinates back (cpp/dist_embed.cpp:346). There is no FFI in the
 * reference; the entry points below are what such an FFI would bind. Every
 * function returns an fe_status; library code never aborts. Host pointers are
 * borrowed for the duration of a call only; the context owns all device memory.
 *
 * Threading: calls on one context are serialised by the caller. Different
 * contexts (or engines) may be used from different threads.
 *
 * The C++ mirror of the reference class (same constructor and method names and
 * argument order) is include/distembed_amd/embedding_algo.hpp.
 */
#ifndef FORCE_EMBED_H
#define FORCE_EMBED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int fe_status;

enum {
  FE_OK = 0,
  FE_ERR_INVALID_ARG = 1, /* bad size, null pointer, index out of range      */
  FE_ERR_OOM = 2,         /* device or host allocation refused               */
  FE_ERR_HIP = 3,         /* a HIP runtime call failed                       */
  FE_ERR_OVER_BUDGET = 4, /* engine admission: job exceeds the HBM budget    */
  FE_ERR_STATE = 5,       /* call out of order (e.g. step before upload)     */
  FE_ERR_NO_DEVICE = 6,   /* no HIP device / device index out of range       */
  FE_ERR_TIMEOUT = 7,     /* a bounded device-side wait expired              */
  FE_ERR_NOT_FINITE = 8   /* loss reduction found a NaN or Inf               */
};

/* Dtype of the rows the gathers read. FP32: read the FP32 master copy.
 * BF16: read a BF16 shadow copy, accumulate in FP32, update the FP32 master
 * and refresh the shadow rows of each batch after its update.              */
enum { FE_GATHER_FP32 = 0, FE_GATHER_BF16 = 1 };

/* Behavioural quirks of the reference, reproduced by default.
 * STICKY_REP   cpp/algo/algo.hpp:135,167-169 -- the repulsive row count is set
 *              to the last batch's size at the final batch and never reset, so
 *              from the second epoch on every batch repels only its first
 *              `last` rows.
 * SKIP_LAST_COL cpp/algo/algo.hpp:401-405,538 -- attractive neighbours are
 *              accepted with `c < N-1`, so edges into vertex N-1 are skipped. */
enum {
  FE_QUIRK_STICKY_REP = 1u,
  FE_QUIRK_SKIP_LAST_COL = 2u,
  FE_QUIRKS_REFERENCE = 3u
};

/* How the reference-mode sampler treats a draw equal to N. The reference draws
 * from the inclusive range [0, N] (cpp/core/common.cpp:44) and an id N then
 * indexes past the end of its row cache (cpp/algo/algo.hpp:605-617), which is
 * undefined behaviour. STRICT reports FE_ERR_INVALID_ARG if any list holds N;
 * REDRAW keeps drawing from the same engine until the value is below N.   */
enum { FE_NEG_STRICT = 0, FE_NEG_REDRAW = 1 };

/* Tiles chosen by the autotuner. A zero field keeps the current value
 * (light: a negative value sets 0). chunk, max_chunks, light and slice
 * change how rows are cut into work and therefore the floating-point
 * summation order; unroll, waves_per_block and items_per_wave do not.      */
typedef struct fe_tuning {
  int chunk;           /* rows of degree <= chunk: one wave per row; up to
                          4*chunk: one block per row (<= 1024). A launch
                          with no block-row runs a kernel without that path
                          (more waves per SIMD)                             */
  int max_chunks;      /* most block chunks (partials) one heavy row may use */
  int unroll;          /* rows in flight per lane group: 1, 2, 4 or 8       */
  int waves_per_block; /* 1, 2, 4 or 8                                      */
  int items_per_wave;  /* wave work units per wave, on average (1 to 64)    */
  int light;           /* rows of degree <= light (<= 32): one row per lane
                          group, 64/(d/slice) rows per wave (rows per
                          wavefront)                                        */
  int slice;           /* components of a row per lane, 8 or 16 (<= d): a
                          row takes d/slice lanes (the dimension-chunk tile)
                                                                            */
} fe_tuning;

typedef struct fe_ctx fe_ctx;

/* ---- lifecycle -------------------------------------------------------- */

/* Creates a context on HIP device `device` for an N x d embedding,
 * 1 <= d <= 512, 0 < n < 2^31. Every host array of rows (X, deltas, SpMM
 * output) has rows of d floats. On the device a row is held at the next
 * width 8 x 2^k >= d, its extra components zero: a zero component adds
 * exactly nothing to a distance and gets exactly zero force, so the results
 * are those of width d (the bytes moved are those of the wider rows).      */
fe_status fe_create(fe_ctx **out, int device, int64_t n, int d,
                    int gather_dtype, uint32_t quirks);
void fe_destroy(fe_ctx *ctx);
/* Message of the last failure on `ctx`; with ctx NULL, of the calling
 * thread's last failed call that takes no context (fe_create and the host
 * helpers below). Never NULL.                                              */
const char *fe_last_error(const fe_ctx *ctx);

/* ---- data (replaces SpMat/CSRLocal native CSR, cpp/core/csr_local.hpp:37-125,
 *      and DenseMat::nCoordinates, cpp/core/dense_mat.hpp:56-64) ---------- */

/* CSR of the local graph: row_ptr[n+1] (monotone, row_ptr[0] = 0,
 * row_ptr[n] = nnz < 2^31), col_idx[nnz] in [0, n).                        */
fe_status fe_upload_csr(fe_ctx *ctx, const int64_t *row_ptr,
                        const int32_t *col_idx, int64_t nnz);
/* Row-major n x d FP32 embedding (also refreshes the BF16 shadow).        */
fe_status fe_upload_x(fe_ctx *ctx, const float *x);
fe_status fe_download_x(fe_ctx *ctx, float *x);

/* Negative lists: list l (l = epoch + batch, cpp/algo/algo.hpp:165) is
 * lists[l*ns .. l*ns+ns-1]; every id must be in [0, n).                    */
fe_status fe_set_negatives(fe_ctx *ctx, const int32_t *lists, int n_lists,
                           int ns);

/* Batch size, learning rate and clamp bound (MAX_BOUND, passed as 5 at
 * cpp/dist_embed.cpp:341). Builds the per-batch work plan on first use or
 * when the batch size or tiling changes.                                   */
fe_status fe_configure(fe_ctx *ctx, int batch_size, float lr, float max_bound);
fe_status fe_set_tuning(fe_ctx *ctx, const fe_tuning *t);
fe_status fe_get_tuning(const fe_ctx *ctx, fe_tuning *t);

/* ---- the hot path ------------------------------------------------------ */

/* One minibatch (epoch i, batch j) of cpp/algo/algo.hpp:165-192: attractive
 * forces over the CSR rows of the batch (:520-584), repulsive forces against
 * negative list i+j (:586-649), then X[batch] += delta (:651-666).         */
fe_status fe_step(fe_ctx *ctx, int epoch, int batch);
/* Same forces as fe_step but X is left unchanged: the rows x d delta of the
 * batch is copied to delta_out (host). rows = min(B, n - batch*B).        */
fe_status fe_step_delta(fe_ctx *ctx, int epoch, int batch, float *delta_out);
/* The step of fe_step_delta without any allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * being captured into a HIP graph, tight serving loops, harnesses that forbid
 * hipMalloc. The rows x d FP32 delta goes to device memory at delta_dev
 * (rows = min(B, n - batch*B)); the launches are enqueued on `stream` (a
 * hipStream_t; NULL: the context's stream) and the call returns without
 * waiting for them. The heavy-row partials and split counters live in one
 * static device array per GPU, held until the process ends: at most
 * 15 MiB of partials (the batch's partial slots x d x 4 bytes) and 262144
 * counters; a batch needing more gets FE_ERR_INVALID_ARG. Every call on a
 * GPU shares that array, so calls whose launches overlap on one GPU (other
 * contexts or streams included) race; callers order them, and order the
 * call after this context's own pending work. The context must
 * be prepared: one fe_step, fe_step_delta or fe_train call after the last
 * change of the CSR, X, negative lists, batch size or tuning prepares it;
 * otherwise FE_ERR_STATE. The call launches three kernels (counter fill,
 * epoch state, force) and nothing else. It needs d = 8 x 2^k (a context of
 * another d holds its rows padded: FE_ERR_INVALID_ARG).                    */
fe_status fe_step_delta_noalloc(fe_ctx *ctx, int epoch, int batch, float *delta_dev,
                                void *stream);
/* algo_force2_vec_ns(iterations, batch_size, ns, lr) (cpp/algo/algo.hpp:74).
 * Uses the lists set with fe_set_negatives when they cover
 * iterations + batches - 1 lists with `ns` ids each; otherwise generates the
 * reference lists (fe_reference_negatives, FE_NEG_REDRAW). use_graph != 0
 * replays one captured HIP graph per epoch.                                */
fe_status fe_train(fe_ctx *ctx, int iterations, int batch_size, int ns,
                   float lr, int use_graph);
fe_status fe_synchronize(fe_ctx *ctx);

/* SpMMAlgo::algo_spmm(iterations, batch_size, lr) at world size 1
 * (cpp/algo/spmm.hpp:54-159; row kernel :341-395, assignment :397-412):
 * Y[r] = sum of lr * X[c] over the CSR entries c of row r, in CSR order,
 * skipping columns c >= N-1 under FE_QUIRK_SKIP_LAST_COL as the reference's
 * dst_id < dst_end_index does. y_out (n x d, row-major) receives Y; X is not
 * changed. Batches are independent, so Y does not depend on batch_size,
 * which only sets the launches; `iterations` repeats the computation as the
 * reference's loop does. Uses the gather dtype of the context.             */
fe_status fe_spmm(fe_ctx *ctx, int iterations, int batch_size, float lr, float *y_out);

/* Force-model loss at the current X, reduced in FP64 on the device:
 *   attr = sum over CSR entries (r,c) taken by the attractive kernel of
 *          log(1 + |x_r - x_c|^2)
 *   rep  = 1/(1-1e-6) * sum over r in [0,n), k in [0,eval_ns) of
 *          log((1 + s) / (s + 1e-6)),  s = |x_r - x_m|^2, m = fe_eval_negative
 * Its negative gradient in x_r is the unclipped update of
 * cpp/algo/algo.hpp:573 and :641. The reference computes no loss.          */
fe_status fe_loss(fe_ctx *ctx, uint64_t eval_seed, int eval_ns,
                  double *attr, double *rep);

/* Work-unit counts of one epoch at the configured batch size: processed CSR
 * entries, (row, negative) pairs in the first epoch and in later epochs.   */
fe_status fe_epoch_units(fe_ctx *ctx, int ns, int64_t *edges,
                         int64_t *negs_epoch0, int64_t *negs_later);
/* Force-kernel timing with HIP events on the context's stream, reset by
 * every call. enable 1: an event pair around every force launch (the
 * per-launch view; the extra event nodes add about 2 us per launch).
 * enable 2: one event pair around each fe_train call, divided by the force
 * launches it ran (the epoch ticks, one short launch per epoch, included).
 * enable 0: off. fe_kernel_times reports the average.                      */
fe_status fe_set_timing(fe_ctx *ctx, int enable);

/* Autotuner (no counterpart in the reference). Trials time `epochs` epochs of
 * the captured chain after one warm-up epoch, each force launch timed with
 * HIP events; X is restored after every trial. A batch's launch time
 * depends only on its own tiling, so each trial serves every batch:
 * (1) every launch shape -- unroll {1,2,4} x waves per block {2,4,8} x
 * items per wave {1,2,4} -- on all batches: each batch keeps its fastest;
 * (2) with search_chunking, every cutting -- slice {8,16} (16 when
 * d >= 16) x chunk {16,32,64,128,256,512} x max_chunks {64,1024} x light
 * {0,4,8,16,32} -- the same way, then (1) again. The
 * context keeps the per-batch result (fe_batch_tuning); *best holds the
 * values most batches chose and *best_ms the resulting ms per epoch. Shapes
 * never change results; the cutting changes the summation order of heavy
 * rows. Up to max_results trials are reported in results / results_ms (the
 * sum of the batches' mean launch times, ms).
 * Needs negative lists covering epochs + batches - 1 lists.                */
fe_status fe_autotune(fe_ctx *ctx, int epochs, int search_chunking,
                      fe_tuning *results, double *results_ms, int max_results,
                      int *n_results, fe_tuning *best, double *best_ms);
/* The tuning one batch's force launch uses (its own after fe_autotune or
 * fe_set_batch_tuning; fe_set_tuning gives every batch the same).          */
fe_status fe_batch_tuning(const fe_ctx *ctx, int batch, fe_tuning *out);
/* Sets one batch's tuning (zero fields keep the current value; light -1
 * sets 0). Needs the plan (CSR, X and batch size set).                     */
fe_status fe_set_batch_tuning(fe_ctx *ctx, int batch, const fe_tuning *t);
fe_status fe_kernel_times(fe_ctx *ctx, double *force_ms, double *fixup_ms,
                          double *apply_ms, int64_t *launches);
/* Rounds n floats to BF16 on device `device` exactly as the BF16 shadow
 * holds them (round to nearest even, subnormals kept, a NaN stays a NaN);
 * out receives the 16-bit patterns. No counterpart in the reference: it
 * exposes the shadow's rounding to callers and tests.                      */
fe_status fe_round_bf16(int device, const float *x, uint16_t *out, int64_t n);

/* ---- host helpers (no device needed) ---------------------------------- */

/* generate_random_numbers(0, n, seed, ns) (cpp/core/common.cpp:37-53): a
 * minimal-standard LCG (multiplier 48271, modulus 2^31-1) seeded with seed,
 * downscaled to the inclusive range [0, n] as libstdc++'s
 * uniform_int_distribution<int> does. Writes lists l = 0..n_lists-1 with
 * seed l into out[l*ns ...]. Returns the number of draws equal to n (before
 * redraw) in *n_hits when n_hits is not NULL.                              */
fe_status fe_reference_negatives(int64_t n, int n_lists, int ns, int policy,
                                 int32_t *out, int64_t *n_hits);
/* Evaluation negative k of row r for the loss: splitmix64 of
 * (seed, r * 2^20 + k) reduced modulo n (a definition of this library).   */
int64_t fe_eval_negative(uint64_t seed, int64_t r, int k, int64_t n);

/* Graph generators (deterministic in seed, independent of thread count).
 * They return a symmetric CSR with sorted, de-duplicated columns and no
 * self-loops. Arrays are malloc'ed by the library and released with
 * fe_free_graph. Vertex order is the generator's natural order.           */
typedef struct fe_graph {
  int64_t n;
  int64_t nnz;
  int64_t *row_ptr; /* n + 1 */
  int32_t *col_idx; /* nnz   */
} fe_graph;
fe_status fe_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                      uint64_t seed, fe_graph *out);
/* Stochastic block model: n vertices in `blocks` near-equal contiguous
 * blocks, pairs inside a block joined with p_in, across blocks with p_out. */
fe_status fe_gen_sbm(int64_t n, int blocks, double p_in, double p_out,
                     uint64_t seed, fe_graph *out);
/* CSR from an edge list (u[i], v[i]) in [0, n): symmetrised, sorted,
 * duplicates merged, self-loops dropped.                                   */
fe_status fe_csr_from_edges(int64_t n, const int64_t *u, const int64_t *v,
                            int64_t m, fe_graph *out);
void fe_free_graph(fe_graph *g);
/* X0 ~ U(lo, hi) from a counter-based generator (seed, element index).    */
fe_status fe_init_uniform(float *x, int64_t count, float lo, float hi,
                          uint64_t seed);

/* MatrixMarket ingest (cpp/io/parrallel_IO.hpp:31-88): square coordinate
 * matrix, 1-based ids; `symmetric` files are expanded to both directions,
 * `general` entries are kept as given; columns sorted, duplicates merged,
 * values ignored, self-loops kept (they add a zero force).                 */
fe_status fe_read_mtx(const char *path, fe_graph *out);
/* Embedding writer (cpp/io/parrallel_IO.hpp:90-146): one line per row,
 * "%lu" 1-based id then " %.5f" per component, no header.                  */
fe_status fe_write_embedding(const char *path, const float *x, int64_t n,
                             int d);

/* ---- multi-job engine: batching and serving forms ------------------------
 * No counterpart in the reference, which embeds one graph per process run
 * (cpp/dist_embed.cpp:330-347); each job below is one such run.
 * An engine owns one device, three streams (rounds, uploads, downloads) and
 * an HBM budget held as one arena. Jobs (independent graphs with their own
 * CSR, X, d, gather dtype, batch size, epochs and negative lists) are
 * admitted while their device state fits the free budget. Every engine round
 * advances each running job by one minibatch, all jobs' work items sharing
 * one fused force launch. A job's chunking is fixed when it is submitted and
 * no job touches another job's memory, so its result does not depend on
 * which other jobs share its rounds.
 * Threads: fe_engine_submit and fe_engine_collect may run on other threads
 * while one thread calls fe_engine_run; they build, upload and download
 * without blocking the rounds. The other calls are serialised internally.
 * fe_engine_last_error reports the calling thread's last failure.          */
typedef struct fe_engine fe_engine;
typedef struct fe_job_desc {
  int64_t n;
  int d;
  int gather_dtype;
  uint32_t quirks;
  int batch_size;
  int iterations;
  int ns;
  float lr;
  float max_bound;
  const int64_t *row_ptr;   /* borrowed during fe_engine_submit only         */
  const int32_t *col_idx;
  int64_t nnz;
  const float *x0;          /* n x d                                          */
  const int32_t *neg_lists; /* (iterations + batches - 1) x ns, or NULL for
                               the reference lists (FE_NEG_REDRAW)            */
} fe_job_desc;

fe_status fe_engine_create(fe_engine **out, int device, int64_t hbm_budget_bytes);
void fe_engine_destroy(fe_engine *e);
const char *fe_engine_last_error(const fe_engine *e);
/* Bytes of the budget the job would hold while admitted (-1: invalid). */
int64_t fe_engine_job_bytes(const fe_engine *e, const fe_job_desc *desc);
/* Copies the job to the device (upload stream) if it fits the free budget,
 * else FE_ERR_OVER_BUDGET with nothing kept. *job_id names it afterwards. */
fe_status fe_engine_submit(fe_engine *e, const fe_job_desc *desc, int *job_id);
/* Enqueues `rounds` rounds (all remaining when rounds <= 0); *running gets
 * the number of jobs with minibatches still to enqueue.                    */
fe_status fe_engine_run(fe_engine *e, int rounds, int *running);
/* Batching form: runs every submitted job to completion from a schedule
 * built up front, its rounds captured in HIP graphs and replayed back to
 * back. Timing of the last call: fe_engine_graph_timing. The schedule
 * (about 232 bytes per job per round) is a device allocation of its own,
 * outside the budget, freed before the call returns.                      */
fe_status fe_engine_run_graph(fe_engine *e);
fe_status fe_engine_graph_timing(const fe_engine *e, double *prepare_ms,
                                 double *exec_ms, int64_t *rounds);
/* *done = 1 once the job's last minibatch has finished on the device (or a
 * round holding the job failed to enqueue: collect then reports it).      */
fe_status fe_engine_job_done(fe_engine *e, int job_id, int *done);
/* Copies a finished job's X (n x d) to host and returns its memory to the
 * budget. Waits only for that job's last round. A job whose round failed in
 * fe_engine_run or fe_engine_run_graph has no result: FE_ERR_STATE, and its
 * memory returns to the budget.                                           */
fe_status fe_engine_collect(fe_engine *e, int job_id, float *x_out);
/* Launch shape for later rounds and chunking for later submissions. One
 * kernel serves a round, so every job runs in the engine's slice: a new
 * slice needs an engine that holds no uncollected job (FE_ERR_STATE), and
 * a job with d below the slice is refused at submission.                 */
fe_status fe_engine_set_tuning(fe_engine *e, const fe_tuning *t);
/* Autotuner of the batching form's tiles (no counterpart in the reference):
 * the n_jobs jobs (their iterations replaced by `epochs`) are submitted once
 * per cutting -- slice {8,16} (16 when every job has d >= 16) x light
 * {8,16,32} at chunk 64, max_chunks 1024 -- and replayed
 * through fe_engine_run_graph once per launch shape -- unroll {2,4,8} x
 * waves per block {4,8} x items per wave {1,2}; the fastest replay (device
 * time, HIP events) becomes the engine's tuning. The engine must hold no
 * uncollected job, and the n_jobs jobs must fit its budget together
 * (FE_ERR_OVER_BUDGET otherwise); the trial jobs' results are discarded. Up to max_results
 * candidates go to results / results_ms (replay ms); *best and *best_ms get
 * the winner.                                                             */
fe_status fe_engine_autotune(fe_engine *e, const fe_job_desc *jobs, int n_jobs,
                             int epochs, fe_tuning *results, double *results_ms,
                             int max_results, int *n_results, fe_tuning *best,
                             double *best_ms);
fe_status fe_engine_stats(const fe_engine *e, int64_t *bytes_in_use,
                          int64_t *rounds_run, int64_t *units_done);

#ifdef __cplusplus
}
#endif
#endif /* FORCE_EMBED_H */
