// Device-resident state of one embedding job (one graph) and the host code
// that fills it. Shared by the single-graph context (fe_capi.cpp) and the
// multi-job engine (fe_engine.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "fe_device.hpp"
#include "fe_host.hpp"

namespace fe {

// Launch wrappers (fe_kernels.hip).
// mode: 0 every step gathers FP32, 1 every step gathers BF16, 2 mixed.
// g_const: lanes per row shared by every step of the launch (0: they differ).
// blk_step[b]: the step of block b (null: blocks binary-search the prefix,
// log2(n_steps) dependent loads). slice: components per lane of every step (each step's
// G = D / slice). blk: some step has block items (false takes the kernel
// without the block-item path; every step must then have none).
hipError_t launch_force(const StepDesc *steps, const int *blk_prefix, const int *blk_step,
                        int n_steps, int n_blocks, int wpb, int ipw, int unroll, size_t lds_bytes,
                        int mode, int g_const, int slice, bool blk, hipStream_t st);
// One step whose descriptor is passed by value (mode 0 or 1); slice: the
// components per lane (8 or 16), sd.G = sd.D / slice.
hipError_t launch_force1(const StepDesc &sd, int n_blocks, int wpb, int unroll, size_t lds_bytes,
                         int mode, int g_const, int slice, hipStream_t st);
hipError_t launch_spmm(const StepDesc *steps, const int *blk_prefix, int n_steps, int n_blocks,
                       int wpb, int unroll, int mode, int g_const, float *Y, hipStream_t st);
hipError_t launch_apply(const StepDesc *steps, const int *blk_prefix, int n_steps, int n_blocks,
                        hipStream_t st);
// set != 0: every job's epoch counter := value; else += 1. Refreshes negcur.
hipError_t launch_epoch(const EpochDesc *jobs, int n, int set, int value, int rot,
                        hipStream_t st);
// One job's state ints {epoch, rot, shift} := the values (no negcur refresh).
hipError_t launch_state(int *state, int epoch, int rot, int shift, hipStream_t st);
// dst (device) := bytes of src (host), bytes <= kRoundBlobBytes and a
// multiple of 16, by a one-block kernel whose arguments carry the bytes.
hipError_t launch_blob(const void *src, size_t bytes, void *dst, hipStream_t st);
hipError_t launch_to_bf16(const float *x, uint16_t *xs, int64_t count, hipStream_t st);
hipError_t launch_fill_lcols(const int32_t *items, int n_items, const int32_t *col,
                             int32_t *lcols, hipStream_t st);
hipError_t launch_loss_edges(const LossDesc &ld, double *block_out, int n_blocks, int ipw,
                             hipStream_t st);
hipError_t launch_loss_negs(const LossDesc &ld, double *block_out, int n_blocks,
                            int pairs_per_group, hipStream_t st);
hipError_t launch_reduce(const double *in, int n, double *out, hipStream_t st);
hipError_t launch_zero_ints(int *p, int n, hipStream_t st);
// This device's address of the static scratch array (kScratchBytes), looked
// up once per device; `device` must be the current device.
hipError_t scratch_address(int device, void **out);

// Defaults: the autotuner's choice for R-MAT scale 20 at d = 128 on MI355X.
struct Tuning {
  int chunk = 64;
  int max_chunks = 1024;
  int unroll = 2;
  int waves_per_block = 4;
  int items_per_wave = 1;
  int light = 16;
  int slice = 8; // components per lane (8 or 16): rows of d / slice lanes
};

// Launch shape of one batch's force launch (does not change results).
struct Shape {
  int unroll;
  int waves_per_block;
  int items_per_wave;
};

// Descriptor variants of one batch.
//   FUSED: a step of a run's epoch chain: new rows go straight into the free
//          slot of the layout (see Layout in fe_device.hpp).
//   SINGLE: the layout at rest is read, new rows go to S and an apply launch
//           writes them back in place (fe_step, engine rounds).
//   DELTA: as SINGLE, but S receives the delta and nothing is applied.
enum DescKind { DESC_FUSED = 0, DESC_SINGLE = 1, DESC_DELTA = 2 };

// Device memory source of a job: hipMalloc by default, or an engine arena.
struct DeviceAllocator {
  virtual ~DeviceAllocator() = default;
  virtual hipError_t alloc(void **p, size_t bytes) = 0;
  virtual void release(void *p) = 0;
};

struct Job {
  DeviceAllocator *allocator = nullptr; // null: hipMalloc / hipFree
  // shape
  int64_t n = 0;
  int d = 0;  // device row width: dx rounded up to 8 x a power of two
  int dx = 0; // row width of the caller's arrays (the d it asked for); the
              // components [dx, d) of every device row hold zeros
  int G = 0;
  int dtype = FE_GATHER_FP32;
  uint32_t quirks = FE_QUIRKS_REFERENCE;
  int64_t c_limit = 0;
  // host copies
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> skip_deg; // per row: entries with c >= c_limit
  int64_t nnz = -1;
  bool x_ready = false;
  // configuration
  int B = 0;
  float lr = 0.02f;
  float bound = 5.0f;
  int ns = 0;
  int n_lists = 0;
  Tuning tun;
  std::vector<Shape> bshape; // per batch (from the autotuner); empty: tun's shape
  std::vector<Cut> bcut;     // per batch cutting; empty: tun's chunk / max_chunks / light
  Plan plan;
  // a plan built ahead (by the engine's admission) for batch size
  // pending_B and the tuning in force; taken by the next plan rebuild
  Plan pending_plan;
  int pending_B = 0;
  std::vector<int> force_blocks; // per batch
  std::vector<int> apply_blocks; // per batch
  // device buffers
  float *X = nullptr;
  uint16_t *Xs = nullptr;
  int32_t *col = nullptr;
  int32_t *items = nullptr;
  float *S = nullptr;        // B x d: new rows or deltas of unfused steps
  float *Y = nullptr;        // n x d: SpMM output (fe_spmm)
  size_t cap_Y = 0;
  // SpMM keeps its own plan over all rows at once (its rows are independent,
  // so one launch covers the graph): items then light columns, one
  // descriptor, its block prefix; rebuilt when the CSR or the cutting changes
  int32_t *sitems = nullptr;
  size_t cap_sitems = 0;
  StepDesc *sdesc = nullptr;
  size_t cap_sdesc = 0;
  int *sprefix = nullptr;
  size_t cap_sprefix = 0;
  int s_items = 0, s_block_items = 0, s_light = 0;
  size_t s_item_ints = 0;
  uint64_t csr_gen = 0, s_gen = ~0ull;
  int s_chunk = 0, s_max_chunks = 0, s_light_cut = -1;
  float *P = nullptr;
  int *counters = nullptr;
  int32_t *negs = nullptr;
  int32_t *negcur = nullptr;   // batches x ns: ids of the current epoch's lists
  EpochDesc *edesc = nullptr;  // one entry: this job's epoch state
  int *epoch = nullptr;       // device state {epoch, rot, shift} (see Layout)
  Layout lay{};               // slot layout for the planned batch size
  int rot = 0;                // layout at rest: batch k in slot (k - rot) mod P
  int lay_B = 0;              // batch size rot refers to
  StepDesc *desc = nullptr;  // 3 * batches + 1: FUSED[b], SINGLE[b], DELTA[b], FIRST
  int *prefix = nullptr;     // per batch: {0, force blocks}, {0, apply blocks}
  std::vector<StepDesc> host_desc;
  size_t cap_items = 0, cap_S = 0, cap_P = 0, cap_ctr = 0, cap_negs = 0, cap_desc = 0;
  size_t cap_prefix = 0, cap_negcur = 0, cap_edesc = 0, cap_col = 0, cap_X = 0, cap_Xs = 0, cap_epoch = 0;
  bool plan_dirty = true;  // plan must be rebuilt (batch size, chunking, CSR)
  bool steps_dirty = true; // descriptors must be rebuilt
  uint64_t desc_gen = 0;    // bumped whenever the descriptors are rebuilt: a
                            // captured launch holds a copy of its descriptor
  int64_t device_bytes = 0;
  std::string err;
};

// All return FE_OK or an error with job.err set. `st` is the stream used.
// slot_B > 0: X (and the shadow) get every slot of the layout of batch size
// slot_B at once, so a fused run later needs no second allocation.
// The device row width for an embedding of d components (1 <= d <= 512):
// the smallest 8 x 2^k >= d. Zero components add exactly nothing to a
// distance and get exactly zero force, so the padding changes no result.
inline int padded_width(int d) {
  int w = 8;
  while (w < d) w *= 2;
  return w;
}
fe_status job_init(Job &j, int64_t n, int d, int dtype, uint32_t quirks, int slot_B = 0);
void job_free(Job &j);
fe_status job_upload_csr(Job &j, const int64_t *row_ptr, const int32_t *col, int64_t nnz,
                         hipStream_t st);
fe_status job_upload_x(Job &j, const float *x, hipStream_t st);
fe_status job_download_x(Job &j, float *x, hipStream_t st);
fe_status job_set_negatives(Job &j, const int32_t *lists, int n_lists, int ns, hipStream_t st);
fe_status job_configure(Job &j, int B, float lr, float bound);
fe_status job_set_tuning(Job &j, const Tuning &t);
// Launch shape and cutting of batch b; per-batch shapes and cuttings
// (both cleared by job_set_tuning).
Shape job_shape(const Job &j, int b);
Cut job_cut(const Job &j, int b);
fe_status job_set_shapes(Job &j, const std::vector<Shape> &shapes);
fe_status job_set_cuts(Job &j, const std::vector<Cut> &cuts);
// Builds the plan / descriptors if dirty and uploads them. slots: X (and the
// shadow) must hold every slot of the layout (fused runs); otherwise only
// the n rows of the layout at rest are needed.
fe_status job_prepare(Job &j, hipStream_t st, bool slots = false);
// Layout at rest after a fused run of `epochs` epochs.
void job_advance_layout(Job &j, int epochs);
// Moves the rows back to rotation 0 (row c at position c).
fe_status job_normalize_layout(Job &j, hipStream_t st);
// Epoch counter := value (and the current negative ids), on stream st.
fe_status job_set_epoch(Job &j, int value, hipStream_t st);
// Epoch counter += 1 (start of an epoch; may be captured in a graph).
fe_status job_tick(Job &j, hipStream_t st);
// SpMMAlgo::algo_spmm at world size 1: Y = lr * A' X (A' without the
// columns the quirk skips) into y_out, `iterations` times; X is unchanged.
fe_status job_spmm(Job &j, int iterations, float lr, float *y_out, hipStream_t st);

// Zeroes the split counters (outside any graph capture).
fe_status job_reset_counters(Job &j, hipStream_t st);
const StepDesc *job_desc(const Job &j, int b, DescKind kind);
// Force launch of batch `b` with descriptor `kind`; ev = {start, end} or null.
fe_status job_launch_force(Job &j, int b, DescKind kind, hipStream_t st, hipEvent_t *ev);
// Apply launch (S -> X) of batch `b` with descriptor `kind`.
fe_status job_launch_apply(Job &j, int b, DescKind kind, hipStream_t st);
fe_status job_loss(Job &j, uint64_t seed, int eval_ns, double *attr, double *rep, hipStream_t st);
// Device bytes a prepared job holds.
int64_t job_bytes_estimate(const Job &j);
fe_status job_check(Job &j, hipError_t e, const char *what);
int rows_of_batch(const Job &j, int b);
// Blocks of a force launch for one step: one per block item, then the wave
// items `items_per_block` (= waves per block x items per wave) at a time.
int force_blocks_for(const StepDesc &sd, int items_per_block);
// Dynamic LDS of a force launch: staged negatives plus kSub sub-range sums.
size_t force_lds_bytes(int ns, int d);
hipError_t record_event(hipEvent_t ev, hipStream_t st);

} // namespace fe
