// Device-side types shared by the HIP kernels (fe_kernels.hip) and the host
// code that fills them (fe_job.cpp). Plain C++ types only.
#pragma once
#include <stdint.h>

namespace fe {

// Work item: kItemInts ints {row, e0, cnt, pslot, sid, p0, k, 0}.
//   row, edges [e0, e0+cnt) of that row;
//   pslot < 0: the item is the whole row and writes the new row;
//   pslot >= 0: the item is chunk pslot-p0 of heavy row `row`, split id sid,
//   whose k partials p0..p0+k-1 the last chunk to finish adds in index order.
// A batch's items are its block items (one per block, cut into kSub
// sub-ranges), then its wave items (one row per wave), then its light rows
// {row, e0, deg, pack_max_deg, 0, 0, 0, 1 + l} (one row per lane group,
// sorted by degree, pack_max_deg = largest degree of the row's pack). Light
// row l of the job (over all batches) has its deg <= kLightCols column ids
// copied to lcols[l * kLightCols, ...), so a light pack reads its columns
// without waiting for its items.
constexpr int kItemInts = 8;
constexpr int kSub = 4;
constexpr int kLightCols = 32;

// Row layout of a job's embedding (X and the BF16 shadow Xs). The rows are
// cut into batches of B; batch k occupies one slot of B rows, and there are
// P = batches + 1 slots, one of them free. Between runs, batch k lives in
// slot (k - rot) mod P (rot = 0: row c at position c). During epoch e of a
// run, batch j writes its new rows into the free slot, so the slot of batch
// k is (k - rot - e - [k < j]) mod P and every batch moves down one slot per
// epoch: no staging copy and no second pass over the rows. The state
// {epoch, rot, shift = (rot + epoch) mod P} sits in three device ints.
struct Layout {
  uint32_t div_m; // k = (c * div_m) >> div_sh = c / B for 0 <= c < 2^31
  int div_sh;
  int B;
  int P;
  int b_shift;    // log2(B) when B is a power of two (k = c >> b_shift), else -1
  int pad;
};

// One minibatch step of one job (graph). Everything here is invariant across
// epochs: the epoch-dependent parts (negative list index epoch+batch,
// cpp/algo/algo.hpp:165, the repulsive row count, :135,167-169, and the slot
// rotation) are derived in the kernels from the job's device state ints.
struct alignas(16) StepDesc {
  const float *X;        // FP32 master copy, P*B x D slot layout
  float *Xw;             // same buffer, written by the step or the apply
  const uint16_t *Xs;    // BF16 shadow (gather source when bf16 != 0)
  uint16_t *Xsw;         // same buffer, refreshed with the master rows
  const int32_t *col;    // CSR column ids (all of the job's graph)
  const int32_t *items;  // this batch's work items
  const int32_t *lcols;  // this batch's light rows' columns, kLightCols each
  float *S;              // rows x D: new rows or deltas (unfused steps)
  float *P;              // partial sums of heavy-row chunks
  int *counters;         // per split row of the batch: chunks finished
  const int32_t *negs;   // negative lists, list l at negs + l*ns
  const int32_t *neg_ids; // this step's ns negative ids (list epoch + batch), read
                          // without waiting for the epoch counter
  const int *state;      // {epoch, rot, shift} of the job (see Layout)
  Layout lay;
  int n_items;           // block items + wave items + light rows
  int n_block_items;
  int n_light;           // light rows (one per lane group), last in items
  int row_begin;         // batch rows [row_begin, row_begin + rows)
  int rows;
  int rep0;              // rows that get repulsion in epoch 0
  int rep1;              // ... in later epochs
  int batch;             // j
  int ns;                // negatives per list
  int n_lists;           // lists available (guards epoch + batch)
  int D;                 // embedding dimension
  int G;                 // lanes per row = D / 8
  int bf16;              // gathers read the BF16 shadow
  int c_limit;           // attractive neighbours must satisfy c < c_limit
  int delta_mode;        // S receives the delta instead of the new row
  int fused;             // 1: rows go straight to the free slot (a run's
                         //    epoch chain); 0: to S, then the apply kernel
  float lr;
  float bound;           // MAX_BOUND of cpp/algo/algo.hpp:65-72
};

// Epoch state of one job and the per-batch copies of the negative ids of
// the current epoch: negcur[b*ns + k] = negs[(epoch + b)*ns + k] (0 where the
// list does not exist), refreshed whenever the counter changes.
struct alignas(16) EpochDesc {
  int *state;           // {epoch, rot, shift}
  const int32_t *negs;
  int32_t *negcur;
  int batches;
  int ns;
  int n_lists;
  int P;                // slots of the layout
};

// Loss over one job's graph (FP64 reduction).
struct alignas(16) LossDesc {
  const float *X;
  const int32_t *col;
  const int32_t *items; // all batches' items
  Layout lay;
  int rot;              // rows between runs: batch k in slot (k - rot) mod P
  int n_items;
  int D;
  int G;
  int c_limit;
  int64_t n;
  uint64_t eval_seed;
  int eval_ns;
  int pad;
};

constexpr int kMaxLdsFloats = 8192; // staged negative rows: ns * D <= 8192
constexpr int kRoundBlobBytes = 3584; // round buffers passed as kernel arguments
constexpr int kMaxNs = 64;

// Scratch array of the step that allocates nothing (fe_step_delta_noalloc):
// kScratchPartialBytes of heavy-row partials, then kScratchCounters split
// counters.
constexpr size_t kScratchPartialBytes = (size_t)15 << 20;
constexpr int kScratchCounters = 1 << 18;
constexpr size_t kScratchBytes = kScratchPartialBytes + (size_t)kScratchCounters * sizeof(int);
constexpr int kScratchDevices = 64;

} // namespace fe
