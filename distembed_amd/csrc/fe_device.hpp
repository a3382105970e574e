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
// {row, e0, deg, -1, pack_max_deg, 0, 0, 0} (one row per lane group, sorted
// by degree, pack_max_deg = largest degree of the row's pack).
constexpr int kItemInts = 8;
constexpr int kSub = 4;

// One minibatch step of one job (graph). Everything here is invariant across
// epochs: the epoch-dependent parts (negative list index epoch+batch,
// cpp/algo/algo.hpp:165, and the repulsive row count, :135,167-169) are
// derived in the kernels from the job's device epoch counter.
struct alignas(16) StepDesc {
  const float *X;        // FP32 master copy, n x D row-major
  float *Xw;             // same buffer, written by the copy phase / apply
  const uint16_t *Xs;    // BF16 shadow (gather source when bf16 != 0)
  uint16_t *Xsw;         // same buffer, refreshed with the master rows
  const int32_t *col;    // CSR column ids (all of the job's graph)
  const int32_t *items;  // this batch's work items
  float *S;              // rows x D: new rows of this batch (or deltas)
  float *P;              // partial sums of heavy-row chunks
  int *counters;         // per split row of the batch: chunks finished
  const int32_t *negs;   // negative lists, list l at negs + l*ns
  const int32_t *neg_ids; // this step's ns negative ids (list epoch + batch), read
                          // without waiting for the epoch counter
  const int *epoch;      // device epoch counter of the job
  const float *Sprev;    // new rows of the previous batch, not yet in X
  uint16_t *Sb;          // BF16 gathers: this batch's new rows rounded, beside S
  const uint16_t *Sbprev; // ... of the previous batch (read instead of Xs)
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
  int prev_begin;        // rows of the previous batch held in Sprev
  int prev_rows;         // 0: no pending previous batch
  float lr;
  float bound;           // MAX_BOUND of cpp/algo/algo.hpp:65-72
};

// Epoch counter of one job and the per-batch copies of the negative ids of
// the current epoch: negcur[b*ns + k] = negs[(epoch + b)*ns + k] (0 where the
// list does not exist), refreshed whenever the counter changes.
struct alignas(16) EpochDesc {
  int *epoch;
  const int32_t *negs;
  int32_t *negcur;
  int batches;
  int ns;
  int n_lists;
  int pad;
};

// Loss over one job's graph (FP64 reduction).
struct alignas(16) LossDesc {
  const float *X;
  const int32_t *col;
  const int32_t *items; // all batches' items
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
constexpr int kMaxNs = 64;

} // namespace fe
