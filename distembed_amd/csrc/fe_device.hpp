// Device-side types shared by the HIP kernels (fe_kernels.hip) and the host
// code that fills them (fe_capi.cpp). Plain C++ types only.
#pragma once
#include <stdint.h>

namespace fe {

// One minibatch step of one job (graph). Everything here is invariant across
// epochs: the epoch-dependent parts (negative list index epoch+batch,
// cpp/algo/algo.hpp:165, and the repulsive row count, :135,167-169) are
// derived in the kernels from the job's device epoch counter.
struct alignas(16) StepDesc {
  const float *X;        // FP32 master copy, n x D row-major
  float *Xw;             // same buffer, written by the apply kernel
  const uint16_t *Xs;    // BF16 shadow (gather source when bf16 != 0)
  uint16_t *Xsw;         // same buffer, refreshed by the apply kernel
  const int32_t *col;    // CSR column ids (all of the job's graph)
  const int32_t *items;  // this batch's work items, 4 ints each: {row, e0, cnt, pslot}
  const int32_t *splits; // this batch's split rows, 4 ints each: {row, p0, k, 0}
  float *S;              // rows x D: new rows (or deltas in delta mode)
  float *P;              // partial sums of heavy-row chunks
  const int32_t *negs;   // negative lists, list l at negs + l*ns
  int *epoch;            // device epoch counter of the job
  int n_items;
  int n_splits;
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
  int delta_mode;        // S receives the delta; apply is not run
  int last_batch;        // apply advances the epoch counter
  float lr;
  float bound;           // MAX_BOUND of cpp/algo/algo.hpp:65-72
  int pad0;
  int pad1;
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
