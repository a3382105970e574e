// Internal host-side declarations shared by the library's translation units.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

#include "force_embed.h"

namespace fe {

void set_host_error(const std::string &s);
const char *host_error();
uint64_t splitmix64_mix(uint64_t z);
uint64_t splitmix64_next(uint64_t &state);
fe_status build_csr(int64_t n, const int64_t *u, const int64_t *v, int64_t m, bool symmetrize,
                    bool drop_self, fe_graph *out);

// Work plan of one graph at one batch size: the rows of batch j
// (cpp/algo/algo.hpp:394-397) cut into work items of kItemInts ints
//   {row, e0, cnt, pslot, sid, p0, k, 0}: edges [e0, e0+cnt) of `row`.
//   pslot < 0: the item is the whole row and writes the new row.
//   pslot >= 0: chunk pslot-p0 of heavy row `row` (split sid of its batch);
//   the last of its k chunks to finish sums partials p0..p0+k-1 in order.
// A row of degree <= chunk is a wave item. A row of degree <= kSub*chunk is
// one block item. A heavier row is cut into
// k = ceil(deg / max(kSub*chunk, ceil(deg / max_chunks))) near-equal block
// items. Each batch lists its block items first, then its wave items
// (medium rows, by decreasing degree), then its light rows (degree <= light,
// by decreasing degree; their columns are also copied to the light-column
// array, see kLightCols).
struct Plan {
  int B = 0;
  int batches = 0;
  int last = 0;
  int chunk = 0;
  int max_chunks = 0;
  std::vector<int32_t> items;
  std::vector<int64_t> item_off;  // batches + 1, in items
  std::vector<int32_t> block_items; // per batch: leading block items
  std::vector<int32_t> light_rows;  // per batch: trailing light rows
  std::vector<int64_t> light_off;   // batches + 1: first light row of batch j
  int light = 0;                    // rows of degree <= light go one per lane group
  int pack = 1;                     // lane groups per wave (64 / (d / slice)) unless
                                    // a batch's cut gives its own
  std::vector<int32_t> splits;    // {row, p0, k, 0} per heavy row
  std::vector<int64_t> split_off; // batches + 1, in splits
  std::vector<int32_t> pslots;    // partial slots used by batch j
  int max_pslots = 0;
  int max_splits = 0;
  std::vector<int64_t> edges_eff; // CSR entries with c < c_limit in batch j
  int64_t edges_eff_total = 0;
};

// Cutting of one batch's rows (chunk, max_chunks, light as above). slice:
// components per lane (8 or 16), so a row is d / slice lanes and a light
// pack holds pack = 64 * slice / d rows (0: build_plan's pack).
struct Cut {
  int chunk;
  int max_chunks;
  int light;
  int slice = 8;
  int pack = 0;
};

// cuts: one per batch, or null for chunk / max_chunks / light everywhere.
fe_status build_plan(int64_t n, const int64_t *row_ptr, const int32_t *col, int B, int chunk,
                     int max_chunks, int light, int pack, int64_t c_limit, Plan *out,
                     const Cut *cuts = nullptr);

// Repulsive row counts of batch j in epoch 0 and in later epochs
// (cpp/algo/algo.hpp:135,167-169,594).
void rep_rows(const Plan &p, uint32_t quirks, int j, int *rep0, int *rep1);

} // namespace fe
