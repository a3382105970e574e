// Work-plan construction (host). See Plan in fe_host.hpp.
#include <algorithm>
#include <new>

#include "fe_device.hpp"
#include "fe_host.hpp"

namespace fe {

fe_status build_plan(int64_t n, const int64_t *row_ptr, const int32_t *col, int B, int chunk,
                     int max_chunks, int light, int pack, int64_t c_limit, Plan *out,
                     const Cut *cuts) {
  if (n <= 0 || !row_ptr || B <= 0 || chunk <= 0 || max_chunks <= 0 || light < 0 || pack < 1 ||
      !out) {
    set_host_error("build_plan: invalid argument");
    return FE_ERR_INVALID_ARG;
  }
  Plan p;
  p.B = B;
  p.chunk = chunk;
  p.max_chunks = max_chunks;
  p.light = std::min(light, chunk);
  p.pack = pack;
  // cpp/algo/algo.hpp:80-88 with proc_row_width = n at world size 1.
  if (n % B == 0) {
    p.batches = (int)(n / B);
    p.last = B;
  } else {
    p.batches = (int)(n / B) + 1;
    p.last = (int)(n - (int64_t)B * (p.batches - 1));
  }
  for (int j = 0; cuts && j < p.batches; ++j) {
    if (cuts[j].chunk <= 0 || cuts[j].max_chunks <= 0 || cuts[j].light < 0 || cuts[j].pack < 0) {
      set_host_error("build_plan: invalid cutting");
      return FE_ERR_INVALID_ARG;
    }
  }
  // row_ptr: from 0, monotone, every row range within the int32 edge ids
  int bad = row_ptr[0] != 0 || row_ptr[n] >= 2147483647LL;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t r = 0; r < n; ++r) bad |= row_ptr[r + 1] < row_ptr[r] ? 1 : 0;
  if (bad) {
    set_host_error("build_plan: row_ptr must start at 0, be monotone and end below 2^31-1");
    return FE_ERR_INVALID_ARG;
  }
  // the batches are planned independently (in parallel), then concatenated
  struct BatchPlan {
    std::vector<int32_t> items;  // block items, then medium rows, then light rows
    std::vector<int32_t> splits; // {row, p0, k, 0} per heavy row
    int32_t nblk = 0, nlight = 0, pcount = 0, scount = 0;
    int64_t eff = 0;
  };
  std::vector<BatchPlan> bp;
  try {
    bp.resize((size_t)p.batches);
  } catch (...) {
    set_host_error("build_plan: out of host memory");
    return FE_ERR_OOM;
  }
  int oom = 0; // an exception cannot leave the parallel region
#pragma omp parallel for schedule(dynamic, 1) reduction(| : oom)
  for (int j = 0; j < p.batches; ++j) {
    try {
      BatchPlan &o = bp[(size_t)j];
      const int chunk_j = cuts ? cuts[j].chunk : chunk;
      const int max_chunks_j = cuts ? cuts[j].max_chunks : max_chunks;
      const int light_j = std::min(std::min(cuts ? cuts[j].light : light, chunk_j), kLightCols);
      const int pack_j = (cuts && cuts[j].pack > 0) ? cuts[j].pack : pack;
      const int64_t block_edges = (int64_t)kSub * chunk_j;
      const int64_t r0 = (int64_t)j * B;
      const int64_t r1 = std::min<int64_t>(r0 + B, n);
      std::vector<std::pair<int32_t, int32_t>> waves;  // (degree, row)
      std::vector<std::pair<int32_t, int32_t>> lights; // (degree, row)
      o.items.reserve((size_t)kItemInts * (size_t)(r1 - r0 + 16));
      for (int64_t r = r0; r < r1; ++r) {
        const int64_t e0 = row_ptr[r];
        const int64_t deg = row_ptr[r + 1] - e0;
        if (col) {
          for (int64_t e = e0; e < e0 + deg; ++e) o.eff += (col[e] < c_limit) ? 1 : 0;
        }
        if (deg <= light_j) {
          lights.push_back({(int32_t)deg, (int32_t)r});
          continue;
        }
        if (deg <= chunk_j) {
          waves.push_back({(int32_t)deg, (int32_t)r});
          continue;
        }
        if (deg <= block_edges) {
          o.items.insert(o.items.end(), {(int32_t)r, (int32_t)e0, (int32_t)deg, -1, 0, 0, 0, 0});
          ++o.nblk;
          continue;
        }
        const int64_t per = std::max<int64_t>(block_edges, (deg + max_chunks_j - 1) / max_chunks_j);
        const int64_t k = (deg + per - 1) / per;
        for (int64_t i = 0; i < k; ++i) {
          const int64_t lo = deg * i / k, hi = deg * (i + 1) / k;
          o.items.insert(o.items.end(), {(int32_t)r, (int32_t)(e0 + lo), (int32_t)(hi - lo),
                                         (int32_t)(o.pcount + i), o.scount, o.pcount, (int32_t)k, 0});
        }
        o.nblk += (int32_t)k;
        o.splits.insert(o.splits.end(), {(int32_t)r, o.pcount, (int32_t)k, 0});
        o.pcount += (int32_t)k;
        ++o.scount;
      }
      // medium and light rows by decreasing degree (row id breaks ties): the
      // grid-stride walk takes the longest units first, and the rows of a
      // light pack have similar degrees (each entry carries its pack's largest)
      auto longer = [](const std::pair<int32_t, int32_t> &a, const std::pair<int32_t, int32_t> &b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
      };
      std::sort(waves.begin(), waves.end(), longer);
      for (const auto &w : waves)
        o.items.insert(o.items.end(),
                       {w.second, (int32_t)row_ptr[w.second], w.first, -1, 0, 0, 0, 0});
      std::sort(lights.begin(), lights.end(), longer);
      for (size_t i = 0; i < lights.size(); ++i) {
        const int32_t r = lights[i].second;
        const int32_t pack_max = lights[(i / pack_j) * pack_j].first;
        // word 7: 1 + light index within the batch (the job-wide index below)
        o.items.insert(o.items.end(),
                       {r, (int32_t)row_ptr[r], lights[i].first, pack_max, 0, 0, 0, 1 + (int32_t)i});
      }
      o.nlight = (int32_t)lights.size();
    } catch (...) {
      oom = 1;
    }
  }
  if (oom) {
    set_host_error("build_plan: out of host memory");
    return FE_ERR_OOM;
  }
  try {
    p.item_off.assign(p.batches + 1, 0);
    p.block_items.assign(p.batches, 0);
    p.light_rows.assign(p.batches, 0);
    p.light_off.assign(p.batches + 1, 0);
    p.split_off.assign(p.batches + 1, 0);
    p.pslots.assign(p.batches, 0);
    p.edges_eff.assign(p.batches, 0);
    size_t n_items = 0, n_splits = 0;
    for (const BatchPlan &o : bp) {
      n_items += o.items.size();
      n_splits += o.splits.size();
    }
    p.items.reserve(n_items);
    p.splits.reserve(n_splits);
    for (int j = 0; j < p.batches; ++j) {
      BatchPlan &o = bp[(size_t)j];
      const size_t first_light = o.items.size() - (size_t)kItemInts * (size_t)o.nlight;
      for (size_t w = first_light + 7; w < o.items.size(); w += kItemInts)
        o.items[w] += (int32_t)p.light_off[j];
      p.items.insert(p.items.end(), o.items.begin(), o.items.end());
      p.splits.insert(p.splits.end(), o.splits.begin(), o.splits.end());
      p.item_off[j + 1] = (int64_t)p.items.size() / kItemInts;
      p.block_items[j] = o.nblk;
      p.light_rows[j] = o.nlight;
      p.light_off[j + 1] = p.light_off[j] + (int64_t)o.nlight;
      p.split_off[j + 1] = (int64_t)p.splits.size() / 4;
      p.pslots[j] = o.pcount;
      p.max_pslots = std::max(p.max_pslots, (int)o.pcount);
      p.max_splits = std::max(p.max_splits, (int)o.scount);
      p.edges_eff[j] = o.eff;
      p.edges_eff_total += o.eff;
      std::vector<int32_t>().swap(o.items);
    }
  } catch (...) {
    set_host_error("build_plan: out of host memory");
    return FE_ERR_OOM;
  }
  *out = std::move(p);
  return FE_OK;
}

void rep_rows(const Plan &p, uint32_t quirks, int j, int *rep0, int *rep1) {
  // considering_batch_size starts at B and becomes `last` at the final batch
  // of every epoch; without the sticky quirk it is B for the other batches.
  const bool final_batch = (j == p.batches - 1);
  *rep0 = final_batch ? p.last : p.B;
  if (final_batch) *rep1 = p.last;
  else *rep1 = (quirks & FE_QUIRK_STICKY_REP) ? p.last : p.B;
}

} // namespace fe
