// C++ mirror of the reference's embedding API over the force_embed.h C ABI.
//
// Same class names, constructor argument order and method signatures as the
// reference, so driver code shaped like cpp/dist_embed.cpp:155-347 compiles
// against this header with only the include lines changed:
//   EmbeddingAlgo<INDEX_TYPE, VALUE_TYPE, embedding_dim>
//       ctor                         cpp/algo/algo.hpp:54-63
//       scale(v)                     cpp/algo/algo.hpp:65-72
//       algo_force2_vec_ns(...)      cpp/algo/algo.hpp:74-262
//   DenseMat(grid, rows), nCoordinates  cpp/core/dense_mat.hpp:23-70
//   SpMat(grid), gRows/gCols/gNNz, batch_size, proc_row_width,
//       proc_col_width, initialize_CSR_blocks   cpp/core/sparse_mat.hpp:30-55
//   Process3DGrid(nr, nc, nl, adjacency)     cpp/net/process_3D_grid.hpp:12-54
//   ParallelIO::parallel_read_MM / parallel_write  cpp/io/parrallel_IO.hpp:31-146
// `distblas` is an alias of `distembed_amd`, so distblas::algo::EmbeddingAlgo
// and the other qualified names of the reference resolve here.
//
// Differences a caller sees:
//   * One GPU per process: a grid with more than one rank is rejected
//     (std::runtime_error); alpha, beta and sync_comm only steer the
//     multi-rank exchange and are accepted and ignored. col_major selects
//     the CSR the step reads, as cpp/algo/algo.hpp:131-133 does: the
//     receiver's (built transposed, cpp/dist_embed.cpp:197-200) or the
//     native one. At world size 1 the kernel is the row-major one either
//     way (cpp/algo/algo.hpp:184-186 passes col_major = false), so the
//     skip of column N-1 applies in both cases; for a symmetric graph the
//     two CSRs are equal.
//   * The arithmetic runs on the GPU in FP32 (VALUE_TYPE = double values are
//     converted on upload and download); set_gather_dtype(FE_GATHER_BF16)
//     selects the BF16-shadow form.
//   * Errors throw std::runtime_error with the library's message (the
//     reference has no error reporting).
//   * DenseMat's initial values are U(-1, 1) like the reference's
//     (cpp/core/dense_mat.hpp:59-64) but from a seeded counter-based
//     generator (seed 0 unless DenseMat::init_seed is changed and
//     reinitialize() called), so runs are reproducible.
#pragma once

#include <force_embed.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

namespace distembed_amd {

namespace detail {
inline void check(fe_status st, const fe_ctx *ctx, const char *what) {
  if (st != FE_OK)
    throw std::runtime_error(std::string(what) + ": " + fe_last_error(ctx));
}

// Resident set size of this process in MB (/proc/self/statm), 0 if unknown.
inline double resident_mb() {
  FILE *f = std::fopen("/proc/self/statm", "r");
  if (!f) return 0.0;
  long pages = 0, resident = 0;
  const int got = std::fscanf(f, "%ld %ld", &pages, &resident);
  std::fclose(f);
  if (got != 2) return 0.0;
  return (double)resident * (double)sysconf(_SC_PAGESIZE) / (1024.0 * 1024.0);
}

// "Memory usage" as the reference records it: the resident set in whole MB
// (cpp/core/common.cpp:55-70), summed after every batch and divided by the
// number of samples (cpp/algo/algo.hpp:256-259). The batches run on the
// device inside one call here, so a helper thread takes the samples, one
// per millisecond while the run lasts and one at its end.
class ResidentSampler {
public:
  void start() {
    run_.store(true);
    th_ = std::thread([this] {
      while (run_.load()) {
        add();
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
      }
    });
  }
  double stop() {
    if (th_.joinable()) {
      run_.store(false);
      th_.join();
    }
    add();
    return (double)(sum_ / count_);
  }

private:
  void add() {
    sum_ += (uint64_t)resident_mb();
    ++count_;
  }
  std::atomic<bool> run_{false};
  std::thread th_;
  uint64_t sum_ = 0, count_ = 0;
};
} // namespace detail

namespace net {
// One process per GPU: the grid only records the world size
// (cpp/net/process_3D_grid.hpp:12-54). device selects the HIP device.
class Process3DGrid {
public:
  Process3DGrid(int nr, int nc, int nl, int adjacency)
      : nr(nr), nc(nc), nl(nl), adjacency(adjacency), world_size(nr * nc * nl),
        col_world_size(nr), row_world_size(nc) {}
  int nr, nc, nl, adjacency;
  int world_size;
  int col_world_size;
  int row_world_size;
  int rank_in_col = 0;
  int rank_in_row = 0;
  int global_rank = 0;
  int device = 0;
};
} // namespace net

namespace core {
using net::Process3DGrid;

// The graph's local CSR (cpp/core/sparse_mat.hpp:30-55 with its CSRLocal,
// cpp/core/csr_local.hpp:37-125): row_ptr[gRows + 1] and col_idx[gNNz],
// columns ascending within each row.
template <typename VALUE_TYPE> class SpMat {
public:
  explicit SpMat(Process3DGrid *grid) : grid(grid) {}
  Process3DGrid *grid;
  uint64_t gRows = 0, gCols = 0, gNNz = 0;
  int batch_size = 0;
  int proc_row_width = 0;
  int proc_col_width = 0;
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col_idx;

  // Takes a CSR (row_ptr of n + 1 entries).
  void set_csr(int64_t n, const int64_t *rp, const int32_t *ci) {
    gRows = gCols = (uint64_t)n;
    gNNz = (uint64_t)rp[n];
    row_ptr.assign(rp, rp + n + 1);
    col_idx.assign(ci, ci + rp[n]);
  }
  // This matrix's CSR transposed into `out` (columns ascending per row), as
  // the reference's receiver matrix holds it (SpMat(..., transpose = true),
  // cpp/dist_embed.cpp:197-200; cpp/core/csr_local.hpp:60-66).
  void transpose_into(SpMat &out) const {
    const int64_t n = (int64_t)gRows;
    out.gRows = out.gCols = gRows;
    out.gNNz = gNNz;
    out.batch_size = batch_size;
    out.proc_row_width = proc_row_width;
    out.proc_col_width = proc_col_width;
    out.row_ptr.assign((size_t)n + 1, 0);
    for (int32_t c : col_idx) ++out.row_ptr[(size_t)c + 1];
    for (int64_t r = 0; r < n; ++r) out.row_ptr[r + 1] += out.row_ptr[r];
    out.col_idx.assign(col_idx.size(), 0);
    std::vector<int64_t> next(out.row_ptr.begin(), out.row_ptr.end() - 1);
    for (int64_t r = 0; r < n; ++r) // rows ascending: each new row's columns come out sorted
      for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
        out.col_idx[(size_t)next[col_idx[e]]++] = (int32_t)r;
  }
  // The CSR is built when the matrix is read (cpp/core/sparse_mat.hpp:405-415).
  void initialize_CSR_blocks(bool = true) {}
};

// The embedding matrix: rows x embedding_dim, row-major, host memory
// (cpp/core/dense_mat.hpp:43-64).
template <typename INDEX_TYPE, typename VALUE_TYPE, size_t embedding_dim> class DenseMat {
public:
  DenseMat(Process3DGrid *grid, INDEX_TYPE rows)
      : grid(grid), rows(rows), storage((size_t)rows * embedding_dim) {
    nCoordinates = storage.data();
    reinitialize();
  }
  void reinitialize() {
    std::vector<float> x(storage.size());
    detail::check(fe_init_uniform(x.data(), (int64_t)x.size(), -1.0f, 1.0f, init_seed), nullptr,
                  "DenseMat init");
    for (size_t i = 0; i < x.size(); ++i) storage[i] = (VALUE_TYPE)x[i];
  }
  Process3DGrid *grid;
  INDEX_TYPE rows;
  uint64_t init_seed = 0;
  std::vector<VALUE_TYPE> storage;
  VALUE_TYPE *nCoordinates;
};
} // namespace core

namespace io {
// MatrixMarket reader and embedding writer (cpp/io/parrallel_IO.hpp:31-146).
class ParallelIO {
public:
  // copy_col_to_value and the value types mirror the reference signature; the
  // CSR holds only the pattern.
  template <typename INDEX_TYPE, typename FILE_VALUE, typename VALUE_TYPE>
  void parallel_read_MM(const std::string &file, core::SpMat<VALUE_TYPE> *mat,
                        bool copy_col_to_value = false, bool = false) {
    (void)copy_col_to_value;
    fe_graph g{};
    detail::check(fe_read_mtx(file.c_str(), &g), nullptr, "parallel_read_MM");
    mat->set_csr(g.n, g.row_ptr, g.col_idx);
    fe_free_graph(&g);
  }

  // Writes `rows` rows of `nCoordinates` to `file`: "%lu" 1-based id then
  // " %.5f" per component, one line per row.
  template <typename VALUE_TYPE, typename SPMAT>
  void parallel_write(const std::string &file, const VALUE_TYPE *nCoordinates, int64_t rows,
                      int dim, core::Process3DGrid *, SPMAT *) {
    std::vector<float> x((size_t)rows * dim);
    for (size_t i = 0; i < x.size(); ++i) x[i] = (float)nCoordinates[i];
    detail::check(fe_write_embedding(file.c_str(), x.data(), rows, dim), nullptr,
                  "parallel_write");
  }
};
} // namespace io

namespace algo {
using core::DenseMat;
using core::Process3DGrid;
using core::SpMat;

template <typename INDEX_TYPE, typename VALUE_TYPE, size_t embedding_dim> class EmbeddingAlgo {
  static_assert(embedding_dim >= 1 && embedding_dim <= 512, "embedding_dim must be in [1, 512]");

protected:
  DenseMat<INDEX_TYPE, VALUE_TYPE, embedding_dim> *dense_local;
  SpMat<VALUE_TYPE> *sp_local_receiver;
  SpMat<VALUE_TYPE> *sp_local_sender;
  SpMat<VALUE_TYPE> *sp_local_native;
  Process3DGrid *grid;
  VALUE_TYPE MAX_BOUND, MIN_BOUND;
  double alpha = 1.0;
  double beta = 1.0;
  bool sync = true;
  bool col_major = true;
  int gather_dtype = FE_GATHER_FP32;
  uint32_t quirks = FE_QUIRKS_REFERENCE;
  fe_ctx *ctx = nullptr;
  std::map<std::string, double> stats;

public:
  EmbeddingAlgo(SpMat<VALUE_TYPE> *sp_local_native, SpMat<VALUE_TYPE> *sp_local_receiver,
                SpMat<VALUE_TYPE> *sp_local_sender,
                DenseMat<INDEX_TYPE, VALUE_TYPE, embedding_dim> *dense_local,
                Process3DGrid *grid, double alpha, double beta, VALUE_TYPE MAX_BOUND,
                VALUE_TYPE MIN_BOUND, bool col_major, bool sync_comm)
      : dense_local(dense_local), sp_local_receiver(sp_local_receiver),
        sp_local_sender(sp_local_sender), sp_local_native(sp_local_native), grid(grid),
        MAX_BOUND(MAX_BOUND), MIN_BOUND(MIN_BOUND), alpha(alpha), beta(beta), sync(sync_comm),
        col_major(col_major) {}
  EmbeddingAlgo(const EmbeddingAlgo &) = delete;
  EmbeddingAlgo &operator=(const EmbeddingAlgo &) = delete;
  ~EmbeddingAlgo() {
    if (ctx) fe_destroy(ctx);
  }

  VALUE_TYPE scale(VALUE_TYPE v) {
    if (v > MAX_BOUND) return MAX_BOUND;
    else if (v < -MAX_BOUND) return -MAX_BOUND;
    else return v;
  }

  // FE_GATHER_FP32 (default) or FE_GATHER_BF16; FE_QUIRKS_REFERENCE (default)
  // or any subset of its bits.
  void set_gather_dtype(int dtype) { gather_dtype = dtype; }
  void set_quirks(uint32_t q) { quirks = q; }
  fe_ctx *context() { return ctx; }
  // "Total Time" (ms, as cpp/algo/algo.hpp:75,256-261 records it),
  // "Memory usage" (mean host resident set in whole MB while the batches
  // run, see detail::ResidentSampler),
  // "Train Time" (ms, device part), "Edges+Negatives/s" and "Achieved GB/s"
  // (the force step's algorithmic bytes: gathered rows and column ids, row
  // pointers, own rows read and new rows written, the shadow rows with BF16
  // gathers and the negative rows, per epoch, over Train Time).
  const std::map<std::string, double> &perf_stats() const { return stats; }

  // iterations epochs of batch_size-row minibatches with ns negatives per
  // batch (lists seeded epoch + batch as cpp/algo/algo.hpp:165,172-173) and
  // learning rate lr; the updated embedding is written back to
  // dense_local->nCoordinates.
  void algo_force2_vec_ns(int iterations, int batch_size, int ns, VALUE_TYPE lr) {
    const auto t0 = std::chrono::steady_clock::now();
    if (grid && grid->world_size != 1)
      throw std::runtime_error("algo_force2_vec_ns: one GPU per process; a grid of " +
                               std::to_string(grid->world_size) + " ranks is not supported");
    const int64_t n = (int64_t)dense_local->rows;
    const int d = (int)embedding_dim;
    // the CSR the reference's step reads (cpp/algo/algo.hpp:131-133)
    const SpMat<VALUE_TYPE> *csr = col_major ? sp_local_receiver : sp_local_native;
    if (!csr || (int64_t)csr->gRows != n || csr->row_ptr.size() != (size_t)n + 1)
      throw std::runtime_error("algo_force2_vec_ns: CSR rows differ from the embedding rows");
    if (!ctx) {
      detail::check(fe_create(&ctx, grid ? grid->device : 0, n, d, gather_dtype, quirks),
                    nullptr, "fe_create");
      detail::check(fe_upload_csr(ctx, csr->row_ptr.data(), csr->col_idx.data(),
                                  (int64_t)csr->col_idx.size()),
                    ctx, "fe_upload_csr");
    }
    std::vector<float> x((size_t)n * d);
    for (size_t i = 0; i < x.size(); ++i) x[i] = (float)dense_local->nCoordinates[i];
    detail::check(fe_upload_x(ctx, x.data()), ctx, "fe_upload_x");
    detail::check(fe_configure(ctx, batch_size, (float)lr, (float)MAX_BOUND), ctx,
                  "fe_configure");
    const auto t1 = std::chrono::steady_clock::now();
    detail::ResidentSampler mem;
    mem.start();
    fe_status st = fe_train(ctx, iterations, batch_size, ns, (float)lr, 1);
    if (st == FE_OK) st = fe_synchronize(ctx);
    const double mem_mb = mem.stop();
    detail::check(st, ctx, "fe_train");
    const auto t2 = std::chrono::steady_clock::now();
    detail::check(fe_download_x(ctx, x.data()), ctx, "fe_download_x");
    for (size_t i = 0; i < x.size(); ++i) dense_local->nCoordinates[i] = (VALUE_TYPE)x[i];
    int64_t edges = 0, negs0 = 0, negs1 = 0;
    detail::check(fe_epoch_units(ctx, ns, &edges, &negs0, &negs1), ctx, "fe_epoch_units");
    const auto t3 = std::chrono::steady_clock::now();
    const double train_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    const double units = iterations > 0
                             ? (double)edges * iterations + negs0 + (double)negs1 * (iterations - 1)
                             : 0.0;
    stats["Total Time"] = std::chrono::duration<double, std::milli>(t3 - t0).count();
    stats["Memory usage"] = mem_mb;
    stats["Train Time"] = train_ms;
    stats["Edges+Negatives/s"] = train_ms > 0 ? units / (train_ms * 1e-3) : 0.0;
    const double s_g = gather_dtype == FE_GATHER_BF16 ? 2.0 : 4.0;
    const double batches = (double)((n + batch_size - 1) / batch_size);
    const double epoch_bytes = (double)edges * (d * s_g + 4) + (double)(n + 1) * 8 +
                               2.0 * n * d * 4 + batches * ns * d * s_g +
                               (gather_dtype == FE_GATHER_BF16 ? (double)n * d * 2 : 0.0);
    stats["Achieved GB/s"] =
        train_ms > 0 ? epoch_bytes * iterations / (train_ms * 1e-3) / 1e9 : 0.0;
  }
};

// SpMMAlgo (cpp/algo/spmm.hpp:13-52): dense_local_output = lr * A' X at
// world size 1, the SpMM sibling of the force step (fe_spmm). X is read from
// dense_local and left unchanged; A' skips the columns the reference's
// dst_id < dst_end_index skips (column N-1).
template <typename INDEX_TYPE, typename VALUE_TYPE, size_t embedding_dim> class SpMMAlgo {
  static_assert(embedding_dim >= 1 && embedding_dim <= 512, "embedding_dim must be in [1, 512]");

  DenseMat<INDEX_TYPE, VALUE_TYPE, embedding_dim> *dense_local_output;
  DenseMat<INDEX_TYPE, VALUE_TYPE, embedding_dim> *dense_local;
  SpMat<VALUE_TYPE> *sp_local_receiver;
  SpMat<VALUE_TYPE> *sp_local_sender;
  SpMat<VALUE_TYPE> *sp_local_native;
  Process3DGrid *grid;
  double alpha = 0;
  double beta = 1.0;
  bool sync = false;
  bool col_major = false;
  int gather_dtype = FE_GATHER_FP32;
  uint32_t quirks = FE_QUIRKS_REFERENCE;
  fe_ctx *ctx = nullptr;
  std::map<std::string, double> stats;

public:
  // per-row time of the reference's row loop (cpp/algo/spmm.hpp:378-388); the
  // device kernel has no per-row clock, so its time is spread evenly
  std::vector<double> timing_info;

  SpMMAlgo(SpMat<VALUE_TYPE> *sp_local_native, SpMat<VALUE_TYPE> *sp_local_receiver,
           SpMat<VALUE_TYPE> *sp_local_sender,
           DenseMat<INDEX_TYPE, VALUE_TYPE, embedding_dim> *dense_local,
           DenseMat<INDEX_TYPE, VALUE_TYPE, embedding_dim> *dense_local_output,
           Process3DGrid *grid, double alpha, double beta, bool col_major, bool sync_comm)
      : dense_local_output(dense_local_output), dense_local(dense_local),
        sp_local_receiver(sp_local_receiver), sp_local_sender(sp_local_sender),
        sp_local_native(sp_local_native), grid(grid), alpha(alpha), beta(beta),
        sync(sync_comm), col_major(col_major) {}
  SpMMAlgo(const SpMMAlgo &) = delete;
  SpMMAlgo &operator=(const SpMMAlgo &) = delete;
  ~SpMMAlgo() {
    if (ctx) fe_destroy(ctx);
  }

  void set_gather_dtype(int dtype) { gather_dtype = dtype; }
  void set_quirks(uint32_t q) { quirks = q; }
  // "Total Time" (ms), "Local SpMM" (ms, device part), "Edges/s".
  const std::map<std::string, double> &perf_stats() const { return stats; }

  void algo_spmm(int iterations, int batch_size, VALUE_TYPE lr) {
    const auto t0 = std::chrono::steady_clock::now();
    if (grid && grid->world_size != 1)
      throw std::runtime_error("algo_spmm: one GPU per process; a grid of " +
                               std::to_string(grid->world_size) + " ranks is not supported");
    const int64_t n = (int64_t)dense_local->rows;
    const int d = (int)embedding_dim;
    if ((int64_t)sp_local_native->gRows != n || sp_local_native->row_ptr.size() != (size_t)n + 1 ||
        (int64_t)dense_local_output->rows != n)
      throw std::runtime_error("algo_spmm: CSR rows differ from the dense rows");
    if (!ctx) {
      detail::check(fe_create(&ctx, grid ? grid->device : 0, n, d, gather_dtype, quirks),
                    nullptr, "fe_create");
      detail::check(fe_upload_csr(ctx, sp_local_native->row_ptr.data(),
                                  sp_local_native->col_idx.data(),
                                  (int64_t)sp_local_native->col_idx.size()),
                    ctx, "fe_upload_csr");
    }
    std::vector<float> x((size_t)n * d), y((size_t)n * d);
    for (size_t i = 0; i < x.size(); ++i) x[i] = (float)dense_local->nCoordinates[i];
    detail::check(fe_upload_x(ctx, x.data()), ctx, "fe_upload_x");
    const auto t1 = std::chrono::steady_clock::now();
    detail::check(fe_spmm(ctx, iterations, batch_size, (float)lr, y.data()), ctx, "fe_spmm");
    const auto t2 = std::chrono::steady_clock::now();
    if (iterations > 0)
      for (size_t i = 0; i < y.size(); ++i) dense_local_output->nCoordinates[i] = (VALUE_TYPE)y[i];
    const double spmm_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    timing_info.assign((size_t)n, n > 0 ? spmm_ms / (double)n : 0.0);
    stats["Total Time"] =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats["Local SpMM"] = spmm_ms;
    stats["Edges/s"] =
        spmm_ms > 0 ? (double)sp_local_native->col_idx.size() * iterations / (spmm_ms * 1e-3) : 0.0;
  }
};
} // namespace algo

} // namespace distembed_amd

namespace distblas = distembed_amd;
