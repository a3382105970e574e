// ORACLE -- test infrastructure only. CPU restatement of the reference's
// world-size-1 force-embedding path, used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, to check the
// GPU results and as the CPU baseline. Never linked into, or called by, the
// product.
//
// Follows (all paths relative to the reference root):
//   cpp/algo/algo.hpp:74-192     driver loop: batches, last batch, seeds i+j,
//                                 considering_batch_size (sticky, :135,167-169)
//   cpp/algo/algo.hpp:388-448    source/destination ranges (:394-405)
//   cpp/algo/algo.hpp:520-584    attractive kernel, `c < dst_end` (:538)
//   cpp/algo/algo.hpp:586-649    repulsive kernel
//   cpp/algo/algo.hpp:651-666    update X[batch] += prev
//   cpp/algo/algo.hpp:65-72      scale() clamp
//   cpp/core/common.cpp:37-53    generate_random_numbers (orc_std_sampler)
//   cpp/algo/spmm.hpp:341-412    SpMMAlgo row kernel and assignment (orc_spmm_*)
// The arithmetic keeps the reference's expressions: the squared distance is a
// serial sum in VALUE_TYPE, d1 is computed from double literals and converted
// to VALUE_TYPE, and each component accumulates sequentially (edges in CSR
// order, then negatives in list order). Built with -ffp-contract=off so no
// fused multiply-add changes the rounding. VALUE_TYPE = double is the
// reference's own precision (cpp/core/common.h:35) and the FP64 oracle.
//
// The BF16-shadow form (orc_train_bf16) has no counterpart in the reference:
// every read of X inside a batch sees X rounded to bfloat16 (round to nearest
// even), the forces are computed from those values in float exactly as above,
// and the batch's update is added to the float master X.
//
// SpMM (orc_spmm_*): out[r][k] = sum over e in row r (CSR order) with
// col[e] < c_limit of lr * X[col[e]][k], accumulated serially in VALUE_TYPE
// from zero (cpp/algo/spmm.hpp:122-127, 379-385), then assigned (:397-412).
// Pinned against scipy.sparse's CSR product and a closed form.
//
// Pinning (tests/test_oracle_*.py): the sampler against std:: compiled with
// g++; the per-batch update against torch autograd of the force-model loss in
// FP64 (an independent route to the same update) and against hand-computed
// closed forms (2-vertex graph, zero distance, saturation); the loop structure
// (Jacobi inside a batch, Gauss-Seidel across batches, sticky repulsive count,
// skipped column n-1, seeds i+j) against invariants each mistake would break.
#include <omp.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace {

template <typename T> inline T scale_clamp(T v, T bound) {
  if (v > bound) return bound;
  else if (v < -bound) return -bound;
  else return v;
}

struct Shape {
  int64_t n;
  int d;
  int B;
  int batches;
  int last;
};

inline Shape make_shape(int64_t n, int d, int B) {
  Shape s{n, d, B, 0, B};
  if (n % B == 0) {
    s.batches = (int)(n / B);
  } else {
    s.batches = (int)(n / B) + 1;
    s.last = (int)(n - (int64_t)B * (s.batches - 1));
  }
  return s;
}

// One batch: zero prev, K1, K2, optional K3. `block_size` is the reference's
// considering_batch_size for this (epoch, batch).
template <typename T>
void batch_step(const Shape &sh, const int64_t *rp, const int32_t *col, T *X, int j,
                int block_size, const int32_t *neg, int ns, T lr, T bound, unsigned quirks,
                T *prev, bool apply) {
  const int d = sh.d;
  const int64_t n = sh.n;
  const int64_t src_start = (int64_t)j * sh.B;
  const int64_t src_end = std::min<int64_t>((int64_t)(j + 1) * sh.B, n) - 1;
  // dst window [0, min(proc_col_width, gCols) - 1] = [0, n-1]; the test is
  // c >= dst_start && c < dst_end, so c = n-1 is skipped (quirk bit 2).
  const int64_t dst_end = (quirks & 2u) ? n - 1 : n;
  for (int64_t k = 0; k < (int64_t)sh.B * d; ++k) prev[k] = 0;

#pragma omp parallel for schedule(static)
  for (int64_t i = src_start; i <= src_end; i++) {
    const int64_t index = i - src_start;
    std::vector<T> forceDiff(d);
    for (int64_t e = rp[i]; e < rp[i + 1]; e++) {
      const int64_t dst_id = col[e];
      if (dst_id >= 0 && dst_id < dst_end) {
        T attrc = 0;
        for (int k = 0; k < d; k++) {
          forceDiff[k] = X[i * d + k] - X[dst_id * d + k];
          attrc += forceDiff[k] * forceDiff[k];
        }
        T d1 = -2.0 / (1.0 + attrc);
        for (int k = 0; k < d; k++) {
          T l = scale_clamp<T>(forceDiff[k] * d1, bound);
          prev[index * d + k] = prev[index * d + k] + (lr)*l;
        }
      }
    }
  }

#pragma omp parallel for schedule(static)
  for (int i = 0; i < block_size; i++) {
    const int64_t row_id = (int64_t)i + src_start;
    std::vector<T> forceDiff(d);
    for (int jn = 0; jn < ns; jn++) {
      const int64_t c = neg[jn];
      T repuls = 0;
      for (int k = 0; k < d; k++) {
        forceDiff[k] = X[row_id * d + k] - X[c * d + k];
        repuls += forceDiff[k] * forceDiff[k];
      }
      T d1 = 2.0 / ((repuls + 0.000001) * (1.0 + repuls));
      for (int k = 0; k < d; k++) {
        forceDiff[k] = scale_clamp<T>(forceDiff[k] * d1, bound);
        prev[i * d + k] += (lr)*forceDiff[k];
      }
    }
  }

  if (apply) {
    const int rows = (int)(src_end - src_start + 1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rows; i++)
      for (int k = 0; k < d; k++) X[(src_start + i) * d + k] += prev[(int64_t)i * d + k];
  }
}

template <typename T>
int check_lists(int64_t n, const int32_t *lists, int64_t count) {
  for (int64_t i = 0; i < count; ++i)
    if (lists[i] < 0 || lists[i] >= n) return -2; // id n is undefined behaviour in the reference
  return 0;
}

template <typename T>
int step_impl(int64_t n, int d, const int64_t *rp, const int32_t *col, T *X, int B, int epoch,
              int batch, const int32_t *lists, int n_lists, int ns, T lr, T bound,
              unsigned quirks, T *delta_out, int apply, int threads) {
  if (n <= 0 || d <= 0 || B <= 0 || epoch < 0 || ns < 0) return -1;
  const Shape sh = make_shape(n, d, B);
  if (batch < 0 || batch >= sh.batches) return -1;
  const int li = epoch + batch;
  if (ns > 0 && li >= n_lists) return -1;
  const int32_t *neg = ns > 0 ? lists + (int64_t)li * ns : nullptr;
  if (ns > 0 && check_lists<T>(n, neg, ns)) return -2;
  // considering_batch_size at (epoch, batch) of the loop in orc_train.
  int block_size = sh.B;
  if (batch == sh.batches - 1) block_size = sh.last;
  else if ((quirks & 1u) && epoch > 0) block_size = sh.last;
  if (threads > 0) omp_set_num_threads(threads);
  std::vector<T> prev((size_t)sh.B * d);
  batch_step<T>(sh, rp, col, X, batch, block_size, neg, ns, lr, bound, quirks, prev.data(),
                apply != 0);
  if (delta_out) {
    const int rows = (int)(std::min<int64_t>((int64_t)(batch + 1) * B, n) - (int64_t)batch * B);
    for (int64_t k = 0; k < (int64_t)rows * d; ++k) delta_out[k] = prev[k];
  }
  return 0;
}

// algo_force2_vec_ns at world size 1 (cpp/algo/algo.hpp:74-192).
template <typename T>
int train_impl(int64_t n, int d, const int64_t *rp, const int32_t *col, T *X, int B,
               int iterations, const int32_t *lists, int n_lists, int ns, T lr, T bound,
               unsigned quirks, int threads) {
  if (n <= 0 || d <= 0 || B <= 0 || iterations < 0 || ns < 0) return -1;
  const Shape sh = make_shape(n, d, B);
  if (ns > 0 && (int64_t)iterations + sh.batches - 1 > n_lists) return -1;
  if (ns > 0 && check_lists<T>(n, lists, (int64_t)(iterations + sh.batches - 1) * ns)) return -2;
  if (threads > 0) omp_set_num_threads(threads);
  std::vector<T> prev((size_t)sh.B * d);
  int considering_batch_size = sh.B;
  for (int i = 0; i < iterations; i++) {
    for (int j = 0; j < sh.batches; j++) {
      const int seed = j + i;
      if (j == sh.batches - 1) considering_batch_size = sh.last;
      else if (!(quirks & 1u)) considering_batch_size = sh.B; // reset when the quirk is off
      const int32_t *neg = ns > 0 ? lists + (int64_t)seed * ns : nullptr;
      batch_step<T>(sh, rp, col, X, j, considering_batch_size, neg, ns, lr, bound, quirks,
                    prev.data(), true);
    }
  }
  return 0;
}

// float -> bfloat16 -> float, round to nearest even; NaN stays NaN.
inline float bf16_round(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) {
    u |= 0x00400000u;
  } else {
    u += 0x7FFFu + ((u >> 16) & 1u);
  }
  u &= 0xFFFF0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

// The BF16-shadow form of train_impl<float>: each batch computes its forces
// on a rounded copy of X and adds them to X.
int train_bf16_impl(int64_t n, int d, const int64_t *rp, const int32_t *col, float *X, int B,
                    int iterations, const int32_t *lists, int n_lists, int ns, float lr,
                    float bound, unsigned quirks, int threads) {
  if (n <= 0 || d <= 0 || B <= 0 || iterations < 0 || ns < 0) return -1;
  const Shape sh = make_shape(n, d, B);
  if (ns > 0 && (int64_t)iterations + sh.batches - 1 > n_lists) return -1;
  if (ns > 0 && check_lists<float>(n, lists, (int64_t)(iterations + sh.batches - 1) * ns))
    return -2;
  if (threads > 0) omp_set_num_threads(threads);
  std::vector<float> prev((size_t)sh.B * d), R((size_t)n * d);
  int considering_batch_size = sh.B;
  for (int i = 0; i < iterations; i++) {
    for (int j = 0; j < sh.batches; j++) {
      if (j == sh.batches - 1) considering_batch_size = sh.last;
      else if (!(quirks & 1u)) considering_batch_size = sh.B;
      const int32_t *neg = ns > 0 ? lists + (int64_t)(i + j) * ns : nullptr;
      for (int64_t k = 0; k < n * d; ++k) R[k] = bf16_round(X[k]);
      batch_step<float>(sh, rp, col, R.data(), j, considering_batch_size, neg, ns, lr, bound,
                        quirks, prev.data(), false);
      const int64_t r0 = (int64_t)j * sh.B;
      const int64_t rows = std::min<int64_t>(r0 + sh.B, n) - r0;
      for (int64_t k = 0; k < rows * d; ++k) X[r0 * d + k] += prev[k];
    }
  }
  return 0;
}

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

} // namespace

// SpMMAlgo::algo_spmm at world size 1: rows are independent, so batching
// does not enter. quirks bit 2 (FE_QUIRK_SKIP_LAST_COL) skips c >= n-1.
template <typename T>
static void spmm_impl(int64_t n, int d, const int64_t *rp, const int32_t *col, const T *X, T lr,
                      unsigned quirks, T *out, int threads) {
  const int64_t c_limit = (quirks & 2u) ? n - 1 : n;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    T *o = out + r * d;
    for (int k = 0; k < d; ++k) o[k] = T(0);
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
      const int64_t c = col[e];
      if (c >= c_limit) continue;
      for (int k = 0; k < d; ++k) o[k] += lr * X[c * d + k];
    }
  }
}

extern "C" {

// The reference's sampler call, verbatim in meaning: std::minstd_rand seeded
// with `seed`, std::uniform_int_distribution<int>(lo, hi), ns draws.
void orc_std_sampler(int lo, int hi, int seed, int ns, int64_t *out) {
  std::minstd_rand generator(seed);
  std::uniform_int_distribution<int> distribution(lo, hi);
  for (int i = 0; i < ns; ++i) out[i] = (int64_t)distribution(generator);
}

int64_t orc_eval_negative(uint64_t seed, int64_t r, int k, int64_t n) {
  const uint64_t x = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)r * 1048576ull + (uint64_t)k + 1ull);
  return (int64_t)(mix64(x) % (uint64_t)n);
}

int orc_step_f32(int64_t n, int d, const int64_t *rp, const int32_t *col, float *X, int B,
                 int epoch, int batch, const int32_t *lists, int n_lists, int ns, float lr,
                 float bound, unsigned quirks, float *delta_out, int apply, int threads) {
  return step_impl<float>(n, d, rp, col, X, B, epoch, batch, lists, n_lists, ns, lr, bound,
                          quirks, delta_out, apply, threads);
}

int orc_step_f64(int64_t n, int d, const int64_t *rp, const int32_t *col, double *X, int B,
                 int epoch, int batch, const int32_t *lists, int n_lists, int ns, double lr,
                 double bound, unsigned quirks, double *delta_out, int apply, int threads) {
  return step_impl<double>(n, d, rp, col, X, B, epoch, batch, lists, n_lists, ns, lr, bound,
                           quirks, delta_out, apply, threads);
}

int orc_train_f32(int64_t n, int d, const int64_t *rp, const int32_t *col, float *X, int B,
                  int iterations, const int32_t *lists, int n_lists, int ns, float lr,
                  float bound, unsigned quirks, int threads) {
  return train_impl<float>(n, d, rp, col, X, B, iterations, lists, n_lists, ns, lr, bound,
                           quirks, threads);
}

int orc_train_f64(int64_t n, int d, const int64_t *rp, const int32_t *col, double *X, int B,
                  int iterations, const int32_t *lists, int n_lists, int ns, double lr,
                  double bound, unsigned quirks, int threads) {
  return train_impl<double>(n, d, rp, col, X, B, iterations, lists, n_lists, ns, lr, bound,
                            quirks, threads);
}

void orc_bf16_round(const float *in, float *out, int64_t count) {
  for (int64_t i = 0; i < count; ++i) out[i] = bf16_round(in[i]);
}

int orc_train_bf16(int64_t n, int d, const int64_t *rp, const int32_t *col, float *X, int B,
                   int iterations, const int32_t *lists, int n_lists, int ns, float lr,
                   float bound, unsigned quirks, int threads) {
  return train_bf16_impl(n, d, rp, col, X, B, iterations, lists, n_lists, ns, lr, bound, quirks,
                         threads);
}

// Force-model loss in FP64: attractive log1p(s) over CSR entries with
// c < c_limit, repulsive log((1+s)/(s+eps))/(1-eps) over the evaluation
// pairs (r, orc_eval_negative(seed, r, k, n)), k < eval_ns.
int orc_loss_f64(int64_t n, int d, const int64_t *rp, const int32_t *col, const double *X,
                 unsigned quirks, uint64_t seed, int eval_ns, double *attr, double *rep,
                 int threads) {
  if (n <= 0 || d <= 0 || eval_ns < 0) return -1;
  if (threads > 0) omp_set_num_threads(threads);
  const int64_t c_limit = (quirks & 2u) ? n - 1 : n;
  const double eps = 1e-6;
  double a = 0.0, r = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : a, r)
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
      const int64_t c = col[e];
      if (c >= c_limit) continue;
      double s = 0.0;
      for (int k = 0; k < d; ++k) {
        const double v = X[i * d + k] - X[c * d + k];
        s += v * v;
      }
      a += std::log1p(s);
    }
    for (int k = 0; k < eval_ns; ++k) {
      const int64_t m = orc_eval_negative(seed, i, k, n);
      double s = 0.0;
      for (int q = 0; q < d; ++q) {
        const double v = X[i * d + q] - X[m * d + q];
        s += v * v;
      }
      r += std::log((1.0 + s) / (s + eps)) / (1.0 - eps);
    }
  }
  *attr = a;
  *rep = r;
  return 0;
}

void orc_spmm_f32(int64_t n, int d, const int64_t *rp, const int32_t *col, const float *X,
                  float lr, unsigned quirks, float *out, int threads) {
  spmm_impl<float>(n, d, rp, col, X, lr, quirks, out, threads);
}

void orc_spmm_f64(int64_t n, int d, const int64_t *rp, const int32_t *col, const double *X,
                  double lr, unsigned quirks, double *out, int threads) {
  spmm_impl<double>(n, d, rp, col, X, lr, quirks, out, threads);
}

} // extern "C"
