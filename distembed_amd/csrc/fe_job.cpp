// Device state of one embedding job: uploads, work plan, step descriptors,
// launches and the FP64 loss.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "fe_job.hpp"

namespace fe {

fe_status job_check(Job &j, hipError_t e, const char *what) {
  if (e == hipSuccess) return FE_OK;
  j.err = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return FE_ERR_OOM;
  return FE_ERR_HIP;
}

#define FE_HIP(j, call, what)                                                                  \
  do {                                                                                         \
    fe_status _s = job_check((j), (call), (what));                                             \
    if (_s != FE_OK) return _s;                                                                \
  } while (0)

static hipError_t dev_alloc(Job &j, void **p, size_t bytes) {
  return j.allocator ? j.allocator->alloc(p, bytes) : hipMalloc(p, bytes);
}

static void dev_free(Job &j, void *p) {
  if (!p) return;
  if (j.allocator) j.allocator->release(p);
  else hipFree(p);
}

template <typename T>
static fe_status ensure_cap(Job &j, T *&ptr, size_t &cap, size_t count, const char *what) {
  if (count <= cap && ptr) return FE_OK;
  if (ptr) {
    dev_free(j, ptr);
    j.device_bytes -= (int64_t)(cap * sizeof(T));
    ptr = nullptr;
    cap = 0;
  }
  const size_t want = std::max<size_t>(count, 1);
  void *p = nullptr;
  hipError_t e = dev_alloc(j, &p, want * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    j.err = std::string(what) + ": allocation of " + std::to_string(want * sizeof(T)) +
            " bytes failed: " + hipGetErrorString(e);
    return FE_ERR_OOM;
  }
  ptr = static_cast<T *>(p);
  cap = want;
  j.device_bytes += (int64_t)(want * sizeof(T));
  return FE_OK;
}

// Grows a buffer to `count` elements keeping its first `keep` elements.
template <typename T>
static fe_status grow_keep(Job &j, T *&ptr, size_t &cap, size_t count, size_t keep,
                           const char *what, hipStream_t st) {
  if (count <= cap && ptr) return FE_OK;
  T *old = ptr;
  const size_t old_cap = cap;
  ptr = nullptr;
  cap = 0;
  fe_status s = ensure_cap(j, ptr, cap, count, what);
  if (s != FE_OK) {
    ptr = old;
    cap = old_cap;
    return s;
  }
  if (old && keep > 0)
    FE_HIP(j, hipMemcpyAsync(ptr, old, sizeof(T) * std::min(keep, old_cap), hipMemcpyDeviceToDevice, st),
           what);
  if (old) {
    FE_HIP(j, hipStreamSynchronize(st), what);
    dev_free(j, old);
    j.device_bytes -= (int64_t)(old_cap * sizeof(T));
  }
  return FE_OK;
}

// Slot layout for batch size B: P = batches + 1 slots and the constants of
// k = floor(c / B) = (c * m) >> (31 + l), l = ceil(log2 B), m = ceil(2^(31+l)/B),
// exact for every 0 <= c < 2^31 (Granlund and Montgomery, PLDI 1994, Thm 4.2).
static Layout make_layout(int B, int batches) {
  Layout L;
  int l = 0;
  while ((1LL << l) < B) ++l;
  const uint64_t num = 1ULL << (31 + l);
  L.div_m = (uint32_t)((num + (uint64_t)B - 1) / (uint64_t)B);
  L.div_sh = 31 + l;
  L.B = B;
  L.P = batches + 1;
  L.b_shift = (B & (B - 1)) == 0 ? l : -1;
  L.pad = 0;
  return L;
}

// Position of row c in the layout at rest with rotation rot.
static int64_t rest_pos(const Job &j, const Layout &L, int rot, int64_t c) {
  const int64_t k = c / L.B;
  int64_t sl = (k - rot) % L.P;
  if (sl < 0) sl += L.P;
  return sl * L.B + (c - k * L.B);
}

// Brings X (and the shadow) into the layout at rest with rotation 0 (row c
// at position c).
fe_status job_normalize_layout(Job &j, hipStream_t st) {
  if (j.rot == 0) return FE_OK;
  const Layout old = make_layout(j.lay_B, (int)((j.n + j.lay_B - 1) / j.lay_B));
  float *nx = nullptr;
  size_t ncap = 0;
  fe_status s = ensure_cap(j, nx, ncap, j.cap_X, "X (layout change)");
  if (s != FE_OK) return s;
  const int nbo = old.P - 1;
  for (int k = 0; k < nbo; ++k) {
    const int64_t r0 = (int64_t)k * old.B;
    const int64_t rows = std::min<int64_t>(r0 + old.B, j.n) - r0;
    const int64_t src = rest_pos(j, old, j.rot, r0);
    FE_HIP(j, hipMemcpyAsync(nx + r0 * j.d, j.X + src * j.d, sizeof(float) * rows * j.d,
                             hipMemcpyDeviceToDevice, st),
           "layout change copy");
  }
  FE_HIP(j, hipStreamSynchronize(st), "layout change sync");
  dev_free(j, j.X);
  j.device_bytes -= (int64_t)(j.cap_X * sizeof(float));
  j.X = nx;
  j.cap_X = ncap;
  if (j.dtype == FE_GATHER_BF16)
    FE_HIP(j, launch_to_bf16(j.X, j.Xs, (int64_t)j.n * j.d, st), "layout change shadow");
  j.rot = 0;
  j.steps_dirty = true;
  return FE_OK;
}

void job_advance_layout(Job &j, int epochs) {
  if (j.lay.P <= 0 || epochs <= 0) return;
  j.rot = (int)(((int64_t)j.rot + epochs) % j.lay.P);
  j.lay_B = j.lay.B;
}

int force_blocks_for(const StepDesc &sd, int items_per_block) {
  const int pack = 64 / sd.G;
  const int units = (sd.n_items - sd.n_block_items - sd.n_light) + (sd.n_light + pack - 1) / pack;
  return std::max(1, sd.n_block_items + (units + items_per_block - 1) / items_per_block);
}

size_t force_lds_bytes(int ns, int d) { return (size_t)(ns + kSub) * d * sizeof(float); }

int rows_of_batch(const Job &j, int b) {
  const int64_t r0 = (int64_t)b * j.plan.B;
  return (int)(std::min<int64_t>(r0 + j.plan.B, j.n) - r0);
}

hipError_t record_event(hipEvent_t ev, hipStream_t st) {
  // during stream capture the record becomes an external event node, so
  // every graph replay stamps the event again
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(st, &cs);
  if (e != hipSuccess) return e;
  if (cs == hipStreamCaptureStatusActive)
    return hipEventRecordWithFlags(ev, st, hipEventRecordExternal);
  return hipEventRecord(ev, st);
}

fe_status job_init(Job &j, int64_t n, int d, int dtype, uint32_t quirks, int slot_B) {
  if (n <= 0 || n >= 2147483647LL) {
    j.err = "n must be in [1, 2^31-1)";
    return FE_ERR_INVALID_ARG;
  }
  if (d < 1 || d > 512) {
    j.err = "d must be in [1, 512]";
    return FE_ERR_INVALID_ARG;
  }
  if (dtype != FE_GATHER_FP32 && dtype != FE_GATHER_BF16) {
    j.err = "gather dtype must be FE_GATHER_FP32 or FE_GATHER_BF16";
    return FE_ERR_INVALID_ARG;
  }
  if (quirks & ~FE_QUIRKS_REFERENCE) {
    j.err = "unknown quirk bits";
    return FE_ERR_INVALID_ARG;
  }
  j.n = n;
  j.dx = d;
  d = padded_width(d);
  j.d = d;
  j.G = d / 8;
  j.dtype = dtype;
  j.quirks = quirks;
  j.c_limit = (quirks & FE_QUIRK_SKIP_LAST_COL) ? n - 1 : n;
  // one allocation at the final size: replacing a smaller block would leave
  // a hole in an engine's first-fit arena, and a job sized to the budget
  // would then never fit
  const int64_t rows = slot_B > 0 ? ((n + slot_B - 1) / slot_B + 1) * (int64_t)slot_B : n;
  fe_status s = ensure_cap(j, j.X, j.cap_X, (size_t)rows * d, "X");
  if (s != FE_OK) return s;
  if (dtype == FE_GATHER_BF16) {
    s = ensure_cap(j, j.Xs, j.cap_Xs, (size_t)rows * d, "BF16 shadow");
    if (s != FE_OK) return s;
  }
  return ensure_cap(j, j.epoch, j.cap_epoch, 4, "epoch state");
}

void job_free(Job &j) {
  void *bufs[] = {j.X, j.Xs, j.col, j.items, j.S, j.P, j.counters, j.negs, j.epoch, j.desc,
                  j.prefix, j.negcur, j.edesc, j.Y, j.sitems, j.sdesc, j.sprefix};
  for (void *p : bufs) dev_free(j, p);
  j.X = nullptr;
  j.Xs = nullptr;
  j.col = j.items = j.negs = j.negcur = nullptr;
  j.edesc = nullptr;
  j.S = j.P = j.Y = nullptr;
  j.cap_Y = 0;
  j.sitems = nullptr;
  j.sdesc = nullptr;
  j.sprefix = nullptr;
  j.cap_sitems = j.cap_sdesc = j.cap_sprefix = 0;
  j.s_gen = ~0ull;
  j.counters = j.epoch = j.prefix = nullptr;
  j.desc = nullptr;
  j.cap_items = j.cap_S = j.cap_P = j.cap_ctr = j.cap_negs = j.cap_desc = 0;
  j.cap_prefix = j.cap_negcur = j.cap_edesc = j.cap_col = j.cap_X = j.cap_Xs = j.cap_epoch = 0;
  j.device_bytes = 0;
}

namespace {
hipError_t h2d_staged(void *dst, const void *src, size_t bytes, hipStream_t st);
hipError_t h2d_rows(void *dst, const void *src, int64_t rows, size_t src_row, size_t dst_row,
                    hipStream_t st);
}

fe_status job_upload_csr(Job &j, const int64_t *row_ptr, const int32_t *col, int64_t nnz,
                         hipStream_t st) {
  if (!row_ptr || (nnz > 0 && !col) || nnz < 0 || nnz >= 2147483647LL) {
    j.err = "fe_upload_csr: null pointer or nnz out of [0, 2^31-1)";
    return FE_ERR_INVALID_ARG;
  }
  if (row_ptr[0] != 0 || row_ptr[j.n] != nnz) {
    j.err = "fe_upload_csr: row_ptr[0] must be 0 and row_ptr[n] must equal nnz";
    return FE_ERR_INVALID_ARG;
  }
  // row_ptr monotone first (then every row range lies in [0, nnz)), then
  // the column ids and each row's count of skipped columns, in parallel
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t r = 0; r < j.n; ++r) bad |= row_ptr[r + 1] < row_ptr[r] ? 1 : 0;
  if (bad) {
    j.err = "fe_upload_csr: row_ptr is not monotone";
    return FE_ERR_INVALID_ARG;
  }
  std::vector<int32_t> skip((size_t)j.n, 0);
  const int64_t n = j.n, c_limit = j.c_limit;
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : bad)
  for (int64_t r = 0; r < n; ++r) {
    int32_t sk = 0;
    for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
      const int32_t c = col[e];
      bad |= (c < 0 || c >= n) ? 1 : 0;
      sk += (c >= c_limit) ? 1 : 0;
    }
    skip[r] = sk;
  }
  if (bad) {
    j.err = "fe_upload_csr: column id out of [0, n)";
    return FE_ERR_INVALID_ARG;
  }
  fe_status s = ensure_cap(j, j.col, j.cap_col, (size_t)nnz, "CSR columns");
  if (s != FE_OK) return s;
  if (nnz > 0) FE_HIP(j, h2d_staged(j.col, col, sizeof(int32_t) * nnz, st), "fe_upload_csr: copy");
  FE_HIP(j, hipStreamSynchronize(st), "fe_upload_csr: sync");
  j.row_ptr.assign(row_ptr, row_ptr + j.n + 1);
  j.skip_deg.swap(skip);
  ++j.csr_gen;
  j.nnz = nnz;
  j.plan_dirty = true;
  j.steps_dirty = true;
  return FE_OK;
}

fe_status job_upload_x(Job &j, const float *x, hipStream_t st) {
  if (!x) {
    j.err = "fe_upload_x: null pointer";
    return FE_ERR_INVALID_ARG;
  }
  const size_t count = (size_t)j.n * j.d;
  // the layout at rest with rotation 0: row c at position c; the caller's
  // rows of dx floats, zero-padded to d
  FE_HIP(j, h2d_rows(j.X, x, j.n, sizeof(float) * (size_t)j.dx, sizeof(float) * (size_t)j.d, st),
         "fe_upload_x: copy");
  if (j.dtype == FE_GATHER_BF16)
    FE_HIP(j, launch_to_bf16(j.X, j.Xs, (int64_t)count, st), "fe_upload_x: shadow refresh");
  FE_HIP(j, hipMemsetAsync(j.epoch, 0, sizeof(int) * 4, st), "fe_upload_x: state");
  FE_HIP(j, hipStreamSynchronize(st), "fe_upload_x: sync");
  j.rot = 0;
  j.x_ready = true;
  return FE_OK;
}

// Copies between host and device through pinned memory: two 16 MiB chunks
// per device and direction (allocated on first use, kept for the process),
// the DMA of one overlapping the parallel host copy of the other. Copies in
// one direction on one device share its chunks and take turns; uploads and
// downloads (an engine's submitting and collecting threads) do not wait for
// each other.
namespace {
struct PinnedStage {
  std::mutex mu;
  char *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};
constexpr size_t kStageChunk = (size_t)16 << 20;
constexpr int kMaxDevices = 64;
PinnedStage g_stage[kMaxDevices];    // device to host
PinnedStage g_stage_up[kMaxDevices]; // host to device

// bytes of device rows from src (device) to dst (host); with src_row > 0,
// rows of src_row bytes of which the first dst_row go to consecutive rows
// of dst_row bytes (the padding of the device rows is dropped)
struct CopySeg {
  void *dst;
  const void *src;
  size_t bytes;
  size_t src_row = 0;
  size_t dst_row = 0;
};

void host_copy(void *dst, const void *src, size_t bytes) {
  const size_t block = (size_t)1 << 20;
  const int64_t nblk = (int64_t)((bytes + block - 1) / block);
#pragma omp parallel for schedule(static) if (nblk > 1)
  for (int64_t b = 0; b < nblk; ++b) {
    const size_t off = (size_t)b * block;
    std::memcpy(static_cast<char *>(dst) + off, static_cast<const char *>(src) + off,
                std::min(block, bytes - off));
  }
}

// `rows` rows of row_bytes each, from src (a row every src_stride bytes) to
// dst (a row every dst_stride bytes); zero_tail > 0 clears that many bytes
// after each dst row (the padding of a device row)
void host_copy_rows(char *dst, size_t dst_stride, const char *src, size_t src_stride,
                    size_t row_bytes, int64_t rows, size_t zero_tail) {
#pragma omp parallel for schedule(static) if (rows > 4096)
  for (int64_t r = 0; r < rows; ++r) {
    std::memcpy(dst + (size_t)r * dst_stride, src + (size_t)r * src_stride, row_bytes);
    if (zero_tail) std::memset(dst + (size_t)r * dst_stride + row_bytes, 0, zero_tail);
  }
}

// The current device's chunks of one direction, allocated on first use
// (the caller holds ps.mu).
hipError_t stage_ready(PinnedStage &ps) {
  hipError_t e = hipSuccess;
  for (int s = 0; s < 2 && e == hipSuccess; ++s) {
    if (!ps.buf[s]) e = hipHostMalloc((void **)&ps.buf[s], kStageChunk, hipHostMallocDefault);
    if (e == hipSuccess && !ps.ev[s]) e = hipEventCreateWithFlags(&ps.ev[s], hipEventDisableTiming);
  }
  return e;
}

hipError_t d2h_staged(const std::vector<CopySeg> &segs, hipStream_t st) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  PinnedStage &ps = g_stage[dev];
  std::lock_guard<std::mutex> lk(ps.mu);
  if ((e = stage_ready(ps)) != hipSuccess) return e;
  CopySeg held[2] = {{nullptr, nullptr, 0}, {nullptr, nullptr, 0}}; // piece in each chunk
  auto drain = [&](int s) -> hipError_t {
    if (!held[s].bytes) return hipSuccess;
    hipError_t r = hipEventSynchronize(ps.ev[s]);
    const CopySeg &h = held[s];
    if (r == hipSuccess) {
      if (h.src_row && h.dst_row != h.src_row)
        host_copy_rows(static_cast<char *>(h.dst), h.dst_row, ps.buf[s], h.src_row, h.dst_row,
                       (int64_t)(h.bytes / h.src_row), 0);
      else
        host_copy(h.dst, ps.buf[s], h.bytes);
    }
    held[s].bytes = 0;
    return r;
  };
  int k = 0;
  for (const CopySeg &sg : segs) {
    // a chunk (16 MiB) holds whole rows of any padded width (8 x 2^k floats)
    const bool strided = sg.src_row && sg.dst_row != sg.src_row;
    for (size_t off = 0; off < sg.bytes && e == hipSuccess; off += kStageChunk) {
      const int s = k & 1;
      e = drain(s);
      const size_t len = std::min(kStageChunk, sg.bytes - off);
      const char *src = static_cast<const char *>(sg.src) + off;
      char *dst = static_cast<char *>(sg.dst) + (strided ? off / sg.src_row * sg.dst_row : off);
      if (e == hipSuccess) e = hipMemcpyAsync(ps.buf[s], src, len, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipEventRecord(ps.ev[s], st);
      if (e == hipSuccess) held[s] = {dst, src, len, sg.src_row, sg.dst_row};
      ++k;
    }
  }
  // the older of the two chunks first
  for (int i = 0; i < 2; ++i) {
    hipError_t r = drain((k + i) & 1);
    if (e == hipSuccess) e = r;
  }
  return e;
}

// dst (device) := src (host), bytes, on stream st; done when it returns, so
// the caller may reuse src. Copies below one chunk go straight from src.
hipError_t h2d_staged(void *dst, const void *src, size_t bytes, hipStream_t st) {
  if (bytes < kStageChunk) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  PinnedStage &ps = g_stage_up[dev];
  std::lock_guard<std::mutex> lk(ps.mu);
  if ((e = stage_ready(ps)) != hipSuccess) return e;
  bool used[2] = {false, false};
  int k = 0;
  for (size_t off = 0; off < bytes && e == hipSuccess; off += kStageChunk, ++k) {
    const int s = k & 1;
    // chunk s is free once its previous DMA has finished
    if (used[s]) e = hipEventSynchronize(ps.ev[s]);
    const size_t len = std::min(kStageChunk, bytes - off);
    if (e == hipSuccess) host_copy(ps.buf[s], static_cast<const char *>(src) + off, len);
    if (e == hipSuccess)
      e = hipMemcpyAsync(static_cast<char *>(dst) + off, ps.buf[s], len, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(ps.ev[s], st);
    if (e == hipSuccess) used[s] = true;
  }
  // the chunks are handed to the next copy only after their DMAs
  for (int s = 0; s < 2; ++s) {
    if (!used[s]) continue;
    hipError_t r = hipEventSynchronize(ps.ev[s]);
    if (e == hipSuccess) e = r;
  }
  return e;
}

// `rows` rows of src_row bytes (host) into rows of dst_row >= src_row bytes
// (device), the rest of each device row zeroed, through the pinned chunks;
// done when it returns. Equal widths take h2d_staged.
hipError_t h2d_rows(void *dst, const void *src, int64_t rows, size_t src_row, size_t dst_row,
                    hipStream_t st) {
  if (src_row == dst_row) return h2d_staged(dst, src, (size_t)rows * src_row, st);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  PinnedStage &ps = g_stage_up[dev];
  std::lock_guard<std::mutex> lk(ps.mu);
  if ((e = stage_ready(ps)) != hipSuccess) return e;
  const int64_t per = (int64_t)(kStageChunk / dst_row); // rows per chunk
  bool used[2] = {false, false};
  int k = 0;
  for (int64_t r0 = 0; r0 < rows && e == hipSuccess; r0 += per, ++k) {
    const int s = k & 1;
    if (used[s]) e = hipEventSynchronize(ps.ev[s]);
    const int64_t m = std::min(per, rows - r0);
    if (e == hipSuccess)
      host_copy_rows(ps.buf[s], dst_row, static_cast<const char *>(src) + (size_t)r0 * src_row,
                     src_row, src_row, m, dst_row - src_row);
    if (e == hipSuccess)
      e = hipMemcpyAsync(static_cast<char *>(dst) + (size_t)r0 * dst_row, ps.buf[s],
                         (size_t)m * dst_row, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(ps.ev[s], st);
    if (e == hipSuccess) used[s] = true;
  }
  for (int s = 0; s < 2; ++s) {
    if (!used[s]) continue;
    hipError_t r = hipEventSynchronize(ps.ev[s]);
    if (e == hipSuccess) e = r;
  }
  return e;
}
} // namespace

fe_status job_download_x(Job &j, float *x, hipStream_t st) {
  if (!x) {
    j.err = "fe_download_x: null pointer";
    return FE_ERR_INVALID_ARG;
  }
  if (!j.x_ready) {
    j.err = "fe_download_x: X was never uploaded";
    return FE_ERR_STATE;
  }
  // device rows of d floats become the caller's rows of dx (the padding
  // is dropped)
  const size_t srow = sizeof(float) * (size_t)j.d, drow = sizeof(float) * (size_t)j.dx;
  std::vector<CopySeg> segs;
  if (j.rot == 0) {
    segs.push_back({x, j.X, srow * (size_t)j.n, srow, drow});
  } else {
    // one piece per batch, from its slot
    const Layout L = make_layout(j.lay_B, (int)((j.n + j.lay_B - 1) / j.lay_B));
    for (int k = 0; k < L.P - 1; ++k) {
      const int64_t r0 = (int64_t)k * L.B;
      const int64_t rows = std::min<int64_t>(r0 + L.B, j.n) - r0;
      segs.push_back({x + r0 * j.dx, j.X + rest_pos(j, L, j.rot, r0) * j.d, srow * (size_t)rows,
                      srow, drow});
    }
  }
  FE_HIP(j, d2h_staged(segs, st), "fe_download_x: copy");
  return FE_OK;
}

fe_status job_set_negatives(Job &j, const int32_t *lists, int n_lists, int ns, hipStream_t st) {
  if (n_lists < 0 || ns < 0 || ns > kMaxNs || (n_lists > 0 && ns > 0 && !lists)) {
    j.err = "fe_set_negatives: invalid argument (ns must be in [0, 64])";
    return FE_ERR_INVALID_ARG;
  }
  if ((int64_t)ns * j.d > kMaxLdsFloats) {
    j.err = "fe_set_negatives: ns * d must be <= 8192";
    return FE_ERR_INVALID_ARG;
  }
  const size_t count = (size_t)n_lists * ns;
  for (size_t i = 0; i < count; ++i) {
    if (lists[i] < 0 || lists[i] >= j.n) {
      j.err = "fe_set_negatives: negative id out of [0, n)";
      return FE_ERR_INVALID_ARG;
    }
  }
  fe_status s = ensure_cap(j, j.negs, j.cap_negs, count, "negative lists");
  if (s != FE_OK) return s;
  if (count > 0)
    FE_HIP(j, hipMemcpyAsync(j.negs, lists, sizeof(int32_t) * count, hipMemcpyHostToDevice, st),
           "fe_set_negatives: copy");
  FE_HIP(j, hipStreamSynchronize(st), "fe_set_negatives: sync");
  j.ns = ns;
  j.n_lists = n_lists;
  j.steps_dirty = true;
  return FE_OK;
}

fe_status job_configure(Job &j, int B, float lr, float bound) {
  if (B <= 0 || !std::isfinite(lr) || !std::isfinite(bound) || bound < 0) {
    j.err = "fe_configure: batch size must be > 0, lr and max_bound finite, max_bound >= 0";
    return FE_ERR_INVALID_ARG;
  }
  if (B != j.B) {
    // per-batch shapes and cuttings name batches of the old size
    j.plan_dirty = true;
    j.bshape.clear();
    j.bcut.clear();
  }
  if (B != j.B || lr != j.lr || bound != j.bound) j.steps_dirty = true;
  j.B = B;
  j.lr = lr;
  j.bound = bound;
  return FE_OK;
}

fe_status job_set_tuning(Job &j, const Tuning &t) {
  auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
  if (t.chunk < 1 || t.chunk > 1024 || t.max_chunks < 1 || !pow2(t.unroll) || t.unroll > 8 ||
      !pow2(t.waves_per_block) || t.waves_per_block > 8 || t.items_per_wave < 1 ||
      t.items_per_wave > 64 || t.light < 0 || t.light > kLightCols) {
    j.err = "tuning out of range";
    return FE_ERR_INVALID_ARG;
  }
  if ((t.slice != 8 && t.slice != 16) || (j.d > 0 && t.slice > j.d)) {
    j.err = "tuning: slice must be 8 or 16 components per lane, and at most d";
    return FE_ERR_INVALID_ARG;
  }
  if (t.chunk != j.tun.chunk || t.max_chunks != j.tun.max_chunks || t.light != j.tun.light ||
      t.slice != j.tun.slice)
    j.plan_dirty = true;
  j.tun = t;
  j.bshape.clear();
  if (!j.bcut.empty()) j.plan_dirty = true;
  j.bcut.clear();
  j.steps_dirty = true;
  return FE_OK;
}

// Light rows per wave of a batch whose lanes hold `slice` components.
static int pack_of(const Job &j, int slice) { return 64 / (j.d / slice); }

// Per-batch values hold only for the plan they were set on (one per batch).
Cut job_cut(const Job &j, int b) {
  if (b >= 0 && b < (int)j.bcut.size() && (int)j.bcut.size() == j.plan.batches) return j.bcut[b];
  return Cut{j.tun.chunk, j.tun.max_chunks, j.tun.light, j.tun.slice, pack_of(j, j.tun.slice)};
}

fe_status job_set_cuts(Job &j, const std::vector<Cut> &cuts) {
  for (const Cut &c : cuts) {
    if (c.chunk < 1 || c.chunk > 1024 || c.max_chunks < 1 || c.light < 0 || c.light > kLightCols ||
        (c.slice != 8 && c.slice != 16) || c.slice > j.d) {
      j.err = "batch cutting out of range";
      return FE_ERR_INVALID_ARG;
    }
  }
  j.bcut = cuts;
  for (Cut &c : j.bcut) c.pack = pack_of(j, c.slice);
  j.plan_dirty = true;
  j.steps_dirty = true;
  return FE_OK;
}

Shape job_shape(const Job &j, int b) {
  if (b >= 0 && b < (int)j.bshape.size() && (int)j.bshape.size() == j.plan.batches)
    return j.bshape[b];
  return Shape{j.tun.unroll, j.tun.waves_per_block, j.tun.items_per_wave};
}

fe_status job_set_shapes(Job &j, const std::vector<Shape> &shapes) {
  auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
  for (const Shape &sh : shapes) {
    if (!pow2(sh.unroll) || sh.unroll > 8 || !pow2(sh.waves_per_block) ||
        sh.waves_per_block > 8 || sh.items_per_wave < 1 || sh.items_per_wave > 64) {
      j.err = "batch launch shape out of range";
      return FE_ERR_INVALID_ARG;
    }
  }
  j.bshape = shapes;
  j.steps_dirty = true;
  return FE_OK;
}


static StepDesc make_desc(const Job &j, int b, DescKind kind) {
  StepDesc sd;
  std::memset(&sd, 0, sizeof sd);
  sd.X = j.X;
  sd.Xw = j.X;
  sd.Xs = j.Xs;
  sd.Xsw = j.Xs;
  sd.col = j.col;
  sd.items = j.items + (size_t)kItemInts * j.plan.item_off[b];
  sd.lcols = j.items + j.plan.items.size() + (size_t)kLightCols * j.plan.light_off[b];
  sd.S = j.S;
  sd.P = j.P;
  sd.counters = j.counters;
  sd.negs = j.negs;
  sd.neg_ids = j.negcur ? j.negcur + (size_t)b * j.ns : nullptr;
  sd.state = j.epoch;
  sd.lay = j.lay;
  sd.n_items = (int)(j.plan.item_off[b + 1] - j.plan.item_off[b]);
  sd.n_block_items = j.plan.block_items[b];
  sd.n_light = j.plan.light_rows[b];
  sd.row_begin = (int)((int64_t)b * j.plan.B);
  sd.rows = rows_of_batch(j, b);
  rep_rows(j.plan, j.quirks, b, &sd.rep0, &sd.rep1);
  sd.batch = b;
  sd.ns = j.ns;
  sd.n_lists = j.n_lists;
  sd.D = j.d;
  sd.G = j.d / job_cut(j, b).slice;
  sd.bf16 = j.dtype == FE_GATHER_BF16 ? 1 : 0;
  sd.c_limit = (int)j.c_limit;
  sd.lr = j.lr;
  sd.bound = j.bound;
  sd.delta_mode = kind == DESC_DELTA ? 1 : 0;
  sd.fused = kind == DESC_FUSED ? 1 : 0;
  return sd;
}

const StepDesc *job_desc(const Job &j, int b, DescKind kind) {
  const int nb = j.plan.batches;
  switch (kind) {
  case DESC_FUSED: return j.desc + b;
  case DESC_SINGLE: return j.desc + nb + b;
  default: return j.desc + 2 * nb + b;
  }
}

fe_status job_reset_counters(Job &j, hipStream_t st) {
  if (!j.counters) return FE_OK;
  FE_HIP(j, hipMemsetAsync(j.counters, 0, sizeof(int) * j.cap_ctr, st), "counter reset");
  return FE_OK;
}

fe_status job_prepare(Job &j, hipStream_t st, bool slots) {
  if (j.nnz < 0) {
    j.err = "CSR not uploaded";
    return FE_ERR_STATE;
  }
  if (!j.x_ready) {
    j.err = "X not uploaded";
    return FE_ERR_STATE;
  }
  if (j.B <= 0) {
    j.err = "batch size not configured";
    return FE_ERR_STATE;
  }
  if (j.plan_dirty) {
    Plan p;
    const int nbp = (int)((j.n + j.B - 1) / j.B);
    const bool per_batch = (int)j.bcut.size() == nbp;
    fe_status s = FE_OK;
    if (j.pending_B == j.B && !per_batch) {
      p = std::move(j.pending_plan);
    } else {
      s = build_plan(j.n, j.row_ptr.data(), nullptr, j.B, j.tun.chunk, j.tun.max_chunks,
                     j.tun.light, pack_of(j, j.tun.slice), j.c_limit, &p,
                     per_batch ? j.bcut.data() : nullptr);
      if (s != FE_OK) {
        j.err = host_error();
        return s;
      }
    }
    j.pending_plan = Plan();
    j.pending_B = 0;
    // processed entries per batch from the per-row skip counts
    p.edges_eff_total = 0;
    for (int b = 0; b < p.batches; ++b) {
      const int64_t r0 = (int64_t)b * p.B, r1 = std::min<int64_t>(r0 + p.B, j.n);
      int64_t eff = j.row_ptr[r1] - j.row_ptr[r0];
      for (int64_t r = r0; r < r1; ++r) eff -= j.skip_deg[r];
      p.edges_eff[b] = eff;
      p.edges_eff_total += eff;
    }
    j.plan = std::move(p);
    // a rotation left by runs with another batch size is undone first
    if (j.rot != 0 && j.lay_B != j.plan.B && (s = job_normalize_layout(j, st)) != FE_OK) return s;
    j.lay = make_layout(j.plan.B, j.plan.batches);
    if (j.rot == 0) j.lay_B = j.plan.B;
    // the items, then the light rows' columns
    const size_t n_lcols = (size_t)kLightCols * j.plan.light_off[j.plan.batches];
    if ((s = ensure_cap(j, j.items, j.cap_items, j.plan.items.size() + n_lcols, "work items")) !=
        FE_OK)
      return s;
    if ((s = ensure_cap(j, j.S, j.cap_S, (size_t)j.B * j.d, "step rows")) != FE_OK) return s;
    if ((s = ensure_cap(j, j.P, j.cap_P, (size_t)std::max(j.plan.max_pslots, 1) * j.d,
                        "partials")) != FE_OK)
      return s;
    if ((s = ensure_cap(j, j.counters, j.cap_ctr, (size_t)std::max(j.plan.max_splits, 1),
                        "split counters")) != FE_OK)
      return s;
    FE_HIP(j, h2d_staged(j.items, j.plan.items.data(), sizeof(int32_t) * j.plan.items.size(), st),
           "work items copy");
    if (n_lcols > 0)
      FE_HIP(j, launch_fill_lcols(j.items, (int)(j.plan.items.size() / kItemInts), j.col,
                                  j.items + j.plan.items.size(), st),
             "light columns");
    if ((s = job_reset_counters(j, st)) != FE_OK) return s;
    FE_HIP(j, hipStreamSynchronize(st), "plan upload sync");
    j.plan_dirty = false;
    j.steps_dirty = true;
  }
  if (slots) {
    // every slot of the layout: P * B rows (the rows at rest, < P * B, kept)
    const size_t need = (size_t)j.lay.P * j.lay.B * j.d;
    const size_t keep = j.rot == 0 ? (size_t)j.n * j.d : j.cap_X;
    float *x0 = j.X;
    uint16_t *xs0 = j.Xs;
    fe_status s = grow_keep(j, j.X, j.cap_X, need, keep, "X slots", st);
    if (s != FE_OK) return s;
    if (j.dtype == FE_GATHER_BF16 &&
        (s = grow_keep(j, j.Xs, j.cap_Xs, need, j.rot == 0 ? (size_t)j.n * j.d : j.cap_Xs,
                       "shadow slots", st)) != FE_OK)
      return s;
    if (j.X != x0 || j.Xs != xs0) j.steps_dirty = true;
  }
  if (j.steps_dirty) {
    const int nb = j.plan.batches;
    fe_status s;
    if ((s = ensure_cap(j, j.desc, j.cap_desc, (size_t)3 * nb, "step descriptors")) != FE_OK)
      return s;
    if ((s = ensure_cap(j, j.prefix, j.cap_prefix, (size_t)nb * 4, "block prefixes")) != FE_OK)
      return s;
    if ((s = ensure_cap(j, j.negcur, j.cap_negcur, (size_t)nb * std::max(j.ns, 1),
                        "current negative ids")) != FE_OK)
      return s;
    if ((s = ensure_cap(j, j.edesc, j.cap_edesc, 1, "epoch state")) != FE_OK) return s;
    EpochDesc ed{};
    ed.state = j.epoch;
    ed.P = j.lay.P;
    ed.negs = j.negs;
    ed.negcur = j.negcur;
    ed.batches = nb;
    ed.ns = j.ns;
    ed.n_lists = j.n_lists;
    FE_HIP(j, hipMemcpyAsync(j.edesc, &ed, sizeof ed, hipMemcpyHostToDevice, st),
           "epoch state copy");
    std::vector<StepDesc> hd((size_t)3 * nb);
    std::vector<int> pf((size_t)nb * 4, 0);
    j.force_blocks.assign(nb, 0);
    j.apply_blocks.assign(nb, 0);
    for (int b = 0; b < nb; ++b) {
      const Shape sh = job_shape(j, b);
      const int ipb = sh.waves_per_block * sh.items_per_wave;
      hd[b] = make_desc(j, b, DESC_FUSED);
      hd[nb + b] = make_desc(j, b, DESC_SINGLE);
      hd[2 * nb + b] = make_desc(j, b, DESC_DELTA);
      j.force_blocks[b] = force_blocks_for(hd[b], ipb);
      const int64_t q4 = (int64_t)hd[b].rows * j.d / 4;
      j.apply_blocks[b] = (int)std::max<int64_t>(1, std::min<int64_t>((q4 + 255) / 256, 4096));
      pf[4 * b + 1] = j.force_blocks[b];
      pf[4 * b + 3] = j.apply_blocks[b];
    }
    FE_HIP(j, hipMemcpyAsync(j.desc, hd.data(), sizeof(StepDesc) * hd.size(),
                             hipMemcpyHostToDevice, st),
           "descriptor copy");
    FE_HIP(j, hipMemcpyAsync(j.prefix, pf.data(), sizeof(int) * pf.size(), hipMemcpyHostToDevice,
                             st),
           "prefix copy");
    FE_HIP(j, hipStreamSynchronize(st), "descriptor sync");
    j.host_desc.swap(hd);
    ++j.desc_gen;
    j.steps_dirty = false;
  }
  return FE_OK;
}

fe_status job_set_epoch(Job &j, int value, hipStream_t st) {
  if (!j.edesc) {
    j.err = "epoch state not prepared";
    return FE_ERR_STATE;
  }
  FE_HIP(j, launch_epoch(j.edesc, 1, 1, value, j.rot, st), "epoch set");
  return FE_OK;
}

fe_status job_tick(Job &j, hipStream_t st) {
  if (!j.edesc) {
    j.err = "epoch state not prepared";
    return FE_ERR_STATE;
  }
  FE_HIP(j, launch_epoch(j.edesc, 1, 0, 0, 0, st), "epoch tick");
  return FE_OK;
}

fe_status job_launch_force(Job &j, int b, DescKind kind, hipStream_t st, hipEvent_t *ev) {
  const size_t lds = force_lds_bytes(j.ns, j.d);
  if (ev) FE_HIP(j, record_event(ev[0], st), "event");
  const Shape sh = job_shape(j, b);
  // the descriptor by value: a graph captured here must be dropped when the
  // descriptors are rebuilt (desc_gen)
  const StepDesc &sd = j.host_desc[(size_t)j.plan.batches * kind + b];
  FE_HIP(j, launch_force1(sd, j.force_blocks[b], sh.waves_per_block, sh.unroll, lds,
                          j.dtype == FE_GATHER_BF16 ? 1 : 0, sd.G, j.d / sd.G, st),
         "force kernel launch");
  if (ev) FE_HIP(j, record_event(ev[1], st), "event");
  return FE_OK;
}

fe_status job_spmm(Job &j, int iterations, float lr, float *y_out, hipStream_t st) {
  if (iterations < 0 || !y_out) {
    j.err = "fe_spmm: negative iterations or NULL output";
    return FE_ERR_INVALID_ARG;
  }
  if (lr != j.lr) {
    j.lr = lr;
    j.steps_dirty = true;
  }
  fe_status s = job_prepare(j, st);
  if (s != FE_OK) return s;
  if ((s = ensure_cap(j, j.Y, j.cap_Y, (size_t)j.n * j.d, "SpMM output")) != FE_OK) return s;
  // the whole-graph plan, as one batch of n rows
  const int light = std::min(j.tun.light, kLightCols);
  if (j.s_gen != j.csr_gen || j.s_chunk != j.tun.chunk || j.s_max_chunks != j.tun.max_chunks ||
      j.s_light_cut != light || !j.sitems) {
    Plan sp;
    const int B = (int)std::min<int64_t>(j.n, 2147483647LL);
    if ((s = build_plan(j.n, j.row_ptr.data(), nullptr, B, j.tun.chunk, j.tun.max_chunks, light,
                        64 / j.G, j.c_limit, &sp)) != FE_OK) {
      j.err = host_error();
      return s;
    }
    if (sp.batches != 1) {
      j.err = "fe_spmm: more than 2^31 - 1 rows";
      return FE_ERR_INVALID_ARG;
    }
    const size_t n_lcols = (size_t)kLightCols * sp.light_off[1];
    if ((s = ensure_cap(j, j.sitems, j.cap_sitems, sp.items.size() + n_lcols, "SpMM items")) !=
        FE_OK)
      return s;
    // partials and split counters are shared with the force step (same stream)
    float *p0 = j.P;
    int *c0 = j.counters;
    if ((s = ensure_cap(j, j.P, j.cap_P, (size_t)std::max(sp.max_pslots, 1) * j.d, "partials")) !=
        FE_OK)
      return s;
    if ((s = ensure_cap(j, j.counters, j.cap_ctr, (size_t)std::max(sp.max_splits, 1),
                        "split counters")) != FE_OK)
      return s;
    if (j.P != p0 || j.counters != c0) j.steps_dirty = true;
    FE_HIP(j, hipMemcpyAsync(j.sitems, sp.items.data(), sizeof(int32_t) * sp.items.size(),
                             hipMemcpyHostToDevice, st),
           "SpMM items copy");
    if (n_lcols > 0)
      FE_HIP(j, launch_fill_lcols(j.sitems, (int)(sp.items.size() / kItemInts), j.col,
                                  j.sitems + sp.items.size(), st),
             "SpMM light columns");
    if ((s = job_reset_counters(j, st)) != FE_OK) return s;
    FE_HIP(j, hipStreamSynchronize(st), "SpMM plan sync");
    j.s_items = (int)sp.item_off[1];
    j.s_block_items = sp.block_items[0];
    j.s_light = sp.light_rows[0];
    j.s_item_ints = sp.items.size();
    j.s_gen = j.csr_gen;
    j.s_chunk = j.tun.chunk;
    j.s_max_chunks = j.tun.max_chunks;
    j.s_light_cut = light;
    // the force step's descriptors may hold the old partials or counters
    if ((s = job_prepare(j, st)) != FE_OK) return s;
  }
  // one descriptor: the force step's batch-0 descriptor over the whole plan
  StepDesc sd = j.host_desc[(size_t)j.plan.batches]; // the unfused descriptor of batch 0
  sd.items = j.sitems;
  sd.lcols = j.sitems + j.s_item_ints;
  sd.n_items = j.s_items;
  sd.n_block_items = j.s_block_items;
  sd.n_light = j.s_light;
  sd.row_begin = 0;
  sd.rows = (int)j.n;
  sd.batch = 0;
  sd.G = j.G; // the SpMM kernel takes slices of 8 components
  const int wpb = j.tun.waves_per_block;
  const int blocks = force_blocks_for(sd, wpb * j.tun.items_per_wave);
  const int pf[2] = {0, blocks};
  if ((s = ensure_cap(j, j.sdesc, j.cap_sdesc, 1, "SpMM descriptor")) != FE_OK) return s;
  if ((s = ensure_cap(j, j.sprefix, j.cap_sprefix, 2, "SpMM prefix")) != FE_OK) return s;
  FE_HIP(j, hipMemcpyAsync(j.sdesc, &sd, sizeof sd, hipMemcpyHostToDevice, st), "SpMM desc");
  FE_HIP(j, hipMemcpyAsync(j.sprefix, pf, sizeof pf, hipMemcpyHostToDevice, st), "SpMM prefix");
  // state[1] = rot: the kernel reads X through the layout at rest
  if ((s = job_set_epoch(j, 0, st)) != FE_OK) return s;
  for (int it = 0; it < iterations; ++it)
    FE_HIP(j, launch_spmm(j.sdesc, j.sprefix, 1, blocks, wpb, j.tun.unroll,
                          j.dtype == FE_GATHER_BF16 ? 1 : 0, j.G, j.Y, st),
           "SpMM kernel launch");
  if (iterations > 0)
    FE_HIP(j,
           d2h_staged({{y_out, j.Y, sizeof(float) * (size_t)j.n * j.d, sizeof(float) * (size_t)j.d,
                        sizeof(float) * (size_t)j.dx}},
                      st),
           "fe_spmm: copy");
  return FE_OK;
}

fe_status job_launch_apply(Job &j, int b, DescKind kind, hipStream_t st) {
  FE_HIP(j, launch_apply(job_desc(j, b, kind), j.prefix + 4 * b + 2, 1, j.apply_blocks[b], st),
         "apply kernel launch");
  return FE_OK;
}

fe_status job_loss(Job &j, uint64_t seed, int eval_ns, double *attr, double *rep,
                   hipStream_t st) {
  if (eval_ns < 0 || eval_ns >= 1048576 || !attr || !rep) {
    j.err = "fe_loss: invalid argument";
    return FE_ERR_INVALID_ARG;
  }
  LossDesc ld;
  std::memset(&ld, 0, sizeof ld);
  ld.X = j.X;
  ld.col = j.col;
  ld.items = j.items;
  ld.lay = j.rot == 0 ? j.lay : make_layout(j.lay_B, (int)((j.n + j.lay_B - 1) / j.lay_B));
  ld.rot = j.rot;
  ld.n_items = (int)j.plan.item_off[j.plan.batches];
  ld.D = j.d;
  ld.G = j.G;
  ld.c_limit = (int)j.c_limit;
  ld.n = j.n;
  ld.eval_seed = seed;
  ld.eval_ns = eval_ns;
  const int ipw = 4, wpb = 4;
  const int eblocks = std::max(1, (ld.n_items + ipw * wpb - 1) / (ipw * wpb));
  const int NG = 64 / j.G;
  const int ppg = 8;
  const int64_t pairs = j.n * (int64_t)eval_ns;
  const int64_t per_block = (int64_t)wpb * NG * ppg;
  const int nblocks = (int)std::max<int64_t>(1, (pairs + per_block - 1) / per_block);
  double *buf = nullptr;
  FE_HIP(j, hipMalloc(&buf, sizeof(double) * ((size_t)eblocks + nblocks + 2)), "loss buffer");
  fe_status s = FE_OK;
  double out[2] = {0.0, 0.0};
  do {
    if ((s = job_check(j, launch_loss_edges(ld, buf + 2, eblocks, ipw, st), "loss edges")) != FE_OK)
      break;
    if ((s = job_check(j, launch_loss_negs(ld, buf + 2 + eblocks, nblocks, ppg, st),
                       "loss negatives")) != FE_OK)
      break;
    if ((s = job_check(j, launch_reduce(buf + 2, eblocks, buf, st), "reduce")) != FE_OK) break;
    if ((s = job_check(j, launch_reduce(buf + 2 + eblocks, nblocks, buf + 1, st), "reduce")) !=
        FE_OK)
      break;
    if ((s = job_check(j, hipMemcpyAsync(out, buf, sizeof out, hipMemcpyDeviceToHost, st),
                       "loss copy")) != FE_OK)
      break;
    s = job_check(j, hipStreamSynchronize(st), "loss sync");
  } while (0);
  hipFree(buf);
  if (s != FE_OK) return s;
  *attr = out[0];
  *rep = eval_ns > 0 ? out[1] : 0.0;
  if (!std::isfinite(out[0]) || !std::isfinite(*rep)) {
    j.err = "fe_loss: the loss is not finite (NaN or Inf in X)";
    return FE_ERR_NOT_FINITE;
  }
  return FE_OK;
}

int64_t job_bytes_estimate(const Job &j) {
  // X (and the shadow) in the slot layout: batches + 1 slots of B rows
  const int64_t rows = j.plan.B > 0 ? (int64_t)(j.plan.batches + 1) * j.plan.B : j.n;
  const int64_t nd = rows * (int64_t)j.d;
  int64_t b = nd * 4 + (j.dtype == FE_GATHER_BF16 ? nd * 2 : 0);
  b += std::max<int64_t>(j.nnz, 1) * 4;
  b += ((int64_t)j.plan.items.size() + (int64_t)kLightCols * j.plan.light_off.back()) * 4;
  b += (int64_t)j.plan.B * j.d * 4;
  b += (int64_t)std::max(j.plan.max_pslots, 1) * j.d * 4;
  b += (int64_t)std::max(j.plan.max_splits, 1) * 4;
  b += (int64_t)j.n_lists * j.ns * 4;
  b += (int64_t)(3 * j.plan.batches) * (int64_t)sizeof(StepDesc) + 16LL * j.plan.batches + 16;
  b += (int64_t)j.plan.batches * std::max(j.ns, 1) * 4 + (int64_t)sizeof(EpochDesc);
  return b;
}

} // namespace fe
