// Multi-job engine: the batching and serving forms of the force step.
//
// Jobs are independent graphs. Their device state lives in one arena carved
// from a single hipMalloc of the HBM budget (first fit, coalescing free
// list), so admission is a pure accounting question and nothing else in the
// process competes for that memory. Every engine round advances each
// running job by one minibatch: the round's step descriptors (one per job)
// go to the device in one copy and all jobs' work items share one fused
// force launch. A job's work items, chunking and summation order are fixed
// when it is submitted, and no job reads or writes another job's memory, so
// its result does not depend on the mix it ran in.
//
// Threads: fe_engine_submit and fe_engine_collect do their host work and
// their copies (on their own streams) without holding the engine's lock, so
// a submitting or collecting thread does not stall the thread that runs
// rounds. Every other call holds the lock for its whole duration.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fe_job.hpp"

using namespace fe;

namespace {

constexpr size_t kAlign = 256;

// First-fit arena over one device allocation (thread-safe).
class Arena : public DeviceAllocator {
public:
  fe_status init(size_t bytes, std::string &err) {
    bytes = (bytes + kAlign - 1) / kAlign * kAlign;
    hipError_t e = hipMalloc(&base_, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      err = std::string("engine arena: hipMalloc of ") + std::to_string(bytes) +
            " bytes failed: " + hipGetErrorString(e);
      return FE_ERR_OOM;
    }
    size_ = bytes;
    free_[0] = bytes;
    return FE_OK;
  }
  ~Arena() override {
    if (base_) hipFree(base_);
  }
  hipError_t alloc(void **p, size_t bytes) override {
    std::lock_guard<std::mutex> lk(mu_);
    bytes = std::max<size_t>((bytes + kAlign - 1) / kAlign * kAlign, kAlign);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_[off] = bytes;
        in_use_ += bytes;
        *p = static_cast<char *>(base_) + off;
        return hipSuccess;
      }
    }
    return hipErrorOutOfMemory;
  }
  void release(void *p) override {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    const size_t off = static_cast<char *>(p) - static_cast<char *>(base_);
    auto u = used_.find(off);
    if (u == used_.end()) return;
    size_t len = u->second;
    used_.erase(u);
    in_use_ -= len;
    size_t start = off;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && next->first == off + len) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    free_[start] = len;
  }
  size_t in_use() const {
    std::lock_guard<std::mutex> lk(mu_);
    return in_use_;
  }
  size_t size() const { return size_; }
  size_t largest_free() const {
    std::lock_guard<std::mutex> lk(mu_);
    size_t m = 0;
    for (auto &f : free_) m = std::max(m, f.second);
    return m;
  }

private:
  mutable std::mutex mu_;
  void *base_ = nullptr;
  size_t size_ = 0;
  size_t in_use_ = 0;
  std::map<size_t, size_t> free_; // offset -> length
  std::map<size_t, size_t> used_;
};

struct EngineJob {
  Job job;
  int id = 0;
  int iterations = 0;
  int epoch = 0;        // next step (epoch, batch)
  int batch = 0;
  bool launched_all = false;
  bool collecting = false; // a collect is downloading the result
  bool collected = false;
  bool failed = false;     // a round of this job failed: it has no result
  hipEvent_t uploaded = nullptr; // uploads finished (on the upload stream)
  hipEvent_t done = nullptr;     // the job's last round finished
  bool waited_upload = false;
  int state_epoch = -1;    // epoch the job's device state ints hold (-1: unset)
  int64_t units = 0;
};

} // namespace

// Pinned round buffers in flight: the host may build this many rounds ahead
// of the device.
constexpr int kStages = 8;

struct fe_engine {
  std::mutex mu;                   // everything below except the streams
  int device = 0;
  hipStream_t stream = nullptr;    // rounds
  hipStream_t upload = nullptr;    // job uploads
  hipStream_t download = nullptr;  // result downloads
  std::unique_ptr<Arena> arena;
  std::vector<std::unique_ptr<EngineJob>> jobs;
  Tuning tun;
  // round staging: pinned host ring of kStages; one device buffer, reused in
  // stream order
  struct Stage {
    char *host = nullptr;
    hipEvent_t copied = nullptr;
  } stage[kStages];
  int stage_next = 0;
  size_t stage_cap = 0; // bytes
  char *dev_round = nullptr;
  std::vector<char> one_host; // a round built on the host only (one job, or
                              // small enough for kernel arguments)
  int64_t rounds_run = 0;
  int64_t units_done = 0;
  double graph_prepare_ms = 0; // last fe_engine_run_graph: schedule, capture, instantiate
  double graph_exec_ms = 0;    // ... and the replay itself (HIP events)
  int64_t graph_rounds = 0;
};

// The calling thread's last engine failure (threads may share an engine).
static thread_local std::string g_engine_error;

static fe_status efail(fe_engine *, fe_status s, const std::string &m) {
  g_engine_error = m;
  return s;
}

static fe_status ehip(fe_engine *e, hipError_t h, const char *what) {
  if (h == hipSuccess) return FE_OK;
  return efail(e, h == hipErrorOutOfMemory ? FE_ERR_OOM : FE_ERR_HIP,
               std::string(what) + ": " + hipGetErrorString(h));
}

// Round layout in the staging buffers for K steps and F force blocks:
//   StepDesc[K] | int state[4K] | int force_prefix[K+1] | int apply_prefix[K+1]
//   | int block_step[F]
// state of step k: {epoch, rot, (rot + epoch) mod P, unused} of its job: a
// fused step reads its job's rotating slot layout and writes the free slot
// (the apply prefix is kept for the layout of the buffer; no apply runs).
// block_step[b] is the step of force block b: a block of a round of many
// jobs finds its descriptor with one load instead of a binary search over
// the prefix (log2 K dependent loads before any work).
constexpr int kStateInts = 4;
static size_t round_bytes(size_t K, size_t F) {
  size_t b = K * sizeof(StepDesc);
  b += K * kStateInts * sizeof(int) + 2 * (K + 1) * sizeof(int) + F * sizeof(int);
  return (b + 255) / 256 * 256;
}

// The bytes of a round's buffer before its block-to-step table, rounded up
// to 16: a round of at most kSearchSteps jobs runs without the table, its
// blocks finding their step by a binary search of the prefix (at most three
// dependent loads).
constexpr int kSearchSteps = 8;
static size_t round_head_bytes(size_t K) {
  const size_t b = K * sizeof(StepDesc) + K * kStateInts * sizeof(int) + 2 * (K + 1) * sizeof(int);
  return (b + 15) / 16 * 16;
}

// Whether a step of the next round of `run` has block items (a round
// without any takes the kernel without the block-item path).
static bool round_blk(const std::vector<EngineJob *> &run) {
  for (const EngineJob *ej : run)
    if (ej->job.host_desc[(size_t)ej->job.plan.batches * DESC_FUSED + ej->batch].n_block_items > 0)
      return true;
  return false;
}

// Force blocks of the next round of `run` (as build_round will lay them out).
static int round_force_blocks(const fe_engine *e, const std::vector<EngineJob *> &run) {
  const int ipb = e->tun.waves_per_block * e->tun.items_per_wave;
  int64_t f = 0;
  for (const EngineJob *ej : run)
    f += force_blocks_for(ej->job.host_desc[(size_t)ej->job.plan.batches * DESC_FUSED + ej->batch],
                          ipb);
  return (int)f;
}

// Staging capacity in bytes (each pinned stage and the device round buffer).
static fe_status ensure_stage(fe_engine *e, size_t bytes) {
  if (bytes <= e->stage_cap) return FE_OK;
  size_t cap = std::max<size_t>(round_bytes(64, 4096), e->stage_cap);
  while (cap < bytes) cap *= 2;
  hipStreamSynchronize(e->stream);
  // a failure below leaves no stage usable: the next call allocates again
  e->stage_cap = 0;
  for (auto &s : e->stage) {
    if (s.host) hipHostFree(s.host);
    s.host = nullptr;
    fe_status st = ehip(e, hipHostMalloc((void **)&s.host, cap), "pinned staging");
    if (st != FE_OK) return st;
  }
  if (e->dev_round) hipFree(e->dev_round);
  e->dev_round = nullptr;
  fe_status st = ehip(e, hipMalloc((void **)&e->dev_round, cap), "round buffer");
  if (st != FE_OK) return st;
  e->stage_cap = cap;
  return FE_OK;
}

static fe_status validate_desc(const fe_job_desc *d, std::string &err) {
  if (!d) {
    err = "job description is NULL";
    return FE_ERR_INVALID_ARG;
  }
  if (!d->row_ptr || (!d->col_idx && d->nnz > 0) || !d->x0 || d->batch_size <= 0 ||
      d->iterations < 0 || d->ns < 0 || d->n <= 0) {
    err = "job description: missing array or non-positive size";
    return FE_ERR_INVALID_ARG;
  }
  return FE_OK;
}

// Host-side plan of a job description (for sizing before any allocation).
static fe_status plan_job(const fe_job_desc *d, const Tuning &tun, Job &tmp, std::string &err) {
  fe_status s = validate_desc(d, err);
  if (s != FE_OK) return s;
  if (d->d < 1 || d->d > 512) {
    err = "job description: d must be in [1, 512]";
    return FE_ERR_INVALID_ARG;
  }
  tmp.n = d->n;
  tmp.dx = d->d;
  tmp.d = padded_width(d->d); // the device width, as job_init sets it
  tmp.dtype = d->gather_dtype;
  tmp.nnz = d->nnz;
  tmp.ns = d->ns;
  tmp.tun = tun;
  const int64_t c_limit = (d->quirks & FE_QUIRK_SKIP_LAST_COL) ? d->n - 1 : d->n;
  if (tmp.d < tun.slice) {
    err = "job description: d below the engine's slice (" + std::to_string(tun.slice) +
          " components per lane)";
    return FE_ERR_INVALID_ARG;
  }
  s = build_plan(d->n, d->row_ptr, nullptr, d->batch_size, tun.chunk, tun.max_chunks, tun.light,
                 64 / (tmp.d / tun.slice), c_limit, &tmp.plan);
  if (s != FE_OK) {
    err = host_error();
    return s;
  }
  tmp.n_lists = (int)(d->iterations + tmp.plan.batches - 1);
  return FE_OK;
}

extern "C" fe_status fe_engine_create(fe_engine **out, int device, int64_t budget) {
  if (!out || budget <= 0) {
    g_engine_error = "fe_engine_create: invalid argument";
    return FE_ERR_INVALID_ARG;
  }
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    g_engine_error = "fe_engine_create: no HIP device available";
    return FE_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= count) {
    g_engine_error = "fe_engine_create: device index out of range";
    return FE_ERR_NO_DEVICE;
  }
  auto *e = new (std::nothrow) fe_engine();
  if (!e) return FE_ERR_OOM;
  e->device = device;
  fe_status s = ehip(e, hipSetDevice(device), "hipSetDevice");
  if (s == FE_OK) s = ehip(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "stream");
  if (s == FE_OK) s = ehip(e, hipStreamCreateWithFlags(&e->upload, hipStreamNonBlocking), "stream");
  if (s == FE_OK) s = ehip(e, hipStreamCreateWithFlags(&e->download, hipStreamNonBlocking), "stream");
  for (auto &st : e->stage)
    if (s == FE_OK) s = ehip(e, hipEventCreateWithFlags(&st.copied, hipEventDisableTiming), "event");
  if (s == FE_OK) {
    e->arena.reset(new Arena());
    s = e->arena->init((size_t)budget, g_engine_error);
  }
  if (s == FE_OK) s = ensure_stage(e, round_bytes(64, 4096));
  if (s != FE_OK) {
    g_engine_error = "fe_engine_create: " + g_engine_error;
    fe_engine_destroy(e);
    return s;
  }
  *out = e;
  return FE_OK;
}

extern "C" void fe_engine_destroy(fe_engine *e) {
  if (!e) return;
  hipSetDevice(e->device);
  if (e->stream) hipStreamSynchronize(e->stream);
  if (e->upload) hipStreamSynchronize(e->upload);
  if (e->download) hipStreamSynchronize(e->download);
  for (auto &jp : e->jobs) {
    job_free(jp->job);
    if (jp->uploaded) hipEventDestroy(jp->uploaded);
    if (jp->done) hipEventDestroy(jp->done);
  }
  e->jobs.clear();
  for (auto &st : e->stage) {
    if (st.host) hipHostFree(st.host);
    if (st.copied) hipEventDestroy(st.copied);
  }
  if (e->dev_round) hipFree(e->dev_round);
  e->arena.reset();
  if (e->stream) hipStreamDestroy(e->stream);
  if (e->upload) hipStreamDestroy(e->upload);
  if (e->download) hipStreamDestroy(e->download);
  delete e;
}

extern "C" const char *fe_engine_last_error(const fe_engine *) {
  return g_engine_error.c_str();
}

static Tuning engine_tuning(const fe_engine *e) {
  if (!e) return Tuning();
  std::lock_guard<std::mutex> lk(const_cast<fe_engine *>(e)->mu);
  return e->tun;
}

static int64_t job_bytes_for(const fe_job_desc *d, const Tuning &tun) {
  Job tmp;
  std::string err;
  if (plan_job(d, tun, tmp, err) != FE_OK) return -1;
  // arena rounding: every buffer is a multiple of 256 bytes
  return job_bytes_estimate(tmp) + 16 * (int64_t)kAlign;
}

extern "C" int64_t fe_engine_job_bytes(const fe_engine *e, const fe_job_desc *d) {
  return job_bytes_for(d, engine_tuning(e));
}

extern "C" fe_status fe_engine_set_tuning(fe_engine *e, const fe_tuning *t) {
  if (!e || !t) return FE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(e->mu);
  Tuning nt = e->tun;
  if (t->chunk) nt.chunk = t->chunk;
  if (t->max_chunks) nt.max_chunks = t->max_chunks;
  if (t->unroll) nt.unroll = t->unroll;
  if (t->waves_per_block) nt.waves_per_block = t->waves_per_block;
  if (t->items_per_wave) nt.items_per_wave = t->items_per_wave;
  if (t->light) nt.light = t->light < 0 ? 0 : t->light;
  if (t->slice) nt.slice = t->slice;
  Job probe;
  if (job_set_tuning(probe, nt) != FE_OK) return efail(e, FE_ERR_INVALID_ARG, probe.err);
  // one kernel serves a round, so the slice (a cutting field, fixed when a
  // job is planned) changes only while the engine holds no job
  if (nt.slice != e->tun.slice) {
    for (auto &jp : e->jobs)
      if (!jp->collected)
        return efail(e, FE_ERR_STATE,
                     "fe_engine_set_tuning: the slice changes only while the engine holds no job");
  }
  // chunking of already submitted jobs is fixed; launch shape applies to all
  e->tun = nt;
  return FE_OK;
}

// The job is built and uploaded without the engine's lock (the arena locks
// itself); only its publication in the job list takes the lock.
extern "C" fe_status fe_engine_submit(fe_engine *e, const fe_job_desc *d, int *job_id) {
  if (!e || !job_id) return FE_ERR_INVALID_ARG;
  fe_status s = ehip(e, hipSetDevice(e->device), "hipSetDevice");
  if (s != FE_OK) return s;
  const Tuning tun = engine_tuning(e);
  // the plan sizes the job and is then handed to it (built once)
  Job sizing;
  std::string perr;
  if (plan_job(d, tun, sizing, perr) != FE_OK) {
    std::string err;
    validate_desc(d, err);
    return efail(e, FE_ERR_INVALID_ARG, "fe_engine_submit: " + (err.empty() ? perr : err));
  }
  const int64_t need = job_bytes_estimate(sizing) + 16 * (int64_t)kAlign;
  if ((size_t)need > e->arena->size() - e->arena->in_use())
    return efail(e, FE_ERR_OVER_BUDGET, "fe_engine_submit: job needs " + std::to_string(need) +
                                            " bytes, " +
                                            std::to_string(e->arena->size() - e->arena->in_use()) +
                                            " free in the budget");
  auto ej = std::make_unique<EngineJob>();
  Job &j = ej->job;
  j.allocator = e->arena.get();
  j.tun = tun;
  // X at its full slot size from the start: no second copy while growing
  s = job_init(j, d->n, d->d, d->gather_dtype, d->quirks, d->batch_size);
  j.pending_plan = std::move(sizing.plan);
  j.pending_B = d->batch_size;
  if (s == FE_OK) s = job_upload_csr(j, d->row_ptr, d->col_idx, d->nnz, e->upload);
  if (s == FE_OK) s = job_upload_x(j, d->x0, e->upload);
  if (s == FE_OK) s = job_configure(j, d->batch_size, d->lr, d->max_bound);
  std::vector<int32_t> lists;
  if (s == FE_OK) {
    const int batches = (int)((d->n + d->batch_size - 1) / d->batch_size);
    const int n_lists = d->iterations + batches - 1;
    if (d->ns > 0 && n_lists > 0) {
      if (d->neg_lists) {
        s = job_set_negatives(j, d->neg_lists, n_lists, d->ns, e->upload);
      } else {
        lists.resize((size_t)n_lists * d->ns);
        s = fe_reference_negatives(d->n, n_lists, d->ns, FE_NEG_REDRAW, lists.data(), nullptr);
        if (s != FE_OK) j.err = host_error();
        else s = job_set_negatives(j, lists.data(), n_lists, d->ns, e->upload);
      }
    }
  }
  // every slot of the layout: rounds write new rows into the free slot
  if (s == FE_OK) s = job_prepare(j, e->upload, true);
  if (s == FE_OK) s = ehip(e, hipEventCreateWithFlags(&ej->uploaded, hipEventDisableTiming), "event");
  if (s == FE_OK) s = ehip(e, hipEventCreateWithFlags(&ej->done, hipEventDisableTiming), "event");
  if (s == FE_OK) s = ehip(e, hipEventRecord(ej->uploaded, e->upload), "event");
  if (s != FE_OK) {
    if (g_engine_error.empty() || s != FE_ERR_HIP) g_engine_error = "fe_engine_submit: " + j.err;
    if (s == FE_ERR_OOM) s = FE_ERR_OVER_BUDGET;
    hipStreamSynchronize(e->upload);
    job_free(j);
    if (ej->uploaded) hipEventDestroy(ej->uploaded);
    if (ej->done) hipEventDestroy(ej->done);
    return s;
  }
  ej->iterations = d->iterations;
  ej->launched_all = d->iterations == 0;
  if (ej->launched_all) hipEventRecord(ej->done, e->upload);
  std::lock_guard<std::mutex> lk(e->mu);
  ej->id = (int)e->jobs.size();
  *job_id = ej->id;
  e->jobs.push_back(std::move(ej));
  return FE_OK;
}

// Step descriptors of the next round of every running job; advances them.
// Kernel variant of a round: 0 all FP32 gathers, 1 all BF16, 2 mixed.
static int round_mode(const std::vector<EngineJob *> &run) {
  bool f = false, h = false;
  for (const EngineJob *ej : run) (ej->job.dtype == FE_GATHER_BF16 ? h : f) = true;
  return (f && h) ? 2 : (h ? 1 : 0);
}

// Lanes per row shared by every step of the round (0 if they differ).
static int round_g(const std::vector<EngineJob *> &run) {
  auto lanes = [](const EngineJob *ej) { return ej->job.d / ej->job.tun.slice; };
  int g = run.empty() ? 0 : lanes(run[0]);
  for (const EngineJob *ej : run)
    if (lanes(ej) != g) return 0;
  return g;
}

// Components per lane of the round's steps: one kernel serves them all, so
// every job must have been planned with the engine's slice (0 otherwise).
static int round_slice(const fe_engine *e, const std::vector<EngineJob *> &run) {
  for (const EngineJob *ej : run)
    if (ej->job.tun.slice != e->tun.slice) return 0;
  return e->tun.slice;
}

static fe_status build_round(fe_engine *e, std::vector<EngineJob *> &run, char *host, int &K,
                             size_t &lds, std::vector<EngineJob *> &finished) {
  K = (int)run.size();
  StepDesc *sd = reinterpret_cast<StepDesc *>(host);
  int *ep = reinterpret_cast<int *>(host + K * sizeof(StepDesc));
  int *pf = ep + kStateInts * K;
  int *pa = pf + K + 1;
  int *dep = reinterpret_cast<int *>(e->dev_round + K * sizeof(StepDesc));
  const int ipb = e->tun.waves_per_block * e->tun.items_per_wave;
  pf[0] = pa[0] = 0;
  lds = 0;
  for (int k = 0; k < K; ++k) {
    EngineJob *ej = run[k];
    Job &j = ej->job;
    const int b = ej->batch;
    // a fused step: the batch's new rows go straight into the free slot of
    // the job's layout (no apply launch), read in place by its next round
    sd[k] = j.host_desc[(size_t)j.plan.batches * DESC_FUSED + b];
    sd[k].state = dep + kStateInts * k;
    ep[kStateInts * k] = ej->epoch;
    ep[kStateInts * k + 1] = j.rot;
    ep[kStateInts * k + 2] = (int)(((int64_t)j.rot + ej->epoch) % j.lay.P);
    ep[kStateInts * k + 3] = 0;
    // this round's negative ids, straight from the job's lists
    const int list = ej->epoch + b;
    sd[k].neg_ids = (j.ns > 0 && list < j.n_lists) ? j.negs + (size_t)list * j.ns : j.negs;
    pf[k + 1] = pf[k] + force_blocks_for(sd[k], ipb);
    pa[k + 1] = pa[k] + j.apply_blocks[b];
    lds = std::max(lds, force_lds_bytes(j.ns, j.d));
    int r0, r1;
    rep_rows(j.plan, j.quirks, b, &r0, &r1);
    ej->units += j.plan.edges_eff[b] + (int64_t)(ej->epoch == 0 ? r0 : r1) * j.ns;
    e->units_done += j.plan.edges_eff[b] + (int64_t)(ej->epoch == 0 ? r0 : r1) * j.ns;
    if (++ej->batch == j.plan.batches) {
      ej->batch = 0;
      if (++ej->epoch == ej->iterations) {
        ej->launched_all = true;
        finished.push_back(ej);
        // after its last round every batch has moved down `iterations` slots
        job_advance_layout(j, ej->iterations);
      }
    }
  }
  int *bs = pa + K + 1;
  for (int k = 0; k < K; ++k)
    for (int b = pf[k]; b < pf[k + 1]; ++b) bs[b] = k;
  return FE_OK;
}

// Jobs of a round that could not be enqueued have lost steps: they end
// without a result (collect reports the failure and frees them).
static void mark_failed(const std::vector<EngineJob *> &run) {
  for (EngineJob *ej : run) {
    ej->failed = true;
    ej->launched_all = true;
  }
}

static fe_status run_rounds(fe_engine *e, int rounds);

extern "C" fe_status fe_engine_run(fe_engine *e, int rounds, int *running) {
  if (!e) return FE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(e->mu);
  fe_status s = ehip(e, hipSetDevice(e->device), "hipSetDevice");
  if (s == FE_OK) s = run_rounds(e, rounds);
  if (running) {
    int n = 0;
    for (auto &jp : e->jobs)
      if (!jp->launched_all && !jp->collected) ++n;
    *running = n;
  }
  return s;
}

// A round of one job: its step descriptor goes by value in the kernel
// arguments of the single-step kernel, reading the epoch state from the
// job's own device ints (set by a one-thread kernel when its epoch changes).
// No staging copy precedes the launch (that copy and its wait cost about
// 8 us a round), and the kernel body is the same, so the bits are those of
// a round of many jobs.
static fe_status run_one(fe_engine *e, std::vector<EngineJob *> &run, size_t rbytes) {
  const int mode = round_mode(run), g = round_g(run), slice = round_slice(e, run);
  if (slice == 0) {
    mark_failed(run);
    return efail(e, FE_ERR_STATE, "fe_engine_run: a job's slice differs from the engine's");
  }
  EngineJob *ej = run[0];
  Job &j = ej->job;
  const int epoch = ej->epoch;
  fe_status s = FE_OK;
  if (ej->state_epoch != epoch) {
    const int shift = (int)(((int64_t)j.rot + epoch) % j.lay.P);
    if ((s = ehip(e, launch_state(j.epoch, epoch, j.rot, shift, e->stream), "state set")) !=
        FE_OK) {
      mark_failed(run);
      return s;
    }
    ej->state_epoch = epoch;
  }
  e->one_host.resize(rbytes);
  int K = 0;
  size_t lds = 0;
  std::vector<EngineJob *> finished;
  build_round(e, run, e->one_host.data(), K, lds, finished);
  // from here on the job has advanced: a failure ends it
  StepDesc sd = *reinterpret_cast<const StepDesc *>(e->one_host.data());
  sd.state = j.epoch;
  const int *pf = reinterpret_cast<const int *>(e->one_host.data() + sizeof(StepDesc)) + kStateInts;
  s = ehip(e, launch_force1(sd, pf[1], e->tun.waves_per_block, e->tun.unroll, lds, mode, g, slice,
                            e->stream),
           "force launch");
  for (EngineJob *f : finished)
    if (s == FE_OK) s = ehip(e, hipEventRecord(f->done, e->stream), "event");
  if (s != FE_OK) mark_failed(run);
  return s;
}

static fe_status run_rounds(fe_engine *e, int rounds) {
  fe_status s = FE_OK;
  int done_rounds = 0;
  while (rounds <= 0 || done_rounds < rounds) {
    std::vector<EngineJob *> run;
    for (auto &jp : e->jobs)
      if (!jp->launched_all && !jp->collected) run.push_back(jp.get());
    if (run.empty()) break;
    const size_t rbytes = round_bytes(run.size(), (size_t)round_force_blocks(e, run));
    if (run.size() > 1 && (s = ensure_stage(e, rbytes)) != FE_OK) return s;
    // jobs whose uploads are still in flight join once they have landed
    for (EngineJob *ej : run) {
      if (!ej->waited_upload) {
        if ((s = ehip(e, hipStreamWaitEvent(e->stream, ej->uploaded, 0), "wait upload")) != FE_OK)
          return s;
        ej->waited_upload = true;
      }
    }
    if (run.size() == 1) {
      if ((s = run_one(e, run, rbytes)) != FE_OK) return s;
      ++e->rounds_run;
      ++done_rounds;
      continue;
    }
    // a round of up to kSearchSteps jobs goes to the device without its
    // block-to-step table (its blocks search the prefix) as the arguments of
    // a one-block kernel; a larger one, whole, by a DMA copy from a pinned
    // stage
    const size_t head = round_head_bytes(run.size());
    const bool small = run.size() <= (size_t)kSearchSteps && head <= (size_t)kRoundBlobBytes;
    auto &stg = e->stage[e->stage_next];
    if (!small) {
      e->stage_next = (e->stage_next + 1) % kStages;
      if ((s = ehip(e, hipEventSynchronize(stg.copied), "staging reuse")) != FE_OK) return s;
    } else {
      e->one_host.resize(rbytes);
    }
    char *host = small ? e->one_host.data() : stg.host;
    int K = 0;
    size_t lds = 0;
    std::vector<EngineJob *> finished;
    const int slice = round_slice(e, run);
    if (slice == 0) {
      mark_failed(run);
      return efail(e, FE_ERR_STATE, "fe_engine_run: a job's slice differs from the engine's");
    }
    const bool blk = round_blk(run);
    build_round(e, run, host, K, lds, finished);
    // from here on the jobs have advanced: a failure ends them all
    const int *hpf = reinterpret_cast<const int *>(host + K * sizeof(StepDesc)) + kStateInts * K;
    const int fblocks = hpf[K];
    if (small) {
      s = ehip(e, launch_blob(host, head, e->dev_round, e->stream), "round upload");
    } else {
      s = ehip(e, hipMemcpyAsync(e->dev_round, host, rbytes, hipMemcpyHostToDevice, e->stream),
               "round copy");
      if (s == FE_OK) s = ehip(e, hipEventRecord(stg.copied, e->stream), "event");
    }
    const StepDesc *dsd = reinterpret_cast<const StepDesc *>(e->dev_round);
    const int *dpf = reinterpret_cast<const int *>(e->dev_round + K * sizeof(StepDesc)) + kStateInts * K;
    if (s == FE_OK)
      s = ehip(e, launch_force(dsd, dpf, small ? nullptr : dpf + 2 * (K + 1), K, fblocks,
                               e->tun.waves_per_block, e->tun.items_per_wave, e->tun.unroll, lds,
                               round_mode(run), round_g(run), slice, blk, e->stream),
               "force launch");
    for (EngineJob *ej : finished)
      if (s == FE_OK) s = ehip(e, hipEventRecord(ej->done, e->stream), "event");
    if (s != FE_OK) {
      mark_failed(run);
      return s;
    }
    ++e->rounds_run;
    ++done_rounds;
  }
  return FE_OK;
}

extern "C" fe_status fe_engine_job_done(fe_engine *e, int job_id, int *done) {
  if (!e || !done) return FE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(e->mu);
  if (job_id < 0 || job_id >= (int)e->jobs.size())
    return efail(e, FE_ERR_INVALID_ARG, "fe_engine_job_done: bad job id");
  EngineJob *ej = e->jobs[job_id].get();
  if (ej->collected || ej->collecting || ej->failed) {
    *done = 1;
    return FE_OK;
  }
  if (!ej->launched_all) {
    *done = 0;
    return FE_OK;
  }
  hipError_t q = hipEventQuery(ej->done);
  if (q == hipErrorNotReady) {
    (void)hipGetLastError();
    *done = 0;
    return FE_OK;
  }
  if (q != hipSuccess) return ehip(e, q, "fe_engine_job_done");
  *done = 1;
  return FE_OK;
}

// The download runs without the engine's lock, on its own stream, after the
// job's last round: rounds of the other jobs go on meanwhile.
extern "C" fe_status fe_engine_collect(fe_engine *e, int job_id, float *x_out) {
  if (!e || !x_out) return e ? efail(e, FE_ERR_INVALID_ARG, "fe_engine_collect: bad argument")
                             : FE_ERR_INVALID_ARG;
  fe_status s = ehip(e, hipSetDevice(e->device), "hipSetDevice");
  if (s != FE_OK) return s;
  EngineJob *ej = nullptr;
  {
    std::lock_guard<std::mutex> lk(e->mu);
    if (job_id < 0 || job_id >= (int)e->jobs.size())
      return efail(e, FE_ERR_INVALID_ARG, "fe_engine_collect: bad job id");
    ej = e->jobs[job_id].get();
    if (ej->collected || ej->collecting)
      return efail(e, FE_ERR_STATE, "fe_engine_collect: already collected");
    if (!ej->launched_all) return efail(e, FE_ERR_STATE, "fe_engine_collect: job still running");
    if (ej->failed) {
      // its memory is freed once the device no longer uses it
      hipStreamSynchronize(e->stream);
      job_free(ej->job);
      hipEventDestroy(ej->uploaded);
      hipEventDestroy(ej->done);
      ej->uploaded = ej->done = nullptr;
      ej->collected = true;
      return efail(e, FE_ERR_STATE, "fe_engine_collect: a round of this job failed; no result");
    }
    ej->collecting = true;
  }
  Job &j = ej->job;
  s = ehip(e, hipStreamWaitEvent(e->download, ej->done, 0), "wait done");
  if (s == FE_OK) {
    s = job_download_x(j, x_out, e->download);
    if (s != FE_OK) efail(e, s, j.err);
  }
  std::lock_guard<std::mutex> lk(e->mu);
  ej->collecting = false;
  if (s != FE_OK) return s;
  job_free(j);
  hipEventDestroy(ej->uploaded);
  hipEventDestroy(ej->done);
  ej->uploaded = ej->done = nullptr;
  ej->collected = true;
  return FE_OK;
}

extern "C" fe_status fe_engine_stats(const fe_engine *e, int64_t *bytes_in_use,
                                     int64_t *rounds_run, int64_t *units_done) {
  if (!e) return FE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(const_cast<fe_engine *>(e)->mu);
  if (bytes_in_use) *bytes_in_use = (int64_t)e->arena->in_use();
  if (rounds_run) *rounds_run = e->rounds_run;
  if (units_done) *units_done = e->units_done;
  return FE_OK;
}

// Batching form: every submitted job runs to completion through rounds whose
// descriptors are all built up front and whose launches are captured in HIP
// graphs (segments of kGraphRounds rounds), each replayed once.
extern "C" fe_status fe_engine_run_graph(fe_engine *e) {
  if (!e) return FE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(e->mu);
  fe_status s = ehip(e, hipSetDevice(e->device), "hipSetDevice");
  if (s != FE_OK) return s;
  std::vector<EngineJob *> all;
  for (auto &jp : e->jobs)
    if (!jp->launched_all && !jp->collected) all.push_back(jp.get());
  if (all.empty()) return FE_OK;
  const auto t0 = std::chrono::steady_clock::now();
  if ((s = ehip(e, hipStreamSynchronize(e->upload), "upload sync")) != FE_OK) return s;
  // schedule: round r holds every job that still has steps at r
  struct RoundInfo {
    size_t offset; // bytes into the schedule buffer
    int K;
    int fblocks, ablocks;
    size_t lds;
    int mode;
    int g_const;
    int slice;
    bool blk;
  };
  std::vector<RoundInfo> rounds;
  std::vector<char> host;
  std::vector<std::pair<int, EngineJob *>> finish_at; // round index -> job
  size_t total = 0;
  while (true) {
    std::vector<EngineJob *> run;
    for (EngineJob *ej : all)
      if (!ej->launched_all) run.push_back(ej);
    if (run.empty()) break;
    const size_t rb = round_bytes(run.size(), (size_t)round_force_blocks(e, run));
    host.resize(total + rb);
    int K = 0;
    size_t lds = 0;
    std::vector<EngineJob *> finished;
    // descriptors point at this round's own epoch slots in the device schedule
    char *h = host.data() + total;
    const int slice = round_slice(e, run);
    if (slice == 0) {
      mark_failed(all);
      return efail(e, FE_ERR_STATE, "fe_engine_run_graph: a job's slice differs from the engine's");
    }
    const bool blk = round_blk(run);
    build_round(e, run, h, K, lds, finished);
    const int *hpf = reinterpret_cast<const int *>(h + K * sizeof(StepDesc)) + kStateInts * K;
    rounds.push_back(
        {total, K, hpf[K], hpf[2 * K + 1], lds, round_mode(run), round_g(run), slice, blk});
    for (EngineJob *ej : finished) finish_at.push_back({(int)rounds.size() - 1, ej});
    total += rb;
  }
  char *dsched = nullptr;
  if ((s = ehip(e, hipMalloc((void **)&dsched, std::max<size_t>(total, 256)), "schedule buffer")) != FE_OK) {
    mark_failed(all);
    return s;
  }
  // rebase the per-round epoch pointers (built against dev_round) onto the schedule
  for (const RoundInfo &ri : rounds) {
    StepDesc *sd = reinterpret_cast<StepDesc *>(host.data() + ri.offset);
    int *dep = reinterpret_cast<int *>(dsched + ri.offset + ri.K * sizeof(StepDesc));
    for (int k = 0; k < ri.K; ++k) sd[k].state = dep + kStateInts * k;
  }
  s = ehip(e, hipMemcpy(dsched, host.data(), total, hipMemcpyHostToDevice), "schedule copy");
  // capture and instantiate every segment first, then replay them back to back
  constexpr int kGraphRounds = 512;
  std::vector<hipGraph_t> graphs;
  std::vector<hipGraphExec_t> execs;
  std::vector<size_t> seg_end;
  for (size_t r0 = 0; s == FE_OK && r0 < rounds.size(); r0 += kGraphRounds) {
    const size_t r1 = std::min(rounds.size(), r0 + kGraphRounds);
    if ((s = ehip(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), "capture")) != FE_OK)
      break;
    fe_status ls = FE_OK;
    for (size_t r = r0; r < r1 && ls == FE_OK; ++r) {
      const RoundInfo &ri = rounds[r];
      const StepDesc *dsd = reinterpret_cast<const StepDesc *>(dsched + ri.offset);
      const int *dpf = reinterpret_cast<const int *>(dsched + ri.offset + ri.K * sizeof(StepDesc)) + kStateInts * ri.K;
      // rounds of one or two steps search the prefix (one load)
      ls = ehip(e, launch_force(dsd, dpf, ri.K > 2 ? dpf + 2 * (ri.K + 1) : nullptr, ri.K, ri.fblocks,
                                e->tun.waves_per_block, e->tun.items_per_wave, e->tun.unroll,
                                ri.lds, ri.mode, ri.g_const, ri.slice, ri.blk, e->stream),
                "force launch");
    }
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (ls != FE_OK) {
      s = ls;
      if (g) hipGraphDestroy(g);
      break;
    }
    if ((s = ehip(e, ce, "end capture")) != FE_OK) break;
    graphs.push_back(g);
    hipGraphExec_t ge = nullptr;
    if ((s = ehip(e, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "graph instantiate")) != FE_OK)
      break;
    execs.push_back(ge);
    seg_end.push_back(r1);
  }
  const auto t1 = std::chrono::steady_clock::now();
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (s == FE_OK) s = ehip(e, hipEventCreate(&ev0), "event");
  if (s == FE_OK) s = ehip(e, hipEventCreate(&ev1), "event");
  if (s == FE_OK) s = ehip(e, hipEventRecord(ev0, e->stream), "event");
  size_t fi = 0;
  for (size_t gi = 0; s == FE_OK && gi < execs.size(); ++gi) {
    s = ehip(e, hipGraphLaunch(execs[gi], e->stream), "graph launch");
    // job completion events after the segment holding their last round
    while (s == FE_OK && fi < finish_at.size() && (size_t)finish_at[fi].first < seg_end[gi]) {
      s = ehip(e, hipEventRecord(finish_at[fi].second->done, e->stream), "event");
      ++fi;
    }
  }
  if (s == FE_OK) s = ehip(e, hipEventRecord(ev1, e->stream), "event");
  if (s == FE_OK) s = ehip(e, hipStreamSynchronize(e->stream), "graph sync");
  float ms = 0.f;
  if (s == FE_OK) s = ehip(e, hipEventElapsedTime(&ms, ev0, ev1), "event time");
  if (s != FE_OK) mark_failed(all);
  e->graph_exec_ms = ms;
  e->graph_prepare_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  e->graph_rounds = (int64_t)rounds.size();
  e->rounds_run += (int64_t)rounds.size();
  hipStreamSynchronize(e->stream);
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  for (auto ge : execs) hipGraphExecDestroy(ge);
  for (auto g : graphs) hipGraphDestroy(g);
  hipFree(dsched);
  return s;
}

extern "C" fe_status fe_engine_graph_timing(const fe_engine *e, double *prepare_ms,
                                            double *exec_ms, int64_t *rounds) {
  if (!e) return FE_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(const_cast<fe_engine *>(e)->mu);
  if (prepare_ms) *prepare_ms = e->graph_prepare_ms;
  if (exec_ms) *exec_ms = e->graph_exec_ms;
  if (rounds) *rounds = e->graph_rounds;
  return FE_OK;
}

// Frees a job that has finished or failed without downloading its result.
static void drop_job(fe_engine *e, int id) {
  std::lock_guard<std::mutex> lk(e->mu);
  if (id < 0 || id >= (int)e->jobs.size()) return;
  EngineJob *ej = e->jobs[id].get();
  if (ej->collected) return;
  hipStreamSynchronize(e->stream);
  hipStreamSynchronize(e->upload);
  job_free(ej->job);
  if (ej->uploaded) hipEventDestroy(ej->uploaded);
  if (ej->done) hipEventDestroy(ej->done);
  ej->uploaded = ej->done = nullptr;
  ej->launched_all = true;
  ej->collected = true;
}

// Makes finished jobs run all their epochs again from their current rows
// (a later replay of the same work, for timing).
static void rewind_jobs(fe_engine *e, const std::vector<int> &ids) {
  std::lock_guard<std::mutex> lk(e->mu);
  for (int id : ids) {
    EngineJob *ej = e->jobs[id].get();
    ej->epoch = 0;
    ej->batch = 0;
    ej->launched_all = ej->iterations == 0;
  }
}

// Tiles of the batching form chosen by timing them on the workload itself:
// a round mixes the batches of many jobs of different sizes, so a tiling
// that suits one graph alone can lose on the mix. The jobs are submitted
// once per cutting (slice and light); every launch shape then replays all their
// epochs through the captured-graph schedule from the rows the previous
// replay left (the work items, and so the time, do not depend on the
// values). The fastest replay (device time, HIP events) becomes the
// engine's tuning.
extern "C" fe_status fe_engine_autotune(fe_engine *e, const fe_job_desc *jobs, int n_jobs,
                                        int epochs, fe_tuning *results, double *results_ms,
                                        int max_results, int *n_results, fe_tuning *best,
                                        double *best_ms) {
  if (!e || !jobs || n_jobs <= 0 || epochs < 1)
    return e ? efail(e, FE_ERR_INVALID_ARG, "fe_engine_autotune: invalid argument")
             : FE_ERR_INVALID_ARG;
  {
    std::lock_guard<std::mutex> lk(e->mu);
    for (auto &jp : e->jobs)
      if (!jp->collected)
        return efail(e, FE_ERR_STATE,
                     "fe_engine_autotune: the engine holds jobs; collect them first");
  }
  const Tuning orig = engine_tuning(e);
  std::vector<fe_job_desc> descs(jobs, jobs + n_jobs);
  for (fe_job_desc &d : descs) d.iterations = epochs;
  fe_status s = FE_OK;
  double best_t = 1e300;
  fe_tuning best_c{};
  int n = 0;
  std::vector<int> ids;
  int min_d = 512;
  for (const fe_job_desc &d : descs)
    min_d = std::min(min_d, (d.d >= 1 && d.d <= 512) ? padded_width(d.d) : d.d);
  std::vector<std::pair<int, int>> cuts; // (slice, light)
  for (int slice : {8, 16})
    for (int light : {8, 16, 32})
      if (slice <= min_d) cuts.push_back({slice, light});
  for (const auto &sl : cuts) {
    const int slice = sl.first, light = sl.second;
    fe_tuning cut{64, 1024, orig.unroll, orig.waves_per_block, orig.items_per_wave, light, slice};
    if ((s = fe_engine_set_tuning(e, &cut)) != FE_OK) break;
    ids.clear();
    for (const fe_job_desc &d : descs) {
      int id = -1;
      if ((s = fe_engine_submit(e, &d, &id)) != FE_OK) break;
      ids.push_back(id);
    }
    for (int u : {2, 4, 8}) {
      for (int w : {4, 8}) {
        for (int ipw : {1, 2}) {
          if (s != FE_OK) break;
          const fe_tuning c{64, 1024, u, w, ipw, light, slice};
          if ((s = fe_engine_set_tuning(e, &c)) != FE_OK) break;
          rewind_jobs(e, ids);
          if ((s = fe_engine_run_graph(e)) != FE_OK) break;
          double ms = 0;
          if ((s = fe_engine_graph_timing(e, nullptr, &ms, nullptr)) != FE_OK) break;
          if (results && n < max_results) {
            results[n] = c;
            if (results_ms) results_ms[n] = ms;
          }
          ++n;
          if (ms < best_t) {
            best_t = ms;
            best_c = c;
          }
        }
      }
    }
    for (int id : ids) drop_job(e, id);
    if (s != FE_OK) break;
  }
  if (s != FE_OK) {
    std::string msg = g_engine_error;
    std::lock_guard<std::mutex> lk(e->mu);
    e->tun = orig;
    g_engine_error = "fe_engine_autotune: " + msg;
    return s;
  }
  if ((s = fe_engine_set_tuning(e, &best_c)) != FE_OK) return s;
  if (n_results) *n_results = n;
  if (best) *best = best_c;
  if (best_ms) *best_ms = best_t;
  return FE_OK;
}
