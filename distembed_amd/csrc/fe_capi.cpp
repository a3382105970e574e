// C ABI of the single-graph context (include/force_embed.h).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fe_job.hpp"

using namespace fe;

struct fe_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Job job;
  // one captured epoch
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_valid = false;
  bool graph_timed = false;
  // per-launch event timing
  bool timing = false;
  std::vector<hipEvent_t> events; // 4 per recorded batch launch
  size_t events_used = 0;
  double t_force = 0, t_fixup = 0, t_apply = 0;
  int64_t n_timed = 0;
  std::string err;
};

static thread_local std::string g_create_error;

static fe_status fail(fe_ctx *c, fe_status s, const std::string &msg) {
  if (c) c->err = msg;
  return s;
}

static fe_status from_job(fe_ctx *c, fe_status s) {
  if (s != FE_OK) c->err = c->job.err;
  return s;
}

static fe_status set_device(fe_ctx *c) {
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return fail(c, FE_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  return FE_OK;
}

static void drop_graph(fe_ctx *c) {
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_valid = false;
}

extern "C" fe_status fe_create(fe_ctx **out, int device, int64_t n, int d, int gather_dtype,
                               uint32_t quirks) {
  if (!out) {
    g_create_error = "fe_create: out is NULL";
    return FE_ERR_INVALID_ARG;
  }
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    g_create_error = "fe_create: no HIP device available";
    return FE_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= count) {
    g_create_error = "fe_create: device index out of range";
    return FE_ERR_NO_DEVICE;
  }
  fe_ctx *c = new (std::nothrow) fe_ctx();
  if (!c) {
    g_create_error = "fe_create: out of host memory";
    return FE_ERR_OOM;
  }
  c->device = device;
  fe_status s = set_device(c);
  if (s == FE_OK) {
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) s = fail(c, FE_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  if (s == FE_OK) s = from_job(c, job_init(c->job, n, d, gather_dtype, quirks));
  if (s != FE_OK) {
    g_create_error = "fe_create: " + c->err;
    job_free(c->job);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return s;
  }
  *out = c;
  return FE_OK;
}

extern "C" void fe_destroy(fe_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (hipEvent_t ev : c->events) hipEventDestroy(ev);
  job_free(c->job);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char *fe_last_error(const fe_ctx *c) {
  if (!c) return g_create_error.c_str();
  return c->err.c_str();
}

#define CTX_GUARD(c)                                                                            \
  do {                                                                                          \
    if (!(c)) return FE_ERR_INVALID_ARG;                                                        \
    fe_status _g = set_device(c);                                                               \
    if (_g != FE_OK) return _g;                                                                 \
  } while (0)

extern "C" fe_status fe_upload_csr(fe_ctx *c, const int64_t *row_ptr, const int32_t *col_idx,
                                   int64_t nnz) {
  CTX_GUARD(c);
  drop_graph(c);
  return from_job(c, job_upload_csr(c->job, row_ptr, col_idx, nnz, c->stream));
}

extern "C" fe_status fe_upload_x(fe_ctx *c, const float *x) {
  CTX_GUARD(c);
  return from_job(c, job_upload_x(c->job, x, c->stream));
}

extern "C" fe_status fe_download_x(fe_ctx *c, float *x) {
  CTX_GUARD(c);
  return from_job(c, job_download_x(c->job, x, c->stream));
}

extern "C" fe_status fe_set_negatives(fe_ctx *c, const int32_t *lists, int n_lists, int ns) {
  CTX_GUARD(c);
  drop_graph(c);
  return from_job(c, job_set_negatives(c->job, lists, n_lists, ns, c->stream));
}

extern "C" fe_status fe_configure(fe_ctx *c, int batch_size, float lr, float max_bound) {
  CTX_GUARD(c);
  Job &j = c->job;
  if (batch_size != j.B || lr != j.lr || max_bound != j.bound) drop_graph(c);
  return from_job(c, job_configure(j, batch_size, lr, max_bound));
}

extern "C" fe_status fe_set_tuning(fe_ctx *c, const fe_tuning *t) {
  CTX_GUARD(c);
  if (!t) return fail(c, FE_ERR_INVALID_ARG, "fe_set_tuning: NULL");
  Tuning nt = c->job.tun;
  if (t->chunk) nt.chunk = t->chunk;
  if (t->max_chunks) nt.max_chunks = t->max_chunks;
  if (t->unroll) nt.unroll = t->unroll;
  if (t->waves_per_block) nt.waves_per_block = t->waves_per_block;
  if (t->items_per_wave) nt.items_per_wave = t->items_per_wave;
  drop_graph(c);
  return from_job(c, job_set_tuning(c->job, nt));
}

extern "C" fe_status fe_get_tuning(const fe_ctx *c, fe_tuning *t) {
  if (!c || !t) return FE_ERR_INVALID_ARG;
  t->chunk = c->job.tun.chunk;
  t->max_chunks = c->job.tun.max_chunks;
  t->unroll = c->job.tun.unroll;
  t->waves_per_block = c->job.tun.waves_per_block;
  t->items_per_wave = c->job.tun.items_per_wave;
  return FE_OK;
}

static fe_status harvest_events(fe_ctx *c) {
  for (size_t i = 0; i + 3 < c->events_used; i += 4) {
    float a = 0, b = 0, d = 0;
    hipError_t e1 = hipEventElapsedTime(&a, c->events[i], c->events[i + 1]);
    hipError_t e2 = hipEventElapsedTime(&b, c->events[i + 1], c->events[i + 2]);
    hipError_t e3 = hipEventElapsedTime(&d, c->events[i + 2], c->events[i + 3]);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
      (void)hipGetLastError(); // do not leave the failure for the next launch check
      c->events_used = 0;
      return fail(c, FE_ERR_HIP, "kernel timing: event elapsed time unavailable");
    }
    c->t_force += a;
    c->t_fixup += b;
    c->t_apply += d;
    c->n_timed += 1;
  }
  c->events_used = 0;
  return FE_OK;
}

static hipEvent_t *next_events(fe_ctx *c) {
  if (!c->timing) return nullptr;
  while (c->events.size() < c->events_used + 4) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    c->events.push_back(ev);
  }
  hipEvent_t *p = &c->events[c->events_used];
  c->events_used += 4;
  return p;
}

static fe_status check_batch(fe_ctx *c, int epoch, int batch) {
  Job &j = c->job;
  if (batch < 0 || batch >= j.plan.batches || epoch < 0)
    return fail(c, FE_ERR_INVALID_ARG, "epoch or batch out of range");
  if (j.ns > 0 && (int64_t)epoch + batch >= j.n_lists)
    return fail(c, FE_ERR_INVALID_ARG, "no negative list for epoch + batch (set more lists)");
  return FE_OK;
}

extern "C" fe_status fe_step(fe_ctx *c, int epoch, int batch) {
  CTX_GUARD(c);
  fe_status s = from_job(c, job_prepare(c->job, c->stream));
  if (s != FE_OK) return s;
  if ((s = check_batch(c, epoch, batch)) != FE_OK) return s;
  if ((s = from_job(c, job_check(c->job, launch_set_int(c->job.epoch, epoch, c->stream),
                                 "epoch set"))) != FE_OK)
    return s;
  if ((s = from_job(c, job_launch_batch(c->job, batch, false, c->stream, next_events(c)))) != FE_OK)
    return s;
  return from_job(c, job_check(c->job, hipStreamSynchronize(c->stream), "fe_step sync"));
}

extern "C" fe_status fe_step_delta(fe_ctx *c, int epoch, int batch, float *delta_out) {
  CTX_GUARD(c);
  if (!delta_out) return fail(c, FE_ERR_INVALID_ARG, "fe_step_delta: NULL output");
  fe_status s = from_job(c, job_prepare(c->job, c->stream));
  if (s != FE_OK) return s;
  if ((s = check_batch(c, epoch, batch)) != FE_OK) return s;
  Job &j = c->job;
  if ((s = from_job(c, job_check(j, launch_set_int(j.epoch, epoch, c->stream), "epoch set"))) != FE_OK)
    return s;
  if ((s = from_job(c, job_launch_batch(j, batch, true, c->stream, nullptr))) != FE_OK) return s;
  const size_t count = (size_t)rows_of_batch(j, batch) * j.d;
  if ((s = from_job(c, job_check(j, hipMemcpyAsync(delta_out, j.S, sizeof(float) * count,
                                                   hipMemcpyDeviceToHost, c->stream),
                                 "delta copy"))) != FE_OK)
    return s;
  return from_job(c, job_check(j, hipStreamSynchronize(c->stream), "fe_step_delta sync"));
}

static fe_status capture_epoch(fe_ctx *c) {
  drop_graph(c);
  Job &j = c->job;
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail(c, FE_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  fe_status s = FE_OK;
  c->events_used = 0;
  for (int b = 0; b < j.plan.batches && s == FE_OK; ++b)
    s = from_job(c, job_launch_batch(j, b, false, c->stream, next_events(c)));
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(c->stream, &g);
  if (s != FE_OK) {
    if (g) hipGraphDestroy(g);
    c->events_used = 0;
    return s;
  }
  if (e != hipSuccess) {
    c->events_used = 0;
    return fail(c, FE_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t ge = nullptr;
  e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    c->events_used = 0;
    return fail(c, FE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  c->graph = g;
  c->graph_exec = ge;
  c->graph_valid = true;
  c->graph_timed = c->timing;
  return FE_OK;
}

extern "C" fe_status fe_train(fe_ctx *c, int iterations, int batch_size, int ns, float lr,
                              int use_graph) {
  CTX_GUARD(c);
  Job &j = c->job;
  if (iterations < 0 || ns < 0) return fail(c, FE_ERR_INVALID_ARG, "fe_train: negative argument");
  fe_status s = fe_configure(c, batch_size, lr, j.bound);
  if (s != FE_OK) return s;
  if ((s = from_job(c, job_prepare(j, c->stream))) != FE_OK) return s;
  const int batches = j.plan.batches;
  const int64_t need = (int64_t)iterations + batches - 1;
  if (ns > 0 && (j.ns != ns || j.n_lists < need)) {
    std::vector<int32_t> lists((size_t)std::max<int64_t>(need, 0) * ns);
    s = fe_reference_negatives(j.n, (int)need, ns, FE_NEG_REDRAW, lists.data(), nullptr);
    if (s != FE_OK) return fail(c, s, host_error());
    if ((s = fe_set_negatives(c, lists.data(), (int)need, ns)) != FE_OK) return s;
  } else if (ns == 0 && j.ns != 0) {
    if ((s = fe_set_negatives(c, nullptr, 0, 0)) != FE_OK) return s;
  }
  if ((s = from_job(c, job_prepare(j, c->stream))) != FE_OK) return s;
  if (iterations == 0) return FE_OK;
  if ((s = from_job(c, job_check(j, launch_set_int(j.epoch, 0, c->stream), "epoch set"))) != FE_OK)
    return s;
  if (use_graph) {
    if (!c->graph_valid || c->graph_timed != c->timing) {
      if ((s = capture_epoch(c)) != FE_OK) return s;
    }
    for (int it = 0; it < iterations; ++it) {
      hipError_t e = hipGraphLaunch(c->graph_exec, c->stream);
      if (e != hipSuccess) return fail(c, FE_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
      if (c->timing) {
        // the captured events are re-recorded by every replay: read them now
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return fail(c, FE_ERR_HIP, std::string("sync: ") + hipGetErrorString(e));
        const size_t used = c->events_used;
        if ((s = harvest_events(c)) != FE_OK) return s;
        c->events_used = used;
      }
    }
    if (c->timing) c->events_used = 0;
  } else {
    for (int it = 0; it < iterations; ++it) {
      for (int b = 0; b < batches; ++b) {
        if ((s = from_job(c, job_launch_batch(j, b, false, c->stream, next_events(c)))) != FE_OK)
          return s;
      }
      if (c->timing && c->events_used >= 4096) {
        hipStreamSynchronize(c->stream);
        harvest_events(c);
      }
    }
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, FE_ERR_HIP, std::string("fe_train sync: ") + hipGetErrorString(e));
  if (c->timing && !use_graph) harvest_events(c);
  return FE_OK;
}

extern "C" fe_status fe_synchronize(fe_ctx *c) {
  CTX_GUARD(c);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, FE_ERR_HIP, std::string("sync: ") + hipGetErrorString(e));
  return FE_OK;
}

extern "C" fe_status fe_loss(fe_ctx *c, uint64_t eval_seed, int eval_ns, double *attr,
                             double *rep) {
  CTX_GUARD(c);
  Job &j = c->job;
  if (j.B <= 0) {
    // the loss walks the work items; any batch size gives the same entries
    fe_status s = from_job(c, job_configure(j, (int)std::min<int64_t>(j.n, 16384), j.lr, j.bound));
    if (s != FE_OK) return s;
  }
  fe_status s = from_job(c, job_prepare(j, c->stream));
  if (s != FE_OK) return s;
  return from_job(c, job_loss(j, eval_seed, eval_ns, attr, rep, c->stream));
}

extern "C" fe_status fe_epoch_units(fe_ctx *c, int ns, int64_t *edges, int64_t *negs0,
                                    int64_t *negs_later) {
  CTX_GUARD(c);
  if (ns < 0) return fail(c, FE_ERR_INVALID_ARG, "fe_epoch_units: ns < 0");
  fe_status s = from_job(c, job_prepare(c->job, c->stream));
  if (s != FE_OK) return s;
  const Job &j = c->job;
  int64_t n0 = 0, n1 = 0;
  for (int b = 0; b < j.plan.batches; ++b) {
    int r0, r1;
    rep_rows(j.plan, j.quirks, b, &r0, &r1);
    n0 += (int64_t)r0 * ns;
    n1 += (int64_t)r1 * ns;
  }
  if (edges) *edges = j.plan.edges_eff_total;
  if (negs0) *negs0 = n0;
  if (negs_later) *negs_later = n1;
  return FE_OK;
}

extern "C" fe_status fe_set_timing(fe_ctx *c, int enable) {
  CTX_GUARD(c);
  hipStreamSynchronize(c->stream);
  if (!c->graph_timed) harvest_events(c);
  c->timing = enable != 0;
  c->t_force = c->t_fixup = c->t_apply = 0;
  c->n_timed = 0;
  c->events_used = 0;
  return FE_OK;
}

extern "C" fe_status fe_kernel_times(fe_ctx *c, double *force_ms, double *fixup_ms,
                                     double *apply_ms, int64_t *launches) {
  CTX_GUARD(c);
  hipStreamSynchronize(c->stream);
  if (!c->graph_valid || !c->graph_timed) harvest_events(c);
  const double n = c->n_timed > 0 ? (double)c->n_timed : 1.0;
  if (force_ms) *force_ms = c->t_force / n;
  if (fixup_ms) *fixup_ms = c->t_fixup / n;
  if (apply_ms) *apply_ms = c->t_apply / n;
  if (launches) *launches = c->n_timed;
  return FE_OK;
}
