// HIP kernels of the force-model embedding step for CDNA4 (gfx950).
//
// Semantics (one minibatch, cpp/algo/algo.hpp:165-192):
//   for every row r of the batch and every CSR neighbour c < c_limit
//     v = x_r - x_c,  s = |v|^2,  d1 = -2/(1+s)                     (:559-573)
//     delta_r += lr * clamp(v*d1, +-bound)                          (:575-579)
//   for the first `rep` rows and every negative n of list epoch+batch
//     v = x_r - x_n,  s = |v|^2,  d1 = 2/((s+1e-6)(1+s))             (:633-641)
//     delta_r += lr * clamp(v*d1, +-bound)                          (:642-645)
//   x_r += delta_r after all reads of the batch                      (:651-666)
//
// Mapping. A lane owns 8 consecutive components of a row, so a row of D
// floats is G = D/8 lanes (16 at D=128: one 512-byte row is 16 lanes x two
// 16-byte loads, or one 16-byte load of BF16) and a wave holds NG = 64/G rows
// at once. |v|^2 is reduced inside the G-lane group with DPP (quad_perm,
// row_half_mirror, row_mirror), which leaves the identical sum in every lane
// of the group. A wave takes one work item (a row, or a bounded chunk of a
// heavy row's edges); its groups take the item's edges round-robin, U rows in
// flight per group, and the group accumulators are combined at the end with
// a fixed xor tree. Heavy rows are split into chunks whose partial deltas
// are summed in index order by whichever chunk finishes last (agent-scope
// release/acquire around a per-row counter), so every result is
// deterministic and independent of the launch shape. The batch's negative
// rows are staged once per block in LDS. Batch j's launch also copies batch
// j-1's new rows from staging into X, reading them from staging meanwhile,
// so one launch per minibatch suffices. No MFMA: the step is a sparse
// gather, not a dense contraction.
#include <hip/hip_runtime.h>
#include "fe_device.hpp"

namespace fe {

template <int CTRL>
__device__ __forceinline__ float dppf(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Sum over the G-lane group (G a power of two, wave-uniform).
__device__ __forceinline__ float group_sum(float v, int G) {
  if (G >= 2) v += dppf<0xB1>(v);   // quad_perm [1,0,3,2]
  if (G >= 4) v += dppf<0x4E>(v);   // quad_perm [2,3,0,1]
  if (G >= 8) v += dppf<0x141>(v);  // row_half_mirror
  if (G >= 16) v += dppf<0x140>(v); // row_mirror
  if (G >= 32) v += __shfl_xor(v, 16);
  if (G >= 64) v += __shfl_xor(v, 32);
  return v;
}

// Device-memory accesses go through address-space-1 pointers so they are
// emitted as global_* instructions (counted on vmcnt only) rather than flat_*
// ones, whose waits also drain LDS and scalar traffic.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const f4v gcf4;
typedef __attribute__((address_space(1))) f4v gf4;
typedef __attribute__((address_space(1))) const u4v gcu4;
typedef __attribute__((address_space(1))) const int gcint;

__device__ __forceinline__ int gld_int(const int32_t *p) { return *(gcint *)p; }

__device__ __forceinline__ u4v gld_u4(const int32_t *p) { return *(gcu4 *)p; }

__device__ __forceinline__ void load8_f32(const float *p, float (&o)[8]) {
  const f4v a = ((gcf4 *)p)[0];
  const f4v b = ((gcf4 *)p)[1];
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
  o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
}

__device__ __forceinline__ void load8_bf16(const uint16_t *p, float (&o)[8]) {
  const u4v w = *(gcu4 *)p;
  o[0] = __uint_as_float(w.x << 16); o[1] = __uint_as_float(w.x & 0xFFFF0000u);
  o[2] = __uint_as_float(w.y << 16); o[3] = __uint_as_float(w.y & 0xFFFF0000u);
  o[4] = __uint_as_float(w.z << 16); o[5] = __uint_as_float(w.z & 0xFFFF0000u);
  o[6] = __uint_as_float(w.w << 16); o[7] = __uint_as_float(w.w & 0xFFFF0000u);
}

// Global memory only (never an LDS pointer: see store8_lds).
__device__ __forceinline__ void store8_f32(float *p, const float (&o)[8]) {
  f4v a = {o[0], o[1], o[2], o[3]};
  f4v b = {o[4], o[5], o[6], o[7]};
  ((gf4 *)p)[0] = a;
  ((gf4 *)p)[1] = b;
}

typedef __attribute__((address_space(3))) f4v lf4;

__device__ __forceinline__ void store8_lds(float *p, const float (&o)[8]) {
  f4v a = {o[0], o[1], o[2], o[3]};
  f4v b = {o[4], o[5], o[6], o[7]};
  ((lf4 *)p)[0] = a;
  ((lf4 *)p)[1] = b;
}

// Partials of heavy rows cross CUs (and XCDs, whose L2s are not coherent with
// each other) inside one launch. They are written and read with agent-scope
// relaxed atomics, which the hardware performs at the device-coherent level
// (sc1), so no L2 writeback or invalidation is needed around them.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) int gint;

__device__ __forceinline__ void store8_coherent(float *p, const float (&o)[8]) {
  gfloat *g = (gfloat *)p;
#pragma unroll
  for (int k = 0; k < 8; ++k) __hip_atomic_store(g + k, o[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void load8_coherent(const float *p, float (&o)[8]) {
  gfloat *g = (gfloat *)p;
#pragma unroll
  for (int k = 0; k < 8; ++k) o[k] = __hip_atomic_load(g + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// s_waitcnt vmcnt(0): this wave's stores have completed.
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// FP32 -> BF16, round to nearest even; NaN stays a quiet NaN.
__device__ __forceinline__ uint32_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return ((u >> 16) | 0x40u) & 0xFFFFu;
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

// scale() of cpp/algo/algo.hpp:65-72 (a NaN passes through unchanged).
__device__ __forceinline__ float clampf(float v, float m) {
  return v > m ? m : (v < -m ? -m : v);
}

__device__ __forceinline__ int find_step(const int *__restrict__ prefix,
                                         int n_steps, int b) {
  int lo = 0, hi = n_steps - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Row c as 8 FP32 components [8q, 8q+8). Rows of the previous batch whose
// new values still sit in Sprev are read from there; with BF16 gathers they
// are rounded exactly as the shadow refresh will round them.
// MODE 0: FP32 gathers, 1: BF16 gathers, 2: per step at run time (a launch
// mixing jobs of both kinds).
template <int MODE> __device__ __forceinline__ bool use_bf16(const StepDesc &sd) {
  return MODE == 2 ? sd.bf16 != 0 : MODE == 1;
}

template <int MODE>
__device__ __forceinline__ void load_row(const StepDesc &sd, int c, int q, float (&o)[8]) {
  const int D = sd.D;
  const unsigned off = (unsigned)(c - sd.prev_begin);
  if (off < (unsigned)sd.prev_rows) {
    load8_f32(sd.Sprev + (size_t)off * D + 8 * q, o);
    if (use_bf16<MODE>(sd)) {
#pragma unroll
      for (int k = 0; k < 8; ++k) o[k] = __uint_as_float(bf16_bits(o[k]) << 16);
    }
  } else if (use_bf16<MODE>(sd)) {
    load8_bf16(sd.Xs + (size_t)c * D + 8 * q, o);
  } else {
    load8_f32(sd.X + (size_t)c * D + 8 * q, o);
  }
}

// A gathered row slice held while its load is in flight: 8 FP32 registers,
// or with BF16 gathers the 8 packed BF16 values (4 registers), unpacked only
// when used -- twice the rows in flight for the same registers.
template <int MODE> struct RowBuf {
  float f[8];
};
template <> struct RowBuf<1> {
  u4v w;
};

template <int MODE>
__device__ __forceinline__ void load_buf(const StepDesc &sd, int c, int q, RowBuf<MODE> &b) {
  if constexpr (MODE == 1) {
    const unsigned off = (unsigned)(c - sd.prev_begin);
    if (off < (unsigned)sd.prev_rows) {
      float o[8];
      load8_f32(sd.Sprev + (size_t)off * sd.D + 8 * q, o);
      b.w.x = bf16_bits(o[0]) | (bf16_bits(o[1]) << 16);
      b.w.y = bf16_bits(o[2]) | (bf16_bits(o[3]) << 16);
      b.w.z = bf16_bits(o[4]) | (bf16_bits(o[5]) << 16);
      b.w.w = bf16_bits(o[6]) | (bf16_bits(o[7]) << 16);
    } else {
      b.w = *(gcu4 *)(sd.Xs + (size_t)c * sd.D + 8 * q);
    }
  } else {
    load_row<MODE>(sd, c, q, b.f);
  }
}

template <int MODE> __device__ __forceinline__ void zero_buf(RowBuf<MODE> &b) {
  if constexpr (MODE == 1) {
    b.w = u4v{0u, 0u, 0u, 0u};
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) b.f[k] = 0.f;
  }
}

template <int MODE>
__device__ __forceinline__ void unpack_buf(const RowBuf<MODE> &b, float (&o)[8]) {
  if constexpr (MODE == 1) {
    o[0] = __uint_as_float(b.w.x << 16); o[1] = __uint_as_float(b.w.x & 0xFFFF0000u);
    o[2] = __uint_as_float(b.w.y << 16); o[3] = __uint_as_float(b.w.y & 0xFFFF0000u);
    o[4] = __uint_as_float(b.w.z << 16); o[5] = __uint_as_float(b.w.z & 0xFFFF0000u);
    o[6] = __uint_as_float(b.w.w << 16); o[7] = __uint_as_float(b.w.w & 0xFFFF0000u);
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = b.f[k];
  }
}

// Stage the `ns` negative rows of list `list` as FP32 in LDS.
template <int MODE>
__device__ __forceinline__ void stage_negatives(const StepDesc &sd, int list, float *lds) {
  const int G = sd.G;
  const int total8 = sd.ns * G; // 8-component pieces
  const int32_t *nl = sd.negs + (size_t)list * sd.ns;
  for (int p = threadIdx.x; p < total8; p += blockDim.x) {
    const int k = p / G;
    const int q = p - k * G;
    float v[8];
    load_row<MODE>(sd, gld_int(nl + k), q, v);
    store8_lds(lds + (size_t)k * sd.D + 8 * q, v);
  }
}

// Repulsive forces of the staged negatives on x_r, spread over the groups.
__device__ __forceinline__ void repulse(const StepDesc &sd, const float *lds,
                                        const float (&xr)[8], float (&acc)[8],
                                        int G, int NG, int g, int q) {
  const int ns = sd.ns;
  const int rounds = (ns + NG - 1) / NG;
  const float lr = sd.lr, bound = sd.bound;
  for (int u = 0; u < rounds; ++u) {
    const int k = u * NG + g;
    const bool ok = k < ns;
    const float *pn = lds + (ok ? k : 0) * sd.D + 8 * q;
    const float4 a = *reinterpret_cast<const float4 *>(pn);
    const float4 b = *reinterpret_cast<const float4 *>(pn + 4);
    const float xn[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    float v[8];
    float p = 0.f;
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) {
      v[k2] = ok ? xr[k2] - xn[k2] : 0.f;
      p = fmaf(v[k2], v[k2], p);
    }
    const float s = group_sum(p, G);
    const float d1 = 2.0f / ((s + 1e-6f) * (1.0f + s));
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) acc[k2] = fmaf(lr, clampf(v[k2] * d1, bound), acc[k2]);
  }
}

__device__ __forceinline__ void combine_groups(float (&acc)[8], int G) {
  for (int m = G; m < 64; m <<= 1) {
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] += __shfl_xor(acc[k], m);
  }
}

// Sum partials p0..p0+k-1: group g takes g, g+NG, ... in order, then the
// fixed xor tree of combine_groups.
template <int U>
__device__ __forceinline__ void sum_partials(const StepDesc &sd, int p0, int k, int NG, int g,
                                             int q, float (&acc)[8]) {
  const int D = sd.D;
  const int rounds = (k + NG - 1) / NG;
  for (int u0 = 0; u0 < rounds; u0 += U) {
    float pv[U][8];
    bool ok[U];
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      const int idx = (u0 + uu) * NG + g;
      ok[uu] = idx < k;
      if (ok[uu]) load8_coherent(sd.P + (size_t)(p0 + idx) * D + 8 * q, pv[uu]);
    }
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      if (ok[uu]) {
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) acc[k2] += pv[uu][k2];
      }
    }
  }
}

// Attractive forces of edges [e0, e0+cnt) of the row on the per-group
// accumulators: group g takes edges g, g+NG, ... in order, U rows in flight.
template <int U, int MODE>
__device__ __forceinline__ void gather_edges(const StepDesc &sd, const float (&xd)[8], int e0,
                                             int cnt, int NG, int g, int q, int lane,
                                             float (&acc)[8]) {
  const int G = sd.G;
  const float lr = sd.lr, bound = sd.bound;
  const int c_limit = sd.c_limit;
  for (int eb = 0; eb < cnt; eb += 64) {
    const int m = min(64, cnt - eb);
    const int cl = lane < m ? gld_int(sd.col + e0 + eb + lane) : 0;
    const int rounds = (m + NG - 1) / NG;
    for (int u0 = 0; u0 < rounds; u0 += U) {
      RowBuf<MODE> xb[U];
      bool ok[U];
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const int k = (u0 + uu) * NG + g;
        const int c = __shfl(cl, k & 63);
        ok[uu] = (k < m) && (c < c_limit);
        if (ok[uu]) load_buf<MODE>(sd, c, q, xb[uu]);
        else zero_buf<MODE>(xb[uu]);
      }
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        float xc[8], v[8];
        unpack_buf<MODE>(xb[uu], xc);
        float p = 0.f;
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) {
          v[k2] = ok[uu] ? xd[k2] - xc[k2] : 0.f;
          p = fmaf(v[k2], v[k2], p);
        }
        const float s = group_sum(p, G);
        const float d1 = -2.0f / (1.0f + s);
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) acc[k2] = fmaf(lr, clampf(v[k2] * d1, bound), acc[k2]);
      }
    }
  }
}

// Row load for the distances: the master row, rounded as the shadow holds it
// when gathers read BF16 (no extra traffic), so a row meeting itself as a
// negative or through a self-loop feels exactly zero force, as in the
// reference; the update still goes to the FP32 master.
template <int MODE>
__device__ __forceinline__ void load_own(const StepDesc &sd, int r, int q, float (&xr)[8],
                                         float (&xd)[8]) {
  load8_f32(sd.X + (size_t)r * sd.D + 8 * q, xr);
#pragma unroll
  for (int k = 0; k < 8; ++k)
    xd[k] = use_bf16<MODE>(sd) ? __uint_as_float(bf16_bits(xr[k]) << 16) : xr[k];
}

// Repulsion (first `rep` rows), group combine, new row (or delta) to S.
__device__ __forceinline__ void finish_row(const StepDesc &sd, const float *lds_neg, bool repel,
                                           int t, const float (&xr)[8], const float (&xd)[8],
                                           float (&acc)[8], int G, int NG, int g, int q) {
  if (repel) repulse(sd, lds_neg, xd, acc, G, NG, g, q);
  combine_groups(acc, G);
  if (g == 0) {
    float o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = sd.delta_mode ? acc[k] : xr[k] + acc[k];
    store8_f32(sd.S + (size_t)t * sd.D + 8 * q, o);
  }
}

// Light rows, one per lane group: group g takes entry pack*NG + g of the
// batch's light rows (sorted by degree), gathers its edges in CSR order (U
// in flight), repels against every staged negative and writes its row. The
// loops run to the pack's largest degree, so control flow stays uniform.
template <int U, int MODE>
__device__ __forceinline__ void light_pack(const StepDesc &sd, const int32_t *lrows, int n_light,
                                           int pack, const float *lds_neg, bool has_neg, int rep,
                                           int NG, int g, int q, int lane) {
  const int G = sd.G;
  const int idx = pack * NG + g;
  const bool valid = idx < n_light;
  const u4v ia = gld_u4(lrows + (size_t)kItemInts * (valid ? idx : pack * NG));
  const int r = (int)ia.x, e0 = (int)ia.y, deg = valid ? (int)ia.z : 0;
  const int kmax = __builtin_amdgcn_readfirstlane((int)gld_u4(lrows + (size_t)kItemInts * pack * NG + 4).x);
  const float lr = sd.lr, bound = sd.bound;
  const int c_limit = sd.c_limit;
  float xr[8], xd[8], acc[8];
  load_own<MODE>(sd, r, q, xr, xd);
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = 0.f;
  for (int kb = 0; kb < kmax; kb += G) {
    const int cl = (kb + q < deg) ? gld_int(sd.col + e0 + kb + q) : 0;
    const int kend = min(kmax, kb + G);
    for (int k0 = kb; k0 < kend; k0 += U) {
      RowBuf<MODE> xb[U];
      bool ok[U];
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const int k = k0 + uu;
        const int c = __shfl(cl, g * G + ((k - kb) & (G - 1)));
        ok[uu] = (k < kend) && (k < deg) && (c < c_limit);
        if (ok[uu]) load_buf<MODE>(sd, c, q, xb[uu]);
        else zero_buf<MODE>(xb[uu]);
      }
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        float xc[8], v[8];
        unpack_buf<MODE>(xb[uu], xc);
        float p = 0.f;
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) {
          v[k2] = ok[uu] ? xd[k2] - xc[k2] : 0.f;
          p = fmaf(v[k2], v[k2], p);
        }
        const float s = group_sum(p, G);
        const float d1 = -2.0f / (1.0f + s);
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) acc[k2] = fmaf(lr, clampf(v[k2] * d1, bound), acc[k2]);
      }
    }
  }
  const int t = r - sd.row_begin;
  if (has_neg) {
    const bool repel = valid && t < rep;
    for (int n = 0; n < sd.ns; ++n) {
      const float *pn = lds_neg + (size_t)n * sd.D + 8 * q;
      const float4 a = *reinterpret_cast<const float4 *>(pn);
      const float4 b = *reinterpret_cast<const float4 *>(pn + 4);
      const float xn[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
      float v[8];
      float p = 0.f;
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) {
        v[k2] = repel ? xd[k2] - xn[k2] : 0.f;
        p = fmaf(v[k2], v[k2], p);
      }
      const float s = group_sum(p, G);
      const float d1 = 2.0f / ((s + 1e-6f) * (1.0f + s));
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) acc[k2] = fmaf(lr, clampf(v[k2] * d1, bound), acc[k2]);
    }
  }
  if (valid) {
    float o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = sd.delta_mode ? acc[k] : xr[k] + acc[k];
    store8_f32(sd.S + (size_t)t * sd.D + 8 * q, o);
  }
}

// One minibatch of one or more jobs. Blocks [0, n_block_items) of a step
// each take one block item: a row of degree <= kSub*chunk, or one chunk of a
// heavier row, cut into kSub fixed sub-ranges that the block's waves gather
// and then add in sub-range order through LDS. A chunk publishes its
// partial (device-coherent stores) and bumps the row's counter; the chunk
// that counts last adds all partials of the row in index order. The
// remaining blocks' waves walk the light rows with a grid stride, one row
// per wave at a time. Every block then copies its share of the previous
// batch's rows from staging into X.
template <int U, int MODE>
__global__ __launch_bounds__(512) void force_kernel(
    const StepDesc *__restrict__ steps, const int *__restrict__ blk_prefix,
    int n_steps, int ipw) {
  extern __shared__ float lds[];
  const int b = blockIdx.x;
  const int si = find_step(blk_prefix, n_steps, b);
  const StepDesc &sd = steps[si];
  const int lb = b - blk_prefix[si];
  const int nb = blk_prefix[si + 1] - blk_prefix[si];
  const int epoch = *sd.epoch;
  const int rep = epoch == 0 ? sd.rep0 : sd.rep1;
  const int list = epoch + sd.batch;
  const bool has_neg = sd.ns > 0 && list < sd.n_lists;
  if (has_neg) stage_negatives<MODE>(sd, list, lds);
  __syncthreads();

  const int D = sd.D, G = sd.G;
  const int NG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int g = lane / G;
  const int q = lane & (G - 1);

  if (lb < sd.n_block_items) {
    const int32_t *item = sd.items + (size_t)kItemInts * lb;
    const u4v ia = gld_u4(item);
    const int r = (int)ia.x, e0 = (int)ia.y, cnt = (int)ia.z, pslot = (int)ia.w;
    const int t = r - sd.row_begin;
    float xr[8], xd[8], acc[8];
    load_own<MODE>(sd, r, q, xr, xd);
    float *lds_sub = lds + (has_neg ? sd.ns * D : 0);
    for (int s = wave; s < kSub; s += wpb) {
      const int lo = (int)((int64_t)cnt * s / kSub), hi = (int)((int64_t)cnt * (s + 1) / kSub);
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = 0.f;
      gather_edges<U, MODE>(sd, xd, e0 + lo, hi - lo, NG, g, q, lane, acc);
      combine_groups(acc, G);
      if (g == 0) store8_lds(lds_sub + (size_t)s * D + 8 * q, acc);
    }
    __syncthreads();
    if (wave == 0) {
      // group 0 holds the sum of the sub-ranges, in sub-range order
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = 0.f;
      if (g == 0) {
        for (int s = 0; s < kSub; ++s) {
          const float *p = lds_sub + (size_t)s * D + 8 * q;
#pragma unroll
          for (int k = 0; k < 8; ++k) acc[k] += p[k];
        }
      }
      bool finish = pslot < 0;
      if (!finish) {
        const u4v ib = gld_u4(item + 4);
        const int sid = (int)ib.x, p0 = (int)ib.y, kparts = (int)ib.z;
        if (g == 0) store8_coherent(sd.P + (size_t)pslot * D + 8 * q, acc);
        wait_stores();
        int old = 0;
        if (lane == 0)
          old = __hip_atomic_fetch_add((gint *)(sd.counters + sid), 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        old = __shfl(old, 0);
        if (old == kparts - 1) {
#pragma unroll
          for (int k = 0; k < 8; ++k) acc[k] = 0.f;
          sum_partials<U>(sd, p0, kparts, NG, g, q, acc);
          if (lane == 0)
            __hip_atomic_store((gint *)(sd.counters + sid), 0, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
          finish = true;
        }
      }
      if (finish) finish_row(sd, lds, has_neg && t < rep, t, xr, xd, acc, G, NG, g, q);
    }
  } else {
    // wave work: packs of light rows (one row per lane group), then the
    // medium rows (one row per wave), walked with a grid stride
    const int n_light = sd.n_light;
    const int n_wave_items = sd.n_items - sd.n_block_items - n_light;
    const int32_t *witems = sd.items + (size_t)kItemInts * sd.n_block_items;
    const int32_t *lrows = witems + (size_t)kItemInts * n_wave_items;
    const int n_packs = (n_light + NG - 1) / NG;
    const int W = (nb - sd.n_block_items) * wpb;
    int u = (lb - sd.n_block_items) * wpb + wave;
    for (; u < n_packs; u += W)
      light_pack<U, MODE>(sd, lrows, n_light, u, lds, has_neg, rep, NG, g, q, lane);
    for (int it = u - n_packs; it < n_wave_items; it += W) {
      const u4v ia = gld_u4(witems + (size_t)kItemInts * it);
      const int r = (int)ia.x, e0 = (int)ia.y, cnt = (int)ia.z;
      const int t = r - sd.row_begin;
      float xr[8], xd[8], acc[8];
      load_own<MODE>(sd, r, q, xr, xd);
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = 0.f;
      gather_edges<U, MODE>(sd, xd, e0, cnt, NG, g, q, lane, acc);
      finish_row(sd, lds, has_neg && t < rep, t, xr, xd, acc, G, NG, g, q);
    }
  }
  (void)ipw;

  // copy phase: the previous batch's new rows into X (and the shadow); no
  // thread of this launch reads those rows of X (load_row redirects them)
  if (sd.prev_rows > 0) {
    const int64_t total4 = (int64_t)sd.prev_rows * D / 4;
    const int64_t per = (total4 + nb - 1) / nb;
    const int64_t lo = (int64_t)lb * per;
    const int64_t hi = min(total4, lo + per);
    const int64_t base4 = (int64_t)sd.prev_begin * D / 4;
    const float4 *S4 = reinterpret_cast<const float4 *>(sd.Sprev);
    float4 *X4 = reinterpret_cast<float4 *>(sd.Xw);
    for (int64_t idx = lo + threadIdx.x; idx < hi; idx += blockDim.x) {
      const float4 v = S4[idx];
      X4[base4 + idx] = v;
      if (use_bf16<MODE>(sd)) {
        uint2 w;
        w.x = bf16_bits(v.x) | (bf16_bits(v.y) << 16);
        w.y = bf16_bits(v.z) | (bf16_bits(v.w) << 16);
        reinterpret_cast<uint2 *>(sd.Xsw)[base4 + idx] = w;
      }
    }
  }
}

// X[batch] = S (cpp/algo/algo.hpp:651-666 with the add already folded into
// S) and the BF16 shadow refresh of the same rows.
__global__ __launch_bounds__(256) void apply_kernel(
    const StepDesc *__restrict__ steps, const int *__restrict__ blk_prefix,
    int n_steps) {
  const int b = blockIdx.x;
  const int si = find_step(blk_prefix, n_steps, b);
  const StepDesc &sd = steps[si];
  const int lb = b - blk_prefix[si];
  const int nb = blk_prefix[si + 1] - blk_prefix[si];
  const int64_t total4 = (int64_t)sd.rows * sd.D / 4;
  const int64_t base4 = (int64_t)sd.row_begin * sd.D / 4;
  const float4 *S4 = reinterpret_cast<const float4 *>(sd.S);
  float4 *X4 = reinterpret_cast<float4 *>(sd.Xw);
  for (int64_t idx = (int64_t)lb * blockDim.x + threadIdx.x; idx < total4;
       idx += (int64_t)nb * blockDim.x) {
    const float4 v = S4[idx];
    X4[base4 + idx] = v;
    if (sd.bf16) {
      uint2 w;
      w.x = bf16_bits(v.x) | (bf16_bits(v.y) << 16);
      w.y = bf16_bits(v.z) | (bf16_bits(v.w) << 16);
      reinterpret_cast<uint2 *>(sd.Xsw)[base4 + idx] = w;
    }
  }
}

__global__ void set_int_kernel(int *p, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// Advances epoch counters (one per job) at the start of an epoch.
__global__ void tick_kernel(int *const *ctrs, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) *ctrs[i] = *ctrs[i] + 1;
}

__global__ void to_bf16_kernel(const float *__restrict__ x, uint16_t *__restrict__ xs,
                               int64_t count2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count2;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = reinterpret_cast<const float2 *>(x)[i];
    reinterpret_cast<uint32_t *>(xs)[i] = bf16_bits(v.x) | (bf16_bits(v.y) << 16);
  }
}

// ---- FP64 loss -----------------------------------------------------------

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ int64_t eval_negative(uint64_t seed, int64_t r, int k,
                                                 int64_t n) {
  const uint64_t x = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)r * 1048576ull + (uint64_t)k + 1ull);
  return (int64_t)(splitmix64(x) % (uint64_t)n);
}

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m);
  return v;
}

__device__ __forceinline__ void block_sum_store(double v, double *out) {
  __shared__ double part[16];
  v = wave_sum_d(v);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) part[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += part[w];
    out[blockIdx.x] = s;
  }
}

// Attractive part: sum of log(1+s) over the CSR entries taken by the step.
__global__ __launch_bounds__(256) void loss_edges_kernel(LossDesc ld, double *block_out,
                                                         int ipw) {
  const int D = ld.D, G = ld.G, NG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int g = lane / G, q = lane & (G - 1);
  double tot = 0.0;
  for (int i = 0; i < ipw; ++i) {
    const int it = (blockIdx.x * ipw + i) * wpb + wave;
    if (it >= ld.n_items) break;
    const int r = ld.items[kItemInts * it + 0];
    const int e0 = ld.items[kItemInts * it + 1];
    const int cnt = ld.items[kItemInts * it + 2];
    float xr[8];
    load8_f32(ld.X + (size_t)r * D + 8 * q, xr);
    for (int eb = 0; eb < cnt; eb += 64) {
      const int m = min(64, cnt - eb);
      const int cl = lane < m ? ld.col[e0 + eb + lane] : 0;
      const int rounds = (m + NG - 1) / NG;
      for (int u = 0; u < rounds; ++u) {
        const int k = u * NG + g;
        const int c = __shfl(cl, k & 63);
        const bool ok = (k < m) && (c < ld.c_limit);
        float xc[8];
        load8_f32(ld.X + (size_t)(ok ? c : r) * D + 8 * q, xc);
        float p = 0.f;
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) {
          const float v = xr[k2] - xc[k2];
          p = fmaf(v, v, p);
        }
        const float s = group_sum(p, G);
        if (ok && q == 0) tot += log1p((double)s);
      }
    }
  }
  block_sum_store(tot, block_out);
}

// Repulsive part over the evaluation pairs (r, eval_negative(seed, r, k)).
__global__ __launch_bounds__(256) void loss_negs_kernel(LossDesc ld, double *block_out,
                                                        int pairs_per_group) {
  const int D = ld.D, G = ld.G, NG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int g = lane / G, q = lane & (G - 1);
  const int64_t group_id = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * NG + g;
  const int64_t total = ld.n * (int64_t)ld.eval_ns;
  double tot = 0.0;
  const double eps = 1e-6;
  for (int i = 0; i < pairs_per_group; ++i) {
    const int64_t p = group_id * pairs_per_group + i;
    const bool ok = p < total;
    const int64_t r = ok ? p / ld.eval_ns : 0;
    const int k = ok ? (int)(p - r * ld.eval_ns) : 0;
    const int64_t m = ok ? eval_negative(ld.eval_seed, r, k, ld.n) : 0;
    float xr[8], xm[8];
    load8_f32(ld.X + (size_t)r * D + 8 * q, xr);
    load8_f32(ld.X + (size_t)m * D + 8 * q, xm);
    float s2 = 0.f;
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) {
      const float v = xr[k2] - xm[k2];
      s2 = fmaf(v, v, s2);
    }
    const float s = group_sum(s2, G);
    if (ok && q == 0) {
      const double sd = (double)s;
      tot += log((1.0 + sd) / (sd + eps)) / (1.0 - eps);
    }
  }
  block_sum_store(tot, block_out);
}

// Fixed-order final reduction (one block).
__global__ __launch_bounds__(1024) void reduce_kernel(const double *__restrict__ in, int n,
                                                      double *out) {
  __shared__ double part[1024];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += in[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

// ---- launch wrappers -------------------------------------------------------

// lds_bytes: staged negative rows and sub-range sums (force_lds_bytes);
// mode: 0 all steps gather FP32, 1 all gather BF16, 2 mixed.
template <int U>
static void launch_force_u(dim3 grid, dim3 block, size_t lds_bytes, hipStream_t st, int mode,
                           const StepDesc *steps, const int *blk_prefix, int n_steps, int ipw) {
  if (mode == 0) hipLaunchKernelGGL((force_kernel<U, 0>), grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw);
  else if (mode == 1) hipLaunchKernelGGL((force_kernel<U, 1>), grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw);
  else hipLaunchKernelGGL((force_kernel<U, 2>), grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw);
}

hipError_t launch_force(const StepDesc *steps, const int *blk_prefix, int n_steps,
                        int n_blocks, int wpb, int ipw, int unroll, size_t lds_bytes, int mode,
                        hipStream_t st) {
  dim3 grid(n_blocks), block(64 * wpb);
  switch (unroll) {
  case 1: launch_force_u<1>(grid, block, lds_bytes, st, mode, steps, blk_prefix, n_steps, ipw); break;
  case 2: launch_force_u<2>(grid, block, lds_bytes, st, mode, steps, blk_prefix, n_steps, ipw); break;
  case 8: launch_force_u<8>(grid, block, lds_bytes, st, mode, steps, blk_prefix, n_steps, ipw); break;
  default: launch_force_u<4>(grid, block, lds_bytes, st, mode, steps, blk_prefix, n_steps, ipw); break;
  }
  return hipGetLastError();
}

hipError_t launch_tick(int *const *ctrs, int n, hipStream_t st) {
  hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(256), 0, st, ctrs, n);
  return hipGetLastError();
}

hipError_t launch_apply(const StepDesc *steps, const int *blk_prefix, int n_steps,
                        int n_blocks, hipStream_t st) {
  hipLaunchKernelGGL(apply_kernel, dim3(n_blocks), dim3(256), 0, st, steps, blk_prefix, n_steps);
  return hipGetLastError();
}

hipError_t launch_set_int(int *p, int v, hipStream_t st) {
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(64), 0, st, p, v);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float *x, uint16_t *xs, int64_t count, hipStream_t st) {
  const int64_t count2 = count / 2;
  int64_t blocks = (count2 + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x, xs, count2);
  return hipGetLastError();
}

hipError_t launch_loss_edges(const LossDesc &ld, double *block_out, int n_blocks, int ipw,
                             hipStream_t st) {
  hipLaunchKernelGGL(loss_edges_kernel, dim3(n_blocks), dim3(256), 0, st, ld, block_out, ipw);
  return hipGetLastError();
}

hipError_t launch_loss_negs(const LossDesc &ld, double *block_out, int n_blocks,
                            int pairs_per_group, hipStream_t st) {
  hipLaunchKernelGGL(loss_negs_kernel, dim3(n_blocks), dim3(256), 0, st, ld, block_out,
                     pairs_per_group);
  return hipGetLastError();
}

hipError_t launch_reduce(const double *in, int n, double *out, hipStream_t st) {
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, st, in, n, out);
  return hipGetLastError();
}

} // namespace fe
