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
// are summed in index order by the fixup kernel, so every result is
// deterministic. The batch's negative rows are staged once per block in LDS.
// No MFMA: the step is a sparse gather, not a dense contraction.
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

__device__ __forceinline__ void load8_f32(const float *p, float (&o)[8]) {
  const float4 *q = reinterpret_cast<const float4 *>(p);
  float4 a = q[0];
  float4 b = q[1];
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
  o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
}

__device__ __forceinline__ void load8_bf16(const uint16_t *p, float (&o)[8]) {
  uint4 w = *reinterpret_cast<const uint4 *>(p);
  o[0] = __uint_as_float(w.x << 16); o[1] = __uint_as_float(w.x & 0xFFFF0000u);
  o[2] = __uint_as_float(w.y << 16); o[3] = __uint_as_float(w.y & 0xFFFF0000u);
  o[4] = __uint_as_float(w.z << 16); o[5] = __uint_as_float(w.z & 0xFFFF0000u);
  o[6] = __uint_as_float(w.w << 16); o[7] = __uint_as_float(w.w & 0xFFFF0000u);
}

__device__ __forceinline__ void store8_f32(float *p, const float (&o)[8]) {
  float4 *q = reinterpret_cast<float4 *>(p);
  q[0] = make_float4(o[0], o[1], o[2], o[3]);
  q[1] = make_float4(o[4], o[5], o[6], o[7]);
}

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

// Stage the `ns` negative rows of list `list` as FP32 in LDS.
__device__ __forceinline__ void stage_negatives(const StepDesc &sd, int list,
                                                float *lds) {
  const int D = sd.D;
  const int total = sd.ns * D;
  const int32_t *nl = sd.negs + (size_t)list * sd.ns;
  for (int idx = threadIdx.x * 4; idx < total; idx += blockDim.x * 4) {
    int k = idx / D;
    int comp = idx - k * D;
    size_t n = (size_t)nl[k];
    float4 v;
    if (sd.bf16) {
      uint2 w = *reinterpret_cast<const uint2 *>(sd.Xs + n * D + comp);
      v = make_float4(__uint_as_float(w.x << 16), __uint_as_float(w.x & 0xFFFF0000u),
                      __uint_as_float(w.y << 16), __uint_as_float(w.y & 0xFFFF0000u));
    } else {
      v = *reinterpret_cast<const float4 *>(sd.X + n * D + comp);
    }
    *reinterpret_cast<float4 *>(lds + idx) = v;
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
      p += v[k2] * v[k2];
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

template <int U>
__global__ __launch_bounds__(512) void force_kernel(
    const StepDesc *__restrict__ steps, const int *__restrict__ blk_prefix,
    int n_steps, int ipw) {
  extern __shared__ float lds_neg[];
  const int b = blockIdx.x;
  const int si = find_step(blk_prefix, n_steps, b);
  const StepDesc &sd = steps[si];
  const int lb = b - blk_prefix[si];
  const int epoch = *sd.epoch;
  const int rep = epoch == 0 ? sd.rep0 : sd.rep1;
  const int list = epoch + sd.batch;
  const bool has_neg = sd.ns > 0 && list < sd.n_lists;
  if (has_neg) stage_negatives(sd, list, lds_neg);
  __syncthreads();

  const int D = sd.D, G = sd.G;
  const int NG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int g = lane / G;
  const int q = lane & (G - 1);
  const float lr = sd.lr, bound = sd.bound;
  const int c_limit = sd.c_limit;

  for (int i = 0; i < ipw; ++i) {
    const int it = (lb * ipw + i) * wpb + wave;
    if (it >= sd.n_items) break;
    const int r = sd.items[4 * it + 0];
    const int e0 = sd.items[4 * it + 1];
    const int cnt = sd.items[4 * it + 2];
    const int pslot = sd.items[4 * it + 3];
    const int t = r - sd.row_begin;

    float xr[8];
    load8_f32(sd.X + (size_t)r * D + 8 * q, xr);
    float acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.f;

    for (int eb = 0; eb < cnt; eb += 64) {
      const int m = min(64, cnt - eb);
      const int cl = lane < m ? sd.col[e0 + eb + lane] : 0;
      const int rounds = (m + NG - 1) / NG;
      for (int u0 = 0; u0 < rounds; u0 += U) {
        float xc[U][8];
        bool ok[U];
#pragma unroll
        for (int uu = 0; uu < U; ++uu) {
          const int k = (u0 + uu) * NG + g;
          const int c = __shfl(cl, k & 63);
          ok[uu] = (k < m) && (c < c_limit);
          if (ok[uu]) {
            if (sd.bf16) load8_bf16(sd.Xs + (size_t)c * D + 8 * q, xc[uu]);
            else load8_f32(sd.X + (size_t)c * D + 8 * q, xc[uu]);
          } else {
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2) xc[uu][k2] = 0.f;
          }
        }
#pragma unroll
        for (int uu = 0; uu < U; ++uu) {
          float v[8];
          float p = 0.f;
#pragma unroll
          for (int k2 = 0; k2 < 8; ++k2) {
            v[k2] = ok[uu] ? xr[k2] - xc[uu][k2] : 0.f;
            p += v[k2] * v[k2];
          }
          const float s = group_sum(p, G);
          const float d1 = -2.0f / (1.0f + s);
#pragma unroll
          for (int k2 = 0; k2 < 8; ++k2)
            acc[k2] = fmaf(lr, clampf(v[k2] * d1, bound), acc[k2]);
        }
      }
    }
    if (pslot < 0 && has_neg && t < rep) repulse(sd, lds_neg, xr, acc, G, NG, g, q);
    combine_groups(acc, G);
    if (g == 0) {
      if (pslot >= 0) {
        store8_f32(sd.P + (size_t)pslot * D + 8 * q, acc);
      } else {
        float o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) o[k] = sd.delta_mode ? acc[k] : xr[k] + acc[k];
        store8_f32(sd.S + (size_t)t * D + 8 * q, o);
      }
    }
  }
}

// Heavy rows: sum the chunk partials in index order, add the repulsion and
// write the new row.
template <int U>
__global__ __launch_bounds__(512) void fixup_kernel(
    const StepDesc *__restrict__ steps, const int *__restrict__ blk_prefix,
    int n_steps, int ipw) {
  extern __shared__ float lds_neg[];
  const int b = blockIdx.x;
  const int si = find_step(blk_prefix, n_steps, b);
  const StepDesc &sd = steps[si];
  const int lb = b - blk_prefix[si];
  const int epoch = *sd.epoch;
  const int rep = epoch == 0 ? sd.rep0 : sd.rep1;
  const int list = epoch + sd.batch;
  const bool has_neg = sd.ns > 0 && list < sd.n_lists;
  if (has_neg) stage_negatives(sd, list, lds_neg);
  __syncthreads();

  const int D = sd.D, G = sd.G;
  const int NG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int g = lane / G;
  const int q = lane & (G - 1);

  for (int i = 0; i < ipw; ++i) {
    const int it = (lb * ipw + i) * wpb + wave;
    if (it >= sd.n_splits) break;
    const int r = sd.splits[4 * it + 0];
    const int p0 = sd.splits[4 * it + 1];
    const int k = sd.splits[4 * it + 2];
    const int t = r - sd.row_begin;
    float xr[8];
    load8_f32(sd.X + (size_t)r * D + 8 * q, xr);
    float acc[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) acc[k2] = 0.f;
    const int rounds = (k + NG - 1) / NG;
    for (int u0 = 0; u0 < rounds; u0 += U) {
      float pv[U][8];
      bool ok[U];
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const int idx = (u0 + uu) * NG + g;
        ok[uu] = idx < k;
        if (ok[uu]) load8_f32(sd.P + (size_t)(p0 + idx) * D + 8 * q, pv[uu]);
      }
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        if (ok[uu]) {
#pragma unroll
          for (int k2 = 0; k2 < 8; ++k2) acc[k2] += pv[uu][k2];
        }
      }
    }
    if (has_neg && t < rep) repulse(sd, lds_neg, xr, acc, G, NG, g, q);
    combine_groups(acc, G);
    if (g == 0) {
      float o[8];
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) o[k2] = sd.delta_mode ? acc[k2] : xr[k2] + acc[k2];
      store8_f32(sd.S + (size_t)t * D + 8 * q, o);
    }
  }
}

// X[batch] = S (cpp/algo/algo.hpp:651-666 with the add already folded into
// S), BF16 shadow refresh of the same rows, epoch counter advance.
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
  if (lb == 0 && threadIdx.x == 0 && sd.last_batch) *sd.epoch = *sd.epoch + 1;
}

__global__ void set_int_kernel(int *p, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
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
    const int r = ld.items[4 * it + 0];
    const int e0 = ld.items[4 * it + 1];
    const int cnt = ld.items[4 * it + 2];
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
          p += v * v;
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
      s2 += v * v;
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

// lds_bytes: staged negative rows, max over the launch's steps of ns * D * 4.
hipError_t launch_force(const StepDesc *steps, const int *blk_prefix, int n_steps,
                        int n_blocks, int wpb, int ipw, int unroll, size_t lds_bytes,
                        hipStream_t st) {
  dim3 grid(n_blocks), block(64 * wpb);
  switch (unroll) {
  case 1: hipLaunchKernelGGL(force_kernel<1>, grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw); break;
  case 2: hipLaunchKernelGGL(force_kernel<2>, grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw); break;
  case 8: hipLaunchKernelGGL(force_kernel<8>, grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw); break;
  default: hipLaunchKernelGGL(force_kernel<4>, grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw); break;
  }
  return hipGetLastError();
}

hipError_t launch_fixup(const StepDesc *steps, const int *blk_prefix, int n_steps,
                        int n_blocks, int wpb, int ipw, int unroll, size_t lds_bytes,
                        hipStream_t st) {
  dim3 grid(n_blocks), block(64 * wpb);
  switch (unroll) {
  case 1: hipLaunchKernelGGL(fixup_kernel<1>, grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw); break;
  case 2: hipLaunchKernelGGL(fixup_kernel<2>, grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw); break;
  case 8: hipLaunchKernelGGL(fixup_kernel<8>, grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw); break;
  default: hipLaunchKernelGGL(fixup_kernel<4>, grid, block, lds_bytes, st, steps, blk_prefix, n_steps, ipw); break;
  }
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
