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
// atomic partials, ordered against a per-row counter by publish_partial),
// so every result is deterministic and independent of the launch shape.
// The batch's negative rows are staged once per block in LDS. X holds one
// slot more than there are batches: a fused launch writes its batch's new
// rows into the free slot, which none of its threads reads, so one launch
// per minibatch suffices.
// No MFMA: the step is a sparse gather, not a dense contraction.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <mutex>
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
typedef __attribute__((address_space(1))) u4v gu4;
typedef __attribute__((address_space(1))) const int gcint;

__device__ __forceinline__ int gld_int(const int32_t *p) { return *(gcint *)p; }

// Streaming load (nontemporal) of data read once per epoch: column ids.
__device__ __forceinline__ int gld_int_nt(const int32_t *p) {
  return __builtin_nontemporal_load((gcint *)p);
}

__device__ __forceinline__ u4v gld_u4(const int32_t *p) { return *(gcu4 *)p; }

// A lane's slice of a row: C consecutive components (C = 8 or 16), so a row
// of D components is G = D / C lanes. The slice helpers below move C floats
// (C/4 16-byte accesses) or C BF16 values (C/8 16-byte accesses).
template <int C>
__device__ __forceinline__ void load_f32(const float *p, float (&o)[C]) {
#pragma unroll
  for (int i = 0; i < C / 4; ++i) {
    const f4v a = ((gcf4 *)p)[i];
    o[4 * i] = a.x; o[4 * i + 1] = a.y; o[4 * i + 2] = a.z; o[4 * i + 3] = a.w;
  }
}

// C BF16 values (C/8 16-byte words) as FP32: the low half of each 32-bit
// lane word is the even component.
template <int C>
__device__ __forceinline__ void unpack_bf16(const u4v (&w)[C / 8], float (&o)[C]) {
#pragma unroll
  for (int i = 0; i < C / 8; ++i) {
    o[8 * i + 0] = __uint_as_float(w[i].x << 16); o[8 * i + 1] = __uint_as_float(w[i].x & 0xFFFF0000u);
    o[8 * i + 2] = __uint_as_float(w[i].y << 16); o[8 * i + 3] = __uint_as_float(w[i].y & 0xFFFF0000u);
    o[8 * i + 4] = __uint_as_float(w[i].z << 16); o[8 * i + 5] = __uint_as_float(w[i].z & 0xFFFF0000u);
    o[8 * i + 6] = __uint_as_float(w[i].w << 16); o[8 * i + 7] = __uint_as_float(w[i].w & 0xFFFF0000u);
  }
}

template <int C>
__device__ __forceinline__ void load_bf16(const uint16_t *p, float (&o)[C]) {
  u4v w[C / 8];
#pragma unroll
  for (int i = 0; i < C / 8; ++i) w[i] = ((gcu4 *)p)[i];
  unpack_bf16<C>(w, o);
}

// Global memory only (never an LDS pointer: see store_lds).
template <int C>
__device__ __forceinline__ void store_f32(float *p, const float (&o)[C]) {
#pragma unroll
  for (int i = 0; i < C / 4; ++i) {
    f4v a = {o[4 * i], o[4 * i + 1], o[4 * i + 2], o[4 * i + 3]};
    ((gf4 *)p)[i] = a;
  }
}

// Streaming store of a new row (nontemporal): the rows are read back only by
// later launches, so they need not stay in this launch's L2.
template <int C>
__device__ __forceinline__ void store_f32_nt(float *p, const float (&o)[C]) {
#pragma unroll
  for (int i = 0; i < C / 4; ++i) {
    f4v a = {o[4 * i], o[4 * i + 1], o[4 * i + 2], o[4 * i + 3]};
    __builtin_nontemporal_store(a, (gf4 *)p + i);
  }
}

// A row's 8G floats, held 8 per lane by the G lanes of a group (lane q:
// components [8q, 8q+8)), stored nontemporally so that each of the two store
// instructions writes one contiguous half of the row: lane q writes
// components [4q, 4q+4) and then [4G+4q, 4G+4q+4), fetched from lanes q/2
// and G/2+q/2 of its group. Stored as each lane's own two 16-byte halves,
// every instruction would write half of each 128-byte line, and a line
// leaving L2 between the two would reach memory twice. `first` is the
// group's first lane; every lane of the group must take part.
__device__ __forceinline__ void store8_f32_nt_lines(float *row, int q, int G, int first,
                                                    const float (&o)[8]) {
  if (G < 2) {
    store_f32_nt(row, o);
    return;
  }
  const int h = q & 1;
  const int la = (first + (q >> 1)) << 2, lb = (first + (G >> 1) + (q >> 1)) << 2;
  f4v a, b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float a0 = __int_as_float(__builtin_amdgcn_ds_bpermute(la, __float_as_int(o[i])));
    const float a1 = __int_as_float(__builtin_amdgcn_ds_bpermute(la, __float_as_int(o[4 + i])));
    const float b0 = __int_as_float(__builtin_amdgcn_ds_bpermute(lb, __float_as_int(o[i])));
    const float b1 = __int_as_float(__builtin_amdgcn_ds_bpermute(lb, __float_as_int(o[4 + i])));
    a[i] = h ? a1 : a0;
    b[i] = h ? b1 : b0;
  }
  __builtin_nontemporal_store(a, (gf4 *)(row + 4 * q));
  __builtin_nontemporal_store(b, (gf4 *)(row + 4 * G + 4 * q));
}

typedef __attribute__((address_space(3))) f4v lf4;

template <int C>
__device__ __forceinline__ void store_lds(float *p, const float (&o)[C]) {
#pragma unroll
  for (int i = 0; i < C / 4; ++i) {
    f4v a = {o[4 * i], o[4 * i + 1], o[4 * i + 2], o[4 * i + 3]};
    ((lf4 *)p)[i] = a;
  }
}

// Partials of heavy rows cross CUs (and XCDs, whose L2s are not coherent with
// each other) inside one launch. They are written and read with agent-scope
// atomics (device-coherent sc1 accesses); publish_partial orders them
// against the per-row counter.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) int gint;

// A heavy-row partial (D = C*G floats) in component-major order: lane q of
// the row's group keeps its component k at base[k*G + q], so each of the C
// device-coherent 4-byte accesses of the group covers one contiguous span
// of 4G bytes. (The row layout, C consecutive floats per lane, would make
// every such access write 4 bytes out of each 4C and pass the whole row to
// memory C times over.)
template <int C>
__device__ __forceinline__ void store_coherent(float *base, int q, int G, const float (&o)[C]) {
  gfloat *g = (gfloat *)base + q;
#pragma unroll
  for (int k = 0; k < C; ++k) __hip_atomic_store(g + k * G, o[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int C>
__device__ __forceinline__ void load_coherent(const float *base, int q, int G, float (&o)[C]) {
  gfloat *g = (gfloat *)base + q;
#pragma unroll
  for (int k = 0; k < C; ++k) o[k] = __hip_atomic_load(g + k * G, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Counts this wave's chunk partial of a heavy row (just stored by its lanes
// with agent-scope atomic stores) and tells, in every lane, whether it was
// the row's last; the last chunk then reads all partials with agent-scope
// atomic loads. Agent-scope atomics are coherent device-wide by themselves
// (one modification order per location across all XCDs); what the protocol
// needs beyond that is order:
//  * the partial before the count: s_waitcnt vmcnt(0) waits until every
//    store of the wave has completed (on this target family vmcnt counts
//    stores), which is the wait the compiler's own agent-scope release
//    emits before the releasing access. Its other part, an L2 writeback,
//    only publishes non-atomic data, which the consumer does not read; as a
//    fence it made the force launch 3.3 times slower (123 us against 37.5 us
//    per launch on R-MAT scale 20).
//  * the count before the partials are read: the loads are issued only after
//    the wave has the counter's old value (a data dependence through the
//    shuffle and the branch); an acquire fence would add an L2 invalidation,
//    again for non-atomic data only.
// The "memory" clobbers keep the compiler from moving any memory access
// across either point.
__device__ __forceinline__ bool publish_partial(int *counter, int kparts, int lane) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  int old = 0;
  if (lane == 0)
    old = __hip_atomic_fetch_add((gint *)counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  old = __shfl(old, 0);
  asm volatile("" ::: "memory");
  return old == kparts - 1;
}

// Two FP32 values to BF16, round to nearest even (one v_cvt_pk_bf16_f32;
// subnormals are kept, a NaN stays a NaN); lo in bits 0-15.
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf16x2(float lo, float hi) {
  const bf16x2v b = __builtin_convertvector((f2v){lo, hi}, bf16x2v);
  return __builtin_bit_cast(uint32_t, b);
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

// MODE 0: FP32 gathers, 1: BF16 gathers, 2: per step at run time (a launch
// mixing jobs of both kinds).
template <int MODE> __device__ __forceinline__ bool use_bf16(const StepDesc &sd) {
  return MODE == 2 ? sd.bf16 != 0 : MODE == 1;
}

// Slot positions of the rows one step reads (see Layout): batch k of row c
// is in slot (k - shift - [k < jcut]) mod P = t or t - P, where
// t = k + (P - shift - 1) + [k >= jcut] lies in [0, 2P).
struct SlotMap {
  uint32_t div_m;
  int div_sh;
  int B;
  int P;
  int b_shift;
  int kadd;    // P - shift - 1
  int jcut;
};

__device__ __forceinline__ SlotMap make_map(const Layout &lay, int shift, int jcut) {
  return SlotMap{lay.div_m, lay.div_sh, lay.B, lay.P, lay.b_shift, lay.P - shift - 1, jcut};
}

// Position (row index in the slot layout) of row c. P2: the batch size is a
// power of two (shifts only); otherwise one multiply-high division.
template <bool P2>
__device__ __forceinline__ uint32_t row_pos_t(const SlotMap &m, uint32_t c) {
  if constexpr (P2) {
    const int k = (int)(c >> m.b_shift);
    int t = k + m.kadd + (k >= m.jcut ? 1 : 0);
    t -= t >= m.P ? m.P : 0;
    return c + ((uint32_t)(t - k) << m.b_shift);
  } else {
    const int k = (int)(uint32_t)(((uint64_t)c * m.div_m) >> m.div_sh);
    int t = k + m.kadd + (k >= m.jcut ? 1 : 0);
    t -= t >= m.P ? m.P : 0;
    return c + (uint32_t)(t - k) * (uint32_t)m.B;
  }
}

__device__ __forceinline__ uint32_t row_pos(const SlotMap &m, uint32_t c) {
  return m.b_shift >= 0 ? row_pos_t<true>(m, c) : row_pos_t<false>(m, c);
}

// Row c as C FP32 components [Cq, Cq+C) (BF16 gathers: the shadow's values).
template <int MODE, int C>
__device__ __forceinline__ void load_row(const StepDesc &sd, const SlotMap &m, int c, int q,
                                         float (&o)[C]) {
  const size_t pos = row_pos(m, (uint32_t)c);
  if (use_bf16<MODE>(sd)) load_bf16(sd.Xs + pos * sd.D + C * q, o);
  else load_f32(sd.X + pos * sd.D + C * q, o);
}

// A gathered row slice held while its load is in flight: C FP32 registers,
// or with BF16 gathers the C packed BF16 values (C/2 registers), unpacked
// only when used -- twice the rows in flight for the same registers.
template <int MODE, int C = 8> struct RowBuf {
  float f[C];
};
template <int C> struct RowBuf<1, C> {
  u4v w[C / 8];
};

// Where a lane reads gathered rows from: its C-component slice of X (or of
// the BF16 shadow) at the rows' slot positions. The address is arithmetic
// only (no branch), so the U loads of a round issue back to back.
struct RowSrc {
  const char *xq;      // X or Xs, plus this lane's slice offset
  int row_shift;       // log2 of the bytes per gathered row (D is 8 x 2^i)
  SlotMap m;
  bool bf;             // gathers read BF16
};

template <int MODE, int C = 8>
__device__ __forceinline__ RowSrc make_src(const StepDesc &sd, const SlotMap &m, int q, int G) {
  RowSrc r;
  r.bf = use_bf16<MODE>(sd);
  // D = C*G floats per row
  const int dshift = (C == 16 ? 4 : 3) + (31 - __builtin_clz((unsigned)G));
  if (r.bf) {
    r.xq = reinterpret_cast<const char *>(sd.Xs + C * q);
    r.row_shift = dshift + 1;
  } else {
    r.xq = reinterpret_cast<const char *>(sd.X + C * q);
    r.row_shift = dshift + 2;
  }
  r.m = m;
  return r;
}

// The lane's slice of the row at slot position pos.
template <int MODE, int C>
__device__ __forceinline__ void load_at(const RowSrc &src, uint32_t pos, RowBuf<MODE, C> &b) {
  const char *p = src.xq + ((uint64_t)pos << src.row_shift);
  if constexpr (MODE == 1) {
#pragma unroll
    for (int i = 0; i < C / 8; ++i) b.w[i] = ((gcu4 *)p)[i];
  } else if constexpr (MODE == 0) {
    load_f32(reinterpret_cast<const float *>(p), b.f);
  } else {
    if (src.bf) load_bf16(reinterpret_cast<const uint16_t *>(p), b.f);
    else load_f32(reinterpret_cast<const float *>(p), b.f);
  }
}

template <int MODE, int C>
__device__ __forceinline__ void unpack_buf(const RowBuf<MODE, C> &b, float (&o)[C]) {
  if constexpr (MODE == 1) {
    unpack_bf16<C>(b.w, o);
  } else {
#pragma unroll
    for (int k = 0; k < C; ++k) o[k] = b.f[k];
  }
}

// acc += lr * clamp(v * d1, +-bound) for one interaction whose squared
// distance s the group has reduced. d1 = -2/(1+s) (attractive,
// cpp/algo/algo.hpp:573) or 2/((s+1e-6)(1+s)) (repulsive, :641).
// While every s in the wave is finite, so is every v*d1: d1 takes the
// hardware reciprocal (1 ulp) and the clamp is one med3, which for finite
// values equals scale(). A non-finite s anywhere in the wave takes the exact
// expressions instead, whose clamp passes a NaN through as scale() does.
// Repulsive forces whose clamp binds nowhere in the wave (s d1^2 <= bound^2,
// so |v_k d1| <= sqrt(s) d1 <= bound, the usual case away from the row
// itself) take one fused multiply-add per component with lr*d1.
template <bool ATTR, int C>
__device__ __forceinline__ void add_force(const float (&v)[C], float s, float lr, float bound,
                                          float (&acc)[C]) {
  if (__builtin_amdgcn_ballot_w64(!__builtin_isfinite(s)) == 0) {
    const float d1 = ATTR ? -2.0f * __builtin_amdgcn_rcpf(1.0f + s)
                          : 2.0f * __builtin_amdgcn_rcpf((s + 1e-6f) * (1.0f + s));
    if (!ATTR && __builtin_amdgcn_ballot_w64(!(s * d1 * d1 <= bound * bound)) == 0) {
      const float c = lr * d1;
#pragma unroll
      for (int k = 0; k < C; ++k) acc[k] = fmaf(c, v[k], acc[k]);
      return;
    }
#pragma unroll
    for (int k = 0; k < C; ++k)
      acc[k] = fmaf(lr, __builtin_amdgcn_fmed3f(v[k] * d1, -bound, bound), acc[k]);
  } else {
    const float d1 = ATTR ? -2.0f / (1.0f + s) : 2.0f / ((s + 1e-6f) * (1.0f + s));
#pragma unroll
    for (int k = 0; k < C; ++k) acc[k] = fmaf(lr, clampf(v[k] * d1, bound), acc[k]);
  }
}

// v = xd - xc and the lane's part of |v|^2: a serial chain of fused
// multiply-adds over the lane's C components.
template <int C>
__device__ __forceinline__ float diff_sq(const float (&xd)[C], const float (&xc)[C],
                                         float (&v)[C]) {
#pragma unroll
  for (int k = 0; k < C; ++k) v[k] = xd[k] - xc[k];
  float p = v[0] * v[0];
#pragma unroll
  for (int k = 1; k < C; ++k) p = fmaf(v[k], v[k], p);
  return p;
}

// One gathered slice against the row's own slice xd. (Written out rather
// than through diff_sq: the helper made the compiler keep 4 more VGPRs in
// the repulsion, which cost the two-rows-in-flight BF16 kernel a wave.)
template <bool ATTR, int C>
__device__ __forceinline__ void interact(const float (&xd)[C], const float (&xc)[C], int G,
                                         float lr, float bound, float (&acc)[C]) {
  float v[C];
#pragma unroll
  for (int k = 0; k < C; ++k) v[k] = xd[k] - xc[k];
  float p = v[0] * v[0];
#pragma unroll
  for (int k = 1; k < C; ++k) p = fmaf(v[k], v[k], p);
  add_force<ATTR>(v, group_sum(p, G), lr, bound, acc);
}

// Attractive forces of U gathered slices (cpp/algo/algo.hpp:559-579) on
// the row's accumulators, one slot after the other. While the clamp bound
// is >= 2 it never binds: |v_k d1| = 2|v_k| / (1 + s) <= 1 because
// s >= v_k^2, so scale() is the identity and each component takes one fused
// multiply-add with c = -2 lr rcp(1 + s) (the hardware reciprocal, 1 ulp).
// The same form reproduces the reference's non-finite cases: s = +inf
// gives c = -0, so finite components gain +-0 and infinite ones NaN
// (inf * -0), as v*d1 with d1 = -2/inf does; s = NaN gives NaN throughout.
// FAST is bound >= 2 (uniform per step); a smaller bound takes the exact
// expressions of add_force.
template <int U, int MODE, bool FAST, int C>
__device__ __forceinline__ void attract_round(const float (&xd)[C],
                                              const RowBuf<MODE, C> (&xb)[U], int G, float lr,
                                              float bound, float (&acc)[C]) {
  if constexpr (FAST) {
    const float c2 = -2.0f * lr;
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      float xc[C], v[C];
      unpack_buf<MODE>(xb[uu], xc);
      const float p = diff_sq(xd, xc, v);
      const float c = c2 * __builtin_amdgcn_rcpf(1.0f + group_sum(p, G));
#pragma unroll
      for (int k = 0; k < C; ++k) acc[k] = fmaf(c, v[k], acc[k]);
    }
  } else {
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      float xc[C], v[C];
      unpack_buf<MODE>(xb[uu], xc);
#pragma unroll
      for (int k = 0; k < C; ++k) v[k] = xd[k] - xc[k];
      float p = v[0] * v[0];
#pragma unroll
      for (int k = 1; k < C; ++k) p = fmaf(v[k], v[k], p);
      add_force<true>(v, group_sum(p, G), lr, bound, acc);
    }
  }
}

// Staging piece p (negative p / G, slice p % G) of the step's negative rows,
// as FP32 in LDS; id is that negative's row.
template <int MODE, int C>
__device__ __forceinline__ void stage_piece(const StepDesc &sd, const SlotMap &m, int p, int id,
                                            int G, float *lds) {
  const int k = p / G;
  const int q = p - k * G;
  float v[C];
  load_row<MODE>(sd, m, id, q, v);
  store_lds(lds + (size_t)k * sd.D + C * q, v);
}

// The lane's slice of staged negative k (LDS).
template <int C>
__device__ __forceinline__ void lds_slice(const float *pn, float (&o)[C]) {
#pragma unroll
  for (int i = 0; i < C / 4; ++i) {
    const float4 a = *reinterpret_cast<const float4 *>(pn + 4 * i);
    o[4 * i] = a.x; o[4 * i + 1] = a.y; o[4 * i + 2] = a.z; o[4 * i + 3] = a.w;
  }
}

// Repulsive forces of the staged negatives on x_r, spread over the groups.
template <int C>
__device__ __forceinline__ void repulse(const StepDesc &sd, const float *lds,
                                        const float (&xr)[C], float (&acc)[C],
                                        int G, int NG, int g, int q) {
  const int ns = sd.ns;
  const int rounds = (ns + NG - 1) / NG;
  const float lr = sd.lr, bound = sd.bound;
  for (int u = 0; u < rounds; ++u) {
    const int k = u * NG + g;
    const bool ok = k < ns;
    // a group without a negative this round meets the row itself: v = 0
    float xn[C];
    if (ok) {
      lds_slice(lds + k * sd.D + C * q, xn);
    } else {
#pragma unroll
      for (int k2 = 0; k2 < C; ++k2) xn[k2] = xr[k2];
    }
    interact<false>(xr, xn, G, lr, bound, acc);
  }
}

template <int C>
__device__ __forceinline__ void combine_groups(float (&acc)[C], int G) {
  for (int m = G; m < 64; m <<= 1) {
#pragma unroll
    for (int k = 0; k < C; ++k) acc[k] += __shfl_xor(acc[k], m);
  }
}

// Sum partials p0..p0+k-1: group g takes g, g+NG, ... in order, then the
// fixed xor tree of combine_groups. UP partials are in flight per group
// (the order of the additions does not depend on it).
template <int U, int C>
__device__ __forceinline__ void sum_partials(const StepDesc &sd, int p0, int k, int NG, int g,
                                             int q, float (&acc)[C]) {
  constexpr int UP = C == 8 ? U : (U > 2 ? 2 : U);
  const int D = sd.D;
  const int G = D / C;
  const int rounds = (k + NG - 1) / NG;
  for (int u0 = 0; u0 < rounds; u0 += UP) {
    float pv[UP][C];
    bool ok[UP];
#pragma unroll
    for (int uu = 0; uu < UP; ++uu) {
      const int idx = (u0 + uu) * NG + g;
      ok[uu] = idx < k;
      if (ok[uu]) load_coherent(sd.P + (size_t)(p0 + idx) * D, q, G, pv[uu]);
    }
#pragma unroll
    for (int uu = 0; uu < UP; ++uu) {
      if (ok[uu]) {
#pragma unroll
        for (int k2 = 0; k2 < C; ++k2) acc[k2] += pv[uu][k2];
      }
    }
  }
}

// Attractive forces of edges [e0, e0+cnt) of row r on the per-group
// accumulators: group g takes edges g, g+NG, ... in order, U rows in flight.
// A slot past the end, or holding a column the quirk skips (c >= c_limit),
// reads row r itself, whose distance to itself is exactly zero: no force.
// Each lane turns its column of the 64-edge block into a slot position once;
// a slot then fetches the position from lane k (ds_bpermute takes the lane
// index modulo 64). Lanes k in [m, 64) hold row r; k >= 64 happens only for
// U > G.
template <int U, int MODE, bool FAST, int C>
__device__ __forceinline__ void gather_edges_t(const StepDesc &sd, const RowSrc &src,
                                             const float (&xd)[C], int r, int e0, int cnt,
                                             int G, int NG, int g, int lane, float (&acc)[C]) {
  const float lr = sd.lr, bound = sd.bound;
  const int c_limit = sd.c_limit;
  const uint32_t pos_r = row_pos(src.m, (uint32_t)r);
  for (int eb = 0; eb < cnt; eb += 64) {
    const int m = min(64, cnt - eb);
    int cl = lane < m ? gld_int_nt(sd.col + e0 + eb + lane) : r;
    cl = cl < c_limit ? cl : r;
    const int pl = (int)row_pos(src.m, (uint32_t)cl);
    const int rounds = (m + NG - 1) / NG;
    for (int u0 = 0; u0 < rounds; u0 += U) {
      RowBuf<MODE, C> xb[U];
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const int k = (u0 + uu) * NG + g;
        uint32_t pos = (uint32_t)__builtin_amdgcn_ds_bpermute(k << 2, pl);
        if (U > G && k >= 64) pos = pos_r;
        load_at<MODE>(src, pos, xb[uu]);
      }
      // keep the U loads ahead of every use of their results
      __builtin_amdgcn_sched_barrier(0);
      attract_round<U, MODE, FAST>(xd, xb, G, lr, bound, acc);
    }
  }
}

// The row's own slice for the distances: the master row, rounded as the
// shadow holds it when gathers read BF16 (no extra traffic), so a row meeting
// itself as a negative or through a self-loop feels exactly zero force, as
// in the reference; the update still goes to the FP32 master.
template <int MODE, int C>
__device__ __forceinline__ void own_slice(const StepDesc &sd, const float (&xr)[C],
                                          float (&xd)[C]) {
  if (use_bf16<MODE>(sd)) {
#pragma unroll
    for (int k = 0; k < C; k += 2) {
      const uint32_t w = bf16x2(xr[k], xr[k + 1]);
      xd[k] = __uint_as_float(w << 16);
      xd[k + 1] = __uint_as_float(w & 0xFFFF0000u);
    }
  } else {
#pragma unroll
    for (int k = 0; k < C; ++k) xd[k] = xr[k];
  }
}

// Row t of the batch: a fused step writes the new row straight into the
// free slot (position out + t) of X and, with BF16 gathers, of the shadow;
// an unfused step writes the new row or the delta to S.
template <int MODE, int C>
__device__ __forceinline__ void write_row(const StepDesc &sd, uint32_t out, int t, int q,
                                          const float (&o)[C]) {
  if (!sd.fused) {
    store_f32(sd.S + (size_t)t * sd.D + C * q, o);
    return;
  }
  const size_t pos = (size_t)(out + (uint32_t)t);
  store_f32_nt(sd.Xw + pos * sd.D + C * q, o);
  if (use_bf16<MODE>(sd)) {
    gu4 *ps = (gu4 *)(sd.Xsw + pos * sd.D + C * q);
#pragma unroll
    for (int i = 0; i < C / 8; ++i) {
      u4v w;
      w.x = bf16x2(o[8 * i], o[8 * i + 1]);
      w.y = bf16x2(o[8 * i + 2], o[8 * i + 3]);
      w.z = bf16x2(o[8 * i + 4], o[8 * i + 5]);
      w.w = bf16x2(o[8 * i + 6], o[8 * i + 7]);
      __builtin_nontemporal_store(w, ps + i);
    }
  }
}

// Repulsion (first `rep` rows), group combine, new row (or delta) to S.
template <int MODE, int C>
__device__ __forceinline__ void finish_row(const StepDesc &sd, uint32_t out, const float *lds_neg,
                                           bool repel, int t, const float (&xr)[C],
                                           const float (&xd)[C], float (&acc)[C], int G, int NG,
                                           int g, int q) {
  if (repel) repulse(sd, lds_neg, xd, acc, G, NG, g, q);
  combine_groups(acc, G);
  if (g == 0) {
    float o[C];
#pragma unroll
    for (int k = 0; k < C; ++k) o[k] = sd.delta_mode ? acc[k] : xr[k] + acc[k];
    write_row<MODE>(sd, out, t, q, o);
  }
}

// Light rows, one per lane group: the group takes light entry ia (row r,
// edges [e0, e0+deg)), gathers its edges in CSR order (U in flight), repels
// against every staged negative and writes its row. The loops run to the
// pack's largest degree kmax, so control flow stays uniform; slots past a
// row's degree read the row itself (no force).
template <int U, int MODE, bool FAST, int C>
__device__ __forceinline__ void light_pack_t(const StepDesc &sd, const RowSrc &src, uint32_t out,
                                           u4v ia, int lidx, int pc, int kmax, bool valid,
                                           const float (&xr)[C], const float *lds_neg,
                                           bool has_neg, int rep, int G, int g, int q,
                                           bool &synced) {
  const int r = (int)ia.x, deg = valid ? (int)ia.z : 0;
  const int32_t *lc = sd.lcols + (size_t)lidx * kLightCols;
  kmax = __builtin_amdgcn_readfirstlane(kmax);
  const float lr = sd.lr, bound = sd.bound;
  const int c_limit = sd.c_limit;
  float xd[C], acc[C];
  own_slice<MODE>(sd, xr, xd);
#pragma unroll
  for (int k = 0; k < C; ++k) acc[k] = 0.f;
  // slot positions as in gather_edges_t: lane q of the group holds the
  // position of edge kb+q (row r past the degree); k - kb >= G only for U > G
  const uint32_t pos_r = row_pos(src.m, (uint32_t)r);
  for (int kb = 0; kb < kmax; kb += G) {
    // the first G columns came with the item (pc); kb + q < deg <= kLightCols
    int cl = (kb + q < deg) ? (kb == 0 ? pc : gld_int_nt(lc + kb + q)) : r;
    cl = cl < c_limit ? cl : r;
    const int pl = (int)row_pos(src.m, (uint32_t)cl);
    const int kend = min(kmax, kb + G);
    for (int k0 = kb; k0 < kend; k0 += U) {
      RowBuf<MODE, C> xb[U];
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const int k = k0 + uu;
        uint32_t pos = (uint32_t)__builtin_amdgcn_ds_bpermute((g * G + (k - kb)) << 2, pl);
        if (U > G && k - kb >= G) pos = pos_r;
        load_at<MODE>(src, pos, xb[uu]);
      }
      __builtin_amdgcn_sched_barrier(0);
      attract_round<U, MODE, FAST>(xd, xb, G, lr, bound, acc);
    }
  }
  // the wave's first unit waits here for the block's staged negative rows
  if (!synced) {
    __syncthreads();
    synced = true;
  }
  const int t = r - sd.row_begin;
  if (has_neg) {
    // rows past the repulsive count (or an empty slot) meet themselves: v = 0
    const bool repel = valid && t < rep;
    for (int n = 0; n < sd.ns; ++n) {
      float xn[C];
      if (repel) {
        lds_slice(lds_neg + (size_t)n * sd.D + C * q, xn);
      } else {
#pragma unroll
        for (int k2 = 0; k2 < C; ++k2) xn[k2] = xd[k2];
      }
      interact<false>(xd, xn, G, lr, bound, acc);
    }
  }
  if (valid) {
    float o[C];
#pragma unroll
    for (int k = 0; k < C; ++k) o[k] = sd.delta_mode ? acc[k] : xr[k] + acc[k];
    write_row<MODE>(sd, out, t, q, o);
  }
}

// One uniform branch per work unit picks the single-FMA attractive form
// (clamp bound >= 2).
template <int U, int MODE, int C>
__device__ __forceinline__ void gather_edges(const StepDesc &sd, const RowSrc &src,
                                             const float (&xd)[C], int r, int e0, int cnt,
                                             int G, int NG, int g, int lane, float (&acc)[C]) {
  if (sd.bound >= 2.0f) gather_edges_t<U, MODE, true>(sd, src, xd, r, e0, cnt, G, NG, g, lane, acc);
  else gather_edges_t<U, MODE, false>(sd, src, xd, r, e0, cnt, G, NG, g, lane, acc);
}

template <int U, int MODE, int C>
__device__ __forceinline__ void light_pack(const StepDesc &sd, const RowSrc &src, uint32_t out,
                                           u4v ia, int lidx, int pc, bool valid,
                                           const float (&xr)[C], const float *lds_neg,
                                           bool has_neg, int rep, int G, int g, int q,
                                           bool &synced) {
  const int kmax = (int)ia.w;
  if (sd.bound >= 2.0f)
    light_pack_t<U, MODE, true>(sd, src, out, ia, lidx, pc, kmax, valid, xr, lds_neg, has_neg,
                                rep, G, g, q, synced);
  else
    light_pack_t<U, MODE, false>(sd, src, out, ia, lidx, pc, kmax, valid, xr, lds_neg, has_neg,
                                 rep, G, g, q, synced);
}

// One minibatch of one or more jobs. Blocks [0, n_block_items) of a step
// each take one block item: a row of degree <= kSub*chunk, or one chunk of a
// heavier row, cut into kSub fixed sub-ranges that the block's waves gather
// and then add in sub-range order through LDS. A chunk publishes its
// partial (device-coherent stores) and bumps the row's counter; the chunk
// that counts last adds all partials of the row in index order. The
// remaining blocks' waves walk the medium rows (longest first) and then the
// light-row packs with a grid stride, so the launch ends on short units. The
// negative rows are staged in LDS without a barrier of their own: every wave
// passes exactly one barrier before its first repulsion (a block item's
// after its gathers, a wave's after its first unit's gathers, or at its end
// if it has no unit), so staging overlaps the first gathers. A fused step
// writes its new rows into the layout's free slot, which no thread of the
// launch reads. C: components per lane (the step's G = D / C lanes per row).
// GT > 0: every step of the launch has G = GT (a compile-time constant);
// 0: per step. BLK false: no step of the launch has block items, and the
// block-item path is compiled out (its registers cost the other paths a
// wave per SIMD). lb, nb: this block's index among the step's blocks and
// their count.
template <int U, int MODE, int GT, int C, bool BLK>
__device__ __forceinline__ void force_body(const StepDesc &sd, int lb, int nb) {
  extern __shared__ float lds[];
  const int D = sd.D;
  const int G = GT > 0 ? GT : sd.G;
  const int NG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int g = lane / G;
  const int q = lane & (G - 1);

  // this wave's first work item, in flight during the staging
  const int nbi = BLK ? sd.n_block_items : 0;
  const int n_light = sd.n_light;
  const int n_wave_items = sd.n_items - nbi - n_light;
  const int32_t *witems = sd.items + (size_t)kItemInts * nbi;
  const int32_t *lrows = witems + (size_t)kItemInts * n_wave_items;
  const int n_packs = (n_light + NG - 1) / NG;
  // wave units: medium rows [0, n_wave_items), then light packs
  int u = (lb - nbi) * wpb + wave;
  const bool blk_item = BLK && lb < nbi; // this block takes a block item
  u4v pre0 = {0u, 0u, 0u, 0u}, pre1 = {0u, 0u, 0u, 0u};
  int pc = 0; // a light row's column kb + q = q (its first G columns)
  if (blk_item) {
    pre0 = gld_u4(sd.items + (size_t)kItemInts * lb);
    pre1 = gld_u4(sd.items + (size_t)kItemInts * lb + 4);
  } else if (u < n_wave_items) {
    pre0 = gld_u4(witems + (size_t)kItemInts * u);
  } else if (u - n_wave_items < n_packs) {
    const int pk = u - n_wave_items;
    const int idx = pk * NG + g;
    const int li = idx < n_light ? idx : pk * NG;
    pre0 = gld_u4(lrows + (size_t)kItemInts * li);
    if (q < kLightCols) pc = gld_int_nt(sd.lcols + (size_t)li * kLightCols + q);
  }
  // this step's negative ids, one per staging piece (first blockDim pieces)
  const int total_pieces = sd.ns * G;
  const int pz = threadIdx.x;
  int nid = 0;
  if (pz < total_pieces) nid = gld_int(sd.neg_ids + pz / G);

  // slot positions: a fused step reads with the run's rotation and writes
  // the free slot; an unfused step reads the layout at rest
  const int epoch = sd.state[0];
  const int shift = sd.fused ? sd.state[2] : sd.state[1];
  const SlotMap m = make_map(sd.lay, shift, sd.fused ? sd.batch : 0);
  int own_sl = sd.batch - shift;
  own_sl += own_sl < 0 ? sd.lay.P : 0;
  const uint32_t own_delta = (uint32_t)(own_sl - sd.batch) * (uint32_t)sd.lay.B;
  int out_sl = own_sl - 1;
  out_sl += out_sl < 0 ? sd.lay.P : 0;
  const uint32_t out = (uint32_t)out_sl * (uint32_t)sd.lay.B;

  // the own row of this wave's first unit, in flight with the negative rows
  const bool unit0 = blk_item || u < n_wave_items + n_packs;
  float xr0[C];
  if (unit0) load_f32(sd.X + (size_t)(pre0.x + own_delta) * D + C * q, xr0);
  const int rep = epoch == 0 ? sd.rep0 : sd.rep1;
  const bool has_neg = sd.ns > 0 && epoch + sd.batch < sd.n_lists;
  if (has_neg) {
    if (pz < total_pieces) stage_piece<MODE, C>(sd, m, pz, nid, G, lds);
    for (int p = pz + blockDim.x; p < total_pieces; p += blockDim.x)
      stage_piece<MODE, C>(sd, m, p, gld_int(sd.neg_ids + p / G), G, lds);
  }
  const RowSrc src = make_src<MODE, C>(sd, m, q, G);

  if (blk_item) {
    const int r = (int)pre0.x, e0 = (int)pre0.y, cnt = (int)pre0.z, pslot = (int)pre0.w;
    const int t = r - sd.row_begin;
    float xd[C], acc[C];
    own_slice<MODE>(sd, xr0, xd);
    float *lds_sub = lds + (has_neg ? sd.ns * D : 0);
    for (int s = wave; s < kSub; s += wpb) {
      const int lo = (int)((int64_t)cnt * s / kSub), hi = (int)((int64_t)cnt * (s + 1) / kSub);
#pragma unroll
      for (int k = 0; k < C; ++k) acc[k] = 0.f;
      gather_edges<U, MODE>(sd, src, xd, r, e0 + lo, hi - lo, G, NG, g, lane, acc);
      combine_groups(acc, G);
      if (g == 0) store_lds(lds_sub + (size_t)s * D + C * q, acc);
    }
    __syncthreads();
    if (wave == 0) {
      // group 0 holds the sum of the sub-ranges, in sub-range order
#pragma unroll
      for (int k = 0; k < C; ++k) acc[k] = 0.f;
      if (g == 0) {
        for (int s = 0; s < kSub; ++s) {
          const float *p = lds_sub + (size_t)s * D + C * q;
#pragma unroll
          for (int k = 0; k < C; ++k) acc[k] += p[k];
        }
      }
      bool finish = pslot < 0;
      if (!finish) {
        const int sid = (int)pre1.x, p0 = (int)pre1.y, kparts = (int)pre1.z;
        if (g == 0) store_coherent(sd.P + (size_t)pslot * D, q, G, acc);
        if (publish_partial(sd.counters + sid, kparts, lane)) {
#pragma unroll
          for (int k = 0; k < C; ++k) acc[k] = 0.f;
          sum_partials<U>(sd, p0, kparts, NG, g, q, acc);
          if (lane == 0)
            __hip_atomic_store((gint *)(sd.counters + sid), 0, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
          finish = true;
        }
      }
      if (finish)
        finish_row<MODE>(sd, out, lds, has_neg && t < rep, t, xr0, xd, acc, G, NG, g, q);
    }
  } else {
    // wave work: the medium rows (one row per wave), then packs of light
    // rows (one row per lane group), walked with a grid stride
    const int W = (nb - nbi) * wpb;
    bool have = true;    // pre0/pre1 and xr0 hold this wave's current unit
    bool synced = false; // this wave has passed the block's one barrier
    for (; u < n_wave_items; u += W) {
      if (!have) {
        pre0 = gld_u4(witems + (size_t)kItemInts * u);
        load_f32(sd.X + (size_t)(pre0.x + own_delta) * D + C * q, xr0);
      }
      have = false;
      const int r = (int)pre0.x, e0 = (int)pre0.y, cnt = (int)pre0.z;
      const int t = r - sd.row_begin;
      float xd[C], acc[C];
      own_slice<MODE>(sd, xr0, xd);
#pragma unroll
      for (int k = 0; k < C; ++k) acc[k] = 0.f;
      gather_edges<U, MODE>(sd, src, xd, r, e0, cnt, G, NG, g, lane, acc);
      if (!synced) {
        __syncthreads();
        synced = true;
      }
      finish_row<MODE>(sd, out, lds, has_neg && t < rep, t, xr0, xd, acc, G, NG, g, q);
    }
    for (int pk = u - n_wave_items; pk < n_packs; pk += W) {
      const int idx = pk * NG + g;
      const int li = idx < n_light ? idx : pk * NG;
      if (!have) {
        pre0 = gld_u4(lrows + (size_t)kItemInts * li);
        if (q < kLightCols) pc = gld_int_nt(sd.lcols + (size_t)li * kLightCols + q);
        load_f32(sd.X + (size_t)(pre0.x + own_delta) * D + C * q, xr0);
      }
      have = false;
      light_pack<U, MODE>(sd, src, out, pre0, li, pc, idx < n_light, xr0, lds, has_neg, rep, G,
                          g, q, synced);
    }
    if (!synced) __syncthreads();
  }
}

// Steps of one or more jobs in one launch (engine rounds): the block finds
// its step (one load from the block-to-step table when the round comes with
// one, else a binary search of the block prefix) and reads the step's
// descriptor from memory. C: components per lane, shared by every step;
// BLK: some step has block items.
template <int U, int MODE, int GT, int C, bool BLK>
__global__ __launch_bounds__(512) void force_kernel(
    const StepDesc *__restrict__ steps, const int *__restrict__ blk_prefix,
    const int *__restrict__ blk_step, int n_steps, int ipw) {
  const int b = blockIdx.x;
  const int si = blk_step ? blk_step[b] : find_step(blk_prefix, n_steps, b);
  force_body<U, MODE, GT, C, BLK>(steps[si], b - blk_prefix[si],
                                  blk_prefix[si + 1] - blk_prefix[si]);
  (void)ipw;
}

// One step (a single graph's minibatch): the descriptor travels in the
// kernel arguments, so the first loads of the items, the columns and the
// epoch state do not wait for a descriptor load. C: components per lane;
// BLK: the step has block items.
template <int U, int MODE, int GT, int C, bool BLK>
__global__ __launch_bounds__(512) void force_kernel1(const StepDesc sd) {
  force_body<U, MODE, GT, C, BLK>(sd, (int)blockIdx.x, (int)gridDim.x);
}

// ---- SpMM sibling: SpMMAlgo::algo_spmm at world size 1 ----------------------
// Y[r] = sum over the CSR entries c of row r with c < c_limit of lr * X[c]
// (cpp/algo/spmm.hpp:341-395: dst_id < dst_end_index skips column N-1 like
// the force step; :415-431 assigns the batch's rows to the output). X is
// read through the slot map of the layout at rest and never written, so the
// batches are independent. The work decomposition is the force step's: block
// items (kSub sub-ranges summed in order through LDS, heavy-row chunks
// summed in index order by the last to finish), medium rows one per wave,
// light rows one per lane group. Slots past a row's end or skipped by the
// quirk add nothing (their loads are masked off).

template <int U, int MODE>
__device__ __forceinline__ void sum_round(const RowBuf<MODE> (&xb)[U], const bool (&ok)[U],
                                          float lr, float (&acc)[8]) {
#pragma unroll
  for (int uu = 0; uu < U; ++uu) {
    float xc[8];
    unpack_buf<MODE>(xb[uu], xc);
    const float w = ok[uu] ? lr : 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = fmaf(w, xc[k], acc[k]);
  }
}

// A masked-off slot holds zeros (no load), so a non-finite value elsewhere
// in X cannot leak into the sum through the zero weight.
template <int MODE>
__device__ __forceinline__ void load_or_zero(const RowSrc &src, uint32_t pos, bool ok,
                                             RowBuf<MODE> &b) {
  if constexpr (MODE == 1) {
    b.w[0] = u4v{0u, 0u, 0u, 0u};
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) b.f[k] = 0.0f;
  }
  if (ok) load_at<MODE>(src, pos, b);
}

// Edges [e0, e0+cnt) of one row: group g takes edges g, g+NG, ... in order.
template <int U, int MODE>
__device__ __forceinline__ void sum_edges(const StepDesc &sd, const RowSrc &src, int e0, int cnt,
                                          int G, int NG, int g, int lane, float (&acc)[8]) {
  const int c_limit = sd.c_limit;
  const float lr = sd.lr;
  for (int eb = 0; eb < cnt; eb += 64) {
    const int m = min(64, cnt - eb);
    int cl = lane < m ? gld_int_nt(sd.col + e0 + eb + lane) : c_limit;
    const bool lv = cl < c_limit;
    const int pl = lv ? (int)row_pos(src.m, (uint32_t)cl) : -1;
    const int rounds = (m + NG - 1) / NG;
    for (int u0 = 0; u0 < rounds; u0 += U) {
      RowBuf<MODE> xb[U];
      bool ok[U];
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const int k = (u0 + uu) * NG + g;
        const int pos = __builtin_amdgcn_ds_bpermute(k << 2, pl);
        ok[uu] = k < m && pos >= 0;
        load_or_zero<MODE>(src, (uint32_t)pos, ok[uu], xb[uu]);
      }
      __builtin_amdgcn_sched_barrier(0);
      sum_round<U, MODE>(xb, ok, lr, acc);
    }
  }
}

// One light pack: group g sums its row's deg <= kLightCols columns (the
// first G of them came with the item, pc), then writes Y.
template <int U, int MODE>
__device__ __forceinline__ void sum_light_pack(const StepDesc &sd, const RowSrc &src, float *Y,
                                               u4v ia, int lidx, int pc, bool valid, int G,
                                               int g, int q) {
  const int r = (int)ia.x, deg = valid ? (int)ia.z : 0;
  const int kmax = __builtin_amdgcn_readfirstlane((int)ia.w);
  const int32_t *lc = sd.lcols + (size_t)lidx * kLightCols;
  const int c_limit = sd.c_limit;
  float acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = 0.f;
  for (int kb = 0; kb < kmax; kb += G) {
    const int cl = (kb + q < deg) ? (kb == 0 ? pc : gld_int_nt(lc + kb + q)) : c_limit;
    const int pl = cl < c_limit ? (int)row_pos(src.m, (uint32_t)cl) : -1;
    const int kend = min(kmax, kb + G);
    for (int k0 = kb; k0 < kend; k0 += U) {
      RowBuf<MODE> xb[U];
      bool ok[U];
#pragma unroll
      for (int uu = 0; uu < U; ++uu) {
        const int k = k0 + uu;
        const int pos = __builtin_amdgcn_ds_bpermute((g * G + (k - kb)) << 2, pl);
        ok[uu] = k < kend && (k - kb) < G && pos >= 0;
        load_or_zero<MODE>(src, (uint32_t)pos, ok[uu], xb[uu]);
      }
      __builtin_amdgcn_sched_barrier(0);
      sum_round<U, MODE>(xb, ok, sd.lr, acc);
    }
  }
  if (valid) store8_f32_nt_lines(Y + (size_t)r * sd.D, q, G, g * G, acc);
}

template <int U, int MODE, int GT>
__global__ __launch_bounds__(512) void spmm_kernel(const StepDesc *__restrict__ steps,
                                                   const int *__restrict__ blk_prefix,
                                                   int n_steps, float *__restrict__ Y) {
  extern __shared__ float lds[];
  const int b = blockIdx.x;
  const int si = find_step(blk_prefix, n_steps, b);
  const StepDesc &sd = steps[si];
  const int lb = b - blk_prefix[si];
  const int nb = blk_prefix[si + 1] - blk_prefix[si];
  const int D = sd.D;
  const int G = GT > 0 ? GT : sd.G;
  const int NG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int g = lane / G;
  const int q = lane & (G - 1);
  const int nbi = sd.n_block_items;
  const int n_light = sd.n_light;
  const int n_wave_items = sd.n_items - nbi - n_light;
  const int32_t *witems = sd.items + (size_t)kItemInts * nbi;
  const int32_t *lrows = witems + (size_t)kItemInts * n_wave_items;
  const int n_packs = (n_light + NG - 1) / NG;
  // X at rest: batch k in slot (k - rot) mod P
  const SlotMap m = make_map(sd.lay, sd.state[1], 0);
  const RowSrc src = make_src<MODE>(sd, m, q, G);

  if (lb < nbi) {
    const u4v it0 = gld_u4(sd.items + (size_t)kItemInts * lb);
    const u4v it1 = gld_u4(sd.items + (size_t)kItemInts * lb + 4);
    const int r = (int)it0.x, e0 = (int)it0.y, cnt = (int)it0.z, pslot = (int)it0.w;
    float acc[8];
    for (int s = wave; s < kSub; s += wpb) {
      const int lo = (int)((int64_t)cnt * s / kSub), hi = (int)((int64_t)cnt * (s + 1) / kSub);
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = 0.f;
      sum_edges<U, MODE>(sd, src, e0 + lo, hi - lo, G, NG, g, lane, acc);
      combine_groups(acc, G);
      if (g == 0) store_lds(lds + (size_t)s * D + 8 * q, acc);
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = 0.f;
      if (g == 0) {
        for (int s = 0; s < kSub; ++s) {
          const float *p = lds + (size_t)s * D + 8 * q;
#pragma unroll
          for (int k = 0; k < 8; ++k) acc[k] += p[k];
        }
      }
      bool finish = pslot < 0;
      if (!finish) {
        const int sid = (int)it1.x, p0 = (int)it1.y, kparts = (int)it1.z;
        if (g == 0) store_coherent(sd.P + (size_t)pslot * D, q, G, acc);
        if (publish_partial(sd.counters + sid, kparts, lane)) {
#pragma unroll
          for (int k = 0; k < 8; ++k) acc[k] = 0.f;
          sum_partials<U>(sd, p0, kparts, NG, g, q, acc);
          combine_groups(acc, G);
          if (lane == 0)
            __hip_atomic_store((gint *)(sd.counters + sid), 0, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
          finish = true;
        }
      }
      if (finish && g == 0) store8_f32_nt_lines(Y + (size_t)r * D, q, G, 0, acc);
    }
  } else {
    const int W = (nb - nbi) * wpb;
    for (int u = (lb - nbi) * wpb + wave; u < n_wave_items + n_packs; u += W) {
      if (u < n_wave_items) {
        const u4v it0 = gld_u4(witems + (size_t)kItemInts * u);
        float acc[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] = 0.f;
        sum_edges<U, MODE>(sd, src, (int)it0.y, (int)it0.z, G, NG, g, lane, acc);
        combine_groups(acc, G);
        if (g == 0) store8_f32_nt_lines(Y + (size_t)it0.x * D, q, G, 0, acc);
      } else {
        const int pk = u - n_wave_items;
        const int idx = pk * NG + g;
        const int li = idx < n_light ? idx : pk * NG;
        const u4v it0 = gld_u4(lrows + (size_t)kItemInts * li);
        const int pc = q < kLightCols ? gld_int_nt(sd.lcols + (size_t)li * kLightCols + q) : 0;
        sum_light_pack<U, MODE>(sd, src, Y, it0, li, pc, idx < n_light, G, g, q);
      }
    }
  }
}

// X[batch] = S (cpp/algo/algo.hpp:651-666 with the add already folded into
// S) and the BF16 shadow refresh of the same rows, at the batch's slot in
// the layout at rest.
__global__ __launch_bounds__(256) void apply_kernel(
    const StepDesc *__restrict__ steps, const int *__restrict__ blk_prefix,
    int n_steps) {
  const int b = blockIdx.x;
  const int si = find_step(blk_prefix, n_steps, b);
  const StepDesc &sd = steps[si];
  const int lb = b - blk_prefix[si];
  const int nb = blk_prefix[si + 1] - blk_prefix[si];
  int sl = sd.batch - sd.state[1];
  sl += sl < 0 ? sd.lay.P : 0;
  const int64_t total4 = (int64_t)sd.rows * sd.D / 4;
  const int64_t base4 = (int64_t)sl * sd.lay.B * sd.D / 4;
  const float4 *S4 = reinterpret_cast<const float4 *>(sd.S);
  float4 *X4 = reinterpret_cast<float4 *>(sd.Xw);
  for (int64_t idx = (int64_t)lb * blockDim.x + threadIdx.x; idx < total4;
       idx += (int64_t)nb * blockDim.x) {
    const float4 v = S4[idx];
    X4[base4 + idx] = v;
    if (sd.bf16) {
      uint2 w;
      w.x = bf16x2(v.x, v.y);
      w.y = bf16x2(v.z, v.w);
      reinterpret_cast<uint2 *>(sd.Xsw)[base4 + idx] = w;
    }
  }
}

// Epoch state of `n` jobs (one block each): set (epoch := value, rot :=
// rot_value) or advance (epoch += 1), keeping shift = (rot + epoch) mod P;
// then each job's current negative ids negcur[b*ns + k] = list (epoch + b),
// id k (0 where no such list exists) are refreshed, so the force launches
// read them without waiting for the counter.
__global__ void epoch_kernel(const EpochDesc *__restrict__ jobs, int set, int value,
                             int rot_value) {
  const EpochDesc ed = jobs[blockIdx.x];
  const int old_epoch = ed.state[0];
  const int old_shift = ed.state[2];
  __syncthreads();
  int e, sh;
  if (set) {
    e = value;
    sh = (rot_value + value) % ed.P;
    sh += sh < 0 ? ed.P : 0;
  } else {
    e = old_epoch + 1;
    sh = old_shift + 1 == ed.P ? 0 : old_shift + 1;
  }
  if (threadIdx.x == 0) {
    ed.state[0] = e;
    if (set) ed.state[1] = rot_value;
    ed.state[2] = sh;
  }
  const int total = ed.batches * ed.ns;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int b = i / ed.ns;
    const int k = i - b * ed.ns;
    const int list = e + b;
    ed.negcur[i] = (list >= 0 && list < ed.n_lists) ? ed.negs[(size_t)list * ed.ns + k] : 0;
  }
}

// Light rows' column ids into lcols (see kLightCols), one thread per item
// and column slot; slots past a row's degree hold 0.
__global__ void fill_lcols_kernel(const int32_t *__restrict__ items, int n_items,
                                  const int32_t *__restrict__ col, int32_t *__restrict__ lcols) {
  const int64_t total = (int64_t)n_items * kLightCols;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t it = i / kLightCols;
    const int k = (int)(i - it * kLightCols);
    const int32_t *item = items + (size_t)kItemInts * it;
    const int l = item[7] - 1;
    if (l < 0) continue;
    lcols[(size_t)l * kLightCols + k] = k < item[2] ? col[item[1] + k] : 0;
  }
}

__global__ void to_bf16_kernel(const float *__restrict__ x, uint16_t *__restrict__ xs,
                               int64_t count2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count2;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = reinterpret_cast<const float2 *>(x)[i];
    reinterpret_cast<uint32_t *>(xs)[i] = bf16x2(v.x, v.y);
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
  const SlotMap mp = make_map(ld.lay, ld.rot, 0);
  double tot = 0.0;
  for (int i = 0; i < ipw; ++i) {
    const int it = (blockIdx.x * ipw + i) * wpb + wave;
    if (it >= ld.n_items) break;
    const int r = ld.items[kItemInts * it + 0];
    const int e0 = ld.items[kItemInts * it + 1];
    const int cnt = ld.items[kItemInts * it + 2];
    float xr[8];
    load_f32(ld.X + (size_t)row_pos(mp, r) * D + 8 * q, xr);
    for (int eb = 0; eb < cnt; eb += 64) {
      const int m = min(64, cnt - eb);
      const int cl = lane < m ? ld.col[e0 + eb + lane] : 0;
      const int rounds = (m + NG - 1) / NG;
      for (int u = 0; u < rounds; ++u) {
        const int k = u * NG + g;
        const int c = __shfl(cl, k & 63);
        const bool ok = (k < m) && (c < ld.c_limit);
        float xc[8];
        load_f32(ld.X + (size_t)row_pos(mp, ok ? c : r) * D + 8 * q, xc);
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
  const SlotMap mp = make_map(ld.lay, ld.rot, 0);
  double tot = 0.0;
  const double eps = 1e-6;
  for (int i = 0; i < pairs_per_group; ++i) {
    const int64_t p = group_id * pairs_per_group + i;
    const bool ok = p < total;
    const int64_t r = ok ? p / ld.eval_ns : 0;
    const int k = ok ? (int)(p - r * ld.eval_ns) : 0;
    const int64_t m = ok ? eval_negative(ld.eval_seed, r, k, ld.n) : 0;
    float xr[8], xm[8];
    load_f32(ld.X + (size_t)row_pos(mp, (uint32_t)r) * D + 8 * q, xr);
    load_f32(ld.X + (size_t)row_pos(mp, (uint32_t)m) * D + 8 * q, xm);
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

// Zeroes n ints (a kernel, not a memset node, so the fill can sit inside a
// captured HIP graph).
__global__ void zero_ints_kernel(int *__restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ---- scratch of the entry point that allocates nothing -----------------------
// fe_step_delta_noalloc keeps a step's heavy-row partials and split counters
// in this array instead of buffers of the context. Every device the library
// runs on holds its own copy for the life of the process. Host code never
// passes the array's name to a kernel: scratch_address() looks up this
// device's address of it once (hipGetSymbolAddress) and every later use
// goes through that address.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices] = {};
} // namespace

hipError_t scratch_address(int device, void **out) {
  *out = nullptr;
  if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice; // the caller makes `device` current
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

// ---- launch wrappers -------------------------------------------------------

// lds_bytes: staged negative rows and sub-range sums (force_lds_bytes);
// mode: 0 all steps gather FP32, 1 all gather BF16, 2 mixed; g_const: the
// lanes per row G shared by every step of the launch, or 0 if they differ.
template <int U, int GT, int C, bool BLK>
static void launch_force_g(dim3 grid, dim3 block, size_t lds_bytes, hipStream_t st, int mode,
                           const StepDesc *steps, const int *blk_prefix, const int *blk_step,
                           int n_steps, int ipw) {
  if (mode == 0) hipLaunchKernelGGL((force_kernel<U, 0, GT, C, BLK>), grid, block, lds_bytes, st, steps, blk_prefix, blk_step, n_steps, ipw);
  else if (mode == 1) hipLaunchKernelGGL((force_kernel<U, 1, GT, C, BLK>), grid, block, lds_bytes, st, steps, blk_prefix, blk_step, n_steps, ipw);
  else hipLaunchKernelGGL((force_kernel<U, 2, GT, C, BLK>), grid, block, lds_bytes, st, steps, blk_prefix, blk_step, n_steps, ipw);
}

template <int U, int GT, int C, bool BLK>
static void launch_force1_g(dim3 grid, dim3 block, size_t lds_bytes, hipStream_t st, int mode,
                            const StepDesc &sd) {
  if (mode == 1) hipLaunchKernelGGL((force_kernel1<U, 1, GT, C, BLK>), grid, block, lds_bytes, st, sd);
  else hipLaunchKernelGGL((force_kernel1<U, 0, GT, C, BLK>), grid, block, lds_bytes, st, sd);
}

// Compile-time G for the common widths: d = 32, 64, 128 in slices of 8 and
// d = 64, 128 in slices of 16. Launches without block items over slices of
// 8 and up to 4 rows in flight take the kernel without the block-item path.
template <int U, int GT>
static void launch_force1_b(dim3 grid, dim3 block, size_t lds_bytes, hipStream_t st, int mode,
                            bool blk, const StepDesc &sd) {
  if constexpr (U <= 4) {
    if (!blk) {
      launch_force1_g<U, GT, 8, false>(grid, block, lds_bytes, st, mode, sd);
      return;
    }
  }
  launch_force1_g<U, GT, 8, true>(grid, block, lds_bytes, st, mode, sd);
}

template <int U>
static void launch_force1_u(dim3 grid, dim3 block, size_t lds_bytes, hipStream_t st, int mode,
                            int g_const, int slice, bool blk, const StepDesc &sd) {
  if constexpr (U <= 4) {
    if (slice == 16) {
      if (g_const == 8) launch_force1_g<U, 8, 16, true>(grid, block, lds_bytes, st, mode, sd);
      else if (g_const == 4) launch_force1_g<U, 4, 16, true>(grid, block, lds_bytes, st, mode, sd);
      else launch_force1_g<U, 0, 16, true>(grid, block, lds_bytes, st, mode, sd);
      return;
    }
  }
  if (g_const == 16) launch_force1_b<U, 16>(grid, block, lds_bytes, st, mode, blk, sd);
  else if (g_const == 8) launch_force1_b<U, 8>(grid, block, lds_bytes, st, mode, blk, sd);
  else if (g_const == 4) launch_force1_b<U, 4>(grid, block, lds_bytes, st, mode, blk, sd);
  else launch_force1_b<U, 0>(grid, block, lds_bytes, st, mode, blk, sd);
}

// Slices of 16 keep at most 4 rows in flight per lane group (8 would spill);
// the unroll never changes results.
hipError_t launch_force1(const StepDesc &sd, int n_blocks, int wpb, int unroll, size_t lds_bytes,
                         int mode, int g_const, int slice, hipStream_t st) {
  dim3 grid(n_blocks), block(64 * wpb);
  const bool blk = sd.n_block_items > 0;
  switch (unroll) {
  case 1: launch_force1_u<1>(grid, block, lds_bytes, st, mode, g_const, slice, blk, sd); break;
  case 2: launch_force1_u<2>(grid, block, lds_bytes, st, mode, g_const, slice, blk, sd); break;
  case 8:
    if (slice == 16) launch_force1_u<4>(grid, block, lds_bytes, st, mode, g_const, slice, blk, sd);
    else launch_force1_u<8>(grid, block, lds_bytes, st, mode, g_const, slice, blk, sd);
    break;
  default: launch_force1_u<4>(grid, block, lds_bytes, st, mode, g_const, slice, blk, sd); break;
  }
  return hipGetLastError();
}

template <int U, int GT>
static void launch_force_b(dim3 grid, dim3 block, size_t lds_bytes, hipStream_t st, int mode,
                           bool blk, const StepDesc *steps, const int *blk_prefix,
                           const int *blk_step, int n_steps, int ipw) {
  if constexpr (U <= 4) {
    if (!blk) {
      launch_force_g<U, GT, 8, false>(grid, block, lds_bytes, st, mode, steps, blk_prefix, blk_step, n_steps, ipw);
      return;
    }
  }
  launch_force_g<U, GT, 8, true>(grid, block, lds_bytes, st, mode, steps, blk_prefix, blk_step, n_steps, ipw);
}

template <int U>
static void launch_force_u(dim3 grid, dim3 block, size_t lds_bytes, hipStream_t st, int mode,
                           int g_const, int slice, bool blk, const StepDesc *steps,
                           const int *blk_prefix, const int *blk_step, int n_steps, int ipw) {
  // compile-time G for the common widths d = 32, 64, 128 (FP32 or BF16):
  // slices of 8, or d = 64, 128 in slices of 16 (2 or 4 rows in flight)
  if constexpr (U == 2 || U == 4) {
    if (slice == 16) {
      if (mode != 2 && g_const == 8) launch_force_g<U, 8, 16, true>(grid, block, lds_bytes, st, mode, steps, blk_prefix, blk_step, n_steps, ipw);
      else if (mode != 2 && g_const == 4) launch_force_g<U, 4, 16, true>(grid, block, lds_bytes, st, mode, steps, blk_prefix, blk_step, n_steps, ipw);
      else launch_force_g<U, 0, 16, true>(grid, block, lds_bytes, st, mode, steps, blk_prefix, blk_step, n_steps, ipw);
      return;
    }
  }
  if (mode != 2 && g_const == 16) launch_force_b<U, 16>(grid, block, lds_bytes, st, mode, blk, steps, blk_prefix, blk_step, n_steps, ipw);
  else if (mode != 2 && g_const == 8) launch_force_b<U, 8>(grid, block, lds_bytes, st, mode, blk, steps, blk_prefix, blk_step, n_steps, ipw);
  else if (mode != 2 && g_const == 4) launch_force_b<U, 4>(grid, block, lds_bytes, st, mode, blk, steps, blk_prefix, blk_step, n_steps, ipw);
  else launch_force_b<U, 0>(grid, block, lds_bytes, st, mode, blk, steps, blk_prefix, blk_step, n_steps, ipw);
}

// Engine rounds in slices of 16 run 2 or 4 rows in flight (1 -> 2, 8 -> 4;
// the unroll never changes results).
hipError_t launch_force(const StepDesc *steps, const int *blk_prefix, const int *blk_step,
                        int n_steps, int n_blocks, int wpb, int ipw, int unroll, size_t lds_bytes,
                        int mode, int g_const, int slice, bool blk, hipStream_t st) {
  dim3 grid(n_blocks), block(64 * wpb);
  if (slice == 16) unroll = unroll <= 2 ? 2 : 4;
  switch (unroll) {
  case 1: launch_force_u<1>(grid, block, lds_bytes, st, mode, g_const, slice, blk, steps, blk_prefix, blk_step, n_steps, ipw); break;
  case 2: launch_force_u<2>(grid, block, lds_bytes, st, mode, g_const, slice, blk, steps, blk_prefix, blk_step, n_steps, ipw); break;
  case 8: launch_force_u<8>(grid, block, lds_bytes, st, mode, g_const, slice, blk, steps, blk_prefix, blk_step, n_steps, ipw); break;
  default: launch_force_u<4>(grid, block, lds_bytes, st, mode, g_const, slice, blk, steps, blk_prefix, blk_step, n_steps, ipw); break;
  }
  return hipGetLastError();
}

hipError_t launch_epoch(const EpochDesc *jobs, int n, int set, int value, int rot,
                        hipStream_t st) {
  hipLaunchKernelGGL(epoch_kernel, dim3(n), dim3(256), 0, st, jobs, set, value, rot);
  return hipGetLastError();
}

// A small round buffer, passed by value and written to device memory by one
// block: in stream order in front of the round's launch, it costs a kernel
// boundary instead of a DMA copy and its wait.
struct RoundBlob {
  u4v w[kRoundBlobBytes / 16];
};

__global__ __launch_bounds__(256) void blob_kernel(const RoundBlob b, int n_vec, u4v *dst) {
  for (int i = threadIdx.x; i < n_vec; i += blockDim.x) dst[i] = b.w[i];
}

hipError_t launch_blob(const void *src, size_t bytes, void *dst, hipStream_t st) {
  if (bytes > (size_t)kRoundBlobBytes || bytes % 16 != 0) return hipErrorInvalidValue;
  RoundBlob b;
  std::memcpy(&b, src, bytes);
  hipLaunchKernelGGL(blob_kernel, dim3(1), dim3(256), 0, st, b, (int)(bytes / 16),
                     static_cast<u4v *>(dst));
  return hipGetLastError();
}

// One job's epoch state {epoch, rot, shift} := the given values.
__global__ void state_kernel(int *state, int epoch, int rot, int shift) {
  if (threadIdx.x == 0) {
    state[0] = epoch;
    state[1] = rot;
    state[2] = shift;
  }
}

hipError_t launch_state(int *state, int epoch, int rot, int shift, hipStream_t st) {
  hipLaunchKernelGGL(state_kernel, dim3(1), dim3(64), 0, st, state, epoch, rot, shift);
  return hipGetLastError();
}

template <int U, int MODE>
static void launch_spmm_m(dim3 grid, dim3 block, size_t lds_bytes, hipStream_t st, int g_const,
                          const StepDesc *steps, const int *blk_prefix, int n_steps, float *Y) {
  if (g_const == 16) hipLaunchKernelGGL((spmm_kernel<U, MODE, 16>), grid, block, lds_bytes, st, steps, blk_prefix, n_steps, Y);
  else hipLaunchKernelGGL((spmm_kernel<U, MODE, 0>), grid, block, lds_bytes, st, steps, blk_prefix, n_steps, Y);
}

hipError_t launch_spmm(const StepDesc *steps, const int *blk_prefix, int n_steps, int n_blocks,
                       int wpb, int unroll, int mode, int g_const, float *Y, hipStream_t st) {
  dim3 grid(n_blocks), block(64 * wpb);
  const size_t lds = (size_t)kSub * 512 * sizeof(float); // kSub sub-range sums of d <= 512
  const bool bf = mode == 1;
  if (unroll <= 2) {
    if (bf) launch_spmm_m<2, 1>(grid, block, lds, st, g_const, steps, blk_prefix, n_steps, Y);
    else launch_spmm_m<2, 0>(grid, block, lds, st, g_const, steps, blk_prefix, n_steps, Y);
  } else {
    if (bf) launch_spmm_m<4, 1>(grid, block, lds, st, g_const, steps, blk_prefix, n_steps, Y);
    else launch_spmm_m<4, 0>(grid, block, lds, st, g_const, steps, blk_prefix, n_steps, Y);
  }
  return hipGetLastError();
}

hipError_t launch_apply(const StepDesc *steps, const int *blk_prefix, int n_steps,
                        int n_blocks, hipStream_t st) {
  hipLaunchKernelGGL(apply_kernel, dim3(n_blocks), dim3(256), 0, st, steps, blk_prefix, n_steps);
  return hipGetLastError();
}

hipError_t launch_fill_lcols(const int32_t *items, int n_items, const int32_t *col,
                             int32_t *lcols, hipStream_t st) {
  const int64_t total = (int64_t)n_items * kLightCols;
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL(fill_lcols_kernel, dim3((unsigned)blocks), dim3(256), 0, st, items, n_items,
                     col, lcols);
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

hipError_t launch_zero_ints(int *p, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int blocks = std::max(1, std::min((n + 255) / 256, 1024));
  hipLaunchKernelGGL(zero_ints_kernel, dim3(blocks), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

} // namespace fe
