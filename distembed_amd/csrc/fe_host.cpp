// Host-side helpers of the force-embedding library: the reference negative
// sampler, graph generators, CSR construction, MatrixMarket ingest and the
// embedding writer. Nothing here touches the GPU.
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>

#include "force_embed.h"
#include "fe_host.hpp"

namespace fe {

thread_local std::string g_host_error;

void set_host_error(const std::string &s) { g_host_error = s; }
const char *host_error() { return g_host_error.c_str(); }

uint64_t splitmix64_next(uint64_t &state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t splitmix64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---- reference sampler (cpp/core/common.cpp:37-53) --------------------------
//
// std::minstd_rand is the linear congruential engine x <- 48271 x mod (2^31-1)
// (C++ [rand.predef]); constructing it with seed s sets x = s mod m, or 1 if
// that is 0. Its outputs lie in [1, 2^31-2]. libstdc++'s
// uniform_int_distribution<int>(lo, hi) downscales such a generator, whose
// range (2^31-3) is neither 2^32-1 nor 2^64-1, by rejection:
//   urange = hi - lo, erange = urange + 1, scaling = (2^31-3) / erange,
//   past = erange * scaling; draw ret = x - 1 until ret < past;
//   return ret / scaling + lo.
// A CPU test checks this restatement against std:: compiled with g++.
struct MinStd {
  uint64_t x;
  explicit MinStd(uint64_t seed) {
    x = seed % 2147483647ull;
    if (x == 0) x = 1;
  }
  uint64_t next() {
    x = (x * 48271ull) % 2147483647ull;
    return x;
  }
};

static inline int64_t minstd_uniform(MinStd &g, int64_t lo, int64_t hi) {
  const uint64_t urngrange = 2147483646ull - 1ull;
  const uint64_t urange = (uint64_t)(hi - lo);
  const uint64_t erange = urange + 1;
  const uint64_t scaling = urngrange / erange;
  const uint64_t past = erange * scaling;
  uint64_t ret;
  do {
    ret = g.next() - 1ull;
  } while (ret >= past);
  return (int64_t)(ret / scaling) + lo;
}

} // namespace fe

using namespace fe;

extern "C" fe_status fe_reference_negatives(int64_t n, int n_lists, int ns, int policy,
                                            int32_t *out, int64_t *n_hits) {
  if (n <= 0 || n >= 2147483647LL || n_lists < 0 || ns < 0 || (n_lists > 0 && ns > 0 && !out) ||
      (policy != FE_NEG_STRICT && policy != FE_NEG_REDRAW)) {
    set_host_error("fe_reference_negatives: invalid argument");
    return FE_ERR_INVALID_ARG;
  }
  // The reference range [0, n] must fit uniform_int_distribution<int>'s
  // downscaling branch: n < 2^31 - 3.
  if (n + 1 > 2147483645LL) {
    set_host_error("fe_reference_negatives: n too large for the reference sampler");
    return FE_ERR_INVALID_ARG;
  }
  int64_t hits = 0;
  for (int l = 0; l < n_lists; ++l) {
    MinStd g((uint64_t)(uint32_t)l);
    for (int k = 0; k < ns; ++k) {
      int64_t v = minstd_uniform(g, 0, n);
      if (v == n) {
        ++hits;
        if (policy == FE_NEG_REDRAW) {
          while (v == n) v = minstd_uniform(g, 0, n);
        }
      }
      out[(size_t)l * ns + k] = (int32_t)v;
    }
  }
  if (n_hits) *n_hits = hits;
  if (hits > 0 && policy == FE_NEG_STRICT) {
    set_host_error("fe_reference_negatives: a draw equals n (out of range) under FE_NEG_STRICT");
    return FE_ERR_INVALID_ARG;
  }
  return FE_OK;
}

extern "C" int64_t fe_eval_negative(uint64_t seed, int64_t r, int k, int64_t n) {
  if (n <= 0) return 0;
  const uint64_t x = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)r * 1048576ull + (uint64_t)k + 1ull);
  return (int64_t)(splitmix64_mix(x) % (uint64_t)n);
}

// ---- CSR construction ----------------------------------------------------------

namespace fe {

// Builds the CSR of the directed entries (u[i], v[i]); with `symmetrize`
// every entry is also added reversed. Columns are sorted and duplicates
// merged; self-loops are dropped when `drop_self`.
fe_status build_csr(int64_t n, const int64_t *u, const int64_t *v, int64_t m, bool symmetrize,
                    bool drop_self, fe_graph *out) {
  if (n <= 0 || n >= 2147483647LL || m < 0 || !out || (m > 0 && (!u || !v))) {
    set_host_error("build_csr: invalid argument");
    return FE_ERR_INVALID_ARG;
  }
  for (int64_t i = 0; i < m; ++i) {
    if (u[i] < 0 || u[i] >= n || v[i] < 0 || v[i] >= n) {
      set_host_error("build_csr: vertex id out of range");
      return FE_ERR_INVALID_ARG;
    }
  }
  std::vector<int64_t> deg(n + 1, 0);
  for (int64_t i = 0; i < m; ++i) {
    if (drop_self && u[i] == v[i]) continue;
    deg[u[i]]++;
    if (symmetrize) deg[v[i]]++;
  }
  std::vector<int64_t> start(n + 1, 0);
  for (int64_t r = 0; r < n; ++r) start[r + 1] = start[r] + deg[r];
  const int64_t total = start[n];
  std::vector<int32_t> tmp;
  try {
    tmp.resize((size_t)total);
  } catch (...) {
    set_host_error("build_csr: out of host memory");
    return FE_ERR_OOM;
  }
  std::vector<int64_t> pos(start.begin(), start.end() - 1);
  for (int64_t i = 0; i < m; ++i) {
    if (drop_self && u[i] == v[i]) continue;
    tmp[pos[u[i]]++] = (int32_t)v[i];
    if (symmetrize) tmp[pos[v[i]]++] = (int32_t)u[i];
  }
  std::vector<int64_t> newdeg(n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t r = 0; r < n; ++r) {
    int32_t *b = tmp.data() + start[r];
    int32_t *e = tmp.data() + start[r + 1];
    std::sort(b, e);
    newdeg[r] = std::unique(b, e) - b;
  }
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  if (!rp) {
    set_host_error("build_csr: out of host memory");
    return FE_ERR_OOM;
  }
  rp[0] = 0;
  for (int64_t r = 0; r < n; ++r) rp[r + 1] = rp[r] + newdeg[r];
  const int64_t nnz = rp[n];
  if (nnz >= 2147483647LL) {
    free(rp);
    set_host_error("build_csr: more than 2^31-1 entries");
    return FE_ERR_INVALID_ARG;
  }
  int32_t *ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
  if (!ci) {
    free(rp);
    set_host_error("build_csr: out of host memory");
    return FE_ERR_OOM;
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t r = 0; r < n; ++r)
    std::memcpy(ci + rp[r], tmp.data() + start[r], sizeof(int32_t) * (size_t)newdeg[r]);
  out->n = n;
  out->nnz = nnz;
  out->row_ptr = rp;
  out->col_idx = ci;
  return FE_OK;
}

} // namespace fe

extern "C" fe_status fe_csr_from_edges(int64_t n, const int64_t *u, const int64_t *v, int64_t m,
                                       fe_graph *out) {
  return build_csr(n, u, v, m, true, true, out);
}

extern "C" void fe_free_graph(fe_graph *g) {
  if (!g) return;
  free(g->row_ptr);
  free(g->col_idx);
  g->row_ptr = nullptr;
  g->col_idx = nullptr;
  g->n = g->nnz = 0;
}

// ---- generators ---------------------------------------------------------------

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004): each of edge_factor * 2^scale
// edges descends `scale` levels of the adjacency matrix, choosing the
// quadrant (row bit, col bit) = (0,0) w.p. a, (0,1) w.p. b, (1,0) w.p. c,
// (1,1) w.p. 1-a-b-c, most significant bit first. Edge e draws its
// uniforms from splitmix64 started at seed-derived state + e, so the output
// does not depend on the thread count. No vertex relabelling.
extern "C" fe_status fe_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                                 uint64_t seed, fe_graph *out) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0 || !out) {
    set_host_error("fe_gen_rmat: invalid argument");
    return FE_ERR_INVALID_ARG;
  }
  const int64_t n = (int64_t)1 << scale;
  const int64_t m = (int64_t)edge_factor * n;
  std::vector<int64_t> u, v;
  try {
    u.resize((size_t)m);
    v.resize((size_t)m);
  } catch (...) {
    set_host_error("fe_gen_rmat: out of host memory");
    return FE_ERR_OOM;
  }
  const double two32 = 4294967296.0;
  const uint64_t ta = (uint64_t)(a * two32);
  const uint64_t tab = (uint64_t)((a + b) * two32);
  const uint64_t tabc = (uint64_t)((a + b + c) * two32);
  const uint64_t base = splitmix64_mix(seed ^ 0x52A7A11C0FFEEull);
#pragma omp parallel for schedule(static, 65536)
  for (int64_t e = 0; e < m; ++e) {
    uint64_t st = base + (uint64_t)e * 0xD1B54A32D192ED03ull;
    int64_t row = 0, col = 0;
    uint64_t word = 0;
    for (int lv = 0; lv < scale; ++lv) {
      uint64_t r32;
      if ((lv & 1) == 0) {
        word = splitmix64_next(st);
        r32 = word & 0xFFFFFFFFull;
      } else {
        r32 = word >> 32;
      }
      const int64_t bit = (int64_t)1 << (scale - 1 - lv);
      if (r32 < ta) {
      } else if (r32 < tab) {
        col |= bit;
      } else if (r32 < tabc) {
        row |= bit;
      } else {
        row |= bit;
        col |= bit;
      }
    }
    u[e] = row;
    v[e] = col;
  }
  return build_csr(n, u.data(), v.data(), m, true, true, out);
}

// Stochastic block model with geometric skipping over the candidate pairs of
// each block pair (Batagelj and Brandes 2005): block pair (bi, bj), bi <= bj,
// draws its own splitmix64 stream from (seed, bi, bj).
extern "C" fe_status fe_gen_sbm(int64_t n, int blocks, double p_in, double p_out, uint64_t seed,
                                fe_graph *out) {
  if (n <= 0 || n >= 2147483647LL || blocks < 1 || blocks > n || p_in < 0 || p_in > 1 ||
      p_out < 0 || p_out > 1 || !out) {
    set_host_error("fe_gen_sbm: invalid argument");
    return FE_ERR_INVALID_ARG;
  }
  std::vector<int64_t> bstart(blocks + 1, 0);
  for (int i = 0; i < blocks; ++i)
    bstart[i + 1] = bstart[i] + n / blocks + (i < n % blocks ? 1 : 0);
  const int64_t npairs = (int64_t)blocks * (blocks + 1) / 2;
  // the expected edge count must fit the CSR (fewer than 2^31 entries)
  double expect = 0.0;
  for (int i = 0; i < blocks; ++i) {
    const double si = (double)(bstart[i + 1] - bstart[i]);
    expect += p_in * si * (si - 1) / 2 + p_out * si * (double)(n - bstart[i + 1]);
  }
  if (2 * expect >= 2147483647.0) {
    set_host_error("fe_gen_sbm: the expected edge count exceeds the CSR limit of 2^31-1 entries");
    return FE_ERR_INVALID_ARG;
  }
  std::vector<std::vector<int64_t>> eu(npairs), ev(npairs);
  int oom = 0; // an exception cannot leave the parallel region
#pragma omp parallel for schedule(dynamic, 1) reduction(| : oom)
  for (int64_t pidx = 0; pidx < npairs; ++pidx) {
    try {
      // decode pidx -> (bi, bj), bi <= bj
      int64_t bi = 0, rem = pidx;
      while (rem >= blocks - bi) {
        rem -= blocks - bi;
        ++bi;
      }
      const int64_t bj = bi + rem;
      const double p = (bi == bj) ? p_in : p_out;
      if (p <= 0.0) continue;
      const int64_t si = bstart[bi + 1] - bstart[bi];
      const int64_t sj = bstart[bj + 1] - bstart[bj];
      const int64_t cand = si * sj; // (i, j) grid; within a block only i < j is kept
      uint64_t st = splitmix64_mix(seed ^ 0x5B3C0DEull) ^ splitmix64_mix((uint64_t)bi * 1000003ull + (uint64_t)bj + 1);
      const double lq = (p < 1.0) ? std::log1p(-p) : 0.0;
      int64_t idx = -1;
      while (true) {
        if (p >= 1.0) {
          idx += 1;
        } else {
          const double uni = ((splitmix64_next(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
          const double skip = std::floor(std::log(uni) / lq);
          if (skip > (double)cand) break;
          idx += 1 + (int64_t)skip;
        }
        if (idx >= cand) break;
        const int64_t i = idx / sj, j = idx % sj;
        if (bi == bj && i >= j) continue;
        eu[pidx].push_back(bstart[bi] + i);
        ev[pidx].push_back(bstart[bj] + j);
      }
    } catch (...) {
      oom = 1;
    }
  }
  std::vector<int64_t> u, v;
  try {
    if (oom) throw std::bad_alloc();
    for (int64_t pidx = 0; pidx < npairs; ++pidx) {
      u.insert(u.end(), eu[pidx].begin(), eu[pidx].end());
      v.insert(v.end(), ev[pidx].begin(), ev[pidx].end());
    }
  } catch (...) {
    set_host_error("fe_gen_sbm: out of host memory");
    return FE_ERR_OOM;
  }
  return build_csr(n, u.data(), v.data(), (int64_t)u.size(), true, true, out);
}

extern "C" fe_status fe_init_uniform(float *x, int64_t count, float lo, float hi, uint64_t seed) {
  if (count < 0 || (count > 0 && !x) || !(hi >= lo)) {
    set_host_error("fe_init_uniform: invalid argument");
    return FE_ERR_INVALID_ARG;
  }
  const uint64_t base = splitmix64_mix(seed ^ 0x1A17C0DEull);
  const double span = (double)hi - (double)lo;
#pragma omp parallel for schedule(static, 65536)
  for (int64_t i = 0; i < count; ++i) {
    const uint64_t z = splitmix64_mix(base + (uint64_t)i * 0x9E3779B97F4A7C15ull);
    const double u01 = (double)(z >> 40) * (1.0 / 16777216.0); // 24-bit uniform in [0,1)
    x[i] = (float)((double)lo + span * u01);
  }
  return FE_OK;
}

// ---- MatrixMarket ingest (cpp/io/parrallel_IO.hpp:31-88) ------------------------
//
// The reference reads through CombBLAS ParallelReadMM(file, onebased=true,
// maximum<>): 1-based ids, duplicates merged. A `symmetric` header stores
// one triangle and is expanded to both directions; `general` entries are
// kept as given. Values are ignored (the reference replaces them with the
// column id, :64-75, and the embedding path never reads them). Self-loops
// are kept: they add a zero force, as in the reference.
namespace {

// One MatrixMarket entry line "i j [value...]" at p (line end excluded):
// 1-based i and j as sscanf("%lld %lld") reads them (leading blanks, an
// optional sign, digits, blanks between the two). False if they do not parse.
bool parse_entry(const char *p, const char *end, long long &i, long long &j) {
  auto blank = [](char c) { return c == ' ' || c == '\t' || c == '\v' || c == '\f' || c == '\r'; };
  auto number = [&](long long &out) {
    while (p < end && blank(*p)) ++p;
    bool neg = false;
    if (p < end && (*p == '+' || *p == '-')) neg = *p++ == '-';
    if (p >= end || *p < '0' || *p > '9') return false;
    long long x = 0;
    while (p < end && *p >= '0' && *p <= '9') {
      if (x > (LLONG_MAX - 9) / 10) return false;
      x = x * 10 + (*p++ - '0');
    }
    out = neg ? -x : x;
    return true;
  };
  return number(i) && number(j);
}

} // namespace

// The entry lines of a MatrixMarket file in buf[off, size): comment lines
// ('%') and lines starting with a line end are skipped, the first nnz entries
// are kept (in file order, 0-based), an unparsable or out-of-range line
// before them is an error, and lines after them are not read, as a serial
// line-by-line reader would. Chunks of whole lines are parsed in parallel.
static fe_status parse_mtx_entries(const char *buf, size_t off, size_t size, long long rows,
                                   long long cols, long long nnz, std::vector<int64_t> &u,
                                   std::vector<int64_t> &v) {
  const size_t len = size > off && buf ? size - off : 0;
  if (len == 0) {
    if (nnz == 0) return FE_OK;
    set_host_error("fe_read_mtx: fewer entries than announced");
    return FE_ERR_INVALID_ARG;
  }
  const char *base = buf + off;
  const int nchunks = (int)std::max<size_t>(1, std::min<size_t>(len / (1 << 20) + 1,
                                                                 (size_t)omp_get_max_threads() * 8));
  std::vector<size_t> cut((size_t)nchunks + 1, len);
  cut[0] = 0;
  for (int c = 1; c < nchunks; ++c) {
    size_t p = std::max(cut[c - 1], len / nchunks * (size_t)c);
    while (p < len && base[p - 1] != '\n') ++p;
    cut[c] = p;
  }
  struct Chunk {
    std::vector<int64_t> u, v;
    bool bad = false; // stopped at a bad line (its entries are those before it)
  };
  std::vector<Chunk> ch((size_t)nchunks);
  int oom = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : oom)
  for (int c = 0; c < nchunks; ++c) {
    try {
      Chunk &k = ch[(size_t)c];
      const char *p = base + cut[c], *e = base + cut[c + 1];
      k.u.reserve((size_t)(e - p) / 8);
      k.v.reserve((size_t)(e - p) / 8);
      while (p < e) {
        const char *nl = static_cast<const char *>(memchr(p, '\n', (size_t)(e - p)));
        const char *le = nl ? nl : e;
        if (*p != '%' && *p != '\n' && *p != '\r') {
          long long i, j;
          if (!parse_entry(p, le, i, j) || i < 1 || j < 1 || i > rows || j > cols) {
            k.bad = true;
            break;
          }
          k.u.push_back(i - 1);
          k.v.push_back(j - 1);
        }
        p = nl ? nl + 1 : e;
      }
    } catch (...) {
      oom = 1;
    }
  }
  if (oom) {
    set_host_error("fe_read_mtx: out of host memory");
    return FE_ERR_OOM;
  }
  long long total = 0;
  try {
    u.reserve((size_t)std::min<long long>(nnz, 1LL << 24));
    v.reserve((size_t)std::min<long long>(nnz, 1LL << 24));
    for (const Chunk &k : ch) {
      const long long take = std::min<long long>((long long)k.u.size(), nnz - total);
      u.insert(u.end(), k.u.begin(), k.u.begin() + take);
      v.insert(v.end(), k.v.begin(), k.v.begin() + take);
      total += take;
      if (total == nnz) break;
      if (k.bad) {
        set_host_error("fe_read_mtx: bad entry line");
        return FE_ERR_INVALID_ARG;
      }
    }
  } catch (...) {
    set_host_error("fe_read_mtx: out of host memory");
    return FE_ERR_OOM;
  }
  if (total != nnz) {
    set_host_error("fe_read_mtx: fewer entries than announced");
    return FE_ERR_INVALID_ARG;
  }
  return FE_OK;
}

extern "C" fe_status fe_read_mtx(const char *path, fe_graph *out) {
  if (!path || !out) {
    set_host_error("fe_read_mtx: invalid argument");
    return FE_ERR_INVALID_ARG;
  }
  FILE *f = fopen(path, "r");
  if (!f) {
    set_host_error(std::string("fe_read_mtx: cannot open ") + path + ": " + strerror(errno));
    return FE_ERR_INVALID_ARG;
  }
  char line[4096];
  if (!fgets(line, sizeof line, f)) {
    fclose(f);
    set_host_error("fe_read_mtx: empty file");
    return FE_ERR_INVALID_ARG;
  }
  std::string hdr(line);
  for (auto &ch : hdr) ch = (char)tolower((unsigned char)ch);
  if (hdr.rfind("%%matrixmarket", 0) != 0 || hdr.find("coordinate") == std::string::npos) {
    fclose(f);
    set_host_error("fe_read_mtx: not a MatrixMarket coordinate file");
    return FE_ERR_INVALID_ARG;
  }
  const bool symmetric = hdr.find("symmetric") != std::string::npos ||
                         hdr.find("hermitian") != std::string::npos;
  if (hdr.find("skew-symmetric") != std::string::npos) {
    fclose(f);
    set_host_error("fe_read_mtx: skew-symmetric matrices are not supported");
    return FE_ERR_INVALID_ARG;
  }
  long long rows = -1, cols = -1, nnz = -1;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '%' || line[0] == '\n' || line[0] == '\r') continue;
    if (sscanf(line, "%lld %lld %lld", &rows, &cols, &nnz) != 3) {
      fclose(f);
      set_host_error("fe_read_mtx: bad size line");
      return FE_ERR_INVALID_ARG;
    }
    break;
  }
  if (rows <= 0 || rows != cols || nnz < 0 || rows >= 2147483647LL) {
    fclose(f);
    set_host_error("fe_read_mtx: the matrix must be square, with 1 to 2^31-2 rows and a valid "
                   "size line");
    return FE_ERR_INVALID_ARG;
  }
  // the entries: the rest of the file, memory-mapped and parsed in parallel
  // chunks cut at line ends, merged in file order
  const long data_off = ftell(f);
  struct stat stt;
  if (data_off < 0 || fstat(fileno(f), &stt) != 0) {
    fclose(f);
    set_host_error("fe_read_mtx: cannot size the file");
    return FE_ERR_INVALID_ARG;
  }
  const size_t fsize = (size_t)stt.st_size;
  const char *map = nullptr;
  if (fsize > (size_t)data_off) {
    void *m = mmap(nullptr, fsize, PROT_READ, MAP_PRIVATE, fileno(f), 0);
    if (m == MAP_FAILED) {
      fclose(f);
      set_host_error(std::string("fe_read_mtx: cannot map the file: ") + strerror(errno));
      return FE_ERR_INVALID_ARG;
    }
    map = static_cast<const char *>(m);
  }
  fclose(f);
  std::vector<int64_t> u, v;
  fe_status st = parse_mtx_entries(map, (size_t)data_off, fsize, rows, cols, nnz, u, v);
  if (map) munmap(const_cast<char *>(map), fsize);
  if (st != FE_OK) return st;
  if (symmetric) return build_csr(rows, u.data(), v.data(), (int64_t)u.size(), true, false, out);
  return build_csr(rows, u.data(), v.data(), (int64_t)u.size(), false, false, out);
}

// Embedding writer (cpp/io/parrallel_IO.hpp:90-146 at world size 1): one
// line per row, "%lu" (1-based id) then " %.5f" per component. Rows are
// formatted in parallel, a group of chunks at a time, and written in order,
// so the file is byte for byte what one formatted stream gives.
extern "C" fe_status fe_write_embedding(const char *path, const float *x, int64_t n, int d) {
  if (!path || (!x && n > 0) || n < 0 || d <= 0) {
    set_host_error("fe_write_embedding: invalid argument");
    return FE_ERR_INVALID_ARG;
  }
  FILE *f = fopen(path, "w");
  if (!f) {
    set_host_error(std::string("fe_write_embedding: cannot open ") + path);
    return FE_ERR_INVALID_ARG;
  }
  constexpr int64_t kRows = 512;   // rows per chunk
  constexpr int kGroup = 64;       // chunks formatted before a write
  const int64_t chunks = (n + kRows - 1) / kRows;
  std::vector<std::string> buf((size_t)std::min<int64_t>(chunks, kGroup));
  bool ok = true;
  int oom = 0; // an exception cannot leave the parallel region
  for (int64_t c0 = 0; c0 < chunks && ok && !oom; c0 += kGroup) {
    const int m = (int)std::min<int64_t>(kGroup, chunks - c0);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : oom)
    for (int k = 0; k < m; ++k) {
      try {
        std::string &s = buf[(size_t)k];
        s.clear();
        char tmp[64]; // " %.5f" of the largest float takes 47 characters
        const int64_t r0 = (c0 + k) * kRows, r1 = std::min(n, r0 + kRows);
        s.reserve((size_t)(r1 - r0) * ((size_t)d * 10 + 9)); // " -x.xxxxx" per component
        for (int64_t i = r0; i < r1; ++i) {
          int len = snprintf(tmp, sizeof tmp, "%lu", (unsigned long)(i + 1));
          s.append(tmp, (size_t)len);
          for (int j = 0; j < d; ++j) {
            len = snprintf(tmp, sizeof tmp, " %.5f", (double)x[i * d + j]);
            s.append(tmp, (size_t)std::min<int>(len, (int)sizeof tmp - 1));
          }
          s.push_back('\n');
        }
      } catch (...) {
        oom = 1;
      }
    }
    for (int k = 0; k < m && ok && !oom; ++k)
      ok = fwrite(buf[(size_t)k].data(), 1, buf[(size_t)k].size(), f) == buf[(size_t)k].size();
  }
  ok = (fclose(f) == 0) && ok;
  if (oom) {
    set_host_error("fe_write_embedding: out of host memory");
    return FE_ERR_OOM;
  }
  if (!ok) {
    set_host_error("fe_write_embedding: write failed");
    return FE_ERR_INVALID_ARG;
  }
  return FE_OK;
}
