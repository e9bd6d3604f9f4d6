/*
 * oracle.c -- CPU restatement of SYgraph's traversal algorithms and of the
 * synthetic graph generators used by the tests and the benchmark.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker (or
 * the timed CPU baseline).  The product never calls into it.
 *
 * Parity pins (see tests/test_oracle.py): BFS/SSSP/CC/TC/BC are checked against
 * the reference's own known answers (tests/algorithms/{bfs,sssp,tc}.cpp), networkx/scipy on
 * small graphs, closed forms (paths, stars, complete graphs, grids) and brute
 * force.  The generators are pinned by closed-form edge counts and by the
 * structural invariants of their output (symmetric, sorted, no duplicates, no
 * self loops) and by committed digests.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define GAMMA 0x9E3779B97F4A7C15ULL

uint64_t or_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Position ctr of the SplitMix64 stream whose state starts at mix64(seed+gamma). */
uint64_t or_rand64(uint64_t seed, uint64_t ctr) {
  uint64_t s0 = or_mix64(seed + GAMMA);
  return or_mix64(s0 + (ctr + 1) * GAMMA);
}

/* ------------------------------------------------------------------------- */
/* Generators                                                                */
/* ------------------------------------------------------------------------- */

/* Seeded bijection on scale-bit integers: three rounds of
 * (multiply by odd constant, xorshift right, add constant), all mod 2^scale. */
uint32_t or_rmat_relabel(uint32_t x0, int scale, uint64_t seed) {
  uint64_t mask = (scale >= 32) ? 0xFFFFFFFFULL : ((1ULL << scale) - 1ULL);
  uint64_t x = x0 & mask;
  int sh = (scale + 1) / 2;
  for (int k = 0; k < 3; ++k) {
    uint64_t mul = or_rand64(seed ^ 0x5EEDF00DULL, 2 * k) | 1ULL;
    uint64_t add = or_rand64(seed ^ 0x5EEDF00DULL, 2 * k + 1);
    x = (x * mul) & mask;
    x ^= (x >> sh);
    x = (x + add) & mask;
  }
  return (uint32_t)x;
}

void or_rmat_pairs(int scale, int64_t nedges, double a, double b, double c, uint64_t seed, int relabel, uint32_t* u,
                   uint32_t* v) {
  /* Quadrant thresholds on a 32-bit uniform draw. */
  const uint64_t two32 = 1ULL << 32;
  uint64_t ta = (uint64_t)(a * (double)two32);
  uint64_t tb = (uint64_t)((a + b) * (double)two32);
  uint64_t tc = (uint64_t)((a + b + c) * (double)two32);
  int draws = (scale + 1) / 2; /* two 32-bit choices per 64-bit draw */
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nedges; ++e) {
    uint32_t su = 0, sv = 0;
    for (int d = 0; d < draws; ++d) {
      uint64_t r = or_rand64(seed, (uint64_t)e * (uint64_t)draws + (uint64_t)d);
      for (int h = 0; h < 2; ++h) {
        int lvl = 2 * d + h;
        if (lvl >= scale) break;
        uint64_t r32 = h == 0 ? (r >> 32) : (r & 0xFFFFFFFFULL);
        uint32_t bu = 0, bv = 0;
        if (r32 < ta) {
        } else if (r32 < tb) {
          bv = 1;
        } else if (r32 < tc) {
          bu = 1;
        } else {
          bu = 1;
          bv = 1;
        }
        /* level 0 decides the most significant bit */
        su |= bu << (scale - 1 - lvl);
        sv |= bv << (scale - 1 - lvl);
      }
    }
    if (relabel) {
      su = or_rmat_relabel(su, scale, seed);
      sv = or_rmat_relabel(sv, scale, seed);
    }
    u[e] = su;
    v[e] = sv;
  }
}

int64_t or_grid_pairs(int rows, int cols, double p_delete, uint64_t seed, uint32_t* u, uint32_t* v) {
  /* Edge index: horizontal edges (r,c)-(r,c+1) first, idx = r*(cols-1)+c;
   * then vertical (r,c)-(r+1,c), idx = H + r*cols + c. */
  int64_t H = (int64_t)rows * (cols - 1);
  int64_t V = (int64_t)(rows - 1) * cols;
  int64_t total = H + V;
  uint64_t thr = (uint64_t)(p_delete * 18446744073709551616.0);
  if (p_delete >= 1.0) thr = UINT64_MAX;
  int64_t k = 0;
  for (int64_t idx = 0; idx < total; ++idx) {
    uint64_t r = or_rand64(seed, (uint64_t)idx);
    if (p_delete > 0.0 && r < thr) continue;
    uint32_t a, b;
    if (idx < H) {
      int64_t rr = idx / (cols - 1), cc = idx % (cols - 1);
      a = (uint32_t)(rr * cols + cc);
      b = a + 1;
    } else {
      int64_t j = idx - H;
      int64_t rr = j / cols, cc = j % cols;
      a = (uint32_t)(rr * cols + cc);
      b = a + (uint32_t)cols;
    }
    u[k] = a;
    v[k] = b;
    ++k;
  }
  return k;
}

static inline uint64_t mulhi64(uint64_t x, uint64_t n) { return (uint64_t)(((__uint128_t)x * (__uint128_t)n) >> 64); }

void or_er_pairs(int64_t n, int64_t m, uint64_t seed, uint32_t* u, uint32_t* v) {
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; ++e) {
    u[e] = (uint32_t)mulhi64(or_rand64(seed, 2 * (uint64_t)e), (uint64_t)n);
    v[e] = (uint32_t)mulhi64(or_rand64(seed, 2 * (uint64_t)e + 1), (uint64_t)n);
  }
}

/* LSD radix sort of 64-bit keys over the low `bits` bits, 8 bits per pass. */
static void radix_sort_u64(uint64_t* keys, uint64_t* tmp, int64_t n, int bits) {
  int nthreads = omp_get_max_threads();
  int64_t* hist = (int64_t*)malloc(sizeof(int64_t) * 256 * (size_t)nthreads);
  uint64_t* src = keys;
  uint64_t* dst = tmp;
  int passes = (bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * p;
    memset(hist, 0, sizeof(int64_t) * 256 * (size_t)nthreads);
#pragma omp parallel num_threads(nthreads)
    {
      int t = omp_get_thread_num();
      int nt = omp_get_num_threads();
      int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
      int64_t* h = hist + 256 * t;
      for (int64_t i = lo; i < hi; ++i) h[(src[i] >> shift) & 255]++;
#pragma omp barrier
#pragma omp single
      {
        int64_t sum = 0;
        for (int d = 0; d < 256; ++d)
          for (int tt = 0; tt < nt; ++tt) {
            int64_t cnt = hist[256 * tt + d];
            hist[256 * tt + d] = sum;
            sum += cnt;
          }
      }
      for (int64_t i = lo; i < hi; ++i) {
        uint64_t k = src[i];
        dst[h[(k >> shift) & 255]++] = k;
      }
    }
    uint64_t* s = src;
    src = dst;
    dst = s;
  }
  if (src != keys) memcpy(keys, src, sizeof(uint64_t) * (size_t)n);
  free(hist);
}

static int bits_for(int64_t n) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < n) ++b;
  return b;
}

int64_t or_build_csr(int64_t n, int64_t npairs, const uint32_t* u, const uint32_t* v, int64_t* offsets, int32_t* cols) {
  int b = bits_for(n);
  uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(2 * npairs + 1));
  int64_t k = 0;
  for (int64_t i = 0; i < npairs; ++i) {
    if (u[i] == v[i]) continue;
    keys[k++] = ((uint64_t)u[i] << b) | v[i];
    keys[k++] = ((uint64_t)v[i] << b) | u[i];
  }
  uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(k + 1));
  radix_sort_u64(keys, tmp, k, 2 * b);
  free(tmp);
  uint64_t mask = ((uint64_t)1 << b) - 1;
  for (int64_t i = 0; i <= n; ++i) offsets[i] = 0;
  int64_t m = 0;
  for (int64_t i = 0; i < k; ++i) {
    if (i > 0 && keys[i] == keys[i - 1]) continue;
    uint64_t s = keys[i] >> b;
    cols[m++] = (int32_t)(keys[i] & mask);
    offsets[s + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) offsets[i + 1] += offsets[i];
  free(keys);
  return m;
}

uint32_t or_edge_weight(uint32_t u, uint32_t v, uint64_t seed) {
  uint32_t lo = u < v ? u : v, hi = u < v ? v : u;
  uint64_t r = or_rand64(seed, ((uint64_t)lo << 32) | hi);
  return 1u + (uint32_t)(((r >> 32) * 255ULL) >> 32);
}

void or_fill_weights(int64_t n, const int64_t* offsets, const int32_t* cols, uint64_t seed, uint32_t* w) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t s = 0; s < n; ++s)
    for (int64_t e = offsets[s]; e < offsets[s + 1]; ++e) w[e] = or_edge_weight((uint32_t)s, (uint32_t)cols[e], seed);
}

uint64_t or_digest_i64(const int64_t* x, int64_t n) {
  uint64_t acc = 0;
#pragma omp parallel for reduction(+ : acc)
  for (int64_t i = 0; i < n; ++i) acc += or_mix64(((uint64_t)i * GAMMA) ^ (uint64_t)x[i]);
  return acc;
}

uint64_t or_digest_i32(const int32_t* x, int64_t n) {
  uint64_t acc = 0;
#pragma omp parallel for reduction(+ : acc)
  for (int64_t i = 0; i < n; ++i) acc += or_mix64(((uint64_t)i * GAMMA) ^ (uint64_t)(uint32_t)x[i]);
  return acc;
}

/* ------------------------------------------------------------------------- */
/* BFS: restates the validator of examples/bfs/bfs.cpp:17-56.                 */
/* ------------------------------------------------------------------------- */
int32_t or_bfs(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t src, int32_t* depth) {
  const int32_t unreached = (int32_t)(n + 1);
  for (int64_t i = 0; i < n; ++i) depth[i] = unreached;
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int64_t head = 0, tail = 0;
  depth[src] = 0;
  q[tail++] = src;
  int32_t levels = 0;
  while (head < tail) {
    int64_t level_end = tail;
    ++levels;
    for (; head < level_end; ++head) {
      int32_t x = q[head];
      for (int64_t e = offsets[x]; e < offsets[x + 1]; ++e) {
        int32_t y = cols[e];
        if (depth[y] == unreached) {
          depth[y] = depth[x] + 1;
          q[tail++] = y;
        }
      }
    }
  }
  free(q);
  return levels;
}

/* ------------------------------------------------------------------------- */
/* Dijkstra: restates examples/sssp/sssp.cpp:24-66 (binary heap, lazy delete). */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint32_t d;
  int32_t v;
} heap_item;

static void heap_push(heap_item* h, int64_t* sz, heap_item it) {
  int64_t i = (*sz)++;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h[p].d <= it.d) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = it;
}

static heap_item heap_pop(heap_item* h, int64_t* sz) {
  heap_item top = h[0];
  heap_item last = h[--(*sz)];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    uint32_t md = last.d;
    if (l < *sz && h[l].d < md) {
      m = l;
      md = h[l].d;
    }
    if (r < *sz && h[r].d < md) m = r;
    if (m == i) break;
    h[i] = h[m];
    i = m;
  }
  if (*sz > 0) h[i] = last;
  return top;
}

void or_sssp(int64_t n, const int64_t* offsets, const int32_t* cols, const uint32_t* w, int32_t src, uint32_t* dist) {
  const uint32_t unreached = (uint32_t)(n + 1);
  for (int64_t i = 0; i < n; ++i) dist[i] = unreached;
  int64_t cap = offsets[n] + 2;
  heap_item* h = (heap_item*)malloc(sizeof(heap_item) * (size_t)cap);
  int64_t sz = 0;
  dist[src] = 0;
  heap_item s0 = {0, src};
  heap_push(h, &sz, s0);
  while (sz > 0) {
    heap_item it = heap_pop(h, &sz);
    if (it.d > dist[it.v]) continue; /* stale entry */
    for (int64_t e = offsets[it.v]; e < offsets[it.v + 1]; ++e) {
      int32_t y = cols[e];
      uint32_t nd = it.d + w[e];
      if (nd < dist[y]) {
        dist[y] = nd;
        heap_item ni = {nd, y};
        heap_push(h, &sz, ni);
      }
    }
  }
  free(h);
}

/* ------------------------------------------------------------------------- */
/* CC: union-find; every vertex gets the largest id of its component, which is */
/* the fixpoint max-label propagation converges to (algorithms/cc.hpp:103-169). */
/* ------------------------------------------------------------------------- */
static int64_t uf_find(int64_t* p, int64_t x) {
  while (p[x] != x) {
    p[x] = p[p[x]];
    x = p[x];
  }
  return x;
}

void or_cc(int64_t n, const int64_t* offsets, const int32_t* cols, uint32_t* labels) {
  int64_t* p = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) p[i] = i;
  for (int64_t s = 0; s < n; ++s)
    for (int64_t e = offsets[s]; e < offsets[s + 1]; ++e) {
      int64_t a = uf_find(p, s), b = uf_find(p, cols[e]);
      if (a != b) {
        if (a < b) p[a] = b; /* root = larger id */
        else p[b] = a;
      }
    }
  for (int64_t i = 0; i < n; ++i) labels[i] = (uint32_t)uf_find(p, i);
  free(p);
}

/* ------------------------------------------------------------------------- */
/* TC: true triangle count.  Orient u->v iff (deg u, u) < (deg v, v); then     */
/* each triangle is counted once by merging N+(u) with N+(v) for v in N+(u).   */
/* Replaces the per-arc merge of algorithms/tc.hpp:80-107 (whose count is     */
/* incremented inside the merge loop) by the exact count.                      */
/* ------------------------------------------------------------------------- */
uint64_t or_tc(int64_t n, const int64_t* offsets, const int32_t* cols) {
  int64_t* ooff = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  ooff[0] = 0;
  for (int64_t s = 0; s < n; ++s) {
    int64_t ds = offsets[s + 1] - offsets[s], c = 0;
    for (int64_t e = offsets[s]; e < offsets[s + 1]; ++e) {
      int64_t t = cols[e], dt = offsets[t + 1] - offsets[t];
      if (ds < dt || (ds == dt && s < t)) ++c;
    }
    ooff[s + 1] = ooff[s] + c;
  }
  int32_t* oc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ooff[n] + 1));
  for (int64_t s = 0; s < n; ++s) {
    int64_t ds = offsets[s + 1] - offsets[s], k = ooff[s];
    for (int64_t e = offsets[s]; e < offsets[s + 1]; ++e) {
      int64_t t = cols[e], dt = offsets[t + 1] - offsets[t];
      if (ds < dt || (ds == dt && s < t)) oc[k++] = (int32_t)t;
    }
  }
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
  for (int64_t s = 0; s < n; ++s)
    for (int64_t e = ooff[s]; e < ooff[s + 1]; ++e) {
      int64_t t = oc[e];
      int64_t i = ooff[s], j = ooff[t];
      while (i < ooff[s + 1] && j < ooff[t + 1]) {
        if (oc[i] == oc[j]) {
          ++total;
          ++i;
          ++j;
        } else if (oc[i] < oc[j]) {
          ++i;
        } else {
          ++j;
        }
      }
    }
  free(ooff);
  free(oc);
  return total;
}

/* ------------------------------------------------------------------------- */
/* BC: Brandes (algorithms/bc.hpp:118-200): per source, a BFS computes depth  */
/* and path counts sigma, then dependencies                                   */
/*   delta(v) = sum_{w in N(v), d(w)=d(v)+1} sigma(v)/sigma(w) * (1+delta(w))  */
/* are accumulated from the deepest level up; bc(v) += delta(v) for v != s.   */
/* Path counts outgrow every floating-point range on long-diameter graphs     */
/* (about 2^8190 on a 4096x4096 grid) and differ by more than 2^1074 within   */
/* one level, so sigma is kept per vertex as m * 2^e (m in [0.5,1), e an      */
/* int64); sigma(v)/sigma(w) = m(v)/m(w) * 2^(e(v)-e(w)).  Powers of two are  */
/* exact, so this equals unscaled FP64 Brandes wherever that is finite.       */
/* Sources are taken in chunks of 16, each chunk summed in source order into  */
/* its own array, the arrays added in chunk order: the result does not depend */
/* on the number of threads.                                                  */
/* ------------------------------------------------------------------------- */
static void bc_one_source(int64_t n, const int64_t* offsets, const int32_t* cols, const int64_t* in_offsets,
                          const int32_t* in_cols, int32_t s, int32_t* depth, int32_t* order, int64_t* lvl_start,
                          double* sm, int64_t* se, double* dl, double* bc) {
  for (int64_t i = 0; i < n; ++i) {
    depth[i] = -1;
    dl[i] = 0.0;
  }
  int64_t head = 0, tail = 0;
  int32_t nl = 0;
  depth[s] = 0;
  sm[s] = 0.5;
  se[s] = 1;
  order[tail++] = s;
  lvl_start[0] = 0;
  while (head < tail) {
    int64_t lend = tail;
    lvl_start[++nl] = lend; /* level nl-1 occupies [lvl_start[nl-1], lend) */
    for (; head < lend; ++head) {
      int32_t x = order[head];
      for (int64_t e = offsets[x]; e < offsets[x + 1]; ++e) {
        int32_t y = cols[e];
        if (depth[y] < 0) {
          depth[y] = depth[x] + 1;
          order[tail++] = y;
        }
      }
    }
    /* sigma of the new level, pulled from the previous level in CSR order:
     * a running sum S * 2^E, rescaled when a larger exponent arrives */
    for (int64_t i = lend; i < tail; ++i) {
      int32_t y = order[i];
      double acc = 0.0;
      int64_t E = INT64_MIN;
      for (int64_t e = in_offsets[y]; e < in_offsets[y + 1]; ++e) {
        int32_t u = in_cols[e];
        if (depth[u] != depth[y] - 1) continue;
        if (se[u] > E) {
          acc = E == INT64_MIN ? 0.0 : ldexp(acc, (int)(E - se[u] < -2000 ? -2000 : E - se[u]));
          E = se[u];
        }
        acc += ldexp(sm[u], (int)(se[u] - E < -2000 ? -2000 : se[u] - E));
      }
      int ex = 0;
      sm[y] = frexp(acc, &ex);
      se[y] = E + ex;
    }
  }
  /* levels 0..nl-1; level L occupies [lvl_start[L], lvl_start[L+1]) */
  for (int32_t L = nl - 2; L >= 0; --L) {
    for (int64_t i = lvl_start[L]; i < lvl_start[L + 1]; ++i) {
      int32_t x = order[i];
      double acc = 0.0;
      for (int64_t e = offsets[x]; e < offsets[x + 1]; ++e) {
        int32_t y = cols[e];
        if (depth[y] == L + 1) {
          int64_t d = se[x] - se[y];
          acc += ldexp((1.0 + dl[y]) / sm[y], (int)(d < -2000 ? -2000 : d));
        }
      }
      dl[x] = sm[x] * acc;
    }
  }
  for (int64_t i = 0; i < tail; ++i)
    if (order[i] != s) bc[order[i]] += dl[order[i]];
}

void or_bc(int64_t n, const int64_t* offsets, const int32_t* cols, const int64_t* in_offsets, const int32_t* in_cols,
           int32_t nsrc, const int32_t* srcs, double* bc) {
  /* sigma is pulled over in-arcs; an undirected graph is its own inverse */
  if (!in_offsets) {
    in_offsets = offsets;
    in_cols = cols;
  }
  const int32_t kChunk = 16;
  const int32_t nchunks = (nsrc + kChunk - 1) / kChunk;
  double* part = (double*)calloc((size_t)(nchunks > 0 ? nchunks : 1) * (size_t)n, sizeof(double));
#pragma omp parallel
  {
    int32_t* depth = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int64_t* lvl_start = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 2));
    double* sm = (double*)malloc(sizeof(double) * (size_t)n);
    int64_t* se = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    double* dl = (double*)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(dynamic, 1)
    for (int32_t c = 0; c < nchunks; ++c)
      for (int32_t si = c * kChunk; si < nsrc && si < (c + 1) * kChunk; ++si)
        bc_one_source(n, offsets, cols, in_offsets, in_cols, srcs[si], depth, order, lvl_start, sm, se, dl,
                      part + (size_t)c * (size_t)n);
    free(depth);
    free(order);
    free(lvl_start);
    free(sm);
    free(se);
    free(dl);
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double acc = 0.0;
    for (int32_t c = 0; c < nchunks; ++c) acc += part[(size_t)c * (size_t)n + i];
    bc[i] = acc;
  }
  free(part);
}

/* Bitmap probes of one pull level (the byte model of the pull kernel).  Each
 * unvisited vertex u with in-arcs first probes hint[u] (when hint != NULL);
 * on a miss it scans its in-list in CSR order up to the first in-neighbour in
 * the frontier (depth == level), k_u probes, or the whole list -- unless the
 * list is the hint alone (one in-arc), which the miss already settled. */
int64_t or_pull_scan_count(int64_t n, const int64_t* offsets, const int32_t* cols, const int32_t* depth, int32_t level,
                           const int32_t* hint, int64_t* n_scanned, int64_t* n_missed) {
  int64_t total = 0, cnt = 0, missed = 0;
#pragma omp parallel for reduction(+ : total, cnt, missed) schedule(static)
  for (int64_t u = 0; u < n; ++u) {
    if (depth[u] <= level) continue;
    int64_t b = offsets[u], e = offsets[u + 1];
    if (e == b) continue;
    cnt += 1;
    if (hint) {
      total += 1;
      if (depth[hint[u]] == level) continue;
      missed += 1;
      if (e - b == 1) continue;
    }
    int64_t k = e - b;
    for (int64_t j = b; j < e; ++j)
      if (depth[cols[j]] == level) {
        k = j - b + 1;
        break;
      }
    total += k;
  }
  if (n_scanned) *n_scanned = cnt;
  if (n_missed) *n_missed = missed;
  return total;
}

/* hint[v] = the in-neighbour of v with the largest out-degree, the lowest id
 * on ties; -1 without in-arcs. */
void or_pull_hint(int64_t n, const int64_t* in_offsets, const int32_t* in_cols, const int64_t* out_offsets,
                  int32_t* hint) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < n; ++v) {
    int32_t best = -1;
    int64_t best_deg = -1;
    for (int64_t j = in_offsets[v]; j < in_offsets[v + 1]; ++j) {
      const int32_t u = in_cols[j];
      const int64_t d = out_offsets[u + 1] - out_offsets[u];
      if (d > best_deg || (d == best_deg && u < best)) {
        best = u;
        best_deg = d;
      }
    }
    hint[v] = best;
  }
}
