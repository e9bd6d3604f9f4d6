/*
 * cpu_baseline.c -- OpenMP restatement of the reference's *plain method*, the
 * CPU baseline reported next to the GPU numbers (labelled "port": it is this
 * repository's restatement, not the reference binary, which needs a SYCL
 * toolchain that is not available).
 *
 * TEST/BENCH INFRASTRUCTURE ONLY (see oracle.c header).
 *
 * It keeps the reference's structure: two 2-level bitmap frontiers
 * (frontier/impls/mlb_frontier.hpp:58-153), a compaction of active level-0
 * words before each advance (mlb_frontier.hpp:388-442, with `invert` for pull),
 * the BFS push/pull functors (algorithms/bfs.hpp:180-209) and the hybrid
 * alpha/beta heuristic evaluated on the new frontier after every step
 * (algorithms/bfs.hpp:284-350).
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
  int64_t n, w0, w1;
  uint64_t* l0;
  uint64_t* l1;
} mlb2;

static void mlb_init(mlb2* f, int64_t n) {
  f->n = n;
  f->w0 = (n + 63) / 64;
  f->w1 = (f->w0 + 63) / 64;
  f->l0 = (uint64_t*)calloc((size_t)f->w0, 8);
  f->l1 = (uint64_t*)calloc((size_t)f->w1, 8);
}

static void mlb_free(mlb2* f) {
  free(f->l0);
  free(f->l1);
}

static void mlb_clear(mlb2* f) {
  memset(f->l0, 0, (size_t)f->w0 * 8);
  memset(f->l1, 0, (size_t)f->w1 * 8);
}

/* MLBDevice::insert: test then atomic OR on each level (mlb_frontier.hpp:86-97) */
static inline void mlb_insert(mlb2* f, int64_t v) {
  int64_t w = v >> 6;
  uint64_t b = 1ULL << (v & 63);
  if (!(__atomic_load_n(&f->l0[w], __ATOMIC_RELAXED) & b)) __atomic_fetch_or(&f->l0[w], b, __ATOMIC_RELAXED);
  uint64_t b1 = 1ULL << (w & 63);
  if (!(__atomic_load_n(&f->l1[w >> 6], __ATOMIC_RELAXED) & b1)) __atomic_fetch_or(&f->l1[w >> 6], b1, __ATOMIC_RELAXED);
}

static inline int mlb_check(const mlb2* f, int64_t v) { return (int)((f->l0[v >> 6] >> (v & 63)) & 1ULL); }

static int mlb_empty(const mlb2* f) {
  for (int64_t i = 0; i < f->w1; ++i)
    if (f->l1[i]) return 0;
  return 1;
}

/* computeActiveFrontier: list level-0 words flagged in level 1 (push) or, with
 * invert, every word that is not all ones (pull). Ascending order. */
static int64_t mlb_compact(const mlb2* f, int invert, int32_t* out) {
  int64_t k = 0;
  for (int64_t w1 = 0; w1 < f->w1; ++w1) {
    uint64_t s = f->l1[w1];
    for (int b = 0; b < 64; ++b) {
      int64_t w = w1 * 64 + b;
      if (w >= f->w0) break;
      int active = (int)((s >> b) & 1ULL);
      if (!invert) {
        if (active) out[k++] = (int32_t)w;
      } else {
        if (!active || f->l0[w] != ~0ULL) out[k++] = (int32_t)w;
      }
    }
  }
  return k;
}

int or_omp_max_threads(void) { return omp_get_max_threads(); }

int32_t or_bfs_plain_omp(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t src, int mode, float alpha,
                         float beta, int nthreads, int32_t* depth, int32_t* dirs, int32_t dirs_cap) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int32_t unreached = (int32_t)(n + 1);
  mlb2 fin, fout;
  mlb_init(&fin, n);
  mlb_init(&fout, n);
  int32_t* words = (int32_t*)malloc(sizeof(int32_t) * (size_t)(fin.w0 + 1));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) depth[i] = unreached;
  depth[src] = 0;
  mlb_insert(&fin, src);
  int push = mode != 1;
  int32_t iter = 0;
  while (!mlb_empty(&fin)) {
    if (dirs && iter < dirs_cap) dirs[iter] = push ? 0 : 1;
    int64_t nw = mlb_compact(&fin, push ? 0 : 1, words);
    if (push) {
#pragma omp parallel for schedule(dynamic, 16)
      for (int64_t i = 0; i < nw; ++i) {
        int64_t w = words[i];
        uint64_t bits = fin.l0[w];
        while (bits) {
          int b = __builtin_ctzll(bits);
          bits &= bits - 1;
          int64_t v = w * 64 + b;
          if (v >= n) break;
          for (int64_t e = offsets[v]; e < offsets[v + 1]; ++e) {
            int32_t y = cols[e];
            /* bfs.hpp:185-191: if depth[dst]==n+1 {depth[dst]=iter+1; return true} */
            if (__atomic_load_n(&depth[y], __ATOMIC_RELAXED) == unreached) {
              __atomic_store_n(&depth[y], iter + 1, __ATOMIC_RELAXED);
              mlb_insert(&fout, y);
            }
          }
        }
      }
    } else {
#pragma omp parallel for schedule(dynamic, 16)
      for (int64_t i = 0; i < nw; ++i) {
        int64_t w = words[i];
        for (int b = 0; b < 64; ++b) {
          int64_t v = w * 64 + b;
          if (v >= n) break;
          if (mlb_check(&fin, v)) continue; /* pull processes non-frontier vertices */
          for (int64_t e = offsets[v]; e < offsets[v + 1]; ++e) {
            int32_t u = cols[e];
            if (!mlb_check(&fin, u)) continue; /* in-neighbour must be in the frontier */
            /* bfs.hpp:201-207 */
            if (depth[v] == unreached && depth[u] == iter) {
              depth[v] = iter + 1;
              mlb_insert(&fout, v);
              break; /* short circuit on first success */
            }
          }
        }
      }
    }
    if (mode == 2) {
      if (push) {
        /* shouldSwitchToPull: frontier degree > unexplored degree / alpha (uint32 sums) */
        uint64_t fdeg = 0, udeg = 0;
#pragma omp parallel for reduction(+ : fdeg, udeg) schedule(static)
        for (int64_t v = 0; v < n; ++v) {
          uint64_t d = (uint64_t)(offsets[v + 1] - offsets[v]);
          if (mlb_check(&fout, v)) fdeg += d;
          if (depth[v] == unreached) udeg += d;
        }
        uint32_t f32 = (uint32_t)fdeg, u32 = (uint32_t)udeg;
        if ((float)f32 > ((float)u32 / alpha)) push = 0;
      } else {
        /* shouldSwitchToPush: |frontier| < n / beta */
        uint64_t cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
        for (int64_t w = 0; w < fout.w0; ++w) cnt += (uint64_t)__builtin_popcountll(fout.l0[w]);
        if ((float)cnt < ((float)n / beta)) push = 1;
      }
    }
    mlb2 t = fin;
    fin = fout;
    fout = t;
    mlb_clear(&fout);
    ++iter;
  }
  free(words);
  mlb_free(&fin);
  mlb_free(&fout);
  return iter;
}

int32_t or_sssp_plain_omp(int64_t n, const int64_t* offsets, const int32_t* cols, const uint32_t* w, int32_t src,
                          int nthreads, uint32_t* dist) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const uint32_t unreached = (uint32_t)(n + 1);
  mlb2 fin, fout;
  mlb_init(&fin, n);
  mlb_init(&fout, n);
  int32_t* words = (int32_t*)malloc(sizeof(int32_t) * (size_t)(fin.w0 + 1));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) dist[i] = unreached;
  dist[src] = 0;
  mlb_insert(&fin, src);
  int32_t iter = 0;
  while (!mlb_empty(&fin)) {
    int64_t nw = mlb_compact(&fin, 0, words);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nw; ++i) {
      int64_t wd = words[i];
      uint64_t bits = fin.l0[wd];
      while (bits) {
        int b = __builtin_ctzll(bits);
        bits &= bits - 1;
        int64_t v = wd * 64 + b;
        if (v >= n) break;
        uint32_t dv = __atomic_load_n(&dist[v], __ATOMIC_RELAXED);
        for (int64_t e = offsets[v]; e < offsets[v + 1]; ++e) {
          int32_t y = cols[e];
          uint32_t nd = dv + w[e];
          /* sssp.hpp:180-189: fetch_min; improved if new < old */
          uint32_t old = __atomic_load_n(&dist[y], __ATOMIC_RELAXED);
          while (nd < old && !__atomic_compare_exchange_n(&dist[y], &old, nd, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
          }
          if (nd < old) mlb_insert(&fout, y);
        }
      }
    }
    /* filter::external(out -> in) keeps every vertex of out (visited[v] != iter
     * holds for each vertex once per round, sssp.hpp:192-196), so in := out. */
    mlb2 t = fin;
    fin = fout;
    fout = t;
    mlb_clear(&fout);
    ++iter;
  }
  free(words);
  mlb_free(&fin);
  mlb_free(&fout);
  return iter;
}
