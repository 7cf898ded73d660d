/*
 * oracle.c -- CPU restatement of Thievory's single-GPU graph algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
 * thievory_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  The product path never calls into it.
 *
 * Contents (reference citations are path:line in qqueke/Thievory):
 *   - readers/writers for the .csr / .csc / .wcsr byte layouts
 *       (src/graph.cuh:641-777, converter/converter.cpp:483-589, 676-808)
 *   - the seeded R-MAT generator used for every synthetic workload
 *       (no reference counterpart; Chakrabarti, Zhan, Faloutsos, SDM 2004,
 *        with a SplitMix64 counter RNG, Steele/Lea/Flood OOPSLA 2014)
 *   - serial oracles with the reference's output conventions
 *       BFS  : val = 1 + hops, unreached = n+1 (src/graph.cuh:240-246,
 *              src/bfs/bfs_kernels.cu:29-44)
 *       CC   : val = min({v} u {u : u reaches v}) (src/graph.cuh:260-264,
 *              src/cc/cc_kernels.cu:50-66)
 *       SSSP : val = min(n+1, 1 + d_w(src, v)) (src/sssp/sssp_kernels.cu:
 *              124-141, with the file's weights, see the weight note below)
 *       PR   : unnormalised PageRank r = (1-a) + a * sum r[u]/outdeg[u]
 *              (src/pr/pr_kernels.cu:45-51, 449-476, 539-542, 846-860)
 *   - multithreaded (OpenMP) versions of the same algorithms: the CPU
 *     baseline timed beside the GPU.
 *
 * Weight note: the reference's SSSP kernels add a constant 20 in the static
 * kernel and read a never-written buffer in the others
 * (src/sssp/sssp_kernels.cu:52, src/sssp/sssp.cu:149); the intended rule
 * nv = val[v] + w(v,w) with the file's weights is what is restated here.
 *
 * Every function is pinned by tests/test_oracle.py: the converter-produced
 * fixture and its hand-derived answers, scipy.sparse.csgraph routines,
 * brute force on tiny graphs and cross-checks between independent
 * algorithms (queue BFS vs scipy, label propagation vs pruned min-reach,
 * Dijkstra vs Bellman-Ford vs scipy, power iteration vs a sparse solve).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* hashing                                                             */
/* ------------------------------------------------------------------ */

/* SplitMix64 output function (Steele, Lea, Flood 2014). */
ORC_EXPORT uint64_t orc_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Order-independent digest of a value array: sum over v with
 * val[v] != sentinel of mix64((v << 32) | val[v]), mod 2^64.  The GPU batch
 * path accumulates the same sum lane by lane. */
ORC_EXPORT uint64_t orc_hash_values(uint64_t n, const uint32_t *val,
                                    uint32_t sentinel, uint64_t *reached,
                                    uint32_t *depth) {
  uint64_t h = 0, r = 0;
  uint32_t d = 0;
  for (uint64_t v = 0; v < n; v++) {
    if (val[v] == sentinel) continue;
    h += orc_mix64((v << 32) | (uint64_t)val[v]);
    r++;
    if (val[v] > d) d = val[v];
  }
  if (reached) *reached = r;
  if (depth) *depth = d;
  return h;
}

/* ------------------------------------------------------------------ */
/* file formats                                                        */
/* ------------------------------------------------------------------ */
/* fmt 0 = .csr  : u64 n, u64 m, u64 off[n], u32 dst[m] [, u32 w[m]]
 *                 (converter.cpp:750-774; graph.cuh:707-754 reads the first
 *                  three parts and sets off[n] = m)
 * fmt 1 = .csc  : u64 n, u64 m, u64 inoff[n], u32 src[m], u32 outdeg[n]
 *                 (converter.cpp:572-584; graph.cuh:773-777)
 * fmt 2 = .wcsr : u64 n, u64 m, u64 off[n], {u32 dst, u32 w}[m]
 *                 (converter.cpp:785-802)                               */

static long long file_size(FILE *f) {
  if (fseeko(f, 0, SEEK_END) != 0) return -1;
  long long s = ftello(f);
  fseeko(f, 0, SEEK_SET);
  return s;
}

ORC_EXPORT int orc_read_header(const char *path, uint64_t *n, uint64_t *m,
                               uint64_t *bytes) {
  FILE *f = fopen(path, "rb");
  if (!f) return -2;
  long long s = file_size(f);
  uint64_t h[2];
  if (s < 16 || fread(h, 8, 2, f) != 2) { fclose(f); return -2; }
  fclose(f);
  *n = h[0]; *m = h[1]; *bytes = (uint64_t)s;
  return 0;
}

/* Reads the body.  off has n+1 slots; adj m; w (nullable) m; outdeg
 * (nullable, fmt 1 only) n.  Returns 1 if weights were present, 0 if not,
 * negative on error. */
ORC_EXPORT int orc_read_graph(const char *path, int fmt, uint64_t n,
                              uint64_t m, uint64_t *off, uint32_t *adj,
                              uint32_t *w, uint32_t *outdeg) {
  FILE *f = fopen(path, "rb");
  if (!f) return -2;
  long long s = file_size(f);
  uint64_t base = 16 + 8 * n;
  int has_w = 0;
  if (fmt == 0) {
    if ((uint64_t)s == base + 8 * m) has_w = 1;
    else if ((uint64_t)s != base + 4 * m) { fclose(f); return -5; }
  } else if (fmt == 1) {
    if ((uint64_t)s != base + 4 * m + 4 * n) { fclose(f); return -5; }
  } else if (fmt == 2) {
    if ((uint64_t)s != base + 8 * m) { fclose(f); return -5; }
    has_w = 1;
  } else { fclose(f); return -1; }
  fseeko(f, 16, SEEK_SET);
  if (fread(off, 8, n, f) != n) { fclose(f); return -2; }
  off[n] = m;
  if (fmt == 2) {
    uint32_t buf[2 * 4096];
    for (uint64_t i = 0; i < m;) {
      uint64_t k = m - i < 4096 ? m - i : 4096;
      if (fread(buf, 8, k, f) != k) { fclose(f); return -2; }
      for (uint64_t j = 0; j < k; j++) {
        adj[i + j] = buf[2 * j];
        if (w) w[i + j] = buf[2 * j + 1];
      }
      i += k;
    }
  } else {
    if (fread(adj, 4, m, f) != m) { fclose(f); return -2; }
    if (fmt == 0 && has_w && w) {
      if (fread(w, 4, m, f) != m) { fclose(f); return -2; }
    }
    if (fmt == 1 && outdeg) {
      if (fread(outdeg, 4, n, f) != n) { fclose(f); return -2; }
    }
  }
  fclose(f);
  return has_w;
}

ORC_EXPORT int orc_write_graph(const char *path, int fmt, uint64_t n,
                               uint64_t m, const uint64_t *off,
                               const uint32_t *adj, const uint32_t *w,
                               const uint32_t *outdeg) {
  FILE *f = fopen(path, "wb");
  if (!f) return -2;
  uint64_t h[2] = {n, m};
  fwrite(h, 8, 2, f);
  fwrite(off, 8, n, f);
  if (fmt == 2) {
    for (uint64_t i = 0; i < m; i++) {
      uint32_t r[2] = {adj[i], w ? w[i] : 20u};
      fwrite(r, 4, 2, f);
    }
  } else {
    fwrite(adj, 4, m, f);
    if (fmt == 0 && w) fwrite(w, 4, m, f);
    if (fmt == 1 && outdeg) fwrite(outdeg, 4, n, f);
  }
  int bad = ferror(f);
  fclose(f);
  return bad ? -2 : 0;
}

/* ------------------------------------------------------------------ */
/* R-MAT generator                                                     */
/* ------------------------------------------------------------------ */
/* Tuple e picks its `scale` quadrant bits from the top 32 bits of
 * mix64(key + l), key = mix64(seed * 0x9E3779B97F4A7C15 + e); bit l of the
 * id pair is the l-th most significant.  Quadrant thresholds are integers
 * (ta, tb, tc) = floor(2^32 * (a, a+b, a+b+c)) so CPU and GPU agree bit for
 * bit. */

ORC_EXPORT void orc_rmat_thresholds(double a, double b, double c,
                                    uint64_t th[3]) {
  th[0] = (uint64_t)floor(a * 4294967296.0);
  th[1] = (uint64_t)floor((a + b) * 4294967296.0);
  th[2] = (uint64_t)floor((a + b + c) * 4294967296.0);
}

static inline void rmat_tuple(int scale, uint64_t seed, const uint64_t th[3],
                              uint64_t e, uint32_t *u, uint32_t *v) {
  uint64_t key = orc_mix64(seed * 0x9E3779B97F4A7C15ull + e);
  uint32_t uu = 0, vv = 0;
  for (int l = 0; l < scale; l++) {
    uint64_t r = orc_mix64(key + (uint64_t)l) >> 32;
    uint32_t bu = 0, bv = 0;
    if (r < th[0]) { bu = 0; bv = 0; }
    else if (r < th[1]) { bu = 0; bv = 1; }
    else if (r < th[2]) { bu = 1; bv = 0; }
    else { bu = 1; bv = 1; }
    uu = (uu << 1) | bu;
    vv = (vv << 1) | bv;
  }
  *u = uu;
  *v = vv;
}

ORC_EXPORT void orc_rmat_tuples(int scale, uint64_t seed, const uint64_t th[3],
                                uint64_t e0, uint64_t e1, uint32_t *u,
                                uint32_t *v) {
#pragma omp parallel for schedule(static)
  for (long long e = (long long)e0; e < (long long)e1; e++)
    rmat_tuple(scale, seed, th, (uint64_t)e, &u[e - e0], &v[e - e0]);
}

/* Symmetric edge weight in [1, 64]. */
ORC_EXPORT uint32_t orc_weight(uint64_t seed, uint32_t u, uint32_t v) {
  uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
  return 1u + (uint32_t)(orc_mix64(seed * 0xD1B54A32D192ED03ull +
                                   ((lo << 32) | hi)) & 63u);
}

/* LSD radix sort of 64-bit keys on their low `bits` bits. */
static void radix_sort_u64(uint64_t *k, uint64_t *tmp, uint64_t cnt, int bits) {
  const int D = 11;
  uint64_t *src = k, *dst = tmp;
  for (int sh = 0; sh < bits; sh += D) {
    static uint64_t hist[1 << 11];
    memset(hist, 0, sizeof(hist));
    for (uint64_t i = 0; i < cnt; i++) hist[(src[i] >> sh) & ((1 << D) - 1)]++;
    uint64_t s = 0;
    for (int i = 0; i < (1 << D); i++) { uint64_t t = hist[i]; hist[i] = s; s += t; }
    for (uint64_t i = 0; i < cnt; i++)
      dst[hist[(src[i] >> sh) & ((1 << D) - 1)]++] = src[i];
    uint64_t *t = src; src = dst; dst = t;
  }
  if (src != k) memcpy(k, src, cnt * sizeof(uint64_t));
}

/* Builds the CSR of an R-MAT graph: ntuples tuples, self-loops dropped,
 * symmetrised when `symmetric`, duplicates removed, adjacency sorted.
 * off must hold 2^scale + 1 entries and adj 2*ntuples (symmetric) or
 * ntuples.  Returns the number of directed edges m. */
ORC_EXPORT uint64_t orc_rmat_csr(int scale, uint64_t ntuples, uint64_t seed,
                                 double a, double b, double c, int symmetric,
                                 uint64_t *off, uint32_t *adj) {
  uint64_t th[3];
  orc_rmat_thresholds(a, b, c, th);
  uint64_t n = 1ull << scale;
  uint64_t cap = symmetric ? 2 * ntuples : ntuples;
  uint64_t *keys = (uint64_t *)malloc(cap * sizeof(uint64_t));
  uint64_t *tmp = (uint64_t *)malloc(cap * sizeof(uint64_t));
  uint64_t cnt = 0;
  for (uint64_t e = 0; e < ntuples; e++) {
    uint32_t u, v;
    rmat_tuple(scale, seed, th, e, &u, &v);
    if (u == v) continue;
    keys[cnt++] = ((uint64_t)u << 32) | v;
    if (symmetric) keys[cnt++] = ((uint64_t)v << 32) | u;
  }
  /* sort on (u, v): u occupies bits 32..32+scale */
  radix_sort_u64(keys, tmp, cnt, 32 + scale);
  free(tmp);
  memset(off, 0, (n + 1) * sizeof(uint64_t));
  uint64_t m = 0;
  for (uint64_t i = 0; i < cnt; i++) {
    if (i > 0 && keys[i] == keys[i - 1]) continue;
    uint32_t u = (uint32_t)(keys[i] >> 32);
    adj[m++] = (uint32_t)keys[i];
    off[u + 1]++;
  }
  for (uint64_t v = 0; v < n; v++) off[v + 1] += off[v];
  free(keys);
  return m;
}

/* Transpose (CSR -> CSC), keeping in-lists sorted by source. */
ORC_EXPORT void orc_transpose(uint64_t n, const uint64_t *off,
                              const uint32_t *adj, uint64_t *toff,
                              uint32_t *tadj) {
  uint64_t m = off[n];
  memset(toff, 0, (n + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; i++) toff[adj[i] + 1]++;
  for (uint64_t v = 0; v < n; v++) toff[v + 1] += toff[v];
  uint64_t *pos = (uint64_t *)malloc(n * sizeof(uint64_t));
  memcpy(pos, toff, n * sizeof(uint64_t));
  for (uint64_t u = 0; u < n; u++)
    for (uint64_t i = off[u]; i < off[u + 1]; i++) tadj[pos[adj[i]]++] = (uint32_t)u;
  free(pos);
}

/* ------------------------------------------------------------------ */
/* serial oracles                                                      */
/* ------------------------------------------------------------------ */

/* BFS: level-synchronous queue BFS; val = 1 + hops, n+1 if unreached
 * (graph.cuh:240-246 init, bfs_kernels.cu:29-44 relaxation). */
ORC_EXPORT void orc_bfs(uint64_t n, const uint64_t *off, const uint32_t *adj,
                        uint32_t src, uint32_t *val) {
  uint32_t inf = (uint32_t)(n + 1);
  for (uint64_t v = 0; v < n; v++) val[v] = inf;
  if (src >= n) return;
  uint32_t *q = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint64_t head = 0, tail = 0;
  val[src] = 1;
  q[tail++] = src;
  while (head < tail) {
    uint32_t v = q[head++];
    uint32_t nv = val[v] + 1;
    for (uint64_t i = off[v]; i < off[v + 1]; i++) {
      uint32_t w = adj[i];
      if (val[w] == inf) { val[w] = nv; q[tail++] = w; }
    }
  }
  free(q);
}

/* CC: min-label propagation to the fixpoint (graph.cuh:260-264 init,
 * cc_kernels.cu:50-66 rule: if val[v] < val[w] then val[w] = val[v]). */
ORC_EXPORT void orc_cc(uint64_t n, const uint64_t *off, const uint32_t *adj,
                       uint32_t *val) {
  uint32_t *q = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  uint8_t *inq = (uint8_t *)malloc(n ? n : 1);
  uint64_t head = 0, cnt = n;
  for (uint64_t v = 0; v < n; v++) { val[v] = (uint32_t)v; q[v] = (uint32_t)v; inq[v] = 1; }
  /* circular work queue */
  while (cnt) {
    uint32_t v = q[head];
    head = (head + 1) % n;
    cnt--;
    inq[v] = 0;
    for (uint64_t i = off[v]; i < off[v + 1]; i++) {
      uint32_t w = adj[i];
      if (val[v] < val[w]) {
        val[w] = val[v];
        if (!inq[w]) { q[(head + cnt) % n] = w; cnt++; inq[w] = 1; }
      }
    }
  }
  free(q);
  free(inq);
}

/* CC, second algorithm: val[v] = least u that reaches v.  Sources are taken
 * in increasing order and each search stops at already-labelled vertices
 * (whatever such a vertex reaches already carries a smaller label). */
ORC_EXPORT void orc_cc_minreach(uint64_t n, const uint64_t *off,
                                const uint32_t *adj, uint32_t *val) {
  const uint32_t none = 0xFFFFFFFFu;
  uint32_t *st = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  for (uint64_t v = 0; v < n; v++) val[v] = none;
  for (uint64_t u = 0; u < n; u++) {
    if (val[u] != none) continue;
    uint64_t sp = 0;
    val[u] = (uint32_t)u;
    st[sp++] = (uint32_t)u;
    while (sp) {
      uint32_t v = st[--sp];
      for (uint64_t i = off[v]; i < off[v + 1]; i++) {
        uint32_t w = adj[i];
        if (val[w] == none) { val[w] = (uint32_t)u; st[sp++] = w; }
      }
    }
  }
  free(st);
}

/* SSSP: Dijkstra with a binary heap; val = min(n+1, 1 + dist), the fixpoint
 * of the reference's strict-less relaxation from val[src] = 1 with INF = n+1
 * (sssp_kernels.cu:124-141). */
typedef struct { uint64_t d; uint32_t v; } hent;
static void heap_push(hent *h, uint64_t *sz, hent x) {
  uint64_t i = (*sz)++;
  while (i > 0) {
    uint64_t p = (i - 1) / 2;
    if (h[p].d <= x.d) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = x;
}
static hent heap_pop(hent *h, uint64_t *sz) {
  hent top = h[0], x = h[--(*sz)];
  uint64_t i = 0;
  for (;;) {
    uint64_t l = 2 * i + 1, r = l + 1, s = i;
    hent best = x;
    if (l < *sz && h[l].d < best.d) { s = l; best = h[l]; }
    if (r < *sz && h[r].d < best.d) { s = r; best = h[r]; }
    if (s == i) break;
    h[i] = h[s];
    i = s;
  }
  h[i] = x;
  return top;
}

ORC_EXPORT void orc_sssp(uint64_t n, const uint64_t *off, const uint32_t *adj,
                         const uint32_t *w, uint32_t src, uint32_t *val) {
  uint64_t inf = n + 1;
  uint64_t *d = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
  for (uint64_t v = 0; v < n; v++) d[v] = UINT64_MAX;
  for (uint64_t v = 0; v < n; v++) val[v] = (uint32_t)inf;
  if (src >= n) { free(d); return; }
  uint64_t m = off[n];
  hent *h = (hent *)malloc((m + 1) * sizeof(hent));
  uint64_t sz = 0;
  d[src] = 0;
  heap_push(h, &sz, (hent){0, src});
  while (sz) {
    hent x = heap_pop(h, &sz);
    if (x.d != d[x.v]) continue;
    for (uint64_t i = off[x.v]; i < off[x.v + 1]; i++) {
      uint64_t nd = x.d + w[i];
      if (nd < d[adj[i]]) { d[adj[i]] = nd; heap_push(h, &sz, (hent){nd, adj[i]}); }
    }
  }
  for (uint64_t v = 0; v < n; v++)
    if (d[v] != UINT64_MAX && 1 + d[v] < inf) val[v] = (uint32_t)(1 + d[v]);
  free(h);
  free(d);
}

/* SSSP, restatement of the reference loop: frontier Bellman-Ford with the
 * strict-less rule on u32 values (sums formed in 64 bits). */
ORC_EXPORT void orc_sssp_bf(uint64_t n, const uint64_t *off,
                            const uint32_t *adj, const uint32_t *w,
                            uint32_t src, uint32_t *val) {
  uint32_t inf = (uint32_t)(n + 1);
  uint8_t *cur = (uint8_t *)calloc(n ? n : 1, 1), *nxt = (uint8_t *)calloc(n ? n : 1, 1);
  for (uint64_t v = 0; v < n; v++) val[v] = inf;
  if (src >= n) { free(cur); free(nxt); return; }
  val[src] = 1;
  cur[src] = 1;
  int any = 1;
  while (any) {
    any = 0;
    for (uint64_t v = 0; v < n; v++) {
      if (!cur[v]) continue;
      for (uint64_t i = off[v]; i < off[v + 1]; i++) {
        uint64_t nv = (uint64_t)val[v] + w[i];
        if (nv < val[adj[i]]) { val[adj[i]] = (uint32_t)nv; nxt[adj[i]] = 1; any = 1; }
      }
    }
    uint8_t *t = cur; cur = nxt; nxt = t;
    memset(nxt, 0, n);
  }
  free(cur);
  free(nxt);
}

/* PageRank, exact fixpoint in FP64 by Jacobi iteration on the CSR (push
 * form of the same sums): r = (1-a) + a * sum_{u->v, outdeg u > 0}
 * r[u]/outdeg[u], from r = 1-a, until max |delta| < tol.  Returns the
 * iteration count. */
ORC_EXPORT int orc_pr_exact(uint64_t n, const uint64_t *off,
                            const uint32_t *adj, double alpha, double tol,
                            int max_iter, double *rank) {
  double *acc = (double *)malloc((n ? n : 1) * sizeof(double));
  for (uint64_t v = 0; v < n; v++) rank[v] = 1.0 - alpha;
  int it = 0;
  for (; it < max_iter; it++) {
    memset(acc, 0, n * sizeof(double));
    for (uint64_t u = 0; u < n; u++) {
      uint64_t d = off[u + 1] - off[u];
      if (!d) continue;
      double c = rank[u] / (double)d;
      for (uint64_t i = off[u]; i < off[u + 1]; i++) acc[adj[i]] += c;
    }
    double md = 0;
    for (uint64_t v = 0; v < n; v++) {
      double nr = (1.0 - alpha) + alpha * acc[v];
      double dd = fabs(nr - rank[v]);
      if (dd > md) md = dd;
      rank[v] = nr;
    }
    if (md < tol) { it++; break; }
  }
  free(acc);
  return it;
}

/* The same FP64 fixpoint, multithreaded, as a Jacobi pull over the CSC
 * (in-offsets, in-sources, out-degrees); used where the serial form is too
 * slow (full-size PageRank workloads).  Returns the iteration count. */
ORC_EXPORT int orc_pr_exact_mt(uint64_t n, const uint64_t *ioff, const uint32_t *isrc,
                               const uint32_t *outdeg, double alpha, double tol,
                               int max_iter, double *rank) {
  double *contrib = (double *)malloc((n ? n : 1) * sizeof(double));
#pragma omp parallel for schedule(static)
  for (long long v = 0; v < (long long)n; v++) rank[v] = 1.0 - alpha;
  int it = 0;
  for (; it < max_iter; it++) {
#pragma omp parallel for schedule(static)
    for (long long u = 0; u < (long long)n; u++)
      contrib[u] = outdeg[u] ? rank[u] / (double)outdeg[u] : 0.0;
    double md = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(max : md)
    for (long long v = 0; v < (long long)n; v++) {
      double s = 0;
      for (uint64_t i = ioff[v]; i < ioff[v + 1]; i++) s += contrib[isrc[i]];
      double nr = (1.0 - alpha) + alpha * s;
      double dd = fabs(nr - rank[v]);
      if (dd > md) md = dd;
      rank[v] = nr;
    }
    if (md < tol) { it++; break; }
  }
  free(contrib);
  return it;
}

/* PageRank pull, the reference's loop in FP64 (pr_kernels.cu:45-51,
 * 449-476; init rank = 1/n, graph.cuh:255-256): active v recomputes
 * sum over in-neighbours, stays active while |new-old| > tol.  Input is the
 * CSC (in-offsets, in-sources) plus out-degrees. */
ORC_EXPORT int orc_pr_pull_ref(uint64_t n, const uint64_t *ioff,
                               const uint32_t *isrc, const uint32_t *outdeg,
                               double alpha, double tol, int max_iter,
                               double *rank) {
  uint8_t *act = (uint8_t *)malloc(n ? n : 1);
  double *sum = (double *)calloc(n ? n : 1, sizeof(double));
  for (uint64_t v = 0; v < n; v++) { rank[v] = 1.0 / (double)n; act[v] = 1; }
  int it = 0;
  for (; it < max_iter; it++) {
    uint64_t nact = 0;
    for (uint64_t v = 0; v < n; v++) {
      if (!act[v]) continue;
      double s = 0;
      for (uint64_t i = ioff[v]; i < ioff[v + 1]; i++)
        if (outdeg[isrc[i]]) s += rank[isrc[i]] / (double)outdeg[isrc[i]];
      sum[v] = s;
    }
    for (uint64_t v = 0; v < n; v++) {
      if (!act[v]) continue;
      double nr = (1.0 - alpha) + alpha * sum[v];
      act[v] = fabs(nr - rank[v]) > tol;
      rank[v] = nr;
      nact += act[v];
    }
    if (!nact) { it++; break; }
  }
  free(act);
  free(sum);
  return it;
}

/* PageRank push, the reference's residual loop in FP64
 * (graph.cuh:248-253 init; pr_kernels.cu:539-542, 846-860). */
ORC_EXPORT int orc_pr_push_ref(uint64_t n, const uint64_t *off,
                               const uint32_t *adj, double alpha, double tol,
                               int max_iter, double *rank) {
  double *delta = (double *)malloc((n ? n : 1) * sizeof(double));
  double *res = (double *)calloc(n ? n : 1, sizeof(double));
  uint8_t *act = (uint8_t *)malloc(n ? n : 1);
  for (uint64_t v = 0; v < n; v++) {
    uint64_t d = off[v + 1] - off[v];
    rank[v] = 1.0 - alpha;
    delta[v] = d ? (1.0 - alpha) * alpha / (double)d : 0.0;
    act[v] = 1;
  }
  int it = 0;
  for (; it < max_iter; it++) {
    for (uint64_t u = 0; u < n; u++) {
      if (!act[u]) continue;
      for (uint64_t i = off[u]; i < off[u + 1]; i++) res[adj[i]] += delta[u];
    }
    uint64_t nact = 0;
    for (uint64_t v = 0; v < n; v++) {
      act[v] = 0;
      if (res[v] > tol) {
        uint64_t d = off[v + 1] - off[v];
        rank[v] += res[v];
        delta[v] = d ? res[v] * alpha / (double)d : 0.0;
        res[v] = 0;
        act[v] = 1;
        nact++;
      }
    }
    if (!nact) { it++; break; }
  }
  free(delta);
  free(res);
  free(act);
  return it;
}

/* ------------------------------------------------------------------ */
/* multithreaded CPU baseline (same results as the serial oracles)     */
/* ------------------------------------------------------------------ */

ORC_EXPORT int orc_threads(void) { return omp_get_max_threads(); }
ORC_EXPORT void orc_set_threads(int t) { if (t > 0) omp_set_num_threads(t); }

/* Top-down level-synchronous BFS with an atomic frontier. */
ORC_EXPORT uint64_t orc_bfs_mt(uint64_t n, const uint64_t *off,
                               const uint32_t *adj, uint32_t src,
                               uint32_t *val) {
  uint32_t inf = (uint32_t)(n + 1);
#pragma omp parallel for schedule(static)
  for (long long v = 0; v < (long long)n; v++) val[v] = inf;
  if (src >= n) return 0;
  uint32_t *cur = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *nxt = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint64_t ncur = 1, ntrav = 0;
  cur[0] = src;
  val[src] = 1;
  uint32_t level = 1;
  while (ncur) {
    uint64_t nnext = 0;
    uint32_t nv = level + 1;
#pragma omp parallel reduction(+ : ntrav)
    {
      uint32_t buf[1024];
      int nb = 0;
#pragma omp for schedule(dynamic, 64)
      for (long long k = 0; k < (long long)ncur; k++) {
        uint32_t v = cur[k];
        ntrav += off[v + 1] - off[v];
        for (uint64_t i = off[v]; i < off[v + 1]; i++) {
          uint32_t w = adj[i];
          if (__atomic_load_n(&val[w], __ATOMIC_RELAXED) != inf) continue;
          uint32_t expect = inf;
          if (__atomic_compare_exchange_n(&val[w], &expect, nv, 0,
                                          __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
            buf[nb++] = w;
            if (nb == 1024) {
              uint64_t p = __atomic_fetch_add(&nnext, (uint64_t)nb, __ATOMIC_RELAXED);
              memcpy(nxt + p, buf, nb * sizeof(uint32_t));
              nb = 0;
            }
          }
        }
      }
      if (nb) {
        uint64_t p = __atomic_fetch_add(&nnext, (uint64_t)nb, __ATOMIC_RELAXED);
        memcpy(nxt + p, buf, nb * sizeof(uint32_t));
      }
    }
    uint32_t *t = cur; cur = nxt; nxt = t;
    ncur = nnext;
    level++;
  }
  free(cur);
  free(nxt);
  return ntrav;
}

static inline int atomic_min_u32(uint32_t *p, uint32_t v) {
  uint32_t old = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (v < old) {
    if (__atomic_compare_exchange_n(p, &old, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      return 1;
  }
  return 0;
}

/* Frontier label propagation (CC) or Bellman-Ford (SSSP, w != NULL) with an
 * atomic min; returns the number of edges processed.  Takes ownership of
 * (and frees) the frontier array `cur`. */
static uint64_t relax_mt(uint64_t n, const uint64_t *off, const uint32_t *adj,
                         const uint32_t *w, uint32_t *val, uint8_t *cur) {
  uint8_t *nxt = (uint8_t *)calloc(n ? n : 1, 1);
  uint64_t nproc = 0;
  int any = 1;
  while (any) {
    any = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : nproc) reduction(| : any)
    for (long long v = 0; v < (long long)n; v++) {
      if (!cur[v]) continue;
      cur[v] = 0;
      uint32_t base = __atomic_load_n(&val[v], __ATOMIC_RELAXED);
      nproc += off[v + 1] - off[v];
      for (uint64_t i = off[v]; i < off[v + 1]; i++) {
        uint64_t nv = w ? (uint64_t)base + w[i] : base;
        if (nv >= 0xFFFFFFFFull) continue;
        if (atomic_min_u32(&val[adj[i]], (uint32_t)nv)) {
          nxt[adj[i]] = 1;
          any = 1;
        }
      }
    }
    uint8_t *t = cur; cur = nxt; nxt = t;
  }
  free(cur);
  free(nxt);
  return nproc;
}

ORC_EXPORT uint64_t orc_cc_mt(uint64_t n, const uint64_t *off,
                              const uint32_t *adj, uint32_t *val) {
  uint8_t *cur = (uint8_t *)malloc(n ? n : 1);
#pragma omp parallel for schedule(static)
  for (long long v = 0; v < (long long)n; v++) { val[v] = (uint32_t)v; cur[v] = 1; }
  return relax_mt(n, off, adj, NULL, val, cur);
}

ORC_EXPORT uint64_t orc_sssp_mt(uint64_t n, const uint64_t *off,
                                const uint32_t *adj, const uint32_t *w,
                                uint32_t src, uint32_t *val) {
  uint8_t *cur = (uint8_t *)calloc(n ? n : 1, 1);
#pragma omp parallel for schedule(static)
  for (long long v = 0; v < (long long)n; v++) val[v] = (uint32_t)(n + 1);
  if (src >= n) { free(cur); return 0; }
  val[src] = 1;
  cur[src] = 1;
  return relax_mt(n, off, adj, w, val, cur);
}

/* SSSP by delta-stepping (Meyer & Sanders 2003) on the host cores: the
 * pending vertices whose value is below the bucket bound thr have their
 * edges relaxed with an atomic min (every improvement makes the target
 * pending again); when none is below thr, thr becomes the least pending
 * value + delta.  Same fixpoint as orc_sssp / orc_sssp_mt (shortest
 * distances are unique), with fewer edges relaxed twice.  Returns the number
 * of edges relaxed. */
ORC_EXPORT uint64_t orc_sssp_delta_mt(uint64_t n, const uint64_t *off,
                                      const uint32_t *adj, const uint32_t *w,
                                      uint32_t src, uint32_t delta, uint32_t *val) {
  uint8_t *pend = (uint8_t *)calloc(n ? n : 1, 1), *cur = (uint8_t *)calloc(n ? n : 1, 1);
#pragma omp parallel for schedule(static)
  for (long long v = 0; v < (long long)n; v++) val[v] = (uint32_t)(n + 1);
  uint64_t nproc = 0;
  if (src < n) {
    val[src] = 1;
    pend[src] = 1;
    uint64_t thr = 1 + (uint64_t)(delta ? delta : 1);
    for (;;) {
      int any = 0;
#pragma omp parallel for schedule(static) reduction(| : any)
      for (long long v = 0; v < (long long)n; v++) {
        if (pend[v] && val[v] < thr) {
          pend[v] = 0;
          cur[v] = 1;
          any = 1;
        }
      }
      if (!any) {
        uint32_t lo = 0xFFFFFFFFu;
#pragma omp parallel for schedule(static) reduction(min : lo)
        for (long long v = 0; v < (long long)n; v++)
          if (pend[v] && val[v] < lo) lo = val[v];
        if (lo == 0xFFFFFFFFu) break;
        thr = (uint64_t)lo + (delta ? delta : 1);
        continue;
      }
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : nproc)
      for (long long v = 0; v < (long long)n; v++) {
        if (!cur[v]) continue;
        cur[v] = 0;
        uint32_t base = __atomic_load_n(&val[v], __ATOMIC_RELAXED);
        nproc += off[v + 1] - off[v];
        for (uint64_t i = off[v]; i < off[v + 1]; i++) {
          uint64_t nv = (uint64_t)base + w[i];
          if (nv >= 0xFFFFFFFFull) continue;
          if (atomic_min_u32(&val[adj[i]], (uint32_t)nv))
            __atomic_store_n(&pend[adj[i]], 1, __ATOMIC_RELAXED);
        }
      }
    }
  }
  free(pend);
  free(cur);
  return nproc;
}

/* k <= 64 BFS queries at once on the host cores (multi-source BFS with one
 * bit per query, Then et al., VLDB 2014): seen[v] holds the queries that
 * reached v, F[v] those with v in their frontier.  A top-down level hands
 * each out-neighbour w the queries F[v] & ~seen[w]; when symmetric != 0 and
 * the frontier's edges exceed 1/8 of the edges still unsettled, a level
 * runs bottom-up instead: every vertex ORs the frontier masks of its
 * neighbours (its in-neighbours, on a symmetric graph) and stops when it has
 * every query it misses.  Either way query q reaches w at level L+1 exactly
 * when w is unreached by q and adjacent to q's level-L frontier, so each
 * query's values are those of orc_bfs.  Per query: the digest of
 * orc_hash_values (sentinel n+1), the reached count and the largest value;
 * out_full (nullable, k x n, query-major) gets the value arrays. */
ORC_EXPORT void orc_bfs_multi(uint64_t n, const uint64_t *off, const uint32_t *adj,
                              const uint32_t *srcs, uint32_t k, int symmetric,
                              uint64_t *out_hash, uint64_t *out_reached,
                              uint32_t *out_depth, uint32_t *out_full) {
  if (k > 64) k = 64;
  uint64_t *seen = (uint64_t *)calloc(n ? n : 1, 8);
  uint64_t *F = (uint64_t *)calloc(n ? n : 1, 8);
  uint64_t *N = (uint64_t *)calloc(n ? n : 1, 8);
  uint64_t hash[64] = {0}, reached[64] = {0};
  uint32_t depth[64] = {0};
  const uint64_t all = k == 64 ? ~0ull : ((1ull << k) - 1);
  if (out_full) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)(k * n); i++) out_full[i] = (uint32_t)(n + 1);
  }
  for (uint32_t q = 0; q < k; q++) {
    if (srcs[q] >= n) continue;
    seen[srcs[q]] |= 1ull << q;
    F[srcs[q]] |= 1ull << q;
    hash[q] = orc_mix64(((uint64_t)srcs[q] << 32) | 1u);
    reached[q] = 1;
    depth[q] = 1;
    if (out_full) out_full[(uint64_t)q * n + srcs[q]] = 1;
  }
  uint64_t m = off[n], settled = 0;  /* edges of vertices every query reached */
  for (uint32_t level = 1;; level++) {
    uint64_t fe = 0;
#pragma omp parallel for schedule(static) reduction(+ : fe)
    for (long long v = 0; v < (long long)n; v++)
      if (F[v]) fe += off[v + 1] - off[v];
    const int bu = symmetric && fe > (m - settled) / 8;
    if (bu) {
#pragma omp parallel for schedule(dynamic, 1024)
      for (long long wv = 0; wv < (long long)n; wv++) {
        const uint64_t miss = all & ~seen[wv];
        uint64_t acc = 0;
        if (miss)
          for (uint64_t i = off[wv]; i < off[wv + 1] && (acc & miss) != miss; i++)
            acc |= F[adj[i]];
        N[wv] = acc & miss;
      }
    } else {
#pragma omp parallel for schedule(dynamic, 1024)
      for (long long v = 0; v < (long long)n; v++) {
        const uint64_t f = F[v];
        if (!f) continue;
        for (uint64_t i = off[v]; i < off[v + 1]; i++) {
          const uint32_t wv = adj[i];
          const uint64_t add = f & ~seen[wv];
          if (add && (__atomic_load_n(&N[wv], __ATOMIC_RELAXED) & add) != add)
            __atomic_fetch_or(&N[wv], add, __ATOMIC_RELAXED);
        }
      }
    }
    int any = 0;
    uint64_t newly_settled = 0;
#pragma omp parallel reduction(| : any) reduction(+ : newly_settled)
    {
      uint64_t h[64] = {0}, r[64] = {0};
#pragma omp for schedule(static)
      for (long long wv = 0; wv < (long long)n; wv++) {
        const uint64_t gain = N[wv] & ~seen[wv];
        N[wv] = 0;
        F[wv] = gain;
        if (!gain) continue;
        any = 1;
        seen[wv] |= gain;
        if ((seen[wv] & all) == all) newly_settled += off[wv + 1] - off[wv];
        const uint64_t hv = orc_mix64(((uint64_t)wv << 32) | (level + 1));
        for (uint64_t g = gain; g; g &= g - 1) {
          const int q = __builtin_ctzll(g);
          h[q] += hv;
          r[q]++;
          if (out_full) out_full[(uint64_t)q * n + wv] = level + 1;
        }
      }
#pragma omp critical
      for (uint32_t q = 0; q < k; q++) {
        hash[q] += h[q];
        reached[q] += r[q];
        if (r[q] && depth[q] < level + 1) depth[q] = level + 1;
      }
    }
    settled += newly_settled;
    if (!any) break;
  }
  for (uint32_t q = 0; q < k; q++) {
    if (out_hash) out_hash[q] = hash[q];
    if (out_reached) out_reached[q] = reached[q];
    if (out_depth) out_depth[q] = depth[q];
  }
  free(seen);
  free(F);
  free(N);
}

/* PageRank pull (Jacobi over the active set) on the CSC, in FP32 ranks like
 * the GPU path; returns the iteration count. */
ORC_EXPORT int orc_pr_pull_mt(uint64_t n, const uint64_t *ioff,
                              const uint32_t *isrc, const uint32_t *outdeg,
                              float alpha, float tol, int max_iter,
                              float *rank) {
  float *contrib = (float *)malloc((n ? n : 1) * sizeof(float));
  float *nr = (float *)malloc((n ? n : 1) * sizeof(float));
  uint8_t *act = (uint8_t *)malloc(n ? n : 1);
#pragma omp parallel for schedule(static)
  for (long long v = 0; v < (long long)n; v++) {
    rank[v] = 1.0f / (float)n;
    act[v] = 1;
  }
  int it = 0;
  for (; it < max_iter; it++) {
#pragma omp parallel for schedule(static)
    for (long long v = 0; v < (long long)n; v++)
      contrib[v] = outdeg[v] ? rank[v] / (float)outdeg[v] : 0.0f;
    uint64_t nact = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : nact)
    for (long long v = 0; v < (long long)n; v++) {
      if (!act[v]) { nr[v] = rank[v]; continue; }
      float s = 0.0f;
      for (uint64_t i = ioff[v]; i < ioff[v + 1]; i++) s += contrib[isrc[i]];
      float x = (1.0f - alpha) + alpha * s;
      act[v] = fabsf(x - rank[v]) > tol;
      nr[v] = x;
      nact += act[v];
    }
    memcpy(rank, nr, n * sizeof(float));
    if (!nact) { it++; break; }
  }
  free(contrib);
  free(nr);
  free(act);
  return it;
}

static inline void atomic_add_f32(float *p, float v) {
  uint32_t *ip = (uint32_t *)p;
  uint32_t old = __atomic_load_n(ip, __ATOMIC_RELAXED), nw;
  do {
    float f;
    memcpy(&f, &old, 4);
    f += v;
    memcpy(&nw, &f, 4);
  } while (!__atomic_compare_exchange_n(ip, &old, nw, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}

/* PageRank push (residual) on the CSR in FP32 with atomic adds. */
ORC_EXPORT int orc_pr_push_mt(uint64_t n, const uint64_t *off,
                              const uint32_t *adj, float alpha, float tol,
                              int max_iter, float *rank) {
  float *delta = (float *)malloc((n ? n : 1) * sizeof(float));
  float *res = (float *)calloc(n ? n : 1, sizeof(float));
  uint8_t *act = (uint8_t *)malloc(n ? n : 1);
#pragma omp parallel for schedule(static)
  for (long long v = 0; v < (long long)n; v++) {
    uint64_t d = off[v + 1] - off[v];
    rank[v] = 1.0f - alpha;
    delta[v] = d ? (1.0f - alpha) * alpha / (float)d : 0.0f;
    act[v] = 1;
  }
  int it = 0;
  for (; it < max_iter; it++) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (long long u = 0; u < (long long)n; u++) {
      if (!act[u]) continue;
      for (uint64_t i = off[u]; i < off[u + 1]; i++) atomic_add_f32(&res[adj[i]], delta[u]);
    }
    uint64_t nact = 0;
#pragma omp parallel for schedule(static) reduction(+ : nact)
    for (long long v = 0; v < (long long)n; v++) {
      act[v] = 0;
      if (res[v] > tol) {
        uint64_t d = off[v + 1] - off[v];
        rank[v] += res[v];
        delta[v] = d ? res[v] * alpha / (float)d : 0.0f;
        res[v] = 0.0f;
        act[v] = 1;
        nact++;
      }
    }
    if (!nact) { it++; break; }
  }
  free(delta);
  free(res);
  free(act);
  return it;
}
