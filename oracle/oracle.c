/*
 * oracle.c — CPU oracle for the QBC / MHL-QBC / RQ betweenness path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker or
 * the timed CPU baseline; the product path never calls it.
 *
 * It restates, with textbook algorithms, what the reference's path computes:
 *  - orc_sssp: distance, shortest-path count and max hop layer from one source
 *    to every vertex.  The reference obtains these from labels
 *    (MHLQBC.cu:334-412: distance, count, layer = my_cnt max-hop) and, for QBC,
 *    from the host Kahn pass (QBC.cu:485-555: sigma, layer = Kahn level).
 *    Here: Dijkstra with counting; a vertex's count is the sum of its shortest
 *    predecessors' counts and its hop the max of theirs plus one, which is the
 *    same quantity as TreeCount's max-hop label and QBC's Kahn level.
 *  - orc_brandes: Brandes dependency accumulation (QBC.cu:379-397,
 *    MHLQBC.cu:452-472): delta[u] += sigma_u * (1 + delta_v) / sigma_v over DAG
 *    arcs u->v, deepest first, bc[v] += 0.5 * delta[v] for v != s.
 *  - orc_rq_pairs: RQ's pair formula (RQ.cu:451-459): for every unordered pair
 *    {s, t} with s < t and s, t != v: if d(s,t) == d(v,s) + d(v,t) then
 *    bc(v) += sigma(v,s) * sigma(v,t) / sigma(s,t).  All-pairs, tiny graphs only.
 * Counts are uint64 with an overflow flag (the reference uses 32-bit counts).
 *
 * Parity pinning (tests/test_oracle.py): closed forms on paths, stars, odd
 * cycles, complete graphs and unit grids (sigma = C(2k-2, k-1), overflow at
 * k = 35), networkx weighted betweenness on random graphs, and the identity
 * rq_pairs(v) == brandes(v).  The graph is given as a symmetric CSR (arcs
 * u->v and v->u with equal weight), 0-based.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF INT64_MAX

typedef struct {
  int64_t *key;
  int32_t *v;
  int64_t size, cap;
} heap_t;

static int heap_push(heap_t *h, int64_t k, int32_t v) {
  if (h->size == h->cap) {
    int64_t nc = h->cap ? 2 * h->cap : 1024;
    int64_t *nk = (int64_t *)realloc(h->key, (size_t)nc * sizeof(int64_t));
    if (!nk) return -1;
    h->key = nk;
    int32_t *nv = (int32_t *)realloc(h->v, (size_t)nc * sizeof(int32_t));
    if (!nv) return -1;
    h->v = nv;
    h->cap = nc;
  }
  int64_t i = h->size++;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h->key[p] < k || (h->key[p] == k && h->v[p] <= v)) break;
    h->key[i] = h->key[p];
    h->v[i] = h->v[p];
    i = p;
  }
  h->key[i] = k;
  h->v[i] = v;
  return 0;
}

static void heap_pop(heap_t *h, int64_t *k, int32_t *v) {
  *k = h->key[0];
  *v = h->v[0];
  int64_t lk = h->key[--h->size];
  int32_t lv = h->v[h->size];
  int64_t i = 0;
  for (;;) {
    int64_t c = 2 * i + 1;
    if (c >= h->size) break;
    if (c + 1 < h->size &&
        (h->key[c + 1] < h->key[c] || (h->key[c + 1] == h->key[c] && h->v[c + 1] < h->v[c])))
      ++c;
    if (lk < h->key[c] || (lk == h->key[c] && lv <= h->v[c])) break;
    h->key[i] = h->key[c];
    h->v[i] = h->v[c];
    i = c;
  }
  h->key[i] = lk;
  h->v[i] = lv;
}

/* Single-source distances, counts, hops; `order` (may be NULL) receives the
 * settled vertices in nondecreasing distance.  Returns the number settled,
 * or -1 on allocation failure.  Unreached vertices get d = -1, sigma = 0. */
int64_t orc_sssp(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w,
                 int32_t s, int64_t *d, uint64_t *sigma, uint32_t *hop, int32_t *order,
                 int32_t *overflow) {
  heap_t h = {0, 0, 0, 0};
  char *done = (char *)calloc((size_t)n, 1);
  if (!done) return -1;
  for (int32_t i = 0; i < n; ++i) {
    d[i] = ORC_INF;
    sigma[i] = 0;
    hop[i] = 0;
  }
  d[s] = 0;
  sigma[s] = 1;
  int64_t settled = 0;
  if (heap_push(&h, 0, s)) goto oom;
  while (h.size) {
    int64_t k;
    int32_t u;
    heap_pop(&h, &k, &u);
    if (done[u] || k != d[u]) continue;
    done[u] = 1;
    if (order) order[settled] = u;
    ++settled;
    for (int64_t a = row[u]; a < row[u + 1]; ++a) {
      int32_t v = col[a];
      int64_t nd = d[u] + w[a];
      if (nd < d[v]) {
        d[v] = nd;
        sigma[v] = sigma[u];
        hop[v] = hop[u] + 1;
        if (heap_push(&h, nd, v)) goto oom;
      } else if (nd == d[v]) {
        uint64_t r;
        if (__builtin_add_overflow(sigma[v], sigma[u], &r)) {
          if (overflow) *overflow = 1;
          r = UINT64_MAX;
        }
        sigma[v] = r;
        if (hop[u] + 1 > hop[v]) hop[v] = hop[u] + 1;
      }
    }
  }
  for (int32_t i = 0; i < n; ++i)
    if (d[i] == ORC_INF) d[i] = -1;
  free(h.key);
  free(h.v);
  free(done);
  return settled;
oom:
  free(h.key);
  free(h.v);
  free(done);
  return -1;
}

/* Adds 0.5 * delta_s(v) for every source in `sources` into bc[n]. */
int orc_brandes(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w,
                const int32_t *sources, int32_t ns, double *bc, int32_t *overflow) {
  int64_t *d = (int64_t *)malloc((size_t)n * sizeof(int64_t));
  uint64_t *sg = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
  uint32_t *hp = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  int32_t *ord = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  double *delta = (double *)malloc((size_t)n * sizeof(double));
  int rc = 0;
  if (!d || !sg || !hp || !ord || !delta) {
    rc = -1;
    goto out;
  }
  for (int32_t k = 0; k < ns; ++k) {
    int32_t s = sources[k];
    int64_t m = orc_sssp(n, row, col, w, s, d, sg, hp, ord, overflow);
    if (m < 0) {
      rc = -1;
      goto out;
    }
    for (int32_t i = 0; i < n; ++i) delta[i] = 0.0;
    for (int64_t k2 = m - 1; k2 >= 0; --k2) {
      int32_t u = ord[k2];
      double acc = 0.0;
      for (int64_t a = row[u]; a < row[u + 1]; ++a) {
        int32_t v = col[a];
        if (d[v] == d[u] + w[a]) acc += (double)sg[u] * (1.0 + delta[v]) / (double)sg[v];
      }
      delta[u] = acc;
    }
    for (int32_t i = 0; i < n; ++i)
      if (i != s) bc[i] += 0.5 * delta[i];
  }
out:
  free(d);
  free(sg);
  free(hp);
  free(ord);
  free(delta);
  return rc;
}

/* orc_brandes over `threads` contiguous slices of the source list, each into
 * its own array, then added into bc[n] in slice order (OpenMP; the result is
 * the same for a given thread count, and within rounding of orc_brandes). */
int orc_brandes_mt(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w,
                   const int32_t *sources, int32_t ns, double *bc, int32_t *overflow,
                   int32_t threads) {
  if (threads < 1) threads = 1;
  if (threads > ns) threads = ns > 0 ? ns : 1;
  double *part = (double *)calloc((size_t)threads * (size_t)n, sizeof(double));
  int32_t *ovf = (int32_t *)calloc((size_t)threads, sizeof(int32_t));
  int *rcs = (int *)calloc((size_t)threads, sizeof(int));
  int rc = 0;
  if (!part || !ovf || !rcs) {
    rc = -1;
    goto out;
  }
#pragma omp parallel for num_threads(threads) schedule(static, 1)
  for (int32_t t = 0; t < threads; ++t) {
    const int32_t a = (int32_t)((int64_t)ns * t / threads), b = (int32_t)((int64_t)ns * (t + 1) / threads);
    rcs[t] = orc_brandes(n, row, col, w, sources + a, b - a, part + (size_t)t * n, &ovf[t]);
  }
  for (int32_t t = 0; t < threads; ++t) {
    if (rcs[t]) rc = -1;
    if (ovf[t] && overflow) *overflow = 1;
    for (int32_t i = 0; i < n; ++i) bc[i] += part[(size_t)t * n + i];
  }
out:
  free(part);
  free(ovf);
  free(rcs);
  return rc;
}

/* Reference-method BC on the CPU, the timed CPU baseline (and checked
 * against orc_brandes).  Per source s, restating MHLQBC.cu's BC_Kernel on the
 * host:
 *  - LCA depth of (s, t) by a range minimum over an Euler tour of the tree
 *    (MHLQBC.cu:302-331; TreeDecomp::MakeRMQIndex, TreeDecomp.cpp:324-345);
 *  - the label query over the common prefix of the two rows, min distance,
 *    summed count of the argmin entries, max hop among them
 *    (MHLQBC.cu:356-404), with a node's own entry (0, 1, 0) standing for the
 *    ancestor case;
 *  - DAG arcs u->v where d[v] == d[u] + w (MHLQBC.cu:415-432), layers by max
 *    hop (MHLQBC.cu:434-450) and the back-propagation deepest layer first,
 *    delta[u] = sum over DAG arcs sigma_u * (1 + delta_v) / sigma_v
 *    (MHLQBC.cu:452-466, pulled per tail instead of pushed per arc);
 *  - bc[v] += 0.5 * delta[v] for v != s (MHLQBC.cu:470).
 * Inputs by vertex id: parent[v] (-1 at the root), label rows over the
 * ancestors root first ([label_off[v], label_off[v + 1]), depth(v) entries:
 * restricted distance, count, hop), the symmetric CSR.  OpenMP over `threads`
 * slices of the sources, summed in slice order.  Returns 0, or -1 on
 * allocation failure or a tree that is not one tree. */
typedef struct {
  int32_t n, levels;
  int32_t *depth, *first, *st; /* st[k * m + i] = min depth over tour[i, i + 2^k) */
  int32_t m;
} orc_tree;

static void orc_tree_free(orc_tree *T) {
  free(T->depth);
  free(T->first);
  free(T->st);
}

static int orc_tree_build(int32_t n, const int32_t *parent, orc_tree *T) {
  memset(T, 0, sizeof *T);
  T->n = n;
  int32_t *cho = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  int32_t *ch = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  int32_t *stk = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  int32_t *nxt = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  T->depth = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  T->first = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  T->m = 2 * n - 1;
  int32_t *tour = (int32_t *)malloc((size_t)T->m * sizeof(int32_t));
  int rc = -1, root = -1;
  if (!cho || !ch || !stk || !nxt || !T->depth || !T->first || !tour) goto out;
  for (int32_t v = 0; v < n; ++v) {
    if (parent[v] < 0) {
      if (root >= 0) goto out;
      root = v;
    } else {
      cho[parent[v] + 1]++;
    }
  }
  if (root < 0) goto out;
  for (int32_t v = 0; v < n; ++v) cho[v + 1] += cho[v];
  for (int32_t v = 0; v < n; ++v) nxt[v] = cho[v];
  for (int32_t v = 0; v < n; ++v)
    if (parent[v] >= 0) ch[nxt[parent[v]]++] = v;
  for (int32_t v = 0; v < n; ++v) nxt[v] = cho[v];
  {
    int32_t sp = 0, len = 0, seen = 0;
    stk[sp++] = root;
    T->depth[root] = 0;
    T->first[root] = 0;
    tour[len++] = root;
    ++seen;
    while (sp > 0) {
      const int32_t v = stk[sp - 1];
      if (nxt[v] < cho[v + 1]) {
        const int32_t c = ch[nxt[v]++];
        T->depth[c] = T->depth[v] + 1;
        T->first[c] = len;
        tour[len++] = c;
        stk[sp++] = c;
        ++seen;
      } else {
        --sp;
        if (sp > 0) tour[len++] = stk[sp - 1];
      }
      if (len > T->m) goto out;
    }
    if (seen != n || len != T->m) goto out;
  }
  T->levels = 1;
  while ((1 << T->levels) <= T->m) ++T->levels;
  T->st = (int32_t *)malloc((size_t)T->levels * (size_t)T->m * sizeof(int32_t));
  if (!T->st) goto out;
  for (int32_t i = 0; i < T->m; ++i) T->st[i] = T->depth[tour[i]];
  for (int32_t k = 1; k < T->levels; ++k) {
    const int32_t h = 1 << (k - 1);
    const int32_t *a = T->st + (size_t)(k - 1) * T->m;
    int32_t *b = T->st + (size_t)k * T->m;
    for (int32_t i = 0; i < T->m; ++i) b[i] = (i + h < T->m && a[i + h] < a[i]) ? a[i + h] : a[i];
  }
  rc = 0;
out:
  free(cho);
  free(ch);
  free(stk);
  free(nxt);
  free(tour);
  if (rc) orc_tree_free(T);
  return rc;
}

static int32_t orc_lca_depth(const orc_tree *T, int32_t s, int32_t t) {
  int32_t a = T->first[s], b = T->first[t];
  if (a > b) {
    int32_t x = a;
    a = b;
    b = x;
  }
  int32_t k = 0;
  while ((2 << k) <= b - a + 1) ++k;
  const int32_t x = T->st[(size_t)k * T->m + a], y = T->st[(size_t)k * T->m + b - (1 << k) + 1];
  return x < y ? x : y;
}

static int orc_label_bc_slice(const orc_tree *T, const int64_t *label_off, const int32_t *dis,
                              const uint64_t *cnt, const uint32_t *hop, const int64_t *row,
                              const int32_t *col, const int32_t *w, const int32_t *sources, int32_t ns,
                              double *bc) {
  const int32_t n = T->n;
  int64_t *d = (int64_t *)malloc((size_t)n * sizeof(int64_t));
  uint64_t *sg = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
  uint32_t *hp = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  int32_t *ord = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  int32_t *bin = (int32_t *)malloc(((size_t)n + 2) * sizeof(int32_t));
  double *delta = (double *)malloc((size_t)n * sizeof(double));
  int rc = 0;
  if (!d || !sg || !hp || !ord || !bin || !delta) {
    rc = -1;
    goto out;
  }
  for (int32_t k = 0; k < ns; ++k) {
    const int32_t s = sources[k];
    const int64_t rs = label_off[s];
    const int32_t ds = T->depth[s];
    uint32_t maxh = 0;
    for (int32_t t = 0; t < n; ++t) {
      const int32_t L = orc_lca_depth(T, s, t), dt = T->depth[t];
      const int64_t rt = label_off[t];
      int64_t best = INT64_MAX;
      uint64_t c = 0;
      uint32_t h = 0;
      for (int32_t i = 0; i <= L; ++i) {
        const int64_t a = i == ds ? 0 : dis[rs + i], b = i == dt ? 0 : dis[rt + i];
        const uint64_t ca = i == ds ? 1 : cnt[rs + i], cb = i == dt ? 1 : cnt[rt + i];
        const uint32_t ha = i == ds ? 0 : hop[rs + i], hb = i == dt ? 0 : hop[rt + i];
        const int64_t v = a + b;
        if (v < best) {
          best = v;
          c = ca * cb;
          h = ha + hb;
        } else if (v == best) {
          c += ca * cb;
          if (ha + hb > h) h = ha + hb;
        }
      }
      d[t] = best;
      sg[t] = c;
      hp[t] = h;
      if (h > maxh) maxh = h;
    }
    /* layers by max hop: counting sort */
    for (uint32_t l = 0; l <= maxh + 1; ++l) bin[l] = 0;
    for (int32_t t = 0; t < n; ++t) bin[hp[t] + 1]++;
    for (uint32_t l = 0; l <= maxh; ++l) bin[l + 1] += bin[l];
    for (int32_t t = 0; t < n; ++t) ord[bin[hp[t]]++] = t;
    for (int32_t t = 0; t < n; ++t) delta[t] = 0.0;
    for (int32_t q = n - 1; q >= 0; --q) { /* deepest layer first */
      const int32_t u = ord[q];
      double acc = 0.0;
      for (int64_t a = row[u]; a < row[u + 1]; ++a) {
        const int32_t v = col[a];
        if (d[v] == d[u] + w[a]) acc += (double)sg[u] * (1.0 + delta[v]) / (double)sg[v];
      }
      delta[u] = acc;
    }
    for (int32_t t = 0; t < n; ++t)
      if (t != s) bc[t] += 0.5 * delta[t];
  }
out:
  free(d);
  free(sg);
  free(hp);
  free(ord);
  free(bin);
  free(delta);
  return rc;
}

int orc_label_bc(int32_t n, const int32_t *parent, const int64_t *label_off, const int32_t *dis,
                 const uint64_t *cnt, const uint32_t *hop, const int64_t *row, const int32_t *col,
                 const int32_t *w, const int32_t *sources, int32_t ns, double *bc, int32_t threads) {
  orc_tree T;
  if (orc_tree_build(n, parent, &T)) return -1;
  if (threads < 1) threads = 1;
  if (threads > ns) threads = ns > 0 ? ns : 1;
  double *part = (double *)calloc((size_t)threads * (size_t)n, sizeof(double));
  int *rcs = (int *)calloc((size_t)threads, sizeof(int));
  int rc = 0;
  if (!part || !rcs) {
    rc = -1;
    goto out;
  }
#pragma omp parallel for num_threads(threads) schedule(static, 1)
  for (int32_t t = 0; t < threads; ++t) {
    const int32_t a = (int32_t)((int64_t)ns * t / threads), b = (int32_t)((int64_t)ns * (t + 1) / threads);
    rcs[t] = orc_label_bc_slice(&T, label_off, dis, cnt, hop, row, col, w, sources + a, b - a,
                                part + (size_t)t * n);
  }
  for (int32_t t = 0; t < threads; ++t) {
    if (rcs[t]) rc = -1;
    for (int32_t i = 0; i < n; ++i) bc[i] += part[(size_t)t * n + i];
  }
out:
  free(part);
  free(rcs);
  orc_tree_free(&T);
  return rc;
}

/* RQ pair formula over all unordered pairs; O(n^2 q) after n Dijkstras. */
int orc_rq_pairs(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w,
                 const int32_t *points, int32_t q, double *out) {
  int64_t *D = (int64_t *)malloc((size_t)n * (size_t)n * sizeof(int64_t));
  uint64_t *S = (uint64_t *)malloc((size_t)n * (size_t)n * sizeof(uint64_t));
  uint32_t *H = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  int rc = 0;
  int32_t ovf = 0;
  if (!D || !S || !H) {
    rc = -1;
    goto out;
  }
  for (int32_t s = 0; s < n; ++s)
    if (orc_sssp(n, row, col, w, s, D + (size_t)s * n, S + (size_t)s * n, H, NULL, &ovf) < 0) {
      rc = -1;
      goto out;
    }
  for (int32_t i = 0; i < q; ++i) {
    int32_t v = points[i];
    double acc = 0.0;
    for (int32_t s = 0; s < n; ++s) {
      if (s == v) continue;
      for (int32_t t = s + 1; t < n; ++t) {
        if (t == v) continue;
        int64_t dst = D[(size_t)s * n + t];
        if (dst < 0 || D[(size_t)v * n + s] < 0 || D[(size_t)v * n + t] < 0) continue;
        if (dst == D[(size_t)v * n + s] + D[(size_t)v * n + t])
          acc += (double)S[(size_t)v * n + s] * (double)S[(size_t)v * n + t] /
                 (double)S[(size_t)s * n + t];
      }
    }
    out[i] = acc;
  }
out:
  free(D);
  free(S);
  free(H);
  return rc;
}
