/*
 * oracle.c -- CPU restatement of Hornet's host-side reference path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product in
 * hornet_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py load it; the product never calls into it.
 *
 * Every function cites the reference file:line it restates.  The pins that
 * check each function independently (golden fixtures, networkx, dense brute
 * force, the reference's own insert/delete multiset identities) live in
 * tests/test_oracle.py.
 *
 *   orc_rmat_pairs / orc_uniform_pairs / orc_pick_sources / orc_chung_lu_pairs
 *        seeded generators.  The reference ships no R-MAT / Chung-Lu code
 *        (its RANDOMIZE and generateBatch are time-seeded,
 *        xlib/src/Graph/GraphStd.cpp:220-224, hornet/src/Util/BatchFunctions.cpp:56-58),
 *        so the generator is specified here; parity unpinned against the
 *        reference, pinned against the GPU generator and by invariants.
 *   orc_coo_to_csr    xlib/src/Graph/GraphStd.cpp:173-341 (symmetrize, dedup, CSR)
 *   orc_directed_by_degree  xlib/src/Graph/GraphStd.cpp:247-304
 *   orc_bfs           xlib/src/Graph/BFS.cpp:54-76
 *   orc_host_tc       hornetsnest/test/TriangleTest2.cu:22-69
 *   orc_tc2_per_vertex  hornetsnest/src/Static/TriangleCounting/triangle2.cu:75-144
 *                     with primitives/Operator++.i.cuh:97-246 (who is credited)
 *   orc_max_degree_id xlib/include/Graph/GraphStd.i.hpp:296-299
 *   orc_insert / orc_erase  host set model of
 *                     hornet/include/Core/BatchUpdate/BatchUpdate.i.cuh:337-392,485-524
 *                     and BatchUpdateKernels.cuh:27-58,263-300,335-393
 *   orc_two_hop       |{v} ∪ N(v) ∪ N(N(v))| (definition owned by this build)
 *   orc_*_mt          the same results split over POSIX threads (generators,
 *                     COO -> CSR, host triangle count), so the configured graph
 *                     sizes (2^22..2^25 vertices) are checked in seconds
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

/* ------------------------------------------------------------------------ */
/* counter-based RNG: SplitMix64 finaliser over a per-seed stream            */
/* ------------------------------------------------------------------------ */
#define GOLDEN 0x9E3779B97F4A7C15ULL

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
/* key of a stream = first SplitMix64 output for state `seed` */
uint64_t orc_stream_key(uint64_t seed) { return mix64(seed + GOLDEN); }
/* draw number `ctr` of the stream */
static inline uint64_t draw(uint64_t key, uint64_t ctr) {
    return mix64(key + (ctr + 1) * GOLDEN);
}
uint64_t orc_draw(uint64_t key, uint64_t ctr) { return draw(key, ctr); }

/* uniform integer in [0, n) by multiply-shift of the high 32 bits */
static inline uint32_t uniform_below(uint64_t r, uint32_t n) {
    return (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
}

/* R-MAT quadrant thresholds a=.57 b=.19 c=.19 d=.05 as integers of 2^32 */
#define RMAT_TA   ((uint32_t)((57ULL << 32) / 100))
#define RMAT_TAB  ((uint32_t)((76ULL << 32) / 100))
#define RMAT_TABC ((uint32_t)((95ULL << 32) / 100))

/*
 * R-MAT pair i: for each level l in [0, scale) draw r = high32(draw(key, i*64+l)).
 * r < TA -> (0,0), r < TAB -> (0,1), r < TABC -> (1,0), else (1,1);
 * the row bit goes to src bit (scale-1-l), the column bit to dst bit.
 */
void orc_rmat_pairs(int scale, int64_t first, int64_t count, uint64_t seed,
                    int32_t* src, int32_t* dst) {
    uint64_t key = orc_stream_key(seed);
    for (int64_t k = 0; k < count; k++) {
        uint64_t i = (uint64_t)(first + k);
        uint32_t s = 0, d = 0;
        for (int l = 0; l < scale; l++) {
            uint32_t r = (uint32_t)(draw(key, i * 64 + (uint64_t)l) >> 32);
            uint32_t row = r >= RMAT_TAB;
            uint32_t col = (r >= RMAT_TA && r < RMAT_TAB) || r >= RMAT_TABC;
            s |= row << (scale - 1 - l);
            d |= col << (scale - 1 - l);
        }
        src[k] = (int32_t)s;
        dst[k] = (int32_t)d;
    }
}

/*
 * Update batch of `count` uniform pairs over [0, n).  Pair i draws
 * r0 = draw(key, 3i); if i > 0 and high32(r0) % 100 == 0 the pair repeats
 * pair j = uniform_below(r0', i) with r0' = draw(key, 3i+1) (the ~1 % in-batch
 * duplicates); otherwise u = uniform_below(draw(key,3i+1), n),
 * v = uniform_below(draw(key,3i+2), n).
 */
void orc_uniform_pairs(int32_t n, int64_t count, uint64_t seed,
                       int32_t* src, int32_t* dst) {
    uint64_t key = orc_stream_key(seed);
    for (int64_t i = 0; i < count; i++) {
        uint64_t r0 = draw(key, 3 * (uint64_t)i);
        if (i > 0 && ((uint32_t)(r0 >> 32)) % 100u == 0u) {
            int64_t j = (int64_t)(((draw(key, 3 * (uint64_t)i + 1) >> 32) *
                                   (uint64_t)i) >> 32);
            src[i] = src[j];
            dst[i] = dst[j];
        } else {
            src[i] = (int32_t)uniform_below(draw(key, 3 * (uint64_t)i + 1), (uint32_t)n);
            dst[i] = (int32_t)uniform_below(draw(key, 3 * (uint64_t)i + 2), (uint32_t)n);
        }
    }
}

/* k sources: draw j = 0,1,... v = uniform_below(draw(key,j), n), keep v with
 * deg(v) > 0.  Returns the number of draws consumed (-1 if none found in 64n). */
int64_t orc_pick_sources(int32_t n, const int64_t* off, int32_t k, uint64_t seed,
                         int32_t* out) {
    uint64_t key = orc_stream_key(seed);
    int32_t got = 0;
    uint64_t j = 0;
    uint64_t limit = 64ULL * (uint64_t)n + 64;
    while (got < k && j < limit) {
        uint32_t v = uniform_below(draw(key, j), (uint32_t)n);
        j++;
        if (off[v + 1] > off[v]) out[got++] = (int32_t)v;
    }
    return got == k ? (int64_t)j : -1;
}

/* ------------------------------------------------------------------------ */
/* sorting helpers (LSD radix on uint64, 16-bit digits)                      */
/* ------------------------------------------------------------------------ */
static void radix_sort_u64(uint64_t* a, int64_t n, int key_bits) {
    if (n <= 1) return;
    uint64_t* tmp = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
    int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * 65536);
    uint64_t* from = a;
    uint64_t* to = tmp;
    for (int shift = 0; shift < key_bits; shift += 16) {
        memset(cnt, 0, sizeof(int64_t) * 65536);
        for (int64_t i = 0; i < n; i++) cnt[(from[i] >> shift) & 0xFFFF]++;
        int64_t s = 0;
        for (int d = 0; d < 65536; d++) { int64_t c = cnt[d]; cnt[d] = s; s += c; }
        for (int64_t i = 0; i < n; i++) to[cnt[(from[i] >> shift) & 0xFFFF]++] = from[i];
        uint64_t* t = from; from = to; to = t;
    }
    if (from != a) memcpy(a, from, (size_t)n * sizeof(uint64_t));
    free(tmp);
    free(cnt);
}

static inline uint64_t pack(int32_t s, int32_t d) {
    return ((uint64_t)(uint32_t)s << 32) | (uint32_t)d;
}

/* keys sorted ascending -> CSR (offsets n+1, adjacency) */
static void keys_to_csr(int32_t n, const uint64_t* keys, int64_t m,
                        int64_t* off, int32_t* adj) {
    memset(off, 0, sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t i = 0; i < m; i++) off[(keys[i] >> 32) + 1]++;
    for (int32_t v = 0; v < n; v++) off[v + 1] += off[v];
    for (int64_t i = 0; i < m; i++) adj[i] = (int32_t)(uint32_t)keys[i];
}

/*
 * COO -> CSR, following GraphStd::COOtoCSR (xlib/src/Graph/GraphStd.cpp:173-341)
 * for a directed file read as UNDIRECTED: every non-self-loop (s,d) also adds
 * (d,s) (:175-190), then sort + unique (:192-206), then the CSR build with
 * ascending adjacency (:333-341).  `drop_self_loops` additionally removes the
 * single copy of a self loop the reference keeps (the generated workloads are
 * specified without self loops).  `symmetrize`=0 keeps arcs as given
 * (DIMACS9 read, :192 not taken) but still sorts when `sort`=1 (:306-310);
 * `dedup` is the sort+unique of :198-200.
 * Returns the number of directed entries; off/adj must hold n+1 / 2m.
 */
int64_t orc_coo_to_csr(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                       int symmetrize, int drop_self_loops, int dedup, int sort,
                       int64_t* off, int32_t* adj) {
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(2 * m + 1));
    int64_t k = 0;
    for (int64_t i = 0; i < m; i++) {
        if (drop_self_loops && src[i] == dst[i]) continue;
        keys[k++] = pack(src[i], dst[i]);
    }
    if (symmetrize) {
        int64_t half = k;
        for (int64_t i = 0; i < half; i++) {
            int32_t s = (int32_t)(keys[i] >> 32), d = (int32_t)(uint32_t)keys[i];
            if (s == d) continue;
            keys[k++] = pack(d, s);
        }
    }
    if (sort || dedup) {
        int bits = 64;
        radix_sort_u64(keys, k, bits);
    }
    if (dedup) {
        int64_t w = 0;
        for (int64_t i = 0; i < k; i++)
            if (w == 0 || keys[i] != keys[w - 1]) keys[w++] = keys[i];
        k = w;
    }
    if (!sort && !dedup) {
        /* stable counting placement, file order inside each vertex */
        memset(off, 0, sizeof(int64_t) * ((size_t)n + 1));
        for (int64_t i = 0; i < k; i++) off[(keys[i] >> 32) + 1]++;
        for (int32_t v = 0; v < n; v++) off[v + 1] += off[v];
        int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        memcpy(pos, off, sizeof(int64_t) * (size_t)n);
        for (int64_t i = 0; i < k; i++) adj[pos[keys[i] >> 32]++] = (int32_t)(uint32_t)keys[i];
        free(pos);
    } else {
        keys_to_csr(n, keys, k, off, adj);
    }
    free(keys);
    return k;
}

/*
 * DIRECTED_BY_DEGREE (xlib/src/Graph/GraphStd.cpp:247-304): keep (u,v) iff
 * deg(u) < deg(v) or (deg(u) == deg(v) and u < v), degrees taken from the input
 * CSR; adjacency order is preserved (sorted input -> sorted output).
 */
int64_t orc_directed_by_degree(int32_t n, const int64_t* off, const int32_t* adj,
                               int64_t* off_out, int32_t* adj_out) {
    int64_t w = 0;
    off_out[0] = 0;
    for (int32_t u = 0; u < n; u++) {
        int64_t du = off[u + 1] - off[u];
        for (int64_t j = off[u]; j < off[u + 1]; j++) {
            int32_t v = adj[j];
            int64_t dv = off[v + 1] - off[v];
            if (du < dv || (du == dv && u < v)) adj_out[w++] = v;
        }
        off_out[u + 1] = w;
    }
    return w;
}

/*
 * Host BFS (xlib/src/Graph/BFS.cpp:54-76): FIFO queue, visited bitmask,
 * dist[src] = 0, unreached = INT32_MAX.  Returns the number of visited
 * vertices; *edges_scanned (optional) = sum of degrees of visited vertices.
 */
int64_t orc_bfs(int32_t n, const int64_t* off, const int32_t* adj, int32_t source,
                int32_t* dist, int64_t* edges_scanned) {
    int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    uint8_t* seen = (uint8_t*)calloc((size_t)n + 1, 1);
    for (int32_t i = 0; i < n; i++) dist[i] = INT32_MAX;
    int64_t head = 0, tail = 0, scanned = 0;
    queue[tail++] = source;
    seen[source] = 1;
    dist[source] = 0;
    while (head < tail) {
        int32_t cur = queue[head++];
        scanned += off[cur + 1] - off[cur];
        for (int64_t j = off[cur]; j < off[cur + 1]; j++) {
            int32_t d = adj[j];
            if (!seen[d]) {
                seen[d] = 1;
                dist[d] = dist[cur] + 1;
                queue[tail++] = d;
            }
        }
    }
    free(queue);
    free(seen);
    if (edges_scanned) *edges_scanned = scanned;
    return tail;
}

/* hostSingleIntersection (hornetsnest/test/TriangleTest2.cu:22-47): merge count */
static int64_t merge_count(const int32_t* a, int64_t alen, const int32_t* b, int64_t blen) {
    if (!alen || !blen || a[alen - 1] < b[0] || b[blen - 1] < a[0]) return 0;
    int64_t i = 0, j = 0, out = 0;
    while (i < alen && j < blen) {
        if (a[i] == b[j]) { i++; j++; out++; }
        else if (a[i] < b[j]) i++;
        else j++;
    }
    return out;
}

/* hostCountTriangles (hornetsnest/test/TriangleTest2.cu:50-69): sum over stored
 * edges (src,dst) of |N(src) ∩ N(dst)|.  On a DIRECTED_BY_DEGREE graph this is
 * the triangle count. */
int64_t orc_host_tc(int32_t n, const int64_t* off, const int32_t* adj) {
    int64_t sum = 0;
    for (int32_t s = 0; s < n; s++)
        for (int64_t j = off[s]; j < off[s + 1]; j++) {
            int32_t d = adj[j];
            sum += merge_count(adj + off[s], off[s + 1] - off[s],
                               adj + off[d], off[d + 1] - off[d]);
        }
    return sum;
}

/*
 * TriangleCounting2 per-vertex counters on an arbitrary stored (simple) graph:
 * forAllAdjUnions visits every stored edge (src,dst)
 * (primitives/Operator++.i.cuh:395-413), skips src==dst or deg(src)<2
 * (:120-122), swaps so u is the endpoint with the smaller degree
 * (sourceSmaller = srcLen < destLen, :125-131), and the balanced operator
 * credits |N(u) ∩ N(v)| to u only (triangle2.cu:121-122).
 */
void orc_tc2_per_vertex(int32_t n, const int64_t* off, const int32_t* adj,
                        uint64_t* tri) {
    memset(tri, 0, sizeof(uint64_t) * (size_t)n);
    for (int32_t s = 0; s < n; s++) {
        int64_t ds = off[s + 1] - off[s];
        for (int64_t j = off[s]; j < off[s + 1]; j++) {
            int32_t d = adj[j];
            int64_t dd = off[d + 1] - off[d];
            if (s == d || ds < 2) continue;
            int32_t u = ds < dd ? s : d;
            int64_t c = merge_count(adj + off[s], ds, adj + off[d], dd);
            tri[u] += (uint64_t)c;
        }
    }
}

/* std::max_element over degrees: first (lowest id) maximum
 * (xlib/include/Graph/GraphStd.i.hpp:296-299) */
int32_t orc_max_degree_id(int32_t n, const int64_t* off) {
    int32_t best = 0;
    int64_t bd = n > 0 ? off[1] - off[0] : 0;
    for (int32_t v = 1; v < n; v++) {
        int64_t d = off[v + 1] - off[v];
        if (d > bd) { bd = d; best = v; }
    }
    return best;
}

/* ------------------------------------------------------------------------ */
/* host set model of BatchUpdate                                             */
/* ------------------------------------------------------------------------ */

/* sort the batch lexicographically by (src,dst) (SoAPtr.i.cuh:560-571) and
 * optionally drop adjacent repeats (thrust::unique_copy, BatchUpdate.i.cuh:250-351) */
static int64_t sorted_batch(const int32_t* src, const int32_t* dst, int64_t b,
                            int dedup, uint64_t** out) {
    uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(b + 1));
    for (int64_t i = 0; i < b; i++) k[i] = pack(src[i], dst[i]);
    radix_sort_u64(k, b, 64);
    if (dedup) {
        int64_t w = 0;
        for (int64_t i = 0; i < b; i++)
            if (w == 0 || k[i] != k[w - 1]) k[w++] = k[i];
        b = w;
    }
    *out = k;
    return b;
}

static int64_t count_in_sorted(const int32_t* a, int64_t len, int32_t x, int64_t* first) {
    int64_t lo = 0, hi = len;
    while (lo < hi) { int64_t mid = (lo + hi) / 2; if (a[mid] < x) lo = mid + 1; else hi = mid; }
    int64_t f = lo;
    hi = len;
    while (lo < hi) { int64_t mid = (lo + hi) / 2; if (a[mid] <= x) lo = mid + 1; else hi = mid; }
    if (first) *first = f;
    return lo - f;
}

/*
 * Hornet::insert (HornetInsert.i.cuh:12-27) on sorted-adjacency CSR input:
 * sort; optional in-batch unique; optional graph dedup, which clears the FIRST
 * batch copy of every (u,v) already stored (mark_duplicate_edges_kernel with
 * lower_bound_left, BatchUpdateKernels.cuh:27-58); the survivors are added as
 * a multiset union.  Output adjacency is sorted.  Returns new nE; *added gets
 * the number of batch entries that were inserted.
 */
int64_t orc_insert(int32_t n, const int64_t* off, const int32_t* adj,
                   const int32_t* bsrc, const int32_t* bdst, int64_t b,
                   int dedup_batch, int dedup_graph,
                   int64_t* off_out, int32_t* adj_out, int64_t* added) {
    uint64_t* k;
    b = sorted_batch(bsrc, bdst, b, dedup_batch, &k);
    if (dedup_graph) {
        int64_t w = 0;
        for (int64_t i = 0; i < b; i++) {
            int32_t s = (int32_t)(k[i] >> 32), d = (int32_t)(uint32_t)k[i];
            int first_copy = (i == 0) || k[i - 1] != k[i];
            if (first_copy && count_in_sorted(adj + off[s], off[s + 1] - off[s], d, NULL) > 0)
                continue;
            k[w++] = k[i];
        }
        b = w;
    }
    if (added) *added = b;
    /* merge per vertex: stored copies first, then batch copies (equal values
     * are indistinguishable, so the order among equals does not matter) */
    int64_t w = 0, bi = 0;
    off_out[0] = 0;
    for (int32_t v = 0; v < n; v++) {
        int64_t i = off[v], e = off[v + 1];
        while (i < e || (bi < b && (int32_t)(k[bi] >> 32) == v)) {
            int takeb = bi < b && (int32_t)(k[bi] >> 32) == v &&
                        (i >= e || (int32_t)(uint32_t)k[bi] < adj[i]);
            if (takeb) adj_out[w++] = (int32_t)(uint32_t)k[bi++];
            else adj_out[w++] = adj[i++];
        }
        off_out[v + 1] = w;
    }
    free(k);
    return w;
}

/*
 * Hornet::erase (HornetInsert.i.cuh:71-89): sort; optional unique; each
 * distinct (u,v) of multiplicity c in the batch removes min(c, stored copies)
 * stored copies (markUniqueOffsets + locate_erased_edges' atomicSub,
 * BatchUpdateKernels.cuh:263-300,335-393); absent edges are ignored.
 * Survivors keep their sorted order.  Returns the new nE; *removed gets the
 * number of removed entries.
 */
int64_t orc_erase(int32_t n, const int64_t* off, const int32_t* adj,
                  const int32_t* bsrc, const int32_t* bdst, int64_t b, int dedup_batch,
                  int64_t* off_out, int32_t* adj_out, int64_t* removed) {
    uint64_t* k;
    b = sorted_batch(bsrc, bdst, b, dedup_batch, &k);
    int64_t w = 0, bi = 0, rem = 0;
    off_out[0] = 0;
    for (int32_t v = 0; v < n; v++) {
        int64_t i = off[v], e = off[v + 1];
        while (i < e) {
            int32_t x = adj[i];
            /* batch multiplicity of (v,x) */
            while (bi < b && k[bi] < pack(v, x)) bi++;
            int64_t c = 0;
            while (bi + c < b && k[bi + c] == pack(v, x)) c++;
            int64_t j = i;
            while (j < e && adj[j] == x) j++;
            int64_t stored = j - i;
            int64_t drop = c < stored ? c : stored;
            rem += drop;
            for (int64_t t = drop; t < stored; t++) adj_out[w++] = x;
            bi += c;
            i = j;
        }
        while (bi < b && (int32_t)(k[bi] >> 32) == v) bi++;
        off_out[v + 1] = w;
    }
    free(k);
    if (removed) *removed = rem;
    return w;
}

/*
 * 2-hop neighbourhood size of v: |{v} ∪ N(v) ∪ N(N(v))| (the serving query the
 * build defines; no reference counterpart).  `mark` is caller scratch of n
 * bytes, all zero on entry and on exit.
 */
int64_t orc_two_hop(int32_t n, const int64_t* off, const int32_t* adj, int32_t v,
                    uint8_t* mark) {
    (void)n;
    int64_t cnt = 0;
    int32_t* touched = (int32_t*)malloc(sizeof(int32_t) * 1);
    int64_t tcap = 1, tn = 0;
#define TOUCH(x) do { if (!mark[x]) { mark[x] = 1; cnt++; \
        if (tn == tcap) { tcap *= 2; touched = (int32_t*)realloc(touched, sizeof(int32_t) * (size_t)tcap); } \
        touched[tn++] = (x); } } while (0)
    TOUCH(v);
    for (int64_t j = off[v]; j < off[v + 1]; j++) {
        int32_t u = adj[j];
        TOUCH(u);
        for (int64_t t = off[u]; t < off[u + 1]; t++) TOUCH(adj[t]);
    }
#undef TOUCH
    for (int64_t t = 0; t < tn; t++) mark[touched[t]] = 0;
    free(touched);
    return cnt;
}

/*
 * Chung-Lu graph: expected degrees ∝ w_i = (i+1)^(-1/(gamma-1)); with
 * avg*n/2 pairs the mean degree is avg.  Pairs are drawn ∝ w_i w_j by
 * inverse-CDF sampling of both endpoints over an integer cumulative weight
 * table (cum, n+1 entries, built with orc_chung_lu_table).  Pair i uses draws
 * 2i and 2i+1; an endpoint is the largest v with cum[v] <= high64(r * cum[n]).
 */
static int32_t cdf_pick(const uint64_t* cum, int32_t n, uint64_t r) {
    /* r in [0, cum[n]) -> largest v with cum[v] <= r */
    int32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        int32_t mid = lo + (hi - lo) / 2;
        if (cum[mid] <= r) lo = mid; else hi = mid;
    }
    return lo;
}
/* cum[v] = sum_{i<v} w_i with w_i = floor(2^24 * (i+1)^(-1/(gamma-1)))
 * (glibc pow in double, evaluated once on the host). */
void orc_chung_lu_table(int32_t n, double gamma, uint64_t* cum) {
    cum[0] = 0;
    double e = -1.0 / (gamma - 1.0);
    for (int32_t i = 0; i < n; i++)
        cum[i + 1] = cum[i] + (uint64_t)(16777216.0 * pow((double)i + 1.0, e));
}
void orc_chung_lu_pairs(int32_t n, const uint64_t* cum, int64_t first, int64_t count,
                        uint64_t seed, int32_t* src, int32_t* dst) {
    uint64_t key = orc_stream_key(seed);
    uint64_t total = cum[n];
    for (int64_t k = 0; k < count; k++) {
        uint64_t i = (uint64_t)(first + k);
        uint64_t r0 = draw(key, 2 * i);
        uint64_t r1 = draw(key, 2 * i + 1);
        src[k] = cdf_pick(cum, n, (uint64_t)(((unsigned __int128)r0 * total) >> 64));
        dst[k] = cdf_pick(cum, n, (uint64_t)(((unsigned __int128)r1 * total) >> 64));
    }
}

/* ------------------------------------------------------------------------ */
/* Parallel forms of the checker.  Each returns exactly what its sequential */
/* counterpart above returns (tests/test_oracle.py compares them); they only */
/* split the work over POSIX threads so that graphs of 2^22..2^25 vertices   */
/* can be checked in seconds.  Not a restatement of anything new.            */
/* ------------------------------------------------------------------------ */
#include <pthread.h>

typedef void (*orc_range_fn)(void* ctx, int64_t lo, int64_t hi);

typedef struct {
    orc_range_fn fn;
    void* ctx;
    int64_t n, chunk;
    int64_t next;   /* shared cursor, taken with an atomic add */
} orc_pool;

static void* orc_pool_worker(void* arg) {
    orc_pool* p = (orc_pool*)arg;
    for (;;) {
        int64_t lo = __atomic_fetch_add(&p->next, p->chunk, __ATOMIC_RELAXED);
        if (lo >= p->n) break;
        int64_t hi = lo + p->chunk < p->n ? lo + p->chunk : p->n;
        p->fn(p->ctx, lo, hi);
    }
    return NULL;
}

/* run fn over [0, n) in chunks taken dynamically by `threads` threads */
static void orc_parallel(int threads, int64_t n, int64_t chunk, orc_range_fn fn, void* ctx) {
    orc_pool p = {fn, ctx, n, chunk > 0 ? chunk : 1, 0};
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    int started = 0;
    for (int t = 1; t < threads; t++)
        if (pthread_create(&tid[started], NULL, orc_pool_worker, &p) == 0) started++;
    orc_pool_worker(&p);
    for (int t = 0; t < started; t++) pthread_join(tid[t], NULL);
}

typedef struct { int scale; int64_t first; uint64_t seed; int32_t* src; int32_t* dst; } rmat_ctx;
static void rmat_range(void* c, int64_t lo, int64_t hi) {
    rmat_ctx* x = (rmat_ctx*)c;
    orc_rmat_pairs(x->scale, x->first + lo, hi - lo, x->seed, x->src + lo, x->dst + lo);
}
/* orc_rmat_pairs over threads (pairs are independent draws of the stream) */
void orc_rmat_pairs_mt(int scale, int64_t first, int64_t count, uint64_t seed, int threads,
                       int32_t* src, int32_t* dst) {
    rmat_ctx c = {scale, first, seed, src, dst};
    orc_parallel(threads, count, 1 << 18, rmat_range, &c);
}

typedef struct { int32_t n; const uint64_t* cum; int64_t first; uint64_t seed; int32_t* src; int32_t* dst; } cl_ctx;
static void cl_range(void* c, int64_t lo, int64_t hi) {
    cl_ctx* x = (cl_ctx*)c;
    orc_chung_lu_pairs(x->n, x->cum, x->first + lo, hi - lo, x->seed, x->src + lo, x->dst + lo);
}
/* orc_chung_lu_pairs over threads */
void orc_chung_lu_pairs_mt(int32_t n, const uint64_t* cum, int64_t first, int64_t count, uint64_t seed,
                           int threads, int32_t* src, int32_t* dst) {
    cl_ctx c = {n, cum, first, seed, src, dst};
    orc_parallel(threads, count, 1 << 18, cl_range, &c);
}

/* COO -> CSR with symmetrize / self-loop drop, always sorted and deduplicated
 * (orc_coo_to_csr with dedup = sort = 1), as a parallel MSD pass: the entries
 * are scattered into buckets of 2^bshift consecutive source ids (per-thread
 * histograms, so each thread writes its own run of every bucket), then every
 * bucket is radix-sorted and deduplicated on its own and written out. */
#define CSR_MAX_THREADS 64
typedef struct {
    const int32_t* src; const int32_t* dst; int sym, drop;
    int64_t m; int T; int bshift; int32_t n; int64_t nb;
    int64_t* hist;        /* T x nb: entries thread t puts in bucket b, then its write cursor */
    int64_t* boff;        /* nb + 1 bucket starts */
    uint64_t* keys;       /* 2m entries in bucket order */
    uint64_t* tmp;        /* radix scratch, 2m */
    int64_t* bcnt;        /* per bucket: distinct entries */
    int64_t* off;
    int32_t* adj;
    int key_bits;
} csrb_ctx;

static inline int64_t csrb_lo(const csrb_ctx* x, int t) { return x->m * t / x->T; }

static void csrb_hist(void* c, int64_t lo, int64_t hi) {
    csrb_ctx* x = (csrb_ctx*)c;
    for (int64_t t = lo; t < hi; t++) {
        int64_t* h = x->hist + t * x->nb;
        for (int64_t i = csrb_lo(x, (int)t); i < csrb_lo(x, (int)t + 1); i++) {
            int32_t s = x->src[i], d = x->dst[i];
            if (s == d && x->drop) continue;
            h[s >> x->bshift]++;
            if (x->sym && s != d) h[d >> x->bshift]++;
        }
    }
}
static void csrb_scatter(void* c, int64_t lo, int64_t hi) {
    csrb_ctx* x = (csrb_ctx*)c;
    for (int64_t t = lo; t < hi; t++) {
        int64_t* h = x->hist + t * x->nb;
        for (int64_t i = csrb_lo(x, (int)t); i < csrb_lo(x, (int)t + 1); i++) {
            int32_t s = x->src[i], d = x->dst[i];
            if (s == d && x->drop) continue;
            x->keys[h[s >> x->bshift]++] = pack(s, d);
            if (x->sym && s != d) x->keys[h[d >> x->bshift]++] = pack(d, s);
        }
    }
}
/* LSD radix sort of a[0..len) on the key bits (source low bits and
 * destination), 11-bit digits, scratch t of len entries */
static void csrb_sort(uint64_t* a, uint64_t* t, int64_t len, int bshift, int dbits) {
    /* sort key: (src & (2^bshift - 1)) << dbits | dst */
    const int bits = bshift + dbits;
    int64_t cnt[2048];
    uint64_t* from = a;
    uint64_t* to = t;
    for (int sh = 0; sh < bits; sh += 11) {
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < len; i++) {
            uint64_t k = ((from[i] >> 32) & ((1ull << bshift) - 1)) << dbits | (from[i] & 0xFFFFFFFFull);
            cnt[(k >> sh) & 2047]++;
        }
        int64_t s = 0;
        for (int d = 0; d < 2048; d++) { int64_t c = cnt[d]; cnt[d] = s; s += c; }
        for (int64_t i = 0; i < len; i++) {
            uint64_t k = ((from[i] >> 32) & ((1ull << bshift) - 1)) << dbits | (from[i] & 0xFFFFFFFFull);
            to[cnt[(k >> sh) & 2047]++] = from[i];
        }
        uint64_t* x = from; from = to; to = x;
    }
    if (from != a) memcpy(a, from, sizeof(uint64_t) * (size_t)len);
}
static void csrb_sort_buckets(void* c, int64_t lo, int64_t hi) {
    csrb_ctx* x = (csrb_ctx*)c;
    for (int64_t b = lo; b < hi; b++) {
        uint64_t* a = x->keys + x->boff[b];
        int64_t len = x->boff[b + 1] - x->boff[b];
        if (len > 1) csrb_sort(a, x->tmp + x->boff[b], len, x->bshift, x->key_bits);
        int64_t w = 0;
        for (int64_t i = 0; i < len; i++)
            if (w == 0 || a[i] != a[w - 1]) a[w++] = a[i];
        x->bcnt[b] = w;
        /* degrees of the bucket's vertices (off[v + 1] is filled here, scanned later) */
        for (int64_t i = 0; i < w; i++) x->off[(a[i] >> 32) + 1]++;
    }
}
static void csrb_write(void* c, int64_t lo, int64_t hi) {
    csrb_ctx* x = (csrb_ctx*)c;
    for (int64_t b = lo; b < hi; b++) {
        const uint64_t* a = x->keys + x->boff[b];
        int64_t v0 = b << x->bshift;
        int64_t o = x->off[v0 < x->n ? v0 : x->n];
        for (int64_t i = 0; i < x->bcnt[b]; i++) x->adj[o + i] = (int32_t)(uint32_t)a[i];
    }
}

/* Returns the number of entries (adj must hold 2m); -1 when out of memory. */
int64_t orc_coo_to_csr_mt(int32_t n, int64_t m, const int32_t* src, const int32_t* dst, int symmetrize,
                          int drop_self_loops, int threads, int64_t* off, int32_t* adj) {
    int T = threads < 1 ? 1 : (threads > CSR_MAX_THREADS ? CSR_MAX_THREADS : threads);
    int nbits = 0;
    while ((1ll << nbits) < (int64_t)n) nbits++;
    int bshift = nbits > 12 ? nbits - 12 : 0;
    int64_t nb = (((int64_t)n + (1ll << bshift) - 1) >> bshift) + 1;
    csrb_ctx c;
    memset(&c, 0, sizeof(c));
    c.src = src; c.dst = dst; c.sym = symmetrize; c.drop = drop_self_loops;
    c.m = m; c.T = T; c.bshift = bshift; c.n = n; c.nb = nb; c.off = off; c.adj = adj;
    c.key_bits = nbits > 0 ? nbits : 1;
    c.hist = (int64_t*)calloc((size_t)(T * nb), sizeof(int64_t));
    c.boff = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
    c.bcnt = (int64_t*)malloc(sizeof(int64_t) * (size_t)nb);
    c.keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(2 * m + 1));
    c.tmp = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(2 * m + 1));
    if (!c.hist || !c.boff || !c.bcnt || !c.keys || !c.tmp) {
        free(c.hist); free(c.boff); free(c.bcnt); free(c.keys); free(c.tmp);
        return -1;
    }
    orc_parallel(T, T, 1, csrb_hist, &c);
    /* bucket starts, and each thread's cursor inside every bucket */
    int64_t s = 0;
    for (int64_t b = 0; b < nb; b++) {
        c.boff[b] = s;
        for (int t = 0; t < T; t++) {
            int64_t h = c.hist[t * nb + b];
            c.hist[t * nb + b] = s;
            s += h;
        }
    }
    c.boff[nb] = s;
    orc_parallel(T, T, 1, csrb_scatter, &c);
    memset(off, 0, sizeof(int64_t) * ((size_t)n + 1));
    orc_parallel(T, nb, 1, csrb_sort_buckets, &c);
    for (int32_t v = 0; v < n; v++) off[v + 1] += off[v];
    orc_parallel(T, nb, 16, csrb_write, &c);
    int64_t k = off[n];
    free(c.hist); free(c.boff); free(c.bcnt); free(c.keys); free(c.tmp);
    return k;
}

/* orc_host_tc (the sum over stored edges (u,v) of |N(u) ∩ N(v)|) with the
 * intersections taken against a per-thread byte mark of N(u) instead of merges:
 * sum over u, over v in N(u), of #{w in N(v) : w in N(u)}.  Same value, fewer
 * operations on lists of very different lengths. */
typedef struct { int32_t n; const int64_t* off; const int32_t* adj; int64_t sum; int64_t next; } tcm_ctx;
static void* tcm_worker(void* arg) {
    tcm_ctx* x = (tcm_ctx*)arg;
    uint8_t* mark = (uint8_t*)calloc((size_t)x->n + 1, 1);
    int64_t s = 0;
    if (!mark) return (void*)1;
    for (;;) {
        int64_t lo = __atomic_fetch_add(&x->next, 256, __ATOMIC_RELAXED);
        if (lo >= x->n) break;
        int64_t hi = lo + 256 < x->n ? lo + 256 : x->n;
        for (int64_t u = lo; u < hi; u++) {
            const int64_t a = x->off[u], b = x->off[u + 1];
            if (b - a < 1) continue;
            for (int64_t j = a; j < b; j++) mark[x->adj[j]] = 1;
            for (int64_t j = a; j < b; j++) {
                const int32_t v = x->adj[j];
                for (int64_t t = x->off[v]; t < x->off[v + 1]; t++) s += mark[x->adj[t]];
            }
            for (int64_t j = a; j < b; j++) mark[x->adj[j]] = 0;
        }
    }
    free(mark);
    __atomic_fetch_add(&x->sum, s, __ATOMIC_RELAXED);
    return NULL;
}
/* Returns -1 when a thread could not allocate its mark array. */
int64_t orc_host_tc_mark_mt(int32_t n, const int64_t* off, const int32_t* adj, int threads) {
    tcm_ctx c = {n, off, adj, 0, 0};
    int T = threads < 1 ? 1 : (threads > 256 ? 256 : threads);
    pthread_t tid[256];
    int started = 0, failed = 0;
    for (int t = 1; t < T; t++)
        if (pthread_create(&tid[started], NULL, tcm_worker, &c) == 0) started++;
    if (tcm_worker(&c)) failed = 1;
    for (int t = 0; t < started; t++) {
        void* r = NULL;
        pthread_join(tid[t], &r);
        if (r) failed = 1;
    }
    return failed ? -1 : c.sum;
}
