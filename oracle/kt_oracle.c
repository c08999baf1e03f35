/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's
 * k-max-truss path (caixiaofan/KC_CUDA), used by tests/, by smoke() and by
 * bench.py's cpu_baseline leg as the checker. Never linked into, called by or
 * shipped with the product library.
 *
 * What each function restates (reference file:line):
 *   kto_parse_tsv          input contract        src/kc_gpu.cpp:243-312, src/kc_cpu.cpp:796-865
 *   kto_csr_from_lines     row build             src/kc_gpu.cpp:38-65 (init_graph)
 *   kto_kcore              fixed-k core peel     src/kc_cpu.cpp:185-251 (Launch_start branch), src/kc.cu:653-694
 *   kto_max_core           max-core              src/kc_cpu.cpp:185-251, src/kc.cu:770-884
 *   kto_rank               degree rank           src/kc_gpu.cpp:27-30 (cmp_link), src/kc.cu:1449-1469
 *   kto_oriented           oriented CSR, sorted  src/kc.cu:1354-1402, 1478-1580
 *   kto_undirected         undirected CSR + fa   src/kc.cu:1583-1673
 *   kto_support            doubled supports      src/kc.cu:505-535 + 193-339, src/kc_cpu.cpp:93-118, 474-488
 *   peel_level             frontier peeling      src/kc.cu:939-1056, 61-159 (the `num` rule), 906-935
 *   kto_solve              two-pass search       src/kc_gpu.cpp:129-208, 336-341; src/kc.cu:1140-1309
 *
 * Deliberate simplifications (result-preserving, so not parity targets of
 * their own): the tc_limit-pruned count (src/kc.cu:537-582) and the galloping
 * k step with backup/restore (src/kc.cu:1169-1305) are replaced by an exact
 * count and a level-by-level upward search. Both yield the unique kmax-truss.
 *
 * Pinning: tests/test_oracle.py checks every function here against
 * independent computations (textbook |N(u) ∩ N(v)| supports, bucket core
 * numbers, brute-force truss on tiny graphs, closed forms, networkx goldens and
 * the reference's own kc_cpu binary built from its sources).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ALIVE 0xFFFFFFFFu

/* ------------------------------------------------------------------ a0 ---- */

/* Parse the reference's edge-list text: each line is x<sep>y[rest], where
 * separators are any byte < '0', digits accumulate as (c - '0'), the rest of
 * the line after y is skipped (src/kc_gpu.cpp:262-291). Fills xs/ys when
 * non-NULL (capacity cap) and returns the number of lines. */
int64_t kto_parse_tsv(const char *buf, uint64_t len, uint32_t *xs, uint32_t *ys, uint64_t cap) {
    uint64_t j = 0, nl = 0;
    while (j < len && !(buf[j] >= '0')) j++;
    while (j < len) {
        uint32_t x = 0, y = 0;
        while (j < len && buf[j] >= '0') { x = x * 10u + (uint32_t)(buf[j] - '0'); j++; }
        while (j < len && !(buf[j] >= '0')) j++;
        while (j < len && buf[j] >= '0') { y = y * 10u + (uint32_t)(buf[j] - '0'); j++; }
        while (j < len && buf[j] != '\n') j++;
        while (j < len && !(buf[j] >= '0')) j++;
        if (xs && nl < cap) { xs[nl] = x; ys[nl] = y; }
        nl++;
    }
    return (int64_t)nl;
}

/* Undirected CSR from the line list: row y receives x, rows in line order.
 * For input grouped by y (the contract) this equals the reference's
 * concatenation of parsed chunks placed by deg prefix sums
 * (src/kc_gpu.cpp:38-65, 283-296). off has n+1 entries, adj has L. */
void kto_csr_from_lines(uint64_t L, const uint32_t *xs, const uint32_t *ys, uint32_t n,
                        uint64_t *off, uint32_t *adj) {
    memset(off, 0, sizeof(uint64_t) * ((size_t)n + 1));
    for (uint64_t i = 0; i < L; i++) off[ys[i] + 1]++;
    for (uint32_t v = 0; v < n; v++) off[v + 1] += off[v];
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)n + 1));
    memcpy(cur, off, sizeof(uint64_t) * ((size_t)n + 1));
    for (uint64_t i = 0; i < L; i++) adj[cur[ys[i]]++] = xs[i];
    free(cur);
}

/* ------------------------------------------------------------------ b1 ---- */

/* k-core: repeatedly remove vertices of remaining degree < k
 * (src/kc_cpu.cpp:212-231; src/kc.cu:653-694). removed[v] is set to 1 for
 * every vertex outside the k-core (caller zero-initialises it or passes a
 * previous, smaller core). Returns the number of vertices in the k-core. */
uint64_t kto_kcore(uint32_t n, const uint64_t *off, const uint32_t *adj, uint32_t k, uint8_t *removed) {
    uint32_t *deg = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t sp = 0, alive = 0;
    for (uint32_t v = 0; v < n; v++) {
        uint32_t d = 0;
        if (!removed[v])
            for (uint64_t j = off[v]; j < off[v + 1]; j++) d += !removed[adj[j]];
        deg[v] = d;
    }
    for (uint32_t v = 0; v < n; v++)
        if (!removed[v] && deg[v] < k) { removed[v] = 1; stack[sp++] = v; }
    while (sp) {
        uint32_t u = stack[--sp];
        for (uint64_t j = off[u]; j < off[u + 1]; j++) {
            uint32_t w = adj[j];
            if (!removed[w] && --deg[w] < k) { removed[w] = 1; stack[sp++] = w; }
        }
    }
    for (uint32_t v = 0; v < n; v++) alive += !removed[v];
    free(deg);
    free(stack);
    return alive;
}

/* Max-core: the k-core for the largest k at which it is non-empty
 * (src/kc_cpu.cpp:185-251 with Launch_start = -1; src/kc.cu:770-884).
 * Returns that k (0 when the graph has no edges); removed[] marks vertices
 * outside it. Uses the core-number peel order, so no k stepping is needed. */
uint32_t kto_max_core(uint32_t n, const uint64_t *off, const uint32_t *adj, uint8_t *removed) {
    uint32_t kbest = 0;
    uint8_t *tmp = (uint8_t *)calloc(n ? n : 1, 1);
    memset(removed, 0, n);
    /* galloping upward with restore, as the reference does, but exact */
    uint32_t k = 0, step = 64;
    for (;;) {
        memcpy(tmp, removed, n);
        uint64_t a = kto_kcore(n, off, adj, k + step, tmp);
        if (a > 0) {
            k += step;
            memcpy(removed, tmp, n);
            kbest = k;
        } else if (step > 1) {
            step /= 2;
        } else {
            break;
        }
    }
    free(tmp);
    if (kbest == 0) memset(removed, 1, n);
    return kbest;
}

/* --------------------------------------------------------------- a1-a3 ---- */

/* Rank of the kept vertices by (degree inside the kept set ascending, id
 * descending) — cmp_link (src/kc_gpu.cpp:27-30) applied to deg_with_mask
 * (src/kc.cu:1311-1352, 1449-1469). to[v] = rank or ALIVE (= not kept);
 * inv[rank] = v; degc[v] = degree inside the kept set. Returns n_C. */
static int cmp_deg_id(const void *a, const void *b) {
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}
uint32_t kto_rank(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint8_t *removed,
                  uint32_t *to, uint32_t *inv, uint32_t *degc) {
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    uint32_t nc = 0;
    for (uint32_t v = 0; v < n; v++) {
        to[v] = ALIVE;
        uint32_t d = 0;
        if (!removed[v]) {
            for (uint64_t j = off[v]; j < off[v + 1]; j++) d += !removed[adj[j]];
            /* ascending degree, then descending id */
            keys[nc++] = ((uint64_t)d << 32) | (uint64_t)(0xFFFFFFFFu - v);
        }
        degc[v] = d;
    }
    qsort(keys, nc, sizeof(uint64_t), cmp_deg_id);
    for (uint32_t r = 0; r < nc; r++) {
        uint32_t v = 0xFFFFFFFFu - (uint32_t)keys[r];
        to[v] = r;
        inv[r] = v;
    }
    free(keys);
    return nc;
}

static int cmp_u32(const void *a, const void *b) {
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* Oriented CSR of the kept subgraph in rank ids: row r holds {to[w] > r},
 * sorted ascending (src/kc.cu:1478-1580; orientation test :1511). off_d has
 * nc+1 entries. When adj_d is NULL only off_d is filled. Returns m_C. */
uint64_t kto_oriented(uint32_t n, const uint64_t *off, const uint32_t *adj, const uint32_t *to,
                      uint32_t nc, uint64_t *off_d, uint32_t *adj_d) {
    memset(off_d, 0, sizeof(uint64_t) * ((size_t)nc + 1));
    for (uint32_t v = 0; v < n; v++) {
        if (to[v] == ALIVE) continue;
        for (uint64_t j = off[v]; j < off[v + 1]; j++) {
            uint32_t w = adj[j];
            if (to[w] != ALIVE && to[w] > to[v]) off_d[to[v] + 1]++;
        }
    }
    for (uint32_t r = 0; r < nc; r++) off_d[r + 1] += off_d[r];
    if (adj_d) {
        uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)nc + 1));
        memcpy(cur, off_d, sizeof(uint64_t) * ((size_t)nc + 1));
        for (uint32_t v = 0; v < n; v++) {
            if (to[v] == ALIVE) continue;
            for (uint64_t j = off[v]; j < off[v + 1]; j++) {
                uint32_t w = adj[j];
                if (to[w] != ALIVE && to[w] > to[v]) adj_d[cur[to[v]]++] = to[w];
            }
        }
        for (uint32_t r = 0; r < nc; r++)
            qsort(adj_d + off_d[r], off_d[r + 1] - off_d[r], sizeof(uint32_t), cmp_u32);
        free(cur);
    }
    return off_d[nc];
}

/* Undirected CSR of an oriented graph, rows sorted, with fa[slot] = oriented
 * edge id, lson[e] / rson[e] = slot of e in the row of its lower / higher
 * endpoint (src/kc.cu:1583-1673, the son fill at :1656-1666). */
void kto_undirected(uint32_t nc, const uint64_t *off_d, const uint32_t *adj_d, uint64_t *off_u,
                    uint32_t *adj_u, uint64_t *fa, uint64_t *lson, uint64_t *rson) {
    const uint64_t m = off_d[nc];
    memset(off_u, 0, sizeof(uint64_t) * ((size_t)nc + 1));
    for (uint32_t r = 0; r < nc; r++) {
        off_u[r + 1] += off_d[r + 1] - off_d[r];
        for (uint64_t j = off_d[r]; j < off_d[r + 1]; j++) off_u[adj_d[j] + 1]++;
    }
    for (uint32_t r = 0; r < nc; r++) off_u[r + 1] += off_u[r];
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)nc + 1));
    memcpy(cur, off_u, sizeof(uint64_t) * ((size_t)nc + 1));
    /* Rows come out sorted without a sort: lower neighbours of v arrive in
     * increasing source order, then v's own (sorted) out-row is appended. */
    for (uint32_t r = 0; r < nc; r++) {
        for (uint64_t j = off_d[r]; j < off_d[r + 1]; j++) {
            uint32_t w = adj_d[j];
            uint64_t p = cur[w]++;
            adj_u[p] = r;
            fa[p] = j;
            rson[j] = p;
        }
    }
    for (uint32_t r = 0; r < nc; r++) {
        for (uint64_t j = off_d[r]; j < off_d[r + 1]; j++) {
            uint64_t p = cur[r]++;
            adj_u[p] = adj_d[j];
            fa[p] = j;
            lson[j] = p;
        }
    }
    free(cur);
    (void)m;
}

/* ------------------------------------------------------------------ c1 ---- */

/* Exact doubled supports per oriented edge: each triangle a<b<c (rank) is
 * found once at a->b as c in out(a) ∩ out(b) and credits (a,b), (a,c), (b,c)
 * with +2 each (src/kc_cpu.cpp:93-118 with tc_limit = 0, :483-485;
 * src/kc.cu:271-316). */
void kto_support(uint32_t nc, const uint64_t *off_d, const uint32_t *adj_d, uint32_t *sup2) {
    const uint64_t m = off_d[nc];
    memset(sup2, 0, sizeof(uint32_t) * (m ? m : 1));
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t a = 0; a < (int64_t)nc; a++) {
        for (uint64_t e = off_d[a]; e < off_d[a + 1]; e++) {
            const uint32_t b = adj_d[e];
            uint64_t i = off_d[a], j = off_d[b];
            const uint64_t ie = off_d[a + 1], je = off_d[b + 1];
            uint32_t cnt = 0;
            while (i < ie && j < je) {
                const uint32_t x = adj_d[i], y = adj_d[j];
                if (x == y) {
                    cnt++;
                    __atomic_fetch_add(&sup2[i], 2u, __ATOMIC_RELAXED);
                    __atomic_fetch_add(&sup2[j], 2u, __ATOMIC_RELAXED);
                }
                i += (x <= y);
                j += (y <= x);
            }
            __atomic_fetch_add(&sup2[e], 2u * cnt, __ATOMIC_RELAXED);
        }
    }
}

/* ------------------------------------------------------------------- d ---- */

typedef struct {
    uint32_t nc;
    uint64_t m;
    uint64_t *off_d; uint32_t *adj_d; uint32_t *src; /* oriented edge list */
    uint64_t *off_u; uint32_t *adj_u; uint64_t *fa, *lson, *rson;
    uint32_t *sup2, *state;
    uint32_t *inv; /* rank -> original id */
} core_graph;

static void cg_free(core_graph *g) {
    free(g->off_d); free(g->adj_d); free(g->src);
    free(g->off_u); free(g->adj_u); free(g->fa); free(g->lson); free(g->rson);
    free(g->sup2); free(g->state); free(g->inv);
    memset(g, 0, sizeof(*g));
}

static void cg_finish(core_graph *g) {
    const uint64_t m = g->m;
    const size_t mm = m ? m : 1;
    g->src = (uint32_t *)malloc(sizeof(uint32_t) * mm);
    for (uint32_t r = 0; r < g->nc; r++)
        for (uint64_t j = g->off_d[r]; j < g->off_d[r + 1]; j++) g->src[j] = r;
    g->off_u = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)g->nc + 1));
    g->adj_u = (uint32_t *)malloc(sizeof(uint32_t) * 2 * mm);
    g->fa = (uint64_t *)malloc(sizeof(uint64_t) * 2 * mm);
    g->lson = (uint64_t *)malloc(sizeof(uint64_t) * mm);
    g->rson = (uint64_t *)malloc(sizeof(uint64_t) * mm);
    kto_undirected(g->nc, g->off_d, g->adj_d, g->off_u, g->adj_u, g->fa, g->lson, g->rson);
    g->sup2 = (uint32_t *)malloc(sizeof(uint32_t) * mm);
    kto_support(g->nc, g->off_d, g->adj_d, g->sup2);
    g->state = (uint32_t *)malloc(sizeof(uint32_t) * mm);
    for (uint64_t e = 0; e < m; e++) g->state[e] = ALIVE;
}

/* Kept subgraph of (n, off, adj) given removed[], ranked and oriented
 * (src/kc_gpu.cpp:68-126 construct_Graph), with exact supports. */
static void cg_build(core_graph *g, uint32_t n, const uint64_t *off, const uint32_t *adj,
                     const uint8_t *removed) {
    memset(g, 0, sizeof(*g));
    uint32_t *to = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *degc = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    g->inv = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    g->nc = kto_rank(n, off, adj, removed, to, g->inv, degc);
    g->off_d = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)g->nc + 1));
    g->m = kto_oriented(n, off, adj, to, g->nc, g->off_d, NULL);
    g->adj_d = (uint32_t *)malloc(sizeof(uint32_t) * (g->m ? g->m : 1));
    kto_oriented(n, off, adj, to, g->nc, g->off_d, g->adj_d);
    free(to);
    free(degc);
    cg_finish(g);
}

/* Drop every edge whose state is not ALIVE and rebuild the CSRs; supports of
 * the survivors are recounted exactly (src/kc.cu:1059-1138 with tc_need = 0,
 * followed by the exact recount at :1229). */
static void cg_compact(core_graph *g) {
    uint64_t *off_d = (uint64_t *)calloc((size_t)g->nc + 1, sizeof(uint64_t));
    uint64_t w = 0;
    for (uint32_t r = 0; r < g->nc; r++) {
        for (uint64_t j = g->off_d[r]; j < g->off_d[r + 1]; j++)
            if (g->state[j] == ALIVE) g->adj_d[w++] = g->adj_d[j];
        off_d[r + 1] = w;
    }
    uint32_t *inv = g->inv, nc = g->nc;
    uint32_t *adj_d = g->adj_d;
    g->inv = NULL; g->adj_d = NULL;
    cg_free(g);
    g->nc = nc; g->inv = inv; g->adj_d = adj_d; g->off_d = off_d; g->m = w;
    cg_finish(g);
}

/* Frontier rounds at doubled threshold k (edges with sup2 < k go), state
 * stamps as in the reference: ALIVE, st = in the current frontier, < st =
 * deleted (src/kc_gpu.h:6-43 encodings; src/kc.cu:975-1029). For a frontier
 * edge (u, v) and common neighbour w whose edges e1 = (u,w), e2 = (v,w) are
 * not deleted, both lose num = 2 - [e1 in frontier] - [e2 in frontier]
 * (src/kc.cu:122-146; src/kc_cpu.cpp:130-138), which keeps every surviving
 * support exact whatever the processing order. A frontier holding at least a
 * third of the live edges is dropped and the rest recounted
 * (src/kc.cu:988, 1206-1233). Returns the number of live edges. */
static uint64_t peel_level(core_graph *g, uint32_t k, uint32_t *st, uint64_t alive) {
    for (;;) {
        uint64_t fsz = 0;
        const uint32_t s = *st;
        for (uint64_t e = 0; e < g->m; e++)
            if (g->state[e] == ALIVE && g->sup2[e] < k) { g->state[e] = s; fsz++; }
        if (fsz == 0) return alive;
        if (fsz == alive) {
            for (uint64_t e = 0; e < g->m; e++) if (g->state[e] == s) g->state[e] = s - 1;
            (*st)++;
            return 0;
        }
        if (3 * fsz >= alive) {
            cg_compact(g);
            *st = 2;
            alive = g->m;
            continue;
        }
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t e = 0; e < (int64_t)g->m; e++) {
            if (g->state[e] != s) continue;
            const uint32_t u = g->src[e], v = g->adj_d[e];
            uint64_t i = g->off_u[u], j = g->off_u[v];
            const uint64_t ie = g->off_u[u + 1], je = g->off_u[v + 1];
            while (i < ie && j < je) {
                const uint32_t x = g->adj_u[i], y = g->adj_u[j];
                if (x == y) {
                    const uint64_t e1 = g->fa[i], e2 = g->fa[j];
                    const uint32_t m1 = g->state[e1], m2 = g->state[e2];
                    if (m1 >= s && m2 >= s) {
                        const uint32_t num = 2u - (m1 == s) - (m2 == s);
                        if (num) {
                            __atomic_fetch_sub(&g->sup2[e1], num, __ATOMIC_RELAXED);
                            __atomic_fetch_sub(&g->sup2[e2], num, __ATOMIC_RELAXED);
                        }
                    }
                }
                i += (x <= y);
                j += (y <= x);
            }
        }
        alive -= fsz;
        (*st)++;
        if (2 * (g->m - alive) >= g->m) { cg_compact(g); *st = 2; alive = g->m; }
    }
}

/* Upward level search on a core graph starting at doubled threshold k0
 * (src/kc.cu:1169-1305 with step 1). If `ans` is non-zero and the first
 * level keeps exactly `ans` edges the search stops there (src/kc.cu:1240).
 * Writes the survivors of the last non-empty level into *out_pairs (malloc,
 * original ids, u < v) when out_pairs is non-NULL. Returns kmax (0 if none). */
static uint32_t truss_search(core_graph *g, uint32_t k0, uint64_t ans, uint64_t *cnt,
                             uint32_t **out_pairs) {
    uint32_t kmax = 0, st = 2;
    uint64_t alive = g->m;
    *cnt = 0;
    for (uint32_t k = k0;; k += 2) {
        alive = peel_level(g, k, &st, alive);
        if (alive == 0) break;
        kmax = (k >> 1) + 2;
        *cnt = alive;
        if (out_pairs) {
            free(*out_pairs);
            *out_pairs = (uint32_t *)malloc(sizeof(uint32_t) * 2 * alive);
            uint64_t w = 0;
            for (uint64_t e = 0; e < g->m; e++) {
                if (g->state[e] != ALIVE) continue;
                uint32_t a = g->inv[g->src[e]], b = g->inv[g->adj_d[e]];
                (*out_pairs)[2 * w] = a < b ? a : b;
                (*out_pairs)[2 * w + 1] = a < b ? b : a;
                w++;
            }
        }
        if (ans && k == k0 && alive == ans) break;
    }
    return kmax;
}

typedef struct {
    uint32_t kmax;
    uint64_t truss_edges;
    uint32_t k_hat;          /* pass-1 kmax on the max-core */
    uint64_t k_hat_edges;
    uint32_t max_core;       /* largest k with a non-empty k-core */
    uint32_t core_k;         /* pass-2 core threshold L - 1 */
    uint32_t n_core;         /* pass-2 core size */
    uint64_t m_core;
} kto_result;

static int cmp_pair(const void *a, const void *b) {
    const uint32_t *x = (const uint32_t *)a, *y = (const uint32_t *)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    return x[1] < y[1] ? -1 : x[1] > y[1];
}

/* Two-pass kmax-truss (src/kc_gpu.cpp:336-341):
 *   pass 1 — max-core C0, its max truss k̂ (a lower bound of kmax);
 *   pass 2 — the (L-1)-core with L = max(k̂, 3), exact supports, levels
 *            upward from L with the pass-1 early exit.
 * The kmax-truss lies in the (kmax-1)-core ⊆ (L-1)-core, so pass 2 is exact.
 * When truss_pairs is non-NULL it receives up to cap edges of the kmax-truss
 * (original ids, u < v, sorted). Returns 0. */
int kto_solve(uint32_t n, const uint64_t *off, const uint32_t *adj, kto_result *res,
              uint32_t *truss_pairs, uint64_t cap) {
    memset(res, 0, sizeof(*res));
    uint8_t *removed = (uint8_t *)calloc(n ? n : 1, 1);
    core_graph g;
    uint64_t cnt = 0;
    res->max_core = kto_max_core(n, off, adj, removed);
    if (res->max_core == 0) { free(removed); return 0; }
    cg_build(&g, n, off, adj, removed);
    res->k_hat = truss_search(&g, 2, 0, &cnt, NULL);
    res->k_hat_edges = cnt;
    cg_free(&g);
    const uint32_t L = res->k_hat >= 3 ? res->k_hat : 3;
    res->core_k = L - 1;
    memset(removed, 0, n);
    kto_kcore(n, off, adj, L - 1, removed);
    cg_build(&g, n, off, adj, removed);
    res->n_core = g.nc;
    res->m_core = g.m;
    uint32_t *pairs = NULL;
    res->kmax = truss_search(&g, 2 * (L - 2), res->k_hat >= 3 ? res->k_hat_edges : 0, &cnt,
                             truss_pairs ? &pairs : NULL);
    res->truss_edges = res->kmax ? cnt : 0;
    if (truss_pairs && pairs) {
        qsort(pairs, cnt, 2 * sizeof(uint32_t), cmp_pair);
        memcpy(truss_pairs, pairs, sizeof(uint32_t) * 2 * (cnt < cap ? cnt : cap));
    }
    free(pairs);
    cg_free(&g);
    free(removed);
    return 0;
}

/* Core numbers by the bucket peel of Batagelj & Zaversnik (2003) — not a
 * reference function; an independent pin for kto_kcore / kto_max_core. */
uint32_t kto_core_numbers(uint32_t n, const uint64_t *off, const uint32_t *adj, uint32_t *core) {
    uint32_t md = 0;
    for (uint32_t v = 0; v < n; v++) {
        core[v] = (uint32_t)(off[v + 1] - off[v]);
        if (core[v] > md) md = core[v];
    }
    uint32_t *bin = (uint32_t *)calloc((size_t)md + 2, sizeof(uint32_t));
    uint32_t *pos = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *vert = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t v = 0; v < n; v++) bin[core[v]]++;
    uint32_t start = 0;
    for (uint32_t d = 0; d <= md; d++) { uint32_t c = bin[d]; bin[d] = start; start += c; }
    for (uint32_t v = 0; v < n; v++) { pos[v] = bin[core[v]]; vert[pos[v]] = v; bin[core[v]]++; }
    for (uint32_t d = md; d >= 1; d--) bin[d] = bin[d - 1];
    bin[0] = 0;
    uint32_t kmax = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t v = vert[i];
        if (core[v] > kmax) kmax = core[v];
        for (uint64_t j = off[v]; j < off[v + 1]; j++) {
            uint32_t u = adj[j];
            if (core[u] > core[v]) {
                uint32_t du = core[u], pu = pos[u], pw = bin[du], w = vert[pw];
                if (u != w) { pos[u] = pw; vert[pu] = w; pos[w] = pu; vert[pw] = u; }
                bin[du]++;
                core[u]--;
            }
        }
    }
    free(bin); free(pos); free(vert);
    return kmax;
}
