/*
 * oracle.c -- CPU oracle for the GROUP BY hot path. TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, as the checker or as the timed CPU baseline. See oracle.h.
 *
 * Pinning (tests/test_oracle.py):
 *   orc_hash_u64           golden vectors from the reference's own
 *                          ankerl::unordered_dense hash, compiled from
 *                          include/third_party/unordered_dense.h into
 *                          oracle/_ref (tests/golden/hash_u64.json), plus
 *                          the closed form lo^hi of the 128-bit product.
 *   orc_groupby_exact      the reference's known-answer tests
 *                          (test_aggregation.cpp:16-120, 196-212;
 *                          test_parallel.cpp:105-160;
 *                          test_performance.cpp:153-171), closed forms
 *                          (SUM of 1..n, arithmetic series per group),
 *                          brute force on tiny inputs, numpy/pandas
 *                          cross-checks, the reference's InlineRowAgg<N>
 *                          and RadixCountAgg built into oracle/_ref, and
 *                          golden fixtures of the plain aggregate workload.
 *                          Double SUM/AVG specials (IEEE result of the
 *                          exact sum) and MIN/MAX over NaN (skipped unless
 *                          a group holds only NaN) against math.fsum and
 *                          brute-force rules on small inputs.
 *   orc_port_groupby       equal to orc_groupby_exact wherever the
 *                          reference's double arithmetic is exact
 *                          (|partial sums| < 2^53).
 *   orc_port_radix_count   equal to the reference RadixCountAgg in oracle/_ref.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Hash: ankerl::unordered_dense::hash<uint64_t> = wyhash::mix(x, C)  */
/* (include/third_party/unordered_dense.h:158-191 mum/mix, :266-267). */
/* ------------------------------------------------------------------ */
uint64_t orc_hash_u64(uint64_t x) {
    unsigned __int128 r = (unsigned __int128)x * 0x9E3779B97F4A7C15ULL;
    return (uint64_t)r ^ (uint64_t)(r >> 64);
}

void orc_free(void *p) { free(p); }

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz ? sz : 1);
    if (!p) abort();
    return p;
}

static void *xrealloc(void *p, size_t sz) {
    void *q = realloc(p, sz ? sz : 1);
    if (!q) abort();
    return q;
}

/* ------------------------------------------------------------------ */
/* Column access helpers                                              */
/* ------------------------------------------------------------------ */
static inline int col_valid(const gbagg_column_view *c, uint64_t r) {
    return c->validity == NULL || c->validity[r] != 0;
}

static inline int64_t col_i64(const gbagg_column_view *c, uint64_t r) {
    switch (c->type) {
    case GBAGG_TYPE_INTEGER: return (int64_t)((const int32_t *)c->data)[r];
    case GBAGG_TYPE_BIGINT: return ((const int64_t *)c->data)[r];
    case GBAGG_TYPE_VARCHAR: return (int64_t)((const uint32_t *)c->data)[r];
    default: return 0;
    }
}

static inline double col_f64(const gbagg_column_view *c, uint64_t r) {
    switch (c->type) {
    case GBAGG_TYPE_INTEGER: return (double)((const int32_t *)c->data)[r];
    case GBAGG_TYPE_BIGINT: return (double)((const int64_t *)c->data)[r];
    case GBAGG_TYPE_DOUBLE: return ((const double *)c->data)[r];
    default: return 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* Predicates: EvalSimplePredicates (physical_planner.cpp:396-496).   */
/* VARCHAR predicates are evaluated once per dictionary entry, as     */
/* BuildTypedKeepMask does (:654-691).                                */
/* ------------------------------------------------------------------ */
typedef struct {
    const gbagg_predicate *p;
    const gbagg_column_view *c;
    uint8_t *dict_match; /* VARCHAR: per dictionary entry */
} pred_rt;

static int like_contains(const char *hs, uint32_t hl, const char *nd, uint32_t nl) {
    if (nl == 0) return 1;
    if (hl < nl) return 0;
    return memmem(hs, hl, nd, nl) != NULL;
}

static void preds_prepare(const gbagg_query_desc *q, const gbagg_column_view *cols, pred_rt *rt) {
    for (uint32_t i = 0; i < q->num_predicates; i++) {
        const gbagg_predicate *p = &q->predicates[i];
        const gbagg_column_view *c = &cols[p->column];
        rt[i].p = p;
        rt[i].c = c;
        rt[i].dict_match = NULL;
        if (c->type == GBAGG_TYPE_VARCHAR) {
            rt[i].dict_match = xcalloc(c->dict_size, 1);
            for (uint32_t d = 0; d < c->dict_size; d++) {
                const char *s = c->dict_bytes + c->dict_offsets[d];
                uint32_t l = c->dict_offsets[d + 1] - c->dict_offsets[d];
                int m;
                if (p->op == GBAGG_CMP_LIKE || p->op == GBAGG_CMP_NOT_LIKE) {
                    m = like_contains(s, l, p->sval, p->slen);
                    if (p->op == GBAGG_CMP_NOT_LIKE) m = !m;
                } else {
                    int eq = (l == p->slen) && (l == 0 || memcmp(s, p->sval, l) == 0);
                    m = (p->op == GBAGG_CMP_EQ) ? eq : (p->op == GBAGG_CMP_NE ? !eq : 0);
                }
                rt[i].dict_match[d] = (uint8_t)m;
            }
        }
    }
}

static void preds_release(const gbagg_query_desc *q, pred_rt *rt) {
    for (uint32_t i = 0; i < q->num_predicates; i++) free(rt[i].dict_match);
}

static int preds_eval(const gbagg_query_desc *q, const pred_rt *rt, uint64_t r) {
    for (uint32_t i = 0; i < q->num_predicates; i++) {
        const gbagg_predicate *p = rt[i].p;
        const gbagg_column_view *c = rt[i].c;
        if (!col_valid(c, r)) return 0;
        if (c->type == GBAGG_TYPE_VARCHAR) {
            uint32_t code = ((const uint32_t *)c->data)[r];
            if (code >= c->dict_size || !rt[i].dict_match[code]) return 0;
            continue;
        }
        if (p->op == GBAGG_CMP_IN) {
            int64_t v = col_i64(c, r);
            int hit = 0;
            for (uint32_t j = 0; j < p->in_count; j++) hit |= (p->in_set[j] == v);
            if (!hit) return 0;
            continue;
        }
        if (c->type == GBAGG_TYPE_INTEGER || c->type == GBAGG_TYPE_BIGINT) {
            /* int64-exact compare against ival (:451-466) */
            int64_t v = col_i64(c, r), lit = p->ival;
            int ok;
            switch (p->op) {
            case GBAGG_CMP_LT: ok = v < lit; break;
            case GBAGG_CMP_LE: ok = v <= lit; break;
            case GBAGG_CMP_GT: ok = v > lit; break;
            case GBAGG_CMP_GE: ok = v >= lit; break;
            case GBAGG_CMP_EQ: ok = v == lit; break;
            case GBAGG_CMP_NE: ok = v != lit; break;
            default: ok = 0;
            }
            if (!ok) return 0;
            continue;
        }
        {
            double v = col_f64(c, r), lit = p->dval;
            int ok;
            switch (p->op) {
            case GBAGG_CMP_LT: ok = v < lit; break;
            case GBAGG_CMP_LE: ok = v <= lit; break;
            case GBAGG_CMP_GT: ok = v > lit; break;
            case GBAGG_CMP_GE: ok = v >= lit; break;
            case GBAGG_CMP_EQ: ok = v == lit; break;
            case GBAGG_CMP_NE: ok = v != lit; break;
            default: ok = 0;
            }
            if (!ok) return 0;
        }
    }
    return 1;
}

/* ------------------------------------------------------------------ */
/* VARCHAR group keys: canonical code = first dictionary entry with    */
/* the same bytes, so duplicate dictionary entries form one group.     */
/* ------------------------------------------------------------------ */
static uint64_t str_hash(const char *s, uint32_t l) {
    uint64_t h = 1469598103934665603ULL; /* FNV-1a, then the wyhash mix */
    for (uint32_t i = 0; i < l; i++) h = (h ^ (uint8_t)s[i]) * 1099511628211ULL;
    return orc_hash_u64(h ^ l);
}

static uint32_t *dict_canonical(const gbagg_column_view *c) {
    uint32_t n = c->dict_size;
    uint32_t *canon = xcalloc(n, sizeof(uint32_t));
    uint64_t cap = 16;
    while (cap < 2ULL * n + 2) cap <<= 1;
    int64_t *slot = xcalloc(cap, sizeof(int64_t));
    for (uint64_t i = 0; i < cap; i++) slot[i] = -1;
    for (uint32_t d = 0; d < n; d++) {
        const char *s = c->dict_bytes + c->dict_offsets[d];
        uint32_t l = c->dict_offsets[d + 1] - c->dict_offsets[d];
        uint64_t pos = str_hash(s, l) & (cap - 1);
        for (;;) {
            if (slot[pos] < 0) { slot[pos] = d; canon[d] = d; break; }
            uint32_t e = (uint32_t)slot[pos];
            const char *es = c->dict_bytes + c->dict_offsets[e];
            uint32_t el = c->dict_offsets[e + 1] - c->dict_offsets[e];
            if (el == l && memcmp(es, s, l) == 0) { canon[d] = e; break; }
            pos = (pos + 1) & (cap - 1);
        }
    }
    free(slot);
    return canon;
}

/* ------------------------------------------------------------------ */
/* Exact oracle                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    int64_t k[GBAGG_MAX_GROUP_COLS];
    uint8_t nullmask;
} gkey;

typedef struct {
    /* per group, per aggregate */
    int64_t *cnt;
    __int128 *isum;
    double *fsum, *fcomp;
    int64_t *imin, *imax;
    double *fmin, *fmax;
    uint8_t *fhas; /* MIN/MAX: a non-NaN value seen; SUM/AVG: 1 = +Inf seen, 2 = -Inf seen (NaN: both) */
} accs;

typedef struct gtable {
    uint32_t na;
    uint64_t ngroups, cap_groups;
    gkey *keys;
    accs a;
    uint64_t tcap;
    int64_t *table; /* group index or -1 */
    struct gtable *dset; /* COUNT(DISTINCT): one entry per (group * 16 + agg, value) seen */
} gtable;

/* Keys hold every group column (unused ones zero). */
static inline uint64_t gkey_hash(const gkey *k) {
    uint64_t h = orc_hash_u64((uint64_t)k->nullmask);
    for (int g = 0; g < GBAGG_MAX_GROUP_COLS; g++) h = (h * 0x9E3779B97F4A7C15ULL) ^ orc_hash_u64((uint64_t)k->k[g]);
    return h;
}

static inline int gkey_eq(const gkey *a, const gkey *b) {
    if (a->nullmask != b->nullmask) return 0;
    for (int g = 0; g < GBAGG_MAX_GROUP_COLS; g++)
        if (a->k[g] != b->k[g]) return 0;
    return 1;
}

/* A DOUBLE group key as the bits of its canonical value: SQL groups -0.0
 * with 0.0 and every NaN together. */
static inline int64_t double_key(double d) {
    uint64_t b;
    memcpy(&b, &d, 8);
    if ((b << 1) == 0) b = 0;  /* -0.0 */
    if ((b & 0x7FF0000000000000ULL) == 0x7FF0000000000000ULL && (b & 0x000FFFFFFFFFFFFFULL))
        b = 0x7FF8000000000000ULL; /* NaN */
    return (int64_t)b;
}

static void gtable_init(gtable *t, uint32_t na) {
    memset(t, 0, sizeof(*t));
    t->na = na;
    t->tcap = 1024;
    t->table = xcalloc(t->tcap, sizeof(int64_t));
    for (uint64_t i = 0; i < t->tcap; i++) t->table[i] = -1;
}

static void gtable_grow_groups(gtable *t) {
    uint64_t nc = t->cap_groups ? t->cap_groups * 2 : 1024;
    uint32_t na = t->na ? t->na : 1;
    t->keys = xrealloc(t->keys, nc * sizeof(gkey));
#define GROW(f, T) t->a.f = xrealloc(t->a.f, nc * na * sizeof(T))
    GROW(cnt, int64_t); GROW(isum, __int128); GROW(fsum, double); GROW(fcomp, double);
    GROW(imin, int64_t); GROW(imax, int64_t); GROW(fmin, double); GROW(fmax, double);
    GROW(fhas, uint8_t);
#undef GROW
    t->cap_groups = nc;
}

static void gtable_rehash(gtable *t) {
    uint64_t nc = t->tcap * 2;
    int64_t *nt = xcalloc(nc, sizeof(int64_t));
    for (uint64_t i = 0; i < nc; i++) nt[i] = -1;
    for (uint64_t g = 0; g < t->ngroups; g++) {
        uint64_t pos = gkey_hash(&t->keys[g]) & (nc - 1);
        while (nt[pos] >= 0) pos = (pos + 1) & (nc - 1);
        nt[pos] = (int64_t)g;
    }
    free(t->table);
    t->table = nt;
    t->tcap = nc;
}

static uint64_t gtable_find(gtable *t, const gkey *k, uint64_t h) {
    uint64_t pos = h & (t->tcap - 1);
    for (;;) {
        int64_t g = t->table[pos];
        if (g < 0) break;
        if (gkey_eq(&t->keys[g], k)) return (uint64_t)g;
        pos = (pos + 1) & (t->tcap - 1);
    }
    /* insert */
    if (t->ngroups == t->cap_groups) gtable_grow_groups(t);
    uint64_t g = t->ngroups++;
    t->keys[g] = *k;
    for (uint32_t a = 0; a < t->na; a++) {
        uint64_t i = g * t->na + a;
        t->a.cnt[i] = 0; t->a.isum[i] = 0; t->a.fsum[i] = 0.0; t->a.fcomp[i] = 0.0;
        t->a.imin[i] = INT64_MAX; t->a.imax[i] = INT64_MIN;
        t->a.fmin[i] = 0.0; t->a.fmax[i] = 0.0; t->a.fhas[i] = 0;
    }
    t->table[pos] = (int64_t)g;
    if (t->ngroups * 2 > t->tcap) gtable_rehash(t);
    return g;
}

static void gtable_free(gtable *t) {
    if (t->dset) { gtable_free(t->dset); free(t->dset); }
    free(t->keys); free(t->table);
    free(t->a.cnt); free(t->a.isum); free(t->a.fsum); free(t->a.fcomp);
    free(t->a.imin); free(t->a.imax); free(t->a.fmin); free(t->a.fmax); free(t->a.fhas);
}

/* Kahan-Neumaier compensated summation (Neumaier 1974). */
static inline void neumaier_add(double *s, double *c, double x) {
    double t = *s + x;
    if (fabs(*s) >= fabs(x)) *c += (*s - t) + x;
    else *c += (x - t) + *s;
    *s = t;
}

typedef struct {
    const gbagg_query_desc *q;
    const gbagg_column_view *cols;
    uint32_t **canon; /* per group col: canonical dictionary codes or NULL */
    uint32_t **acanon; /* per aggregate: canonical codes of a VARCHAR COUNT(DISTINCT) argument */
    pred_rt *preds;
} exact_ctx;

/* Byte order of two entries of a VARCHAR column's dictionary (memcmp, then
 * the shorter string first). */
static int dict_cmp(const gbagg_column_view *c, uint32_t a, uint32_t b) {
    const uint32_t la = c->dict_offsets[a + 1] - c->dict_offsets[a], lb = c->dict_offsets[b + 1] - c->dict_offsets[b];
    const int m = memcmp(c->dict_bytes + c->dict_offsets[a], c->dict_bytes + c->dict_offsets[b], la < lb ? la : lb);
    return m ? m : (la < lb ? -1 : la > lb ? 1 : 0);
}

/* Builds the group key of row r; returns 0 if the row is dropped. */
static inline int row_key(const exact_ctx *x, uint64_t r, gkey *k) {
    const gbagg_query_desc *q = x->q;
    memset(k, 0, sizeof(*k));
    for (uint32_t g = 0; g < q->num_group_cols; g++) {
        const gbagg_column_view *c = &x->cols[q->group_cols[g]];
        if (!col_valid(c, r)) {
            if (q->drop_null_keys) return 0;
            k->nullmask |= (uint8_t)(1u << g);
            continue;
        }
        int64_t v = c->type == GBAGG_TYPE_DOUBLE ? double_key(((const double *)c->data)[r]) : col_i64(c, r);
        if (x->canon[g]) v = x->canon[g][(uint32_t)v];
        k->k[g] = v;
    }
    return 1;
}

static inline void row_update(const exact_ctx *x, gtable *t, uint64_t g, uint64_t r) {
    const gbagg_query_desc *q = x->q;
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        const gbagg_agg *ag = &q->aggs[a];
        uint64_t i = g * t->na + a;
        if (ag->kind == GBAGG_AGG_COUNT_STAR) { t->a.cnt[i]++; continue; }
        const gbagg_column_view *c = &x->cols[ag->column];
        if (!col_valid(c, r)) continue;
        if (ag->kind == GBAGG_AGG_COUNT_DISTINCT) {
            /* distinct non-NULL values per group (physical_planner.cpp:12325-12330) */
            gkey dk;
            memset(&dk, 0, sizeof(dk));
            /* DOUBLE arguments by value: -0.0 as 0.0, one NaN */
            int64_t v = c->type == GBAGG_TYPE_DOUBLE ? double_key(((const double *)c->data)[r]) : col_i64(c, r);
            if (x->acanon[a]) v = x->acanon[a][(uint32_t)v];
            dk.k[0] = (int64_t)(g * 16 + a);
            dk.k[1] = v;
            dk.nullmask = 0;
            if (!t->dset) {
                t->dset = xcalloc(1, sizeof(gtable));
                gtable_init(t->dset, 0);
            }
            uint64_t before = t->dset->ngroups;
            gtable_find(t->dset, &dk, gkey_hash(&dk));
            if (t->dset->ngroups > before) t->a.cnt[i]++;
            continue;
        }
        t->a.cnt[i]++;
        if (ag->kind == GBAGG_AGG_COUNT) continue;
        if (c->type == GBAGG_TYPE_DOUBLE) {
            double v = ((const double *)c->data)[r];
            if (ag->kind == GBAGG_AGG_SUM || ag->kind == GBAGG_AGG_AVG) {
                /* IEEE semantics of the exact sum: a NaN, or +Inf together
                 * with -Inf, gives NaN; otherwise an infinity gives itself;
                 * finite values are summed compensated */
                if (isnan(v)) t->a.fhas[i] |= 3;
                else if (isinf(v)) t->a.fhas[i] |= v > 0 ? 1 : 2;
                else neumaier_add(&t->a.fsum[i], &t->a.fcomp[i], v);
            } else if (isnan(v)) {
                /* MIN/MAX skip NaN unless every input of the group is NaN */
                if (!t->a.fhas[i]) *(ag->kind == GBAGG_AGG_MIN ? &t->a.fmin[i] : &t->a.fmax[i]) = NAN;
            } else if (ag->kind == GBAGG_AGG_MIN) {
                if (!t->a.fhas[i] || v < t->a.fmin[i]) t->a.fmin[i] = v;
                t->a.fhas[i] = 1;
            } else {
                if (!t->a.fhas[i] || v > t->a.fmax[i]) t->a.fmax[i] = v;
                t->a.fhas[i] = 1;
            }
        } else if (c->type == GBAGG_TYPE_VARCHAR) {
            /* MIN/MAX of strings: byte order of the dictionary entries
             * (the generic path's Value comparison) */
            const int64_t v = col_i64(c, r);
            int64_t *cur = ag->kind == GBAGG_AGG_MIN ? &t->a.imin[i] : &t->a.imax[i];
            const int d = t->a.cnt[i] == 1 ? 0 : dict_cmp(c, (uint32_t)v, (uint32_t)*cur);
            if (t->a.cnt[i] == 1 || (ag->kind == GBAGG_AGG_MIN ? d < 0 : d > 0)) *cur = v;
        } else {
            int64_t v = col_i64(c, r);
            if (ag->kind == GBAGG_AGG_SUM || ag->kind == GBAGG_AGG_AVG) t->a.isum[i] += v;
            else if (ag->kind == GBAGG_AGG_MIN) { if (v < t->a.imin[i]) t->a.imin[i] = v; }
            else { if (v > t->a.imax[i]) t->a.imax[i] = v; }
        }
    }
}

#define NPART 64

typedef struct {
    uint32_t *rows;
    uint64_t n, cap;
} rowlist;

typedef struct {
    const exact_ctx *x;
    uint64_t begin, end;
    rowlist lists[NPART];
} scan_job;

static void *scan_worker(void *arg) {
    scan_job *j = arg;
    gkey k;
    for (uint64_t r = j->begin; r < j->end; r++) {
        if (j->x->q->num_predicates && !preds_eval(j->x->q, j->x->preds, r)) continue;
        if (!row_key(j->x, r, &k)) continue;
        uint32_t p = (uint32_t)(gkey_hash(&k) >> 58);
        rowlist *l = &j->lists[p];
        if (l->n == l->cap) {
            l->cap = l->cap ? l->cap * 2 : 1024;
            l->rows = xrealloc(l->rows, l->cap * sizeof(uint32_t));
        }
        l->rows[l->n++] = (uint32_t)(r - j->begin);
    }
    return NULL;
}

typedef struct {
    const exact_ctx *x;
    scan_job *jobs;
    int njobs;
    gtable parts[NPART];
    atomic_int next;
} agg_job;

static void *agg_worker(void *arg) {
    agg_job *a = arg;
    gkey k;
    for (;;) {
        int p = atomic_fetch_add(&a->next, 1);
        if (p >= NPART) break;
        gtable *t = &a->parts[p];
        gtable_init(t, a->x->q->num_aggs);
        for (int jj = 0; jj < a->njobs; jj++) {
            rowlist *l = &a->jobs[jj].lists[p];
            for (uint64_t i = 0; i < l->n; i++) {
                uint64_t r = a->jobs[jj].begin + l->rows[i];
                row_key(a->x, r, &k);
                uint64_t g = gtable_find(t, &k, gkey_hash(&k));
                row_update(a->x, t, g, r);
            }
        }
    }
    return NULL;
}

static int out_type_of(const gbagg_query_desc *q, const gbagg_column_view *cols, uint32_t a) {
    const gbagg_agg *ag = &q->aggs[a];
    switch (ag->kind) {
    case GBAGG_AGG_COUNT_STAR:
    case GBAGG_AGG_COUNT:
    case GBAGG_AGG_COUNT_DISTINCT: return GBAGG_TYPE_BIGINT;
    case GBAGG_AGG_AVG: return GBAGG_TYPE_DOUBLE;
    case GBAGG_AGG_SUM:
        return cols[ag->column].type == GBAGG_TYPE_DOUBLE ? GBAGG_TYPE_DOUBLE : GBAGG_TYPE_BIGINT;
    default: return cols[ag->column].type;
    }
}

/* The double SUM of (group, aggregate) i: the specials' IEEE result, else
 * the compensated sum. */
static double fsum_value(const accs *a, uint64_t i) {
    if (a->fhas[i] == 3) return NAN;
    if (a->fhas[i] == 1) return INFINITY;
    if (a->fhas[i] == 2) return -INFINITY;
    return a->fsum[i] + a->fcomp[i];
}

static void result_alloc(orc_result *out, const gbagg_query_desc *q, const gbagg_column_view *cols,
                         uint64_t n) {
    memset(out, 0, sizeof(*out));
    out->num_rows = n;
    out->num_columns = q->num_group_cols + q->num_aggs;
    for (uint32_t c = 0; c < out->num_columns; c++) {
        out->types[c] = c < q->num_group_cols ? (int32_t)cols[q->group_cols[c]].type
                                              : out_type_of(q, cols, c - q->num_group_cols);
        out->cells[c] = xcalloc(n, sizeof(int64_t));
        out->valid[c] = xcalloc(n, 1);
    }
}

static void emit_exact(const gbagg_query_desc *q, const gbagg_column_view *cols, const gtable *t,
                       uint64_t g, orc_result *out, uint64_t row) {
    for (uint32_t c = 0; c < q->num_group_cols; c++) {
        int isnull = (t->keys[g].nullmask >> c) & 1;
        out->valid[c][row] = (uint8_t)!isnull;
        out->cells[c][row] = isnull ? 0 : t->keys[g].k[c];
    }
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        uint32_t c = q->num_group_cols + a;
        uint64_t i = g * t->na + a;
        const gbagg_agg *ag = &q->aggs[a];
        int64_t cnt = t->a.cnt[i];
        int isdbl = ag->column >= 0 && cols[ag->column].type == GBAGG_TYPE_DOUBLE;
        double d;
        switch (ag->kind) {
        case GBAGG_AGG_COUNT_STAR:
        case GBAGG_AGG_COUNT:
        case GBAGG_AGG_COUNT_DISTINCT:
            out->cells[c][row] = cnt;
            out->valid[c][row] = 1;
            break;
        case GBAGG_AGG_SUM:
            out->valid[c][row] = cnt > 0;
            if (isdbl) {
                d = fsum_value(&t->a, i);
                memcpy(&out->cells[c][row], &d, 8);
            } else {
                __int128 s = t->a.isum[i];
                if (s > INT64_MAX || s < INT64_MIN) out->overflow = 1;
                out->cells[c][row] = (int64_t)s;
            }
            break;
        case GBAGG_AGG_AVG:
            out->valid[c][row] = cnt > 0;
            if (cnt > 0) {
                d = isdbl ? fsum_value(&t->a, i) / (double)cnt
                          : (double)t->a.isum[i] / (double)cnt;
                memcpy(&out->cells[c][row], &d, 8);
            }
            break;
        case GBAGG_AGG_MIN:
        case GBAGG_AGG_MAX:
            out->valid[c][row] = cnt > 0;
            if (cnt > 0) {
                if (isdbl) {
                    d = ag->kind == GBAGG_AGG_MIN ? t->a.fmin[i] : t->a.fmax[i];
                    memcpy(&out->cells[c][row], &d, 8);
                } else {
                    out->cells[c][row] = ag->kind == GBAGG_AGG_MIN ? t->a.imin[i] : t->a.imax[i];
                }
            }
            break;
        }
    }
}

int orc_groupby_exact(const gbagg_query_desc *q, const gbagg_column_view *cols, uint64_t nrows,
                      int threads, orc_result *out) {
    if (q->num_group_cols > GBAGG_MAX_GROUP_COLS || q->num_aggs > GBAGG_MAX_AGGS) return 2;
    if (threads < 1) threads = 1;
    exact_ctx x;
    uint32_t *canon[GBAGG_MAX_GROUP_COLS] = {0};
    uint32_t *acanon[GBAGG_MAX_AGGS] = {0};
    pred_rt rt[GBAGG_MAX_PREDICATES];
    x.q = q;
    x.cols = cols;
    for (uint32_t g = 0; g < q->num_group_cols; g++)
        if (cols[q->group_cols[g]].type == GBAGG_TYPE_VARCHAR) canon[g] = dict_canonical(&cols[q->group_cols[g]]);
    x.canon = canon;
    for (uint32_t a = 0; a < q->num_aggs; a++)
        if (q->aggs[a].kind == GBAGG_AGG_COUNT_DISTINCT && cols[q->aggs[a].column].type == GBAGG_TYPE_VARCHAR)
            acanon[a] = dict_canonical(&cols[q->aggs[a].column]);
    x.acanon = acanon;
    preds_prepare(q, cols, rt);
    x.preds = rt;

    /* Rows are bucketed by key hash into NPART partitions, each partition
     * aggregated by one thread in input row order, so every group sees its
     * rows in the same order as a sequential scan. */
    int nj = threads;
    uint64_t per = (nrows + nj - 1) / (nj ? nj : 1);
    if (per < 1) per = 1;
    scan_job *jobs = xcalloc(nj, sizeof(scan_job));
    pthread_t *th = xcalloc(nj, sizeof(pthread_t));
    for (int j = 0; j < nj; j++) {
        jobs[j].x = &x;
        jobs[j].begin = per * j < nrows ? per * j : nrows;
        jobs[j].end = per * (j + 1) < nrows ? per * (j + 1) : nrows;
        pthread_create(&th[j], NULL, scan_worker, &jobs[j]);
    }
    for (int j = 0; j < nj; j++) pthread_join(th[j], NULL);

    agg_job aj;
    memset(&aj, 0, sizeof(aj));
    aj.x = &x;
    aj.jobs = jobs;
    aj.njobs = nj;
    atomic_init(&aj.next, 0);
    for (int j = 0; j < nj; j++) pthread_create(&th[j], NULL, agg_worker, &aj);
    for (int j = 0; j < nj; j++) pthread_join(th[j], NULL);

    uint64_t total = 0;
    for (int p = 0; p < NPART; p++) total += aj.parts[p].ngroups;
    /* Empty input with no GROUP BY: one row (COUNT = 0, others NULL). */
    int scalar_agg = (q->num_group_cols == 0 && total == 0);
    result_alloc(out, q, cols, scalar_agg ? 1 : total);
    uint64_t row = 0;
    for (int p = 0; p < NPART; p++) {
        gtable *t = &aj.parts[p];
        for (uint64_t g = 0; g < t->ngroups; g++) emit_exact(q, cols, t, g, out, row++);
        gtable_free(t);
    }
    if (scalar_agg) {
        for (uint32_t a = 0; a < q->num_aggs; a++) {
            int k = q->aggs[a].kind;
            out->valid[a][0] = (k == GBAGG_AGG_COUNT_STAR || k == GBAGG_AGG_COUNT || k == GBAGG_AGG_COUNT_DISTINCT);
        }
    }
    for (int j = 0; j < nj; j++)
        for (int p = 0; p < NPART; p++) free(jobs[j].lists[p].rows);
    free(jobs);
    free(th);
    preds_release(q, rt);
    for (uint32_t g = 0; g < q->num_group_cols; g++) free(canon[g]);
    for (uint32_t a = 0; a < q->num_aggs; a++) free(acanon[a]);
    return 0;
}

void orc_result_free(orc_result *r) {
    for (uint32_t c = 0; c < r->num_columns; c++) {
        free(r->cells[c]);
        free(r->valid[c]);
    }
    memset(r, 0, sizeof(*r));
}

/* ------------------------------------------------------------------ */
/* Exact TopN with deterministic tie-break                             */
/* ------------------------------------------------------------------ */
typedef struct {
    const orc_result *r;
    uint32_t nk;
    int32_t col;
    int32_t asc;
} topn_ctx;

static topn_ctx g_tctx; /* qsort has no context argument; orc_topn is not reentrant */

static int cmp_cell(const orc_result *r, uint32_t c, uint64_t a, uint64_t b) {
    int va = r->valid[c][a], vb = r->valid[c][b];
    if (va != vb) return va ? 1 : -1; /* NULL first */
    if (!va) return 0;
    if (r->types[c] == GBAGG_TYPE_DOUBLE) {
        double x, y;
        memcpy(&x, &r->cells[c][a], 8);
        memcpy(&y, &r->cells[c][b], 8);
        return x < y ? -1 : (x > y ? 1 : 0);
    }
    int64_t x = r->cells[c][a], y = r->cells[c][b];
    return x < y ? -1 : (x > y ? 1 : 0);
}

static int topn_cmp(const void *pa, const void *pb) {
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    int c = cmp_cell(g_tctx.r, (uint32_t)g_tctx.col, a, b);
    if (c) return g_tctx.asc ? c : -c;
    for (uint32_t k = 0; k < g_tctx.nk; k++) {
        c = cmp_cell(g_tctx.r, k, a, b);
        if (c) return c;
    }
    return 0;
}

uint64_t orc_topn(const orc_result *r, uint32_t num_key_cols, int32_t col, int32_t ascending,
                  uint64_t k, uint64_t *idx) {
    uint64_t n = r->num_rows;
    uint64_t *all = xcalloc(n, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) all[i] = i;
    g_tctx.r = r;
    g_tctx.nk = num_key_cols;
    g_tctx.col = col;
    g_tctx.asc = ascending;
    qsort(all, n, sizeof(uint64_t), topn_cmp);
    uint64_t m = k < n ? k : n;
    memcpy(idx, all, m * sizeof(uint64_t));
    free(all);
    return m;
}

/* ------------------------------------------------------------------ */
/* Port of the reference's single-column fused path                    */
/* (physical_planner.cpp:7613-7931, emit :8027-8173).                  */
/* ------------------------------------------------------------------ */
#define PORT_RG_ROWS 122880 /* ROW_GROUP_SIZE, constants.hpp:18 */
#define PORT_MERGE_SHARDS 4 /* physical_planner.cpp:7120 */

enum { AK_COUNT_STAR, AK_COUNT, AK_SUM, AK_MIN, AK_MAX };

typedef struct { /* AggState (physical_planner.cpp:6080-6091), numeric fields */
    int64_t count;
    double sum;
    uint8_t has_min, has_max;
    double sum_min, sum_max;
} pstate;

typedef struct { /* per-thread map: key -> group, insertion-ordered like ankerl */
    uint64_t n, cap, tcap;
    int64_t *keys;
    pstate *st;
    int64_t *table;
    uint32_t na;
} pmap;

static void pmap_init(pmap *m, uint32_t na, uint64_t reserve) {
    memset(m, 0, sizeof(*m));
    m->na = na ? na : 1;
    m->tcap = 64;
    while (m->tcap < reserve * 2) m->tcap <<= 1;
    m->table = xcalloc(m->tcap, sizeof(int64_t));
    for (uint64_t i = 0; i < m->tcap; i++) m->table[i] = -1;
}

static void pmap_free(pmap *m) {
    free(m->keys);
    free(m->st);
    free(m->table);
}

static pstate *pmap_get(pmap *m, int64_t k) {
    uint64_t pos = orc_hash_u64((uint64_t)k) & (m->tcap - 1);
    for (;;) {
        int64_t g = m->table[pos];
        if (g < 0) break;
        if (m->keys[g] == k) return &m->st[(uint64_t)g * m->na];
        pos = (pos + 1) & (m->tcap - 1);
    }
    if (m->n == m->cap) {
        m->cap = m->cap ? m->cap * 2 : 256;
        m->keys = xrealloc(m->keys, m->cap * sizeof(int64_t));
        m->st = xrealloc(m->st, m->cap * m->na * sizeof(pstate));
    }
    uint64_t g = m->n++;
    m->keys[g] = k;
    memset(&m->st[g * m->na], 0, m->na * sizeof(pstate));
    m->table[pos] = (int64_t)g;
    if (m->n * 2 > m->tcap) { /* grow */
        uint64_t nc = m->tcap * 2;
        int64_t *nt = xcalloc(nc, sizeof(int64_t));
        for (uint64_t i = 0; i < nc; i++) nt[i] = -1;
        for (uint64_t j = 0; j < m->n; j++) {
            uint64_t p2 = orc_hash_u64((uint64_t)m->keys[j]) & (nc - 1);
            while (nt[p2] >= 0) p2 = (p2 + 1) & (nc - 1);
            nt[p2] = (int64_t)j;
        }
        free(m->table);
        m->table = nt;
        m->tcap = nc;
    }
    return &m->st[g * m->na];
}

typedef struct {
    const gbagg_query_desc *q;
    const gbagg_column_view *cols;
    uint64_t nrows;
    int kinds[GBAGG_MAX_AGGS];
    pred_rt preds[GBAGG_MAX_PREDICATES];
    atomic_ulong next_rg;
    pmap *tls;
    pmap shards[PORT_MERGE_SHARDS];
    int nt;
} port_ctx;

/* BuildTypedKeepMask (physical_planner.cpp:505-720): a uint8 mask per row
 * group, AND-ed predicate by predicate. */
static void port_keep_mask(port_ctx *c, uint64_t r0, uint64_t n, uint8_t *mask) {
    memset(mask, 1, n);
    for (uint32_t i = 0; i < c->q->num_predicates; i++) {
        for (uint64_t r = 0; r < n; r++) {
            if (!mask[r]) continue;
            gbagg_query_desc one = *c->q;
            one.num_predicates = 1;
            one.predicates = c->preds[i].p;
            mask[r] = (uint8_t)preds_eval(&one, &c->preds[i], r0 + r);
        }
    }
}

static void *port_worker(void *arg) {
    void **pa = arg;
    port_ctx *c = pa[0];
    int tid = (int)(intptr_t)pa[1];
    pmap *m = &c->tls[tid];
    const gbagg_query_desc *q = c->q;
    const gbagg_column_view *kc = &c->cols[q->group_cols[0]];
    uint8_t *mask = q->num_predicates ? xcalloc(PORT_RG_ROWS, 1) : NULL;
    uint64_t nrg = (c->nrows + PORT_RG_ROWS - 1) / PORT_RG_ROWS;
    for (;;) {
        uint64_t rg = atomic_fetch_add(&c->next_rg, 1);
        if (rg >= nrg) break;
        uint64_t r0 = rg * PORT_RG_ROWS;
        uint64_t n = c->nrows - r0 < PORT_RG_ROWS ? c->nrows - r0 : PORT_RG_ROWS;
        if (mask) port_keep_mask(c, r0, n, mask);
        for (uint64_t i = 0; i < n; i++) {
            if (mask && !mask[i]) continue;
            uint64_t r = r0 + i;
            pstate *s = pmap_get(m, col_i64(kc, r));
            for (uint32_t a = 0; a < q->num_aggs; a++) {
                pstate *st = &s[a];
                if (c->kinds[a] == AK_COUNT_STAR) { st->count++; continue; }
                const gbagg_column_view *ac = &c->cols[q->aggs[a].column];
                if (!col_valid(ac, r)) continue;
                double val = col_f64(ac, r);
                switch (c->kinds[a]) {
                case AK_COUNT: st->count++; break;
                case AK_SUM: st->sum += val; st->count++; break;
                case AK_MIN:
                    if (!st->has_min || val < st->sum_min) { st->sum_min = val; st->has_min = 1; }
                    break;
                case AK_MAX:
                    if (!st->has_max || val > st->sum_max) { st->sum_max = val; st->has_max = 1; }
                    break;
                }
            }
        }
    }
    free(mask);
    return NULL;
}

static void *port_merge_worker(void *arg) {
    void **pa = arg;
    port_ctx *c = pa[0];
    int s = (int)(intptr_t)pa[1];
    pmap *out = &c->shards[s];
    uint32_t na = c->q->num_aggs;
    for (int t = 0; t < c->nt; t++) {
        pmap *m = &c->tls[t];
        for (uint64_t g = 0; g < m->n; g++) {
            int64_t k = m->keys[g];
            uint64_t h = (uint64_t)k * 0x9E3779B97F4A7C15ULL;
            if ((int)(h & (PORT_MERGE_SHARDS - 1)) != s) continue;
            pstate *d = pmap_get(out, k);
            const pstate *sr = &m->st[g * m->na];
            for (uint32_t a = 0; a < na; a++) { /* merge_states (:7830-7863) */
                d[a].count += sr[a].count;
                d[a].sum += sr[a].sum;
                if (sr[a].has_min) {
                    if (!d[a].has_min) { d[a].has_min = 1; d[a].sum_min = sr[a].sum_min; }
                    else if (sr[a].sum_min < d[a].sum_min) d[a].sum_min = sr[a].sum_min;
                }
                if (sr[a].has_max) {
                    if (!d[a].has_max) { d[a].has_max = 1; d[a].sum_max = sr[a].sum_max; }
                    else if (sr[a].sum_max > d[a].sum_max) d[a].sum_max = sr[a].sum_max;
                }
            }
        }
    }
    return NULL;
}

/* EmitAggValue (agg_emit_helpers.cpp:26-98) for one group. */
static void port_emit(port_ctx *c, int64_t key, const pstate *s, orc_result *out, uint64_t row) {
    const gbagg_query_desc *q = c->q;
    out->cells[0][row] = key;
    out->valid[0][row] = 1;
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        uint32_t col = 1 + a;
        const pstate *st = &s[a];
        int rt = out->types[col];
        double d;
        switch (q->aggs[a].kind) {
        case GBAGG_AGG_COUNT_STAR:
        case GBAGG_AGG_COUNT:
            out->cells[col][row] = st->count;
            out->valid[col][row] = 1;
            break;
        case GBAGG_AGG_SUM:
            out->valid[col][row] = st->count != 0;
            if (rt == GBAGG_TYPE_BIGINT) out->cells[col][row] = (int64_t)st->sum;
            else memcpy(&out->cells[col][row], &st->sum, 8);
            break;
        case GBAGG_AGG_AVG:
            out->valid[col][row] = st->count > 0;
            d = st->count > 0 ? st->sum / st->count : 0.0;
            memcpy(&out->cells[col][row], &d, 8);
            break;
        case GBAGG_AGG_MIN:
        case GBAGG_AGG_MAX: {
            int has = q->aggs[a].kind == GBAGG_AGG_MIN ? st->has_min : st->has_max;
            d = q->aggs[a].kind == GBAGG_AGG_MIN ? st->sum_min : st->sum_max;
            out->valid[col][row] = (uint8_t)has;
            if (rt == GBAGG_TYPE_INTEGER) out->cells[col][row] = (int32_t)d;
            else if (rt == GBAGG_TYPE_BIGINT) out->cells[col][row] = (int64_t)d;
            else memcpy(&out->cells[col][row], &d, 8);
            break;
        }
        }
    }
}

typedef struct {
    int64_t sk, gk;
    int sh;
} heap_ent;

static int heap_less(const heap_ent *a, const heap_ent *b, int asc) {
    /* priority: the heap's front is the worst kept element */
    return asc ? a->sk < b->sk : a->sk > b->sk;
}

static void heap_sift_up(heap_ent *h, uint64_t i, int asc) {
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (!heap_less(&h[p], &h[i], asc)) break;
        heap_ent t = h[p]; h[p] = h[i]; h[i] = t;
        i = p;
    }
}

static void heap_sift_down(heap_ent *h, uint64_t n, int asc) {
    uint64_t i = 0;
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < n && heap_less(&h[m], &h[l], asc)) m = l;
        if (r < n && heap_less(&h[m], &h[r], asc)) m = r;
        if (m == i) break;
        heap_ent t = h[m]; h[m] = h[i]; h[i] = t;
        i = m;
    }
}

int orc_port_groupby(const gbagg_query_desc *q, const gbagg_column_view *cols, uint64_t nrows,
                     int threads, orc_result *out) {
    if (q->num_group_cols != 1) return 2;
    int kt = cols[q->group_cols[0]].type;
    if (kt != GBAGG_TYPE_INTEGER && kt != GBAGG_TYPE_BIGINT) return 2;
    port_ctx *c = xcalloc(1, sizeof(port_ctx));
    c->q = q;
    c->cols = cols;
    c->nrows = nrows;
    for (uint32_t a = 0; a < q->num_aggs; a++) {
        switch (q->aggs[a].kind) {
        case GBAGG_AGG_COUNT_STAR: c->kinds[a] = AK_COUNT_STAR; break;
        case GBAGG_AGG_COUNT: c->kinds[a] = AK_COUNT; break;
        case GBAGG_AGG_SUM:
        case GBAGG_AGG_AVG: c->kinds[a] = AK_SUM; break;
        case GBAGG_AGG_MIN: c->kinds[a] = AK_MIN; break;
        default: c->kinds[a] = AK_MAX; break;
        }
    }
    preds_prepare(q, cols, c->preds);
    int nt = threads < 1 ? 1 : threads;
    c->nt = nt;
    c->tls = xcalloc(nt, sizeof(pmap));
    for (int t = 0; t < nt; t++) pmap_init(&c->tls[t], q->num_aggs, 0);
    atomic_init(&c->next_rg, 0);
    pthread_t *th = xcalloc(nt > PORT_MERGE_SHARDS ? nt : PORT_MERGE_SHARDS, sizeof(pthread_t));
    void *args[64][2];
    for (int t = 0; t < nt; t++) {
        args[t][0] = c;
        args[t][1] = (void *)(intptr_t)t;
        pthread_create(&th[t], NULL, port_worker, args[t]);
    }
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);

    uint64_t est = 0;
    for (int t = 0; t < nt; t++) est += c->tls[t].n;
    for (int s = 0; s < PORT_MERGE_SHARDS; s++) {
        pmap_init(&c->shards[s], q->num_aggs, est / PORT_MERGE_SHARDS + 64);
        args[s][0] = c;
        args[s][1] = (void *)(intptr_t)s;
        pthread_create(&th[s], NULL, port_merge_worker, args[s]);
    }
    for (int s = 0; s < PORT_MERGE_SHARDS; s++) pthread_join(th[s], NULL);

    uint64_t total = 0;
    for (int s = 0; s < PORT_MERGE_SHARDS; s++) total += c->shards[s].n;

    /* TopN pushdown into the direct emit (:8067-8152): group col or COUNT. */
    int topn = 0, key_is_group = 0;
    int32_t topn_agg = -1;
    if (q->topn_col >= 0 && q->topn_limit > 0 && q->topn_limit < total) {
        if (q->topn_col == 0) { topn = 1; key_is_group = 1; }
        else if ((uint32_t)q->topn_col - 1 < q->num_aggs) {
            int k = q->aggs[q->topn_col - 1].kind;
            if (k == GBAGG_AGG_COUNT || k == GBAGG_AGG_COUNT_STAR) { topn = 1; topn_agg = q->topn_col - 1; }
        }
    }
    if (topn) {
        uint64_t K = q->topn_limit, hn = 0;
        int asc = q->topn_ascending;
        heap_ent *h = xcalloc(K + 1, sizeof(heap_ent));
        for (int s = 0; s < PORT_MERGE_SHARDS; s++) {
            pmap *m = &c->shards[s];
            for (uint64_t g = 0; g < m->n; g++) {
                int64_t sk = key_is_group ? m->keys[g] : m->st[g * m->na + topn_agg].count;
                heap_ent e = {sk, m->keys[g], s};
                if (hn < K) {
                    h[hn] = e;
                    heap_sift_up(h, hn++, asc);
                } else if (asc ? sk < h[0].sk : sk > h[0].sk) {
                    h[0] = e;
                    heap_sift_down(h, hn, asc);
                }
            }
        }
        result_alloc(out, q, cols, hn);
        for (uint64_t i = 0; i < hn; i++)
            port_emit(c, h[i].gk, pmap_get(&c->shards[h[i].sh], h[i].gk), out, i);
        free(h);
    } else {
        result_alloc(out, q, cols, total);
        uint64_t row = 0;
        for (int s = 0; s < PORT_MERGE_SHARDS; s++) {
            pmap *m = &c->shards[s];
            for (uint64_t g = 0; g < m->n; g++) port_emit(c, m->keys[g], &m->st[g * m->na], out, row++);
        }
    }
    for (int t = 0; t < nt; t++) pmap_free(&c->tls[t]);
    for (int s = 0; s < PORT_MERGE_SHARDS; s++) pmap_free(&c->shards[s]);
    preds_release(q, c->preds);
    free(c->tls);
    free(th);
    free(c);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Port of RadixCountAgg (radix_count_agg.cpp:17-106) with             */
/* SimpleI64CountMap (simple_i64_count_map.hpp:20-109).                */
/* ------------------------------------------------------------------ */
#define RC_RADIX 16

typedef struct {
    int64_t *key, *cnt;
    uint64_t cap, mask, n;
    int has_zero;
    int64_t zero_count;
} cmap;

static void cmap_reserve(cmap *m, uint64_t expected) {
    uint64_t target = 64;
    while ((expected + 1) * 2 > target) target <<= 1;
    m->cap = target;
    m->mask = target - 1;
    m->key = xcalloc(target, sizeof(int64_t));
    m->cnt = xcalloc(target, sizeof(int64_t));
}

static void cmap_grow(cmap *m) {
    uint64_t nc = m->cap * 2;
    int64_t *nk = xcalloc(nc, sizeof(int64_t)), *ncnt = xcalloc(nc, sizeof(int64_t));
    for (uint64_t i = 0; i < m->cap; i++) {
        if (m->key[i] == 0) continue;
        uint64_t p = orc_hash_u64((uint64_t)m->key[i]) & (nc - 1);
        while (nk[p] != 0) p = (p + 1) & (nc - 1);
        nk[p] = m->key[i];
        ncnt[p] = m->cnt[i];
    }
    free(m->key); free(m->cnt);
    m->key = nk; m->cnt = ncnt; m->cap = nc; m->mask = nc - 1;
}

static void cmap_increment(cmap *m, int64_t k) {
    if (k == 0) { m->zero_count++; m->has_zero = 1; return; }
    if ((m->n + 1) * 2 > m->cap) cmap_grow(m);
    uint64_t i = orc_hash_u64((uint64_t)k) & m->mask;
    for (;;) {
        if (m->key[i] == 0) { m->key[i] = k; m->cnt[i] = 1; m->n++; return; }
        if (m->key[i] == k) { m->cnt[i]++; return; }
        i = (i + 1) & m->mask;
    }
}

typedef struct {
    const int64_t *keys;
    uint64_t b, e;
    int64_t *buf[RC_RADIX];
    uint64_t n[RC_RADIX], cap[RC_RADIX];
} rc_thread;

static void *rc_scatter(void *arg) {
    rc_thread *t = arg;
    for (uint64_t r = t->b; r < t->e; r++) {
        int64_t k = t->keys[r];
        int radix = (int)(orc_hash_u64((uint64_t)k) >> 60);
        if (t->n[radix] == t->cap[radix]) {
            t->cap[radix] = t->cap[radix] ? t->cap[radix] * 2 : 1024;
            t->buf[radix] = xrealloc(t->buf[radix], t->cap[radix] * sizeof(int64_t));
        }
        t->buf[radix][t->n[radix]++] = k;
    }
    return NULL;
}

typedef struct {
    rc_thread *ts;
    int nt;
    cmap maps[RC_RADIX];
    atomic_int next;
} rc_merge_ctx;

static void *rc_merge(void *arg) {
    rc_merge_ctx *c = arg;
    for (;;) {
        int radix = atomic_fetch_add(&c->next, 1);
        if (radix >= RC_RADIX) break;
        uint64_t total = 0;
        for (int t = 0; t < c->nt; t++) total += c->ts[t].n[radix];
        cmap_reserve(&c->maps[radix], total / 4 + 64);
        for (int t = 0; t < c->nt; t++)
            for (uint64_t i = 0; i < c->ts[t].n[radix]; i++) cmap_increment(&c->maps[radix], c->ts[t].buf[radix][i]);
    }
    return NULL;
}

int orc_port_radix_count(const int64_t *keys, uint64_t nrows, int threads, int k,
                         int64_t **out_keys, int64_t **out_counts, uint64_t *out_n) {
    int nt = threads < 1 ? 1 : threads;
    rc_thread *ts = xcalloc(nt, sizeof(rc_thread));
    pthread_t *th = xcalloc(nt, sizeof(pthread_t));
    uint64_t per = (nrows + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        ts[t].keys = keys;
        ts[t].b = per * t < nrows ? per * t : nrows;
        ts[t].e = per * (t + 1) < nrows ? per * (t + 1) : nrows;
        pthread_create(&th[t], NULL, rc_scatter, &ts[t]);
    }
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    rc_merge_ctx mc;
    memset(&mc, 0, sizeof(mc));
    mc.ts = ts;
    mc.nt = nt;
    atomic_init(&mc.next, 0);
    int nm = nt < RC_RADIX ? nt : RC_RADIX;
    for (int t = 0; t < nm; t++) pthread_create(&th[t], NULL, rc_merge, &mc);
    for (int t = 0; t < nm; t++) pthread_join(th[t], NULL);

    uint64_t total = 0;
    for (int r = 0; r < RC_RADIX; r++) total += mc.maps[r].n + (mc.maps[r].has_zero ? 1 : 0);
    uint64_t cap = (k > 0 && (uint64_t)k < total) ? (uint64_t)k : total;
    int64_t *ok = xcalloc(cap, sizeof(int64_t)), *oc = xcalloc(cap, sizeof(int64_t));
    uint64_t n = 0;
    if (k <= 0) {
        for (int r = 0; r < RC_RADIX; r++) {
            cmap *m = &mc.maps[r];
            if (m->has_zero) { ok[n] = 0; oc[n++] = m->zero_count; }
            for (uint64_t i = 0; i < m->cap; i++)
                if (m->key[i] != 0) { ok[n] = m->key[i]; oc[n++] = m->cnt[i]; }
        }
    } else {
        /* EmitTopK: min-heap of size k; replace the min only on a strictly
         * larger count (radix_count_agg.cpp:76-100). */
        heap_ent *h = xcalloc(cap + 1, sizeof(heap_ent));
        uint64_t hn = 0;
        for (int r = 0; r < RC_RADIX; r++) {
            cmap *m = &mc.maps[r];
            for (uint64_t i = 0; i <= m->cap; i++) {
                int64_t key, cnt;
                if (i == 0) { if (!m->has_zero) continue; key = 0; cnt = m->zero_count; }
                else { if (m->key[i - 1] == 0) continue; key = m->key[i - 1]; cnt = m->cnt[i - 1]; }
                heap_ent e = {cnt, key, 0};
                if (hn < cap) { h[hn] = e; heap_sift_up(h, hn++, 0); }
                else if (cnt > h[0].sk) { h[0] = e; heap_sift_down(h, hn, 0); }
            }
        }
        /* pop in ascending order, then reverse: DESC by count */
        while (hn > 0) {
            ok[hn - 1] = h[0].gk;
            oc[hn - 1] = h[0].sk;
            h[0] = h[hn - 1];
            hn--;
            heap_sift_down(h, hn, 0);
            n++;
        }
        free(h);
    }
    for (int t = 0; t < nt; t++)
        for (int r = 0; r < RC_RADIX; r++) free(ts[t].buf[r]);
    for (int r = 0; r < RC_RADIX; r++) { free(mc.maps[r].key); free(mc.maps[r].cnt); }
    free(ts);
    free(th);
    *out_keys = ok;
    *out_counts = oc;
    *out_n = n;
    return 0;
}
