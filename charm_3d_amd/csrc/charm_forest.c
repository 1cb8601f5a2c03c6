/*
 * charm_forest.c -- host linear octree on the unit cube (replaces the part
 * of p4est this path uses: p4est_new_ext uniform fill, refine/coarsen,
 * 2:1 face balance, face neighbour queries) and the flattening of the leaves
 * into the element / face / chunk arrays the device kernels consume.
 *
 * Leaves are kept as a Morton-sorted array of anchors; a point query is a
 * search on the anchors' Morton keys (a leaf at level l owns the key
 * interval [key, key + 8^(19-l))), galloping outward from a nearby leaf when
 * one is known.  The per-leaf scans (balance, face listing) run on host
 * threads over contiguous leaf ranges; their results do not depend on the
 * thread count.
 */
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include "charm_internal.h"

/* ------------------------------------------------------------ errors */
static __thread char charm_err_buf[512];

void charm_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(charm_err_buf, sizeof(charm_err_buf), fmt, ap);
    va_end(ap);
}

const char *charm_host_last_error(void) { return charm_err_buf; }

/* ------------------------------------------------------------ Morton */
static uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8)  & 0x100f00f00f00f00fULL;
    x = (x | x << 4)  & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2)  & 0x1249249249249249ULL;
    return x;
}

static uint32_t compact3(uint64_t x)
{
    x &= 0x1249249249249249ULL;
    x = (x ^ (x >> 2))  & 0x10c30c30c30c30c3ULL;
    x = (x ^ (x >> 4))  & 0x100f00f00f00f00fULL;
    x = (x ^ (x >> 8))  & 0x1f0000ff0000ffULL;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
    x = (x ^ (x >> 32)) & 0x1fffff;
    return (uint32_t) x;
}

uint64_t charm_morton(int32_t x, int32_t y, int32_t z)
{
    return spread3((uint32_t) x) | (spread3((uint32_t) y) << 1) | (spread3((uint32_t) z) << 2);
}

static uint64_t span(int level) { return 1ULL << (3 * (CHARM_MAXLEVEL - level)); }

int64_t charm_forest_find(const charm_forest_t *f, int32_t x, int32_t y, int32_t z)
{
    uint64_t k = charm_morton(x, y, z);
    int64_t lo = 0, hi = f->n;            /* first index with key > k */
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (f->key[mid] <= k) lo = mid + 1; else hi = mid;
    }
    lo -= 1;
    if (lo < 0) return -1;
    if (k >= f->key[lo] + span(f->q[4*lo+3])) return -1;
    return lo;
}

/* leaf containing the point with Morton key k, searched outward from leaf
 * `hint` (exponential steps, then bisection): O(log distance) */
static int64_t find_near(const charm_forest_t *f, uint64_t k, int64_t hint)
{
    int64_t lo, hi, step = 1;
    if (f->key[hint] <= k) {             /* key[lo] <= k < key[hi] (hi = n: none) */
        lo = hint;
        for (;;) {
            hi = lo + step;
            if (hi >= f->n) { hi = f->n; break; }
            if (f->key[hi] > k) break;
            lo = hi;
            step <<= 1;
        }
    } else {
        hi = hint;
        for (;;) {
            lo = hi - step;
            if (lo < 0) { lo = -1; break; }
            if (f->key[lo] <= k) break;
            hi = lo;
            step <<= 1;
        }
    }
    while (hi - lo > 1) {
        int64_t mid = lo + (hi - lo) / 2;
        if (f->key[mid] <= k) lo = mid; else hi = mid;
    }
    if (lo < 0 || k >= f->key[lo] + span(f->q[4*lo+3])) return -1;
    return lo;
}

/* ------------------------------------------------------------ host threads */

/* CHARM_HOST_THREADS, else OMP_NUM_THREADS, else the online cores; [1, 64] */
static int host_threads(void)
{
    const char *s = getenv("CHARM_HOST_THREADS");
    long t = s ? atol(s) : 0;
    if (t <= 0) { s = getenv("OMP_NUM_THREADS"); t = s ? atol(s) : 0; }
    if (t <= 0) t = sysconf(_SC_NPROCESSORS_ONLN);
    return t < 1 ? 1 : (t > 64 ? 64 : (int) t);
}

typedef void (*range_fn)(void *arg, int64_t lo, int64_t hi, int part);
typedef struct { range_fn fn; void *arg; int64_t lo, hi; int part; } range_job_t;

static void *range_run(void *p)
{
    range_job_t *j = (range_job_t *) p;
    j->fn(j->arg, j->lo, j->hi, j->part);
    return NULL;
}

/* number of contiguous parts [0, n) is split into (>= min_per_part each) */
static int par_parts(int64_t n, int64_t min_per_part)
{
    int64_t t = host_threads();
    if (n / min_per_part < t) t = n / min_per_part;
    return t < 1 ? 1 : (int) t;
}

/* fn over the `parts` contiguous ranges of [0, n), part 0 on the caller */
static void par_for(int64_t n, int parts, range_fn fn, void *arg)
{
    range_job_t job[64];
    pthread_t th[64];
    int ok[64], p;
    for (p = 0; p < parts; p++) {
        job[p].fn = fn; job[p].arg = arg; job[p].part = p;
        job[p].lo = n * p / parts; job[p].hi = n * (p + 1) / parts;
    }
    for (p = 1; p < parts; p++) ok[p] = pthread_create(&th[p], NULL, range_run, &job[p]) == 0;
    range_run(&job[0]);
    for (p = 1; p < parts; p++) {
        if (ok[p]) pthread_join(th[p], NULL);
        else range_run(&job[p]);          /* no thread: run it here */
    }
}

/* ------------------------------------------------------------ lifecycle */
static int forest_alloc(charm_forest_t *f, int64_t n)
{
    f->q = (int32_t *) malloc((size_t) n * 4 * sizeof(int32_t));
    f->key = (uint64_t *) malloc((size_t) n * sizeof(uint64_t));
    if (!f->q || !f->key) {
        free(f->q); free(f->key); f->q = NULL; f->key = NULL;
        charm_set_error("forest: out of memory for %lld leaves", (long long) n);
        return CHARM_ERR_MEM;
    }
    f->n = n;
    return CHARM_OK;
}

int charm_forest_new_uniform(int level, const int periodic[3], charm_forest_t **out)
{
    charm_forest_t *f;
    int64_t n, i;
    int32_t len;
    if (!out || level < 0 || level > 10) {
        charm_set_error("forest_new_uniform: level %d out of range [0, 10]", level);
        return CHARM_ERR_ARG;
    }
    f = (charm_forest_t *) calloc(1, sizeof(*f));
    if (!f) { charm_set_error("forest: out of memory"); return CHARM_ERR_MEM; }
    n = 1LL << (3 * level);
    if (forest_alloc(f, n)) { free(f); return CHARM_ERR_MEM; }
    len = CHARM_ROOT_LEN >> level;
    for (i = 0; i < n; i++) {
        uint64_t m = (uint64_t) i;
        f->q[4*i+0] = (int32_t) compact3(m) * len;
        f->q[4*i+1] = (int32_t) compact3(m >> 1) * len;
        f->q[4*i+2] = (int32_t) compact3(m >> 2) * len;
        f->q[4*i+3] = level;
        f->key[i] = charm_morton(f->q[4*i], f->q[4*i+1], f->q[4*i+2]);
    }
    for (i = 0; i < 3; i++) f->periodic[i] = periodic ? (periodic[i] != 0) : 0;
    f->balanced = 1;
    *out = f;
    return CHARM_OK;
}

typedef struct { uint64_t k; int64_t i; } kidx_t;
static int cmp_kidx(const void *a, const void *b)
{
    uint64_t x = ((const kidx_t *) a)->k, y = ((const kidx_t *) b)->k;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static int forest_validate(const charm_forest_t *f)
{
    int64_t i;
    uint64_t expect = 0;
    for (i = 0; i < f->n; i++) {
        int l = f->q[4*i+3];
        int32_t len = CHARM_ROOT_LEN >> l;
        int a;
        if (l < 0 || l > CHARM_MAXLEVEL) { charm_set_error("forest: leaf %lld has level %d", (long long) i, l); return CHARM_ERR_MESH; }
        for (a = 0; a < 3; a++)
            if (f->q[4*i+a] < 0 || f->q[4*i+a] % len || f->q[4*i+a] + len > CHARM_ROOT_LEN) {
                charm_set_error("forest: leaf %lld is not aligned to its level", (long long) i);
                return CHARM_ERR_MESH;
            }
        if (f->key[i] != expect) {
            charm_set_error("forest: leaves do not tile the cube (gap or overlap at leaf %lld)", (long long) i);
            return CHARM_ERR_MESH;
        }
        expect += span(l);
    }
    if (expect != span(0)) { charm_set_error("forest: leaves do not cover the cube"); return CHARM_ERR_MESH; }
    return CHARM_OK;
}

int charm_forest_new_leaves(int64_t n, const int32_t *leaves, const int periodic[3],
                            charm_forest_t **out)
{
    charm_forest_t *f;
    kidx_t *ord;
    int64_t i;
    int rc;
    if (!out || n <= 0 || !leaves) { charm_set_error("forest_new_leaves: bad arguments"); return CHARM_ERR_ARG; }
    f = (charm_forest_t *) calloc(1, sizeof(*f));
    ord = (kidx_t *) malloc((size_t) n * sizeof(kidx_t));
    if (!f || !ord || forest_alloc(f, n)) { free(f); free(ord); charm_set_error("forest: out of memory"); return CHARM_ERR_MEM; }
    for (i = 0; i < n; i++) {
        ord[i].k = charm_morton(leaves[4*i], leaves[4*i+1], leaves[4*i+2]);
        ord[i].i = i;
    }
    qsort(ord, (size_t) n, sizeof(kidx_t), cmp_kidx);
    for (i = 0; i < n; i++) {
        memcpy(&f->q[4*i], &leaves[4*ord[i].i], 4 * sizeof(int32_t));
        f->key[i] = ord[i].k;
    }
    free(ord);
    for (i = 0; i < 3; i++) f->periodic[i] = periodic ? (periodic[i] != 0) : 0;
    rc = forest_validate(f);
    if (rc) { charm_forest_destroy(f); return rc; }
    *out = f;
    return CHARM_OK;
}

int charm_forest_destroy(charm_forest_t *f)
{
    if (!f) return CHARM_OK;
    free(f->q); free(f->key); free(f);
    return CHARM_OK;
}

int64_t charm_forest_num_leaves(const charm_forest_t *f) { return f ? f->n : 0; }

int charm_forest_get_leaves(const charm_forest_t *f, int32_t *out)
{
    if (!f || !out) { charm_set_error("forest_get_leaves: bad arguments"); return CHARM_ERR_ARG; }
    memcpy(out, f->q, (size_t) f->n * 4 * sizeof(int32_t));
    return CHARM_OK;
}

void *charm_forest_user_pointer(const charm_forest_t *f) { return f ? f->user : NULL; }
void charm_forest_set_user_pointer(charm_forest_t *f, void *p) { if (f) f->user = p; }

/* ------------------------------------------------------------ refinement */

/* rebuild with per-leaf actions: +1 refine into 8 children, -1 coarsen (the
 * caller guarantees whole families, first child carries the action), 0 keep,
 * -2 drop (other members of a coarsened family).  `fresh` (may be NULL)
 * receives for every new leaf the balance checks it needs: child_code for a
 * child of a refined leaf, FRESH_BOTH for a coarsened parent, 0 when kept.
 * Host threads: a counting pass and a filling pass over the same contiguous
 * ranges of old leaves. */
#define FRESH_FINE 1        /* check the leaf as the fine side of a violation */
#define FRESH_BOTH 2        /* ... and as the coarse side */
typedef struct {
    const charm_forest_t *f;
    const int8_t *act;
    int32_t *q;
    uint64_t *key;
    int64_t *map;
    int8_t *fresh;
    int8_t child_code;
    int64_t cnt[64], off[64];
} rebuild_arg_t;

static void rebuild_count(void *arg, int64_t lo, int64_t hi, int part)
{
    rebuild_arg_t *r = (rebuild_arg_t *) arg;
    int64_t i, c = 0;
    for (i = lo; i < hi; i++) c += r->act[i] == 1 ? 8 : (r->act[i] == -2 ? 0 : 1);
    r->cnt[part] = c;
}

static void rebuild_fill(void *arg, int64_t lo, int64_t hi, int part)
{
    rebuild_arg_t *r = (rebuild_arg_t *) arg;
    const charm_forest_t *f = r->f;
    int32_t *q = r->q;
    int64_t i, o = r->off[part];
    for (i = lo; i < hi; i++) {
        const int32_t *p = &f->q[4*i];
        const int8_t act = r->act[i];
        if (act == 1) {
            int32_t h = (CHARM_ROOT_LEN >> p[3]) / 2;
            int c;
            for (c = 0; c < 8; c++) {
                q[4*o+0] = p[0] + (c & 1) * h;
                q[4*o+1] = p[1] + ((c >> 1) & 1) * h;
                q[4*o+2] = p[2] + ((c >> 2) & 1) * h;
                q[4*o+3] = p[3] + 1;
                r->key[o] = charm_morton(q[4*o], q[4*o+1], q[4*o+2]);
                if (r->map) r->map[o] = i;
                if (r->fresh) r->fresh[o] = r->child_code;
                o++;
            }
        } else if (act == -1) {
            int32_t h = (CHARM_ROOT_LEN >> (p[3] - 1));
            q[4*o+0] = (p[0] / h) * h;
            q[4*o+1] = (p[1] / h) * h;
            q[4*o+2] = (p[2] / h) * h;
            q[4*o+3] = p[3] - 1;
            r->key[o] = charm_morton(q[4*o], q[4*o+1], q[4*o+2]);
            if (r->map) r->map[o] = -1 - i;
            if (r->fresh) r->fresh[o] = FRESH_BOTH;
            o++;
        } else if (act == 0) {
            memcpy(&q[4*o], p, 4 * sizeof(int32_t));
            r->key[o] = f->key[i];
            if (r->map) r->map[o] = i;
            if (r->fresh) r->fresh[o] = 0;
            o++;
        }
    }
}

static int forest_rebuild(charm_forest_t *f, const int8_t *act, int64_t **map_out, int64_t *n_map,
                          int8_t **fresh_out, int8_t child_code)
{
    rebuild_arg_t r;
    int64_t n_new = 0;
    int p, parts = par_parts(f->n, 16384);
    memset(&r, 0, sizeof(r));
    r.f = f; r.act = act; r.child_code = child_code;
    par_for(f->n, parts, rebuild_count, &r);
    for (p = 0; p < parts; p++) { r.off[p] = n_new; n_new += r.cnt[p]; }
    r.q = (int32_t *) malloc((size_t) (n_new ? n_new : 1) * 4 * sizeof(int32_t));
    r.key = (uint64_t *) malloc((size_t) (n_new ? n_new : 1) * sizeof(uint64_t));
    if (map_out) r.map = (int64_t *) malloc((size_t) (n_new ? n_new : 1) * sizeof(int64_t));
    if (fresh_out) r.fresh = (int8_t *) malloc((size_t) (n_new ? n_new : 1));
    if (!r.q || !r.key || (map_out && !r.map) || (fresh_out && !r.fresh)) {
        free(r.q); free(r.key); free(r.map); free(r.fresh);
        charm_set_error("forest: out of memory");
        return CHARM_ERR_MEM;
    }
    par_for(f->n, parts, rebuild_fill, &r);
    free(f->q); free(f->key);
    f->q = r.q; f->key = r.key; f->n = n_new;
    if (map_out) { *map_out = r.map; *n_map = n_new; }
    if (fresh_out) *fresh_out = r.fresh;
    return CHARM_OK;
}

/* 2:1 face balance by ripple refinement: a leaf at level l forces the leaf
 * across each of its faces to level >= l - 1.  A pass marks the coarse side
 * of every violation for refinement.  Without `fresh` it scans every leaf
 * from the fine side: the leaf containing the point just across each face
 * at the leaf's anchor corner (inside the coarser neighbour's face).  With
 * `fresh` it scans only the leaves a rebuild has just made (a violation of
 * a forest that was balanced, or of one whose previous pass refined every
 * coarse side, involves such a leaf): from the fine side as above, and for
 * FRESH_BOTH leaves also from the coarse side, marking the leaf itself when
 * a leaf across a face is two levels finer.  Such a leaf fills one of the
 * four quarters of the face's neighbour region, and the leaf containing that
 * quarter's centre point is then at least as fine; the four centres are
 * checked.  A child made by the adaptation's refinement (FRESH_FINE) cannot
 * be the coarse side: a leaf two levels finer than it was three levels finer
 * than its balanced parent, or is itself a new child, one level finer.
 * Coarsened parents and the children of a balance pass can (a parent's
 * neighbour may have been refined in the same adaptation).  Both forms mark
 * the same leaves. */
typedef struct { const charm_forest_t *f; const int8_t *fresh; int8_t *act; int8_t found[64]; } bal_arg_t;

/* the point (x) just across face (a, s) of the leaf at p with size len,
 * wrapped on a periodic axis; 0 when it leaves a non-periodic domain */
static int across(const charm_forest_t *f, const int32_t *p, int32_t len, int a, int s, int32_t x[3])
{
    x[a] = s ? p[a] + len : p[a] - 1;
    if (x[a] < 0 || x[a] >= CHARM_ROOT_LEN) {
        if (!f->periodic[a]) return 0;
        x[a] = (x[a] + CHARM_ROOT_LEN) % CHARM_ROOT_LEN;
    }
    return 1;
}

static void balance_scan(void *arg, int64_t lo, int64_t hi, int part)
{
    bal_arg_t *b = (bal_arg_t *) arg;
    const charm_forest_t *f = b->f;
    int64_t i;
    for (i = lo; i < hi; i++) {
        const int32_t *p = &f->q[4*i];
        const int32_t len = CHARM_ROOT_LEN >> p[3];
        int a, s;
        if (b->fresh && !b->fresh[i]) continue;
        for (a = 0; a < 3; a++)
            for (s = 0; s < 2; s++) {
                int32_t x[3] = {p[0], p[1], p[2]};
                int64_t nb;
                if (p[3] >= 2 && across(f, p, len, a, s, x)) {
                    nb = find_near(f, charm_morton(x[0], x[1], x[2]), i);
                    if (nb >= 0 && f->q[4*nb+3] < p[3] - 1) {
                        __atomic_store_n(&b->act[nb], (int8_t) 1, __ATOMIC_RELAXED);
                        b->found[part] = 1;
                    }
                }
                if (b->fresh && b->fresh[i] == FRESH_BOTH && len >= 4) {
                    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
                    int j;
                    for (j = 0; j < 4; j++) {
                        int32_t y[3] = {p[0], p[1], p[2]};
                        if (!across(f, p, len, a, s, y)) break;
                        y[t1] = p[t1] + ((j & 1) ? 3 * (len / 4) : len / 4);
                        y[t2] = p[t2] + ((j & 2) ? 3 * (len / 4) : len / 4);
                        nb = find_near(f, charm_morton(y[0], y[1], y[2]), i);
                        if (nb >= 0 && f->q[4*nb+3] > p[3] + 1) {
                            __atomic_store_n(&b->act[i], (int8_t) 1, __ATOMIC_RELAXED);
                            b->found[part] = 1;
                            break;
                        }
                    }
                }
            }
    }
}

typedef struct { int64_t *m2; const int64_t *m; } compose_arg_t;

static void compose_maps(void *arg, int64_t lo, int64_t hi, int part)
{
    compose_arg_t *c = (compose_arg_t *) arg;
    int64_t j;
    (void) part;
    for (j = lo; j < hi; j++) c->m2[j] = c->m[c->m2[j]];
}

/* fresh: the leaves the caller's rebuild made (taken over and freed), or
 * NULL for a full first pass */
static int forest_balance_map(charm_forest_t *f, int64_t **map, int8_t *fresh)
{
    for (;;) {
        bal_arg_t b;
        int p, parts = par_parts(f->n, 4096), any = 0;
        memset(&b, 0, sizeof(b));
        b.f = f;
        b.fresh = fresh;
        b.act = (int8_t *) calloc((size_t) (f->n ? f->n : 1), 1);
        if (!b.act) { free(fresh); charm_set_error("forest: out of memory"); return CHARM_ERR_MEM; }
        par_for(f->n, parts, balance_scan, &b);
        free(fresh);
        fresh = NULL;
        for (p = 0; p < parts; p++) any |= b.found[p];
        if (!any) { free(b.act); f->balanced = 1; return CHARM_OK; }
        {
            int64_t *m2 = NULL, nm2;
            int rc = forest_rebuild(f, b.act, map ? &m2 : NULL, &nm2, &fresh, FRESH_BOTH);
            free(b.act);
            if (rc) return rc;
            if (map && *map) {           /* compose maps: new -> old original */
                compose_arg_t c = {m2, *map};
                par_for(nm2, par_parts(nm2, 16384), compose_maps, &c);
                free(*map);
            }
            if (map) *map = m2;
        }
    }
}

int charm_forest_balance(charm_forest_t *f)
{
    if (!f) { charm_set_error("forest_balance: null forest"); return CHARM_ERR_ARG; }
    return forest_balance_map(f, NULL, NULL);
}

int charm_forest_adapt(charm_forest_t *f, const int8_t *flags, int max_level, int min_level,
                       int64_t **map, int64_t *n_map)
{
    int64_t i, nm = 0;
    int8_t *act, *fresh = NULL;
    int64_t *m = NULL;
    int rc;
    if (!f || !flags) { charm_set_error("forest_adapt: bad arguments"); return CHARM_ERR_ARG; }
    act = (int8_t *) calloc((size_t) f->n, 1);
    if (!act) { charm_set_error("forest: out of memory"); return CHARM_ERR_MEM; }
    for (i = 0; i < f->n; i++) {
        const int32_t *p = &f->q[4*i];
        if (act[i]) continue;
        if (flags[i] > 0 && p[3] < max_level) { act[i] = 1; continue; }
        if (flags[i] < 0 && p[3] > min_level && p[3] > 0) {
            /* a family is 8 consecutive leaves of the same level whose first
             * member is child 0 of the parent */
            int32_t ph = CHARM_ROOT_LEN >> (p[3] - 1);
            int c, ok = (i + 8 <= f->n) && p[0] % ph == 0 && p[1] % ph == 0 && p[2] % ph == 0;
            for (c = 0; ok && c < 8; c++)
                ok = f->q[4*(i+c)+3] == p[3] && flags[i+c] < 0 &&
                     f->key[i+c] == f->key[i] + (uint64_t) c * span(p[3]);
            if (ok) {
                act[i] = -1;
                for (c = 1; c < 8; c++) act[i+c] = -2;
                i += 7;
            }
        }
    }
    {
        const int was_balanced = f->balanced;
        f->balanced = 0;
        rc = forest_rebuild(f, act, map ? &m : NULL, &nm, &fresh, FRESH_FINE);
        free(act);
        if (rc) return rc;
        /* the fresh-leaf passes need a forest that was balanced before */
        if (!was_balanced) { free(fresh); fresh = NULL; }
    }
    rc = forest_balance_map(f, map ? &m : NULL, fresh);
    if (rc) { free(m); return rc; }
    if (map) { *map = m; *n_map = f->n; }
    return CHARM_OK;
}

/* ------------------------------------------------------------ flattening */

/* faces owned by leaf e (its lower domain faces and every face on its upper
 * side along each axis), at most 15, written to out[4 * k]; returns the
 * count or -1 (no leaf across a face) / -2 (not 2:1 balanced) */
static int leaf_faces(const charm_forest_t *f, int64_t e, int32_t *out)
{
    const int32_t *p = &f->q[4*e];
    int l = p[3], a, nf = 0;
    int32_t len = CHARM_ROOT_LEN >> l;
#define PUT(E0, E1, INFO) do { out[4*nf] = (E0); out[4*nf+1] = (E1); out[4*nf+2] = (INFO); out[4*nf+3] = 0; nf++; } while (0)
    for (a = 0; a < 3; a++) {
        int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
        int32_t x[3] = {p[0], p[1], p[2]};
        int64_t nb;
        const int32_t *r;
        if (p[a] == 0 && !f->periodic[a])
            PUT((int32_t) e, -1 - 2*a, CHARM_FACE_INFO(2*a, CHARM_FACE_BND, 0));
        x[a] += len;
        if (x[a] == CHARM_ROOT_LEN) {
            if (!f->periodic[a]) {
                PUT((int32_t) e, -1 - (2*a+1), CHARM_FACE_INFO(2*a+1, CHARM_FACE_BND, 0));
                continue;
            }
            x[a] = 0;
        }
        nb = find_near(f, charm_morton(x[0], x[1], x[2]), e);
        if (nb < 0) return -1;
        r = &f->q[4*nb];
        if (r[3] == l) {
            PUT((int32_t) e, (int32_t) nb, CHARM_FACE_INFO(2*a+1, CHARM_FACE_CONF, 0));
        } else if (r[3] == l - 1) {
            int sub = ((p[t1] - r[t1]) / len) | (((p[t2] - r[t2]) / len) << 1);
            PUT((int32_t) e, (int32_t) nb, CHARM_FACE_INFO(2*a+1, CHARM_FACE_HANG0, sub));
        } else if (r[3] == l + 1) {
            int j;
            for (j = 0; j < 4; j++) {
                int32_t y[3] = {x[0], x[1], x[2]};
                int64_t sm;
                y[t1] += (j & 1) * (len / 2);
                y[t2] += (j >> 1) * (len / 2);
                sm = find_near(f, charm_morton(y[0], y[1], y[2]), nb);
                if (sm < 0 || f->q[4*sm+3] != l + 1) return -2;
                PUT((int32_t) e, (int32_t) sm, CHARM_FACE_INFO(2*a+1, CHARM_FACE_HANG1, j));
            }
        } else {
            return -2;
        }
    }
#undef PUT
    return nf;
}

typedef struct {
    const charm_forest_t *f;
    int32_t *buf[64];
    int64_t cnt[64];
    int err[64];
} faces_arg_t;

static void faces_scan(void *arg, int64_t lo, int64_t hi, int part)
{
    faces_arg_t *a = (faces_arg_t *) arg;
    int64_t e, cap = (hi - lo) * 3 + 64, n = 0;
    int32_t *buf = (int32_t *) malloc((size_t) cap * 4 * sizeof(int32_t));
    a->buf[part] = NULL; a->cnt[part] = 0; a->err[part] = 0;
    if (!buf) { a->err[part] = -3; return; }
    for (e = lo; e < hi; e++) {
        int k;
        if (n + 15 > cap) {
            int32_t *t;
            cap *= 2;
            t = (int32_t *) realloc(buf, (size_t) cap * 4 * sizeof(int32_t));
            if (!t) { a->err[part] = -3; break; }
            buf = t;
        }
        k = leaf_faces(a->f, e, &buf[4*n]);
        if (k < 0) { a->err[part] = k; break; }
        n += k;
    }
    a->buf[part] = buf;
    a->cnt[part] = n;
}

/* every (sub)face once, in leaf order, side 0 = lower-coordinate leaf */
static int mesh_faces(const charm_forest_t *f, charm_mesh_t *m)
{
    faces_arg_t a;
    int parts = par_parts(f->n, 8192), p, err = 0;
    int64_t nf = 0, o = 0;
    memset(&a, 0, sizeof(a));
    a.f = f;
    par_for(f->n, parts, faces_scan, &a);
    for (p = 0; p < parts; p++) { nf += a.cnt[p]; if (a.err[p] && !err) err = a.err[p]; }
    if (!err) {
        m->f4 = (int32_t *) malloc((size_t) (nf ? nf : 1) * 4 * sizeof(int32_t));
        if (!m->f4) err = -3;
    }
    for (p = 0; p < parts; p++) {
        if (!err) memcpy(&m->f4[4*o], a.buf[p], (size_t) a.cnt[p] * 4 * sizeof(int32_t));
        o += a.cnt[p];
        free(a.buf[p]);
    }
    if (err == -1) { charm_set_error("flatten: no leaf across face"); return CHARM_ERR_MESH; }
    if (err == -2) { charm_set_error("flatten: forest is not 2:1 face balanced"); return CHARM_ERR_MESH; }
    if (err) { charm_set_error("flatten: out of memory"); return CHARM_ERR_MEM; }
    m->nf = nf;
    return CHARM_OK;
}

/* Neighbour table from the face list, on host threads: (1) every face
 * writes the slots it fills (each slot has one writer; the four faces of a
 * split face mark the big side's slot with the same sentinel); (2) the
 * split slots are numbered as hanging groups in (leaf, direction) order;
 * (3) every hanging face writes its small leaf into its group. */
#define NBR_HANG_MARK (INT32_MIN + 1)

typedef struct {
    charm_mesh_t *m;
    int64_t cnt[64], off[64];
    int bad[64];
} nbr_arg_t;

static void nbr_slots(void *arg, int64_t lo, int64_t hi, int part)
{
    charm_mesh_t *m = ((nbr_arg_t *) arg)->m;
    int64_t i;
    (void) part;
    for (i = lo; i < hi; i++) {
        const int32_t e0 = m->f4[4*i], e1 = m->f4[4*i+1], info = m->f4[4*i+2];
        const int d0 = info & 7, kind = (info >> 3) & 3;
        switch (kind) {
        case CHARM_FACE_BND:  m->nbr[6*e0 + d0] = e1; break;
        case CHARM_FACE_CONF: m->nbr[6*e0 + d0] = e1; m->nbr[6*e1 + (d0 ^ 1)] = e0; break;
        case CHARM_FACE_HANG0:
            m->nbr[6*e0 + d0] = e1;
            __atomic_store_n(&m->nbr[6*e1 + (d0 ^ 1)], NBR_HANG_MARK, __ATOMIC_RELAXED);
            break;
        case CHARM_FACE_HANG1:
            m->nbr[6*e1 + (d0 ^ 1)] = e0;
            __atomic_store_n(&m->nbr[6*e0 + d0], NBR_HANG_MARK, __ATOMIC_RELAXED);
            break;
        }
    }
}

static void nbr_count(void *arg, int64_t lo, int64_t hi, int part)
{
    nbr_arg_t *r = (nbr_arg_t *) arg;
    int64_t i, c = 0;
    for (i = 6 * lo; i < 6 * hi; i++) c += r->m->nbr[i] == NBR_HANG_MARK;
    r->cnt[part] = c;
}

static void nbr_number(void *arg, int64_t lo, int64_t hi, int part)
{
    nbr_arg_t *r = (nbr_arg_t *) arg;
    charm_mesh_t *m = r->m;
    int64_t i, g = r->off[part];
    for (i = 6 * lo; i < 6 * hi; i++)
        if (m->nbr[i] == NBR_HANG_MARK) {
            m->aux[4*g] = m->aux[4*g+1] = m->aux[4*g+2] = m->aux[4*g+3] = -1;
            m->nbr[i] = -CHARM_NBR_AUX_BASE - (int32_t) g++;
        }
}

static void nbr_groups(void *arg, int64_t lo, int64_t hi, int part)
{
    charm_mesh_t *m = ((nbr_arg_t *) arg)->m;
    int64_t i;
    (void) part;
    for (i = lo; i < hi; i++) {
        const int32_t e0 = m->f4[4*i], e1 = m->f4[4*i+1], info = m->f4[4*i+2];
        const int d0 = info & 7, kind = (info >> 3) & 3, sub = (info >> 5) & 3;
        int32_t slot;
        if (kind == CHARM_FACE_HANG0) slot = m->nbr[6*e1 + (d0 ^ 1)];
        else if (kind == CHARM_FACE_HANG1) slot = m->nbr[6*e0 + d0];
        else continue;
        m->aux[4 * (-CHARM_NBR_AUX_BASE - slot) + sub] = kind == CHARM_FACE_HANG0 ? e0 : e1;
    }
}

static void nbr_check(void *arg, int64_t lo, int64_t hi, int part)
{
    nbr_arg_t *r = (nbr_arg_t *) arg;
    int64_t i;
    for (i = 6 * lo; i < 6 * hi; i++)
        if (r->m->nbr[i] == INT32_MIN) { r->bad[part] = 1; return; }
}

static int mesh_neighbors(charm_mesh_t *m)
{
    nbr_arg_t r;
    int64_t i, naux = 0;
    int p, fparts = par_parts(m->nf, 16384), lparts = par_parts(m->n, 8192);
    memset(&r, 0, sizeof(r));
    r.m = m;
    m->nbr = (int32_t *) malloc((size_t) (m->n ? m->n : 1) * 6 * sizeof(int32_t));
    if (!m->nbr) { charm_set_error("flatten: out of memory"); return CHARM_ERR_MEM; }
    for (i = 0; i < m->n * 6; i++) m->nbr[i] = INT32_MIN;
    par_for(m->nf, fparts, nbr_slots, &r);
    par_for(m->n, lparts, nbr_count, &r);
    for (p = 0; p < lparts; p++) { r.off[p] = naux; naux += r.cnt[p]; }
    m->aux = (int32_t *) malloc((size_t) (naux ? naux : 1) * 4 * sizeof(int32_t));
    if (!m->aux) { charm_set_error("flatten: out of memory"); return CHARM_ERR_MEM; }
    par_for(m->n, lparts, nbr_number, &r);
    par_for(m->nf, fparts, nbr_groups, &r);
    par_for(m->n, lparts, nbr_check, &r);
    for (p = 0; p < lparts; p++)
        if (r.bad[p]) { charm_set_error("flatten: an element face is not covered"); return CHARM_ERR_MESH; }
    for (i = 0; i < naux * 4; i++)
        if (m->aux[i] < 0) { charm_set_error("flatten: hanging face group %lld incomplete", (long long) (i / 4)); return CHARM_ERR_MESH; }
    m->naux = naux;
    return CHARM_OK;
}

/* ------------------------------------------------------------ chunks
 *
 * Workgroup chunks: Morton-consecutive leaves, at most ch_e leaves and ch_f
 * distinct faces per chunk, grown greedily.  The leaves are cut into blocks
 * of CHUNK_BLOCK leaves that are chunked independently on host threads (a
 * chunk never spans two blocks), so the layout depends on the leaf count
 * and the chunk limits only, not on the thread count.  Each chunk's faces
 * are ordered by (axis, kind) so that a wave of the face phase mostly sees
 * one face orientation; the slots of a leaf are its face incidences in
 * incidence order, (local face << 1 | side). */
#define CHUNK_BLOCK 32768

typedef struct {
    const charm_mesh_t *m;
    const int32_t *inc_ptr, *inc;
    int ch_e, ch_f;
    int64_t n;
    uint16_t *slots;            /* global, at incidence positions */
    /* per block */
    int32_t **ch_ebeg, **ch_nf, **cf;
    int64_t *nch, *ncf;
    int *err;
} chunk_job_t;

typedef struct {
    int32_t *key, *val;
    int cap;
} face_map_t;

static int map_find(const face_map_t *h, int32_t k)
{
    unsigned i = ((unsigned) k * 2654435761u) & (unsigned) (h->cap - 1);
    while (h->key[i] != -1) {
        if (h->key[i] == k) return h->val[i];
        i = (i + 1) & (unsigned) (h->cap - 1);
    }
    return -1;
}

static void map_put(face_map_t *h, int32_t k, int32_t v)
{
    unsigned i = ((unsigned) k * 2654435761u) & (unsigned) (h->cap - 1);
    while (h->key[i] != -1) i = (i + 1) & (unsigned) (h->cap - 1);
    h->key[i] = k;
    h->val[i] = v;
}

static void map_clear(face_map_t *h, const int32_t *keys, int64_t nk)
{
    int64_t j;
    for (j = 0; j < nk; j++) {
        unsigned i = ((unsigned) keys[j] * 2654435761u) & (unsigned) (h->cap - 1);
        while (h->key[i] != keys[j]) i = (i + 1) & (unsigned) (h->cap - 1);
        h->key[i] = -1;
    }
    /* re-insert displaced entries is unnecessary: the whole chunk is
     * cleared, so the table is empty again */
}

/* the side of leaf e in face fi at incidence k: a leaf that is its own
 * periodic neighbour holds both sides, its two incidences adjacent, side 0
 * first */
static int slot_side(const charm_mesh_t *m, const int32_t *inc_ptr, const int32_t *inc, int64_t e,
                     int64_t k, int32_t fi)
{
    if (m->f4[4*fi] == e && m->f4[4*fi+1] == e) return (k > inc_ptr[e] && inc[k-1] == fi) ? 1 : 0;
    return m->f4[4*fi] == e ? 0 : 1;
}

/* close the chunk [eb, ee) whose faces are cf[0..nf): sort them by (axis,
 * kind), then write its leaves' slots with the sorted local indices */
static void close_chunk(const chunk_job_t *J, face_map_t *h, int32_t *cf, int64_t nf, int64_t eb, int64_t ee,
                        int32_t *tmp, int32_t *rank)
{
    const charm_mesh_t *m = J->m;
    int64_t cnt[13] = {0}, i, e, k;
    for (i = 0; i < nf; i++) {
        const int info = m->f4[4 * cf[i] + 2];
        cnt[1 + ((info & 7) >> 1) * 4 + ((info >> 3) & 3)]++;
    }
    for (i = 1; i < 13; i++) cnt[i] += cnt[i - 1];
    for (i = 0; i < nf; i++) {
        const int info = m->f4[4 * cf[i] + 2];
        const int64_t to = cnt[((info & 7) >> 1) * 4 + ((info >> 3) & 3)]++;
        tmp[to] = cf[i];
        rank[i] = (int32_t) to;             /* provisional local index i -> sorted */
    }
    for (e = eb; e < ee; e++)
        for (k = J->inc_ptr[e]; k < J->inc_ptr[e + 1]; k++) {
            const int32_t fi = J->inc[k];
            J->slots[k] = (uint16_t) ((rank[map_find(h, fi)] << 1) | slot_side(m, J->inc_ptr, J->inc, e, k, fi));
        }
    map_clear(h, cf, nf);
    memcpy(cf, tmp, (size_t) nf * sizeof(int32_t));
}

static void chunk_blocks(void *arg, int64_t lo, int64_t hi, int part)
{
    chunk_job_t *J = (chunk_job_t *) arg;
    const int32_t *inc_ptr = J->inc_ptr, *inc = J->inc;
    face_map_t h;
    int32_t *tmp = NULL, *rank = NULL;
    int64_t b;
    (void) part;
    h.cap = 1;
    while (h.cap < 4 * (J->ch_f + 1)) h.cap <<= 1;
    h.key = (int32_t *) malloc((size_t) h.cap * sizeof(int32_t));
    h.val = (int32_t *) malloc((size_t) h.cap * sizeof(int32_t));
    tmp = (int32_t *) malloc((size_t) (J->ch_f + 1) * sizeof(int32_t));
    rank = (int32_t *) malloc((size_t) (J->ch_f + 1) * sizeof(int32_t));
    if (!h.key || !h.val || !tmp || !rank) {
        for (b = lo; b < hi; b++) J->err[b] = CHARM_ERR_MEM;
        goto done;
    }
    memset(h.key, 0xff, (size_t) h.cap * sizeof(int32_t));
    for (b = lo; b < hi; b++) {
        const int64_t e0 = b * CHUNK_BLOCK, e1 = (e0 + CHUNK_BLOCK < J->n) ? e0 + CHUNK_BLOCK : J->n;
        const int64_t nfb = inc_ptr[e1] - inc_ptr[e0];     /* incidences: bound on the face copies */
        int32_t *cf = (int32_t *) malloc((size_t) (nfb + 1) * sizeof(int32_t));
        int32_t *ceb = (int32_t *) malloc((size_t) (e1 - e0 + 1) * sizeof(int32_t));
        int32_t *cnf = (int32_t *) malloc((size_t) (e1 - e0 + 1) * sizeof(int32_t));
        int64_t e, k, eb = e0, fb = 0, cur = 0, nch = 0;
        J->cf[b] = cf; J->ch_ebeg[b] = ceb; J->ch_nf[b] = cnf;
        if (!cf || !ceb || !cnf) { J->err[b] = CHARM_ERR_MEM; continue; }
        for (e = e0; e < e1; e++) {
            int64_t nnew = 0;
            for (k = inc_ptr[e]; k < inc_ptr[e + 1]; k++)
                if (map_find(&h, inc[k]) < 0 && !(k > inc_ptr[e] && inc[k - 1] == inc[k])) nnew++;
            if (e > eb && (e - eb >= J->ch_e || cur + nnew > J->ch_f)) {
                close_chunk(J, &h, cf + fb, cur, eb, e, tmp, rank);
                ceb[nch] = (int32_t) eb; cnf[nch] = (int32_t) cur; nch++;
                fb += cur; eb = e; cur = 0;
                nnew = 0;
                for (k = inc_ptr[e]; k < inc_ptr[e + 1]; k++)
                    if (!(k > inc_ptr[e] && inc[k - 1] == inc[k])) nnew++;
            }
            if (nnew > J->ch_f) { J->err[b] = CHARM_ERR_ARG; break; }
            for (k = inc_ptr[e]; k < inc_ptr[e + 1]; k++)
                if (map_find(&h, inc[k]) < 0) {
                    map_put(&h, inc[k], (int32_t) cur);
                    cf[fb + cur] = inc[k];
                    cur++;
                }
        }
        if (J->err[b]) { map_clear(&h, cf + fb, cur); continue; }
        close_chunk(J, &h, cf + fb, cur, eb, e1, tmp, rank);
        ceb[nch] = (int32_t) eb; cnf[nch] = (int32_t) cur; nch++;
        J->nch[b] = nch;
        J->ncf[b] = fb + cur;
    }
done:
    free(h.key); free(h.val); free(tmp); free(rank);
}

typedef struct {
    charm_mesh_t *m;
    const chunk_job_t *J;
    const int64_t *ch_off, *cf_off;
} chunk_merge_t;

static void chunk_merge(void *arg, int64_t lo, int64_t hi, int part)
{
    chunk_merge_t *M = (chunk_merge_t *) arg;
    charm_mesh_t *m = M->m;
    int64_t b, c, i;
    (void) part;
    for (b = lo; b < hi; b++) {
        int64_t f = M->cf_off[b];
        for (c = 0; c < M->J->nch[b]; c++) {
            m->ch_ebeg[M->ch_off[b] + c] = M->J->ch_ebeg[b][c];
            m->ch_fbeg[M->ch_off[b] + c] = (int32_t) f;
            f += M->J->ch_nf[b][c];
        }
        for (i = 0; i < M->J->ncf[b]; i++)
            memcpy(&m->cf4[4 * (M->cf_off[b] + i)], &m->f4[4 * M->J->cf[b][i]], 4 * sizeof(int32_t));
    }
}

typedef struct { const charm_mesh_t *m; int32_t *inc_ptr, *inc, *fill; } inc_arg_t;

static void inc_count(void *arg, int64_t lo, int64_t hi, int part)
{
    inc_arg_t *I = (inc_arg_t *) arg;
    int64_t i;
    (void) part;
    for (i = lo; i < hi; i++) {
        __atomic_fetch_add(&I->inc_ptr[I->m->f4[4*i] + 1], 1, __ATOMIC_RELAXED);
        if (I->m->f4[4*i+1] >= 0) __atomic_fetch_add(&I->inc_ptr[I->m->f4[4*i+1] + 1], 1, __ATOMIC_RELAXED);
    }
}

static void inc_fill(void *arg, int64_t lo, int64_t hi, int part)
{
    inc_arg_t *I = (inc_arg_t *) arg;
    int64_t i;
    (void) part;
    for (i = lo; i < hi; i++) {
        I->inc[__atomic_fetch_add(&I->fill[I->m->f4[4*i]], 1, __ATOMIC_RELAXED)] = (int32_t) i;
        if (I->m->f4[4*i+1] >= 0)
            I->inc[__atomic_fetch_add(&I->fill[I->m->f4[4*i+1]], 1, __ATOMIC_RELAXED)] = (int32_t) i;
    }
}

/* insertion sort of each leaf's incident faces (at most a few dozen) */
static void inc_sort(void *arg, int64_t lo, int64_t hi, int part)
{
    inc_arg_t *I = (inc_arg_t *) arg;
    int64_t e;
    (void) part;
    for (e = lo; e < hi; e++) {
        int32_t *a = I->inc + I->inc_ptr[e];
        const int32_t k = I->inc_ptr[e+1] - I->inc_ptr[e];
        for (int32_t j = 1; j < k; j++) {
            const int32_t v = a[j];
            int32_t t = j - 1;
            while (t >= 0 && a[t] > v) { a[t+1] = a[t]; t--; }
            a[t+1] = v;
        }
    }
}

static int mesh_chunks(charm_mesh_t *m, int ch_e, int ch_f)
{
    const int64_t n = m->n, nf = m->nf, nblk = (n + CHUNK_BLOCK - 1) / CHUNK_BLOCK;
    int32_t *inc_ptr = NULL, *inc = NULL, *fill = NULL;
    int64_t e, b, *ch_off = NULL, *cf_off = NULL;
    chunk_job_t J;
    int rc = CHARM_ERR_MEM;
    memset(&J, 0, sizeof(J));

    /* leaf -> incident faces, in face order (a face joining a leaf to
     * itself appears twice, adjacent): counted and filled by host threads
     * over the faces, then each leaf's short list sorted */
    inc_ptr = (int32_t *) calloc((size_t) n + 1, sizeof(int32_t));
    if (!inc_ptr) goto done;
    {
        inc_arg_t I = {m, inc_ptr, NULL, NULL};
        par_for(nf, par_parts(nf, 16384), inc_count, &I);
        for (e = 0; e < n; e++) inc_ptr[e+1] += inc_ptr[e];
        inc = (int32_t *) malloc((size_t) (inc_ptr[n] + 1) * sizeof(int32_t));
        fill = (int32_t *) malloc((size_t) (n + 1) * sizeof(int32_t));
        if (!inc || !fill) goto done;
        memcpy(fill, inc_ptr, (size_t) n * sizeof(int32_t));
        I.inc = inc; I.fill = fill;
        par_for(nf, par_parts(nf, 16384), inc_fill, &I);
        par_for(n, par_parts(n, 8192), inc_sort, &I);
    }

    m->slot_ptr = (int32_t *) malloc((size_t) (n + 1) * sizeof(int32_t));
    m->slots = (uint16_t *) malloc((size_t) (inc_ptr[n] + 1) * sizeof(uint16_t));
    J.ch_ebeg = (int32_t **) calloc((size_t) nblk + 1, sizeof(int32_t *));
    J.ch_nf = (int32_t **) calloc((size_t) nblk + 1, sizeof(int32_t *));
    J.cf = (int32_t **) calloc((size_t) nblk + 1, sizeof(int32_t *));
    J.nch = (int64_t *) calloc((size_t) nblk + 1, sizeof(int64_t));
    J.ncf = (int64_t *) calloc((size_t) nblk + 1, sizeof(int64_t));
    J.err = (int *) calloc((size_t) nblk + 1, sizeof(int));
    ch_off = (int64_t *) malloc((size_t) (nblk + 1) * sizeof(int64_t));
    cf_off = (int64_t *) malloc((size_t) (nblk + 1) * sizeof(int64_t));
    if (!m->slot_ptr || !m->slots || !J.ch_ebeg || !J.ch_nf || !J.cf || !J.nch || !J.ncf || !J.err ||
        !ch_off || !cf_off)
        goto done;
    memcpy(m->slot_ptr, inc_ptr, (size_t) (n + 1) * sizeof(int32_t));
    J.m = m; J.inc_ptr = inc_ptr; J.inc = inc; J.ch_e = ch_e; J.ch_f = ch_f; J.n = n; J.slots = m->slots;
    par_for(nblk, par_parts(nblk, 1), chunk_blocks, &J);
    for (b = 0; b < nblk; b++)
        if (J.err[b]) {
            rc = J.err[b];
            if (rc == CHARM_ERR_ARG) charm_set_error("flatten: chunk_faces %d below one element's faces", ch_f);
            goto done;
        }
    ch_off[0] = cf_off[0] = 0;
    for (b = 0; b < nblk; b++) { ch_off[b + 1] = ch_off[b] + J.nch[b]; cf_off[b + 1] = cf_off[b] + J.ncf[b]; }
    m->nch = ch_off[nblk];
    m->ncf = cf_off[nblk];
    m->ch_ebeg = (int32_t *) malloc((size_t) (m->nch + 1) * sizeof(int32_t));
    m->ch_fbeg = (int32_t *) malloc((size_t) (m->nch + 1) * sizeof(int32_t));
    m->cf4 = (int32_t *) malloc((size_t) (m->ncf + 1) * 4 * sizeof(int32_t));
    if (!m->ch_ebeg || !m->ch_fbeg || !m->cf4) goto done;
    {
        chunk_merge_t M;
        M.m = m; M.J = &J; M.ch_off = ch_off; M.cf_off = cf_off;
        par_for(nblk, par_parts(nblk, 1), chunk_merge, &M);
    }
    m->ch_ebeg[m->nch] = (int32_t) n;
    m->ch_fbeg[m->nch] = (int32_t) m->ncf;
    m->nslot = inc_ptr[n];
    m->chunk_elems = ch_e;
    m->chunk_faces = ch_f;
    rc = CHARM_OK;
done:
    if (rc == CHARM_ERR_MEM) charm_set_error("flatten: out of memory");
    for (b = 0; b < nblk && J.cf; b++) { free(J.cf[b]); free(J.ch_ebeg[b]); free(J.ch_nf[b]); }
    free(J.ch_ebeg); free(J.ch_nf); free(J.cf); free(J.nch); free(J.ncf); free(J.err);
    free(ch_off); free(cf_off); free(inc_ptr); free(inc); free(fill);
    return rc;
}

int charm_forest_flatten(const charm_forest_t *f, int chunk_elems, int chunk_faces,
                         charm_mesh_t **out)
{
    charm_mesh_t *m;
    int64_t i;
    int rc;
    if (!f || !out) { charm_set_error("flatten: bad arguments"); return CHARM_ERR_ARG; }
    if (chunk_elems <= 0) chunk_elems = 128;
    if (chunk_faces <= 0) chunk_faces = 512;
    if (chunk_faces > 32767 || chunk_elems > 1024) { charm_set_error("flatten: chunk sizes too large"); return CHARM_ERR_ARG; }
    m = (charm_mesh_t *) calloc(1, sizeof(*m));
    if (!m) { charm_set_error("flatten: out of memory"); return CHARM_ERR_MEM; }
    m->n = f->n;
    for (i = 0; i < 3; i++) m->periodic[i] = f->periodic[i];
    m->level = (uint8_t *) malloc((size_t) f->n);
    m->q = (int32_t *) malloc((size_t) f->n * 4 * sizeof(int32_t));
    if (!m->level || !m->q) { charm_mesh_destroy(m); charm_set_error("flatten: out of memory"); return CHARM_ERR_MEM; }
    memcpy(m->q, f->q, (size_t) f->n * 4 * sizeof(int32_t));
    for (i = 0; i < f->n; i++) m->level[i] = (uint8_t) f->q[4*i+3];
    if ((rc = mesh_faces(f, m)) || (rc = mesh_neighbors(m)) || (rc = mesh_chunks(m, chunk_elems, chunk_faces))) {
        charm_mesh_destroy(m);
        return rc;
    }
    *out = m;
    return CHARM_OK;
}

int charm_mesh_destroy(charm_mesh_t *m)
{
    if (!m) return CHARM_OK;
    free(m->level); free(m->q); free(m->f4); free(m->nbr); free(m->aux);
    free(m->ch_ebeg); free(m->ch_fbeg); free(m->cf4); free(m->slot_ptr); free(m->slots);
    free(m);
    return CHARM_OK;
}

int64_t charm_mesh_num_faces(const charm_mesh_t *m) { return m ? m->nf : 0; }
int64_t charm_mesh_num_chunks(const charm_mesh_t *m) { return m ? m->nch : 0; }

int64_t charm_mesh_num_chunk_faces(const charm_mesh_t *m) { return m ? m->ncf : 0; }

int charm_mesh_get_chunks(const charm_mesh_t *m, int32_t *ch_ebeg, int32_t *ch_fbeg, int32_t *cf4,
                          int32_t *slot_ptr, uint16_t *slots)
{
    if (!m || !ch_ebeg || !ch_fbeg || !cf4 || !slot_ptr || !slots) {
        charm_set_error("mesh_get_chunks: bad arguments");
        return CHARM_ERR_ARG;
    }
    memcpy(ch_ebeg, m->ch_ebeg, (size_t) (m->nch + 1) * sizeof(int32_t));
    memcpy(ch_fbeg, m->ch_fbeg, (size_t) (m->nch + 1) * sizeof(int32_t));
    memcpy(cf4, m->cf4, (size_t) m->ncf * 4 * sizeof(int32_t));
    memcpy(slot_ptr, m->slot_ptr, (size_t) (m->n + 1) * sizeof(int32_t));
    memcpy(slots, m->slots, (size_t) m->nslot * sizeof(uint16_t));
    return CHARM_OK;
}

int charm_mesh_get_faces(const charm_mesh_t *m, int32_t *out)
{
    if (!m || !out) { charm_set_error("mesh_get_faces: bad arguments"); return CHARM_ERR_ARG; }
    memcpy(out, m->f4, (size_t) m->nf * 4 * sizeof(int32_t));
    return CHARM_OK;
}

int charm_mesh_get_neighbors(const charm_mesh_t *m, int32_t *out)
{
    if (!m || !out) { charm_set_error("mesh_get_neighbors: bad arguments"); return CHARM_ERR_ARG; }
    memcpy(out, m->nbr, (size_t) m->n * 6 * sizeof(int32_t));
    return CHARM_OK;
}
