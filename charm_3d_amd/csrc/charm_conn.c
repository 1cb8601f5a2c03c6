/* charm_conn.c -- Gmsh 2.x ASCII hexahedral meshes as multi-tree connectivity.
 *
 * What charm_conn_reader_msh (src/charm_connectivity.c:315-519) builds for
 * p4est: tree vertices, the 8 corners of every hexahedron in z-order, the
 * tree-to-tree face links with the face orientation of p4est's connectivity
 * (p4est a9910fd, p8est_connectivity.h: "the first face corner of the lower
 * of the two face numbers connects to face corner r of the higher one",
 * tree_to_face = neighbour face + 6 r), the physical patch of every boundary
 * quadrangle and the physical region of every hexahedron.  Faces are matched
 * by their sorted corner ids in a hash map, as the reference matches boundary
 * quadrangles (charm_fhash_*, src/charm_connectivity.c:17-178) and as
 * p4est_connectivity_complete matches shared faces.
 *
 * Binding the patches and regions to a task's boundaries and regions by name
 * follows charm_conn_bnd_find_by_id / charm_conn_reg_find_by_id
 * (src/charm_connectivity.c:178-230).
 */
#include <ctype.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "charm_gpu.h"
#include "charm_internal.h"

/* corners of each face in z-order corner numbering (src/charm_connectivity.c:9-15) */
static const int conn_face_corners[6][4] = {
    {0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};

struct charm_conn {
    int64_t nv, nt;
    double *vert;              /* [nv][3] */
    int32_t *ttv;              /* [nt][8] */
    int32_t *ttt;              /* [nt][6] */
    int8_t *ttf;               /* [nt][6] */
    int32_t *patch;            /* [nt][6] physical id of the boundary patch, -1 none */
    int32_t *region;           /* [nt] physical id of the hexahedron, -1 none */
    int np;
    charm_conn_patch_t *pn;    /* $PhysicalNames */
};

/* --------------------------------------------------------- face hash map */
typedef struct { int32_t k[4]; int64_t val; } fh_slot_t;
typedef struct { fh_slot_t *s; size_t cap; } fh_t;

static void sort4(const int32_t in[4], int32_t o[4])
{
    memcpy(o, in, 4 * sizeof(int32_t));
    for (int i = 1; i < 4; i++)
        for (int j = i; j > 0 && o[j - 1] > o[j]; j--) { int32_t t = o[j]; o[j] = o[j - 1]; o[j - 1] = t; }
}

static size_t fh_hash(const int32_t k[4], size_t mask)
{
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 4; i++) { h ^= (uint64_t) (uint32_t) k[i]; h *= 1099511628211ull; }
    return (size_t) (h ^ (h >> 29)) & mask;
}

static int fh_init(fh_t *m, size_t n)
{
    size_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    m->s = (fh_slot_t *) malloc(cap * sizeof(fh_slot_t));
    if (!m->s) return 0;
    for (size_t i = 0; i < cap; i++) m->s[i].val = -1;
    m->cap = cap;
    return 1;
}

/* the slot of key k (sorted corners): existing, or empty (val == -1) */
static fh_slot_t *fh_find(fh_t *m, const int32_t k[4])
{
    size_t i = fh_hash(k, m->cap - 1);
    for (;;) {
        fh_slot_t *s = &m->s[i];
        if (s->val < 0 || memcmp(s->k, k, sizeof(s->k)) == 0) return s;
        i = (i + 1) & (m->cap - 1);
    }
}

/* ----------------------------------------------------------------- lines */
typedef struct { FILE *f; char *buf; size_t cap; int64_t line; } rd_t;

static const char *rd_line(rd_t *r)
{
    ssize_t n = getline(&r->buf, &r->cap, r->f);
    if (n < 0) return NULL;
    r->line++;
    while (n > 0 && (r->buf[n - 1] == '\n' || r->buf[n - 1] == '\r')) r->buf[--n] = 0;
    return r->buf;
}

static int starts(const char *s, const char *key)
{
    while (isspace((unsigned char) *s)) s++;
    return strncmp(s, key, strlen(key)) == 0;
}

void charm_conn_destroy(charm_conn_t *c)
{
    if (!c) return;
    free(c->vert); free(c->ttv); free(c->ttt); free(c->ttf); free(c->patch); free(c->region); free(c->pn);
    free(c);
}

/* Gmsh hexahedron corner order 0..7 -> z-order (the reference reads nodes
 * 0 1 3 2 4 5 7 6 into corners 0..7, src/charm_connectivity.c:413-416) */
static const int gmsh_hex_to_z[8] = {0, 1, 3, 2, 4, 5, 7, 6};

int charm_conn_read_msh(const char *path, charm_conn_t **out)
{
    if (!path || !out) { charm_set_error("charm_conn_read_msh: null argument"); return CHARM_ERR_ARG; }
    *out = NULL;
    rd_t r = {fopen(path, "rb"), NULL, 0, 0};
    if (!r.f) { charm_set_error("Failed to open %s", path); return CHARM_ERR_ARG; }
    charm_conn_t *c = (charm_conn_t *) calloc(1, sizeof(*c));
    int64_t *node_id = NULL, n_quad = 0, max_id = 0;
    int32_t *quads = NULL, *quad_tag = NULL, *idmap = NULL;
    int rc = CHARM_OK, have_nodes = 0, have_elems = 0;
    const char *s;
#define FAIL(code, ...) do { charm_set_error(__VA_ARGS__); rc = (code); goto done; } while (0)
    if (!c) FAIL(CHARM_ERR_MEM, "charm_conn_read_msh: out of memory");
    while ((s = rd_line(&r))) {
        if (starts(s, "$MeshFormat")) {
            double ver = 0.0;
            int ftype = -1, dsize = 0;
            if (!(s = rd_line(&r)) || sscanf(s, "%lf %d %d", &ver, &ftype, &dsize) != 3)
                FAIL(CHARM_ERR_MESH, "%s:%lld: bad $MeshFormat", path, (long long) r.line);
            if (ver < 2.0 || ver >= 3.0)
                FAIL(CHARM_ERR_UNSUPPORTED, "%s: Gmsh format %g (only 2.x is read)", path, ver);
            if (ftype != 0) FAIL(CHARM_ERR_UNSUPPORTED, "%s: binary Gmsh files are not read", path);
        } else if (starts(s, "$PhysicalNames")) {
            int n = 0;
            if (!(s = rd_line(&r)) || sscanf(s, "%d", &n) != 1 || n < 0)
                FAIL(CHARM_ERR_MESH, "%s:%lld: bad $PhysicalNames", path, (long long) r.line);
            c->pn = (charm_conn_patch_t *) calloc((size_t) (n ? n : 1), sizeof(charm_conn_patch_t));
            if (!c->pn) FAIL(CHARM_ERR_MEM, "charm_conn_read_msh: out of memory");
            for (int i = 0; i < n; i++) {
                charm_conn_patch_t *p = &c->pn[i];
                if (!(s = rd_line(&r)) || sscanf(s, "%d %d \"%63[^\"]", &p->dim, &p->id, p->name) != 3)
                    FAIL(CHARM_ERR_MESH, "%s:%lld: bad physical name", path, (long long) r.line);
            }
            c->np = n;
        } else if (starts(s, "$Nodes")) {
            long long n = 0;
            if (!(s = rd_line(&r)) || sscanf(s, "%lld", &n) != 1 || n <= 0 || n > (1ll << 31))
                FAIL(CHARM_ERR_MESH, "%s:%lld: bad $Nodes count", path, (long long) r.line);
            c->nv = n;
            c->vert = (double *) malloc((size_t) n * 3 * sizeof(double));
            node_id = (int64_t *) malloc((size_t) n * sizeof(int64_t));
            if (!c->vert || !node_id) FAIL(CHARM_ERR_MEM, "charm_conn_read_msh: out of memory");
            for (long long i = 0; i < n; i++) {
                long long id;
                if (!(s = rd_line(&r)) ||
                    sscanf(s, "%lld %lf %lf %lf", &id, &c->vert[3 * i], &c->vert[3 * i + 1], &c->vert[3 * i + 2]) != 4)
                    FAIL(CHARM_ERR_MESH, "Premature end of file (%s:%lld)", path, (long long) r.line);
                if (id <= 0) FAIL(CHARM_ERR_MESH, "%s:%lld: node id %lld", path, (long long) r.line, id);
                node_id[i] = id;
                if (id > max_id) max_id = id;
            }
            if (max_id > 16 * n + 1024) FAIL(CHARM_ERR_MESH, "%s: node ids up to %lld for %lld nodes", path, (long long) max_id, n);
            idmap = (int32_t *) malloc((size_t) (max_id + 1) * sizeof(int32_t));
            if (!idmap) FAIL(CHARM_ERR_MEM, "charm_conn_read_msh: out of memory");
            for (int64_t i = 0; i <= max_id; i++) idmap[i] = -1;
            for (int64_t i = 0; i < n; i++) {
                if (idmap[node_id[i]] >= 0) FAIL(CHARM_ERR_MESH, "%s: node id %lld repeated", path, (long long) node_id[i]);
                idmap[node_id[i]] = (int32_t) i;
            }
            have_nodes = 1;
        } else if (starts(s, "$Elements")) {
            long long n = 0;
            if (!have_nodes) FAIL(CHARM_ERR_MESH, "%s: $Elements before $Nodes", path);
            if (!(s = rd_line(&r)) || sscanf(s, "%lld", &n) != 1 || n < 0)
                FAIL(CHARM_ERR_MESH, "%s:%lld: bad $Elements count", path, (long long) r.line);
            c->ttv = (int32_t *) malloc((size_t) (n ? n : 1) * 8 * sizeof(int32_t));
            c->region = (int32_t *) malloc((size_t) (n ? n : 1) * sizeof(int32_t));
            quads = (int32_t *) malloc((size_t) (n ? n : 1) * 4 * sizeof(int32_t));
            quad_tag = (int32_t *) malloc((size_t) (n ? n : 1) * sizeof(int32_t));
            if (!c->ttv || !c->region || !quads || !quad_tag) FAIL(CHARM_ERR_MEM, "charm_conn_read_msh: out of memory");
            for (long long i = 0; i < n; i++) {
                if (!(s = rd_line(&r))) FAIL(CHARM_ERR_MESH, "Premature end of file (%s)", path);
                /* elm-number elm-type number-of-tags <tags> node-number-list */
                char *p = (char *) s, *q;
                long long v[3 + 64 + 8];
                int k = 0;
                for (;;) {
                    long long x = strtoll(p, &q, 10);
                    if (q == p) break;
                    if (k == (int) (sizeof(v) / sizeof(v[0]))) FAIL(CHARM_ERR_MESH, "%s:%lld: element line too long", path, (long long) r.line);
                    v[k++] = x;
                    p = q;
                }
                if (k < 3 || v[2] < 0 || k < 3 + v[2]) FAIL(CHARM_ERR_MESH, "Premature end of file (%s:%lld)", path, (long long) r.line);
                const int type = (int) v[1], ntag = (int) v[2], nn = k - 3 - ntag;
                const long long *nodes = v + 3 + ntag;
                const int tag = ntag > 0 ? (int) v[3] : -1;
                if (type == 5 || type == 3) {
                    const int want = type == 5 ? 8 : 4;
                    if (nn != want) FAIL(CHARM_ERR_MESH, "%s:%lld: element type %d with %d nodes", path, (long long) r.line, type, nn);
                    int32_t idx[8];
                    for (int j = 0; j < want; j++) {
                        if (nodes[j] <= 0 || nodes[j] > max_id || idmap[nodes[j]] < 0)
                            FAIL(CHARM_ERR_MESH, "%s:%lld: unknown node %lld", path, (long long) r.line, nodes[j]);
                        idx[j] = idmap[nodes[j]];
                    }
                    if (type == 5) {
                        for (int j = 0; j < 8; j++) c->ttv[8 * c->nt + j] = idx[gmsh_hex_to_z[j]];
                        c->region[c->nt] = tag;
                        c->nt++;
                    } else {
                        /* quadrangle corners 0 1 3 2 (src/charm_connectivity.c:266-268): only their set is used */
                        memcpy(&quads[4 * n_quad], idx, 4 * sizeof(int32_t));
                        quad_tag[n_quad++] = tag;
                    }
                }
            }
            have_elems = 1;
        }
    }
    if (!have_nodes || !have_elems) FAIL(CHARM_ERR_MESH, "%s: no $Nodes or $Elements section", path);
    if (c->nt == 0) FAIL(CHARM_ERR_MESH, "%s: no hexahedra (element type 5)", path);
    {
        const int64_t nt = c->nt;
        c->ttt = (int32_t *) malloc((size_t) nt * 6 * sizeof(int32_t));
        c->ttf = (int8_t *) malloc((size_t) nt * 6);
        c->patch = (int32_t *) malloc((size_t) nt * 6 * sizeof(int32_t));
        fh_t m = {NULL, 0};
        if (!c->ttt || !c->ttf || !c->patch || !fh_init(&m, (size_t) nt * 6))
            FAIL(CHARM_ERR_MEM, "charm_conn_read_msh: out of memory");
        for (int64_t t = 0; t < nt; t++)
            for (int f = 0; f < 6; f++) {
                c->ttt[6 * t + f] = (int32_t) t;
                c->ttf[6 * t + f] = (int8_t) f;
                c->patch[6 * t + f] = -1;
            }
        /* shared faces: the second visit of a key links the two; a third is an error */
        for (int64_t t = 0; t < nt && rc == CHARM_OK; t++)
            for (int f = 0; f < 6; f++) {
                int32_t fc[4], key[4];
                for (int j = 0; j < 4; j++) fc[j] = c->ttv[8 * t + conn_face_corners[f][j]];
                sort4(fc, key);
                fh_slot_t *sl = fh_find(&m, key);
                if (sl->val < 0) { memcpy(sl->k, key, sizeof(key)); sl->val = 6 * t + f; continue; }
                if (sl->val >= 6 * nt) { rc = CHARM_ERR_MESH; charm_set_error("%s: a face shared by more than two hexahedra (tree %lld)", path, (long long) t); break; }
                const int64_t o = sl->val / 6;
                const int of = (int) (sl->val % 6);
                int32_t oc[4];
                for (int j = 0; j < 4; j++) oc[j] = c->ttv[8 * o + conn_face_corners[of][j]];
                /* orientation: corner 0 of the lower face number in the higher one's corners */
                const int32_t *lo = (f <= of) ? fc : oc, *hi = (f <= of) ? oc : fc;
                int rr = -1;
                for (int j = 0; j < 4; j++) if (hi[j] == lo[0]) rr = j;
                if (rr < 0) { rc = CHARM_ERR_MESH; charm_set_error("%s: face corner mismatch (tree %lld)", path, (long long) t); break; }
                c->ttt[6 * t + f] = (int32_t) o; c->ttf[6 * t + f] = (int8_t) (of + 6 * rr);
                c->ttt[6 * o + of] = (int32_t) t; c->ttf[6 * o + of] = (int8_t) (f + 6 * rr);
                sl->val = 6 * nt + 6 * t + f;          /* both sides seen */
            }
        /* boundary patches: every tree face with the quadrangle's corner set */
        for (int64_t i = 0; i < n_quad && rc == CHARM_OK; i++) {
            int32_t key[4];
            sort4(&quads[4 * i], key);
            fh_slot_t *sl = fh_find(&m, key);
            if (sl->val < 0) { rc = CHARM_ERR_MESH; charm_set_error("%s: boundary quadrangle %lld matches no hexahedron face", path, (long long) i); break; }
            const int64_t tf = sl->val >= 6 * nt ? sl->val - 6 * nt : sl->val;
            c->patch[tf] = quad_tag[i];
            if (sl->val >= 6 * nt) {                   /* an interior face: tag both sides */
                const int64_t t = tf / 6;
                const int f = (int) (tf % 6);
                c->patch[6 * (int64_t) c->ttt[6 * t + f] + c->ttf[6 * t + f] % 6] = quad_tag[i];
            }
        }
        free(m.s);
    }
done:
#undef FAIL
    if (r.f) fclose(r.f);
    free(r.buf); free(node_id); free(idmap); free(quads); free(quad_tag);
    if (rc != CHARM_OK) { charm_conn_destroy(c); return rc; }
    *out = c;
    return CHARM_OK;
}

int charm_conn_info(const charm_conn_t *c, int64_t *num_vertices, int64_t *num_trees, int *num_patches)
{
    if (!c) { charm_set_error("charm_conn_info: null connectivity"); return CHARM_ERR_ARG; }
    if (num_vertices) *num_vertices = c->nv;
    if (num_trees) *num_trees = c->nt;
    if (num_patches) *num_patches = c->np;
    return CHARM_OK;
}

int charm_conn_get(const charm_conn_t *c, double *vertices, int32_t *tree_to_vertex, int32_t *tree_to_tree,
                   int8_t *tree_to_face, int32_t *tree_to_patch, int32_t *tree_region)
{
    if (!c) { charm_set_error("charm_conn_get: null connectivity"); return CHARM_ERR_ARG; }
    if (vertices) memcpy(vertices, c->vert, (size_t) c->nv * 3 * sizeof(double));
    if (tree_to_vertex) memcpy(tree_to_vertex, c->ttv, (size_t) c->nt * 8 * sizeof(int32_t));
    if (tree_to_tree) memcpy(tree_to_tree, c->ttt, (size_t) c->nt * 6 * sizeof(int32_t));
    if (tree_to_face) memcpy(tree_to_face, c->ttf, (size_t) c->nt * 6);
    if (tree_to_patch) memcpy(tree_to_patch, c->patch, (size_t) c->nt * 6 * sizeof(int32_t));
    if (tree_region) memcpy(tree_region, c->region, (size_t) c->nt * sizeof(int32_t));
    return CHARM_OK;
}

int charm_conn_patch(const charm_conn_t *c, int i, charm_conn_patch_t *p)
{
    if (!c || !p || i < 0 || i >= c->np) { charm_set_error("charm_conn_patch: bad argument"); return CHARM_ERR_ARG; }
    *p = c->pn[i];
    return CHARM_OK;
}

static const char *patch_name(const charm_conn_t *c, int id)
{
    for (int i = 0; i < c->np; i++) if (c->pn[i].id == id) return c->pn[i].name;
    return NULL;
}

int charm_conn_bind_task(const charm_conn_t *c, const charm_task_t *task, int32_t *tree_bnd, int32_t *tree_reg)
{
    if (!c || !task) { charm_set_error("charm_conn_bind_task: null argument"); return CHARM_ERR_ARG; }
    for (int64_t t = 0; t < c->nt; t++) {
        for (int f = 0; f < 6 && tree_bnd; f++) {
            const int id = c->patch[6 * t + f];
            int b = -1;
            if (id >= 0) {
                const char *name = patch_name(c, id);
                if (!name) { charm_set_error("Patch (id=%d) has no physical name", id); return CHARM_ERR_MESH; }
                for (int s = 0; s < 6; s++) if (strcmp(task->bnd_name[s], name) == 0) b = s;
                if (b < 0) { charm_set_error("Patch (id=%d, name='%s') is nod defined in task.yaml", id, name); return CHARM_ERR_MESH; }
            }
            tree_bnd[6 * t + f] = b;
        }
        if (tree_reg) {
            const int id = c->region[t];
            const char *name = id >= 0 ? patch_name(c, id) : NULL;
            int g = -1;
            if (!name) { charm_set_error("Region (id=%d) has no physical name", id); return CHARM_ERR_MESH; }
            for (int i = 0; i < task->n_regions; i++) if (strcmp(task->region[i].name, name) == 0) g = i;
            if (g < 0) { charm_set_error("Region (id=%d, name='%s') is nod defined in task.yaml", id, name); return CHARM_ERR_MESH; }
            tree_reg[t] = g;
        }
    }
    return CHARM_OK;
}
