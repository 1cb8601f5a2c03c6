/*
 * charm_yaml.c -- task.yaml loader: the keys charm_init_context_yaml reads
 * (src/charm_init_yaml.cpp:333-426 and the fetch functions above it,
 * src/models/ns/charm_model_ns_init.cpp:14-88) into charm_task_t, the
 * unit-cube forest at MIN_LEVEL and the initial condition of
 * charm_init_initial_condition (src/charm_init.c:12-73).
 *
 * yaml-cpp is not available, so the file is read by a small parser of the
 * YAML subset task files use: block mappings and sequences by indentation,
 * "- key: value" items, flow mappings {a: b, c: {d: e}} and flow sequences
 * [1, 2], plain or quoted scalars, and # comments.
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "charm_internal.h"

#define YAML_VERSION_MIN "0.1.1"        /* include/charm_version.h:9 */

/* ---------------------------------------------------------------- nodes */

enum { YN_NULL = 0, YN_SCALAR, YN_MAP, YN_SEQ };

typedef struct yn {
    int         type;
    char       *key;       /* key within a mapping */
    char       *val;       /* scalar text */
    struct yn **kid;
    int         n, cap;
} yn_t;

static yn_t *yn_new(int type)
{
    yn_t *y = (yn_t *) calloc(1, sizeof(yn_t));
    if (y) y->type = type;
    return y;
}

static void yn_free(yn_t *y)
{
    int i;
    if (!y) return;
    for (i = 0; i < y->n; i++) yn_free(y->kid[i]);
    free(y->kid); free(y->key); free(y->val); free(y);
}

static int yn_add(yn_t *p, yn_t *c)
{
    if (!p || !c) return 0;
    if (p->n == p->cap) {
        int cap = p->cap ? 2 * p->cap : 8;
        yn_t **k = (yn_t **) realloc(p->kid, (size_t) cap * sizeof(yn_t *));
        if (!k) return 0;
        p->kid = k; p->cap = cap;
    }
    p->kid[p->n++] = c;
    return 1;
}

static char *dup_range(const char *a, const char *b)
{
    char *s;
    while (a < b && isspace((unsigned char) *a)) a++;
    while (b > a && isspace((unsigned char) b[-1])) b--;
    if (b - a >= 2 && ((*a == '"' && b[-1] == '"') || (*a == '\'' && b[-1] == '\''))) { a++; b--; }
    s = (char *) malloc((size_t) (b - a) + 1);
    if (!s) return NULL;
    memcpy(s, a, (size_t) (b - a));
    s[b - a] = 0;
    return s;
}

/* ---------------------------------------------------------------- lines */

typedef struct { int indent; char *txt; int lineno; } yline_t;

typedef struct {
    yline_t *l;
    int      n;
    char     err[256];
} yparse_t;

static void yerr(yparse_t *P, int lineno, const char *msg)
{
    if (!P->err[0]) snprintf(P->err, sizeof(P->err), "task.yaml line %d: %s", lineno, msg);
}

/* end of the text before a comment: '#' at the start or after white
 * space, outside quotes */
static char *strip_comment(char *s)
{
    char q = 0, *p;
    for (p = s; *p; p++) {
        if (q) { if (*p == q) q = 0; continue; }
        if (*p == '"' || *p == '\'') { q = *p; continue; }
        if (*p == '#' && (p == s || isspace((unsigned char) p[-1]))) { *p = 0; break; }
    }
    p = s + strlen(s);
    while (p > s && isspace((unsigned char) p[-1])) *--p = 0;
    return s;
}

/* ----------------------------------------------------------- flow style */

static yn_t *parse_flow(yparse_t *P, const char **ps, int lineno);

static void skip_ws(const char **ps) { while (**ps && isspace((unsigned char) **ps)) (*ps)++; }

/* a scalar up to one of the stop characters (outside quotes) */
static char *flow_scalar(const char **ps, const char *stop)
{
    const char *a = *ps, *p = a;
    char q = 0;
    for (; *p; p++) {
        if (q) { if (*p == q) q = 0; continue; }
        if (*p == '"' || *p == '\'') { q = *p; continue; }
        if (strchr(stop, *p)) break;
    }
    *ps = p;
    return dup_range(a, p);
}

static yn_t *parse_flow(yparse_t *P, const char **ps, int lineno)
{
    yn_t *y;
    skip_ws(ps);
    if (**ps == '{' || **ps == '[') {
        const int map = **ps == '{';
        const char close = map ? '}' : ']';
        y = yn_new(map ? YN_MAP : YN_SEQ);
        if (!y) return NULL;
        (*ps)++;
        skip_ws(ps);
        if (**ps == close) { (*ps)++; return y; }
        for (;;) {
            yn_t *c;
            char *key = NULL;
            skip_ws(ps);
            if (map) {
                key = flow_scalar(ps, ":,}]");
                if (**ps != ':') { free(key); yerr(P, lineno, "expected ':' in a flow mapping"); yn_free(y); return NULL; }
                (*ps)++;
            }
            skip_ws(ps);
            if (**ps == '{' || **ps == '[') c = parse_flow(P, ps, lineno);
            else { c = yn_new(YN_SCALAR); if (c) c->val = flow_scalar(ps, map ? ",}" : ",]"); }
            if (!c) { free(key); yn_free(y); return NULL; }
            c->key = key;
            if (!yn_add(y, c)) { yn_free(c); yn_free(y); return NULL; }
            skip_ws(ps);
            if (**ps == ',') { (*ps)++; continue; }
            if (**ps == close) { (*ps)++; return y; }
            yerr(P, lineno, "unterminated flow collection");
            yn_free(y);
            return NULL;
        }
    }
    y = yn_new(YN_SCALAR);
    if (y) { y->val = dup_range(*ps, *ps + strlen(*ps)); *ps += strlen(*ps); }
    return y;
}

static yn_t *parse_value_text(yparse_t *P, const char *txt, int lineno)
{
    const char *p = txt;
    yn_t *y;
    skip_ws(&p);
    if (!*p) return yn_new(YN_NULL);
    y = parse_flow(P, &p, lineno);
    skip_ws(&p);
    if (y && *p) { yerr(P, lineno, "text after a flow collection"); yn_free(y); return NULL; }
    return y;
}

/* ---------------------------------------------------------- block style */

static int is_item(const char *t) { return t[0] == '-' && (t[1] == 0 || t[1] == ' '); }

/* position of the ':' that ends a mapping key on this line, or NULL */
static char *key_colon(char *t)
{
    char q = 0, *p;
    if (*t == '{' || *t == '[') return NULL;
    for (p = t; *p; p++) {
        if (q) { if (*p == q) q = 0; continue; }
        if (*p == '"' || *p == '\'') { q = *p; continue; }
        if (*p == ':' && (p[1] == 0 || p[1] == ' ')) return p;
    }
    return NULL;
}

static yn_t *parse_block(yparse_t *P, int *i, int indent);

/* the value of "key:" with nothing after it: the deeper block below, or a
 * sequence at the key's own indentation, or null */
static yn_t *block_below(yparse_t *P, int *i, int indent)
{
    if (*i < P->n && (P->l[*i].indent > indent || (P->l[*i].indent == indent && is_item(P->l[*i].txt))))
        return parse_block(P, i, P->l[*i].indent);
    return yn_new(YN_NULL);
}

static yn_t *parse_block(yparse_t *P, int *i, int indent)
{
    yn_t *y;
    if (*i >= P->n) return yn_new(YN_NULL);
    if (is_item(P->l[*i].txt)) {
        y = yn_new(YN_SEQ);
        if (!y) return NULL;
        while (*i < P->n && P->l[*i].indent == indent && is_item(P->l[*i].txt)) {
            yline_t *L = &P->l[*i];
            char *rest = L->txt + 1;
            yn_t *c;
            while (*rest == ' ') rest++;
            if (!*rest) {
                (*i)++;
                c = (*i < P->n && P->l[*i].indent > indent) ? parse_block(P, i, P->l[*i].indent) : yn_new(YN_NULL);
            } else if (key_colon(rest)) {
                /* "- key: ..." opens a mapping whose keys sit at the column of
                 * its first key */
                L->indent += (int) (rest - L->txt);
                memmove(L->txt, rest, strlen(rest) + 1);
                c = parse_block(P, i, L->indent);
            } else {
                c = parse_value_text(P, rest, L->lineno);
                (*i)++;
            }
            if (!c || !yn_add(y, c)) { yn_free(c); yn_free(y); return NULL; }
        }
        return y;
    }
    y = yn_new(YN_MAP);
    if (!y) return NULL;
    while (*i < P->n && P->l[*i].indent == indent && !is_item(P->l[*i].txt)) {
        yline_t *L = &P->l[*i];
        char *colon = key_colon(L->txt);
        yn_t *c;
        if (!colon) { yerr(P, L->lineno, "expected 'key: value'"); yn_free(y); return NULL; }
        (*i)++;
        if (colon[1] == 0) c = block_below(P, i, indent);
        else c = parse_value_text(P, colon + 1, L->lineno);
        if (!c) { yn_free(y); return NULL; }
        c->key = dup_range(L->txt, colon);
        if (!yn_add(y, c)) { yn_free(c); yn_free(y); return NULL; }
    }
    if (*i < P->n && P->l[*i].indent > indent) {
        yerr(P, P->l[*i].lineno, "unexpected indentation");
        yn_free(y);
        return NULL;
    }
    return y;
}

static yn_t *yaml_load(const char *path, char *err, size_t errlen)
{
    FILE *fp = fopen(path, "rb");
    yparse_t P;
    char *buf = NULL, *s, *line;
    long len;
    int i = 0, cap = 0, lineno = 0;
    yn_t *root = NULL;
    memset(&P, 0, sizeof(P));
    if (!fp) { snprintf(err, errlen, "cannot open %s", path); return NULL; }
    if (fseek(fp, 0, SEEK_END) || (len = ftell(fp)) < 0 || fseek(fp, 0, SEEK_SET)) {
        fclose(fp); snprintf(err, errlen, "cannot read %s", path); return NULL;
    }
    buf = (char *) malloc((size_t) len + 1);
    if (!buf || fread(buf, 1, (size_t) len, fp) != (size_t) len) {
        fclose(fp); free(buf); snprintf(err, errlen, "cannot read %s", path); return NULL;
    }
    fclose(fp);
    buf[len] = 0;
    for (s = buf; s && *s; ) {
        char *nl = strchr(s, '\n');
        int ind = 0;
        if (nl) *nl = 0;
        line = s;
        s = nl ? nl + 1 : NULL;
        lineno++;
        {
            char *r = strchr(line, '\r');
            if (r) *r = 0;
        }
        strip_comment(line);
        while (line[ind] == ' ') ind++;
        if (line[ind] == '\t') { snprintf(err, errlen, "task.yaml line %d: tab indentation", lineno); goto fail; }
        if (!line[ind] || !strcmp(line + ind, "---")) continue;
        if (P.n == cap) {
            yline_t *l;
            cap = cap ? 2 * cap : 128;
            l = (yline_t *) realloc(P.l, (size_t) cap * sizeof(yline_t));
            if (!l) { snprintf(err, errlen, "out of memory"); goto fail; }
            P.l = l;
        }
        P.l[P.n].indent = ind;
        P.l[P.n].txt = line + ind;
        P.l[P.n].lineno = lineno;
        P.n++;
    }
    root = P.n ? parse_block(&P, &i, P.l[0].indent) : yn_new(YN_MAP);
    if (root && i < P.n) { yerr(&P, P.l[i].lineno, "unexpected indentation"); yn_free(root); root = NULL; }
    /* keys and scalars are copies: buf can go */
    if (!root) snprintf(err, errlen, "%s", P.err[0] ? P.err : "out of memory");
fail:
    free(P.l);
    free(buf);
    return root;
}

/* ---------------------------------------------------------- node access */

static const yn_t *yget(const yn_t *m, const char *key)
{
    int i;
    if (!m || m->type != YN_MAP) return NULL;
    for (i = 0; i < m->n; i++)
        if (m->kid[i]->key && !strcmp(m->kid[i]->key, key)) return m->kid[i];
    return NULL;
}

typedef struct {
    char err[256];
} yctx_t;

static void cerr(yctx_t *C, const char *fmt, const char *a, const char *b)
{
    if (!C->err[0]) snprintf(C->err, sizeof(C->err), fmt, a, b);
}

static const char *ystr(yctx_t *C, const yn_t *m, const char *key, const char *where)
{
    const yn_t *y = yget(m, key);
    if (!y || y->type != YN_SCALAR || !y->val) { cerr(C, "task.yaml: missing '%s' in %s", key, where); return ""; }
    return y->val;
}

static double ynum(yctx_t *C, const yn_t *m, const char *key, const char *where)
{
    const char *s = ystr(C, m, key, where);
    char *end;
    double v;
    if (!*s) return 0.0;
    v = strtod(s, &end);
    while (*end && isspace((unsigned char) *end)) end++;
    if (end == s || *end) cerr(C, "task.yaml: '%s' in %s is not a number", key, where);
    return v;
}

static int yint(yctx_t *C, const yn_t *m, const char *key, const char *where)
{
    const double v = ynum(C, m, key, where);
    if (v != floor(v)) cerr(C, "task.yaml: '%s' in %s is not an integer", key, where);
    return (int) v;
}

static double ynum_node(yctx_t *C, const yn_t *y, const char *what)
{
    char *end;
    double v = 0.0;
    if (!y || y->type != YN_SCALAR || !y->val) { cerr(C, "task.yaml: %s%s is not a number", what, ""); return 0.0; }
    v = strtod(y->val, &end);
    if (end == y->val || *end) cerr(C, "task.yaml: %s%s is not a number", what, "");
    return v;
}

/* ----------------------------------------------------------- the task */

/* the version check of _charm_init_yaml_check_version
 * (src/charm_init_yaml.cpp:276-327): three dot-separated integers, at
 * least the loader's */
static int version_ok(const char *v)
{
    int a[3], b[3], i;
    if (sscanf(v, "%d.%d.%d", &a[0], &a[1], &a[2]) != 3) return 0;
    if (sscanf(YAML_VERSION_MIN, "%d.%d.%d", &b[0], &b[1], &b[2]) != 3) return 0;
    for (i = 0; i < 3; i++) {
        if (a[i] > b[i]) return 1;
        if (a[i] < b[i]) return 0;
    }
    return 1;
}

static int type_of(const char *s, const char *const *names, int n)
{
    int i;
    for (i = 0; i < n; i++)
        if (!strcmp(s, names[i])) return i;
    return -1;
}

/* cube side of a boundary name on the unit-cube mesh */
static int cube_side(const char *name)
{
    static const char *const sides[6] = {"LEFT", "RIGHT", "FRONT", "BACK", "BOTTOM", "TOP"};
    return type_of(name, sides, 6);
}

static void load_task(yctx_t *C, const yn_t *root, charm_task_t *T)
{
    charm_config_t *cfg = &T->cfg;
    const yn_t *ctl, *model, *comps, *mats, *regs, *bnds, *mesh, *rx;
    const char *s;
    int i, j, comp_id[CHARM_MAX_COMP];
    static const char *const flux_names[4] = {"LF", "HLLC", "GODUNOV", "CD"};
    static const int flux_ids[4] = {CHARM_FLUX_LF, CHARM_FLUX_HLLC, CHARM_FLUX_GODUNOV, CHARM_FLUX_CD};

    if (strcmp(ystr(C, root, "method", "the file"), "CHARM_3D")) { cerr(C, "task.yaml: %s%s", "method name must be 'CHARM_3D'", ""); return; }
    if (!version_ok(ystr(C, root, "version", "the file"))) {
        cerr(C, "task.yaml: wrong version; must be %s or higher%s", YAML_VERSION_MIN, "");
        return;
    }
    ctl = yget(root, "control");
    if (!ctl || ctl->type != YN_MAP) { cerr(C, "task.yaml: missing '%s'%s", "control", ""); return; }

    s = ystr(C, ctl, "FLUX_TYPE", "control");
    i = type_of(s, flux_names, 4);
    if (i < 0) { cerr(C, "task.yaml: unknown flux type '%s'%s", s, ""); return; }
    cfg->flux = flux_ids[i];
    s = ystr(C, ctl, "LIMITER", "control");
    if (!strcmp(s, "NONE")) cfg->limiter = 0;
    else if (!strcmp(s, "BJ")) cfg->limiter = 1;
    else { cerr(C, "task.yaml: unknown limiter type '%s'%s", s, ""); return; }
    cfg->max_err = ynum(C, ctl, "MAX_ERROR", "control");
    cfg->refine_period = yint(C, ctl, "REFINE_PERIOD", "control");
    T->repartition_period = yint(C, ctl, "REPARTITION_PERIOD", "control");
    cfg->min_level = yint(C, ctl, "MIN_LEVEL", "control");
    cfg->max_level = yint(C, ctl, "MAX_LEVEL", "control");
    cfg->write_period = yint(C, ctl, "FILE_OUTPUT_STEP", "control");
    cfg->log_period = yint(C, ctl, "LOG_OUTPUT_STEP", "control");
    cfg->dt = ynum(C, ctl, "TAU", "control");
    cfg->cfl = ynum(C, ctl, "CFL", "control");
    cfg->tmax = ynum(C, ctl, "TMAX", "control");
    /* this build's own key: the bulk arithmetic (default FP64, the reference's) */
    cfg->precision = CHARM_FP64;
    if (yget(ctl, "PRECISION")) {
        static const char *const pn[3] = {"FP64", "FP32", "FP32_MIX"};
        s = ystr(C, ctl, "PRECISION", "control");
        i = type_of(s, pn, 3);
        if (i < 0) { cerr(C, "task.yaml: unknown precision '%s'%s", s, ""); return; }
        cfg->precision = i;
    }

    /* components (_charm_init_fetch_comp) */
    comps = yget(root, "components");
    if (!comps || comps->type != YN_SEQ || comps->n < 1) { cerr(C, "task.yaml: missing '%s'%s", "components", ""); return; }
    if (comps->n > CHARM_MAX_COMP) { cerr(C, "task.yaml: more than %s components%s", "8", ""); return; }
    cfg->n_comp = comps->n;
    for (i = 0; i < comps->n; i++) {
        const yn_t *q = comps->kid[i], *cp;
        charm_comp_desc_t *d = &cfg->comp[i];
        static const char *const ct[2] = {"CONST", "POLYNOM"}, *const tt[2] = {"CONST", "SATHERLAND"};
        comp_id[i] = yint(C, q, "id", "a component");
        snprintf(cfg->comp_name[i], sizeof(cfg->comp_name[i]), "%s", ystr(C, q, "name", "a component"));
        s = ystr(C, q, "cp_type", "a component");
        if ((d->cp_type = type_of(s, ct, 2)) < 0) { cerr(C, "task.yaml: unknown Cp type '%s'%s", s, ""); return; }
        s = ystr(C, q, "kp_type", "a component");
        if ((d->kp_type = type_of(s, tt, 2)) < 0) { cerr(C, "task.yaml: unknown KP type '%s'%s", s, ""); return; }
        s = ystr(C, q, "ml_type", "a component");
        if ((d->ml_type = type_of(s, tt, 2)) < 0) { cerr(C, "task.yaml: unknown ML type '%s'%s", s, ""); return; }
        d->m = ynum(C, q, "M", "a component");
        d->ml0 = ynum(C, q, "ML0", "a component");
        d->kp0 = ynum(C, q, "KP0", "a component");
        d->t0 = ynum(C, q, "T0", "a component");
        d->ts = ynum(C, q, "TS", "a component");
        d->sig = ynum(C, q, "Sig", "a component");
        d->ek = ynum(C, q, "ek", "a component");
        d->h0 = ynum(C, q, "h0", "a component");
        cp = yget(q, "Cp");
        if (!cp || cp->type != YN_SEQ || cp->n < 1 || cp->n > 8) { cerr(C, "task.yaml: component '%s' needs 1..8 Cp values%s", cfg->comp_name[i], ""); return; }
        d->n_cp = cp->n;
        for (j = 0; j < cp->n; j++) d->cp[j] = ynum_node(C, cp->kid[j], "a Cp value");
    }

    /* regions (_charm_init_fetch_reg): concentrations by component id */
    regs = yget(root, "regions");
    if (!regs || regs->type != YN_SEQ || regs->n < 1) { cerr(C, "task.yaml: missing '%s'%s", "regions", ""); return; }
    if (regs->n > 16) { cerr(C, "task.yaml: more than %s regions%s", "16", ""); return; }
    T->n_regions = regs->n;
    for (i = 0; i < regs->n; i++) {
        const yn_t *r = regs->kid[i], *par = yget(r, "parameters"), *cs = yget(r, "components");
        charm_region_desc_t *g = &T->region[i];
        double sum = 0.0;
        g->id = yint(C, r, "id", "a region");
        snprintf(g->name, sizeof(g->name), "%s", ystr(C, r, "name", "a region"));
        g->cell_type = yget(r, "cell_type") ? yint(C, r, "cell_type", "a region") : 0;
        g->mat_id = yint(C, r, "material_id", "a region");
        g->v[0] = ynum(C, yget(par, "V"), "x", "a region's V");
        g->v[1] = ynum(C, yget(par, "V"), "y", "a region's V");
        g->v[2] = ynum(C, yget(par, "V"), "z", "a region's V");
        g->t = ynum(C, par, "T", "a region's parameters");
        g->p = ynum(C, par, "P", "a region's parameters");
        g->grav[0] = ynum(C, yget(par, "G"), "x", "a region's G");
        g->grav[1] = ynum(C, yget(par, "G"), "y", "a region's G");
        g->grav[2] = ynum(C, yget(par, "G"), "z", "a region's G");
        if (!cs || cs->type != YN_SEQ) { cerr(C, "task.yaml: region '%s' has no components%s", g->name, ""); return; }
        for (j = 0; j < cs->n; j++) {
            const int id = yint(C, cs->kid[j], "id", "a region's component");
            int k, idx = -1;
            for (k = 0; k < cfg->n_comp; k++) if (comp_id[k] == id) idx = k;
            if (idx < 0) { cerr(C, "task.yaml: unknown component id in region '%s'%s", g->name, ""); return; }
            g->c[idx] = ynum(C, cs->kid[j], "concentration", "a region's component");
        }
        for (j = 0; j < CHARM_MAX_COMP; j++) sum += g->c[j];
        /* the reference's test, src/charm_init_yaml.cpp:245-248 */
        if (fabs(sum) - 1. > 1.e-12) { cerr(C, "task.yaml: sum of concentrations for region '%s' is not equal to 1%s", g->name, ""); return; }
    }

    /* materials (_charm_init_fetch_mat): the EOS of the first region's material */
    mats = yget(root, "materials");
    if (!mats || mats->type != YN_SEQ || mats->n < 1) { cerr(C, "task.yaml: missing '%s'%s", "materials", ""); return; }
    {
        int found = 0;
        for (i = 0; i < mats->n; i++) {
            const yn_t *m = mats->kid[i];
            const int id = yint(C, m, "id", "a material");
            s = ystr(C, m, "eof_type", "a material");
            if (strcmp(s, "IDEAL") && strcmp(s, "MIX")) {
                cerr(C, "task.yaml: EOS '%s' is not supported (IDEAL, MIX)%s", s, "");
                return;
            }
            if (id == T->region[0].mat_id) { cfg->eos = strcmp(s, "IDEAL") ? CHARM_EOS_MIX : CHARM_EOS_IDEAL; found = 1; }
        }
        if (!found) { cerr(C, "task.yaml: region '%s' names an unknown material%s", T->region[0].name, ""); return; }
    }
    for (i = 0; i < 3; i++) cfg->grav[i] = T->region[0].grav[i];

    /* mesh (_charm_init_mesh_info) */
    mesh = yget(root, "mesh");
    if (!mesh) { cerr(C, "task.yaml: missing '%s'%s", "mesh", ""); return; }
    snprintf(T->mesh_files_type, sizeof(T->mesh_files_type), "%s", ystr(C, mesh, "files_type", "mesh"));
    if (yget(mesh, "name")) snprintf(T->mesh_name, sizeof(T->mesh_name), "%s", ystr(C, mesh, "name", "mesh"));
    if (!strcmp(T->mesh_files_type, "unit_cube") && yget(mesh, "periodic")) {
        const yn_t *pr = yget(mesh, "periodic");
        if (pr->type != YN_SEQ || pr->n != 3) { cerr(C, "task.yaml: mesh.periodic needs three values%s%s", "", ""); return; }
        for (i = 0; i < 3; i++) T->mesh_periodic[i] = ynum_node(C, pr->kid[i], "a periodic flag") != 0.0;
    }

    /* boundaries (_charm_init_fetch_bnd): by name onto the cube sides of a
     * unit-cube mesh; a Gmsh mesh's boundaries take the slots 0-5 in the
     * order of the file (the mesh's patches bind to them by name) */
    for (i = 0; i < 6; i++) cfg->bnd_type[i] = T->mesh_periodic[i / 2] ? CHARM_BND_PERIODIC : -1;
    bnds = yget(root, "boundaries");
    if (bnds && bnds->type == YN_SEQ) {
        const int cube = !strcmp(T->mesh_files_type, "unit_cube");
        if (!cube && bnds->n > 6) { cerr(C, "task.yaml: %d boundaries (at most 6)%s", bnds->n, ""); return; }
        static const char *const bn[5] = {"BOUND_INLET", "BOUND_OUTLET", "BOUND_WALL_SLIP", "BOUND_WALL_NO_SLIP", "BOUND_MASS_FLOW"};
        for (i = 0; i < bnds->n; i++) {
            const yn_t *b = bnds->kid[i], *par = yget(b, "parameters");
            const char *name = ystr(C, b, "name", "a boundary");
            const int side = cube ? cube_side(name) : i;
            int type;
            s = ystr(C, b, "type", "a boundary");
            if ((type = type_of(s, bn, 5)) < 0) { cerr(C, "task.yaml: unknown boundary type '%s'%s", s, ""); return; }
            if (side < 0) {
                if (!strcmp(T->mesh_files_type, "unit_cube")) {
                    cerr(C, "task.yaml: boundary '%s' is not a cube side (LEFT RIGHT FRONT BACK BOTTOM TOP)%s", name, "");
                    return;
                }
                continue;
            }
            if (T->mesh_periodic[side / 2]) { cerr(C, "task.yaml: boundary '%s' lies on a periodic axis%s", name, ""); return; }
            snprintf(T->bnd_name[side], sizeof(T->bnd_name[side]), "%s", name);
            cfg->bnd_type[side] = type;
            if (type == CHARM_BND_INLET) {
                cfg->bnd_param[side][0] = ynum(C, yget(par, "V"), "x", "an inlet's V");
                cfg->bnd_param[side][1] = ynum(C, yget(par, "V"), "y", "an inlet's V");
                cfg->bnd_param[side][2] = ynum(C, yget(par, "V"), "z", "an inlet's V");
                cfg->bnd_param[side][3] = ynum(C, par, "T", "an inlet's parameters");
                cfg->bnd_param[side][4] = ynum(C, par, "P", "an inlet's parameters");
            } else if (type == CHARM_BND_WALL_NO_SLIP) {
                cfg->bnd_param[side][0] = ynum(C, par, "T", "a wall's parameters");
            } else if (type == CHARM_BND_MASS_FLOW) {
                static const char *const mk[7] = {"M", "P", "T", "CosX", "CosY", "CosZ", "P"};
                const yn_t *cs = yget(par, "components");
                for (j = 0; j < 7; j++) cfg->bnd_param[side][j] = ynum(C, par, mk[j], "a mass-flow boundary's parameters");
                if (cs && cs->type == YN_SEQ)
                    for (j = 0; j < cs->n && 7 + j < 16; j++) cfg->bnd_param[side][7 + j] = ynum_node(C, cs->kid[j], "a mass-flow concentration");
            }
        }
    }
    if (!strcmp(T->mesh_files_type, "unit_cube"))
        for (i = 0; i < 6; i++)
            if (cfg->bnd_type[i] < 0) { cerr(C, "task.yaml: no boundary for cube side %s%s", (const char *[]){"LEFT", "RIGHT", "FRONT", "BACK", "BOTTOM", "TOP"}[i], ""); return; }

    /* model (charm_model_euler_init / charm_model_ns_init) */
    model = yget(ctl, "MODEL");
    s = ystr(C, model, "name", "control.MODEL");
    if (!strcmp(s, "EULER")) {
        cfg->model = CHARM_MODEL_EULER;
    } else if (!strcmp(s, "NS")) {
        cfg->model = CHARM_MODEL_NS;
        cfg->use_visc = yint(C, model, "use_visc", "control.MODEL");
        cfg->use_diff = yint(C, model, "use_diffusion", "control.MODEL");
        cfg->t_ref = ynum(C, model, "t_ref", "control.MODEL");
        rx = yget(root, "reactions");
        if (rx && rx->type == YN_SEQ) {
            if (rx->n > 4) { cerr(C, "task.yaml: more than %s reactions%s", "4", ""); return; }
            cfg->n_react = rx->n;
            for (i = 0; i < rx->n; i++) {
                const yn_t *r = rx->kid[i];
                const yn_t *side[2] = {yget(r, "left"), yget(r, "right")};
                charm_reaction_desc_t *d = &cfg->react[i];
                int k;
                d->a = ynum(C, r, "a", "a reaction");
                d->e = ynum(C, r, "e", "a reaction");
                d->n = ynum(C, r, "n", "a reaction");
                for (k = 0; k < 2; k++) {
                    int *dst = k ? d->right : d->left;
                    dst[0] = dst[1] = dst[2] = -1;
                    if (!side[k] || side[k]->type != YN_SEQ || side[k]->n > 3) {
                        cerr(C, "task.yaml: a reaction needs at most 3 components %s%s", k ? "right" : "left", "");
                        return;
                    }
                    for (j = 0; j < side[k]->n; j++) {
                        const int comp = (int) ynum_node(C, side[k]->kid[j], "a reaction component");
                        if (!(-1 <= comp && comp < cfg->n_comp)) { cerr(C, "task.yaml: wrong component number in a reaction%s%s", "", ""); return; }
                        dst[j] = comp;
                    }
                }
            }
        }
    } else {
        cerr(C, "task.yaml: unknown model type '%s'%s", s, "");
    }
}

int charm_task_load_yaml(const char *path, charm_task_t *task)
{
    char err[256] = {0};
    yn_t *root;
    yctx_t C;
    if (!path || !task) { charm_set_error("task_load_yaml: null argument"); return CHARM_ERR_ARG; }
    root = yaml_load(path, err, sizeof(err));
    if (!root) { charm_set_error("%s", err); return CHARM_ERR_ARG; }
    memset(task, 0, sizeof(*task));
    memset(&C, 0, sizeof(C));
    load_task(&C, root, task);
    yn_free(root);
    if (C.err[0]) { charm_set_error("%s", C.err); return CHARM_ERR_ARG; }
    return CHARM_OK;
}

int charm_task_forest(const charm_task_t *task, charm_forest_t **out)
{
    if (!task || !out) { charm_set_error("task_forest: null argument"); return CHARM_ERR_ARG; }
    if (strcmp(task->mesh_files_type, "unit_cube")) {
        charm_set_error("task_forest: mesh type '%s' is not supported (this path runs on the unit-cube "
                        "octree: files_type unit_cube)", task->mesh_files_type);
        return CHARM_ERR_UNSUPPORTED;
    }
    /* p4est_new_ext(..., min_level, fill_uniform = 1, ...), src/charm_3d.c:36-37 */
    return charm_forest_new_uniform(task->cfg.min_level, task->mesh_periodic, out);
}

/* one region's state as coefficients [V][4]: (T, p) -> (r, cz, e) by EOS
 * flag 3 (src/charm_eos.c:57-96), then charm_param_prim_to_cons
 * (src/charm_globals.c:211-223) into coefficient 0, the slopes zero
 * (src/charm_init.c:47-66) */
static void region_coeffs(const charm_config_t *cfg, const charm_region_desc_t *g, double *w)
{
    const int nv = 4 + cfg->n_comp;
    charm_prim_t p;
    int i;
    memset(&p, 0, sizeof(p));
    p.mat_id = g->mat_id;
    p.p = g->p;
    p.t = g->t;
    p.u = g->v[0]; p.v = g->v[1]; p.w = g->v[2];
    for (i = 0; i < cfg->n_comp; i++) p.c[i] = g->c[i];
    if (cfg->eos == CHARM_EOS_IDEAL) charm_eos_ideal_cfg(cfg, &p, 3);
    else charm_eos_mix_cfg(cfg, &p, 3);
    for (i = 0; i < nv * 4; i++) w[i] = 0.0;
    w[0] = p.r * p.u;
    w[4] = p.r * p.v;
    w[8] = p.r * p.w;
    w[12] = p.r * (p.e + 0.5 * (p.u * p.u + p.v * p.v + p.w * p.w));
    for (i = 0; i < cfg->n_comp; i++) w[4 * (4 + i)] = p.r * p.c[i];
}

int charm_task_gmesh(const charm_task_t *task, const char *dir, charm_gmesh_t **out)
{
    char path[600];
    charm_conn_t *conn = NULL;
    int64_t nv = 0, nt = 0;
    int np = 0, rc;
    double *v = NULL;
    int32_t *ttv = NULL, *tb = NULL, *tr = NULL;
    if (!task || !out) { charm_set_error("task_gmesh: null argument"); return CHARM_ERR_ARG; }
    *out = NULL;
    if (strcmp(task->mesh_files_type, "gmsh_msh")) {
        charm_set_error("task_gmesh: mesh type '%s' is not a Gmsh mesh (files_type gmsh_msh)", task->mesh_files_type);
        return CHARM_ERR_ARG;
    }
    if (task->mesh_name[0] == '/' || !dir || !dir[0]) snprintf(path, sizeof(path), "%s", task->mesh_name);
    else snprintf(path, sizeof(path), "%.300s/%s", dir, task->mesh_name);
    if ((rc = charm_conn_read_msh(path, &conn))) return rc;
    if ((rc = charm_conn_info(conn, &nv, &nt, &np))) goto done;
    v = (double *) malloc((size_t) (3 * nv + 1) * sizeof(double));
    ttv = (int32_t *) malloc((size_t) (8 * nt + 1) * sizeof(int32_t));
    tb = (int32_t *) malloc((size_t) (6 * nt + 1) * sizeof(int32_t));
    tr = (int32_t *) malloc((size_t) (nt + 1) * sizeof(int32_t));
    if (!v || !ttv || !tb || !tr) { charm_set_error("task_gmesh: out of memory"); rc = CHARM_ERR_MEM; goto done; }
    if ((rc = charm_conn_get(conn, v, ttv, NULL, NULL, NULL, NULL))) goto done;
    if ((rc = charm_conn_bind_task(conn, task, tb, tr))) goto done;
    /* p4est_new_ext(..., min_level, fill_uniform = 1, ...), src/charm_3d.c:36-37 */
    if ((rc = charm_gmesh_build(nv, v, nt, ttv, tb, task->cfg.min_level, out))) goto done;
    rc = charm_gmesh_set_regions(*out, tr);
    if (rc) { charm_gmesh_destroy(*out); *out = NULL; }
done:
    charm_conn_destroy(conn);
    free(v); free(ttv); free(tb); free(tr);
    return rc;
}

int charm_task_initial_state_gmesh(const charm_task_t *task, const charm_gmesh_t *m, double *u, int64_t n)
{
    int64_t nl = 0, q;
    int32_t *lv = NULL;
    const int32_t *treg;
    int nv;
    if (!task || !m || !u) { charm_set_error("task_initial_state_gmesh: null argument"); return CHARM_ERR_ARG; }
    charm_gmesh_info(m, &nl, NULL);
    if (n != nl) {
        charm_set_error("task_initial_state_gmesh: %lld leaves, state for %lld", (long long) nl, (long long) n);
        return CHARM_ERR_ARG;
    }
    treg = charm_gmesh_regions(m);
    if (!treg) { charm_set_error("task_initial_state_gmesh: the mesh has no regions (charm_task_gmesh binds them)"); return CHARM_ERR_ARG; }
    lv = (int32_t *) malloc((size_t) (5 * n + 1) * sizeof(int32_t));
    if (!lv) { charm_set_error("task_initial_state_gmesh: out of memory"); return CHARM_ERR_MEM; }
    charm_gmesh_leaves(m, lv);
    nv = 4 + task->cfg.n_comp;
    /* each leaf takes the region of its tree (the tree attribute's region,
     * src/charm_init.c:26-27) */
    for (q = 0; q < n; q++) {
        const int g = treg[lv[5 * q]];
        if (g < 0 || g >= task->n_regions) {
            free(lv);
            charm_set_error("task_initial_state_gmesh: tree %d has region %d of %d", lv[5 * q], g, task->n_regions);
            return CHARM_ERR_ARG;
        }
        region_coeffs(&task->cfg, &task->region[g], u + (size_t) q * nv * 4);
    }
    free(lv);
    return CHARM_OK;
}

int charm_task_initial_state(const charm_task_t *task, const charm_forest_t *f, double *u, int64_t n)
{
    const charm_config_t *cfg;
    const charm_region_desc_t *g;
    int nv, i;
    int64_t q;
    if (!task || !f || !u) { charm_set_error("task_initial_state: null argument"); return CHARM_ERR_ARG; }
    if (n != charm_forest_num_leaves(f)) {
        charm_set_error("task_initial_state: %lld leaves, state for %lld",
                        (long long) charm_forest_num_leaves(f), (long long) n);
        return CHARM_ERR_ARG;
    }
    cfg = &task->cfg;
    nv = 4 + cfg->n_comp;
    /* one tree: the region of cell type 0 (else the first region) */
    g = &task->region[0];
    for (i = 0; i < task->n_regions; i++)
        if (task->region[i].cell_type == 0) { g = &task->region[i]; break; }
    for (q = 0; q < n; q++) region_coeffs(cfg, g, u + (size_t) q * nv * 4);
    return CHARM_OK;
}
