/*
 * charm_output.c -- VTK cell-data output of the leaves and their state:
 * charm_write_solution (src/charm_output.c:56-126 of the reference) and the
 * plain-array writer charm_write_vtk it is built on.
 *
 * Format: one VTK XML UnstructuredGrid piece (<name>_0000.vtu, raw appended
 * data, UInt64 headers, little endian) and its parallel index <name>.pvtu,
 * as the reference's p4est-style writer names them.  Points are the leaf
 * corners shared between leaves (continuous mode), cells are hexahedra;
 * cell data: treeid, level, mpirank, then the cell-centre fields
 * R P E E_TOT U V W T C_<species> of the reference
 * (charm_interpolate_cell_solution, src/charm_output.c:10-52): the Taylor
 * basis vanishes at the centre except phi_0, so the centre state is the
 * coefficient-0 state, turned into primitive fields by
 * charm_param_cons_to_prim with EOS flag 4 (src/charm_globals.c:183-208,
 * src/charm_eos.c:57-148).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "charm_internal.h"

/* centre-state primitive fields: out = R P E E_TOT U V W T C_0..C_{nc-1},
 * charm_param_cons_to_prim of the coefficient-0 state */
static void cell_fields(const charm_config_t *cfg, const double *u /* [V][4] */, double *out)
{
    const int nc = cfg->n_comp;
    charm_prim_t p;
    int i;
    memset(&p, 0, sizeof(p));
    for (i = 0; i < nc; i++) p.r += u[4 * (4 + i)];
    p.u = u[0] / p.r;
    p.v = u[4] / p.r;
    p.w = u[8] / p.r;
    p.e_tot = u[12] / p.r;
    p.e = p.e_tot - 0.5 * (p.u * p.u + p.v * p.v + p.w * p.w);
    for (i = 0; i < nc; i++) {
        p.c[i] = u[4 * (4 + i)] / p.r;
        if (p.c[i] < 0.0) p.c[i] = 0.0;
        if (p.c[i] > 1.0) p.c[i] = 1.0;
    }
    if (cfg->eos == CHARM_EOS_IDEAL) charm_eos_ideal_cfg(cfg, &p, 4);
    else charm_eos_mix_cfg(cfg, &p, 4);
    out[0] = p.r; out[1] = p.p; out[2] = p.e; out[3] = p.e_tot;
    out[4] = p.u; out[5] = p.v; out[6] = p.w; out[7] = p.t;
    for (i = 0; i < nc; i++) out[8 + i] = p.c[i];
}

static int cmp_u64(const void *a, const void *b)
{
    const uint64_t x = *(const uint64_t *) a, y = *(const uint64_t *) b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static uint64_t corner_key(int32_t x, int32_t y, int32_t z)
{
    return ((uint64_t) x << 42) | ((uint64_t) y << 21) | (uint64_t) z;
}

static int64_t find_key(const uint64_t *keys, int64_t n, uint64_t k)
{
    int64_t lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int64_t mid = (lo + hi) / 2;
        if (keys[mid] == k) return mid;
        if (keys[mid] < k) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* one appended block: UInt64 byte count, then the bytes */
static int put_block(FILE *fp, const void *data, uint64_t bytes)
{
    return fwrite(&bytes, sizeof(bytes), 1, fp) == 1 &&
           (bytes == 0 || fwrite(data, 1, (size_t) bytes, fp) == (size_t) bytes);
}

/* the piece <name>_0000.vtu and its index <name>.pvtu: np points [np][3],
 * n hexahedra conn [n][8] (VTK corner order), cell data treeid, level,
 * mpirank (zero) and the nf fields fld [nf][n] */
static int write_pieces(const char *name, int64_t n, int64_t np, const double *pts, const int64_t *conn,
                        const int32_t *itree, const int32_t *ilev, const double *fld, int nf,
                        const charm_config_t *cfg)
{
    int64_t *offs = (int64_t *) malloc((size_t) (n + 1) * sizeof(int64_t));
    uint8_t *types = (uint8_t *) malloc((size_t) (n + 1));
    int32_t *izero = (int32_t *) calloc((size_t) (n + 1), sizeof(int32_t));
    char path[4096];
    FILE *fp = NULL;
    int rc = CHARM_ERR_MEM, k, j;
    uint64_t off = 0;
    int64_t i;
    if (!offs || !types || !izero) { charm_set_error("write_vtk: out of memory"); goto done; }
    for (i = 0; i < n; i++) {
        offs[i] = 8 * (i + 1);
        types[i] = 12;                       /* VTK_HEXAHEDRON */
    }
    /* piece */
    snprintf(path, sizeof(path), "%s_0000.vtu", name);
    fp = fopen(path, "wb");
    if (!fp) { charm_set_error("write_vtk: cannot open %s", path); rc = CHARM_ERR_ARG; goto done; }
    fprintf(fp, "<?xml version=\"1.0\"?>\n"
                "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
                "header_type=\"UInt64\">\n  <UnstructuredGrid>\n"
                "    <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n",
            (long long) np, (long long) n);
    fprintf(fp, "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" "
                "format=\"appended\" offset=\"%llu\"/>\n      </Points>\n", (unsigned long long) off);
    off += 8 + (uint64_t) np * 3 * sizeof(double);
    fprintf(fp, "      <Cells>\n        <DataArray type=\"Int64\" Name=\"connectivity\" "
                "format=\"appended\" offset=\"%llu\"/>\n", (unsigned long long) off);
    off += 8 + (uint64_t) n * 8 * sizeof(int64_t);
    fprintf(fp, "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"appended\" "
                "offset=\"%llu\"/>\n", (unsigned long long) off);
    off += 8 + (uint64_t) n * sizeof(int64_t);
    fprintf(fp, "        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" "
                "offset=\"%llu\"/>\n      </Cells>\n", (unsigned long long) off);
    off += 8 + (uint64_t) n;
    fprintf(fp, "      <CellData Scalars=\"R\">\n");
    {
        static const char *iname[3] = {"treeid", "level", "mpirank"};
        static const char *fname[8] = {"R", "P", "E", "E_TOT", "U", "V", "W", "T"};
        for (k = 0; k < 3; k++) {
            fprintf(fp, "        <DataArray type=\"Int32\" Name=\"%s\" format=\"appended\" "
                        "offset=\"%llu\"/>\n", iname[k], (unsigned long long) off);
            off += 8 + (uint64_t) n * sizeof(int32_t);
        }
        for (j = 0; j < nf; j++) {
            char cn[48];
            if (j < 8) snprintf(cn, sizeof(cn), "%s", fname[j]);
            else if (cfg->comp_name[j - 8][0]) snprintf(cn, sizeof(cn), "C_%.31s", cfg->comp_name[j - 8]);
            else snprintf(cn, sizeof(cn), "C_%d", j - 8);
            fprintf(fp, "        <DataArray type=\"Float64\" Name=\"%s\" format=\"appended\" "
                        "offset=\"%llu\"/>\n", cn, (unsigned long long) off);
            off += 8 + (uint64_t) n * sizeof(double);
        }
    }
    fprintf(fp, "      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n_");
    if (!put_block(fp, pts, (uint64_t) np * 3 * sizeof(double)) ||
        !put_block(fp, conn, (uint64_t) n * 8 * sizeof(int64_t)) ||
        !put_block(fp, offs, (uint64_t) n * sizeof(int64_t)) ||
        !put_block(fp, types, (uint64_t) n) ||
        !put_block(fp, itree ? itree : izero, (uint64_t) n * sizeof(int32_t)) ||
        !put_block(fp, ilev, (uint64_t) n * sizeof(int32_t)) ||
        !put_block(fp, izero, (uint64_t) n * sizeof(int32_t))) {
        charm_set_error("write_vtk: write failed (%s)", path);
        rc = CHARM_ERR_ARG;
        goto done;
    }
    for (j = 0; j < nf; j++)
        if (!put_block(fp, fld + (size_t) j * n, (uint64_t) n * sizeof(double))) {
            charm_set_error("write_vtk: write failed (%s)", path);
            rc = CHARM_ERR_ARG;
            goto done;
        }
    fprintf(fp, "\n  </AppendedData>\n</VTKFile>\n");
    if (fclose(fp) != 0) { fp = NULL; charm_set_error("write_vtk: close failed (%s)", path); rc = CHARM_ERR_ARG; goto done; }
    fp = NULL;

    /* parallel index of the single piece */
    snprintf(path, sizeof(path), "%s.pvtu", name);
    fp = fopen(path, "w");
    if (!fp) { charm_set_error("write_vtk: cannot open %s", path); rc = CHARM_ERR_ARG; goto done; }
    {
        static const char *fname[8] = {"R", "P", "E", "E_TOT", "U", "V", "W", "T"};
        const char *base = strrchr(name, '/');
        base = base ? base + 1 : name;
        fprintf(fp, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"1.0\" "
                    "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
                    "  <PUnstructuredGrid GhostLevel=\"0\">\n"
                    "    <PPoints>\n      <PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n"
                    "    </PPoints>\n    <PCellData Scalars=\"R\">\n"
                    "      <PDataArray type=\"Int32\" Name=\"treeid\"/>\n"
                    "      <PDataArray type=\"Int32\" Name=\"level\"/>\n"
                    "      <PDataArray type=\"Int32\" Name=\"mpirank\"/>\n");
        for (j = 0; j < nf; j++) {
            if (j < 8) fprintf(fp, "      <PDataArray type=\"Float64\" Name=\"%s\"/>\n", fname[j]);
            else if (cfg->comp_name[j - 8][0])
                fprintf(fp, "      <PDataArray type=\"Float64\" Name=\"C_%.31s\"/>\n", cfg->comp_name[j - 8]);
            else fprintf(fp, "      <PDataArray type=\"Float64\" Name=\"C_%d\"/>\n", j - 8);
        }
        fprintf(fp, "    </PCellData>\n    <Piece Source=\"%s_0000.vtu\"/>\n"
                    "  </PUnstructuredGrid>\n</VTKFile>\n", base);
    }
    if (fclose(fp) != 0) { fp = NULL; charm_set_error("write_vtk: close failed (%s)", path); rc = CHARM_ERR_ARG; goto done; }
    fp = NULL;
    rc = CHARM_OK;
done:
    if (fp) fclose(fp);
    free(offs); free(types); free(izero);
    return rc;
}

int charm_write_vtk(const char *name, int64_t n, const int32_t *leaves, const double *u,
                    const charm_config_t *cfg)
{
    const double inv = 1.0 / (double) CHARM_ROOT_LEN;
    const int nc = cfg ? cfg->n_comp : 0, nv = 4 + nc, nf = 8 + nc;
    /* hexahedron corner order (VTK_HEXAHEDRON): x fastest on the bottom
     * face counter-clockwise, then the top face */
    static const int hx[8] = {0, 1, 1, 0, 0, 1, 1, 0}, hy[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                     hz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    uint64_t *keys = NULL, *ukeys = NULL;
    int64_t *conn = NULL, np = 0, i;
    double *pts = NULL, *fld = NULL;
    int32_t *ilev = NULL;
    int rc = CHARM_ERR_MEM, k, j;
    if (!name || !cfg || n < 0 || (n > 0 && (!leaves || !u)) || nc < 1 || nc > CHARM_MAX_COMP) {
        charm_set_error("write_vtk: bad arguments");
        return CHARM_ERR_ARG;
    }
    keys = (uint64_t *) malloc((size_t) (8 * n + 1) * sizeof(uint64_t));
    ukeys = (uint64_t *) malloc((size_t) (8 * n + 1) * sizeof(uint64_t));
    conn = (int64_t *) malloc((size_t) (8 * n + 1) * sizeof(int64_t));
    ilev = (int32_t *) malloc((size_t) (n + 1) * sizeof(int32_t));
    fld = (double *) malloc((size_t) (n * nf + 1) * sizeof(double));
    if (!keys || !ukeys || !conn || !ilev || !fld) {
        charm_set_error("write_vtk: out of memory");
        goto done;
    }
    for (i = 0; i < n; i++) {
        const int32_t *q = leaves + 4 * i;
        const int32_t len = CHARM_ROOT_LEN >> q[3];
        for (k = 0; k < 8; k++)
            keys[8 * i + k] = corner_key(q[0] + hx[k] * len, q[1] + hy[k] * len, q[2] + hz[k] * len);
        ilev[i] = q[3];
    }
    memcpy(ukeys, keys, (size_t) (8 * n) * sizeof(uint64_t));
    qsort(ukeys, (size_t) (8 * n), sizeof(uint64_t), cmp_u64);
    for (i = 0; i < 8 * n; i++)
        if (np == 0 || ukeys[i] != ukeys[np - 1]) ukeys[np++] = ukeys[i];
    for (i = 0; i < 8 * n; i++) conn[i] = find_key(ukeys, np, keys[i]);
    pts = (double *) malloc((size_t) (3 * np + 1) * sizeof(double));
    if (!pts) { charm_set_error("write_vtk: out of memory"); goto done; }
    for (i = 0; i < np; i++) {
        pts[3 * i + 0] = (double) (ukeys[i] >> 42) * inv;
        pts[3 * i + 1] = (double) ((ukeys[i] >> 21) & 0x1FFFFF) * inv;
        pts[3 * i + 2] = (double) (ukeys[i] & 0x1FFFFF) * inv;
    }
    for (i = 0; i < n; i++) {
        double row[8 + CHARM_MAX_COMP];
        cell_fields(cfg, u + (size_t) i * nv * 4, row);
        for (j = 0; j < nf; j++) fld[(size_t) j * n + i] = row[j];
    }

    rc = write_pieces(name, n, np, pts, conn, NULL, ilev, fld, nf, cfg);
done:
    free(keys); free(ukeys); free(conn); free(ilev); free(fld); free(pts);
    return rc;
}

int charm_gmesh_write_vtk(const char *name, const charm_gmesh_t *m, const double *u, const charm_config_t *cfg)
{
    /* VTK_HEXAHEDRON corner k is the z-order corner zk[k] */
    static const int zk[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    const int nc = cfg ? cfg->n_comp : 0, nv = 4 + nc, nf = 8 + nc;
    int64_t n = 0, i;
    double *cor = NULL, *pts = NULL, *fld = NULL;
    int64_t *conn = NULL;
    int32_t *lv = NULL, *itree = NULL, *ilev = NULL;
    int rc = CHARM_ERR_MEM, k, j;
    if (!name || !m || !cfg || !u || nc < 1 || nc > CHARM_MAX_COMP) {
        charm_set_error("gmesh_write_vtk: bad arguments");
        return CHARM_ERR_ARG;
    }
    charm_gmesh_info(m, &n, NULL);
    cor = (double *) malloc((size_t) (24 * n + 1) * sizeof(double));
    pts = (double *) malloc((size_t) (24 * n + 1) * sizeof(double));
    conn = (int64_t *) malloc((size_t) (8 * n + 1) * sizeof(int64_t));
    lv = (int32_t *) malloc((size_t) (5 * n + 1) * sizeof(int32_t));
    itree = (int32_t *) malloc((size_t) (n + 1) * sizeof(int32_t));
    ilev = (int32_t *) malloc((size_t) (n + 1) * sizeof(int32_t));
    fld = (double *) malloc((size_t) (n * nf + 1) * sizeof(double));
    if (!cor || !pts || !conn || !lv || !itree || !ilev || !fld) {
        charm_set_error("gmesh_write_vtk: out of memory");
        goto done;
    }
    if ((rc = charm_gmesh_leaf_corners(m, cor)) || (rc = charm_gmesh_leaves(m, lv))) goto done;
    /* each leaf its own 8 points (the trees' corners are not merged) */
    for (i = 0; i < n; i++) {
        double row[8 + CHARM_MAX_COMP];
        for (k = 0; k < 8; k++) {
            for (j = 0; j < 3; j++) pts[3 * (8 * i + k) + j] = cor[3 * (8 * i + zk[k]) + j];
            conn[8 * i + k] = 8 * i + k;
        }
        itree[i] = lv[5 * i];
        ilev[i] = lv[5 * i + 4];
        cell_fields(cfg, u + (size_t) i * nv * 4, row);
        for (j = 0; j < nf; j++) fld[(size_t) j * n + i] = row[j];
    }
    rc = write_pieces(name, n, 8 * n, pts, conn, itree, ilev, fld, nf, cfg);
done:
    free(cor); free(pts); free(conn); free(lv); free(itree); free(ilev); free(fld);
    return rc;
}

/* src/charm_output.c:56-126: <prefix>_<timestep %08d> of the current state
 * (prefix "charm_3d" = CHARM_STRING, include/charm_def.h:153, unless the
 * configuration names one) */
void charm_write_solution(charm_forest_t *p4est)
{
    charm_ctx_t *c = charm_ctx_of(p4est);
    const int64_t n = charm_forest_num_leaves(p4est);
    int32_t *q = NULL;
    double *u = NULL;
    char name[512];
    int rc;
    if (!c) return;
    q = (int32_t *) malloc((size_t) (n + 1) * 4 * sizeof(int32_t));
    u = (double *) malloc((size_t) (n + 1) * 4 * (4 + c->cfg.n_comp) * sizeof(double));
    if (!q || !u) {
        charm_set_error("write_solution: out of memory");
        c->err = CHARM_ERR_MEM;
    } else if ((rc = charm_forest_get_leaves(p4est, q)) || (rc = charm_gpu_get_state(c->gpu, u, n))) {
        c->err = rc;
    } else {
        snprintf(name, sizeof(name), "%.480s_%08d", c->cfg.out_prefix[0] ? c->cfg.out_prefix : "charm_3d",
                 c->timestep);
        if ((rc = charm_write_vtk(name, n, q, u, &c->cfg))) c->err = rc;
    }
    free(q);
    free(u);
}
