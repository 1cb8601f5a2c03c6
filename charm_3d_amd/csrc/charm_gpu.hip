// charm_gpu.hip -- device context and the charm_gpu_* C-ABI entry points.
//
// A context owns one HIP stream (default priority), the flattened mesh on
// the device and three state buffers B0 (u^n, the state the caller sees),
// B1, B2 (stage results).  Every call returns CHARM_OK or a negative code;
// the message of a failure is kept per host thread (charm_gpu_last_error).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "charm_gpu.h"
#include "charm_internal.h"
#include "charm_kernels.h"

using namespace charm;

extern "C" const char *charm_host_last_error(void);

struct charm_gpu_ctx {
    charm_config_t cfg;
    int prec, nc, nv;
    int sprec;                // storage: 1 FP32 state, 0 FP64 state (FP64 and mixed modes)
    int ph;                   // 1: Navier-Stokes / real-gas kernels (PH = 1)
    size_t esz;
    DevParams P;
    DevParams *d_P;           // device copy of P (kernels that index it at run time)
    hipStream_t st;
    hipEvent_t ev0, ev1;
    float last_ms;
    // mesh
    int64_t n, nch, ncf, nslot, naux;
    int max_faces, chunk_elems;
    int n_xcd;
    uint8_t *d_level;
    int4 *d_cf4, *d_aux;
    int *d_ch_ebeg, *d_ch_fbeg, *d_slot_ptr, *d_nbr;
    uint16_t *d_slots;
    int4 *d_q;                // leaves (x, y, z, level): adaptation
    int32_t *hq;              // host copy of the leaves of the uploaded mesh
    // Navier-Stokes buffers (allocated when ph)
    void *d_auxf;             // [36][n] tau, q coefficients
    void *d_vi;               // [16][n] viscous integrals
    void *d_auxo;             // [36][n] BR2: the own-trace tau, q (allocated with visc_br2_eta > 0)
    void *d_lift;             // [24][28][n] BR2: the (sub)face liftings
    double *d_dcoef;          // [nc][n] diffusion coefficients
    double *d_chem;           // [n] chemistry energy source
    // point EOS closures and transport (allocated when ph and a species has
    // a polynomial cp): [kMemoPlanes][kMemoRows][n] in the state's precision,
    // planes m, cv, gamma, mu, k, rows the own face points then the volume
    // points; temps_of is the state they hold, for the stage in progress
    // (reset after it)
    void *d_tpt;
    const void *temps_of;
    // multi-tree mesh (charm_gpu_upload_gmesh; gm = 1): the general kernels
    int gm;
    int64_t gnf;
    void *d_glg, *d_gfg, *d_gfq, *d_gmb;
    double *d_gfn;
    int *d_gfe, *d_gff, *d_glnb;
    const void *gmb_of;             // the state whose cell means d_gmb holds (multi-tree)
    // state
    void *B[3];
    double *d_stage64;        // FP64 staging for host transfers
    double *d_partial, *d_result;
    int *d_err;
};

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            charm_set_error("%s failed: %s", #x, hipGetErrorString(e_));           \
            return CHARM_ERR_HIP;                                                  \
        }                                                                          \
    } while (0)

extern "C" const char *charm_gpu_last_error(void) { return charm_host_last_error(); }

extern "C" int charm_gpu_device_count(int *n)
{
    if (!n) { charm_set_error("device_count: null argument"); return CHARM_ERR_ARG; }
    HIPCHK(hipGetDeviceCount(n));
    return CHARM_OK;
}

static void free_mesh(charm_gpu_ctx_t *c)
{
    void **p[] = {(void **) &c->d_level, (void **) &c->d_cf4, (void **) &c->d_aux,
                  (void **) &c->d_ch_ebeg, (void **) &c->d_ch_fbeg, (void **) &c->d_slot_ptr,
                  (void **) &c->d_nbr, (void **) &c->d_slots, &c->B[0], &c->B[1], &c->B[2],
                  (void **) &c->d_stage64, (void **) &c->d_partial, (void **) &c->d_q,
                  &c->d_auxf, &c->d_vi, &c->d_auxo, &c->d_lift, (void **) &c->d_dcoef, (void **) &c->d_chem,
                  (void **) &c->d_tpt,
                  &c->d_glg, &c->d_gfg, &c->d_gfq, &c->d_gmb, (void **) &c->d_gfn, (void **) &c->d_gfe,
                  (void **) &c->d_gff, (void **) &c->d_glnb};
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++)
        if (*p[i]) { (void) hipFree(*p[i]); *p[i] = NULL; }
    c->gm = 0; c->gnf = 0; c->gmb_of = NULL; c->temps_of = NULL;
    free(c->hq);
    c->hq = NULL;
    c->n = 0;
}

extern "C" int charm_gpu_create(const charm_config_t *cfg, charm_gpu_ctx_t **out)
{
    if (!cfg || !out) { charm_set_error("gpu_create: null argument"); return CHARM_ERR_ARG; }
    if (cfg->n_comp < 1 || cfg->n_comp > 4) {
        charm_set_error("gpu_create: n_comp %d not in [1, 4]", cfg->n_comp);
        return CHARM_ERR_UNSUPPORTED;
    }
    if (cfg->flux < 0 || cfg->flux > 3 ||
        (cfg->precision != CHARM_FP64 && cfg->precision != CHARM_FP32 && cfg->precision != CHARM_FP32_MIX)) {
        charm_set_error("gpu_create: bad flux %d or precision %d", cfg->flux, cfg->precision);
        return CHARM_ERR_ARG;
    }
    for (int i = 0; i < cfg->n_comp; i++)
        if (cfg->comp[i].n_cp < 1 || cfg->comp[i].n_cp > 8 || !(cfg->comp[i].m > 0)) {
            charm_set_error("gpu_create: species %d needs 1..8 cp coefficients and a positive molar mass", i);
            return CHARM_ERR_ARG;
        }
    if (cfg->n_react < 0 || cfg->n_react > 4) { charm_set_error("gpu_create: %d reactions (at most 4)", cfg->n_react); return CHARM_ERR_ARG; }
    for (int r = 0; r < cfg->n_react; r++)
        for (int m = 0; m < 3; m++) {
            const int l = cfg->react[r].left[m], q = cfg->react[r].right[m];
            if (l < -1 || l >= cfg->n_comp || q < -1 || q >= cfg->n_comp) {
                charm_set_error("gpu_create: reaction %d names species outside [0, %d)", r, cfg->n_comp);
                return CHARM_ERR_ARG;
            }
        }
    for (int i = 0; i < 6; i++) {
        int t = cfg->bnd_type[i];
        if (t != CHARM_BND_PERIODIC && (t < 0 || t > 4)) { charm_set_error("gpu_create: bad boundary type %d", t); return CHARM_ERR_ARG; }
    }
    if (!(cfg->visc_br2_eta >= 0.0 && cfg->visc_br2_eta < 1e6)) {
        charm_set_error("gpu_create: BR2 penalty factor %g (0 for BR1, else positive)", cfg->visc_br2_eta);
        return CHARM_ERR_ARG;
    }
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) { charm_set_error("gpu_create: no HIP device"); return CHARM_ERR_HIP; }
    charm_gpu_ctx_t *c = (charm_gpu_ctx_t *) calloc(1, sizeof(charm_gpu_ctx_t));
    if (!c) { charm_set_error("gpu_create: out of memory"); return CHARM_ERR_MEM; }
    c->cfg = *cfg;
    c->prec = cfg->precision;
    c->sprec = cfg->precision == CHARM_FP32;
    c->nc = cfg->n_comp;
    c->nv = 4 + cfg->n_comp;
    c->esz = c->sprec ? sizeof(float) : sizeof(double);
    // MI355X: 8 XCDs; only the workgroup-to-L2 placement depends on it
    c->n_xcd = 8;
    if (getenv("CHARM_NUM_XCD")) c->n_xcd = atoi(getenv("CHARM_NUM_XCD"));
    if (c->n_xcd < 1 || c->n_xcd > 64) c->n_xcd = 8;
    DevParams &P = c->P;
    memset(&P, 0, sizeof(P));
    P.nc = c->nc;
    P.eos = cfg->eos;
    // ideal gas from component 0 (src/charm_eos.c:113-116)
    const double cp0 = cfg->comp[0].cp[0];
    P.m = cfg->comp[0].m;
    P.cv = cp0 - kGasR / P.m;
    P.gam = cp0 / P.cv;
    for (int i = 0; i < c->nc; i++) {
        P.comp_m[i] = cfg->comp[i].m; P.comp_cp[i] = cfg->comp[i].cp[0];
        P.comp_minv[i] = 1.0 / cfg->comp[i].m;
    }
    for (int f = 0; f < 6; f++) {
        P.bnd_type[f] = cfg->bnd_type[f];
        for (int j = 0; j < kBndParams; j++) P.bnd_param[f][j] = cfg->bnd_param[f][j];
    }
    for (int d = 0; d < 3; d++) P.grav[d] = cfg->grav[d];
    P.cfl = cfg->cfl;
    // Navier-Stokes / real-gas parameters
    P.newton_t = cfg->comp[0].cp_type != 0;                  // src/charm_eos.c:104,127
    int any_poly = 0;
    for (int i = 0; i < c->nc; i++) {
        const charm_comp_desc_t &q = cfg->comp[i];
        // charm_comp_calc_cp reads the power series when cp_type is not
        // constant and ml_type is 1 (src/charm_globals.c:609-613)
        P.cp_poly[i] = q.cp_type != 0 && q.ml_type == 1;
        any_poly |= P.cp_poly[i];
        P.n_cp[i] = q.n_cp;
        for (int k = 0; k < 8; k++) P.cp_c[i][k] = k < q.n_cp ? q.cp[k] : 0.0;
        // its cp is cp[0], while dcp/dT is its series' derivative
        // (src/charm_globals.c:628-649)
        P.cp_dt_rest |= !P.cp_poly[i] && q.n_cp > 1;
        P.ml_type[i] = q.ml_type; P.kp_type[i] = q.kp_type;
        P.ml0[i] = q.ml0; P.kp0[i] = q.kp0; P.t0[i] = q.t0; P.ts[i] = q.ts; P.h0[i] = q.h0;
        P.t0inv[i] = q.t0 != 0.0 ? 1.0 / q.t0 : 0.0;
        P.sig[i] = q.sig; P.ek[i] = q.ek;
        // enthalpy polynomial: hc_k = cp_k / (k+1), hH0 = h0 - sum_k hc_k T_ref^(k+1)
        double hr = 0.0, tp = cfg->t_ref;
        for (int k = 0; k < 8; k++) {
            P.hc[i][k] = k < q.n_cp ? q.cp[k] / (double) (k + 1) : 0.0;
            hr += P.hc[i][k] * tp;
            tp *= cfg->t_ref;
        }
        P.hH0[i] = q.h0 - hr;
    }
    P.t_ref = cfg->t_ref;
    P.use_visc = cfg->use_visc != 0;
    P.use_diff = cfg->use_diff != 0;
    P.n_react = cfg->n_react;
    P.chem_parity = cfg->chem_parity != 0;
    P.chem_energy_molar = cfg->chem_energy_molar != 0;
    P.br2_eta = cfg->visc_br2_eta;
    P.br2_eta_f = (float) cfg->visc_br2_eta;
    for (int r = 0; r < cfg->n_react; r++) {
        for (int m = 0; m < 3; m++) { P.r_left[r][m] = cfg->react[r].left[m]; P.r_right[r][m] = cfg->react[r].right[m]; }
        P.r_a[r] = cfg->react[r].a; P.r_e[r] = cfg->react[r].e; P.r_n[r] = cfg->react[r].n;
    }
    c->ph = (P.newton_t || any_poly || P.use_visc || P.n_react > 0) ? 1 : 0;
    if (c->ph && c->prec == CHARM_FP32_MIX) {
        free(c);
        charm_set_error("gpu_create: the mixed precision mode covers the constant-cp Euler terms only");
        return CHARM_ERR_UNSUPPORTED;
    }
    {
        // FP64 mass-matrix inverse (level 0); the octree's matrices are
        // diagonal, which is checked here
        double a[4][4], ai[4][4];
        if (charm_level_mass_inverse(0, a, ai)) { free(c); return CHARM_ERR_ARG; }
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                if (i != j && fabs(ai[i][j]) > 1e-14 * fabs(ai[i][i])) {
                    charm_set_error("gpu_create: mass matrix inverse is not diagonal");
                    free(c);
                    return CHARM_ERR_ARG;
                }
        for (int k = 0; k < 4; k++) { P.ainv0[k] = ai[k][k]; P.ainv0_f[k] = (float) ai[k][k]; }
    }
    P.gam_f = (float) P.gam; P.cv_f = (float) P.cv; P.m_f = (float) P.m;
    for (int i = 0; i < kMaxComp; i++) { P.comp_m_f[i] = (float) P.comp_m[i]; P.comp_cp_f[i] = (float) P.comp_cp[i]; }
    for (int f = 0; f < 6; f++)
        for (int j = 0; j < kBndParams; j++) P.bnd_param_f[f][j] = (float) P.bnd_param[f][j];
    for (int d = 0; d < 3; d++) P.grav_f[d] = (float) P.grav[d];
    hipError_t e;
    if ((e = hipStreamCreate(&c->st)) != hipSuccess || (e = hipEventCreate(&c->ev0)) != hipSuccess ||
        (e = hipEventCreate(&c->ev1)) != hipSuccess || (e = hipMalloc(&c->d_err, sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&c->d_result, 16 * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&c->d_P, sizeof(DevParams))) != hipSuccess ||
        (e = hipMemcpyAsync(c->d_P, &c->P, sizeof(DevParams), hipMemcpyHostToDevice, c->st)) != hipSuccess ||
        (e = hipMemsetAsync(c->d_err, 0, sizeof(int), c->st)) != hipSuccess ||
        (e = hipStreamSynchronize(c->st)) != hipSuccess) {
        charm_set_error("gpu_create: %s", hipGetErrorString(e));
        charm_gpu_destroy(c);
        return CHARM_ERR_HIP;
    }
    *out = c;
    return CHARM_OK;
}

extern "C" int charm_gpu_destroy(charm_gpu_ctx_t *c)
{
    if (!c) return CHARM_OK;
    if (c->st) (void) hipStreamSynchronize(c->st);
    free_mesh(c);
    if (c->d_err) (void) hipFree(c->d_err);
    if (c->d_result) (void) hipFree(c->d_result);
    if (c->d_P) (void) hipFree(c->d_P);
    if (c->ev0) (void) hipEventDestroy(c->ev0);
    if (c->ev1) (void) hipEventDestroy(c->ev1);
    if (c->st) (void) hipStreamDestroy(c->st);
    free(c);
    return CHARM_OK;
}

#define DMALLOC(ptr, bytes)                                                         \
    do {                                                                            \
        size_t b_ = (bytes) ? (bytes) : 16;                                         \
        hipError_t e_ = hipMalloc((void **) &(ptr), b_);                            \
        if (e_ != hipSuccess) {                                                     \
            charm_set_error("upload_mesh: hipMalloc(%zu) failed: %s", b_, hipGetErrorString(e_)); \
            free_mesh(c);                                                           \
            return CHARM_ERR_MEM;                                                   \
        }                                                                           \
    } while (0)

extern "C" int charm_gpu_upload_mesh(charm_gpu_ctx_t *c, const charm_mesh_t *m)
{
    if (!c || !m) { charm_set_error("upload_mesh: null argument"); return CHARM_ERR_ARG; }
    if (m->chunk_elems < 1 || m->chunk_elems > 256) {
        charm_set_error("upload_mesh: chunk_elems %d outside [1, 256] (one thread per element)", m->chunk_elems);
        return CHARM_ERR_ARG;
    }
    if (4 * (int64_t) m->n * (int64_t) c->esz >= ((int64_t) 1 << 31)) {
        // the stage kernel addresses one variable's 4 coefficient planes with
        // 32-bit byte offsets
        charm_set_error("upload_mesh: %lld leaves: one variable of the state exceeds 2^31 bytes",
                        (long long) m->n);
        return CHARM_ERR_ARG;
    }
    const size_t smem = stage_smem_bytes(c->prec, c->nc, m->chunk_faces);
    if (smem > 65536) {
        charm_set_error("upload_mesh: chunk_faces %d needs %zu B of LDS per workgroup (limit 65536)",
                        m->chunk_faces, smem);
        return CHARM_ERR_ARG;
    }
    for (int a = 0; a < 3; a++) {
        const bool per = c->cfg.bnd_type[2 * a] == CHARM_BND_PERIODIC;
        if (per != (m->periodic[a] != 0) || per != (c->cfg.bnd_type[2 * a + 1] == CHARM_BND_PERIODIC)) {
            charm_set_error("upload_mesh: periodic links of the forest do not match the boundary types");
            return CHARM_ERR_ARG;
        }
    }
    HIPCHK(hipStreamSynchronize(c->st));
    free_mesh(c);
    const size_t ssz = (size_t) m->n * 4 * c->nv * c->esz;
    DMALLOC(c->d_level, (size_t) m->n);
    DMALLOC(c->d_cf4, (size_t) m->ncf * sizeof(int4));
    DMALLOC(c->d_ch_ebeg, (size_t) (m->nch + 1) * sizeof(int));
    DMALLOC(c->d_ch_fbeg, (size_t) (m->nch + 1) * sizeof(int));
    DMALLOC(c->d_slot_ptr, (size_t) (m->n + 1) * sizeof(int));
    DMALLOC(c->d_slots, (size_t) m->nslot * sizeof(uint16_t));
    DMALLOC(c->d_nbr, (size_t) m->n * 6 * sizeof(int));
    DMALLOC(c->d_aux, (size_t) m->naux * sizeof(int4));
    for (int i = 0; i < 3; i++) DMALLOC(c->B[i], ssz);
    // FP64 staging: the state, or the 36 auxiliary coefficients of the
    // Navier-Stokes test hook, whichever is larger
    DMALLOC(c->d_stage64, (size_t) m->n * (4 * c->nv > 36 || !c->ph ? 4 * c->nv : 36) * sizeof(double));
    DMALLOC(c->d_partial, (size_t) ((m->n + 255) / 256 + 1) * (c->nv + 1) * sizeof(double));
    DMALLOC(c->d_q, (size_t) m->n * sizeof(int4));
    if (c->ph) {
        DMALLOC(c->d_auxf, (size_t) m->n * 36 * c->esz);
        DMALLOC(c->d_vi, (size_t) m->n * 16 * c->esz);
        if (c->cfg.use_visc && c->cfg.visc_br2_eta > 0.0) {
            DMALLOC(c->d_auxo, (size_t) m->n * 36 * c->esz);
            DMALLOC(c->d_lift, (size_t) m->n * kLiftSlots * kLiftFields * c->esz);
        }
        DMALLOC(c->d_dcoef, (size_t) m->n * c->nc * sizeof(double));
        DMALLOC(c->d_chem, (size_t) m->n * sizeof(double));
        if (c->P.newton_t) {
            DMALLOC(c->d_tpt, (size_t) m->n * kMemoPlanes * kMemoRows * c->esz);
        }
    }
    c->hq = (int32_t *) malloc((size_t) (m->n ? m->n : 1) * 4 * sizeof(int32_t));
    if (!c->hq) { charm_set_error("upload_mesh: out of memory"); free_mesh(c); return CHARM_ERR_MEM; }
    memcpy(c->hq, m->q, (size_t) m->n * 4 * sizeof(int32_t));
    HIPCHK(hipMemcpyAsync(c->d_q, m->q, (size_t) m->n * sizeof(int4), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_level, m->level, (size_t) m->n, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_cf4, m->cf4, (size_t) m->ncf * sizeof(int4), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_ch_ebeg, m->ch_ebeg, (size_t) (m->nch + 1) * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_ch_fbeg, m->ch_fbeg, (size_t) (m->nch + 1) * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_slot_ptr, m->slot_ptr, (size_t) (m->n + 1) * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_slots, m->slots, (size_t) m->nslot * sizeof(uint16_t), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_nbr, m->nbr, (size_t) m->n * 6 * sizeof(int), hipMemcpyHostToDevice, c->st));
    if (m->naux)
        HIPCHK(hipMemcpyAsync(c->d_aux, m->aux, (size_t) m->naux * sizeof(int4), hipMemcpyHostToDevice, c->st));
    for (int i = 0; i < 3; i++) HIPCHK(hipMemsetAsync(c->B[i], 0, ssz, c->st));
    if (c->ph) {
        HIPCHK(hipMemsetAsync(c->d_dcoef, 0, (size_t) m->n * c->nc * sizeof(double), c->st));
        HIPCHK(hipMemsetAsync(c->d_chem, 0, (size_t) m->n * sizeof(double), c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->n = m->n; c->nch = m->nch; c->ncf = m->ncf; c->nslot = m->nslot; c->naux = m->naux;
    c->max_faces = m->chunk_faces;
    c->chunk_elems = m->chunk_elems;
    return CHARM_OK;
}

// device copy of a FP64 host array in the state precision
static hipError_t upload_real(charm_gpu_ctx_t *c, void *dst, const double *src, size_t count)
{
    if (!c->sprec) return hipMemcpy(dst, src, count * sizeof(double), hipMemcpyHostToDevice);
    float *tmp = (float *) malloc(count * sizeof(float) + 1);
    if (!tmp) return hipErrorOutOfMemory;
    for (size_t i = 0; i < count; i++) tmp[i] = (float) src[i];
    const hipError_t e = hipMemcpy(dst, tmp, count * sizeof(float), hipMemcpyHostToDevice);
    free(tmp);
    return e;
}

extern "C" int charm_gpu_upload_gmesh(charm_gpu_ctx_t *c, const charm_gmesh_t *m)
{
    charm_gmesh_arrays_t g;
    if (!c || !m) { charm_set_error("upload_gmesh: null argument"); return CHARM_ERR_ARG; }
    if (c->ph || c->prec == CHARM_FP32_MIX || c->cfg.use_visc || c->cfg.n_react > 0) {
        charm_set_error("upload_gmesh: multi-tree meshes run the constant-cp Euler terms without viscosity or "
                        "chemistry, in FP32 or FP64");
        return CHARM_ERR_UNSUPPORTED;
    }
    int rc = charm_gmesh_arrays(m, &g);
    if (rc) return rc;
    for (int64_t i = 0; i < g.nf; i++)
        if (g.fe[g.nf + i] < 0 && (-1 - g.fe[g.nf + i] > 5 || c->cfg.bnd_type[-1 - g.fe[g.nf + i]] == CHARM_BND_PERIODIC)) {
            charm_set_error("upload_gmesh: boundary slot %d is not a boundary type of the configuration",
                            -1 - g.fe[g.nf + i]);
            return CHARM_ERR_ARG;
        }
    HIPCHK(hipStreamSynchronize(c->st));
    free_mesh(c);
    const int64_t n = g.n, nf = g.nf;
    const size_t ssz = (size_t) n * 4 * c->nv * c->esz;
    for (int i = 0; i < 3; i++) DMALLOC(c->B[i], ssz);
    DMALLOC(c->d_stage64, (size_t) n * 4 * c->nv * sizeof(double));
    DMALLOC(c->d_partial, (size_t) ((n + 255) / 256 + 1) * (c->nv + 1) * sizeof(double));
    DMALLOC(c->d_glg, (size_t) n * CHARM_GM_LEAFQ * c->esz);
    DMALLOC(c->d_gfg, (size_t) nf * CHARM_GM_FACEQ * c->esz);
    DMALLOC(c->d_gfq, (size_t) n * 6 * 4 * c->nv * c->esz);
    DMALLOC(c->d_gmb, (size_t) n * c->nv * c->esz);
    DMALLOC(c->d_gfn, (size_t) nf * 3 * sizeof(double));
    DMALLOC(c->d_gfe, (size_t) nf * 2 * sizeof(int));
    DMALLOC(c->d_gff, (size_t) nf * 2 * sizeof(int));
    DMALLOC(c->d_glnb, (size_t) n * 6 * sizeof(int));
    hipError_t e = upload_real(c, c->d_glg, g.lg, (size_t) n * CHARM_GM_LEAFQ);
    if (e == hipSuccess) e = upload_real(c, c->d_gfg, g.fg, (size_t) nf * CHARM_GM_FACEQ);
    if (e == hipSuccess) e = hipMemcpy(c->d_gfn, g.fn, (size_t) nf * 3 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_gfe, g.fe, (size_t) nf * 2 * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_glnb, g.lnb, (size_t) n * 6 * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_gff, g.ff, (size_t) nf * 2 * sizeof(int), hipMemcpyHostToDevice);
    for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipMemsetAsync(c->B[i], 0, ssz, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) {
        charm_set_error("upload_gmesh: %s", hipGetErrorString(e));
        free_mesh(c);
        return CHARM_ERR_HIP;
    }
    c->n = n; c->gnf = nf; c->gm = 1;
    c->nch = 0; c->ncf = 0; c->nslot = 0; c->naux = 0; c->max_faces = 0; c->chunk_elems = 0;
    return CHARM_OK;
}

// the entry points that need the cube octree's arrays
static int need_octree(const charm_gpu_ctx_t *c, const char *what)
{
    if (c && c->gm) {
        charm_set_error("%s: not available on a multi-tree mesh (charm_gpu_upload_gmesh)", what);
        return CHARM_ERR_UNSUPPORTED;
    }
    return CHARM_OK;
}

static int need_mesh(const charm_gpu_ctx_t *c, int64_t n)
{
    if (!c) { charm_set_error("null context"); return CHARM_ERR_ARG; }
    if (c->n == 0) { charm_set_error("no mesh uploaded"); return CHARM_ERR_ARG; }
    if (n >= 0 && n != c->n) { charm_set_error("element count %lld != mesh %lld", (long long) n, (long long) c->n); return CHARM_ERR_ARG; }
    return CHARM_OK;
}

extern "C" int charm_gpu_set_state(charm_gpu_ctx_t *c, const double *u, int64_t n)
{
    int rc = need_mesh(c, n);
    if (rc) return rc;
    if (!u) { charm_set_error("set_state: null state"); return CHARM_ERR_ARG; }
    const int64_t cnt = n * 4 * c->nv;
    HIPCHK(hipMemcpyAsync(c->d_stage64, u, (size_t) cnt * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(launch_convert(c->sprec, 0, c->d_stage64, c->B[0], n, 4 * c->nv, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return CHARM_OK;
}

static int copy_out(charm_gpu_ctx_t *c, const void *src, double *u, int64_t n)
{
    const int64_t cnt = n * 4 * c->nv;
    HIPCHK(launch_convert(c->sprec, 1, src, c->d_stage64, n, 4 * c->nv, c->st));
    HIPCHK(hipMemcpyAsync(u, c->d_stage64, (size_t) cnt * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return CHARM_OK;
}

extern "C" int charm_gpu_get_state(charm_gpu_ctx_t *c, double *u, int64_t n)
{
    int rc = need_mesh(c, n);
    if (rc) return rc;
    if (!u) { charm_set_error("get_state: null state"); return CHARM_ERR_ARG; }
    return copy_out(c, c->B[0], u, n);
}

static StageArgs stage_args(charm_gpu_ctx_t *c, const void *U, const void *U0, void *out, int mode,
                            int combine, double dt)
{
    StageArgs A;
    A.U = U; A.U0 = U0; A.out = out;
    A.level = c->d_level; A.cf4 = c->d_cf4; A.ch_ebeg = c->d_ch_ebeg; A.ch_fbeg = c->d_ch_fbeg;
    A.slot_ptr = c->d_slot_ptr; A.slots = c->d_slots; A.err = c->d_err;
    A.n_elem = c->n; A.n_chunks = (int) c->nch; A.max_faces = c->max_faces; A.chunk_elems = c->chunk_elems;
    A.n_xcd = c->n_xcd;
    A.mode = mode; A.combine = combine; A.dt = dt;
    A.visc = (c->ph && c->cfg.use_visc) ? c->d_vi : NULL;
    A.prm = c->d_P;
    A.tpt = NULL;
    return A;
}

static LimArgs lim_args(charm_gpu_ctx_t *c, const void *W, const void *U0, void *out, int limit, int combine)
{
    LimArgs A;
    A.W = W; A.U0 = U0; A.out = out; A.nbr = c->d_nbr; A.aux = c->d_aux;
    A.n = c->n; A.limit = limit; A.combine = combine; A.n_xcd = c->n_xcd;
    return A;
}

static int check_flag(charm_gpu_ctx_t *c)
{
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (flag) {
        HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(int), c->st));
        charm_set_error("HLLC: inequality SL <= S* <= SR is FALSE (NaN or invalid state)");
        return CHARM_ERR_STATE;
    }
    return CHARM_OK;
}

static NsArgs ns_args(charm_gpu_ctx_t *c, const void *U, void *out, double dt)
{
    NsArgs A;
    A.U = U; A.out = out; A.level = c->d_level; A.nbr = c->d_nbr; A.aux = c->d_aux;
    A.chem_rhs = c->d_chem;
    A.dcoef = (c->cfg.use_diff && c->nc >= 2) ? c->d_dcoef : NULL;
    A.aux_in = c->d_auxf; A.aux_out = c->d_auxf; A.visc_out = c->d_vi; A.aux_own = c->d_auxo;
    A.lift = c->d_lift;
    const bool fresh = c->d_tpt && c->temps_of == U;
    A.tpt = fresh ? c->d_tpt : NULL;
    A.n = c->n; A.dt = dt;
    A.prm = c->d_P;
    return A;
}

// the point EOS closures of the state U (launch code 6), for the viscous
// passes and the stage that read U next; a no-op without a polynomial cp
static hipError_t ns_temps(charm_gpu_ctx_t *c, const void *U)
{
    if (!c->d_tpt || c->temps_of == U) return hipSuccess;
    NsArgs A = ns_args(c, U, NULL, 0.0);
    A.tpt = c->d_tpt;
    hipError_t e = launch_ns(c->sprec, 6, c->nc, A, c->P, c->st);
    c->temps_of = e == hipSuccess ? U : NULL;
    return e;
}

// The viscous passes on the state A.U: diffusion coefficients (DIFFUSION and
// two or more species), the auxiliary tau / q, then the viscous integrals.
// Launch codes of launch_ns: 1 diffusion; 2 / 4 auxiliary field (BR1 / BR2,
// which also writes each (sub)face's lifting and the own-trace field); 3 / 5
// viscous integrals (BR1 / BR2 surface flux).
static hipError_t ns_viscous(charm_gpu_ctx_t *c, const NsArgs &A)
{
    hipError_t e = hipSuccess;
    const bool br2 = c->cfg.visc_br2_eta > 0.0;
    if (A.dcoef && (e = launch_ns(c->sprec, 1, c->nc, A, c->P, c->st)) != hipSuccess) return e;
    if ((e = launch_ns(c->sprec, br2 ? 4 : 2, c->nc, A, c->P, c->st)) != hipSuccess) return e;
    return launch_ns(c->sprec, br2 ? 5 : 3, c->nc, A, c->P, c->st);
}

// The Navier-Stokes pieces in front of a stage (src/models/ns/charm_model.c:
// 226-238): chemistry on `in` into `chem_out` (may be `in`; the energy source
// goes to d_chem), then with viscosity the diffusion coefficients, the
// auxiliary tau / q and the viscous integrals the stage adds.  Returns the
// state the stage reads.
static hipError_t ns_prepare(charm_gpu_ctx_t *c, const void *in, void *chem_out, double dt, const void **stage_in)
{
    hipError_t e = hipSuccess;
    *stage_in = in;
    if (c->cfg.model != CHARM_MODEL_NS || !c->ph) return hipSuccess;
    if (c->cfg.n_react > 0) {
        NsArgs A = ns_args(c, in, chem_out, dt);
        if ((e = launch_ns(c->sprec, 0, c->nc, A, c->P, c->st)) != hipSuccess) return e;
        *stage_in = chem_out;
    }
    if (c->cfg.use_visc) {
        if ((e = ns_temps(c, *stage_in)) != hipSuccess) return e;
        NsArgs A = ns_args(c, *stage_in, NULL, dt);
        if ((e = ns_viscous(c, A)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// one stage / one limiter pass on the context's mesh: the chunked octree
// kernels, or the multi-tree face + element passes
static GArgs gargs(charm_gpu_ctx_t *c, const void *U, const void *U0, void *out, int mode, int combine, double dt)
{
    GArgs G;
    G.U = U; G.U0 = U0; G.out = out; G.fq = c->d_gfq; G.mb = NULL;
    G.lg = c->d_glg; G.fg = c->d_gfg; G.fn = c->d_gfn;
    G.fe = c->d_gfe; G.ff = c->d_gff; G.lnb = c->d_glnb;
    G.err = c->d_err; G.n = c->n; G.nf = c->gnf;
    G.mode = mode; G.combine = combine; G.dt = dt;
    G.n_xcd = c->n_xcd;
    return G;
}

static hipError_t stage_launch(charm_gpu_ctx_t *c, const StageArgs &A)
{
    if (!c->gm) {
        if (!c->d_tpt) return launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, c->st);
        // PH = 1 with a polynomial cp: the stage reads the point EOS closures
        // of its state (computed here unless the viscous passes just did);
        // they are valid for this stage only, as the buffers are rewritten
        StageArgs B = A;
        hipError_t e = ns_temps(c, A.U);
        if (e == hipSuccess) {
            B.tpt = c->d_tpt;
            e = launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, B, c->P, c->st);
        }
        c->temps_of = NULL;
        return e;
    }
    GArgs G = gargs(c, A.U, A.U0, A.out, A.mode, A.combine, A.dt);
    // a stage that a limiter may follow (new state, no combine) also writes
    // the limiter's cell means of that state
    const bool means = A.mode != 0 && A.combine == 0;
    G.mb = means ? c->d_gmb : NULL;
    c->gmb_of = means ? A.out : NULL;
    return launch_gstage(c->sprec, c->nc, c->cfg.flux, G, c->P, c->st);
}

static hipError_t limiter_launch(charm_gpu_ctx_t *c, const LimArgs &L)
{
    if (!c->gm) return launch_limiter(c->sprec, c->nc, c->ph, L, c->P, c->st);
    if (!L.limit) return hipErrorInvalidValue;      // every multi-tree call limits
    GArgs G = gargs(c, L.W, L.U0, L.out, 1, L.combine, 0.0);
    G.mb = c->gmb_of == L.W ? c->d_gmb : NULL;
    return launch_glimiter(c->sprec, c->nc, G, c->P, c->st);
}

extern "C" int charm_gpu_residual(charm_gpu_ctx_t *c, double *out, int64_t n)
{
    int rc = need_mesh(c, n);
    if (rc) return rc;
    if (c->ph && c->cfg.model == CHARM_MODEL_NS && c->cfg.use_visc) {
        // viscous passes on the current state, without the chemistry
        HIPCHK(ns_temps(c, c->B[0]));
        HIPCHK(ns_viscous(c, ns_args(c, c->B[0], NULL, 0.0)));
    }
    StageArgs A = stage_args(c, c->B[0], c->B[0], c->B[1], 0, 0, 0.0);
    HIPCHK(stage_launch(c, A));
    rc = check_flag(c);
    if (rc) return rc;
    return copy_out(c, c->B[1], out, n);
}

extern "C" int charm_gpu_limiter_bj(charm_gpu_ctx_t *c)
{
    int rc = need_mesh(c, -1);
    if (rc) return rc;
    if ((rc = need_octree(c, "limiter_bj"))) return rc;
    LimArgs A = lim_args(c, c->B[0], c->B[0], c->B[1], 1, 0);
    HIPCHK(launch_limiter(c->sprec, c->nc, c->ph, A, c->P, c->st));
    void *t = c->B[0]; c->B[0] = c->B[1]; c->B[1] = t;
    HIPCHK(hipStreamSynchronize(c->st));
    return CHARM_OK;
}

// one time step of the configured model, state in and out of B[0]
static int step_once(charm_gpu_ctx_t *c, double dt)
{
    const int lim = c->cfg.limiter;
    void *B0 = c->B[0], *B1 = c->B[1], *B2 = c->B[2];
    if (c->cfg.model == CHARM_MODEL_EULER) {
        // forward Euler + limiter, src/models/euler/charm_model.c:519-535
        StageArgs A = stage_args(c, B0, B0, B1, 1, 0, dt);
        HIPCHK(stage_launch(c, A));
        if (lim) {
            LimArgs L = lim_args(c, B1, B0, B0, 1, 0);
            HIPCHK(limiter_launch(c, L));
        } else {
            c->B[0] = B1; c->B[1] = B0;
        }
        return CHARM_OK;
    }
    // SSP-RK3 with the chemistry and viscous passes in front of each stage,
    // the limiter inside it and the combines after it,
    // src/models/ns/charm_model.c:226-266.  The first stage's chemistry
    // writes a copy (B2): u^n in B0 stays the combine's base.
    const void *X;
    if (lim) {
        HIPCHK(ns_prepare(c, B0, B2, dt, &X));
        StageArgs A = stage_args(c, X, B0, B1, 1, 0, dt);
        HIPCHK(stage_launch(c, A));
        LimArgs L = lim_args(c, B1, B0, B2, 1, 0);
        HIPCHK(limiter_launch(c, L));
        HIPCHK(ns_prepare(c, B2, B2, dt, &X));
        A = stage_args(c, X, B0, B1, 1, 0, dt);
        HIPCHK(stage_launch(c, A));
        L = lim_args(c, B1, B0, B2, 1, 1);
        HIPCHK(limiter_launch(c, L));
        HIPCHK(ns_prepare(c, B2, B2, dt, &X));
        A = stage_args(c, X, B0, B1, 1, 0, dt);
        HIPCHK(stage_launch(c, A));
        L = lim_args(c, B1, B0, B0, 1, 2);
        HIPCHK(limiter_launch(c, L));
    } else {
        HIPCHK(ns_prepare(c, B0, B2, dt, &X));
        StageArgs A = stage_args(c, X, B0, B1, 1, 0, dt);
        HIPCHK(stage_launch(c, A));
        HIPCHK(ns_prepare(c, B1, B1, dt, &X));
        A = stage_args(c, X, B0, B2, 1, 1, dt);
        HIPCHK(stage_launch(c, A));
        HIPCHK(ns_prepare(c, B2, B2, dt, &X));
        A = stage_args(c, X, B0, B0, 1, 2, dt);
        HIPCHK(stage_launch(c, A));
    }
    return CHARM_OK;
}

extern "C" int charm_gpu_step(charm_gpu_ctx_t *c, int model, double dt, int nsteps)
{
    int rc = need_mesh(c, -1);
    if (rc) return rc;
    if (nsteps < 0 || !(dt >= 0) || (model != CHARM_MODEL_EULER && model != CHARM_MODEL_NS)) {
        charm_set_error("step: bad model, dt or step count");
        return CHARM_ERR_ARG;
    }
    const int keep = c->cfg.model;
    c->cfg.model = model;
    HIPCHK(hipEventRecord(c->ev0, c->st));
    for (int s = 0; s < nsteps && !rc; s++) rc = step_once(c, dt);
    c->cfg.model = keep;
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->st));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    return check_flag(c);
}

extern "C" int charm_gpu_run(charm_gpu_ctx_t *c, double dt, int nsteps)
{
    if (!c) { charm_set_error("null context"); return CHARM_ERR_ARG; }
    return charm_gpu_step(c, c->cfg.model, dt, nsteps);
}

// physics variant of a configuration: 1 when any cp is not constant, or
// viscosity or reactions are on (the Navier-Stokes stage kernels)
static int config_physics(const charm_config_t *cfg)
{
    int ph = cfg->comp[0].cp_type != 0 || cfg->use_visc != 0 || cfg->n_react > 0;
    for (int i = 0; i < cfg->n_comp; i++) ph |= cfg->comp[i].cp_type != 0 && cfg->comp[i].ml_type == 1;
    return ph ? 1 : 0;
}

// Workgroups per CU that the stage kernel's registers allow (4 SIMDs of 512
// VGPRs per lane, at most 8 waves each), or 0 when it cannot be queried
static int stage_groups_per_cu(const charm_config_t *cfg, int bs)
{
    hipFuncAttributes at;
    const void *fn = stage_kernel(cfg->precision, cfg->n_comp, cfg->flux, config_physics(cfg), bs);
    if (!fn || hipFuncGetAttributes(&at, fn) != hipSuccess || at.numRegs <= 0) {
        (void) hipGetLastError();
        return 0;
    }
    int waves = 512 / (((at.numRegs + 7) / 8) * 8);
    if (waves > 8) waves = 8;
    return waves * 4 / (bs / 64);
}

extern "C" int charm_gpu_recommended_chunks(const charm_config_t *cfg, int *ce, int *cf)
{
    if (!cfg || !ce || !cf || cfg->n_comp < 1 || cfg->n_comp > 4) {
        charm_set_error("recommended_chunks: bad arguments");
        return CHARM_ERR_ARG;
    }
    // Chunk LDS sized to the occupancy the kernel's registers allow: the
    // CU's LDS divided among that many workgroups (larger chunks evaluate
    // fewer faces twice; more workgroups than the registers allow gain
    // nothing).  Without a device to query: 40 KB, four workgroups of 160 KB.
    size_t lds_cu = 160 * 1024;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.maxSharedMemoryPerMultiProcessor > 0)
        lds_cu = prop.maxSharedMemoryPerMultiProcessor;
    else
        (void) hipGetLastError();
    const size_t per_face = stage_smem_bytes(cfg->precision, cfg->n_comp, 1);
    int faces = 0;
    for (int bs = 256; bs >= 128; bs -= 128) {
        const int groups = stage_groups_per_cu(cfg, bs);
        size_t budget = groups > 0 ? lds_cu / groups : 40960;
        if (budget > 65536) budget = 65536;
        // whole KiB: LDS is allocated in blocks, and a share that rounds
        // up past lds_cu / groups leaves room for one workgroup fewer
        budget &= ~(size_t) 1023;
        faces = (int) (budget / per_face);
        if (faces > 1024) faces = 1024;
        // at most three passes of the workgroup over the faces (a fourth,
        // partial pass runs with most lanes idle)
        if (faces > 3 * bs) faces = 3 * bs;
        // the FP32 stage adjoint runs on the same chunks with wider LDS
        // records: keep two of its workgroups per CU (one halves the
        // adjoint's occupancy; the forward stage time does not change
        // between 640 and 768 faces)
        if (cfg->precision == CHARM_FP32) {
            const int adj_faces = (int) (lds_cu / 2 / adj_chunk_smem_bytes(1, cfg->n_comp, 1));
            if (faces > adj_faces) faces = adj_faces;
        }
        // 128-thread workgroups serve chunks of at most 128 elements
        if (bs == 128 || faces / 3 > 128 || config_physics(cfg)) break;
    }
    if (faces < 96) faces = 96;
    *cf = faces;
    *ce = faces / 3 > 256 ? 256 : faces / 3;
    return CHARM_OK;
}

extern "C" int charm_gpu_time_copy(charm_gpu_ctx_t *c, int reps, double *copy_ms, double *bytes)
{
    int rc = need_mesh(c, -1);
    if (rc) return rc;
    if (reps < 1) { charm_set_error("time_copy: reps < 1"); return CHARM_ERR_ARG; }
    const size_t ssz = (size_t) c->n * 4 * c->nv * c->esz;
    float ms = 0.f;
    HIPCHK(hipEventRecord(c->ev0, c->st));
    for (int r = 0; r < reps; r++) HIPCHK(launch_copy(c->B[0], c->B[2], ssz, c->st));
    HIPCHK(hipEventRecord(c->ev1, c->st));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (copy_ms) *copy_ms = ms / reps;
    if (bytes) *bytes = 2.0 * (double) ssz;
    return CHARM_OK;
}

extern "C" int charm_gpu_time_kernels(charm_gpu_ctx_t *c, double dt, int reps, double *stage_ms,
                                      double *limiter_ms)
{
    int rc = need_mesh(c, -1);
    if (rc) return rc;
    if (reps < 1) { charm_set_error("time_kernels: reps < 1"); return CHARM_ERR_ARG; }
    float ms = 0.f;
    StageArgs A = stage_args(c, c->B[0], c->B[0], c->B[1], 1, 0, dt);
    HIPCHK(hipEventRecord(c->ev0, c->st));
    for (int r = 0; r < reps; r++) HIPCHK(stage_launch(c, A));
    HIPCHK(hipEventRecord(c->ev1, c->st));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stage_ms) *stage_ms = ms / reps;
    LimArgs L = lim_args(c, c->B[1], c->B[0], c->B[2], 1, 1);
    HIPCHK(hipEventRecord(c->ev0, c->st));
    for (int r = 0; r < reps; r++) HIPCHK(limiter_launch(c, L));
    HIPCHK(hipEventRecord(c->ev1, c->st));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (limiter_ms) *limiter_ms = ms / reps;
    return check_flag(c);
}

extern "C" int charm_gpu_hllc_branch_flips(charm_gpu_ctx_t *a, charm_gpu_ctx_t *b, int64_t *points,
                                           int64_t *flips, int64_t *per_bit)
{
    int rc = need_mesh(a, -1);
    if (rc) return rc;
    if ((rc = need_mesh(b, -1))) return rc;
    if ((rc = need_octree(a, "hllc_branch_flips")) || (rc = need_octree(b, "hllc_branch_flips"))) return rc;
    if (!points || !flips) { charm_set_error("hllc_branch_flips: null output"); return CHARM_ERR_ARG; }
    if (a->prec != CHARM_FP64 || b->prec != CHARM_FP64 || a->ph || b->ph || a->nc != b->nc || a->n != b->n ||
        a->nch != b->nch || a->ncf != b->ncf) {
        charm_set_error("hllc_branch_flips: two FP64 constant-cp contexts on the same mesh");
        return CHARM_ERR_ARG;
    }
    unsigned long long host[2 + kHllcBits] = {0}, *dev = NULL;
    HIPCHK(hipStreamSynchronize(b->st));
    HIPCHK(hipMalloc((void **) &dev, sizeof(host)));
    FlipArgs F;
    F.Ua = (const double *) a->B[0]; F.Ub = (const double *) b->B[0];
    F.cf4 = a->d_cf4; F.ch_ebeg = a->d_ch_ebeg; F.ch_fbeg = a->d_ch_fbeg;
    F.n = a->n; F.n_chunks = (int) a->nch; F.out = dev;
    hipError_t e = hipMemcpyAsync(dev, host, sizeof(host), hipMemcpyHostToDevice, a->st);
    if (e == hipSuccess) e = launch_hllc_flips(a->nc, F, a->P, a->st);
    if (e == hipSuccess) e = hipMemcpyAsync(host, dev, sizeof(host), hipMemcpyDeviceToHost, a->st);
    if (e == hipSuccess) e = hipStreamSynchronize(a->st);
    (void) hipFree(dev);
    if (e != hipSuccess) { charm_set_error("hllc_branch_flips: %s", hipGetErrorString(e)); return CHARM_ERR_HIP; }
    *points = (int64_t) host[0];
    *flips = (int64_t) host[1];
    if (per_bit)
        for (int i = 0; i < kHllcBits; i++) per_bit[i] = (int64_t) host[2 + i];
    return CHARM_OK;
}

extern "C" int charm_gpu_last_timing(charm_gpu_ctx_t *c, double *stage_ms, double *limiter_ms)
{
    if (!c) { charm_set_error("null context"); return CHARM_ERR_ARG; }
    if (stage_ms) *stage_ms = c->last_ms;
    if (limiter_ms) *limiter_ms = 0.0;
    return CHARM_OK;
}

static int reduce(charm_gpu_ctx_t *c, int mode, double *host)
{
    RedArgs A;
    A.U = c->B[0]; A.level = c->d_level; A.partial = c->d_partial; A.result = c->d_result;
    A.n = c->n; A.mode = mode;
    HIPCHK(launch_reduce(c->sprec, c->nc, c->ph, A, c->P, c->st));
    HIPCHK(hipMemcpyAsync(host, c->d_result, (size_t) (c->nv + 1) * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return CHARM_OK;
}

extern "C" int charm_gpu_get_dt(charm_gpu_ctx_t *c, int mode, double *dt)
{
    int rc = need_mesh(c, -1);
    if (rc) return rc;
    if ((rc = need_octree(c, "get_dt"))) return rc;
    if (!dt || (mode != 0 && mode != 1)) { charm_set_error("get_dt: bad arguments"); return CHARM_ERR_ARG; }
    double r[16];
    if ((rc = reduce(c, mode, r))) return rc;
    *dt = r[c->nv];
    return CHARM_OK;
}

extern "C" int charm_gpu_conservation_sums(charm_gpu_ctx_t *c, double *out)
{
    int rc = need_mesh(c, -1);
    if (rc) return rc;
    if ((rc = need_octree(c, "conservation_sums"))) return rc;
    if (!out) { charm_set_error("conservation_sums: null output"); return CHARM_ERR_ARG; }
    double r[16];
    if ((rc = reduce(c, 1, r))) return rc;
    memcpy(out, r, (size_t) c->nv * sizeof(double));
    return CHARM_OK;
}

// charm_gpu_conservation_sums for callers that cannot allocate: the same two
// kernels, the block partials in the library's static per-GPU array instead
// of the context's buffer, the V sums written to device memory out_dev, the
// work only enqueued on the context's stream (no allocation, memset, copy or
// synchronisation).  Every block partial the final pass reads is written by
// the first kernel of the same call, so nothing needs clearing; calls that
// overlap on one GPU race on the shared array.
extern "C" int charm_gpu_conservation_sums_noalloc(charm_gpu_ctx_t *c, double *out_dev)
{
    int rc = need_mesh(c, -1);
    if (rc) return rc;
    if ((rc = need_octree(c, "conservation_sums_noalloc"))) return rc;
    if (!out_dev) { charm_set_error("conservation_sums_noalloc: null output"); return CHARM_ERR_ARG; }
    void *scr = NULL;
    size_t bytes = 0;
    HIPCHK(noalloc_scratch(&scr, &bytes));
    const int64_t nblk = (c->n + 255) / 256;
    const size_t need = (size_t) nblk * (size_t) (c->nv + 1) * sizeof(double);
    if (need > bytes) {
        charm_set_error("conservation_sums_noalloc: %lld leaves need %zu bytes of block partials, the static "
                        "array holds %zu (at most %lld leaves)", (long long) c->n, need, bytes,
                        (long long) (bytes / ((size_t) (c->nv + 1) * sizeof(double)) * 256));
        return CHARM_ERR_ARG;
    }
    RedArgs A;
    A.U = c->B[0]; A.level = c->d_level; A.partial = (double *) scr; A.result = out_dev;
    A.n = c->n; A.mode = 1;
    HIPCHK(launch_reduce_n(c->sprec, c->nc, c->ph, A, c->P, c->st, c->nv));
    return CHARM_OK;
}

extern "C" int charm_gpu_state_flag(charm_gpu_ctx_t *c, int *flag)
{
    if (!c || !flag) { charm_set_error("state_flag: null argument"); return CHARM_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(flag, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(int), c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return CHARM_OK;
}

extern "C" int charm_gpu_synchronize(charm_gpu_ctx_t *c)
{
    if (!c) { charm_set_error("null context"); return CHARM_ERR_ARG; }
    HIPCHK(hipStreamSynchronize(c->st));
    return CHARM_OK;
}

extern "C" int charm_gpu_flux_batch(charm_gpu_ctx_t *c, int flux, int64_t n, const double *prim_l,
                                    const double *prim_r, const double *normals, double *out)
{
    if (!c || n < 0 || (n > 0 && (!prim_l || !prim_r || !normals || !out)) || flux < 0 || flux > 3) {
        charm_set_error("flux_batch: bad arguments");
        return CHARM_ERR_ARG;
    }
    if (n == 0) return CHARM_OK;
    const size_t np = (size_t) n * (5 + c->nc) * sizeof(double);
    const size_t nn = (size_t) n * 3 * sizeof(double), no = (size_t) n * c->nv * sizeof(double);
    double *dl = NULL, *dr = NULL, *dn = NULL, *dq = NULL;
    hipError_t e = hipMalloc(&dl, np);
    if (e == hipSuccess) e = hipMalloc(&dr, np);
    if (e == hipSuccess) e = hipMalloc(&dn, nn);
    if (e == hipSuccess) e = hipMalloc(&dq, no);
    if (e == hipSuccess) e = hipMemcpyAsync(dl, prim_l, np, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, prim_r, np, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dn, normals, nn, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) {
        FluxArgs A;
        memset(&A, 0, sizeof(A));
        A.prim_l = dl; A.prim_r = dr; A.normals = dn; A.out = dq; A.n = n;
        e = launch_flux_points(c->sprec, c->nc, flux, A, c->P, c->st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dq, no, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    (void) hipFree(dl); (void) hipFree(dr); (void) hipFree(dn); (void) hipFree(dq);
    if (e != hipSuccess) { charm_set_error("flux_batch: %s", hipGetErrorString(e)); return CHARM_ERR_HIP; }
    return CHARM_OK;
}

extern "C" int charm_gpu_flux_vjp_batch(charm_gpu_ctx_t *c, int flux, int64_t n, const double *prim_l,
                                        const double *prim_r, const double *normals, const double *qbar,
                                        double *prim_l_bar, double *prim_r_bar)
{
    if (!c || n < 0 || (n > 0 && (!prim_l || !prim_r || !normals || !qbar || !prim_l_bar || !prim_r_bar))) {
        charm_set_error("flux_vjp_batch: bad arguments");
        return CHARM_ERR_ARG;
    }
    if ((flux != CHARM_FLUX_LF && flux != CHARM_FLUX_HLLC && flux != CHARM_FLUX_CD) || c->nc > 2 || c->ph) {
        charm_set_error("flux_vjp_batch: LF, HLLC or CD flux, constant cp, 1-2 species");
        return CHARM_ERR_UNSUPPORTED;
    }
    if (n == 0) return CHARM_OK;
    const size_t np = (size_t) n * (5 + c->nc) * sizeof(double);
    const size_t nn = (size_t) n * 3 * sizeof(double), nq = (size_t) n * c->nv * sizeof(double);
    double *dl = NULL, *dr = NULL, *dn = NULL, *dq = NULL, *bl = NULL, *br = NULL;
    hipError_t e = hipMalloc(&dl, np);
    if (e == hipSuccess) e = hipMalloc(&dr, np);
    if (e == hipSuccess) e = hipMalloc(&dn, nn);
    if (e == hipSuccess) e = hipMalloc(&dq, nq);
    if (e == hipSuccess) e = hipMalloc(&bl, np);
    if (e == hipSuccess) e = hipMalloc(&br, np);
    if (e == hipSuccess) e = hipMemcpyAsync(dl, prim_l, np, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, prim_r, np, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dn, normals, nn, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dq, qbar, nq, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) {
        FluxArgs A;
        memset(&A, 0, sizeof(A));
        A.prim_l = dl; A.prim_r = dr; A.normals = dn; A.n = n;
        A.qbar = dq; A.bar_l = bl; A.bar_r = br;
        e = launch_flux_points_vjp(c->nc, flux, A, c->P, c->st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(prim_l_bar, bl, np, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(prim_r_bar, br, np, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    (void) hipStreamSynchronize(c->st);
    (void) hipFree(dl); (void) hipFree(dr); (void) hipFree(dn); (void) hipFree(dq); (void) hipFree(bl); (void) hipFree(br);
    if (e != hipSuccess) { charm_set_error("flux_vjp_batch: %s", hipGetErrorString(e)); return CHARM_ERR_HIP; }
    return CHARM_OK;
}

// ------------------------------------------------------------ adaptation

extern "C" int charm_gpu_amr_indicator(charm_gpu_ctx_t *c, double *err2, double *rho_c, int64_t n)
{
    int rc = need_mesh(c, n);
    if (rc) return rc;
    if ((rc = need_octree(c, "amr_indicator"))) return rc;
    if (!err2 || !rho_c) { charm_set_error("amr_indicator: null output"); return CHARM_ERR_ARG; }
    AmrArgs A;
    A.U = c->B[0]; A.level = c->d_level; A.nbr = c->d_nbr; A.aux = c->d_aux;
    A.err2 = c->d_stage64; A.rho_c = c->d_stage64 + n; A.n = n;
    HIPCHK(launch_amr_indicator(c->sprec, c->nc, c->ph, A, c->P, c->st));
    HIPCHK(hipMemcpyAsync(err2, A.err2, (size_t) n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(rho_c, A.rho_c, (size_t) n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return CHARM_OK;
}

// leaf q (x, y, z, level) lies inside leaf p
static bool leaf_inside(const int32_t *q, const int32_t *p)
{
    if (q[3] < p[3] || p[3] < 0 || q[3] > CHARM_MAXLEVEL) return false;
    const int32_t len = CHARM_ROOT_LEN >> p[3];
    for (int d = 0; d < 3; d++)
        if (q[d] < p[d] || q[d] >= p[d] + len) return false;
    return true;
}

// every index the projection kernel reads is checked here, on the host
static int check_map(const int32_t *oq, int64_t n_old, const int32_t *nq, int64_t n_new, const int64_t *map)
{
    for (int64_t i = 0; i < n_new; i++) {
        const int64_t m = map[i];
        if (m >= 0) {
            if (m >= n_old || !leaf_inside(&nq[4 * i], &oq[4 * m])) {
                charm_set_error("remesh: map[%lld] = %lld does not contain the new leaf", (long long) i, (long long) m);
                return CHARM_ERR_ARG;
            }
            continue;
        }
        const int64_t f = -1 - m;
        if (f + 8 > n_old || oq[4 * f + 3] < 1) {
            charm_set_error("remesh: map[%lld] names no family of 8 old leaves", (long long) i);
            return CHARM_ERR_ARG;
        }
        const int32_t par[4] = {oq[4 * f], oq[4 * f + 1], oq[4 * f + 2], oq[4 * f + 3] - 1};
        const int32_t plen = CHARM_ROOT_LEN >> par[3];
        bool ok = par[0] % plen == 0 && par[1] % plen == 0 && par[2] % plen == 0 && leaf_inside(&nq[4 * i], par);
        int seen = 0;
        for (int ch = 0; ok && ch < 8; ch++) {
            const int32_t *cq = &oq[4 * (f + ch)];
            ok = cq[3] == par[3] + 1 && leaf_inside(cq, par);
            if (ok) {
                const int32_t h = plen / 2;
                const int bit = ((cq[0] - par[0]) / h) | (((cq[1] - par[1]) / h) << 1) | (((cq[2] - par[2]) / h) << 2);
                ok = !(seen & (1 << bit));
                seen |= 1 << bit;
            }
        }
        if (!ok) {
            charm_set_error("remesh: old leaves %lld..%lld are not the family of the new leaf %lld",
                            (long long) f, (long long) (f + 7), (long long) i);
            return CHARM_ERR_ARG;
        }
    }
    return CHARM_OK;
}

extern "C" int charm_gpu_remesh(charm_gpu_ctx_t *c, const charm_mesh_t *m, const int64_t *map, int64_t n_map)
{
    int rc = need_mesh(c, -1);
    if (rc) return rc;
    if ((rc = need_octree(c, "remesh"))) return rc;
    if (!m || !map || n_map != m->n) { charm_set_error("remesh: map length must equal the new leaf count"); return CHARM_ERR_ARG; }
    if ((rc = check_map(c->hq, c->n, m->q, m->n, map))) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    // keep the old state and leaves while the new mesh is uploaded
    void *u_old = c->B[0];
    int4 *q_old = c->d_q;
    const int64_t n_old = c->n;
    c->B[0] = NULL; c->d_q = NULL;
    int64_t *d_map = NULL;
    hipError_t e = hipSuccess;
    rc = charm_gpu_upload_mesh(c, m);
    if (!rc) {
        e = hipMalloc((void **) &d_map, (size_t) (m->n ? m->n : 1) * sizeof(int64_t));
        if (e == hipSuccess) e = hipMemcpyAsync(d_map, map, (size_t) m->n * sizeof(int64_t), hipMemcpyHostToDevice, c->st);
        if (e == hipSuccess) {
            ProjArgs A;
            A.Uold = u_old; A.Unew = c->B[0]; A.old_q = q_old; A.new_q = c->d_q; A.map = d_map;
            A.n_old = n_old; A.n_new = m->n; A.nvk = 4 * c->nv;
            e = launch_amr_project(c->sprec, A, c->st);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    }
    (void) hipStreamSynchronize(c->st);
    if (d_map) (void) hipFree(d_map);
    (void) hipFree(u_old);
    (void) hipFree(q_old);
    if (rc) return rc;
    if (e != hipSuccess) { charm_set_error("remesh: %s", hipGetErrorString(e)); return CHARM_ERR_HIP; }
    return CHARM_OK;
}

// ------------------------------------------------ Navier-Stokes test hooks

extern "C" int charm_gpu_ns_aux(charm_gpu_ctx_t *c, double *aux, int64_t n)
{
    int rc = need_mesh(c, n);
    if (rc) return rc;
    if ((rc = need_octree(c, "ns_aux"))) return rc;
    if (!c->ph || !aux) { charm_set_error("ns_aux: no Navier-Stokes terms in this configuration"); return CHARM_ERR_ARG; }
    NsArgs N = ns_args(c, c->B[0], NULL, 0.0);
    if (N.dcoef) HIPCHK(launch_ns(c->sprec, 1, c->nc, N, c->P, c->st));
    HIPCHK(launch_ns(c->sprec, 2, c->nc, N, c->P, c->st));
    // [36][n] -> [n][9][4] FP64
    HIPCHK(launch_convert(c->sprec, 1, c->d_auxf, c->d_stage64, n, 36, c->st));
    HIPCHK(hipMemcpyAsync(aux, c->d_stage64, (size_t) n * 36 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return CHARM_OK;
}

extern "C" int charm_gpu_ns_chem(charm_gpu_ctx_t *c, double dt, double *chem_rhs, int64_t n)
{
    int rc = need_mesh(c, n);
    if (rc) return rc;
    if ((rc = need_octree(c, "ns_chem"))) return rc;
    if (!c->ph || c->cfg.n_react <= 0) { charm_set_error("ns_chem: no reactions in this configuration"); return CHARM_ERR_ARG; }
    NsArgs N = ns_args(c, c->B[0], c->B[0], dt);
    HIPCHK(launch_ns(c->sprec, 0, c->nc, N, c->P, c->st));
    if (chem_rhs) HIPCHK(hipMemcpyAsync(chem_rhs, c->d_chem, (size_t) n * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return CHARM_OK;
}

// ------------------------------------------------------------ adjoint
//
// Discrete adjoint of nsteps SSP-RK3 steps (Euler terms, no limiter) for
// J = sum_{n=1..N} dt Q(u^n), Q(u) = sum over the leaves whose centre lies in
// the box [lo, hi] of int p.  The forward pass stores u^n every `ckpt` steps;
// the reverse pass recomputes each segment's stage states and runs the
// transposed stages:
//   l += dt dQ/du(u^{n+1});  ub2 = S^T_{u2}(2/3 l);  a = l/3 + 3/4 ub2;
//   ub1 = S^T_{u1}(1/4 ub2);  l = S^T_{u0}(ub1) + a,
// S^T_U(y) = y + (d int / d U)^T (-dt A^-1 y).  The inlet parameters
// (velocity x, T, p of face theta_face) accumulate their gradient per leaf.

// the boundary types the adjoint's boundary linearisation covers (all but
// the mass-flow condition)
static int adj_bnd_supported(const charm_gpu_ctx_t *c)
{
    for (int f = 0; f < 6; f++)
        if (c->cfg.bnd_type[f] == CHARM_BND_MASS_FLOW) {
            charm_set_error("adjoint: the mass-flow boundary has no adjoint here");
            return CHARM_ERR_UNSUPPORTED;
        }
    return CHARM_OK;
}

static int adj_check(charm_gpu_ctx_t *c, int nsteps, int ckpt, int theta_face)
{
    int rc = need_mesh(c, -1);
    if (rc) return rc;
    if ((rc = need_octree(c, "adjoint"))) return rc;
    if ((rc = adj_bnd_supported(c))) return rc;
    if (c->cfg.model != CHARM_MODEL_NS || c->cfg.limiter || c->ph || c->nc > 2 || c->prec == CHARM_FP32_MIX ||
        c->cfg.flux == CHARM_FLUX_GODUNOV) {
        charm_set_error("adjoint: needs the SSP-RK3 model without limiter, constant cp, 1-2 species and "
                        "the LF, HLLC or CD flux");
        return CHARM_ERR_UNSUPPORTED;
    }
    if (nsteps < 1 || ckpt < 0 || theta_face < -1 || theta_face > 5 ||
        (theta_face >= 0 && c->cfg.bnd_type[theta_face] != CHARM_BND_INLET)) {
        charm_set_error("adjoint: bad step count, checkpoint period or inlet face");
        return CHARM_ERR_ARG;
    }
    return CHARM_OK;
}

static ObjArgs obj_args(charm_gpu_ctx_t *c, const void *U, const double lo[3], const double hi[3])
{
    ObjArgs O;
    O.U = U; O.lambda = NULL; O.q = c->d_q; O.coef = 0.0; O.partial = NULL; O.mode = 0; O.n = c->n;
    for (int d = 0; d < 3; d++) { O.lo[d] = lo[d]; O.hi[d] = hi[d]; }
    return O;
}

// forward pass: nsteps steps from B[0]; per-step objective partials into
// part[nsteps][nblk]; optional checkpoints every ckpt steps
// forward pass with checkpoints every ckpt steps.  With seg_last, the steps
// of the last segment keep their stage states there, in the segment layout
// of the reverse sweep (slot j of S0 = u^(start+j), S1 / S2 the stage
// states), so that segment needs no recompute; the stage launches are the
// recompute's, so the states are the same bits.
static int adj_forward(charm_gpu_ctx_t *c, double dt, int nsteps, int ckpt, const double lo[3],
                       const double hi[3], char *cp, char *seg_last, double *part, int64_t nblk, double *J)
{
    const size_t ssz = (size_t) c->n * 4 * c->nv * c->esz;
    const int last = ((nsteps - 1) / ckpt) * ckpt;
    char *S0 = seg_last, *S1 = seg_last ? S0 + (size_t) (ckpt + 1) * ssz : NULL;
    char *S2 = seg_last ? S1 + (size_t) ckpt * ssz : NULL;
    for (int s = 0; s < nsteps; s++) {
        // the last segment's checkpoint is not stored: its states stay in seg_last
        if (cp && s % ckpt == 0 && s < last)
            HIPCHK(hipMemcpyAsync(cp + (size_t) (s / ckpt) * ssz, c->B[0], ssz, hipMemcpyDeviceToDevice, c->st));
        const void *ucur;
        if (seg_last && s >= last) {
            const int j = s - last;
            void *u0 = S0 + (size_t) j * ssz, *u1 = S1 + (size_t) j * ssz, *u2 = S2 + (size_t) j * ssz;
            void *un = S0 + (size_t) (j + 1) * ssz;
            if (j == 0) HIPCHK(hipMemcpyAsync(u0, c->B[0], ssz, hipMemcpyDeviceToDevice, c->st));
            StageArgs A = stage_args(c, u0, u0, u1, 1, 0, dt);
            HIPCHK(launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, c->st));
            A = stage_args(c, u1, u0, u2, 1, 1, dt);
            HIPCHK(launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, c->st));
            A = stage_args(c, u2, u0, un, 1, 2, dt);
            HIPCHK(launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, c->st));
            ucur = un;
        } else {
            int rc = step_once(c, dt);
            if (rc) return rc;
            ucur = c->B[0];
        }
        ObjArgs O = obj_args(c, ucur, lo, hi);
        O.partial = part + (size_t) s * nblk;
        HIPCHK(launch_adj_objective(c->sprec, c->nc, O, c->P, c->st));
    }
    // the context's state ends at u^nsteps, as after step_once
    if (seg_last)
        HIPCHK(hipMemcpyAsync(c->B[0], S0 + (size_t) (nsteps - last) * ssz, ssz, hipMemcpyDeviceToDevice, c->st));
    double *h = (double *) malloc((size_t) nsteps * nblk * sizeof(double));
    if (!h) { charm_set_error("adjoint: out of memory"); return CHARM_ERR_MEM; }
    hipError_t e = hipMemcpyAsync(h, part, (size_t) nsteps * nblk * sizeof(double), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) { free(h); charm_set_error("adjoint: %s", hipGetErrorString(e)); return CHARM_ERR_HIP; }
    double j = 0.0;
    for (int s = 0; s < nsteps; s++) {
        double q = 0.0;
        for (int64_t b = 0; b < nblk; b++) q += h[(size_t) s * nblk + b];
        j += dt * q;
    }
    free(h);
    *J = j;
    return check_flag(c);
}

// forward pass that keeps every stage state: slot j of S0 = u^j (slot 0 the
// initial state), S1 / S2 the stage states of step j, in the segment layout
// the reverse sweep reads; the same stage launches as the recompute
static int adj_forward_store(charm_gpu_ctx_t *c, double dt, int nsteps, const double lo[3],
                             const double hi[3], char *S0, double *part, int64_t nblk, double *J)
{
    const size_t ssz = (size_t) c->n * 4 * c->nv * c->esz;
    char *S1 = S0 + (size_t) (nsteps + 1) * ssz, *S2 = S1 + (size_t) nsteps * ssz;
    HIPCHK(hipMemcpyAsync(S0, c->B[0], ssz, hipMemcpyDeviceToDevice, c->st));
    for (int j = 0; j < nsteps; j++) {
        void *u0 = S0 + (size_t) j * ssz, *u1 = S1 + (size_t) j * ssz, *u2 = S2 + (size_t) j * ssz;
        void *un = S0 + (size_t) (j + 1) * ssz;
        StageArgs A = stage_args(c, u0, u0, u1, 1, 0, dt);
        HIPCHK(launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, c->st));
        A = stage_args(c, u1, u0, u2, 1, 1, dt);
        HIPCHK(launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, c->st));
        A = stage_args(c, u2, u0, un, 1, 2, dt);
        HIPCHK(launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, c->st));
        ObjArgs O = obj_args(c, un, lo, hi);
        O.partial = part + (size_t) j * nblk;
        HIPCHK(launch_adj_objective(c->sprec, c->nc, O, c->P, c->st));
    }
    HIPCHK(hipMemcpyAsync(c->B[0], S0 + (size_t) nsteps * ssz, ssz, hipMemcpyDeviceToDevice, c->st));
    double *h = (double *) malloc((size_t) nsteps * nblk * sizeof(double));
    if (!h) { charm_set_error("adjoint: out of memory"); return CHARM_ERR_MEM; }
    hipError_t e = hipMemcpyAsync(h, part, (size_t) nsteps * nblk * sizeof(double), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) { free(h); charm_set_error("adjoint: %s", hipGetErrorString(e)); return CHARM_ERR_HIP; }
    double j = 0.0;
    for (int s = 0; s < nsteps; s++) {
        double q = 0.0;
        for (int64_t b = 0; b < nblk; b++) q += h[(size_t) s * nblk + b];
        j += dt * q;
    }
    free(h);
    *J = j;
    return check_flag(c);
}

extern "C" int charm_gpu_objective(charm_gpu_ctx_t *c, double dt, int nsteps, const double lo[3],
                                   const double hi[3], double *J)
{
    int rc = adj_check(c, nsteps, 1, -1);
    if (rc) return rc;
    if (!lo || !hi || !J) { charm_set_error("objective: null argument"); return CHARM_ERR_ARG; }
    const int64_t nblk = (c->n + 255) / 256;
    double *part = NULL;
    HIPCHK(hipMalloc((void **) &part, (size_t) nsteps * nblk * sizeof(double)));
    rc = adj_forward(c, dt, nsteps, 1, lo, hi, NULL, NULL, part, nblk, J);
    (void) hipFree(part);
    return rc;
}

// Boundary faces of the mesh for the adjoint's boundary kernels (from the
// neighbour table: codes -1..-6 are boundary faces): the leaves with one,
// their face ranges, the faces (leaf * 8 + direction, ordered by leaf), and
// the per-point scratch of their contributions.
struct AdjBoundary {
    int *elems, *ptr, *faces, *of;
    void *part;
    double *theta;
    int n_elems, n_faces;
};

static void adj_boundary_free(AdjBoundary &b)
{
    void *bufs[] = {b.elems, b.ptr, b.faces, b.of, b.part, b.theta};
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++)
        if (bufs[i]) (void) hipFree(bufs[i]);
    memset(&b, 0, sizeof(b));
}

static int adj_boundary(charm_gpu_ctx_t *c, AdjBoundary &b)
{
    memset(&b, 0, sizeof(b));
    const int64_t n = c->n;
    int *nbr = (int *) malloc((size_t) n * 6 * sizeof(int));
    int *elems = (int *) malloc((size_t) n * sizeof(int));
    int *ptr = (int *) malloc(((size_t) n + 1) * sizeof(int));
    int *faces = (int *) malloc((size_t) n * 6 * sizeof(int));
    int *of = (int *) malloc((size_t) (n ? n : 1) * sizeof(int));
    if (!nbr || !elems || !ptr || !faces || !of) {
        free(nbr); free(elems); free(ptr); free(faces); free(of);
        charm_set_error("adjoint: out of memory");
        return CHARM_ERR_MEM;
    }
    hipError_t e = hipMemcpyAsync(nbr, c->d_nbr, (size_t) n * 6 * sizeof(int), hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    int ne = 0, nf = 0;
    if (e == hipSuccess) {
        ptr[0] = 0;
        for (int64_t i = 0; i < n; i++) {
            const int f0 = nf;
            for (int d = 0; d < 6; d++)
                if (nbr[6 * i + d] < 0 && nbr[6 * i + d] > -8) faces[nf++] = (int) (i * 8 + d);
            of[i] = -1;
            if (nf > f0) { of[i] = ne; elems[ne] = (int) i; ptr[++ne] = nf; }
        }
    }
    if (e == hipSuccess && ne > 0) {
        e = hipMalloc((void **) &b.elems, (size_t) ne * sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void **) &b.ptr, ((size_t) ne + 1) * sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void **) &b.faces, (size_t) nf * sizeof(int));
        if (e == hipSuccess) e = hipMalloc(&b.part, (size_t) nf * 6 * 4 * c->nv * c->esz);
        if (e == hipSuccess) e = hipMalloc((void **) &b.theta, (size_t) nf * 6 * 3 * sizeof(double));
        if (e == hipSuccess) e = hipMemcpy(b.elems, elems, (size_t) ne * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b.ptr, ptr, ((size_t) ne + 1) * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b.faces, faces, (size_t) nf * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc((void **) &b.of, (size_t) n * sizeof(int));
        if (e == hipSuccess) e = hipMemcpy(b.of, of, (size_t) n * sizeof(int), hipMemcpyHostToDevice);
    }
    free(nbr); free(elems); free(ptr); free(faces); free(of);
    if (e != hipSuccess) {
        adj_boundary_free(b);
        charm_set_error("adjoint: %s", hipGetErrorString(e));
        return CHARM_ERR_HIP;
    }
    b.n_elems = ne;
    b.n_faces = nf;
    return CHARM_OK;
}

// One forward-Euler stage S(u) = u - dt A^-1 int(u) (the Euler model's step
// without limiter) and its vector-Jacobian product out = S_u^T x = x +
// (d int / d u)^T (-dt A^-1 x): the stage kernel the adjoint sweeps with,
// on host FP64 arrays [n][V][4].
extern "C" int charm_gpu_stage_vjp(charm_gpu_ctx_t *c, double dt, const double *u, const double *x,
                                   double *out, int64_t n)
{
    int rc = need_mesh(c, n);
    if (rc) return rc;
    if ((rc = need_octree(c, "stage_vjp"))) return rc;
    if (!u || !x || !out || !(dt >= 0)) { charm_set_error("stage_vjp: bad argument"); return CHARM_ERR_ARG; }
    if (c->cfg.limiter || c->ph || c->nc > 2 || c->prec == CHARM_FP32_MIX || c->cfg.flux == CHARM_FLUX_GODUNOV) {
        charm_set_error("stage_vjp: needs no limiter, constant cp, 1-2 species and the LF, HLLC or CD flux");
        return CHARM_ERR_UNSUPPORTED;
    }
    if ((rc = adj_bnd_supported(c))) return rc;
    const size_t ssz = (size_t) n * 4 * c->nv * c->esz, dsz = (size_t) n * 4 * c->nv * sizeof(double);
    AdjBoundary bnd;
    if ((rc = adj_boundary(c, bnd))) return rc;
    int dev = 0, lds_max = 0;
    const bool chunked = c->sprec && hipGetDevice(&dev) == hipSuccess &&
                         hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
                         adj_chunk_smem_bytes(1, c->nc, c->max_faces) <= (size_t) lds_max;
    (void) hipGetLastError();
    char *dU = NULL, *dX = NULL, *dO = NULL;
    hipError_t e = hipMalloc((void **) &dU, ssz);
    if (e == hipSuccess) e = hipMalloc((void **) &dX, ssz);
    if (e == hipSuccess) e = hipMalloc((void **) &dO, ssz);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_stage64, u, dsz, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = launch_convert(c->sprec, 0, c->d_stage64, dU, n, 4 * c->nv, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_stage64, x, dsz, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = launch_convert(c->sprec, 0, c->d_stage64, dX, n, 4 * c->nv, c->st);
    if (e == hipSuccess) {
        AdjArgs A;
        memset(&A, 0, sizeof(A));
        A.level = c->d_level; A.nbr = c->d_nbr; A.aux = c->d_aux; A.n = n; A.dt = dt;
        A.theta_bar = NULL; A.theta_face = -1;
        A.bnd_elems = bnd.elems; A.bnd_ptr = bnd.ptr; A.bnd_faces = bnd.faces;
        A.n_bnd = bnd.n_elems; A.n_bface = bnd.n_faces; A.bnd_part = bnd.part; A.bnd_theta = bnd.theta;
        A.cf4 = chunked ? c->d_cf4 : NULL;
        A.ch_ebeg = c->d_ch_ebeg; A.ch_fbeg = c->d_ch_fbeg; A.slot_ptr = c->d_slot_ptr; A.slots = c->d_slots;
        A.n_chunks = (int) c->nch; A.max_faces = c->max_faces; A.chunk_elems = c->chunk_elems;
        A.n_xcd = c->n_xcd;
        A.bnd_of = chunked ? bnd.of : NULL;
        A.U = dU; A.X = dX; A.ACC = NULL; A.ACC2 = NULL; A.OUT = dO;
        A.c_in = 1.0; A.c_out = 1.0; A.c_acc = 0.0; A.c_acc2 = 0.0;
        e = launch_adj_stage(c->sprec, c->nc, c->cfg.flux, A, c->P, c->st);
    }
    if (e == hipSuccess) e = launch_convert(c->sprec, 1, dO, c->d_stage64, n, 4 * c->nv, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, c->d_stage64, dsz, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    (void) hipStreamSynchronize(c->st);
    void *bufs[] = {dU, dX, dO};
    for (size_t i = 0; i < 3; i++)
        if (bufs[i]) (void) hipFree(bufs[i]);
    adj_boundary_free(bnd);
    if (e != hipSuccess) { charm_set_error("stage_vjp: %s", hipGetErrorString(e)); return CHARM_ERR_HIP; }
    return check_flag(c);
}

extern "C" int charm_gpu_adjoint(charm_gpu_ctx_t *c, double dt, int nsteps, int ckpt, const double lo[3],
                                 const double hi[3], int theta_face, double *J, double *grad_u0,
                                 double *grad_theta, double *fwd_ms, double *adj_ms)
{
    int rc = adj_check(c, nsteps, ckpt, theta_face);
    if (rc) return rc;
    if (!lo || !hi || !J) { charm_set_error("adjoint: null argument"); return CHARM_ERR_ARG; }
    const int64_t n = c->n, nblk = (n + 255) / 256;
    const size_t ssz = (size_t) n * 4 * c->nv * c->esz;
    const int64_t m = n * 4 * c->nv;
    // ckpt = 0: every stage state of the forward pass is kept (one segment
    // of nsteps steps filled by the forward pass, no recompute); else
    // checkpoints every ckpt steps and segments recomputed in two buffers
    const bool store_all = ckpt == 0;
    // a period beyond the run is one segment of the whole run
    if (ckpt > nsteps) ckpt = nsteps;
    const int seg_len = store_all ? nsteps : ckpt;
    const int ncp = store_all ? 1 : (nsteps + ckpt - 1) / ckpt;
    // checkpoints 0 .. ncp - 2 are stored (the last segment's states are kept
    // by the forward pass); two segment buffers only when a segment is recomputed
    const int nseg_buf = ncp > 1 ? 2 : 1;
    char *cp = NULL, *seg = NULL, *la = NULL, *lb = NULL, *lc = NULL, *segb[2] = {NULL, NULL};
    hipStream_t st2 = NULL;
    hipEvent_t ev_ready[2] = {NULL, NULL}, ev_done[2] = {NULL, NULL};
    double *part = NULL, *theta = NULL;
    AdjBoundary bnd;
    if ((rc = adj_boundary(c, bnd))) return rc;
    // FP32: the chunked face-centric stage adjoint when its LDS fits a
    // workgroup (else, and in FP64, the element-centric one)
    int dev = 0, lds_max = 0;
    const bool chunked = c->sprec && hipGetDevice(&dev) == hipSuccess &&
                         hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
                         adj_chunk_smem_bytes(1, c->nc, c->max_faces) <= (size_t) lds_max;
    (void) hipGetLastError();
    hipError_t e = ncp > 1 ? hipMalloc((void **) &cp, (size_t) (ncp - 1) * ssz) : hipSuccess;
    if (e == hipSuccess) e = hipMalloc((void **) &seg, (size_t) nseg_buf * (3 * seg_len + 1) * ssz);
    if (e == hipSuccess) e = hipStreamCreate(&st2);
    for (int i = 0; i < 2 && e == hipSuccess; i++) {
        e = hipEventCreateWithFlags(&ev_ready[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_done[i], hipEventDisableTiming);
    }
    segb[0] = seg;
    segb[1] = (seg && nseg_buf == 2) ? seg + (size_t) (3 * seg_len + 1) * ssz : seg;
    if (e == hipSuccess) e = hipMalloc((void **) &la, ssz);
    if (e == hipSuccess) e = hipMalloc((void **) &lb, ssz);
    if (e == hipSuccess) e = hipMalloc((void **) &lc, ssz);
    if (e == hipSuccess) e = hipMalloc((void **) &part, (size_t) nsteps * nblk * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **) &theta, (size_t) n * 3 * sizeof(double));
    float ms_f = 0.f, ms_a = 0.f;
    if (e != hipSuccess) {
        charm_set_error("adjoint: device memory: %s", hipGetErrorString(e));
        rc = CHARM_ERR_MEM;
        goto done;
    }
#define ADJ_CHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { charm_set_error("adjoint: %s", hipGetErrorString(e_)); rc = CHARM_ERR_HIP; goto done; } } while (0)
    ADJ_CHK(hipEventRecord(c->ev0, c->st));
    if (store_all) rc = adj_forward_store(c, dt, nsteps, lo, hi, segb[0], part, nblk, J);
    else rc = adj_forward(c, dt, nsteps, ckpt, lo, hi, cp, segb[(ncp - 1) & 1], part, nblk, J);
    if (rc) goto done;
    ADJ_CHK(hipEventRecord(c->ev1, c->st));
    ADJ_CHK(hipEventSynchronize(c->ev1));
    ADJ_CHK(hipEventElapsedTime(&ms_f, c->ev0, c->ev1));

    // Reverse sweep.  The last segment's stage states were kept by the
    // forward pass (in segment buffer (ncp - 1) & 1, on the context stream).
    // Segment sg's are recomputed from its
    // checkpoint into one of two segment buffers on a second stream while
    // the adjoint of segment sg + 1 runs on the context stream from the
    // other buffer (events order the reuse of a buffer); per segment buffer
    // slot j of S0 holds u^(start+j) (slot 0 = the checkpoint, slot L the
    // segment's end state), S1 / S2 the stage states.
    ADJ_CHK(hipEventRecord(c->ev0, c->st));
    ADJ_CHK(hipMemsetAsync(la, 0, ssz, c->st));
    ADJ_CHK(hipMemsetAsync(theta, 0, (size_t) n * 3 * sizeof(double), c->st));
    ADJ_CHK(hipEventRecord(ev_done[0], c->st));
    ADJ_CHK(hipEventRecord(ev_done[1], c->st));
    for (int sg = ncp - 1; sg >= -1; sg--) {
        // recompute segment sg - 1 ahead (the last segment's states come
        // from the forward pass)
        if (!store_all && sg >= 1) {
            const int r = sg - 1;
            const int start = r * ckpt, end = (start + ckpt < nsteps) ? start + ckpt : nsteps, L = end - start;
            char *S0 = segb[r & 1], *S1 = S0 + (size_t) (ckpt + 1) * ssz, *S2 = S1 + (size_t) ckpt * ssz;
            // the buffer's previous user (the adjoint of segment r + 2) is done
            ADJ_CHK(hipStreamWaitEvent(st2, ev_done[r & 1], 0));
            ADJ_CHK(hipMemcpyAsync(S0, cp + (size_t) r * ssz, ssz, hipMemcpyDeviceToDevice, st2));
            for (int j = 0; j < L; j++) {
                void *u0 = S0 + (size_t) j * ssz, *u1 = S1 + (size_t) j * ssz, *u2 = S2 + (size_t) j * ssz;
                void *un = S0 + (size_t) (j + 1) * ssz;
                StageArgs A = stage_args(c, u0, u0, u1, 1, 0, dt);
                ADJ_CHK(launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, st2));
                A = stage_args(c, u1, u0, u2, 1, 1, dt);
                ADJ_CHK(launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, st2));
                A = stage_args(c, u2, u0, un, 1, 2, dt);
                ADJ_CHK(launch_stage(c->prec, c->nc, c->cfg.flux, c->ph, A, c->P, st2));
            }
            ADJ_CHK(hipEventRecord(ev_ready[r & 1], st2));
        }
        if (sg < 0) break;
        const int start = sg * seg_len, end = (start + seg_len < nsteps) ? start + seg_len : nsteps, L = end - start;
        char *S0 = segb[sg & 1], *S1 = S0 + (size_t) (seg_len + 1) * ssz, *S2 = S1 + (size_t) seg_len * ssz;
        if (!store_all && sg != ncp - 1) ADJ_CHK(hipStreamWaitEvent(c->st, ev_ready[sg & 1], 0));
        for (int j = L - 1; j >= 0; j--) {
            void *u0 = S0 + (size_t) j * ssz, *u1 = S1 + (size_t) j * ssz, *u2 = S2 + (size_t) j * ssz;
            void *un = S0 + (size_t) (j + 1) * ssz;
            ObjArgs O = obj_args(c, un, lo, hi);
            O.mode = 1; O.lambda = la; O.coef = dt;
            ADJ_CHK(launch_adj_objective(c->sprec, c->nc, O, c->P, c->st));
            AdjArgs A;
            A.level = c->d_level; A.nbr = c->d_nbr; A.aux = c->d_aux; A.n = n; A.dt = dt;
            A.theta_bar = theta_face >= 0 ? theta : NULL; A.theta_face = theta_face;
            A.bnd_elems = bnd.elems; A.bnd_ptr = bnd.ptr; A.bnd_faces = bnd.faces;
            A.n_bnd = bnd.n_elems; A.n_bface = bnd.n_faces; A.bnd_part = bnd.part; A.bnd_theta = bnd.theta;
            A.cf4 = chunked ? c->d_cf4 : NULL;
            A.ch_ebeg = c->d_ch_ebeg; A.ch_fbeg = c->d_ch_fbeg; A.slot_ptr = c->d_slot_ptr; A.slots = c->d_slots;
            A.n_chunks = (int) c->nch; A.max_faces = c->max_faces; A.chunk_elems = c->chunk_elems;
            A.n_xcd = c->n_xcd;
            A.bnd_of = chunked ? bnd.of : NULL;
            // l3 = S^T(2/3 l); l2 = S^T(1/4 l3); l = S^T(l2) + l / 3 + 3/4 l3
            A.U = u2; A.X = la; A.ACC = NULL; A.ACC2 = NULL; A.OUT = lb;
            A.c_in = 2.0 / 3.0; A.c_out = 1.0; A.c_acc = 0.0; A.c_acc2 = 0.0;
            ADJ_CHK(launch_adj_stage(c->sprec, c->nc, c->cfg.flux, A, c->P, c->st));
            A.U = u1; A.X = lb; A.OUT = lc; A.c_in = 0.25;
            ADJ_CHK(launch_adj_stage(c->sprec, c->nc, c->cfg.flux, A, c->P, c->st));
            A.U = u0; A.X = lc; A.ACC = la; A.ACC2 = lb; A.OUT = la; A.c_in = 1.0;
            A.c_acc = 1.0 / 3.0; A.c_acc2 = 0.75;
            ADJ_CHK(launch_adj_stage(c->sprec, c->nc, c->cfg.flux, A, c->P, c->st));
        }
        ADJ_CHK(hipEventRecord(ev_done[sg & 1], c->st));
    }
    ADJ_CHK(hipEventRecord(c->ev1, c->st));
    ADJ_CHK(hipEventSynchronize(c->ev1));
    ADJ_CHK(hipEventElapsedTime(&ms_a, c->ev0, c->ev1));
    if (grad_u0) {
        ADJ_CHK(launch_convert(c->sprec, 1, la, c->d_stage64, n, 4 * c->nv, c->st));
        ADJ_CHK(hipMemcpyAsync(grad_u0, c->d_stage64, (size_t) m * sizeof(double), hipMemcpyDeviceToHost, c->st));
    }
    if (grad_theta) {
        double *h = (double *) malloc((size_t) n * 3 * sizeof(double));
        if (!h) { charm_set_error("adjoint: out of memory"); rc = CHARM_ERR_MEM; goto done; }
        e = hipMemcpyAsync(h, theta, (size_t) n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        grad_theta[0] = grad_theta[1] = grad_theta[2] = 0.0;
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) grad_theta[k] += h[3 * i + k];
        free(h);
        ADJ_CHK(e);
    }
    ADJ_CHK(hipStreamSynchronize(c->st));
    if (fwd_ms) *fwd_ms = ms_f;
    if (adj_ms) *adj_ms = ms_a;
    rc = check_flag(c);
#undef ADJ_CHK
done:
    (void) hipStreamSynchronize(c->st);
    if (st2) { (void) hipStreamSynchronize(st2); (void) hipStreamDestroy(st2); }
    for (int i = 0; i < 2; i++) {
        if (ev_ready[i]) (void) hipEventDestroy(ev_ready[i]);
        if (ev_done[i]) (void) hipEventDestroy(ev_done[i]);
    }
    void *bufs[] = {cp, seg, la, lb, lc, part, theta};
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++)
        if (bufs[i]) (void) hipFree(bufs[i]);
    adj_boundary_free(bnd);
    return rc;
}
