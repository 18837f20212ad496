/*
 * oracle.c -- CPU restatement of the reference's range-limited force path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h for the rules and for what pins each
 * function).  Built with -O2 -ffp-contract=off and no -ffast-math so that every
 * float operation below rounds exactly where it is written; fused operations are
 * written as explicit fmaf() calls.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* tables                                                                     */
/* ------------------------------------------------------------------------- */

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* D/RL_LJ_Evaluate_Pairs_1st_Order.v:176-199: rdaddr = {segment_id, bin_id},
 * bin_id = r2[22:15], segment_id = r2[30:23] - offset when that lies in
 * [0, SEGMENT_NUM), else 0 (the subtraction is unsigned in the HDL, so a
 * negative difference is also out of range). */
int or_lut_index(float r2, int exp_offset, int nseg)
{
    uint32_t u = f2u(r2);
    int e = (int)((u >> 23) & 0xffu);
    int seg = e - exp_offset;
    if (seg < 0 || seg >= nseg) seg = 0;
    int bin = (int)((u >> 15) & 0xffu);
    return seg * 256 + bin;
}

/* M/Simulation_Verification_Raw_ApoA1_LJ_no_smooth.m:120-198. */
void or_pair_golden(const float *c0_14, const float *c1_14, const float *c0_8,
                    const float *c1_8, int exp_offset, int nseg, float rc2, int strict,
                    const float *ref, const float *nbr, int64_t n, float *out)
{
    for (int64_t p = 0; p < n; ++p) {
        float dx = ref[3 * p + 0] - nbr[3 * p + 0];             /* :132-134 */
        float dy = ref[3 * p + 1] - nbr[3 * p + 1];
        float dz = ref[3 * p + 2] - nbr[3 * p + 2];
        float r2 = dx * dx + dy * dy;                           /* :135, left to right */
        r2 = r2 + dz * dz;
        int in = strict ? (r2 < rc2 && r2 > 0.0f) : (r2 <= rc2 && r2 > 0.0f);
        float F = 0.0f, fx = 0.0f, fy = 0.0f, fz = 0.0f;
        if (in) {
            int idx = or_lut_index(r2, exp_offset, nseg);       /* :139-153 */
            float t14 = c1_14[idx] * r2;                        /* :174 polyval, unfused */
            t14 = t14 + c0_14[idx];
            float t8 = c1_8[idx] * r2;                          /* :175 */
            t8 = t8 + c0_8[idx];
            F = t14 - t8;                                       /* :184 */
            fx = F * dx;                                        /* :185-187 */
            fy = F * dy;
            fz = F * dz;
        }
        float *o = out + 8 * p;
        o[0] = F; o[1] = fx; o[2] = fy; o[3] = fz;
        o[4] = r2; o[5] = dx; o[6] = dy; o[7] = dz;
    }
}

/* ------------------------------------------------------------------------- */
/* C model                                                                    */
/* ------------------------------------------------------------------------- */

/* R/C_Model_ForceEval/MD_Evaluation_Model.cpp:11-29, with the same promotions:
 * double literals rounded to float on assignment, float/float divisions. */
void or_cmodel_default_params(or_cmodel_params *p)
{
    float Nav = (float)6.022e23;                       /* :15 */
    float Ar_weight = (float)39.95;                    /* :16 */
    p->eps = (float)(1.995996 * 1.995996);             /* :17 */
    p->sigma = (float)(0.1675 * 2);                    /* :18 */
    p->mass = Ar_weight / Nav / 1000.0f;               /* :21 */
    p->dt = (float)2E-15;                              /* :22 */
    p->rc = 8.0f;                                      /* :23 */
    p->box[0] = p->box[1] = p->box[2] = 24.0f;         /* :26-28 */
    p->iterations = 100;                               /* :11 */
}

/* One PBC component, :81-98: dx -= L * (int)(dx/L +- 0.5), the +-0.5 in double. */
static inline float cm_pbc(float d, float L)
{
    if (d >= 0) {
        float q = d / L;
        int k = (int)((double)q + 0.5);
        d -= L * (float)k;
    } else {
        float q = d / L;
        int k = (int)((double)q - 0.5);
        d -= L * (float)k;
    }
    return d;
}

typedef struct { float s12, s6, c4, c48, c24, rc2; } cm_consts;

static void cm_setup(const or_cmodel_params *p, cm_consts *c)
{
    c->s12 = (float)pow((double)p->sigma, 12.0);       /* :19 */
    c->s6 = (float)pow((double)p->sigma, 6.0);         /* :20 */
    c->c4 = 4.0f * p->eps;
    c->c48 = 48.0f * p->eps;
    c->c24 = 24.0f * p->eps;
    c->rc2 = p->rc * p->rc;                            /* :24 */
}

/* The :72-124 neighbour loop for one reference particle. */
static void cm_force_one(const float *particle, int n, int ref, const or_cmodel_params *p,
                         const cm_consts *c, float *fx_o, float *fy_o, float *fz_o,
                         float *e_o, int64_t *npairs)
{
    float Evdw_acc = 0, Fx_acc = 0, Fy_acc = 0, Fz_acc = 0;
    float ref_x = particle[11 * ref + 0];
    float ref_y = particle[11 * ref + 1];
    float ref_z = particle[11 * ref + 2];
    int64_t cnt = 0;
    for (int nb = 0; nb < n; ++nb) {
        float dx = ref_x - particle[11 * nb + 0];
        float dy = ref_y - particle[11 * nb + 1];
        float dz = ref_z - particle[11 * nb + 2];
        dx = cm_pbc(dx, p->box[0]);
        dy = cm_pbc(dy, p->box[1]);
        dz = cm_pbc(dz, p->box[2]);
        float r2 = dx * dx + dy * dy;                  /* :99 */
        r2 = r2 + dz * dz;
        if (r2 > 0 && r2 <= c->rc2) {                  /* :101 */
            ++cnt;
            float inv_r2 = 1.0f / r2;                  /* :104 */
            float inv_r8 = (float)pow((double)inv_r2, 4.0);
            float inv_r14 = (float)pow((double)inv_r2, 7.0);
            float inv_r6 = (float)pow((double)inv_r2, 3.0);
            float inv_r12 = (float)pow((double)inv_r2, 6.0);
            float vdw12 = c->c4 * c->s12 * inv_r12;    /* :109-112, left to right */
            float vdw6 = c->c4 * c->s6 * inv_r6;
            float vdw14 = c->c48 * c->s12 * inv_r14;
            float vdw8 = c->c24 * c->s6 * inv_r8;
            float Fvdw = vdw14 - vdw8;                 /* :114-115 */
            float Evdw = vdw12 - vdw6;
            float t;
            t = Fvdw * dx; Fx_acc += t;                /* :117-121 */
            t = Fvdw * dy; Fy_acc += t;
            t = Fvdw * dz; Fz_acc += t;
            Evdw_acc += Evdw;
        }
    }
    *fx_o = Fx_acc; *fy_o = Fy_acc; *fz_o = Fz_acc; *e_o = Evdw_acc;
    if (npairs) *npairs += cnt;
}

int or_cmodel_run(float *particle, int n, const or_cmodel_params *p, float *energies,
                  char *lines, int stop_after_force)
{
    cm_consts c;
    cm_setup(p, &c);
    const float MASS = p->mass, DT = p->dt;
    size_t lpos = 0;
    if (lines) lines[0] = 0;
    for (int it = 0; it < p->iterations; ++it) {
        float System_Energy = 0;
        float System_Kinetic_Energy = 0;
        for (int ref = 0; ref < n; ++ref) {
            float Fx, Fy, Fz, E;
            cm_force_one(particle, n, ref, p, &c, &Fx, &Fy, &Fz, &E, NULL);
            float *pr = particle + 11 * ref;
            pr[9] = E;                                 /* :126 */
            pr[6] = Fx; pr[7] = Fy; pr[8] = Fz;        /* :128-130 */
            float ax = Fx / MASS, ay = Fy / MASS, az = Fz / MASS;   /* :133-135 */
            float vx = pr[3], vy = pr[4], vz = pr[5];
            float t;
            t = ax * DT; vx += t;                      /* :140-142 */
            t = ay * DT; vy += t;
            t = az * DT; vz += t;
            float v2 = vx * vx + vy * vy;              /* :144: 0.5*MASS in double */
            v2 = v2 + vz * vz;
            float Ek = (float)((0.5 * (double)MASS) * (double)v2);
            pr[10] = Ek;
            pr[3] = vx; pr[4] = vy; pr[5] = vz;
            System_Energy += E;                        /* :154-155 */
            System_Kinetic_Energy += Ek;
        }
        System_Energy = (float)((double)System_Energy * 0.5);      /* :158 */
        if (energies) {
            energies[2 * it + 0] = System_Energy;
            energies[2 * it + 1] = System_Kinetic_Energy;
        }
        if (lines) {
            lpos += (size_t)sprintf(lines + lpos,
                                    "Iteration %d, System energy is %f, Kinetic Energy is %f\n",
                                    it, (double)System_Energy, (double)System_Kinetic_Energy);
        }
        if (stop_after_force) return 1;
        for (int i = 0; i < n; ++i) {                  /* :164-202 */
            float *pr = particle + 11 * i;
            float t;
            float posx = pr[0], posy = pr[1], posz = pr[2];
            t = pr[3] * DT; posx += t;
            t = pr[4] * DT; posy += t;
            t = pr[5] * DT; posz += t;
            if (posx < 0) posx += p->box[0]; else if (posx >= p->box[0]) posx -= p->box[0];
            if (posy < 0) posy += p->box[1]; else if (posy >= p->box[1]) posy -= p->box[1];
            if (posz < 0) posz += p->box[2]; else if (posz >= p->box[2]) posz -= p->box[2];
            pr[0] = posx; pr[1] = posy; pr[2] = posz;
            pr[6] = 0; pr[7] = 0; pr[8] = 0;
        }
    }
    return 1;
}

int64_t or_cmodel_force_sample(const float *particle, int n, const or_cmodel_params *p,
                               int nref, float *fout)
{
    cm_consts c;
    cm_setup(p, &c);
    int64_t cnt = 0;
    if (nref > n) nref = n;
    for (int ref = 0; ref < nref; ++ref) {
        float Fx, Fy, Fz, E;
        cm_force_one(particle, n, ref, p, &c, &Fx, &Fy, &Fz, &E, &cnt);
        if (fout) { fout[3 * ref] = Fx; fout[3 * ref + 1] = Fy; fout[3 * ref + 2] = Fz; }
    }
    return cnt;
}

/* ------------------------------------------------------------------------- */
/* MATLAB binning and the pair stream                                         */
/* ------------------------------------------------------------------------- */

/* M/Simulation_Verification_LJArgon.m:143-175. */
int or_bin_ceil(const float *pos, int n, float rc, const int ncell[3], int *cell_xyz_out)
{
    int out = 0;
    for (int i = 0; i < n; ++i) {
        int c[3];
        for (int d = 0; d < 3; ++d) {
            float v = pos[3 * i + d];
            c[d] = (v != 0.0f) ? (int)ceilf(v / rc) : 1;   /* single / single, then ceil */
        }
        int ok = c[0] > 0 && c[0] <= ncell[0] && c[1] > 0 && c[1] <= ncell[1] &&
                 c[2] > 0 && c[2] <= ncell[2];
        for (int d = 0; d < 3; ++d) cell_xyz_out[3 * i + d] = ok ? c[d] : 0;
        if (!ok) ++out;
    }
    return out;
}

typedef struct { char *buf; int64_t cap, len; } sbuf;

static void sb_put(sbuf *s, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));
#include <stdarg.h>
static void sb_put(sbuf *s, const char *fmt, ...)
{
    char tmp[512];
    va_list ap;
    va_start(ap, fmt);
    int k = vsnprintf(tmp, sizeof tmp, fmt, ap);
    va_end(ap);
    if (s->len + k <= s->cap) memcpy(s->buf + s->len, tmp, (size_t)k);
    s->len += k;
}

/* M/Simulation_Verification_LJArgon.m:125-175 (load, min-shift, bin) and
 * :250-572 (filter reservoirs and the pair stream text). */
int64_t or_pair_stream_text(const double *raw_xyz, int n, float rc, const int ncell[3],
                            const int home[3], int cell_particle_max, char *buf,
                            int64_t cap)
{
    /* :136-143 min-shift in double, stored as single */
    double mn[3] = {raw_xyz[0], raw_xyz[1], raw_xyz[2]};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d)
            if (raw_xyz[3 * i + d] < mn[d]) mn[d] = raw_xyz[3 * i + d];
    float *pos = (float *)calloc(3 * (size_t)n, sizeof(float));
    int *cxyz = (int *)malloc(sizeof(int) * 3 * (size_t)n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) pos[3 * i + d] = (float)(raw_xyz[3 * i + d] - mn[d]);
    or_bin_ceil(pos, n, rc, ncell, cxyz);

    /* cell_particle(cell_id, k, 1:3): stable input order */
    int ncells = ncell[0] * ncell[1] * ncell[2];
    int *count = (int *)calloc((size_t)ncells, sizeof(int));
    float *cellpos = (float *)calloc((size_t)ncells * (size_t)cell_particle_max * 3, sizeof(float));
    for (int i = 0; i < n; ++i) {
        if (cxyz[3 * i] == 0) continue;
        int cid = ((cxyz[3 * i] - 1) * ncell[1] + (cxyz[3 * i + 1] - 1)) * ncell[2] + (cxyz[3 * i + 2] - 1);
        int k = count[cid]++;
        if (k < cell_particle_max) memcpy(cellpos + ((size_t)cid * cell_particle_max + k) * 3, pos + 3 * i, 12);
    }

    /* filter -> cells, :250-510 (offsets relative to the home cell) */
    static const int fmap[8][2][3] = {
        {{0, 0, 0}, {9, 9, 9}},   {{0, 0, 1}, {9, 9, 9}},
        {{0, 1, -1}, {0, 1, 0}},  {{0, 1, 1}, {1, -1, -1}},
        {{1, -1, 0}, {1, -1, 1}}, {{1, 0, -1}, {1, 0, 0}},
        {{1, 0, 1}, {1, 1, -1}},  {{1, 1, 0}, {1, 1, 1}},
    };
    int RES = 2 * cell_particle_max;
    float *res = (float *)calloc((size_t)8 * RES * 7, sizeof(float));
    int fnum[8] = {0};
    for (int f = 0; f < 8; ++f) {
        int base = 0;
        for (int part = 0; part < 2; ++part) {
            if (fmap[f][part][0] == 9) break;
            int cx = home[0] + fmap[f][part][0], cy = home[1] + fmap[f][part][1],
                cz = home[2] + fmap[f][part][2];
            int cid = ((cx - 1) * ncell[1] + (cy - 1)) * ncell[2] + (cz - 1);
            int m = count[cid];
            for (int k = 0; k < m; ++k) {
                float *r = res + ((size_t)f * RES + base + k) * 7;
                memcpy(r, cellpos + ((size_t)cid * cell_particle_max + k) * 3, 12);
                r[3] = (float)cx; r[4] = (float)cy; r[5] = (float)cz; r[6] = (float)(k + 1);
            }
            base += m;
        }
        fnum[f] = base;
    }
    int fmax = 0;
    for (int f = 0; f < 8; ++f) if (fnum[f] > fmax) fmax = fnum[f];

    sbuf s = {buf, cap, 0};
    sb_put(&s, "Ref ID\tNeighbor ID\t\t\t\t\t\t\t\t\tValid\tReference Particle Position\t\tNeighbor Particle Position\r\n");
    int hid = ((home[0] - 1) * ncell[1] + (home[1] - 1)) * ncell[2] + (home[2] - 1);
    for (int r = 0; r < count[hid]; ++r) {
        const float *rp = cellpos + ((size_t)hid * cell_particle_max + r) * 3;
        for (int k = 0; k < fmax; ++k) {
            sb_put(&s, "%d%d%d%2X\t", home[0], home[1], home[2], r + 1);
            float tmp[8][3];
            int valid[8];
            for (int f = 7; f >= 0; --f) {
                int src = (k < fnum[f]) ? k : fnum[f] - 1;
                valid[f] = k < fnum[f];
                memcpy(tmp[f], res + ((size_t)f * RES + src) * 7, 12);
                const float *idr = res + ((size_t)f * RES + k) * 7;
                sb_put(&s, "%d%d%d%2X", (int)idr[3], (int)idr[4], (int)idr[5], (unsigned)idr[6]);
            }
            sb_put(&s, "\t");
            for (int f = 7; f >= 0; --f) sb_put(&s, "%d", valid[f]);
            sb_put(&s, "\t%08X%08X%08X\t", f2u(rp[2]), f2u(rp[1]), f2u(rp[0]));
            for (int f = 7; f >= 0; --f)
                sb_put(&s, "%08X%08X%08X", f2u(tmp[f][2]), f2u(tmp[f][1]), f2u(tmp[f][0]));
            sb_put(&s, "\r\n");
        }
    }
    free(pos); free(cxyz); free(count); free(cellpos); free(res);
    return s.len <= cap ? s.len : -s.len;
}

/* ------------------------------------------------------------------------- */
/* half-shell table forces, int64 fixed point                                 */
/* ------------------------------------------------------------------------- */

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n >= 1) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_cell_of(float x, float inv_edge, int nc)
{
    float t = x * inv_edge;
    t = fminf(fmaxf(t, 0.0f), (float)(nc - 1));
    return (int)t;
}

int64_t or_fix(float f)
{
    double d = (double)f * 4294967296.0;   /* exact: 24-bit mantissa times 2^32 */
    if (d != d) return 0;
    if (d > 0x1p62) d = 0x1p62;
    if (d < -0x1p62) d = -0x1p62;
    return (int64_t)llrint(d);
}

/* Half-shell neighbour offsets (dx, dy, dz) of the reference's 13 neighbour cells
 * {223,231,232,233,311,312,313,321,322,323,331,332,333} around home 222,
 * D/Particle_Pair_Gen_HalfShell_64_Cells.v:213-226, preceded by the home cell. */
static const int HS[14][3] = {
    {0, 0, 0},  {0, 0, 1},  {0, 1, -1}, {0, 1, 0},  {0, 1, 1},
    {1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0},
    {1, 0, 1},  {1, 1, -1}, {1, 1, 0},  {1, 1, 1},
};

typedef struct {
    int ncells;
    int *start, *count, *list;
} cell_list;

static void build_cells(int n, const float *pos4, const float box[3], const int nc[3],
                        cell_list *cl)
{
    float inv[3];
    for (int d = 0; d < 3; ++d) inv[d] = (float)((double)nc[d] / (double)box[d]);
    cl->ncells = nc[0] * nc[1] * nc[2];
    cl->start = (int *)calloc((size_t)cl->ncells + 1, sizeof(int));
    cl->count = (int *)calloc((size_t)cl->ncells, sizeof(int));
    cl->list = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *cid = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        int cx = or_cell_of(pos4[4 * i + 0], inv[0], nc[0]);
        int cy = or_cell_of(pos4[4 * i + 1], inv[1], nc[1]);
        int cz = or_cell_of(pos4[4 * i + 2], inv[2], nc[2]);
        cid[i] = (cx * nc[1] + cy) * nc[2] + cz;
        cl->count[cid[i]]++;
    }
    for (int c = 0; c < cl->ncells; ++c) cl->start[c + 1] = cl->start[c] + cl->count[c];
    int *fill = (int *)calloc((size_t)(cl->ncells > 0 ? cl->ncells : 1), sizeof(int));
    for (int i = 0; i < n; ++i) cl->list[cl->start[cid[i]] + fill[cid[i]]++] = i;
    free(fill);
    free(cid);
}

static void free_cells(cell_list *cl) { free(cl->start); free(cl->count); free(cl->list); }

/* Scalar force of one accepted pair (the kernel's per-pair arithmetic). */
static inline float pair_scalar(const or_rl_system *s, float r2, int ti, int tj, float qq)
{
    float A = s->pair_a[ti * s->ntypes + tj];
    float B = s->pair_b[ti * s->ntypes + tj];
    const int par = s->mode & OR_MODE_PARITY;
    const int coul = s->mode & OR_MODE_COULOMB;
    if (s->mode & OR_MODE_ANALYTIC) {
        float ir2 = 1.0f / r2;
        float ir6 = ir2 * ir2;
        ir6 = ir6 * ir2;
        float ir8 = ir6 * ir2;
        float ir14 = ir8 * ir6;
        float F = fmaf(A, ir14, -(B * ir8));
        if (coul) {
            float ir3 = ir2 * sqrtf(ir2);
            ir3 = ir3 * 0.0796178343949044585987261f;   /* 1/(4*3.14) as float */
            F = fmaf(qq, ir3, F);
        }
        return F;
    }
    int idx = or_lut_index(r2, s->exp_offset, s->nseg);
    const float *c = s->tab + 6 * (size_t)idx;
    float F;
    if (par) {
        float t14 = c[1] * r2; t14 = t14 + c[0];
        float t8 = c[3] * r2;  t8 = t8 + c[2];
        float a = A * t14, b = B * t8;
        F = a - b;
        if (coul) {
            float t3 = c[5] * r2; t3 = t3 + c[4];
            float q3 = qq * t3;
            F = F + q3;
        }
    } else {
        float t14 = fmaf(c[1], r2, c[0]);
        float t8 = fmaf(c[3], r2, c[2]);
        F = fmaf(A, t14, -(B * t8));
        if (coul) {
            float t3 = fmaf(c[5], r2, c[4]);
            F = fmaf(qq, t3, F);
        }
    }
    return F;
}

/* Exclusion window: bit d of m[i] excludes the pair (i, i + d - 16); windows are
 * symmetric, so testing either atom's window gives the same answer. */
static inline int or_excluded(const uint32_t *m, int i, int j)
{
    const int d = j - i + 16;
    return (unsigned)d < 32u && ((m[i] >> d) & 1u);
}

/* The analytic mode's cutoff decided exactly: D = (x_i - t) - e per axis in FP64, with
 * t = fl(x_j + S) the staged image and e = (x_j + S) - t its exact error, so D is the exact
 * separation x_i - (x_j + S) (both differences exact in FP64 for positions in [0, L) and S
 * in {-L, 0, L}); r^2 = (Dx^2 + Dy^2) + Dz^2 against rc^2 in FP64.  This is the membership of
 * the FP64 direct evaluation, so the pair set carries no FP32 rounding. */
static inline int or_within_cutoff_exact(const float *pi, const float t[3], const float e[3],
                                         double rc2e)
{
    double D[3];
    for (int d = 0; d < 3; ++d) D[d] = ((double)pi[d] - (double)t[d]) - (double)e[d];
    double e2 = D[0] * D[0] + D[1] * D[1];
    e2 = e2 + D[2] * D[2];
    return e2 <= rc2e;
}

int64_t or_rl_forces_fix(const or_rl_system *s, int64_t *fix_out)
{
    const int n = s->n;
    cell_list cl;
    build_cells(n, s->pos4, s->box, s->ncell, &cl);
    memset(fix_out, 0, sizeof(int64_t) * 3 * (size_t)n);
    const int par = s->mode & OR_MODE_PARITY;
    const int ana = s->mode & OR_MODE_ANALYTIC;
    const double rc2e = (double)s->rc * (double)s->rc;
    int64_t npairs = 0;
    const int *nc = s->ncell;
    /* Home cells run in parallel, each thread into its own int64 accumulators; integer
     * sums are order-free, so the result does not depend on the thread count. */
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    uint64_t *priv = NULL;
    if (nthr > 1) priv = (uint64_t *)calloc((size_t)nthr * 3 * (size_t)n, sizeof(uint64_t));
    if (!priv) nthr = 1;
    #pragma omp parallel for schedule(dynamic, 1) reduction(+:npairs) num_threads(nthr) if(nthr > 1)
    for (int h = 0; h < cl.ncells; ++h) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        uint64_t *acc = priv ? priv + (size_t)tid * 3 * (size_t)n : (uint64_t *)fix_out;
        const int hx = h / (nc[1] * nc[2]), hy = (h / nc[2]) % nc[1], hz = h % nc[2];
        for (int o = 0; o < 14; ++o) {
            int raw[3] = {hx + HS[o][0], hy + HS[o][1], hz + HS[o][2]};
            float S[3];
            int nb[3];
            for (int d = 0; d < 3; ++d) {
                if (raw[d] < 0) { nb[d] = raw[d] + nc[d]; S[d] = -s->box[d]; }
                else if (raw[d] >= nc[d]) { nb[d] = raw[d] - nc[d]; S[d] = s->box[d]; }
                else { nb[d] = raw[d]; S[d] = 0.0f; }
            }
            int c = (nb[0] * nc[1] + nb[1]) * nc[2] + nb[2];
            for (int a = 0; a < cl.count[h]; ++a) {
                int i = cl.list[cl.start[h] + a];
                const float *pi = s->pos4 + 4 * (size_t)i;
                int b0 = (o == 0) ? a + 1 : 0;
                for (int b = b0; b < cl.count[c]; ++b) {
                    int j = cl.list[cl.start[c] + b];
                    const float *pj = s->pos4 + 4 * (size_t)j;
                    float dx, dy, dz, r2;
                    if (ana) {   /* analytic takes precedence over parity, as on the GPU */
                        /* t = fl(x_j + S) as staged, e its exact error, dx = fl((x_i - t) - e) */
                        float t[3], e[3], d[3];
                        for (int k = 0; k < 3; ++k) {
                            t[k] = pj[k] + S[k];
                            const float z = t[k] - S[k];
                            e[k] = pj[k] - z;
                            d[k] = pi[k] - t[k];
                            d[k] = d[k] - e[k];
                        }
                        dx = d[0]; dy = d[1]; dz = d[2];
                        r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        if (!(r2 > 0.0f && or_within_cutoff_exact(pi, t, e, rc2e))) continue;
                    } else if (par) {
                        dx = pi[0] - pj[0]; dx = dx - S[0];
                        dy = pi[1] - pj[1]; dy = dy - S[1];
                        dz = pi[2] - pj[2]; dz = dz - S[2];
                        r2 = dx * dx + dy * dy;
                        r2 = r2 + dz * dz;
                    } else {
                        float xj = pj[0] + S[0], yj = pj[1] + S[1], zj = pj[2] + S[2];
                        dx = pi[0] - xj; dy = pi[1] - yj; dz = pi[2] - zj;
                        r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    }
                    if (!ana && !(r2 > 0.0f && r2 <= s->rc2)) continue;
                    if (s->excl && or_excluded(s->excl, i, j)) continue;
                    ++npairs;
                    float qq = pi[3] * pj[3];
                    float F = pair_scalar(s, r2, s->type[i], s->type[j], qq);
                    int64_t kx = or_fix(F * dx), ky = or_fix(F * dy), kz = or_fix(F * dz);
                    acc[3 * i + 0] += (uint64_t)kx; acc[3 * j + 0] -= (uint64_t)kx;
                    acc[3 * i + 1] += (uint64_t)ky; acc[3 * j + 1] -= (uint64_t)ky;
                    acc[3 * i + 2] += (uint64_t)kz; acc[3 * j + 2] -= (uint64_t)kz;
                }
            }
        }
    }
    if (priv) {
        uint64_t *out = (uint64_t *)fix_out;
        #pragma omp parallel for schedule(static)
        for (long long k = 0; k < 3LL * n; ++k) {
            uint64_t v = 0;
            for (int t = 0; t < nthr; ++t) v += priv[(size_t)t * 3 * (size_t)n + (size_t)k];
            out[k] = v;
        }
        free(priv);
    }
    free_cells(&cl);
    return npairs;
}

/* FP64 evaluation of the same table model: FP32 coefficients, everything else in
 * double, the same pair set (exact images). */
int64_t or_table_fp64(const or_rl_system *s, double *f_out)
{
    const int n = s->n;
    cell_list cl;
    build_cells(n, s->pos4, s->box, s->ncell, &cl);
    memset(f_out, 0, sizeof(double) * 3 * (size_t)n);
    int64_t npairs = 0;
    const int *nc = s->ncell;
    for (int h = 0; h < cl.ncells; ++h) {
        int hx = h / (nc[1] * nc[2]), hy = (h / nc[2]) % nc[1], hz = h % nc[2];
        for (int o = 0; o < 14; ++o) {
            int raw[3] = {hx + HS[o][0], hy + HS[o][1], hz + HS[o][2]};
            double S[3];
            int nb[3];
            for (int d = 0; d < 3; ++d) {
                if (raw[d] < 0) { nb[d] = raw[d] + nc[d]; S[d] = -(double)s->box[d]; }
                else if (raw[d] >= nc[d]) { nb[d] = raw[d] - nc[d]; S[d] = s->box[d]; }
                else { nb[d] = raw[d]; S[d] = 0.0; }
            }
            int c = (nb[0] * nc[1] + nb[1]) * nc[2] + nb[2];
            for (int a = 0; a < cl.count[h]; ++a) {
                int i = cl.list[cl.start[h] + a];
                const float *pi = s->pos4 + 4 * (size_t)i;
                int b0 = (o == 0) ? a + 1 : 0;
                for (int b = b0; b < cl.count[c]; ++b) {
                    int j = cl.list[cl.start[c] + b];
                    const float *pj = s->pos4 + 4 * (size_t)j;
                    double dx = (double)pi[0] - (double)pj[0] - S[0];
                    double dy = (double)pi[1] - (double)pj[1] - S[1];
                    double dz = (double)pi[2] - (double)pj[2] - S[2];
                    double r2 = dx * dx + dy * dy + dz * dz;
                    if (!(r2 > 0.0 && r2 <= (double)s->rc2)) continue;
                    if (s->excl && or_excluded(s->excl, i, j)) continue;
                    ++npairs;
                    int idx = or_lut_index((float)r2, s->exp_offset, s->nseg);
                    const float *cf = s->tab + 6 * (size_t)idx;
                    int ti = s->type[i], tj = s->type[j];
                    double A = s->pair_a[ti * s->ntypes + tj], B = s->pair_b[ti * s->ntypes + tj];
                    double F = A * ((double)cf[1] * r2 + cf[0]) - B * ((double)cf[3] * r2 + cf[2]);
                    if (s->mode & OR_MODE_COULOMB)
                        F += (double)pi[3] * (double)pj[3] * ((double)cf[5] * r2 + cf[4]);
                    F *= 0.1;   /* kJ/mol/nm multipliers -> kJ/mol/A */
                    f_out[3 * i] += F * dx; f_out[3 * j] -= F * dx;
                    f_out[3 * i + 1] += F * dy; f_out[3 * j + 1] -= F * dy;
                    f_out[3 * i + 2] += F * dz; f_out[3 * j + 2] -= F * dz;
                }
            }
        }
    }
    free_cells(&cl);
    return npairs;
}

/* ------------------------------------------------------------------------- */
/* FP64 analytic direct oracle                                                */
/* ------------------------------------------------------------------------- */

static inline void lj_coul_fp64(double r2, double eps, double sig, double qq, double ke,
                                int coulomb, double *F, double *E)
{
    double ir2 = 1.0 / r2;
    double s2 = sig * sig * ir2;
    double s6 = s2 * s2 * s2;
    double s12 = s6 * s6;
    *F = (48.0 * eps * s12 - 24.0 * eps * s6) * ir2;   /* -dU/dr / r */
    *E = 4.0 * eps * (s12 - s6);
    if (coulomb) {
        double ir = sqrt(ir2);
        *F += ke * qq * ir * ir2;
        *E += ke * qq * ir;
    }
}

int64_t or_fp64_direct(int n, const double *xyz, const int *type, const double *q,
                       int ntypes, const double *eps_ij, const double *sig_ij,
                       double ke, int coulomb, const double box[3], double rc,
                       int use_cells, double *f_out, double *energy, double *virial,
                       const uint32_t *excl)
{
    const double rc2 = rc * rc;
    int64_t npairs = 0;
    double etot = 0.0, vir[9] = {0};
    memset(f_out, 0, sizeof(double) * 3 * (size_t)n);
    if (!use_cells) {
        /* every image within rc: image range per axis */
        int im[3];
        for (int d = 0; d < 3; ++d) im[d] = (int)ceil(rc / box[d]) + 1;
        for (int i = 0; i < n; ++i) {
            double fi[3] = {0, 0, 0};
            for (int j = 0; j < n; ++j) {
                for (int ax = -im[0]; ax <= im[0]; ++ax)
                for (int ay = -im[1]; ay <= im[1]; ++ay)
                for (int az = -im[2]; az <= im[2]; ++az) {
                    if (i == j && ax == 0 && ay == 0 && az == 0) continue;
                    if (excl && or_excluded(excl, i, j)) continue;
                    double dx = xyz[3 * i] - xyz[3 * j] - ax * box[0];
                    double dy = xyz[3 * i + 1] - xyz[3 * j + 1] - ay * box[1];
                    double dz = xyz[3 * i + 2] - xyz[3 * j + 2] - az * box[2];
                    double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 > rc2 || r2 <= 0.0) continue;
                    int tt = type[i] * ntypes + type[j];
                    double F, E;
                    lj_coul_fp64(r2, eps_ij[tt], sig_ij[tt], q ? q[i] * q[j] : 0.0, ke,
                                 coulomb, &F, &E);
                    fi[0] += F * dx; fi[1] += F * dy; fi[2] += F * dz;
                    etot += 0.5 * E;
                    double d3[3] = {dx, dy, dz};
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) vir[3 * a + b] += 0.5 * F * d3[a] * d3[b];
                    ++npairs;
                }
            }
            f_out[3 * i] = fi[0]; f_out[3 * i + 1] = fi[1]; f_out[3 * i + 2] = fi[2];
        }
        npairs /= 2;
    } else {
        int nc[3];
        for (int d = 0; d < 3; ++d) { nc[d] = (int)floor(box[d] / rc); if (nc[d] < 1) nc[d] = 1; }
        int ncells = nc[0] * nc[1] * nc[2];
        int *start = (int *)calloc((size_t)ncells + 1, sizeof(int));
        int *cid = (int *)malloc(sizeof(int) * (size_t)n);
        int *list = (int *)malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; ++i) {
            int c3[3];
            for (int d = 0; d < 3; ++d) {
                double x = xyz[3 * i + d] - box[d] * floor(xyz[3 * i + d] / box[d]);
                int c = (int)(x / box[d] * nc[d]);
                if (c >= nc[d]) c = nc[d] - 1;
                if (c < 0) c = 0;
                c3[d] = c;
            }
            cid[i] = (c3[0] * nc[1] + c3[1]) * nc[2] + c3[2];
            start[cid[i] + 1]++;
        }
        for (int c = 0; c < ncells; ++c) start[c + 1] += start[c];
        int *fill = (int *)calloc((size_t)ncells, sizeof(int));
        for (int i = 0; i < n; ++i) list[start[cid[i]] + fill[cid[i]]++] = i;
        free(fill);
        double et = 0.0;
        int64_t np = 0;
        /* full shell: every atom gathers from the 27 neighbour cell images */
        #pragma omp parallel for schedule(dynamic, 64) reduction(+:et, np, vir[:9])
        for (int i = 0; i < n; ++i) {
            int c = cid[i];
            int cx = c / (nc[1] * nc[2]), cy = (c / nc[2]) % nc[1], cz = c % nc[2];
            double xi[3];
            for (int d = 0; d < 3; ++d) xi[d] = xyz[3 * i + d] - box[d] * floor(xyz[3 * i + d] / box[d]);
            double fi[3] = {0, 0, 0};
            for (int ox = -1; ox <= 1; ++ox)
            for (int oy = -1; oy <= 1; ++oy)
            for (int oz = -1; oz <= 1; ++oz) {
                int r[3] = {cx + ox, cy + oy, cz + oz}, nb[3];
                double S[3];
                for (int d = 0; d < 3; ++d) {
                    if (r[d] < 0) { nb[d] = r[d] + nc[d]; S[d] = -box[d]; }
                    else if (r[d] >= nc[d]) { nb[d] = r[d] - nc[d]; S[d] = box[d]; }
                    else { nb[d] = r[d]; S[d] = 0.0; }
                }
                int cc = (nb[0] * nc[1] + nb[1]) * nc[2] + nb[2];
                for (int k = start[cc]; k < start[cc + 1]; ++k) {
                    int j = list[k];
                    double xj[3];
                    for (int d = 0; d < 3; ++d) xj[d] = xyz[3 * j + d] - box[d] * floor(xyz[3 * j + d] / box[d]) + S[d];
                    double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
                    double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 > rc2 || r2 <= 0.0) continue;
                    if (excl && or_excluded(excl, i, j)) continue;
                    int tt = type[i] * ntypes + type[j];
                    double F, E;
                    lj_coul_fp64(r2, eps_ij[tt], sig_ij[tt], q ? q[i] * q[j] : 0.0, ke,
                                 coulomb, &F, &E);
                    fi[0] += F * dx; fi[1] += F * dy; fi[2] += F * dz;
                    et += 0.5 * E;
                    const double d3[3] = {dx, dy, dz};
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) vir[3 * a + b] += 0.5 * F * d3[a] * d3[b];
                    ++np;
                }
            }
            f_out[3 * i] = fi[0]; f_out[3 * i + 1] = fi[1]; f_out[3 * i + 2] = fi[2];
        }
        etot = et;
        npairs = np / 2;
        free(start); free(cid); free(list);
    }
    if (energy) *energy = etot;
    if (virial) memcpy(virial, vir, sizeof vir);
    return npairs;
}

/* ------------------------------------------------------------------------- */
/* Motion_Update                                                              */
/* ------------------------------------------------------------------------- */

static inline float wrap1(float x, float L)
{
    if (x < 0.0f) x = x + L;
    if (x >= L) x = x - L;
    if (!(x >= 0.0f && x < L)) {
        x = x - L * floorf(x / L);
        if (!(x >= 0.0f && x < L)) x = 0.0f;
    }
    return x;
}

void or_minimize_update(int n, float *pos4, const int64_t *fix, double gamma, double dmax,
                        const float box[3])
{
    for (int i = 0; i < n; ++i) {
        double f[3];
        for (int d = 0; d < 3; ++d) f[d] = (double)fix[3 * i + d] * (1.0 / 4294967296.0);
        double fm = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        double sc = (fm * gamma > dmax) ? dmax / fm : gamma;
        for (int d = 0; d < 3; ++d) {
            float step = (float)(f[d] * sc);
            float x = pos4[4 * i + d] + step;
            pos4[4 * i + d] = wrap1(x, box[d]);
        }
    }
}

void or_motion_update(int n, float *pos4, float *vel3, const int64_t *fix,
                      const int *type, const double *hk, const double *mass, float dt,
                      const float box[3], int first, int kick_drift, double *ke_out)
{
    double ke = 0.0;
    for (int i = 0; i < n; ++i) {
        double k = hk[type[i]];
        float v2 = 0.0f;
        double vv = 0.0;
        for (int d = 0; d < 3; ++d) {
            float dv = (float)((double)fix[3 * i + d] * k);
            float v = vel3[3 * i + d];
            float vn;
            if (kick_drift) {
                float dv2 = dv + dv;
                vn = v + dv2;
                vv += (double)vn * (double)vn;
            } else {
                float vf = first ? v : v + dv;
                vv += (double)vf * (double)vf;
                vn = vf + dv;
            }
            float step = vn * dt;
            float x = pos4[4 * i + d] + step;
            pos4[4 * i + d] = wrap1(x, box[d]);
            vel3[3 * i + d] = vn;
        }
        (void)v2;
        ke += 0.5 * mass[type[i]] * vv * 1.0e4;
    }
    if (ke_out) *ke_out = ke;
}

/* ---- bonded terms ---------------------------------------------------------------- */

/* Minimum image of one FP32 component (positions are wrapped into [0, L)). */
static inline float img_f(float d, float L)
{
    const float h = 0.5f * L;
    if (d > h) return d - L;
    if (d < -h) return d + L;
    return d;
}
static inline float dot3f(const float *a, const float *b)
{
    float r = a[0] * b[0] + a[1] * b[1];
    return r + a[2] * b[2];
}
static inline void cross3f(const float *a, const float *b, float *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* atan2 for y >= 0: reduce to a = min/max in [0, 1]; above tan(pi/12) use
 * atan a = pi/6 + atan((a sqrt3 - 1)/(a + sqrt3)), leaving |t| <= tan(pi/12) = 0.268 for
 * the Taylor series t - t^3/3 + ... - t^11/11 (next term < 3e-9). */
float or_atan2f(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    const float a = mn / mx;
    float t = a, off = 0.0f;
    if (a > 0.26794919f) {
        const float num = a * 1.7320508f - 1.0f;
        const float den = a + 1.7320508f;
        t = num / den;
        off = 0.52359878f;
    }
    const float t2 = t * t;
    float p = -0.09090909f;          /* -1/11 */
    p = p * t2 + 0.11111111f;        /*  1/9  */
    p = p * t2 - 0.14285715f;        /* -1/7  */
    p = p * t2 + 0.2f;               /*  1/5  */
    p = p * t2 - 0.33333334f;        /* -1/3  */
    p = p * t2 + 1.0f;
    float r = p * t;
    r = off + r;
    if (ay > ax) r = 1.5707964f - r;
    if (x < 0.0f) r = 3.1415927f - r;
    return y < 0.0f ? -r : r;
}

static inline void to_fix(const float *f, int64_t *k)
{
    k[0] = or_fix(f[0]); k[1] = or_fix(f[1]); k[2] = or_fix(f[2]);
}
static inline void add_k(int64_t *fix, int a, const int64_t *k)
{
    uint64_t *u = (uint64_t *)fix;
    u[3 * (size_t)a + 0] += (uint64_t)k[0];
    u[3 * (size_t)a + 1] += (uint64_t)k[1];
    u[3 * (size_t)a + 2] += (uint64_t)k[2];
}
static inline void add_fix(int64_t *fix, int a, const float *f)
{
    int64_t k[3];
    to_fix(f, k);
    add_k(fix, a, k);
}

void or_bonded_fix(int n, const float *pos4, const float box[3], int nterms,
                   const int32_t *atoms, const float *params, int64_t *fix_out,
                   double *term_energy)
{
    (void)n;
    for (int t = 0; t < nterms; ++t) {
        const int32_t *at = atoms + 4 * (size_t)t;
        const float *pp = params + 4 * (size_t)t;
        const float *x1 = pos4 + 4 * (size_t)at[0], *x2 = pos4 + 4 * (size_t)at[1];
        double e = 0.0;
        if (at[2] < 0) {
            /* bond, Bonded/calc_bond_v2.v:83-110: r12 = x1 - x2, dpot/dr = 2 k (r - r0),
             * f1 = (-dpot/dr / r) r12, f2 = -f1 */
            float r12[3];
            for (int d = 0; d < 3; ++d) r12[d] = img_f(x1[d] - x2[d], box[d]);
            const float r = sqrtf(dot3f(r12, r12));
            const float dr = r - pp[0];
            float dp = 2.0f * pp[1];
            dp = dp * dr;
            const float c = -dp / r;
            float f1[3], f2[3];
            for (int d = 0; d < 3; ++d) { f1[d] = c * r12[d]; f2[d] = -f1[d]; }
            add_fix(fix_out, at[0], f1);
            add_fix(fix_out, at[1], f2);
            const double ddr = (double)r - (double)pp[0];
            e = (double)pp[1] / 10.0 * ddr * ddr;
        } else if (at[3] < 0) {
            /* angle, Bonded/calc_angle_v2.v:113-147 (no Urey-Bradley term):
             * theta = atan2(|r12 x r32|, r12 . r32), dpot/dtheta = 2 k (theta - theta0),
             * dtheta1 = (r^12 cos - r^32) / (sin d12), dtheta3 = (r^32 cos - r^12) / (sin d32),
             * f1 = -dpot dtheta1, f3 = -dpot dtheta3, f2 = -f1 - f3 */
            const float *x3 = pos4 + 4 * (size_t)at[2];
            float r12[3], r32[3], cr[3];
            for (int d = 0; d < 3; ++d) {
                r12[d] = img_f(x1[d] - x2[d], box[d]);
                r32[d] = img_f(x3[d] - x2[d], box[d]);
            }
            const float d12 = sqrtf(dot3f(r12, r12)), d32 = sqrtf(dot3f(r32, r32));
            cross3f(r12, r32, cr);
            const float y = sqrtf(dot3f(cr, cr)), xx = dot3f(r12, r32);
            const float th = or_atan2f(y, xx);
            const float dd = d12 * d32;
            const float cs = xx / dd, sn = y / dd;
            {
                double dth = (double)th - (double)pp[0];
                e = (double)pp[1] / 10.0 * dth * dth;
            }
            if (sn > 1e-6f) {
                float dp = 2.0f * pp[1];
                dp = dp * (th - pp[0]);
                const float c = -dp;
                const float a1 = sn * d12, a3 = sn * d32;
                float f1[3], f2[3], f3[3];
                for (int d = 0; d < 3; ++d) {
                    const float u12 = r12[d] / d12, u32 = r32[d] / d32;
                    float g1 = u12 * cs - u32;
                    g1 = g1 / a1;
                    float g3 = u32 * cs - u12;
                    g3 = g3 / a3;
                    f1[d] = c * g1;
                    f3[d] = c * g3;
                    f2[d] = -f1[d] - f3[d];
                }
                /* the vertex takes minus the sum of the converted end forces, so the
                 * term's fixed-point forces sum to zero exactly */
                int64_t k1[3], k3[3], k2[3];
                to_fix(f1, k1);
                to_fix(f3, k3);
                for (int d = 0; d < 3; ++d) k2[d] = (int64_t)(0ull - (uint64_t)k1[d] - (uint64_t)k3[d]);
                (void)f2;
                add_k(fix_out, at[0], k1);
                add_k(fix_out, at[1], k2);
                add_k(fix_out, at[2], k3);
            }
        } else {
            /* torsion, Bonded/calcTorsion_v2.v:70-274 potential k (1 + cos(n phi + d)):
             * rij = xi - xj, rkj = xk - xj, rkl = xk - xl, m = rij x rkj, v = rkj x rkl,
             * cos phi = m.v / (|m||v|), sin phi = |rkj| (rij . v) / (|m||v|),
             * dU/dphi = -k n sin(n phi + d); fi = -dU/dphi |rkj| / |m|^2 m,
             * fl = dU/dphi |rkj| / |v|^2 v, s = (rij.rkj / |rkj|^2) fi - (rkl.rkj / |rkj|^2) fl,
             * Fi += fi, Fj -= fi - s, Fk -= fl + s, Fl += fl */
            const float *x3 = pos4 + 4 * (size_t)at[2], *x4 = pos4 + 4 * (size_t)at[3];
            float rij[3], rkj[3], rkl[3], m[3], v[3];
            for (int d = 0; d < 3; ++d) {
                rij[d] = img_f(x1[d] - x2[d], box[d]);
                rkj[d] = img_f(x3[d] - x2[d], box[d]);
                rkl[d] = img_f(x3[d] - x4[d], box[d]);
            }
            cross3f(rij, rkj, m);
            cross3f(rkj, rkl, v);
            const float m2 = dot3f(m, m), v2 = dot3f(v, v);
            if (m2 > 1e-12f && v2 > 1e-12f) {
                const float kj2 = dot3f(rkj, rkj);
                const float kj = sqrtf(kj2);
                const float mv = sqrtf(m2) * sqrtf(v2);
                const float cp = dot3f(m, v) / mv;
                float sp = kj * dot3f(rij, v);
                sp = sp / mv;
                const int nper = (int)pp[1];
                float cn = cp, sn = sp;
                for (int k = 1; k < nper; ++k) {
                    const float c2 = cn * cp - sn * sp;
                    const float s2 = sn * cp + cn * sp;
                    cn = c2;
                    sn = s2;
                }
                float sarg = sn * pp[2] + cn * pp[3];      /* sin(n phi + d) */
                float carg = cn * pp[2] - sn * pp[3];      /* cos(n phi + d) */
                float ddphi = -pp[0] * pp[1];
                ddphi = ddphi * sarg;
                float a = -ddphi * kj;
                a = a / m2;
                float b = ddphi * kj;
                b = b / v2;
                const float p = dot3f(rij, rkj) / kj2, q = dot3f(rkl, rkj) / kj2;
                float fi[3], fl[3], fj[3], fk[3];
                for (int d = 0; d < 3; ++d) {
                    fi[d] = a * m[d];
                    fl[d] = b * v[d];
                    const float sv = p * fi[d] - q * fl[d];
                    fj[d] = -(fi[d] - sv);
                    fk[d] = -(fl[d] + sv);
                }
                /* k takes minus the sum of the other three converted forces */
                int64_t qi[3], qj[3], qk[3], ql[3];
                to_fix(fi, qi);
                to_fix(fj, qj);
                to_fix(fl, ql);
                for (int d = 0; d < 3; ++d)
                    qk[d] = (int64_t)(0ull - (uint64_t)qi[d] - (uint64_t)qj[d] - (uint64_t)ql[d]);
                (void)fk;
                add_k(fix_out, at[0], qi);
                add_k(fix_out, at[1], qj);
                add_k(fix_out, at[2], qk);
                add_k(fix_out, at[3], ql);
                e = (double)pp[0] / 10.0 * (1.0 + (double)carg);
            }
        }
        if (term_energy) term_energy[t] = e;
    }
}

/* ---- FP64 reference of the bonded potentials ---- */
static inline double img_d(double d, double L)
{
    return d - L * floor(d / L + 0.5);
}
static inline double dot3d(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3d(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

double or_bonded_fp64(int n, const double *xyz, const double box[3], int nterms,
                      const int32_t *atoms, const double *params, double *f_out)
{
    (void)n;
    double etot = 0.0;
    for (int t = 0; t < nterms; ++t) {
        const int32_t *at = atoms + 4 * (size_t)t;
        const double *pp = params + 4 * (size_t)t;
        const double *x1 = xyz + 3 * (size_t)at[0], *x2 = xyz + 3 * (size_t)at[1];
        if (at[2] < 0) {
            double r12[3];
            for (int d = 0; d < 3; ++d) r12[d] = img_d(x1[d] - x2[d], box[d]);
            const double r = sqrt(dot3d(r12, r12));
            const double dp = 2.0 * pp[1] * (r - pp[0]);
            for (int d = 0; d < 3; ++d) {
                f_out[3 * at[0] + d] += -dp / r * r12[d];
                f_out[3 * at[1] + d] -= -dp / r * r12[d];
            }
            etot += pp[1] * (r - pp[0]) * (r - pp[0]);
        } else if (at[3] < 0) {
            const double *x3 = xyz + 3 * (size_t)at[2];
            double r12[3], r32[3], cr[3];
            for (int d = 0; d < 3; ++d) {
                r12[d] = img_d(x1[d] - x2[d], box[d]);
                r32[d] = img_d(x3[d] - x2[d], box[d]);
            }
            const double d12 = sqrt(dot3d(r12, r12)), d32 = sqrt(dot3d(r32, r32));
            cross3d(r12, r32, cr);
            const double y = sqrt(dot3d(cr, cr)), xx = dot3d(r12, r32);
            const double th = atan2(y, xx);
            const double cs = cos(th), sn = sin(th);
            const double dp = 2.0 * pp[1] * (th - pp[0]);
            etot += pp[1] * (th - pp[0]) * (th - pp[0]);
            if (sn > 1e-12) {
                for (int d = 0; d < 3; ++d) {
                    const double g1 = (r12[d] / d12 * cs - r32[d] / d32) / (sn * d12);
                    const double g3 = (r32[d] / d32 * cs - r12[d] / d12) / (sn * d32);
                    f_out[3 * at[0] + d] += -dp * g1;
                    f_out[3 * at[2] + d] += -dp * g3;
                    f_out[3 * at[1] + d] += dp * g1 + dp * g3;
                }
            }
        } else {
            const double *x3 = xyz + 3 * (size_t)at[2], *x4 = xyz + 3 * (size_t)at[3];
            double rij[3], rkj[3], rkl[3], m[3], v[3];
            for (int d = 0; d < 3; ++d) {
                rij[d] = img_d(x1[d] - x2[d], box[d]);
                rkj[d] = img_d(x3[d] - x2[d], box[d]);
                rkl[d] = img_d(x3[d] - x4[d], box[d]);
            }
            cross3d(rij, rkj, m);
            cross3d(rkj, rkl, v);
            const double m2 = dot3d(m, m), v2 = dot3d(v, v), kj2 = dot3d(rkj, rkj);
            const double kj = sqrt(kj2);
            const double phi = atan2(kj * dot3d(rij, v), dot3d(m, v));
            const double nper = pp[1], ph = pp[2];
            etot += pp[0] * (1.0 + cos(nper * phi + ph));
            const double ddphi = -pp[0] * nper * sin(nper * phi + ph);
            const double p = dot3d(rij, rkj) / kj2, q = dot3d(rkl, rkj) / kj2;
            for (int d = 0; d < 3; ++d) {
                const double fi = -ddphi * kj / m2 * m[d], fl = ddphi * kj / v2 * v[d];
                const double sv = p * fi - q * fl;
                f_out[3 * at[0] + d] += fi;
                f_out[3 * at[1] + d] -= fi - sv;
                f_out[3 * at[2] + d] -= fl + sv;
                f_out[3 * at[3] + d] += fl;
            }
        }
    }
    return etot;
}
