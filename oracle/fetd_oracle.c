/*
 * fetd_oracle.c -- CPU restatement of TEMF3DT's FETD forward path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see the header.
 *
 * Reference files (yxddyxzh2016/TEMF3DT):
 *   src/FETD_FWD.f90      model_mesh, model_ctl, mesh_process, adap_euler,
 *                         sys_right, solution_output, p_in_tet, e/f_kind_iso
 *   src/FETD_UTILITY.f90  det_mkl_3/4, two_error, interp_lin, parseCode,
 *                         pardiso_multi_real_mpi, source_mod
 */
#define _GNU_SOURCE
#include "fetd_oracle.h"

#include <ctype.h>
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* src/FETD_FWD.f90:14-16 */
static const double PI = 3.141592653589793238462643383279502884197;
#define MU0 (PI * 4.0e-7)
static const double EPS0 = 8.8541878176e-12;

/* local edge -> local vertex pair, src/FETD_FWD.f90:23 (1-based there) */
static const int IND_TET[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};

static char g_err[512];
const char *fo_error(void) { return g_err; }
static void set_err(const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* PARDISO through libmkl_rt (src/FETD_UTILITY.f90:826-1028)           */
/* ------------------------------------------------------------------ */
typedef void (*pardiso_fn_t)(void *pt, const int *maxfct, const int *mnum, const int *mtype,
                             const int *phase, const int *n, const double *a, const int *ia,
                             const int *ja, int *perm, const int *nrhs, int *iparm,
                             const int *msglvl, double *b, double *x, int *error);
typedef void (*mkl_set_threads_t)(int);
static pardiso_fn_t g_pardiso = NULL;
static mkl_set_threads_t g_mkl_threads = NULL;
static int g_mkl_tried = 0;

static int load_mkl(void) {
    if (g_mkl_tried) return g_pardiso != NULL;
    g_mkl_tried = 1;
    /* MKL shares the GNU OpenMP runtime with this library */
    setenv("MKL_THREADING_LAYER", "GNU", 0);
    const char *cands[] = {"libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so",
                           "/opt/conda/lib/libmkl_rt.so.1", NULL};
    void *h = NULL;
    for (int i = 0; cands[i] && !h; i++) h = dlopen(cands[i], RTLD_NOW | RTLD_GLOBAL);
    if (!h) return 0;
    g_pardiso = (pardiso_fn_t)dlsym(h, "pardiso");
    g_mkl_threads = (mkl_set_threads_t)dlsym(h, "MKL_Set_Num_Threads");
    return g_pardiso != NULL;
}

int fo_pardiso_available(void) { return load_mkl(); }

void fo_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
    if (load_mkl() && g_mkl_threads && n > 0) g_mkl_threads(n);
}

typedef struct {
    void *pt[64];
    int iparm[64];
    int mtype, n, live;
    const int *ia, *ja;
    const double *va;
    long long nnzL;
} pslot;

/* pardiso_direct_1: phase 12, mtype 2 then -2 (src/FETD_UTILITY.f90:879-936) */
static int pslot_factor(pslot *s, int n, const int *ia, const int *ja, const double *va) {
    int maxfct = 1, mnum = 1, phase = 12, nrhs = 1, msglvl = 0, error = 0, idum = 0;
    double ddum = 0;
    memset(s->pt, 0, sizeof s->pt);
    memset(s->iparm, 0, sizeof s->iparm);
    s->mtype = 2;
    s->n = n;
    s->ia = ia;
    s->ja = ja;
    s->va = va;
    g_pardiso(s->pt, &maxfct, &mnum, &s->mtype, &phase, &n, va, ia, ja, &idum, &nrhs, s->iparm,
              &msglvl, &ddum, &ddum, &error);
    if (error != 0) {
        int ph = -1, e2 = 0;
        g_pardiso(s->pt, &maxfct, &mnum, &s->mtype, &ph, &n, &ddum, &idum, &idum, &idum, &nrhs,
                  s->iparm, &msglvl, &ddum, &ddum, &e2);
        memset(s->pt, 0, sizeof s->pt);
        memset(s->iparm, 0, sizeof s->iparm);
        s->mtype = -2;
        g_pardiso(s->pt, &maxfct, &mnum, &s->mtype, &phase, &n, va, ia, ja, &idum, &nrhs,
                  s->iparm, &msglvl, &ddum, &ddum, &error);
        if (error != 0) return error;
        fprintf(stderr, " WARNING - SYSTEM MATRIX IS INDEFINITE, CHANGE TO INDEFINITE MODE.\n");
    }
    s->live = 1;
    s->nnzL = (long long)s->iparm[17];
    return 0;
}

/* pardiso_direct_2: phase 33 (src/FETD_UTILITY.f90:938-957) */
static int pslot_solve(pslot *s, double *b, double *x, int nrhs) {
    int maxfct = 1, mnum = 1, phase = 33, msglvl = 0, error = 0, idum = 0;
    g_pardiso(s->pt, &maxfct, &mnum, &s->mtype, &phase, &s->n, s->va, s->ia, s->ja, &idum, &nrhs,
              s->iparm, &msglvl, b, x, &error);
    return error;
}

/* pardiso_direct_3: phase -1 (src/FETD_UTILITY.f90:959-975) */
static void pslot_release(pslot *s) {
    if (!s->live) return;
    int maxfct = 1, mnum = 1, phase = -1, nrhs = 1, msglvl = 0, error = 0, idum = 0;
    double ddum = 0;
    g_pardiso(s->pt, &maxfct, &mnum, &s->mtype, &phase, &s->n, &ddum, &idum, &idum, &idum, &nrhs,
              s->iparm, &msglvl, &ddum, &ddum, &error);
    s->live = 0;
}

int fo_pardiso_solve_once(int n, const int *ia, const int *ja, const double *va, const double *b,
                          double *x, int nrhs) {
    if (!load_mkl()) return -9999;
    pslot s;
    memset(&s, 0, sizeof s);
    int err = pslot_factor(&s, n, ia, ja, va);
    if (err) return err;
    double *bb = malloc(sizeof(double) * (size_t)n * nrhs);
    memcpy(bb, b, sizeof(double) * (size_t)n * nrhs);
    err = pslot_solve(&s, bb, x, nrhs);
    free(bb);
    pslot_release(&s);
    return err;
}

/* ------------------------------------------------------------------ */
/* model                                                               */
/* ------------------------------------------------------------------ */
struct fo_model {
    int verbose;
    /* mesh (src/FETD_FWD.f90:33-106 'type model') */
    int nN;
    double *xyz; /* [N][3] */
    int nseg;
    int *seg_vex; /* [seg][2] */
    int *seg_phy;
    int nT;
    int *tet_vex; /* [T][4] */
    int *tet_phy;
    int nE;
    int *edg_vex; /* [E][2] */
    double *edg_len;
    int *tet_edg; /* [T][6] 1-based */
    int *tet_drc; /* [T][6] */
    double *ca, *cb, *cc, *cd; /* [T][4] */
    double *vol;
    double *tet_len; /* [T][6] */
    /* control */
    int have_ctl;
    int nR;
    double rec_charac;
    double *rec_xyz;
    int *rec_tet; /* 1-based */
    double *n_arr, *d_arr; /* [R][6][3] */
    int ndom;
    double *sig, *per, *mag; /* per tet */
    int aniso;               /* anisotropic media: per-tet 3x3 tensors (row-major) */
    double *sig9, *nu9;      /* conductivity tensor; curl weight = elementwise 1/mu tensor */
    int bound_cond;
    int src_num, src_type, pulse_type;
    double pulse_width, current;
    int *src_edg, *src_drc; /* [seg] */
    int step_mthd, step_dvsi, step_dble, step_dblesize;
    double time_maxm;
    int o_f[9];
    int n_gate;
    double *gates;
    /* assembled upper CSR, 1-based, shared pattern (src/FETD_FWD.f90:696-709) */
    int nnz;
    int *ia, *ja;
    double *av, *bv, *cv; /* eps-mass, sigma-mass, curl-curl */
    int *ia0, *ja0; double *bfull; /* full symmetric copy of B for threaded SpMV */
    /* output rows */
    int nrows, cap_rows;
    int *row_step;
    double *row_t, *row_dt, *row_f; /* row_f [row][R][9] */
    int nchk, cap_chk;
    int *chk_step, *chk_keep;
    double *chk_err;
    double timers[6];
    int nfact;
    long long factor_nnz;
    int nreplay;
    int *replay; /* forced keep/change decisions of the doubling checks */
    /* small-loop variant (src_small/FETD_FWD.f90:314-328,1526,1540) */
    int sig_rescale;
    double sig_parameter;
    /* per-tet field slices (vtkout, src/FETD_FWD.f90:476-484,1642-1711) */
    int nslice;
    double *slice_t;      /* requested times */
    int *slice_done, *slice_step;
    double *slice_tt;     /* time of the output row that took the slice */
    double *slice_f;      /* [slice][T][9]: E, curl E, running sum of curl E * dt */
    double *tet_ed;       /* [T][6] current E and curl E */
    double *tet_b;        /* [T][3] */
};

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) {
        fprintf(stderr, "fetd_oracle: out of memory\n");
        abort();
    }
    return p;
}

void fo_free(fo_model *m) {
    if (!m) return;
    void *ptrs[] = {m->xyz, m->seg_vex, m->seg_phy, m->tet_vex, m->tet_phy, m->edg_vex,
                    m->edg_len, m->tet_edg, m->tet_drc, m->ca, m->cb, m->cc, m->cd, m->vol,
                    m->tet_len, m->rec_xyz, m->rec_tet, m->n_arr, m->d_arr, m->sig, m->per,
                    m->mag, m->src_edg, m->src_drc, m->gates, m->ia, m->ja, m->av, m->bv,
                    m->cv, m->ia0, m->ja0, m->bfull, m->row_step, m->row_t, m->row_dt,
                    m->row_f, m->chk_step, m->chk_keep, m->chk_err, m->replay, m->sig9, m->nu9,
                    m->slice_t, m->slice_done, m->slice_step, m->slice_tt, m->slice_f, m->tet_ed, m->tet_b};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
    free(m);
}

/* ------------------------------------------------------------------ */
/* small dense kernels                                                  */
/* ------------------------------------------------------------------ */

/* det_mkl_3 / det_mkl_4 (src/FETD_UTILITY.f90:65-103): LU with partial
 * pivoting (first maximum), reciprocal scaling, det = prod(diag) with a
 * sign flip per row interchange; a zero pivot (info>0) gives 0. */
double fo_det_lu(const double *acm, int n) {
    double a[4][4];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) a[i][j] = acm[j * n + i];
    int piv[4];
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = fabs(a[k][k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(a[i][k]) > mx) {
                mx = fabs(a[i][k]);
                p = i;
            }
        piv[k] = p;
        if (a[p][k] != 0.0) {
            if (p != k)
                for (int j = 0; j < n; j++) {
                    double t = a[k][j];
                    a[k][j] = a[p][j];
                    a[p][j] = t;
                }
            double r = 1.0 / a[k][k];
            for (int i = k + 1; i < n; i++) a[i][k] *= r;
        } else if (info == 0) {
            info = k + 1;
        }
        for (int i = k + 1; i < n; i++)
            for (int j = k + 1; j < n; j++) a[i][j] -= a[i][k] * a[k][j];
    }
    if (info != 0) return 0.0;
    double det = 1.0;
    for (int i = 0; i < n; i++) det = (piv[i] != i) ? -det * a[i][i] : det * a[i][i];
    return det;
}

static double det3_rows(const double r0[3], const double r1[3], const double r2[3]) {
    double cm[9] = {r0[0], r1[0], r2[0], r0[1], r1[1], r2[1], r0[2], r1[2], r2[2]};
    return fo_det_lu(cm, 3);
}

/* cofactors a,b,c,d and signed volume (src/FETD_FWD.f90:245-277,
 * tet_volume src/FETD_UTILITY.f90:105-112) */
void fo_tet_cofactors(const double *x, const double *y, const double *z, double *a, double *b,
                      double *c, double *d, double *vol) {
    static const int others[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    static const double sa[4] = {1, -1, 1, -1};
    for (int k = 0; k < 4; k++) {
        const int *o = others[k];
        double r[3][3];
        for (int q = 0; q < 3; q++) { r[q][0] = x[o[q]]; r[q][1] = y[o[q]]; r[q][2] = z[o[q]]; }
        a[k] = det3_rows(r[0], r[1], r[2]) * sa[k];
        for (int q = 0; q < 3; q++) { r[q][0] = 1.0; r[q][1] = y[o[q]]; r[q][2] = z[o[q]]; }
        b[k] = det3_rows(r[0], r[1], r[2]) * (-sa[k]);
        for (int q = 0; q < 3; q++) { r[q][0] = 1.0; r[q][1] = x[o[q]]; r[q][2] = z[o[q]]; }
        c[k] = det3_rows(r[0], r[1], r[2]) * sa[k];
        for (int q = 0; q < 3; q++) { r[q][0] = 1.0; r[q][1] = x[o[q]]; r[q][2] = y[o[q]]; }
        d[k] = det3_rows(r[0], r[1], r[2]) * (-sa[k]);
    }
    double cm[16] = {1, 1, 1, 1, x[0], x[1], x[2], x[3], y[0], y[1], y[2], y[3], z[0], z[1], z[2], z[3]};
    *vol = fo_det_lu(cm, 4) / 6.0;
}

static double dist3(const double *p, const double *q) {
    double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* f_kind_iso (src/FETD_FWD.f90:826-851) and e_kind_iso (:797-824) from
 * cofactors; F is the scalar-weighted Whitney mass, K the curl-curl. */
static void elem_mats(const double *a, const double *b, const double *c, const double *d,
                      double vol, const double *len, double s_mass, double s_curl, double *F,
                      double *K) {
    (void)a;
    double g[6][3];
    for (int i = 0; i < 6; i++) {
        int i1 = IND_TET[i][0], i2 = IND_TET[i][1];
        g[i][0] = c[i1] * d[i2] - c[i2] * d[i1];
        g[i][1] = b[i2] * d[i1] - b[i1] * d[i2];
        g[i][2] = b[i1] * c[i2] - b[i2] * c[i1];
    }
    double v6 = 6.0 * vol;
    double v64 = v6 * v6 * v6 * v6;
    for (int i = 0; i < 6; i++) {
        for (int j = 0; j < 6; j++) {
            int im[2] = {IND_TET[i][0], IND_TET[i][1]};
            int jn[2] = {IND_TET[j][0], IND_TET[j][1]};
            double f[2][2];
            for (int mm = 0; mm < 2; mm++)
                for (int nn = 0; nn < 2; nn++)
                    f[mm][nn] = b[im[mm]] * b[jn[nn]] + c[im[mm]] * c[jn[nn]] + d[im[mm]] * d[jn[nn]];
            double v = f[1][1] * (1 + (im[0] == jn[0])) - f[0][1] * (1 + (im[1] == jn[0])) -
                       f[1][0] * (1 + (im[0] == jn[1])) + f[0][0] * (1 + (im[1] == jn[1]));
            F[i * 6 + j] = len[i] * len[j] / (720.0 * vol) * v * s_mass;
            double dot = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
            K[i * 6 + j] = 4.0 * len[i] * len[j] * vol / v64 * dot * s_curl;
        }
    }
}

void fo_tet_matrices(const double *x, const double *y, const double *z, double s_mass,
                     double s_curl, double *F, double *K) {
    double a[4], b[4], c[4], d[4], vol, len[6];
    fo_tet_cofactors(x, y, z, a, b, c, d, &vol);
    for (int j = 0; j < 6; j++) {
        int p = IND_TET[j][0], q = IND_TET[j][1];
        double P[3] = {x[p], y[p], z[p]}, Q[3] = {x[q], y[q], z[q]};
        len[j] = dist3(P, Q);
    }
    elem_mats(a, b, c, d, vol, len, s_mass, s_curl, F, K);
}

/* f_kind_ani (src/FETD_FWD.f90:884-912) and e_kind_ani (:853-882): the
 * same Whitney integrals with a 3x3 tensor A (row-major) between the
 * gradients / curls.  Operation order follows the reference (scalar 1). */
static void elem_mats_ani(const double *b, const double *c, const double *d, double vol,
                          const double *len, const double *Am, const double *Ak, double *F,
                          double *K) {
    double g[6][3];
    double v6 = 6.0 * vol;
    for (int i = 0; i < 6; i++) {
        int i1 = IND_TET[i][0], i2 = IND_TET[i][1];
        double s = 2.0 * len[i] / (v6 * v6);
        g[i][0] = (c[i1] * d[i2] - c[i2] * d[i1]) * s;
        g[i][1] = (b[i2] * d[i1] - b[i1] * d[i2]) * s;
        g[i][2] = (b[i1] * c[i2] - b[i2] * c[i1]) * s;
    }
    for (int i = 0; i < 6; i++) {
        for (int j = 0; j < 6; j++) {
            int im[2] = {IND_TET[i][0], IND_TET[i][1]};
            int jn[2] = {IND_TET[j][0], IND_TET[j][1]};
            double f[2][2];
            for (int mm = 0; mm < 2; mm++)
                for (int nn = 0; nn < 2; nn++) {
                    double v1[3] = {b[im[mm]], c[im[mm]], d[im[mm]]};
                    double v2[3] = {b[jn[nn]], c[jn[nn]], d[jn[nn]]};
                    double w[3];
                    for (int r = 0; r < 3; r++) w[r] = Am[3 * r] * v2[0] + Am[3 * r + 1] * v2[1] + Am[3 * r + 2] * v2[2];
                    f[mm][nn] = v1[0] * w[0] + v1[1] * w[1] + v1[2] * w[2];
                }
            double v = f[1][1] * (1 + (im[0] == jn[0])) - f[0][1] * (1 + (im[1] == jn[0])) -
                       f[1][0] * (1 + (im[0] == jn[1])) + f[0][0] * (1 + (im[1] == jn[1]));
            F[i * 6 + j] = len[i] * len[j] / (vol * 720.0) * v;
            double w[3];
            for (int r = 0; r < 3; r++) w[r] = Ak[3 * r] * g[j][0] + Ak[3 * r + 1] * g[j][1] + Ak[3 * r + 2] * g[j][2];
            K[i * 6 + j] = vol * (g[i][0] * w[0] + g[i][1] * w[1] + g[i][2] * w[2]);
        }
    }
}

void fo_tet_matrices_ani(const double *x, const double *y, const double *z, const double *A_mass,
                         const double *A_curl, double *F, double *K) {
    double a[4], b[4], c[4], d[4], vol, len[6];
    fo_tet_cofactors(x, y, z, a, b, c, d, &vol);
    for (int j = 0; j < 6; j++) {
        int p = IND_TET[j][0], q = IND_TET[j][1];
        double P[3] = {x[p], y[p], z[p]}, Q[3] = {x[q], y[q], z[q]};
        len[j] = dist3(P, Q);
    }
    elem_mats_ani(b, c, d, vol, len, A_mass, A_curl, F, K);
}

/* model_ctl anisotropic tensors (src/FETD_FWD.f90:368-410): principal
 * values v[3] rotated by the angles ang[3] (degrees) as R diag(v) R^T with
 * R = Rx Ry Rz.  As in the reference, the (2,2) entry of Rz takes the cosine
 * of the second angle. */
void fo_aniso_tensor(const double *v, const double *ang_deg, double scale, double *A) {
    const double d2r = 3.141592653589793238462643383279502884197 / 180.0;
    double a1 = ang_deg[0] * d2r, a2 = ang_deg[1] * d2r, a3 = ang_deg[2] * d2r;
    double dx[3][3] = {{1.0, 0.0, 0.0}, {0.0, cos(a1), -sin(a1)}, {0.0, sin(a1), cos(a1)}};
    double dy[3][3] = {{cos(a2), 0.0, sin(a2)}, {0.0, 1.0, 0.0}, {-sin(a2), 0.0, cos(a2)}};
    double dz[3][3] = {{cos(a3), -sin(a3), 0.0}, {sin(a3), cos(a2), 0.0}, {0.0, 0.0, 1.0}};
    double t[3][3], dm[3][3], u[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int k = 0; k < 3; k++) acc += dx[i][k] * dy[k][j];
            t[i][j] = acc;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int k = 0; k < 3; k++) acc += t[i][k] * dz[k][j];
            dm[i][j] = acc;
        }
    for (int i = 0; i < 3; i++)  /* dm * diag(v) */
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int k = 0; k < 3; k++) acc += dm[i][k] * (k == j ? v[j] : 0.0);
            u[i][j] = acc;
        }
    for (int i = 0; i < 3; i++)  /* (dm diag(v)) dm^T, times scale */
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int k = 0; k < 3; k++) acc += u[i][k] * dm[j][k];
            A[3 * i + j] = acc * scale;
        }
}

/* source_mod (src/FETD_UTILITY.f90:1030-1166) */
static double erfstepon1(double t, double t0) {
    double u = t - t0 / 2.0;
    return (1.0 / sqrt(3.141592653589793) * exp(1.0 / (t0 * t0) * (u * u) * (-25.0)) * 5.0) / t0;
}
static double deltafunc1(double t, double t0) {
    double w = t * 5.436563656918091 - t0 * 2.718281828459046;
    return -1.0 / (t0 * t0) * exp(-1.0 / (t0 * t0) * (w * w)) *
           (t * 59.11244879144521 - t0 * 29.55622439572261);
}
double fo_source_didt(int source_type, int pulse_type, double t, double t0, double cur) {
    if (source_type == 1 && pulse_type == 1) return erfstepon1(t, t0) * cur;
    if (source_type == 1 && pulse_type == 2) return deltafunc1(t, t0) * cur;
    if (source_type == 1 && pulse_type == 3) return -erfstepon1(t, t0) * cur;
    return NAN;
}

/* ------------------------------------------------------------------ */
/* input parsing                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    FILE *f;
    char line[4096];
} lreader;

static int next_line(lreader *r) {
    if (!fgets(r->line, sizeof r->line, r->f)) return 0;
    for (char *p = r->line; *p; p++) {
        if (*p == '\r' || *p == '\n') { *p = 0; break; }
        if (*p == '\t') *p = ' ';
    }
    return 1;
}

static double fnum(const char *s, char **end) {
    char buf[64];
    size_t k = 0;
    while (*s == ' ' || *s == ',') s++;
    const char *st = s;
    while (*s && *s != ' ' && *s != ',' && k < sizeof buf - 1) {
        char ch = *s++;
        buf[k++] = (ch == 'd' || ch == 'D') ? 'e' : ch;
    }
    buf[k] = 0;
    char *e2;
    double v = strtod(buf, &e2);
    if (end) *end = (e2 == buf) ? (char *)st : (char *)s;
    return v;
}

/* Fortran list-directed read of n numbers from the next record(s). */
static int read_nums(lreader *r, double *out, int n) {
    int got = 0;
    while (got < n) {
        if (!next_line(r)) return 0;
        char *p = r->line;
        while (got < n) {
            char *e;
            while (*p == ' ' || *p == ',') p++;
            if (!*p) break;
            double v = fnum(p, &e);
            if (e == p) break;
            out[got++] = v;
            p = e;
        }
    }
    return 1;
}

/* MEDIT reader (src/FETD_FWD.f90:126-167): sections keyed by the first 5
 * characters of the first token. */
static int read_mesh(fo_model *m, const char *path) {
    lreader r;
    r.f = fopen(path, "r");
    if (!r.f) {
        snprintf(g_err, sizeof g_err, "cannot open mesh file %s", path);
        return -1;
    }
    while (next_line(&r)) {
        char word[6] = {0};
        const char *p = r.line;
        while (*p == ' ') p++;
        if (!*p) continue;
        int k = 0;
        while (p[k] && p[k] != ' ' && k < 5) { word[k] = p[k]; k++; }
        double v[6];
        if (!strcmp(word, "Verti")) {
            if (!read_nums(&r, v, 1)) break;
            m->nN = (int)v[0];
            m->xyz = xcalloc((size_t)m->nN * 3, sizeof(double));
            for (int i = 0; i < m->nN; i++) {
                if (!read_nums(&r, v, 3)) goto bad;
                m->xyz[3 * i] = v[0]; m->xyz[3 * i + 1] = v[1]; m->xyz[3 * i + 2] = v[2];
            }
        } else if (!strcmp(word, "Edges")) {
            if (!read_nums(&r, v, 1)) break;
            int ns = (int)v[0];
            int *tmp = xcalloc((size_t)ns * 3, sizeof(int));
            int keep = 0;
            for (int i = 0; i < ns; i++) {
                if (!read_nums(&r, v, 3)) goto bad;
                tmp[3 * i] = (int)v[0]; tmp[3 * i + 1] = (int)v[1]; tmp[3 * i + 2] = (int)v[2];
                if (tmp[3 * i + 2] != 0) keep++;
            }
            m->nseg = keep;
            m->seg_vex = xcalloc((size_t)keep * 2, sizeof(int));
            m->seg_phy = xcalloc((size_t)keep, sizeof(int));
            keep = 0;
            for (int i = 0; i < ns; i++)
                if (tmp[3 * i + 2] != 0) {
                    m->seg_vex[2 * keep] = tmp[3 * i];
                    m->seg_vex[2 * keep + 1] = tmp[3 * i + 1];
                    m->seg_phy[keep] = tmp[3 * i + 2];
                    keep++;
                }
            free(tmp);
        } else if (!strcmp(word, "Trian")) {
            if (!read_nums(&r, v, 1)) break;
            int nt = (int)v[0];
            for (int i = 0; i < nt; i++)
                if (!read_nums(&r, v, 3)) goto bad;
        } else if (!strcmp(word, "Tetra")) {
            if (!read_nums(&r, v, 1)) break;
            m->nT = (int)v[0];
            m->tet_vex = xcalloc((size_t)m->nT * 4, sizeof(int));
            m->tet_phy = xcalloc((size_t)m->nT, sizeof(int));
            for (int i = 0; i < m->nT; i++) {
                if (!read_nums(&r, v, 5)) goto bad;
                for (int q = 0; q < 4; q++) m->tet_vex[4 * i + q] = (int)v[q];
                m->tet_phy[i] = (int)v[4];
            }
        }
    }
    fclose(r.f);
    if (m->nN == 0 || m->nT == 0) {
        set_err("mesh file has no Vertices or Tetrahedra");
        return -1;
    }
    return 0;
bad:
    fclose(r.f);
    set_err("truncated mesh file");
    return -1;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* edge numbering + tet_edg/tet_drc + geometry (src/FETD_FWD.f90:170-281) */
static void build_topology(fo_model *m) {
    size_t np = (size_t)m->nT * 6;
    uint64_t *key = xcalloc(np, sizeof(uint64_t));
    for (int i = 0; i < m->nT; i++)
        for (int j = 0; j < 6; j++) {
            uint32_t v1 = (uint32_t)m->tet_vex[4 * i + IND_TET[j][0]];
            uint32_t v2 = (uint32_t)m->tet_vex[4 * i + IND_TET[j][1]];
            if (v1 > v2) { uint32_t t = v1; v1 = v2; v2 = t; }
            key[(size_t)i * 6 + j] = ((uint64_t)v1 << 32) | v2;
        }
    qsort(key, np, sizeof(uint64_t), cmp_u64);
    size_t ne = 0;
    for (size_t i = 0; i < np; i++)
        if (i == 0 || key[i] != key[i - 1]) key[ne++] = key[i];
    m->nE = (int)ne;
    m->edg_vex = xcalloc(ne * 2, sizeof(int));
    m->edg_len = xcalloc(ne, sizeof(double));
    for (size_t e = 0; e < ne; e++) {
        int v1 = (int)(key[e] >> 32), v2 = (int)(key[e] & 0xffffffffu);
        m->edg_vex[2 * e] = v1;
        m->edg_vex[2 * e + 1] = v2;
        m->edg_len[e] = dist3(&m->xyz[3 * (v1 - 1)], &m->xyz[3 * (v2 - 1)]);
    }
    m->tet_edg = xcalloc(np, sizeof(int));
    m->tet_drc = xcalloc(np, sizeof(int));
    m->ca = xcalloc((size_t)m->nT * 4, sizeof(double));
    m->cb = xcalloc((size_t)m->nT * 4, sizeof(double));
    m->cc = xcalloc((size_t)m->nT * 4, sizeof(double));
    m->cd = xcalloc((size_t)m->nT * 4, sizeof(double));
    m->vol = xcalloc((size_t)m->nT, sizeof(double));
    m->tet_len = xcalloc(np, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m->nT; i++) {
        const int *tv = &m->tet_vex[4 * i];
        for (int j = 0; j < 6; j++) {
            uint32_t va = (uint32_t)tv[IND_TET[j][0]], vb = (uint32_t)tv[IND_TET[j][1]];
            uint32_t v1 = va < vb ? va : vb, v2 = va < vb ? vb : va;
            uint64_t k = ((uint64_t)v1 << 32) | v2;
            uint64_t *hit = bsearch(&k, key, ne, sizeof(uint64_t), cmp_u64);
            m->tet_edg[6 * i + j] = (int)(hit - key) + 1;
            m->tet_drc[6 * i + j] = (va < vb) ? 1 : -1;
        }
        double x[4], y[4], z[4];
        for (int q = 0; q < 4; q++) {
            x[q] = m->xyz[3 * (tv[q] - 1)];
            y[q] = m->xyz[3 * (tv[q] - 1) + 1];
            z[q] = m->xyz[3 * (tv[q] - 1) + 2];
        }
        fo_tet_cofactors(x, y, z, &m->ca[4 * i], &m->cb[4 * i], &m->cc[4 * i], &m->cd[4 * i],
                         &m->vol[i]);
        for (int j = 0; j < 6; j++) {
            int p = IND_TET[j][0], q = IND_TET[j][1];
            double P[3] = {x[p], y[p], z[p]}, Q[3] = {x[q], y[q], z[q]};
            m->tet_len[6 * i + j] = dist3(P, Q);
        }
    }
    free(key);
}

/* parseCode (src/FETD_UTILITY.f90:678-731): returns 1 with lower-cased key
 * and value when the line holds "key: value", 0 otherwise. */
static int parse_code(const char *line, char *code, char *value) {
    const char *p = line;
    while (*p == ' ') p++;
    if (!*p || *p == '%' || *p == '!') return 0;
    const char *colon = strchr(line, ':');
    if (!colon || colon < p) return 0;
    size_t n = (size_t)(colon - p);
    memcpy(code, p, n);
    code[n] = 0;
    while (n > 0 && code[n - 1] == ' ') code[--n] = 0;
    for (char *q = code; *q; q++) *q = (char)tolower((unsigned char)*q);
    const char *v = colon + 1;
    while (*v == ' ') v++;
    strcpy(value, v);
    char *c1 = strpbrk(value, "%!");
    if (c1) *c1 = 0;
    return 1;
}

/* p_in_tet (src/FETD_FWD.f90:1009-1049) */
static int p_in_tet(const fo_model *m, const double *p) {
    for (int i = 0; i < m->nT; i++) {
        const int *tv = &m->tet_vex[4 * i];
        double x[4], y[4], z[4];
        for (int q = 0; q < 4; q++) {
            x[q] = m->xyz[3 * (tv[q] - 1)];
            y[q] = m->xyz[3 * (tv[q] - 1) + 1];
            z[q] = m->xyz[3 * (tv[q] - 1) + 2];
        }
        double mx = (x[0] + x[1] + x[2] + x[3]) / 4.0, my = (y[0] + y[1] + y[2] + y[3]) / 4.0,
               mz = (z[0] + z[1] + z[2] + z[3]) / 4.0;
        double dd = sqrt((mx - p[0]) * (mx - p[0]) + (my - p[1]) * (my - p[1]) +
                         (mz - p[2]) * (mz - p[2]));
        if (dd > m->rec_charac) continue;
        double det[5];
        for (int s = 0; s < 5; s++) {
            double X[4] = {x[0], x[1], x[2], x[3]}, Y[4] = {y[0], y[1], y[2], y[3]},
                   Z[4] = {z[0], z[1], z[2], z[3]};
            if (s > 0) { X[s - 1] = p[0]; Y[s - 1] = p[1]; Z[s - 1] = p[2]; }
            double cm[16] = {X[0], X[1], X[2], X[3], Y[0], Y[1], Y[2], Y[3],
                             Z[0], Z[1], Z[2], Z[3], 1, 1, 1, 1};
            det[s] = fo_det_lu(cm, 4);
        }
        int allp = 1, alln = 1;
        for (int s = 0; s < 5; s++) {
            if (!(det[s] >= 0.0)) allp = 0;
            if (!(det[s] <= 0.0)) alln = 0;
        }
        if (allp || alln) return i + 1;
    }
    return 0;
}

static int read_cntl(fo_model *m, const char *path) {
    lreader r;
    r.f = fopen(path, "r");
    if (!r.f) {
        snprintf(g_err, sizeof g_err, "cannot open control file %s", path);
        return -1;
    }
    m->rec_charac = 10.0;
    int *dom_ind = NULL;
    double *dsig = NULL, *dper = NULL, *dmag = NULL;
    char code[4096], value[4096];
    double v[16];
    while (next_line(&r)) {
        if (!parse_code(r.line, code, value)) continue;
        if (!strcmp(code, "time_points")) {
            m->n_gate = (int)fnum(value, NULL);
            m->gates = xcalloc((size_t)m->n_gate, sizeof(double));
            for (int i = 0; i < m->n_gate; i++) {
                if (!read_nums(&r, v, 1)) goto bad;
                m->gates[i] = v[0];
            }
        } else if (!strcmp(code, "rec_charac")) {
            m->rec_charac = fnum(value, NULL);
        } else if (!strcmp(code, "receivers")) {
            m->nR = (int)fnum(value, NULL);
            m->rec_xyz = xcalloc((size_t)m->nR * 3, sizeof(double));
            for (int i = 0; i < m->nR; i++) {
                if (!read_nums(&r, v, 3)) goto bad;
                memcpy(&m->rec_xyz[3 * i], v, 3 * sizeof(double));
            }
        } else if (!strcmp(code, "isotropic")) {
            m->ndom = (int)fnum(value, NULL);
            dom_ind = xcalloc((size_t)m->ndom, sizeof(int));
            dsig = xcalloc((size_t)m->ndom, sizeof(double));
            dper = xcalloc((size_t)m->ndom, sizeof(double));
            dmag = xcalloc((size_t)m->ndom, sizeof(double));
            for (int i = 0; i < m->ndom; i++) {
                if (!read_nums(&r, v, 4)) goto bad;
                dom_ind[i] = (int)v[0]; dsig[i] = v[1]; dper[i] = v[2]; dmag[i] = v[3];
            }
            if (m->sig_rescale) {
                /* src_small/FETD_FWD.f90:314-328: the parameter is the last
                 * domain's sigma (the branch meant to take the minimum is
                 * unreachable), inverted; domains 2..n are scaled by it */
                double sp = 1.0;
                for (int i = 1; i < m->ndom; i++) sp = dsig[i];
                sp = 1.0 / sp;
                for (int i = 1; i < m->ndom; i++) dsig[i] = dsig[i] * sp;
                m->sig_parameter = sp;
            }
        } else if (!strcmp(code, "anisotropic")) {
            /* three records per domain: sigma, permittivity, permeability,
             * each "domain v1 v2 v3 angle1 angle2 angle3" */
            m->aniso = 1;
            m->ndom = (int)fnum(value, NULL);
            dom_ind = xcalloc((size_t)m->ndom, sizeof(int));
            dsig = xcalloc((size_t)m->ndom * 18, sizeof(double)); /* sig v+ang, mag v+ang, per v+ang */
            for (int i = 0; i < m->ndom; i++) {
                for (int q = 0; q < 3; q++) {
                    if (!read_nums(&r, v, 7)) goto bad;
                    dom_ind[i] = (int)v[0];
                    memcpy(&dsig[18 * i + 6 * q], &v[1], 6 * sizeof(double));
                }
            }
        } else if (!strcmp(code, "boundary condition")) {
            m->bound_cond = (int)fnum(value, NULL);
        } else if (!strcmp(code, "source")) {
            m->src_num = (int)fnum(value, NULL);
            if (!read_nums(&r, v, 4)) goto bad;
            m->src_type = (int)v[0]; m->pulse_type = (int)v[1];
            m->pulse_width = v[2]; m->current = v[3];
        } else if (!strcmp(code, "sp_mthd")) {
            m->step_mthd = (int)fnum(value, NULL);
        } else if (!strcmp(code, "sp_division")) {
            m->step_dvsi = (int)fnum(value, NULL);
        } else if (!strcmp(code, "time_maximum")) {
            m->time_maxm = fnum(value, NULL);
        } else if (!strcmp(code, "sp_double")) {
            m->step_dble = (int)fnum(value, NULL);
        } else if (!strcmp(code, "sp_dblesize")) {
            m->step_dblesize = (int)fnum(value, NULL);
        } else if (!strcmp(code, "output")) {
            if (!read_nums(&r, v, 9)) goto bad;
            for (int i = 0; i < 9; i++) m->o_f[i] = ((int)v[i] == 1);
        } else if (!strcmp(code, "vtkout")) {
            int nv = (int)fnum(value, NULL);
            for (int i = 0; i < nv; i++)
                if (!read_nums(&r, v, 1)) goto bad;
        }
    }
    fclose(r.f);
    /* required keys (src/FETD_FWD.f90:490-497) */
    int any = 0;
    for (int i = 0; i < 9; i++) any |= m->o_f[i];
    if (m->nR == 0) { set_err("cmod%rec_num error"); goto bad3; }
    if (m->ndom == 0) { set_err("cmod%dom_num error"); goto bad3; }
    if (m->src_num == 0) { set_err("cmod%src_num error"); goto bad3; }
    if (m->step_mthd == 0) { set_err("cmod%step_mthd error"); goto bad3; }
    if (m->step_dvsi == 0) { set_err("cmod%step_dvsi error"); goto bad3; }
    if (fabs(m->time_maxm) < 1e-20) { set_err("cmod%time_maxm error"); goto bad3; }
    if (m->step_dble == 0) { set_err("cmod%step_dble error"); goto bad3; }
    if (!any) { set_err("cmod%o_f error"); goto bad3; }
    if (m->aniso) { /* tensors per tet (src/FETD_FWD.f90:381-408) */
        m->sig9 = xcalloc((size_t)m->nT * 9, sizeof(double));
        m->nu9 = xcalloc((size_t)m->nT * 9, sizeof(double));
        m->sig = xcalloc((size_t)m->nT, sizeof(double));
        m->per = xcalloc((size_t)m->nT, sizeof(double));
        m->mag = xcalloc((size_t)m->nT, sizeof(double));
        for (int i = 0; i < m->nT; i++) {
            int found = 0;
            for (int j = 0; j < m->ndom; j++)
                if (m->tet_phy[i] == dom_ind[j]) {
                    found = 1;
                    double mg[9];
                    fo_aniso_tensor(&dsig[18 * j], &dsig[18 * j + 3], 1.0, &m->sig9[9 * i]);
                    fo_aniso_tensor(&dsig[18 * j + 12], &dsig[18 * j + 15], MU0, mg);
                    /* the curl weight is the entrywise reciprocal, zero entries kept (:624-632) */
                    for (int q = 0; q < 9; q++) m->nu9[9 * i + q] = (mg[q] == 0.0) ? 0.0 : 1.0 / mg[q];
                    m->sig[i] = m->sig9[9 * i];
                    m->mag[i] = mg[0];
                }
            if (!found) { set_err("error - physical domain not found!"); goto bad3; }
        }
        free(dom_ind); free(dsig);
        dom_ind = NULL; dsig = NULL;
        goto materials_done;
    }
    /* materials per tet (src/FETD_FWD.f90:347-364) */
    m->sig = xcalloc((size_t)m->nT, sizeof(double));
    m->per = xcalloc((size_t)m->nT, sizeof(double));
    m->mag = xcalloc((size_t)m->nT, sizeof(double));
    for (int i = 0; i < m->nT; i++) {
        int found = 0;
        for (int j = 0; j < m->ndom; j++)
            if (m->tet_phy[i] == dom_ind[j]) {
                found = 1;
                m->sig[i] = dsig[j];
                m->per[i] = dper[j] * EPS0;
                m->mag[i] = dmag[j] * MU0;
            }
        if (!found) { set_err("error - physical domain not found!"); goto bad3; }
    }
    free(dom_ind); free(dsig); free(dper); free(dmag);
    dom_ind = NULL;
materials_done:
    /* source edges and signs (src/FETD_FWD.f90:436-448) */
    m->src_edg = xcalloc((size_t)m->nseg, sizeof(int));
    m->src_drc = xcalloc((size_t)m->nseg, sizeof(int));
    for (int i = 0; i < m->nseg; i++) {
        int a = m->seg_vex[2 * i], b = m->seg_vex[2 * i + 1];
        int lo = a < b ? a : b, hi = a < b ? b : a;
        uint64_t k = ((uint64_t)lo << 32) | (uint32_t)hi;
        int found = 0, L = 0, H = m->nE - 1;
        while (L <= H) {
            int mid = (L + H) / 2;
            uint64_t km = ((uint64_t)m->edg_vex[2 * mid] << 32) | (uint32_t)m->edg_vex[2 * mid + 1];
            if (km == k) { found = mid + 1; break; }
            if (km < k) L = mid + 1; else H = mid - 1;
        }
        if (!found) { set_err("source segment is not a mesh edge"); goto bad3; }
        m->src_edg[i] = found;
        m->src_drc[i] = (i == m->nseg - 1) ? -1 : 1;
    }
    /* receivers (src/FETD_FWD.f90:500-520) */
    m->rec_tet = xcalloc((size_t)m->nR, sizeof(int));
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < m->nR; i++) m->rec_tet[i] = p_in_tet(m, &m->rec_xyz[3 * i]);
    for (int i = 0; i < m->nR; i++)
        if (m->rec_tet[i] == 0) {
            snprintf(g_err, sizeof g_err, "no rec tet No. %d", i + 1);
            return -1;
        }
    m->have_ctl = 1;
    return 0;
bad:
    set_err("truncated control file");
    fclose(r.f);
bad3:
    free(dom_ind); free(dsig); free(dper); free(dmag);
    return -1;
}

fo_model *fo_load_files(const char *mesh_path, const char *cntl_path, int verbose) {
    return fo_load_files_variant(mesh_path, cntl_path, verbose, 0);
}

fo_model *fo_load_files_variant(const char *mesh_path, const char *cntl_path, int verbose, int sig_rescale) {
    fo_model *m = xcalloc(1, sizeof(fo_model));
    m->verbose = verbose;
    m->sig_rescale = sig_rescale;
    m->sig_parameter = 1.0;
    if (read_mesh(m, mesh_path)) { fo_free(m); return NULL; }
    build_topology(m);
    if (cntl_path && read_cntl(m, cntl_path)) { fo_free(m); return NULL; }
    if (verbose) {
        printf(" number of nodes          : %d\n", m->nN);
        printf(" number of edges (unknows): %d\n", m->nE);
        printf(" number of tets           : %d\n", m->nT);
    }
    return m;
}

fo_model *fo_load(const char *root, int verbose) {
    char a[1024], b[1024];
    snprintf(a, sizeof a, "%s.mesh", root);
    snprintf(b, sizeof b, "%s.cntl", root);
    return fo_load_files(a, b, verbose);
}

/* ------------------------------------------------------------------ */
/* assembly (src/FETD_FWD.f90:588-732)                                 */
/* ------------------------------------------------------------------ */
typedef struct { int row, col; double a, b, c; } coo_t;

int fo_assemble(fo_model *m) {
    if (!m->have_ctl) { set_err("no control data"); return -1; }
    if (m->bound_cond != 0) { set_err("absorbing boundary condition is not supported"); return -1; }
    double t0 = now_s();
    size_t cap = (size_t)m->nT * 21;
    int *cnt = xcalloc((size_t)m->nE + 2, sizeof(int));
    /* COO in the reference's order: tet i, local j, local k, kept if
     * tet_edg(j) <= tet_edg(k); duplicates are summed in that order. */
    coo_t *coo = xcalloc(cap, sizeof(coo_t));
    size_t nc = 0;
    for (int i = 0; i < m->nT; i++) {
        double Fa[36], Fb[36], Kc[36], len[6];
        for (int j = 0; j < 6; j++) len[j] = m->tet_len[6 * i + j];
        double *a = &m->ca[4 * i], *b = &m->cb[4 * i], *c = &m->cc[4 * i], *d = &m->cd[4 * i];
        double dummy[36];
        if (m->aniso) {
            /* the permittivity mass is multiplied by zero in P (:1288) */
            for (int q = 0; q < 36; q++) Fa[q] = 0.0;
            elem_mats_ani(b, c, d, m->vol[i], len, &m->sig9[9 * i], &m->nu9[9 * i], Fb, Kc);
        } else {
            elem_mats(a, b, c, d, m->vol[i], len, m->per[i], 1.0 / m->mag[i], Fa, Kc);
            elem_mats(a, b, c, d, m->vol[i], len, m->sig[i], 0.0, Fb, dummy);
        }
        const int *te = &m->tet_edg[6 * i];
        const int *td = &m->tet_drc[6 * i];
        for (int j = 0; j < 6; j++)
            for (int k = 0; k < 6; k++)
                if (te[j] <= te[k]) {
                    double sgn = (double)(td[j] * td[k]);
                    coo[nc].row = te[j];
                    coo[nc].col = te[k];
                    coo[nc].a = Fa[6 * j + k] * sgn;
                    coo[nc].b = Fb[6 * j + k] * sgn;
                    coo[nc].c = Kc[6 * j + k] * sgn;
                    cnt[te[j]]++;
                    nc++;
                }
    }
    /* COO -> CSR with sorted columns, duplicates summed (stable by COO order) */
    int *start = xcalloc((size_t)m->nE + 2, sizeof(int));
    for (int r = 1; r <= m->nE; r++) start[r + 1] = start[r] + cnt[r];
    int *fill = xcalloc((size_t)m->nE + 2, sizeof(int));
    size_t *perm = xcalloc(nc, sizeof(size_t));
    for (size_t q = 0; q < nc; q++) {
        int r = coo[q].row;
        perm[start[r] + fill[r]++] = q;
    }
    m->ia = xcalloc((size_t)m->nE + 1, sizeof(int));
    int *colbuf = xcalloc(nc, sizeof(int));
    double *ab = xcalloc(nc, sizeof(double)), *bb = xcalloc(nc, sizeof(double)),
           *cb = xcalloc(nc, sizeof(double));
    size_t nnz = 0;
    for (int r = 1; r <= m->nE; r++) {
        m->ia[r - 1] = (int)nnz + 1;
        size_t s = (size_t)start[r], e = (size_t)start[r + 1];
        /* insertion sort by column, stable (keeps COO order among equal columns) */
        for (size_t u = s + 1; u < e; u++) {
            size_t key = perm[u];
            size_t w = u;
            while (w > s && coo[perm[w - 1]].col > coo[key].col) { perm[w] = perm[w - 1]; w--; }
            perm[w] = key;
        }
        for (size_t u = s; u < e; u++) {
            coo_t *q = &coo[perm[u]];
            if (u > s && coo[perm[u - 1]].col == q->col) {
                ab[nnz - 1] += q->a; bb[nnz - 1] += q->b; cb[nnz - 1] += q->c;
            } else {
                colbuf[nnz] = q->col; ab[nnz] = q->a; bb[nnz] = q->b; cb[nnz] = q->c; nnz++;
            }
        }
    }
    m->ia[m->nE] = (int)nnz + 1;
    m->nnz = (int)nnz;
    m->ja = xcalloc(nnz, sizeof(int));
    m->av = xcalloc(nnz, sizeof(double));
    m->bv = xcalloc(nnz, sizeof(double));
    m->cv = xcalloc(nnz, sizeof(double));
    memcpy(m->ja, colbuf, nnz * sizeof(int));
    memcpy(m->av, ab, nnz * sizeof(double));
    memcpy(m->bv, bb, nnz * sizeof(double));
    memcpy(m->cv, cb, nnz * sizeof(double));
    free(colbuf); free(ab); free(bb); free(cb); free(perm); free(fill); free(start); free(cnt);
    free(coo);
    m->timers[0] = now_s() - t0;
    /* full symmetric copy of B for a row-parallel SpMV */
    {
        int n = m->nE;
        int *deg = xcalloc((size_t)n + 1, sizeof(int));
        for (int r = 0; r < n; r++)
            for (int q = m->ia[r] - 1; q < m->ia[r + 1] - 1; q++) {
                deg[r]++;
                if (m->ja[q] - 1 != r) deg[m->ja[q] - 1]++;
            }
        m->ia0 = xcalloc((size_t)n + 1, sizeof(int));
        for (int r = 0; r < n; r++) m->ia0[r + 1] = m->ia0[r] + deg[r];
        m->ja0 = xcalloc((size_t)m->ia0[n], sizeof(int));
        m->bfull = xcalloc((size_t)m->ia0[n], sizeof(double));
        memset(deg, 0, sizeof(int) * (size_t)n);
        for (int r = 0; r < n; r++)
            for (int q = m->ia[r] - 1; q < m->ia[r + 1] - 1; q++) {
                int c = m->ja[q] - 1;
                int p = m->ia0[r] + deg[r]++;
                m->ja0[p] = c; m->bfull[p] = m->bv[q];
                if (c != r) {
                    p = m->ia0[c] + deg[c]++;
                    m->ja0[p] = r; m->bfull[p] = m->bv[q];
                }
            }
        free(deg);
    }
    /* receiver vectors n_arr, d_arr (src/FETD_FWD.f90:713-731) */
    m->n_arr = xcalloc((size_t)m->nR * 18, sizeof(double));
    m->d_arr = xcalloc((size_t)m->nR * 18, sizeof(double));
    for (int i = 0; i < m->nR; i++) {
        int it = m->rec_tet[i] - 1;
        const double *a = &m->ca[4 * it], *b = &m->cb[4 * it], *c = &m->cc[4 * it],
                     *d = &m->cd[4 * it];
        const double *p = &m->rec_xyz[3 * i];
        double v6 = 6.0 * m->vol[it];
        for (int j = 0; j < 6; j++) {
            int j1 = IND_TET[j][0], j2 = IND_TET[j][1];
            double L1 = a[j1] + b[j1] * p[0] + c[j1] * p[1] + d[j1] * p[2];
            double L2 = a[j2] + b[j2] * p[0] + c[j2] * p[1] + d[j2] * p[2];
            double sc = m->tet_len[6 * it + j] / (v6 * v6);
            double *na = &m->n_arr[18 * i + 3 * j], *da = &m->d_arr[18 * i + 3 * j];
            na[0] = (L1 * b[j2] - L2 * b[j1]) * sc;
            na[1] = (L1 * c[j2] - L2 * c[j1]) * sc;
            na[2] = (L1 * d[j2] - L2 * d[j1]) * sc;
            da[0] = (-2.0) * (c[j1] * d[j2] - c[j2] * d[j1]) * sc;
            da[1] = (-2.0) * (b[j2] * d[j1] - b[j1] * d[j2]) * sc;
            da[2] = (-2.0) * (b[j1] * c[j2] - b[j2] * c[j1]) * sc;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* time loop (src/FETD_FWD.f90:1174-1464)                              */
/* ------------------------------------------------------------------ */

/* Fortran Ew.d edit descriptor: [-]0.ddddE+xx right-justified in w. */
static void fort_e(char *out, double v, int w, int d) {
    char mant[64], buf[96];
    int ex;
    const char *sign = (v < 0 || (v == 0 && signbit(v))) ? "-" : "";
    if (v == 0.0) {
        memset(mant, '0', (size_t)d);
        mant[d] = 0;
        ex = 0;
    } else {
        char tmp[64];
        snprintf(tmp, sizeof tmp, "%.*e", d - 1, fabs(v));
        int k = 0;
        for (const char *p = tmp; *p && *p != 'e'; p++)
            if (isdigit((unsigned char)*p)) mant[k++] = *p;
        mant[k] = 0;
        ex = atoi(strchr(tmp, 'e') + 1) + 1;
    }
    if (v == 0.0) sign = "";
    if (abs(ex) <= 99)
        snprintf(buf, sizeof buf, "%s0.%sE%c%02d", sign, mant, ex < 0 ? '-' : '+', abs(ex));
    else
        snprintf(buf, sizeof buf, "%s0.%s%c%03d", sign, mant, ex < 0 ? '-' : '+', abs(ex));
    int len = (int)strlen(buf);
    if (len > w) {
        memset(out, '*', (size_t)w);
        out[w] = 0;
    } else {
        snprintf(out, (size_t)w + 1, "%*s", w, buf);
    }
}

static void write_row(FILE *f, const fo_model *m, int step, double t, const double *vals9) {
    char s[64];
    fprintf(f, "%6d", step);
    fort_e(s, t, 15, 5);
    fprintf(f, "%s  ", s);
    /* small-loop variant: value / sp**2 * pi**log10(sp) (src_small/FETD_FWD.f90:1526,1540) */
    const double sp2 = pow(m->sig_parameter, 2.0), spk = pow(PI, log10(m->sig_parameter));
    for (int r = 0; r < m->nR; r++) {
        for (int k = 0; k < 9; k++)
            if (m->o_f[k]) {
                const double val = m->sig_rescale ? vals9[9 * r + k] / sp2 * spk : vals9[9 * r + k];
                fort_e(s, val, 15, 5);
                fputs(s, f);
            }
        if (r < m->nR - 1) fputs("  ", f);
    }
    fputc('\n', f);
}

typedef struct {
    fo_model *m;
    FILE *f1, *f2;
    int *gate_ok;
    int out_enabled;
} outctx;

/* vtk_field_save (src/FETD_FWD.f90:1659-1711): E and curl E at every tet's
 * centroid, B += curl E * dt; on every output row while a slice is pending
 * (:1642-1656).  A slice is taken at the first row with t >= its time. */
void fo_tet_fields(const fo_model *m, const double *e, double *ed) {
#pragma omp parallel for schedule(static)
    for (int it = 0; it < m->nT; it++) {
        const int *tv = &m->tet_vex[4 * it];
        double cp[3];
        for (int k = 0; k < 3; k++)
            cp[k] = (m->xyz[3 * (tv[0] - 1) + k] + m->xyz[3 * (tv[1] - 1) + k] + m->xyz[3 * (tv[2] - 1) + k] +
                     m->xyz[3 * (tv[3] - 1) + k]) / 4.0;
        const double *a = &m->ca[4 * it], *b = &m->cb[4 * it], *c = &m->cc[4 * it], *d = &m->cd[4 * it];
        const double v6 = 6.0 * m->vol[it], v62 = v6 * v6;
        double ef[3] = {0, 0, 0}, df[3] = {0, 0, 0};
        for (int j = 0; j < 6; j++) {
            const int j1 = IND_TET[j][0], j2 = IND_TET[j][1];
            const double L1 = a[j1] + b[j1] * cp[0] + c[j1] * cp[1] + d[j1] * cp[2];
            const double L2 = a[j2] + b[j2] * cp[0] + c[j2] * cp[1] + d[j2] * cp[2];
            const double len = m->tet_len[6 * it + j];
            const double nv[3] = {(L1 * b[j2] - L2 * b[j1]) * len / v62, (L1 * c[j2] - L2 * c[j1]) * len / v62,
                                  (L1 * d[j2] - L2 * d[j1]) * len / v62};
            const double dv[3] = {2.0 * (c[j1] * d[j2] - c[j2] * d[j1]) * len / v62,
                                  2.0 * (b[j2] * d[j1] - b[j1] * d[j2]) * len / v62,
                                  2.0 * (b[j1] * c[j2] - b[j2] * c[j1]) * len / v62};
            const double ev = e[m->tet_edg[6 * it + j] - 1];
            const double drc = (double)m->tet_drc[6 * it + j];
            for (int k = 0; k < 3; k++) {
                ef[k] += nv[k] * drc * ev;
                df[k] += dv[k] * drc * ev;
            }
        }
        for (int k = 0; k < 3; k++) {
            ed[6 * (size_t)it + k] = ef[k];
            ed[6 * (size_t)it + 3 + k] = df[k];
        }
    }
}

static void slice_output(fo_model *m, int step, double t, double dt, const double *e) {
    int pending = 0;
    for (int i = 0; i < m->nslice; i++) pending |= !m->slice_done[i];
    if (!pending) return;
    if (step == 0) memset(m->tet_b, 0, sizeof(double) * 3 * (size_t)m->nT);
    fo_tet_fields(m, e, m->tet_ed);
    for (size_t it = 0; it < (size_t)m->nT; it++)
        for (int k = 0; k < 3; k++) m->tet_b[3 * it + k] = m->tet_b[3 * it + k] + m->tet_ed[6 * it + 3 + k] * dt;
    for (int i = 0; i < m->nslice; i++)
        if (!m->slice_done[i] && t >= m->slice_t[i]) {
            double *o = &m->slice_f[(size_t)i * m->nT * 9];
            for (size_t it = 0; it < (size_t)m->nT; it++) {
                memcpy(&o[9 * it], &m->tet_ed[6 * it], 6 * sizeof(double));
                memcpy(&o[9 * it + 6], &m->tet_b[3 * it], 3 * sizeof(double));
            }
            m->slice_done[i] = 1;
            m->slice_step[i] = step;
            m->slice_tt[i] = t;
        }
}

/* The .field2 gate rule of solution_output (src/FETD_FWD.f90:1624-1641):
 * scanning the gates in order, the first one not yet filled whose time has
 * come (t >= gate) is filled by interp_lin (src/FETD_UTILITY.f90:21-31)
 * between the previous output row (tp, fp) and this one (t, f), marked and
 * returned (its values in out[nq]); -1 when none is left for this row.  On
 * the first output row (r == 0) the reference copies the row and leaves the
 * gate loop without writing or marking anything (`exit`), so nothing is
 * returned there. */
static int gate_next(int ngate, const double *gates, int *ok, int r, double tp, const double *fp, double t,
                     const double *f, int nq, double *out) {
    for (int g = 0; g < ngate; g++) {
        if (t >= gates[g] && !ok[g]) {
            if (r == 0) return -1;
            for (int q = 0; q < nq; q++) {
                double x1 = tp, y1 = fp[q], x2 = t, y2 = f[q];
                double a = (y1 - y2) / (x1 - x2);
                double b = y1 - a * x1;
                out[q] = a * gates[g] + b;
            }
            ok[g] = 1;
            return g;
        }
    }
    return -1;
}

/* solution_output (src/FETD_FWD.f90:1513-1657), receiver and file parts */
static void solution_output(outctx *o, int step, double t, double dt, const double *e) {
    fo_model *m = o->m;
    double t0 = now_s();
    if (m->nrows == m->cap_rows) {
        m->cap_rows = m->cap_rows ? 2 * m->cap_rows : 256;
        m->row_step = realloc(m->row_step, sizeof(int) * m->cap_rows);
        m->row_t = realloc(m->row_t, sizeof(double) * m->cap_rows);
        m->row_dt = realloc(m->row_dt, sizeof(double) * m->cap_rows);
        m->row_f = realloc(m->row_f, sizeof(double) * (size_t)m->cap_rows * m->nR * 9);
    }
    int r = m->nrows++;
    double *f = &m->row_f[(size_t)r * m->nR * 9];
    for (int i = 0; i < m->nR; i++) {
        int it = m->rec_tet[i] - 1;
        double ep[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
        for (int j = 0; j < 6; j++) {
            double ev = e[m->tet_edg[6 * it + j] - 1];
            double drc = (double)m->tet_drc[6 * it + j];
            for (int k = 0; k < 3; k++) {
                ep[k] += m->n_arr[18 * i + 3 * j + k] * drc * ev;
                dp[k] += m->d_arr[18 * i + 3 * j + k] * drc * ev;
            }
        }
        if (step == 0) dp[0] = dp[1] = dp[2] = 0.0;
        for (int k = 0; k < 3; k++) {
            f[9 * i + k] = ep[k];
            f[9 * i + 3 + k] = dp[k];
            f[9 * i + 6 + k] = (r == 0) ? dp[k] * dt : m->row_f[(size_t)(r - 1) * m->nR * 9 + 9 * i + 6 + k] + dp[k] * dt;
        }
    }
    m->row_step[r] = step;
    m->row_t[r] = t;
    m->row_dt[r] = dt;
    if (o->f1) write_row(o->f1, m, step, t, f);
    if (m->nslice > 0) slice_output(m, step, t, dt, e);
    {
        double ch[9 * 4096];
        double *chp = (m->nR * 9 <= 9 * 4096) ? ch : malloc(sizeof(double) * m->nR * 9);
        const double *fp = r > 0 ? &m->row_f[(size_t)(r - 1) * m->nR * 9] : NULL;
        const double tp = r > 0 ? m->row_t[r - 1] : 0.0;
        int g;
        while ((g = gate_next(m->n_gate, m->gates, o->gate_ok, r, tp, fp, t, f, m->nR * 9, chp)) >= 0)
            if (o->f2) write_row(o->f2, m, g + 1, m->gates[g], chp);
        if (chp != ch) free(chp);
    }
    m->timers[4] += now_s() - t0;
}

void fo_gate_rows(int nrows, int nq, const double *t, const double *f, int ngate, const double *gates,
                  double *out, int *row_of_gate) {
    int *ok = xcalloc((size_t)ngate + 1, sizeof(int));
    for (int g = 0; g < ngate; g++) row_of_gate[g] = -1;
    for (int r = 0; r < nrows; r++) {
        int g;
        double *buf = xcalloc((size_t)nq, sizeof(double));
        while ((g = gate_next(ngate, gates, ok, r, r > 0 ? t[r - 1] : 0.0, r > 0 ? f + (size_t)(r - 1) * nq : NULL,
                              t[r], f + (size_t)r * nq, nq, buf)) >= 0) {
            memcpy(out + (size_t)g * nq, buf, sizeof(double) * nq);
            row_of_gate[g] = r;
        }
        free(buf);
    }
    free(ok);
}

static void sys_right(const fo_model *m, int step, double t, double *r, int negate) {
    memset(r, 0, sizeof(double) * m->nE);
    if (step >= m->step_dvsi) return;
    double didt = fo_source_didt(m->src_type, m->pulse_type, t, m->pulse_width, m->current);
    for (int i = 0; i < m->nseg; i++) {
        int e = m->src_edg[i] - 1;
        r[e] = didt * m->src_drc[i] * m->edg_len[e];
    }
    if (negate)
        for (int i = 0; i < m->nseg; i++) r[m->src_edg[i] - 1] = -r[m->src_edg[i] - 1];
}

/* r = B*(4 e2 - e3) - 2 dt * s */
static void build_rhs(fo_model *m, const double *e2, const double *e3, double dt, double *r) {
    double t0 = now_s();
    int n = m->nE;
#pragma omp parallel for schedule(static, 4096)
    for (int i = 0; i < n; i++) {
        double acc = 0.0;
        for (int q = m->ia0[i]; q < m->ia0[i + 1]; q++) {
            int c = m->ja0[q];
            acc += m->bfull[q] * (4.0 * e2[c] - e3[c]);
        }
        r[i] = acc - 2.0 * dt * r[i];
    }
    m->timers[3] += now_s() - t0;
}

static void form_p(const fo_model *m, double dt, double *pv) {
    /* mkl_add_3 (src/FETD_FWD.f90:2006-2039): P = 0*A + 3*B + 2*dt*C */
    double cdt = 2.0 * dt;
    for (int q = 0; q < m->nnz; q++) pv[q] = cdt * m->cv[q] + (3.0 * m->bv[q] + 0.0 * m->av[q]);
}

static double two_error(const double *a, const double *b, int n) {
    double s1 = 0, s2 = 0;
    for (int i = 0; i < n; i++) {
        double d = a[i] - b[i];
        s1 += d * d;
        s2 += b[i] * b[i];
    }
    return sqrt(s1) / sqrt(s2);
}

static void log_check(fo_model *m, int step, double err, int keep) {
    if (m->nchk == m->cap_chk) {
        m->cap_chk = m->cap_chk ? 2 * m->cap_chk : 64;
        m->chk_step = realloc(m->chk_step, sizeof(int) * m->cap_chk);
        m->chk_keep = realloc(m->chk_keep, sizeof(int) * m->cap_chk);
        m->chk_err = realloc(m->chk_err, sizeof(double) * m->cap_chk);
    }
    m->chk_step[m->nchk] = step;
    m->chk_err[m->nchk] = err;
    m->chk_keep[m->nchk] = keep;
    m->nchk++;
}

/* Cost sample for the CPU baseline: the pieces of one time step and one
 * factorisation of this model at time step dt, timed separately.  P is
 * formed as in mkl_add_3 and analysed (phase 11) and factored (phase 22)
 * with PARDISO's defaults, mtype 2, falling back to -2 like pardiso_direct_1
 * (src/FETD_UTILITY.f90:879-936); a phase-12 call costs phase 11 + 22, the
 * reference's path at every factorisation, while a caller that keeps the
 * analysis for the unchanged pattern pays phase 22 only.  Then phase 33
 * with one right-hand side (pardiso_direct_2) and with nrhs_block of them
 * on the same factor, and the right-hand side sys_right + SpMV of one
 * source (src/FETD_FWD.f90:1425-1428).  out: [0] phase 11 s, [1] phase 22 s,
 * [2] phase 33 (1 rhs) s, [3] phase 33 (nrhs_block) s, [4] rhs + SpMV s,
 * [5] factor nnz (iparm(18)), [6] mtype used.  Returns a PARDISO error. */
int fo_cpu_costs(fo_model *m, double dt, int nrhs_block, double *out) {
    if (!load_mkl()) return -9999;
    const int n = m->nE;
    double *pv = malloc(sizeof(double) * m->nnz);
    form_p(m, dt, pv);
    pslot s;
    memset(&s, 0, sizeof s);
    int maxfct = 1, mnum = 1, msglvl = 0, error = 0, idum = 0, nrhs = 1;
    double ddum = 0;
    s.n = n; s.ia = m->ia; s.ja = m->ja; s.va = pv;
    double t11 = 0.0, t22 = 0.0;
    for (int attempt = 0; attempt < 2; attempt++) {
        memset(s.pt, 0, sizeof s.pt);
        memset(s.iparm, 0, sizeof s.iparm);
        s.mtype = attempt == 0 ? 2 : -2;
        int ph = 11;
        double t0 = now_s();
        g_pardiso(s.pt, &maxfct, &mnum, &s.mtype, &ph, &s.n, pv, m->ia, m->ja, &idum, &nrhs, s.iparm, &msglvl,
                  &ddum, &ddum, &error);
        t11 = now_s() - t0;
        if (error == 0) {
            ph = 22;
            t0 = now_s();
            g_pardiso(s.pt, &maxfct, &mnum, &s.mtype, &ph, &s.n, pv, m->ia, m->ja, &idum, &nrhs, s.iparm, &msglvl,
                      &ddum, &ddum, &error);
            t22 = now_s() - t0;
        }
        s.live = 1;
        if (error == 0) break;
        pslot_release(&s);
    }
    if (error != 0) { free(pv); return error; }
    s.nnzL = (long long)s.iparm[17];
    const int K = nrhs_block > 1 ? nrhs_block : 1;
    double *b = malloc(sizeof(double) * (size_t)n * K), *x = malloc(sizeof(double) * (size_t)n * K);
    double *e2 = calloc((size_t)n, sizeof(double)), *e3 = calloc((size_t)n, sizeof(double));
    for (size_t i = 0; i < (size_t)n * K; i++) b[i] = sin(0.37 * (double)i);
    for (int i = 0; i < n; i++) { e2[i] = cos(0.11 * i); e3[i] = sin(0.23 * i); }
    double t0 = now_s();
    pslot_solve(&s, b, x, 1);
    const double t33 = now_s() - t0;
    t0 = now_s();
    pslot_solve(&s, b, x, K);
    const double t33k = now_s() - t0;
    t0 = now_s();
    sys_right(m, 1, 0.5 * m->pulse_width, x, 0);
    build_rhs(m, e2, e3, dt, x);
    const double trhs = now_s() - t0;
    out[0] = t11; out[1] = t22; out[2] = t33; out[3] = t33k; out[4] = trhs;
    out[5] = (double)s.nnzL; out[6] = (double)s.mtype;
    pslot_release(&s);
    free(b); free(x); free(e2); free(e3); free(pv);
    return 0;
}

/* adap_euler (src/FETD_FWD.f90:1229-1464).  inout: 0 = plain run with
 * output, 1 = first pass of a step-off run (negated source, no output,
 * result in res_e), 2 = second pass starting from res_e. */
static int adap_euler(fo_model *m, outctx *o, int inout, double *res_e, int max_steps) {
    int n = m->nE;
    int dbl = m->step_dblesize;
    if (m->step_dble < dbl + 2) { set_err("sp_double must be at least sp_dblesize+2"); return -1; }
    int step = 0, not_changed = 0, idx = 0;
    double time_cur = 0.0, step_cur = m->pulse_width / m->step_dvsi;
    double time_limit = m->time_maxm;
    int negate = (inout == 1);
    int outp = (inout != 1);
    size_t nb = sizeof(double) * (size_t)n;
    double *e1 = calloc(n, sizeof(double)), *e2 = calloc(n, sizeof(double)),
           *e3 = calloc(n, sizeof(double)), *rh = calloc(n, sizeof(double));
    double *te1 = calloc(n, sizeof(double)), *te2 = calloc(n, sizeof(double)),
           *te3 = calloc(n, sizeof(double)), *erec = calloc(n, sizeof(double)),
           *ev1 = calloc(n, sizeof(double)), *ev2 = calloc(n, sizeof(double));
    int ring = dbl + 1;
    double *alle = calloc((size_t)n * ring, sizeof(double));
    double *pv[2] = {malloc(sizeof(double) * m->nnz), malloc(sizeof(double) * m->nnz)};
    pslot slot[2];
    memset(slot, 0, sizeof slot);
    int cur = 0;
    int rc = 0;
    double loop0 = now_s();
    while (time_cur < time_limit) {
        if (max_steps > 0 && step >= max_steps) break;
        step++;
        time_cur = time_cur + step_cur;
        if (step == 1) {
            form_p(m, step_cur, pv[cur]);
            if (inout == 2) { memcpy(e2, res_e, nb); memcpy(e3, res_e, nb); }
            if (outp) solution_output(o, 0, time_cur - step_cur, step_cur, e2);
            double t0 = now_s();
            int err = pslot_factor(&slot[cur], n, m->ia, m->ja, pv[cur]);
            m->timers[1] += now_s() - t0;
            m->nfact++;
            m->factor_nnz = slot[cur].nnzL;
            if (err) { snprintf(g_err, sizeof g_err, "PARDISO factorisation error %d", err); rc = -1; break; }
        }
        if (step % m->step_dble == 0) {
            memcpy(te2, e2, nb);
            memcpy(te3, e3, nb);
            double ttime = time_cur, tstep = step_cur;
            for (int i = 1; i <= dbl; i++) {
                sys_right(m, step, ttime, rh, negate);
                build_rhs(m, te2, te3, tstep, rh);
                double t0 = now_s();
                pslot_solve(&slot[cur], rh, te1, 1);
                m->timers[2] += now_s() - t0;
                if (i == dbl) break;
                if (i == 1) memcpy(erec, te1, nb);
                memcpy(te3, te2, nb);
                memcpy(te2, te1, nb);
                ttime = ttime + tstep;
            }
            memcpy(ev1, te1, nb);
            tstep = step_cur * dbl;
            ttime = time_cur - step_cur + tstep;
            int other = 1 - cur;
            form_p(m, tstep, pv[other]);
            double t0 = now_s();
            int err = pslot_factor(&slot[other], n, m->ia, m->ja, pv[other]);
            m->timers[1] += now_s() - t0;
            m->nfact++;
            if (err) { snprintf(g_err, sizeof g_err, "PARDISO factorisation error %d", err); rc = -1; break; }
            sys_right(m, step, ttime, rh, negate);
            const double *ea = &alle[(size_t)((idx - 1) % ring) * n];
            const double *eb = &alle[(size_t)((idx - 1 - dbl) % ring) * n];
            build_rhs(m, ea, eb, tstep, rh);
            t0 = now_s();
            pslot_solve(&slot[other], rh, ev2, 1);
            m->timers[2] += now_s() - t0;
            double error = two_error(ev2, ev1, n);
            int keep = (error > 1.0e-4 && not_changed < 2);
            if (m->nreplay > 0) {
                if (m->nchk >= m->nreplay) { set_err("replayed schedule is too short"); rc = -1; break; }
                keep = m->replay[m->nchk] != 0;
            }
            log_check(m, step, error, keep);
            if (m->verbose)
                printf(" error: %.6e %s\n", error, keep ? "keep step" : "change step");
            if (keep) {
                pslot_release(&slot[other]);
                memcpy(e1, erec, nb);
                if (outp) solution_output(o, step, time_cur, step_cur, e1);
                memcpy(e3, e2, nb);
                memcpy(e2, e1, nb);
                not_changed++;
                idx = 0;
            } else {
                pslot_release(&slot[cur]);
                cur = other;
                memcpy(e3, ea, nb);
                memcpy(e2, ev1, nb);
                time_cur = ttime;
                step_cur = tstep;
                if (outp) solution_output(o, step, time_cur, step_cur, e2);
                not_changed = 0;
                idx = 0;
            }
            continue;
        }
        sys_right(m, step, time_cur, rh, negate);
        build_rhs(m, e2, e3, step_cur, rh);
        double t0 = now_s();
        pslot_solve(&slot[cur], rh, e1, 1);
        m->timers[2] += now_s() - t0;
        if (outp) solution_output(o, step, time_cur, step_cur, e1);
        memcpy(&alle[(size_t)(idx % ring) * n], e1, nb);
        idx++;
        memcpy(e3, e2, nb);
        memcpy(e2, e1, nb);
    }
    m->timers[5] += now_s() - loop0;
    if (inout == 1 && res_e) memcpy(res_e, e1, nb);
    pslot_release(&slot[0]);
    pslot_release(&slot[1]);
    free(e1); free(e2); free(e3); free(rh); free(te1); free(te2); free(te3); free(erec);
    free(ev1); free(ev2); free(alle); free(pv[0]); free(pv[1]);
    return rc;
}

int fo_run(fo_model *m, const char *field_path, const char *field2_path, int max_steps) {
    if (!load_mkl()) { set_err("libmkl_rt.so (PARDISO) is not loadable"); return -1; }
    if (!m->ia) { set_err("call fo_assemble first"); return -1; }
    if (m->step_mthd != 2) { set_err("error!"); return -1; }
    outctx o;
    memset(&o, 0, sizeof o);
    o.m = m;
    o.gate_ok = xcalloc((size_t)m->n_gate + 1, sizeof(int));
    if (field_path) o.f1 = fopen(field_path, "w");
    if (field2_path && m->n_gate > 0) o.f2 = fopen(field2_path, "w");
    m->nrows = 0;
    m->nchk = 0;
    for (int i = 0; i < m->nslice; i++) m->slice_done[i] = 0;
    int rc;
    if (m->pulse_type == 3) {
        double *res = xcalloc((size_t)m->nE, sizeof(double));
        rc = adap_euler(m, &o, 1, res, max_steps);
        if (!rc) rc = adap_euler(m, &o, 2, res, max_steps);
        free(res);
    } else {
        rc = adap_euler(m, &o, 0, NULL, max_steps);
    }
    if (o.f1) fclose(o.f1);
    if (o.f2) fclose(o.f2);
    free(o.gate_ok);
    if (m->verbose) {
        printf(" time of assembling: %12.2f seconds.\n", m->timers[0]);
        printf(" time of fractorization: %12.2f seconds.\n", m->timers[1]);
        printf(" time of stepping: %12.2f seconds.\n", m->timers[2]);
    }
    return rc;
}

/* ------------------------------------------------------------------ */
/* accessors                                                            */
/* ------------------------------------------------------------------ */
void fo_counts(const fo_model *m, int *nN, int *nT, int *nE, int *nnzU, int *nR, int *nseg) {
    if (nN) *nN = m->nN;
    if (nT) *nT = m->nT;
    if (nE) *nE = m->nE;
    if (nnzU) *nnzU = m->nnz;
    if (nR) *nR = m->nR;
    if (nseg) *nseg = m->nseg;
}
void fo_get_topology(const fo_model *m, int *edg_vex, int *tet_edg, int *tet_drc, double *edg_len) {
    if (edg_vex) memcpy(edg_vex, m->edg_vex, sizeof(int) * 2 * (size_t)m->nE);
    if (tet_edg) memcpy(tet_edg, m->tet_edg, sizeof(int) * 6 * (size_t)m->nT);
    if (tet_drc) memcpy(tet_drc, m->tet_drc, sizeof(int) * 6 * (size_t)m->nT);
    if (edg_len) memcpy(edg_len, m->edg_len, sizeof(double) * (size_t)m->nE);
}
void fo_get_geometry(const fo_model *m, double *a, double *b, double *c, double *d, double *vol,
                     double *tet_len) {
    size_t T = (size_t)m->nT;
    if (a) memcpy(a, m->ca, sizeof(double) * 4 * T);
    if (b) memcpy(b, m->cb, sizeof(double) * 4 * T);
    if (c) memcpy(c, m->cc, sizeof(double) * 4 * T);
    if (d) memcpy(d, m->cd, sizeof(double) * 4 * T);
    if (vol) memcpy(vol, m->vol, sizeof(double) * T);
    if (tet_len) memcpy(tet_len, m->tet_len, sizeof(double) * 6 * T);
}
void fo_get_csr(const fo_model *m, int *ia, int *ja, double *ms, double *cc, double *me) {
    if (ia) memcpy(ia, m->ia, sizeof(int) * ((size_t)m->nE + 1));
    if (ja) memcpy(ja, m->ja, sizeof(int) * (size_t)m->nnz);
    if (ms) memcpy(ms, m->bv, sizeof(double) * (size_t)m->nnz);
    if (cc) memcpy(cc, m->cv, sizeof(double) * (size_t)m->nnz);
    if (me) memcpy(me, m->av, sizeof(double) * (size_t)m->nnz);
}
void fo_get_receivers(const fo_model *m, int *rec_tet, double *n_arr, double *d_arr) {
    if (rec_tet) memcpy(rec_tet, m->rec_tet, sizeof(int) * (size_t)m->nR);
    if (n_arr && m->n_arr) memcpy(n_arr, m->n_arr, sizeof(double) * 18 * (size_t)m->nR);
    if (d_arr && m->d_arr) memcpy(d_arr, m->d_arr, sizeof(double) * 18 * (size_t)m->nR);
}
void fo_get_source(const fo_model *m, int *src_edg, int *src_drc) {
    if (src_edg) memcpy(src_edg, m->src_edg, sizeof(int) * (size_t)m->nseg);
    if (src_drc) memcpy(src_drc, m->src_drc, sizeof(int) * (size_t)m->nseg);
}
int fo_get_tensors(const fo_model *m, double *sig9, double *nu9) {
    if (!m->aniso) return 0;
    if (sig9) memcpy(sig9, m->sig9, sizeof(double) * 9 * (size_t)m->nT);
    if (nu9) memcpy(nu9, m->nu9, sizeof(double) * 9 * (size_t)m->nT);
    return 1;
}

void fo_get_material(const fo_model *m, double *sigma, double *inv_mu) {
    for (int i = 0; i < m->nT; i++) {
        if (sigma) sigma[i] = m->sig[i];
        if (inv_mu) inv_mu[i] = 1.0 / m->mag[i];
    }
}
void fo_get_control(const fo_model *m, int *ictl, double *dctl) {
    if (ictl) {
        ictl[0] = m->step_dvsi; ictl[1] = m->step_dble; ictl[2] = m->step_dblesize;
        ictl[3] = m->pulse_type; ictl[4] = m->src_type; ictl[5] = m->n_gate;
        for (int k = 0; k < 9; k++) ictl[6 + k] = m->o_f[k];
    }
    if (dctl) { dctl[0] = m->pulse_width; dctl[1] = m->current; dctl[2] = m->time_maxm; }
}
void fo_set_control(fo_model *m, const int *ictl, const double *dctl) {
    if (ictl) {
        m->step_dvsi = ictl[0]; m->step_dble = ictl[1]; m->step_dblesize = ictl[2];
        m->pulse_type = ictl[3]; m->src_type = ictl[4];
    }
    if (dctl) { m->pulse_width = dctl[0]; m->current = dctl[1]; m->time_maxm = dctl[2]; }
}
void fo_set_replay(fo_model *m, int n, const int *keep) {
    free(m->replay);
    m->replay = NULL;
    m->nreplay = 0;
    if (n > 0) {
        m->replay = xcalloc((size_t)n, sizeof(int));
        memcpy(m->replay, keep, sizeof(int) * (size_t)n);
        m->nreplay = n;
    }
}
void fo_set_sigma(fo_model *m, const double *sigma) {
    memcpy(m->sig, sigma, sizeof(double) * (size_t)m->nT);
}
void fo_set_slices(fo_model *m, int n, const double *times) {
    free(m->slice_t); free(m->slice_done); free(m->slice_step); free(m->slice_tt); free(m->slice_f);
    free(m->tet_ed); free(m->tet_b);
    m->slice_t = NULL; m->slice_done = NULL; m->slice_step = NULL; m->slice_tt = NULL; m->slice_f = NULL;
    m->tet_ed = NULL; m->tet_b = NULL;
    m->nslice = n > 0 ? n : 0;
    if (!m->nslice) return;
    m->slice_t = xcalloc((size_t)n, sizeof(double));
    memcpy(m->slice_t, times, sizeof(double) * (size_t)n);
    m->slice_done = xcalloc((size_t)n, sizeof(int));
    m->slice_step = xcalloc((size_t)n, sizeof(int));
    m->slice_tt = xcalloc((size_t)n, sizeof(double));
    m->slice_f = xcalloc((size_t)n * m->nT * 9, sizeof(double));
    m->tet_ed = xcalloc((size_t)m->nT * 6, sizeof(double));
    m->tet_b = xcalloc((size_t)m->nT * 3, sizeof(double));
}
int fo_get_slice(const fo_model *m, int i, double *f9, int *step, double *t) {
    if (i < 0 || i >= m->nslice || !m->slice_done[i]) return -1;
    if (f9) memcpy(f9, &m->slice_f[(size_t)i * m->nT * 9], sizeof(double) * 9 * (size_t)m->nT);
    if (step) *step = m->slice_step[i];
    if (t) *t = m->slice_tt[i];
    return 0;
}
double fo_sig_parameter(const fo_model *m) { return m->sig_parameter; }
int fo_num_rows(const fo_model *m) { return m->nrows; }
void fo_get_rows(const fo_model *m, int *step, double *t, double *dt, double *fields) {
    if (step) memcpy(step, m->row_step, sizeof(int) * (size_t)m->nrows);
    if (t) memcpy(t, m->row_t, sizeof(double) * (size_t)m->nrows);
    if (dt) memcpy(dt, m->row_dt, sizeof(double) * (size_t)m->nrows);
    if (fields) memcpy(fields, m->row_f, sizeof(double) * (size_t)m->nrows * m->nR * 9);
}
int fo_num_checks(const fo_model *m) { return m->nchk; }
void fo_get_checks(const fo_model *m, int *step, double *err, int *keep) {
    if (step) memcpy(step, m->chk_step, sizeof(int) * (size_t)m->nchk);
    if (err) memcpy(err, m->chk_err, sizeof(double) * (size_t)m->nchk);
    if (keep) memcpy(keep, m->chk_keep, sizeof(int) * (size_t)m->nchk);
}
void fo_get_timers(const fo_model *m, double *t6) { memcpy(t6, m->timers, sizeof m->timers); }
int fo_num_factorizations(const fo_model *m) { return m->nfact; }
long long fo_factor_nnz(const fo_model *m) { return m->factor_nnz; }
