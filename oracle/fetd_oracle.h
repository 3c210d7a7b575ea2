/*
 * fetd_oracle.h -- CPU restatement of TEMF3DT's FETD forward path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in temf3dt_amd/: only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.  The product never links or
 * calls it.
 *
 * Every stage follows the reference Fortran (src/FETD_FWD.f90,
 * src/FETD_UTILITY.f90); each function in fetd_oracle.c cites the lines it
 * restates.  The sparse factor/solve is Intel MKL PARDISO (the library the
 * reference calls through cluster_sparse_solver), loaded at run time from
 * libmkl_rt.so with the reference's settings (iparm all zero = defaults,
 * mtype 2 with a -2 fallback, src/FETD_UTILITY.f90:879-936).
 *
 * Index conventions: vertex, edge and tet ids are 1-based as in the
 * reference files; arrays are row-major [item][component].
 */
#ifndef FETD_ORACLE_H
#define FETD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct fo_model fo_model;

/* Load ROOT.mesh and ROOT.cntl (src/FETD_FWD.f90:112-586).  Returns NULL on
 * error; fo_error() gives the message. */
fo_model *fo_load(const char *root, int verbose);
/* Same, from explicit file names (cntl may be NULL to skip control parsing). */
fo_model *fo_load_files(const char *mesh_path, const char *cntl_path, int verbose);
/* sig_rescale = 1: the small-loop variant (src_small/FETD_FWD.f90:314-328,
 * 1526,1540): sigma of domains 2..n scaled by sp = 1 / sigma of the last
 * domain, written values scaled by pi**log10(sp) / sp**2. */
fo_model *fo_load_files_variant(const char *mesh_path, const char *cntl_path, int verbose, int sig_rescale);
double fo_sig_parameter(const fo_model *m);
const char *fo_error(void);
void fo_free(fo_model *m);

/* Element matrices / global CSR (src/FETD_FWD.f90:588-732). */
int fo_assemble(fo_model *m);

/* Time loop (src/FETD_FWD.f90:1174-1464).  field paths may be NULL.
 * max_steps > 0 stops after that many loop iterations (bounded sample). */
int fo_run(fo_model *m, const char *field_path, const char *field2_path, int max_steps);

/* Per-tet field slices (vtkout, src/FETD_FWD.f90:1642-1711): before fo_run,
 * n slice times; after it, slice i as [T][9] (E, curl E at the centroid, the
 * running sum of curl E * dt), with the step and time of the output row
 * that took it.  fo_get_slice returns -1 for a slice never reached. */
void fo_set_slices(fo_model *m, int n, const double *times);
int fo_get_slice(const fo_model *m, int i, double *f9, int *step, double *t);
/* vtk_field_save's per-tet evaluation for an edge vector e[nE]: ed[T][6] =
 * E and curl E at the tet centroids (src/FETD_FWD.f90:1672-1705). */
void fo_tet_fields(const fo_model *m, const double *e, double *ed);

/* ---- accessors for checks ---- */
void fo_counts(const fo_model *m, int *nN, int *nT, int *nE, int *nnzU, int *nR, int *nseg);
/* copies out (any pointer may be NULL) */
void fo_get_topology(const fo_model *m, int *edg_vex, int *tet_edg, int *tet_drc, double *edg_len);
void fo_get_geometry(const fo_model *m, double *a, double *b, double *c, double *d, double *vol, double *tet_len);
void fo_get_csr(const fo_model *m, int *ia, int *ja, double *mass_sigma, double *curl, double *mass_eps);
void fo_get_receivers(const fo_model *m, int *rec_tet, double *n_arr, double *d_arr);
void fo_get_source(const fo_model *m, int *src_edg, int *src_drc);
void fo_get_material(const fo_model *m, double *sigma, double *inv_mu);
/* control values: sp_division, sp_double, sp_dblesize, pulse_type, source_type */
void fo_get_control(const fo_model *m, int *ictl, double *dctl);
void fo_set_control(fo_model *m, const int *ictl, const double *dctl);
/* force the keep (1) / change (0) decisions of the doubling checks */
void fo_set_replay(fo_model *m, int n, const int *keep);
/* override per-tet conductivity (for perturbation tests) */
void fo_set_sigma(fo_model *m, const double *sigma);

/* per output row: step, t, dt, and R*9 fields (ex ey ez dx dy dz bx by bz) */
int fo_num_rows(const fo_model *m);
void fo_get_rows(const fo_model *m, int *step, double *t, double *dt, double *fields);
/* doubling-check log: step, err, decision (1 keep / 0 change) */
int fo_num_checks(const fo_model *m);
void fo_get_checks(const fo_model *m, int *step, double *err, int *keep);
/* timers: assemble, factorisation, stepping (solve), spmv+rhs, readout, loop wall */
void fo_get_timers(const fo_model *m, double *t6);
int fo_num_factorizations(const fo_model *m);
long long fo_factor_nnz(const fo_model *m);

/* The .field2 gate rule (src/FETD_FWD.f90:1624-1641) applied to given output
 * rows t[nrows], f[nrows][nq]: out[ngate][nq] receives each gate's
 * interpolated values and row_of_gate[ngate] the row that filled it (-1:
 * never filled). */
void fo_gate_rows(int nrows, int nq, const double *t, const double *f, int ngate, const double *gates,
                  double *out, int *row_of_gate);

/* ---- pointwise kernels (for known-answer tests) ---- */
/* cofactors of one tet: x[4],y[4],z[4] -> a,b,c,d[4], vol (signed) */
void fo_tet_cofactors(const double *x, const double *y, const double *z,
                      double *a, double *b, double *c, double *d, double *vol);
/* 6x6 mass (scalar s) and curl-curl (scalar s) matrices of one tet, row-major */
void fo_tet_matrices(const double *x, const double *y, const double *z,
                     double s_mass, double s_curl, double *F, double *K);
/* anisotropic variants (f_kind_ani / e_kind_ani): 3x3 row-major tensors
 * between the gradients (mass) and the curls (curl-curl) */
void fo_tet_matrices_ani(const double *x, const double *y, const double *z, const double *A_mass,
                         const double *A_curl, double *F, double *K);
/* principal values v[3] rotated by three angles (degrees) as the reference
 * builds its tensors, times scale; A row-major 3x3 */
void fo_aniso_tensor(const double *v, const double *ang_deg, double scale, double *A);
/* per-tet tensors of an anisotropic model ([nT][9], row-major; NULL-safe);
 * returns 0 for an isotropic model */
int fo_get_tensors(const fo_model *m, double *sig9, double *nu9);
/* 4x4 determinant by LU with partial pivoting, column-major input */
double fo_det_lu(const double *a_colmajor, int n);
/* source time function d I/dt (src/FETD_UTILITY.f90:1045-1078) */
double fo_source_didt(int source_type, int pulse_type, double t, double t0, double cur);

/* ---- solver backend (MKL PARDISO via dlopen) ---- */
/* Timed pieces of one step and one factorisation at time step dt (CPU
 * baseline): out[0] phase 11, [1] phase 22, [2] phase 33 with 1 rhs,
 * [3] phase 33 with nrhs_block, [4] sys_right + SpMV (seconds), [5] factor
 * nnz, [6] mtype used.  Returns a PARDISO error code (0 on success). */
int fo_cpu_costs(fo_model *m, double dt, int nrhs_block, double *out);
int fo_pardiso_available(void);
void fo_set_threads(int n);
/* Solve the upper-CSR (1-based) symmetric system once; returns PARDISO error. */
int fo_pardiso_solve_once(int n, const int *ia, const int *ja, const double *va,
                          const double *b, double *x, int nrhs);

#ifdef __cplusplus
}
#endif
#endif
