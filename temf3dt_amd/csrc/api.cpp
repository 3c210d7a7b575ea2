// api.cpp -- libtemfetd C ABI: context, device memory, execution plan and the
// BDF2 time loop with the reference's doubling controller.
//
// Reference call structure replaced here (yxddyxzh2016/TEMF3DT):
//   mesh_process   src/FETD_FWD.f90:588-732    tf_assemble
//   adap_euler     src/FETD_FWD.f90:1229-1464  tf_run (run_loop below)
//   sys_right      src/FETD_FWD.f90:1466-1511  source term in rhs()
//   solution_output src/FETD_FWD.f90:1513-1623 readout rows
//   pardiso_direct_1/2/3 src/FETD_UTILITY.f90:879-975  tf_factor/tf_solve/tf_release
#include <algorithm>
#include <array>
#include <map>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/temfetd.h"
#include "tf_device.h"
#include "tf_internal.h"

namespace tf {
// assembly.hip
void launch_element(int nT, const double *xyz, const int *tet_vex, const int *tet_drc, const uint8_t *orient,
                    const double *sigma, const double *inv_mu, double *vm, double *vk, hipStream_t st);
void launch_element_ani(int nT, const double *xyz, const int *tet_vex, const int *tet_drc, const uint8_t *orient,
                        const double *As, const double *Ak, double *vm, double *vk, hipStream_t st);
void launch_gather(int64_t nnz, const int64_t *cptr, const int *ctr, const double *vm, const double *vk,
                   double *M, double *K, hipStream_t st);
void launch_recvec(int nR, const double *xyz, const int *tet_vex, const int *rec_tet, const double *rec_xyz,
                   double *n_arr, double *d_arr, hipStream_t st);
void launch_slice(int nT, const double *xyz, const int *tet_vex, const int *tet_edg, const int *tet_drc,
                  const double *e, double dt, int zero_b, double *B, double *out, hipStream_t st);
// factor.hip
void launch_fill0(double *p, int64_t n, hipStream_t st);
void launch_pinv_w21(double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                     int cb, hipStream_t st);
void launch_flags_init(FactorFlags *f, hipStream_t st);
void launch_pivot_scan(const double *Dn, const FrontDesc *fd, int nsn, FactorFlags *f, hipStream_t st);
void launch_transpose_l21(const double *L, double *LT, const FrontDesc *fd, const int *units, int64_t nunits,
                          hipStream_t st);
void launch_scatter_p(int64_t nnz, const int64_t *amap, const double *M, const double *K, double cdt, double *L,
                      hipStream_t st);
void launch_scatter_vals(int64_t nnz, const int64_t *amap, const double *va, double *L, hipStream_t st);
void launch_extend_add(double *L, double *U, const FrontDesc *fd, const int *parent, const int *list,
                       const int64_t *pref, int cnt, int64_t nwg, const int *rel, hipStream_t st);
void launch_factor_diag(double *L, const FrontDesc *fd, const int *list, int cnt, int p0, double *Dn, int *flag,
                        double *W, const int64_t *woff, hipStream_t st);
void launch_factor_trsm(double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                        int p0, const double *Dn, const double *W, const int64_t *woff, hipStream_t st);
void launch_factor_update(double *L, double *U, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                          int64_t nwg, int p0, int mode, const double *Dn, int band, hipStream_t st);
// factor.hip (inversion)
void launch_inv_diag(double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                     hipStream_t st);
void launch_inv_merge(int step, double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                      int64_t nwg, int b, double *W, const int64_t *woff, hipStream_t st);
// solve.hip
void launch_fwd_small(const double *L, const FrontDesc *fd, const int *list, int cnt, const double *b,
                      const int *perm, const int *chp, const int *ch, const int *rel, double *V, double *Y,
                      hipStream_t st);
void launch_bwd_small(const double *L, const FrontDesc *fd, const int *list, int cnt, const double *Y,
                      const double *Dn, const int *sidx, double *xw, hipStream_t st);
void launch_fwd_own(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                    int64_t nwg, const double *V, double *Y, hipStream_t st);
void launch_fwd_struct(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                       int64_t nwg, double *V, const double *Y, hipStream_t st);
void launch_bwd_t(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                  const double *Y, const double *Dn, const double *V, double *T, hipStream_t st);
void launch_bwd_x(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                  const double *T, double *xw, hipStream_t st);
void launch_unpermute(int n, const int *perm, const double *xw, double *x, hipStream_t st);
hipError_t noalloc_scratch(int device, double **ptr);
// msolve.hip
void launch_fill_pattern(int64_t len, double *b, hipStream_t st);
// step.hip: independent models
void launch_source_models(int ns, const int *edge, const int *model, const double *w, const double *coef, double *r,
                          hipStream_t st);
void launch_scatter_p_models(int64_t nnz, const int64_t *amap, const int *qmodel, const double *M, const double *K,
                             const double *cdt, const int64_t *slot_off, double *L, hipStream_t st);
void launch_range_norms(int nm, const int *range, const double *a, const double *b, double *out, hipStream_t st);
// adjoint.hip
void launch_adj_seed(int nR, int K, const int *rec_edge, const double *rec_drc, const double *n_arr,
                     const double *d_arr, const double *w, double *g, hipStream_t st);
void launch_sens(int nT, int K, const int *tet_edg, const uint8_t *orient, const double *vm1, const double *lam,
                 const double *y, const double *u, const double *v, double *grad, hipStream_t st);
void launch_axpy(int64_t len, double a, const double *x, double *y, hipStream_t st);
void launch_scale(int n, const double *s, double *g, hipStream_t st);
// mtile.hip
void launch_mtile(int mode, int ks, int nb, int na, const double *L, const FrontDesc *fd, const TileUnit *units, int nunits,
                  int K, const double *in0, const double *in1, const double *Dn, const int *sidx, double *out,
                  double *part, uint32_t gbytes, const int *perm, double *x1, double *x2, const double *LT,
                  hipStream_t st, double *z = nullptr);
void launch_spull(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg, const double *b,
                  const int *perm, const int *gptr, const int *gidx, const double *Z, double *Yin, hipStream_t st);
void launch_mpull(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg, const double *b,
                  const int *perm, const int *gptr, const int *gidx, const double *Z, double *Yin, int K,
                  hipStream_t st);
void launch_stile(int mode, int nc, const double *L, const FrontDesc *fd, const TileUnit *units, int nunits,
                  const double *in0, const double *in1, const double *Dn, const int *sidx, double *out, double *part,
                  const int *perm, double *x1, double *x2, hipStream_t st, double *z = nullptr);
void launch_stile_lt(int mode, const double *LT, const FrontDesc *fd, const TileUnit *units, int nunits,
                     const double *in0, const double *in1, const double *Dn, const int *sidx, double *out,
                     double *part, const int *perm, double *x1, double *x2, hipStream_t st);
void launch_stile_red(int mode, const FrontDesc *fd, const TileUnit *units, int nunits, const double *in0,
                      const double *in1, const double *Dn, double *out, const double *part, const int *perm,
                      double *x1, double *x2, hipStream_t st, double *z = nullptr);
void launch_mtile_red(int mode, int nb, int na, const FrontDesc *fd, const TileUnit *units, int nunits, int K,
                      const double *in0, const double *in1, const double *Dn, double *out, const double *part,
                      const int *perm, double *x1, double *x2, hipStream_t st, double *z = nullptr);
void launch_mspmv(int n, int K, const int *ia, const int *ja, const double *val, const double *e, double *u, int xcd,
                  hipStream_t st);
void launch_mspmv_chain(int n, int K, const int *ia, const int *ja, const double *val, const double *e2,
                        const double *u3, double *u2, double *r, int xcd, hipStream_t st);
void launch_mspmv_rhs(int n, int K, const int *ia, const int *ja, const double *val, const double *e2,
                      const double *e3, double *w, double *r, int xcd, hipStream_t st);
void launch_msource(int ns, int K, const int *edge, const int *src, const double *w, double didt, double cdt,
                    double *r, hipStream_t st);
void launch_mreadout(int nR, int K, const int *rec_edge, const double *rec_drc, const double *n_arr,
                     const double *d_arr, const double *e, double *out, hipStream_t st);
void launch_mdiffnorm(int64_t n, int K, const double *a, const double *b, double *part, double *out,
                      hipStream_t st);
void launch_mg_init(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg, const double *b,
                    const int *perm, double *V, int K, hipStream_t st);
void launch_mg_child(const FrontDesc *fd, const int *parent, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                     const int *rel, double *V, int K, hipStream_t st);
void launch_mg_xs(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg, const int *sidx,
                  const double *xw, double *V, int K, hipStream_t st);
// step.hip
void launch_spmv_rhs(int n, const int *ia, const int *ja, const double *val, const double *e2, const double *e3,
                     double *r, hipStream_t st);
void launch_source(int ns, const int *edge, const double *w, double didt, double cdt, double *r, hipStream_t st);
void launch_readout(int nR, const int *rec_edge, const double *rec_drc, const double *n_arr, const double *d_arr,
                    const double *e, double *out, hipStream_t st);
void launch_diffnorm(int64_t n, const double *a, const double *b, double *part, double *out, hipStream_t st);
void launch_gather_vals(int64_t n, const int64_t *map, const double *src, double *dst, hipStream_t st);
}  // namespace tf

namespace {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// a factorisation whose pivots the caller did not accept (TF_ERR_NUMERIC)
struct NumericError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                                          \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess)                                                                              \
            throw HipError(std::string("HIP error ") + hipGetErrorString(e_) + " at " + #x);               \
    } while (0)

template <class T>
T *dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    return (T *)p;
}
template <class T>
T *dupload(const std::vector<T> &v) {
    T *p = dalloc<T>(v.size());
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}
template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

enum LaunchKind {
    L_EXTADD, L_FDIAG, L_FTRSM, L_FUPD, L_IMERGE_T, L_IMERGE_X, L_IDIAG,
    L_SSMALL, L_SOWN, L_SSTRUCT,
    L_BSMALL, L_BT, L_BX,
    L_GINIT, L_GCHILD, L_GXS, L_MT, L_MTRED, L_MPULL,
    L_ST, L_STRED, L_SPULL,  // single right-hand side through the tile kernels
    L_PINV                   // factorisation: partitioned inverse of a level's fronts (W = -L21 L11^-1)
};
struct Launch {
    int kind, p0, off, cnt, poff;
    int64_t nwg;
    int aux;  // factor updates: mode; backward X tiles: 1 = no front-order copy of x (lowest level)
    int nb = 4;  // tile launches: 16-row output blocks per wave (2 or 4)
};

int prof_category(int kind) {
    switch (kind) {
    case L_MT: return TF_PROF_TILE;
    case L_MTRED: return TF_PROF_TILE_RED;
    case L_MPULL: return TF_PROF_PULL;
    case L_FUPD: return TF_PROF_FACTOR_UPDATE;
    case L_EXTADD: case L_FDIAG: case L_FTRSM: case L_IMERGE_T: case L_IMERGE_X: case L_IDIAG:
    case L_PINV:
        return TF_PROF_FACTOR_OTHER;
    default: return TF_PROF_SOLVE1;
    }
}

// Execution plan over the front tree: which fronts each launch touches.
struct Plan {
    std::vector<Launch> factor, fwd, bwd, mfwd, mbwd, sfwd, sbwd;
    std::vector<int> list;
    std::vector<int64_t> pref;
    std::vector<int64_t> woff;  // per supernode: offset in the inversion scratch
    int64_t wsize = 0;
    int64_t part_max = 0;       // partial-sum tiles (4096 doubles each) of one launch
    int64_t spart_max = 0;      // the same for the single right-hand-side plan (1024 doubles each)
    std::vector<TileUnit> units;  // block-solve tile units (L_MT / L_MTRED launches)
};

void free_native(tf_ctx *c);

// partial-sum storage for both plans' split tiles
size_t part_doubles(const Plan &p) {
    return (size_t)std::max<int64_t>({p.part_max * 4096, p.spart_max * 1024, 4096});
}

static const double PI = 3.141592653589793;

// source_mod (src/FETD_UTILITY.f90:1030-1166): d I / dt of the waveform
double source_didt(int source_type, int pulse_type, double t, double t0, double cur) {
    if (source_type == 1 && pulse_type == 1) {
        double u = t - t0 / 2.0;
        return (1.0 / std::sqrt(PI) * std::exp(1.0 / (t0 * t0) * (u * u) * (-25.0)) * 5.0) / t0 * cur;
    }
    if (source_type == 1 && pulse_type == 2) {
        double w = t * 5.436563656918091 - t0 * 2.718281828459046;
        return -1.0 / (t0 * t0) * std::exp(-1.0 / (t0 * t0) * (w * w)) *
               (t * 59.11244879144521 - t0 * 29.55622439572261) * cur;
    }
    if (source_type == 1 && pulse_type == 3) {
        double u = t - t0 / 2.0;
        return -((1.0 / std::sqrt(PI) * std::exp(1.0 / (t0 * t0) * (u * u) * (-25.0)) * 5.0) / t0) * cur;
    }
    return NAN;
}

}  // namespace

struct tf_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    hipStream_t st2 = nullptr;  // the doubling check's factorisation, overlapped with the steps before it
    std::string err;
    Topology topo;
    bool have_mesh = false, have_mat = false, assembled = false, analyzed = false;
    std::vector<double> sigma, inv_mu;
    bool aniso = false;  // tensors set by tf_set_materials_tensor
    double *d_sig9 = nullptr, *d_nu9 = nullptr;
    // sources: nsrc loops sharing one waveform; entries (edge, source, weight)
    int nsrc = 1;
    int nseg = 0;
    std::vector<std::vector<int>> src_edges;
    std::vector<std::vector<double>> src_ws;
    std::vector<int> src_edge, src_id;
    std::vector<double> src_w;
    std::vector<int> replay;  // forced keep/change decisions of the doubling checks
    // per-tet field slices (vtkout): requested times, which were taken, when
    std::vector<double> slice_t, slice_tt;
    std::vector<int> slice_done, slice_step;
    double *d_slice = nullptr, *d_sliceB = nullptr;  // [slice][T][9], [T][3]
    int slice_nT = 0;
    int stype = 1, ptype = 1;
    double pulse_width = 0.0, current = 1.0;
    bool have_src = false;
    // receivers
    int nR = 0;
    std::vector<double> rec_xyz;
    std::vector<int> rec_tet;
    double rec_charac = 10.0;
    // schedule
    int sp_division = 0, sp_double = 0, sp_dblesize = 2;
    double time_max = 0.0;
    // device: mesh and assembly
    double *d_xyz = nullptr, *d_sigma = nullptr, *d_invmu = nullptr, *d_vm = nullptr, *d_vk = nullptr;
    int *d_tet_vex = nullptr, *d_tet_drc = nullptr, *d_ctr = nullptr;
    uint8_t *d_orient = nullptr;
    int64_t *d_cptr = nullptr;
    double *d_M = nullptr, *d_K = nullptr;
    // full symmetric Msigma for the right-hand side
    int *d_fia = nullptr, *d_fja = nullptr;
    int64_t *d_fmap = nullptr;
    double *d_fval = nullptr;
    // receivers / source on device
    double *d_rec_xyz = nullptr, *d_narr = nullptr, *d_darr = nullptr, *d_rec_drc = nullptr;
    int *d_rec_tet = nullptr, *d_rec_edge = nullptr;
    int *d_src_edge = nullptr, *d_src_id = nullptr;
    double *d_src_w = nullptr;
    // block-solve work (K right-hand sides)
    double *d_mV = nullptr, *d_mY = nullptr, *d_mT = nullptr, *d_mb = nullptr, *d_mx = nullptr;
    double *d_part = nullptr;
    int mK = 0;
    int64_t vrows = 0;
    // solver
    Symbolic sym;
    Plan plan;
    FrontDesc *d_fd = nullptr;
    int *d_perm = nullptr, *d_parent = nullptr, *d_chp = nullptr, *d_ch = nullptr, *d_rel = nullptr,
        *d_sidx = nullptr, *d_plist = nullptr;
    int64_t *d_amap = nullptr, *d_ppref = nullptr;
    TileUnit *d_units = nullptr;
    int *d_gptr = nullptr, *d_gidx = nullptr;
    double *d_mZ = nullptr;
    // the defaults are the autotuner's choice on the 1.6M-unknown batched mesh
    int mt_units = 8192, mt_chmin = 512;  // tile-launch sizing (units per launch, min inner chunk)
    int mt_wide = 256;  // levels of at most this many fronts use 64-row tiles, the others 32-row tiles
    int mt_ks_n = 2, mt_ks_t = 2;         // k-steps per stage of the N- and T-form tile kernels
    // fronts on the lowest pinv_levels levels keep the partitioned inverse
    // (FrontDesc::pinv): one tile product per front and direction
    int pinv_levels = 10;
    int *d_frows = nullptr;  // per partitioned-inverse front (at voff): its f rows in the [v; x] block vector
    // single right-hand-side solve: 2 through the vector tile kernels, 1
    // through the matrix-core tile kernels with one right-hand-side block, 0
    // through the matrix-vector kernels of solve.hip; launch sizing of its tiles
    int s_tiles = 2;
    int s_units = 2048, s_chmin = 512;
    int spmv_xcd = 1;  // block SpMV: contiguous row ranges per XCD
    // block time loop: regular steps form M (4 e2 - e3) as 4 M e2 - M e3 from
    // the product M e2 kept from the step before (tuning key "rhs_chain")
    bool rhs_chain = true;
    // The time loops (tf_run, tf_run_models, tf_adjoint) keep their vectors
    // in the factor's column order ("native" order: row i is edge perm[i]),
    // so the solves read b and write x in place instead of through perm, and
    // the SpMV's neighbouring rows gather neighbouring rows (the nested
    // dissection keeps subdomains contiguous).  The edge-indexed maps of the
    // step (full CSR of Msigma, source and receiver edges, tet edges, model
    // ranges) are translated once per analysis (tuning key "native").
    bool native = true, native_ok = false;
    int *d_ident = nullptr;
    int *d_nfia = nullptr, *d_nfja = nullptr;
    int64_t *d_nfmap = nullptr;
    double *d_nfval = nullptr;
    int64_t nfnnz = 0;
    int *d_nsrc_edge = nullptr, *d_nrec_edge = nullptr, *d_ntet_edg = nullptr, *d_nmrange = nullptr;
    std::vector<int> nm_e0;
    // maps of the loop being run, and the permutation the solves map b and x
    // through (nullptr: perm, edge-order vectors)
    struct RunMaps {
        const int *fia, *fja;
        const double *fval;
        const int *src_edge, *rec_edge, *tet_edg, *mrange, *perm;
        const int *e0;  // host: first row of each model
    } rm{};
    const int *solve_perm = nullptr;
    int s_tcols = 32;  // output columns per pass of the vector T-form tiles (16 or 32)
    // single-source contexts keep the row-major copy of L21 as well, and the
    // single solve's backward products read it as N-form tiles (tuning key s_lt)
    bool s_lt = true;
    // trailing-update tiles walked in bands of this many tile rows (0: row by
    // row; tuning key "f_band")
    int f_band = 16;
    double *d_U = nullptr, *d_V = nullptr, *d_xw = nullptr, *d_Y = nullptr, *d_T = nullptr, *d_W = nullptr;
    int64_t *d_woff = nullptr;
    double *d_L[2] = {nullptr, nullptr}, *d_D[2] = {nullptr, nullptr};
    // row-major copies of every front's L21 per slot (block solves of K > 1
    // sources read them in the backward product); lt_total doubles each
    double *d_LT[2] = {nullptr, nullptr};
    int64_t lt_total = 0;
    int *d_tunits = nullptr;  // (front, first struct row, first own column) per 64 x 64 copy tile
    int64_t n_tunits = 0;
    bool slot_live[2] = {false, false};
    FactorFlags *d_flag = nullptr;
    // pivots of the last factorisation: negative count, smallest and largest
    // magnitude; allow_indefinite accepts negative pivots (tuning key)
    long long last_nneg = 0;
    double last_min_abs = 0.0, last_max_abs = 0.0;
    bool allow_indefinite = false;
    double *d_tmpb = nullptr, *d_tmpx = nullptr;
    double *d_scratch = nullptr;  // this device's scratch array of tf_solve_noalloc (not owned)
    // stepping work vectors
    std::vector<double *> vecs;
    int64_t vec_len = 0;
    double *d_norm_part = nullptr, *d_norm_out = nullptr;
    // output rows
    std::vector<int> row_step;
    std::vector<double> row_t, row_dt;
    double *d_rows = nullptr;
    int64_t rows_cap = 0;
    int nrows = 0;
    std::vector<int> chk_step, chk_keep;
    std::vector<double> chk_err;
    double timers[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // adjoint: tape of the last tf_run (solve nodes, output rows, history
    // checkpoints every ckpt_every steps)
    struct TNode {
        int u, v, step, row;  // history nodes stepped from; row written from this node or -1
        double dt, t;
    };
    int ckpt_every = 0;
    bool tape_valid = false;
    std::vector<TNode> tape;
    struct Ckpt {
        int first, id2, id3;
        double *s2, *s3;
    };
    std::vector<Ckpt> ckpts;
    std::vector<int> row_node;
    int *d_tet_edg = nullptr;
    double *d_vm1 = nullptr;
    double slot_dt[2] = {-1.0, -1.0};  // dt a slot was factored for (-1: none / caller matrix)
    // factors of a taped run kept for its adjoint sweep (tuning key
    // "keep_factors": at most this many besides the two slots; 0: none)
    struct KeptFactor {
        double dt;
        double *L, *D, *LT;
    };
    std::vector<KeptFactor> kept;
    int keep_factors = 8;
    bool slot_taped[2] = {false, false};  // the slot was factored during the current taped run
    bool taping = false;                   // a taped tf_run is in progress
    // solutions of a taped run's nodes kept for its adjoint sweep, which then
    // skips recomputing them (tuning key "keep_states": while the kept
    // states take at most a quarter of the device memory and a fifth of it
    // stays free; 0: none), in chunks of TF_STATE_CHUNK states
    std::vector<double *> tstate;         // by node; nullptr: recomputed from the checkpoints
    std::vector<double *> tstate_chunks;
    int tstate_used = 0;                  // states taken from the last chunk
    size_t tstate_bytes = 0;
    bool keep_states = true;
    // independent models in one context (tf_set_models): block-diagonal
    // system, one tree of fronts per model, per-model controller state
    std::vector<FrontDesc> fd_host;       // front descriptors, offsets of slot 0
    int nmodels = 0;                      // 0: a single model
    std::vector<int> m_tet0, m_ntet, m_rec0, m_nrec, m_e0, m_ne;
    std::vector<int64_t> m_llo, m_lhi;    // each model's range of L storage (within a slot)
    std::vector<int> front_model;
    FrontDesc *d_fd_cur = nullptr, *d_fd_oth = nullptr;  // fronts in each model's current / other slot
    double *d_Lmm = nullptr, *d_Dmm = nullptr;          // both slots of every model
    int *d_qmodel = nullptr, *d_mrange = nullptr;
    int64_t *d_slot_off = nullptr;
    double *d_mcdt = nullptr, *d_mcoef = nullptr, *d_mnorm = nullptr, *d_mstage = nullptr;
    bool use_graphs = true;
    bool models_factored = false;  // every model's current slot holds a factor
    bool overlap_factor = true;  // tf_run: factor P(2 dt) on st2 ahead of the check
    // rows of the last tf_run_models: all receivers per step, per-model metadata
    std::vector<int> mm_step;                  // [steps]
    std::vector<double> mm_t, mm_dt;           // [steps][M]
    std::vector<char> mm_on;                   // [steps][M]
    double mm_graph_steps = 0.0;
    // optional per-category kernel timing with HIP events on the context's stream
    bool prof_on = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<int> prof_rec;  // triples: category, start event, end event
    double prof_ms[TF_PROF_NCAT] = {0};
    int64_t prof_cnt[TF_PROF_NCAT] = {0};
    int leaf_size = 192;  // nested-dissection leaf: larger leaves give fewer, larger fronts but more fill (fastest of 128-384 on a 1.35M-tet mesh with 64 right-hand sides)
};

namespace {

void free_models(tf_ctx *c);
void drop_kept(tf_ctx *c);

void free_solver(tf_ctx *c) {
    free_models(c);
    dfree(c->d_fd); dfree(c->d_perm); dfree(c->d_parent); dfree(c->d_chp); dfree(c->d_ch); dfree(c->d_rel);
    dfree(c->d_sidx); dfree(c->d_plist); dfree(c->d_amap); dfree(c->d_ppref); dfree(c->d_units); dfree(c->d_gptr); dfree(c->d_gidx); dfree(c->d_U); dfree(c->d_V);
    dfree(c->d_xw); dfree(c->d_Y); dfree(c->d_T); dfree(c->d_W); dfree(c->d_woff);
    for (int s = 0; s < 2; s++) { dfree(c->d_L[s]); dfree(c->d_D[s]); dfree(c->d_LT[s]); c->slot_live[s] = false; }
    drop_kept(c);
    dfree(c->d_tunits);
    dfree(c->d_frows);
    c->n_tunits = 0;
    dfree(c->d_tmpb); dfree(c->d_tmpx);
    dfree(c->d_mV); dfree(c->d_mY); dfree(c->d_mT); dfree(c->d_mb); dfree(c->d_mx);
    dfree(c->d_mZ); dfree(c->d_part);
    c->mK = 0;
    c->analyzed = false;
}

void free_vecs(tf_ctx *c) {
    for (auto &p : c->vecs) dfree(p);
    c->vecs.clear();
    dfree(c->d_norm_part); dfree(c->d_norm_out); dfree(c->d_rows);
    c->rows_cap = 0;
}

void free_models(tf_ctx *c) {
    dfree(c->d_fd_cur); dfree(c->d_fd_oth); dfree(c->d_Lmm); dfree(c->d_Dmm); dfree(c->d_qmodel);
    dfree(c->d_mrange); dfree(c->d_slot_off); dfree(c->d_mcdt); dfree(c->d_mcoef); dfree(c->d_mnorm);
    dfree(c->d_mstage);
    c->models_factored = false;
}

void drop_kept(tf_ctx *c) {
    for (auto &k : c->kept) { dfree(k.L); dfree(k.D); dfree(k.LT); }
    c->kept.clear();
}

void drop_states(tf_ctx *c) {
    for (auto &p : c->tstate_chunks) dfree(p);
    c->tstate_chunks.clear();
    c->tstate.clear();
    c->tstate_used = 0;
    c->tstate_bytes = 0;
}

void free_tape(tf_ctx *c) {
    drop_kept(c);
    drop_states(c);
    for (auto &k : c->ckpts) { dfree(k.s2); dfree(k.s3); }
    c->ckpts.clear();
    c->tape.clear();
    c->row_node.clear();
    c->tape_valid = false;
}

void free_slices(tf_ctx *c) {
    dfree(c->d_slice); dfree(c->d_sliceB);
    c->d_slice = nullptr; c->d_sliceB = nullptr;
    c->slice_nT = 0;
}

void free_native(tf_ctx *c) {
    dfree(c->d_ident); dfree(c->d_nfia); dfree(c->d_nfja); dfree(c->d_nfmap); dfree(c->d_nfval);
    dfree(c->d_nsrc_edge); dfree(c->d_nrec_edge); dfree(c->d_ntet_edg); dfree(c->d_nmrange);
    c->nm_e0.clear();
    c->native_ok = false;
}

void free_all(tf_ctx *c) {
    free_native(c);
    free_slices(c);
    free_models(c);
    free_solver(c);
    free_vecs(c);
    free_tape(c);
    dfree(c->d_tet_edg); dfree(c->d_vm1); dfree(c->d_sig9); dfree(c->d_nu9);
    dfree(c->d_xyz); dfree(c->d_sigma); dfree(c->d_invmu); dfree(c->d_vm); dfree(c->d_vk);
    dfree(c->d_tet_vex); dfree(c->d_tet_drc); dfree(c->d_ctr); dfree(c->d_orient); dfree(c->d_cptr);
    dfree(c->d_M); dfree(c->d_K); dfree(c->d_fia); dfree(c->d_fja); dfree(c->d_fmap); dfree(c->d_fval);
    dfree(c->d_rec_xyz); dfree(c->d_narr); dfree(c->d_darr); dfree(c->d_rec_drc); dfree(c->d_rec_tet);
    dfree(c->d_rec_edge); dfree(c->d_src_edge); dfree(c->d_src_id); dfree(c->d_src_w); dfree(c->d_flag);
}

int fail(tf_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

// Kernel timing: prof_mark records an event before a launch, prof_close one
// after it and files the pair under a category.  Pairs are folded into the
// per-category totals when the pool fills up and on tf_get_profile.
void prof_fold(tf_ctx *c) {
    if (c->prof_rec.empty()) return;
    HIPCHK(hipStreamSynchronize(c->st));  // events are recorded on both streams
    HIPCHK(hipStreamSynchronize(c->st2));
    for (size_t i = 0; i + 3 <= c->prof_rec.size(); i += 3) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev_pool[c->prof_rec[i + 1]], c->ev_pool[c->prof_rec[i + 2]]));
        c->prof_ms[c->prof_rec[i]] += ms;
        c->prof_cnt[c->prof_rec[i]] += 1;
    }
    c->prof_rec.clear();
    c->ev_used = 0;
}
int prof_event(tf_ctx *c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    HIPCHK(hipEventRecord(c->ev_pool[c->ev_used], c->st));
    return (int)c->ev_used++;
}
inline int prof_mark(tf_ctx *c) {
    if (!c->prof_on) return -1;
    if (c->ev_used + 2 > 8192) prof_fold(c);
    return prof_event(c);
}
inline void prof_close(tf_ctx *c, int cat, int mark) {
    if (mark < 0) return;
    const int e = prof_event(c);
    c->prof_rec.push_back(cat);
    c->prof_rec.push_back(mark);
    c->prof_rec.push_back(e);
}

template <class F>
int guarded(tf_ctx *c, F &&f) {
    if (!c) return TF_ERR_ARG;
    try {
        HIPCHK(hipSetDevice(c->device));
        return f();
    } catch (const HipError &e) {
        return fail(c, TF_ERR_HIP, e.what());
    } catch (const NumericError &e) {
        return fail(c, TF_ERR_NUMERIC, e.what());
    } catch (const std::exception &e) {
        return fail(c, TF_ERR_INPUT, e.what());
    }
}

// ---- plan --------------------------------------------------------------------
void build_plan(tf_ctx *c) {
    const Symbolic &s = c->sym;
    Plan &p = c->plan;
    p = Plan();
    auto f_of = [&](int k) { return s.nc[k] + (int)(s.sptr[k + 1] - s.sptr[k]); };
    auto m_of = [&](int k) { return (int)(s.sptr[k + 1] - s.sptr[k]); };
    auto pinv_of = [&](int k) { return s.level[k] < c->pinv_levels; };
    auto add = [&](std::vector<Launch> &dst, int kind, int p0, const std::vector<int> &fr,
                   const std::vector<int64_t> &units) {
        Launch L{kind, p0, (int)p.list.size(), (int)fr.size(), (int)p.pref.size(), 0, 0};
        p.list.insert(p.list.end(), fr.begin(), fr.end());
        int64_t acc = 0;
        for (size_t i = 0; i < fr.size(); i++) {
            p.pref.push_back(acc);
            acc += units.empty() ? 1 : units[i];
        }
        p.pref.push_back(acc);
        L.nwg = acc;
        if (!fr.empty() && acc > 0) dst.push_back(L);
    };
    // v = [b(own); 0] + extend-add of the children's update vectors, for the
    // fronts `fr` of one level: one init launch, then one launch per child
    // rank (children of one parent never share a launch)
    auto gathers = [&](std::vector<Launch> &dst, const std::vector<int> &fr) {
        std::vector<int64_t> ui;
        int maxch = 0;
        for (int k : fr) {
            ui.push_back((f_of(k) + 31) / 32);
            maxch = std::max(maxch, s.children_ptr[k + 1] - s.children_ptr[k]);
        }
        add(dst, L_GINIT, 0, fr, ui);
        for (int r = 0; r < maxch; r++) {
            std::vector<int> chl;
            std::vector<int64_t> uc;
            for (int k : fr)
                if (s.children_ptr[k + 1] - s.children_ptr[k] > r) {
                    int ch = s.children[s.children_ptr[k] + r];
                    if (m_of(ch) > 0) { chl.push_back(ch); uc.push_back((m_of(ch) + 31) / 32); }
                }
            add(dst, L_GCHILD, 0, chl, uc);
        }
    };
    // the single-RHS matrix-vector solve (s_tiles 0) takes small fronts one
    // workgroup each and large fronts in grid launches
    std::vector<std::vector<int>> small(s.nlevels), grid(s.nlevels);
    for (int l = 0; l < s.nlevels; l++)
        for (int q = s.lev_ptr[l]; q < s.lev_ptr[l + 1]; q++) {
            int k = s.lev_sn[q];
            if (s.nc[k] <= TF_PB && f_of(k) <= 512) small[l].push_back(k);
            else grid[l].push_back(k);
        }
    p.woff.assign(s.nsn, 0);
    // factorisation: every front of a level in the same grid launches (the
    // fronts of at most 64 columns too: one workgroup per such front, with
    // LDS loops and a barrier per pivot, made the factorisation of the
    // batched mesh 10 ms slower)
    for (int l = 0; l < s.nlevels; l++) {
        const std::vector<int> lev(s.lev_sn.begin() + s.lev_ptr[l], s.lev_sn.begin() + s.lev_ptr[l + 1]);
        int maxch = 0;
        for (int q = s.lev_ptr[l]; q < s.lev_ptr[l + 1]; q++) {
            int k = s.lev_sn[q];
            maxch = std::max(maxch, s.children_ptr[k + 1] - s.children_ptr[k]);
        }
        for (int r = 0; r < maxch; r++) {
            std::vector<int> fr;
            std::vector<int64_t> un;
            for (int q = s.lev_ptr[l]; q < s.lev_ptr[l + 1]; q++) {
                int k = s.lev_sn[q];
                if (s.children_ptr[k + 1] - s.children_ptr[k] > r) {
                    int ch = s.children[s.children_ptr[k] + r];
                    if (m_of(ch) > 0) { fr.push_back(ch); un.push_back(m_of(ch)); }
                }
            }
            add(p.factor, L_EXTADD, 0, fr, un);
        }
        int kmax = 0;
        int64_t wlev = 0;
        for (int k : lev) {
            kmax = std::max(kmax, (s.nc[k] + TF_PB - 1) / TF_PB);
            // scratch: the pivot block's inverse during the panels (64 x 64),
            // then the largest merge level's T blocks of the inversion
            int64_t need = (int64_t)TF_PB * TF_PB;
            for (int b = TF_PB; b < s.nc[k]; b *= 2) {
                int64_t npair = (s.nc[k] - b + 2 * b - 1) / (2 * b);
                need = std::max(need, npair * inv_ldt(s.nc[k], b) * (int64_t)b);
            }
            p.woff[k] = wlev;
            wlev += need;
        }
        p.wsize = std::max(p.wsize, wlev);
        // panels of TF_PB columns grouped into super-panels of TF_SP columns:
        // after each panel, its update of the rest of its super-panel
        // (mode 0); after a super-panel's last panel, one rank-TF_SP update
        // of everything to its right (mode 1)
        for (int pk = 0; pk < kmax; pk++) {
            int p0 = pk * TF_PB, P0 = p0 - p0 % TF_SP;
            std::vector<int> act, a0, a1;
            std::vector<int64_t> ut, u0, u1;
            for (int k : lev)
                if (s.nc[k] > p0) {
                    int p1 = std::min(s.nc[k], p0 + TF_PB), pe = std::min(s.nc[k], P0 + TF_SP);
                    int rows = f_of(k) - p1;
                    act.push_back(k);
                    ut.push_back((rows + TF_TRSM_ROWS - 1) / TF_TRSM_ROWS);
                    if (p1 < pe) {  // 64 x 64 update tiles, one per wave
                        int64_t nt = (rows + 63) / 64, ncj = (pe - p1 + 63) / 64;
                        a0.push_back(k);
                        u0.push_back(ncj * nt - ncj * (ncj - 1) / 2);
                    } else if (f_of(k) > pe) {
                        int64_t nt = (f_of(k) - pe + 63) / 64;
                        a1.push_back(k);
                        u1.push_back(nt * (nt + 1) / 2);
                    }
                }
            add(p.factor, L_FDIAG, p0, act, {});
            add(p.factor, L_FTRSM, p0, act, ut);
            add(p.factor, L_FUPD, p0, a0, u0);
            if (!a1.empty()) {
                add(p.factor, L_FUPD, P0, a1, u1);
                p.factor.back().aux = 1;
            }
        }
        // in-place inversion of every front's L11 by doubling: all 64 x 64
        // diagonal blocks, then merges of block pairs of size b = 64, 128, ...
        {
            std::vector<int64_t> ud;
            for (int k : lev) ud.push_back((s.nc[k] + TF_PB - 1) / TF_PB);
            add(p.factor, L_IDIAG, 0, lev, ud);
        }
        for (int b = TF_PB;; b *= 2) {
            std::vector<int> act;
            std::vector<int64_t> um;
            const int64_t ntb = (b + TF_UT - 1) / TF_UT;
            for (int k : lev)
                if (s.nc[k] > b) {  // ntb x ntb regions per pair; the last pair only those of its J rows
                    const int64_t npair = (s.nc[k] - b + 2 * b - 1) / (2 * b);
                    const int64_t jlast = std::min<int64_t>(b, s.nc[k] - (2 * b * (npair - 1) + b));
                    act.push_back(k);
                    um.push_back((npair - 1) * ntb * ntb + (jlast + TF_UT - 1) / TF_UT * ntb);
                }
            if (act.empty()) break;
            add(p.factor, L_IMERGE_T, b, act, um);
            add(p.factor, L_IMERGE_X, b, act, um);
        }
        // partitioned inverse of the level's fronts that keep it (after their
        // inversion): one launch per 64-column block, left to right
        for (int cb = 0;; cb += 64) {
            std::vector<int> pv;
            std::vector<int64_t> up;
            for (int q = s.lev_ptr[l]; q < s.lev_ptr[l + 1]; q++) {
                const int k = s.lev_sn[q];
                if (pinv_of(k) && m_of(k) > 0 && s.nc[k] > cb) { pv.push_back(k); up.push_back((m_of(k) + 63) / 64); }
            }
            if (pv.empty()) break;
            add(p.factor, L_PINV, cb, pv, up);
        }
    }
    // forward solve, bottom-up
    for (int l = 0; l < s.nlevels; l++) {
        add(p.fwd, L_SSMALL, 0, small[l], {});
        std::vector<int64_t> uo, us;
        for (int k : grid[l]) {
            uo.push_back((s.nc[k] + 63) / 64);
            us.push_back((m_of(k) + 63) / 64);
        }
        gathers(p.fwd, grid[l]);
        add(p.fwd, L_SOWN, 0, grid[l], uo);
        add(p.fwd, L_SSTRUCT, 0, grid[l], us);
    }
    // backward solve, top-down
    for (int l = s.nlevels - 1; l >= 0; l--) {
        std::vector<int64_t> uc;
        for (int k : grid[l]) uc.push_back((s.nc[k] + 15) / 16);
        std::vector<int64_t> ux;
        for (int k : grid[l]) ux.push_back((m_of(k) + 31) / 32);
        add(p.bwd, L_GXS, 0, grid[l], ux);
        add(p.bwd, L_BT, 0, grid[l], uc);
        add(p.bwd, L_BX, 0, grid[l], uc);
        add(p.bwd, L_BSMALL, 0, small[l], {});
    }
    // block solve for K right-hand sides: every front through the tile kernels
    // mode 0 own forward, 1 struct forward, 2 backward T, 3 backward X
    struct Tile { int k, t0, lo, hi; };
    // 64-row tiles on levels of at most mt_wide fronts (few large fronts),
    // 32-row tiles below
    auto nb_of = [&](const std::vector<int> &fr) { return (int)fr.size() <= c->mt_wide ? 4 : 2; };
    auto tiles_of = [&](int mode, const std::vector<int> &fr, int TW = 0) {
        std::vector<Tile> t;
        if (TW == 0) TW = 16 * nb_of(fr);  // output rows (columns) per tile
        for (int k : fr) {
            const int nc = s.nc[k], m = m_of(k);
            if (mode == 0)
                for (int r = 0; r < nc; r += TW) t.push_back({k, r, 0, std::min(nc, r + TW)});
            else if (mode == 1) {  // row tiles start on 16-row lines of the column (rows nc + r)
                if (m > 0)
                    for (int r = -(nc & 15); r < m; r += TW) t.push_back({k, r, 0, nc});
            }
            else if (mode == 2)
                for (int j = 0; j < nc; j += TW) t.push_back({k, j, 0, m});
            else if (mode == 3)
                for (int j = 0; j < nc; j += TW) t.push_back({k, j, j, nc});
            else if (mode == 5)  // partitioned forward: all f rows, row r reads columns < min(r, nc)
                for (int r = 0; r < nc + m; r += TW) t.push_back({k, r, 0, std::min(nc, r + TW)});
            else  // partitioned backward (7): column j reads rows >= j of all f rows
                for (int j = 0; j < nc; j += TW) t.push_back({k, j, j, nc + m});
        }
        return t;
    };
    // one tile launch (+ its reduction): tiles longer than the chunk are split
    // so that a launch has about mt_units units
    auto add_tiles = [&](std::vector<Launch> &dst, int mode, const std::vector<Tile> &tl, int nb, int64_t utarget = 0,
                         int chmin = 0, int kt = L_MT, int kr = L_MTRED) {
        if (tl.empty()) return;
        if (utarget == 0) utarget = c->mt_units;
        if (chmin == 0) chmin = c->mt_chmin;
        int64_t work = 0;
        for (const Tile &t : tl) work += t.hi - t.lo;
        int64_t ch = (work + utarget - 1) / utarget;
        ch = std::max<int64_t>(chmin, (ch + 15) / 16 * 16);
        // units of one front are emitted chunk-major (all tiles of inner chunk
        // 0, then chunk 1, ...): neighbouring waves then share the right-hand
        // side rows of their inner chunk in L1/L2
        std::vector<TileUnit> tu, ru;
        int slots = 0;
        size_t t0 = 0;
        while (t0 < tl.size()) {
            size_t t1 = t0;
            while (t1 < tl.size() && tl[t1].k == tl[t0].k) t1++;
            std::vector<int> nk(t1 - t0), base(t1 - t0);
            int maxnk = 0;
            for (size_t j = t0; j < t1; j++) {
                const int len = tl[j].hi - tl[j].lo;
                nk[j - t0] = len <= ch ? 1 : (int)((len + ch - 1) / ch);
                maxnk = std::max(maxnk, nk[j - t0]);
                if (nk[j - t0] > 1) {
                    base[j - t0] = slots;
                    ru.push_back({tl[j].k, tl[j].t0, slots, nk[j - t0], 0, 0});
                    slots += nk[j - t0];
                }
            }
            for (int q = 0; q < maxnk; q++)
                for (size_t j = t0; j < t1; j++) {
                    const Tile &t = tl[j];
                    const int n = nk[j - t0], len = t.hi - t.lo;
                    if (q >= n) continue;
                    if (n == 1) { tu.push_back({t.k, t.t0, t.lo, t.hi, -1, 0}); continue; }
                    const int a = t.lo + (int)((int64_t)len * q / n / 16 * 16);
                    const int b = (q + 1 == n) ? t.hi : t.lo + (int)((int64_t)len * (q + 1) / n / 16 * 16);
                    tu.push_back({t.k, t.t0, a, b, base[j - t0] + q, 0});
                }
            t0 = t1;
        }
        dst.push_back(Launch{kt, mode, (int)p.units.size(), (int)tu.size(), 0, (int64_t)(tu.size() + 3) / 4, 0, nb});
        p.units.insert(p.units.end(), tu.begin(), tu.end());
        if (!ru.empty()) {
            dst.push_back(Launch{kr, mode, (int)p.units.size(), (int)ru.size(), 0, (int64_t)ru.size(), 0, nb});
            p.units.insert(p.units.end(), ru.begin(), ru.end());
        }
        int64_t &pmax = (kt == L_ST) ? p.spart_max : p.part_max;
        pmax = std::max<int64_t>(pmax, slots);
    };
    for (int l = 0; l < s.nlevels; l++) {
        std::vector<int> all(s.lev_sn.begin() + s.lev_ptr[l], s.lev_sn.begin() + s.lev_ptr[l + 1]);
        std::vector<int64_t> ur;
        for (int k : all) ur.push_back((s.nc[k] + 4 * TF_RPW - 1) / (4 * TF_RPW));
        if (l < c->pinv_levels) {  // partitioned inverse: one product per front
            if (l > 0) add(p.mfwd, L_MPULL, 0, all, ur);
            add_tiles(p.mfwd, l == 0 ? 6 : 5, tiles_of(5, all), nb_of(all));
            continue;
        }
        if (l == 0) {  // fronts without children: Yin is b itself, read through perm by the tiles
            add_tiles(p.mfwd, 4, tiles_of(0, all), nb_of(all));
        } else {
            add(p.mfwd, L_MPULL, 0, all, ur);
            add_tiles(p.mfwd, 0, tiles_of(0, all), nb_of(all));
        }
        add_tiles(p.mfwd, 1, tiles_of(1, all), nb_of(all));
    }
    // the backward sweep, top level first
    for (int l = s.nlevels - 1; l >= 0; l--) {
        std::vector<int> all(s.lev_sn.begin() + s.lev_ptr[l], s.lev_sn.begin() + s.lev_ptr[l + 1]);
        if (l < c->pinv_levels) {
            const size_t nb = p.mbwd.size();
            add_tiles(p.mbwd, 7, tiles_of(7, all), nb_of(all));
            if (l == 0)
                for (size_t i = nb; i < p.mbwd.size(); i++) p.mbwd[i].aux = 1;
            continue;
        }
        add_tiles(p.mbwd, 2, tiles_of(2, all), nb_of(all));
        const size_t nb = p.mbwd.size();
        add_tiles(p.mbwd, 3, tiles_of(3, all), nb_of(all));
        if (l == 0)  // no lower level gathers these rows: edge order only
            for (size_t i = nb; i < p.mbwd.size(); i++) p.mbwd[i].aux = 1;
    }
    // One right-hand side through the same four products, with one 16-wide
    // right-hand-side block per tile: 64-row tiles on every level (a tile's
    // matrix-core work is a quarter of the K = 64 tile's, so the factor
    // stream sets its pace), own launch sizing (s_units, s_chmin), and the
    // pull as a thread per row.
    for (int l = 0; l < s.nlevels; l++) {
        std::vector<int> all(s.lev_sn.begin() + s.lev_ptr[l], s.lev_sn.begin() + s.lev_ptr[l + 1]);
        if (l < c->pinv_levels) {  // partitioned inverse: one product per front
            if (l > 0) {
                std::vector<int64_t> ur;
                for (int k : all) ur.push_back((s.nc[k] + 63) / 64);
                add(p.sfwd, L_SPULL, 0, all, ur);
            }
            add_tiles(p.sfwd, l == 0 ? 6 : 5, tiles_of(5, all, 64), 4, c->s_units, c->s_chmin, L_ST, L_STRED);
            continue;
        }
        if (l == 0) {
            add_tiles(p.sfwd, 4, tiles_of(0, all, 64), 4, c->s_units, c->s_chmin, L_ST, L_STRED);
        } else {
            std::vector<int64_t> ur;
            for (int k : all) ur.push_back((s.nc[k] + 63) / 64);
            add(p.sfwd, L_SPULL, 0, all, ur);
            add_tiles(p.sfwd, 0, tiles_of(0, all, 64), 4, c->s_units, c->s_chmin, L_ST, L_STRED);
        }
        add_tiles(p.sfwd, 1, tiles_of(1, all, 64), 4, c->s_units, c->s_chmin, L_ST, L_STRED);
    }
    for (int l = s.nlevels - 1; l >= 0; l--) {
        std::vector<int> all(s.lev_sn.begin() + s.lev_ptr[l], s.lev_sn.begin() + s.lev_ptr[l + 1]);
        if (l < c->pinv_levels) {
            const size_t nb = p.sbwd.size();
            add_tiles(p.sbwd, 7, tiles_of(7, all, 64), 4, c->s_units, c->s_chmin, L_ST, L_STRED);
            if (l == 0)
                for (size_t i = nb; i < p.sbwd.size(); i++) p.sbwd[i].aux = 1;
            continue;
        }
        add_tiles(p.sbwd, 2, tiles_of(2, all, 64), 4, c->s_units, c->s_chmin, L_ST, L_STRED);
        const size_t nb = p.sbwd.size();
        add_tiles(p.sbwd, 3, tiles_of(3, all, 64), 4, c->s_units, c->s_chmin, L_ST, L_STRED);
        if (l == 0)
            for (size_t i = nb; i < p.sbwd.size(); i++) p.sbwd[i].aux = 1;
    }
}

void do_analyze(tf_ctx *c, int leaf) {
    free_solver(c);
    free_native(c);
    auto t0 = std::chrono::steady_clock::now();
    tf::analyze(c->sym, c->topo, leaf);
    Symbolic &s = c->sym;
    // front descriptors
    std::vector<FrontDesc> fd(s.nsn);
    int64_t voff = 0;
    for (int k = 0; k < s.nsn; k++) {
        int m = (int)(s.sptr[k + 1] - s.sptr[k]);
        fd[k].loff = s.loff[k];
        fd[k].uoff = s.uoff[k];
        fd[k].soff = s.sptr[k];
        fd[k].voff = voff;
        fd[k].nc = s.nc[k];
        fd[k].m = m;
        fd[k].f = s.nc[k] + m;
        fd[k].ld = (int)tf::front_ld(fd[k].f);
        fd[k].doff = s.c0[k];
        fd[k].c0 = s.c0[k];
        fd[k].ldt = (s.nc[k] + 15) & ~15;
        fd[k].toff = 0;
        fd[k].pinv = s.level[k] < c->pinv_levels ? 1 : 0;
        voff += fd[k].f;
    }
    c->vrows = voff;
    {  // row-major L21 store (all f rows of a partitioned-inverse front): offsets and the copy tiles
        int64_t toff = 0;
        std::vector<int> tu;
        for (int k = 0; k < s.nsn; k++) {
            const int rows = fd[k].pinv ? fd[k].f : fd[k].m;
            fd[k].toff = toff;
            toff += (int64_t)rows * fd[k].ldt;
            for (int r0 = 0; r0 < rows; r0 += 64)
                for (int c0 = 0; c0 < fd[k].nc; c0 += 64) {
                    tu.push_back(k);
                    tu.push_back(r0);
                    tu.push_back(c0);
                }
        }
        c->lt_total = toff + TF_READ_SLACK;
        c->n_tunits = (int64_t)tu.size() / 3;
        c->d_tunits = dupload(tu);
    }
    {  // the rows of every partitioned-inverse front in the [v; x] block vector:
        // own rows c0 + i (v), struct rows n + sidx (x)
        std::vector<int> fr(std::max<int64_t>(voff, 1), 0);
        for (int k = 0; k < s.nsn; k++)
            if (fd[k].pinv) {
                for (int i = 0; i < fd[k].nc; i++) fr[fd[k].voff + i] = fd[k].c0 + i;
                for (int i = 0; i < fd[k].m; i++) fr[fd[k].voff + fd[k].nc + i] = s.n + s.sidx[fd[k].soff + i];
            }
        c->d_frows = dupload(fr);
    }
    build_plan(c);
    c->d_fd = dupload(fd);
    c->fd_host = fd;
    c->d_perm = dupload(s.perm);
    c->d_parent = dupload(s.parent);
    c->d_chp = dupload(s.children_ptr);
    c->d_ch = dupload(s.children);
    c->d_rel = dupload(s.rel);
    c->d_sidx = dupload(s.sidx);
    c->d_amap = dupload(s.amap);
    c->d_plist = dupload(c->plan.list);
    c->d_ppref = dupload(c->plan.pref);
    c->d_units = dupload(c->plan.units);
    {  // rows receiving each struct entry's update: inverse of sidx, in entry order
        std::vector<int> gptr(s.n + 1, 0), gidx(s.sidx.size());
        for (int r : s.sidx) gptr[r + 1]++;
        for (int i = 0; i < s.n; i++) gptr[i + 1] += gptr[i];
        std::vector<int> fill(gptr.begin(), gptr.end() - 1);
        for (size_t pz = 0; pz < s.sidx.size(); pz++) gidx[fill[s.sidx[pz]]++] = (int)pz;
        c->d_gptr = dupload(gptr);
        c->d_gidx = dupload(gidx);
    }
    c->d_U = dalloc<double>(s.utotal);
    // (the tile kernels' unmasked reads may run up to 63 values past a vector)
    c->d_V = dalloc<double>(voff + TF_READ_SLACK);
    c->d_xw = dalloc<double>(s.n + TF_READ_SLACK);
    c->d_Y = dalloc<double>(s.n + TF_READ_SLACK);
    c->d_T = dalloc<double>(s.n + TF_READ_SLACK);
    c->d_W = dalloc<double>((size_t)std::max<int64_t>(c->plan.wsize, 1));
    c->d_woff = dupload(c->plan.woff);
    c->d_part = dalloc<double>(part_doubles(c->plan));
    c->d_tmpb = dalloc<double>(s.n);
    c->d_tmpx = dalloc<double>(s.n);
    if (!c->d_flag) c->d_flag = dalloc<FactorFlags>(1);
    c->analyzed = true;
    c->timers[6] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The step's edge-indexed maps in native order (see tf_ctx::native).
void ensure_native(tf_ctx *c) {
    if (c->native_ok) return;
    const Topology &t = c->topo;
    const Symbolic &s = c->sym;
    const int n = t.nE;
    const std::vector<int> &iperm = s.iperm;
    std::vector<int> ident(n);
    for (int i = 0; i < n; i++) ident[i] = i;
    // full symmetric pattern of Msigma with rows and columns renumbered,
    // columns sorted within each row; nfmap points at the upper-CSR value
    std::vector<int> deg(n, 0);
    for (int r = 0; r < n; r++)
        for (int q = t.ia[r]; q < t.ia[r + 1]; q++) {
            deg[iperm[r]]++;
            if (t.ja[q] != r) deg[iperm[t.ja[q]]]++;
        }
    std::vector<int> fia(n + 1, 0);
    for (int i = 0; i < n; i++) fia[i + 1] = fia[i] + deg[i];
    std::vector<std::pair<int, int64_t>> ent(fia[n]);
    std::fill(deg.begin(), deg.end(), 0);
    for (int r = 0; r < n; r++)
        for (int q = t.ia[r]; q < t.ia[r + 1]; q++) {
            const int a = iperm[r], b = iperm[t.ja[q]];
            ent[fia[a] + deg[a]++] = {b, q};
            if (a != b) ent[fia[b] + deg[b]++] = {a, q};
        }
    std::vector<int> fja(fia[n]);
    std::vector<int64_t> fmap(fia[n]);
    for (int i = 0; i < n; i++) {
        std::sort(ent.begin() + fia[i], ent.begin() + fia[i + 1]);
        for (int p = fia[i]; p < fia[i + 1]; p++) { fja[p] = ent[p].first; fmap[p] = ent[p].second; }
    }
    free_native(c);
    c->d_ident = dupload(ident);
    c->d_nfia = dupload(fia);
    c->d_nfja = dupload(fja);
    c->d_nfmap = dupload(fmap);
    c->nfnnz = fia[n];
    c->d_nfval = dalloc<double>(std::max<int64_t>(c->nfnnz, 1));
    if (c->assembled) tf::launch_gather_vals(c->nfnnz, c->d_nfmap, c->d_M, c->d_nfval, c->st);
    if (!c->src_edge.empty()) {
        std::vector<int> se(c->src_edge.size());
        for (size_t i = 0; i < se.size(); i++) se[i] = iperm[c->src_edge[i]];
        c->d_nsrc_edge = dupload(se);
    }
    if (c->nR > 0) {
        std::vector<int> re(6 * (size_t)c->nR);
        for (int i = 0; i < c->nR; i++)
            for (int j = 0; j < 6; j++) re[6 * i + j] = iperm[t.tet_edg[6 * (size_t)(c->rec_tet[i] - 1) + j]];
        c->d_nrec_edge = dupload(re);
    }
    {
        std::vector<int> te(t.tet_edg.size());
        for (size_t i = 0; i < te.size(); i++) te[i] = iperm[t.tet_edg[i]];
        c->d_ntet_edg = dupload(te);
    }
    if (c->nmodels > 0) {  // each model's fronts are one run of the postorder: one range of rows
        std::vector<int> rg(2 * (size_t)c->nmodels);
        c->nm_e0.assign(c->nmodels, 0);
        for (int m = 0; m < c->nmodels; m++) {
            int lo = n, hi = -1;
            for (int e = c->m_e0[m]; e < c->m_e0[m] + c->m_ne[m]; e++) {
                lo = std::min(lo, iperm[e]);
                hi = std::max(hi, iperm[e]);
            }
            if (hi - lo + 1 != c->m_ne[m]) throw std::runtime_error("model rows are not contiguous in factor order");
            c->nm_e0[m] = lo;
            rg[2 * m] = lo;
            rg[2 * m + 1] = lo + c->m_ne[m];
        }
        c->d_nmrange = dupload(rg);
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->native_ok = true;
}

// the maps a time loop runs with, and its solves' permutation of b and x
tf_ctx::RunMaps run_maps(tf_ctx *c) {
    if (!c->d_tet_edg) c->d_tet_edg = dupload(c->topo.tet_edg);
    if (c->native) {
        ensure_native(c);
        return tf_ctx::RunMaps{c->d_nfia, c->d_nfja, c->d_nfval, c->d_nsrc_edge, c->d_nrec_edge, c->d_ntet_edg,
                               c->d_nmrange, c->d_ident, c->nm_e0.data()};
    }
    return tf_ctx::RunMaps{c->d_fia, c->d_fja, c->d_fval, c->d_src_edge, c->d_rec_edge, c->d_tet_edg,
                           c->d_mrange, c->d_perm, c->m_e0.data()};
}

// Sets a time loop's maps for the lifetime of the guard: the step kernels
// index through c->rm, the solves map b and x through c->solve_perm.
struct RunScope {
    tf_ctx *c;
    explicit RunScope(tf_ctx *cc) : c(cc) {
        c->rm = run_maps(c);
        c->solve_perm = c->rm.perm;
    }
    ~RunScope() { c->solve_perm = nullptr; }
};

// the permutation the solves map b and x through
inline const int *sperm(const tf_ctx *c) { return c->solve_perm ? c->solve_perm : c->d_perm; }

// numeric factorisation of the values already scattered into slot's L
// numeric factorisation into L / Dn, fronts addressed through fd
void run_factor(tf_ctx *c, double *L, double *Dn, const FrontDesc *fd) {
    const Plan &p = c->plan;
    const int *lst = c->d_plist;
    const int64_t *pr = c->d_ppref;
    for (const Launch &l : p.factor) {
        const int pm = prof_mark(c);
        switch (l.kind) {
        case L_EXTADD:
            tf::launch_extend_add(L, c->d_U, fd, c->d_parent, lst + l.off, pr + l.poff, l.cnt, l.nwg, c->d_rel,
                                  c->st);
            break;
        case L_FDIAG:
            tf::launch_factor_diag(L, fd, lst + l.off, l.cnt, l.p0, Dn, &c->d_flag->bad, c->d_W, c->d_woff, c->st);
            break;
        case L_FTRSM:
            tf::launch_factor_trsm(L, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, l.p0, Dn, c->d_W, c->d_woff, c->st);
            break;
        case L_FUPD:
            tf::launch_factor_update(L, c->d_U, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, l.p0, l.aux, Dn, c->f_band,
                                     c->st);
            break;
        case L_IMERGE_T:
            tf::launch_inv_merge(0, L, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, l.p0, c->d_W, c->d_woff, c->st);
            break;
        case L_IMERGE_X:
            tf::launch_inv_merge(1, L, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, l.p0, c->d_W, c->d_woff, c->st);
            break;
        case L_IDIAG:
            tf::launch_inv_diag(L, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, c->st);
            break;
        case L_PINV:
            tf::launch_pinv_w21(L, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, l.p0, c->st);
            break;
        }
        prof_close(c, prof_category(l.kind), pm);
    }
    tf::launch_pivot_scan(Dn, fd, c->sym.nsn, c->d_flag, c->st);
    HIPCHK(hipGetLastError());
}

// the slot's row-major L21 copy, when block solves use one
void copy_l21(tf_ctx *c, int slot) {
    if (!c->d_LT[slot]) return;
    const int pm = prof_mark(c);
    tf::launch_transpose_l21(c->d_L[slot], c->d_LT[slot], c->d_fd, c->d_tunits, c->n_tunits, c->st);
    prof_close(c, TF_PROF_FACTOR_OTHER, pm);
}

void prepare_slot(tf_ctx *c, int slot) {
    if (!c->d_L[slot]) c->d_L[slot] = dalloc<double>(c->sym.ltotal + TF_READ_SLACK);
    if (!c->d_D[slot]) c->d_D[slot] = dalloc<double>(c->sym.n);
    if ((c->nsrc > 1 || c->s_lt) && !c->d_LT[slot]) c->d_LT[slot] = dalloc<double>(c->lt_total);
    tf::launch_fill0(c->d_L[slot], c->sym.ltotal, c->st);
    tf::launch_fill0(c->d_U, c->sym.utotal, c->st);
    tf::launch_flags_init(c->d_flag, c->st);
}

// The pivot report of the factorisation just finished.  P = 3 Msigma +
// 2 dt K is symmetric positive definite in exact arithmetic; the reference
// factors it as SPD (PARDISO mtype 2) and, when that fails, again as
// symmetric indefinite with pivoting (mtype -2, src/FETD_UTILITY.f90:919-935).
// This factorisation has no pivoting: a negative pivot is reported as
// TF_ERR_NUMERIC unless the caller accepts indefinite factors (tuning key
// "indefinite"); a zero or non-finite pivot always is.
void check_flag(tf_ctx *c) {
    FactorFlags f;
    HIPCHK(hipMemcpyAsync(&f, c->d_flag, sizeof(f), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->last_nneg = (long long)f.nneg;
    c->last_min_abs = __builtin_bit_cast(double, f.min_abs);
    c->last_max_abs = __builtin_bit_cast(double, f.max_abs);
    if (f.bad) throw NumericError("zero or non-finite pivot in LDL^T factorisation");
    if (f.nneg && !c->allow_indefinite)
        throw NumericError("P is not positive definite: " + std::to_string(f.nneg) +
                           " negative pivots in LDL^T (the reference would refactor it as symmetric indefinite, "
                           "PARDISO mtype -2; set tuning key \"indefinite\" to 1 to accept the unpivoted factor)");
}

// A taped run's factor about to be overwritten is kept for the adjoint
// sweep instead, when the cap and the device's free memory allow (a tenth
// of the device stays free): its buffers move to the kept list, no copy, and
// the slot is given new ones.
void keep_slot(tf_ctx *c, int slot) {
    if ((int)c->kept.size() >= c->keep_factors || !c->d_L[slot] || !c->d_D[slot]) return;
    for (const auto &k : c->kept)
        if (k.dt == c->slot_dt[slot]) return;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return;
    const size_t need = sizeof(double) * ((size_t)c->sym.ltotal + TF_READ_SLACK + (size_t)c->sym.n +
                                          (c->d_LT[slot] ? (size_t)c->lt_total : 0));
    if (fr < need + tot / 10) return;
    // the slot's new buffers first, without throwing: if they cannot be had,
    // nothing is kept and the slot is overwritten as usual
    void *L = nullptr, *D = nullptr, *LT = nullptr;
    const bool ok = hipMalloc(&L, sizeof(double) * ((size_t)c->sym.ltotal + TF_READ_SLACK)) == hipSuccess &&
                    hipMalloc(&D, sizeof(double) * (size_t)c->sym.n) == hipSuccess &&
                    (!c->d_LT[slot] || hipMalloc(&LT, sizeof(double) * (size_t)c->lt_total) == hipSuccess);
    if (!ok) {
        (void)hipGetLastError();
        if (L) (void)hipFree(L);
        if (D) (void)hipFree(D);
        return;
    }
    c->kept.push_back(tf_ctx::KeptFactor{c->slot_dt[slot], c->d_L[slot], c->d_D[slot], c->d_LT[slot]});
    c->d_L[slot] = (double *)L;
    c->d_D[slot] = (double *)D;
    c->d_LT[slot] = (double *)LT;
}

void factor_dt(tf_ctx *c, int slot, double dt) {
    if (c->taping && c->slot_taped[slot]) keep_slot(c, slot);
    c->slot_taped[slot] = c->taping;
    prepare_slot(c, slot);
    tf::launch_scatter_p(c->topo.nnz, c->d_amap, c->d_M, c->d_K, 2.0 * dt, c->d_L[slot], c->st);
    run_factor(c, c->d_L[slot], c->d_D[slot], c->d_fd);
    copy_l21(c, slot);
    c->slot_live[slot] = true;
    c->slot_dt[slot] = dt;
}

// Work vectors of the single right-hand-side solve: V (every front's
// f-row vector, vrows doubles), Y, T and the permuted solution xw (n each).
// With the tile kernels V holds Yin (n) followed by Z (one value per struct
// row), and part the split tiles' partial sums.
struct SolveWork {
    double *V, *Y, *T, *xw, *part;
};
SolveWork ctx_work(tf_ctx *c) { return SolveWork{c->d_V, c->d_Y, c->d_T, c->d_xw, c->d_part}; }

void solve_dev_gemv(tf_ctx *c, const double *L, const double *Dn, const FrontDesc *fd, const double *b, double *x,
                    const SolveWork &w);

// x = P^-1 b (device vectors, edge order), work vectors w; x2, if given,
// receives a second copy of x.  The four products of every front (own and
// struct forward, backward T and X) run as tile launches per level with one
// right-hand side; the forward pull forms each level's Yin from b and the
// descendants' Z, the backward X epilogue writes x in edge order.
// LT: the factor's row-major L21 copy (backward products as N-form tiles), or
// nullptr (T-form tiles over L itself)
void solve_dev(tf_ctx *c, const double *L, const double *Dn, const FrontDesc *fd, const double *b, double *x,
               const SolveWork &w, double *x2 = nullptr, const double *LT = nullptr) {
    if (!c->s_tiles) {
        solve_dev_gemv(c, L, Dn, fd, b, x, w);
        if (x2) HIPCHK(hipMemcpyAsync(x2, x, sizeof(double) * c->sym.n, hipMemcpyDeviceToDevice, c->st));
        return;
    }
    const Plan &p = c->plan;
    const int *lst = c->d_plist;
    const int64_t *pr = c->d_ppref;
    const int64_t n = c->sym.n;
    const uint32_t gbytes = (uint32_t)(n * 8);
    double *Yin = w.V, *Z = w.V + n;
    // b already in the factor's column order (time loops): the lowest level
    // reads its rows contiguously, as the products of a pulled Yin do
    const bool ident = sperm(c) == c->d_ident;
    for (const Launch &l : p.sfwd) {
        const int pm = prof_mark(c);
        const bool fromb = l.p0 == 4 || l.p0 == 6;
        const int md = fromb && ident ? (l.p0 == 6 ? 5 : 0) : l.p0;
        const bool pin = l.p0 == 5 || l.p0 == 6;  // partitioned inverse: v = Y_own / d into T, Z rows
        const double *in0 = fromb ? b : ((l.p0 == 0 || l.p0 == 5) ? Yin : w.Y);
        double *out = pin ? w.T : (l.p0 == 1 ? Z : w.Y);
        const bool gath = md == 4 || md == 6;
        switch (l.kind) {
        case L_SPULL:
            tf::launch_spull(fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, b, sperm(c), c->d_gptr, c->d_gidx, Z, Yin,
                             c->st);
            break;
        case L_ST:
            if (c->s_tiles == 2)
                tf::launch_stile(md, c->s_tcols, L, fd, c->d_units + l.off, l.cnt, in0, nullptr, Dn, nullptr, out,
                                 w.part, sperm(c), nullptr, nullptr, c->st, Z);
            else
                tf::launch_mtile(md, c->mt_ks_n, 4, 1, L, fd, c->d_units + l.off, l.cnt, 1, in0, nullptr, Dn, nullptr,
                                 out, w.part, gath ? gbytes : 0u, gath ? sperm(c) : nullptr, nullptr, nullptr,
                                 nullptr, c->st, Z);
            break;
        case L_STRED:
            if (c->s_tiles == 2)
                tf::launch_stile_red(md, fd, c->d_units + l.off, l.cnt, in0, nullptr, Dn, out, w.part,
                                     gath ? sperm(c) : nullptr, nullptr, nullptr, c->st, Z);
            else
                tf::launch_mtile_red(md, 4, 1, fd, c->d_units + l.off, l.cnt, 1, in0, nullptr, Dn, out, w.part,
                                     gath ? sperm(c) : nullptr, nullptr, nullptr, c->st, Z);
            break;
        }
        prof_close(c, prof_category(l.kind), pm);
    }
    for (const Launch &l : p.sbwd) {
        const int pm = prof_mark(c);
        const bool bx = l.p0 == 3 || l.p0 == 7;
        double *out = bx ? (l.aux == 1 ? nullptr : w.xw) : w.T;
        const bool nt = LT && c->s_lt && c->s_tiles == 2;  // N form over the row-major copy
        if (l.p0 == 7) {  // partitioned inverse: [v; x(struct)] from T and xw
            if (l.kind == L_ST && nt)
                tf::launch_stile_lt(7, LT, fd, c->d_units + l.off, l.cnt, w.xw, w.T, Dn, c->d_sidx, out, w.part,
                                    sperm(c), x, x2, c->st);
            else if (l.kind == L_ST)
                tf::launch_stile(7, c->s_tcols, L, fd, c->d_units + l.off, l.cnt, w.xw, w.T, Dn, c->d_sidx, out,
                                 w.part, sperm(c), x, x2, c->st);
            else
                tf::launch_stile_red(7, fd, c->d_units + l.off, l.cnt, w.xw, w.T, Dn, out, w.part, sperm(c), x, x2,
                                     c->st);
            prof_close(c, prof_category(l.kind), pm);
            continue;
        }
        switch (l.kind) {
        case L_ST:
            if (nt && l.p0 == 2)
                tf::launch_stile_lt(2, LT, fd, c->d_units + l.off, l.cnt, w.xw, w.Y, Dn, c->d_sidx, out, w.part,
                                    sperm(c), nullptr, nullptr, c->st);
            else if (c->s_tiles == 2)
                tf::launch_stile(l.p0, c->s_tcols, L, fd, c->d_units + l.off, l.cnt, bx ? w.T : w.xw, bx ? w.T : w.Y,
                                 Dn, c->d_sidx, out, w.part, sperm(c), bx ? x : nullptr, bx ? x2 : nullptr, c->st);
            else
                tf::launch_mtile(l.p0, c->mt_ks_t, 4, 1, L, fd, c->d_units + l.off, l.cnt, 1, bx ? w.T : w.xw,
                                 bx ? w.T : w.Y, Dn, c->d_sidx, out, w.part, gbytes, sperm(c), bx ? x : nullptr,
                                 bx ? x2 : nullptr, nullptr, c->st);
            break;
        case L_STRED:
            if (c->s_tiles == 2)
                tf::launch_stile_red(l.p0, fd, c->d_units + l.off, l.cnt, bx ? w.T : w.xw, bx ? w.T : w.Y, Dn, out,
                                     w.part, sperm(c), bx ? x : nullptr, bx ? x2 : nullptr, c->st);
            else
                tf::launch_mtile_red(l.p0, 4, 1, fd, c->d_units + l.off, l.cnt, 1, bx ? w.T : w.xw, bx ? w.T : w.Y, Dn,
                                     out, w.part, sperm(c), bx ? x : nullptr, bx ? x2 : nullptr, c->st);
            break;
        }
        prof_close(c, prof_category(l.kind), pm);
    }
}

// The same solve through the matrix-vector kernels of solve.hip (tuning key
// "s_tiles" = 0): every front's update vector is extend-added into its
// parent's front vector in V.
void solve_dev_gemv(tf_ctx *c, const double *L, const double *Dn, const FrontDesc *fd, const double *b, double *x,
                    const SolveWork &w) {
    const Plan &p = c->plan;
    const int *lst = c->d_plist;
    const int64_t *pr = c->d_ppref;
    for (const Launch &l : p.fwd) {
        const int pm = prof_mark(c);
        switch (l.kind) {
        case L_SSMALL:
            tf::launch_fwd_small(L, fd, lst + l.off, l.cnt, b, sperm(c), c->d_chp, c->d_ch, c->d_rel, w.V,
                                 w.Y, c->st);
            break;
        case L_GINIT:
            tf::launch_mg_init(fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, b, sperm(c), w.V, 1, c->st);
            break;
        case L_GCHILD:
            tf::launch_mg_child(fd, c->d_parent, lst + l.off, pr + l.poff, l.cnt, l.nwg, c->d_rel, w.V, 1,
                                c->st);
            break;
        case L_SOWN:
            tf::launch_fwd_own(L, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, w.V, w.Y, c->st);
            break;
        case L_SSTRUCT:
            tf::launch_fwd_struct(L, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, w.V, w.Y, c->st);
            break;
        }
        prof_close(c, prof_category(l.kind), pm);
    }
    for (const Launch &l : p.bwd) {
        const int pm = prof_mark(c);
        switch (l.kind) {
        case L_BSMALL:
            tf::launch_bwd_small(L, fd, lst + l.off, l.cnt, w.Y, Dn, c->d_sidx, w.xw, c->st);
            break;
        case L_GXS:
            tf::launch_mg_xs(fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, c->d_sidx, w.xw, w.V, 1, c->st);
            break;
        case L_BT:
            tf::launch_bwd_t(L, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, w.Y, Dn, w.V, w.T, c->st);
            break;
        case L_BX:
            tf::launch_bwd_x(L, fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, w.T, w.xw, c->st);
            break;
        }
        prof_close(c, prof_category(l.kind), pm);
    }
    tf::launch_unpermute(c->sym.n, sperm(c), w.xw, x, c->st);
}
void solve_dev(tf_ctx *c, const double *L, const double *Dn, const FrontDesc *fd, const double *b, double *x) {
    solve_dev(c, L, Dn, fd, b, x, ctx_work(c));
}

void ensure_block(tf_ctx *c, int K) {
    if (c->mK >= K) return;
    dfree(c->d_mV); dfree(c->d_mY); dfree(c->d_mT); dfree(c->d_mb); dfree(c->d_mx); dfree(c->d_mZ);
    const int64_t n = c->sym.n;
    // the tile kernels' unmasked reads of the last rows may run up to 63
    // values past an array (see TF_READ_SLACK)
    c->d_mV = dalloc<double>((size_t)n * K + TF_READ_SLACK);  // Yin: own rows after the descendants' updates
    c->d_mZ = dalloc<double>(std::max<size_t>(c->sym.sidx.size(), 1) * K);  // per-front L21 Y
    c->d_mY = dalloc<double>((size_t)n * K + TF_READ_SLACK);
    // T (v = Y_own / d on the rows of partitioned-inverse fronts) and the
    // front-order solution, adjacent: the [v; x] block vector that a
    // partitioned-inverse front's backward product gathers its rows from
    c->d_mT = dalloc<double>(2 * (size_t)n * K + TF_READ_SLACK);
    c->d_mb = dalloc<double>((size_t)n * K);
    c->d_mx = dalloc<double>((size_t)n * K);
    c->mK = K;
}

// X = P^-1 B for K <= 64 right-hand sides, [n][K] device arrays in edge order;
// x2, if given, receives a second copy of X
void solve_block(tf_ctx *c, const double *L, const double *Dn, const FrontDesc *fd, int K, const double *b,
                 double *x, double *x2 = nullptr, const double *LT = nullptr) {
    ensure_block(c, K);
    // the tile kernels address the gathered solution vector with 32-bit byte offsets
    const uint64_t gb = (uint64_t)c->sym.n * K * 8;
    if (2 * gb >= (1ull << 32)) throw std::runtime_error("block solve: 2 n x K too large for the tile kernels (>= 4 GiB)");
    const uint32_t gbytes = (uint32_t)gb, gbytes2 = (uint32_t)(2 * gb);
    // the front-order solution right after T (rows n .. 2n of the [v; x] block vector at this K)
    double *const mxw = c->d_mT + (size_t)c->sym.n * K;
    const int na = K <= 16 ? 1 : 4;  // 16-wide right-hand-side blocks per tile
    const Plan &p = c->plan;
    // b already in the factor's column order (time loops): the lowest level
    // reads its rows contiguously (modes 5 and 0 on b instead of 6 and 4)
    const bool ident = sperm(c) == c->d_ident;
    const int *lst = c->d_plist;
    const int64_t *pr = c->d_ppref;
    for (const Launch &l : p.mfwd) {
        const int pm = prof_mark(c);
        switch (l.kind) {
        case L_MPULL:
            tf::launch_mpull(fd, lst + l.off, pr + l.poff, l.cnt, l.nwg, b, sperm(c), c->d_gptr, c->d_gidx,
                             c->d_mZ, c->d_mV, K, c->st);
            break;
        case L_MT:  // own forward (Y from Yin, or from b), struct forward (Z from Y)
            if (l.p0 == 5 || l.p0 == 6)  // partitioned inverse: v rows into T, Z rows
                tf::launch_mtile(l.p0 == 6 && !ident ? 6 : 5, c->mt_ks_n, l.nb, na, L, fd, c->d_units + l.off, l.cnt,
                                 K, l.p0 == 6 ? b : c->d_mV, nullptr, Dn, nullptr, c->d_mT, c->d_part,
                                 l.p0 == 6 && !ident ? gbytes : 0u, sperm(c), nullptr, nullptr, nullptr, c->st,
                                 c->d_mZ);
            else if (l.p0 == 4)
                tf::launch_mtile(ident ? 0 : 4, c->mt_ks_n, l.nb, na, L, fd, c->d_units + l.off, l.cnt, K, b, nullptr,
                                 Dn, nullptr, c->d_mY, c->d_part, ident ? 0u : gbytes, sperm(c), nullptr, nullptr,
                                 nullptr, c->st);
            else
                tf::launch_mtile(l.p0, c->mt_ks_n, l.nb, na, L, fd, c->d_units + l.off, l.cnt, K,
                                 l.p0 == 0 ? c->d_mV : c->d_mY, nullptr, Dn, nullptr, l.p0 == 0 ? c->d_mY : c->d_mZ,
                                 c->d_part, 0u, nullptr, nullptr, nullptr, nullptr, c->st);
            break;
        case L_MTRED:
            if (l.p0 == 5 || l.p0 == 6)
                tf::launch_mtile_red(l.p0 == 6 && !ident ? 6 : 5, l.nb, na, fd, c->d_units + l.off, l.cnt, K,
                                     l.p0 == 6 ? b : c->d_mV, nullptr, Dn, c->d_mT, c->d_part, sperm(c), nullptr,
                                     nullptr, c->st, c->d_mZ);
            else if (l.p0 == 4)
                tf::launch_mtile_red(ident ? 0 : 4, l.nb, na, fd, c->d_units + l.off, l.cnt, K, b, nullptr, Dn, c->d_mY,
                                     c->d_part, sperm(c), nullptr, nullptr, c->st);
            else
                tf::launch_mtile_red(l.p0, l.nb, na, fd, c->d_units + l.off, l.cnt, K, l.p0 == 0 ? c->d_mV : c->d_mY, nullptr,
                                     Dn, l.p0 == 0 ? c->d_mY : c->d_mZ, c->d_part, nullptr, nullptr, nullptr, c->st);
            break;
        }
        prof_close(c, prof_category(l.kind), pm);
    }
    // backward T (from x at the struct rows, Y, D), backward X (from T); the
    // backward X epilogues write the finished rows in edge order as well
    for (const Launch &l : p.mbwd) {
        const int pm = prof_mark(c);
        const bool bx = l.p0 == 3 || l.p0 == 7;
        double *out = bx ? (l.aux == 1 ? nullptr : mxw) : c->d_mT;
        if (l.p0 == 7) {  // partitioned inverse: rows of [v; x] through the front's row list
            if (l.kind == L_MT)
                tf::launch_mtile(7, LT ? c->mt_ks_n : c->mt_ks_t, l.nb, na, L, fd, c->d_units + l.off, l.cnt, K,
                                 c->d_mT, c->d_mT, Dn, c->d_frows, out, c->d_part, gbytes2, sperm(c), x, x2, LT, c->st);
            else
                tf::launch_mtile_red(7, l.nb, na, fd, c->d_units + l.off, l.cnt, K, c->d_mT, c->d_mT, Dn, out,
                                     c->d_part, sperm(c), x, x2, c->st);
            prof_close(c, prof_category(l.kind), pm);
            continue;
        }
        switch (l.kind) {
        case L_MT:
            tf::launch_mtile(l.p0, c->mt_ks_t, l.nb, na, L, fd, c->d_units + l.off, l.cnt, K, bx ? c->d_mT : mxw,
                             bx ? c->d_mT : c->d_mY, Dn, c->d_sidx, out, c->d_part, gbytes, sperm(c),
                             bx ? x : nullptr, bx ? x2 : nullptr, LT, c->st);
            break;
        case L_MTRED:
            tf::launch_mtile_red(l.p0, l.nb, na, fd, c->d_units + l.off, l.cnt, K, bx ? c->d_mT : mxw,
                                 bx ? c->d_mT : c->d_mY, Dn, out, c->d_part, sperm(c), bx ? x : nullptr,
                                 bx ? x2 : nullptr, c->st);
            break;
        }
        prof_close(c, prof_category(l.kind), pm);
    }
}

// ---- time loop -------------------------------------------------------------------
// Executable graphs of a run, keyed by the buffers they were captured on.
struct GraphCache {
    std::map<std::array<const void *, 6>, hipGraphExec_t> m;
    ~GraphCache() {
        for (auto &kv : m) (void)hipGraphExecDestroy(kv.second);
    }
};

struct StepVecs {
    double *e1, *e2, *e3, *rh, *te1, *te2, *te3, *erec, *ev1, *ev2;
    double *u[3];  // M e of the history buffers vecs[0..2], in that order (see run_loop)
    std::vector<double *> ring;
};

void ensure_vecs(tf_ctx *c, StepVecs &v) {
    const int64_t len = (int64_t)c->topo.nE * c->nsrc;
    int need = 13 + c->sp_dblesize + 1;
    if ((int)c->vecs.size() != need || c->vec_len != len) {
        for (auto &p : c->vecs) dfree(p);
        c->vecs.clear();
        for (int i = 0; i < need; i++) c->vecs.push_back(dalloc<double>(len));
        c->vec_len = len;
    }
    v.e1 = c->vecs[0]; v.e2 = c->vecs[1]; v.e3 = c->vecs[2]; v.rh = c->vecs[3];
    v.te1 = c->vecs[4]; v.te2 = c->vecs[5]; v.te3 = c->vecs[6]; v.erec = c->vecs[7];
    v.ev1 = c->vecs[8]; v.ev2 = c->vecs[9];
    for (int i = 0; i < 3; i++) v.u[i] = c->vecs[10 + i];
    v.ring.assign(c->vecs.begin() + 13, c->vecs.end());
    if (!c->d_norm_part) c->d_norm_part = dalloc<double>(256 * 64 * 2);
    if (!c->d_norm_out) c->d_norm_out = dalloc<double>(2 * 64);
}

// r -= 2 dt s(t) while the source is on (sys_right: step < sp_division)
void rhs_source(tf_ctx *c, int step, double t, double dt, bool negate, double *r) {
    if (step >= c->sp_division) return;
    const int K = c->nsrc;
    const tf_ctx::RunMaps &m = c->rm;
    double didt = source_didt(c->stype, c->ptype, t, c->pulse_width, c->current);
    if (negate) didt = -didt;
    if (K == 1) tf::launch_source(c->nseg, m.src_edge, c->d_src_w, didt, 2.0 * dt, r, c->st);
    else tf::launch_msource(c->nseg, K, m.src_edge, c->d_src_id, c->d_src_w, didt, 2.0 * dt, r, c->st);
}

// r = Msigma (4 e2 - e3) - 2 dt s(t), K sources side by side
void rhs(tf_ctx *c, int step, double t, const double *e2, const double *e3, double dt, bool negate, double *r) {
    const int K = c->nsrc;
    const int pm = prof_mark(c);
    const tf_ctx::RunMaps &m = c->rm;
    if (K == 1) tf::launch_spmv_rhs(c->topo.nE, m.fia, m.fja, m.fval, e2, e3, r, c->st);
    else tf::launch_mspmv_rhs(c->topo.nE, K, m.fia, m.fja, m.fval, e2, e3, c->d_mb, r, c->spmv_xcd, c->st);
    rhs_source(c, step, t, dt, negate, r);
    prof_close(c, TF_PROF_STEP, pm);
}

void dcopy(tf_ctx *c, double *dst, const double *src) {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * c->topo.nE * c->nsrc, hipMemcpyDeviceToDevice, c->st));
}

#define TF_STATE_CHUNK 32
// Keeps node k's solution x for the adjoint sweep when the budget allows.
void keep_state(tf_ctx *c, int k, const double *x) {
    if (!c->keep_states) return;
    const size_t len = (size_t)c->topo.nE * c->nsrc;
    if (c->tstate_chunks.empty() || c->tstate_used == TF_STATE_CHUNK) {
        const size_t bytes = sizeof(double) * len * TF_STATE_CHUNK;
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return;
        if (c->tstate_bytes + bytes > tot / 4 || fr < bytes + tot / 5) return;
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {  // not kept: the sweep recomputes this state
            (void)hipGetLastError();
            return;
        }
        c->tstate_chunks.push_back((double *)p);
        c->tstate_used = 0;
        c->tstate_bytes += bytes;
    }
    if ((int)c->tstate.size() <= k) c->tstate.resize(k + 1, nullptr);
    double *p = c->tstate_chunks.back() + len * (size_t)c->tstate_used++;
    dcopy(c, p, x);
    c->tstate[k] = p;
}

// x2, if given, receives a copy of the solution
void step_solve(tf_ctx *c, int slot, const double *b, double *x, double *x2 = nullptr) {
    if (c->nsrc == 1) {
        solve_dev(c, c->d_L[slot], c->d_D[slot], c->d_fd, b, x, ctx_work(c), x2, c->d_LT[slot]);
    } else {
        solve_block(c, c->d_L[slot], c->d_D[slot], c->d_fd, c->nsrc, b, x, x2, c->d_LT[slot]);
    }
}

void grow_rows(tf_ctx *c) {
    const int64_t per = (int64_t)c->nR * 6 * c->nsrc;
    if (c->nrows + 1 > c->rows_cap) {
        int64_t cap = std::max<int64_t>(256, 2 * c->rows_cap);
        double *nb = dalloc<double>(cap * per);
        if (c->d_rows && c->nrows > 0)
            HIPCHK(hipMemcpyAsync(nb, c->d_rows, sizeof(double) * per * c->nrows, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        dfree(c->d_rows);
        c->d_rows = nb;
        c->rows_cap = cap;
    }
}

// a readout staged by a replayed step graph becomes the next row
void output_row_from(tf_ctx *c, const double *stage) {
    const int64_t per = (int64_t)c->nR * 6 * c->nsrc;
    grow_rows(c);
    HIPCHK(hipMemcpyAsync(c->d_rows + per * c->nrows, stage, sizeof(double) * per, hipMemcpyDeviceToDevice, c->st));
    c->row_step.push_back(0);
    c->row_t.push_back(0.0);
    c->row_dt.push_back(0.0);
    c->nrows++;
}

// Field slices on an output row (solution_output's vtk part,
// src/FETD_FWD.f90:1642-1656): while a slice is pending, every row updates
// the per-tet fields; a pending slice whose time has come takes them.
void slice_row(tf_ctx *c, int step, double t, double dt, const double *e) {
    const int n = (int)c->slice_t.size(), T = c->topo.nT;
    std::vector<int> take;
    bool pending = false;
    for (int i = 0; i < n; i++) {
        if (c->slice_done[i]) continue;
        pending = true;
        if (t >= c->slice_t[i]) take.push_back(i);
    }
    if (!pending) return;
    double *out = take.empty() ? nullptr : c->d_slice + (int64_t)take[0] * T * 9;
    tf::launch_slice(T, c->d_xyz, c->d_tet_vex, c->rm.tet_edg, c->d_tet_drc, e, dt, step == 0 ? 1 : 0, c->d_sliceB,
                     out, c->st);
    for (size_t q = 0; q < take.size(); q++) {
        const int i = take[q];
        if (q > 0)
            HIPCHK(hipMemcpyAsync(c->d_slice + (int64_t)i * T * 9, out, sizeof(double) * 9 * T,
                                  hipMemcpyDeviceToDevice, c->st));
        c->slice_done[i] = 1;
        c->slice_step[i] = step;
        c->slice_tt[i] = t;
    }
}

// slice storage for a run that writes output rows
void prepare_slices(tf_ctx *c) {
    const int n = (int)c->slice_t.size(), T = c->topo.nT;
    std::fill(c->slice_done.begin(), c->slice_done.end(), 0);
    if (n == 0) return;
    if (c->nsrc != 1) throw std::runtime_error("field slices need a single-source context");
    if (c->slice_nT != T || !c->d_slice) {
        free_slices(c);
        c->d_slice = dalloc<double>((size_t)n * T * 9);
        c->d_sliceB = dalloc<double>((size_t)T * 3);
        c->slice_nT = T;
    }
    if (!c->d_tet_edg) c->d_tet_edg = dupload(c->topo.tet_edg);
}

void output_row(tf_ctx *c, int step, double t, double dt, const double *e) {
    const int64_t per = (int64_t)c->nR * 6 * c->nsrc;
    if (c->nrows + 1 > c->rows_cap) {
        int64_t cap = std::max<int64_t>(256, 2 * c->rows_cap);
        double *nb = dalloc<double>(cap * per);
        if (c->d_rows && c->nrows > 0)
            HIPCHK(hipMemcpyAsync(nb, c->d_rows, sizeof(double) * per * c->nrows, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        dfree(c->d_rows);
        c->d_rows = nb;
        c->rows_cap = cap;
    }
    double *o = c->d_rows + per * c->nrows;
    if (c->nsrc == 1) tf::launch_readout(c->nR, c->rm.rec_edge, c->d_rec_drc, c->d_narr, c->d_darr, e, o, c->st);
    else tf::launch_mreadout(c->nR, c->nsrc, c->rm.rec_edge, c->d_rec_drc, c->d_narr, c->d_darr, e, o, c->st);
    if (!c->slice_t.empty()) slice_row(c, step, t, dt, e);
    c->row_step.push_back(step);
    c->row_t.push_back(t);
    c->row_dt.push_back(dt);
    c->nrows++;
}

// two_error (src/FETD_UTILITY.f90:578-599): ||a-b|| / ||b||; for K sources
// the largest of the K per-source values decides
double two_error(tf_ctx *c, const double *a, const double *b) {
    const int K = c->nsrc;
    std::vector<double> h(2 * K);
    if (K == 1) tf::launch_diffnorm(c->topo.nE, a, b, c->d_norm_part, c->d_norm_out, c->st);
    else tf::launch_mdiffnorm(c->topo.nE, K, a, b, c->d_norm_part, c->d_norm_out, c->st);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_norm_out, sizeof(double) * 2 * K, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double e = 0.0;
    for (int q = 0; q < K; q++) e = std::max(e, std::sqrt(h[2 * q]) / std::sqrt(h[2 * q + 1]));
    return e;
}

double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    return ms * 1e-3;
}

// adap_euler (src/FETD_FWD.f90:1229-1464).  inout: 0 plain run with output,
// 1 first pass of a step-off run (negated source, no output, result kept in
// res), 2 second pass starting from res.
void run_loop(tf_ctx *c, int inout, double *res, int max_steps) {
    StepVecs v;
    ensure_vecs(c, v);
    if (c->nsrc > 1) ensure_block(c, c->nsrc);
    const int n = c->topo.nE;
    const int dbl = c->sp_dblesize;
    if (c->sp_double < dbl + 2) throw std::runtime_error("sp_double must be at least sp_dblesize+2");
    int step = 0, not_changed = 0, idx = 0;
    // the reference's e1 (handed to a step-off second pass) is the last
    // regular or kept solution: after a check that changed dt it is the
    // regular step before the check (now e3 here), otherwise e2
    bool last_changed = false;
    double time_cur = 0.0, step_cur = c->pulse_width / c->sp_division;
    const double time_limit = c->time_max;
    const bool negate = (inout == 1), outp = (inout != 1);
    const int ring = dbl + 1;
    int cur = 0;
    for (double *p : {v.e1, v.e2, v.e3}) HIPCHK(hipMemsetAsync(p, 0, sizeof(double) * n * c->nsrc, c->st));
    // Regular steps of a block run take M (4 e2 - e3) = 4 M e2 - M e3 with M e3
    // kept from the step before (the history buffers rotate, so M e of each of
    // the three buffers lives in the product buffer paired with it; anything
    // written into a history buffer other than by rotation marks its product
    // stale, and a stale M e3 is formed again before the step).  Taped runs
    // keep the direct form, which the adjoint's recomputation repeats.
    const bool tape_run = c->ckpt_every > 0 && inout == 0;
    const bool chain = c->rhs_chain && c->nsrc > 1 && c->nsrc % 2 == 0 && !tape_run;
    double *const ebuf[3] = {v.e1, v.e2, v.e3};
    bool ufresh[3] = {false, false, false};
    auto uslot = [&](const double *p) {
        for (int i = 0; i < 3; i++)
            if (ebuf[i] == p) return i;
        throw std::logic_error("history buffer not in the product table");
    };
    auto touch = [&](const double *p) { ufresh[uslot(p)] = false; };
    // regular step's right-hand side into r (the chain's M e3 formed already)
    auto rhs_regular = [&](double *r) {
        if (!chain) {
            rhs(c, step, time_cur, v.e2, v.e3, step_cur, negate, r);
            return;
        }
        const int pm = prof_mark(c);
        const int i2 = uslot(v.e2), i3 = uslot(v.e3);
        tf::launch_mspmv_chain(n, c->nsrc, c->rm.fia, c->rm.fja, c->rm.fval, v.e2, v.u[i3], v.u[i2], r,
                               c->spmv_xcd, c->st);
        rhs_source(c, step, time_cur, step_cur, negate, r);
        prof_close(c, TF_PROF_STEP, pm);
    };
    auto chain_ready = [&] {  // M e3 for the next regular step, formed eagerly when stale
        if (!chain) return;
        const int i3 = uslot(v.e3);
        if (ufresh[i3]) return;
        const int pm = prof_mark(c);
        tf::launch_mspmv(n, c->nsrc, c->rm.fia, c->rm.fja, c->rm.fval, v.e3, v.u[i3], c->spmv_xcd, c->st);
        prof_close(c, TF_PROF_STEP, pm);
        ufresh[i3] = true;
    };
    if (outp) prepare_slices(c);
    // factors from before this run may be of other materials: none is reused
    c->slot_live[0] = c->slot_live[1] = false;
    c->slot_taped[0] = c->slot_taped[1] = false;
    drop_kept(c);    // what the last taped run kept is of no use once this run starts
    drop_states(c);
    // adjoint tape: node ids of the history states e2, e3 and of the ring
    const bool tape = c->ckpt_every > 0 && inout == 0;
    c->taping = tape && c->keep_factors > 0;
    struct TapingOff {
        tf_ctx *c;
        ~TapingOff() { c->taping = false; }
    } taping_off{c};
    int id2 = 0, id3 = 0;
    std::vector<int> ring_id(ring, 0);
    if (tape) {
        for (auto &k : c->ckpts) { dfree(k.s2); dfree(k.s3); }
        c->ckpts.clear();
        c->tape.assign(1, tf_ctx::TNode{-1, -1, 0, 0, 0.0, 0.0});  // node 0: the zero initial state
        c->row_node.assign(1, 0);
    }
    c->tape_valid = false;
    auto node = [&](int u, int v, int st, double dt, double t) {
        c->tape.push_back(tf_ctx::TNode{u, v, st, -1, dt, t});
        return (int)c->tape.size() - 1;
    };
    auto row_of = [&](int k) {  // the output row just written comes from node k
        c->tape[k].row = c->nrows - 1;
        c->row_node.push_back(k);
    };
    GraphCache graphs;
    struct StreamDrain {  // no factorisation stays in flight past this run
        hipStream_t s;
        ~StreamDrain() { (void)hipStreamSynchronize(s); }
    } drain{c->st2};
    hipEvent_t f0, f1, a0, a1, a2;
    for (hipEvent_t *e : {&f0, &f1, &a0, &a1, &a2}) HIPCHK(hipEventCreate(e));
    auto timed_factor = [&](int slot, double dt) {
        HIPCHK(hipEventRecord(f0, c->st));
        factor_dt(c, slot, dt);
        HIPCHK(hipEventRecord(f1, c->st));
        check_flag(c);
        c->timers[1] += elapsed(f0, f1);
        c->timers[7] += 1.0;
    };
    // The check at step S needs P(2 dt) in the other slot.  It is factored on
    // the second stream as soon as dt is known (step 1, or a check that
    // changed dt), while the steps before S run on the first: the two touch
    // disjoint buffers (the slot being factored, U and the inversion scratch
    // against the current slot and the solve vectors).  The check waits for
    // it.  Only started when the loop will reach the next check.
    int pend = -1;
    double pend_dt = 0.0;
    auto reaches_next_check = [&]() {
        const int S = (step / c->sp_double + 1) * c->sp_double;
        if (max_steps > 0 && S > max_steps) return false;
        double t = time_cur;
        for (int k = step + 1; k < S; k++) t = t + step_cur;
        return t < time_limit;
    };
    auto start_factor = [&](int slot, double dt) {
        if (!c->overlap_factor || !reaches_next_check()) return;
        HIPCHK(hipEventRecord(a0, c->st));  // the slot's last solve is enqueued before this point
        HIPCHK(hipStreamWaitEvent(c->st2, a0, 0));
        std::swap(c->st, c->st2);
        HIPCHK(hipEventRecord(a1, c->st));
        factor_dt(c, slot, dt);
        HIPCHK(hipEventRecord(a2, c->st));
        std::swap(c->st, c->st2);
        pend = slot;
        pend_dt = dt;
    };
    auto finish_factor = [&] {
        HIPCHK(hipStreamWaitEvent(c->st, a2, 0));
        check_flag(c);
        c->timers[1] += elapsed(a1, a2);
        c->timers[7] += 1.0;
        pend = -1;
    };
    while (time_cur < time_limit) {
        if (max_steps > 0 && step >= max_steps) break;
        step++;
        time_cur = time_cur + step_cur;
        if (tape && (step - 1) % c->ckpt_every == 0) {  // history checkpoint before this step
            const size_t bytes = sizeof(double) * n * c->nsrc;
            tf_ctx::Ckpt k{(int)c->tape.size(), id2, id3, dalloc<double>(n * c->nsrc), dalloc<double>(n * c->nsrc)};
            HIPCHK(hipMemcpyAsync(k.s2, v.e2, bytes, hipMemcpyDeviceToDevice, c->st));
            HIPCHK(hipMemcpyAsync(k.s3, v.e3, bytes, hipMemcpyDeviceToDevice, c->st));
            c->ckpts.push_back(k);
        }
        if (step == 1) {
            if (inout == 2) { dcopy(c, v.e2, res); dcopy(c, v.e3, res); touch(v.e2); touch(v.e3); }
            if (outp) output_row(c, 0, time_cur - step_cur, step_cur, v.e2);
            timed_factor(cur, step_cur);
            start_factor(1 - cur, step_cur * dbl);
        }
        if (step % c->sp_double == 0) {
            dcopy(c, v.te2, v.e2);
            dcopy(c, v.te3, v.e3);
            double ttime = time_cur, tstep = step_cur;
            int fid[3] = {0, id2, id3};  // nodes of te1, te2, te3
            int f1id = 0;
            for (int i = 1; i <= dbl; i++) {
                rhs(c, step, ttime, v.te2, v.te3, tstep, negate, v.rh);
                step_solve(c, cur, v.rh, v.te1);
                if (tape) {
                    fid[0] = node(fid[1], fid[2], step, tstep, ttime);
                    keep_state(c, fid[0], v.te1);
                    if (i == 1) f1id = fid[0];
                    fid[2] = fid[1];
                    fid[1] = fid[0];
                }
                if (i == dbl) break;
                if (i == 1) dcopy(c, v.erec, v.te1);
                std::swap(v.te3, v.te2);   // te3 <- te2
                dcopy(c, v.te2, v.te1);    // te2 <- te1
                ttime = ttime + tstep;
            }
            dcopy(c, v.ev1, v.te1);
            tstep = step_cur * dbl;
            ttime = time_cur - step_cur + tstep;
            const int other = 1 - cur;
            // the other slot still holds P(tstep) when the previous check kept
            // dt: the same matrix, so the same factor (no refactorisation)
            if (pend >= 0) {
                if (pend != other || pend_dt != tstep) throw std::runtime_error("factorisation in flight for another slot");
                finish_factor();
            } else if (!(c->slot_live[other] && c->slot_dt[other] == tstep)) {
                timed_factor(other, tstep);
            }
            double *ea = v.ring[(idx - 1) % ring];
            double *eb = v.ring[(idx - 1 - dbl) % ring];
            rhs(c, step, ttime, ea, eb, tstep, negate, v.rh);
            step_solve(c, other, v.rh, v.ev2);
            const double error = two_error(c, v.ev2, v.ev1);
            // keep the step iff err > 1e-4 and fewer than 2 kept in a row
            // (src/FETD_FWD.f90:1381); a replayed schedule forces the decision
            bool keep = (error > 1.0e-4 && not_changed < 2);
            if (!c->replay.empty()) {  // indexed over all checks of the run (both passes of a step-off)
                const size_t ic = c->chk_step.size();
                if (ic >= c->replay.size()) throw std::runtime_error("replayed schedule is too short");
                keep = c->replay[ic] != 0;
            }
            c->chk_step.push_back(step);
            c->chk_err.push_back(error);
            c->chk_keep.push_back(keep ? 1 : 0);
            last_changed = !keep;
            if (keep) {
                dcopy(c, v.e1, v.erec);
                touch(v.e1);
                if (outp) output_row(c, step, time_cur, step_cur, v.e1);
                if (tape) { row_of(f1id); id3 = id2; id2 = f1id; }
                std::swap(v.e3, v.e2);  // e3 <- e2
                dcopy(c, v.e2, v.e1);   // e2 <- e1
                touch(v.e2);
                not_changed++;
                idx = 0;
            } else {
                c->slot_live[cur] = false;
                cur = other;
                dcopy(c, v.e3, ea);
                dcopy(c, v.e2, v.ev1);
                touch(v.e3);
                touch(v.e2);
                time_cur = ttime;
                step_cur = tstep;
                if (outp) output_row(c, step, time_cur, step_cur, v.e2);
                if (tape) { row_of(fid[0]); id3 = ring_id[(idx - 1) % ring]; id2 = fid[0]; }
                not_changed = 0;
                idx = 0;
                start_factor(1 - cur, step_cur * dbl);
            }
            continue;
        }
        if (step >= c->sp_division && c->use_graphs && !c->prof_on) {
            // regular step, sources off: the right-hand side and the solve
            // (with its ring copy) replay a graph captured for this factor slot and
            // arrangement of the history vectors (at most 2 x 3 x ring of them)
            // (the product buffers follow from the history buffers, so the key covers them)
            const std::array<const void *, 6> key = {c->d_L[cur], v.e1, v.e2, v.e3, v.rh, v.ring[idx % ring]};
            chain_ready();
            auto it = graphs.m.find(key);
            if (it == graphs.m.end()) {
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
                rhs_regular(v.rh);
                step_solve(c, cur, v.rh, v.e1, v.ring[idx % ring]);
                HIPCHK(hipStreamEndCapture(c->st, &g));
                hipGraphExec_t ex;
                HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
                HIPCHK(hipGraphDestroy(g));
                it = graphs.m.emplace(key, ex).first;
            }
            HIPCHK(hipGraphLaunch(it->second, c->st));
        } else {
            chain_ready();
            rhs_regular(v.rh);
            step_solve(c, cur, v.rh, v.e1, v.ring[idx % ring]);
        }
        if (chain) ufresh[uslot(v.e2)] = true;  // M e2 formed by this step
        touch(v.e1);
        if (outp) output_row(c, step, time_cur, step_cur, v.e1);
        if (tape) {
            const int k = node(id2, id3, step, step_cur, time_cur);
            keep_state(c, k, v.e1);
            row_of(k);
            id3 = id2;
            id2 = k;
            ring_id[idx % ring] = k;
        }
        idx++;
        last_changed = false;
        // e3 <- e2 <- e1 by rotation; e1 is overwritten next step
        double *t3 = v.e3;
        v.e3 = v.e2;
        v.e2 = v.e1;
        v.e1 = t3;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipGetLastError());
    if (tape) c->tape_valid = ((int)c->row_node.size() == c->nrows);
    // adap_euler's res_e = e1 (src/FETD_FWD.f90:1459)
    if (inout == 1 && res) dcopy(c, res, last_changed ? v.e3 : v.e2);
    if (pend >= 0) {  // not reached by construction; drained all the same
        HIPCHK(hipStreamSynchronize(c->st2));
        check_flag(c);
    }
    for (hipEvent_t e : {f0, f1, a0, a1, a2}) HIPCHK(hipEventDestroy(e));
    // keep the rotated pointers in the context's vector table
    c->vecs[0] = v.e1; c->vecs[1] = v.e2; c->vecs[2] = v.e3; c->vecs[5] = v.te2; c->vecs[6] = v.te3;
}

}  // namespace

// ======================================================================
// C ABI
// ======================================================================
extern "C" {

// ---- independent models in one context ------------------------------------------
// Device set-up after the symbolic analysis: model of every front and CSR
// entry, both factor slots of every model in one buffer, the two front views.
void setup_models(tf_ctx *c) {
    const Symbolic &s = c->sym;
    const int M = c->nmodels, n = s.n;
    auto model_of_edge = [&](int e) {
        int m = (int)(std::upper_bound(c->m_e0.begin(), c->m_e0.end(), e) - c->m_e0.begin()) - 1;
        return m;
    };
    c->front_model.assign(s.nsn, 0);
    for (int k = 0; k < s.nsn; k++) {
        const int m = model_of_edge(s.perm[s.c0[k]]);
        for (int j = 0; j < s.nc[k]; j++)
            if (model_of_edge(s.perm[s.c0[k] + j]) != m) throw std::runtime_error("a front spans two models");
        c->front_model[k] = m;
    }
    // a model's fronts are one run of the postorder (one tree per model), so
    // its factor occupies one range of L storage in either slot
    c->m_llo.assign(M, INT64_MAX);
    c->m_lhi.assign(M, 0);
    for (int k = 0; k < s.nsn; k++) {
        const int m = c->front_model[k];
        c->m_llo[m] = std::min<int64_t>(c->m_llo[m], s.loff[k]);
        c->m_lhi[m] = std::max<int64_t>(c->m_lhi[m], s.loff[k + 1]);
    }
    {
        int64_t span = 0;
        for (int m = 0; m < M; m++) span += c->m_lhi[m] - c->m_llo[m];
        if (span != s.ltotal) throw std::runtime_error("a model's fronts are not contiguous");
    }
    std::vector<int> qm(c->topo.nnz);
    for (int r = 0; r < n; r++) {
        const int m = model_of_edge(r);
        for (int q = c->topo.ia[r]; q < c->topo.ia[r + 1]; q++) qm[q] = m;
    }
    std::vector<int> rg(2 * (size_t)M);
    for (int m = 0; m < M; m++) { rg[2 * m] = c->m_e0[m]; rg[2 * m + 1] = c->m_e0[m] + c->m_ne[m]; }
    free_models(c);
    c->d_qmodel = dupload(qm);
    c->d_mrange = dupload(rg);
    c->d_Lmm = dalloc<double>(2 * (size_t)s.ltotal + TF_READ_SLACK);
    c->d_Dmm = dalloc<double>(2 * (size_t)n);
    c->d_fd_cur = dalloc<FrontDesc>(s.nsn);
    c->d_fd_oth = dalloc<FrontDesc>(s.nsn);
    c->d_slot_off = dalloc<int64_t>(M);
    c->d_mcdt = dalloc<double>(M);
    c->d_mcoef = dalloc<double>(2 * (size_t)M);
    c->d_mnorm = dalloc<double>(2 * (size_t)M);
    c->d_mstage = dalloc<double>(6 * (size_t)c->nR);
}

// front views: every model's fronts in its slot cur[m] (which = 0) or the other one (which = 1)
void upload_views(tf_ctx *c, const std::vector<int> &cur) {
    const Symbolic &s = c->sym;
    std::vector<FrontDesc> a(c->fd_host), b(c->fd_host);
    for (int k = 0; k < s.nsn; k++) {
        const int sa = cur[c->front_model[k]], sb = 1 - sa;
        a[k].loff += (int64_t)sa * s.ltotal;
        a[k].doff += sa * s.n;
        b[k].loff += (int64_t)sb * s.ltotal;
        b[k].doff += sb * s.n;
    }
    HIPCHK(hipMemcpyAsync(c->d_fd_cur, a.data(), sizeof(FrontDesc) * a.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_fd_oth, b.data(), sizeof(FrontDesc) * b.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
}

// factorise every model at its own dt into the slot `target[m]`
void factor_models(tf_ctx *c, const std::vector<double> &dt, const std::vector<int> &target, const FrontDesc *fd) {
    const int M = c->nmodels;
    std::vector<double> cdt(M);
    std::vector<int64_t> off(M);
    for (int m = 0; m < M; m++) { cdt[m] = 2.0 * dt[m]; off[m] = (int64_t)target[m] * c->sym.ltotal; }
    HIPCHK(hipMemcpyAsync(c->d_mcdt, cdt.data(), sizeof(double) * M, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_slot_off, off.data(), sizeof(int64_t) * M, hipMemcpyHostToDevice, c->st));
    for (int m = 0; m < M; m++)  // clear the target slot of every model (not its current factor)
        tf::launch_fill0(c->d_Lmm + (int64_t)target[m] * c->sym.ltotal + c->m_llo[m], c->m_lhi[m] - c->m_llo[m],
                         c->st);
    tf::launch_fill0(c->d_U, c->sym.utotal, c->st);
    tf::launch_flags_init(c->d_flag, c->st);
    tf::launch_scatter_p_models(c->topo.nnz, c->d_amap, c->d_qmodel, c->d_M, c->d_K, c->d_mcdt, c->d_slot_off,
                                c->d_Lmm, c->st);
    run_factor(c, c->d_Lmm, c->d_Dmm, fd);
    check_flag(c);
    c->models_factored = true;
    c->timers[7] += 1.0;
}

// right-hand side for every model: Msigma (4 e2 - e3) minus each model's
// source at its own time and step (sys_right, src/FETD_FWD.f90:1466-1511)
void rhs_models(tf_ctx *c, int step, const std::vector<double> &t, const std::vector<double> &dt, const double *e2,
                const double *e3, double *r) {
    const int M = c->nmodels;
    tf::launch_spmv_rhs(c->topo.nE, c->rm.fia, c->rm.fja, c->rm.fval, e2, e3, r, c->st);
    if (step < c->sp_division) {
        std::vector<double> coef(2 * (size_t)M);
        for (int m = 0; m < M; m++) {
            coef[2 * m] = source_didt(c->stype, c->ptype, t[m], c->pulse_width, c->current);
            coef[2 * m + 1] = 2.0 * dt[m];
        }
        HIPCHK(hipMemcpyAsync(c->d_mcoef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));  // coef lives on the host stack
        tf::launch_source_models(c->nseg, c->rm.src_edge, c->d_src_id, c->d_src_w, c->d_mcoef, r, c->st);
    }
}

// The BDF2 loop with the doubling controller of adap_euler
// (src/FETD_FWD.f90:1229-1464) run for every model at once: one block-
// diagonal solve per step, each model with its own time, dt, keep/change
// decisions and factor slot.  Regular steps with the sources off replay a
// captured hipGraph (SpMV, level-scheduled solve, readout).
void run_models(tf_ctx *c, int max_steps) {
    const int M = c->nmodels, n = c->topo.nE, R = c->nR;
    StepVecs v;
    ensure_vecs(c, v);
    const int dbl = c->sp_dblesize, ring = dbl + 1;
    if (c->sp_double < dbl + 2) throw std::runtime_error("sp_double must be at least sp_dblesize+2");
    std::vector<double> tcur(M, 0.0), dt(M, c->pulse_width / c->sp_division);
    std::vector<int> cur(M, 0), not_changed(M, 0);
    std::vector<char> on(M, 1);
    for (double *p : {v.e1, v.e2, v.e3}) HIPCHK(hipMemsetAsync(p, 0, sizeof(double) * n, c->st));
    upload_views(c, cur);
    c->mm_step.clear(); c->mm_t.clear(); c->mm_dt.clear(); c->mm_on.clear();
    c->nrows = 0;
    c->mm_graph_steps = 0.0;
    std::map<std::array<const void *, 4>, hipGraphExec_t> graphs;
    auto record = [&](int step, const std::vector<double> &tt, const double *e) {
        // one row per step: every receiver, and per model its time, dt and whether it is still running
        output_row(c, step, 0.0, 0.0, e);
        c->mm_step.push_back(step);
        c->mm_t.insert(c->mm_t.end(), tt.begin(), tt.end());
        c->mm_dt.insert(c->mm_dt.end(), dt.begin(), dt.end());
        c->mm_on.insert(c->mm_on.end(), on.begin(), on.end());
    };
    hipEvent_t f0, f1;
    HIPCHK(hipEventCreate(&f0));
    HIPCHK(hipEventCreate(&f1));
    auto timed = [&](auto &&fn) {
        HIPCHK(hipEventRecord(f0, c->st));
        fn();
        HIPCHK(hipEventRecord(f1, c->st));
        HIPCHK(hipEventSynchronize(f1));
        c->timers[1] += elapsed(f0, f1);
    };
    int step = 0, idx = 0;
    std::vector<double> t0v(M, 0.0);
    while (true) {
        bool any = false;
        for (int m = 0; m < M; m++) {
            on[m] = on[m] && tcur[m] < c->time_max;
            any = any || on[m];
        }
        if (!any || (max_steps > 0 && step >= max_steps)) break;
        step++;
        for (int m = 0; m < M; m++) tcur[m] = tcur[m] + dt[m];
        if (step == 1) {
            record(0, t0v, v.e2);
            timed([&] { factor_models(c, dt, cur, c->d_fd_cur); });
        }
        if (step % c->sp_double == 0) {
            dcopy(c, v.te2, v.e2);
            dcopy(c, v.te3, v.e3);
            std::vector<double> tt(tcur);
            for (int i = 1; i <= dbl; i++) {
                rhs_models(c, step, tt, dt, v.te2, v.te3, v.rh);
                solve_dev(c, c->d_Lmm, c->d_Dmm, c->d_fd_cur, v.rh, v.te1);
                if (i == dbl) break;
                if (i == 1) dcopy(c, v.erec, v.te1);
                std::swap(v.te3, v.te2);
                dcopy(c, v.te2, v.te1);
                for (int m = 0; m < M; m++) tt[m] = tt[m] + dt[m];
            }
            dcopy(c, v.ev1, v.te1);
            std::vector<double> tstep(M), ttime(M);
            std::vector<int> other(M);
            for (int m = 0; m < M; m++) {
                tstep[m] = dt[m] * dbl;
                ttime[m] = tcur[m] - dt[m] + tstep[m];
                other[m] = 1 - cur[m];
            }
            timed([&] { factor_models(c, tstep, other, c->d_fd_oth); });
            double *ea = v.ring[(idx - 1) % ring];
            double *eb = v.ring[(idx - 1 - dbl) % ring];
            rhs_models(c, step, ttime, tstep, ea, eb, v.rh);
            solve_dev(c, c->d_Lmm, c->d_Dmm, c->d_fd_oth, v.rh, v.ev2);
            tf::launch_range_norms(M, c->rm.mrange, v.ev2, v.ev1, c->d_mnorm, c->st);
            std::vector<double> nrm(2 * (size_t)M);
            HIPCHK(hipMemcpyAsync(nrm.data(), c->d_mnorm, sizeof(double) * nrm.size(), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            // e3 <- e2 for every model; e2 <- first fine step (keep) or the
            // two fine steps' end state (change), model by model
            std::swap(v.e3, v.e2);
            dcopy(c, v.e2, v.erec);
            bool flip = false;
            for (int m = 0; m < M; m++) {
                const double err = std::sqrt(nrm[2 * m]) / std::sqrt(nrm[2 * m + 1]);
                const bool keep = !on[m] || (err > 1.0e-4 && not_changed[m] < 2);
                if (keep) {
                    not_changed[m]++;
                } else {
                    HIPCHK(hipMemcpyAsync(v.e2 + c->rm.e0[m], v.ev1 + c->rm.e0[m], sizeof(double) * c->m_ne[m],
                                          hipMemcpyDeviceToDevice, c->st));
                    cur[m] = other[m];
                    tcur[m] = ttime[m];
                    dt[m] = tstep[m];
                    not_changed[m] = 0;
                    flip = true;
                }
            }
            if (flip) upload_views(c, cur);
            record(step, tcur, v.e2);
            idx = 0;
            continue;
        }
        if (step >= c->sp_division && c->use_graphs) {
            // regular step, sources off: replay the captured step for this
            // arrangement of the history vectors
            const std::array<const void *, 4> key = {v.e1, v.e2, v.e3, v.ring[idx % ring]};
            auto it = graphs.find(key);
            if (it == graphs.end()) {
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
                tf::launch_spmv_rhs(n, c->rm.fia, c->rm.fja, c->rm.fval, v.e2, v.e3, v.rh, c->st);
                solve_dev(c, c->d_Lmm, c->d_Dmm, c->d_fd_cur, v.rh, v.e1);
                tf::launch_readout(R, c->rm.rec_edge, c->d_rec_drc, c->d_narr, c->d_darr, v.e1, c->d_mstage, c->st);
                HIPCHK(hipMemcpyAsync(v.ring[idx % ring], v.e1, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
                HIPCHK(hipStreamEndCapture(c->st, &g));
                hipGraphExec_t ex;
                HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
                HIPCHK(hipGraphDestroy(g));
                it = graphs.emplace(key, ex).first;
            }
            HIPCHK(hipGraphLaunch(it->second, c->st));
            // the row from the staging buffer
            output_row_from(c, c->d_mstage);
            c->mm_step.push_back(step);
            c->mm_t.insert(c->mm_t.end(), tcur.begin(), tcur.end());
            c->mm_dt.insert(c->mm_dt.end(), dt.begin(), dt.end());
            c->mm_on.insert(c->mm_on.end(), on.begin(), on.end());
            c->mm_graph_steps += 1.0;
        } else {
            rhs_models(c, step, tcur, dt, v.e2, v.e3, v.rh);
            solve_dev(c, c->d_Lmm, c->d_Dmm, c->d_fd_cur, v.rh, v.e1);
            record(step, tcur, v.e1);
            dcopy(c, v.ring[idx % ring], v.e1);
        }
        idx++;
        double *t3 = v.e3;
        v.e3 = v.e2;
        v.e2 = v.e1;
        v.e1 = t3;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipGetLastError());
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    HIPCHK(hipEventDestroy(f0));
    HIPCHK(hipEventDestroy(f1));
    c->vecs[0] = v.e1; c->vecs[1] = v.e2; c->vecs[2] = v.e3; c->vecs[5] = v.te2; c->vecs[6] = v.te3;
}

// ---- adjoint ------------------------------------------------------------------
// Backward-in-time adjoint over the tape of the last tf_run.  Segment by
// segment from the last checkpoint to the first: recompute the segment's
// forward states from its checkpoint (same factors, same right-hand sides),
// then walk its nodes backwards solving the adjoint systems with the same
// symmetric factors and reducing the per-element sigma sensitivities.
// w: [row][src][rec][9] weights of J = sum w * fields.
void run_adjoint(tf_ctx *c, const double *w, double *grad_host) {
    const int n = c->topo.nE, K = c->nsrc, R = c->nR, T = c->topo.nT;
    const int64_t len = (int64_t)n * K;
    const int nrows = c->nrows;
    // effective per-row weights on E (n_arr) and dB/dt (d_arr): B_r is the
    // running sum of dB/dt_i dt_i, so its weights fold into the dB/dt weights
    // of every earlier row; dB/dt is zero at step 0 and row 0 is constant
    std::vector<double> hw((size_t)nrows * K * R * 6, 0.0);
    {
        std::vector<double> bsum((size_t)K * R * 3, 0.0);
        for (int r = nrows - 1; r >= 0; r--)
            for (int q = 0; q < K; q++)
                for (int i = 0; i < R; i++) {
                    const double *wi = w + (((int64_t)r * K + q) * R + i) * 9;
                    double *o = &hw[(((size_t)r * K + q) * R + i) * 6];
                    for (int k = 0; k < 3; k++) {
                        bsum[((size_t)q * R + i) * 3 + k] += wi[6 + k];
                        o[k] = wi[k];
                        o[3 + k] = (c->row_step[r] == 0) ? 0.0 : wi[3 + k] + c->row_dt[r] * bsum[((size_t)q * R + i) * 3 + k];
                    }
                }
    }
    double *d_w = dupload(hw);
    if (!c->d_tet_edg) c->d_tet_edg = dupload(c->topo.tet_edg);
    if (!c->d_vm1) {  // unit-conductivity element mass entries
        std::vector<double> ones(T, 1.0);
        double *d_ones = dupload(ones), *d_vk = dalloc<double>((size_t)T * 21);
        c->d_vm1 = dalloc<double>((size_t)T * 21);
        tf::launch_element(T, c->d_xyz, c->d_tet_vex, c->d_tet_drc, c->d_orient, d_ones, c->d_invmu, c->d_vm1, d_vk,
                           c->st);
        HIPCHK(hipStreamSynchronize(c->st));
        dfree(d_ones);
        dfree(d_vk);
    }
    // live nodes: those a row was written from, and their history
    const int nn = (int)c->tape.size();
    std::vector<char> live(nn, 0);
    for (int r = 1; r < nrows; r++) live[c->row_node[r]] = 1;
    for (int k = nn - 1; k > 0; k--)
        if (live[k]) {
            if (c->tape[k].u > 0) live[c->tape[k].u] = 1;
            if (c->tape[k].v > 0) live[c->tape[k].v] = 1;
        }
    // vector pool
    std::vector<double *> pool, owned;
    auto get = [&]() {
        double *p;
        if (pool.empty()) { p = dalloc<double>(len); owned.push_back(p); }
        else { p = pool.back(); pool.pop_back(); }
        return p;
    };
    auto zeroed = [&]() {
        double *p = get();
        tf::launch_fill0(p, len, c->st);
        return p;
    };
    double *zero = zeroed(), *rh = get(), *lam = get(), *grad = dalloc<double>(T);
    tf::launch_fill0(grad, T, c->st);
    std::vector<double *> state(nn, nullptr);
    std::vector<double *> accA(nn, nullptr), accB(nn, nullptr);
    state[0] = zero;
    // The sweep's factor slots: the dt of every live node in processing
    // order (segments from the last; recompute forward, adjoint backward) is
    // known from the tape, so the next dt to come is factored on the second
    // stream while the current one is in use, as in tf_run.
    auto stored = [&](int k) { return k < (int)c->tstate.size() && c->tstate[k] != nullptr; };
    std::vector<double> dts;
    for (int g = (int)c->ckpts.size() - 1; g >= 0; g--) {
        const int first = c->ckpts[g].first, end = (g + 1 < (int)c->ckpts.size()) ? c->ckpts[g + 1].first : nn;
        for (int k = first; k < end; k++)
            if (live[k] && !stored(k)) dts.push_back(c->tape[k].dt);
        for (int k = end - 1; k >= first; k--)
            if (live[k]) dts.push_back(c->tape[k].dt);
    }
    size_t pos = 0;  // next entry of dts
    int last_used = 0, pend = -1;
    hipEvent_t a0, a2;
    HIPCHK(hipEventCreate(&a0));
    HIPCHK(hipEventCreate(&a2));
    struct StreamDrain {
        hipStream_t s;
        ~StreamDrain() { (void)hipStreamSynchronize(s); }
    } drain{c->st2};
    auto kept_index = [&](double dt) {
        for (size_t q = 0; q < c->kept.size(); q++)
            if (c->kept[q].dt == dt) return (int)q;
        return -1;
    };
    auto slot_for = [&](double dt) {
        if (pos < dts.size() && dts[pos] == dt) pos++;
        int s = -1;
        for (int q = 0; q < 2; q++)
            if (c->slot_live[q] && c->slot_dt[q] == dt) s = q;
        if (s >= 0 && s == pend) {  // factored ahead: wait for it
            HIPCHK(hipStreamWaitEvent(c->st, a2, 0));
            check_flag(c);
            pend = -1;
        }
        const int ki = kept_index(dt);
        if (s < 0 && ki >= 0) {  // kept by the taped run: swapped into a slot, the slot's factor kept instead
            if (pend >= 0) {
                HIPCHK(hipStreamSynchronize(c->st2));
                check_flag(c);
                pend = -1;
            }
            s = 1 - last_used;
            tf_ctx::KeptFactor &k = c->kept[ki];
            std::swap(k.L, c->d_L[s]);
            std::swap(k.D, c->d_D[s]);
            std::swap(k.LT, c->d_LT[s]);
            if (c->slot_live[s] && k.L && k.D && kept_index(c->slot_dt[s]) < 0) {
                k.dt = c->slot_dt[s];
            } else {
                dfree(k.L); dfree(k.D); dfree(k.LT);
                c->kept.erase(c->kept.begin() + ki);
            }
            c->slot_live[s] = true;
            c->slot_dt[s] = dt;
        }
        if (s < 0) {
            if (pend >= 0) {  // not the dt it was factored for
                HIPCHK(hipStreamSynchronize(c->st2));
                check_flag(c);
                pend = -1;
            }
            s = 1 - last_used;
            factor_dt(c, s, dt);
            check_flag(c);
            c->timers[7] += 1.0;
        }
        last_used = s;
        // the next different dt, into the other slot, when nothing before it needs that slot
        if (c->overlap_factor && pend < 0) {
            size_t j = pos;
            while (j < dts.size() && dts[j] == dt) j++;
            const int o = 1 - s;
            if (j < dts.size() && !(c->slot_live[o] && c->slot_dt[o] == dts[j]) && kept_index(dts[j]) < 0) {
                HIPCHK(hipEventRecord(a0, c->st));
                HIPCHK(hipStreamWaitEvent(c->st2, a0, 0));
                std::swap(c->st, c->st2);
                factor_dt(c, o, dts[j]);
                HIPCHK(hipEventRecord(a2, c->st));
                std::swap(c->st, c->st2);
                c->timers[7] += 1.0;
                pend = o;
            }
        }
        return s;
    };
    auto S = [&](int k) -> double * {
        if (k == 0) return zero;
        if (state[k]) return state[k];
        if (stored(k)) return c->tstate[k];
        throw std::runtime_error("adjoint: history state missing from the tape");
    };
    const int nseg = (int)c->ckpts.size();
    for (int g = nseg - 1; g >= 0; g--) {
        const tf_ctx::Ckpt &ck = c->ckpts[g];
        const int first = ck.first, end = (g + 1 < nseg) ? c->ckpts[g + 1].first : nn;
        if (ck.id2 > 0) state[ck.id2] = ck.s2;
        if (ck.id3 > 0) state[ck.id3] = ck.s3;
        // recompute the segment's live states that the run did not keep
        for (int k = first; k < end; k++) {
            if (!live[k] || stored(k)) continue;
            const tf_ctx::TNode &nd = c->tape[k];
            const int slot = slot_for(nd.dt);
            rhs(c, nd.step, nd.t, S(nd.u), S(nd.v), nd.dt, false, rh);
            state[k] = get();
            step_solve(c, slot, rh, state[k]);
        }
        // adjoint sweep
        for (int k = end - 1; k >= first; k--) {
            if (!live[k]) continue;
            const tf_ctx::TNode &nd = c->tape[k];
            const int slot = slot_for(nd.dt);
            if (accA[k] || accB[k]) {
                const double *a = accA[k] ? accA[k] : zero, *b = accB[k] ? accB[k] : zero;
                if (K == 1) tf::launch_spmv_rhs(n, c->rm.fia, c->rm.fja, c->rm.fval, a, b, rh, c->st);
                else tf::launch_mspmv_rhs(n, K, c->rm.fia, c->rm.fja, c->rm.fval, a, b, c->d_mb, rh, c->spmv_xcd, c->st);
            } else {
                tf::launch_fill0(rh, len, c->st);
            }
            if (nd.row > 0)
                tf::launch_adj_seed(R, K, c->rm.rec_edge, c->d_rec_drc, c->d_narr, c->d_darr,
                                    d_w + (int64_t)nd.row * K * R * 6, rh, c->st);
            step_solve(c, slot, rh, lam);
            tf::launch_sens(T, K, c->rm.tet_edg, c->d_orient, c->d_vm1, lam, S(k), S(nd.u), S(nd.v), grad, c->st);
            if (nd.u > 0) {
                if (!accA[nd.u]) accA[nd.u] = zeroed();
                tf::launch_axpy(len, 1.0, lam, accA[nd.u], c->st);
            }
            if (nd.v > 0) {
                if (!accB[nd.v]) accB[nd.v] = zeroed();
                tf::launch_axpy(len, 1.0, lam, accB[nd.v], c->st);
            }
            if (accA[k]) { pool.push_back(accA[k]); accA[k] = nullptr; }
            if (accB[k]) { pool.push_back(accB[k]); accB[k] = nullptr; }
        }
        // the segment's states are no longer needed (earlier segments
        // recompute their own; checkpoint buffers are not pooled)
        for (int k = first; k < end; k++)
            if (state[k]) { pool.push_back(state[k]); state[k] = nullptr; }
        if (ck.id2 > 0) state[ck.id2] = nullptr;
        if (ck.id3 > 0) state[ck.id3] = nullptr;
    }
    tf::launch_scale(T, c->d_sigma, grad, c->st);
    HIPCHK(hipMemcpyAsync(grad_host, grad, sizeof(double) * T, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipStreamSynchronize(c->st2));
    if (pend >= 0) check_flag(c);  // factored ahead for nothing (the sweep ended)
    HIPCHK(hipGetLastError());
    (void)hipEventDestroy(a0);
    (void)hipEventDestroy(a2);
    for (double *p : owned) dfree(p);
    dfree(grad);
    dfree(d_w);
}

// message of the last tf_create failure (no context exists to hold it)
static thread_local std::string g_create_err = "no context";

int tf_create(tf_ctx **out, int device) {
    if (!out) return TF_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_err = "no HIP device available (libtemfetd runs on the GPU only)";
        return TF_ERR_HIP;
    }
    if (device < 0 || device >= ndev) {
        g_create_err = "device index out of range";
        return TF_ERR_ARG;
    }
    tf_ctx *c = new tf_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess ||
        hipStreamCreate(&c->st2) != hipSuccess) {
        if (c->st) (void)hipStreamDestroy(c->st);
        delete c;
        g_create_err = "could not create a HIP stream on the device";
        return TF_ERR_HIP;
    }
    // the allocation-free solve's scratch array on this device (resolved once per device)
    if (tf::noalloc_scratch(device, &c->d_scratch) != hipSuccess) {
        (void)hipStreamDestroy(c->st);
        (void)hipStreamDestroy(c->st2);
        delete c;
        g_create_err = "could not resolve the device address of the solve scratch array";
        return TF_ERR_HIP;
    }
    *out = c;
    return TF_OK;
}

void tf_destroy(tf_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    (void)hipStreamSynchronize(c->st2);
    free_all(c);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->st);
    (void)hipStreamDestroy(c->st2);
    delete c;
}

const char *tf_last_error(const tf_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int tf_set_mesh(tf_ctx *c, int nN, const double *nod_cor, int nT, const int *tet_vex) {
    return guarded(c, [&] {
        c->tape_valid = false;
        if (nN <= 0 || nT <= 0 || !nod_cor || !tet_vex) return fail(c, TF_ERR_ARG, "empty mesh");
        free_all(c);
        c->assembled = c->have_mat = c->have_src = false;
        c->nR = 0;
        tf::build_topology(c->topo, nN, nod_cor, nT, tet_vex);
        const Topology &t = c->topo;
        c->d_xyz = dupload(t.xyz);
        c->d_tet_vex = dupload(t.tet_vex);
        c->d_tet_drc = dupload(t.tet_drc);
        c->d_orient = dupload(t.pair_orient);
        c->d_cptr = dupload(t.cptr);
        std::vector<int> ctr32(t.ctr.begin(), t.ctr.end());
        c->d_ctr = dupload(ctr32);
        // full symmetric pattern of the mass matrix for the right-hand side
        int n = t.nE;
        std::vector<int> deg(n + 1, 0);
        for (int r = 0; r < n; r++)
            for (int q = t.ia[r]; q < t.ia[r + 1]; q++) {
                deg[r]++;
                if (t.ja[q] != r) deg[t.ja[q]]++;
            }
        std::vector<int> fia(n + 1, 0);
        for (int r = 0; r < n; r++) fia[r + 1] = fia[r] + deg[r];
        std::vector<int> fja(fia[n]);
        std::vector<int64_t> fmap(fia[n]);
        std::fill(deg.begin(), deg.end(), 0);
        for (int r = 0; r < n; r++)
            for (int q = t.ia[r]; q < t.ia[r + 1]; q++) {
                int cc = t.ja[q];
                int p = fia[r] + deg[r]++;
                fja[p] = cc;
                fmap[p] = q;
                if (cc != r) {
                    p = fia[cc] + deg[cc]++;
                    fja[p] = r;
                    fmap[p] = q;
                }
            }
        c->d_fia = dupload(fia);
        c->d_fja = dupload(fja);
        c->d_fmap = dupload(fmap);
        c->d_fval = dalloc<double>(fja.size());
        c->d_M = dalloc<double>(t.nnz);
        c->d_K = dalloc<double>(t.nnz);
        c->d_vm = dalloc<double>((size_t)t.nT * 21);
        c->d_vk = dalloc<double>((size_t)t.nT * 21);
        c->have_mesh = true;
        return TF_OK;
    });
}

int tf_get_sizes(const tf_ctx *c, int *nE, int *nnzU) {
    if (!c || !c->have_mesh) return TF_ERR_STATE;
    if (nE) *nE = c->topo.nE;
    if (nnzU) *nnzU = (int)c->topo.nnz;
    return TF_OK;
}

int tf_get_topology(const tf_ctx *c, int *edg_vex, int *tet_edg, int *tet_drc, int *ia, int *ja) {
    if (!c || !c->have_mesh) return TF_ERR_STATE;
    const Topology &t = c->topo;
    if (edg_vex) std::copy(t.edg_vex.begin(), t.edg_vex.end(), edg_vex);
    if (tet_edg)
        for (size_t i = 0; i < t.tet_edg.size(); i++) tet_edg[i] = t.tet_edg[i] + 1;
    if (tet_drc) std::copy(t.tet_drc.begin(), t.tet_drc.end(), tet_drc);
    if (ia)
        for (size_t i = 0; i < t.ia.size(); i++) ia[i] = t.ia[i] + 1;
    if (ja)
        for (size_t i = 0; i < t.ja.size(); i++) ja[i] = t.ja[i] + 1;
    return TF_OK;
}

int tf_set_materials(tf_ctx *c, const double *sigma, const double *inv_mu) {
    return guarded(c, [&] {
        c->tape_valid = false;
        if (!c->have_mesh) return fail(c, TF_ERR_STATE, "tf_set_mesh first");
        if (!sigma || !inv_mu) return fail(c, TF_ERR_ARG, "null material array");
        c->sigma.assign(sigma, sigma + c->topo.nT);
        c->inv_mu.assign(inv_mu, inv_mu + c->topo.nT);
        dfree(c->d_sigma);
        dfree(c->d_invmu);
        c->d_sigma = dupload(c->sigma);
        c->d_invmu = dupload(c->inv_mu);
        c->have_mat = true;
        c->assembled = false;
        c->aniso = false;
        return TF_OK;
    });
}

int tf_set_materials_tensor(tf_ctx *c, const double *sigma9, const double *nu9) {
    return guarded(c, [&] {
        c->tape_valid = false;
        if (!c->have_mesh) return fail(c, TF_ERR_STATE, "tf_set_mesh first");
        if (!sigma9 || !nu9) return fail(c, TF_ERR_ARG, "null material array");
        const int T = c->topo.nT;
        std::vector<double> s9(sigma9, sigma9 + 9 * (size_t)T), n9(nu9, nu9 + 9 * (size_t)T);
        // scalar stand-ins (first diagonal entries) for what reads sigma per tet
        c->sigma.resize(T);
        c->inv_mu.resize(T);
        for (int i = 0; i < T; i++) { c->sigma[i] = s9[9 * (size_t)i]; c->inv_mu[i] = n9[9 * (size_t)i]; }
        dfree(c->d_sig9); dfree(c->d_nu9); dfree(c->d_sigma); dfree(c->d_invmu);
        c->d_sig9 = dupload(s9);
        c->d_nu9 = dupload(n9);
        c->d_sigma = dupload(c->sigma);
        c->d_invmu = dupload(c->inv_mu);
        c->have_mat = true;
        c->assembled = false;
        c->aniso = true;
        return TF_OK;
    });
}

int tf_set_num_sources(tf_ctx *c, int nsrc) {
    if (!c) return TF_ERR_ARG;
    c->tape_valid = false;
    if (nsrc < 1 || nsrc > 64) return fail(c, TF_ERR_ARG, "number of sources must be 1..64");
    c->nsrc = nsrc;
    c->src_edges.assign(nsrc, {});
    c->src_ws.assign(nsrc, {});
    c->have_src = false;
    return TF_OK;
}

int tf_set_source(tf_ctx *c, int isrc, int nseg, const int *seg_vex, int source_type, int pulse_type,
                  double pulse_width, double current) {
    return guarded(c, [&] {
        c->tape_valid = false;
        if (!c->have_mesh) return fail(c, TF_ERR_STATE, "tf_set_mesh first");
        const int nslots = c->nmodels > 0 ? c->nmodels : c->nsrc;  // one wire per model, or per source
        if ((int)c->src_edges.size() != nslots) {
            c->src_edges.assign(nslots, {});
            c->src_ws.assign(nslots, {});
        }
        if (isrc < 0 || isrc >= nslots) return fail(c, TF_ERR_ARG, "source index out of range");
        if (nseg <= 0 || !seg_vex) return fail(c, TF_ERR_ARG, "no source segments");
        if (source_type != 1 || pulse_type < 1 || pulse_type > 3)
            return fail(c, TF_ERR_INPUT, "no such kind of source");
        bool any = false;
        for (auto &e : c->src_edges) any |= !e.empty();
        if (any && (source_type != c->stype || pulse_type != c->ptype || pulse_width != c->pulse_width ||
                    current != c->current))
            return fail(c, TF_ERR_INPUT, "all sources of one run share the waveform");
        std::vector<int> edges(nseg);
        std::vector<double> ws(nseg);
        for (int i = 0; i < nseg; i++) {
            int e = tf::find_edge(c->topo, seg_vex[2 * i], seg_vex[2 * i + 1]);
            if (e < 0) return fail(c, TF_ERR_INPUT, "source segment " + std::to_string(i + 1) + " is not a mesh edge");
            edges[i] = e;
            const double drc = (i == nseg - 1) ? -1.0 : 1.0;  // src/FETD_FWD.f90:446-448
            ws[i] = drc * c->topo.edg_len[e];
        }
        c->src_edges[isrc] = edges;
        c->src_ws[isrc] = ws;
        c->stype = source_type;
        c->ptype = pulse_type;
        c->pulse_width = pulse_width;
        c->current = current;
        bool all = true;
        for (auto &e : c->src_edges) all &= !e.empty();
        c->have_src = all;
        if (all) {
            c->src_edge.clear(); c->src_id.clear(); c->src_w.clear();
            for (int q = 0; q < (int)c->src_edges.size(); q++)
                for (size_t i = 0; i < c->src_edges[q].size(); i++) {
                    c->src_edge.push_back(c->src_edges[q][i]);
                    c->src_id.push_back(q);
                    c->src_w.push_back(c->src_ws[q][i]);
                }
            c->nseg = (int)c->src_edge.size();
            dfree(c->d_src_edge); dfree(c->d_src_id); dfree(c->d_src_w);
            c->d_src_edge = dupload(c->src_edge);
            c->d_src_id = dupload(c->src_id);
            c->d_src_w = dupload(c->src_w);
        }
        free_native(c);  // the native-order maps follow the new sources, receivers or models
        return TF_OK;
    });
}

int tf_set_slices(tf_ctx *c, int nslice, const double *times) {
    return guarded(c, [&] {
        if (nslice < 0 || (nslice > 0 && !times)) return fail(c, TF_ERR_ARG, "bad slice times");
        for (int i = 0; i < nslice; i++)
            if (!std::isfinite(times[i])) return fail(c, TF_ERR_ARG, "slice times must be finite");
        free_slices(c);
        c->slice_t.assign(times, times + nslice);
        c->slice_tt.assign(nslice, 0.0);
        c->slice_done.assign(nslice, 0);
        c->slice_step.assign(nslice, 0);
        return TF_OK;
    });
}

int tf_get_slice(tf_ctx *c, int i, double *fields9, int *step, double *t) {
    return guarded(c, [&] {
        if (i < 0 || i >= (int)c->slice_t.size()) return fail(c, TF_ERR_ARG, "no such slice");
        if (!c->slice_done[i]) return fail(c, TF_ERR_STATE, "the run did not reach this slice's time");
        if (fields9)
            HIPCHK(hipMemcpy(fields9, c->d_slice + (int64_t)i * c->slice_nT * 9,
                             sizeof(double) * 9 * c->slice_nT, hipMemcpyDeviceToHost));
        if (step) *step = c->slice_step[i];
        if (t) *t = c->slice_tt[i];
        return TF_OK;
    });
}

int tf_set_replay(tf_ctx *c, int nchecks, const int *keep) {
    if (!c) return TF_ERR_ARG;
    if (nchecks < 0 || (nchecks > 0 && !keep)) return fail(c, TF_ERR_ARG, "bad replay schedule");
    c->replay.assign(keep, keep + nchecks);
    return TF_OK;
}

int tf_set_receivers(tf_ctx *c, int nR, const double *xyz, double rec_charac, int *rec_tet_out) {
    return guarded(c, [&] {
        c->tape_valid = false;
        if (!c->have_mesh) return fail(c, TF_ERR_STATE, "tf_set_mesh first");
        if (nR <= 0 || !xyz) return fail(c, TF_ERR_ARG, "no receivers");
        c->nR = nR;
        c->rec_xyz.assign(xyz, xyz + 3 * (size_t)nR);
        c->rec_charac = rec_charac;
        c->rec_tet.assign(nR, 0);
        if (c->nmodels > 0 && c->m_rec0.back() + c->m_nrec.back() != nR)
            return fail(c, TF_ERR_ARG, "receiver count differs from the models' total");
        for (int i = 0; i < nR; i++) {
            int t0 = 0, t1 = -1;
            if (c->nmodels > 0) {  // a receiver lies in its own model's tets
                const int m = (int)(std::upper_bound(c->m_rec0.begin(), c->m_rec0.end(), i) - c->m_rec0.begin()) - 1;
                t0 = c->m_tet0[m];
                t1 = c->m_tet0[m] + c->m_ntet[m];
            }
            c->rec_tet[i] = tf::point_in_tet(c->topo, &xyz[3 * i], rec_charac, t0, t1);
            if (c->rec_tet[i] == 0) return fail(c, TF_ERR_INPUT, "no rec tet No. " + std::to_string(i + 1));
        }
        if (rec_tet_out) std::copy(c->rec_tet.begin(), c->rec_tet.end(), rec_tet_out);
        std::vector<int> redge(6 * (size_t)nR);
        std::vector<double> rdrc(6 * (size_t)nR);
        for (int i = 0; i < nR; i++)
            for (int j = 0; j < 6; j++) {
                redge[6 * i + j] = c->topo.tet_edg[6 * (size_t)(c->rec_tet[i] - 1) + j];
                rdrc[6 * i + j] = (double)c->topo.tet_drc[6 * (size_t)(c->rec_tet[i] - 1) + j];
            }
        dfree(c->d_rec_xyz); dfree(c->d_rec_tet); dfree(c->d_rec_edge); dfree(c->d_rec_drc);
        dfree(c->d_narr); dfree(c->d_darr);
        c->d_rec_xyz = dupload(c->rec_xyz);
        c->d_rec_tet = dupload(c->rec_tet);
        c->d_rec_edge = dupload(redge);
        c->d_rec_drc = dupload(rdrc);
        c->d_narr = dalloc<double>(18 * (size_t)nR);
        c->d_darr = dalloc<double>(18 * (size_t)nR);
        tf::launch_recvec(nR, c->d_xyz, c->d_tet_vex, c->d_rec_tet, c->d_rec_xyz, c->d_narr, c->d_darr, c->st);
        HIPCHK(hipStreamSynchronize(c->st));
        free_native(c);  // the native-order maps follow the new sources, receivers or models
        return TF_OK;
    });
}

int tf_set_schedule(tf_ctx *c, int sp_division, int sp_double, int sp_dblesize, double time_maximum) {
    if (!c) return TF_ERR_ARG;
    c->tape_valid = false;
    if (sp_division <= 0) return fail(c, TF_ERR_INPUT, "cmod%step_dvsi error");
    if (sp_double <= 0) return fail(c, TF_ERR_INPUT, "cmod%step_dble error");
    if (std::fabs(time_maximum) < 1e-20) return fail(c, TF_ERR_INPUT, "cmod%time_maxm error");
    if (sp_dblesize < 2) return fail(c, TF_ERR_INPUT, "sp_dblesize must be at least 2");
    c->sp_division = sp_division;
    c->sp_double = sp_double;
    c->sp_dblesize = sp_dblesize;
    c->time_max = time_maximum;
    return TF_OK;
}

int tf_assemble(tf_ctx *c) {
    return guarded(c, [&] {
        c->tape_valid = false;
        if (!c->have_mesh || !c->have_mat) return fail(c, TF_ERR_STATE, "mesh and materials first");
        auto t0 = std::chrono::steady_clock::now();
        const Topology &t = c->topo;
        if (c->aniso)
            tf::launch_element_ani(t.nT, c->d_xyz, c->d_tet_vex, c->d_tet_drc, c->d_orient, c->d_sig9, c->d_nu9,
                                   c->d_vm, c->d_vk, c->st);
        else
            tf::launch_element(t.nT, c->d_xyz, c->d_tet_vex, c->d_tet_drc, c->d_orient, c->d_sigma, c->d_invmu,
                               c->d_vm, c->d_vk, c->st);
        tf::launch_gather(t.nnz, c->d_cptr, c->d_ctr, c->d_vm, c->d_vk, c->d_M, c->d_K, c->st);
        tf::launch_gather_vals((int64_t)t.ia.size() > 0 ? (int64_t)(2 * t.nnz - t.nE) : 0, c->d_fmap, c->d_M,
                               c->d_fval, c->st);
        if (c->native_ok) tf::launch_gather_vals(c->nfnnz, c->d_nfmap, c->d_M, c->d_nfval, c->st);
        HIPCHK(hipStreamSynchronize(c->st));
        HIPCHK(hipGetLastError());
        c->timers[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        c->assembled = true;
        return TF_OK;
    });
}

int tf_get_matrices(const tf_ctx *c, double *mass_sigma, double *curl) {
    if (!c || !c->assembled) return TF_ERR_STATE;
    if (mass_sigma && hipMemcpy(mass_sigma, c->d_M, sizeof(double) * c->topo.nnz, hipMemcpyDeviceToHost))
        return TF_ERR_HIP;
    if (curl && hipMemcpy(curl, c->d_K, sizeof(double) * c->topo.nnz, hipMemcpyDeviceToHost)) return TF_ERR_HIP;
    return TF_OK;
}

int tf_get_receiver_vectors(const tf_ctx *c, double *n_arr, double *d_arr) {
    if (!c || c->nR == 0) return TF_ERR_STATE;
    if (n_arr && hipMemcpy(n_arr, c->d_narr, sizeof(double) * 18 * c->nR, hipMemcpyDeviceToHost)) return TF_ERR_HIP;
    if (d_arr && hipMemcpy(d_arr, c->d_darr, sizeof(double) * 18 * c->nR, hipMemcpyDeviceToHost)) return TF_ERR_HIP;
    return TF_OK;
}

int tf_analyze(tf_ctx *c, int leaf_size) {
    return guarded(c, [&] {
        if (!c->have_mesh) return fail(c, TF_ERR_STATE, "tf_set_mesh first");
        if (leaf_size > 0) c->leaf_size = leaf_size;
        do_analyze(c, c->leaf_size);
        return TF_OK;
    });
}

int tf_set_tuning(tf_ctx *c, const char *key, int value) {
    return guarded(c, [&] {
        if (!key) return fail(c, TF_ERR_ARG, "null key");
        const std::string k(key);
        bool replan = false;
        if (k == "mt_units") {
            if (value < 64) return fail(c, TF_ERR_ARG, "mt_units must be >= 64");
            c->mt_units = value;
            replan = true;
        } else if (k == "mt_chmin") {
            if (value < 16 || value % 16) return fail(c, TF_ERR_ARG, "mt_chmin must be a positive multiple of 16");
            c->mt_chmin = value;
            replan = true;
        } else if (k == "mt_wide") {
            if (value < 0) return fail(c, TF_ERR_ARG, "mt_wide must be >= 0");
            c->mt_wide = value;
            replan = true;
        } else if (k == "s_lt") {
            if (value != 0 && value != 1) return fail(c, TF_ERR_ARG, "s_lt must be 0 or 1");
            c->s_lt = value != 0;
        } else if (k == "f_band") {
            if (value < 0 || value > 1024) return fail(c, TF_ERR_ARG, "f_band must be in 0..1024");
            c->f_band = value;
        } else if (k == "keep_states") {
            if (value != 0 && value != 1) return fail(c, TF_ERR_ARG, "keep_states must be 0 or 1");
            c->keep_states = value != 0;
        } else if (k == "keep_factors") {
            if (value < 0) return fail(c, TF_ERR_ARG, "keep_factors must be >= 0");
            c->keep_factors = value;
            if ((int)c->kept.size() > value) drop_kept(c);
        } else if (k == "rhs_chain") {
            if (value != 0 && value != 1) return fail(c, TF_ERR_ARG, "rhs_chain must be 0 or 1");
            c->rhs_chain = value != 0;
        } else if (k == "graphs") {
            if (value != 0 && value != 1) return fail(c, TF_ERR_ARG, "graphs must be 0 or 1");
            c->use_graphs = value != 0;
        } else if (k == "overlap") {
            if (value != 0 && value != 1) return fail(c, TF_ERR_ARG, "overlap must be 0 or 1");
            c->overlap_factor = value != 0;
        } else if (k == "s_tiles") {
            if (value < 0 || value > 2) return fail(c, TF_ERR_ARG, "s_tiles must be 0, 1 or 2");
            if (value != 2 && c->pinv_levels > 0)
                return fail(c, TF_ERR_ARG, "s_tiles 0 and 1 read L21 itself: set pinv to 0 first");
            c->s_tiles = value;
        } else if (k == "pinv") {
            if (value < 0) return fail(c, TF_ERR_ARG, "pinv must be >= 0");
            if (value > 0 && c->s_tiles != 2) return fail(c, TF_ERR_ARG, "pinv needs s_tiles 2");
            if (value != c->pinv_levels) {  // the front layout changes: analyse again (drops the factors)
                c->pinv_levels = value;
                if (c->analyzed) {
                    HIPCHK(hipDeviceSynchronize());
                    do_analyze(c, c->leaf_size);
                }
            }
        } else if (k == "native") {
            if (value != 0 && value != 1) return fail(c, TF_ERR_ARG, "native must be 0 or 1");
            if (c->native != (value != 0)) c->tape_valid = false;
            c->native = value != 0;
        } else if (k == "spmv_xcd") {
            if (value != 0 && value != 1) return fail(c, TF_ERR_ARG, "spmv_xcd must be 0 or 1");
            c->spmv_xcd = value;
        } else if (k == "s_tcols") {
            if (value != 16 && value != 32) return fail(c, TF_ERR_ARG, "s_tcols must be 16 or 32");
            c->s_tcols = value;
        } else if (k == "s_units") {
            if (value < 64) return fail(c, TF_ERR_ARG, "s_units must be >= 64");
            c->s_units = value;
            replan = true;
        } else if (k == "s_chmin") {
            if (value < 16 || value % 16) return fail(c, TF_ERR_ARG, "s_chmin must be a positive multiple of 16");
            c->s_chmin = value;
            replan = true;
        } else if (k == "indefinite") {
            if (value != 0 && value != 1) return fail(c, TF_ERR_ARG, "indefinite must be 0 or 1");
            c->allow_indefinite = value != 0;
        } else if (k == "mt_ks_n" || k == "mt_ks_t") {
            if (value != 2 && value != 4) return fail(c, TF_ERR_ARG, "k-steps per stage must be 2 or 4");
            (k == "mt_ks_n" ? c->mt_ks_n : c->mt_ks_t) = value;
        } else {
            return fail(c, TF_ERR_ARG, "unknown tuning key " + k);
        }
        if (replan && c->analyzed) {  // the tile units and partial buffer follow the plan
            HIPCHK(hipStreamSynchronize(c->st));
            HIPCHK(hipStreamSynchronize(c->st2));
            build_plan(c);
            dfree(c->d_plist); dfree(c->d_ppref); dfree(c->d_units); dfree(c->d_part);
            c->d_plist = dupload(c->plan.list);
            c->d_ppref = dupload(c->plan.pref);
            c->d_units = dupload(c->plan.units);
            c->d_part = dalloc<double>(part_doubles(c->plan));
            dfree(c->d_W); dfree(c->d_woff);  // the panel-inverse scratch follows the small/grid split
            c->d_W = dalloc<double>((size_t)std::max<int64_t>(c->plan.wsize, 1));
            c->d_woff = dupload(c->plan.woff);
        }
        return TF_OK;
    });
}

int tf_set_checkpointing(tf_ctx *c, int every) {
    return guarded(c, [&] {
        if (every < 0) return fail(c, TF_ERR_ARG, "checkpoint interval must be >= 0");
        c->ckpt_every = every;
        if (every == 0) free_tape(c);
        return TF_OK;
    });
}

int tf_adjoint(tf_ctx *c, int nrows, const double *w, double *grad_logsigma) {
    return guarded(c, [&] {
        if (c->ptype == 3) return fail(c, TF_ERR_STATE, "the adjoint covers step-on and impulse sources (pulse types 1, 2)");
        if (c->aniso) return fail(c, TF_ERR_STATE, "the adjoint is per-tet scalar log-conductivity (isotropic media)");
        if (!c->tape_valid) return fail(c, TF_ERR_STATE, "no taped run: tf_set_checkpointing(ctx, k > 0) then tf_run");
        if (nrows != c->nrows) return fail(c, TF_ERR_ARG, "weights must cover every row of the taped run");
        if (!w || !grad_logsigma) return fail(c, TF_ERR_ARG, "null buffer");
        auto t0 = std::chrono::steady_clock::now();
        RunScope scope(c);  // the taped vectors are in the order tf_run kept them (the "native" key resets the tape)
        run_adjoint(c, w, grad_logsigma);
        c->timers[5] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return TF_OK;
    });
}

int tf_set_profiling(tf_ctx *c, int on) {
    return guarded(c, [&] {
        prof_fold(c);
        for (int k = 0; k < TF_PROF_NCAT; k++) { c->prof_ms[k] = 0.0; c->prof_cnt[k] = 0; }
        c->prof_on = on != 0;
        return TF_OK;
    });
}

int tf_get_profile(tf_ctx *c, int category, int64_t *launches, double *ms) {
    return guarded(c, [&] {
        if (category < 0 || category >= TF_PROF_NCAT) return fail(c, TF_ERR_ARG, "bad profile category");
        prof_fold(c);
        if (launches) *launches = c->prof_cnt[category];
        if (ms) *ms = c->prof_ms[category];
        return TF_OK;
    });
}

int tf_factor_stats(const tf_ctx *c, int64_t *nnzL, double *flops, int *nsuper, int *nlevels) {
    if (!c || !c->analyzed) return TF_ERR_STATE;
    if (nnzL) *nnzL = c->sym.nnzL;
    if (flops) *flops = c->sym.flops;
    if (nsuper) *nsuper = c->sym.nsn;
    if (nlevels) *nlevels = c->sym.nlevels;
    return TF_OK;
}

int tf_factor_layout(const tf_ctx *c, int64_t *struct_rows, int64_t *l_doubles, int64_t *u_doubles) {
    if (!c || !c->analyzed) return TF_ERR_STATE;
    if (struct_rows) *struct_rows = (int64_t)c->sym.sidx.size();
    if (l_doubles) *l_doubles = c->sym.ltotal;
    if (u_doubles) *u_doubles = c->sym.utotal;
    return TF_OK;
}

int tf_factor_pivots(const tf_ctx *c, int64_t *n_negative, double *min_abs, double *max_abs) {
    if (!c || !c->analyzed) return TF_ERR_STATE;
    if (n_negative) *n_negative = c->last_nneg;
    if (min_abs) *min_abs = c->last_min_abs;
    if (max_abs) *max_abs = c->last_max_abs;
    return TF_OK;
}

int tf_factor(tf_ctx *c, int slot, double dt) {
    return guarded(c, [&] {
        if (slot < 0 || slot > 1) return fail(c, TF_ERR_ARG, "slot must be 0 or 1");
        if (!c->assembled) return fail(c, TF_ERR_STATE, "tf_assemble first");
        if (!c->analyzed) do_analyze(c, c->leaf_size);
        factor_dt(c, slot, dt);
        try {
            check_flag(c);
        } catch (...) {
            c->slot_live[slot] = false;
            throw;
        }
        return TF_OK;
    });
}

int tf_factor_csr(tf_ctx *c, int slot, int n, const int *ia, const int *ja, const double *va) {
    return guarded(c, [&] {
        if (slot < 0 || slot > 1) return fail(c, TF_ERR_ARG, "slot must be 0 or 1");
        if (!c->have_mesh) return fail(c, TF_ERR_STATE, "tf_set_mesh first");
        const Topology &t = c->topo;
        if (n != t.nE || !ia || !ja || !va) return fail(c, TF_ERR_ARG, "matrix size differs from the mesh");
        for (int i = 0; i <= n; i++)
            if (ia[i] != t.ia[i] + 1) return fail(c, TF_ERR_INPUT, "row pointer differs from the assembled pattern");
        for (int64_t q = 0; q < t.nnz; q++)
            if (ja[q] != t.ja[q] + 1) return fail(c, TF_ERR_INPUT, "column index differs from the assembled pattern");
        if (!c->analyzed) do_analyze(c, c->leaf_size);
        prepare_slot(c, slot);
        double *dv = dalloc<double>(t.nnz);
        HIPCHK(hipMemcpyAsync(dv, va, sizeof(double) * t.nnz, hipMemcpyHostToDevice, c->st));
        tf::launch_scatter_vals(t.nnz, c->d_amap, dv, c->d_L[slot], c->st);
        run_factor(c, c->d_L[slot], c->d_D[slot], c->d_fd);
        copy_l21(c, slot);
        c->slot_live[slot] = true;
        c->slot_dt[slot] = -1.0;
        HIPCHK(hipStreamSynchronize(c->st));
        dfree(dv);
        try {
            check_flag(c);
        } catch (...) {
            c->slot_live[slot] = false;
            throw;
        }
        return TF_OK;
    });
}

int tf_solve(tf_ctx *c, int slot, int nrhs, const double *b, double *x) {
    return guarded(c, [&] {
        if (slot < 0 || slot > 1 || !c->slot_live[slot]) return fail(c, TF_ERR_STATE, "slot not factored");
        if (nrhs <= 0 || !b || !x) return fail(c, TF_ERR_ARG, "bad right-hand side");
        const int n = c->topo.nE;
        if (nrhs == 1) {
            HIPCHK(hipMemcpyAsync(c->d_tmpb, b, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
            solve_dev(c, c->d_L[slot], c->d_D[slot], c->d_fd, c->d_tmpb, c->d_tmpx, ctx_work(c), nullptr, c->d_LT[slot]);
            HIPCHK(hipMemcpyAsync(x, c->d_tmpx, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
        } else {
            // blocks of up to 64 right-hand sides through the block solve
            for (int q0 = 0; q0 < nrhs; q0 += 64) {
                const int K = std::min(64, nrhs - q0);
                ensure_block(c, K);
                std::vector<double> h((size_t)n * K);
                for (int i = 0; i < n; i++)
                    for (int q = 0; q < K; q++) h[(size_t)i * K + q] = b[(size_t)(q0 + q) * n + i];
                HIPCHK(hipMemcpyAsync(c->d_mb, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->st));
                solve_block(c, c->d_L[slot], c->d_D[slot], c->d_fd, K, c->d_mb, c->d_mx, nullptr, c->d_LT[slot]);
                HIPCHK(hipMemcpyAsync(h.data(), c->d_mx, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
                HIPCHK(hipStreamSynchronize(c->st));
                for (int i = 0; i < n; i++)
                    for (int q = 0; q < K; q++) x[(size_t)(q0 + q) * n + i] = h[(size_t)i * K + q];
            }
        }
        HIPCHK(hipStreamSynchronize(c->st));
        HIPCHK(hipGetLastError());
        return TF_OK;
    });
}

int tf_solve_noalloc(tf_ctx *c, int slot, const double *b, double *x) {
    return guarded(c, [&] {
        if (slot < 0 || slot > 1 || !c->slot_live[slot]) return fail(c, TF_ERR_STATE, "slot not factored");
        if (c->nmodels > 0) return fail(c, TF_ERR_STATE, "independent-model contexts have no single-system solve");
        if (!b || !x) return fail(c, TF_ERR_ARG, "bad right-hand side");
        for (const void *p : {(const void *)b, (const void *)x}) {
            hipPointerAttribute_t at;
            if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type == hipMemoryTypeHost ||
                at.type == hipMemoryTypeUnregistered) {
                (void)hipGetLastError();
                return fail(c, TF_ERR_ARG, "tf_solve_noalloc takes device pointers for b and x");
            }
        }
        const int64_t n = c->sym.n;
        auto up = [](int64_t v) { return (v + 31) / 32 * 32; };  // 256-byte aligned vectors
        const int64_t nv = up(c->vrows + TF_READ_SLACK), nn = up(n + TF_READ_SLACK);
        const int64_t npart = up(c->plan.spart_max * 1024);
        const int64_t need = nv + 3 * nn + npart;
        if (need > TF_NOALLOC_DOUBLES)
            return fail(c, TF_ERR_ARG,
                        "system too large for the allocation-free solve: its work vectors need " +
                            std::to_string(need) + " doubles, the scratch array holds " +
                            std::to_string((int64_t)TF_NOALLOC_DOUBLES) + " (64 MiB)");
        double *p = c->d_scratch;
        const SolveWork w{p, p + nv, p + nv + nn, p + nv + 2 * nn, p + nv + 3 * nn};
        const bool prof = c->prof_on;  // no event pool growth inside this call
        c->prof_on = false;
        solve_dev(c, c->d_L[slot], c->d_D[slot], c->d_fd, b, x, w, nullptr, c->d_LT[slot]);
        c->prof_on = prof;
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
        return TF_OK;
    });
}

int tf_time_solve(tf_ctx *c, int slot, int nrhs, int reps, double *ms_per_solve) {
    return guarded(c, [&] {
        const bool models = c->nmodels > 0;
        if (models ? !c->models_factored : (slot < 0 || slot > 1 || !c->slot_live[slot]))
            return fail(c, TF_ERR_STATE, models ? "models not factored (tf_run_models first)" : "slot not factored");
        if (nrhs < 1 || nrhs > 64 || reps < 1 || !ms_per_solve) return fail(c, TF_ERR_ARG, "bad arguments");
        if (models && nrhs != 1) return fail(c, TF_ERR_ARG, "independent models solve one right-hand side each");
        const int n = c->topo.nE;
        double *b, *x;
        if (nrhs == 1) { b = c->d_tmpb; x = c->d_tmpx; }
        else { ensure_block(c, nrhs); b = c->d_mb; x = c->d_mx; }
        tf::launch_fill_pattern((int64_t)n * nrhs, b, c->st);
        auto one = [&] {
            if (models) solve_dev(c, c->d_Lmm, c->d_Dmm, c->d_fd_cur, b, x);  // every model's current factor
            else if (nrhs == 1) solve_dev(c, c->d_L[slot], c->d_D[slot], c->d_fd, b, x, ctx_work(c), nullptr, c->d_LT[slot]);
            else solve_block(c, c->d_L[slot], c->d_D[slot], c->d_fd, nrhs, b, x, nullptr, c->d_LT[slot]);
        };
        one();  // warm-up
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, c->st));
        for (int r = 0; r < reps; r++) one();
        HIPCHK(hipEventRecord(e1, c->st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIPCHK(hipGetLastError());
        *ms_per_solve = ms / reps;
        return TF_OK;
    });
}

int tf_release(tf_ctx *c, int slot) {
    if (!c || slot < 0 || slot > 1) return TF_ERR_ARG;
    c->slot_live[slot] = false;
    return TF_OK;
}

int tf_set_models(tf_ctx *c, int nmodels, const int *ntet, const int *nrec) {
    return guarded(c, [&] {
        c->tape_valid = false;
        if (!c->have_mesh) return fail(c, TF_ERR_STATE, "tf_set_mesh first");
        if (nmodels < 1 || !ntet || !nrec) return fail(c, TF_ERR_ARG, "bad model counts");
        const Topology &t = c->topo;
        std::vector<int> tet0(nmodels), rec0(nmodels), vlo(nmodels, INT32_MAX), vhi(nmodels, -1);
        int64_t ts = 0, rs = 0;
        for (int m = 0; m < nmodels; m++) {
            if (ntet[m] <= 0 || nrec[m] < 0) return fail(c, TF_ERR_ARG, "bad model counts");
            tet0[m] = (int)ts;
            rec0[m] = (int)rs;
            ts += ntet[m];
            rs += nrec[m];
        }
        if (ts != t.nT) return fail(c, TF_ERR_ARG, "model tet counts do not add up to the mesh");
        for (int m = 0; m < nmodels; m++)
            for (int i = tet0[m]; i < tet0[m] + ntet[m]; i++)
                for (int q = 0; q < 4; q++) {
                    const int vx = t.tet_vex[4 * (size_t)i + q];
                    vlo[m] = std::min(vlo[m], vx);
                    vhi[m] = std::max(vhi[m], vx);
                }
        for (int m = 1; m < nmodels; m++)
            if (vlo[m] <= vhi[m - 1]) return fail(c, TF_ERR_INPUT, "models must use disjoint, increasing vertex blocks");
        // edges are numbered by (min, max) vertex: each model's edges form one range
        std::vector<int> e0(nmodels, -1), ne(nmodels, 0);
        for (int e = 0; e < t.nE; e++) {
            const int v1 = t.edg_vex[2 * (size_t)e];
            int m = 0;
            while (m + 1 < nmodels && v1 >= vlo[m + 1]) m++;
            if (e0[m] < 0) e0[m] = e;
            ne[m]++;
        }
        for (int m = 0; m < nmodels; m++)
            if (e0[m] < 0) return fail(c, TF_ERR_INPUT, "a model without edges");
        c->nmodels = nmodels;
        c->m_tet0 = tet0;
        c->m_ntet.assign(ntet, ntet + nmodels);
        c->m_rec0 = rec0;
        c->m_nrec.assign(nrec, nrec + nmodels);
        c->m_e0 = e0;
        c->m_ne = ne;
        c->nsrc = 1;
        c->src_edges.assign(nmodels, {});
        c->src_ws.assign(nmodels, {});
        c->have_src = false;
        free_native(c);  // the native-order maps follow the new sources, receivers or models
        return TF_OK;
    });
}

int tf_run_models(tf_ctx *c, int max_steps, int use_graphs) {
    return guarded(c, [&] {
        if (c->nmodels <= 0) return fail(c, TF_ERR_STATE, "tf_set_models first");
        if (!c->assembled) return fail(c, TF_ERR_STATE, "tf_assemble first");
        if (!c->have_src) return fail(c, TF_ERR_STATE, "tf_set_source for every model first");
        if (c->nR == 0) return fail(c, TF_ERR_STATE, "tf_set_receivers first");
        if (c->sp_division == 0) return fail(c, TF_ERR_STATE, "tf_set_schedule first");
        if (c->ptype == 3) return fail(c, TF_ERR_STATE, "independent models run step-on and impulse sources");
        if (!c->analyzed) do_analyze(c, c->leaf_size);
        if (!c->d_Lmm) setup_models(c);
        for (int k = 1; k < 8; k++) c->timers[k] = 0.0;
        c->row_step.clear(); c->row_t.clear(); c->row_dt.clear();
        c->use_graphs = use_graphs != 0;
        auto t0 = std::chrono::steady_clock::now();
        RunScope scope(c);
        run_models(c, max_steps);
        c->timers[5] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return TF_OK;
    });
}

int tf_get_model_rows(tf_ctx *c, int model, int max_rows, int *nrows, int *row_step, double *row_t, double *row_dt,
                      double *fields) {
    return guarded(c, [&] {
        if (c->nmodels <= 0 || model < 0 || model >= c->nmodels) return fail(c, TF_ERR_ARG, "no such model");
        const int M = c->nmodels, R = c->nR, r0 = c->m_rec0[model], nr = c->m_nrec[model];
        const int64_t per = (int64_t)R * 6;
        std::vector<double> h(per * c->nrows);
        if (c->nrows > 0)
            HIPCHK(hipMemcpy(h.data(), c->d_rows, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
        int k = 0;
        std::vector<double> bsum(3 * (size_t)nr, 0.0);
        for (int r = 0; r < c->nrows; r++) {
            if (!c->mm_on[(size_t)r * M + model]) continue;
            const double dtm = c->mm_dt[(size_t)r * M + model];
            if (k < max_rows) {
                if (row_step) row_step[k] = c->mm_step[r];
                if (row_t) row_t[k] = c->mm_t[(size_t)r * M + model];
                if (row_dt) row_dt[k] = dtm;
                for (int i = 0; i < nr; i++)
                    for (int q = 0; q < 3; q++) {
                        const double *src = &h[(size_t)r * per + (size_t)(r0 + i) * 6];
                        const double d = (c->mm_step[r] == 0) ? 0.0 : src[3 + q];
                        bsum[3 * i + q] += d * dtm;
                        if (fields) {
                            double *o = fields + ((int64_t)k * nr + i) * 9;
                            o[q] = src[q];
                            o[3 + q] = d;
                            o[6 + q] = bsum[3 * i + q];
                        }
                    }
            }
            k++;
        }
        if (nrows) *nrows = k;
        return TF_OK;
    });
}

int tf_run(tf_ctx *c, int max_steps, int max_rows, int *nrows, int *row_step, double *row_t, double *row_dt,
           double *fields) {
    return guarded(c, [&] {
        if (c->nmodels > 0) return fail(c, TF_ERR_STATE, "independent models run with tf_run_models");
        if (!c->assembled) return fail(c, TF_ERR_STATE, "tf_assemble first");
        if (!c->have_src) return fail(c, TF_ERR_STATE, "tf_set_source first");
        if (c->nR == 0) return fail(c, TF_ERR_STATE, "tf_set_receivers first");
        if (c->sp_division == 0) return fail(c, TF_ERR_STATE, "tf_set_schedule first");
        if (!c->analyzed) do_analyze(c, c->leaf_size);
        for (int k = 1; k < 6; k++) c->timers[k] = 0.0;
        c->timers[7] = 0.0;
        c->nrows = 0;
        c->row_step.clear(); c->row_t.clear(); c->row_dt.clear();
        c->chk_step.clear(); c->chk_err.clear(); c->chk_keep.clear();
        auto t0 = std::chrono::steady_clock::now();
        RunScope scope(c);
        if (c->ptype == 3) {
            double *res = dalloc<double>((size_t)c->topo.nE * c->nsrc);
            run_loop(c, 1, res, max_steps);
            c->nrows = 0;
            c->row_step.clear(); c->row_t.clear(); c->row_dt.clear();
            run_loop(c, 2, res, max_steps);
            dfree(res);
        } else {
            run_loop(c, 0, nullptr, max_steps);
        }
        HIPCHK(hipStreamSynchronize(c->st));
        c->timers[5] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (nrows) *nrows = c->nrows;
        int keep = std::min(c->nrows, std::max(max_rows, 0));
        if (keep > 0) {
            if (row_step) std::copy(c->row_step.begin(), c->row_step.begin() + keep, row_step);
            if (row_t) std::copy(c->row_t.begin(), c->row_t.begin() + keep, row_t);
            if (row_dt) std::copy(c->row_dt.begin(), c->row_dt.begin() + keep, row_dt);
            if (fields) {
                const int K = c->nsrc, R = c->nR;
                const int64_t per = (int64_t)R * 6 * K;
                std::vector<double> h(per * keep);
                HIPCHK(hipMemcpy(h.data(), c->d_rows, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
                // fields[row][src][rec][9]: E and dB/dt from the device; dB/dt is
                // zero at step 0 (src/FETD_FWD.f90:1541); B is the running sum of
                // dB/dt * dt (:1602-1621)
                for (int r = 0; r < keep; r++)
                    for (int q = 0; q < K; q++)
                        for (int i = 0; i < R; i++) {
                            double *o = fields + (((int64_t)r * K + q) * R + i) * 9;
                            const double *src = &h[(r * (int64_t)K + q) * R * 6 + (int64_t)i * 6];
                            for (int k = 0; k < 3; k++) {
                                o[k] = src[k];
                                o[3 + k] = (c->row_step[r] == 0) ? 0.0 : src[3 + k];
                                const double prev =
                                    (r == 0) ? 0.0 : fields[(((int64_t)(r - 1) * K + q) * R + i) * 9 + 6 + k];
                                o[6 + k] = (r == 0) ? o[3 + k] * c->row_dt[r] : prev + o[3 + k] * c->row_dt[r];
                            }
                        }
            }
        }
        return TF_OK;
    });
}

int tf_get_checks(const tf_ctx *c, int max, int *n, int *step, double *err, int *keep) {
    if (!c) return TF_ERR_ARG;
    int cnt = (int)c->chk_step.size();
    if (n) *n = cnt;
    int k = std::min(cnt, std::max(max, 0));
    if (step) std::copy(c->chk_step.begin(), c->chk_step.begin() + k, step);
    if (err) std::copy(c->chk_err.begin(), c->chk_err.begin() + k, err);
    if (keep) std::copy(c->chk_keep.begin(), c->chk_keep.begin() + k, keep);
    return TF_OK;
}

int tf_get_timers(const tf_ctx *c, double *t8) {
    if (!c || !t8) return TF_ERR_ARG;
    std::copy(c->timers, c->timers + 8, t8);
    return TF_OK;
}

}  // extern "C"
