// msolve.hip -- block solve for K right-hand sides (K <= 64 sources sharing
// one LDL^T factor), FP64 on matrix cores.
//
// The reference solves one source per run (pardiso_direct_2 with nrhs=1,
// src/FETD_UTILITY.f90:938-957).  Here K sources are time-stepped together:
// every vector is [row][K] (the K values of a row contiguous) and each
// front's solve becomes four tile products of 64 rows x 64 right-hand sides
// (v_mfma_f64_16x16x4f64 via mfma64_acc):
//   forward   Y_own = X11 V_own          V_struct -= L21 Y_own
//   backward  T = D^-1 Y_own - L21^T Xs  X_own = X11^T T
// with X11 = L11^-1 stored in place by the factorisation.  The factor is
// read once per step for all K sources.
//
// Result layout of mfma64_acc here: the right-hand-side index plays the
// "row" role (fast index of a result register), so every store writes
// 16 consecutive right-hand sides (128 bytes) of 4 matrix rows.
#include "tf_device.h"

#define TLD 65  // odd leading dimension for operands staged transposed

// ---- gathers ---------------------------------------------------------------------

// ---- tile products ----------------------------------------------------------------

// stage 64 rows x 64 rhs of a [row][K] block starting at row r0 (rows < rend)
__device__ __forceinline__ void stage_rows(double *dst, const double *src, int r0, int rend, int K) {
#pragma unroll 4
    for (int idx = threadIdx.x; idx < 64 * 64; idx += TF_NT) {
        const int r = idx >> 6, q = idx & 63;
        dst[r * MLD + q] = (r0 + r < rend && q < K) ? src[(int64_t)(r0 + r) * K + q] : 0.0;
    }
}

// forward own rows R0..R0+63 (< nc): Y = V + X11 V (strict lower part)
__global__ void __launch_bounds__(TF_NT) k_mfwd_own(const double *L, const FrontDesc *fd, const int *list,
                                                    const int64_t *pref, int cnt, const double *V, double *Y,
                                                    int K) {
    __shared__ double Bt[64 * MLD];  // [j][rhs]
    __shared__ double Xt[64 * MLD];  // [j][row]
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int R0 = (int)(blockIdx.x - pref[i]) * 64;
    const double *A = L + d.loff;
    const double *vf = V + d.voff * K;
    tf_d4 acc[2][2];
    acc_zero(acc);
    for (int j0 = 0; j0 <= R0; j0 += 64) {
        stage_rows(Bt, vf, j0, d.nc, K);
#pragma unroll 4
        for (int idx = threadIdx.x; idx < 64 * 64; idx += TF_NT) {
            const int jj = idx >> 6, r = idx & 63;
            const int j = j0 + jj, row = R0 + r;
            Xt[jj * MLD + r] = (j < row && row < d.nc) ? A[(int64_t)j * d.f + row] : 0.0;
        }
        __syncthreads();
        mfma64_acc(Bt, MLD, Xt, MLD, 64, acc);
        __syncthreads();
    }
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = R0 + ACC_COL(bj, r);
#pragma unroll
            for (int bi = 0; bi < 2; bi++) {
                const int q = ACC_ROW(bi);
                if (row < d.nc && q < K)
                    Y[(int64_t)(d.c0 + row) * K + q] = vf[(int64_t)row * K + q] + acc[bi][bj][r];
            }
        }
}

// backward columns J0..J0+63: X = T + X11^T T (strict lower part of X11)
__global__ void __launch_bounds__(TF_NT) k_mbwd_x(const double *L, const FrontDesc *fd, const int *list,
                                                  const int64_t *pref, int cnt, const double *T, double *xw,
                                                  int K) {
    __shared__ double Bt[64 * MLD];  // [i][rhs]
    __shared__ double Xt[64 * TLD];  // [i][col]
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int J0 = (int)(blockIdx.x - pref[i]) * 64;
    const double *A = L + d.loff;
    const double *t = T + (int64_t)d.c0 * K;
    tf_d4 acc[2][2];
    acc_zero(acc);
    for (int i0 = J0; i0 < d.nc; i0 += 64) {
        stage_rows(Bt, t, i0, d.nc, K);
#pragma unroll 4
        for (int idx = threadIdx.x; idx < 64 * 64; idx += TF_NT) {
            const int c = idx >> 6, ii = idx & 63;
            const int j = J0 + c, row = i0 + ii;
            Xt[ii * TLD + c] = (row > j && row < d.nc) ? A[(int64_t)j * d.f + row] : 0.0;
        }
        __syncthreads();
        mfma64_acc(Bt, MLD, Xt, TLD, 64, acc);
        __syncthreads();
    }
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int j = J0 + ACC_COL(bj, r);
            if (j >= d.nc) continue;
#pragma unroll
            for (int bi = 0; bi < 2; bi++) {
                const int q = ACC_ROW(bi);
                if (q < K) {
                    const int64_t o = (int64_t)(d.c0 + j) * K + q;
                    xw[o] = t[(int64_t)j * K + q] + acc[bi][bj][r];
                }
            }
        }
}

__global__ void k_munpermute(int n, int K, const int *perm, const double *xw, double *x) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < (int64_t)n * K) {
        const int i = (int)(idx / K), q = (int)(idx % K);
        x[(int64_t)perm[i] * K + q] = xw[idx];
    }
}

// ---- K-wide step kernels ------------------------------------------------------------

// w = 4 e2 - e3 (the BDF2 history combination), once per entry
__global__ void k_mhist(int64_t len, const double *e2, const double *e3, double *w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) w[i] = 4.0 * e2[i] - e3[i];
}

// r[row][q] = sum_j M[row][j] w[j][q]
__global__ void __launch_bounds__(256) k_mspmv_rhs(int n, int K, const int *ia, const int *ja, const double *val,
                                                   const double *w, double *r) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * K) return;
    const int row = (int)(idx / K), q = (int)(idx % K);
    double acc = 0.0;
    for (int p = ia[row]; p < ia[row + 1]; p++) acc += val[p] * w[(int64_t)ja[p] * K + q];
    r[idx] = acc;
}

// r[edge][q] -= (2 dt) (didt w) for every source-wire entry (edge, q, w)
__global__ void k_msource(int ns, int K, const int *edge, const int *src, const double *w, double didt, double cdt,
                          double *r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) {
        const int64_t o = (int64_t)edge[i] * K + src[i];
        const double s = didt * w[i];
        r[o] = r[o] - cdt * s;
    }
}

// out[(q*nR + rec)*6 + 0..5] = E, dB/dt of source q at receiver rec
__global__ void k_mreadout(int nR, int K, const int *rec_edge, const double *rec_drc, const double *n_arr,
                           const double *d_arr, const double *e, double *out) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nR * K) return;
    const int rec = idx / K, q = idx % K;
    double ep[3] = {0.0, 0.0, 0.0}, dp[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 6; j++) {
        const double ev = e[(int64_t)rec_edge[6 * rec + j] * K + q];
        const double drc = rec_drc[6 * rec + j];
        for (int k = 0; k < 3; k++) {
            ep[k] += n_arr[18 * rec + 3 * j + k] * drc * ev;
            dp[k] += d_arr[18 * rec + 3 * j + k] * drc * ev;
        }
    }
    double *o = out + ((int64_t)q * nR + rec) * 6;
    for (int k = 0; k < 3; k++) {
        o[k] = ep[k];
        o[3 + k] = dp[k];
    }
}

// per-column sums of (a-b)^2 and b^2: part[blk][q][2]
__global__ void __launch_bounds__(256) k_mdiffnorm(int64_t n, int K, const double *a, const double *b, double *part) {
    __shared__ double s1[256], s2[256];
    const int q = threadIdx.x % 64, lane_row = threadIdx.x / 64;
    double x1 = 0.0, x2 = 0.0;
    if (q < K) {
        for (int64_t row = (int64_t)blockIdx.x * 4 + lane_row; row < n; row += (int64_t)gridDim.x * 4) {
            const double bv = b[row * K + q];
            const double dv = a[row * K + q] - bv;
            x1 += dv * dv;
            x2 += bv * bv;
        }
    }
    s1[threadIdx.x] = x1;
    s2[threadIdx.x] = x2;
    __syncthreads();
    if (threadIdx.x < 64) {
        const double t1 = s1[threadIdx.x] + s1[threadIdx.x + 64] + s1[threadIdx.x + 128] + s1[threadIdx.x + 192];
        const double t2 = s2[threadIdx.x] + s2[threadIdx.x + 64] + s2[threadIdx.x + 128] + s2[threadIdx.x + 192];
        part[((int64_t)blockIdx.x * 64 + threadIdx.x) * 2] = t1;
        part[((int64_t)blockIdx.x * 64 + threadIdx.x) * 2 + 1] = t2;
    }
}

__global__ void k_mreduce(int nb, int K, const double *part, double *out) {
    const int q = threadIdx.x;
    if (q >= K) return;
    double t1 = 0.0, t2 = 0.0;
    for (int b = 0; b < nb; b++) {
        t1 += part[((int64_t)b * 64 + q) * 2];
        t2 += part[((int64_t)b * 64 + q) * 2 + 1];
    }
    out[2 * q] = t1;
    out[2 * q + 1] = t2;
}

// ---- grid-parallel gathers (work units of TF_GCH row*rhs elements) -----------------
#define TF_GROWS 32  // rows per workgroup in the gathers (x K <= 64 values each)

// vf = [b(own); 0] for every front of a level
__global__ void __launch_bounds__(TF_NT) k_mg_init(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                                                   const double *b, const int *perm, double *V, int K) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    double *vf = V + d.voff * K;
    const int r0 = (int)(blockIdx.x - pref[i]) * TF_GROWS;
    const int64_t e0 = (int64_t)r0 * K, e1 = (int64_t)min(r0 + TF_GROWS, d.f) * K;
    for (int64_t idx = e0 + threadIdx.x; idx < e1; idx += TF_NT) {
        const int j = (int)(idx / K), q = (int)(idx % K);
        vf[idx] = (j < d.nc) ? b[(int64_t)perm[d.c0 + j] * K + q] : 0.0;
    }
}

// extend-add of one child per parent: parent rows rel[a] += child update row a
__global__ void __launch_bounds__(TF_NT) k_mg_child(const FrontDesc *fd, const int *parent, const int *list,
                                                    const int64_t *pref, int cnt, const int *rel, double *V, int K) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc dc = fd[list[i]];
    const FrontDesc dp = fd[parent[list[i]]];
    const double *uc = V + (dc.voff + dc.nc) * K;
    double *vp = V + dp.voff * K;
    const int *rc = rel + dc.soff;
    const int r0 = (int)(blockIdx.x - pref[i]) * TF_GROWS;
    const int64_t e0 = (int64_t)r0 * K, e1 = (int64_t)min(r0 + TF_GROWS, dc.m) * K;
    for (int64_t idx = e0 + threadIdx.x; idx < e1; idx += TF_NT) {
        const int a = (int)(idx / K), q = (int)(idx % K);
        vp[(int64_t)rc[a] * K + q] += uc[idx];
    }
}

// xs rows of every front of a level: V[nc + r] = xw[sidx[r]]
__global__ void __launch_bounds__(TF_NT) k_mg_xs(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                                                 const int *sidx, const double *xw, double *V, int K) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    double *xs = V + (d.voff + d.nc) * K;
    const int r0 = (int)(blockIdx.x - pref[i]) * TF_GROWS;
    const int64_t e0 = (int64_t)r0 * K, e1 = (int64_t)min(r0 + TF_GROWS, d.m) * K;
    for (int64_t idx = e0 + threadIdx.x; idx < e1; idx += TF_NT) {
        const int r = (int)(idx / K), q = (int)(idx % K);
        xs[idx] = xw[(int64_t)sidx[d.soff + r] * K + q];
    }
}

// ---- split tile products with partial sums ----------------------------------------
#define TF_SCH 512  // inner-dimension chunk per workgroup

// partial (row tile, column chunk) of L21 Y_own: part[u][rhs*64 + row]
__global__ void __launch_bounds__(TF_NT) k_mfwd_struct_part(const double *L, const FrontDesc *fd, const int *list,
                                                            const int64_t *pref, int cnt, const double *Y, int K,
                                                            double *part, double *V) {
    __shared__ double Bt[64 * MLD];
    __shared__ double Lt[64 * MLD];
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int u = (int)(blockIdx.x - pref[i]);
    const int nch = (d.nc + TF_SCH - 1) / TF_SCH;
    const int R0 = (u / nch) * 64, jc = (u % nch) * TF_SCH;
    const double *A = L + d.loff;
    const double *y = Y + (int64_t)d.c0 * K;
    tf_d4 acc[2][2];
    acc_zero(acc);
    const int jend = min(d.nc, jc + TF_SCH);
    for (int j0 = jc; j0 < jend; j0 += 64) {
        stage_rows(Bt, y, j0, jend, K);
#pragma unroll 4
        for (int idx = threadIdx.x; idx < 64 * 64; idx += TF_NT) {
            const int jj = idx >> 6, r = idx & 63;
            const int j = j0 + jj, row = R0 + r;
            Lt[jj * MLD + r] = (j < jend && row < d.m) ? A[(int64_t)j * d.f + d.nc + row] : 0.0;
        }
        __syncthreads();
        mfma64_acc(Bt, MLD, Lt, MLD, 64, acc);
        __syncthreads();
    }
    if (nch == 1) {  // single chunk: finish here, no partial tile
        double *vs = V + (d.voff + d.nc) * K;
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = R0 + ACC_COL(bj, r);
#pragma unroll
                for (int bi = 0; bi < 2; bi++) {
                    const int q = ACC_ROW(bi);
                    if (row < d.m && q < K) vs[(int64_t)row * K + q] -= acc[bi][bj][r];
                }
            }
        return;
    }
    double *pp = part + (int64_t)blockIdx.x * 4096;
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int bi = 0; bi < 2; bi++) pp[ACC_ROW(bi) * 64 + ACC_COL(bj, r)] = acc[bi][bj][r];
}

// V_struct[row] -= sum over the row tile's chunks
__global__ void __launch_bounds__(TF_NT) k_mfwd_struct_red(const FrontDesc *fd, const int *list, const int64_t *pref,
                                                           const int64_t *ppref, int cnt, const double *part, double *V,
                                                           int K) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int rt = (int)(blockIdx.x - pref[i]);
    const int nch = (d.nc + TF_SCH - 1) / TF_SCH;
    if (nch == 1) return;
    const double *p0 = part + (ppref[i] + (int64_t)rt * nch) * 4096;
    double *vs = V + (d.voff + d.nc) * K;
    for (int e = threadIdx.x; e < 64 * 64; e += TF_NT) {
        const int q = e >> 6, rr = e & 63;
        const int row = rt * 64 + rr;
        if (row >= d.m || q >= K) continue;
        double s = 0.0;
        for (int c = 0; c < nch; c++) s += p0[(int64_t)c * 4096 + e];
        vs[(int64_t)row * K + q] -= s;
    }
}

// partial (column tile, row chunk) of L21^T Xs: part[u][rhs*64 + col]
__global__ void __launch_bounds__(TF_NT) k_mbwd_t_part(const double *L, const FrontDesc *fd, const int *list,
                                                       const int64_t *pref, int cnt, const double *V, int K,
                                                       double *part, const double *Y, const double *Dn, double *T) {
    __shared__ double Bt[64 * MLD];
    __shared__ double Lt[64 * TLD];
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int u = (int)(blockIdx.x - pref[i]);
    const int nch = (d.m + TF_SCH - 1) / TF_SCH;
    const int J0 = (u / nch) * 64, rc = (u % nch) * TF_SCH;
    const double *A = L + d.loff;
    const double *xs = V + (d.voff + d.nc) * K;
    tf_d4 acc[2][2];
    acc_zero(acc);
    const int rend = min(d.m, rc + TF_SCH);
    for (int r0 = rc; r0 < rend; r0 += 64) {
        stage_rows(Bt, xs, r0, rend, K);
#pragma unroll 4
        for (int idx = threadIdx.x; idx < 64 * 64; idx += TF_NT) {
            const int c = idx >> 6, rr = idx & 63;
            const int j = J0 + c, row = r0 + rr;
            Lt[rr * TLD + c] = (j < d.nc && row < rend) ? A[(int64_t)j * d.f + d.nc + row] : 0.0;
        }
        __syncthreads();
        mfma64_acc(Bt, MLD, Lt, TLD, 64, acc);
        __syncthreads();
    }
    if (nch == 1) {  // single chunk: finish here, no partial tile
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int j = J0 + ACC_COL(bj, r);
                if (j >= d.nc) continue;
                const double dj = Dn[d.c0 + j];
#pragma unroll
                for (int bi = 0; bi < 2; bi++) {
                    const int q = ACC_ROW(bi);
                    if (q < K) {
                        const int64_t o = (int64_t)(d.c0 + j) * K + q;
                        T[o] = Y[o] / dj - acc[bi][bj][r];
                    }
                }
            }
        return;
    }
    double *pp = part + (int64_t)blockIdx.x * 4096;
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int bi = 0; bi < 2; bi++) pp[ACC_ROW(bi) * 64 + ACC_COL(bj, r)] = acc[bi][bj][r];
}

// T[col] = Y/d - sum over the column tile's chunks
__global__ void __launch_bounds__(TF_NT) k_mbwd_t_red(const FrontDesc *fd, const int *list, const int64_t *pref,
                                                      const int64_t *ppref, int cnt, const double *part, const double *Y,
                                                      const double *Dn, double *T, int K) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int jt = (int)(blockIdx.x - pref[i]);
    const int nch = (d.m + TF_SCH - 1) / TF_SCH;
    if (nch == 1) return;
    const double *p0 = part + (ppref[i] + (int64_t)jt * nch) * 4096;
    for (int e = threadIdx.x; e < 64 * 64; e += TF_NT) {
        const int q = e >> 6, cc = e & 63;
        const int j = jt * 64 + cc;
        if (j >= d.nc || q >= K) continue;
        double s = 0.0;
        for (int c = 0; c < nch; c++) s += p0[(int64_t)c * 4096 + e];
        const int64_t o = (int64_t)(d.c0 + j) * K + q;
        T[o] = Y[o] / Dn[d.c0 + j] - s;
    }
}

// ---- host launchers --------------------------------------------------------------
namespace tf {
void launch_mfwd_own(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                     int64_t nwg, const double *V, double *Y, int K, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_mfwd_own, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, V, Y, K);
}
void launch_mbwd_x(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                   const double *T, double *xw, int K, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_mbwd_x, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, T, xw, K);
}
void launch_munpermute(int n, int K, const int *perm, const double *xw, double *x, hipStream_t st) {
    const int64_t tot = (int64_t)n * K;
    hipLaunchKernelGGL(k_munpermute, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, K, perm, xw, x);
}
void launch_mspmv_rhs(int n, int K, const int *ia, const int *ja, const double *val, const double *e2,
                      const double *e3, double *w, double *r, hipStream_t st) {
    const int64_t tot = (int64_t)n * K;
    hipLaunchKernelGGL(k_mhist, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, tot, e2, e3, w);
    hipLaunchKernelGGL(k_mspmv_rhs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, K, ia, ja, val, w, r);
}
void launch_msource(int ns, int K, const int *edge, const int *src, const double *w, double didt, double cdt,
                    double *r, hipStream_t st) {
    if (ns > 0)
        hipLaunchKernelGGL(k_msource, dim3((ns + 63) / 64), dim3(64), 0, st, ns, K, edge, src, w, didt, cdt, r);
}
void launch_mreadout(int nR, int K, const int *rec_edge, const double *rec_drc, const double *n_arr,
                     const double *d_arr, const double *e, double *out, hipStream_t st) {
    const int tot = nR * K;
    hipLaunchKernelGGL(k_mreadout, dim3((tot + 63) / 64), dim3(64), 0, st, nR, K, rec_edge, rec_drc, n_arr, d_arr, e,
                       out);
}
void launch_mdiffnorm(int64_t n, int K, const double *a, const double *b, double *part, double *out,
                      hipStream_t st) {
    const int nb = 256;
    hipLaunchKernelGGL(k_mdiffnorm, dim3(nb), dim3(256), 0, st, n, K, a, b, part);
    hipLaunchKernelGGL(k_mreduce, dim3(1), dim3(64), 0, st, nb, K, part, out);
}
void launch_mg_init(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg, const double *b,
                    const int *perm, double *V, int K, hipStream_t st) {
    if (nwg > 0) hipLaunchKernelGGL(k_mg_init, dim3((unsigned)nwg), dim3(TF_NT), 0, st, fd, list, pref, cnt, b, perm, V, K);
}
void launch_mg_child(const FrontDesc *fd, const int *parent, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                     const int *rel, double *V, int K, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_mg_child, dim3((unsigned)nwg), dim3(TF_NT), 0, st, fd, parent, list, pref, cnt, rel, V, K);
}
void launch_mg_xs(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg, const int *sidx,
                  const double *xw, double *V, int K, hipStream_t st) {
    if (nwg > 0) hipLaunchKernelGGL(k_mg_xs, dim3((unsigned)nwg), dim3(TF_NT), 0, st, fd, list, pref, cnt, sidx, xw, V, K);
}
void launch_mfwd_struct_part(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                             int64_t nwg, const double *Y, int K, double *part, double *V, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_mfwd_struct_part, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, Y, K, part,
                           V);
}
void launch_mfwd_struct_red(const FrontDesc *fd, const int *list, const int64_t *pref, const int64_t *ppref, int cnt,
                            int64_t nwg, const double *part, double *V, int K, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_mfwd_struct_red, dim3((unsigned)nwg), dim3(TF_NT), 0, st, fd, list, pref, ppref, cnt, part, V,
                           K);
}
void launch_mbwd_t_part(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                        const double *V, int K, double *part, const double *Y, const double *Dn, double *T,
                        hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_mbwd_t_part, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, V, K, part, Y,
                           Dn, T);
}
void launch_mbwd_t_red(const FrontDesc *fd, const int *list, const int64_t *pref, const int64_t *ppref, int cnt,
                       int64_t nwg, const double *part, const double *Y, const double *Dn, double *T, int K,
                       hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_mbwd_t_red, dim3((unsigned)nwg), dim3(TF_NT), 0, st, fd, list, pref, ppref, cnt, part, Y, Dn,
                           T, K);
}
}  // namespace tf
