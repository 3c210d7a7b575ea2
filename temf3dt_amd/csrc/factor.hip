// factor.hip -- multifrontal LDL^T of P = 3*Msigma + 2*dt*K on the GPU (FP64).
//
// Replaces the numeric part of PARDISO phase 12 called by the reference
// (pardiso_direct_1, src/FETD_UTILITY.f90:879-936; called at
// src/FETD_FWD.f90:1307,1361).  No pivoting (P is symmetric positive
// definite in exact arithmetic, PARDISO mtype 2); zero or non-finite pivots
// raise a flag that the host turns into an error.
//
// Work is organised by levels of the supernode tree (all fronts of a level
// are independent).  Levels with many fronts run one workgroup per front
// ("fused"); levels with few, large fronts split every 64-column panel into
// three grid launches: diagonal-block factor, panel solve, trailing update.
#include "tf_device.h"

#define LDS_LD 65

// ---- workgroup-level building blocks ------------------------------------

// Factor the pw x pw diagonal block of panel p0 in LDS: A = L D L^T with
// unit-lower L; writes L (strict lower), d on the diagonal and into D.
__device__ void blk_diag_factor(double *L, const FrontDesc &d, int p0, double *Dn, int *flag,
                                double *lds) {
    const int pw = min(TF_PB, d.nc - p0);
    double *base = L + d.loff;
    for (int idx = threadIdx.x; idx < pw * pw; idx += blockDim.x) {
        int i = idx % pw, j = idx / pw;
        lds[j * LDS_LD + i] = (i >= j) ? base[(int64_t)(p0 + j) * d.f + p0 + i] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < pw; j++) {
        double dj = lds[j * LDS_LD + j];
        if (!(fabs(dj) > 1e-300) || !isfinite(dj)) dj = (dj < 0.0) ? -1e-300 : 1e-300;
        const int nt = pw - j - 1;
        for (int idx = threadIdx.x; idx < nt * nt; idx += blockDim.x) {
            int i = j + 1 + idx % nt, k = j + 1 + idx / nt;
            if (i >= k) lds[k * LDS_LD + i] -= (lds[j * LDS_LD + i] / dj) * lds[j * LDS_LD + k];
        }
        __syncthreads();
        for (int i = j + 1 + threadIdx.x; i < pw; i += blockDim.x) lds[j * LDS_LD + i] /= dj;
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < pw * pw; idx += blockDim.x) {
        int i = idx % pw, j = idx / pw;
        if (i >= j) base[(int64_t)(p0 + j) * d.f + p0 + i] = lds[j * LDS_LD + i];
    }
    for (int j = threadIdx.x; j < pw; j += blockDim.x) {
        double dj = lds[j * LDS_LD + j];
        Dn[d.c0 + p0 + j] = dj;
        if (!(fabs(dj) > 1e-300) || !isfinite(dj)) atomicOr(flag, 1);
    }
    __syncthreads();
}

// Panel solve for rows [r0, r0+TF_TRSM_ROWS): L21 = A21 L11^-T D^-1, one
// thread per row with the row held in registers.
__device__ void blk_trsm(double *L, const FrontDesc &d, int p0, int r0, double *lds) {
    const int pw = min(TF_PB, d.nc - p0);
    double *base = L + d.loff;
    for (int idx = threadIdx.x; idx < pw * pw; idx += blockDim.x) {
        int i = idx % pw, j = idx / pw;
        lds[j * LDS_LD + i] = (i >= j) ? base[(int64_t)(p0 + j) * d.f + p0 + i] : 0.0;
    }
    __syncthreads();
    const int r = r0 + threadIdx.x;
    const bool act = r < d.f;
    double a[TF_PB];
#pragma unroll
    for (int j = 0; j < TF_PB; j++) a[j] = (act && j < pw) ? base[(int64_t)(p0 + j) * d.f + r] : 0.0;
#pragma unroll
    for (int j = 0; j < TF_PB; j++) {
        const double yj = a[j];
#pragma unroll
        for (int i = j + 1; i < TF_PB; i++) a[i] -= lds[j * LDS_LD + i] * yj;
    }
    if (act) {
#pragma unroll
        for (int j = 0; j < TF_PB; j++)
            if (j < pw) base[(int64_t)(p0 + j) * d.f + r] = a[j] / lds[j * LDS_LD + j];
    }
    __syncthreads();
}

// Trailing update of tile (ti, tj), ti >= tj, of the rows/columns after the
// panel: C -= L_i D L_j^T on the lower triangle.
__device__ void blk_update_tile(double *L, double *U, const FrontDesc &d, int p0, const double *Dn,
                                int ti, int tj, double *lds) {
    const int pw = min(TF_PB, d.nc - p0);
    const int p1 = p0 + pw;
    const int rb = p1 + ti * TF_TB, cb = p1 + tj * TF_TB;
    const double *base = L + d.loff;
    double *A = lds;               // [k][64]
    double *B = lds + TF_PB * TF_TB;  // [k][64], scaled by d_k
    for (int idx = threadIdx.x; idx < TF_PB * TF_TB; idx += blockDim.x) {
        int k = idx / TF_TB, i = idx % TF_TB;
        double va = 0.0, vb = 0.0;
        if (k < pw) {
            const double *col = base + (int64_t)(p0 + k) * d.f;
            if (rb + i < d.f) va = col[rb + i];
            if (cb + i < d.f) vb = col[cb + i] * Dn[d.c0 + p0 + k];
        }
        A[idx] = va;
        B[idx] = vb;
    }
    __syncthreads();
    const int tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[u][v] = 0.0;
    for (int k = 0; k < pw; k++) {
        double a4[4], b4[4];
#pragma unroll
        for (int u = 0; u < 4; u++) a4[u] = A[k * TF_TB + tr + 16 * u];
#pragma unroll
        for (int v = 0; v < 4; v++) b4[v] = B[k * TF_TB + tc + 16 * v];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[u][v] += a4[u] * b4[v];
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int c = cb + tc + 16 * v;
        if (c >= d.f) continue;
        double *col = front_col(const_cast<double *>(L), U, d, c);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int r = rb + tr + 16 * u;
            if (r < d.f && r >= c) col[r] -= acc[u][v];
        }
    }
    __syncthreads();
}

// ---- kernels ---------------------------------------------------------------

__global__ void k_fill0(double *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

// P = 3*M + 2*dt*K scattered into the L storage (src/FETD_FWD.f90:1288-1289).
__global__ void k_scatter_p(int64_t nnz, const int64_t *amap, const double *M, const double *K,
                            double cdt, double *L) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nnz) L[amap[q]] = cdt * K[q] + 3.0 * M[q];
}

// caller-provided values in the assembled CSR order
__global__ void k_scatter_vals(int64_t nnz, const int64_t *amap, const double *va, double *L) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nnz) L[amap[q]] = va[q];
}

// Extend-add of child update blocks into their parents' fronts; one
// workgroup per (child, column); at most one child per parent per launch.
__global__ void __launch_bounds__(TF_NT) k_extend_add(double *L, double *U, const FrontDesc *fd,
                                                      const int *parent, const int *list,
                                                      const int64_t *pref, int cnt, const int *rel) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const int c = list[i];
    const int b = (int)(blockIdx.x - pref[i]);
    const FrontDesc dc = fd[c];
    const FrontDesc dp = fd[parent[c]];
    const double *uc = U + dc.uoff + (int64_t)b * dc.m;
    const int *rc = rel + dc.soff;
    double *dst = front_col(L, U, dp, rc[b]);
    for (int a = b + threadIdx.x; a < dc.m; a += blockDim.x) dst[rc[a]] += uc[a];
}

__global__ void __launch_bounds__(TF_NT) k_factor_fused(double *L, double *U, const FrontDesc *fd,
                                                        const int *list, double *Dn, int *flag) {
    __shared__ double lds[2 * TF_PB * TF_TB + 64];
    const FrontDesc d = fd[list[blockIdx.x]];
    for (int p0 = 0; p0 < d.nc; p0 += TF_PB) {
        blk_diag_factor(L, d, p0, Dn, flag, lds);
        const int p1 = min(d.nc, p0 + TF_PB);
        for (int r0 = p1; r0 < d.f; r0 += TF_TRSM_ROWS) blk_trsm(L, d, p0, r0, lds);
        const int nt = (d.f - p1 + TF_TB - 1) / TF_TB;
        const int ntile = nt * (nt + 1) / 2;
        for (int t = 0; t < ntile; t++) {
            int ti, tj;
            tri_index(t, ti, tj);
            blk_update_tile(L, U, d, p0, Dn, ti, tj, lds);
        }
    }
}

__global__ void __launch_bounds__(TF_NT) k_factor_diag(double *L, const FrontDesc *fd, const int *list,
                                                       int p0, double *Dn, int *flag) {
    __shared__ double lds[TF_PB * LDS_LD];
    const FrontDesc d = fd[list[blockIdx.x]];
    blk_diag_factor(L, d, p0, Dn, flag, lds);
}

__global__ void __launch_bounds__(TF_NT) k_factor_trsm(double *L, const FrontDesc *fd, const int *list,
                                                       const int64_t *pref, int cnt, int p0) {
    __shared__ double lds[TF_PB * LDS_LD];
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int p1 = min(d.nc, p0 + TF_PB);
    const int r0 = p1 + (int)(blockIdx.x - pref[i]) * TF_TRSM_ROWS;
    blk_trsm(L, d, p0, r0, lds);
}

__global__ void __launch_bounds__(TF_NT) k_factor_update(double *L, double *U, const FrontDesc *fd,
                                                         const int *list, const int64_t *pref,
                                                         int cnt, int p0, const double *Dn) {
    __shared__ double lds[2 * TF_PB * TF_TB];
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    int ti, tj;
    tri_index((int64_t)blockIdx.x - pref[i], ti, tj);
    blk_update_tile(L, U, d, p0, Dn, ti, tj, lds);
}

// ---- host launchers --------------------------------------------------------
namespace tf {

void launch_fill0(double *p, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill0, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
}
void launch_scatter_p(int64_t nnz, const int64_t *amap, const double *M, const double *K, double cdt,
                      double *L, hipStream_t st) {
    hipLaunchKernelGGL(k_scatter_p, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, nnz, amap,
                       M, K, cdt, L);
}
void launch_scatter_vals(int64_t nnz, const int64_t *amap, const double *va, double *L,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_scatter_vals, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, nnz,
                       amap, va, L);
}
void launch_extend_add(double *L, double *U, const FrontDesc *fd, const int *parent, const int *list,
                       const int64_t *pref, int cnt, int64_t nwg, const int *rel, hipStream_t st) {
    if (nwg <= 0) return;
    hipLaunchKernelGGL(k_extend_add, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, U, fd, parent, list,
                       pref, cnt, rel);
}
void launch_factor_fused(double *L, double *U, const FrontDesc *fd, const int *list, int cnt, double *Dn,
                         int *flag, hipStream_t st) {
    if (cnt <= 0) return;
    hipLaunchKernelGGL(k_factor_fused, dim3(cnt), dim3(TF_NT), 0, st, L, U, fd, list, Dn, flag);
}
void launch_factor_diag(double *L, const FrontDesc *fd, const int *list, int cnt, int p0, double *Dn,
                        int *flag, hipStream_t st) {
    if (cnt <= 0) return;
    hipLaunchKernelGGL(k_factor_diag, dim3(cnt), dim3(TF_NT), 0, st, L, fd, list, p0, Dn, flag);
}
void launch_factor_trsm(double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                        int64_t nwg, int p0, hipStream_t st) {
    if (nwg <= 0) return;
    hipLaunchKernelGGL(k_factor_trsm, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt,
                       p0);
}
void launch_factor_update(double *L, double *U, const FrontDesc *fd, const int *list,
                          const int64_t *pref, int cnt, int64_t nwg, int p0, const double *Dn,
                          hipStream_t st) {
    if (nwg <= 0) return;
    hipLaunchKernelGGL(k_factor_update, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, U, fd, list, pref,
                       cnt, p0, Dn);
}

}  // namespace tf
