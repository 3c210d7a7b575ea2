// solve.hip -- forward / diagonal / backward substitution with the
// multifrontal LDL^T factors (FP64), one right-hand side.
//
// Replaces PARDISO phase 33 (pardiso_direct_2, src/FETD_UTILITY.f90:938-957;
// called per time step at src/FETD_FWD.f90:1334,1371,1431).  The forward
// pass runs bottom-up over the front tree and hands every front's update
// vector to its parent (extend-add, no atomics, so results do not depend on
// scheduling); the backward pass runs top-down and reads the already solved
// rows of ancestor fronts.  Input b and output x are in edge order; the
// permutation is applied while gathering/scattering.
#include "tf_device.h"

#define LDS_LD 65

// ---- forward ------------------------------------------------------------------

// vf = [b(own columns); 0] + extend-add of the children's update vectors
__device__ void blk_fwd_gather(const FrontDesc *fd, int s, const double *b, const int *perm,
                               const int *chp, const int *ch, const int *rel, double *V) {
    const FrontDesc d = fd[s];
    double *vf = V + d.voff;
    for (int j = threadIdx.x; j < d.f; j += blockDim.x) vf[j] = (j < d.nc) ? b[perm[d.c0 + j]] : 0.0;
    __syncthreads();
    for (int q = chp[s]; q < chp[s + 1]; q++) {
        const FrontDesc dc = fd[ch[q]];
        const double *uc = V + dc.voff + dc.nc;
        const int *rc = rel + dc.soff;
        for (int a = threadIdx.x; a < dc.m; a += blockDim.x) vf[rc[a]] += uc[a];
        __syncthreads();
    }
}

// unit-lower solve of the panel's diagonal block
__device__ void blk_fwd_diag(const double *L, const FrontDesc &d, int p0, double *V, double *xs) {
    const int pw = min(TF_PB, d.nc - p0);
    const double *base = L + d.loff;
    double *vf = V + d.voff;
    for (int i = threadIdx.x; i < pw; i += blockDim.x) xs[i] = vf[p0 + i];
    __syncthreads();
    for (int j = 0; j < pw; j++) {
        const double xj = xs[j];
        const double *col = base + (int64_t)(p0 + j) * d.f + p0;
        for (int i = j + 1 + threadIdx.x; i < pw; i += blockDim.x) xs[i] -= col[i] * xj;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < pw; i += blockDim.x) vf[p0 + i] = xs[i];
    __syncthreads();
}

// vf[r] -= L[r, panel] * vf[panel] for rows [r0, r0+NT)
__device__ void blk_fwd_update(const double *L, const FrontDesc &d, int p0, int r0, double *V,
                               double *xs) {
    const int pw = min(TF_PB, d.nc - p0);
    const double *base = L + d.loff;
    double *vf = V + d.voff;
    for (int i = threadIdx.x; i < pw; i += blockDim.x) xs[i] = vf[p0 + i];
    __syncthreads();
    const int r = r0 + threadIdx.x;
    if (r < d.f) {
        double acc = 0.0;
        for (int j = 0; j < pw; j++) acc += base[(int64_t)(p0 + j) * d.f + r] * xs[j];
        vf[r] -= acc;
    }
    __syncthreads();
}

// ---- backward ----------------------------------------------------------------

__device__ __forceinline__ double xrow(const FrontDesc &d, int r, const double *xw, const int *sidx) {
    return (r < d.nc) ? xw[d.c0 + r] : xw[sidx[d.soff + r - d.nc]];
}

// partial[j] = sum over rows [r0, r0+NT) of L[r, p0+j] * x(r)
__device__ void blk_bwd_partial(const double *L, const FrontDesc &d, int p0, int r0, const double *xw,
                                const int *sidx, double *out, double *lds) {
    const int pw = min(TF_PB, d.nc - p0);
    const int p1 = p0 + pw;
    const double *base = L + d.loff;
    double *xr = lds;  // [TF_NT]
    const int r = r0 + threadIdx.x;
    xr[threadIdx.x] = (r < d.f && r >= p1) ? xrow(d, r, xw, sidx) : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    for (int j = wave; j < pw; j += nw) {
        const double *col = base + (int64_t)(p0 + j) * d.f;
        double acc = 0.0;
        for (int t = lane; t < TF_NT; t += 64) {
            int rr = r0 + t;
            if (rr < d.f && rr >= p1) acc += col[rr] * xr[t];
        }
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if (lane == 0) out[j] = acc;
    }
    __syncthreads();
}

// x[panel] = L11^-T (D^-1 y[panel] - sum of partials)
__device__ void blk_bwd_diag(const double *L, const FrontDesc &d, int p0, const double *V,
                             const double *Dn, const double *partial, int nparts, double *xw,
                             double *lds) {
    const int pw = min(TF_PB, d.nc - p0);
    const double *base = L + d.loff;
    const double *vf = V + d.voff;
    double *L11 = lds;                  // [64][65]
    double *t = lds + TF_PB * LDS_LD;   // [64]
    for (int idx = threadIdx.x; idx < pw * pw; idx += blockDim.x) {
        int i = idx % pw, j = idx / pw;
        L11[j * LDS_LD + i] = (i > j) ? base[(int64_t)(p0 + j) * d.f + p0 + i] : 0.0;
    }
    for (int j = threadIdx.x; j < pw; j += blockDim.x) {
        double s = vf[p0 + j] / Dn[d.c0 + p0 + j];
        for (int q = 0; q < nparts; q++) s -= partial[(int64_t)q * TF_PB + j];
        t[j] = s;
    }
    __syncthreads();
    // L11^T x = t, reverse column sweep: x_j final, then t_i -= L[j][i] x_j
    for (int j = pw - 1; j >= 0; j--) {
        const double xj = t[j];
        for (int i = threadIdx.x; i < j; i += blockDim.x) t[i] -= L11[i * LDS_LD + j] * xj;
        __syncthreads();
    }
    for (int j = threadIdx.x; j < pw; j += blockDim.x) xw[d.c0 + p0 + j] = t[j];
    __syncthreads();
}

// ---- kernels -------------------------------------------------------------------

__global__ void __launch_bounds__(TF_NT) k_fwd_fused(const double *L, const FrontDesc *fd, const int *list,
                                                     const double *b, const int *perm, const int *chp,
                                                     const int *ch, const int *rel, double *V) {
    __shared__ double xs[TF_PB];
    const int s = list[blockIdx.x];
    blk_fwd_gather(fd, s, b, perm, chp, ch, rel, V);
    const FrontDesc d = fd[s];
    for (int p0 = 0; p0 < d.nc; p0 += TF_PB) {
        blk_fwd_diag(L, d, p0, V, xs);
        const int p1 = min(d.nc, p0 + TF_PB);
        for (int r0 = p1; r0 < d.f; r0 += TF_NT) blk_fwd_update(L, d, p0, r0, V, xs);
    }
}

__global__ void __launch_bounds__(TF_NT) k_fwd_gather(const FrontDesc *fd, const int *list, const double *b,
                                                      const int *perm, const int *chp, const int *ch,
                                                      const int *rel, double *V) {
    blk_fwd_gather(fd, list[blockIdx.x], b, perm, chp, ch, rel, V);
}

__global__ void __launch_bounds__(TF_NT) k_fwd_diag(const double *L, const FrontDesc *fd, const int *list,
                                                    int p0, double *V) {
    __shared__ double xs[TF_PB];
    const FrontDesc d = fd[list[blockIdx.x]];
    blk_fwd_diag(L, d, p0, V, xs);
}

__global__ void __launch_bounds__(TF_NT) k_fwd_update(const double *L, const FrontDesc *fd, const int *list,
                                                      const int64_t *pref, int cnt, int p0, double *V) {
    __shared__ double xs[TF_PB];
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int p1 = min(d.nc, p0 + TF_PB);
    blk_fwd_update(L, d, p0, p1 + (int)(blockIdx.x - pref[i]) * TF_NT, V, xs);
}

__global__ void __launch_bounds__(TF_NT) k_bwd_fused(const double *L, const FrontDesc *fd, const int *list,
                                                     const double *V, const double *Dn, const int *sidx,
                                                     double *xw) {
    __shared__ double lds[TF_PB * LDS_LD + TF_PB];
    __shared__ double xr[TF_NT];
    __shared__ double part[TF_PB];
    const FrontDesc d = fd[list[blockIdx.x]];
    const int npan = (d.nc + TF_PB - 1) / TF_PB;
    for (int pk = npan - 1; pk >= 0; pk--) {
        const int p0 = pk * TF_PB;
        const int pw = min(TF_PB, d.nc - p0);
        const int p1 = p0 + pw;
        for (int j = threadIdx.x; j < TF_PB; j += blockDim.x) part[j] = 0.0;
        __syncthreads();
        for (int r0 = p1; r0 < d.f; r0 += TF_NT) {
            // partial sums for this row chunk, accumulated in `part`
            const int r = r0 + threadIdx.x;
            xr[threadIdx.x] = (r < d.f) ? xrow(d, r, xw, sidx) : 0.0;
            __syncthreads();
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            const int nw = blockDim.x >> 6;
            for (int j = wave; j < pw; j += nw) {
                const double *col = L + d.loff + (int64_t)(p0 + j) * d.f;
                double acc = 0.0;
                for (int t = lane; t < TF_NT; t += 64) {
                    int rr = r0 + t;
                    if (rr < d.f) acc += col[rr] * xr[t];
                }
                for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
                if (lane == 0) part[j] += acc;
            }
            __syncthreads();
        }
        blk_bwd_diag(L, d, p0, V, Dn, part, 1, xw, lds);
    }
}

__global__ void __launch_bounds__(TF_NT) k_bwd_partial(const double *L, const FrontDesc *fd, const int *list,
                                                       const int64_t *pref, int cnt, int p0, const double *xw,
                                                       const int *sidx, double *partial) {
    __shared__ double xr[TF_NT];
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int p1 = min(d.nc, p0 + TF_PB);
    blk_bwd_partial(L, d, p0, p1 + (int)(blockIdx.x - pref[i]) * TF_NT, xw, sidx,
                    partial + (int64_t)blockIdx.x * TF_PB, xr);
}

__global__ void __launch_bounds__(TF_NT) k_bwd_diag(const double *L, const FrontDesc *fd, const int *list,
                                                    const int64_t *pref, int p0, const double *V,
                                                    const double *Dn, const double *partial, double *xw) {
    __shared__ double lds[TF_PB * LDS_LD + TF_PB];
    const FrontDesc d = fd[list[blockIdx.x]];
    const int64_t q0 = pref[blockIdx.x], q1 = pref[blockIdx.x + 1];
    blk_bwd_diag(L, d, p0, V, Dn, partial + q0 * TF_PB, (int)(q1 - q0), xw, lds);
}

__global__ void k_unpermute(int n, const int *perm, const double *xw, double *x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[perm[i]] = xw[i];
}

// ---- host launchers ----------------------------------------------------------
namespace tf {

void launch_fwd_fused(const double *L, const FrontDesc *fd, const int *list, int cnt, const double *b,
                      const int *perm, const int *chp, const int *ch, const int *rel, double *V,
                      hipStream_t st) {
    if (cnt > 0)
        hipLaunchKernelGGL(k_fwd_fused, dim3(cnt), dim3(TF_NT), 0, st, L, fd, list, b, perm, chp, ch, rel, V);
}
void launch_fwd_gather(const FrontDesc *fd, const int *list, int cnt, const double *b, const int *perm,
                       const int *chp, const int *ch, const int *rel, double *V, hipStream_t st) {
    if (cnt > 0)
        hipLaunchKernelGGL(k_fwd_gather, dim3(cnt), dim3(TF_NT), 0, st, fd, list, b, perm, chp, ch, rel, V);
}
void launch_fwd_diag(const double *L, const FrontDesc *fd, const int *list, int cnt, int p0, double *V,
                     hipStream_t st) {
    if (cnt > 0) hipLaunchKernelGGL(k_fwd_diag, dim3(cnt), dim3(TF_NT), 0, st, L, fd, list, p0, V);
}
void launch_fwd_update(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                       int64_t nwg, int p0, double *V, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_fwd_update, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, p0, V);
}
void launch_bwd_fused(const double *L, const FrontDesc *fd, const int *list, int cnt, const double *V,
                      const double *Dn, const int *sidx, double *xw, hipStream_t st) {
    if (cnt > 0)
        hipLaunchKernelGGL(k_bwd_fused, dim3(cnt), dim3(TF_NT), 0, st, L, fd, list, V, Dn, sidx, xw);
}
void launch_bwd_partial(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                        int64_t nwg, int p0, const double *xw, const int *sidx, double *partial,
                        hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_bwd_partial, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, p0,
                           xw, sidx, partial);
}
void launch_bwd_diag(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                     int p0, const double *V, const double *Dn, const double *partial, double *xw,
                     hipStream_t st) {
    if (cnt > 0)
        hipLaunchKernelGGL(k_bwd_diag, dim3(cnt), dim3(TF_NT), 0, st, L, fd, list, pref, p0, V, Dn, partial, xw);
}
void launch_unpermute(int n, const int *perm, const double *xw, double *x, hipStream_t st) {
    hipLaunchKernelGGL(k_unpermute, dim3((n + 255) / 256), dim3(256), 0, st, n, perm, xw, x);
}

}  // namespace tf
