// solve.hip -- forward / diagonal / backward substitution with the
// multifrontal LDL^T factors (FP64), one right-hand side.
//
// Replaces PARDISO phase 33 (pardiso_direct_2, src/FETD_UTILITY.f90:938-957;
// called per time step at src/FETD_FWD.f90:1334,1371,1431).
//
// Every front's unit-lower diagonal block L11 is stored inverted (X11 =
// L11^-1, computed once per factorisation), so each front's solve is a pair
// of matrix-vector products instead of a sequential substitution:
//   forward  (bottom-up):  v = [b_own; 0] + children's update vectors
//                          y_own = X11 v_own,  v_struct -= L21 y_own
//   backward (top-down):   t = D^-1 y_own - L21^T x_struct,  x_own = X11^T t
// The forward pass hands every front's update vector to its parent
// (extend-add without atomics, so results do not depend on scheduling).
// Vectors Y (forward result) and T are kept in the factor's column order;
// b and x are in edge order.
#include "tf_device.h"

#include <mutex>
#include <vector>

// ---- forward ------------------------------------------------------------------

// vf = [b(own columns); 0] + extend-add of the children's update vectors
__device__ void blk_fwd_gather(const FrontDesc *fd, int s, const double *b, const int *perm, const int *chp,
                               const int *ch, const int *rel, double *V) {
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

// y[r] (rows r0..r0+63 of the front) = sum_j A[r][j] x[j] over columns
// [0, jend(r)), A column-major with leading dimension f: 64 rows x 4 column
// phases per workgroup, phases reduced through LDS.
__device__ double blk_gemv_rows(const double *A, int f, int r, int jend, const double *x, double *red) {
    const int ph = threadIdx.x >> 6;
    double acc = 0.0;
    int j = ph;
    for (; j + 12 < jend; j += 16) {
        double a0 = A[(int64_t)j * f + r], a1 = A[(int64_t)(j + 4) * f + r];
        double a2 = A[(int64_t)(j + 8) * f + r], a3 = A[(int64_t)(j + 12) * f + r];
        acc += a0 * x[j] + a1 * x[j + 4] + a2 * x[j + 8] + a3 * x[j + 12];
    }
    for (; j < jend; j += 4) acc += A[(int64_t)j * f + r] * x[j];
    red[threadIdx.x] = acc;
    __syncthreads();
    double s = 0.0;
    if (ph == 0) s = red[threadIdx.x] + red[threadIdx.x + 64] + red[threadIdx.x + 128] + red[threadIdx.x + 192];
    __syncthreads();
    return s;
}

// ---- kernels: small fronts (nc <= 64), one workgroup per front ------------------

__global__ void __launch_bounds__(TF_NT) k_fwd_small(const double *L, const FrontDesc *fd, const int *list,
                                                     const double *b, const int *perm, const int *chp,
                                                     const int *ch, const int *rel, double *V, double *Y) {
    __shared__ double xs[TF_PB];
    __shared__ double red[TF_NT];
    const int s = list[blockIdx.x];
    blk_fwd_gather(fd, s, b, perm, chp, ch, rel, V);
    const FrontDesc d = fd[s];
    const double *A = L + d.loff;
    double *vf = V + d.voff;
    for (int i = threadIdx.x; i < d.nc; i += blockDim.x) xs[i] = vf[i];
    __syncthreads();
    // y_own = X11 v_own (unit lower): rows < nc
    {
        const int r = threadIdx.x & 63;
        double acc = blk_gemv_rows(A, d.ld, r, r < d.nc ? r : 0, xs, red);
        if (threadIdx.x < d.nc) {
            double y = xs[r] + acc;
            vf[r] = y;
            Y[d.c0 + r] = y;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < d.nc; i += blockDim.x) xs[i] = vf[i];
    __syncthreads();
    // v_struct -= L21 y_own
    for (int r0 = d.nc; r0 < d.f; r0 += 64) {
        const int r = r0 + (threadIdx.x & 63);
        const int rr = r < d.f ? r : d.f - 1;
        double acc = blk_gemv_rows(A, d.ld, rr, d.nc, xs, red);
        if (threadIdx.x < 64 && r < d.f) vf[r] -= acc;
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// t = D^-1 y_own - L21^T x_struct ; x_own = X11^T t, for small fronts
__global__ void __launch_bounds__(TF_NT) k_bwd_small(const double *L, const FrontDesc *fd, const int *list,
                                                     const double *Y, const double *Dn, const int *sidx,
                                                     double *xw) {
    __shared__ double ws[TF_NT];
    __shared__ double ts[TF_PB];
    const FrontDesc d = fd[list[blockIdx.x]];
    const double *A = L + d.loff;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[16];
#pragma unroll
    for (int k = 0; k < 16; k++) acc[k] = 0.0;
    for (int r0 = d.nc; r0 < d.f; r0 += TF_NT) {
        const int r = r0 + threadIdx.x;
        ws[threadIdx.x] = (r < d.f) ? xw[sidx[d.soff + r - d.nc]] : 0.0;
        __syncthreads();
        for (int c = 0; c < 4; c++) {
            const int rr = r0 + 64 * c + lane;
            if (rr < d.f) {
                const double w = ws[64 * c + lane];
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    const int j = wave * 16 + k;
                    if (j < d.nc) acc[k] += A[(int64_t)j * d.ld + rr] * w;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 16; k++) {
        double s = wave_sum(acc[k]);
        const int j = wave * 16 + k;
        if (lane == 0 && j < d.nc) ts[j] = Y[d.c0 + j] / Dn[d.doff + j] - s;
    }
    __syncthreads();
    // x_j = t_j + sum_{i>j} X[i][j] t_i
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int j = wave * 16 + k;
        double v = 0.0;
        if (j < d.nc && lane > j && lane < d.nc) v = A[(int64_t)j * d.ld + lane] * ts[lane];
        v = wave_sum(v);
        if (lane == 0 && j < d.nc) xw[d.c0 + j] = ts[j] + v;
    }
}

// ---- kernels: large fronts, several workgroups per front --------------------------

// y_own rows [64w, 64w+64): y = v + sum_{j<r} X[r][j] v[j]
__global__ void __launch_bounds__(TF_NT) k_fwd_own(const double *L, const FrontDesc *fd, const int *list,
                                                   const int64_t *pref, int cnt, const double *V, double *Y) {
    __shared__ double red[TF_NT];
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int r = (int)(blockIdx.x - pref[i]) * 64 + (threadIdx.x & 63);
    const int rr = r < d.nc ? r : d.nc - 1;
    const double *vf = V + d.voff;
    double acc = blk_gemv_rows(L + d.loff, d.ld, rr, rr, vf, red);
    if (threadIdx.x < 64 && r < d.nc) Y[d.c0 + r] = vf[r] + acc;
}

// v_struct rows: v[nc + r] -= sum_j L21[r][j] y_own[j]
__global__ void __launch_bounds__(TF_NT) k_fwd_struct(const double *L, const FrontDesc *fd, const int *list,
                                                      const int64_t *pref, int cnt, double *V, const double *Y) {
    __shared__ double red[TF_NT];
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int r = d.nc + (int)(blockIdx.x - pref[i]) * 64 + (threadIdx.x & 63);
    const int rr = r < d.f ? r : d.f - 1;
    double acc = blk_gemv_rows(L + d.loff, d.ld, rr, d.nc, Y + d.c0, red);
    if (threadIdx.x < 64 && r < d.f) V[d.voff + r] -= acc;
}

// dot(col[lo:hi], x[lo:hi]) by one wave, four independent accumulators
__device__ __forceinline__ double wave_dot(const double *col, const double *x, int lo, int hi, int lane) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int r = lo + lane;
    for (; r + 192 < hi; r += 256) {
        a0 += col[r] * x[r];
        a1 += col[r + 64] * x[r + 64];
        a2 += col[r + 128] * x[r + 128];
        a3 += col[r + 192] * x[r + 192];
    }
    for (; r < hi; r += 64) a0 += col[r] * x[r];
    return wave_sum((a0 + a1) + (a2 + a3));
}

// t[j] = y[j]/d_j - L21[:,j] . x_struct, one wave per column, 16 columns
// per workgroup
__global__ void __launch_bounds__(TF_NT) k_bwd_t(const double *L, const FrontDesc *fd, const int *list,
                                                 const int64_t *pref, int cnt, const double *Y, const double *Dn,
                                                 const double *V, double *T) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double *xs = V + d.voff;  // rows >= nc hold x_struct
    for (int q = 0; q < 4; q++) {
        const int j = (int)(blockIdx.x - pref[i]) * 16 + wave * 4 + q;
        if (j >= d.nc) break;
        const double *col = L + d.loff + (int64_t)j * d.ld;
        double s = wave_dot(col, xs, d.nc, d.f, lane);
        if (lane == 0) T[d.c0 + j] = Y[d.c0 + j] / Dn[d.doff + j] - s;
    }
}

// x[j] = t[j] + X11[j+1:nc, j] . t[j+1:nc], one wave per column
__global__ void __launch_bounds__(TF_NT) k_bwd_x(const double *L, const FrontDesc *fd, const int *list,
                                                 const int64_t *pref, int cnt, const double *T, double *xw) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double *t = T + d.c0;
    for (int q = 0; q < 4; q++) {
        const int j = (int)(blockIdx.x - pref[i]) * 16 + wave * 4 + q;
        if (j >= d.nc) break;
        const double *col = L + d.loff + (int64_t)j * d.ld;
        double s = wave_dot(col, t, j + 1, d.nc, lane);
        if (lane == 0) xw[d.c0 + j] = t[j] + s;
    }
}

__global__ void k_unpermute(int n, const int *perm, const double *xw, double *x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[perm[i]] = xw[i];
}

// ---- workspace of the allocation-free solve ---------------------------------------
// The work vectors of tf_solve_noalloc (V, Y, T and the permuted solution)
// live in this file-scope array: TF_NOALLOC_DOUBLES doubles on every GPU the
// library runs on, held until the process ends, shared by every call on that
// GPU (calls that overlap on one GPU race; every vector is written before it
// is read, so nothing needs clearing between calls).
static __device__ __align__(256) double scratch[TF_NOALLOC_DOUBLES];

// ---- host launchers ----------------------------------------------------------
namespace tf {

// Device address of this GPU's scratch array, resolved once per device with
// that device current (the name `scratch` is only a host-side stand-in);
// nullptr and the HIP error if the lookup fails.
hipError_t noalloc_scratch(int device, double **ptr) {
    static std::mutex mu;
    static std::vector<double *> addr;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) return hipErrorInvalidDevice;
    if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
    if (!addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            *ptr = nullptr;
            return e;
        }
        addr[device] = static_cast<double *>(p);
    }
    *ptr = addr[device];
    return hipSuccess;
}

void launch_fwd_small(const double *L, const FrontDesc *fd, const int *list, int cnt, const double *b,
                      const int *perm, const int *chp, const int *ch, const int *rel, double *V, double *Y,
                      hipStream_t st) {
    if (cnt > 0)
        hipLaunchKernelGGL(k_fwd_small, dim3(cnt), dim3(TF_NT), 0, st, L, fd, list, b, perm, chp, ch, rel, V, Y);
}
void launch_bwd_small(const double *L, const FrontDesc *fd, const int *list, int cnt, const double *Y,
                      const double *Dn, const int *sidx, double *xw, hipStream_t st) {
    if (cnt > 0) hipLaunchKernelGGL(k_bwd_small, dim3(cnt), dim3(TF_NT), 0, st, L, fd, list, Y, Dn, sidx, xw);
}
void launch_fwd_own(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                    int64_t nwg, const double *V, double *Y, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_fwd_own, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, V, Y);
}
void launch_fwd_struct(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                       int64_t nwg, double *V, const double *Y, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_fwd_struct, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, V, Y);
}
void launch_bwd_t(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                  const double *Y, const double *Dn, const double *V, double *T, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_bwd_t, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, Y, Dn, V, T);
}
void launch_bwd_x(const double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                  const double *T, double *xw, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_bwd_x, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, T, xw);
}
void launch_unpermute(int n, const int *perm, const double *xw, double *x, hipStream_t st) {
    hipLaunchKernelGGL(k_unpermute, dim3((n + 255) / 256), dim3(256), 0, st, n, perm, xw, x);
}

}  // namespace tf
