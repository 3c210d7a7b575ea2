// step.hip -- per-time-step kernels: right-hand side, receiver readout and the
// doubling-check norms (FP64).
//
// Right-hand side of the BDF2 step (src/FETD_FWD.f90:1425-1428):
//   r = Msigma (4 e2 - e3) - 2 dt s(t),
// where s is the loop source of sys_right (:1466-1511).  Msigma is kept as a
// full symmetric CSR so each row is one independent dot product.
// Readout (solution_output, :1532-1541):
//   E_r = sum_j n_arr(r,j,:) drc_j e(tet_edg_j),  dB/dt_r = sum_j d_arr(...).
#include "tf_device.h"

// one wavefront-quarter (16 lanes) per row
__global__ void __launch_bounds__(256) k_spmv_rhs(int n, const int *ia, const int *ja, const double *val,
                                                  const double *e2, const double *e3, double *r) {
    const int lane = threadIdx.x & 15;
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    double acc = 0.0;
    if (row < n) {
        for (int q = ia[row] + lane; q < ia[row + 1]; q += 16) {
            int c = ja[q];
            acc += val[q] * (4.0 * e2[c] - e3[c]);
        }
    }
    for (int off = 8; off > 0; off >>= 1) acc += __shfl_down(acc, off, 16);
    if (row < n && lane == 0) r[row] = acc;
}

// r[e] -= (2 dt) * (didt * w_e) on the source edges
__global__ void k_source(int ns, const int *edge, const double *w, double didt, double cdt, double *r) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) {
        double s = didt * w[i];
        r[edge[i]] = r[edge[i]] - cdt * s;
    }
}

// out[r*6 + 0..2] = E, out[r*6 + 3..5] = dB/dt at receiver r
__global__ void k_readout(int nR, const int *rec_edge, const double *rec_drc, const double *n_arr,
                          const double *d_arr, const double *e, double *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nR) return;
    double ep[3] = {0.0, 0.0, 0.0}, dp[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 6; j++) {
        double ev = e[rec_edge[6 * i + j]];
        double drc = rec_drc[6 * i + j];
        for (int k = 0; k < 3; k++) {
            ep[k] += n_arr[18 * i + 3 * j + k] * drc * ev;
            dp[k] += d_arr[18 * i + 3 * j + k] * drc * ev;
        }
    }
    for (int k = 0; k < 3; k++) {
        out[6 * i + k] = ep[k];
        out[6 * i + 3 + k] = dp[k];
    }
}

// partial sums of (a-b)^2 and b^2 per block
__global__ void __launch_bounds__(256) k_diffnorm(int64_t n, const double *a, const double *b, double *part) {
    __shared__ double s1[256], s2[256];
    double x1 = 0.0, x2 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double d = a[i] - b[i];
        x1 += d * d;
        x2 += b[i] * b[i];
    }
    s1[threadIdx.x] = x1;
    s2[threadIdx.x] = x2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1[0];
        part[2 * blockIdx.x + 1] = s2[0];
    }
}

__global__ void k_reduce2(int nb, const double *part, double *out) {
    __shared__ double s1[256], s2[256];
    double x1 = 0.0, x2 = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        x1 += part[2 * i];
        x2 += part[2 * i + 1];
    }
    s1[threadIdx.x] = x1;
    s2[threadIdx.x] = x2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s1[0];
        out[1] = s2[0];
    }
}

__global__ void k_gather_vals(int64_t n, const int64_t *map, const double *src, double *dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[map[i]];
}

// ---- independent models in one context ---------------------------------------

// sys_right for several models: r[edge] -= (2 dt_m) (dI/dt(t_m) w) for
// every wire entry, with dt_m and t_m of the entry's model; coef holds
// (dI/dt, 2 dt) per model.  Same rounding as the single-model k_source.
__global__ void k_source_models(int ns, const int *edge, const int *model, const double *w, const double *coef,
                                double *r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) {
        const int e = edge[i], m = model[i];
        const double s = coef[2 * m] * w[i];
        r[e] = r[e] - coef[2 * m + 1] * s;
    }
}

// P = 3 Msigma + 2 dt_m K into each model's target factor slot:
// L[amap[q] + slot_off[model]] with the entry's model
__global__ void k_scatter_p_models(int64_t nnz, const int64_t *amap, const int *qmodel, const double *M,
                                   const double *K, const double *cdt, const int64_t *slot_off, double *L) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nnz) {
        const int m = qmodel[q];
        L[amap[q] + slot_off[m]] = cdt[m] * K[q] + 3.0 * M[q];
    }
}

// per-model sums of (a-b)^2 and b^2 over the model's edge range: one
// workgroup per model, fixed reduction order
__global__ void __launch_bounds__(256) k_range_norms(const int *range, const double *a, const double *b, double *out) {
    __shared__ double s1[256], s2[256];
    const int m = blockIdx.x;
    const int e0 = range[2 * m], e1 = range[2 * m + 1];
    double x1 = 0.0, x2 = 0.0;
    for (int e = e0 + threadIdx.x; e < e1; e += 256) {
        const double d = a[e] - b[e];
        x1 += d * d;
        x2 += b[e] * b[e];
    }
    s1[threadIdx.x] = x1;
    s2[threadIdx.x] = x2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            s1[threadIdx.x] += s1[threadIdx.x + h];
            s2[threadIdx.x] += s2[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * m] = s1[0];
        out[2 * m + 1] = s2[0];
    }
}

namespace tf {
void launch_source_models(int ns, const int *edge, const int *model, const double *w, const double *coef, double *r,
                          hipStream_t st) {
    if (ns > 0) hipLaunchKernelGGL(k_source_models, dim3((ns + 63) / 64), dim3(64), 0, st, ns, edge, model, w, coef, r);
}
void launch_scatter_p_models(int64_t nnz, const int64_t *amap, const int *qmodel, const double *M, const double *K,
                             const double *cdt, const int64_t *slot_off, double *L, hipStream_t st) {
    hipLaunchKernelGGL(k_scatter_p_models, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, nnz, amap, qmodel, M,
                       K, cdt, slot_off, L);
}
void launch_range_norms(int nm, const int *range, const double *a, const double *b, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_range_norms, dim3(nm), dim3(256), 0, st, range, a, b, out);
}
void launch_spmv_rhs(int n, const int *ia, const int *ja, const double *val, const double *e2, const double *e3,
                     double *r, hipStream_t st) {
    int64_t threads = (int64_t)n * 16;
    hipLaunchKernelGGL(k_spmv_rhs, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, n, ia, ja, val, e2,
                       e3, r);
}
void launch_source(int ns, const int *edge, const double *w, double didt, double cdt, double *r, hipStream_t st) {
    if (ns > 0) hipLaunchKernelGGL(k_source, dim3((ns + 63) / 64), dim3(64), 0, st, ns, edge, w, didt, cdt, r);
}
void launch_readout(int nR, const int *rec_edge, const double *rec_drc, const double *n_arr, const double *d_arr,
                    const double *e, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_readout, dim3((nR + 63) / 64), dim3(64), 0, st, nR, rec_edge, rec_drc, n_arr, d_arr, e,
                       out);
}
void launch_diffnorm(int64_t n, const double *a, const double *b, double *part, double *out, hipStream_t st) {
    const int nb = 512;
    hipLaunchKernelGGL(k_diffnorm, dim3(nb), dim3(256), 0, st, n, a, b, part);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, st, nb, part, out);
}
void launch_gather_vals(int64_t n, const int64_t *map, const double *src, double *dst, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL(k_gather_vals, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, map, src, dst);
}
}  // namespace tf
