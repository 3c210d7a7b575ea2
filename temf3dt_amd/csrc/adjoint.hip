// adjoint.hip -- kernels of the backward-in-time adjoint of the BDF2 loop.
//
// Forward node k of the loop (src/FETD_FWD.f90:1281-1443) solves
//   P_k y_k = Msigma (4 u_k - v_k) - 2 dt_k s_k,   P_k = 3 Msigma + 2 dt_k K
// with (u_k, v_k) the two history states it was stepped from.  For a
// functional J of the receiver rows the adjoint states solve
//   P_k lambda_k = dJ/dy_k + Msigma (4 A_k - B_k)
// where A_k (B_k) sums the lambda of the nodes that used y_k as u (as v),
// and with Msigma = sum_t sigma_t M1_t (M1_t the unit-conductivity element
// mass matrix of tet t),
//   dJ/dsigma_t = - sum_k lambda_k^T M1_t (3 y_k - 4 u_k + v_k).
// These kernels: the seed dJ/dy_k from the receiver weights, the per-element
// sensitivity reduction, and vector updates.  Vectors are [edge][K].
#include "tf_device.h"

// g[edge][q] += sum over receivers and local edges of the readout
// weights: E weights on n_arr, dB/dt weights on d_arr (src/FETD_FWD.f90:713-731
// and the readout :1532-1541).  w[(q*nR + rec)*6 + 0..5]; one thread per
// source walks the receivers in order, so shared edges add up
// deterministically.
__global__ void k_adj_seed(int nR, int K, const int *rec_edge, const double *rec_drc, const double *n_arr,
                           const double *d_arr, const double *w, double *g) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= K) return;
    for (int r = 0; r < nR; r++) {
        const double *wr = w + ((int64_t)q * nR + r) * 6;
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++)
                s += wr[k] * n_arr[18 * r + 3 * j + k] + wr[3 + k] * d_arr[18 * r + 3 * j + k];
            g[(int64_t)rec_edge[6 * r + j] * K + q] += rec_drc[6 * r + j] * s;
        }
    }
}

// grad[t] -= sum_q lambda^T M1_t (3 y - 4 u + v) over the tet's 21 oriented
// entries (vm1 as written by k_element with unit conductivity: signed by
// tet_drc, pair p couples local edges orient&15 and orient>>4).
__global__ void __launch_bounds__(256) k_sens(int nT, int K, const int *tet_edg, const uint8_t *orient,
                                              const double *vm1, const double *lam, const double *y,
                                              const double *u, const double *v, double *grad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    int e[6];
    for (int j = 0; j < 6; j++) e[j] = tet_edg[6 * (int64_t)t + j];
    double m[21];
    int pa[21], pb[21];
    for (int p = 0; p < 21; p++) {
        m[p] = vm1[21 * (int64_t)t + p];
        const int o = orient[21 * (int64_t)t + p];
        pa[p] = o & 15;
        pb[p] = o >> 4;
    }
    double acc = 0.0;
    for (int q = 0; q < K; q++) {
        double l[6], w[6];
        for (int j = 0; j < 6; j++) {
            const int64_t o = (int64_t)e[j] * K + q;
            l[j] = lam[o];
            w[j] = 3.0 * y[o] - 4.0 * u[o] + v[o];
        }
        for (int p = 0; p < 21; p++) {
            const int a = pa[p], b = pb[p];
            acc += m[p] * (a == b ? l[a] * w[a] : l[a] * w[b] + l[b] * w[a]);
        }
    }
    grad[t] -= acc;
}

// y += a x
__global__ void k_axpy(int64_t len, double a, const double *x, double *y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) y[i] += a * x[i];
}

// g *= s (dJ/dsigma -> dJ/dlog sigma)
__global__ void k_scale(int n, const double *s, double *g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[i] *= s[i];
}

namespace tf {
void launch_adj_seed(int nR, int K, const int *rec_edge, const double *rec_drc, const double *n_arr,
                     const double *d_arr, const double *w, double *g, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_seed, dim3((K + 63) / 64), dim3(64), 0, st, nR, K, rec_edge, rec_drc, n_arr, d_arr, w, g);
}
void launch_sens(int nT, int K, const int *tet_edg, const uint8_t *orient, const double *vm1, const double *lam,
                 const double *y, const double *u, const double *v, double *grad, hipStream_t st) {
    hipLaunchKernelGGL(k_sens, dim3((nT + 255) / 256), dim3(256), 0, st, nT, K, tet_edg, orient, vm1, lam, y, u, v,
                       grad);
}
void launch_axpy(int64_t len, double a, const double *x, double *y, hipStream_t st) {
    hipLaunchKernelGGL(k_axpy, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, len, a, x, y);
}
void launch_scale(int n, const double *s, double *g, hipStream_t st) {
    hipLaunchKernelGGL(k_scale, dim3((n + 255) / 256), dim3(256), 0, st, n, s, g);
}
}  // namespace tf
