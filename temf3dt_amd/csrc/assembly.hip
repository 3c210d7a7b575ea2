// assembly.hip -- Whitney edge-element matrices and their CSR assembly (FP64).
//
// Restates mesh_process (src/FETD_FWD.f90:588-732): per tet the cofactors
// a,b,c,d and signed volume from LU determinants (src/FETD_FWD.f90:245-277,
// det_mkl_3/4 src/FETD_UTILITY.f90:65-112), the sigma-mass matrix f_kind_iso
// (:826-851) and the curl-curl matrix e_kind_iso (:797-824), each entry
// signed by tet_drc and oriented as the reference keeps it
// (tet_edg(j) <= tet_edg(k)).  The CSR values are then gathered per entry
// in the reference's COO order, so the result does not depend on thread
// scheduling.  Also the receiver interpolation vectors n_arr / d_arr
// (:713-731).
//
// Floating-point contraction is off in this file: every product and sum is
// rounded as written, which keeps the values bit-identical to a plain
// IEEE restatement of the same expressions.
#pragma clang fp contract(off)

#include "tf_device.h"

__constant__ int c_ind_tet[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};

// determinant by LU with partial pivoting (first maximum), reciprocal
// scaling, zero pivot -> 0
__device__ double dev_det_lu(double a[4][4], int n) {
    int piv[4];
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = fabs(a[k][k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(a[i][k]) > mx) { mx = fabs(a[i][k]); p = i; }
        piv[k] = p;
        if (a[p][k] != 0.0) {
            if (p != k)
                for (int j = 0; j < n; j++) { double t = a[k][j]; a[k][j] = a[p][j]; a[p][j] = t; }
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

__device__ void dev_cofactors(const double *x, const double *y, const double *z, double *a, double *b,
                              double *c, double *d, double *vol) {
    const int others[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    const double sa[4] = {1, -1, 1, -1};
    for (int k = 0; k < 4; k++) {
        const int *o = others[k];
        double m[4][4];
        for (int q = 0; q < 3; q++) { m[q][0] = x[o[q]]; m[q][1] = y[o[q]]; m[q][2] = z[o[q]]; }
        a[k] = dev_det_lu(m, 3) * sa[k];
        for (int q = 0; q < 3; q++) { m[q][0] = 1.0; m[q][1] = y[o[q]]; m[q][2] = z[o[q]]; }
        b[k] = dev_det_lu(m, 3) * (-sa[k]);
        for (int q = 0; q < 3; q++) { m[q][0] = 1.0; m[q][1] = x[o[q]]; m[q][2] = z[o[q]]; }
        c[k] = dev_det_lu(m, 3) * sa[k];
        for (int q = 0; q < 3; q++) { m[q][0] = 1.0; m[q][1] = x[o[q]]; m[q][2] = y[o[q]]; }
        d[k] = dev_det_lu(m, 3) * (-sa[k]);
    }
    double m4[4][4];
    for (int q = 0; q < 4; q++) { m4[q][0] = 1.0; m4[q][1] = x[q]; m4[q][2] = y[q]; m4[q][3] = z[q]; }
    *vol = dev_det_lu(m4, 4) / 6.0;
}

__device__ void load_tet(const double *xyz, const int *tv, double *x, double *y, double *z) {
    for (int q = 0; q < 4; q++) {
        const double *p = xyz + 3 * (int64_t)(tv[q] - 1);
        x[q] = p[0]; y[q] = p[1]; z[q] = p[2];
    }
}

__device__ double dev_len(const double *x, const double *y, const double *z, int j) {
    int p = c_ind_tet[j][0], q = c_ind_tet[j][1];
    double dx = x[p] - x[q], dy = y[p] - y[q], dz = z[p] - z[q];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// One thread per tet: 21 oriented, drc-signed entries of the sigma-mass and
// curl-curl element matrices.
__global__ void k_element(int nT, const double *xyz, const int *tet_vex, const int *tet_drc,
                          const uint8_t *orient, const double *sigma, const double *inv_mu, double *vm,
                          double *vk) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nT) return;
    double x[4], y[4], z[4], a[4], b[4], c[4], d[4], vol, len[6], g[6][3];
    load_tet(xyz, tet_vex + 4 * (int64_t)it, x, y, z);
    dev_cofactors(x, y, z, a, b, c, d, &vol);
    for (int j = 0; j < 6; j++) len[j] = dev_len(x, y, z, j);
    for (int i = 0; i < 6; i++) {
        int i1 = c_ind_tet[i][0], i2 = c_ind_tet[i][1];
        g[i][0] = c[i1] * d[i2] - c[i2] * d[i1];
        g[i][1] = b[i2] * d[i1] - b[i1] * d[i2];
        g[i][2] = b[i1] * c[i2] - b[i2] * c[i1];
    }
    const double v6 = 6.0 * vol;
    const double v64 = v6 * v6 * v6 * v6;
    const double sm = sigma[it], sk = inv_mu[it];
    const int *drc = tet_drc + 6 * (int64_t)it;
    for (int p = 0; p < 21; p++) {
        const int o = orient[21 * (int64_t)it + p];
        const int ia = o & 15, ib = o >> 4;
        const int im0 = c_ind_tet[ia][0], im1 = c_ind_tet[ia][1];
        const int jn0 = c_ind_tet[ib][0], jn1 = c_ind_tet[ib][1];
        double f00 = b[im0] * b[jn0] + c[im0] * c[jn0] + d[im0] * d[jn0];
        double f01 = b[im0] * b[jn1] + c[im0] * c[jn1] + d[im0] * d[jn1];
        double f10 = b[im1] * b[jn0] + c[im1] * c[jn0] + d[im1] * d[jn0];
        double f11 = b[im1] * b[jn1] + c[im1] * c[jn1] + d[im1] * d[jn1];
        double v = f11 * (double)(1 + (im0 == jn0)) - f01 * (double)(1 + (im1 == jn0)) -
                   f10 * (double)(1 + (im0 == jn1)) + f00 * (double)(1 + (im1 == jn1));
        double F = len[ia] * len[ib] / (720.0 * vol) * v * sm;
        double dot = g[ia][0] * g[ib][0] + g[ia][1] * g[ib][1] + g[ia][2] * g[ib][2];
        double K = 4.0 * len[ia] * len[ib] * vol / v64 * dot * sk;
        double sgn = (double)(drc[ia] * drc[ib]);
        vm[21 * (int64_t)it + p] = F * sgn;
        vk[21 * (int64_t)it + p] = K * sgn;
    }
}

// Anisotropic media (f_kind_ani src/FETD_FWD.f90:884-912, e_kind_ani
// :853-882): per tet a conductivity tensor As (between the gradients of the
// Whitney mass) and a curl weight tensor Ak (between the curls), row-major
// [nT][9].  Same 21 oriented, drc-signed entries as k_element.
__global__ void k_element_ani(int nT, const double *xyz, const int *tet_vex, const int *tet_drc,
                              const uint8_t *orient, const double *As, const double *Ak, double *vm, double *vk) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nT) return;
    double x[4], y[4], z[4], a[4], b[4], c[4], d[4], vol, len[6], g[6][3], S[9], N[9];
    load_tet(xyz, tet_vex + 4 * (int64_t)it, x, y, z);
    dev_cofactors(x, y, z, a, b, c, d, &vol);
    for (int j = 0; j < 6; j++) len[j] = dev_len(x, y, z, j);
    for (int q = 0; q < 9; q++) { S[q] = As[9 * (int64_t)it + q]; N[q] = Ak[9 * (int64_t)it + q]; }
    const double v6 = 6.0 * vol;
    for (int i = 0; i < 6; i++) {
        int i1 = c_ind_tet[i][0], i2 = c_ind_tet[i][1];
        const double sc = 2.0 * len[i] / (v6 * v6);
        g[i][0] = (c[i1] * d[i2] - c[i2] * d[i1]) * sc;
        g[i][1] = (b[i2] * d[i1] - b[i1] * d[i2]) * sc;
        g[i][2] = (b[i1] * c[i2] - b[i2] * c[i1]) * sc;
    }
    const int *drc = tet_drc + 6 * (int64_t)it;
    for (int p = 0; p < 21; p++) {
        const int o = orient[21 * (int64_t)it + p];
        const int ia = o & 15, ib = o >> 4;
        const int im[2] = {c_ind_tet[ia][0], c_ind_tet[ia][1]};
        const int jn[2] = {c_ind_tet[ib][0], c_ind_tet[ib][1]};
        double f[2][2];
        for (int mm = 0; mm < 2; mm++)
            for (int nn = 0; nn < 2; nn++) {
                const double v1[3] = {b[im[mm]], c[im[mm]], d[im[mm]]};
                const double v2[3] = {b[jn[nn]], c[jn[nn]], d[jn[nn]]};
                double w[3];
                for (int r = 0; r < 3; r++) w[r] = S[3 * r] * v2[0] + S[3 * r + 1] * v2[1] + S[3 * r + 2] * v2[2];
                f[mm][nn] = v1[0] * w[0] + v1[1] * w[1] + v1[2] * w[2];
            }
        double v = f[1][1] * (double)(1 + (im[0] == jn[0])) - f[0][1] * (double)(1 + (im[1] == jn[0])) -
                   f[1][0] * (double)(1 + (im[0] == jn[1])) + f[0][0] * (double)(1 + (im[1] == jn[1]));
        double F = len[ia] * len[ib] / (vol * 720.0) * v;
        double w[3];
        for (int r = 0; r < 3; r++) w[r] = N[3 * r] * g[ib][0] + N[3 * r + 1] * g[ib][1] + N[3 * r + 2] * g[ib][2];
        double K = vol * (g[ia][0] * w[0] + g[ia][1] * w[1] + g[ia][2] * w[2]);
        double sgn = (double)(drc[ia] * drc[ib]);
        vm[21 * (int64_t)it + p] = F * sgn;
        vk[21 * (int64_t)it + p] = K * sgn;
    }
}

// CSR entry q = sum of its element contributions in COO (tet) order.
__global__ void k_gather(int64_t nnz, const int64_t *cptr, const int *ctr, const double *vm, const double *vk,
                         double *M, double *K) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    int64_t s = cptr[q], e = cptr[q + 1];
    double am = vm[ctr[s]], ak = vk[ctr[s]];
    for (int64_t u = s + 1; u < e; u++) {
        am += vm[ctr[u]];
        ak += vk[ctr[u]];
    }
    M[q] = am;
    K[q] = ak;
}

// receiver vectors (src/FETD_FWD.f90:713-731): one thread per receiver
__global__ void k_recvec(int nR, const double *xyz, const int *tet_vex, const int *rec_tet,
                         const double *rec_xyz, double *n_arr, double *d_arr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nR) return;
    int it = rec_tet[i] - 1;
    double x[4], y[4], z[4], a[4], b[4], c[4], d[4], vol;
    load_tet(xyz, tet_vex + 4 * (int64_t)it, x, y, z);
    dev_cofactors(x, y, z, a, b, c, d, &vol);
    const double *p = rec_xyz + 3 * i;
    const double v6 = 6.0 * vol;
    for (int j = 0; j < 6; j++) {
        int j1 = c_ind_tet[j][0], j2 = c_ind_tet[j][1];
        double L1 = a[j1] + b[j1] * p[0] + c[j1] * p[1] + d[j1] * p[2];
        double L2 = a[j2] + b[j2] * p[0] + c[j2] * p[1] + d[j2] * p[2];
        double sc = dev_len(x, y, z, j) / (v6 * v6);
        double *na = n_arr + 18 * i + 3 * j, *da = d_arr + 18 * i + 3 * j;
        na[0] = (L1 * b[j2] - L2 * b[j1]) * sc;
        na[1] = (L1 * c[j2] - L2 * c[j1]) * sc;
        na[2] = (L1 * d[j2] - L2 * d[j1]) * sc;
        da[0] = (-2.0) * (c[j1] * d[j2] - c[j2] * d[j1]) * sc;
        da[1] = (-2.0) * (b[j2] * d[j1] - b[j1] * d[j2]) * sc;
        da[2] = (-2.0) * (b[j1] * c[j2] - b[j2] * c[j1]) * sc;
    }
}

// Per-tet fields at the centroid for the field slices (vtk_field_save,
// src/FETD_FWD.f90:1659-1711), one thread per tet: E = sum_j N_j drc_j e_j,
// curl E = sum_j curl N_j drc_j e_j (+2 l_j grad L1 x grad L2 / (6V)^2, the
// opposite sign of the receivers' d_arr) and the running sum B += curl E * dt
// (B is cleared first on the step-0 row).  `out` ([T][9]: E, curl E, B) is
// written only on a row that takes a slice.
__global__ void k_slice(int nT, const double *xyz, const int *tet_vex, const int *tet_edg, const int *tet_drc,
                        const double *e, double dt, int zero_b, double *B, double *out) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nT) return;
    double x[4], y[4], z[4], a[4], b[4], c[4], d[4], vol;
    load_tet(xyz, tet_vex + 4 * (int64_t)it, x, y, z);
    dev_cofactors(x, y, z, a, b, c, d, &vol);
    const double cp[3] = {(x[0] + x[1] + x[2] + x[3]) / 4.0, (y[0] + y[1] + y[2] + y[3]) / 4.0,
                          (z[0] + z[1] + z[2] + z[3]) / 4.0};
    const double v6 = 6.0 * vol, v62 = v6 * v6;
    double ef[3] = {0.0, 0.0, 0.0}, df[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 6; j++) {
        const int j1 = c_ind_tet[j][0], j2 = c_ind_tet[j][1];
        const double L1 = a[j1] + b[j1] * cp[0] + c[j1] * cp[1] + d[j1] * cp[2];
        const double L2 = a[j2] + b[j2] * cp[0] + c[j2] * cp[1] + d[j2] * cp[2];
        const double len = dev_len(x, y, z, j);
        const double nv[3] = {(L1 * b[j2] - L2 * b[j1]) * len / v62, (L1 * c[j2] - L2 * c[j1]) * len / v62,
                              (L1 * d[j2] - L2 * d[j1]) * len / v62};
        const double dv[3] = {2.0 * (c[j1] * d[j2] - c[j2] * d[j1]) * len / v62,
                              2.0 * (b[j2] * d[j1] - b[j1] * d[j2]) * len / v62,
                              2.0 * (b[j1] * c[j2] - b[j2] * c[j1]) * len / v62};
        const double ev = e[tet_edg[6 * (int64_t)it + j]];
        const double drc = (double)tet_drc[6 * (int64_t)it + j];
        for (int k = 0; k < 3; k++) {
            ef[k] += nv[k] * drc * ev;
            df[k] += dv[k] * drc * ev;
        }
    }
    double *bp = B + 3 * (int64_t)it;
    for (int k = 0; k < 3; k++) bp[k] = (zero_b ? 0.0 : bp[k]) + df[k] * dt;
    if (out) {
        double *o = out + 9 * (int64_t)it;
        for (int k = 0; k < 3; k++) {
            o[k] = ef[k];
            o[3 + k] = df[k];
            o[6 + k] = bp[k];
        }
    }
}

namespace tf {
void launch_slice(int nT, const double *xyz, const int *tet_vex, const int *tet_edg, const int *tet_drc,
                  const double *e, double dt, int zero_b, double *B, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_slice, dim3((nT + 127) / 128), dim3(128), 0, st, nT, xyz, tet_vex, tet_edg, tet_drc, e, dt,
                       zero_b, B, out);
}
void launch_element(int nT, const double *xyz, const int *tet_vex, const int *tet_drc, const uint8_t *orient,
                    const double *sigma, const double *inv_mu, double *vm, double *vk, hipStream_t st) {
    hipLaunchKernelGGL(k_element, dim3((nT + 127) / 128), dim3(128), 0, st, nT, xyz, tet_vex, tet_drc, orient,
                       sigma, inv_mu, vm, vk);
}
void launch_element_ani(int nT, const double *xyz, const int *tet_vex, const int *tet_drc, const uint8_t *orient,
                        const double *As, const double *Ak, double *vm, double *vk, hipStream_t st) {
    hipLaunchKernelGGL(k_element_ani, dim3((nT + 127) / 128), dim3(128), 0, st, nT, xyz, tet_vex, tet_drc, orient,
                       As, Ak, vm, vk);
}
void launch_gather(int64_t nnz, const int64_t *cptr, const int *ctr, const double *vm, const double *vk,
                   double *M, double *K, hipStream_t st) {
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, nnz, cptr, ctr, vm, vk,
                       M, K);
}
void launch_recvec(int nR, const double *xyz, const int *tet_vex, const int *rec_tet, const double *rec_xyz,
                   double *n_arr, double *d_arr, hipStream_t st) {
    hipLaunchKernelGGL(k_recvec, dim3((nR + 63) / 64), dim3(64), 0, st, nR, xyz, tet_vex, rec_tet, rec_xyz,
                       n_arr, d_arr);
}
}  // namespace tf
