// msolve.hip -- K right-hand-side (K <= 64 sources sharing one LDL^T factor)
// vector kernels of the block solve and time step: gathers between fronts,
// permutation, BDF2 history, SpMV, source, readout and the doubling-check
// norms.  The tile products themselves are in mtile.hip.
//
// The reference solves one source per run (pardiso_direct_2 with nrhs=1,
// src/FETD_UTILITY.f90:938-957).  Here K sources are time-stepped together:
// every vector is [row][K] (the K values of a row contiguous), so the factor
// is read once per step for all K sources.
#include "tf_device.h"

// ---- K-wide step kernels ------------------------------------------------------------

// w = 4 e2 - e3 (the BDF2 history combination), once per entry
__global__ void k_mhist(int64_t len, const double *e2, const double *e3, double *w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) w[i] = 4.0 * e2[i] - e3[i];
}

// r[row][q] = sum_j M[row][j] w[j][q], one wave per row at a time (lane = q,
// K <= 64).  The row's column indices and values are loaded once across the
// lanes and broadcast one at a time from a scalar register, so the gathers
// of the rows of w do not wait on each other and eight of them are in flight.
// With xcd, every XCD takes one contiguous range of rows, so that the rows of
// w that neighbouring rows gather (edges of nearby vertices) meet in one L2.
__global__ void __launch_bounds__(256) k_mspmv_rhs(int n, int K, const int *ia, const int *ja, const double *val,
                                                   const double *w, double *r, int xcd) {
    const int lane = threadIdx.x & 63;
    const int blk = xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int row0 = blk * 4 * TF_RPW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int rr = 0; rr < TF_RPW; rr++) {
        const int row = row0 + 4 * rr;
        if (row >= n) return;
        const int p0 = ia[row], p1 = ia[row + 1];
        const int q = min(lane, K - 1);
        double acc = 0.0;
        for (int pb = p0; pb < p1; pb += 64) {
            const int cnt = min(64, p1 - pb);
            const int myj = lane < cnt ? ja[pb + lane] : 0;
            const double myv = lane < cnt ? val[pb + lane] : 0.0;
            const int myjlo = myj;
            const long long myvb = __builtin_bit_cast(long long, myv);
            int t = 0;
            for (; t + 8 <= cnt; t += 8) {
                double x[8], v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int j = __builtin_amdgcn_readlane(myjlo, t + u);
                    const int vlo = __builtin_amdgcn_readlane((int)myvb, t + u);
                    const int vhi = __builtin_amdgcn_readlane((int)(myvb >> 32), t + u);
                    v[u] = __builtin_bit_cast(double, ((long long)vhi << 32) | (unsigned)vlo);
                    x[u] = w[(int64_t)j * K + q];
                }
#pragma unroll
                for (int u = 0; u < 8; u++) acc += v[u] * x[u];
            }
            for (; t < cnt; t++) {
                const int j = __builtin_amdgcn_readlane(myjlo, t);
                const int vlo = __builtin_amdgcn_readlane((int)myvb, t);
                const int vhi = __builtin_amdgcn_readlane((int)(myvb >> 32), t);
                acc += __builtin_bit_cast(double, ((long long)vhi << 32) | (unsigned)vlo) * w[(int64_t)j * K + q];
            }
        }
        if (lane < K) r[(int64_t)row * K + lane] = acc;
    }
}

// The same product with two rows per wave at a time and two right-hand sides
// per lane (K even): lanes 0-31 take one row and lanes 32-63 the next, each
// lane the columns 2j, 2j + 1 of every gathered row of w as one 16-byte load
// (half the load instructions of k_mspmv_rhs).  Each half holds up to 32 of
// its row's column indices and values and shares them by lane shuffles; the
// terms are summed in CSR order, as in k_mspmv_rhs.
__global__ void __launch_bounds__(256) k_mspmv_rhs2(int n, int K, const int *ia, const int *ja, const double *val,
                                                    const double *w, double *r, int xcd) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    constexpr int D = 4;  // gathers in flight per lane
    const int lane = threadIdx.x & 63, h = lane >> 5, j = lane & 31;
    const int q = min(2 * j, K - 2);
    const int blk = xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int row0 = blk * 4 * TF_RPW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int rr = 0; rr < TF_RPW; rr += 2) {
        if (row0 + 4 * rr >= n) return;
        const int row = row0 + 4 * (rr + h);
        const bool live = row < n;
        const int p0 = live ? ia[row] : 0, cnt = live ? ia[row + 1] - p0 : 0;
        const int cm = max(cnt, __shfl_xor(cnt, 32, 64));
        d2 acc = (d2){0.0, 0.0};
        for (int pb = 0; pb < cm; pb += 32) {
            const int myj = pb + j < cnt ? ja[p0 + pb + j] : 0;
            const double myv = pb + j < cnt ? val[p0 + pb + j] : 0.0;
            const int m = min(32, cm - pb);
            for (int t = 0; t < m; t += D) {
                d2 x[D];
                double v[D];
#pragma unroll
                for (int u = 0; u < D; u++) {
                    const int src = (lane & 32) + min(t + u, m - 1);
                    const int c = __shfl(myj, src, 64);
                    v[u] = __shfl(myv, src, 64);
                    x[u] = *(const d2 *)(w + (int64_t)c * K + q);
                }
#pragma unroll
                for (int u = 0; u < D; u++) {  // the shorter row's half skips the terms past its end
                    const d2 a = acc + v[u] * x[u];
                    acc = t + u < m && pb + t + u < cnt ? a : acc;
                }
            }
        }
        if (live && 2 * j < K) *(d2 *)(r + (int64_t)row * K + 2 * j) = acc;
    }
}

// The same product of the newest history state alone, with the BDF2
// combination taken from the product kept for the state before it (the SpMV
// is linear: M (4 e2 - e3) = 4 M e2 - M e3): u2 = M e2 and r = 4 u2 - u3, as
// 16-byte pairs (K even).  u3 was written by this kernel one step earlier
// (or by k_mspmv_rhs2 from e3 after the history was reset).
__global__ void __launch_bounds__(256) k_mspmv_chain(int n, int K, const int *ia, const int *ja, const double *val,
                                                     const double *e2, const double *u3, double *u2, double *r,
                                                     int xcd) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    constexpr int D = 4;
    const int lane = threadIdx.x & 63, h = lane >> 5, j = lane & 31;
    const int q = min(2 * j, K - 2);
    const int blk = xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int row0 = blk * 4 * TF_RPW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int rr = 0; rr < TF_RPW; rr += 2) {
        if (row0 + 4 * rr >= n) return;
        const int row = row0 + 4 * (rr + h);
        const bool live = row < n;
        const int p0 = live ? ia[row] : 0, cnt = live ? ia[row + 1] - p0 : 0;
        const int cm = max(cnt, __shfl_xor(cnt, 32, 64));
        // the older product's row, fetched ahead of the gathers
        const d2 old = live ? *(const d2 *)(u3 + (int64_t)row * K + q) : (d2){0.0, 0.0};
        d2 acc = (d2){0.0, 0.0};
        for (int pb = 0; pb < cm; pb += 32) {
            const int myj = pb + j < cnt ? ja[p0 + pb + j] : 0;
            const double myv = pb + j < cnt ? val[p0 + pb + j] : 0.0;
            const int m = min(32, cm - pb);
            for (int t = 0; t < m; t += D) {
                d2 x[D];
                double v[D];
#pragma unroll
                for (int u = 0; u < D; u++) {
                    const int src = (lane & 32) + min(t + u, m - 1);
                    const int c = __shfl(myj, src, 64);
                    v[u] = __shfl(myv, src, 64);
                    x[u] = *(const d2 *)(e2 + (int64_t)c * K + q);
                }
#pragma unroll
                for (int u = 0; u < D; u++) {
                    const d2 a = acc + v[u] * x[u];
                    acc = t + u < m && pb + t + u < cnt ? a : acc;
                }
            }
        }
        if (live && 2 * j < K) {
            *(d2 *)(u2 + (int64_t)row * K + 2 * j) = acc;
            *(d2 *)(r + (int64_t)row * K + 2 * j) = 4.0 * acc - old;
        }
    }
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

// ---- forward pull -----------------------------------------------------------------
// Own rows of the fronts of one level: Yin[row] = b[perm[row]] - sum of the Z
// rows (descendant updates) that land on row, in the fixed order of gidx.
// One wave per row at a time, TF_RPW rows per wave (lane = right-hand side);
// the row's Z indices are loaded
// once across the lanes and broadcast from scalar registers, so eight Z rows
// are in flight at a time (summed in order).
__global__ void __launch_bounds__(256) k_mpull(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                                               const double *b, const int *perm, const int *gptr, const int *gidx,
                                               const double *Z, double *Yin, int K) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int lane = threadIdx.x & 63;
    const int r0 = (int)(blockIdx.x - pref[i]) * 4 * TF_RPW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int rr = 0; rr < TF_RPW; rr++) {
        const int r = r0 + 4 * rr;
        if (r >= d.nc) return;
        const int row = d.c0 + r, q = min(lane, K - 1);
        double s = b[(int64_t)perm[row] * K + q];
        const int p0 = gptr[row], p1 = gptr[row + 1];
        for (int pb = p0; pb < p1; pb += 64) {
            const int n = min(64, p1 - pb);
            const int mine = lane < n ? gidx[pb + lane] : 0;
            int t = 0;
            for (; t + 8 <= n; t += 8) {
                double z[8];
#pragma unroll
                for (int u = 0; u < 8; u++) z[u] = Z[(int64_t)__builtin_amdgcn_readlane(mine, t + u) * K + q];
#pragma unroll
                for (int u = 0; u < 8; u++) s -= z[u];
            }
            for (; t < n; t++) s -= Z[(int64_t)__builtin_amdgcn_readlane(mine, t) * K + q];
        }
        if (lane < K) Yin[(int64_t)row * K + lane] = s;
    }
}

// The same pull with two rows per wave at a time and two right-hand sides per
// lane (K even): lanes 0-31 take one row and lanes 32-63 the next, each lane
// the columns 2j, 2j + 1 as one 16-byte load of every Z row it subtracts.
// Each half walks its own row's Z indices (held across its 32 lanes and
// fetched with a lane shuffle) in gidx order, so the sums are those of
// k_mpull bit for bit; a half whose row has fewer entries subtracts zeros.
__global__ void __launch_bounds__(256) k_mpull2(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                                                const double *b, const int *perm, const int *gptr, const int *gidx,
                                                const double *Z, double *Yin, int K) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int lane = threadIdx.x & 63, h = lane >> 5, j = lane & 31;
    const int q = min(2 * j, K - 2);  // K even: q, q + 1 < K
    const int r0 = (int)(blockIdx.x - pref[i]) * 4 * TF_RPW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int rr = 0; rr < TF_RPW; rr += 2) {
        const int r = r0 + 4 * (rr + h);
        if (r0 + 4 * rr >= d.nc) return;
        const bool live = r < d.nc;
        const int row = d.c0 + min(r, d.nc - 1);
        d2 s = *(const d2 *)(b + (int64_t)perm[row] * K + q);
        const int p0 = gptr[row], p1 = live ? gptr[row + 1] : p0;
        const int n = p1 - p0;
        const int nmax = max(n, __shfl_xor(n, 32, 64));
        for (int pb = 0; pb < nmax; pb += 32) {
            const int mine = pb + j < n ? gidx[p0 + pb + j] : 0;
            const int m = min(32, nmax - pb);
            int t = 0;
            for (; t + 8 <= m; t += 8) {
                d2 z[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int g = __shfl(mine, (lane & 32) + t + u, 64);
                    const d2 v = *(const d2 *)(Z + (int64_t)g * K + q);
                    z[u] = pb + t + u < n ? v : (d2){0.0, 0.0};
                }
#pragma unroll
                for (int u = 0; u < 8; u++) s -= z[u];
            }
            for (; t < m; t++) {
                const int g = __shfl(mine, (lane & 32) + t, 64);
                const d2 v = *(const d2 *)(Z + (int64_t)g * K + q);
                s -= pb + t < n ? v : (d2){0.0, 0.0};
            }
        }
        if (live && 2 * j < K) *(d2 *)(Yin + (int64_t)row * K + 2 * j) = s;
    }
}

// The same pull for one right-hand side: a thread per own row, 64 rows per
// workgroup, the row's Z entries subtracted in gidx order (the order of
// k_mpull, so one column of a block solve and this give the same bits).
__global__ void __launch_bounds__(64) k_spull(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                                               const double *b, const int *perm, const int *gptr, const int *gidx,
                                               const double *Z, double *Yin) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    const int r = (int)(blockIdx.x - pref[i]) * 64 + threadIdx.x;
    if (r >= d.nc) return;
    const int row = d.c0 + r;
    double s = b[perm[row]];
    const int p1 = gptr[row + 1];
    int p = gptr[row];
    for (; p + 8 <= p1; p += 8) {  // eight index loads, then eight values, in flight
        int g[8];
        double z[8];
#pragma unroll
        for (int u = 0; u < 8; u++) g[u] = gidx[p + u];
#pragma unroll
        for (int u = 0; u < 8; u++) z[u] = Z[g[u]];
#pragma unroll
        for (int u = 0; u < 8; u++) s -= z[u];
    }
    for (; p < p1; p++) s -= Z[gidx[p]];
    Yin[row] = s;
}

// ---- host launchers --------------------------------------------------------------
// deterministic right-hand sides for solve timing
__global__ void k_fill_pattern(int64_t len, double *b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) b[i] = 1.0 + 0.125 * (double)(i % 7) - 0.0625 * (double)(i % 5);
}

namespace tf {
void launch_fill_pattern(int64_t len, double *b, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, len, b);
}
void launch_mpull(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg, const double *b,
                  const int *perm, const int *gptr, const int *gidx, const double *Z, double *Yin, int K,
                  hipStream_t st) {
    if (nwg <= 0) return;
    if (K % 2 == 0)
        hipLaunchKernelGGL(k_mpull2, dim3((unsigned)nwg), dim3(256), 0, st, fd, list, pref, cnt, b, perm, gptr, gidx, Z,
                           Yin, K);
    else
        hipLaunchKernelGGL(k_mpull, dim3((unsigned)nwg), dim3(256), 0, st, fd, list, pref, cnt, b, perm, gptr, gidx, Z, Yin,
                           K);
}
void launch_spull(const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg, const double *b,
                  const int *perm, const int *gptr, const int *gidx, const double *Z, double *Yin, hipStream_t st) {
    if (nwg > 0)
        hipLaunchKernelGGL(k_spull, dim3((unsigned)nwg), dim3(64), 0, st, fd, list, pref, cnt, b, perm, gptr, gidx, Z,
                           Yin);
}
void launch_mspmv_rhs(int n, int K, const int *ia, const int *ja, const double *val, const double *e2,
                      const double *e3, double *w, double *r, int xcd, hipStream_t st) {
    const int64_t tot = (int64_t)n * K;
    hipLaunchKernelGGL(k_mhist, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, tot, e2, e3, w);
    const dim3 g((unsigned)((n + 4 * TF_RPW - 1) / (4 * TF_RPW)));
    if (K % 2 == 0)
        hipLaunchKernelGGL(k_mspmv_rhs2, g, dim3(256), 0, st, n, K, ia, ja, val, w, r, xcd);
    else
        hipLaunchKernelGGL(k_mspmv_rhs, g, dim3(256), 0, st, n, K, ia, ja, val, w, r, xcd);
}
// u = M e (no history combination), K even
void launch_mspmv(int n, int K, const int *ia, const int *ja, const double *val, const double *e, double *u, int xcd,
                  hipStream_t st) {
    hipLaunchKernelGGL(k_mspmv_rhs2, dim3((unsigned)((n + 4 * TF_RPW - 1) / (4 * TF_RPW))), dim3(256), 0, st, n, K,
                       ia, ja, val, e, u, xcd);
}
// u2 = M e2 and r = 4 u2 - u3, K even
void launch_mspmv_chain(int n, int K, const int *ia, const int *ja, const double *val, const double *e2,
                        const double *u3, double *u2, double *r, int xcd, hipStream_t st) {
    hipLaunchKernelGGL(k_mspmv_chain, dim3((unsigned)((n + 4 * TF_RPW - 1) / (4 * TF_RPW))), dim3(256), 0, st, n, K,
                       ia, ja, val, e2, u3, u2, r, xcd);
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
}  // namespace tf
