// factor.hip -- multifrontal LDL^T of P = 3*Msigma + 2*dt*K on the GPU (FP64).
//
// Replaces the numeric part of PARDISO phase 12 called by the reference
// (pardiso_direct_1, src/FETD_UTILITY.f90:879-936; called at
// src/FETD_FWD.f90:1307,1361).  No pivoting (P is symmetric positive
// definite in exact arithmetic, PARDISO mtype 2); zero or non-finite pivots
// raise a flag that the host turns into an error.
//
// Work is organised by levels of the supernode tree (all fronts of a level
// are independent).  Every 64-column panel of a level's fronts is one set of
// grid launches (pivot block, panel solve, trailing updates grouped by
// super-panels); then L11 is inverted by recursive doubling.
#include "tf_device.h"

#define LDS_LD 65


// ---- workgroup-level building blocks ------------------------------------

// Invert the unit-lower pw x pw diagonal block of panel p0 into LDS
// (column-major, ld 65, strict lower part; the diagonal is implicit).
// Column j of X = L^-1 follows from columns > j:
//   X[i][j] = -L[i][j] - sum_{j<k<i} X[i][k] L[k][j].
__device__ void blk_invert_to_lds(const double *L, const FrontDesc &d, int p0, double *lds) {
    const int pw = min(TF_PB, d.nc - p0);
    const double *base = L + d.loff;
    for (int idx = threadIdx.x; idx < TF_PB * TF_PB; idx += blockDim.x) {
        int i = idx % TF_PB, j = idx / TF_PB;
        lds[j * LDS_LD + i] = (i > j && i < pw && j < pw) ? base[(int64_t)(p0 + j) * d.ld + p0 + i] : 0.0;
    }
    __syncthreads();
    for (int j = pw - 2; j >= 0; j--) {
        const int i = j + 1 + threadIdx.x;
        double v = 0.0;
        if (i < pw) {
            double s = lds[j * LDS_LD + i];
            for (int k = j + 1; k < i; k++) s += lds[k * LDS_LD + i] * lds[j * LDS_LD + k];
            v = -s;
        }
        __syncthreads();
        if (i < pw) lds[j * LDS_LD + i] = v;
        __syncthreads();
    }
}

// Diagonal block of panel p0 replaced by its inverse's strict lower part; the
// diagonal positions, whose pivots are kept in D, are zeroed, so that the
// whole upper part of L11's storage reads as zero and the solves can run the
// strict-triangle products of X11 = L11^-1 as plain dense products.
__device__ void blk_invert_diag(double *L, const FrontDesc &d, int p0, double *lds) {
    const int pw = min(TF_PB, d.nc - p0);
    blk_invert_to_lds(L, d, p0, lds);
    double *base = L + d.loff;
    for (int idx = threadIdx.x; idx < pw * pw; idx += blockDim.x) {
        int i = idx % pw, j = idx / pw;
        if (i > j) base[(int64_t)(p0 + j) * d.ld + p0 + i] = lds[j * LDS_LD + i];
        else if (i == j) base[(int64_t)(p0 + j) * d.ld + p0 + i] = 0.0;
    }
    __syncthreads();
}

// In-place inversion of the unit-lower L11 of a large front by recursive
// doubling (the diagonal positions hold D and are left alone; the inverse's
// unit diagonal stays implicit).  Every 64 x 64 diagonal block is inverted
// first; then blocks of size b = 64, 128, 256, ... are merged in pairs
// I = [s, s+b), J = [s+b, min(s+2b, nc)), s = 2b*pair, with
//   inv([L_II 0; L_JI L_JJ]) = [X_II 0; -X_JJ L_JI X_II  X_JJ]:
//   step 0: T    = L_JI X_II    (scratch, |J| x b, column-major, ld ldT)
//   step 1: X_JI = -X_JJ T      (in place over L_JI)
// Each step is a grid of 128 x 128 output regions, a 64 x 64 matrix-core
// tile per wave, so the top merges of a front with thousands of columns fill
// the whole chip in a handful of launches.

// Wave-level trailing update on matrix cores: one wave updates the 64 x 64
// tile C(rb.., cb..) -= L_rows D L_cols^T over the panel columns
// [kbeg, kend), on the lower triangle of the trailing block (rows < f,
// columns < cmax).  Both operands are read straight from the panel columns
// into the MFMA lane layout with raw buffer loads (stage offset in a scalar
// register, one lane offset per operand, block offsets as immediates; rows
// past the front read into the next column or the storage slack and are
// never stored), four inner values per k-step, KS k-steps per stage, the
// next stage in flight while the current one is multiplied; a last partial
// stage zeroes the columns past kend.  The column operand is scaled by the
// pivots and passed as the MFMA's A operand.  Each operand load fetches two
// consecutive rows of a panel column per lane (one 16-byte load for two
// 16-wide blocks), so block 2p holds the rows 32p + 2i and block 2p + 1 the
// rows 32p + 2i + 1 (i = 0..15) of the tile, on either side: result
// register r of block (a, b) holds column cb + 32(a/2) + 2(4r + l/16) + a%2,
// row rb + 32(b/2) + 2(l%16) + b%2, and the blocks b = 2p, 2p + 1 of one
// register are two consecutive rows, updated with one 16-byte access.
template <int KS>
__device__ __forceinline__ void upd_wave(double *L, double *U, const FrontDesc &d, const double *Dn, int kbeg,
                                         int kend, int rb, int cb, int cmax) {
    constexpr int SW = 4 * KS;
    const double *base = L + d.loff;
    const __amdgpu_buffer_rsrc_t rL = mt_rsrc(base, (uint32_t)((int64_t)d.nc * d.ld * 8));
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
    const uint32_t lr = (uint32_t)(lk * d.ld + rb + 2 * l16) * 8u, lc = (uint32_t)(lk * d.ld + cb + 2 * l16) * 8u;
    const double *Dk = Dn + d.doff;
    tf_d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (tf_d4){0.0, 0.0, 0.0, 0.0};
    // The column operand is scaled by D when its stage is multiplied, not
    // when it is loaded, so that a stage's loads stay in flight while the
    // stage before it runs on the matrix cores.
    auto mult = [&](const double xr[KS][4], double yc[KS][4], const double dk[KS]) {
#pragma unroll
        for (int s = 0; s < KS; s++) {
#pragma unroll
            for (int a = 0; a < 4; a++) yc[s][a] *= dk[s];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(yc[s][a], xr[s][b], acc[a][b], 0, 0, 0);
        }
    };
    auto load = [&](int ib, double xr[KS][4], double yc[KS][4], double dk[KS]) {
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const uint32_t sl = (uint32_t)((ib + 4 * s) * d.ld) * 8u;
            dk[s] = Dk[ib + 4 * s + lk];
#pragma unroll
            for (int p = 0; p < 2; p++) mt_bld2(rL, lr + 256u * p, sl, xr[s][2 * p], xr[s][2 * p + 1]);
#pragma unroll
            for (int p = 0; p < 2; p++) mt_bld2(rL, lc + 256u * p, sl, yc[s][2 * p], yc[s][2 * p + 1]);
        }
    };
    const int nfull = (kend - kbeg) / SW;
    int ib = kbeg;
    if (nfull > 0) {
        double xr0[KS][4], yc0[KS][4], xr1[KS][4], yc1[KS][4], dk0[KS], dk1[KS];
        int j = 0;
        if (nfull & 1) {  // odd count: the first stage on its own, the rest in pairs
            load(ib, xr0, yc0, dk0);
            mult(xr0, yc0, dk0);
            ib += SW;
            j = 1;
        }
        if (j < nfull) {
            const int last = kbeg + (nfull - 1) * SW;
            load(ib, xr0, yc0, dk0);
            for (; j < nfull; j += 2) {  // ib + SW <= last
                load(ib + SW, xr1, yc1, dk1);
                __builtin_amdgcn_sched_barrier(0);
                mult(xr0, yc0, dk0);
                __builtin_amdgcn_sched_barrier(0);
                load(min(ib + 2 * SW, last), xr0, yc0, dk0);
                __builtin_amdgcn_sched_barrier(0);
                mult(xr1, yc1, dk1);
                __builtin_amdgcn_sched_barrier(0);
                ib += 2 * SW;
            }
        }
    }
    if (ib < kend) {  // partial stage: columns past kend read as zero
        double xr[KS][4], yc[KS][4], dk[KS];
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int k = ib + 4 * s + lk;
            const bool kv = k < kend;
            const int kc = kv ? k : kend - 1;
            const uint32_t sl = (uint32_t)(kc * d.ld) * 8u;
            const uint32_t l0 = (uint32_t)(rb + 2 * l16) * 8u, c0 = (uint32_t)(cb + 2 * l16) * 8u;
            dk[s] = kv ? Dk[kc] : 0.0;
#pragma unroll
            for (int p = 0; p < 2; p++) mt_bld2(rL, sl + l0 + 256u * p, 0u, xr[s][2 * p], xr[s][2 * p + 1]);
#pragma unroll
            for (int p = 0; p < 2; p++) mt_bld2(rL, sl + c0 + 256u * p, 0u, yc[s][2 * p], yc[s][2 * p + 1]);
        }
        mult(xr, yc, dk);
    }
    // C -= acc on the stored lower triangle; all loads of a column block
    // first.  A row pair whose rows are both stored and whose address is
    // 16-byte aligned (L columns always; update-stack columns when their
    // offset is even) is read and written as one 16-byte access.
    typedef double u_d2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int a = 0; a < 4; a++) {
        u_d2 cv[4][2];
        double *cp[4];
        bool both[4][2];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int c = cb + 32 * (a >> 1) + 2 * (4 * r + lk) + (a & 1);
            cp[r] = (c < cmax) ? front_col(L, U, d, c) : nullptr;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int row = rb + 32 * p + 2 * l16;
                const bool v0 = cp[r] && row < d.f && row >= c, v1 = cp[r] && row + 1 < d.f && row + 1 >= c;
                both[r][p] = v0 && v1 && (((uintptr_t)(cp[r] + row)) & 15) == 0;
                if (both[r][p]) {
                    cv[r][p] = *(const u_d2 *)(cp[r] + row);
                } else {
                    cv[r][p].x = v0 ? cp[r][row] : 0.0;
                    cv[r][p].y = v1 ? cp[r][row + 1] : 0.0;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int c = cb + 32 * (a >> 1) + 2 * (4 * r + lk) + (a & 1);
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int row = rb + 32 * p + 2 * l16;
                const u_d2 nv = cv[r][p] - (u_d2){acc[a][2 * p][r], acc[a][2 * p + 1][r]};
                if (both[r][p]) {
                    *(u_d2 *)(cp[r] + row) = nv;
                } else {
                    if (cp[r] && row < d.f && row >= c) cp[r][row] = nv.x;
                    if (cp[r] && row + 1 < d.f && row + 1 >= c) cp[r][row + 1] = nv.y;
                }
            }
        }
    }
}

// Pivot block of panel p0 of a large front by one wave, in registers: lane i
// holds row i of the pw x pw block; the entries of column j that the
// elimination and the inversion need from other rows are read with
// readlane.  LDL^T is written back with D; then X = L11^-1 of the block,
// column by column from the right (X[i][j] = -L[i][j] - sum_{j<k<i}
// X[i][k] L[k][j]), is stored column-major (ld 64, zero on and above the
// diagonal) at Xg for the panel solve.
__device__ __forceinline__ double lane_bcast(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ void wave_diag_factor(double *L, const FrontDesc &d, int p0, double *Dn, int *flag, double *Xg) {
    const int pw = min(TF_PB, d.nc - p0);
    const int i = threadIdx.x & 63;  // lane i holds row i of the pivot block
    double *base = L + d.loff + (int64_t)p0 * d.ld + p0;
    double a[TF_PB];
#pragma unroll
    for (int j = 0; j < TF_PB; j++) a[j] = (j <= i && i < pw) ? base[(int64_t)j * d.ld + i] : 0.0;
    // right-looking LDL^T: a_ik -= (a_ij / d_j) a_kj for j < k <= i, then
    // a_ij /= d_j; a_kj (row k, column j) is broadcast from lane k
    double di = 0.0;
#pragma unroll
    for (int j = 0; j < TF_PB; j++) {  // columns j >= pw: all zero, no effect
        double dj = lane_bcast(a[j], j);
        if (i == j) di = dj;
        if (!(fabs(dj) > 1e-300) || !isfinite(dj)) dj = (dj < 0.0) ? -1e-300 : 1e-300;
        const double lij = (i > j) ? a[j] / dj : 0.0;
#pragma unroll
        for (int k = j + 1; k < TF_PB; k++) a[k] -= lij * lane_bcast(a[j], k);
        if (i > j) a[j] = lij;
    }
    if (i < pw) {
#pragma unroll
        for (int j = 0; j < TF_PB; j++)
            if (j <= i) base[(int64_t)j * d.ld + i] = a[j];
        Dn[d.doff + p0 + i] = di;
        if (!(fabs(di) > 1e-300) || !isfinite(di)) atomicOr(flag, 1);
    }
    // X = L11^-1 (unit lower), strict part, by columns from the right:
    // x_ij = -(l_ij + sum_{j<k<i} x_ik l_kj); x_ij overwrites l_ij once
    // column j of L has been broadcast
#pragma unroll
    for (int j = TF_PB - 2; j >= 0; j--) {
        double s = a[j];
#pragma unroll
        for (int k = j + 1; k < TF_PB - 1; k++) s += a[k] * lane_bcast(a[j], k);  // a[k] = 0 for k >= i
        a[j] = (i > j && i < pw) ? -s : 0.0;
    }
    a[TF_PB - 1] = 0.0;
#pragma unroll
    for (int j = 0; j < TF_PB; j++) Xg[j * TF_PB + i] = a[j];
}

// ---- kernels ---------------------------------------------------------------

// Row-major copy of every front's L21 (the block solve's backward product
// then reads it with the output index running along rows): one 64 x 64
// tile per workgroup through LDS; units[u] = (front, first struct row,
// first own column).
__global__ void __launch_bounds__(256) k_transpose_l21(const double *L, double *LT, const FrontDesc *fd,
                                                       const int *units) {
    __shared__ double t[64][65];
    const int *u = units + 3 * (int64_t)blockIdx.x;
    const FrontDesc d = fd[u[0]];
    const int r0 = u[1], c0 = u[2];
    const int x = threadIdx.x & 63, y = threadIdx.x >> 6;
    // rows of the copy: the m L21 rows, or all f rows of a partitioned-inverse front
    const int rows = d.pinv ? d.f : d.m, rb = d.pinv ? 0 : d.nc;
    const double *src = L + d.loff + (int64_t)c0 * d.ld + rb + r0;
    for (int c = y; c < 64; c += 4)  // column c0 + c, rows r0 + x
        t[c][x] = (c0 + c < d.nc && r0 + x < rows) ? src[(int64_t)c * d.ld + x] : 0.0;
    __syncthreads();
    double *dst = LT + d.toff + (int64_t)r0 * d.ldt + c0;
    for (int r = y; r < 64; r += 4)  // row r0 + r, columns c0 + x
        if (r0 + r < rows && c0 + x < d.ldt) dst[(int64_t)r * d.ldt + x] = t[x][r];
}

// p[0 .. n) = 0 with 16-byte stores (p itself may be 8-byte aligned only:
// the leading and trailing odd elements are written separately)
__global__ void k_fill0(double *p, int64_t n) {
    const int64_t h = ((reinterpret_cast<uintptr_t>(p) & 15) != 0 && n > 0) ? 1 : 0;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    if (t == 0) {
        if (h) p[0] = 0.0;
        if ((n - h) & 1) p[n - 1] = 0.0;
    }
    double2 *q = reinterpret_cast<double2 *>(p + h);
    const int64_t n2 = (n - h) / 2;
    for (int64_t i = t; i < n2; i += stride) q[i] = make_double2(0.0, 0.0);
}

__global__ void k_flags_init(FactorFlags *f) {
    if (threadIdx.x == 0) {
        f->bad = 0;
        f->pad = 0;
        f->nneg = 0ull;
        f->min_abs = 0x7ff0000000000000ull;  // +inf
        f->max_abs = 0ull;
    }
}

// Pivot scan: one workgroup per front over its nc pivots D[doff .. doff+nc).
// P is symmetric positive definite in exact arithmetic (PARDISO mtype 2);
// a negative pivot is where the reference's SPD factorisation would fail and
// fall back to the indefinite one (src/FETD_UTILITY.f90:919-935).
__global__ void __launch_bounds__(256) k_pivot_scan(const double *Dn, const FrontDesc *fd, int nsn,
                                                    FactorFlags *f) {
    __shared__ unsigned long long s_neg[256], s_min[256], s_max[256];
    unsigned long long neg = 0ull, mn = 0x7ff0000000000000ull, mx = 0ull;
    for (int k = blockIdx.x; k < nsn; k += gridDim.x) {
        const FrontDesc d = fd[k];
        for (int j = threadIdx.x; j < d.nc; j += blockDim.x) {
            const double v = Dn[d.doff + j];
            if (v < 0.0) neg++;
            const unsigned long long a = (unsigned long long)__double_as_longlong(fabs(v));
            mn = a < mn ? a : mn;
            mx = a > mx ? a : mx;
        }
    }
    s_neg[threadIdx.x] = neg;
    s_min[threadIdx.x] = mn;
    s_max[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_neg[threadIdx.x] += s_neg[threadIdx.x + w];
            s_min[threadIdx.x] = min(s_min[threadIdx.x], s_min[threadIdx.x + w]);
            s_max[threadIdx.x] = max(s_max[threadIdx.x], s_max[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (s_neg[0]) atomicAdd(&f->nneg, s_neg[0]);
        atomicMin(&f->min_abs, s_min[0]);
        atomicMax(&f->max_abs, s_max[0]);
    }
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
// Extend-add of the children's update stacks into their parents, one wave
// per child column b: rows a >= b of the column go to the parent's column
// rel[b], rows rel[a].  pref[] counts columns per child; four loads of each
// kind are issued before their adds.
__global__ void __launch_bounds__(TF_NT) k_extend_add(double *L, double *U, const FrontDesc *fd,
                                                      const int *parent, const int *list,
                                                      const int64_t *pref, int cnt, const int *rel) {
    const int64_t g = (int64_t)blockIdx.x * (TF_NT / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (g >= pref[cnt]) return;
    const int lane = threadIdx.x & 63;
    const int i = seg_find(pref, cnt, g);
    const int c = list[i];
    const int b = (int)(g - pref[i]);
    const FrontDesc dc = fd[c];
    const FrontDesc dp = fd[parent[c]];
    const double *uc = U + dc.uoff + ucol(dc.m, b) - b;  // uc[a], a >= b: row a of column b
    const int *rc = rel + dc.soff;
    double *dst = front_col(L, U, dp, rc[b]);
    int a = b + lane;
    for (; a + 192 < dc.m; a += 256) {
        int r[4];
        double u[4], v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            r[q] = rc[a + 64 * q];
            u[q] = uc[a + 64 * q];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = dst[r[q]];
#pragma unroll
        for (int q = 0; q < 4; q++) dst[r[q]] = v[q] + u[q];
    }
    for (; a < dc.m; a += 64) dst[rc[a]] += uc[a];
}

// One wave per front: pivot block of panel p0 and its inverse (for the panel
// solve) at W + woff[front].
__global__ void __launch_bounds__(64) k_factor_diag(double *L, const FrontDesc *fd, const int *list, int p0,
                                                    double *Dn, int *flag, double *W, const int64_t *woff) {
    const int sn = list[blockIdx.x];
    const FrontDesc d = fd[sn];
    wave_diag_factor(L, d, p0, Dn, flag, W + woff[sn]);
}

// Panel solve of a large front by one wave per 64 rows [rb, rb + 64) below
// the panel: L21 = A21 X^T D^-1 with X = L11^-1 of the pivot block (unit
// lower; the pivot kernel leaves its strict part column-major, ld 64, at Xg):
//   L21[row][c] = (A21[row][c] + sum_{k < c} A21[row][k] X[c][k]) / d_c.
// Both operands stream from memory into the MFMA lane layout with 16-byte
// loads, two rows (or two columns of X) per lane, as in the trailing update:
// result register r of block (a, b) holds column 32(a/2) + 2(4r + l/16) + a%2
// of the panel, row rb + 32(b/2) + 2(l%16) + b%2.  Rows past the front read
// whatever follows (buffer range check at the block's end) and are never
// stored; columns past the panel width pw multiply zeros of Xg.
template <int KS>
__device__ __forceinline__ void trsm_wave(double *L, const FrontDesc &d, int p0, int rb, const double *Dn,
                                          const double *Xg) {
    constexpr int SW = 4 * KS;
    const int pw = min(TF_PB, d.nc - p0);
    double *base = L + d.loff;
    const __amdgpu_buffer_rsrc_t rL = mt_rsrc(base, (uint32_t)((int64_t)d.nc * d.ld * 8));
    const __amdgpu_buffer_rsrc_t rX = mt_rsrc(Xg, (uint32_t)(TF_PB * TF_PB * 8));
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
    const uint32_t lx = (uint32_t)(lk * TF_PB + 2 * l16) * 8u;
    tf_d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (tf_d4){0.0, 0.0, 0.0, 0.0};
    auto mult = [&](const double xr[KS][4], const double yc[KS][4]) {
#pragma unroll
        for (int s = 0; s < KS; s++)
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(yc[s][a], xr[s][b], acc[a][b], 0, 0, 0);
    };
    // inner index k of stage value s: ib + 4 s + l/16.  X is zero for k >= pw,
    // so the panel column is clamped to pw - 1 there (every read stays in
    // the front's block; the whole offset is per lane)
    auto load = [&](int ib, double xr[KS][4], double yc[KS][4]) {
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int k = min(ib + 4 * s + lk, pw - 1);
            const uint32_t lr = (uint32_t)((p0 + k) * d.ld + rb + 2 * l16) * 8u;
#pragma unroll
            for (int p = 0; p < 2; p++) mt_bld2(rL, lr + 256u * p, 0u, xr[s][2 * p], xr[s][2 * p + 1]);
#pragma unroll
            for (int p = 0; p < 2; p++)
                mt_bld2(rX, lx + 256u * p, (uint32_t)((ib + 4 * s) * TF_PB) * 8u, yc[s][2 * p], yc[s][2 * p + 1]);
        }
    };
    const int nst = (pw + SW - 1) / SW;  // X rows k >= pw are zero, so whole stages suffice
    double xr0[KS][4], yc0[KS][4], xr1[KS][4], yc1[KS][4];
    int js = 0;
    if (nst & 1) {
        load(0, xr0, yc0);
        mult(xr0, yc0);
        js = 1;
    }
    if (js < nst) {
        load(js * SW, xr0, yc0);
        for (; js < nst; js += 2) {
            load((js + 1) * SW, xr1, yc1);
            __builtin_amdgcn_sched_barrier(0);
            mult(xr0, yc0);
            __builtin_amdgcn_sched_barrier(0);
            load(min(js + 2, nst - 1) * SW, xr0, yc0);
            __builtin_amdgcn_sched_barrier(0);
            mult(xr1, yc1);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // L21 = (A21 + acc) / d on rows < f, columns < pw (row pairs as one 16-byte access)
    typedef double t_d2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int c = 32 * (a >> 1) + 2 * (4 * r + lk) + (a & 1);
            if (c >= pw) continue;
            const double dc = Dn[d.doff + p0 + c];
            double *col = base + (int64_t)(p0 + c) * d.ld;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int row = rb + 32 * p + 2 * l16;
                const t_d2 s = (t_d2){acc[a][2 * p][r], acc[a][2 * p + 1][r]};
                if (row + 1 < d.f) {
                    *(t_d2 *)(col + row) = (*(const t_d2 *)(col + row) + s) / dc;
                } else if (row < d.f) {
                    col[row] = (col[row] + s.x) / dc;
                }
            }
        }
}

// one workgroup of two waves per 128 rows below the panel
__global__ void __launch_bounds__(128, 2) k_factor_trsm(double *L, const FrontDesc *fd, const int *list,
                                                     const int64_t *pref, int cnt, int p0, const double *Dn,
                                                     const double *W, const int64_t *woff) {
    const int i = seg_find(pref, cnt, blockIdx.x);
    const int sn = list[i];
    const FrontDesc d = fd[sn];
    const int p1 = min(d.nc, p0 + TF_PB);
    const int rb = p1 + (int)(blockIdx.x - pref[i]) * TF_TRSM_ROWS + 64 * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (rb >= d.f) return;
    trsm_wave<2>(L, d, p0, rb, Dn, W + woff[sn]);
}

// Tile (ti, tj), tj <= ti, of a lower triangle of nt tile rows for index u
// in band order: bands of g tile rows, each walked column by column (the g
// tiles of a column in a row, then the triangle's part of the band).  The
// tiles that run at once on one XCD then share a few row panels and column
// panels in its L2 instead of streaming every column panel once per tile
// row.  g = 0 walks the triangle row by row.
__device__ __forceinline__ void band_index(int64_t u, int nt, int g, int &ti, int &tj) {
    tri_index(u, ti, tj);
    if (g <= 0) return;
    const int r0 = ti - ti % g, h = min(g, nt - r0);
    int64_t v = u - (int64_t)r0 * (r0 + 1) / 2;
    if (v < (int64_t)r0 * h) {
        tj = (int)(v / h);
        ti = r0 + (int)(v % h);
        return;
    }
    v -= (int64_t)r0 * h;
    int c = 0;
    while (v >= h - c) {
        v -= h - c;
        c++;
    }
    tj = r0 + c;
    ti = tj + (int)v;
}

// Updates of a super-panel [P0, P0 + TF_SP) of TF_PB-column panels, one
// 64 x 64 tile per wave:
//   mode 0 (after panel p0): the rest of the super-panel's own columns,
//          rows from p1 down, with panel p0 only;
//   mode 1 (p0 = P0, after its last panel): everything right of the
//          super-panel with all of its columns at once.
// The trailing block is read and written once per super-panel instead of
// once per panel.  pref[] counts tiles per front.
__global__ void __launch_bounds__(TF_NT, 2) k_factor_update(double *L, double *U, const FrontDesc *fd,
                                                            const int *list, const int64_t *pref, int cnt,
                                                            int p0, int mode, const double *Dn, int band) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t ug = (int64_t)xcd_block(blockIdx.x, gridDim.x) * 4 + w;
    if (ug >= pref[cnt]) return;
    const int i = seg_find(pref, cnt, ug);
    const FrontDesc d = fd[list[i]];
    const int64_t u = ug - pref[i];
    if (mode == 0) {
        const int P0 = p0 - p0 % TF_SP;
        const int p1 = min(d.nc, p0 + TF_PB), pe = min(d.nc, P0 + TF_SP);
        const int nt = (d.f - p1 + 63) / 64;
        int tj = 0, ti = (int)u;
        while (ti >= nt - tj) { ti -= nt - tj; tj++; }
        ti += tj;
        upd_wave<2>(L, U, d, Dn, p0, p1, p1 + 64 * ti, p1 + 64 * tj, pe);
    } else {
        const int pe = min(d.nc, p0 + TF_SP);
        int ti, tj;
        band_index(u, (d.f - pe + 63) / 64, band, ti, tj);
        upd_wave<2>(L, U, d, Dn, p0, pe, pe + 64 * ti, pe + 64 * tj, d.f);
    }
}

// All 64 x 64 diagonal blocks of the listed fronts' L11, one per workgroup.
__global__ void __launch_bounds__(TF_NT) k_inv_diag(double *L, const FrontDesc *fd, const int *list,
                                                    const int64_t *pref, int cnt) {
    __shared__ double lds[TF_PB * LDS_LD];
    const int i = seg_find(pref, cnt, blockIdx.x);
    const FrontDesc d = fd[list[i]];
    blk_invert_diag(L, d, (int)(blockIdx.x - pref[i]) * TF_PB, lds);
}

// One wave's 64 x 64 tile of a merge step, straight from memory into the
// MFMA lane layout (no LDS, no barriers):
//   acc(c, r) = sum_{k0 <= k < k1} Cop(k, c) Rop(r, k),
//   Cop(k, c) = C[c * ldc + k]  (a lane reads two consecutive k of column c),
//   Rop(r, k) = R[k * ldr + r]  (a lane reads two consecutive rows r of column k),
// with UC / UR making that operand unit lower triangular (stored values below
// the diagonal, 1 on it, 0 above).  The inner index of k-step s of a stage is
// ib + 2 l/16 + s for both operands.  Result register q of block (a, b) holds
// column cb + 16a + 4q + l/16 and row rb + 32(b/2) + 2(l%16) + b%2, so the
// blocks b = 2p, 2p + 1 of one register are two consecutive rows.  Reads past
// k1 are clamped to k1 - 1 and zeroed; rows past the operand's valid range
// read neighbouring storage (the results are not stored).
template <bool UC, bool UR>
__device__ __forceinline__ void inv_wave(const double *R, int ldr, uint32_t rbytes, const double *C, int ldc,
                                         uint32_t cbytes, int rb, int cb, int k0, int k1, tf_d4 (&acc)[4][4]) {
    constexpr int SW = 8;  // inner values per stage (two k-steps)
    const __amdgpu_buffer_rsrc_t rR = mt_rsrc(R, rbytes), rC = mt_rsrc(C, cbytes);
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (tf_d4){0.0, 0.0, 0.0, 0.0};
    if (k0 >= k1) return;
    // A stage's values are loaded raw and masked when the stage is
    // multiplied, so that its loads stay in flight while the stage before it
    // runs on the matrix cores.
    auto load = [&](int ib, double cv[2][4], double rv[2][4]) {
        const int kp = ib + 2 * lk;  // the lane's two inner values kp, kp + 1
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int c = cb + 16 * a + l16;
            mt_bld2(rC, (uint32_t)((int64_t)c * ldc + kp) * 8u, 0u, cv[0][a], cv[1][a]);  // zeroed past k1
        }
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const int k = min(kp + s2, k1 - 1);
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int r0 = rb + 32 * p + 2 * l16;
                mt_bld2(rR, (uint32_t)((int64_t)k * ldr + r0) * 8u, 0u, rv[s2][2 * p], rv[s2][2 * p + 1]);
            }
        }
    };
    auto mult = [&](int ib, double cv[2][4], double rv[2][4]) {
        const int kp = ib + 2 * lk;
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
            const int k = kp + s2;
            const bool kv = k < k1;
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int c = cb + 16 * a + l16;
                double y = kv ? cv[s2][a] : 0.0;
                if (UC) y = k > c ? y : (k == c ? 1.0 : 0.0);
                cv[s2][a] = y;
            }
            const int kc = min(k, k1 - 1);
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int r0 = rb + 32 * p + 2 * l16;
                double y0 = rv[s2][2 * p], y1 = rv[s2][2 * p + 1];
                if (UR) {
                    y0 = r0 > kc ? y0 : (r0 == kc ? 1.0 : 0.0);
                    y1 = r0 + 1 > kc ? y1 : (r0 + 1 == kc ? 1.0 : 0.0);
                }
                rv[s2][2 * p] = kv ? y0 : 0.0;
                rv[s2][2 * p + 1] = kv ? y1 : 0.0;
            }
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(cv[s2][a], rv[s2][b], acc[a][b], 0, 0, 0);
        }
    };
    const int nst = (k1 - k0 + SW - 1) / SW;
    double c0v[2][4], r0v[2][4], c1v[2][4], r1v[2][4];
    int js = 0;
    if (nst & 1) {
        load(k0, c0v, r0v);
        mult(k0, c0v, r0v);
        js = 1;
    }
    if (js < nst) {
        load(k0 + js * SW, c0v, r0v);
        for (; js < nst; js += 2) {
            load(k0 + (js + 1) * SW, c1v, r1v);
            __builtin_amdgcn_sched_barrier(0);
            mult(k0 + js * SW, c0v, r0v);
            __builtin_amdgcn_sched_barrier(0);
            load(k0 + min(js + 2, nst - 1) * SW, c0v, r0v);
            __builtin_amdgcn_sched_barrier(0);
            mult(k0 + (js + 1) * SW, c1v, r1v);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// One merge step at block size b; work unit u of a front is (pair, ti, tj)
// with ntb x ntb output regions of 128 x 128 per pair, ntb = ceil(b / 128)
// (the last pair only the row regions of its J rows), one 64 x 64 quadrant
// per wave:
//   step 0: T    = L_JI X_II   (k >= c: X_II(k, c) = 0 above its diagonal)
//   step 1: X_JI = -X_JJ T     (k <= r: X_JJ(r, k) = 0 above its diagonal)
template <int STEP>
__global__ void __launch_bounds__(TF_NT, 2) k_inv_merge(double *L, const FrontDesc *fd, const int *list,
                                                        const int64_t *pref, int cnt, int b, double *W,
                                                        const int64_t *woff) {
    // Step 0 runs consecutive work units on one XCD (each XCD then takes whole
    // row regions, whose tiles share L_JI in its L2 and mix long and short
    // inner ranges, k >= c); step 1 keeps the round-robin order, since its
    // inner range grows with the row region and whole row regions per XCD
    // would leave the last XCDs the longest tiles.
    const int64_t gb = STEP == 0 ? (int64_t)xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int i = seg_find(pref, cnt, gb);
    const int sn = list[i];
    const FrontDesc d = fd[sn];
    const int64_t u = gb - pref[i];
    const int ntb = (b + TF_UT - 1) / TF_UT;
    const int pair = (int)(u / (ntb * ntb)), t = (int)(u % (ntb * ntb));
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int I0 = 2 * b * pair, J0 = I0 + b, Jn = min(b, d.nc - J0);
    const int rb = (t / ntb) * TF_UT + 64 * (w >> 1), cb = (t % ntb) * TF_UT + 64 * (w & 1);
    if (rb >= Jn || cb >= b) return;  // this wave's quadrant lies past the pair
    const int ldT = inv_ldt(d.nc, b);
    double *T = W + woff[sn] + (int64_t)pair * ldT * b;
    double *base = L + d.loff;
    const int64_t lend = (int64_t)d.nc * d.ld;  // the front's L block
    tf_d4 acc[4][4];
    if (STEP == 0) {
        const int64_t ro = (int64_t)I0 * d.ld + J0, co = (int64_t)I0 * d.ld + I0;
        inv_wave<true, false>(base + ro, d.ld, (uint32_t)((lend - ro) * 8), base + co, d.ld,
                              (uint32_t)((lend - co) * 8), rb, cb, cb, b, acc);
    } else {
        const int64_t ro = (int64_t)J0 * d.ld + J0;
        inv_wave<false, true>(base + ro, d.ld, (uint32_t)((lend - ro) * 8), T, ldT, (uint32_t)((int64_t)ldT * b * 8),
                              rb, cb, 0, min(rb + 64, Jn), acc);
    }
    typedef double i_d2 __attribute__((ext_vector_type(2)));
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int c = cb + 16 * a + 4 * q + lk;
            if (c >= b) continue;
            double *col = (STEP == 0) ? T + (int64_t)c * ldT : base + (int64_t)(I0 + c) * d.ld + J0;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int r0 = rb + 32 * p + 2 * l16;
                const i_d2 v = (STEP == 0) ? (i_d2){acc[a][2 * p][q], acc[a][2 * p + 1][q]}
                                           : (i_d2){-acc[a][2 * p][q], -acc[a][2 * p + 1][q]};
                if (r0 + 1 < Jn) *(i_d2 *)(col + r0) = v;
                else if (r0 < Jn) col[r0] = v.x;
            }
        }
}

// Partitioned inverse of a front (fronts with pinv set, after their
// factorisation and the in-place inversion of L11): the L21 rows become
// W21 = -L21 X11, X11 = L11^-1 (unit lower; its strict part stored in L11's
// place, whose diagonal and upper part read as zero), so that the solve's
// forward product of the front is one product of the whole block with the
// front's right-hand-side rows ([Y_own - Yin; -Z] = [strict(X11); W21] Yin)
// and its backward product one product with [Y_own / d; x(struct)].
// One launch per 64-column block cb, left to right; one wave per 64-row
// slab of L21: the block is the product of the slab's columns j >= cb with
// X11's rows j >= cb (X11[j][c] = 0 for j < c; its unit diagonal takes L21
// itself into the product) and overwrites L21's columns [cb, cb + 64),
// which no later launch reads.
__global__ void __launch_bounds__(TF_NT, 2) k_pinv_w21(double *L, const FrontDesc *fd, const int *list,
                                                       const int64_t *pref, int cnt, int cb) {
    const int64_t g = (int64_t)blockIdx.x * (TF_NT / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (g >= pref[cnt]) return;
    const int i = seg_find(pref, cnt, g);
    const FrontDesc d = fd[list[i]];
    const int r0 = (int)(g - pref[i]) * 64;
    double *base = L + d.loff;
    const int nc = d.nc, m = d.m, ld = d.ld;
    const int64_t lend = (int64_t)nc * ld;  // the front's L block
    tf_d4 acc[4][4];
    inv_wave<true, false>(base + nc, ld, (uint32_t)((lend - nc) * 8), base, ld, (uint32_t)(lend * 8), r0, cb, cb,
                          nc, acc);
    typedef double i_d2 __attribute__((ext_vector_type(2)));
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int c = cb + 16 * a + 4 * q + lk;
            if (c >= nc) continue;
            double *col = base + (int64_t)c * ld + nc;
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int r = r0 + 32 * p + 2 * l16;
                const i_d2 v = (i_d2){-acc[a][2 * p][q], -acc[a][2 * p + 1][q]};
                if (r + 1 < m && ((reinterpret_cast<uintptr_t>(col + r) & 15) == 0)) {
                    *(i_d2 *)(col + r) = v;
                } else {
                    if (r < m) col[r] = v.x;
                    if (r + 1 < m) col[r + 1] = v.y;
                }
            }
        }
}

// ---- host launchers --------------------------------------------------------
namespace tf {

void launch_transpose_l21(const double *L, double *LT, const FrontDesc *fd, const int *units, int64_t nunits,
                          hipStream_t st) {
    if (nunits > 0) hipLaunchKernelGGL(k_transpose_l21, dim3((unsigned)nunits), dim3(256), 0, st, L, LT, fd, units);
}
void launch_pinv_w21(double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                     int cb, hipStream_t st) {
    // nwg counts 64-row slabs: one wave each
    if (nwg > 0)
        hipLaunchKernelGGL(k_pinv_w21, dim3((unsigned)((nwg + TF_NT / 64 - 1) / (TF_NT / 64))), dim3(TF_NT), 0, st,
                           L, fd, list, pref, cnt, cb);
}
void launch_fill0(double *p, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    int64_t blocks = std::min<int64_t>((n / 2 + 255) / 256 + 1, 16384);
    hipLaunchKernelGGL(k_fill0, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
}
void launch_flags_init(FactorFlags *f, hipStream_t st) {
    hipLaunchKernelGGL(k_flags_init, dim3(1), dim3(64), 0, st, f);
}
void launch_pivot_scan(const double *Dn, const FrontDesc *fd, int nsn, FactorFlags *f, hipStream_t st) {
    if (nsn <= 0) return;
    const int blocks = std::min(nsn, 2048);
    hipLaunchKernelGGL(k_pivot_scan, dim3(blocks), dim3(256), 0, st, Dn, fd, nsn, f);
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
    // nwg counts child columns: one wave each
    hipLaunchKernelGGL(k_extend_add, dim3((unsigned)((nwg + TF_NT / 64 - 1) / (TF_NT / 64))), dim3(TF_NT), 0, st, L,
                       U, fd, parent, list, pref, cnt, rel);
}
void launch_factor_diag(double *L, const FrontDesc *fd, const int *list, int cnt, int p0, double *Dn, int *flag,
                        double *W, const int64_t *woff, hipStream_t st) {
    if (cnt <= 0) return;
    hipLaunchKernelGGL(k_factor_diag, dim3(cnt), dim3(64), 0, st, L, fd, list, p0, Dn, flag, W, woff);
}
void launch_factor_trsm(double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                        int64_t nwg, int p0, const double *Dn, const double *W, const int64_t *woff,
                        hipStream_t st) {
    if (nwg <= 0) return;
    hipLaunchKernelGGL(k_factor_trsm, dim3((unsigned)nwg), dim3(128), 0, st, L, fd, list, pref, cnt, p0, Dn, W,
                       woff);
}
void launch_factor_update(double *L, double *U, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                          int64_t ntiles, int p0, int mode, const double *Dn, int band, hipStream_t st) {
    if (ntiles > 0)
        hipLaunchKernelGGL(k_factor_update, dim3((unsigned)((ntiles + 3) / 4)), dim3(TF_NT), 0, st, L, U, fd, list,
                           pref, cnt, p0, mode, Dn, band);
}

void launch_inv_diag(double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt, int64_t nwg,
                     hipStream_t st) {
    if (nwg > 0) hipLaunchKernelGGL(k_inv_diag, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt);
}
void launch_inv_merge(int step, double *L, const FrontDesc *fd, const int *list, const int64_t *pref, int cnt,
                      int64_t nwg, int b, double *W, const int64_t *woff, hipStream_t st) {
    if (nwg <= 0) return;
    if (step == 0)
        hipLaunchKernelGGL(k_inv_merge<0>, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, b, W, woff);
    else
        hipLaunchKernelGGL(k_inv_merge<1>, dim3((unsigned)nwg), dim3(TF_NT), 0, st, L, fd, list, pref, cnt, b, W, woff);
}

}  // namespace tf
