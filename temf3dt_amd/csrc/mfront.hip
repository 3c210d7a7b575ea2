// mfront.hip -- fused per-front products of the K right-hand-side block solve
// for the small fronts of the lowest levels (nc <= 16 * MF_MAXB columns).
//
// One workgroup per front, one wave per 16-wide block of right-hand sides.
// The factor block is streamed once per front through LDS in slabs of 16
// rows (forward, backward T) or 16 columns (backward X): all 256 threads load
// a slab with 16-byte loads into registers while the waves multiply the
// previous slab out of LDS, then store it and meet at one barrier, so each
// slab costs one memory latency shared by the four waves.  A wave keeps the
// front's own rows (nc x 16 values) in registers, in the
// v_mfma_f64_16x16x4f64 result layout: register r of row block b holds row
// 16b + 4r + l/16, right-hand side l%16.  That is also the B-operand layout of
// a k-step over rows 16b + 4r .. 16b + 4r + 3, so the rows computed by one
// product feed the next product of the same front without moving.
//
//   forward  (pull, own and struct products of one front in one pass)
//     Yin    = b[perm] - sum of the descendants' Z rows     (k_mpull's order)
//     Y_own  = Yin + strict(X11) Yin        row slabs of X11 bottom first,
//                                           in place
//     Z      = L21 Y_own                    row slabs of L21
//   backward (T and X products of one front in one pass)
//     T      = Y_own / d - L21^T x(struct)  row slabs of the row-major L21 copy
//     X      = T + strict(X11)^T T          column slabs of X11 left first,
//                                           in place
//
// The forward products and the backward T product accumulate their k-steps in
// the order of the tile kernels (k_mtile, unsplit tiles), so they give the
// same bits; the backward X product groups its k-steps by consecutive rows
// and agrees with the tile kernel to rounding.
//
// Slab loads are raw buffer loads whose whole offset is per lane, so the
// range check zeroes every read past the front's block (columns >= nc, rows
// of the L21 copy >= m); right-hand-side rows >= nc are zero in registers, so
// rows of A past the front (L21 rows read with the triangle, rows past f)
// only reach outputs that are never stored.  Every wave takes part in every
// slab load and barrier, also one whose right-hand-side block lies past K
// (its results are not stored).
#include "tf_device.h"

#define MF_MAXB 12  // largest front handled: 12 row blocks = 192 own columns
// workgroups per CU the register budget is set for (one wave per SIMD each):
// the own-row registers take 8 VGPRs per row block
#define MF_OCC(ncb) ((ncb) <= 4 ? 4 : ((ncb) <= 7 ? 3 : 2))

typedef double mf_d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ mf_d2 mf_ld2(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    double x0, x1;
    mt_bld2(r, off, 0u, x0, x1);
    return (mf_d2){x0, x1};
}

template <int NCB, bool LEAF>
__global__ void __launch_bounds__(256, MF_OCC(NCB)) k_mfront_fwd(const double *L, const FrontDesc *fd, const int *list, int K,
                                                    const double *b, const int *perm, const int *gptr,
                                                    const int *gidx, double *Z, double *Y) {
    constexpr int SC = 16 * NCB;             // columns of a slab (at most)
    constexpr int RND = (SC * 8 + 255) / 256;  // 16-byte chunks per thread and slab (8 per column)
    __shared__ __attribute__((aligned(16))) double sl[2][16 * SC];  // [column][16 rows]
    const int tid = threadIdx.x;
    const int a = __builtin_amdgcn_readfirstlane(tid >> 6);
    const FrontDesc d = fd[list[blockIdx.x]];
    const int lane = tid & 63, l16 = lane & 15, lk = lane >> 4;
    const int q = 16 * a + l16, qc = min(q, K - 1);
    const int nc = d.nc, m = d.m, ld = d.ld;
    const int nob = (m + (nc & 15) + 15) >> 4;  // 16-row output blocks of the struct product
    const int rz = nc - (nc & 15);               // block row of the first one
    const __amdgpu_buffer_rsrc_t rL = mt_rsrc(L + d.loff, (uint32_t)((int64_t)nc * ld * 8));
    // slab (row0, cols): rows row0 .. row0 + 15 of columns 0 .. cols - 1; rows
    // >= rlim read as zero (the triangle's slabs stop at row nc, so that the
    // own-row registers past nc stay zero)
    mf_d2 stg[RND];
    auto gload = [&](int row0, int cols, int rlim) {
#pragma unroll
        for (int i = 0; i < RND; i++) {
            const int ch = tid + 256 * i, c = ch >> 3, rr = row0 + 2 * (ch & 7);
            if (c < cols) {
                const mf_d2 v = mf_ld2(rL, ((uint32_t)c * (uint32_t)ld + (uint32_t)rr) * 8u);
                stg[i] = (mf_d2){rr < rlim ? v.x : 0.0, rr + 1 < rlim ? v.y : 0.0};
            }
        }
    };
    auto lstore = [&](int buf, int cols) {
#pragma unroll
        for (int i = 0; i < RND; i++) {
            const int ch = tid + 256 * i, c = ch >> 3;
            if (c < cols) *(mf_d2 *)&sl[buf][c * 16 + 2 * (ch & 7)] = stg[i];
        }
    };
    // first slab: the bottom row block of X11
    gload(16 * (NCB - 1), SC, nc);
    // Yin: b through perm, minus the descendants' Z rows in gidx order
    double y[NCB][4];
#pragma unroll
    for (int bb = 0; bb < NCB; bb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 16 * bb + 4 * r + lk;
            const double v = b[(int64_t)perm[d.c0 + min(row, nc - 1)] * K + qc];
            y[bb][r] = row < nc ? v : 0.0;
        }
    if (!LEAF) {
#pragma unroll
        for (int bb = 0; bb < NCB; bb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * bb + 4 * r + lk;
                if (row >= nc) continue;
                double s = y[bb][r];
                const int p1 = gptr[d.c0 + row + 1];
                int p = gptr[d.c0 + row];
                for (; p + 4 <= p1; p += 4) {
                    int g[4];
                    double z[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) g[u] = gidx[p + u];
#pragma unroll
                    for (int u = 0; u < 4; u++) z[u] = Z[(int64_t)g[u] * K + qc];
#pragma unroll
                    for (int u = 0; u < 4; u++) s -= z[u];
                }
                for (; p < p1; p++) s -= Z[(int64_t)gidx[p] * K + qc];
                y[bb][r] = s;
            }
    }
    lstore(0, SC);
    __syncthreads();
    // own product, bottom row block first: row block bb reads Yin of blocks
    // <= bb, which are still unchanged
#pragma unroll
    for (int bb = NCB - 1; bb >= 0; bb--) {
        const int buf = (NCB - 1 - bb) & 1;
        const bool next = bb > 0 || nob > 0;
        if (bb > 0) gload(16 * (bb - 1), 16 * bb, nc);
        else if (nob > 0) gload(rz, SC, 1 << 30);
        tf_d4 t = (tf_d4){0.0, 0.0, 0.0, 0.0};
        // LDS reads one column group (16 columns) ahead of the multiplies
        const double *sa = &sl[buf][lk * 16 + l16];
        double ac[4], an[4];
#pragma unroll
        for (int r = 0; r < 4; r++) ac[r] = sa[(4 * r) * 16];
#pragma unroll
        for (int bk = 0; bk <= bb; bk++) {
            if (bk < bb)
#pragma unroll
                for (int r = 0; r < 4; r++) an[r] = sa[(16 * (bk + 1) + 4 * r) * 16];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 4; r++) t = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[r], y[bk][r], t, 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 4; r++) ac[r] = an[r];
        }
#pragma unroll
        for (int r = 0; r < 4; r++) y[bb][r] += t[r];
        if (next) {
            lstore(buf ^ 1, bb > 0 ? 16 * bb : SC);
            __syncthreads();
        }
    }
    // Y_own out (the backward sweep reads it).  The row index is taken from
    // an opaque copy of lane / 16 so that the row tests of the gather are
    // recomputed here rather than kept live through the triangle.
    int lk2 = lk;
    asm volatile("" : "+v"(lk2));
    if (q < K)
#pragma unroll
        for (int bb = 0; bb < NCB; bb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * bb + 4 * r + lk2;
                if (row < nc) Y[(int64_t)(d.c0 + row) * K + q] = y[bb][r];
            }
    // struct product, one 16-row slab of L21 per output block: output rows
    // o = 16 ob + j - (nc & 15) (row nc + o of the block)
    for (int ob = 0; ob < nob; ob++) {
        const int buf = (NCB + ob) & 1;
        const bool next = ob + 1 < nob;
        if (next) gload(rz + 16 * (ob + 1), SC, 1 << 30);
        tf_d4 z = (tf_d4){0.0, 0.0, 0.0, 0.0};
        const double *sa = &sl[buf][lk * 16 + l16];
        double ac[4], an[4];
#pragma unroll
        for (int r = 0; r < 4; r++) ac[r] = sa[(4 * r) * 16];
#pragma unroll
        for (int bk = 0; bk < NCB; bk++) {
            if (bk + 1 < NCB)
#pragma unroll
                for (int r = 0; r < 4; r++) an[r] = sa[(16 * (bk + 1) + 4 * r) * 16];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 4; r++) z = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[r], y[bk][r], z, 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 4; r++) ac[r] = an[r];
        }
        if (q < K)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int o = 16 * ob + 4 * r + lk - (nc & 15);
                if (o >= 0 && o < m) Z[(d.soff + o) * K + q] = z[r];
            }
        if (next) {
            lstore(buf ^ 1, SC);
            __syncthreads();
        }
    }
}

template <int NCB>
__global__ void __launch_bounds__(256, MF_OCC(NCB)) k_mfront_bwd(const double *L, const double *LT, const FrontDesc *fd,
                                                    const int *list, int K, const double *Y, const double *Dn,
                                                    const int *sidx, double *xw, const int *perm, double *x1,
                                                    double *x2, int own_xw) {
    // T slabs: 16 struct rows of the row-major L21 copy x 16 NCB columns, row
    // stride SJ (an odd multiple of 16 doubles: the four rows a k-step reads
    // fall in different banks); X slabs: 16 columns of X11 x rows 16 bb ..
    // 16 NCB - 1, column stride SI (2 more than a multiple of 32 doubles: the
    // 16 columns a k-step reads fall in different banks)
    constexpr int SC = 16 * NCB;
    constexpr int SJ = 16 * (NCB | 1);
    constexpr int SI = (SC + 31) / 32 * 32 + 2;
    constexpr int SLAB = (16 * SJ > 16 * SI) ? 16 * SJ : 16 * SI;
    constexpr int RND = (SC * 8 + 255) / 256;  // 16-byte chunks per thread and slab
    __shared__ __attribute__((aligned(16))) double sl[2][SLAB];
    const int tid = threadIdx.x;
    const int a = __builtin_amdgcn_readfirstlane(tid >> 6);
    const FrontDesc d = fd[list[blockIdx.x]];
    const int lane = tid & 63, l16 = lane & 15, lk = lane >> 4;
    const int q = 16 * a + l16, qc = min(q, K - 1);
    const int nc = d.nc, m = d.m, ld = d.ld;
    const int ns1 = (m + 15) >> 4;  // T slabs
    const __amdgpu_buffer_rsrc_t rT = mt_rsrc(LT + d.toff, (uint32_t)((int64_t)m * d.ldt * 8));
    const __amdgpu_buffer_rsrc_t rL = mt_rsrc(L + d.loff, (uint32_t)((int64_t)nc * ld * 8));
    const int *si = sidx + d.soff;
    mf_d2 stg[RND];
    // T slab s: struct rows 16 s .. 16 s + 15, columns 0 .. SC - 1 (8 chunks per 16 columns of a row)
    auto gload_t = [&](int s) {
#pragma unroll
        for (int i = 0; i < RND; i++) {
            const int ch = tid + 256 * i, row = ch / (SC / 2), col = 2 * (ch % (SC / 2));
            if (row < 16)
                stg[i] = mf_ld2(rT, ((uint32_t)(16 * s + row) * (uint32_t)d.ldt + (uint32_t)col) * 8u);
        }
    };
    auto lstore_t = [&](int buf) {
#pragma unroll
        for (int i = 0; i < RND; i++) {
            const int ch = tid + 256 * i, row = ch / (SC / 2), col = 2 * (ch % (SC / 2));
            if (row < 16) *(mf_d2 *)&sl[buf][row * SJ + col] = stg[i];
        }
    };
    // X slab bb: columns 16 bb .. 16 bb + 15, rows 16 bb .. SC - 1
    auto gload_x = [&](int bb) {
        const int h = (SC - 16 * bb) / 2;  // chunks per column
#pragma unroll
        for (int i = 0; i < RND; i++) {
            const int ch = tid + 256 * i, jj = ch / h, rr = 2 * (ch % h);
            if (jj < 16)
                stg[i] = mf_ld2(rL, ((uint32_t)(16 * bb + jj) * (uint32_t)ld + (uint32_t)(16 * bb + rr)) * 8u);
        }
    };
    auto lstore_x = [&](int buf, int bb) {
        const int h = (SC - 16 * bb) / 2;
#pragma unroll
        for (int i = 0; i < RND; i++) {
            const int ch = tid + 256 * i, jj = ch / h, rr = 2 * (ch % h);
            if (jj < 16) *(mf_d2 *)&sl[buf][jj * SI + rr] = stg[i];
        }
    };
    // B operands of a T slab: x at struct rows 16 s + 4 r + l/16
    double bx[4], bn[4];
    auto gload_b = [&](int s, double (&o)[4]) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 16 * s + 4 * r + lk;
            const double v = xw[(int64_t)si[min(row, m - 1)] * K + qc];
            o[r] = row < m ? v : 0.0;
        }
    };
    tf_d4 t[NCB];
#pragma unroll
    for (int bb = 0; bb < NCB; bb++) t[bb] = (tf_d4){0.0, 0.0, 0.0, 0.0};
    int buf = 0;
    if (ns1 > 0) {
        gload_t(0);
        gload_b(0, bx);
        lstore_t(0);
    } else {
        gload_x(0);
        lstore_x(0, 0);
    }
    __syncthreads();
    // T product over the struct rows: A(j, o) = LT[o][j], B = x at the struct rows
    for (int s = 0; s < ns1; s++) {
        const bool more = s + 1 < ns1;
        if (more) {
            gload_t(s + 1);
            gload_b(s + 1, bn);
        } else {
            gload_x(0);
        }
        const double *sa = &sl[buf][lk * SJ + l16];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            double ac[NCB];
#pragma unroll
            for (int bb = 0; bb < NCB; bb++) ac[bb] = sa[4 * r * SJ + 16 * bb];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int bb = 0; bb < NCB; bb++) t[bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[bb], bx[r], t[bb], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (more) {
            lstore_t(buf ^ 1);
#pragma unroll
            for (int r = 0; r < 4; r++) bx[r] = bn[r];
        } else {
            lstore_x(buf ^ 1, 0);
        }
        buf ^= 1;
        __syncthreads();
    }
    // T = Y_own / d - s; rows >= nc to zero
    double x[NCB][4];
#pragma unroll
    for (int bb = 0; bb < NCB; bb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int j = 16 * bb + 4 * r + lk;
            x[bb][r] = j < nc ? Y[(int64_t)(d.c0 + j) * K + qc] / Dn[d.doff + j] - t[bb][r] : 0.0;
        }
    // X product, left column slab first: block bb reads T of blocks >= bb.
    // A(j, i) = L[j][i] (column j = 16 bb + l16, row i = 16 bi + 4 r + l/16)
#pragma unroll
    for (int bb = 0; bb < NCB; bb++) {
        const bool next = bb + 1 < NCB;
        if (next) gload_x(bb + 1);
        tf_d4 s = (tf_d4){0.0, 0.0, 0.0, 0.0};
        const double *sa = &sl[buf][l16 * SI + lk];
        double ac[4], an[4];
#pragma unroll
        for (int r = 0; r < 4; r++) ac[r] = sa[4 * r];
#pragma unroll
        for (int bi = bb; bi < NCB; bi++) {
            if (bi + 1 < NCB)
#pragma unroll
                for (int r = 0; r < 4; r++) an[r] = sa[16 * (bi + 1 - bb) + 4 * r];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 4; r++) s = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[r], x[bi][r], s, 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < 4; r++) ac[r] = an[r];
        }
#pragma unroll
        for (int r = 0; r < 4; r++) x[bb][r] += s[r];
        if (next) {
            lstore_x(buf ^ 1, bb + 1);
            buf ^= 1;
            __syncthreads();
        }
    }
    int lk2 = lk;
    asm volatile("" : "+v"(lk2));
    if (q < K)
#pragma unroll
        for (int bb = 0; bb < NCB; bb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int j = 16 * bb + 4 * r + lk2;
                if (j >= nc) continue;
                const double v = x[bb][r];
                if (own_xw) xw[(int64_t)(d.c0 + j) * K + q] = v;
                const int64_t e = (int64_t)perm[d.c0 + j] * K + q;
                x1[e] = v;
                if (x2) x2[e] = v;
            }
}

namespace tf {
int mfront_max_blocks() { return MF_MAXB; }

// ncb: row blocks of the launch's fronts (ceil(nc / 16) <= ncb); leaf: fronts
// without descendants (no pull)
void launch_mfront_fwd(int ncb, bool leaf, const double *L, const FrontDesc *fd, const int *list, int cnt, int K,
                       const double *b, const int *perm, const int *gptr, const int *gidx, double *Z, double *Y,
                       hipStream_t st) {
    if (cnt <= 0) return;
    const dim3 g(cnt), blk(256);
#define TF_MFF(N)                                                                                                 \
    case N:                                                                                                       \
        if (leaf) hipLaunchKernelGGL((k_mfront_fwd<N, true>), g, blk, 0, st, L, fd, list, K, b, perm, gptr, gidx, Z, Y); \
        else hipLaunchKernelGGL((k_mfront_fwd<N, false>), g, blk, 0, st, L, fd, list, K, b, perm, gptr, gidx, Z, Y); \
        break;
    switch (ncb) {
        TF_MFF(1) TF_MFF(2) TF_MFF(3) TF_MFF(4) TF_MFF(5) TF_MFF(6)
        TF_MFF(7) TF_MFF(8) TF_MFF(9) TF_MFF(10) TF_MFF(11) TF_MFF(12)
    default: break;
    }
#undef TF_MFF
}

void launch_mfront_bwd(int ncb, const double *L, const double *LT, const FrontDesc *fd, const int *list, int cnt,
                       int K, const double *Y, const double *Dn, const int *sidx, double *xw, const int *perm,
                       double *x1, double *x2, int own_xw, hipStream_t st) {
    if (cnt <= 0) return;
    const dim3 g(cnt), blk(256);
#define TF_MFB(N)                                                                                                  \
    case N:                                                                                                        \
        hipLaunchKernelGGL((k_mfront_bwd<N>), g, blk, 0, st, L, LT, fd, list, K, Y, Dn, sidx, xw, perm, x1, x2, own_xw); \
        break;
    switch (ncb) {
        TF_MFB(1) TF_MFB(2) TF_MFB(3) TF_MFB(4) TF_MFB(5) TF_MFB(6)
        TF_MFB(7) TF_MFB(8) TF_MFB(9) TF_MFB(10) TF_MFB(11) TF_MFB(12)
    default: break;
    }
#undef TF_MFB
}
}  // namespace tf
