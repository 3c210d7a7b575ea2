// mtile.hip -- wave-level FP64 matrix-core tile products of the K-RHS block
// solve (K <= 64 sources sharing one LDL^T factor).
//
// Each wave owns one work unit: a (16 * NB) x 64 output tile (NB = 2: 32
// matrix rows or columns x 64 right-hand sides; NB = 4: 64) accumulated over
// an inner range [k0, k1).  Operands go straight from memory into the MFMA lane
// layout -- no LDS, no barriers -- so many waves per CU stream the factor at
// once.  32-row tiles take about 120 VGPRs (four waves per SIMD) and suit
// the many small fronts of the lower levels; 64-row tiles (about 200
// VGPRs, two waves per SIMD) read each right-hand-side row for twice the
// output and suit the few large fronts of the upper levels.  The four
// products of a front (X11 = L11^-1 stored in place by the factorisation):
//   forward   Y_own = Yin + strict(X11) Yin           (N form, triangle)
//             Z = L21 Y_own                          (N form)
//   backward  T = Y_own / d - L21^T x(struct rows)   (T form, gathered B)
//             X_own = T + strict(X11)^T T            (T form, triangle)
// Yin (own rows of the right-hand side after the descendants' updates) is
// formed by k_mpull from b and the Z rows of the descendants: the forward
// update is written once per front and pulled once, not accumulated front
// to front.  N form: output index runs down a column of L (16 consecutive
// rows per lane group, 128-byte loads).  T form: output index runs across
// columns; a lane reads two consecutive rows of one column per stage.
//
// v_mfma_f64_16x16x4f64 lane layout (measured): lane l supplies A[l%16][l/16]
// and B[l/16][l%16] and receives D[4r + l/16][l%16] in result register r.
// A carries the matrix (output index = l%16), B the right-hand sides.  The
// operands are fetched two values per lane (16-byte loads, half the load
// instructions of 8-byte ones): in the N form a lane reads two consecutive
// rows of a column, which go to the output blocks 2p and 2p + 1, and with
// NA even a lane reads two consecutive right-hand sides, which go to the
// blocks 2c and 2c + 1.  Output block b therefore holds the rows (columns in
// the T form, which keeps 16 consecutive outputs per block)
//   N form: 32 (b/2) + 2 i + b%2,   T form: 16 b + i       (i = 0..15)
// and right-hand-side block a the columns 32 (a/2) + 2 j + a%2 (NA even) or
// j (NA = 1), so result register r of block (a, b) is output i = 4r + l/16,
// right-hand side j = l%16 of those blocks.
//
// Units whose tile needs several inner chunks write a partial tile
// (part[slot][out][rhs]); a reduction launch sums the chunks in a fixed
// order, so the result does not depend on scheduling.
#include "tf_device.h"

// MT_OWN_FWD_B: the own forward product of a front without children, whose
// Yin is the right-hand side itself: B rows are read from b through perm
// (no pull launch for the lowest level).
// Fronts that keep the partitioned inverse (FrontDesc::pinv; L block
// [strict(L11^-1); W], W = -L21 L11^-1) take one product per direction:
//   MT_PIN_FWD    rows 0 .. f of the block times Yin: rows < nc give
//                 Y_own = Yin + s, stored as Y_own / d; rows >= nc give
//                 Z = -s (the update of the ancestors' rows)
//   MT_PIN_FWD_B  the same for a front without children (Yin = b through perm)
//   MT_PIN_BWD    X_own = v + [strict(L11^-1); W]^T [v; x(struct)], v = Y_own / d,
//                 the inner rows gathered through the front's row list (own
//                 rows from v, struct rows from x)
enum { MT_OWN_FWD = 0, MT_STR_FWD = 1, MT_BWD_T = 2, MT_BWD_X = 3, MT_OWN_FWD_B = 4, MT_PIN_FWD = 5,
       MT_PIN_FWD_B = 6, MT_PIN_BWD = 7 };

template <int NA, int NB>
__device__ __forceinline__ void mt_acc_zero(tf_d4 acc[NA][NB]) {
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) acc[a][b] = (tf_d4){0.0, 0.0, 0.0, 0.0};
}

// Output-row/column count, matrix offset inside the L block, and whether the
// (strict) triangle applies, per mode.
template <int MODE>
__device__ __forceinline__ void mt_geom(const FrontDesc &d, int &nout, int &roff) {
    if (MODE == MT_OWN_FWD || MODE == MT_OWN_FWD_B || MODE == MT_BWD_X || MODE == MT_PIN_BWD) { nout = d.nc; roff = 0; }
    else if (MODE == MT_PIN_FWD || MODE == MT_PIN_FWD_B) { nout = d.f; roff = 0; }
    else if (MODE == MT_STR_FWD) { nout = d.m; roff = d.nc; }
    else { nout = d.nc; roff = d.nc; }
}

// Epilogue for one output element (o = output row/column of the front,
// q = rhs); `s` is the accumulated product.
// Backward X also scatters the finished rows to edge order: x1 (and x2, if
// given) at row perm[c0 + o].
struct MtEdgeOut {
    const int *perm;
    double *x1, *x2;
    double *z;  // partitioned forward: the front's Z rows (at soff)
};

template <int MODE>
__device__ __forceinline__ void mt_store(const FrontDesc &d, int o, int q, double s, int K, const double *in0,
                                         const double *in1, const double *Dn, double *out, const MtEdgeOut &xo) {
    if (MODE == MT_OWN_FWD) {  // Y[c0+o] = Yin[c0+o] + s
        const int64_t oo = (int64_t)(d.c0 + o) * K + q;
        out[oo] = in0[oo] + s;
    } else if (MODE == MT_OWN_FWD_B) {  // Y[c0+o] = b[perm[c0+o]] + s
        out[(int64_t)(d.c0 + o) * K + q] = in0[(int64_t)xo.perm[d.c0 + o] * K + q] + s;
    } else if (MODE == MT_STR_FWD) {  // Z[soff+o] = s (subtracted where it is pulled)
        out[(d.soff + o) * K + q] = s;
    } else if (MODE == MT_BWD_T) {  // T = Y/d - s
        const int64_t oo = (int64_t)(d.c0 + o) * K + q;
        out[oo] = in1[oo] / Dn[d.doff + o] - s;
    } else if (MODE == MT_PIN_FWD || MODE == MT_PIN_FWD_B) {  // v = (Yin + s) / d, or Z = -s
        if (o < d.nc) {
            const int64_t oo = (int64_t)(d.c0 + o) * K + q;
            const double yin = MODE == MT_PIN_FWD ? in0[oo] : in0[(int64_t)xo.perm[d.c0 + o] * K + q];
            out[oo] = (yin + s) / Dn[d.doff + o];
        } else {
            xo.z[(d.soff + o - d.nc) * K + q] = -s;
        }
    } else {  // X = T + s (front order unless out is null: rows no lower level gathers); PIN: X = v + s
        const int64_t oo = (int64_t)(d.c0 + o) * K + q;
        const double v = in1[oo] + s;
        if (out) out[oo] = v;
        if (xo.x1) {
            const int64_t e = (int64_t)xo.perm[d.c0 + o] * K + q;
            xo.x1[e] = v;
            if (xo.x2) xo.x2[e] = v;
        }
    }
}

// The same epilogue for the right-hand sides q and q + 1 at once (q even, K
// even: every vector offset is even, so the pair is one 16-byte access).
typedef double mt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ mt_d2 mt_ld2(const double *p) { return *(const mt_d2 *)p; }
__device__ __forceinline__ void mt_st2(double *p, mt_d2 v) { *(mt_d2 *)p = v; }

template <int MODE>
__device__ __forceinline__ void mt_store2(const FrontDesc &d, int o, int q, mt_d2 s, int K, const double *in0,
                                          const double *in1, const double *Dn, double *out, const MtEdgeOut &xo) {
    if (MODE == MT_OWN_FWD) {
        const int64_t oo = (int64_t)(d.c0 + o) * K + q;
        mt_st2(out + oo, mt_ld2(in0 + oo) + s);
    } else if (MODE == MT_OWN_FWD_B) {
        mt_st2(out + (int64_t)(d.c0 + o) * K + q, mt_ld2(in0 + (int64_t)xo.perm[d.c0 + o] * K + q) + s);
    } else if (MODE == MT_STR_FWD) {
        mt_st2(out + (d.soff + o) * K + q, s);
    } else if (MODE == MT_BWD_T) {
        const int64_t oo = (int64_t)(d.c0 + o) * K + q;
        mt_st2(out + oo, mt_ld2(in1 + oo) / Dn[d.doff + o] - s);
    } else if (MODE == MT_PIN_FWD || MODE == MT_PIN_FWD_B) {
        if (o < d.nc) {
            const int64_t oo = (int64_t)(d.c0 + o) * K + q;
            const mt_d2 yin = MODE == MT_PIN_FWD ? mt_ld2(in0 + oo) : mt_ld2(in0 + (int64_t)xo.perm[d.c0 + o] * K + q);
            mt_st2(out + oo, (yin + s) / Dn[d.doff + o]);
        } else {
            mt_st2(xo.z + (d.soff + o - d.nc) * K + q, -s);
        }
    } else {
        const int64_t oo = (int64_t)(d.c0 + o) * K + q;
        const mt_d2 v = mt_ld2(in1 + oo) + s;
        if (out) mt_st2(out + oo, v);
        if (xo.x1) {
            const int64_t e = (int64_t)xo.perm[d.c0 + o] * K + q;
            mt_st2(xo.x1 + e, v);
            if (xo.x2) mt_st2(xo.x2 + e, v);
        }
    }
}

// Byte-offset load from a wave-uniform base: the base stays in scalar
// registers and each lane adds a 32-bit offset (no 64-bit address math per
// load).
__device__ __forceinline__ double mt_ld(const double *base, uint32_t off) {
    return *(const double *)((const char *)base + off);
}

// One wave: acc(out, rhs) = sum_{k0 <= i < k1} A(out, i) B(i, rhs) where
//   N form: A(out, i) = Lf[i*ld + roff + out]   (column i, row out)
//   T form: A(out, i) = Lf[out*ld + roff + i]   (column out, row i)
// B(i, rhs) = Bm[row(i)*K + rhs] with row(i) = i, or bidx[i] when GATH.
// The strict triangle of X11 needs no mask (the diagonal and upper part of
// L11's storage are zero).  The inner range is walked in stages of KS
// k-steps (4*KS inner values; a T-form stage starts on a 16-row line of the
// column).
//
// The factor block and the front's own right-hand-side rows are read with
// raw buffer loads: per-lane offsets in vector registers, the stage offset
// in a scalar register, and a read whose per-lane offset lies past the end
// of the block returns zero.  Output rows/columns and right-hand-side
// columns outside the tile's valid range are therefore read without masks
// (they land in neighbouring rows or past the block) and their results are
// never stored.  Stages inside [k0, k1) form one contiguous, unmasked run,
// pipelined in pairs of stages: the next stage's loads are issued (the last
// pair re-reads the final stage) before the current one is multiplied, so
// each wait covers only the stage being multiplied; gathered rows are
// fetched before the multiply that precedes their stage's loads.  The
// stages before and after that run (a T-form stage aligned before k0, a
// partial last stage) clamp the inner index and zero what lies outside.
template <bool TFORM, bool GATH, int KS, int NB, int NA>
__device__ __forceinline__ void mt_wave(const double *Lf, uint32_t lbytes, int f, int roff, int o0, int k0, int k1,
                                        const double *Bm, uint32_t bbytes, const int *bidx, int K, tf_d4 acc[NA][NB]) {
    constexpr int SW = 4 * KS;  // inner values per stage
    if (k0 >= k1) return;
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
    const __amdgpu_buffer_rsrc_t rL = mt_rsrc(Lf, lbytes), rR = mt_rsrc(Bm, bbytes);
    // lane offsets (bytes) at the stage origin.  Only the per-lane offset and
    // the immediate take part in a buffer load's range check, so whatever
    // can leave the block stays there: the T-form output column (one offset
    // per output block b), while an N-form row past the block's last column
    // reads at most 62 values beyond it (the factor storage ends with slack).
    constexpr bool PB = NA % 2 == 0;  // right-hand sides in pairs
    const int jq = PB ? 2 * l16 : l16;  // first right-hand side of the lane
    uint32_t la[TFORM ? NB : 1];
#pragma unroll
    for (int b = 0; b < (TFORM ? NB : 1); b++)
        la[b] = TFORM ? (uint32_t)((int64_t)(o0 + 16 * b + l16) * f + roff + KS * lk) * 8u
                      : (uint32_t)(lk * f + roff + o0 + 2 * l16) * 8u;
    const uint32_t ba = GATH ? (uint32_t)jq * 8u : (uint32_t)((TFORM ? KS * lk : lk) * K + jq) * 8u;
    const int kb = TFORM ? (k0 - ((roff + k0) & 15)) : k0;
    const int nst = (k1 - kb + SW - 1) / SW;
    // the unmasked run of stages [ja, jb)
    const int ja = min(nst, (k0 - kb + SW - 1) / SW);
    const int jb = max(ja, (k1 - kb) / SW);
    auto inner = [&](int ib, int s) { return TFORM ? (ib + KS * lk + s) : (ib + 4 * s + lk); };
    auto mult = [&](const double xa[KS][NB], const double yb[KS][NA]) {
#pragma unroll
        for (int s = 0; s < KS; s++)
#pragma unroll
            for (int a = 0; a < NA; a++)
#pragma unroll
                for (int b = 0; b < NB; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[s][b], yb[s][a], acc[a][b], 0, 0, 0);
    };
    // a masked stage, loaded and multiplied on its own
    auto masked_stage = [&](int js) {
        const int ib = kb + js * SW;
        double xa[KS][NB], yb[KS][NA];
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int i = inner(ib, s);
            const bool iv = i >= k0 && i < k1;
            const int ic = min(max(i, k0), k1 - 1);
            if (TFORM) {
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    const double x = mt_bld(rL, (uint32_t)((int64_t)(o0 + 16 * b + l16) * f + roff + ic) * 8u, 0u);
                    xa[s][b] = iv ? x : 0.0;
                }
            } else {  // the rows of a pair share the inner index, hence the mask
#pragma unroll
                for (int p = 0; p < NB / 2; p++) {
                    double x0, x1;
                    mt_bld2(rL, (uint32_t)((int64_t)ic * f + roff + o0 + 32 * p + 2 * l16) * 8u, 0u, x0, x1);
                    xa[s][2 * p] = iv ? x0 : 0.0;
                    xa[s][2 * p + 1] = iv ? x1 : 0.0;
                }
            }
            const uint32_t row = (uint32_t)(GATH ? bidx[ic] : ic);
            if (PB) {
#pragma unroll
                for (int c = 0; c < NA / 2; c++) {
                    double y0, y1;
                    mt_bld2(rR, (row * (uint32_t)K + (uint32_t)(32 * c + jq)) * 8u, 0u, y0, y1);
                    yb[s][2 * c] = iv ? y0 : 0.0;
                    yb[s][2 * c + 1] = iv ? y1 : 0.0;
                }
            } else {
#pragma unroll
                for (int a = 0; a < NA; a++) {
                    const double y = mt_bld(rR, (row * (uint32_t)K + (uint32_t)(16 * a + l16)) * 8u, 0u);
                    yb[s][a] = iv ? y : 0.0;
                }
            }
        }
        mult(xa, yb);
    };
    if (TFORM)  // an N-form run starts at k0 itself (ja = 0)
        for (int js = 0; js < ja; js++) masked_stage(js);
    if (ja < jb) {
        auto ldidx = [&](int js, int r[KS]) {
            const int ib = kb + js * SW;
#pragma unroll
            for (int s = 0; s < KS; s++) r[s] = bidx[inner(ib, s)];
        };
        auto load = [&](int js, const int r[KS], double xa[KS][NB], double yb[KS][NA]) {
            const int ib = kb + js * SW;
            if (TFORM && KS % 2 == 0) {  // a lane's KS rows of a column are contiguous: 16-byte loads
#pragma unroll
                for (int s = 0; s < KS; s += 2)
#pragma unroll
                    for (int b = 0; b < NB; b++) mt_bld2(rL, la[b], (uint32_t)(ib + s) * 8u, xa[s][b], xa[s + 1][b]);
            }
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const uint32_t sl = (uint32_t)(TFORM ? (ib + s) : (ib + 4 * s) * f) * 8u;
                if (TFORM && KS % 2 != 0)
#pragma unroll
                    for (int b = 0; b < NB; b++) xa[s][b] = mt_bld(rL, la[b], sl);
                if (!TFORM)  // rows 2 l16 and 2 l16 + 1 of each 32-row pair of blocks
#pragma unroll
                    for (int p = 0; p < NB / 2; p++) mt_bld2(rL, la[0] + 256u * p, sl, xa[s][2 * p], xa[s][2 * p + 1]);
                const uint32_t vo = GATH ? ((uint32_t)r[s] * (uint32_t)K) * 8u + ba : ba;
                const uint32_t sb = GATH ? 0u : (uint32_t)((TFORM ? ib + s : ib + 4 * s) * K) * 8u;
                if (PB)
#pragma unroll
                    for (int c = 0; c < NA / 2; c++) mt_bld2(rR, vo + 256u * c, sb, yb[s][2 * c], yb[s][2 * c + 1]);
                else
#pragma unroll
                    for (int a = 0; a < NA; a++) yb[s][a] = mt_bld(rR, vo + 128u * a, sb);
            }
        };
        double xa0[KS][NB], yb0[KS][NA], xa1[KS][NB], yb1[KS][NA];
        int rA[KS], rB[KS];
        const int last = jb - 1;
        int js = ja;
        if ((jb - ja) & 1) {  // odd run: its first stage on its own, the rest in pairs
            if (GATH) ldidx(js, rA);
            load(js, rA, xa0, yb0);
            mult(xa0, yb0);
            js++;
        }
        if (js < jb) {
            if (GATH) {
                ldidx(js, rA);
                ldidx(js + 1, rB);
            }
            load(js, rA, xa0, yb0);
            for (; js < jb; js += 2) {  // one exit per pair; js + 1 <= last
                load(js + 1, rB, xa1, yb1);
                if (GATH) ldidx(min(js + 2, last), rA);
                __builtin_amdgcn_sched_barrier(0);
                mult(xa0, yb0);
                __builtin_amdgcn_sched_barrier(0);
                load(min(js + 2, last), rA, xa0, yb0);
                if (GATH) ldidx(min(js + 3, last), rB);
                __builtin_amdgcn_sched_barrier(0);
                mult(xa1, yb1);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    for (int js = jb; js < nst; js++) masked_stage(js);
}

template <int MODE, int KS, int NB, int NA>
__global__ void __launch_bounds__(256, 2) k_mtile(const double *L, const FrontDesc *fd, const TileUnit *units,
                                                  int nunits, int K, const double *in0, const double *in1,
                                                  const double *Dn, const int *sidx, double *out, double *part,
                                                  uint32_t gbytes, MtEdgeOut xo, const double *LT) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int u = xcd_block(blockIdx.x, gridDim.x) * 4 + w;
    if (u >= nunits) return;
    const TileUnit tu = units[u];
    const FrontDesc d = fd[tu.front];
    int nout, roff;
    mt_geom<MODE>(d, nout, roff);
    tf_d4 acc[NA][NB];
    mt_acc_zero<NA, NB>(acc);
    const double *Lf = L + d.loff;
    const uint32_t lbytes = (uint32_t)((int64_t)d.nc * d.ld * 8);
    // B operand rows: own right-hand sides (Yin, Y_own, T) at c0, or the
    // solution rows of the struct indices (backward T)
    const double *Bm = in0 + (int64_t)d.c0 * K;
    const uint32_t bbytes = (uint32_t)((int64_t)d.nc * K * 8);  // the front's own right-hand-side rows
    // the strict triangle of X11 needs no mask: the factorisation leaves the
    // diagonal and upper part of L11's storage zero
    // T form (output index across the columns of L): the backward products
    // that read L itself rather than its row-major copy
    const bool tform = (MODE == MT_BWD_T || MODE == MT_PIN_BWD) ? LT == nullptr : MODE == MT_BWD_X;
    if (MODE == MT_OWN_FWD || MODE == MT_STR_FWD || MODE == MT_PIN_FWD)
        mt_wave<false, false, KS, NB, NA>(Lf, lbytes, d.ld, roff, tu.t0, tu.k0, tu.k1, Bm, bbytes, nullptr, K, acc);
    else if (MODE == MT_OWN_FWD_B || MODE == MT_PIN_FWD_B)
        mt_wave<false, true, KS, NB, NA>(Lf, lbytes, d.ld, roff, tu.t0, tu.k0, tu.k1, in0, gbytes, xo.perm + d.c0, K, acc);
    else if (MODE == MT_PIN_BWD && LT)  // all f rows from the row-major copy, gathered [v; x] rows (sidx: row lists)
        mt_wave<false, true, KS, NB, NA>(LT + d.toff, (uint32_t)((int64_t)d.f * d.ldt * 8), d.ldt, 0, tu.t0, tu.k0,
                                     tu.k1, in0, gbytes, sidx + d.voff, K, acc);
    else if (MODE == MT_PIN_BWD)
        mt_wave<true, true, KS, NB, NA>(Lf, lbytes, d.ld, 0, tu.t0, tu.k0, tu.k1, in0, gbytes, sidx + d.voff, K, acc);
    else if (MODE == MT_BWD_T && LT)  // L21 read from its row-major copy: N form over struct rows
        mt_wave<false, true, KS, NB, NA>(LT + d.toff, (uint32_t)((int64_t)d.m * d.ldt * 8), d.ldt, 0, tu.t0, tu.k0,
                                     tu.k1, in0, gbytes, sidx + d.soff, K, acc);
    else if (MODE == MT_BWD_T)
        mt_wave<true, true, KS, NB, NA>(Lf, lbytes, d.ld, roff, tu.t0, tu.k0, tu.k1, in0, gbytes, sidx + d.soff, K, acc);
    else
        mt_wave<true, false, KS, NB, NA>(Lf, lbytes, d.ld, roff, tu.t0, tu.k0, tu.k1, Bm, bbytes, nullptr, K, acc);
    const int lane = threadIdx.x & 63, l16 = lane & 15, lk = lane >> 4;
    // output (relative to t0) of register r of output block b, and right-hand
    // side of block a: the operand pairing of mt_wave
    auto orel = [&](int b, int r) { return tform ? 16 * b + 4 * r + lk : 32 * (b >> 1) + 2 * (4 * r + lk) + (b & 1); };
    auto qof = [&](int a) { return NA % 2 == 0 ? 32 * (a >> 1) + 2 * l16 + (a & 1) : 16 * a + l16; };
    if (tu.slot < 0) {
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int o = tu.t0 + orel(b, r);
                if (o < 0 || o >= nout) continue;
                if (NA % 2 == 0 && (K & 1) == 0) {  // blocks a, a + 1 hold right-hand sides q, q + 1
#pragma unroll
                    for (int a = 0; a < NA; a += 2) {
                        const int q = qof(a);
                        if (q < K)
                            mt_store2<MODE>(d, o, q, (mt_d2){acc[a][b][r], acc[a + 1][b][r]}, K, in0, in1, Dn, out, xo);
                    }
                } else {
#pragma unroll
                    for (int a = 0; a < NA; a++) {
                        const int q = qof(a);
                        if (q < K) mt_store<MODE>(d, o, q, acc[a][b][r], K, in0, in1, Dn, out, xo);
                    }
                }
            }
    } else {
        // partial tile: 64 outputs x 16 NA right-hand sides
        double *pp = part + (int64_t)tu.slot * (64 * 16 * NA);
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (NA % 2 == 0)
#pragma unroll
                    for (int a = 0; a < NA; a += 2)
                        mt_st2(pp + orel(b, r) * (16 * NA) + qof(a), (mt_d2){acc[a][b][r], acc[a + 1][b][r]});
                else
#pragma unroll
                    for (int a = 0; a < NA; a++) pp[orel(b, r) * (16 * NA) + qof(a)] = acc[a][b][r];
    }
}

// Sum of a tile's chunk partials (slot0 .. slot0+nk-1, in order) + epilogue.
// Reduction units reuse TileUnit: front, t0, k0 = slot0, k1 = nk.  A partial
// tile holds 64 outputs x PS right-hand sides (PS = 16 NA); 64 PS / 256
// workgroups per tile, one element per thread.
// With P2 (PS and K even) a thread sums two neighbouring right-hand sides
// with 16-byte loads (64 PS / 512 workgroups per tile).
template <int MODE, int PS, bool P2>
__global__ void __launch_bounds__(256) k_mtile_red(const FrontDesc *fd, const TileUnit *units, int K,
                                                   const double *in0, const double *in1, const double *Dn,
                                                   double *out, const double *part, MtEdgeOut xo, int nb) {
    constexpr int EPT = P2 ? 2 : 1;               // elements per thread
    constexpr int WPT = 64 * PS / (256 * EPT);    // workgroups per tile
    const TileUnit tu = units[blockIdx.x / WPT];
    const FrontDesc d = fd[tu.front];
    int nout, roff;
    mt_geom<MODE>(d, nout, roff);
    const int e = ((blockIdx.x % WPT) * 256 + threadIdx.x) * EPT;
    const int o = tu.t0 + e / PS, q = e % PS;
    if (e / PS >= 16 * nb || o < 0 || o >= nout || q >= K) return;
    const double *p0 = part + (int64_t)tu.k0 * (64 * PS) + e;
    if (P2) {
        mt_d2 s = (mt_d2){0.0, 0.0};
        for (int c = 0; c < tu.k1; c++) s += mt_ld2(p0 + (int64_t)c * (64 * PS));
        mt_store2<MODE>(d, o, q, s, K, in0, in1, Dn, out, xo);
    } else {
        double s = 0.0;
        for (int c = 0; c < tu.k1; c++) s += p0[(int64_t)c * (64 * PS)];
        mt_store<MODE>(d, o, q, s, K, in0, in1, Dn, out, xo);
    }
}

// ---- one right-hand side on the vector ALUs ------------------------------------------
// The same work units, products and epilogues as k_mtile with K = 1, as
// matrix-vector tiles (one wave per unit, 16-byte loads, no LDS, no
// barriers).  N form (own and struct forward): a tile of 64 output rows; a
// lane reads two consecutive rows of one column per load, lanes 0-31 the even
// columns of a step and lanes 32-63 the odd ones, which are folded together
// at the end.  T form (backward T and X): a tile of 64 output columns, taken
// 16 at a time; a lane reads two consecutive rows of each of the 16 columns
// (128 rows per block), and the 16 column sums are reduced across the wave by
// butterfly shuffles.  Every matrix value outside the unit's inner range is
// zeroed before use; per-lane offsets past the L block read zero (buffer
// range check), so columns past the front read nothing.
template <int MODE>
__device__ __forceinline__ double st_b(const FrontDesc &d, int k, const double *in0, const double *in1,
                                       const int *sidx, const int *perm) {
    if (MODE == MT_OWN_FWD_B || MODE == MT_PIN_FWD_B) return in0[perm[d.c0 + k]];
    if (MODE == MT_BWD_T) return in0[sidx[d.soff + k]];
    if (MODE == MT_PIN_BWD) return k < d.nc ? in1[d.c0 + k] : in0[sidx[d.soff + k - d.nc]];  // [v; x(struct)]
    return in0[d.c0 + k];
}

template <int MODE>
__device__ __forceinline__ void st_out(const FrontDesc &d, const TileUnit &tu, int o, double s, int nout,
                                       const double *in0, const double *in1, const double *Dn, double *out,
                                       double *part, const MtEdgeOut &xo) {
    if (o < 0 || o >= nout) return;
    if (tu.slot < 0) mt_store<MODE>(d, o, 0, s, 1, in0, in1, Dn, out, xo);
    else part[(int64_t)tu.slot * 1024 + (o - tu.t0) * 16] = s;  // partial tile layout of NA = 1
}

template <int MODE>
__global__ void __launch_bounds__(256) k_stile_n(const double *L, const FrontDesc *fd, const TileUnit *units,
                                                 int nunits, const double *in0, const double *Dn, double *out,
                                                 double *part, MtEdgeOut xo) {
    constexpr int U = 8;  // column pairs per group (four groups per 64 columns)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int u = xcd_block(blockIdx.x, gridDim.x) * 4 + w;
    if (u >= nunits) return;
    const TileUnit tu = units[u];
    const FrontDesc d = fd[tu.front];
    int nout, roff;
    mt_geom<MODE>(d, nout, roff);
    const int lane = threadIdx.x & 63, cp = lane >> 5;
    const int o = tu.t0 + 2 * (lane & 31);  // rows o, o + 1 (of the block's rows from roff)
    const __amdgpu_buffer_rsrc_t rL = mt_rsrc(L + d.loff, (uint32_t)((int64_t)d.nc * d.ld * 8));
    double acc0 = 0.0, acc1 = 0.0;
    for (int k = tu.k0; k < tu.k1; k += 64) {
        // the right-hand-side values of the next 64 columns, one per lane,
        // handed to the lanes that use them by shuffles
        const int kl = k + lane;
        const double bl = kl < tu.k1 ? st_b<MODE>(d, kl, in0, nullptr, nullptr, xo.perm) : 0.0;
#pragma unroll
        for (int g = 0; g < 4; g++) {  // U column pairs per group
            if (k + 2 * U * g >= tu.k1) break;
            double x0[U], x1[U];
#pragma unroll
            for (int q = 0; q < U; q++)
                mt_bld2(rL, (uint32_t)((k + 2 * U * g + 2 * q + cp) * d.ld + roff + o) * 8u, 0u, x0[q], x1[q]);
#pragma unroll
            for (int q = 0; q < U; q++) {
                const bool kv = k + 2 * U * g + 2 * q + cp < tu.k1;
                const double b = __shfl(bl, 2 * U * g + 2 * q + cp, 64);
                acc0 += (kv ? x0[q] : 0.0) * b;
                acc1 += (kv ? x1[q] : 0.0) * b;
            }
        }
    }
    acc0 += __shfl_xor(acc0, 32, 64);
    acc1 += __shfl_xor(acc1, 32, 64);
    if (cp == 0) {
        st_out<MODE>(d, tu, o, acc0, nout, in0, nullptr, Dn, out, part, xo);
        st_out<MODE>(d, tu, o + 1, acc1, nout, in0, nullptr, Dn, out, part, xo);
    }
}

// The backward products (T: L21^T x(struct), partitioned: [strict(L11^-1); W]^T
// [v; x(struct)]) of one right-hand side as N-form tiles over the row-major
// copy of the front's block (LT: the m rows of L21, or all f rows of a
// partitioned-inverse front, row stride ldt): a tile of 64 output columns,
// a lane reading two neighbouring columns of one row per load, lanes 0-31
// and 32-63 taking alternate rows (folded at the end).  The gathered
// right-hand-side values are fetched one per lane for 64 rows at a time and
// handed to the lanes that use them by shuffles.  Same units, partial sums
// and epilogue as the T-form tiles (k_stile_t).
template <int MODE>
__global__ void __launch_bounds__(256) k_stile_nt(const double *LT, const FrontDesc *fd, const TileUnit *units,
                                                  int nunits, const double *in0, const double *in1, const double *Dn,
                                                  const int *sidx, double *out, double *part, MtEdgeOut xo) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int u = xcd_block(blockIdx.x, gridDim.x) * 4 + w;
    if (u >= nunits) return;
    const TileUnit tu = units[u];
    const FrontDesc d = fd[tu.front];
    int nout, roff;
    mt_geom<MODE>(d, nout, roff);
    const int lane = threadIdx.x & 63, cp = lane >> 5;
    const int o = tu.t0 + 2 * (lane & 31);  // columns o, o + 1 of the block
    const int rows = MODE == MT_PIN_BWD ? d.f : d.m;
    const __amdgpu_buffer_rsrc_t rL = mt_rsrc(LT + d.toff, (uint32_t)((int64_t)rows * d.ldt * 8));
    double acc0 = 0.0, acc1 = 0.0;
    for (int k = tu.k0; k < tu.k1; k += 64) {
        const int kl = k + lane;  // this lane's row of the 64: its right-hand-side value (zero past k1)
        const double bl = kl < tu.k1 ? st_b<MODE>(d, kl, in0, in1, sidx, nullptr) : 0.0;
#pragma unroll
        for (int g = 0; g < 4; g++) {  // 16 rows per group, a row pair per step
            double x0[8], x1[8];
#pragma unroll
            for (int q = 0; q < 8; q++)
                mt_bld2(rL, (uint32_t)((k + 16 * g + 2 * q + cp) * d.ldt + o) * 8u, 0u, x0[q], x1[q]);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const double b = __shfl(bl, 16 * g + 2 * q + cp, 64);  // rows past k1 (or past the block) have b = 0
                acc0 += x0[q] * b;
                acc1 += x1[q] * b;
            }
        }
    }
    acc0 += __shfl_xor(acc0, 32, 64);
    acc1 += __shfl_xor(acc1, 32, 64);
    if (cp == 0) {
        st_out<MODE>(d, tu, o, acc0, nout, in0, in1, Dn, out, part, xo);
        st_out<MODE>(d, tu, o + 1, acc1, nout, in0, in1, Dn, out, part, xo);
    }
}

template <int MODE, int NC>
__global__ void __launch_bounds__(256) k_stile_t(const double *L, const FrontDesc *fd, const TileUnit *units,
                                                 int nunits, const double *in0, const double *in1, const double *Dn,
                                                 const int *sidx, double *out, double *part, MtEdgeOut xo) {
    // NC output columns per pass (16 or 32)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int u = xcd_block(blockIdx.x, gridDim.x) * 4 + w;
    if (u >= nunits) return;
    const TileUnit tu = units[u];
    const FrontDesc d = fd[tu.front];
    int nout, roff;
    mt_geom<MODE>(d, nout, roff);
    const int lane = threadIdx.x & 63;
    const __amdgpu_buffer_rsrc_t rL = mt_rsrc(L + d.loff, (uint32_t)((int64_t)d.nc * d.ld * 8));
    // row blocks of 128 start where roff + row is even (16-byte loads)
    const int kb0 = tu.k0 - ((roff + tu.k0) & 1);
    for (int j0 = tu.t0; j0 < tu.t0 + 64 && j0 < nout; j0 += NC) {
        double acc[NC];
#pragma unroll
        for (int j = 0; j < NC; j++) acc[j] = 0.0;
        for (int kb = kb0; kb < tu.k1; kb += 128) {
            const int r0 = kb + 2 * lane;
            const bool v0 = r0 >= tu.k0 && r0 < tu.k1, v1 = r0 + 1 >= tu.k0 && r0 + 1 < tu.k1;
            const double b0 = v0 ? st_b<MODE>(d, r0, in0, in1, sidx, nullptr) : 0.0;
            const double b1 = v1 ? st_b<MODE>(d, r0 + 1, in0, in1, sidx, nullptr) : 0.0;
            double x0[NC], x1[NC];
#pragma unroll
            for (int j = 0; j < NC; j++)
                mt_bld2(rL, (uint32_t)((int64_t)(j0 + j) * d.ld + roff + r0) * 8u, 0u, x0[j], x1[j]);
#pragma unroll
            for (int j = 0; j < NC; j++) acc[j] += (v0 ? x0[j] : 0.0) * b0 + (v1 ? x1[j] : 0.0) * b1;
        }
        // sum over the 64 lanes: fold lanes 32 (and for NC = 16 also 16)
        // apart, then halve the set of columns a lane keeps at each of the
        // distances NC/2, ..., 2, 1; lane l ends with column j0 + (l % NC)
#pragma unroll
        for (int j = 0; j < NC; j++) {
            acc[j] += __shfl_xor(acc[j], 32, 64);
            if (NC == 16) acc[j] += __shfl_xor(acc[j], 16, 64);
        }
#pragma unroll
        for (int dist = NC / 2; dist >= 1; dist >>= 1) {
            const bool up = (lane & dist) != 0;
#pragma unroll
            for (int j = 0; j < dist; j++) {
                const double keep = up ? acc[j + dist] : acc[j];
                const double give = up ? acc[j] : acc[j + dist];
                acc[j] = keep + __shfl_xor(give, dist, 64);
            }
        }
        if (lane < NC) st_out<MODE>(d, tu, j0 + lane, acc[0], nout, in0, in1, Dn, out, part, xo);
    }
}

namespace tf {
// mode: 0 own forward, 1 struct forward, 2 backward T, 3 backward X,
// 4 own forward of a childless front straight from b (through perm)
// gbytes: size of the gathered vector (backward T form), < 4 GiB
// perm, x1, x2: backward X also writes its rows to x1 (and x2) in edge order
// LT: row-major L21 copies for the backward T product (nullptr: read L21 itself)
void launch_mtile(int mode, int ks, int nb, int na, const double *L, const FrontDesc *fd, const TileUnit *units, int nunits,
                  int K, const double *in0, const double *in1, const double *Dn, const int *sidx, double *out,
                  double *part, uint32_t gbytes, const int *perm, double *x1, double *x2, const double *LT,
                  hipStream_t st, double *z) {
    if (nunits <= 0) return;
    const dim3 g((nunits + 3) / 4), b(256);
    const MtEdgeOut xo{perm, x1, x2, z};
#define TF_MT(M, S, N)                                                                                     \
    if (na == 1)                                                                                           \
        hipLaunchKernelGGL((k_mtile<M, S, N, 1>), g, b, 0, st, L, fd, units, nunits, K, in0, in1, Dn, sidx, \
                           out, part, gbytes, xo, LT);                                                     \
    else                                                                                                   \
        hipLaunchKernelGGL((k_mtile<M, S, N, 4>), g, b, 0, st, L, fd, units, nunits, K, in0, in1, Dn, sidx, \
                           out, part, gbytes, xo, LT)
#define TF_MT_MODES(S, N)                                        \
    switch (mode) {                                              \
    case MT_OWN_FWD: TF_MT(MT_OWN_FWD, S, N); break;             \
    case MT_OWN_FWD_B: TF_MT(MT_OWN_FWD_B, S, N); break;         \
    case MT_STR_FWD: TF_MT(MT_STR_FWD, S, N); break;             \
    case MT_BWD_T: TF_MT(MT_BWD_T, S, N); break;                 \
    case MT_PIN_FWD: TF_MT(MT_PIN_FWD, S, N); break;             \
    case MT_PIN_FWD_B: TF_MT(MT_PIN_FWD_B, S, N); break;         \
    case MT_PIN_BWD: TF_MT(MT_PIN_BWD, S, N); break;             \
    default: TF_MT(MT_BWD_X, S, N); break;                       \
    }
    if (nb == 4) {
        if (ks == 4) { TF_MT_MODES(4, 4) } else { TF_MT_MODES(2, 4) }
    } else {
        if (ks == 4) { TF_MT_MODES(4, 2) } else { TF_MT_MODES(2, 2) }
    }
#undef TF_MT_MODES
#undef TF_MT
}
template <int PS, bool P2>
void launch_red_ps(int mode, int nb, const FrontDesc *fd, const TileUnit *units, int nunits, int K, const double *in0,
                   const double *in1, const double *Dn, double *out, const double *part, const MtEdgeOut &xo,
                   hipStream_t st) {
    const dim3 g(nunits * (64 * PS / (256 * (P2 ? 2 : 1)))), b(256);
#define TF_RED(M) hipLaunchKernelGGL((k_mtile_red<M, PS, P2>), g, b, 0, st, fd, units, K, in0, in1, Dn, out, part, xo, nb)
    switch (mode) {
    case MT_OWN_FWD: TF_RED(MT_OWN_FWD); break;
    case MT_OWN_FWD_B: TF_RED(MT_OWN_FWD_B); break;
    case MT_STR_FWD: TF_RED(MT_STR_FWD); break;
    case MT_BWD_T: TF_RED(MT_BWD_T); break;
    case MT_PIN_FWD: TF_RED(MT_PIN_FWD); break;
    case MT_PIN_FWD_B: TF_RED(MT_PIN_FWD_B); break;
    case MT_PIN_BWD: TF_RED(MT_PIN_BWD); break;
    default: TF_RED(MT_BWD_X); break;
    }
#undef TF_RED
}
// backward products of one right-hand side over the row-major copy LT (modes 2 and 7)
void launch_stile_lt(int mode, const double *LT, const FrontDesc *fd, const TileUnit *units, int nunits,
                     const double *in0, const double *in1, const double *Dn, const int *sidx, double *out,
                     double *part, const int *perm, double *x1, double *x2, hipStream_t st) {
    if (nunits <= 0) return;
    const dim3 g((nunits + 3) / 4), b(256);
    const MtEdgeOut xo{perm, x1, x2, nullptr};
    if (mode == MT_PIN_BWD)
        hipLaunchKernelGGL(k_stile_nt<MT_PIN_BWD>, g, b, 0, st, LT, fd, units, nunits, in0, in1, Dn, sidx, out, part, xo);
    else
        hipLaunchKernelGGL(k_stile_nt<MT_BWD_T>, g, b, 0, st, LT, fd, units, nunits, in0, in1, Dn, sidx, out, part, xo);
}
void launch_stile(int mode, int nc, const double *L, const FrontDesc *fd, const TileUnit *units, int nunits,
                  const double *in0, const double *in1, const double *Dn, const int *sidx, double *out, double *part,
                  const int *perm, double *x1, double *x2, hipStream_t st, double *z) {
    if (nunits <= 0) return;
    const dim3 g((nunits + 3) / 4), b(256);
    const MtEdgeOut xo{perm, x1, x2, z};
    switch (mode) {
    case MT_OWN_FWD: hipLaunchKernelGGL(k_stile_n<MT_OWN_FWD>, g, b, 0, st, L, fd, units, nunits, in0, Dn, out, part, xo); break;
    case MT_OWN_FWD_B: hipLaunchKernelGGL(k_stile_n<MT_OWN_FWD_B>, g, b, 0, st, L, fd, units, nunits, in0, Dn, out, part, xo); break;
    case MT_STR_FWD: hipLaunchKernelGGL(k_stile_n<MT_STR_FWD>, g, b, 0, st, L, fd, units, nunits, in0, Dn, out, part, xo); break;
    case MT_PIN_FWD: hipLaunchKernelGGL(k_stile_n<MT_PIN_FWD>, g, b, 0, st, L, fd, units, nunits, in0, Dn, out, part, xo); break;
    case MT_PIN_FWD_B: hipLaunchKernelGGL(k_stile_n<MT_PIN_FWD_B>, g, b, 0, st, L, fd, units, nunits, in0, Dn, out, part, xo); break;
    case MT_PIN_BWD:
        if (nc == 32) hipLaunchKernelGGL((k_stile_t<MT_PIN_BWD, 32>), g, b, 0, st, L, fd, units, nunits, in0, in1, Dn, sidx, out, part, xo);
        else hipLaunchKernelGGL((k_stile_t<MT_PIN_BWD, 16>), g, b, 0, st, L, fd, units, nunits, in0, in1, Dn, sidx, out, part, xo);
        break;
    case MT_BWD_T:
        if (nc == 32) hipLaunchKernelGGL((k_stile_t<MT_BWD_T, 32>), g, b, 0, st, L, fd, units, nunits, in0, in1, Dn, sidx, out, part, xo);
        else hipLaunchKernelGGL((k_stile_t<MT_BWD_T, 16>), g, b, 0, st, L, fd, units, nunits, in0, in1, Dn, sidx, out, part, xo);
        break;
    default:
        if (nc == 32) hipLaunchKernelGGL((k_stile_t<MT_BWD_X, 32>), g, b, 0, st, L, fd, units, nunits, in0, in1, Dn, sidx, out, part, xo);
        else hipLaunchKernelGGL((k_stile_t<MT_BWD_X, 16>), g, b, 0, st, L, fd, units, nunits, in0, in1, Dn, sidx, out, part, xo);
        break;
    }
}
// one right-hand side: a thread per output of each reduced tile (64 per
// tile, four tiles per workgroup), the partials summed in chunk order
template <int MODE>
__global__ void __launch_bounds__(256) k_stile_red(const FrontDesc *fd, const TileUnit *units, int nunits,
                                                   const double *in0, const double *in1, const double *Dn, double *out,
                                                   const double *part, MtEdgeOut xo) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= nunits) return;
    const TileUnit tu = units[t];
    const FrontDesc d = fd[tu.front];
    int nout, roff;
    mt_geom<MODE>(d, nout, roff);
    const int e = threadIdx.x & 63, o = tu.t0 + e;
    if (o < 0 || o >= nout) return;
    const double *p0 = part + (int64_t)tu.k0 * 1024 + e * 16;
    double s = 0.0;
    for (int c = 0; c < tu.k1; c++) s += p0[(int64_t)c * 1024];
    mt_store<MODE>(d, o, 0, s, 1, in0, in1, Dn, out, xo);
}
void launch_stile_red(int mode, const FrontDesc *fd, const TileUnit *units, int nunits, const double *in0,
                      const double *in1, const double *Dn, double *out, const double *part, const int *perm,
                      double *x1, double *x2, hipStream_t st, double *z) {
    if (nunits <= 0) return;
    const dim3 g((nunits + 3) / 4), b(256);
    const MtEdgeOut xo{perm, x1, x2, z};
    switch (mode) {
    case MT_OWN_FWD: hipLaunchKernelGGL(k_stile_red<MT_OWN_FWD>, g, b, 0, st, fd, units, nunits, in0, in1, Dn, out, part, xo); break;
    case MT_OWN_FWD_B: hipLaunchKernelGGL(k_stile_red<MT_OWN_FWD_B>, g, b, 0, st, fd, units, nunits, in0, in1, Dn, out, part, xo); break;
    case MT_STR_FWD: hipLaunchKernelGGL(k_stile_red<MT_STR_FWD>, g, b, 0, st, fd, units, nunits, in0, in1, Dn, out, part, xo); break;
    case MT_BWD_T: hipLaunchKernelGGL(k_stile_red<MT_BWD_T>, g, b, 0, st, fd, units, nunits, in0, in1, Dn, out, part, xo); break;
    case MT_PIN_FWD: hipLaunchKernelGGL(k_stile_red<MT_PIN_FWD>, g, b, 0, st, fd, units, nunits, in0, in1, Dn, out, part, xo); break;
    case MT_PIN_FWD_B: hipLaunchKernelGGL(k_stile_red<MT_PIN_FWD_B>, g, b, 0, st, fd, units, nunits, in0, in1, Dn, out, part, xo); break;
    case MT_PIN_BWD: hipLaunchKernelGGL(k_stile_red<MT_PIN_BWD>, g, b, 0, st, fd, units, nunits, in0, in1, Dn, out, part, xo); break;
    default: hipLaunchKernelGGL(k_stile_red<MT_BWD_X>, g, b, 0, st, fd, units, nunits, in0, in1, Dn, out, part, xo); break;
    }
}
void launch_mtile_red(int mode, int nb, int na, const FrontDesc *fd, const TileUnit *units, int nunits, int K,
                      const double *in0, const double *in1, const double *Dn, double *out, const double *part,
                      const int *perm, double *x1, double *x2, hipStream_t st, double *z) {
    if (nunits <= 0) return;
    const MtEdgeOut xo{perm, x1, x2, z};
    if (na == 1)
        launch_red_ps<16, false>(mode, nb, fd, units, nunits, K, in0, in1, Dn, out, part, xo, st);
    else if (K % 2 == 0)
        launch_red_ps<64, true>(mode, nb, fd, units, nunits, K, in0, in1, Dn, out, part, xo, st);
    else
        launch_red_ps<64, false>(mode, nb, fd, units, nunits, K, in0, in1, Dn, out, part, xo, st);
}
}  // namespace tf
