// tf_device.h -- device-side layouts shared by the HIP kernels of libtemfetd.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define TF_PB 64       // panel width (columns per elimination panel)
#define TF_TB 64       // tile edge of trailing-update tiles
#define TF_NT 256      // threads per workgroup
#define TF_TRSM_ROWS 128
#define TF_UT 128      // tile edge of the matrix-core trailing update
#define TF_SP 512      // super-panel: panel columns whose trailing update is applied at once

// Values allocated past the end of the factor storage and of the block-solve
// vectors: the tile kernels' unmasked stage loads may read up to 63 values
// beyond the last row of an array (results of those lanes are discarded).
#define TF_READ_SLACK 4096
// Doubles of the file-scope scratch array of the allocation-free solve
// (64 MiB per GPU; TF_NOALLOC_SCRATCH_BYTES in include/temfetd.h)
#define TF_NOALLOC_DOUBLES (8 * 1024 * 1024)
#define TF_RPW 4  // rows per wave of the row-per-wave kernels (SpMV, pull): 4 * TF_RPW rows per workgroup

// L11 inversion by doubling: leading dimension of the scratch block T (|J| x b)
// of a merge at block size b in a front of nc columns (|J| <= min(b, nc - b)).
__host__ __device__ inline int inv_ldt(int nc, int b) { return ((nc - b < b ? nc - b : b) + 15) & ~15; }

// One multifrontal front (= supernode).  Its columns are the contiguous new
// indices [c0, c0+nc); its rows are those columns followed by the m struct
// rows.  L block: f x nc column-major (ld f) at L + loff; update block
// (Schur complement passed to the parent): its lower triangle at U + uoff,
// column j (struct-relative) holding rows j .. m-1 from offset ucol(m, j).
struct FrontDesc {
    int64_t loff;  // multiple of 16 doubles (128-byte aligned columns)
    int64_t uoff;
    int64_t soff;  // offset of struct rows / relative indices
    int64_t voff;  // offset of the front's slice in solve work vectors
    int64_t toff;  // offset of the row-major copy of L21 (m x nc, row stride ldt) in the transposed store
    int f, nc, m, c0;
    int ld;        // column stride of the L block: f rounded up to 16
    int doff;      // offset of the front's pivots in the D array (c0 for one model)
    int ldt;       // row stride of the L21 copy: nc rounded up to 16
    // 1: the front keeps the partitioned inverse -- its L block holds
    // [strict(L11^-1); -L21 L11^-1] and its row-major copy (f x nc, row
    // stride ldt) all f rows of that block; 0: [strict(L11^-1); L21], copy of
    // the m L21 rows
    int pinv;
};

// Pivot report of one factorisation.  The factor kernels set `bad` for a
// zero or non-finite pivot (bad must stay the first member: they take it as
// an int *); the pivot scan after the factorisation counts the negative
// pivots and records the smallest and largest pivot magnitudes as bit
// patterns (non-negative doubles order like their bits).
struct FactorFlags {
    int bad;
    int pad;
    unsigned long long nneg;
    unsigned long long min_abs;
    unsigned long long max_abs;
};

// Work unit of the block-solve tile kernels: a 64-wide output tile starting
// at t0 of front `front`, accumulated over the inner range [k0, k1); slot >= 0
// names its partial tile when the tile is split over several units.  A
// reduction unit reuses the fields as (front, t0, first slot, #slots).
struct TileUnit {
    int front, t0, k0, k1, slot, pad;
};

// Offset of column j of an m x m packed lower triangle (rows j .. m-1).
__host__ __device__ __forceinline__ int64_t ucol(int64_t m, int64_t j) { return j * m - j * (j - 1) / 2; }

// Column c of a front addressed uniformly: columns < nc live in the L block,
// the rest in the update block.  Element (r, c), r >= c, is colptr(c)[r];
// the update block stores no element above its diagonal.
__device__ __forceinline__ double *front_col(double *L, double *U, const FrontDesc &d, int c) {
    if (c < d.nc) return L + d.loff + (int64_t)c * d.ld;
    const int64_t j = c - d.nc;
    return U + d.uoff + ucol(d.m, j) - j - d.nc;
}

// Work segment of a grid launch: fronts[i] is a supernode id and
// pref[i]..pref[i+1] its range of workgroups.
__device__ __forceinline__ int seg_find(const int64_t *pref, int cnt, int64_t w) {
    int lo = 0, hi = cnt - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): workgroup i runs on XCD i % 8.  This maps the hardware id to a
// logical block so that every XCD processes one contiguous run of logical
// blocks, keeping neighbouring work (which shares operands) in one L2.
__device__ __forceinline__ int xcd_block(int i, int n) {
    const int q = n >> 3, r = n & 7, x = i & 7;
    return x * q + min(x, r) + (i >> 3);
}

__device__ __forceinline__ void tri_index(int64_t t, int &ti, int &tj) {
    int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((int64_t)i * (i + 1) / 2 > t) i--;
    while ((int64_t)(i + 1) * (i + 2) / 2 <= t) i++;
    ti = i;
    tj = (int)(t - (int64_t)i * (i + 1) / 2);
}

typedef double tf_d4 __attribute__((ext_vector_type(4)));

// Raw buffer loads: resource (base, byte size) in scalar registers, a
// per-lane 32-bit offset and a wave-uniform offset added by the load unit.
// 0x00020000: the gfx9 buffer-resource word 3 (raw 32-bit data).
typedef unsigned int mt_u2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mt_rsrc(const double *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double mt_bld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
typedef unsigned int mt_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void mt_bld2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double &x0,
                                        double &x1) {
    const mt_u4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
    x0 = __builtin_bit_cast(double, (mt_u2){v.x, v.y});
    x1 = __builtin_bit_cast(double, (mt_u2){v.z, v.w});
}

