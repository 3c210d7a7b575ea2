// tf_device.h -- device-side layouts shared by the HIP kernels of libtemfetd.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define TF_PB 64       // panel width (columns per elimination panel)
#define TF_TB 64       // tile edge of trailing-update tiles
#define TF_NT 256      // threads per workgroup
#define TF_TRSM_ROWS 256

// One multifrontal front (= supernode).  Its columns are the contiguous new
// indices [c0, c0+nc); its rows are those columns followed by the m struct
// rows.  L block: f x nc column-major (ld f) at L + loff; update block
// (Schur complement passed to the parent): m x m column-major at U + uoff.
struct FrontDesc {
    int64_t loff;
    int64_t uoff;
    int64_t soff;  // offset of struct rows / relative indices
    int64_t voff;  // offset of the front's slice in solve work vectors
    int f, nc, m, c0;
};

// Column c of a front addressed uniformly: columns < nc live in the L block,
// the rest in the update block.  Element (r, c), r >= c, is colptr(c)[r].
__device__ __forceinline__ double *front_col(double *L, double *U, const FrontDesc &d, int c) {
    return (c < d.nc) ? (L + d.loff + (int64_t)c * d.f)
                      : (U + d.uoff + (int64_t)(c - d.nc) * d.m - d.nc);
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

__device__ __forceinline__ void tri_index(int64_t t, int &ti, int &tj) {
    int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((int64_t)i * (i + 1) / 2 > t) i--;
    while ((int64_t)(i + 1) * (i + 2) / 2 <= t) i++;
    ti = i;
    tj = (int)(t - (int64_t)i * (i + 1) / 2);
}
