// topology.cpp -- mesh topology and the upper CSR pattern (host, once per mesh).
//
// Restates reference model_mesh (src/FETD_FWD.f90:170-244): global edges are
// the sorted unique (min,max) vertex pairs of all tets' six local edges, the
// edge id is the rank in lexicographic order (unique_ii / bSearch_Ii,
// src/FETD_UTILITY.f90:409-429,601-676), and tet_drc = +1 when the local
// edge's first vertex has the smaller id.  The CSR pattern and the order in
// which element contributions are summed follow mesh_process
// (src/FETD_FWD.f90:642-709): entries (tet_edg(j), tet_edg(k)) with
// tet_edg(j) <= tet_edg(k), tet-major.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "tf_internal.h"

namespace tf {

// Stable LSD radix sort of (key, val) by key, 16-bit digits; passes whose
// digit is constant are skipped.
static void radix_sort_kv(std::vector<uint64_t> &key, std::vector<int64_t> &val) {
    size_t n = key.size();
    std::vector<uint64_t> k2(n);
    std::vector<int64_t> v2(n);
    uint64_t allor = 0, alland = ~0ull;
    for (size_t i = 0; i < n; i++) { allor |= key[i]; alland &= key[i]; }
    std::vector<size_t> cnt(65537);
    for (int shift = 0; shift < 64; shift += 16) {
        if ((((allor ^ alland) >> shift) & 0xffffull) == 0) continue;
        std::fill(cnt.begin(), cnt.end(), 0);
        for (size_t i = 0; i < n; i++) cnt[((key[i] >> shift) & 0xffff) + 1]++;
        for (int d = 0; d < 65536; d++) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; i++) {
            size_t p = cnt[(key[i] >> shift) & 0xffff]++;
            k2[p] = key[i];
            v2[p] = val[i];
        }
        key.swap(k2);
        val.swap(v2);
    }
}

double det_lu(const double *acm, int n) {
    // LU with partial pivoting (first maximum), reciprocal scaling; a zero
    // pivot gives 0 (det_mkl_3/4, src/FETD_UTILITY.f90:65-103).
    double a[4][4];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) a[i][j] = acm[j * n + i];
    int piv[4];
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = std::fabs(a[k][k]);
        for (int i = k + 1; i < n; i++)
            if (std::fabs(a[i][k]) > mx) { mx = std::fabs(a[i][k]); p = i; }
        piv[k] = p;
        if (a[p][k] != 0.0) {
            if (p != k)
                for (int j = 0; j < n; j++) std::swap(a[k][j], a[p][j]);
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

void build_topology(Topology &t, int nN, const double *xyz, int nT, const int *tet_vex) {
    t = Topology();
    t.nN = nN;
    t.nT = nT;
    t.xyz.assign(xyz, xyz + (size_t)nN * 3);
    t.tet_vex.assign(tet_vex, tet_vex + (size_t)nT * 4);
    for (size_t i = 0; i < (size_t)nT * 4; i++)
        if (tet_vex[i] < 1 || tet_vex[i] > nN) throw std::runtime_error("tet vertex id out of range");
    size_t np = (size_t)nT * 6;
    std::vector<uint64_t> key(np);
    std::vector<int64_t> idx(np);
    for (int i = 0; i < nT; i++)
        for (int j = 0; j < 6; j++) {
            uint32_t va = (uint32_t)tet_vex[4 * i + TF_IND_TET[j][0]];
            uint32_t vb = (uint32_t)tet_vex[4 * i + TF_IND_TET[j][1]];
            uint32_t v1 = std::min(va, vb), v2 = std::max(va, vb);
            key[(size_t)i * 6 + j] = ((uint64_t)v1 << 32) | v2;
            idx[(size_t)i * 6 + j] = (int64_t)i * 6 + j;
        }
    radix_sort_kv(key, idx);
    t.tet_edg.assign(np, 0);
    t.tet_drc.assign(np, 0);
    int ne = -1;
    for (size_t q = 0; q < np; q++) {
        if (q == 0 || key[q] != key[q - 1]) {
            ne++;
            t.edg_vex.push_back((int)(key[q] >> 32));
            t.edg_vex.push_back((int)(key[q] & 0xffffffffu));
        }
        t.tet_edg[idx[q]] = ne;
    }
    t.nE = ne + 1;
    t.edg_len.resize(t.nE);
    for (int e = 0; e < t.nE; e++) {
        const double *p = &xyz[3 * (size_t)(t.edg_vex[2 * e] - 1)];
        const double *q = &xyz[3 * (size_t)(t.edg_vex[2 * e + 1] - 1)];
        double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
        t.edg_len[e] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    for (int i = 0; i < nT; i++)
        for (int j = 0; j < 6; j++) {
            int va = tet_vex[4 * i + TF_IND_TET[j][0]], vb = tet_vex[4 * i + TF_IND_TET[j][1]];
            t.tet_drc[6 * (size_t)i + j] = (va < vb) ? 1 : -1;
        }
    // Upper CSR pattern and the contributor lists in COO (tet-major) order.
    size_t nc = (size_t)nT * 21;
    std::vector<uint64_t> ck(nc);
    std::vector<int64_t> cv(nc);
    t.pair_orient.assign(nc, 0);
    for (int i = 0; i < nT; i++) {
        const int *te = &t.tet_edg[6 * (size_t)i];
        int p = 0;
        for (int j = 0; j < 6; j++)
            for (int k = j; k < 6; k++, p++) {
                int a = j, b = k;
                if (te[j] > te[k]) { a = k; b = j; }
                t.pair_orient[(size_t)i * 21 + p] = (uint8_t)(a | (b << 4));
                ck[(size_t)i * 21 + p] = ((uint64_t)(uint32_t)te[a] << 32) | (uint32_t)te[b];
                cv[(size_t)i * 21 + p] = (int64_t)i * 21 + p;
            }
    }
    // Within one tet the reference visits (j,k) row-major over the 6x6
    // matrix; a tet contributes at most once per (row,col), so the summation
    // order of one CSR entry is the tet order, which the stable sort keeps.
    radix_sort_kv(ck, cv);
    t.ia.assign((size_t)t.nE + 1, 0);
    t.cptr.clear();
    t.ctr.swap(cv);
    int64_t nnz = 0;
    for (size_t q = 0; q < nc; q++) {
        if (q == 0 || ck[q] != ck[q - 1]) {
            int r = (int)(ck[q] >> 32);
            t.ja.push_back((int)(ck[q] & 0xffffffffu));
            t.ia[r + 1]++;
            t.cptr.push_back((int64_t)q);
            nnz++;
        }
    }
    t.cptr.push_back((int64_t)nc);
    for (int r = 0; r < t.nE; r++) t.ia[r + 1] += t.ia[r];
    t.nnz = nnz;
}

int find_edge(const Topology &t, int v1, int v2) {
    int lo = std::min(v1, v2), hi = std::max(v1, v2);
    int L = 0, H = t.nE - 1;
    while (L <= H) {
        int mid = (L + H) / 2;
        int a = t.edg_vex[2 * mid], b = t.edg_vex[2 * mid + 1];
        if (a == lo && b == hi) return mid;
        if (a < lo || (a == lo && b < hi)) L = mid + 1;
        else H = mid - 1;
    }
    return -1;
}

// p_in_tet (src/FETD_FWD.f90:1009-1049): first tet, in index order, whose
// centroid lies within rec_charac of p and whose five determinants share a
// sign (zero counts for both signs).
int point_in_tet(const Topology &t, const double *p, double rec_charac, int t0, int t1) {
    if (t1 < 0) t1 = t.nT;
    for (int i = t0; i < t1; i++) {
        double x[4], y[4], z[4];
        for (int q = 0; q < 4; q++) {
            const double *v = &t.xyz[3 * (size_t)(t.tet_vex[4 * (size_t)i + q] - 1)];
            x[q] = v[0]; y[q] = v[1]; z[q] = v[2];
        }
        double mx = (x[0] + x[1] + x[2] + x[3]) / 4.0, my = (y[0] + y[1] + y[2] + y[3]) / 4.0,
               mz = (z[0] + z[1] + z[2] + z[3]) / 4.0;
        double dd = std::sqrt((mx - p[0]) * (mx - p[0]) + (my - p[1]) * (my - p[1]) +
                              (mz - p[2]) * (mz - p[2]));
        if (dd > rec_charac) continue;
        double det[5];
        for (int s = 0; s < 5; s++) {
            double X[4] = {x[0], x[1], x[2], x[3]}, Y[4] = {y[0], y[1], y[2], y[3]},
                   Z[4] = {z[0], z[1], z[2], z[3]};
            if (s > 0) { X[s - 1] = p[0]; Y[s - 1] = p[1]; Z[s - 1] = p[2]; }
            double cm[16] = {X[0], X[1], X[2], X[3], Y[0], Y[1], Y[2], Y[3],
                             Z[0], Z[1], Z[2], Z[3], 1, 1, 1, 1};
            det[s] = det_lu(cm, 4);
        }
        bool allp = true, alln = true;
        for (int s = 0; s < 5; s++) {
            if (!(det[s] >= 0.0)) allp = false;
            if (!(det[s] <= 0.0)) alln = false;
        }
        if (allp || alln) return i + 1;
    }
    return 0;
}

}  // namespace tf
