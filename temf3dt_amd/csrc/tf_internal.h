// tf_internal.h -- host-side structures of libtemfetd (not part of the ABI).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

// local edge -> local vertex pair (reference src/FETD_FWD.f90:23, 0-based here)
static const int TF_IND_TET[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};

// Mesh topology (reference model_mesh, src/FETD_FWD.f90:170-244) and the
// upper CSR pattern shared by the system matrices (mesh_process :642-709).
struct Topology {
    int nN = 0, nT = 0, nE = 0;
    int64_t nnz = 0;
    std::vector<double> xyz;   // [N][3]
    std::vector<int> tet_vex;  // [T][4], 1-based vertex ids
    std::vector<int> edg_vex;  // [E][2], 1-based, v1 < v2
    std::vector<double> edg_len;
    std::vector<int> tet_edg;  // [T][6], 0-based edge ids
    std::vector<int> tet_drc;  // [T][6], +-1
    std::vector<int> ia;       // [E+1], 0-based
    std::vector<int> ja;       // [nnz], 0-based, sorted per row
    // For every CSR entry q, its element contributions in the reference's COO
    // order (tet-major): ctr[cptr[q] .. cptr[q+1]) = tet*21 + local pair id.
    std::vector<int64_t> cptr;
    std::vector<int64_t> ctr;
    // local pair id p (0..20, j <= k) -> oriented (a,b) per tet: the pair is
    // taken as (a,b) with tet_edg[a] <= tet_edg[b]; stored as a | b<<4.
    std::vector<uint8_t> pair_orient;  // [T][21]
};

// Supernodal multifrontal structure from nested dissection (symbolic.cpp).
struct Symbolic {
    int n = 0;
    int leaf_size = 0;
    std::vector<int> perm;   // new -> old
    std::vector<int> iperm;  // old -> new
    int nsn = 0;
    std::vector<int> c0, nc;        // first column (new index), #columns
    std::vector<int> parent;        // supernode parent, -1 root
    std::vector<int> level;         // 0 = leaves; parent level > child level
    std::vector<int64_t> sptr;      // struct rows beyond own columns
    std::vector<int> sidx;          // new indices, sorted
    std::vector<int64_t> loff;      // L block offset: f x nc column-major, ld = f
    std::vector<int64_t> uoff;      // U block offset: m x m lower triangle, packed by columns
    std::vector<int64_t> rptr;      // relative indices of struct in parent front
    std::vector<int> rel;
    std::vector<int> children_ptr, children;  // CSR of children per supernode
    int nlevels = 0;
    std::vector<int> lev_ptr, lev_sn;        // supernodes grouped by level
    int64_t ltotal = 0, utotal = 0;
    std::vector<int64_t> amap;  // original upper-CSR entry -> L storage offset
    int64_t nnzL = 0;           // incl. diagonal, excl. upper half of diag blocks
    double flops = 0.0;         // LDL^T flops of the dense fronts
    int max_f = 0, max_nc = 0;
};

namespace tf {
// column stride of a front's L block: rows rounded up to 16 doubles, so that
// with 16-aligned offsets every column starts on a 128-byte line
inline int64_t front_ld(int64_t f) { return (f + 15) / 16 * 16; }
// The factor and solve kernels read a front's L block (nc columns of stride
// ld) and the row-major copy of its L21 (m rows of stride nc rounded up to
// 16) through buffer resources with 32-bit byte sizes and offsets; a tile
// may address up to 64 columns past the block (those reads return zero), so
// (nc + 64) columns and the copy must stay below 4 GiB.
inline bool front_fits_32(int64_t nc, int64_t m) {
    const int64_t lim = int64_t(1) << 32;
    return (nc + 64) * front_ld(nc + m) * 8 < lim && m * ((nc + 15) / 16 * 16) * 8 < lim;
}
// topology.cpp
void build_topology(Topology &t, int nN, const double *xyz, int nT, const int *tet_vex);
int find_edge(const Topology &t, int v1, int v2);  // 0-based edge id or -1
// first tet in [t0, t1) (all tets by default) holding p; 1-based or 0
int point_in_tet(const Topology &t, const double *p, double rec_charac, int t0 = 0, int t1 = -1);
double det_lu(const double *a_colmajor, int n);
// symbolic.cpp
void analyze(Symbolic &s, const Topology &t, int leaf_size);
}  // namespace tf
