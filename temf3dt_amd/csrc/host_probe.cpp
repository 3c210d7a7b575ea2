// host_probe.cpp -- host-only entry points over the topology and symbolic
// analysis (no GPU needed).  Built into libtemfetd_host.so for CPU-side checks
// of the edge numbering, CSR pattern and supernodal structure, and for
// predicting factor sizes before a run.
#include <algorithm>
#include <cstring>
#include <exception>
#include <string>

#include "tf_internal.h"

namespace {
struct Probe {
    Topology t;
    Symbolic s;
    std::string err;
};
Probe g;
}  // namespace

extern "C" {

const char *tfh_last_error(void) { return g.err.c_str(); }

// 1 if a front of nc columns and m struct rows fits the kernels' 32-bit
// buffer addressing (the check analyze applies to every front)
int tfh_front_fits(long long nc, long long m) { return tf::front_fits_32(nc, m) ? 1 : 0; }

int tfh_set_mesh(int nN, const double *xyz, int nT, const int *tet_vex, int *nE, long long *nnz) {
    try {
        tf::build_topology(g.t, nN, xyz, nT, tet_vex);
        if (nE) *nE = g.t.nE;
        if (nnz) *nnz = g.t.nnz;
        return 0;
    } catch (const std::exception &e) {
        g.err = e.what();
        return -1;
    }
}

void tfh_get_topology(int *edg_vex, int *tet_edg, int *tet_drc, int *ia, int *ja) {
    const Topology &t = g.t;
    if (edg_vex) std::copy(t.edg_vex.begin(), t.edg_vex.end(), edg_vex);
    if (tet_edg) for (size_t i = 0; i < t.tet_edg.size(); i++) tet_edg[i] = t.tet_edg[i] + 1;
    if (tet_drc) std::copy(t.tet_drc.begin(), t.tet_drc.end(), tet_drc);
    if (ia) for (size_t i = 0; i < t.ia.size(); i++) ia[i] = t.ia[i] + 1;
    if (ja) for (size_t i = 0; i < t.ja.size(); i++) ja[i] = t.ja[i] + 1;
}

int tfh_analyze(int leaf, int *nsn, int *nlevels, long long *nnzL, double *flops, long long *ltotal,
                long long *utotal, long long *nsidx) {
    try {
        tf::analyze(g.s, g.t, leaf);
        const Symbolic &s = g.s;
        if (nsn) *nsn = s.nsn;
        if (nlevels) *nlevels = s.nlevels;
        if (nnzL) *nnzL = s.nnzL;
        if (flops) *flops = s.flops;
        if (ltotal) *ltotal = s.ltotal;
        if (utotal) *utotal = s.utotal;
        if (nsidx) *nsidx = (long long)s.sidx.size();
        return 0;
    } catch (const std::exception &e) {
        g.err = e.what();
        return -1;
    }
}

// perm[n] (new->old, 0-based); per supernode c0, nc, parent, level; struct
// pointers [nsn+1] and indices; relative indices; amap[nnz]
void tfh_get_symbolic(int *perm, int *c0, int *nc, int *parent, int *level, long long *sptr, int *sidx,
                      int *rel, long long *amap, long long *loff) {
    const Symbolic &s = g.s;
    if (perm) std::copy(s.perm.begin(), s.perm.end(), perm);
    if (c0) std::copy(s.c0.begin(), s.c0.end(), c0);
    if (nc) std::copy(s.nc.begin(), s.nc.end(), nc);
    if (parent) std::copy(s.parent.begin(), s.parent.end(), parent);
    if (level) std::copy(s.level.begin(), s.level.end(), level);
    if (sptr) for (size_t i = 0; i < s.sptr.size(); i++) sptr[i] = s.sptr[i];
    if (sidx) std::copy(s.sidx.begin(), s.sidx.end(), sidx);
    if (rel) std::copy(s.rel.begin(), s.rel.end(), rel);
    if (amap) for (size_t i = 0; i < s.amap.size(); i++) amap[i] = s.amap[i];
    if (loff) for (size_t i = 0; i < s.loff.size(); i++) loff[i] = s.loff[i];
}
}
