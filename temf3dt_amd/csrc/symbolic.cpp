// symbolic.cpp -- fill-reducing ordering and the multifrontal supernode tree.
//
// The reference hands P to PARDISO phase 12, which re-runs METIS ordering
// and symbolic factorisation at every call (src/FETD_UTILITY.f90:879-936,
// called once per doubling check, src/FETD_FWD.f90:1307,1361).  The pattern
// of P never changes, so this library analyses it once per mesh:
//
//  * geometric nested dissection on the edge graph, using edge midpoints as
//    coordinates: each set is split by a plane (axis and quantile chosen for
//    the smallest separator); the separator is the minimum vertex cover of
//    the edges that cross the plane, and is ordered last;
//  * every dissection node with columns becomes one supernode whose diagonal
//    block is dense (separators fill in completely; leaves are small);
//  * row structures, relative indices into the parent front, L/U storage
//    offsets, and the map from assembled CSR entries into L storage.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <numeric>
#include <tuple>
#include <queue>
#include <stdexcept>

#include "tf_internal.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace tf {

namespace {

struct NDNode {
    std::vector<int> cols;
    int child[2] = {-1, -1};
};

struct NDState {
    int n;
    const std::vector<int> *xadj;
    const std::vector<int> *adj;
    const std::vector<double> *co;  // [n][3]
    std::vector<int64_t> mark;     // set id * 2 + side
    std::vector<int> loc;          // scratch: local index of a boundary vertex (sets are disjoint)
    std::vector<int8_t> part;      // side 0 / 1, or 2 for the separator, of the set being split
    std::vector<int> ver;          // version of a vertex's queued gains
    std::vector<int64_t> stamp;    // refinement pass that moved the vertex
    std::vector<NDNode> nodes;
    std::atomic<int> nnodes{0};
    std::atomic<int64_t> setid{1};
    int leaf;
};

// Minimum vertex cover of a bipartite graph (left vertices 0..nl-1 with
// adjacency lptr/ladj into right vertices 0..nr-1): maximum matching by
// Hopcroft-Karp, then Koenig's construction -- the left vertices not
// reached and the right vertices reached by alternating paths from the
// free left vertices.
static void bip_min_cover(int nl, int nr, const std::vector<int> &lptr, const std::vector<int> &ladj,
                          std::vector<char> &coverL, std::vector<char> &coverR) {
    const int INF = std::numeric_limits<int>::max();
    std::vector<int> mL(nl, -1), mR(nr, -1), dist(nl), it(nl), q(nl), stk;
    for (int u = 0; u < nl; u++)  // greedy start
        for (int e = lptr[u]; e < lptr[u + 1]; e++)
            if (mR[ladj[e]] < 0) { mL[u] = ladj[e]; mR[ladj[e]] = u; break; }
    while (true) {
        int qh = 0, qt = 0;
        bool found = false;
        for (int u = 0; u < nl; u++) {
            if (mL[u] < 0) { dist[u] = 0; q[qt++] = u; }
            else dist[u] = INF;
        }
        while (qh < qt) {
            const int u = q[qh++];
            for (int e = lptr[u]; e < lptr[u + 1]; e++) {
                const int w = mR[ladj[e]];
                if (w < 0) found = true;
                else if (dist[w] == INF) { dist[w] = dist[u] + 1; q[qt++] = w; }
            }
        }
        if (!found) break;
        // augmenting paths along the BFS layers, depth first without recursion
        for (int u = 0; u < nl; u++) it[u] = lptr[u];
        for (int s0 = 0; s0 < nl; s0++) {
            if (mL[s0] >= 0) continue;
            stk.assign(1, s0);
            while (!stk.empty()) {
                const int u = stk.back();
                bool moved = false;
                while (it[u] < lptr[u + 1]) {
                    const int w = mR[ladj[it[u]]];
                    if (w < 0) {  // free right vertex: flip the path held on the stack
                        for (int x : stk) { const int r = ladj[it[x]]; mL[x] = r; mR[r] = x; }
                        stk.clear();
                        moved = true;
                        break;
                    }
                    if (dist[w] == dist[u] + 1) { stk.push_back(w); moved = true; break; }
                    it[u]++;
                }
                if (!moved) {  // dead end
                    dist[u] = INF;
                    stk.pop_back();
                    if (!stk.empty()) it[stk.back()]++;
                }
            }
        }
    }
    // alternating reachability from the free left vertices
    std::vector<char> visL(nl, 0), visR(nr, 0);
    int qh = 0, qt = 0;
    for (int u = 0; u < nl; u++)
        if (mL[u] < 0) { visL[u] = 1; q[qt++] = u; }
    while (qh < qt) {
        const int u = q[qh++];
        for (int e = lptr[u]; e < lptr[u + 1]; e++) {
            const int r = ladj[e];
            if (visR[r] || mL[u] == r) continue;
            visR[r] = 1;
            const int w = mR[r];
            if (w >= 0 && !visL[w]) { visL[w] = 1; q[qt++] = w; }
        }
    }
    coverL.assign(nl, 0);
    coverR.assign(nr, 0);
    for (int u = 0; u < nl; u++) coverL[u] = !visL[u];
    for (int r = 0; r < nr; r++) coverR[r] = visR[r];
}

// Separator of `v` split at position `cut` of the sorted order `ord`
// (first `cut` nodes form side 0): the smallest set of boundary vertices
// that covers every edge between the two sides.
static void eval_split(NDState &st, const std::vector<int> &ord, size_t cut, int64_t sid,
                       std::vector<int> &sep, int &sepside) {
    auto &mark = st.mark;
    for (size_t i = 0; i < ord.size(); i++) mark[ord[i]] = sid * 2 + (i < cut ? 0 : 1);
    std::vector<int> b0, b1;
    const auto &xadj = *st.xadj;
    const auto &adj = *st.adj;
    for (size_t i = 0; i < ord.size(); i++) {
        int v = ord[i];
        int64_t mv = mark[v];
        for (int q = xadj[v]; q < xadj[v + 1]; q++) {
            int64_t mu = mark[adj[q]];
            if ((mu >> 1) == sid && mu != mv) {
                if ((mv & 1) == 0) b0.push_back(v); else b1.push_back(v);
                break;
            }
        }
    }
    sepside = b0.size() <= b1.size() ? 0 : 1;
    const int nl = (int)b0.size(), nr = (int)b1.size();
    for (int r = 0; r < nr; r++) st.loc[b1[r]] = r;
    std::vector<int> lptr(nl + 1, 0), ladj;
    for (int u = 0; u < nl; u++) {
        const int v = b0[u];
        for (int q = xadj[v]; q < xadj[v + 1]; q++)
            if (mark[adj[q]] == sid * 2 + 1) ladj.push_back(st.loc[adj[q]]);
        lptr[u + 1] = (int)ladj.size();
    }
    std::vector<char> cL, cR;
    bip_min_cover(nl, nr, lptr, ladj, cL, cR);
    sep.clear();
    for (int u = 0; u < nl; u++)
        if (cL[u]) sep.push_back(b0[u]);
    for (int r = 0; r < nr; r++)
        if (cR[r]) sep.push_back(b1[r]);
}

// Fiduccia-Mattheyses refinement of the vertex separator of the set `vs`
// (members have mark == code; part[] gives side 0 / 1 or 2 = separator).
// Moving a separator vertex into side X pulls its neighbours on the other
// side into the separator: gain = 1 - their number.  Each pass makes
// best-first moves, every vertex at most once, accepting losses for up to
// 64 moves past the best state, and rolls back to the smallest separator
// seen that keeps both sides within maxfrac of the set.
static void fm_refine(NDState &st, const std::vector<int> &vs, int64_t code, double maxfrac) {
    auto &part = st.part;
    auto &ver = st.ver;
    auto &stamp = st.stamp;
    const auto &xadj = *st.xadj;
    const auto &adj = *st.adj;
    const auto &mark = st.mark;
    int64_t cnt[3] = {0, 0, 0};
    for (int v : vs) cnt[part[v]]++;
    const int64_t maxside = (int64_t)(maxfrac * (double)vs.size());
    auto gain = [&](int v, int X) {
        int g = 1;
        for (int q = xadj[v]; q < xadj[v + 1]; q++) {
            const int u = adj[q];
            if (mark[u] == code && part[u] == 1 - X) g--;
        }
        return g;
    };
    using Entry = std::tuple<int, int, int, int>;  // gain, vertex, side, version
    for (int pass = 0; pass < 8; pass++) {
        const int64_t pst = st.setid.fetch_add(1);
        std::priority_queue<Entry> pq;
        auto queue = [&](int w) {
            if (part[w] != 2 || stamp[w] == pst) return;
            ver[w]++;
            pq.emplace(gain(w, 0), w, 0, ver[w]);
            pq.emplace(gain(w, 1), w, 1, ver[w]);
        };
        for (int v : vs)
            if (part[v] == 2) queue(v);
        std::vector<std::pair<int, int>> moves;
        std::vector<int> pulled;
        std::vector<size_t> pptr(1, 0);
        int64_t best = cnt[2], best_max = std::max(cnt[0], cnt[1]);
        size_t best_len = 0;
        int since = 0;
        while (!pq.empty()) {
            const auto [g, v, X, vv] = pq.top();
            pq.pop();
            if (vv != ver[v] || part[v] != 2 || stamp[v] == pst) continue;
            if (cnt[X] + 1 > maxside) continue;
            stamp[v] = pst;
            part[v] = (int8_t)X;
            cnt[X]++;
            cnt[2]--;
            moves.emplace_back(v, X);
            const size_t p0 = pulled.size();
            for (int q = xadj[v]; q < xadj[v + 1]; q++) {
                const int u = adj[q];
                if (mark[u] == code && part[u] == 1 - X) {
                    part[u] = 2;
                    cnt[1 - X]--;
                    cnt[2]++;
                    pulled.push_back(u);
                }
            }
            pptr.push_back(pulled.size());
            for (int q = xadj[v]; q < xadj[v + 1]; q++)
                if (mark[adj[q]] == code) queue(adj[q]);
            for (size_t k = p0; k < pulled.size(); k++) {
                const int u = pulled[k];
                queue(u);
                for (int q = xadj[u]; q < xadj[u + 1]; q++)
                    if (mark[adj[q]] == code) queue(adj[q]);
            }
            const int64_t mx = std::max(cnt[0], cnt[1]);
            if (cnt[2] < best || (cnt[2] == best && mx < best_max)) {
                best = cnt[2];
                best_max = mx;
                best_len = moves.size();
                since = 0;
            } else if (++since > 64) {
                break;
            }
        }
        for (size_t i = moves.size(); i-- > best_len;) {  // roll back past the best state
            const auto [v, X] = moves[i];
            for (size_t k = pptr[i]; k < pptr[i + 1]; k++) {
                part[pulled[k]] = (int8_t)(1 - X);
                cnt[1 - X]++;
                cnt[2]--;
            }
            part[v] = 2;
            cnt[X]--;
            cnt[2]++;
        }
        if (best_len == 0) break;
    }
}

static int nd_rec(NDState &st, std::vector<int> v) {
    int id = st.nnodes.fetch_add(1);
    NDNode &nd = st.nodes[id];
    if ((int)v.size() <= st.leaf) {
        nd.cols.swap(v);
        return id;
    }
    const auto &co = *st.co;
    // planes normal to each axis at count quantiles 0.4..0.6, then the best
    // axis at the quantiles in between; the smallest separator wins, with a
    // mild penalty for imbalance
    std::vector<int> best_sep;
    std::vector<int> best_ord;
    size_t best_cut = 0;
    int best_side = 0;
    double best_score = 1e300;
    static const std::vector<double> qs0 = {0.4, 0.45, 0.5, 0.55, 0.6}, qs1 = {0.425, 0.475, 0.525, 0.575};
    int best_axis = -1;
    for (int pass = 0; pass < 2; pass++) {
        for (int axis = 0; axis < 3; axis++) {
            if (pass == 1 && axis != best_axis) continue;
            std::vector<int> ord(v);
            std::sort(ord.begin(), ord.end(), [&](int a, int b) {
                double xa = co[3 * (size_t)a + axis], xb = co[3 * (size_t)b + axis];
                return xa < xb || (xa == xb && a < b);
            });
            const std::vector<double> &qs = pass == 0 ? qs0 : qs1;
            for (size_t qi = 0; qi < qs.size(); qi++) {
                size_t cut = (size_t)(qs[qi] * ord.size());
                if (cut == 0 || cut >= ord.size()) continue;
                int64_t sid = st.setid.fetch_add(1);
                std::vector<int> sep;
                int side;
                eval_split(st, ord, cut, sid, sep, side);
                double bal = std::fabs(qs[qi] - 0.5);
                double score = (double)sep.size() * (1.0 + 0.5 * bal);
                if (score < best_score) {
                    best_score = score;
                    best_sep.swap(sep);
                    best_ord = ord;
                    best_cut = cut;
                    best_side = side;
                    best_axis = axis;
                }
            }
        }
    }
    // refine the separator, then form the two remaining sides
    const int64_t code = st.setid.fetch_add(1) * 2;
    for (size_t i = 0; i < best_ord.size(); i++) {
        const int u = best_ord[i];
        st.mark[u] = code;
        st.part[u] = i < best_cut ? 0 : 1;
    }
    for (int u : best_sep) st.part[u] = 2;
    fm_refine(st, best_ord, code, 0.6);
    std::vector<int> a, b;
    best_sep.clear();
    for (int u : best_ord) (st.part[u] == 2 ? best_sep : st.part[u] == 0 ? a : b).push_back(u);
    (void)best_side;
    nd.cols = best_sep;
    std::vector<int>().swap(best_ord);
    int ca = -1, cb = -1;
    bool par = v.size() > 20000;
    std::vector<int>().swap(v);
#pragma omp task shared(ca, st) if (par)
    { if (!a.empty()) ca = nd_rec(st, std::move(a)); }
#pragma omp task shared(cb, st) if (par)
    { if (!b.empty()) cb = nd_rec(st, std::move(b)); }
#pragma omp taskwait
    st.nodes[id].child[0] = ca;
    st.nodes[id].child[1] = cb;
    return id;
}

// Disconnected graph (independent models in one context): one dissection
// per connected component, joined under column-free nodes, so the fronts
// form one tree per component and no front mixes components.
static int nd_components(NDState &st, std::vector<std::vector<int>> &comps, size_t a, size_t b) {
    if (b - a == 1) return nd_rec(st, std::move(comps[a]));
    const int id = st.nnodes.fetch_add(1);
    const size_t mid = (a + b) / 2;
    int ca = -1, cb = -1;
#pragma omp task shared(ca, st, comps)
    { ca = nd_components(st, comps, a, mid); }
#pragma omp task shared(cb, st, comps)
    { cb = nd_components(st, comps, mid, b); }
#pragma omp taskwait
    st.nodes[id].child[0] = ca;
    st.nodes[id].child[1] = cb;
    return id;
}

}  // namespace

void analyze(Symbolic &s, const Topology &t, int leaf_size) {
    s = Symbolic();
    int n = t.nE;
    s.n = n;
    s.leaf_size = leaf_size;
    // full adjacency without diagonal
    std::vector<int> deg(n + 1, 0);
    for (int r = 0; r < n; r++)
        for (int q = t.ia[r]; q < t.ia[r + 1]; q++)
            if (t.ja[q] != r) { deg[r]++; deg[t.ja[q]]++; }
    std::vector<int> xadj(n + 1, 0);
    for (int r = 0; r < n; r++) xadj[r + 1] = xadj[r] + deg[r];
    std::vector<int> adj(xadj[n]);
    std::fill(deg.begin(), deg.end(), 0);
    for (int r = 0; r < n; r++)
        for (int q = t.ia[r]; q < t.ia[r + 1]; q++) {
            int c = t.ja[q];
            if (c == r) continue;
            adj[xadj[r] + deg[r]++] = c;
            adj[xadj[c] + deg[c]++] = r;
        }
    std::vector<double> co((size_t)n * 3);
    for (int e = 0; e < n; e++)
        for (int k = 0; k < 3; k++)
            co[3 * (size_t)e + k] = 0.5 * (t.xyz[3 * (size_t)(t.edg_vex[2 * e] - 1) + k] +
                                           t.xyz[3 * (size_t)(t.edg_vex[2 * e + 1] - 1) + k]);
    NDState st;
    st.n = n;
    st.xadj = &xadj;
    st.adj = &adj;
    st.co = &co;
    st.mark.assign(n, 0);
    st.loc.assign(n, 0);
    st.part.assign(n, 0);
    st.ver.assign(n, 0);
    st.stamp.assign(n, 0);
    st.leaf = std::max(leaf_size, 4);
    st.nodes.resize((size_t)(2 * (n / std::max(1, st.leaf / 2)) + 64));
    // connected components (breadth-first, in order of the first vertex)
    std::vector<int> comp(n, -1);
    std::vector<std::vector<int>> comps;
    for (int v0 = 0; v0 < n; v0++) {
        if (comp[v0] >= 0) continue;
        const int cid = (int)comps.size();
        comps.emplace_back();
        std::vector<int> &cv = comps.back();
        comp[v0] = cid;
        cv.push_back(v0);
        for (size_t h = 0; h < cv.size(); h++) {
            const int v = cv[h];
            for (int q = xadj[v]; q < xadj[v + 1]; q++)
                if (comp[adj[q]] < 0) { comp[adj[q]] = cid; cv.push_back(adj[q]); }
        }
        std::sort(cv.begin(), cv.end());
    }
    st.nodes.resize(st.nodes.size() + comps.size());
    int root = -1;
#pragma omp parallel
#pragma omp single
    root = nd_components(st, comps, 0, comps.size());
    if ((size_t)st.nnodes.load() > st.nodes.size()) throw std::runtime_error("nd node overflow");
    // postorder: assign columns; non-empty nodes become supernodes
    s.perm.reserve(n);
    std::vector<int> stack, sn_of_node(st.nnodes.load(), -1);
    std::vector<int> nd_parent_sn(st.nnodes.load(), -1);
    // iterative postorder with "nearest non-empty ancestor" tracking
    struct Fr { int node; int state; int anc; };
    std::vector<Fr> stk;
    stk.push_back({root, 0, -1});
    std::vector<int> sn_parent_node;  // supernode -> nd node of nearest non-empty ancestor
    std::vector<int> sn_node;
    std::vector<int> pending_parent;  // per supernode: nearest non-empty ancestor nd node
    while (!stk.empty()) {
        Fr &f = stk.back();
        NDNode &nd = st.nodes[f.node];
        if (f.state < 2) {
            int c = nd.child[f.state];
            int anc = nd.cols.empty() ? f.anc : f.node;
            f.state++;
            if (c >= 0) stk.push_back({c, 0, anc});
            continue;
        }
        if (!nd.cols.empty()) {
            int sid = (int)s.c0.size();
            s.c0.push_back((int)s.perm.size());
            s.nc.push_back((int)nd.cols.size());
            for (int u : nd.cols) s.perm.push_back(u);
            sn_of_node[f.node] = sid;
            sn_node.push_back(f.node);
            pending_parent.push_back(f.anc);
        }
        stk.pop_back();
    }
    if ((int)s.perm.size() != n) throw std::runtime_error("ordering lost columns");
    s.nsn = (int)s.c0.size();
    s.iperm.assign(n, 0);
    for (int i = 0; i < n; i++) s.iperm[s.perm[i]] = i;
    s.parent.assign(s.nsn, -1);
    for (int k = 0; k < s.nsn; k++)
        s.parent[k] = pending_parent[k] >= 0 ? sn_of_node[pending_parent[k]] : -1;
    // children lists
    std::vector<int> nch(s.nsn + 1, 0);
    for (int k = 0; k < s.nsn; k++)
        if (s.parent[k] >= 0) nch[s.parent[k] + 1]++;
    s.children_ptr.assign(s.nsn + 1, 0);
    for (int k = 0; k < s.nsn; k++) s.children_ptr[k + 1] = s.children_ptr[k] + nch[k + 1];
    s.children.assign(s.children_ptr[s.nsn], 0);
    std::vector<int> fillc(s.nsn, 0);
    for (int k = 0; k < s.nsn; k++)
        if (s.parent[k] >= 0) {
            int p = s.parent[k];
            s.children[s.children_ptr[p] + fillc[p]++] = k;
        }
    // row structures (children precede parents in supernode order)
    std::vector<int> mark(n, -1);
    s.sptr.assign(s.nsn + 1, 0);
    std::vector<std::vector<int>> rows(s.nsn);
    for (int k = 0; k < s.nsn; k++) {
        int c0 = s.c0[k], c1 = c0 + s.nc[k];
        auto &rk = rows[k];
        for (int j = c0; j < c1; j++) {
            int old = s.perm[j];
            for (int q = xadj[old]; q < xadj[old + 1]; q++) {
                int i = s.iperm[adj[q]];
                if (i >= c1 && mark[i] != k) { mark[i] = k; rk.push_back(i); }
            }
        }
        for (int q = s.children_ptr[k]; q < s.children_ptr[k + 1]; q++) {
            int ch = s.children[q];
            for (int i : rows[ch]) {
                if (i < c0) throw std::runtime_error("child structure below parent columns");
                if (i >= c1 && mark[i] != k) { mark[i] = k; rk.push_back(i); }
            }
        }
        std::sort(rk.begin(), rk.end());
        s.sptr[k + 1] = s.sptr[k] + (int64_t)rk.size();
        if (s.parent[k] >= 0 && !rk.empty() && rk.front() < s.c0[s.parent[k]])
            throw std::runtime_error("structure not covered by parent");
        if (s.parent[k] < 0 && !rk.empty()) throw std::runtime_error("root with nonempty structure");
    }
    s.sidx.resize(s.sptr[s.nsn]);
    for (int k = 0; k < s.nsn; k++) std::copy(rows[k].begin(), rows[k].end(), s.sidx.begin() + s.sptr[k]);
    // storage, levels, statistics
    s.loff.assign(s.nsn + 1, 0);
    s.uoff.assign(s.nsn + 1, 0);
    s.level.assign(s.nsn, 0);
    s.nnzL = 0;
    s.flops = 0.0;
    for (int k = 0; k < s.nsn; k++) {
        int64_t nc = s.nc[k], m = s.sptr[k + 1] - s.sptr[k], f = nc + m;
        s.loff[k + 1] = s.loff[k] + front_ld(f) * nc;
        s.uoff[k + 1] = s.uoff[k] + m * (m + 1) / 2;  // packed lower triangle
        s.nnzL += nc * (nc + 1) / 2 + m * nc;
        for (int64_t j = 0; j < nc; j++) {
            double c = (double)(f - j - 1);
            s.flops += c * c + 2.0 * c;
        }
        if (!front_fits_32(nc, m))
            throw std::runtime_error("front of " + std::to_string(nc) + " columns and " + std::to_string(m) +
                                     " struct rows exceeds the kernels' 4 GiB per-front buffer limit; "
                                     "use a smaller nested-dissection leaf or a finer partition");
        s.max_f = std::max<int>(s.max_f, (int)f);
        s.max_nc = std::max<int>(s.max_nc, (int)nc);
        int lv = 0;
        for (int q = s.children_ptr[k]; q < s.children_ptr[k + 1]; q++)
            lv = std::max(lv, s.level[s.children[q]] + 1);
        s.level[k] = lv;
    }
    s.ltotal = s.loff[s.nsn];
    s.utotal = s.uoff[s.nsn];
    s.nlevels = 0;
    for (int k = 0; k < s.nsn; k++) s.nlevels = std::max(s.nlevels, s.level[k] + 1);
    s.lev_ptr.assign(s.nlevels + 1, 0);
    for (int k = 0; k < s.nsn; k++) s.lev_ptr[s.level[k] + 1]++;
    for (int l = 0; l < s.nlevels; l++) s.lev_ptr[l + 1] += s.lev_ptr[l];
    s.lev_sn.assign(s.nsn, 0);
    std::vector<int> lf(s.nlevels, 0);
    for (int k = 0; k < s.nsn; k++) s.lev_sn[s.lev_ptr[s.level[k]] + lf[s.level[k]]++] = k;
    // relative indices of each struct in the parent's front list
    s.rptr.assign(s.nsn + 1, 0);
    s.rel.resize(s.sidx.size());
    for (int k = 0; k < s.nsn; k++) {
        s.rptr[k + 1] = s.sptr[k + 1];
        int p = s.parent[k];
        if (p < 0) continue;
        int pc0 = s.c0[p], pc1 = pc0 + s.nc[p];
        const int *ps = s.sidx.data() + s.sptr[p];
        int64_t pm = s.sptr[p + 1] - s.sptr[p];
        for (int64_t q = s.sptr[k]; q < s.sptr[k + 1]; q++) {
            int i = s.sidx[q];
            int pos;
            if (i < pc1) pos = i - pc0;
            else {
                const int *it = std::lower_bound(ps, ps + pm, i);
                if (it == ps + pm || *it != i) throw std::runtime_error("relind miss");
                pos = s.nc[p] + (int)(it - ps);
            }
            s.rel[q] = pos;
        }
    }
    // assembled CSR entry -> L storage offset
    std::vector<int> col2sn(n);
    for (int k = 0; k < s.nsn; k++)
        for (int j = s.c0[k]; j < s.c0[k] + s.nc[k]; j++) col2sn[j] = k;
    s.amap.resize(t.nnz);
    for (int r = 0; r < n; r++)
        for (int q = t.ia[r]; q < t.ia[r + 1]; q++) {
            int pr = s.iperm[r], pc = s.iperm[t.ja[q]];
            int lo = std::min(pr, pc), hi = std::max(pr, pc);
            int k = col2sn[lo];
            int c0 = s.c0[k], c1 = c0 + s.nc[k];
            int64_t m = s.sptr[k + 1] - s.sptr[k], f = s.nc[k] + m;
            int64_t row;
            if (hi < c1) row = hi - c0;
            else {
                const int *ps = s.sidx.data() + s.sptr[k];
                const int *it = std::lower_bound(ps, ps + m, hi);
                if (it == ps + m || *it != hi) throw std::runtime_error("amap miss");
                row = s.nc[k] + (it - ps);
            }
            s.amap[q] = s.loff[k] + (int64_t)(lo - c0) * front_ld(f) + row;
        }
}

}  // namespace tf
