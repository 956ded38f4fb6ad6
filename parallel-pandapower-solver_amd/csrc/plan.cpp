// plan.cpp -- host-side topology, Jacobian pattern, ordering, symbolic LU and schedule.
// See plan.hpp for the layout conventions.
#include "plan.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <iterator>
#include <limits>
#include <numeric>
#include <queue>
#include <tuple>
#include <stdexcept>

namespace p3smi {

namespace {

// Longest-processing-time dealing: tasks (weights w) of one level go, heaviest first, to
// the currently least-loaded of `ns` lists. `load` carries over between calls so that two
// task kinds of one level share one balance. Returns the slot of each task.
std::vector<int> lpt_deal(const std::vector<int64_t>& w, int ns, std::vector<int64_t>& load) {
    std::vector<int> idx(w.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return w[a] > w[b]; });
    using Key = std::pair<int64_t, int>;
    std::priority_queue<Key, std::vector<Key>, std::greater<Key>> pq;
    for (int s = 0; s < ns; ++s) pq.push({load[s], s});
    std::vector<int> slot(w.size());
    for (int i : idx) {
        Key k = pq.top();
        pq.pop();
        slot[i] = k.second;
        k.first += w[i];
        load[k.second] = k.first;
        pq.push(k);
    }
    return slot;
}

// Group items by (level, slot): returns ptr (nlev*ns + 1) and the item order.
std::vector<int> group_by_slot(const std::vector<int>& level, const std::vector<int>& slot, int nlev, int ns,
                               std::vector<int>& ptr) {
    ptr.assign((size_t)nlev * ns + 1, 0);
    for (size_t i = 0; i < level.size(); ++i) ptr[(size_t)level[i] * ns + slot[i] + 1]++;
    for (size_t q = 1; q < ptr.size(); ++q) ptr[q] += ptr[q - 1];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1), order(level.size());
    for (size_t i = 0; i < level.size(); ++i) order[fill[(size_t)level[i] * ns + slot[i]]++] = (int)i;
    return order;
}

}  // namespace

// ------------------------------------------------------------------ topology
static void build_polar(Topology& t, const int* Yp, const int* Yj, const double* Yx, int n, int nnzY, const int* pv,
                        int npv, const int* pq, int npq) {
    if (n <= 0 || nnzY <= 0) throw std::runtime_error("empty Ybus");
    if (Yp[0] != 0 || Yp[n] != nnzY) throw std::runtime_error("Ybus CSR indptr does not match nnz");
    t.n = n; t.nnzY = nnzY; t.npv = npv; t.npq = npq; t.npvpq = npv + npq; t.m = npv + 2 * npq;
    t.Yp.assign(Yp, Yp + n + 1);
    t.Yj.assign(Yj, Yj + nnzY);
    for (int i = 0; i < n; ++i) {
        if (Yp[i + 1] < Yp[i]) throw std::runtime_error("Ybus CSR indptr is not monotone");
        for (int k = Yp[i]; k < Yp[i + 1]; ++k) {
            if (Yj[k] < 0 || Yj[k] >= n) throw std::runtime_error("Ybus column index out of range");
            if (k > Yp[i] && Yj[k] <= Yj[k - 1]) throw std::runtime_error("Ybus columns must be sorted and unique per row");
        }
    }
    t.G.resize(nnzY); t.B.resize(nnzY);
    for (int k = 0; k < nnzY; ++k) { t.G[k] = Yx[2 * k]; t.B[k] = Yx[2 * k + 1]; }
    t.Ydiag.assign(n, -1);
    for (int i = 0; i < n; ++i)
        for (int k = Yp[i]; k < Yp[i + 1]; ++k)
            if (Yj[k] == i) { t.Ydiag[i] = k; break; }
    t.pv.assign(pv, pv + npv);
    t.pq.assign(pq, pq + npq);
    t.pvpq = t.pv;
    t.pvpq.insert(t.pvpq.end(), t.pq.begin(), t.pq.end());
    t.row_p.assign(n, -1);
    t.row_q.assign(n, -1);
    for (int r = 0; r < t.npvpq; ++r) {
        int b = t.pvpq[r];
        if (b < 0 || b >= n) throw std::runtime_error("pv/pq bus index out of range");
        if (t.row_p[b] != -1) throw std::runtime_error("a bus appears twice in pv/pq");
        t.row_p[b] = r;
    }
    for (int r = 0; r < npq; ++r) t.row_q[t.pq[r]] = t.npvpq + r;
    for (int i = 0; i < n; ++i)
        if (t.row_p[i] >= 0 && t.Ydiag[i] < 0) throw std::runtime_error("a pv/pq bus has no Ybus diagonal entry");
}

// Jacobian pattern, CSR by equation row. Row r < npvpq: P equation of bus pvpq[r];
// row r >= npvpq: Q equation of bus pq[r - npvpq]. Column c < npvpq: Va of pvpq[c];
// column c >= npvpq: Vm of pq[c - npvpq]. J(r, c) is structural iff Ybus(bus(r), bus(c))
// is (p3s/cpp/nr_klu.cpp:218-229).
static void build_jacobian_pattern(Topology& t) {
    std::vector<std::vector<std::pair<int, int>>> rows(t.m);  // (column, 4k + block)
    for (int i = 0; i < t.n; ++i) {
        const int rp = t.row_p[i], rq = t.row_q[i];
        if (rp < 0 && rq < 0) continue;
        for (int k = t.Yp[i]; k < t.Yp[i + 1]; ++k) {
            const int j = t.Yj[k];
            const int cva = t.row_p[j], cvm = t.row_q[j];
            if (rp >= 0 && cva >= 0) rows[rp].push_back({cva, 4 * k + 0});
            if (rp >= 0 && cvm >= 0) rows[rp].push_back({cvm, 4 * k + 1});
            if (rq >= 0 && cva >= 0) rows[rq].push_back({cva, 4 * k + 2});
            if (rq >= 0 && cvm >= 0) rows[rq].push_back({cvm, 4 * k + 3});
        }
    }
    t.Jp.assign(t.m + 1, 0);
    for (int r = 0; r < t.m; ++r) {
        std::sort(rows[r].begin(), rows[r].end());
        t.Jp[r + 1] = t.Jp[r] + (int)rows[r].size();
    }
    t.nnzJ = t.Jp[t.m];
    t.Jj.resize(t.nnzJ); t.Jk.resize(t.nnzJ); t.Jb.resize(t.nnzJ);
    t.posJ.assign(4 * (size_t)t.nnzY, -1);
    for (int r = 0; r < t.m; ++r)
        for (size_t q = 0; q < rows[r].size(); ++q) {
            const int e = t.Jp[r] + (int)q;
            t.Jj[e] = rows[r][q].first;
            t.Jk[e] = rows[r][q].second >> 2;
            t.Jb[e] = rows[r][q].second & 3;
            t.posJ[rows[r][q].second] = e;
        }
    // by column: column c of bus j takes, for the buses i of Ybus column j in ascending order,
    // first the P rows of those with one, then their Q rows
    std::vector<std::vector<int>> ycol(t.n);
    for (int i = 0; i < t.n; ++i)
        for (int k = t.Yp[i]; k < t.Yp[i + 1]; ++k) ycol[t.Yj[k]].push_back(i);
    auto entry = [&](int r, int c) {
        const auto b = t.Jj.begin() + t.Jp[r], e = t.Jj.begin() + t.Jp[r + 1];
        const auto it = std::lower_bound(b, e, c);
        if (it == e || *it != c) throw std::runtime_error("Jacobian pattern: missing column entry");
        return (int)(it - t.Jj.begin());
    };
    t.Cp.assign(t.m + 1, 0);
    t.Ci.clear();
    t.Ce.clear();
    t.Ci.reserve(t.nnzJ);
    t.Ce.reserve(t.nnzJ);
    for (int c = 0; c < t.m; ++c) {
        const int j = c < t.npvpq ? t.pvpq[c] : t.pq[c - t.npvpq];
        for (int pass = 0; pass < 2; ++pass)
            for (int i : ycol[j]) {
                const int r = pass == 0 ? t.row_p[i] : t.row_q[i];
                if (r < 0) continue;
                t.Ci.push_back(r);
                t.Ce.push_back(entry(r, c));
            }
        t.Cp[c + 1] = (int)t.Ci.size();
    }
    if ((int)t.Ci.size() != t.nnzJ) throw std::runtime_error("Jacobian pattern: column and row views differ");
}

// ------------------------------------------------------------------ ordering
// Minimum fill on a weighted graph: eliminate the node whose clique adds the least weight
// (sum of w_a w_b over its non-adjacent live neighbour pairs), ties by weighted degree (sum of
// its live neighbours' w), then lowest index. After an elimination only the clique and its
// neighbours change keys; they are pushed again with a new version, older entries skipped.
std::vector<int> minimum_fill(int m, std::vector<std::vector<int>> adj, const std::vector<int>& wt) {
    std::vector<char> dead(m, 0);
    std::vector<int> mark(m, -1);
    int tick = 0;
    auto cost = [&](int v, int64_t& fill, int64_t& deg) {
        fill = 0;
        deg = 0;
        const auto& nv = adj[v];
        for (int a : nv) deg += wt[a];
        for (size_t i = 0; i < nv.size(); ++i) {
            const int a = nv[i];
            ++tick;
            for (int x : adj[a]) mark[x] = tick;
            for (size_t j = i + 1; j < nv.size(); ++j)
                if (mark[nv[j]] != tick) fill += (int64_t)wt[a] * wt[nv[j]];
        }
    };
    using Key = std::tuple<int64_t, int64_t, int, int>;  // fill, degree, node, version
    std::priority_queue<Key, std::vector<Key>, std::greater<Key>> pq;
    std::vector<int> ver(m, 0);
    for (int v = 0; v < m; ++v) {
        int64_t f, d;
        cost(v, f, d);
        pq.push({f, d, v, 0});
    }
    std::vector<int> order;
    order.reserve(m);
    std::vector<int> stamp(m, -1);
    int st = 0;
    while ((int)order.size() < m) {
        auto [f, d, v, vv] = pq.top();
        pq.pop();
        if (dead[v] || vv != ver[v]) continue;
        dead[v] = 1;
        order.push_back(v);
        std::vector<int> nb;
        nb.swap(adj[v]);
        std::vector<int> touched;
        for (int u : nb) {
            if (dead[u]) continue;
            ++st;
            auto& au = adj[u];
            size_t w = 0;
            for (size_t r = 0; r < au.size(); ++r) {
                int x = au[r];
                if (dead[x]) continue;
                au[w++] = x;
                stamp[x] = st;
            }
            au.resize(w);
            stamp[u] = st;
            for (int x : nb)
                if (!dead[x] && stamp[x] != st) {
                    stamp[x] = st;
                    au.push_back(x);
                }
        }
        // fill changes for the clique and its neighbours
        ++st;
        for (int u : nb) {
            if (dead[u]) continue;
            if (stamp[u] != st) { stamp[u] = st; touched.push_back(u); }
            for (int x : adj[u]) if (stamp[x] != st) { stamp[x] = st; touched.push_back(x); }
        }
        for (int u : touched) {
            int64_t f2, d2;
            cost(u, f2, d2);
            pq.push({f2, d2, u, ++ver[u]});
        }
    }
    return order;
}

// Minimum fill on the buses, then each bus's P equation before its Q equation. The two
// equations of a bus couple to the same neighbours, so eliminated side by side they form a
// pivot pair (parent p + 1, S_p = {p + 1} + S_{p+1}) at every PQ bus, and the 2x2-block
// schedule loads one row of operands for both. A bus weighs its equations (2 for a PQ bus, 1
// for a PV bus): a fill edge between buses a and b is w_a w_b fill entries of the Jacobian's
// factors. The batch LU per launch, each step measured against the previous one on one
// machine (the LU is bound by its operand reads, so fewer operand pairs pay):
//   - minimum degree over single equations (the two side by side at 89% of the PQ buses of
//     the 10k grid) -> minimum degree on the buses counting neighbour buses: n1_10k
//     40.7 -> 34.5 ms, ts2000 4.62 -> 4.21 ms, with 0.7-2% more fill;
//   - counting neighbour equations in the degree: n1_10k 34.5 -> 32.0 ms, ts2000 4.04 ->
//     3.91 ms; on n1_10k 4.5% less fill, 17% fewer operand pairs, 176 levels against 204;
//   - minimum fill instead of minimum degree: n1_10k 31.8 -> 29.2 ms, ts2000 3.84 -> 3.73 ms;
//     on n1_10k 3.2% less fill, 15% fewer operand pairs, 168 levels. Multiple elimination
//     (an independent set of minimum-degree buses per round) gave 2% fewer pairs and more
//     levels, and was not kept.
std::vector<int> bus_minimum_fill(const Topology& t, const std::vector<std::vector<int>>& adj) {
    const int m = t.m;
    std::vector<int> bus_of(m, -1);
    for (int i = 0; i < t.n; ++i) {
        if (t.row_p[i] >= 0) bus_of[t.row_p[i]] = i;
        if (t.row_q[i] >= 0) bus_of[t.row_q[i]] = i;
    }
    std::vector<int> bidx(t.n, -1), buses;  // buses with equations, in first-equation order
    for (int v = 0; v < m; ++v)
        if (bidx[bus_of[v]] < 0) {
            bidx[bus_of[v]] = (int)buses.size();
            buses.push_back(bus_of[v]);
        }
    const int nb = (int)buses.size();
    std::vector<std::vector<int>> badj(nb);
    for (int v = 0; v < m; ++v)
        for (int u : adj[v]) {
            const int a = bidx[bus_of[v]], c = bidx[bus_of[u]];
            if (a != c) badj[a].push_back(c);
        }
    for (auto& l : badj) {
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
    }
    std::vector<int> perm;
    perm.reserve(m);
    std::vector<int> wt(nb);
    for (int k = 0; k < nb; ++k) wt[k] = t.row_q[buses[k]] >= 0 ? 2 : 1;
    for (int k : minimum_fill(nb, std::move(badj), wt)) {
        const int b = buses[k];
        if (t.row_p[b] >= 0) perm.push_back(t.row_p[b]);
        if (t.row_q[b] >= 0) perm.push_back(t.row_q[b]);
    }
    return perm;
}

// Symmetric static-pivot LU: order, elimination tree, strict column patterns S_p,
// positions, operand pairs. The Jacobian pattern is structurally symmetric with a full
// diagonal, so the pattern of U(p,:) is S_p as well.
static void build_lu(Topology& t, int ordering) {
    const int m = t.m;
    std::vector<std::vector<int>> adj(m);
    for (int r = 0; r < m; ++r)
        for (int e = t.Jp[r]; e < t.Jp[r + 1]; ++e)
            if (t.Jj[e] != r) adj[r].push_back(t.Jj[e]);
    // symmetrise (the pattern already is; this guards malformed inputs)
    for (int r = 0; r < m; ++r)
        for (int c : adj[r]) if (!std::binary_search(t.Jj.begin() + t.Jp[c], t.Jj.begin() + t.Jp[c + 1], r))
            throw std::runtime_error("Jacobian pattern is not structurally symmetric (Ybus pattern must be symmetric)");
    t.ordering = ordering;
    if (ordering == 2) {
        t.perm.resize(m);
        std::iota(t.perm.begin(), t.perm.end(), 0);
    } else {
        t.perm = bus_minimum_fill(t, adj);
    }
    t.pinv.assign(m, -1);
    for (int p = 0; p < m; ++p) t.pinv[t.perm[p]] = p;

    // symbolic factorisation on the elimination tree: S_p = (higher neighbours of p)
    // united with S_c \ {p} of every child c.
    t.parent.assign(m, -1);
    std::vector<int> head(m, -1), next(m, -1), mark(m, -1);
    t.Lp.assign(m + 1, 0);
    t.Li.clear();
    for (int p = 0; p < m; ++p) {
        const size_t start = t.Li.size();
        mark[p] = p;
        for (int c : adj[t.perm[p]]) {
            int i = t.pinv[c];
            if (i > p && mark[i] != p) { mark[i] = p; t.Li.push_back(i); }
        }
        for (int ch = head[p]; ch != -1; ch = next[ch])
            for (int q = t.Lp[ch]; q < t.Lp[ch + 1]; ++q) {
                int i = t.Li[q];
                if (i > p && mark[i] != p) { mark[i] = p; t.Li.push_back(i); }
            }
        std::sort(t.Li.begin() + start, t.Li.end());
        t.Lp[p + 1] = (int)t.Li.size();
        if (t.Li.size() > start) {
            const int par = t.Li[start];
            t.parent[p] = par;
            next[p] = head[par];
            head[par] = p;
        }
    }
    t.height.assign(m, 0);
    for (int p = 0; p < m; ++p)
        if (t.parent[p] >= 0) t.height[t.parent[p]] = std::max(t.height[t.parent[p]], t.height[p] + 1);
    t.depth.assign(m, 0);
    for (int p = m - 1; p >= 0; --p)
        if (t.parent[p] >= 0) t.depth[p] = t.depth[t.parent[p]] + 1;

    // positions
    t.base.assign(m + 1, 0);
    for (int p = 0; p < m; ++p) {
        int64_t nxt = (int64_t)t.base[p] + 2 * (int64_t)(t.Lp[p + 1] - t.Lp[p]) + 1;
        if (nxt >= std::numeric_limits<int>::max()) throw std::runtime_error("factor pattern too large for int32 positions");
        t.base[p + 1] = (int)nxt;
    }
    t.nnz_lu = t.base[m];
    // position of (a, b) in pivot space: diagonal, U(a, b) for a < b, L(a, b) for a > b
    auto pos = [&](int a, int b) -> int {
        if (a == b) return t.base[a];
        const int lo = std::min(a, b), hi = std::max(a, b);
        const int* beg = t.Li.data() + t.Lp[lo];
        const int* end = t.Li.data() + t.Lp[lo + 1];
        const int* it = std::lower_bound(beg, end, hi);
        if (it == end || *it != hi) throw std::runtime_error("internal: entry outside the factor pattern");
        const int r = (int)(it - beg);
        const int cnt = t.Lp[lo + 1] - t.Lp[lo];
        return a < b ? t.base[lo] + 1 + r : t.base[lo] + 1 + cnt + r;
    };
    t.apos.resize(t.nnzJ);
    for (int r = 0; r < m; ++r)
        for (int e = t.Jp[r]; e < t.Jp[r + 1]; ++e) t.apos[e] = pos(t.pinv[r], t.pinv[t.Jj[e]]);
    t.xva.assign(t.n, -1);
    t.xvm.assign(t.n, -1);
    for (int i = 0; i < t.n; ++i) {
        if (t.row_p[i] >= 0) t.xva[i] = t.pinv[t.row_p[i]];
        if (t.row_q[i] >= 0) t.xvm[i] = t.pinv[t.row_q[i]];
    }

    // operand pairs, grouped by target position: pivot t contributes L(a,t) U(t,b) to
    // every entry (a, b) with a, b in S_t
    std::vector<int64_t> cnt(t.nnz_lu + 1, 0);
    int64_t npairs = 0;
    for (int p = 0; p < m; ++p) {
        const int c = t.Lp[p + 1] - t.Lp[p];
        npairs += (int64_t)c * c;
    }
    if (npairs >= std::numeric_limits<int>::max()) throw std::runtime_error("too many LU operand pairs for int32");
    t.n_pairs = npairs;
    std::vector<int> tgt((size_t)npairs), opa((size_t)npairs), opb((size_t)npairs);
    int64_t q = 0;
    for (int p = 0; p < m; ++p) {
        const int c = t.Lp[p + 1] - t.Lp[p];
        for (int ia = 0; ia < c; ++ia) {
            const int a = t.Li[t.Lp[p] + ia];
            for (int ib = 0; ib < c; ++ib) {
                const int b = t.Li[t.Lp[p] + ib];
                tgt[q] = pos(a, b);
                opa[q] = t.base[p] + 1 + c + ia;  // L(a, p)
                opb[q] = t.base[p] + 1 + ib;      // U(p, b)
                ++q;
            }
        }
    }
    for (int64_t e = 0; e < npairs; ++e) cnt[tgt[e] + 1]++;
    for (int64_t e = 0; e < t.nnz_lu; ++e) cnt[e + 1] += cnt[e];
    t.pa.resize((size_t)npairs);
    t.pb.resize((size_t)npairs);
    {
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        // pivots ascending within a target: the order a left-looking sweep would use
        for (int64_t e = 0; e < npairs; ++e) {
            const int64_t d = fill[tgt[e]]++;
            t.pa[d] = opa[e];
            t.pb[d] = opb[e];
        }
    }
    // forward-substitution rows: L(i, j) entries of row i
    t.Fp.assign(m + 1, 0);
    for (int p = 0; p < m; ++p)
        for (int q2 = t.Lp[p]; q2 < t.Lp[p + 1]; ++q2) t.Fp[t.Li[q2] + 1]++;
    for (int i = 0; i < m; ++i) t.Fp[i + 1] += t.Fp[i];
    t.Fpos.resize(t.Fp[m]);
    t.Fcol.resize(t.Fp[m]);
    {
        std::vector<int> fill(t.Fp.begin(), t.Fp.end() - 1);
        for (int p = 0; p < m; ++p) {
            const int c = t.Lp[p + 1] - t.Lp[p];
            for (int r = 0; r < c; ++r) {
                const int i = t.Li[t.Lp[p] + r];
                const int d = fill[i]++;
                t.Fpos[d] = t.base[p] + 1 + c + r;
                t.Fcol[d] = p;
            }
        }
    }
    t.prange.assign(t.nnz_lu + 1, 0);
    for (int64_t e = 0; e <= t.nnz_lu; ++e) t.prange[e] = (int)cnt[e];
}

void build_schedule(Topology& t, int nslots) { t.sched = make_schedule(t, nslots, 0); }

Schedule make_schedule(const Topology& t, int nslots, int group) {
    if (nslots < 1) throw std::runtime_error("nslots must be >= 1");
    const int m = t.m;
    const std::vector<int>& prange = t.prange;
    std::vector<int> asrc(t.nnz_lu, -1);
    for (int e = 0; e < t.nnzJ; ++e) asrc[t.apos[e]] = e;

    Schedule s;
    s.nslots = nslots;
    s.nlev = 0;
    for (int p = 0; p < m; ++p) s.nlev = std::max(s.nlev, t.height[p] + 1);
    s.nblev = 0;
    for (int p = 0; p < m; ++p) s.nblev = std::max(s.nblev, t.depth[p] + 1);

    // items per level: A = diagonal + U entries + forward rows (one balance), B = L entries;
    // an item is an entry's sum or a chunk of it: operand pairs [e0, e1), partial slots in / out
    struct Item {
        int d, e0, e1;
        int pin = -1, pout = -1;  // part ids
    };
    std::vector<std::vector<Item>> levA(s.nlev), levB(s.nlev);
    std::vector<std::vector<int>> levF(s.nlev), levW(s.nblev);
    for (int p = 0; p < m; ++p) {
        const int h = t.height[p];
        const int c = t.Lp[p + 1] - t.Lp[p];
        auto item = [&](int d) { return Item{d, prange[d], prange[d + 1]}; };
        levA[h].push_back(item(t.base[p]));
        for (int r = 0; r < c; ++r) levA[h].push_back(item(t.base[p] + 1 + r));
        for (int r = 0; r < c; ++r) levB[h].push_back(item(t.base[p] + 1 + c + r));
        levF[h].push_back(p);
        levW[t.depth[p]].push_back(p);
    }
    // ---- long sums cut into chunks (ONE_SPLIT_MIN): a chunk can run in sub-level index sl
    // (2 l for A, 2 l + 1 for B) once the L and U entries of all its pairs' pivots are final,
    // i.e. after sub-level B of the pivot's level
    std::vector<int> ready_pos(t.nnz_lu, 0);
    for (int p = 0; p < m; ++p)
        for (int q = t.base[p]; q < t.base[p + 1]; ++q) ready_pos[q] = 2 * t.height[p] + 2;
    std::vector<int> part_sub;  // sub-level that writes each part (read in the next one)
    std::vector<std::pair<int, Item>> chunks;
    auto cut = [&](Item& it, int sub) {
        const int len = it.e1 - it.e0;
        if (len <= ONE_SPLIT_MIN) return;
        std::vector<std::array<int, 3>> ch;  // (sub-level, beg, end), the last chunk first
        int end = len, sl = sub;
        while (end > 0) {
            int beg = std::max(0, end - ONE_CHUNK);
            if (beg < ONE_TAIL) beg = 0;
            int need = 0;
            for (int e = it.e0 + beg; e < it.e0 + end; ++e) need = std::max(need, ready_pos[t.pa[e]]);
            if (sl < 0 || need > sl) return;  // cannot run this early: the entry stays whole
            ch.push_back({sl, beg, end});
            end = beg;
            --sl;
        }
        if (ch.size() < 2) return;
        int prev = -1;
        for (size_t k = ch.size(); k-- > 1;) {  // earlier chunks first
            Item c{it.d, it.e0 + ch[k][1], it.e0 + ch[k][2]};
            c.pin = prev;
            c.pout = (int)part_sub.size();
            part_sub.push_back(ch[k][0]);
            prev = c.pout;
            chunks.push_back({ch[k][0], c});
        }
        it.e1 = it.e0 + ch[0][2];
        it.e0 = it.e0 + ch[0][1];
        it.pin = prev;
        ++s.ncut;
    };
    for (int l = 0; l < s.nlev; ++l) {
        for (Item& it : levA[l]) cut(it, 2 * l);
        for (Item& it : levB[l]) cut(it, 2 * l + 1);
    }
    for (auto& c : chunks) (c.first % 2 == 0 ? levA : levB)[c.first / 2].push_back(c.second);
    // partial slots: parts written in sub-levels of one parity share slots (read in the next)
    std::vector<int> nsub(2 * (size_t)s.nlev + 1, 0), prank(part_sub.size());
    for (size_t q = 0; q < part_sub.size(); ++q) prank[q] = nsub[part_sub[q]]++;
    int nE = 0, nO = 0;
    for (size_t w = 0; w < nsub.size(); ++w) (w % 2 == 0 ? nE : nO) = std::max(w % 2 == 0 ? nE : nO, nsub[w]);
    s.nparts = nE + nO;
    const int zero = (int)t.nnz_lu, one = zero + 1;
    auto part_pos = [&](int q) { return zero + 2 + (part_sub[q] % 2 == 0 ? prank[q] : nE + prank[q]); };

    // ---- forward and backward rows cut the same way, their partial sums in solution-vector
    // slots (rows m, m + 1, ...; a chunk's first pair (one row, slot) brings one back). Forward
    // rows run in sub-level A: a chunk over columns j can run in sub-level A of level l once
    // L(i, j) and y_j are final, l > height(j). Backward rows (columns descending) run one
    // level per depth: a chunk can run at depth level d once every x_j is, d > depth(j).
    struct SItem {
        int p, e0, e1;            // forward: entries [e0, e1) of Fp; backward: S_p columns [e0, e1)
        int pin = -1, pout = -1;  // slot ids
    };
    std::vector<std::vector<SItem>> levFi(s.nlev), levWi(s.nblev);
    for (int l = 0; l < s.nlev; ++l)
        for (int p : levF[l]) levFi[l].push_back(SItem{p, t.Fp[p], t.Fp[p + 1]});
    for (int l = 0; l < s.nblev; ++l)
        for (int p : levW[l]) levWi[l].push_back(SItem{p, t.Lp[p], t.Lp[p + 1]});
    std::vector<int> fslot_lev, bslot_lev;  // level that writes each slot (read at the next)
    std::vector<std::pair<int, SItem>> fchunks, bchunks;
    // forward: chunks of consecutive entries at levels l, l - 1, ... (sub-level A of each)
    auto cut_fwd = [&](SItem& it, int l) {
        const int len = it.e1 - it.e0;
        if (len <= ONE_SPLIT_MIN) return;
        std::vector<std::array<int, 3>> ch;
        int end = len, lv = l;
        while (end > 0) {
            int beg = std::max(0, end - ONE_CHUNK);
            if (beg < ONE_TAIL) beg = 0;
            int need = 0;
            for (int e = it.e0 + beg; e < it.e0 + end; ++e) need = std::max(need, t.height[t.Fcol[e]] + 1);
            if (lv < 0 || need > lv) return;
            ch.push_back({lv, beg, end});
            end = beg;
            --lv;
        }
        if (ch.size() < 2) return;
        int prev = -1;
        for (size_t k = ch.size(); k-- > 1;) {
            SItem c{it.p, it.e0 + ch[k][1], it.e0 + ch[k][2]};
            c.pin = prev;
            c.pout = (int)fslot_lev.size();
            fslot_lev.push_back(ch[k][0]);
            prev = c.pout;
            fchunks.push_back({ch[k][0], c});
        }
        it.e1 = it.e0 + ch[0][2];
        it.e0 = it.e0 + ch[0][1];
        it.pin = prev;
    };
    // backward: the row's columns in descending order, chunks of that sequence at depths d, d - 1,
    // ...; a chunk is the ascending range [e0, e1) of S_p walked downwards
    auto cut_bwd = [&](SItem& it, int d) {
        const int len = it.e1 - it.e0;
        if (len <= ONE_SPLIT_MIN) return;
        std::vector<std::array<int, 3>> ch;  // (depth, k0, k1): positions k in the descending walk
        int end = len, lv = d;
        while (end > 0) {
            int beg = std::max(0, end - ONE_CHUNK);
            if (beg < ONE_TAIL) beg = 0;
            int need = 0;
            for (int k = beg; k < end; ++k) need = std::max(need, t.depth[t.Li[it.e1 - 1 - k]] + 1);
            if (lv < 0 || need > lv) return;
            ch.push_back({lv, beg, end});
            end = beg;
            --lv;
        }
        if (ch.size() < 2) return;
        const int e1 = it.e1;
        int prev = -1;
        for (size_t k = ch.size(); k-- > 1;) {
            SItem c{it.p, e1 - ch[k][2], e1 - ch[k][1]};
            c.pin = prev;
            c.pout = (int)bslot_lev.size();
            bslot_lev.push_back(ch[k][0]);
            prev = c.pout;
            bchunks.push_back({ch[k][0], c});
        }
        it.e0 = e1 - ch[0][2];
        it.e1 = e1 - ch[0][1];
        it.pin = prev;
    };
    for (int l = 0; l < s.nlev; ++l)
        for (SItem& it : levFi[l]) cut_fwd(it, l);
    for (int l = 0; l < s.nblev; ++l)
        for (SItem& it : levWi[l]) cut_bwd(it, l);
    for (auto& c : fchunks) levFi[c.first].push_back(c.second);
    for (auto& c : bchunks) levWi[c.first].push_back(c.second);
    // slots: written at level v and read at v + 1, so levels of one parity share them; forward
    // slots are all read before the backward rows start, which reuse the same rows
    auto slot_map = [&](const std::vector<int>& lev_of, int nlevels, std::vector<int>& slot) {
        std::vector<int> cnt((size_t)nlevels + 1, 0);
        slot.assign(lev_of.size(), 0);
        for (size_t q = 0; q < lev_of.size(); ++q) slot[q] = cnt[lev_of[q]]++;
        int e = 0, o = 0;
        for (int v = 0; v <= nlevels; ++v) (v % 2 == 0 ? e : o) = std::max(v % 2 == 0 ? e : o, cnt[v]);
        for (size_t q = 0; q < lev_of.size(); ++q) slot[q] = t.m + (lev_of[q] % 2 == 0 ? slot[q] : e + slot[q]);
        return e + o;
    };
    std::vector<int> fslot, bslot;
    s.nxparts = std::max(slot_map(fslot_lev, s.nlev, fslot), slot_map(bslot_lev, s.nblev, bslot));

    std::vector<int> lA, sA, lB, sB, lF, sF, lW, sW;
    std::vector<Item> itemsA, itemsB;
    std::vector<SItem> itemsF, itemsW;
    // dealing units: single entries (group 0) or runs of up to `group` consecutive entries
    // of one pivot's U row / L column, which share one operand of every pair (L(p,t) along a
    // U row, U(t,p) down an L column) and so reuse it from cache within one task list; a
    // chunk of a cut sum is a unit of its own
    auto whole = [&](const Item& it) { return it.pin < 0 && it.pout < 0; };
    auto units_of = [&](const std::vector<Item>& items, std::vector<std::vector<Item>>& units) {
        units.clear();
        for (size_t i = 0; i < items.size(); ++i) {
            const Item& it = items[i];
            const bool cont = group > 0 && !units.empty() && (int)units.back().size() < group && whole(it) &&
                              whole(units.back().back()) && units.back().back().d + 1 == it.d;
            if (cont) units.back().push_back(it);
            else units.push_back({it});
        }
    };
    auto weight = [&](const Item& it) { return (int64_t)2 + (it.e1 - it.e0) + (it.pin >= 0 ? 1 : 0); };
    std::vector<std::vector<Item>> ua, ub;
    for (int l = 0; l < s.nlev; ++l) {
        units_of(levA[l], ua);
        std::vector<int64_t> load(nslots, 0);
        std::vector<int64_t> w;
        for (auto& u : ua) {
            int64_t wt = 0;
            for (const Item& it : u) wt += weight(it);
            w.push_back(wt);
        }
        for (const SItem& it : levFi[l]) w.push_back(2 + (it.e1 - it.e0) + (it.pin >= 0 ? 1 : 0));
        std::vector<int> slot = lpt_deal(w, nslots, load);
        for (size_t i = 0; i < ua.size(); ++i)
            for (const Item& it : ua[i]) { itemsA.push_back(it); lA.push_back(l); sA.push_back(slot[i]); }
        for (size_t i = 0; i < levFi[l].size(); ++i) {
            itemsF.push_back(levFi[l][i]); lF.push_back(l); sF.push_back(slot[ua.size() + i]);
        }
        units_of(levB[l], ub);
        std::vector<int64_t> loadB(nslots, 0), wb;
        for (auto& u : ub) {
            int64_t wt = 0;
            for (const Item& it : u) wt += weight(it);
            wb.push_back(wt);
        }
        std::vector<int> slotB = lpt_deal(wb, nslots, loadB);
        for (size_t i = 0; i < ub.size(); ++i)
            for (const Item& it : ub[i]) { itemsB.push_back(it); lB.push_back(l); sB.push_back(slotB[i]); }
    }
    for (int l = 0; l < s.nblev; ++l) {
        std::vector<int64_t> load(nslots, 0), w;
        for (const SItem& it : levWi[l]) w.push_back(2 + (it.e1 - it.e0) + (it.pin >= 0 ? 1 : 0));
        std::vector<int> slot = lpt_deal(w, nslots, load);
        for (size_t i = 0; i < levWi[l].size(); ++i) {
            itemsW.push_back(levWi[l][i]); lW.push_back(l); sW.push_back(slot[i]);
        }
    }
    // pivot of each factor position (for L entries: the diagonal they are scaled by)
    std::vector<int> diag_of(t.nnz_lu, -1);
    std::vector<char> is_diag(t.nnz_lu, 0);
    for (int p = 0; p < m; ++p) {
        const int c = t.Lp[p + 1] - t.Lp[p];
        is_diag[t.base[p]] = 1;
        for (int r = 0; r < c; ++r) diag_of[t.base[p] + 1 + c + r] = t.base[p];
    }
    auto task_of = [&](const Item& it) {
        Task k{it.d, asrc[it.d], it.e0, it.e1};
        if (it.pin >= 0) k.pin = part_pos(it.pin);
        if (it.pout >= 0) {
            k.dst = part_pos(it.pout);
            k.asrc = -1;
            k.part = 1;
        }
        return k;
    };
    auto ordA = group_by_slot(lA, sA, s.nlev, nslots, s.ptrA);
    for (int i : ordA) {
        const Item& it = itemsA[i];
        const Task k = task_of(it);
        s.taskA.push_back(k);
        s.isdiagA.push_back((unsigned char)(k.part ? 0 : is_diag[it.d]));
    }
    auto ordB = group_by_slot(lB, sB, s.nlev, nslots, s.ptrB);
    for (int i : ordB) {
        const Item& it = itemsB[i];
        const Task k = task_of(it);
        s.taskB.push_back(k);
        s.diagB.push_back(k.part ? zero : diag_of[it.d]);
    }
    (void)one;
    auto ordF = group_by_slot(lF, sF, s.nlev, nslots, s.ptrF);
    for (int i : ordF) {
        const SItem& it = itemsF[i];
        FwdRow f{it.p, t.perm[it.p], it.e0, it.e1};
        if (it.pin >= 0) f.pin = fslot[it.pin];
        if (it.pout >= 0) f.dst = fslot[it.pout];
        s.fwd.push_back(f);
    }
    auto ordW = group_by_slot(lW, sW, s.nblev, nslots, s.ptrW);
    for (int i : ordW) {
        const SItem& it = itemsW[i];
        BwdRow b{it.p, t.base[it.p], it.e0, it.e1};
        if (it.pin >= 0) b.pin = bslot[it.pin];
        if (it.pout >= 0) b.dst = bslot[it.pout];
        s.bwd.push_back(b);
    }
    return s;
}

BlockSchedule make_block_schedule(const Topology& t, int nslots) {
    if (nslots < 1) throw std::runtime_error("nslots must be >= 1");
    if (t.nnz_lu >= std::numeric_limits<int>::max()) throw std::runtime_error("factor pattern too large");
    const int m = t.m;
    const int zero = (int)t.nnz_lu;  // the zero row
    auto S_beg = [&](int p) { return t.Li.data() + t.Lp[p]; };
    auto S_cnt = [&](int p) { return t.Lp[p + 1] - t.Lp[p]; };
    auto in_S = [&](int p, int i) { return std::binary_search(S_beg(p), S_beg(p) + S_cnt(p), i); };
    // rank of i in S_p (must be there)
    auto rank = [&](int p, int i) {
        const int* b = S_beg(p);
        const int* e = b + S_cnt(p);
        const int* it = std::lower_bound(b, e, i);
        if (it == e || *it != i) throw std::runtime_error("internal: block entry outside the factor pattern");
        return (int)(it - b);
    };
    auto posU = [&](int p, int j) { return t.base[p] + 1 + rank(p, j); };             // U(p, j), j > p
    auto posL = [&](int i, int p) { return t.base[p] + 1 + S_cnt(p) + rank(p, i); };  // L(i, p), i > p
    std::vector<int> asrc(t.nnz_lu, -1);
    for (int e = 0; e < t.nnzJ; ++e) asrc[t.apos[e]] = e;

    // row patterns of L: rows[i] = pivots t < i with i in S_t, ascending
    std::vector<std::vector<int>> rows(m);
    for (int p = 0; p < m; ++p)
        for (int q = t.Lp[p]; q < t.Lp[p + 1]; ++q) rows[t.Li[q]].push_back(p);
    // pivot pairs, greedily from the first pivot
    std::vector<int> gsize(m, 1);  // 2 = first of a pair, 0 = second of a pair
    BlockSchedule s;
    s.nslots = nslots;
    for (int p = 0; p + 1 < m; ++p) {
        if (gsize[p] != 1 || t.parent[p] != p + 1) continue;
        const int c0 = S_cnt(p), c1 = S_cnt(p + 1);
        if (c0 != c1 + 1 || S_beg(p)[0] != p + 1 || !std::equal(S_beg(p) + 1, S_beg(p) + c0, S_beg(p + 1))) continue;
        gsize[p] = 2;
        gsize[p + 1] = 0;
        ++s.npairs;
    }
    // groups of a pattern S (ascending): (lead, size); a pair with one member in S counts
    // as a single pivot there
    auto groups_of = [&](const int* b, int cnt, std::vector<std::pair<int, int>>& out) {
        out.clear();
        for (int k = 0; k < cnt; ++k) {
            const int i = b[k];
            if (gsize[i] == 2 && k + 1 < cnt && b[k + 1] == i + 1) {
                out.push_back({i, 2});
                ++k;
            } else {
                out.push_back({i, 1});
            }
        }
    };
    s.nlev = 0;
    for (int p = 0; p < m; ++p) s.nlev = std::max(s.nlev, t.height[p] + 1);
    s.nblev = 0;
    for (int p = 0; p < m; ++p) s.nblev = std::max(s.nblev, t.depth[p] + 1);

    // blocks per level and sub-level, each with its shared pivots
    struct Blk {
        BlockTask k;
        int r0, c0, p0;        // lead row, lead column, lead pivot of the owning group
        int tb = 0, te = -1;   // chunk: shared pivots [tb, te) of the block's list (te -1: all)
        int pin = -1;          // chunk: partial sums to start from (part id), -1: from zero
        int pout = -1;         // chunk other than the last: part id its finish stores (BLK_PART)
    };
    // the pairs' second pivots never start a block; their own rows and columns are in the
    // pair's blocks
    std::vector<std::vector<Blk>> levA(s.nlev), levB(s.nlev);
    std::vector<std::vector<int>> levF(s.nlev), levW(s.nblev);
    std::vector<std::pair<int, int>> cg;
    for (int p0 = 0; p0 < m; ++p0) {
        if (gsize[p0] == 0) continue;
        const int sz = gsize[p0] == 2 ? 2 : 1;
        const int last = p0 + sz - 1;
        const int lev = t.height[last];
        auto src = [&](int pos) { return pos >= 0 ? asrc[pos] : -1; };
        {  // diagonal block
            Blk b{};
            b.r0 = b.c0 = b.p0 = p0;
            b.k.nr = b.k.nc = sz;
            b.k.tgt0 = t.base[p0];
            b.k.src[0] = src(t.base[p0]);
            b.k.src[1] = b.k.src[2] = b.k.src[3] = -1;
            if (sz == 2) {
                b.k.kind = BLK_DIAG2;
                b.k.tgt1 = t.base[p0 + 1];
                b.k.aux0 = posU(p0, p0 + 1);
                b.k.aux1 = posL(p0 + 1, p0);
                b.k.src[1] = src(b.k.aux0);
                b.k.src[2] = src(b.k.aux1);
                b.k.src[3] = src(t.base[p0 + 1]);
                if (b.k.src[3] < 0) throw std::runtime_error("internal: pivot without a Jacobian entry");
            } else {
                b.k.kind = BLK_DIAG1;
                b.k.tgt1 = b.k.aux0 = b.k.aux1 = 0;
            }
            if (b.k.src[0] < 0) throw std::runtime_error("internal: pivot without a Jacobian entry");
            levA[lev].push_back(b);
        }
        groups_of(S_beg(last), S_cnt(last), cg);
        for (auto [g0, gs] : cg) {
            // U block P x C
            Blk u{};
            u.r0 = u.p0 = p0;
            u.c0 = g0;
            u.k.nr = sz;
            u.k.nc = gs;
            u.k.tgt0 = posU(p0, g0);
            u.k.src[0] = src(u.k.tgt0);
            u.k.src[1] = gs == 2 ? src(u.k.tgt0 + 1) : -1;
            if (sz == 2) {
                u.k.kind = BLK_U2;
                u.k.tgt1 = posU(p0 + 1, g0);
                u.k.aux0 = posL(p0 + 1, p0);
                u.k.src[2] = src(u.k.tgt1);
                u.k.src[3] = gs == 2 ? src(u.k.tgt1 + 1) : -1;
                levB[lev].push_back(u);
            } else {
                u.k.kind = BLK_U1;
                u.k.src[2] = u.k.src[3] = -1;
                levA[lev].push_back(u);
            }
            // L block C x P
            Blk l{};
            l.r0 = g0;
            l.c0 = l.p0 = p0;
            l.k.nr = gs;
            l.k.nc = sz;
            l.k.tgt0 = posL(g0, p0);
            l.k.aux0 = t.base[p0];
            l.k.src[0] = src(l.k.tgt0);
            l.k.src[2] = gs == 2 ? src(l.k.tgt0 + 1) : -1;
            if (sz == 2) {
                l.k.kind = BLK_L2;
                l.k.tgt1 = posL(g0, p0 + 1);
                l.k.aux1 = t.base[p0 + 1];
                l.k.src[1] = src(l.k.tgt1);
                l.k.src[3] = gs == 2 ? src(l.k.tgt1 + 1) : -1;
            } else {
                l.k.kind = BLK_L1;
                l.k.src[1] = l.k.src[3] = -1;
            }
            levB[lev].push_back(l);
        }
        levW[t.depth[last]].push_back(p0);  // backward rows by pivot group, at the group's depth
    }
    for (int p0 = 0; p0 < m; ++p0)  // forward rows by pivot group, at the group's level
        if (gsize[p0] != 0) levF[t.height[p0 + (gsize[p0] == 2 ? 1 : 0)]].push_back(p0);

    // pivots of a block (R, C) of group P: t < p0 with a row of R and a column of C in S_t
    std::vector<int> tl, ra, ca;
    auto merge = [](const std::vector<int>& x, const std::vector<int>* y, std::vector<int>& out) {
        out.clear();
        if (!y) { out = x; return; }
        std::set_union(x.begin(), x.end(), y->begin(), y->end(), std::back_inserter(out));
    };
    auto shared = [&](const Blk& b) {
        merge(rows[b.r0], b.k.nr == 2 ? &rows[b.r0 + 1] : nullptr, ra);
        merge(rows[b.c0], b.k.nc == 2 ? &rows[b.c0 + 1] : nullptr, ca);
        tl.clear();
        size_t i = 0, j = 0;
        while (i < ra.size() && j < ca.size()) {
            if (ra[i] < ca[j]) ++i;
            else if (ca[j] < ra[i]) ++j;
            else {
                if (ra[i] < b.p0) tl.push_back(ra[i]);
                ++i;
                ++j;
            }
        }
    };
    // ---- long sums cut into chunks (see BLK_SPLIT_MIN): a chunk of shared pivots can run in
    // sub-level index sl (2 l for A, 2 l + 1 for B) when every pivot's L and U rows are final
    // before it, i.e. after sub-level B of the level of that pivot's group
    const int one = zero + 1;
    std::vector<int> ready(m);
    for (int p = 0; p < m; ++p) {
        const int lead = gsize[p] == 0 ? p - 1 : p;
        ready[p] = 2 * t.height[lead + (gsize[lead] == 2 ? 1 : 0)] + 2;
    }
    struct Part { int sub; };  // a partial slot use: written in sub-level `sub`, read in sub + 1
    std::vector<Part> parts;
    std::vector<std::pair<int, Blk>> chunks;  // (sub-level index, chunk task)
    auto fine_cut = [](int len, int64_t total) {  // see BLK_FINE_MIN
        return len > BLK_FINE_MIN && total < (int64_t)BLK_REF_LISTS * BLK_FINE_MIN &&
               (int64_t)len * BLK_REF_LISTS > total;
    };
    // pivots summed per sub-level, before any cut (BLK_FINE_MIN)
    std::vector<int64_t> subtot(2 * (size_t)s.nlev, 0);
    for (int l = 0; l < s.nlev; ++l) {
        for (Blk& b : levA[l]) subtot[2 * l] += (shared(b), (int64_t)tl.size());
        for (Blk& b : levB[l]) subtot[2 * l + 1] += (shared(b), (int64_t)tl.size());
    }
    auto cut = [&](Blk& b, int sub) {
        shared(b);
        const int len = (int)tl.size();
        const bool fine = fine_cut(len, subtot[sub]);
        const int chunk = fine ? BLK_FINE_CHUNK : BLK_CHUNK, tail = fine ? BLK_FINE_TAIL : BLK_TAIL;
        if (len <= (fine ? BLK_FINE_MIN : BLK_SPLIT_MIN)) return;
        std::vector<std::array<int, 3>> ch;  // (sub-level, beg, end), the last chunk first
        int end = len, sl = sub;
        while (end > 0) {
            int beg = std::max(0, end - chunk);
            if (beg < tail) beg = 0;
            int need = 0;
            for (int i = beg; i < end; ++i) need = std::max(need, ready[tl[i]]);
            if (sl < 0 || need > sl) return;  // cannot run this early: the block stays whole
            ch.push_back({sl, beg, end});
            end = beg;
            --sl;
        }
        if (ch.size() < 2) return;
        int prev = -1;
        for (size_t k = ch.size(); k-- > 1;) {  // earlier chunks first
            Blk c = b;
            c.k.kind = BLK_PART;
            c.k.tgt1 = c.k.aux0 = c.k.aux1 = 0;
            c.k.src[0] = c.k.src[1] = c.k.src[2] = c.k.src[3] = -1;
            c.tb = ch[k][1];
            c.te = ch[k][2];
            c.pin = prev;
            c.pout = (int)parts.size();
            parts.push_back({ch[k][0]});
            prev = c.pout;
            chunks.push_back({ch[k][0], c});
        }
        b.tb = ch[0][1];
        b.te = ch[0][2];
        b.pin = prev;
        ++s.ncut;
    };
    for (int l = 0; l < s.nlev; ++l) {
        for (Blk& b : levA[l]) cut(b, 2 * l);
        for (Blk& b : levB[l]) cut(b, 2 * l + 1);
    }
    for (auto& c : chunks) (c.first % 2 == 0 ? levA : levB)[c.first / 2].push_back(c.second);
    // partial slots: a part written in sub-level w is read in w + 1, so parts written in
    // sub-levels of one parity can share slots; slots [0, nE) for even, [nE, nE + nO) for odd
    std::vector<int> nsub(2 * (size_t)s.nlev + 1, 0), prank(parts.size());
    for (size_t q = 0; q < parts.size(); ++q) prank[q] = nsub[parts[q].sub]++;
    int nE = 0, nO = 0;
    for (size_t w = 0; w < nsub.size(); ++w) (w % 2 == 0 ? nE : nO) = std::max(w % 2 == 0 ? nE : nO, nsub[w]);
    s.nparts = nE + nO;
    auto part_pos = [&](int q) {
        const int slot = parts[q].sub % 2 == 0 ? prank[q] : nE + prank[q];
        return zero + 2 + 4 * slot;
    };

    auto weight = [&](const Blk& b) {  // row loads of the task
        shared(b);
        const int64_t nt = (b.te < 0 ? (int64_t)tl.size() : (int64_t)(b.te - b.tb)) +
                           (b.pin >= 0 ? (b.k.nc == 2 ? 2 : 1) : 0);
        return (int64_t)2 * b.k.nr * b.k.nc + nt * (b.k.nr + b.k.nc);
    };
    auto emit = [&](Blk b, std::vector<BlockTask>& out) {
        shared(b);
        b.k.sbeg = (int)(s.steps.size() / 4);
        if (b.pin >= 0) {  // the partial sums of the previous chunk, back exactly (see BLK_SPLIT_MIN)
            const int P = part_pos(b.pin);
            s.steps.push_back(P);
            s.steps.push_back(b.k.nr == 2 ? P + 2 : -1);
            s.steps.push_back(one);
            s.steps.push_back(b.k.nc == 2 ? zero : -1);
            if (b.k.nc == 2) {
                s.steps.push_back(P + 1);
                s.steps.push_back(b.k.nr == 2 ? P + 3 : -1);
                s.steps.push_back(zero);
                s.steps.push_back(one);
            }
        }
        const int i1 = b.te < 0 ? (int)tl.size() : b.te;
        for (int i = b.tb; i < i1; ++i) {
            const int tp = tl[i];
            s.steps.push_back(in_S(tp, b.r0) ? posL(b.r0, tp) : zero);
            s.steps.push_back(b.k.nr == 2 ? (in_S(tp, b.r0 + 1) ? posL(b.r0 + 1, tp) : zero) : -1);
            s.steps.push_back(in_S(tp, b.c0) ? posU(tp, b.c0) : zero);
            s.steps.push_back(b.k.nc == 2 ? (in_S(tp, b.c0 + 1) ? posU(tp, b.c0 + 1) : zero) : -1);
        }
        if (b.pout >= 0) b.k.tgt0 = part_pos(b.pout);
        b.k.send = (int)(s.steps.size() / 4);
        out.push_back(b.k);
    };
    // ---- forward and backward rows of a pivot group, cut like the blocks: a row's steps (the
    // columns j < p of its L rows ascending; the columns of S_last descending) in chunks at the
    // levels before the row's own, each where its y_j (forward: after the level of j's group)
    // or x_j (backward: after the depth level of j's group) are final
    struct RItem {
        int p, k0, k1;            // group lead, steps [k0, k1) of the row's column list
        int pin = -1, pout = -1;  // slot ids
    };
    std::vector<int> lastg(m);  // the last pivot of each pivot's group
    for (int p = 0; p < m; ++p) lastg[p] = gsize[p] == 2 ? p + 1 : p;
    std::vector<std::vector<int>> fcols(m), bcols(m);
    for (int p = 0; p < m; ++p) {
        if (gsize[p] == 0) continue;
        merge(rows[p], gsize[p] == 2 ? &rows[p + 1] : nullptr, ra);
        for (int j : ra)
            if (j < p) fcols[p].push_back(j);
        const int last = lastg[p];
        for (int q = t.Lp[last + 1] - 1; q >= t.Lp[last]; --q) bcols[p].push_back(t.Li[q]);
    }
    std::vector<std::vector<RItem>> levFi(s.nlev), levWi(s.nblev);
    for (int l = 0; l < s.nlev; ++l)
        for (int p : levF[l]) levFi[l].push_back(RItem{p, 0, (int)fcols[p].size()});
    for (int l = 0; l < s.nblev; ++l)
        for (int p : levW[l]) levWi[l].push_back(RItem{p, 0, (int)bcols[p].size()});
    std::vector<int> fslot_lev, bslot_lev;  // level that writes each slot (read at the next)
    // steps per level before any cut: forward rows with the blocks of their sub-level B
    std::vector<int64_t> ftot(s.nlev, 0), wtot(s.nblev, 0);
    for (int l = 0; l < s.nlev; ++l) {
        ftot[l] = subtot[2 * l + 1];
        for (const RItem& it : levFi[l]) ftot[l] += it.k1 - it.k0;
    }
    for (int l = 0; l < s.nblev; ++l)
        for (const RItem& it : levWi[l]) wtot[l] += it.k1 - it.k0;
    auto cut_row = [&](RItem& it, int l, const std::vector<int>& cols, const std::vector<int>& lev_of,
                       std::vector<int>& slot_lev, std::vector<std::pair<int, RItem>>& out,
                       const std::vector<int64_t>& tot) {
        const int len = it.k1 - it.k0;
        const bool fine = fine_cut(len, tot[l]);
        const int chunk = fine ? BLK_FINE_CHUNK : BLK_CHUNK, tail = fine ? BLK_FINE_TAIL : BLK_TAIL;
        if (len <= (fine ? BLK_FINE_MIN : BLK_SPLIT_MIN)) return false;
        std::vector<std::array<int, 3>> ch;  // (level, beg, end), the last piece first
        int end = len, lv = l;
        while (end > 0) {
            int beg = std::max(0, end - chunk);
            if (beg < tail) beg = 0;
            int need = 0;
            for (int k = beg; k < end; ++k) need = std::max(need, lev_of[lastg[cols[it.k0 + k]]] + 1);
            if (lv < 0 || need > lv) return false;  // cannot run this early: the row stays whole
            ch.push_back({lv, beg, end});
            end = beg;
            --lv;
        }
        if (ch.size() < 2) return false;
        int prev = -1;
        for (size_t k = ch.size(); k-- > 1;) {  // earlier chunks first
            RItem c{it.p, it.k0 + ch[k][1], it.k0 + ch[k][2]};
            c.pin = prev;
            c.pout = (int)slot_lev.size();
            slot_lev.push_back(ch[k][0]);
            prev = c.pout;
            out.push_back({ch[k][0], c});
        }
        it.k1 = it.k0 + ch[0][2];
        it.k0 = it.k0 + ch[0][1];
        it.pin = prev;
        return true;
    };
    {
        std::vector<std::pair<int, RItem>> fch, bch;
        for (int l = 0; l < s.nlev; ++l)
            for (RItem& it : levFi[l]) s.nfcut += cut_row(it, l, fcols[it.p], t.height, fslot_lev, fch, ftot);
        for (int l = 0; l < s.nblev; ++l)
            for (RItem& it : levWi[l]) s.nbcut += cut_row(it, l, bcols[it.p], t.depth, bslot_lev, bch, wtot);
        for (auto& c : fch) levFi[c.first].push_back(c.second);
        for (auto& c : bch) levWi[c.first].push_back(c.second);
    }
    // slots of two solution-vector rows each: written at level v and read at v + 1, so levels of
    // one parity share them; the forward slots are all read before the backward rows start,
    // which reuse the same rows
    auto xslots = [&](const std::vector<int>& lev_of, int nlevels, std::vector<int>& row) {
        std::vector<int> cnt((size_t)nlevels + 1, 0), rank(lev_of.size());
        for (size_t q = 0; q < lev_of.size(); ++q) rank[q] = cnt[lev_of[q]]++;
        int e = 0, o = 0;
        for (int v = 0; v <= nlevels; ++v) (v % 2 == 0 ? e : o) = std::max(v % 2 == 0 ? e : o, cnt[v]);
        row.resize(lev_of.size());
        for (size_t q = 0; q < lev_of.size(); ++q) row[q] = m + 2 * (lev_of[q] % 2 == 0 ? rank[q] : e + rank[q]);
        return 2 * (e + o);
    };
    std::vector<int> fslot, bslot;
    s.nxparts = std::max(xslots(fslot_lev, s.nlev, fslot), xslots(bslot_lev, s.nblev, bslot));

    std::vector<int> lA, sA, lB, sB, lF, sF, lW, sW;
    std::vector<const Blk*> itA, itB;
    std::vector<const RItem*> itF, itW;
    for (int l = 0; l < s.nlev; ++l) {
        std::vector<int64_t> load(nslots, 0), w;
        for (const Blk& b : levA[l]) w.push_back(weight(b));
        std::vector<int> slot = lpt_deal(w, nslots, load);
        for (size_t i = 0; i < levA[l].size(); ++i) { itA.push_back(&levA[l][i]); lA.push_back(l); sA.push_back(slot[i]); }
        std::vector<int64_t> loadB(nslots, 0), wb;
        for (const Blk& b : levB[l]) wb.push_back(weight(b));
        for (const RItem& it : levFi[l]) {
            const int nr = gsize[it.p] == 2 ? 2 : 1;
            wb.push_back(2 + (int64_t)(it.k1 - it.k0) * (nr + 1));
        }
        std::vector<int> slotB = lpt_deal(wb, nslots, loadB);
        for (size_t i = 0; i < levB[l].size(); ++i) { itB.push_back(&levB[l][i]); lB.push_back(l); sB.push_back(slotB[i]); }
        for (size_t i = 0; i < levFi[l].size(); ++i) {
            itF.push_back(&levFi[l][i]); lF.push_back(l); sF.push_back(slotB[levB[l].size() + i]);
        }
    }
    for (int l = 0; l < s.nblev; ++l) {
        std::vector<int64_t> load(nslots, 0), w;
        for (const RItem& it : levWi[l]) w.push_back(2 + (int64_t)(it.k1 - it.k0) * (gsize[it.p] == 2 ? 3 : 2));
        std::vector<int> slot = lpt_deal(w, nslots, load);
        for (size_t i = 0; i < levWi[l].size(); ++i) { itW.push_back(&levWi[l][i]); lW.push_back(l); sW.push_back(slot[i]); }
    }
    // execution order: per level, sub-level A of every list, then sub-level B; steps follow it
    std::vector<int> ordA = group_by_slot(lA, sA, s.nlev, nslots, s.ptrA);
    std::vector<int> ordB = group_by_slot(lB, sB, s.nlev, nslots, s.ptrB);
    for (int l = 0; l < s.nlev; ++l) {
        for (int q = s.ptrA[(size_t)l * nslots]; q < s.ptrA[(size_t)(l + 1) * nslots]; ++q) emit(*itA[ordA[q]], s.taskA);
        for (int q = s.ptrB[(size_t)l * nslots]; q < s.ptrB[(size_t)(l + 1) * nslots]; ++q) emit(*itB[ordB[q]], s.taskB);
    }
    for (int i : group_by_slot(lF, sF, s.nlev, nslots, s.ptrF)) {
        const RItem& it = *itF[i];
        const int p = it.p;
        const int nr = gsize[p] == 2 ? 2 : 1;
        FwdBlock f{};
        f.row = p;
        f.nr = nr;
        f.rsrc0 = t.perm[p];
        f.rsrc1 = nr == 2 ? t.perm[p + 1] : -1;
        f.lpos = nr == 2 ? posL(p + 1, p) : -1;
        f.pin = it.pin >= 0 ? fslot[it.pin] : -1;
        f.pout = it.pout >= 0 ? fslot[it.pout] : -1;
        f.sbeg = (int)(s.fsteps.size() / 3);
        for (int k = it.k0; k < it.k1; ++k) {  // columns j < p of the group's L rows, ascending
            const int j = fcols[p][k];
            s.fsteps.push_back(in_S(j, p) ? posL(p, j) : zero);
            s.fsteps.push_back(nr == 2 ? (in_S(j, p + 1) ? posL(p + 1, j) : zero) : -1);
            s.fsteps.push_back(j);
        }
        f.send = (int)(s.fsteps.size() / 3);
        s.fwd.push_back(f);
    }
    for (int i : group_by_slot(lW, sW, s.nblev, nslots, s.ptrW)) {
        const RItem& it = *itW[i];
        const int p = it.p;
        const int nr = gsize[p] == 2 ? 2 : 1;
        BwdBlock b{};
        b.row = p;
        b.nr = nr;
        b.base0 = t.base[p];
        b.base1 = nr == 2 ? t.base[p + 1] : -1;
        b.pin = it.pin >= 0 ? bslot[it.pin] : -1;
        b.pout = it.pout >= 0 ? bslot[it.pout] : -1;
        b.sbeg = (int)(s.bsteps.size() / 3);
        for (int k = it.k0; k < it.k1; ++k) {  // columns of S_last, descending
            const int j = bcols[p][k];
            s.bsteps.push_back(posU(p, j));
            s.bsteps.push_back(nr == 2 ? posU(p + 1, j) : -1);
            s.bsteps.push_back(j);
        }
        b.send = (int)(s.bsteps.size() / 3);
        s.bwd.push_back(b);
    }
    return s;
}

// ------------------------------------------------------------------ host numerics
// Jacobian and residual of one system at (Vm, Va), rectangular form: with the unit phasor
// (cos Va, sin Va) of every bus, Ym cos(delta) = G cu + B su and Ym sin(delta) = G su - B cu
// where cu = cos(Va_i - Va_k), su = sin(Va_i - Va_k) -- the same quantities as
// p3s/cpp/nr_klu.cpp:344-378 without a sincos per nonzero.
void host_eval(const Topology& t, const double* Vm, const double* Va, const double* G, const double* B,
               const double* Ssp, const unsigned char* pin, double* R, double* J) {
    if (!G) G = t.G.data();
    if (!B) B = t.B.data();
    std::vector<double> cs(t.n), sn(t.n);
    for (int i = 0; i < t.n; ++i) { cs[i] = std::cos(Va[i]); sn[i] = std::sin(Va[i]); }
    for (int i = 0; i < t.n; ++i) {
        const double vi = Vm[i], ci = cs[i], si = sn[i];
        const bool pinned = pin && pin[i];
        double p = 0, q = 0, dPVa = 0, dQVa = 0, dPVm = 0, dQVm = 0;
        for (int k = t.Yp[i]; k < t.Yp[i + 1]; ++k) {
            const int j = t.Yj[k];
            const double g = G[k], b = B[k];
            if (j != i) {
                const double vk = Vm[j];
                const double cu = ci * cs[j] + si * sn[j];
                const double su = si * cs[j] - ci * sn[j];
                const double gc = g * cu + b * su, gs = g * su - b * cu;
                const double a = vi * gc * vk, bt = vi * gs * vk;
                p += a; q += bt;
                dPVa -= bt; dQVa += a; dPVm += vk * gc; dQVm += vk * gs;
                if (J) {
                    const double vals[4] = {bt, vi * gc, -a, vi * gs};
                    for (int blk = 0; blk < 4; ++blk) {
                        int e = t.posJ[4 * (size_t)k + blk];
                        if (e >= 0) J[e] = pinned ? 0.0 : vals[blk];
                    }
                }
            } else {
                p += vi * vi * g; q -= vi * vi * b;
                dPVm += 2.0 * vi * g; dQVm -= 2.0 * vi * b;
            }
        }
        const int d = t.Ydiag[i];
        if (J && d >= 0) {
            const double vals[4] = {pinned ? 1.0 : dPVa, pinned ? 0.0 : dPVm, pinned ? 0.0 : dQVa, pinned ? 1.0 : dQVm};
            for (int blk = 0; blk < 4; ++blk) {
                int e = t.posJ[4 * (size_t)d + blk];
                if (e >= 0) J[e] = vals[blk];
            }
        }
        if (t.row_p[i] >= 0) R[t.row_p[i]] = pinned ? 0.0 : Ssp[2 * i] - p;
        if (t.row_q[i] >= 0) R[t.row_q[i]] = pinned ? 0.0 : Ssp[2 * i + 1] - q;
    }
}

void host_factor(const Topology& t, const double* J, double* fac, double* pmin, double* pmax, double* psum) {
    const Schedule& s = t.sched;
    double lo = std::numeric_limits<double>::infinity(), hi = 0.0, sum = 0.0;
    // every entry's whole sum, at the level of its task (the chunks of a cut sum are skipped:
    // the last one sits where the whole entry would)
    auto run = [&](const Task& tk) {
        double acc = tk.asrc >= 0 ? J[tk.asrc] : 0.0;
        for (int e = t.prange[tk.dst]; e < t.prange[tk.dst + 1]; ++e) acc -= fac[t.pa[e]] * fac[t.pb[e]];
        return acc;
    };
    for (int l = 0; l < s.nlev; ++l) {
        for (int q = s.ptrA[(size_t)l * s.nslots]; q < s.ptrA[(size_t)(l + 1) * s.nslots]; ++q) {
            const Task& tk = s.taskA[q];
            if (tk.part) continue;
            double acc = run(tk);
            if (s.isdiagA[q]) {
                const double a = std::fabs(acc);
                lo = std::min(lo, a); hi = std::max(hi, a); sum += a;
                acc = 1.0 / acc;
            }
            fac[tk.dst] = acc;
        }
        for (int q = s.ptrB[(size_t)l * s.nslots]; q < s.ptrB[(size_t)(l + 1) * s.nslots]; ++q) {
            const Task& tk = s.taskB[q];
            if (tk.part) continue;
            fac[tk.dst] = run(tk) * fac[s.diagB[q]];
        }
    }
    *pmin = lo; *pmax = hi; *psum = sum;
}

void host_solve(const Topology& t, const double* fac, const double* R, double* x) {
    // every row's whole sum at the level of its last chunk
    const Schedule& s = t.sched;
    for (int l = 0; l < s.nlev; ++l)
        for (int q = s.ptrF[(size_t)l * s.nslots]; q < s.ptrF[(size_t)(l + 1) * s.nslots]; ++q) {
            const FwdRow& f = s.fwd[q];
            if (f.dst >= 0) continue;
            double acc = R[f.rsrc];
            for (int e = t.Fp[f.row]; e < t.Fp[f.row + 1]; ++e) acc -= fac[t.Fpos[e]] * x[t.Fcol[e]];
            x[f.row] = acc;
        }
    for (int l = 0; l < s.nblev; ++l)
        for (int q = s.ptrW[(size_t)l * s.nslots]; q < s.ptrW[(size_t)(l + 1) * s.nslots]; ++q) {
            const BwdRow& b = s.bwd[q];
            if (b.dst >= 0) continue;
            double acc = x[b.row];
            const int l0 = t.Lp[b.row];
            for (int e = t.Lp[b.row + 1] - 1; e >= l0; --e) acc -= fac[b.base + 1 + (e - l0)] * x[t.Li[e]];
            x[b.row] = acc * fac[b.base];
        }
}

// ------------------------------------------------------------------ LU emulation
namespace {

template <typename real>
struct Emu {
    const Topology& t;
    std::vector<real> fac;
    std::vector<char> written;
    std::vector<double> x;
    std::vector<char> solved;
    std::vector<int> wsub;  // partial slots: sub-level of the last write
    int sub = 0;            // the sub-level running (2 level + 0 for A, 1 for B)
    int64_t bad = 0;
    double lo = std::numeric_limits<double>::max(), hi = 0.0;
    std::vector<int> xw;    // solution-vector slots: level of the last write (-1 none)
    int xlev = 0;           // the substitution level running
    explicit Emu(const Topology& tt)
        : t(tt), fac(std::max<int64_t>(tt.fac_rows, tt.nnz_lu + 2), (real)0), written(fac.size(), 0),
          x(std::max<int64_t>(tt.x_rows, tt.m), 0.0), solved(x.size(), 0), wsub(fac.size(), -1),
          xw(x.size(), -1) {
        written[t.nnz_lu] = 1;  // the zero row is readable
        written[t.nnz_lu + 1] = 1;  // and the one row
        fac[t.nnz_lu + 1] = (real)1;
    }
    bool partial(int64_t pos) const { return pos >= t.nnz_lu + 2 && pos < (int64_t)fac.size(); }
    real rd(int pos) {
        if (pos < 0 || pos >= (int64_t)fac.size() || !written[pos]) { ++bad; return (real)0; }
        if (partial(pos) && wsub[pos] >= sub) ++bad;  // a partial sum read in the sub-level that writes it
        return fac[pos];
    }
    void wr(int pos, real v) {
        if (partial(pos)) {  // partial slots are reused once read
            if (wsub[pos] >= sub) { ++bad; return; }
            written[pos] = 1;
            wsub[pos] = sub;
            fac[pos] = v;
            return;
        }
        if (pos < 0 || pos >= t.nnz_lu || written[pos]) { ++bad; return; }
        written[pos] = 1;
        fac[pos] = v;
    }
    real pivot(real v) {
        const double ab = std::fabs((double)v);
        lo = std::min(lo, ab);
        hi = std::max(hi, ab);
        return (real)1 / v;
    }
    double xv(int i) {
        if (!solved[i]) ++bad;
        return x[i];
    }
    double xslot(int i) {  // a cut row's partial sum, written at an earlier level
        if (i < t.m || i >= (int)x.size() || xw[i] < 0 || xw[i] >= xlev) { ++bad; return 0.0; }
        return x[i];
    }
    void set_slot(int i, double v) {
        if (i < t.m || i >= (int)x.size() || xw[i] >= xlev) { ++bad; return; }
        xw[i] = xlev;
        x[i] = v;
    }
    template <typename F>
    void lists(const std::vector<int>& ptr, int l, int ns, F f) {  // lists of one sub-level, last first
        for (int sl = ns - 1; sl >= 0; --sl)
            for (int q = ptr[(size_t)l * ns + sl]; q < ptr[(size_t)l * ns + sl + 1]; ++q) f(q);
    }
};

template <typename real>
void emulate_one(const Topology& t, int ns, const double* J, const double* R, Emu<real>& E, int64_t* checks) {
    const Schedule s = make_schedule(t, ns, 0);
    const int one = (int)t.nnz_lu + 1;
    auto dot = [&](const Task& k) {
        real acc = 0;
        if (k.pin >= 0) acc = std::fma(-E.rd(k.pin), E.rd(one), acc);  // the previous chunk's sum
        for (int e = k.pbeg; e < k.pend; ++e) acc = std::fma(-E.rd(t.pa[e]), E.rd(t.pb[e]), acc);
        return acc;
    };
    for (int l = 0; l < s.nlev; ++l) {
        E.sub = 2 * l;
        E.lists(s.ptrA, l, ns, [&](int q) {
            const Task& k = s.taskA[q];
            if (k.part) { E.wr(k.dst, -dot(k)); return; }
            real v = (real)(k.asrc >= 0 ? J[k.asrc] : 0.0) + dot(k);
            E.wr(k.dst, s.isdiagA[q] ? E.pivot(v) : v);
        });
        E.sub = 2 * l + 1;
        E.lists(s.ptrB, l, ns, [&](int q) {
            const Task& k = s.taskB[q];
            if (k.part) { E.wr(k.dst, -dot(k)); return; }
            E.wr(k.dst, ((real)(k.asrc >= 0 ? J[k.asrc] : 0.0) + dot(k)) * (real)(double)E.rd(s.diagB[q]));
        });
    }
    for (int l = 0; l < s.nlev; ++l) {
        E.xlev = l;
        E.lists(s.ptrF, l, ns, [&](int q) {
            const FwdRow& f = s.fwd[q];
            double acc = 0.0;
            if (f.pin >= 0) acc = std::fma(-(double)E.rd(one), E.xslot(f.pin), acc);
            for (int e = f.ebeg; e < f.eend; ++e) acc = std::fma(-(double)E.rd(t.Fpos[e]), E.xv(t.Fcol[e]), acc);
            if (f.dst >= 0) { E.set_slot(f.dst, -acc); return; }
            E.x[f.row] = R[f.rsrc] + acc;
            E.solved[f.row] = 1;
        });
    }
    std::fill(E.solved.begin(), E.solved.end(), 0);
    std::fill(E.xw.begin(), E.xw.end(), -1);
    for (int l = 0; l < s.nblev; ++l) {
        E.xlev = l;
        E.lists(s.ptrW, l, ns, [&](int q) {
            const BwdRow& b = s.bwd[q];
            double acc = 0.0;
            if (b.pin >= 0) acc = std::fma(-(double)E.rd(one), E.xslot(b.pin), acc);
            const int l0 = t.Lp[b.row];
            for (int e = b.lend - 1; e >= b.lbeg; --e)
                acc = std::fma(-(double)E.rd(b.base + 1 + (e - l0)), E.xv(t.Li[e]), acc);
            if (b.dst >= 0) { E.set_slot(b.dst, -acc); return; }
            E.x[b.row] = (E.x[b.row] + acc) * (double)E.rd(b.base);
            E.solved[b.row] = 1;
        });
    }
    checks[1] = 0;
    checks[2] = t.n_pairs;
}

template <typename real>
void emulate_blocks(const Topology& t, int ns, const double* J, const double* R, Emu<real>& E, int64_t* checks) {
    const BlockSchedule s = make_block_schedule(t, ns);
    auto run = [&](const BlockTask& k) {
        real acc[4] = {0, 0, 0, 0};
        for (int e = k.sbeg; e < k.send; ++e) {
            const int* st = &s.steps[4 * (size_t)e];
            if ((st[1] < 0) != (k.nr == 1) || (st[3] < 0) != (k.nc == 1)) ++E.bad;
            const real a0 = E.rd(st[0]), a1 = st[1] >= 0 ? E.rd(st[1]) : (real)0;
            const real b0 = E.rd(st[2]), b1 = st[3] >= 0 ? E.rd(st[3]) : (real)0;
            acc[0] = std::fma(-a0, b0, acc[0]);
            acc[1] = std::fma(-a0, b1, acc[1]);
            acc[2] = std::fma(-a1, b0, acc[2]);
            acc[3] = std::fma(-a1, b1, acc[3]);
        }
        real j[4];
        for (int i = 0; i < 4; ++i) j[i] = (real)(k.src[i] >= 0 ? J[k.src[i]] : 0.0);
        switch (k.kind) {
            case BLK_DIAG1: E.wr(k.tgt0, E.pivot(j[0] + acc[0])); break;
            case BLK_DIAG2: {
                const real inv0 = E.pivot(j[0] + acc[0]), u01 = j[1] + acc[1], l10 = (j[2] + acc[2]) * inv0;
                E.wr(k.tgt0, inv0);
                E.wr(k.aux0, u01);
                E.wr(k.aux1, l10);
                E.wr(k.tgt1, E.pivot(j[3] + std::fma(-l10, u01, acc[3])));
                break;
            }
            case BLK_U1:
                E.wr(k.tgt0, j[0] + acc[0]);
                if (k.nc == 2) E.wr(k.tgt0 + 1, j[1] + acc[1]);
                break;
            case BLK_U2: {
                const real l10 = E.rd(k.aux0), u00 = j[0] + acc[0];
                E.wr(k.tgt0, u00);
                E.wr(k.tgt1, j[2] + std::fma(-l10, u00, acc[2]));
                if (k.nc == 2) {
                    const real u01 = j[1] + acc[1];
                    E.wr(k.tgt0 + 1, u01);
                    E.wr(k.tgt1 + 1, j[3] + std::fma(-l10, u01, acc[3]));
                }
                break;
            }
            case BLK_L1: {
                const real inv = E.rd(k.aux0);
                E.wr(k.tgt0, (j[0] + acc[0]) * inv);
                if (k.nr == 2) E.wr(k.tgt0 + 1, (j[2] + acc[2]) * inv);
                break;
            }
            case BLK_PART:  // the chunk's sums, negated, for the next chunk of the block
                E.wr(k.tgt0, -acc[0]);
                if (k.nc == 2) E.wr(k.tgt0 + 1, -acc[1]);
                if (k.nr == 2) E.wr(k.tgt0 + 2, -acc[2]);
                if (k.nr == 2 && k.nc == 2) E.wr(k.tgt0 + 3, -acc[3]);
                break;
            default: {
                const real inv0 = E.rd(k.aux0), u01 = E.rd(k.aux0 + 1), inv1 = E.rd(k.aux1);
                const real l00 = (j[0] + acc[0]) * inv0;
                E.wr(k.tgt0, l00);
                E.wr(k.tgt1, (j[1] + std::fma(-l00, u01, acc[1])) * inv1);
                if (k.nr == 2) {
                    const real l10 = (j[2] + acc[2]) * inv0;
                    E.wr(k.tgt0 + 1, l10);
                    E.wr(k.tgt1 + 1, (j[3] + std::fma(-l10, u01, acc[3])) * inv1);
                }
            }
        }
    };
    for (int l = 0; l < s.nlev; ++l) {
        E.sub = 2 * l;
        E.lists(s.ptrA, l, ns, [&](int q) { run(s.taskA[q]); });
        E.sub = 2 * l + 1;
        E.lists(s.ptrB, l, ns, [&](int q) { run(s.taskB[q]); });
    }
    for (int l = 0; l < s.nlev; ++l) {
        E.xlev = l;
        E.lists(s.ptrF, l, ns, [&](int q) {
            const FwdBlock& f = s.fwd[q];
            double a0 = 0.0, a1 = 0.0;
            if (f.pin >= 0) {  // the previous chunk's sums
                a0 = E.xslot(f.pin);
                if (f.nr == 2) a1 = E.xslot(f.pin + 1);
            }
            for (int e = f.sbeg; e < f.send; ++e) {
                const int* st = &s.fsteps[3 * (size_t)e];
                if ((st[1] < 0) != (f.nr == 1)) ++E.bad;
                const double y = E.xv(st[2]);
                a0 = std::fma(-(double)E.rd(st[0]), y, a0);
                if (st[1] >= 0) a1 = std::fma(-(double)E.rd(st[1]), y, a1);
            }
            if (f.pout >= 0) {  // a chunk: its sums for the next piece
                E.set_slot(f.pout, a0);
                if (f.nr == 2) E.set_slot(f.pout + 1, a1);
                return;
            }
            const double y0 = R[f.rsrc0] + a0;
            E.x[f.row] = y0;
            E.solved[f.row] = 1;
            if (f.nr == 2) {
                E.x[f.row + 1] = R[f.rsrc1] + std::fma(-(double)E.rd(f.lpos), y0, a1);
                E.solved[f.row + 1] = 1;
            }
        });
    }
    std::fill(E.solved.begin(), E.solved.end(), 0);
    std::fill(E.xw.begin(), E.xw.end(), -1);
    for (int l = 0; l < s.nblev; ++l) {
        E.xlev = l;
        E.lists(s.ptrW, l, ns, [&](int q) {
            const BwdBlock& b = s.bwd[q];
            double a0 = 0.0, a1 = 0.0;
            if (b.pin >= 0) {
                a0 = E.xslot(b.pin);
                if (b.nr == 2) a1 = E.xslot(b.pin + 1);
            }
            for (int e = b.sbeg; e < b.send; ++e) {
                const int* st = &s.bsteps[3 * (size_t)e];
                if ((st[1] < 0) != (b.nr == 1)) ++E.bad;
                const double xj = E.xv(st[2]);
                a0 = std::fma(-(double)E.rd(st[0]), xj, a0);
                if (st[1] >= 0) a1 = std::fma(-(double)E.rd(st[1]), xj, a1);
            }
            if (b.pout >= 0) {
                E.set_slot(b.pout, a0);
                if (b.nr == 2) E.set_slot(b.pout + 1, a1);
                return;
            }
            if (b.nr == 2) {
                const double x1 = (E.x[b.row + 1] + a1) * (double)E.rd(b.base1);
                E.x[b.row + 1] = x1;
                E.solved[b.row + 1] = 1;
                E.x[b.row] = (E.x[b.row] + std::fma(-(double)E.rd(b.base0 + 1), x1, a0)) * (double)E.rd(b.base0);
            } else {
                E.x[b.row] = (E.x[b.row] + a0) * (double)E.rd(b.base0);
            }
            E.solved[b.row] = 1;
        });
    }
    checks[1] = s.npairs;
    checks[2] = (int64_t)(s.steps.size() / 4);
}

template <typename real>
void emulate(const Topology& t, bool blocks, int ns, const double* J, const double* R, double* x, double* pmin,
             double* pmax, int64_t* checks) {
    Emu<real> E(t);
    if (blocks) emulate_blocks<real>(t, ns, J, R, E, checks);
    else emulate_one<real>(t, ns, J, R, E, checks);
    for (int64_t i = 0; i < t.nnz_lu; ++i) E.bad += !E.written[i];
    if (E.fac[t.nnz_lu] != (real)0 || E.fac[t.nnz_lu + 1] != (real)1) ++E.bad;
    for (int i = 0; i < t.m; ++i) E.bad += !E.solved[i];
    std::copy(E.x.begin(), E.x.begin() + t.m, x);
    *pmin = E.lo;
    *pmax = E.hi;
    checks[0] = E.bad;
}

}  // namespace

void host_lu_emulate(const Topology& t, bool blocks, int nslots, bool single, const double* J, const double* R,
                     double* x, double* pmin, double* pmax, int64_t* checks) {
    if (nslots < 1) throw std::runtime_error("nslots must be >= 1");
    if (single) emulate<float>(t, blocks, nslots, J, R, x, pmin, pmax, checks);
    else emulate<double>(t, blocks, nslots, J, R, x, pmin, pmax, checks);
}

// ------------------------------------------------------------------ entry points
// min/max |pivot| of the static-pivot LU of the Jacobian at flat start, or 0 when a pivot is zero
// or non-finite (the frozen order is then not used: p3s/cpp/nr_klu.cpp:718-729 leaves the lean plan
// invalid and every solve factors with pivoting)
static double reference_pivot_ratio(const Topology& t) {
    std::vector<double> Vm(t.n, 1.0), Va(t.n, 0.0), S(2 * (size_t)t.n, 0.0), R(t.m), J(t.nnzJ), fac(t.nnz_lu);
    host_eval(t, Vm.data(), Va.data(), nullptr, nullptr, S.data(), nullptr, R.data(), J.data());
    double lo, hi, sum;
    host_factor(t, J.data(), fac.data(), &lo, &hi, &sum);
    if (!std::isfinite(sum) || !(lo > 0.0) || !(hi > 0.0)) return 0.0;
    return lo / hi;
}

void build_topology(Topology& t, const int* Yp, const int* Yj, const double* Yx, int n, int nnzY, const int* pv,
                    int npv, const int* pq, int npq, int ordering, int nslots) {
    build_polar(t, Yp, Yj, Yx, n, nnzY, pv, npv, pq, npq);
    if (t.m == 0) throw std::runtime_error("no pv/pq buses: nothing to solve");
    build_jacobian_pattern(t);
    build_lu(t, ordering);
    build_schedule(t, nslots);
    // the factor arena: entries, the zero and one rows, the block schedule's partial slots (the
    // cuts and slots do not depend on the number of task lists)
    const Schedule one = make_schedule(t, 1, 0);
    const BlockSchedule blk = make_block_schedule(t, 1);
    t.fac_rows = t.nnz_lu + 2 + std::max<int64_t>(4 * (int64_t)blk.nparts, one.nparts);
    t.x_rows = (int64_t)t.m + std::max(one.nxparts, blk.nxparts);
    t.rcond0 = reference_pivot_ratio(t);
    t.lean_valid = t.rcond0 > 0.0;
}

void update_values(Topology& t, const double* Yx) {
    for (int k = 0; k < t.nnzY; ++k) { t.G[k] = Yx[2 * k]; t.B[k] = Yx[2 * k + 1]; }
    t.rcond0 = reference_pivot_ratio(t);
    t.lean_valid = t.rcond0 > 0.0;
}

}  // namespace p3smi
