// fg_reorder.hip -- reverse Cuthill-McKee numbering of a CSR pattern (host).
//
// Reference: 00_CommonFiles/Reordering/symrcm.c:4-242 (REORDERING_SYMRCM, an
// array-heap Cuthill-McKee), called from csr_Initialization.c when the
// parameter file asks for SYMRCM.  The permutation is integer work done once
// per mesh, so it runs on the host; it follows the reference's tie-breaking
// exactly (pseudo-peripheral start node by repeated BFS, neighbours of a node
// visited in ascending order and queued by ascending degree through a binary
// min-heap with the reference's sift rules, components taken in order of
// their lowest unvisited equation) so that the device solves a system
// numbered exactly as the reference numbers it.
//
// Host-parallel form (fg_host.h): the transposed pattern and the degrees are
// built row by row on all workers; every breadth-first search runs level by
// level.  A level's nodes claim their unvisited neighbours with an atomic
// minimum of their queue positions, so each new node goes to the first node
// in queue order that reaches it, as in the sequential FIFO search; then
// every node writes the neighbours it won -- ascending, or in the order the
// degree heap returns them for the Cuthill-McKee pass -- at its offset in
// the next level.  The queue order, and so the permutation, is the
// sequential algorithm's whatever the thread count.
#include <limits.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fg_host.h"
#include "fg_internal.h"

namespace {

struct Node {
    int id, deg, dist;
};

class Graph {
public:
    Graph(FgPool &pool, int n, const int *ia, const int *ja) : ia_(ia), ja_(ja), ta_(n + 1, 0)
    {
        // transposed pattern (columns of A as rows), rows ascending.  Worker
        // t owns the columns [n t / T, n (t+1) / T): it streams the whole
        // pattern and counts, then places, only its own columns, rows in
        // ascending order -- no atomics, no sort.
        const int nz = ia[n];
        tj_.resize(nz > 0 ? nz : 1);
        const int T = pool.size();
        auto owned = [&](int t, int *c0, int *c1) {
            *c0 = (int)((long long)n * t / T);
            *c1 = (int)((long long)n * (t + 1) / T);
        };
        pool.run([&](int t) {
            int c0, c1;
            owned(t, &c0, &c1);
            for (int q = 0; q < nz; q++)
                if (ja[q] >= c0 && ja[q] < c1) ta_[ja[q]]++;
        });
        ta_[n] = 0;
        fg_exclusive_scan(ta_);
        ta_[n] = nz;
        {
            std::vector<int> cur(ta_.begin(), ta_.end() - 1);
            pool.run([&](int t) {
                int c0, c1;
                owned(t, &c0, &c1);
                for (int j = 0; j < n; j++)
                    for (int k = ia[j]; k < ia[j + 1]; k++)
                        if (ja[k] >= c0 && ja[k] < c1) tj_[cur[ja[k]]++] = j;
            });
        }
        // degree of j: the entries of column j, plus one for each entry (j,k),
        // k != j, whose mirror (k,j) the scan of row k (ascending, stopping
        // past j) does not find (symmetrised pattern).  With every row
        // strictly ascending that scan finds (k,j) exactly when k is in
        // column j, so the count is a merge of row j with column j (no
        // access to other rows); otherwise the scans are done as written.
        std::vector<char> wsorted(pool.size(), 1);
        pool.range(n, [&](long long b, long long e, int t) {
            for (long long j = b; j < e && wsorted[t]; j++)
                for (int q = ia[j] + 1; q < ia[j + 1]; q++)
                    if (ja[q] <= ja[q - 1]) { wsorted[t] = 0; break; }
        });
        const bool sorted = std::all_of(wsorted.begin(), wsorted.end(), [](char c) { return c; });
        deg_.assign(n, 0);
        std::vector<int> wmax(pool.size(), 0);
        pool.range(n, [&](long long b, long long e, int t) {
            int m = 0;
            for (long long j = b; j < e; j++) {
                int d = ta_[j + 1] - ta_[j];
                if (sorted) {
                    int c = ta_[j];
                    const int ce = ta_[j + 1];
                    for (int q = ia[j]; q < ia[j + 1]; q++) {
                        const int k = ja[q];
                        if (k == j) continue;
                        while (c < ce && tj_[c] < k) c++;
                        d += !(c < ce && tj_[c] == k);
                    }
                } else {
                    for (int q = ia[j]; q < ia[j + 1]; q++) {
                        const int k = ja[q];
                        if (k == j) continue;
                        bool found = false;
                        for (int l = ia[k]; l < ia[k + 1]; l++) {
                            if (ja[l] == j) { found = true; break; }
                            if (ja[l] > j) break;
                        }
                        d += !found;
                    }
                }
                deg_[j] = d;
                m = std::max(m, d);
            }
            wmax[t] = std::max(wmax[t], m);
        });
        maxdeg_ = *std::max_element(wmax.begin(), wmax.end());
        // the neighbour lists the searches walk, stored once: the merge of
        // row i and column i, each index at its first occurrence (a search
        // ignores later ones, the node being seen by then).  Strictly
        // ascending rows give ascending lists without repeats.
        aoff_.assign((size_t)n + 1, 0);
        auto list = [&](int i, std::vector<int> &out) {
            out.clear();
            merge(i, [&](int r) {
                if (sorted || std::find(out.begin(), out.end(), r) == out.end()) out.push_back(r);
            });
        };
        pool.range(n, [&](long long b, long long e, int) {
            std::vector<int> l;
            for (long long i = b; i < e; i++) {
                list((int)i, l);
                aoff_[i] = (long long)l.size();
            }
        });
        aoff_[n] = 0;
        const long long na = fg_exclusive_scan(aoff_);
        aoff_[n] = na;
        adj_.resize(na > 0 ? na : 1);
        pool.range(n, [&](long long b, long long e, int) {
            std::vector<int> l;
            for (long long i = b; i < e; i++) {
                list((int)i, l);
                std::copy(l.begin(), l.end(), adj_.begin() + aoff_[i]);
            }
        });
        std::vector<int>().swap(tj_);
    }

    int maxdeg() const { return maxdeg_; }
    int deg(int i) const { return deg_[i]; }

    // neighbours of i, ascending, each index once
    template <typename F>
    void each_neighbour(int i, F &&f) const
    {
        for (long long q = aoff_[i]; q < aoff_[i + 1]; q++) f(adj_[q]);
    }

private:
    template <typename F>
    void merge(int i, F &&f) const
    {
        int a = ia_[i], b = ta_[i];
        const int ae = ia_[i + 1], be = ta_[i + 1];
        while (a < ae || b < be) {
            if (a == ae) f(tj_[b++]);
            else if (b == be) f(ja_[a++]);
            else if (ja_[a] <= tj_[b]) {
                const int r = ja_[a++];
                if (tj_[b] == r) b++;
                f(r);
            } else
                f(tj_[b++]);
        }
    }

    const int *ia_, *ja_;
    std::vector<int> ta_, tj_, deg_, adj_;
    std::vector<long long> aoff_;
    int maxdeg_ = 0;
};

// binary min-heap on degree with strict comparisons
struct DegHeap {
    std::vector<Node> h;
    int size = 0;
    explicit DegHeap(int cap) : h(cap > 0 ? cap : 1) {}
    void push(const Node &x)
    {
        int i = size++;
        h[i] = x;
        while (i > 0) {
            const int p = (i - 1) >> 1;
            if (!(h[i].deg < h[p].deg)) break;
            std::swap(h[i], h[p]);
            i = p;
        }
    }
    Node pop()
    {
        Node r = h[0];
        h[0] = h[--size];
        for (int j = 0;;) {
            const int l = 2 * j + 1, rr = 2 * j + 2;
            int s = (l < size && h[l].deg < h[j].deg) ? l : j;
            if (rr < size && h[rr].deg < h[s].deg) s = rr;
            if (s == j) break;
            std::swap(h[j], h[s]);
            j = s;
        }
        return r;
    }
};

// Breadth-first search from `root` over the nodes not yet seen.  Appends
// the queue order to order[pos..] and returns the new end; marks the nodes
// it queues seen; levels receives the level boundaries (positions) when
// non-null.  claim[] must be INT_MAX on every unseen node (it is left so for
// the nodes the search does not reach).  heap: each node's new neighbours
// enter the queue in the order the degree heap returns them (Cuthill-McKee)
// instead of ascending.
int bfs(FgPool &pool, const Graph &g, int root, int pos, bool heap, std::vector<int> &order,
        std::vector<char> &seen, std::vector<int> &claim, std::vector<int> *levels)
{
    // levels narrower than this run on the calling thread
    constexpr long long kSerial = 1024;
    order[pos] = root;
    seen[root] = 1;
    int a = pos, b = pos + 1;
    if (levels) levels->assign(1, a);
    std::vector<int> cnt;
    std::vector<DegHeap> heaps;
    if (heap) heaps.assign(pool.size(), DegHeap(g.maxdeg()));
    while (a < b) {
        if (levels) levels->push_back(b);
        const long long w = b - a;
        // an unseen neighbour goes to the level node of smallest queue position
        pool.range(w, [&](long long p0, long long p1, int) {
            for (long long p = p0; p < p1; p++) {
                const int me = a + (int)p;
                g.each_neighbour(order[me], [&](int r) {
                    if (seen[r]) return;
                    int cur = __atomic_load_n(&claim[r], __ATOMIC_RELAXED);
                    while (me < cur && !__atomic_compare_exchange_n(&claim[r], &cur, me, true,
                                                                    __ATOMIC_RELAXED,
                                                                    __ATOMIC_RELAXED)) {
                    }
                });
            }
        }, kSerial);
        cnt.assign(w + 1, 0);
        pool.range(w, [&](long long p0, long long p1, int) {
            for (long long p = p0; p < p1; p++) {
                const int me = a + (int)p;
                int c = 0;
                g.each_neighbour(order[me], [&](int r) { c += !seen[r] && claim[r] == me; });
                cnt[p] = c;
            }
        }, kSerial);
        cnt[w] = 0;
        const int nw = (int)fg_exclusive_scan(cnt);
        pool.range(w, [&](long long p0, long long p1, int t) {
            for (long long p = p0; p < p1; p++) {
                const int me = a + (int)p;
                int o = b + cnt[p];
                if (heap) {
                    DegHeap &hp = heaps[t];
                    g.each_neighbour(order[me], [&](int r) {
                        if (!seen[r] && claim[r] == me) hp.push(Node{r, g.deg(r), 0});
                    });
                    while (hp.size > 0) order[o++] = hp.pop().id;
                } else {
                    g.each_neighbour(order[me], [&](int r) {
                        if (!seen[r] && claim[r] == me) order[o++] = r;
                    });
                }
            }
        }, kSerial);
        pool.range(nw, [&](long long q0, long long q1, int) {
            for (long long q = q0; q < q1; q++) seen[order[b + q]] = 1;
        }, kSerial);
        a = b;
        b += nw;
    }
    return b;
}

// pseudo-peripheral node: search from `start`, then take the node the
// sequential scan of the queue keeps (a node replaces the kept one when it
// lies farther out, or is another node of larger degree); search again from
// it while the eccentricity grows
int start_node(FgPool &pool, const Graph &g, int start, std::vector<int> &order,
               std::vector<char> &seen, std::vector<int> &claim)
{
    int best = 0;
    Node x{start, g.deg(start), 0};
    std::vector<int> levels;
    for (;;) {
        std::fill(seen.begin(), seen.end(), 0);
        std::fill(claim.begin(), claim.end(), INT_MAX);
        bfs(pool, g, x.id, 0, false, order, seen, claim, &levels);
        Node y{x.id, x.deg, 0};
        for (size_t L = 0; L + 1 < levels.size(); L++)
            for (int q = levels[L]; q < levels[L + 1]; q++) {
                const Node v{order[q], g.deg(order[q]), (int)L};
                if (v.dist > y.dist || (v.id != y.id && v.deg > y.deg)) y = v;
            }
        if (best < y.dist) {
            best = y.dist;
            x = Node{y.id, y.deg, 0};
        } else {
            return y.id;
        }
    }
}

}  // namespace

extern "C" int fg_rcm(int neq, const int *IA, const int *JA, int *perm)
{
    FG_ARG(neq >= 0 && IA && JA && perm, "rcm args");
    FG_ARG(IA[0] == 0, "IA[0] must be 0");
    for (int r = 0; r < neq; r++) FG_ARG(IA[r + 1] >= IA[r], "IA must be non-decreasing");
    for (int q = 0; q < IA[neq]; q++)
        FG_ARG(JA[q] >= 0 && JA[q] < neq, "JA entry out of range");
    FgPool pool;
    const Graph g(pool, neq, IA, JA);
    if (g.maxdeg() == 0) {
        for (int i = 0; i < neq; i++) perm[i] = i;
        return FG_OK;
    }
    std::vector<int> order(neq), claim(neq, INT_MAX), sorder(neq), sclaim(neq);
    std::vector<char> seen(neq, 0), sseen(neq);
    int c = 0, lowest = 0;
    while (c < neq) {
        while (lowest < neq && seen[lowest]) lowest++;
        const int s = start_node(pool, g, lowest, sorder, sseen, sclaim);
        c = bfs(pool, g, s, c, true, order, seen, claim, nullptr);
    }
    for (int i = 0; i < neq; i++) perm[i] = order[neq - 1 - i];
    return FG_OK;
}

extern "C" int fg_csr_bandwidth(int neq, const int *IA, const int *JA, int *bandwidth)
{
    FG_ARG(neq >= 0 && IA && JA && bandwidth, "bandwidth args");
    int bw = 0;
    for (int i = 0; i < neq; i++)
        for (int q = IA[i]; q < IA[i + 1]; q++) bw = std::max(bw, std::abs(JA[q] - i));
    *bandwidth = bw;
    return FG_OK;
}
