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
#include <string.h>

#include <vector>

#include "fg_internal.h"

namespace {

struct Node {
    int id, deg, dist;
};

class Graph {
public:
    Graph(int n, const int *ia, const int *ja) : ia_(ia), ja_(ja), ta_(n + 1, 0)
    {
        // transposed pattern (columns of A as rows), rows ascending
        const int nz = ia[n];
        tj_.resize(nz > 0 ? nz : 1);
        std::vector<int> w(n + 1, 0);
        for (int q = 0; q < nz; q++) w[ja[q]]++;
        int acc = 0;
        for (int i = 0; i < n; i++) {
            ta_[i] = acc;
            acc += w[i];
            w[i] = ta_[i];
        }
        ta_[n] = acc;
        for (int j = 0; j < n; j++)
            for (int k = ia[j]; k < ia[j + 1]; k++) tj_[w[ja[k]]++] = j;
        // degree: entries of each column, plus the mirror of each entry that
        // has none (symmetrised pattern)
        deg_.assign(n, 0);
        maxdeg_ = 0;
        for (int j = 0; j < n; j++)
            for (int q = ia[j]; q < ia[j + 1]; q++) {
                const int k = ja[q];
                maxdeg_ = std::max(maxdeg_, ++deg_[k]);
                if (k == j) continue;
                bool found = false;
                for (int l = ia[k]; l < ia[k + 1]; l++) {
                    if (ja[l] == j) { found = true; break; }
                    if (ja[l] > j) break;
                }
                if (!found) maxdeg_ = std::max(maxdeg_, ++deg_[j]);
            }
    }

    int maxdeg() const { return maxdeg_; }
    int deg(int i) const { return deg_[i]; }

    // neighbours of i, ascending, each index once
    template <typename F>
    void each_neighbour(int i, F &&f) const
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

private:
    const int *ia_, *ja_;
    std::vector<int> ta_, tj_, deg_;
    int maxdeg_ = 0;
};

// circular queue of capacity n+1
struct Queue {
    std::vector<Node> q;
    int head = 0, tail = 0;
    explicit Queue(int n) : q(n + 1) {}
    bool empty() const { return head == tail; }
    void push(const Node &x) { q[tail] = x; tail = (tail + 1) % (int)q.size(); }
    Node pop() { Node x = q[head]; head = (head + 1) % (int)q.size(); return x; }
    void clear() { head = tail = 0; }
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

// pseudo-peripheral node: BFS from `start`, move to the farthest node (ties:
// a different node of larger degree), repeat while the eccentricity grows
int start_node(const Graph &g, int n, int start)
{
    Queue q(n);
    std::vector<char> seen(n, 0);
    Node x{start, g.deg(start), 0};
    q.push(x);
    seen[start] = 1;
    int best = 0;
    for (;;) {
        while (!q.empty()) {
            const Node v = q.pop();
            if (v.dist > x.dist || (v.id != x.id && v.deg > x.deg)) x = v;
            g.each_neighbour(v.id, [&](int r) {
                if (seen[r]) return;
                seen[r] = 1;
                q.push(Node{r, g.deg(r), v.dist + 1});
            });
        }
        if (best < x.dist) {
            best = x.dist;
            std::fill(seen.begin(), seen.end(), 0);
            seen[x.id] = 1;
            x.dist = 0;
            q.clear();
            q.push(x);
        } else {
            return x.id;
        }
    }
}

}  // namespace

extern "C" int fg_rcm(int neq, const int *IA, const int *JA, int *perm)
{
    FG_ARG(neq >= 0 && IA && JA && perm, "rcm args");
    for (int q = 0; q < IA[neq]; q++)
        FG_ARG(JA[q] >= 0 && JA[q] < neq, "JA entry out of range");
    const Graph g(neq, IA, JA);
    if (g.maxdeg() == 0) {
        for (int i = 0; i < neq; i++) perm[i] = i;
        return FG_OK;
    }
    std::vector<char> seen(neq, 0);
    Queue q(neq);
    DegHeap heap(g.maxdeg());
    int c = 0, lowest = 0;
    while (c < neq) {
        while (lowest < neq && seen[lowest]) lowest++;
        const int s = start_node(g, neq, lowest);
        seen[s] = 1;
        q.push(Node{s, g.deg(s), 0});
        while (!q.empty()) {
            const Node v = q.pop();
            perm[c++] = v.id;
            g.each_neighbour(v.id, [&](int r) {
                if (seen[r]) return;
                seen[r] = 1;
                heap.push(Node{r, g.deg(r), v.dist + 1});
            });
            while (heap.size > 0) q.push(heap.pop());
        }
    }
    for (int i = 0, j = neq - 1; i < j; i++, j--) std::swap(perm[i], perm[j]);
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
