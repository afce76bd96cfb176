// fg_ede4.hip -- the edge-by-edge structure of the 4-dof (Euler 2D) driver.
//
// 03_Euler_Equations/2D/01_CommonFiles/ede_Initialization.c:4-190 numbers
// the edges of a triangle mesh in order of first appearance (element
// ascending, within an element the pairs (I1,I2), (I1,I3), (I2,I3) of its
// sorted vertices), records each element's vertex order and edge numbers,
// and gives every edge the 8 equation numbers of its two nodes (smaller node
// first, DOF fastest; prescribed dofs -> neq).  ede_assembly.c:3-113 of that
// driver adds each element's 12x12 matrix into its three 8x8 edge matrices:
// the off-diagonal node blocks whole, the node-diagonal blocks halved, in
// element order into zeroed edge matrices.
//
// Host: the numbering (an ordered map per node, the reference's sorted edge
// lists).  Device: the assembly, one thread per (edge, row, node block); an interior
// edge has exactly two elements, and the thread adds the first one's
// contribution and then the second's, as the reference's element loop does,
// so the result is bitwise the reference's without atomics.
#include <string.h>

#include <map>
#include <vector>

#include "fg_internal.h"

struct fg_ede4 {
    fg_ctx *ctx = nullptr;
    int nnodes = 0, nel = 0, neq = 0, nedge = 0;
    std::vector<int> order, ebe, lm2, enodes;   // host copies
    int *d_order = nullptr;                     // [nel*3]
    int *d_einc = nullptr;                      // [nedge*2] 3*e + slot of the first / second element, -1 none
};

namespace {

// the 8x8 contribution of element e's edge `slot` (local vertex order o):
// row r of the edge = (node a or b, dof r % 4); thread (J, r, hb) writes the
// four entries of row r in column block hb (node a: hb = 0, node b: 1), from
// four contiguous entries of a row of each element's row-major 12x12 matrix
__global__ void k_ede4_assemble(int nedge, const int *__restrict__ order,
                                const int *__restrict__ einc, const double *__restrict__ A12,
                                double *__restrict__ Aede)
{
    const long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x;
    if (t >= 16LL * nedge) return;
    const int J = (int)(t >> 4), r = (int)((t >> 1) & 7), hb = (int)(t & 1);
    const bool diag = (r < 4) == (hb == 0);   // node-diagonal block: halved
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int h = 0; h < 2; h++) {
        const int q = einc[2 * J + h];
        if (q < 0) break;
        const int e = q / 3, slot = q - 3 * e;
        const int *o = order + 3 * e;
        // edges: (I1,I2) = slot 0, (I1,I3) = slot 1, (I2,I3) = slot 2
        const int la = slot == 2 ? o[1] : o[0];
        const int lb = slot == 0 ? o[1] : o[2];
        const int row = 4 * (r < 4 ? la : lb) + (r & 3);
        const double *src = A12 + 144LL * e + 12 * row + 4 * (hb ? lb : la);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const double v = src[c];
            m[c] += diag ? 0.5 * v : v;
        }
    }
    double *dst = Aede + 64LL * J + 8 * r + 4 * hb;
#pragma unroll
    for (int c = 0; c < 4; c++) dst[c] = m[c];
}

}  // namespace

extern "C" int fg_ede4_create(fg_ctx *c, int nnodes, int nel, int neq, const int *vertex,
                              const int *ids, fg_ede4 **out)
{
    FG_ARG(c && vertex && ids && out && nnodes > 0 && nel >= 0 && neq >= 0, "ede4 args");
    for (long long q = 0; q < 3LL * nel; q++)
        FG_ARG(vertex[q] >= 0 && vertex[q] < nnodes, "vertex out of range");
    // ede_Initialization.c:153-178 maps id -1 (prescribed) to neq and keeps any other id
    for (long long q = 0; q < 4LL * nnodes; q++)
        FG_ARG(ids[q] >= -1 && ids[q] < neq, "node dof id outside [-1, neq)");
    fg_ede4 *g = new fg_ede4();
    g->ctx = c;
    g->nnodes = nnodes;
    g->nel = nel;
    g->neq = neq;
    g->order.resize(3 * (size_t)nel);
    g->ebe.resize(3 * (size_t)nel);
    std::vector<std::map<int, int>> lists(nnodes);   // node I -> (J > I -> edge number)
    std::vector<int> a_node, b_node;
    auto edge = [&](int I, int J) {
        auto it = lists[I].find(J);
        if (it != lists[I].end()) return it->second;
        const int id = (int)a_node.size();
        lists[I].emplace(J, id);
        a_node.push_back(I);
        b_node.push_back(J);
        return id;
    };
    for (long long e = 0; e < nel; e++) {
        const int J1 = vertex[3 * e], J2 = vertex[3 * e + 1], J3 = vertex[3 * e + 2];
        int o0, o1, o2;   // local indices of the smallest, middle, largest vertex
        if (J1 < J2 && J1 < J3) {
            o0 = 0;
            if (J2 < J3) { o1 = 1; o2 = 2; } else { o1 = 2; o2 = 1; }
        } else if (J2 < J3) {
            o0 = 1;
            if (J1 < J3) { o1 = 0; o2 = 2; } else { o1 = 2; o2 = 0; }
        } else {
            o0 = 2;
            if (J1 < J2) { o1 = 0; o2 = 1; } else { o1 = 1; o2 = 0; }
        }
        int *o = &g->order[3 * e];
        o[0] = o0; o[1] = o1; o[2] = o2;
        const int I1 = vertex[3 * e + o0], I2 = vertex[3 * e + o1], I3 = vertex[3 * e + o2];
        g->ebe[3 * e + 0] = edge(I1, I2);
        g->ebe[3 * e + 1] = edge(I1, I3);
        g->ebe[3 * e + 2] = edge(I2, I3);
    }
    const int nedge = g->nedge = (int)a_node.size();
    g->lm2.resize(8 * (size_t)nedge);
    g->enodes.resize(2 * (size_t)nedge);
    for (int J = 0; J < nedge; J++) {
        g->enodes[2 * J] = a_node[J];
        g->enodes[2 * J + 1] = b_node[J];
        for (int k = 0; k < 4; k++) {
            const int ia = ids[4 * (size_t)a_node[J] + k], ib = ids[4 * (size_t)b_node[J] + k];
            g->lm2[8 * (size_t)J + k] = ia == -1 ? neq : ia;
            g->lm2[8 * (size_t)J + 4 + k] = ib == -1 ? neq : ib;
        }
    }
    // the elements of each edge, element order
    std::vector<int> einc(2 * (size_t)nedge, -1);
    for (long long e = 0; e < nel; e++)
        for (int s = 0; s < 3; s++) {
            const int J = g->ebe[3 * e + s];
            if (einc[2 * J] < 0) einc[2 * J] = (int)(3 * e + s);
            else if (einc[2 * J + 1] < 0) einc[2 * J + 1] = (int)(3 * e + s);
            else {
                delete g;
                fg_set_error("edge %d belongs to more than two elements", J);
                return FG_ERR_ARG;
            }
        }
    FG_HIP(hipSetDevice(c->device));
    if (hipMalloc(&g->d_order, sizeof(int) * (g->order.empty() ? 1 : g->order.size())) !=
            hipSuccess ||
        hipMalloc(&g->d_einc, sizeof(int) * (einc.empty() ? 1 : einc.size())) != hipSuccess) {
        hipFree(g->d_order);
        delete g;
        fg_set_error("ede4: out of device memory");
        return FG_ERR_NOMEM;
    }
    if (!g->order.empty())
        FG_HIP(hipMemcpy(g->d_order, g->order.data(), sizeof(int) * g->order.size(),
                         hipMemcpyHostToDevice));
    if (!einc.empty())
        FG_HIP(hipMemcpy(g->d_einc, einc.data(), sizeof(int) * einc.size(),
                         hipMemcpyHostToDevice));
    *out = g;
    return FG_OK;
}

void fg_ede4_sizes(const fg_ede4 *g, fg_ctx **c, int *nel, int *nedge)
{
    *c = g->ctx;
    *nel = g->nel;
    *nedge = g->nedge;
}

extern "C" int fg_ede4_info(const fg_ede4 *g, int *nedge)
{
    FG_ARG(g && nedge, "ede4/nedge");
    *nedge = g->nedge;
    return FG_OK;
}

extern "C" int fg_ede4_get(const fg_ede4 *g, int *order, int *edge_by_elem, int *lm2,
                           int *edge_nodes)
{
    FG_ARG(g, "ede4");
    if (order) memcpy(order, g->order.data(), sizeof(int) * g->order.size());
    if (edge_by_elem) memcpy(edge_by_elem, g->ebe.data(), sizeof(int) * g->ebe.size());
    if (lm2) memcpy(lm2, g->lm2.data(), sizeof(int) * g->lm2.size());
    if (edge_nodes) memcpy(edge_nodes, g->enodes.data(), sizeof(int) * g->enodes.size());
    return FG_OK;
}

extern "C" int fg_ede4_assemble(fg_ede4 *g, const double *A12_dev, double *Aede_dev)
{
    FG_ARG(g && A12_dev && Aede_dev, "ede4/A12/Aede");
    fg_ctx *c = g->ctx;
    if (g->nedge == 0) return FG_OK;
    const long long n = 16LL * g->nedge;
    const int grid = (int)((n + FG_TPB - 1) / FG_TPB);
    FG_LAUNCH(k_ede4_assemble, grid, FG_TPB, 0, c->stream, g->nedge, g->d_order, g->d_einc,
              A12_dev, Aede_dev);
    return FG_OK;
}

extern "C" int fg_ede4_destroy(fg_ede4 *g)
{
    if (!g) return FG_OK;
    hipStreamSynchronize(g->ctx->stream);
    hipFree(g->d_order);
    hipFree(g->d_einc);
    delete g;
    return FG_OK;
}
