// fg_mesh.hip -- mesh structures (host) and SUPG assembly (device).
//
// Host structure builders restate what the SS-2D driver does once per mesh:
//   node numbering  Preprocess.c:73-81
//   Fill_LM         Fill_LM.c:3-24
//   CSR pattern     csr_Initialization.c:5-218 (reordering NOT)
//   EDE pattern     ede_Initialization.c:4-171
// with counting sorts instead of the reference's sorted linked lists (same
// sets, same numbering; O(nnz) instead of O(nnz * row length)), each pass
// split over the host cores (fg_host.h).
//
// Device assembly:
//   k_supg2d     one thread per triangle: Build_K_F_SUPG.c:35-149 (ke, Fe,
//                CAU shock capturing) -> ke[nel*9] row-major, Fe[nel*3]
//   k_gather_F   one thread per equation: F_assembly.c:5-27 as a gather over
//                the equation's (element, vertex) incidences in element order,
//                so F is summed exactly in the reference's order
//   k_gather_csr / k_gather_ede: csr_assembly.c / ede_assembly.c the same way.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fg_ebe_body.h"
#include "fg_host.h"
#include "fg_internal.h"
#include "fg_sell.h"


// ---------------------------------------------------------------------------
// host structure builders
// ---------------------------------------------------------------------------
extern "C" int fg_number_nodes(int nnodes, const int *type, int *id)
{
    int neq = 0;
    for (int i = 0; i < nnodes; i++) id[i] = type[i] >= 1 ? neq++ : -1;
    return neq;
}

extern "C" int fg_fill_lm(int neq, int nel, int nnoel, const int *vertex, const int *type,
                          const int *id, int *lm)
{
    FG_ARG(vertex && type && id && lm && nnoel > 0, "fill_lm arguments");
    for (long e = 0; e < nel; e++)
        for (int j = 0; j < nnoel; j++) {
            int v = vertex[e * nnoel + j];
            lm[e * nnoel + j] = type[v] >= 1 ? id[v] : neq;
        }
    return FG_OK;
}

namespace {
// (row, col) pairs -> sorted unique columns of every row (rows 0..nrows-1).
// pair(q, &r, &c) gives the pair at position q in [0, nq), or false when
// position q contributes none.  Parallel: counts per row (atomic), columns
// placed into their row's segment (atomic cursors: any order), each segment
// sorted and made unique, the segments compacted.  The result is the same
// whatever the thread count or timing.
template <class Pair>
int build_pattern(FgPool &pool, int nrows, long long nq, Pair &&pair, std::vector<long long> &off,
                  std::vector<int> &cols, bool need_cols)
{
    std::vector<long long> seg((size_t)nrows + 1, 0);
    pool.range(nq, [&](long long b, long long e, int) {
        for (long long q = b; q < e; q++) {
            int r, c;
            if (pair(q, &r, &c)) __atomic_fetch_add(&seg[r], 1LL, __ATOMIC_RELAXED);
        }
    });
    seg[nrows] = 0;
    const long long total = fg_exclusive_scan(seg);
    seg[nrows] = total;
    std::vector<int> tmp(total > 0 ? total : 1);
    {
        std::vector<long long> cur(seg.begin(), seg.end() - 1);
        pool.range(nq, [&](long long b, long long e, int) {
            for (long long q = b; q < e; q++) {
                int r, c;
                if (pair(q, &r, &c)) tmp[__atomic_fetch_add(&cur[r], 1LL, __ATOMIC_RELAXED)] = c;
            }
        });
    }
    off.assign((size_t)nrows + 1, 0);
    pool.range(nrows, [&](long long b, long long e, int) {
        for (long long r = b; r < e; r++) {
            int *p0 = tmp.data() + seg[r], *p1 = tmp.data() + seg[r + 1];
            std::sort(p0, p1);
            off[r] = std::unique(p0, p1) - p0;
        }
    });
    off[nrows] = 0;
    const long long nnz = fg_exclusive_scan(off);
    off[nrows] = nnz;
    cols.clear();
    if (!need_cols) return FG_OK;
    cols.resize(nnz > 0 ? nnz : 1);
    pool.range(nrows, [&](long long b, long long e, int) {
        for (long long r = b; r < e; r++)
            std::copy(tmp.data() + seg[r], tmp.data() + seg[r] + (off[r + 1] - off[r]),
                      cols.data() + off[r]);
    });
    cols.resize(nnz);
    return FG_OK;
}

int find_col(const int *cols, long long lo, long long hi, int c)
{
    return (int)(std::lower_bound(cols + lo, cols + hi, c) - cols);
}

void ede_order(const int *l, int *o)
{
    // ede_Initialization.c:22-83 comparison chain (ties resolved identically)
    if (l[0] < l[1] && l[0] < l[2]) {
        if (l[1] < l[2]) { o[0] = 0; o[1] = 1; o[2] = 2; }
        else             { o[0] = 0; o[1] = 2; o[2] = 1; }
    } else if (l[1] < l[2]) {
        if (l[0] < l[2]) { o[0] = 1; o[1] = 0; o[2] = 2; }
        else             { o[0] = 1; o[1] = 2; o[2] = 0; }
    } else {
        if (l[0] < l[1]) { o[0] = 2; o[1] = 0; o[2] = 1; }
        else             { o[0] = 2; o[1] = 1; o[2] = 0; }
    }
}
}  // namespace

extern "C" int fg_csr_pattern(int neq, int nel, const int *lm, int *IA, int *JA,
                              int *csr_by_elem, long long *nnz_out)
{
    return fg_csr_pattern_k(neq, nel, 3, lm, IA, JA, csr_by_elem, nnz_out);
}

extern "C" int fg_csr_pattern_k(int neq, int nel, int k, const int *lm, int *IA, int *JA,
                                int *csr_by_elem, long long *nnz_out)
{
    FG_ARG(lm && nnz_out && k >= 1 && k <= 16 && neq >= 0 && nel >= 0, "lm/nnz/k");
    FG_ARG((long long)nel * k * k < (1LL << 31), "nel * k * k must fit in an int");
    const long long nq = (long long)nel * k;
    for (long long q = 0; q < nq; q++)
        FG_ARG(lm[q] >= 0 && lm[q] <= neq, "lm entries must lie in [0, neq]");
    FgPool pool;
    // equation -> incident elements, ascending.  Worker t owns the
    // equations [neq t / T, neq (t+1) / T): it streams lm and counts, then
    // places, only its own equations (no atomics)
    std::vector<long long> ioff((size_t)neq + 1, 0);
    const int T = pool.size();
    auto owned = [&](int t, int *r0, int *r1) {
        *r0 = (int)((long long)neq * t / T);
        *r1 = (int)((long long)neq * (t + 1) / T);
    };
    pool.run([&](int t) {
        int r0, r1;
        owned(t, &r0, &r1);
        for (long long q = 0; q < nq; q++)
            if (lm[q] >= r0 && lm[q] < r1) ioff[lm[q]]++;
    });
    ioff[neq] = 0;
    const long long ninc = fg_exclusive_scan(ioff);
    ioff[neq] = ninc;
    // incidences as lm positions q = e k + a (element e, local row a)
    std::vector<int> inc(ninc > 0 ? ninc : 1);
    {
        std::vector<long long> cur(ioff.begin(), ioff.end() - 1);
        pool.run([&](int t) {
            int r0, r1;
            owned(t, &r0, &r1);
            for (long long q = 0; q < nq; q++)
                if (lm[q] >= r0 && lm[q] < r1) inc[cur[lm[q]]++] = (int)q;
        });
    }
    // row r: the sorted, distinct non-sink dofs of its elements.  One pass:
    // every worker keeps the columns of its contiguous rows, which are then
    // copied to their offsets.
    auto row_cols = [&](long long r, std::vector<int> &buf) {
        buf.clear();
        for (long long q = ioff[r]; q < ioff[r + 1]; q++) {
            const int *le = lm + (long long)k * (inc[q] / k);
            for (int b = 0; b < k; b++)
                if (le[b] != neq) buf.push_back(le[b]);
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    };
    std::vector<long long> off((size_t)neq + 1, 0);
    std::vector<std::vector<int>> tcols(T);
    std::vector<long long> tfirst(T, -1);
    pool.range(neq, [&](long long b, long long e, int t) {
        std::vector<int> buf;
        tfirst[t] = b;
        for (long long r = b; r < e; r++) {
            row_cols(r, buf);
            off[r] = (long long)buf.size();
            if (JA) tcols[t].insert(tcols[t].end(), buf.begin(), buf.end());
        }
    });
    off[neq] = 0;
    const long long nnz = fg_exclusive_scan(off);
    off[neq] = nnz;
    *nnz_out = nnz;
    if (!JA) return FG_OK;
    FG_ARG(nnz < 2147483647LL, "nnz fits int32");
    FG_ARG(IA, "IA");
    for (int r = 0; r <= neq; r++) IA[r] = (int)off[r];
    pool.run([&](int t) {
        if (tfirst[t] >= 0) std::copy(tcols[t].begin(), tcols[t].end(), JA + off[tfirst[t]]);
    });
    std::vector<std::vector<int>>().swap(tcols);
    if (csr_by_elem) {
        // rows of the sink: the sink slot nnz
        pool.range(nq, [&](long long b, long long e, int) {
            for (long long q = b; q < e; q++)
                if (lm[q] == neq)
                    for (int bb = 0; bb < k; bb++) csr_by_elem[(long long)k * q + bb] = (int)nnz;
        });
        // the other rows from their equation's incidences: the slot of each
        // column in the (cache-resident) row
        pool.range(neq, [&](long long b, long long e, int) {
            for (long long r = b; r < e; r++)
                for (long long q = ioff[r]; q < ioff[r + 1]; q++) {
                    const long long qq = inc[q];
                    const int *le = lm + (long long)k * (qq / k);
                    for (int bb = 0; bb < k; bb++)
                        csr_by_elem[(long long)k * qq + bb] =
                            le[bb] != neq ? find_col(JA, off[r], off[r + 1], le[bb]) : (int)nnz;
                }
        });
    }
    return FG_OK;
}

extern "C" int fg_ede_pattern(int neq, int nel, const int *lm, int *order, int *edge_by_elem,
                              int *lm2, int *nedge_out)
{
    FG_ARG(lm && nedge_out && neq >= 0 && nel >= 0, "lm/nedge");
    FG_ARG((long long)nel * 3 < (1LL << 31), "nel * 3 must fit in an int");
    for (long long q = 0; q < 3LL * nel; q++)
        FG_ARG(lm[q] >= 0 && lm[q] <= neq, "lm entries must lie in [0, neq]");
    FgPool pool;
    // each element's vertices in the comparison order, then its three edges
    // (I1,I2), (I1,I3), (I2,I3) as (row, column) pairs
    std::vector<int> ord((size_t)nel * 3);
    pool.range(nel, [&](long long b, long long e1, int) {
        for (long long e = b; e < e1; e++) ede_order(lm + 3 * e, &ord[3 * e]);
    });
    auto vtx = [&](long long e, int j) { return lm[3 * e + ord[3 * e + j]]; };
    auto pair = [&](long long q, int *r, int *c) {
        const long long e = q / 3;
        const int j = (int)(q - 3 * e);
        *r = vtx(e, j == 2 ? 1 : 0);
        *c = vtx(e, j == 0 ? 1 : 2);
        return true;
    };
    std::vector<long long> off;
    std::vector<int> cols;
    // rows 0..neq (row neq: the sink edge)
    FG_CHECK(build_pattern(pool, neq + 1, 3LL * nel, pair, off, cols, lm2 != nullptr));
    const int nedge = (int)off[neq + 1] - 1;   // ede_Initialization.c:103
    *nedge_out = nedge;
    if (!lm2) return FG_OK;
    if (order) memcpy(order, ord.data(), sizeof(int) * ord.size());
    if (edge_by_elem)
        pool.range(nel, [&](long long b, long long e1, int) {
            for (long long e = b; e < e1; e++) {
                const int I1 = vtx(e, 0), I2 = vtx(e, 1), I3 = vtx(e, 2);
                edge_by_elem[3 * e + 0] = find_col(cols.data(), off[I1], off[I1 + 1], I2);
                edge_by_elem[3 * e + 1] = find_col(cols.data(), off[I1], off[I1 + 1], I3);
                edge_by_elem[3 * e + 2] = find_col(cols.data(), off[I2], off[I2 + 1], I3);
            }
        });
    pool.range(neq, [&](long long b, long long e1, int) {
        for (long long r = b; r < e1; r++)
            for (long long q = off[r]; q < off[r + 1]; q++) {
                lm2[2 * q] = (int)r;
                lm2[2 * q + 1] = cols[q];
            }
    });
    return FG_OK;
}

// ---------------------------------------------------------------------------
// mesh object
// ---------------------------------------------------------------------------
struct fg_mesh2d {
    fg_ctx *ctx = nullptr;
    int nnodes = 0, nel = 0, neq = 0;
    double *d_x = nullptr, *d_y = nullptr;
    int *d_vertex = nullptr, *d_id = nullptr, *d_type = nullptr;
    int *d_finc_off = nullptr, *d_finc = nullptr;   // equation -> e*3+r, element order
    // the same incidences as sliced rows (fg_sell.h) for the residual gather
    int nslice = 0;
    int *d_isoff = nullptr, *d_isinc = nullptr;
    double *d_Fe = nullptr;                          // [nel*3]
    double *d_Tr = nullptr;                          // [nel*12] transient residual terms
    std::vector<int> h_lm;
};

template <typename T>
static int up(const T *h, size_t n, T **d)
{
    FG_HIP(hipMalloc(d, sizeof(T) * (n ? n : 1)));
    if (n) FG_HIP(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
    return FG_OK;
}

extern "C" int fg_mesh2d_create(fg_ctx *c, int nnodes, int nel, const double *x,
                                const double *y, const int *type, const int *vertex,
                                fg_mesh2d **out)
{
    FG_ARG(c && out && x && y && type && vertex && nnodes > 0 && nel >= 0, "mesh2d args");
    for (long q = 0; q < 3L * nel; q++)
        FG_ARG(vertex[q] >= 0 && vertex[q] < nnodes, "vertex index out of range");
    FG_HIP(hipSetDevice(c->device));
    fg_mesh2d *m = new fg_mesh2d();
    m->ctx = c;
    m->nnodes = nnodes;
    m->nel = nel;
    std::vector<int> id(nnodes);
    m->neq = fg_number_nodes(nnodes, type, id.data());
    m->h_lm.resize((size_t)nel * 3);
    fg_fill_lm(m->neq, nel, 3, vertex, type, id.data(), m->h_lm.data());
    // F incidence: equation -> (element, vertex), element order
    std::vector<int> off(m->neq + 1, 0), inc;
    for (long q = 0; q < 3L * nel; q++)
        if (m->h_lm[q] != m->neq) off[m->h_lm[q] + 1]++;
    for (int g = 0; g < m->neq; g++) off[g + 1] += off[g];
    inc.resize(off[m->neq]);
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (long q = 0; q < 3L * nel; q++)
            if (m->h_lm[q] != m->neq) inc[fill[m->h_lm[q]]++] = (int)q;
    }
    FG_CHECK(up(x, nnodes, &m->d_x));
    FG_CHECK(up(y, nnodes, &m->d_y));
    FG_CHECK(up(vertex, (size_t)nel * 3, &m->d_vertex));
    FG_CHECK(up(id.data(), nnodes, &m->d_id));
    FG_CHECK(up(type, nnodes, &m->d_type));
    FG_CHECK(up(off.data(), off.size(), &m->d_finc_off));
    FG_CHECK(up(inc.data(), inc.size(), &m->d_finc));
    {
        const fg_sell_layout L = fg_sell_plan(m->neq, off.data());
        std::vector<int> sinc(L.nslot, -1);
        for (int g = 0; g < m->neq; g++)
            for (int q = off[g]; q < off[g + 1]; q++) sinc[L.slot(g, q - off[g])] = inc[q];
        m->nslice = L.nslice;
        FG_CHECK(up(L.soff.data(), L.soff.size(), &m->d_isoff));
        FG_CHECK(up(sinc.data(), sinc.size(), &m->d_isinc));
    }
    FG_HIP(hipMalloc(&m->d_Fe, sizeof(double) * 3 * (nel ? nel : 1)));
    *out = m;
    return FG_OK;
}

// eval_U_dU.c of the 02 driver: nodal U, dU from u and du on the free nodes
// (Type > 0), the prescribed value and 0 on the others
__global__ void k_eval_U_dU_tr(int nn, const int *__restrict__ type, const int *__restrict__ id,
                               const double *__restrict__ u, const double *__restrict__ du,
                               const double *__restrict__ upresc, double *__restrict__ U,
                               double *__restrict__ dU)
{
    for (int i = blockIdx.x * FG_TPB + threadIdx.x; i < nn; i += gridDim.x * FG_TPB) {
        if (type[i] > 0) {
            U[i] = u[id[i]];
            dU[i] = du[id[i]];
        } else {
            U[i] = upresc[i];
            dU[i] = 0.0;
        }
    }
}

int fg_mesh2d_eval_U_dU(const fg_mesh2d *m, const double *u, const double *du,
                        const double *upresc, double *U, double *dU)
{
    fg_ctx *c = m->ctx;
    FG_LAUNCH(k_eval_U_dU_tr, fg_vec_grid(c, m->nnodes), FG_TPB, 0, c->stream, m->nnodes,
              m->d_type, m->d_id, u, du, upresc, U, dU);
    return FG_OK;
}

fg_ctx *fg_mesh2d_ctx(const fg_mesh2d *m) { return m->ctx; }

extern "C" int fg_mesh2d_info(const fg_mesh2d *m, int *neq, int *nel, int *nnodes)
{
    FG_ARG(m, "mesh");
    if (neq) *neq = m->neq;
    if (nel) *nel = m->nel;
    if (nnodes) *nnodes = m->nnodes;
    return FG_OK;
}

extern "C" int fg_mesh2d_get_lm(const fg_mesh2d *m, int *lm)
{
    FG_ARG(m && lm, "mesh/lm");
    memcpy(lm, m->h_lm.data(), sizeof(int) * m->h_lm.size());
    return FG_OK;
}

extern "C" int fg_mesh2d_destroy(fg_mesh2d *m)
{
    if (!m) return FG_OK;
    hipStreamSynchronize(m->ctx->stream);
    hipFree(m->d_x); hipFree(m->d_y); hipFree(m->d_vertex); hipFree(m->d_id);
    hipFree(m->d_type); hipFree(m->d_finc_off); hipFree(m->d_finc); hipFree(m->d_Fe);
    hipFree(m->d_isoff); hipFree(m->d_isinc);
    hipFree(m->d_Tr);
    delete m;
    return FG_OK;
}

// ---------------------------------------------------------------------------
// SUPG element kernel
// ---------------------------------------------------------------------------
__device__ __forceinline__ void prob_velocity(const fg_supg2d_problem &pb, double X, double Y,
                                              double &bx, double &by)
{
    if (pb.kind == FG_P_ROTATING) {
        bx = (-Y + 0.5) * pb.scale;
        by = (X - 0.5) * pb.scale;
    } else if (pb.kind == FG_P_ROTATE_ABOUT) {   // 02 CONE/CONE_Velocity.c:3-12
        bx = -(Y - pb.by) * pb.scale;
        by = (X - pb.bx) * pb.scale;
    } else if (pb.kind == FG_P_VALIDACAO) {
        bx = 1;
        by = 20 * Y;
    } else {
        bx = pb.bx;
        by = pb.by;
    }
}

// VALIDACAO/VALIDACAO_Font.c:4-17
__device__ double validacao_font(double X, double Y, double k, double gamma, double Be_x,
                                 double Be_y)
{
    double E = exp(pow(X, 4.5));
    double u = 10 * X * Y * (1 - X) * (1 - Y) * E;
    double u_x = 10 * Y * (1 - Y) * (E - 2 * E * X + 4.5 * E * pow(X, 4.5) - 4.5 * E * pow(X, 5.5));
    double u_x2 = 10 * Y * (1 - Y) * (-20.25 * E * pow(X, 9.0) + 20.25 * E * pow(X, 8.0)
                  - 33.75 * E * pow(X, 4.5) + 24.75 * E * pow(X, 3.5) - 2 * E);
    double u_y = 10 * E * X * (1 - X) * (1 - 2 * Y);
    double u_y2 = -20 * E * X * (1 - X);
    return -k * (u_x2 + u_y2) + Be_x * u_x + Be_y * u_y + gamma * u;
}

__device__ __forceinline__ double prob_font(const fg_supg2d_problem &pb, double X, double Y,
                                            double k, double g, double bx, double by)
{
    if (pb.kind == FG_P_VALIDACAO) return validacao_font(X, Y, k, g, bx, by);
    return pb.f;
}

__global__ __launch_bounds__(FG_TPB) void k_supg2d(
    int nel, const double *__restrict__ x, const double *__restrict__ y,
    const int *__restrict__ vertex, const int *__restrict__ id, const double *__restrict__ u,
    fg_supg2d_problem pb, double *__restrict__ ke, double *__restrict__ Fe)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    double kx, ky, gamma;
    if (pb.kind == FG_P_VALIDACAO) { kx = 1.0; ky = 1.0; gamma = 1; }
    else { kx = pb.kappa; ky = pb.kappa; gamma = pb.gamma; }
    for (int E = blockIdx.x * FG_TPB + threadIdx.x; E < nel; E += gridDim.x * FG_TPB) {
        int J1 = vertex[3 * E], J2 = vertex[3 * E + 1], J3 = vertex[3 * E + 2];
        double X0 = x[J1], X1 = x[J2], X2 = x[J3];
        double Y0 = y[J1], Y1 = y[J2], Y2 = y[J3];
        double y23 = Y1 - Y2, y31 = Y2 - Y0, y12 = Y0 - Y1;
        double x32 = X2 - X1, x13 = X0 - X2, x21 = X1 - X0;
        double TwoA = x21 * y31 - x13 * y12;
        double Area = 0.5 * TwoA;
        double invArea = 1.0 / Area;
        double h = sqrt(fabs(TwoA));
        double XB = third * (X0 + X1 + X2);
        double YB = third * (Y0 + Y1 + Y2);
        double Be_x, Be_y;
        prob_velocity(pb, XB, YB, Be_x, Be_y);
        double norm = sqrt(Be_x * Be_x + Be_y * Be_y);
        double tau = norm > 0 ? 0.5 * h / norm : 0;
        double fe1 = prob_font(pb, X0, Y0, kx, gamma, Be_x, Be_y);
        double fe2 = prob_font(pb, X1, Y1, kx, gamma, Be_x, Be_y);
        double fe3 = prob_font(pb, X2, Y2, kx, gamma, Be_x, Be_y);
        double feb = third * (fe1 + fe2 + fe3);
        int i1 = id[J1], i2 = id[J2], i3 = id[J3];
        double ue1 = (u && i1 >= 0) ? u[i1] : 0.0;
        double ue2 = (u && i2 >= 0) ? u[i2] : 0.0;
        double ue3 = (u && i3 >= 0) ? u[i3] : 0.0;
        double ueb = third * (ue1 + ue2 + ue3);

        double D12 = 0.25 * invArea * ((kx * y23 * y31) + (ky * x32 * x13));
        double D13 = 0.25 * invArea * ((kx * y23 * y12) + (ky * x32 * x21));
        double D23 = 0.25 * invArea * ((kx * y31 * y12) + (ky * x13 * x21));
        double C11 = sixth * (Be_x * y23 + Be_y * x32);
        double C12 = sixth * (Be_x * y31 + Be_y * x13);
        double C13 = sixth * (Be_x * y12 + Be_y * x21);
        double R11 = sixth * (gamma * Area);
        double R12 = twelfth * (gamma * Area);
        double Cpg12 = 0.25 * invArea * tau * (y23 * (Be_x * Be_x * y31 + Be_x * Be_y * x13) +
                                               x32 * (Be_x * Be_y * y31 + Be_y * Be_y * x13));
        double Cpg13 = 0.25 * invArea * tau * (y23 * (Be_x * Be_x * y12 + Be_x * Be_y * x21) +
                                               x32 * (Be_x * Be_y * y12 + Be_y * Be_y * x21));
        double Cpg23 = 0.25 * invArea * tau * (y31 * (Be_x * Be_x * y12 + Be_x * Be_y * x21) +
                                               x13 * (Be_x * Be_y * y12 + Be_y * Be_y * x21));
        double Rpg11 = sixth * tau * gamma * (Be_x * y23 + Be_y * x32);
        double Rpg21 = sixth * tau * gamma * (Be_x * y31 + Be_y * x13);
        double Rpg31 = sixth * tau * gamma * (Be_x * y12 + Be_y * x21);

        // h_shock (h_shock_2sqrtArea.c / h_shock_sqrtArea.c) and CAU (CAU_ShockCapture.c:3-30)
        double hs = pb.h_shock == FG_H_SQRTAREA ? 2 * sqrt(fabs(Area)) : sqrt(fabs(2 * Area));
        double Eu = 0.0;
        {
            double BetaGradu = 0.5 * invArea * (ue1 * (Be_x * y23 + Be_y * x32) +
                                                ue2 * (Be_x * y31 + Be_y * x13) +
                                                ue3 * (Be_x * y12 + Be_y * x21));
            double Ru = BetaGradu + gamma * ueb - feb;
            double g0 = 0.5 * invArea * (y23 * ue1 + y31 * ue2 + y12 * ue3);
            double g1 = 0.5 * invArea * (x32 * ue1 + x13 * ue2 + x21 * ue3);
            double ng = sqrt(g0 * g0 + g1 * g1);
            if (ng > 0) {
                double bxp = BetaGradu * g0 / (ng * ng);
                double byp = BetaGradu * g1 / (ng * ng);
                double kt = bxp * bxp * kx + byp * byp * ky;
                double nbp = sqrt(bxp * bxp + byp * byp);
                double Pee = 0.5 * hs * nbp * nbp * nbp / kt;
                if (third * Pee < 0.7) Eu = 0.5 * hs * third * Pee * fabs(Ru) / ng;
                else Eu = 0.5 * hs * 0.7 * fabs(Ru) / ng;
            }
        }
        double Csc12 = Eu * 0.25 * invArea * (y23 * y31 + x32 * x13);
        double Csc13 = Eu * 0.25 * invArea * (y23 * y12 + x32 * x21);
        double Csc23 = Eu * 0.25 * invArea * (y31 * y12 + x13 * x21);

        double *k9 = ke + 9L * E;
        k9[0] = -(D12 + D13) + C11 - (Cpg12 + Cpg13) + R11 + Rpg11 - (Csc12 + Csc13);
        k9[1] = D12 + C12 + Cpg12 + R12 + Rpg11 + Csc12;
        k9[2] = D13 + C13 + Cpg13 + R12 + Rpg11 + Csc13;
        k9[3] = D12 + C11 + Cpg12 + R12 + Rpg21 + Csc12;
        k9[4] = -(D12 + D23) + C12 - (Cpg12 + Cpg23) + R11 + Rpg21 - (Csc12 + Csc23);
        k9[5] = D23 + C13 + Cpg23 + R12 + Rpg21 + Csc23;
        k9[6] = D13 + C11 + Cpg13 + R12 + Rpg31 + Csc13;
        k9[7] = D23 + C12 + Cpg23 + R12 + Rpg31 + Csc23;
        k9[8] = -(D13 + D23) + C13 - (Cpg13 + Cpg23) + R11 + Rpg31 - (Csc13 + Csc23);
        double *f3 = Fe + 3L * E;
        f3[0] = Area * twelfth * (2 * fe1 + fe2 + fe3) + sixth * tau * (fe1 + fe2 + fe3) * (Be_x * y23 + Be_y * x32);
        f3[1] = Area * twelfth * (fe1 + 2 * fe2 + fe3) + sixth * tau * (fe1 + fe2 + fe3) * (Be_x * y31 + Be_y * x13);
        f3[2] = Area * twelfth * (fe1 + fe2 + 2 * fe3) + sixth * tau * (fe1 + fe2 + fe3) * (Be_x * y12 + Be_y * x21);
    }
}

// F[g] = sum of Fe over g's incidences in element order (F_assembly.c:5-27, u_D = 0)
__global__ void k_gather_F(int neq, const int *__restrict__ off, const int *__restrict__ inc,
                           const double *__restrict__ Fe, double *__restrict__ F)
{
    for (int g = blockIdx.x * FG_TPB + threadIdx.x; g < neq; g += gridDim.x * FG_TPB) {
        double s = 0;
        for (int q = off[g]; q < off[g + 1]; q++) s += Fe[inc[q]];
        F[g] = s;
    }
}

extern "C" int fg_supg2d_build(fg_mesh2d *m, const fg_supg2d_problem *pb, const double *u,
                               double *ke, double *F)
{
    FG_ARG(m && pb && ke, "mesh/problem/ke");
    FG_ARG(pb->kind >= FG_P_CONST && pb->kind <= FG_P_ROTATE_ABOUT, "problem kind");
    fg_ctx *c = m->ctx;
    if (m->nel > 0)
        FG_LAUNCH(k_supg2d, fg_vec_grid(c, m->nel), FG_TPB, 0, c->stream, m->nel, m->d_x, m->d_y,
                  m->d_vertex, m->d_id, u, *pb, ke, m->d_Fe);
    if (F && m->neq > 0)
        FG_LAUNCH(k_gather_F, fg_vec_grid(c, m->neq), FG_TPB, 0, c->stream, m->neq,
                  m->d_finc_off, m->d_finc, m->d_Fe, F);
    return FG_OK;
}

// ---------------------------------------------------------------------------
// transient SUPG (02_Transport_Equation/2D/01_CommonFiles/Build_M_K_R_SUPG.c)
// ---------------------------------------------------------------------------
// One thread per element: Ae = me + alpha dt ke (:157-159) into A[9E..], and
// for each row i the four terms the reference applies to R[lm[E][i]] in
// order -- the load (:162-164), then me[i][j] du_j + ke[i][j] u_j for
// j = 0, 1, 2 (:166-171) -- into Tr[12E + 4i ..]; a per-equation gather then
// replays them in element order.  Shock capture: CAU with the residual's time
// derivative (02 CAU_ShockCapture.c:3-30), h = |sqrt(2 Area)|
// (02 h_shock_2sqrtArea.c:3-6).
// where an element's 3x3 matrix goes: the AoS array A[9E..] (the reference's
// MatrixData->A), or straight into an operator's storage -- its blocked
// values (FP64 and FP32) and, when it has them, its element rows -- the
// same places fg_op_set_values_dev puts A[9E..]
struct ElemSink {
    double *A;
    const long long *kpos;
    const int *kstride;
    double *K64;
    float *K32;
    const long long *vpos;
    const int *wst;
    double *V64;
    float *V32;
    __device__ __forceinline__ void put(int E, const double (&v)[9]) const
    {
        if (A) {
            double *a = A + 9L * E;
#pragma unroll
            for (int q = 0; q < 9; q++) a[q] = v[q];
            return;
        }
        const long long p = kpos[E];
        if (p < 0) return;   // every dof prescribed: not stored
        const int st = kstride[E];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const long long d = fg_ebe_vidx(p, st, q, 1);
            K64[d] = v[q];
            K32[d] = (float)v[q];
        }
        if (!vpos) return;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const long long vp = vpos[3L * E + r];
            if (vp < 0) continue;
            const long long ws = fg_rows_wst_stride(wst[3L * E + r]);
            const int sh = fg_rows_wst_shift(wst[3L * E + r]);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                V64[vp + c * ws + sh] = v[3 * r + c];
                V32[vp + c * ws] = (float)v[3 * r + c];
            }
        }
    }
};

__global__ __launch_bounds__(FG_TPB) void k_supg2d_tr(
    int nel, const double *__restrict__ x, const double *__restrict__ y,
    const int *__restrict__ vertex, const double *__restrict__ U, const double *__restrict__ dU,
    fg_supg2d_problem pb, double alpha, double dt, ElemSink out, double *__restrict__ Tr)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    double kx, ky, gamma;
    if (pb.kind == FG_P_VALIDACAO) { kx = 1.0; ky = 1.0; gamma = 1; }
    else { kx = pb.kappa; ky = pb.kappa; gamma = pb.gamma; }
    for (int E = blockIdx.x * FG_TPB + threadIdx.x; E < nel; E += gridDim.x * FG_TPB) {
        const int J1 = vertex[3 * E], J2 = vertex[3 * E + 1], J3 = vertex[3 * E + 2];
        const double X0 = x[J1], X1 = x[J2], X2 = x[J3];
        const double Y0 = y[J1], Y1 = y[J2], Y2 = y[J3];
        const double y23 = Y1 - Y2, y31 = Y2 - Y0, y12 = Y0 - Y1;
        const double x32 = X2 - X1, x13 = X0 - X2, x21 = X1 - X0;
        const double TwoA = x21 * y31 - x13 * y12;
        const double Area = 0.5 * TwoA;
        const double invArea = 1.0 / Area;
        const double h = sqrt(fabs(TwoA));
        const double XB = third * (X0 + X1 + X2);
        const double YB = third * (Y0 + Y1 + Y2);
        double Be_x, Be_y;
        prob_velocity(pb, XB, YB, Be_x, Be_y);
        const double norm = sqrt(Be_x * Be_x + Be_y * Be_y);
        const double tau = norm > 0 ? 0.5 * h / norm : 0;
        const double fe1 = prob_font(pb, X0, Y0, kx, gamma, Be_x, Be_y);
        const double fe2 = prob_font(pb, X1, Y1, kx, gamma, Be_x, Be_y);
        const double fe3 = prob_font(pb, X2, Y2, kx, gamma, Be_x, Be_y);
        const double feb = third * (fe1 + fe2 + fe3);
        const double ue[3] = {U[J1], U[J2], U[J3]};
        const double due[3] = {dU[J1], dU[J2], dU[J3]};
        const double ueb = third * (ue[0] + ue[1] + ue[2]);
        const double dueb = third * (due[0] + due[1] + due[2]);

        const double D12 = 0.25 * invArea * ((kx * y23 * y31) + (ky * x32 * x13));
        const double D13 = 0.25 * invArea * ((kx * y23 * y12) + (ky * x32 * x21));
        const double D23 = 0.25 * invArea * ((kx * y31 * y12) + (ky * x13 * x21));
        const double C11 = sixth * (Be_x * y23 + Be_y * x32);
        const double C12 = sixth * (Be_x * y31 + Be_y * x13);
        const double C13 = sixth * (Be_x * y12 + Be_y * x21);
        const double R11 = sixth * (gamma * Area);
        const double R12 = twelfth * (gamma * Area);
        const double Cpg12 = 0.25 * invArea * tau * (y23 * (Be_x * Be_x * y31 + Be_x * Be_y * x13) +
                                                     x32 * (Be_x * Be_y * y31 + Be_y * Be_y * x13));
        const double Cpg13 = 0.25 * invArea * tau * (y23 * (Be_x * Be_x * y12 + Be_x * Be_y * x21) +
                                                     x32 * (Be_x * Be_y * y12 + Be_y * Be_y * x21));
        const double Cpg23 = 0.25 * invArea * tau * (y31 * (Be_x * Be_x * y12 + Be_x * Be_y * x21) +
                                                     x13 * (Be_x * Be_y * y12 + Be_y * Be_y * x21));
        const double Rpg11 = sixth * tau * gamma * (Be_x * y23 + Be_y * x32);
        const double Rpg21 = sixth * tau * gamma * (Be_x * y31 + Be_y * x13);
        const double Rpg31 = sixth * tau * gamma * (Be_x * y12 + Be_y * x21);

        const double hs = fabs(sqrt(2 * Area));
        double Eu = 0.0;
        {
            const double BetaGradu = 0.5 * invArea * (ue[0] * (Be_x * y23 + Be_y * x32) +
                                                      ue[1] * (Be_x * y31 + Be_y * x13) +
                                                      ue[2] * (Be_x * y12 + Be_y * x21));
            const double Ru = dueb + BetaGradu + gamma * ueb - feb;
            const double g0 = 0.5 * invArea * (y23 * ue[0] + y31 * ue[1] + y12 * ue[2]);
            const double g1 = 0.5 * invArea * (x32 * ue[0] + x13 * ue[1] + x21 * ue[2]);
            const double ng = sqrt(g0 * g0 + g1 * g1);
            if (ng > 0) {
                const double bxp = BetaGradu * g0 / (ng * ng);
                const double byp = BetaGradu * g1 / (ng * ng);
                const double kt = bxp * bxp * kx + byp * byp * ky;
                const double nbp = sqrt(bxp * bxp + byp * byp);
                const double Pee = 0.5 * hs * nbp * nbp * nbp / kt;
                if (third * Pee < 0.7) Eu = 0.5 * hs * third * Pee * fabs(Ru) / ng;
                else Eu = 0.5 * hs * 0.7 * fabs(Ru) / ng;
            }
        }
        const double Csc12 = Eu * 0.25 * invArea * (y23 * y31 + x32 * x13);
        const double Csc13 = Eu * 0.25 * invArea * (y23 * y12 + x32 * x21);
        const double Csc23 = Eu * 0.25 * invArea * (y31 * y12 + x13 * x21);

        double me[9], ke[9];
        me[0] = sixth * Area + tau * C11;
        me[1] = twelfth * Area + tau * C11;
        me[2] = twelfth * Area + tau * C11;
        me[3] = twelfth * Area + tau * C12;
        me[4] = sixth * Area + tau * C12;
        me[5] = twelfth * Area + tau * C12;
        me[6] = twelfth * Area + tau * C13;
        me[7] = twelfth * Area + tau * C13;
        me[8] = sixth * Area + tau * C13;
        ke[0] = -(D12 + D13) + C11 - (Cpg12 + Cpg13) + R11 + Rpg11 - (Csc12 + Csc13);
        ke[1] = D12 + C12 + Cpg12 + R12 + Rpg11 + Csc12;
        ke[2] = D13 + C13 + Cpg13 + R12 + Rpg11 + Csc13;
        ke[3] = D12 + C11 + Cpg12 + R12 + Rpg21 + Csc12;
        ke[4] = -(D12 + D23) + C12 - (Cpg12 + Cpg23) + R11 + Rpg21 - (Csc12 + Csc23);
        ke[5] = D23 + C13 + Cpg23 + R12 + Rpg21 + Csc23;
        ke[6] = D13 + C11 + Cpg13 + R12 + Rpg31 + Csc13;
        ke[7] = D23 + C12 + Cpg23 + R12 + Rpg31 + Csc23;
        ke[8] = -(D13 + D23) + C13 - (Cpg13 + Cpg23) + R11 + Rpg31 - (Csc13 + Csc23);
        double ae[9];
#pragma unroll
        for (int q = 0; q < 9; q++) ae[q] = me[q] + alpha * dt * ke[q];
        out.put(E, ae);
        const double fs = fe1 + fe2 + fe3;
        double load[3];
        load[0] = Area * twelfth * (2 * fe1 + fe2 + fe3) + sixth * tau * fs * (Be_x * y23 + Be_y * x32);
        load[1] = Area * twelfth * (fe1 + 2 * fe2 + fe3) + sixth * tau * fs * (Be_x * y31 + Be_y * x13);
        load[2] = Area * twelfth * (fe1 + fe2 + 2 * fe3) + sixth * tau * fs * (Be_x * y12 + Be_y * x21);
        double *t = Tr + 12L * E;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            t[4 * i] = load[i];
#pragma unroll
            for (int j = 0; j < 3; j++) t[4 * i + 1 + j] = me[3 * i + j] * due[j] + ke[3 * i + j] * ue[j];
        }
    }
}

// R[g] = sum over g's incidences (e, i) in element order of
// ((+load) - t0) - t1 - t2, from zero (Build_M_K_R_SUPG.c:39, :162-171)
__global__ void k_gather_R(int neq, const int *__restrict__ off, const int *__restrict__ inc,
                           const double *__restrict__ Tr, double *__restrict__ R)
{
    for (int g = blockIdx.x * FG_TPB + threadIdx.x; g <= neq; g += gridDim.x * FG_TPB) {
        double s = 0.0;
        if (g < neq)
            for (int q = off[g]; q < off[g + 1]; q++) {
                const double *t = Tr + 4L * inc[q];
                s += t[0];
                s -= t[1];
                s -= t[2];
                s -= t[3];
            }
        R[g] = s;
    }
}

// the same sums with the incidences as sliced rows: one equation per lane,
// its incidences' four terms (32 bytes) gathered together, folded in the
// order above
typedef double fg_d4 __attribute__((ext_vector_type(4)));
struct ResidualRow {
    using T = double;
    struct Rec {
        int c;
        fg_d4 t;
    };
    const int *__restrict__ inc;
    const fg_d4 *__restrict__ Tr4;
    __device__ __forceinline__ Rec key(long long q) const
    {
        Rec r;
        r.c = __builtin_nontemporal_load(&inc[q]);
        r.t = fg_d4{0.0, 0.0, 0.0, 0.0};
        return r;
    }
    __device__ __forceinline__ void stream(Rec &, long long) const {}
    __device__ __forceinline__ void gather(Rec &r) const { r.t = Tr4[r.c < 0 ? 0 : r.c]; }
    __device__ __forceinline__ double add(double acc, const Rec &r) const
    {
        double a = acc + r.t.x;
        a = a - r.t.y;
        a = a - r.t.z;
        a = a - r.t.w;
        return r.c < 0 ? acc : a;
    }
};

__global__ __launch_bounds__(FG_TPB) void k_gather_R_sliced(int nslice, const int *__restrict__ soff,
                                                            const int *__restrict__ sinc,
                                                            const double *__restrict__ Tr,
                                                            double *__restrict__ R, int neq)
{
    constexpr int SPW = FG_TPB / FG_SELL_W;
    const int s = fg_xcd_remap(blockIdx.x, (nslice + SPW - 1) / SPW) * SPW +
                  (int)(threadIdx.x / FG_SELL_W);
    const int lane = (int)(threadIdx.x % FG_SELL_W);
    if (s < nslice) {
        const int o0 = soff[s];
        const ResidualRow row{sinc, reinterpret_cast<const fg_d4 *>(Tr)};
        const double acc = fg_sell_row(row, o0, soff[s + 1] - o0, lane);
        const int r = s * FG_SELL_W + lane;
        if (r < neq) R[r] = acc;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) R[neq] = 0.0;
}

static int gather_R(fg_mesh2d *m, double *R)
{
    fg_ctx *c = m->ctx;
    if (m->nslice == 0) {
        FG_LAUNCH(k_gather_R, 1, FG_TPB, 0, c->stream, m->neq, m->d_finc_off, m->d_finc,
                  m->d_Tr, R);
        return FG_OK;
    }
    constexpr int SPW = FG_TPB / FG_SELL_W;
    FG_LAUNCH(k_gather_R_sliced, (m->nslice + SPW - 1) / SPW, FG_TPB, 0, c->stream, m->nslice,
              m->d_isoff, m->d_isinc, m->d_Tr, R, m->neq);
    return FG_OK;
}

extern "C" int fg_supg2d_transient_build(fg_mesh2d *m, const fg_supg2d_problem *pb,
                                         double alpha, double dt, const double *U,
                                         const double *dU, double *A, double *R)
{
    FG_ARG(m && pb && U && dU && A && R, "mesh/problem/U/dU/A/R");
    FG_ARG(pb->kind >= FG_P_CONST && pb->kind <= FG_P_ROTATE_ABOUT, "problem kind");
    fg_ctx *c = m->ctx;
    if (!m->d_Tr) FG_HIP(hipMalloc(&m->d_Tr, sizeof(double) * 12 * (size_t)(m->nel ? m->nel : 1)));
    ElemSink out{A, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (m->nel > 0)
        FG_LAUNCH(k_supg2d_tr, fg_vec_grid(c, m->nel), FG_TPB, 0, c->stream, m->nel, m->d_x,
                  m->d_y, m->d_vertex, U, dU, *pb, alpha, dt, out, m->d_Tr);
    return gather_R(m, R);
}

// the same build with the element matrices written straight into the k = 3
// EBE operator of the mesh's lm (its element e is the mesh's element e):
// the time loop's correction, without the AoS array and its scatter pass
int fg_supg2d_transient_build_op(fg_mesh2d *m, const fg_supg2d_problem *pb, double alpha,
                                 double dt, const double *U, const double *dU, fg_op *op,
                                 double *R)
{
    FG_ARG(m && pb && U && dU && op && R, "mesh/problem/U/dU/op/R");
    FG_ARG(op->kind == FG_OP_EBE && op->k == 3 && op->nitems == m->nel && op->neq == m->neq,
           "the k = 3 EBE operator of the mesh");
    fg_ctx *c = m->ctx;
    if (!m->d_Tr) FG_HIP(hipMalloc(&m->d_Tr, sizeof(double) * 12 * (size_t)(m->nel ? m->nel : 1)));
    ElemSink out{nullptr, op->d_kpos, op->d_kstride, op->d_K64, op->d_K32, op->d_rvpos,
                 op->d_rwst, op->d_sv64, op->d_sv32};
    if (m->nel > 0)
        FG_LAUNCH(k_supg2d_tr, fg_vec_grid(c, m->nel), FG_TPB, 0, c->stream, m->nel, m->d_x,
                  m->d_y, m->d_vertex, U, dU, *pb, alpha, dt, out, m->d_Tr);
    // the rows, when the operator has them, were written with the values
    op->rows_stale = op->d_rvpos ? 0 : 1;
    return gather_R(m, R);
}

// ---------------------------------------------------------------------------
// CSR / EDE assembly from element matrices (gathers in element order)
// ---------------------------------------------------------------------------
// target t receives sum over its list of (half ? 0.5 : 1) * ke[idx]
__global__ void k_gather_targets(int nt, const int *__restrict__ off,
                                 const long long *__restrict__ src, const double *__restrict__ ke,
                                 double *__restrict__ out)
{
    for (int t = blockIdx.x * FG_TPB + threadIdx.x; t < nt; t += gridDim.x * FG_TPB) {
        double s = 0;
        for (int q = off[t]; q < off[t + 1]; q++) {
            long long v = src[q];
            double a = ke[v >> 1];
            s += (v & 1) ? 0.5 * a : a;
        }
        out[t] = s;
    }
}

// contributions (target, ke index, half) in the reference's sequence order ->
// per-target lists (stable counting sort)
static int upload_targets(int nt, const std::vector<int> &tg, const std::vector<long long> &src,
                          int **d_off, long long **d_src)
{
    std::vector<int> off((size_t)nt + 1, 0);
    for (int t : tg) off[t + 1]++;
    for (int t = 0; t < nt; t++) off[t + 1] += off[t];
    std::vector<long long> lst(off[nt]);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (size_t q = 0; q < tg.size(); q++) lst[fill[tg[q]]++] = src[q];
    FG_CHECK(up(off.data(), off.size(), d_off));
    FG_CHECK(up(lst.data(), lst.size(), d_src));
    return FG_OK;
}

extern "C" int fg_op_assemble_csr(fg_op *op, int nel, const int *cbe, const double *ke)
{
    return fg_op_assemble_csr_k(op, nel, 3, cbe, ke);
}

extern "C" int fg_op_assemble_csr_k(fg_op *op, int nel, int k, const int *cbe, const double *ke)
{
    FG_ARG(op && op->kind == FG_OP_CSR && cbe && ke && k >= 1 && k <= 16,
           "CSR operator / map / ke / k");
    fg_ctx *c = op->ctx;
    int nnz = (int)op->nnz;
    const int kk = k * k;
    std::vector<int> tg;
    std::vector<long long> src;
    tg.reserve((size_t)nel * kk);
    src.reserve((size_t)nel * kk);
    for (long e = 0; e < nel; e++)
        for (int q = 0; q < kk; q++) {
            int s = cbe[(long)kk * e + q];
            FG_ARG(s >= 0 && s <= nnz, "csr_by_elem entry out of range");
            if (s < nnz) { tg.push_back(s); src.push_back(((long long)kk * e + q) << 1); }
        }
    int *d_off = nullptr;
    long long *d_src = nullptr;
    FG_CHECK(upload_targets(nnz, tg, src, &d_off, &d_src));
    double *AA = nullptr;
    FG_HIP(hipMalloc(&AA, sizeof(double) * (nnz ? nnz : 1)));
    if (nnz) FG_LAUNCH(k_gather_targets, fg_vec_grid(c, nnz), FG_TPB, 0, c->stream, nnz, d_off,
                       d_src, ke, AA);
    int rc = fg_op_set_values_dev(op, AA);
    FG_HIP(hipStreamSynchronize(c->stream));
    hipFree(AA); hipFree(d_off); hipFree(d_src);
    return rc;
}

extern "C" int fg_op_assemble_ede(fg_op *op, int nel, const int *order, const int *ebe,
                                  const double *ke)
{
    FG_ARG(op && op->kind == FG_OP_EDE && order && ebe && ke, "EDE operator / maps / ke");
    fg_ctx *c = op->ctx;
    int nedge = op->nitems, nt = 4 * nedge;
    std::vector<int> tg;
    std::vector<long long> src;
    tg.reserve((size_t)nel * 12);
    src.reserve((size_t)nel * 12);
    // ede_assembly.c:15-26: per element the 12 updates in this sequence
    for (long e = 0; e < nel; e++) {
        const int *o = order + 3 * e;
        int I[3] = {o[0], o[1], o[2]};
        int J[3] = {ebe[3 * e], ebe[3 * e + 1], ebe[3 * e + 2]};
        const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};   // (I1,I2), (I1,I3), (I2,I3)
        for (int m = 0; m < 3; m++) {
            int a = I[pa[m]], b = I[pb[m]];
            long long base = 9LL * e;
            long long s4[4] = {((base + 3 * a + a) << 1) | 1, (base + 3 * a + b) << 1,
                               (base + 3 * b + a) << 1, ((base + 3 * b + b) << 1) | 1};
            FG_ARG(J[m] >= 0 && J[m] <= nedge, "edge_by_elem entry out of range");
            if (J[m] == nedge) continue;   // the sink edge row, zeroed by the reference
            for (int q = 0; q < 4; q++) { tg.push_back(4 * J[m] + q); src.push_back(s4[q]); }
        }
    }
    int *d_off = nullptr;
    long long *d_src = nullptr;
    FG_CHECK(upload_targets(nt, tg, src, &d_off, &d_src));
    double *A = nullptr;
    FG_HIP(hipMalloc(&A, sizeof(double) * (nt ? nt : 1)));
    if (nt) FG_LAUNCH(k_gather_targets, fg_vec_grid(c, nt), FG_TPB, 0, c->stream, nt, d_off,
                      d_src, ke, A);
    int rc = fg_op_set_values_dev(op, A);
    FG_HIP(hipStreamSynchronize(c->stream));
    hipFree(A); hipFree(d_off); hipFree(d_src);
    return rc;
}
