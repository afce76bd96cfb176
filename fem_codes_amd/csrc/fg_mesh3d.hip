// fg_mesh3d.hip -- 3D steady transport: device copy of a P1 tetrahedral mesh,
// SUPG element assembly, and the k-dof CSR structures.
//
// Reference: 01_SS_Transport_Equation/3D/01_CommonFiles/Build_SUPG.c:22-196,
// Element_Configuration.c:4-58, font_assembly.c:3-57, ebe_assembly.c:3-17,
// csr_Initialization.c:80-256, csr_assembly.c:3-30, Fill_ID.c, Fill_LM.c.
//
// One thread per element computes the 4x4 element matrix and load vector
// with the reference's expressions in the reference's order (the library is
// built without FMA contraction); F is then gathered per equation over its
// incident (element, vertex) pairs in element order, which is the order in
// which font_assembly.c adds them.  No atomics.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fg_internal.h"

struct fg_mesh3d {
    fg_ctx *ctx = nullptr;
    int nnodes = 0, nel = 0, neq = 0;
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
    int *d_vertex = nullptr;
    int *d_finc_off = nullptr, *d_finc = nullptr;   // equation -> e*4+r, element order
    double *d_Fe = nullptr;                          // [nel*4]
    std::vector<int> h_lm;
};

namespace {

template <typename T>
int up(const T *h, size_t n, T **d)
{
    FG_HIP(hipMalloc(d, sizeof(T) * (n ? n : 1)));
    if (n) FG_HIP(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
    return FG_OK;
}

constexpr double PI3 = 3.141592654;   // SSTransportEquation3D.h:15

// DIFCONV/DIFCONV_f.c
__device__ double difconv_f(double x, double y, double z, double A)
{
    double f;
    f = A * 3.0 * PI3 * PI3 * sin(PI3 * x) * sin(PI3 * y) * sin(PI3 * z) +
        PI3 * cos(PI3 * x) * sin(PI3 * y) * sin(PI3 * z) +
        PI3 * sin(PI3 * x) * cos(PI3 * y) * sin(PI3 * z) +
        PI3 * sin(PI3 * x) * sin(PI3 * y) * cos(PI3 * z);
    return f;
}

__global__ __launch_bounds__(FG_TPB) void k_supg3d(int nel, const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ z,
                                                   const int *__restrict__ vertex,
                                                   fg_supg3d_problem pb, double *__restrict__ ke,
                                                   double *__restrict__ Fe)
{
    const bool dc = pb.kind == FG_P3_DIFCONV;
    // DIFCONV_kappa.c / DIFCONV_beta.c / DIFCONV_sigma.c or constants
    const double kX = dc ? pb.A : pb.kappa, kY = kX, kZ = kX;
    const double bX = dc ? 1.0 : pb.bx, bY = dc ? 1.0 : pb.by, bZ = dc ? 1.0 : pb.bz;
    const double sigma = dc ? 0.0 : pb.sigma;
    for (int e = blockIdx.x * FG_TPB + threadIdx.x; e < nel; e += gridDim.x * FG_TPB) {
        double X[4], Y[4], Z[4], b[4], c[4], d[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int v = vertex[4 * e + i];
            X[i] = x[v]; Y[i] = y[v]; Z[i] = z[v];
        }
        // Element_Configuration.c:32-58
        const double a0 = (X[1]*Y[2]*Z[3] + X[3]*Y[1]*Z[2] + X[2]*Y[3]*Z[1]) -
                          (X[3]*Y[2]*Z[1] + X[1]*Y[3]*Z[2] + X[2]*Y[1]*Z[3]);
        b[0] = -(Y[1]-Y[3])*(Z[2]-Z[3]) + (Y[2]-Y[3])*(Z[1]-Z[3]);
        b[1] =  (Y[0]-Y[3])*(Z[2]-Z[3]) - (Y[2]-Y[3])*(Z[0]-Z[3]);
        b[2] = -(Y[0]-Y[3])*(Z[1]-Z[3]) + (Y[1]-Y[3])*(Z[0]-Z[3]);
        b[3] =  (Y[0]-Y[2])*(Z[1]-Z[2]) - (Y[1]-Y[2])*(Z[0]-Z[2]);
        c[0] =  (X[1]-X[3])*(Z[2]-Z[3]) - (X[2]-X[3])*(Z[1]-Z[3]);
        c[1] = -(X[0]-X[3])*(Z[2]-Z[3]) + (X[2]-X[3])*(Z[0]-Z[3]);
        c[2] =  (X[0]-X[3])*(Z[1]-Z[3]) - (X[1]-X[3])*(Z[0]-Z[3]);
        c[3] = -(X[0]-X[2])*(Z[1]-Z[2]) + (X[1]-X[2])*(Z[0]-Z[2]);
        d[0] = -(X[1]-X[3])*(Y[2]-Y[3]) + (X[2]-X[3])*(Y[1]-Y[3]);
        d[1] =  (X[0]-X[3])*(Y[2]-Y[3]) - (X[2]-X[3])*(Y[0]-Y[3]);
        d[2] = -(X[0]-X[3])*(Y[1]-Y[3]) + (X[1]-X[3])*(Y[0]-Y[3]);
        d[3] =  (X[0]-X[2])*(Y[1]-Y[2]) - (X[1]-X[2])*(Y[0]-Y[2]);
        const double Volume = (a0 + X[0]*b[0] + Y[0]*c[0] + Z[0]*d[0])/6.0;
        // Build_SUPG.c:64-172
        const double h = pow(Volume, 1.0/3.0);
        const double ConstD = 1.0/(36.0*Volume);
        const double normK = sqrt(kX*kX + kY*kY + kZ*kZ);
        const double ConstC = 1.0/24.0;
        const double normB = sqrt(bX*bX + bY*bY + bZ*bZ);
        double C[4];
#pragma unroll
        for (int r = 0; r < 4; r++) C[r] = bX*b[r] + bY*c[r] + bZ*d[r];
        const double ConstR = (sigma*Volume)/20.0;
        const double peclet = (normB*h)/(2.0*normK);
        const double xi = fmax(0.0, 1-1/peclet);
        const double tau = (h*xi)/(2.0*normB);
        const double ConstCsupg = tau/(36.0*Volume);
        double Cs[4][4];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int s = r + 1; s < 4; s++)
                Cs[r][s] = Cs[s][r] =
                    bX*bX*b[r]*b[s] + bX*bY*c[r]*b[s] + bX*bZ*d[r]*b[s] + bX*bY*b[r]*c[s] +
                    bY*bY*c[r]*c[s] + bY*bZ*d[r]*c[s] + bX*bZ*b[r]*d[s] + bY*bZ*c[r]*d[s] +
                    bZ*bZ*d[r]*d[s];
        Cs[0][0] = -(Cs[0][1] + Cs[0][2] + Cs[0][3]);
        Cs[1][1] = -(Cs[0][1] + Cs[1][2] + Cs[1][3]);
        Cs[2][2] = -(Cs[0][2] + Cs[1][2] + Cs[2][3]);
        Cs[3][3] = -(Cs[0][3] + Cs[1][3] + Cs[2][3]);
        const double ConstRsupg = (sigma*tau)/24.0;
        double *k16 = ke + 16L * e;
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const double E = kX*(b[r]*b[s]) + kY*(c[r]*c[s]) + kZ*(d[r]*d[s]);
                k16[4 * r + s] = ConstD*E + ConstC*C[s] + ConstCsupg*Cs[r][s] + ConstRsupg*C[r] +
                                 (r == s ? 2*ConstR : ConstR);
            }
        // Build_SUPG.c:182-195 and font_assembly.c:45-48 (prescribed values 0)
        double f[4];
#pragma unroll
        for (int i = 0; i < 4; i++) f[i] = dc ? difconv_f(X[i], Y[i], Z[i], pb.A) : pb.f;
        const double ConstF = Volume/20.0;
        const double Faux = f[0] + f[1] + f[2] + f[3];
        const double ConstFsupg = tau/24.0;
        double *f4 = Fe + 4L * e;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            double F0 = ConstF*(Faux + f[r]) + ConstFsupg*C[r]*Faux;
            f4[r] = F0 - (k16[4 * r] * 0.0 + k16[4 * r + 1] * 0.0 + k16[4 * r + 2] * 0.0 +
                          k16[4 * r + 3] * 0.0);
        }
    }
}

// F[g] = sum of Fe over g's incidences in element order
__global__ void k_gather_F4(int neq, const int *__restrict__ off, const int *__restrict__ inc,
                            const double *__restrict__ Fe, double *__restrict__ F)
{
    for (int g = blockIdx.x * FG_TPB + threadIdx.x; g < neq; g += gridDim.x * FG_TPB) {
        double s = 0;
        for (int q = off[g]; q < off[g + 1]; q++) s += Fe[inc[q]];
        F[g] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) F[neq] = 0.0;
}

}  // namespace

extern "C" int fg_mesh3d_create(fg_ctx *c, int nnodes, int nel, const double *x,
                                const double *y, const double *z, const int *utype,
                                const int *vertex, fg_mesh3d **out)
{
    FG_ARG(c && out && x && y && z && utype && vertex && nnodes > 0 && nel >= 0, "mesh3d args");
    for (long q = 0; q < 4L * nel; q++)
        FG_ARG(vertex[q] >= 0 && vertex[q] < nnodes, "vertex index out of range");
    FG_HIP(hipSetDevice(c->device));
    fg_mesh3d *m = new fg_mesh3d();
    m->ctx = c;
    m->nnodes = nnodes;
    m->nel = nel;
    // Fill_ID.c: uType 1 = unknown (the same rule as fg_number_nodes, type >= 1)
    std::vector<int> id(nnodes), type01(nnodes);
    for (int i = 0; i < nnodes; i++) type01[i] = utype[i] == 1 ? 1 : 0;
    m->neq = fg_number_nodes(nnodes, type01.data(), id.data());
    m->h_lm.resize((size_t)nel * 4);
    fg_fill_lm(m->neq, nel, 4, vertex, type01.data(), id.data(), m->h_lm.data());
    std::vector<int> off(m->neq + 1, 0), inc;
    for (long q = 0; q < 4L * nel; q++)
        if (m->h_lm[q] != m->neq) off[m->h_lm[q] + 1]++;
    for (int g = 0; g < m->neq; g++) off[g + 1] += off[g];
    inc.resize(off[m->neq]);
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (long q = 0; q < 4L * nel; q++)
            if (m->h_lm[q] != m->neq) inc[fill[m->h_lm[q]]++] = (int)q;
    }
    FG_CHECK(up(x, nnodes, &m->d_x));
    FG_CHECK(up(y, nnodes, &m->d_y));
    FG_CHECK(up(z, nnodes, &m->d_z));
    FG_CHECK(up(vertex, (size_t)nel * 4, &m->d_vertex));
    FG_CHECK(up(off.data(), off.size(), &m->d_finc_off));
    FG_CHECK(up(inc.data(), inc.size(), &m->d_finc));
    FG_HIP(hipMalloc(&m->d_Fe, sizeof(double) * 4 * (nel ? nel : 1)));
    *out = m;
    return FG_OK;
}

extern "C" int fg_mesh3d_info(const fg_mesh3d *m, int *neq, int *nel, int *nnodes)
{
    FG_ARG(m, "mesh");
    if (neq) *neq = m->neq;
    if (nel) *nel = m->nel;
    if (nnodes) *nnodes = m->nnodes;
    return FG_OK;
}

extern "C" int fg_mesh3d_get_lm(const fg_mesh3d *m, int *lm)
{
    FG_ARG(m && lm, "mesh/lm");
    memcpy(lm, m->h_lm.data(), sizeof(int) * m->h_lm.size());
    return FG_OK;
}

extern "C" int fg_mesh3d_destroy(fg_mesh3d *m)
{
    if (!m) return FG_OK;
    hipStreamSynchronize(m->ctx->stream);
    hipFree(m->d_x); hipFree(m->d_y); hipFree(m->d_z); hipFree(m->d_vertex);
    hipFree(m->d_finc_off); hipFree(m->d_finc); hipFree(m->d_Fe);
    delete m;
    return FG_OK;
}

extern "C" int fg_supg3d_build(fg_mesh3d *m, const fg_supg3d_problem *pb, double *ke,
                               double *F)
{
    FG_ARG(m && pb && ke, "mesh/problem/ke");
    FG_ARG(pb->kind == FG_P3_DIFCONV || pb->kind == FG_P3_CONST, "problem kind");
    fg_ctx *c = m->ctx;
    if (m->nel > 0)
        FG_LAUNCH(k_supg3d, fg_vec_grid(c, m->nel), FG_TPB, 0, c->stream, m->nel, m->d_x,
                  m->d_y, m->d_z, m->d_vertex, *pb, ke, m->d_Fe);
    if (F)
        FG_LAUNCH(k_gather_F4, fg_vec_grid(c, m->neq > 0 ? m->neq : 1), FG_TPB, 0, c->stream,
                  m->neq, m->d_finc_off, m->d_finc, m->d_Fe, F);
    return FG_OK;
}
