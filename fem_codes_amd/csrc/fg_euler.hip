// fg_euler.hip -- SUPG element matrices of the 2D compressible Euler
// equations (4 conserved variables per node, 12x12 element matrices).
//
// Reference: 03_Euler_Equations/2D/01_CommonFiles/Build_M_K_F_SUPG.c:6-1073
// with the CAU discontinuity-capturing operator (Delta_CAU.c:4-78), the flux
// Jacobians (Ax_Ay_calculations.c:3-95, dimensionless or dimensional) and
// the slip-wall rotation (rotation.c:3-45).
//
// One thread per element.  The element's barycentric state, Jacobians,
// A0^{-1} metric, CAU coefficient and SUPG tau are formed once; then the
// 12x12 matrices are produced one node block row a at a time: the SUPG
// coefficients a_a = y_a [AxAx | AxAy] + x_a [AyAx | AyAy] of that row
// (products formed on the fly, k ascending), then each 4x4 block (a, b) of
// Me and Ke entry by entry, written straight into Ae = Me + alpha dt Ke,
// while Me dU and Ke U accumulate in column order for the residual.  The
// expressions keep the reference's operation order; the reference's
// sparsity-specialised products equal the k-ascending sums because the
// omitted terms are exact zeros and ones of the Jacobians.  pow(v, 2) is
// evaluated as v*v (correctly rounded), which the host's pow matches except
// in the last bit of rare arguments.  The residual terms go to a scratch
// array and a per-equation gather adds them in element order.
#include <math.h>

#include <vector>

#include "fg_internal.h"

struct fg_euler2d {
    fg_ctx *ctx = nullptr;
    int nnodes = 0, nel = 0, neq = 0;
    double *d_x = nullptr, *d_y = nullptr;
    int *d_vertex = nullptr, *d_roff = nullptr, *d_rinc = nullptr;
    int *d_id4 = nullptr;     // [nnodes*4] equation of each dof, neq: prescribed
    double *d_Re = nullptr;   // [nel*12]
};

namespace {

template <typename T>
int up(const T *h, size_t n, T **d)
{
    FG_HIP(hipMalloc(d, sizeof(T) * (n ? n : 1)));
    if (n) FG_HIP(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
    return FG_OK;
}

__device__ __forceinline__ void eu_jac(double gamma, double Mach, int dimless, const double Ub[4],
                                       double Ax[4][4], double Ay[4][4])
{
    const double v1 = Ub[1] / Ub[0], v2 = Ub[2] / Ub[0], E = Ub[3] / Ub[0];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) { Ax[i][j] = 0.0; Ay[i][j] = 0.0; }
    Ax[0][1] = 1.0;
    Ay[0][2] = 1.0;
    if (dimless) {
        const double nv = v1 * v1 + v2 * v2, e2 = Mach * Mach;
        Ax[1][0] = -(v1 * v1) + ((gamma - 1) * nv * 0.5);
        Ax[1][1] = (3.0 - gamma) * v1;
        Ax[1][2] = -(gamma - 1.0) * v2;
        Ax[1][3] = (gamma - 1.0) / e2;
        Ax[2][0] = -v1 * v2;
        Ax[2][1] = v2;
        Ax[2][2] = v1;
        Ax[3][0] = -v1 * (gamma * E - ((gamma - 1.0) * e2 * nv));
        Ax[3][1] = gamma * E - (gamma - 1.0) * e2 * (0.5 * nv + v1 * v1);
        Ax[3][2] = -(gamma - 1.0) * e2 * v1 * v2;
        Ax[3][3] = gamma * v1;
        Ay[2][0] = -(v2 * v2) + (((gamma - 1.0) * nv * 0.5));
        Ay[3][0] = -v2 * (gamma * E - (gamma - 1.0) * e2 * nv);
        Ay[3][2] = gamma * E - (gamma - 1.0) * e2 * (0.5 * nv + v2 * v2);
    } else {
        const double rho = Ub[0], nU = (Ub[1] * Ub[1]) + (Ub[2] * Ub[2]);
        Ax[1][0] = -(v1 * v1) + (((gamma - 1) * nU * 0.5) / (rho * rho));
        Ax[1][1] = (3.0 - gamma) * v1;
        Ax[1][2] = -(gamma - 1.0) * v2;
        Ax[1][3] = gamma - 1.0;
        Ax[2][0] = -v1 * v2;
        Ax[2][1] = v2;
        Ax[2][2] = v1;
        Ax[3][0] = (-gamma * v1 * E) + ((gamma - 1.0) * v1 * (nU / (rho * rho)));
        Ax[3][1] = (gamma * E) - ((gamma - 1.0) * 0.5 * ((nU / (rho * rho)) + (2.0 * v1 * v1)));
        Ax[3][2] = -((gamma - 1.0) * v1 * v2);
        Ax[3][3] = gamma * v1;
        Ay[2][0] = -(v2 * v2) + (((gamma - 1.0) * nU * 0.5) / (rho * rho));
        Ay[3][0] = (-gamma * v2 * E) + (((gamma - 1.0) * v2 * nU) / (rho * rho));
        Ay[3][2] = (gamma * E) - ((gamma - 1.0) * 0.5 * ((nU / (rho * rho)) + (2.0 * v2 * v2)));
    }
    Ay[1][0] = Ax[2][0];
    Ay[1][1] = v2;
    Ay[1][2] = v1;
    Ay[2][1] = -(gamma - 1.0) * v1;
    Ay[2][2] = (3.0 - gamma) * v2;
    Ay[2][3] = Ax[1][3];
    Ay[3][1] = Ax[3][2];
    Ay[3][3] = gamma * v2;
}

// x^T A0 x as Delta_CAU.c:35-44 groups it (upper triangle, doubled off-diagonals)
__device__ __forceinline__ double eu_quad(const double A0[4][4], const double v[4])
{
    return v[0] * (A0[0][0] * v[0] + 2.0 * A0[0][1] * v[1] + 2.0 * A0[0][2] * v[2] + 2.0 * A0[0][3] * v[3]) +
           v[1] * (A0[1][1] * v[1] + 2.0 * A0[1][2] * v[2] + 2.0 * A0[1][3] * v[3]) +
           v[2] * (A0[2][2] * v[2] + 2.0 * A0[2][3] * v[3]) +
           v[3] * A0[3][3] * v[3];
}

// product entry (L R)_{ij}, k ascending
__device__ __forceinline__ double eu_mm(const double L[4][4], const double R[4][4], int i, int j)
{
    double s = L[i][0] * R[0][j];
#pragma unroll
    for (int k = 1; k < 4; k++) s = s + L[i][k] * R[k][j];
    return s;
}

// C_b[i][j] = (y_b Ax + x_b Ay)_{ij} / 6 with the reference's direct forms
// for row 0 and the structural zeros (Build_M_K_F_SUPG.c:273-332)
__device__ __forceinline__ double eu_C(const double Ax[4][4], const double Ay[4][4], double yb,
                                       double xb, int i, int j)
{
    const double sixth = 1.0 / 6.0;
    if (i == 0) return j == 1 ? yb * sixth : j == 2 ? xb * sixth : 0.0;
    if (i == 1 && j == 3) return yb * Ax[1][3] * sixth;
    if (i == 2 && j == 3) return xb * Ay[2][3] * sixth;
    return (yb * Ax[i][j] + xb * Ay[i][j]) * sixth;
}

struct EuParams {
    double gamma, mach, alpha, dt, tol;
    int dimless;
};

__global__ __launch_bounds__(FG_TPB) void k_euler_build(
    int nel, const double *__restrict__ x, const double *__restrict__ y,
    const int *__restrict__ vertex, const double *__restrict__ U, const double *__restrict__ dU,
    const int *__restrict__ rot, const double *__restrict__ theta, EuParams prm,
    double *__restrict__ A, double *__restrict__ Re)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    const double gamma = prm.gamma;
    const long long nt = 3LL * nel;
    for (long long t = (long long)blockIdx.x * FG_TPB + threadIdx.x; t < nt;
         t += (long long)gridDim.x * FG_TPB) {
        const int e = (int)(t / 3), a = (int)(t - 3LL * e);
        const int Jn[3] = {vertex[3 * e], vertex[3 * e + 1], vertex[3 * e + 2]};
        const double x1 = x[Jn[0]], x2 = x[Jn[1]], x3 = x[Jn[2]];
        const double y1 = y[Jn[0]], y2 = y[Jn[1]], y3 = y[Jn[2]];
        const double y23 = y2 - y3, y31 = y3 - y1, y12 = y1 - y2;
        const double x32 = x3 - x2, x13 = x1 - x3, x21 = x2 - x1;
        const double twoArea = fabs(x21 * y31 - x13 * y12);
        const double Area = 0.5 * twoArea;
        double Ue[12], dUe[12], Ub[4], dUb[4], gx[4], gy[4];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int d = 0; d < 4; d++) {
                Ue[4 * a + d] = U[4 * Jn[a] + d];
                dUe[4 * a + d] = dU[4 * Jn[a] + d];
            }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            Ub[i] = (Ue[i] + Ue[i + 4] + Ue[i + 8]) * third;
            dUb[i] = (dUe[i] + dUe[i + 4] + dUe[i + 8]) * third;
            gx[i] = (Ue[i] * y23 + Ue[i + 4] * y31 + Ue[i + 8] * y12) / twoArea;
            gy[i] = (Ue[i] * x32 + Ue[i + 4] * x13 + Ue[i + 8] * x21) / twoArea;
        }
        const double v1 = Ub[1] / Ub[0], v2 = Ub[2] / Ub[0], E = Ub[3] / Ub[0];
        const double nU23 = (Ub[1] * Ub[1]) + (Ub[2] * Ub[2]);
        double Ax[4][4], Ay[4][4], A0[4][4];
        eu_jac(gamma, prm.mach, prm.dimless, Ub, Ax, Ay);
        // A0^{-1} (:131-159)
        const double rho_i = Ub[3] - ((nU23 * 0.5) / Ub[0]);
        const double V2 = Ub[1] / rho_i, V3 = Ub[2] / rho_i, V4 = (-Ub[0]) / rho_i;
        const double re = -1.0 / (rho_i * V4);
        const double k1 = (V2 * V2 + V3 * V3) * 0.5 / V4;
        A0[0][0] = re * (k1 * k1 + gamma);
        A0[0][1] = re * (k1 * V2);
        A0[0][2] = re * (k1 * V3);
        A0[0][3] = re * ((k1 + 1.0) * V4);
        A0[1][1] = re * (V2 * V2 - V4);
        A0[1][2] = re * (V2 * V3);
        A0[1][3] = re * (V2 * V4);
        A0[2][2] = re * (V3 * V3 - V4);
        A0[2][3] = re * (V3 * V4);
        A0[3][3] = re * V4 * V4;
        A0[1][0] = A0[0][1]; A0[2][0] = A0[0][2]; A0[2][1] = A0[1][2];
        A0[3][0] = A0[0][3]; A0[3][1] = A0[1][3]; A0[3][2] = A0[2][3];
        // CAU (Delta_CAU.c)
        double delta;
        {
            double kx[4], ky[4], L[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                double sx = Ax[i][0] * gx[0], sy = Ay[i][0] * gy[0];
#pragma unroll
                for (int k = 1; k < 4; k++) { sx = sx + Ax[i][k] * gx[k]; sy = sy + Ay[i][k] * gy[k]; }
                L[i] = dUb[i] + sx + sy;
                kx[i] = (y23 * gx[i] + x32 * gy[i]) / twoArea;
                ky[i] = (y31 * gx[i] + x13 * gy[i]) / twoArea;
            }
            const double ngk = sqrt(eu_quad(A0, kx) + eu_quad(A0, ky));
            const double nL = sqrt(eu_quad(A0, L));
            const double ngu = sqrt(eu_quad(A0, gx) + eu_quad(A0, gy));
            delta = ngu >= prm.tol ? nL / ngk : 0.0;
        }
        // SUPG tau (:243-269)
        double betax = 2.0 * (gx[0] * Ub[0] + gx[1] * Ub[1] + gx[2] * Ub[2] + gx[3] * Ub[3]);
        double betay = 2.0 * (gy[0] * Ub[0] + gy[1] * Ub[1] + gy[2] * Ub[2] + gy[3] * Ub[3]);
        const double betaxy = sqrt(betax * betax + betay * betay);
        if (betaxy != 0.0) { betax = betax / betaxy; betay = betay / betaxy; }
        const double h = sqrt(twoArea);
        const double avb = fabs(v1 * betax + v2 * betay);
        const double c = sqrt(gamma * (gamma - 1.0) * (E - 0.5 * (v1 * v1 + v2 * v2)));
        const double CFL = ((c + avb) * prm.dt) / h;
        const double psi = (2.0 * prm.alpha * CFL) * (1.0 / (1.0 + 2.0 * prm.alpha * CFL));
        const double tau_d = delta / ((c + avb) * (c + avb));
        const double tau_a = h / (2.0 * (c + avb));
        const double tau_t = (2.0 / (3.0 * (1.0 + 2.0 * prm.alpha * CFL))) * tau_a;
        double tau = tau_t + psi * (tau_a - tau_d);
        if (tau < 0.0) tau = 0.0;

        const double Mg1 = Area * sixth, Mg2 = Area / 12.0;
        const double t4a = tau / (4 * Area);
        const double d4a = delta / (4.0 * Area);
        const double Kc01 = d4a * (y23 * y31 + x32 * x13);
        const double Kc02 = d4a * (y23 * y12 + x32 * x21);
        const double Kc12 = d4a * (y31 * y12 + x13 * x21);
        const double adt = prm.alpha * prm.dt;
        double *Ael = A + 144L * e;
        {
            // SUPG coefficients of node row a (:618-737) from AxAx, AyAx,
            // AxAy, AyAy (:161-237)
            double axa[4][4], aya[4][4];
            const double yA = a == 0 ? y23 : a == 1 ? y31 : y12;
            const double xA = a == 0 ? x32 : a == 1 ? x13 : x21;
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    axa[i][j] = yA * eu_mm(Ax, Ax, i, j) + xA * eu_mm(Ay, Ax, i, j);
                    aya[i][j] = yA * eu_mm(Ax, Ay, i, j) + xA * eu_mm(Ay, Ay, i, j);
                }
            const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
            const double yB1 = b1 == 0 ? y23 : y31, xB1 = b1 == 0 ? x32 : x13;
            const double yB2 = b2 == 1 ? y31 : y12, xB2 = b2 == 1 ? x13 : x21;
            const double kcb1 = a == 0 ? Kc01 : a == 1 ? Kc01 : Kc02;
            const double kcb2 = a == 0 ? Kc02 : Kc12;
            double md[4] = {0, 0, 0, 0}, ku[4] = {0, 0, 0, 0};
            const bool rt = rot && rot[Jn[a]];
            const double cs = rt ? cos(theta[Jn[a]]) : 1.0, sn = rt ? sin(theta[Jn[a]]) : 0.0;
            for (int b = 0; b < 3; b++) {
                const double yB = b == 0 ? y23 : b == 1 ? y31 : y12;
                const double xB = b == 0 ? x32 : b == 1 ? x13 : x21;
                const double kcab = (a + b == 1) ? Kc01 : (a + b == 2) ? Kc02 : Kc12;
                double M[4][4];
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const double mpg = eu_C(Ax, Ay, yA, xA, i, j) * tau;
                        const double me = i == j ? (a == b ? Mg1 : Mg2) + mpg : mpg;
                        double ke;
                        if (a != b) {
                            const double kpg = t4a * (yB * axa[i][j] + xB * aya[i][j]);
                            ke = eu_C(Ax, Ay, yB, xB, i, j) + kpg;
                            if (i == j) ke = ke + kcab;
                        } else {
                            const double p1 = t4a * (yB1 * axa[i][j] + xB1 * aya[i][j]);
                            const double p2 = t4a * (yB2 * axa[i][j] + xB2 * aya[i][j]);
                            ke = eu_C(Ax, Ay, yA, xA, i, j) - (p1 + p2);
                            if (i == j) ke = ke - (kcb1 + kcb2);
                        }
                        md[i] += me * dUe[4 * b + j];
                        ku[i] += ke * Ue[4 * b + j];
                        M[i][j] = me + adt * ke;
                    }
                if (rt) {   // rotation.c:13-37, block (a, b)
                    const double bb = M[0][1], cc = M[0][2], ee = M[1][0], ff = M[1][1];
                    const double gg = M[1][2], hh = M[1][3], ii = M[2][0], jj = M[2][1];
                    const double kk = M[2][2], ll = M[2][3], nn = M[3][1], oo = M[3][2];
                    M[0][1] = bb * cs - cc * sn;
                    M[0][2] = bb * sn + cc * cs;
                    M[1][0] = ee * cs - ii * sn;
                    M[1][1] = cs * (ff * cs - jj * sn) - sn * (gg * cs - kk * sn);
                    M[1][2] = cs * (gg * cs - kk * sn) + sn * (ff * cs - jj * sn);
                    M[1][3] = hh * cs - ll * sn;
                    M[2][0] = ee * sn + ii * cs;
                    M[2][1] = cs * (ff * sn + jj * cs) - sn * (gg * sn + kk * cs);
                    M[2][2] = cs * (gg * sn + kk * cs) + sn * (ff * sn + jj * cs);
                    M[2][3] = hh * sn + ll * cs;
                    M[3][1] = nn * cs - oo * sn;
                    M[3][2] = nn * sn + oo * cs;
                }
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) Ael[12 * (4 * a + i) + 4 * b + j] = M[i][j];
            }
            double r[4];
#pragma unroll
            for (int i = 0; i < 4; i++) r[i] = -md[i] - ku[i];
            if (rt) {   // rotation.c:39-42
                const double bb = r[1], cc = r[2];
                r[1] = bb * cs - cc * sn;
                r[2] = bb * sn + cc * cs;
            }
#pragma unroll
            for (int i = 0; i < 4; i++) Re[12L * e + 4 * a + i] = r[i];
        }
    }
}

// R[g] = sum of the element residual entries of g in element order, from 0
__global__ void k_euler_gather(int neq, const int *__restrict__ off, const int *__restrict__ inc,
                               const double *__restrict__ Re, double *__restrict__ R)
{
    for (int g = blockIdx.x * FG_TPB + threadIdx.x; g <= neq; g += gridDim.x * FG_TPB) {
        double s = 0.0;
        if (g < neq)
            for (int q = off[g]; q < off[g + 1]; q++) s += Re[inc[q]];
        R[g] = s;
    }
}

}  // namespace

// eval_U_dU.c of the 03 driver: U, dU [nnodes*4] from u and du on the free
// dofs, the prescribed values and 0 on the others
__global__ void k_eval_U_dU_eu(int n4, int neq, const int *__restrict__ id4,
                               const double *__restrict__ u, const double *__restrict__ du,
                               const double *__restrict__ presc, double *__restrict__ U,
                               double *__restrict__ dU)
{
    for (int q = blockIdx.x * FG_TPB + threadIdx.x; q < n4; q += gridDim.x * FG_TPB) {
        const int id = id4[q];
        U[q] = id < neq ? u[id] : presc[q];
        dU[q] = id < neq ? du[id] : 0.0;
    }
}

int fg_euler2d_eval_U_dU(const fg_euler2d *m, const double *u, const double *du,
                         const double *presc, double *U, double *dU)
{
    fg_ctx *c = m->ctx;
    const int n4 = 4 * m->nnodes;
    FG_LAUNCH(k_eval_U_dU_eu, fg_vec_grid(c, n4), FG_TPB, 0, c->stream, n4, m->neq, m->d_id4,
              u, du, presc, U, dU);
    return FG_OK;
}

void fg_euler2d_sizes(const fg_euler2d *m, fg_ctx **c, int *nnodes, int *nel, int *neq)
{
    *c = m->ctx;
    *nnodes = m->nnodes;
    *nel = m->nel;
    *neq = m->neq;
}

extern "C" int fg_euler2d_create(fg_ctx *c, int nnodes, int nel, int neq, const double *x,
                                 const double *y, const int *vertex, const int *lm,
                                 fg_euler2d **out)
{
    FG_ARG(c && x && y && vertex && lm && out, "ctx/x/y/vertex/lm/out");
    FG_ARG(nnodes > 0 && nel >= 0 && neq >= 0, "sizes");
    FG_ARG((long long)nel * 12 < (1LL << 31), "too many elements");
    for (long long t = 0; t < 3LL * nel; t++)
        FG_ARG(vertex[t] >= 0 && vertex[t] < nnodes, "vertex out of range");
    std::vector<int> off(neq + 1, 0);
    for (long long t = 0; t < 12LL * nel; t++) {
        FG_ARG(lm[t] >= 0 && lm[t] <= neq, "lm entry outside [0, neq]");
        if (lm[t] < neq) off[lm[t] + 1]++;
    }
    for (int g = 0; g < neq; g++) off[g + 1] += off[g];
    std::vector<int> inc(off[neq]);
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (long long t = 0; t < 12LL * nel; t++)
            if (lm[t] < neq) inc[fill[lm[t]]++] = (int)t;
    }
    // equation of every node dof, from the elements (Node[].id; a node in no
    // element, or a prescribed dof, reads neq)
    std::vector<int> id4(4 * (size_t)nnodes, neq);
    for (long long e = 0; e < nel; e++)
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 4; k++) id4[4 * (size_t)vertex[3 * e + a] + k] = lm[12 * e + 4 * a + k];
    FG_HIP(hipSetDevice(c->device));
    fg_euler2d *m = new fg_euler2d();
    m->ctx = c;
    m->nnodes = nnodes;
    m->nel = nel;
    m->neq = neq;
    int rc = FG_OK;
    if ((rc = up(x, nnodes, &m->d_x)) || (rc = up(y, nnodes, &m->d_y)) ||
        (rc = up(vertex, 3 * (size_t)nel, &m->d_vertex)) ||
        (rc = up(off.data(), off.size(), &m->d_roff)) ||
        (rc = up(inc.data(), inc.size(), &m->d_rinc)) ||
        (rc = up(id4.data(), id4.size(), &m->d_id4))) {
        fg_euler2d_destroy(m);
        return rc;
    }
    if (hipMalloc(&m->d_Re, sizeof(double) * 12 * (size_t)(nel ? nel : 1)) != hipSuccess) {
        fg_euler2d_destroy(m);
        fg_set_error("Euler mesh: out of device memory");
        return FG_ERR_NOMEM;
    }
    *out = m;
    return FG_OK;
}

extern "C" int fg_euler2d_destroy(fg_euler2d *m)
{
    if (!m) return FG_OK;
    hipStreamSynchronize(m->ctx->stream);
    hipFree(m->d_x); hipFree(m->d_y); hipFree(m->d_vertex); hipFree(m->d_roff);
    hipFree(m->d_rinc); hipFree(m->d_Re); hipFree(m->d_id4);
    delete m;
    return FG_OK;
}

extern "C" int fg_euler2d_build(fg_euler2d *m, const fg_euler2d_params *p, const double *U,
                                const double *dU, const int *rot, const double *theta,
                                double *A, double *R)
{
    FG_ARG(m && p && U && dU && A && R, "mesh/params/U/dU/A/R");
    FG_ARG(!rot || theta, "rot needs theta");
    fg_ctx *c = m->ctx;
    EuParams prm{p->gamma, p->mach, p->alpha, p->dt, p->tolerance, p->dimensionless ? 1 : 0};
    if (m->nel > 0)
        FG_LAUNCH(k_euler_build, fg_vec_grid(c, 3LL * m->nel), FG_TPB, 0, c->stream, m->nel, m->d_x,
                  m->d_y, m->d_vertex, U, dU, rot, theta, prm, A, m->d_Re);
    FG_LAUNCH(k_euler_gather, fg_vec_grid(c, m->neq + 1), FG_TPB, 0, c->stream, m->neq,
              m->d_roff, m->d_rinc, m->d_Re, R);
    return FG_OK;
}
