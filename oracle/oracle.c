/*
 * oracle.c -- CPU restatement of the FEM_CODES 2D solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libfemgpu, never the product.
 * Build: oracle/Makefile, flags -O2 -ffp-contract=off (order-faithful: no FMA
 * contraction, no reassociation), so that each routine reproduces the
 * reference compiled with the same flags bit for bit.
 *
 * Pinning (how each function is checked; tests/test_oracle_*.py):
 *   BLAS-1, csrmv, ebemv, edemv, Diag setups, Jacobi EBE, pgmres,
 *   mesh numbering, Fill_LM, csr/ede initialisation, SUPG build,
 *   ede/csr assembly:
 *     - bitwise against oracle/_ref (the reference sources compiled from
 *       the FEM_CODES tree by oracle/Makefile) on seeded meshes, and against
 *       the fixtures in tests/golden/ that _ref wrote;
 *     - independently of the reference: dense assembly of the three storage
 *       schemes must agree, row sums of ke vanish for gamma=0 and u=0,
 *       sum(Fe)=Area*f for constant f, GMRES against a dense LU solve, and
 *       GMRES on diag(1..k) converges in k steps.
 *   nnz / nedge closed forms for the structured mesh.
 *   ILU(0) apply, FP64: bitwise against the reference's fixtures and the
 *   identity (LU)_ij = A_ij on the pattern; FP32 (orc_ilu0_apply_f32): bit
 *   for bit the pinned FP64 apply on inputs whose arithmetic is exact in
 *   both precisions, and the componentwise backward-error bound of a
 *   triangular solve on real factors (tests/test_oracle_fp32.py).
 */
#include "oracle.h"
#include "dense_solve.h"
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* BLAS-1                                                                    */
/* ======================================================================== */

/* ddot.c:4-19 : blocks of five products, summed left to right inside the
 * block, then added to one running accumulator; scalar clean-up loop. */
double orc_ddot(int n, const double *x, const double *y)
{
    double acc = 0.0;
    long i = 0;
    long last5 = (long)n - 4;
    for (; i < last5; i += 5) {
        double blk = x[i] * y[i];
        blk = blk + x[i + 1] * y[i + 1];
        blk = blk + x[i + 2] * y[i + 2];
        blk = blk + x[i + 3] * y[i + 3];
        blk = blk + x[i + 4] * y[i + 4];
        acc = acc + blk;
    }
    for (; i < n; i++)
        acc = acc + x[i] * y[i];
    return acc;
}

/* daxpy.c:4-21 (the 4-way unroll does not change any result) */
void orc_daxpy(int n, double a, const double *x, double *y)
{
    for (long i = 0; i < n; i++)
        y[i] = y[i] + a * x[i];
}

/* dscal.c:4-21 */
void orc_dscal(int n, double a, double *x)
{
    for (long i = 0; i < n; i++)
        x[i] = a * x[i];
}

/* dcopy.c:4-19 */
void orc_dcopy(int n, const double *x, double *y)
{
    if (x == y) return;
    memmove(y, x, sizeof(double) * (size_t)(n > 0 ? n : 0));
}

/* dzero.c:3-18 */
void orc_dzero(int n, double *x)
{
    for (long i = 0; i < n; i++)
        x[i] = 0.0;
}

/* ======================================================================== */
/* matrix-vector products                                                    */
/* ======================================================================== */

/* CSR/csrmv.c:3-24 : Q zeroed, then each row accumulated in slot order. */
void orc_csrmv(int neq, const double *AA, const int *JA, const int *IA,
               const double *P, double *Q)
{
    for (int r = 0; r < neq; r++)
        Q[r] = 0.0;
    for (int r = 0; r < neq; r++)
        for (int j = IA[r]; j < IA[r + 1]; j++)
            Q[r] = Q[r] + AA[j] * P[JA[j]];
}

/* EBE/ebemv.c:4-43 and its k=4/12/16 twins: for each element in order, row
 * r of the element product is summed over columns left to right and added
 * to Q at the row's equation; the sink slot is reset after every element. */
void orc_ebemv(int neq, int nel, int k, const double *A, const int *lm,
               double *P, double *Q)
{
    double p[16];
    for (int r = 0; r < neq; r++)
        Q[r] = 0.0;
    P[neq] = 0.0;
    Q[neq] = 0.0;
    for (long e = 0; e < nel; e++) {
        const int *le = lm + e * k;
        const double *Ae = A + e * (long)k * k;
        for (int c = 0; c < k; c++)
            p[c] = P[le[c]];
        for (int r = 0; r < k; r++) {
            double row = Ae[r * k] * p[0];
            for (int c = 1; c < k; c++)
                row = row + Ae[r * k + c] * p[c];
            Q[le[r]] = Q[le[r]] + row;
        }
        Q[neq] = 0.0;
    }
}

/* EDE/edemv.c:4-39 */
void orc_edemv(int neq, int nedge, const double *A, const int *lm2,
               double *P, double *Q)
{
    for (int r = 0; r < neq; r++)
        Q[r] = 0.0;
    P[neq] = 0.0;
    Q[neq] = 0.0;
    for (long g = 0; g < nedge; g++) {
        int a = lm2[2 * g], b = lm2[2 * g + 1];
        double pa = P[a], pb = P[b];
        const double *Ag = A + 4 * g;
        Q[a] = Q[a] + (Ag[0] * pa + Ag[1] * pb);
        Q[b] = Q[b] + (Ag[2] * pa + Ag[3] * pb);
        Q[neq] = 0.0;
    }
}

/* ======================================================================== */
/* preconditioners                                                           */
/* ======================================================================== */

/* Diag_precond.c:3-14 */
void orc_diag_apply(int neq, const double *invDiag, const double *p, double *z)
{
    for (int r = 0; r < neq; r++)
        z[r] = p[r] * invDiag[r];
}

static void diag_finish(int neq, const double *Diag, double *invDiag, double *F)
{
    for (int r = 0; r < neq; r++)
        invDiag[r] = 1.0 / Diag[r];
    for (int r = 0; r < neq; r++)
        F[r] = F[r] * invDiag[r];
}

/* EBE/Diag_precond_EBE_setup.c:3-35 : element diagonals summed in element
 * order (stride k+1 walks the diagonal of the row-major element matrix). */
void orc_diag_setup_ebe(int neq, int nel, int k, const double *A, const int *lm,
                        double *Diag, double *invDiag, double *F)
{
    for (int r = 0; r < neq; r++)
        Diag[r] = 0.0;
    for (long e = 0; e < nel; e++) {
        for (int j = 0; j < k; j++) {
            int g = lm[e * k + j];
            Diag[g] = Diag[g] + A[e * (long)k * k + (long)j * (k + 1)];
        }
        Diag[neq] = 0.0;
    }
    diag_finish(neq, Diag, invDiag, F);
    invDiag[neq] = 1.0;                       /* EBE variant only, :29 */
}

/* EDE/Diag_precond_EDE_setup.c:3-33 (size = 2*NDOF = 2, stride 3) */
void orc_diag_setup_ede(int neq, int nedge, const double *A, const int *lm2,
                        double *Diag, double *invDiag, double *F)
{
    for (int r = 0; r < neq; r++)
        Diag[r] = 0.0;
    for (long g = 0; g < nedge; g++) {
        Diag[lm2[2 * g]] = Diag[lm2[2 * g]] + A[4 * g];
        Diag[lm2[2 * g + 1]] = Diag[lm2[2 * g + 1]] + A[4 * g + 3];
        Diag[neq] = 0.0;
    }
    diag_finish(neq, Diag, invDiag, F);
}

/* CSR/Diag_precond_CSR_setup.c:3-29 : Diag is not cleared first; the last
 * slot of the row whose column equals the row wins. */
void orc_diag_setup_csr(int neq, const double *AA, const int *JA, const int *IA,
                        double *Diag, double *invDiag, double *F)
{
    for (int r = 0; r < neq; r++)
        for (int j = IA[r]; j < IA[r + 1]; j++)
            if (JA[j] == r)
                Diag[r] = AA[j];
    diag_finish(neq, Diag, invDiag, F);
}

/* EBE/Jacobi_precond_EBE.c:8-47 */
void orc_jacobi_ebe_apply(int neq, int nel, const double *invDe, const int *lm,
                          const double *p, double *z)
{
    for (int r = 0; r < neq; r++)
        z[r] = 0.0;
    z[neq] = 0.0;
    for (long e = 0; e < nel; e++) {
        const int *le = lm + 3 * e;
        double z0 = p[le[0]] * invDe[3 * e + 0];
        double z1 = p[le[1]] * invDe[3 * e + 1];
        double z2 = p[le[2]] * invDe[3 * e + 2];
        z[le[0]] = z[le[0]] + z0;
        z[le[1]] = z[le[1]] + z1;
        z[le[2]] = z[le[2]] + z2;
        z[neq] = 0.0;
    }
}

/* EBE/Jacobi_precond_EBE_setup.c:4-43 */
void orc_jacobi_ebe_setup(int neq, int nel, const double *A, const int *lm,
                          double *invDe, double *F)
{
    for (long e = 0; e < nel; e++) {
        invDe[3 * e + 0] = 1.0 / A[9 * e + 0];
        invDe[3 * e + 1] = 1.0 / A[9 * e + 4];
        invDe[3 * e + 2] = 1.0 / A[9 * e + 8];
    }
    double *faux = (double *)calloc((size_t)neq + 1, sizeof(double));
    for (int r = 0; r < neq; r++)
        faux[r] = F[r];
    faux[neq] = 0.0;
    F[neq] = 0.0;
    orc_jacobi_ebe_apply(neq, nel, invDe, lm, faux, F);
    free(faux);
}

/* EBE/JacobiDOF4_precond_EBE.c:7-83 for k = 12 (and any k dofs per
 * element): z = 0, then per element every z_j = p[lm_j] invDe[e][j] first and
 * the k sums z[lm_j] += z_j in j order (:30-78); z[neq] = 0 */
void orc_jacobi_ebe_apply_k(int neq, int nel, int k, const double *invDe, const int *lm,
                            const double *p, double *z)
{
    double zj[16];
    for (int r = 0; r < neq; r++)
        z[r] = 0.0;
    z[neq] = 0.0;
    for (long e = 0; e < nel; e++) {
        const int *le = lm + (long)k * e;
        for (int j = 0; j < k; j++)
            zj[j] = p[le[j]] * invDe[(long)k * e + j];
        for (int j = 0; j < k; j++)
            z[le[j]] = z[le[j]] + zj[j];
    }
    z[neq] = 0.0;
}

/* EBE/JacobiDOF4_precond_EBE_setup.c:4-64: invDe[e][j] = 1 / A_e[j][j]
 * (A[I][13 j] for k = 12, :19-31), then F <- Jacobi(copy of F) (:52-61) */
void orc_jacobi_ebe_setup_k(int neq, int nel, int k, const double *A, const int *lm,
                            double *invDe, double *F)
{
    for (long e = 0; e < nel; e++)
        for (int j = 0; j < k; j++)
            invDe[(long)k * e + j] = 1.0 / A[(long)k * k * e + (long)j * (k + 1)];
    double *faux = (double *)calloc((size_t)neq + 1, sizeof(double));
    for (int r = 0; r < neq; r++)
        faux[r] = F[r];
    faux[neq] = 0.0;
    F[neq] = 0.0;
    orc_jacobi_ebe_apply_k(neq, nel, k, invDe, lm, faux, F);
    free(faux);
}

/* ======================================================================== */
/* operator callbacks                                                        */
/* ======================================================================== */
void orc_op_csr(void *ctx, double *x, double *y)
{
    orc_csr_ctx *c = (orc_csr_ctx *)ctx;
    orc_csrmv(c->neq, c->AA, c->JA, c->IA, x, y);
}
void orc_op_ebe(void *ctx, double *x, double *y)
{
    orc_ebe_ctx *c = (orc_ebe_ctx *)ctx;
    orc_ebemv(c->neq, c->nel, c->k, c->A, c->lm, x, y);
}
void orc_op_ede(void *ctx, double *x, double *y)
{
    orc_ede_ctx *c = (orc_ede_ctx *)ctx;
    orc_edemv(c->neq, c->nedge, c->A, c->lm2, x, y);
}
void orc_op_diag(void *ctx, double *x, double *y)
{
    orc_diag_ctx *c = (orc_diag_ctx *)ctx;
    orc_diag_apply(c->neq, c->invDiag, x, y);
}

void orc_op_ilu0(void *ctx, double *x, double *y)
{
    orc_ilu0_ctx *c = (orc_ilu0_ctx *)ctx;
    orc_ilu0_apply(c->neq, c->LU, c->JA, c->IA, c->D, x, y);
}

/* ======================================================================== */
/* restarted GMRES                                                           */
/* ======================================================================== */

/* pgmres.c:4-211.  Left preconditioner `precond`, right preconditioner
 * `precondR`; modified Gram-Schmidt Arnoldi; Givens rotations on the
 * Hessenberg column; back substitution; X update with the basis index as the
 * outer loop.  A NULL callback is NO_precond (NO_precond.c:3-6). */
int orc_pgmres(int neq, int kmax, int lmax, double tol,
               const double *B, double *X, const orc_ops *ops,
               double *rho_hist, int hist_cap, double *e0_hist, int *n_cycles)
{
    size_t len = (size_t)neq + 1;
    double **u = (double **)calloc((size_t)kmax + 1, sizeof(double *));
    double **h = (double **)calloc((size_t)kmax + 1, sizeof(double *));
    for (int i = 0; i <= kmax; i++) {
        u[i] = (double *)calloc(len, sizeof(double));
        h[i] = (double *)calloc((size_t)kmax + 1, sizeof(double));
    }
    double *e = (double *)calloc((size_t)kmax + 1, sizeof(double));
    double *c = (double *)calloc((size_t)kmax, sizeof(double));
    double *s = (double *)calloc((size_t)kmax, sizeof(double));
    double *y = (double *)calloc((size_t)kmax, sizeof(double));
    double *v = (double *)calloc(len, sizeof(double));

    orc_dzero(neq, X);
    double normb = sqrt(orc_ddot(neq, B, B));
    double eps = tol * normb;
    int cycles = 0, steps = 0;
    double rho;

    do {
        int i = 0;
        for (int j = 0; j < kmax; j++)
            orc_dzero(neq, u[j]);
        orc_dcopy(neq, X, v);
        if (ops->precondR) ops->precondR(ops->precondR_ctx, v, v);
        ops->matvec(ops->matvec_ctx, v, u[0]);
        if (ops->precond) ops->precond(ops->precond_ctx, u[0], u[0]);
        orc_daxpy(neq, -1.0, B, u[0]);
        orc_dscal(neq, -1.0, u[0]);
        e[0] = sqrt(orc_ddot(neq, u[0], u[0]));
        if (e0_hist) e0_hist[cycles] = e[0];
        orc_dscal(neq, 1.0 / e[0], u[0]);
        rho = e[0];

        do {
            steps++;
            orc_dcopy(neq, u[i], v);
            if (ops->precondR) ops->precondR(ops->precondR_ctx, v, v);
            ops->matvec(ops->matvec_ctx, v, u[i + 1]);
            if (ops->precond) ops->precond(ops->precond_ctx, u[i + 1], u[i + 1]);
            for (int j = 0; j <= i; j++) {
                h[j][i] = orc_ddot(neq, u[j], u[i + 1]);
                orc_daxpy(neq, -h[j][i], u[j], u[i + 1]);
            }
            h[i + 1][i] = sqrt(orc_ddot(neq, u[i + 1], u[i + 1]));
            orc_dscal(neq, 1.0 / h[i + 1][i], u[i + 1]);
            for (int j = 0; j <= i - 1; j++) {
                double h1 = c[j] * h[j][i] + s[j] * h[j + 1][i];
                double h2 = -s[j] * h[j][i] + c[j] * h[j + 1][i];
                h[j][i] = h1;
                h[j + 1][i] = h2;
            }
            double r = sqrt(h[i][i] * h[i][i] + h[i + 1][i] * h[i + 1][i]);
            c[i] = h[i][i] / r;
            s[i] = h[i + 1][i] / r;
            h[i][i] = r;
            h[i + 1][i] = 0.0;
            e[i + 1] = -s[i] * e[i];
            e[i] = c[i] * e[i];
            rho = fabs(e[i + 1]);
            if (rho_hist && steps - 1 < hist_cap) rho_hist[steps - 1] = rho;
            i++;
        } while (rho > eps && i < kmax);

        i--;
        y[i] = e[i] / h[i][i];
        for (int j = i - 1; j >= 0; j--) {
            double acc = 0.0;
            for (int q = j + 1; q <= i; q++)
                acc = acc + h[j][q] * y[q];
            y[j] = (e[j] - acc) / h[j][j];
        }
        for (int j = 0; j <= i; j++)
            for (long q = 0; q < neq; q++)
                X[q] = X[q] + u[j][q] * y[j];
        cycles++;
    } while (rho > eps && cycles < lmax);

    if (ops->precondR) ops->precondR(ops->precondR_ctx, X, X);

    for (int i = 0; i <= kmax; i++) { free(u[i]); free(h[i]); }
    free(u); free(h); free(e); free(c); free(s); free(y); free(v);
    if (n_cycles) *n_cycles = cycles;
    return steps;
}

/* ======================================================================== */
/* 2D mesh and data structures                                               */
/* ======================================================================== */

/* geraMalha.c:112-176 ("direita"): nodes row by row, x fastest; each cell
 * (i,j) with lower-left node a gives the triangles (a, a+1, a+1+W) and
 * (a, a+1+W, a+W), W = n+1 nodes per row; cells row by row. */
void orc_mesh2d_structured(int n, double *x, double *y, int *type, int *vertex)
{
    int W = n + 1;
    double hx = 1.0 / n;
    for (int j = 0, k = 0; j < W; j++)
        for (int i = 0; i < W; i++, k++) {
            x[k] = 0.0 + i * hx;
            y[k] = 0.0 + j * hx;
            int bnd = fabs(x[k] - 1.0) <= 1e-15 || fabs(x[k]) <= 1e-15 ||
                      fabs(y[k] - 1.0) <= 1e-15 || fabs(y[k]) <= 1e-15;
            type[k] = bnd ? 0 : 1;
        }
    for (int j = 0, e = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            int a = j * W + i;
            vertex[3 * e + 0] = a; vertex[3 * e + 1] = a + 1; vertex[3 * e + 2] = a + 1 + W; e++;
            vertex[3 * e + 0] = a; vertex[3 * e + 1] = a + 1 + W; vertex[3 * e + 2] = a + W; e++;
        }
}

/* Preprocess.c:73-81 */
int orc_number_nodes(int nnodes, const int *type, int *id)
{
    int neq = 0;
    for (int i = 0; i < nnodes; i++)
        id[i] = type[i] >= 1 ? neq++ : -1;
    return neq;
}

/* Fill_LM.c:3-24 */
void orc_fill_lm(int neq, int nel, int nnoel, const int *vertex, const int *type,
                 const int *id, int *lm)
{
    for (long e = 0; e < nel; e++)
        for (int j = 0; j < nnoel; j++) {
            int v = vertex[e * nnoel + j];
            lm[e * nnoel + j] = type[v] >= 1 ? id[v] : neq;
        }
}

/* (row, col) pair list sorted lexicographically and deduplicated: the same
 * sets that the reference's sorted linked lists hold (csr_List_insertA,
 * csr_Initialization.c:221-261; ede_List_insertA, ede_Initialization.c:172-212). */
typedef struct { int r, c; } pair_t;
static int pair_cmp(const void *a, const void *b)
{
    const pair_t *p = (const pair_t *)a, *q = (const pair_t *)b;
    if (p->r != q->r) return p->r < q->r ? -1 : 1;
    if (p->c != q->c) return p->c < q->c ? -1 : 1;
    return 0;
}
static long pair_unique(pair_t *v, long n)
{
    if (n == 0) return 0;
    qsort(v, (size_t)n, sizeof(pair_t), pair_cmp);
    long m = 1;
    for (long i = 1; i < n; i++)
        if (v[i].r != v[m - 1].r || v[i].c != v[m - 1].c)
            v[m++] = v[i];
    return m;
}
/* rank of column c inside the sorted row [lo,hi) */
static int pair_find(const pair_t *v, long lo, long hi, int c)
{
    while (lo < hi) {
        long mid = (lo + hi) / 2;
        if (v[mid].c < c) lo = mid + 1; else hi = mid;
    }
    return (int)lo;
}

/* csr_Initialization.c:5-218 with reordering NOT (reordering.c:8-10) */
int orc_csr_init_k(int neq, int nel, int k, const int *lm, int *IA, int *JA, int *csr_by_elem)
{
    pair_t *pr = (pair_t *)malloc(sizeof(pair_t) * (size_t)nel * k * k + 1);
    long np = 0;
    for (long e = 0; e < nel; e++)
        for (int a = 0; a < k; a++) {
            int I = lm[k * e + a];
            if (I == neq) continue;
            for (int b = 0; b < k; b++) {
                int J = lm[k * e + b];
                if (J == neq) continue;
                pr[np].r = I; pr[np].c = J; np++;
            }
        }
    long nnz = pair_unique(pr, np);
    if (JA == NULL) { free(pr); return (int)nnz; }
    for (int r = 0; r <= neq; r++) IA[r] = 0;
    for (long q = 0; q < nnz; q++) { IA[pr[q].r + 1]++; JA[q] = pr[q].c; }
    for (int r = 0; r < neq; r++) IA[r + 1] += IA[r];
    for (long e = 0; e < nel; e++)
        for (int a = 0; a < k; a++)
            for (int b = 0; b < k; b++) {
                int I = lm[k * e + a], J = lm[k * e + b];
                int slot = (int)nnz;
                if (I != neq && J != neq)
                    slot = pair_find(pr, IA[I], IA[I + 1], J);
                csr_by_elem[(long)k * k * e + k * a + b] = slot;
            }
    free(pr);
    return (int)nnz;
}

int orc_csr_init(int neq, int nel, const int *lm, int *IA, int *JA, int *csr_by_elem)
{
    return orc_csr_init_k(neq, nel, 3, lm, IA, JA, csr_by_elem);
}

/* ede_Initialization.c:4-171.  The three local vertices of every element are
 * ordered by equation number (ties broken exactly as the reference's
 * comparison chain, :22-83); edges (I1,I2),(I1,I3),(I2,I3) are numbered row
 * by row in ascending column order; the final count is decremented once
 * (:103) so that the edge (neq,neq) becomes the sink slot. */
static void ede_order(const int *l, int *o)
{
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

int orc_ede_init(int neq, int nel, const int *lm, int *order, int *edge_by_elem, int *lm2)
{
    pair_t *pr = (pair_t *)malloc(sizeof(pair_t) * (size_t)nel * 3 + 1);
    long np = 0;
    for (long e = 0; e < nel; e++) {
        int o[3];
        ede_order(lm + 3 * e, o);
        int I1 = lm[3 * e + o[0]], I2 = lm[3 * e + o[1]], I3 = lm[3 * e + o[2]];
        pr[np].r = I1; pr[np].c = I2; np++;
        pr[np].r = I1; pr[np].c = I3; np++;
        pr[np].r = I2; pr[np].c = I3; np++;
    }
    long ntot = pair_unique(pr, np);
    int nedge = (int)ntot - 1;
    if (lm2 == NULL) { free(pr); return nedge; }
    /* row offsets over rows 0..neq */
    long *off = (long *)calloc((size_t)neq + 2, sizeof(long));
    for (long q = 0; q < ntot; q++) off[pr[q].r + 1]++;
    for (int r = 0; r <= neq; r++) off[r + 1] += off[r];
    for (long e = 0; e < nel; e++) {
        int *o = order + 3 * e;
        ede_order(lm + 3 * e, o);
        int I1 = lm[3 * e + o[0]], I2 = lm[3 * e + o[1]], I3 = lm[3 * e + o[2]];
        edge_by_elem[3 * e + 0] = pair_find(pr, off[I1], off[I1 + 1], I2);
        edge_by_elem[3 * e + 1] = pair_find(pr, off[I1], off[I1 + 1], I3);
        edge_by_elem[3 * e + 2] = pair_find(pr, off[I2], off[I2 + 1], I3);
    }
    for (long q = 0; q < off[neq]; q++) {
        lm2[2 * q] = pr[q].r;
        lm2[2 * q + 1] = pr[q].c;
    }
    free(off);
    free(pr);
    return nedge;
}

/* ======================================================================== */
/* problem callbacks                                                         */
/* ======================================================================== */
void orc_prob2d_velocity(const orc_prob2d *pb, double x, double y, double beta[2])
{
    switch (pb->kind) {
    case ORC_P_ROTATING:
        beta[0] = (-y + 0.5) * pb->scale;
        beta[1] = (x - 0.5) * pb->scale;
        break;
    case ORC_P_VALIDACAO:                       /* VALIDACAO_Velocity.c:3-12 */
        beta[0] = 1;
        beta[1] = 20 * y;
        break;
    case ORC_P_ROTATE_ABOUT:                    /* 02 CONE_Velocity.c:3-12 */
        beta[0] = -(y - pb->by) * pb->scale;
        beta[1] = (x - pb->bx) * pb->scale;
        break;
    default:
        beta[0] = pb->bx;
        beta[1] = pb->by;
    }
}

/* VALIDACAO_Font.c:4-17: forcing of u = 10xy(1-x)(1-y)exp(x^4.5) */
static double validacao_font(double X, double Y, double k, double gamma,
                             double Be_x, double Be_y)
{
    double E = exp(pow(X, 4.5));
    double u = 10 * X * Y * (1 - X) * (1 - Y) * E;
    double u_x = 10 * Y * (1 - Y) * (E - 2 * E * X + 4.5 * E * pow(X, 4.5) - 4.5 * E * pow(X, 5.5));
    double u_x2 = 10 * Y * (1 - Y) * (-20.25 * E * pow(X, 9) + 20.25 * E * pow(X, 8)
                  - 33.75 * E * pow(X, 4.5) + 24.75 * E * pow(X, 3.5) - 2 * E);
    double u_y = 10 * E * X * (1 - X) * (1 - 2 * Y);
    double u_y2 = -20 * E * X * (1 - X);
    return -k * (u_x2 + u_y2) + Be_x * u_x + Be_y * u_y + gamma * u;
}

double orc_prob2d_font(const orc_prob2d *pb, double x, double y, double k, double gamma,
                       double bx, double by)
{
    if (pb->kind == ORC_P_VALIDACAO)
        return validacao_font(x, y, k, gamma, bx, by);
    return pb->f;
}

/* CAU_ShockCapture.c:3-30 */
static double cau_shock(double kx, double ky, double Be_x, double Be_y, double gamma,
                        double ue1, double ue2, double ue3, double ueb, double feb,
                        double y23, double y31, double y12, double x32, double x13,
                        double x21, double invArea, double h)
{
    const double third = 1. / 3.;
    double BetaGradu = 0.5 * invArea * (ue1 * (Be_x * y23 + Be_y * x32) +
                                        ue2 * (Be_x * y31 + Be_y * x13) +
                                        ue3 * (Be_x * y12 + Be_y * x21));
    double Ru = BetaGradu + gamma * ueb - feb;
    double g0 = 0.5 * invArea * (y23 * ue1 + y31 * ue2 + y12 * ue3);
    double g1 = 0.5 * invArea * (x32 * ue1 + x13 * ue2 + x21 * ue3);
    double ng = sqrt(g0 * g0 + g1 * g1);
    double bxp = BetaGradu * g0 / (ng * ng);
    double byp = BetaGradu * g1 / (ng * ng);
    double kt = bxp * bxp * kx + byp * byp * ky;
    double nbp = sqrt(bxp * bxp + byp * byp);
    double Pee = 0.5 * h * nbp * nbp * nbp / kt;
    if (ng > 0) {
        if (third * Pee < 0.7)
            return 0.5 * h * third * Pee * fabs(Ru) / ng;
        return 0.5 * h * 0.7 * fabs(Ru) / ng;
    }
    return 0;
}

/* Build_K_F_SUPG.c:5-163.  Every expression keeps the reference's operand
 * order so that the restatement is bitwise faithful at -ffp-contract=off. */
void orc_build_supg2d(int nnodes, int nel, int neq, const double *x, const double *y,
                      const int *type, const int *id, const int *vertex,
                      const double *u, const orc_prob2d *pb, double *ke, double *F)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    double kx, ky, gamma;
    if (pb->kind == ORC_P_VALIDACAO) { kx = 1.0; ky = 1.0; gamma = 1; }
    else { kx = pb->kappa; ky = pb->kappa; gamma = pb->gamma; }

    /* eval_U.c:3-25 : nodal field, Dirichlet nodes take u_D = 0 */
    double *U = (double *)calloc((size_t)nnodes, sizeof(double));
    for (int i = 0; i < nnodes; i++)
        U[i] = id[i] >= 0 ? u[id[i]] : 0.0;
    for (int r = 0; r < neq; r++)
        F[r] = 0;

    for (long E = 0; E < nel; E++) {
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
        double Beta[2];
        orc_prob2d_velocity(pb, XB, YB, Beta);
        double Be_x = Beta[0], Be_y = Beta[1];
        double norm = sqrt(Be_x * Be_x + Be_y * Be_y);
        double tau = norm > 0 ? 0.5 * h / norm : 0;

        double fe1 = orc_prob2d_font(pb, X0, Y0, kx, gamma, Be_x, Be_y);
        double fe2 = orc_prob2d_font(pb, X1, Y1, kx, gamma, Be_x, Be_y);
        double fe3 = orc_prob2d_font(pb, X2, Y2, kx, gamma, Be_x, Be_y);
        double feb = third * (fe1 + fe2 + fe3);
        double ue1 = U[J1], ue2 = U[J2], ue3 = U[J3];
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

        double hs = pb->h_shock == ORC_H_SQRTAREA ? 2 * sqrt(fabs(Area))
                                                  : sqrt(fabs(2 * Area));
        double Eu = cau_shock(kx, ky, Be_x, Be_y, gamma, ue1, ue2, ue3, ueb, feb,
                              y23, y31, y12, x32, x13, x21, invArea, hs);
        double Csc12 = Eu * 0.25 * invArea * (y23 * y31 + x32 * x13);
        double Csc13 = Eu * 0.25 * invArea * (y23 * y12 + x32 * x21);
        double Csc23 = Eu * 0.25 * invArea * (y31 * y12 + x13 * x21);

        double *k9 = ke + 9 * E;
        k9[0] = -(D12 + D13) + C11 - (Cpg12 + Cpg13) + R11 + Rpg11 - (Csc12 + Csc13);
        k9[1] = D12 + C12 + Cpg12 + R12 + Rpg11 + Csc12;
        k9[2] = D13 + C13 + Cpg13 + R12 + Rpg11 + Csc13;
        k9[3] = D12 + C11 + Cpg12 + R12 + Rpg21 + Csc12;
        k9[4] = -(D12 + D23) + C12 - (Cpg12 + Cpg23) + R11 + Rpg21 - (Csc12 + Csc23);
        k9[5] = D23 + C13 + Cpg23 + R12 + Rpg21 + Csc23;
        k9[6] = D13 + C11 + Cpg13 + R12 + Rpg31 + Csc13;
        k9[7] = D23 + C12 + Cpg23 + R12 + Rpg31 + Csc23;
        k9[8] = -(D13 + D23) + C13 - (Cpg13 + Cpg23) + R11 + Rpg31 - (Csc13 + Csc23);

        double Fe1 = Area * twelfth * (2 * fe1 + fe2 + fe3) + sixth * tau * (fe1 + fe2 + fe3) * (Be_x * y23 + Be_y * x32);
        double Fe2 = Area * twelfth * (fe1 + 2 * fe2 + fe3) + sixth * tau * (fe1 + fe2 + fe3) * (Be_x * y31 + Be_y * x13);
        double Fe3 = Area * twelfth * (fe1 + fe2 + 2 * fe3) + sixth * tau * (fe1 + fe2 + fe3) * (Be_x * y12 + Be_y * x21);

        /* F_assembly.c:5-27 with u_D = 0 for every supported problem: the
         * lifting terms subtract ke*0.  The Neumann term (:29-50) multiplies
         * by sixth = 1/6 in integer arithmetic, i.e. 0, and needs Type 2
         * nodes, which no supported mesh has. */
        if (type[J1] >= 1) {
            F[id[J1]] += Fe1;
            if (type[J2] == 0) F[id[J1]] -= k9[1] * 0.0;
            if (type[J3] == 0) F[id[J1]] -= k9[2] * 0.0;
        }
        if (type[J2] >= 1) {
            F[id[J2]] += Fe2;
            if (type[J1] == 0) F[id[J2]] -= k9[3] * 0.0;
            if (type[J3] == 0) F[id[J2]] -= k9[5] * 0.0;
        }
        if (type[J3] >= 1) {
            F[id[J3]] += Fe3;
            if (type[J1] == 0) F[id[J3]] -= k9[6] * 0.0;
            if (type[J2] == 0) F[id[J3]] -= k9[7] * 0.0;
        }
    }
    free(U);
}

/* ede_assembly.c:3-35 after setzeros.c:8-11 */
void orc_ede_assembly(int nel, int nedge, const double *ke, const int *order,
                      const int *edge_by_elem, double *Aede)
{
    memset(Aede, 0, sizeof(double) * 4 * ((size_t)nedge + 1));
    for (long E = 0; E < nel; E++) {
        const double *k9 = ke + 9 * E;
        int I1 = order[3 * E], I2 = order[3 * E + 1], I3 = order[3 * E + 2];
        double *K1 = Aede + 4L * edge_by_elem[3 * E];
        double *K2 = Aede + 4L * edge_by_elem[3 * E + 1];
        double *K3 = Aede + 4L * edge_by_elem[3 * E + 2];
        K1[0] += 0.5 * k9[3 * I1 + I1]; K1[1] += k9[3 * I1 + I2];
        K1[2] += k9[3 * I2 + I1];       K1[3] += 0.5 * k9[3 * I2 + I2];
        K2[0] += 0.5 * k9[3 * I1 + I1]; K2[1] += k9[3 * I1 + I3];
        K2[2] += k9[3 * I3 + I1];       K2[3] += 0.5 * k9[3 * I3 + I3];
        K3[0] += 0.5 * k9[3 * I2 + I2]; K3[1] += k9[3 * I2 + I3];
        K3[2] += k9[3 * I3 + I2];       K3[3] += 0.5 * k9[3 * I3 + I3];
        double *Ks = Aede + 4L * nedge;
        Ks[0] = Ks[1] = Ks[2] = Ks[3] = 0;
    }
}

/* csr_assembly.c:3-24 after setzeros.c:12-15 */
void orc_csr_assembly(int nel, int nnz, const double *ke, const int *csr_by_elem, double *AA)
{
    orc_csr_assembly_k(nel, 3, nnz, ke, csr_by_elem, AA);
}

/* 2D csr_assembly.c:3-24 / 3D csr_assembly.c:3-30: slot by slot in row-major
 * element order, sink reset after every element */
void orc_csr_assembly_k(int nel, int k, int nnz, const double *ke, const int *csr_by_elem,
                        double *AA)
{
    const int kk = k * k;
    memset(AA, 0, sizeof(double) * ((size_t)nnz + 1));
    for (long E = 0; E < nel; E++) {
        for (int q = 0; q < kk; q++)
            AA[csr_by_elem[kk * E + q]] += ke[kk * E + q];
        AA[nnz] = 0;
    }
}

/* ======================================================================== */
/* ILU(0)                                                                    */
/* ======================================================================== */

/* ILUp_precond_setup.c:8-60 with p = 0 and ILUp :66-167.  With p = 0 the
 * symbolic phase (ilup.c LEVEL_OF_FILL_lu) adds no fill-in, so L and U have
 * the pattern of A below and above the diagonal, each row in ascending
 * column order (the CSR order).  Row i: L, D, U copied from A; for each L
 * entry in order, the multiplier L_ij *= D_j (D already inverted) and row j
 * of U is merged into row i where the patterns meet; finally D_i = 1/D_i.
 * LU[nnz] holds L and U in A's slots (the diagonal slot unused), D[neq] the
 * inverted pivots.  Returns 0, or 1 + i for a zero pivot at row i (the
 * reference prints and exits, :154-163). */
int orc_ilu0_factor(int neq, const double *AA, const int *JA, const int *IA, double *LU,
                    double *D)
{
    int *jw = (int *)malloc(sizeof(int) * (size_t)(neq > 0 ? neq : 1));
    for (int j = 0; j < neq; j++) jw[j] = -1;
    int rc = 0;
    for (int i = 0; i < neq && !rc; i++) {
        for (int q = IA[i]; q < IA[i + 1]; q++) {
            jw[JA[q]] = q;
            LU[q] = 0;
        }
        D[i] = 0;
        for (int q = IA[i]; q < IA[i + 1]; q++) {
            if (JA[q] == i) D[i] = AA[q];
            else LU[q] = AA[q];
        }
        for (int q = IA[i]; q < IA[i + 1] && JA[q] < i; q++) {
            int jrow = JA[q];
            LU[q] *= D[jrow];
            for (int t = IA[jrow]; t < IA[jrow + 1]; t++) {
                int col = JA[t];
                if (col <= jrow) continue;            /* U part of row jrow */
                int pos = jw[col];
                if (pos == -1) continue;
                if (col == i) D[i] -= LU[q] * LU[t];
                else LU[pos] -= LU[q] * LU[t];
            }
        }
        for (int q = IA[i]; q < IA[i + 1]; q++) jw[JA[q]] = -1;
        if (D[i] == 0) rc = 1 + i;
        else D[i] = 1.0 / D[i];
    }
    free(jw);
    return rc;
}

/* ILUp_precond.c:26-46: forward solve with unit L, backward with U and the
 * inverted pivots; z may alias p */
void orc_ilu0_apply(int neq, const double *LU, const int *JA, const int *IA, const double *D,
                    const double *p, double *z)
{
    for (int i = 0; i < neq; i++) {
        double x = p[i];
        for (int q = IA[i]; q < IA[i + 1] && JA[q] < i; q++) x -= z[JA[q]] * LU[q];
        z[i] = x;
    }
    for (int i = neq - 1; i >= 0; i--) {
        double x = z[i];
        for (int q = IA[i]; q < IA[i + 1]; q++)
            if (JA[q] > i) x -= z[JA[q]] * LU[q];
        z[i] = x * D[i];
    }
}

void orc_ilu0_apply_f32(int neq, const float *LU, const int *JA, const int *IA, const float *D,
                        const float *p, float *z)
{
    for (int i = 0; i < neq; i++) {
        float x = p[i];
        for (int q = IA[i]; q < IA[i + 1] && JA[q] < i; q++) x -= z[JA[q]] * LU[q];
        z[i] = x;
    }
    for (int i = neq - 1; i >= 0; i--) {
        float x = z[i];
        for (int q = IA[i]; q < IA[i + 1]; q++)
            if (JA[q] > i) x -= z[JA[q]] * LU[q];
        z[i] = x * D[i];
    }
}

/* ======================================================================== */
/* 3D steady transport (01_SS_Transport_Equation/3D)                         */
/* ======================================================================== */

/* Unit cube, n^3 cells, each split into the 6 tetrahedra around its main
 * diagonal (Kuhn split, conforming across cells): for every permutation
 * (a,b,c) of the axes the tetrahedron (v0, v0+e_a, v0+e_a+e_b, v0+e_a+e_b+e_c),
 * with the last two vertices swapped for odd permutations so that every
 * element has positive volume in Element_Configuration.c:58.
 * Node (i,j,k) has number i + (n+1)(j + (n+1)k); boundary nodes get uType 0
 * (prescribed), inner nodes uType 1 (unknown), as Fill_ID.c reads them.
 * Elements cell by cell (i fastest), the 6 permutations in lexicographic
 * order.  x,y,z,utype: (n+1)^3 ; vertex: 6 n^3 * 4.  This generator is the
 * build's own (the reference reads meshes from files). */
void orc_mesh3d_structured(int n, double *x, double *y, double *z, int *utype, int *vertex)
{
    const long N = n + 1;
    for (long k = 0; k < N; k++)
        for (long j = 0; j < N; j++)
            for (long i = 0; i < N; i++) {
                long v = i + N * (j + N * k);
                x[v] = (double)i / n;
                y[v] = (double)j / n;
                z[v] = (double)k / n;
                utype[v] = (i > 0 && i < n && j > 0 && j < n && k > 0 && k < n) ? 1 : 0;
            }
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const long stride[3] = {1, N, N * N};
    long e = 0;
    for (long k = 0; k < n; k++)
        for (long j = 0; j < n; j++)
            for (long i = 0; i < n; i++) {
                long v0 = i + N * (j + N * k);
                for (int p = 0; p < 6; p++) {
                    long v = v0;
                    vertex[4 * e] = (int)v;
                    for (int s = 0; s < 3; s++) {
                        v += stride[perm[p][s]];
                        vertex[4 * e + 1 + s] = (int)v;
                    }
                    if (p == 1 || p == 2 || p == 5) {   /* odd permutation: keep volume > 0 */
                        int tmp = vertex[4 * e + 2];
                        vertex[4 * e + 2] = vertex[4 * e + 3];
                        vertex[4 * e + 3] = tmp;
                    }
                    e++;
                }
            }
}

#define ORC_PI3 3.141592654   /* SSTransportEquation3D.h:15 */

/* DIFCONV/DIFCONV_*.c (kappa = A, beta = (1,1,1), sigma = 0, manufactured f,
 * u = 0 on the boundary) or constant coefficients */
static void prob3d_kappa(const orc_prob3d *pb, double *kx, double *ky, double *kz)
{
    if (pb->kind == ORC_P3_DIFCONV) { *kx = pb->A; *ky = pb->A; *kz = pb->A; }
    else { *kx = pb->kappa; *ky = pb->kappa; *kz = pb->kappa; }
}
static void prob3d_beta(const orc_prob3d *pb, double *bx, double *by, double *bz)
{
    if (pb->kind == ORC_P3_DIFCONV) { *bx = 1.0; *by = 1.0; *bz = 1.0; }
    else { *bx = pb->bx; *by = pb->by; *bz = pb->bz; }
}
static double prob3d_sigma(const orc_prob3d *pb)
{
    return pb->kind == ORC_P3_DIFCONV ? 0.0 : pb->sigma;
}
/* DIFCONV_f.c */
double orc_prob3d_f(const orc_prob3d *pb, double x, double y, double z)
{
    if (pb->kind != ORC_P3_DIFCONV) return pb->f;
    const double A = pb->A, PI = ORC_PI3;
    double f;
    f = A * 3.0 * PI * PI * sin(PI * x) * sin(PI * y) * sin(PI * z) +
        PI * cos(PI * x) * sin(PI * y) * sin(PI * z) + PI * sin(PI * x) * cos(PI * y) * sin(PI * z) +
        PI * sin(PI * x) * sin(PI * y) * cos(PI * z);
    return f;
}

/* Build_SUPG.c:22-196 with Element_Configuration.c:4-58 and
 * font_assembly.c:3-57 (prescribed values are 0 for both problem kinds).
 * ke[16e + 4r + c] = Ke[r][c] (ebe_assembly.c:11-15 stores row-major). */
void orc_build_supg3d(int nel, int neq, const double *x, const double *y, const double *z,
                      const int *utype, const int *vertex, const int *lm,
                      const orc_prob3d *pb, double *ke, double *F)
{
    for (int i = 0; i <= neq; i++) F[i] = 0.0;
    for (long e = 0; e < nel; e++) {
        const int *V = vertex + 4 * e;
        double X[4], Y[4], Z[4], a[4], b[4], c[4], d[4];
        for (int i = 0; i < 4; i++) { X[i] = x[V[i]]; Y[i] = y[V[i]]; Z[i] = z[V[i]]; }
        a[0] =  (X[1]*Y[2]*Z[3] + X[3]*Y[1]*Z[2] + X[2]*Y[3]*Z[1]) - (X[3]*Y[2]*Z[1] + X[1]*Y[3]*Z[2] + X[2]*Y[1]*Z[3]);
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
        const double Volume = (a[0] + X[0]*b[0] + Y[0]*c[0] + Z[0]*d[0])/6.0;
        const double h = pow(Volume, 1.0/3.0);
        const double xb = (X[0] + X[1] + X[2] + X[3])/4.0;
        const double yb = (Y[0] + Y[1] + Y[2] + Y[3])/4.0;
        const double zb = (Z[0] + Z[1] + Z[2] + Z[3])/4.0;
        (void)xb; (void)yb; (void)zb;   /* coefficients here are constant in space */
        double bb[4][4], cc[4][4], dd[4][4];
        for (int r = 0; r < 4; r++)
            for (int s = 0; s < 4; s++) { bb[r][s] = b[r]*b[s]; cc[r][s] = c[r]*c[s]; dd[r][s] = d[r]*d[s]; }
        const double ConstD = 1.0/(36.0*Volume);
        double kX, kY, kZ, bX, bY, bZ;
        prob3d_kappa(pb, &kX, &kY, &kZ);
        const double normK = sqrt(kX*kX + kY*kY + kZ*kZ);
        double E[4][4];
        for (int r = 0; r < 4; r++)
            for (int s = 0; s < 4; s++) E[r][s] = kX*bb[r][s] + kY*cc[r][s] + kZ*dd[r][s];
        const double ConstC = 1.0/24.0;
        prob3d_beta(pb, &bX, &bY, &bZ);
        const double normB = sqrt(bX*bX + bY*bY + bZ*bZ);
        double C[4];
        for (int r = 0; r < 4; r++) C[r] = bX*b[r] + bY*c[r] + bZ*d[r];
        const double sigma = prob3d_sigma(pb);
        const double ConstR = (sigma*Volume)/20.0;
        const double peclet = (normB*h)/(2.0*normK);
        const double xi = fmax(0, 1-1/peclet);
        const double tau = (h*xi)/(2.0*normB);
        const double ConstCsupg = tau/(36.0*Volume);
        double Cs[4][4];
        for (int r = 0; r < 4; r++)
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
        double Ke[4][4];
        for (int r = 0; r < 4; r++)
            for (int s = 0; s < 4; s++)
                Ke[r][s] = ConstD*E[r][s] + ConstC*C[s] + ConstCsupg*Cs[r][s] + ConstRsupg*C[r] +
                           (r == s ? 2*ConstR : ConstR);
        double f[4], Fe[4];
        const double ConstF = Volume/20.0;
        for (int i = 0; i < 4; i++) f[i] = orc_prob3d_f(pb, X[i], Y[i], Z[i]);
        const double Faux = f[0] + f[1] + f[2] + f[3];
        const double ConstFsupg = tau/24.0;
        for (int i = 0; i < 4; i++) Fe[i] = ConstF*(Faux + f[i]) + ConstFsupg*C[i]*Faux;
        for (int r = 0; r < 4; r++)
            for (int s = 0; s < 4; s++) ke[16 * e + 4 * r + s] = Ke[r][s];
        /* font_assembly.c:21-56: g = 0 at unknowns (uType 1) and upresc at
         * prescribed nodes, which is 0 for both problem kinds here */
        double g[4];
        for (int i = 0; i < 4; i++) g[i] = 0.0;
        (void)utype;
        for (int r = 0; r < 4; r++)
            Fe[r] = Fe[r] - (Ke[r][0]*g[0] + Ke[r][1]*g[1] + Ke[r][2]*g[2] + Ke[r][3]*g[3]);
        const int *L = lm + 4 * e;
        for (int r = 0; r < 4; r++) F[L[r]] = F[L[r]] + Fe[r];
        F[neq] = 0.0;
    }
}

/* ======================================================================== */
/* reverse Cuthill-McKee (Reordering/symrcm.c)                               */
/* ======================================================================== */

/* a BFS node: equation, degree, distance to the root (symrcm.h:8-16) */
typedef struct { int id, deg, dist; } rcm_node;

static void rcm_heap_down(rcm_node *A, int j, int size)      /* H_heapify_min */
{
    for (;;) {
        int l = 2 * j + 1, r = 2 * j + 2, s = j;
        if (l < size && A[l].deg < A[j].deg) s = l;
        if (r < size && A[r].deg < A[s].deg) s = r;
        if (s == j) break;
        rcm_node t = A[j]; A[j] = A[s]; A[s] = t;
        j = s;
    }
}

static void rcm_heap_push(rcm_node *H, int *h, rcm_node o)    /* H_insert */
{
    int i = (*h)++;
    H[i] = o;
    while (i > 0) {
        int p = (i - 1) >> 1;
        if (!(H[i].deg < H[p].deg)) break;
        rcm_node t = H[i]; H[i] = H[p]; H[p] = t;
        i = p;
    }
}

static rcm_node rcm_heap_pop(rcm_node *H, int *h)             /* H_remove_min */
{
    rcm_node r = H[0];
    H[0] = H[--(*h)];
    rcm_heap_down(H, 0, *h);
    return r;
}

/* the neighbours of i in ascending order: column i of A (ridx) merged with
 * column i of A^T (ridx2), a shared index visited once (symrcm.c:117-177) */
typedef void (*rcm_visit_fn)(void *st, int r);
static void rcm_neighbours(int i, const int *cidx, const int *ridx, const int *cidx2,
                           const int *ridx2, rcm_visit_fn fn, void *st)
{
    int j1 = cidx[i], j2 = cidx2[i];
    while (j1 < cidx[i + 1] || j2 < cidx2[i + 1]) {
        if (j1 == cidx[i + 1]) fn(st, ridx2[j2++]);
        else if (j2 == cidx2[i + 1]) fn(st, ridx[j1++]);
        else {
            int r1 = ridx[j1], r2 = ridx2[j2];
            if (r1 <= r2) {
                fn(st, r1);
                j1++;
                if (r1 == r2) j2++;
            } else {
                fn(st, r2);
                j2++;
            }
        }
    }
}

typedef struct {
    rcm_node *Q;  int N, *qt;       /* queue (find_starting_node) or heap (main BFS) */
    rcm_node *H;  int *h;
    int *visit;   const int *D;  int dist;
} rcm_state;

static void rcm_enqueue_visit(void *s_, int r)
{
    rcm_state *s = (rcm_state *)s_;
    if (s->visit[r]) return;
    rcm_node w = {r, s->D[r], s->dist};
    s->Q[*s->qt] = w;
    *s->qt = (*s->qt + 1) % (s->N + 1);
    s->visit[r] = 1;
}

static void rcm_heap_visit(void *s_, int r)
{
    rcm_state *s = (rcm_state *)s_;
    if (s->visit[r]) return;
    rcm_node w = {r, s->D[r], s->dist};
    rcm_heap_push(s->H, s->h, w);
    s->visit[r] = 1;
}

/* find_starting_node (symrcm.c:357-493): repeated BFS from the farthest node
 * found, preferring larger distance, then (other node) larger degree */
static int rcm_start(int N, const int *cidx, const int *ridx, const int *cidx2, const int *ridx2,
                     const int *D, int start)
{
    rcm_node *Q = (rcm_node *)calloc((size_t)N + 1, sizeof(rcm_node));
    int *visit = (int *)calloc((size_t)N, sizeof(int));
    int qh = 0, qt = 0, max_dist = 0;
    rcm_node x = {start, D[start], 0};
    Q[qt] = x; qt = (qt + 1) % (N + 1);
    visit[start] = 1;
    for (;;) {
        while (qh != qt) {
            rcm_node v = Q[qh]; qh = (qh + 1) % (N + 1);
            if (v.dist > x.dist || (v.id != x.id && v.deg > x.deg)) x = v;
            rcm_state st = {Q, N, &qt, NULL, NULL, visit, D, v.dist + 1};
            rcm_neighbours(v.id, cidx, ridx, cidx2, ridx2, rcm_enqueue_visit, &st);
        }
        if (max_dist < x.dist) {
            max_dist = x.dist;
            memset(visit, 0, sizeof(int) * (size_t)N);
            visit[x.id] = 1;
            x.dist = 0;
            qt = qh = 0;
            Q[qt] = x; qt = (qt + 1) % (N + 1);
        } else
            break;
    }
    free(Q);
    free(visit);
    return x.id;
}

void orc_symrcm(int N, const int *ia, const int *ja, int *P)
{
    const int *cidx = ia, *ridx = ja;   /* the pattern read as CCS (symrcm.c:23-27) */
    const int nz = ia[N];
    int *cidx2 = (int *)calloc((size_t)N + 1, sizeof(int));
    int *ridx2 = (int *)calloc((size_t)(nz > 0 ? nz : 1), sizeof(int));
    int *w = (int *)calloc((size_t)N + 1, sizeof(int));
    /* transpose (symrcm.c:277-308) */
    for (int q = 0; q < nz; q++) w[ridx[q]]++;
    int acc = 0;
    for (int i = 0; i < N; i++) { cidx2[i] = acc; acc += w[i]; w[i] = cidx2[i]; }
    cidx2[N] = acc;
    for (int j = 0; j < N; j++)
        for (int k = cidx[j]; k < cidx[j + 1]; k++) ridx2[w[ridx[k]]++] = j;
    free(w);
    /* calc_degrees (symrcm.c:310-355) */
    int *D = (int *)calloc((size_t)(N > 0 ? N : 1), sizeof(int));
    int max_deg = 0;
    for (int j = 0; j < N; j++)
        for (int q = cidx[j]; q < cidx[j + 1]; q++) {
            int k = ridx[q];
            if (++D[k] > max_deg) max_deg = D[k];
            if (k != j) {
                int found = 0;
                for (int l = cidx[k]; l < cidx[k + 1]; l++) {
                    if (ridx[l] == j) { found = 1; break; }
                    if (ridx[l] > j) break;
                }
                if (!found && ++D[j] > max_deg) max_deg = D[j];
            }
        }
    if (max_deg == 0) {
        for (int i = 0; i < N; i++) P[i] = i;
    } else {
        rcm_node *S = (rcm_node *)calloc((size_t)max_deg, sizeof(rcm_node));
        rcm_node *Q = (rcm_node *)calloc((size_t)N + 1, sizeof(rcm_node));
        int *visit = (int *)calloc((size_t)N, sizeof(int));
        int c = -1, qh = 0, qt = 0;
        do {
            int i = 0;
            while (i < N && visit[i]) i++;
            rcm_node v = {0, 0, 0};
            v.id = rcm_start(N, cidx, ridx, cidx2, ridx2, D, i);
            v.deg = D[v.id];
            v.dist = 0;
            visit[v.id] = 1;
            Q[qt] = v; qt = (qt + 1) % (N + 1);
            while (qh != qt) {
                v = Q[qh]; qh = (qh + 1) % (N + 1);
                P[++c] = v.id;
                int s = 0;
                rcm_state st = {NULL, N, NULL, S, &s, visit, D, v.dist + 1};
                rcm_neighbours(v.id, cidx, ridx, cidx2, ridx2, rcm_heap_visit, &st);
                while (s > 0) {
                    rcm_node u = rcm_heap_pop(S, &s);
                    Q[qt] = u; qt = (qt + 1) % (N + 1);
                }
            }
        } while (c + 1 < N);
        free(S); free(Q); free(visit);
        /* reverse (symrcm.c:232-235) */
        for (int i = 0, j = N - 1; i <= N / 2 - 1; i++, j--) { int t = P[i]; P[i] = P[j]; P[j] = t; }
    }
    free(D); free(cidx2); free(ridx2);
}

/* ======================================================================== */
/* 4-dof (Euler) element products and node block-diagonal preconditioners   */
/* ======================================================================== */

/* EBE/ebe2mvNDOF4.c:4-79 (2D): the 12x12 element product without the
 * same-node 4x4 blocks, then BlockDiag2_precond.c (invBD != NULL) or
 * NOBlockDiag2_precond.c (BD != NULL).  Id[nnodes*4]: equation of dof d of
 * node I (sink = neq). */
void orc_ebe2mv_ndof4(int neq, int nel, const double *A, const int *lm, int nnodes,
                      const int *Id, const double *invBD, const double *BD, double *P,
                      double *Q)
{
    orc_ebe2mv_items(neq, nel, 12, A, lm, nnodes, Id, invBD, BD, P, Q);
}

/* the off-block product of 4-dof items of k = 8 (edges, ede2mvNDOF4.c:23-53:
 * each node's rows against the other node's columns) or k = 12 (elements),
 * then the node pass below */
void orc_ebe2mv_items(int neq, int nitems, int k, const double *A, const int *lm, int nnodes,
                      const int *Id, const double *invBD, const double *BD, double *P,
                      double *Q)
{
    for (int r = 0; r < neq; r++) Q[r] = 0;
    P[neq] = 0;
    Q[neq] = 0;
    for (long e = 0; e < nitems; e++) {
        const int *le = lm + (long)k * e;
        const double *Ae = A + (long)k * k * e;
        double p[12];
        for (int c = 0; c < k; c++) p[c] = P[le[c]];
        for (int r = 0; r < k; r++) {
            const int nb = r / 4;     /* node block of the row: skip its columns */
            double s = 0;
            int first = 1;
            for (int c = 0; c < k; c++) {
                if (c / 4 == nb) continue;
                double t = Ae[k * r + c] * p[c];
                if (first) { s = t; first = 0; }
                else s = s + t;
            }
            Q[le[r]] += s;
        }
        Q[neq] = 0.0;
    }
    for (int I = 0; I < nnodes; I++) {
        const int *id = Id + 4 * I;
        if (invBD) {   /* BlockDiag2_precond.c:11-20 (z = Q here) */
            const double *M = invBD + 16 * I;
            double a = Q[id[0]], b = Q[id[1]], c = Q[id[2]], d = Q[id[3]];
            for (int r = 0; r < 4; r++)
                Q[id[r]] = P[id[r]] + M[4 * r] * a + M[4 * r + 1] * b + M[4 * r + 2] * c +
                           M[4 * r + 3] * d;
        } else if (BD) {   /* NOBlockDiag2_precond.c:11-20 */
            const double *M = BD + 16 * I;
            double a = P[id[0]], b = P[id[1]], c = P[id[2]], d = P[id[3]];
            for (int r = 0; r < 4; r++)
                Q[id[r]] += M[4 * r] * a + M[4 * r + 1] * b + M[4 * r + 2] * c + M[4 * r + 3] * d;
        }
        Q[neq] = 0;
    }
}

/* BlockDiag_precond_EBE_setup.c:88-145 matrix_Inverse: cofactor inverse of a
 * row-major 4x4 block, the reference's expression grouping */
void orc_matrix_inverse4(const double *B, double *inv)
{
    const double a = B[0], b = B[1], c = B[2], d = B[3], e = B[4], f = B[5], g = B[6],
                 h = B[7], i = B[8], j = B[9], k = B[10], l = B[11], m = B[12], n = B[13],
                 o = B[14], p = B[15];
    const double det = (a*f*k*p - a*f*l*o - a*g*j*p + a*g*l*n + a*h*j*o - a*h*k*n
                        -b*e*k*p  + b*e*l*o + b*g*i*p - b*g*l*m - b*h*i*o + b*h*k*m +
                        c*(e*j*p - e*l*n - f*i*p + f*l*m + h*i*n - h*j*m) +
                        d*(-e*j*o + e*k*n + f*i*o - f*k*m - g*i*n + g*j*m));
    const double invdet = 1.0/det;
    inv[0] = invdet*(-h*k*n + g*l*n + h*j*o - f*l*o - g*j*p + f*k*p);
    inv[1] = invdet*(d*k*n - c*l*n - d*j*o + b*l*o + c*j*p - b*k*p);
    inv[2] = invdet*(-d*g*n + c*h*n + d*f*o - b*h*o - c*f*p + b*g*p);
    inv[3] = invdet*(d*g*j - c*h*j - d*f*k + b*h*k + c*f*l - b*g*l);
    inv[4] = invdet*(h*k*m - g*l*m - h*i*o + e*l*o + g*i*p - e*k*p);
    inv[5] = invdet*(-d*k*m + c*l*m + d*i*o - a*l*o - c*i*p + a*k*p);
    inv[6] = invdet*(d*g*m - c*h*m - d*e*o + a*h*o + c*e*p - a*g*p);
    inv[7] = invdet*(-d*g*i + c*h*i + d*e*k - a*h*k - c*e*l + a*g*l);
    inv[8] = invdet*(-h*j*m + f*l*m + h*i*n - e*l*n - f*i*p + e*j*p);
    inv[9] = invdet*(d*j*m - b*l*m - d*i*n + a*l*n + b*i*p - a*j*p);
    inv[10] = invdet*(-d*f*m + b*h*m + d*e*n - a*h*n - b*e*p + a*f*p);
    inv[11] = invdet*(d*f*i - b*h*i - d*e*j + a*h*j + b*e*l - a*f*l);
    inv[12] = invdet*(g*j*m - f*k*m - g*i*n + e*k*n + f*i*o - e*j*o);
    inv[13] = invdet*(-c*j*m + b*k*m + c*i*n - a*k*n - b*i*o + a*j*o);
    inv[14] = invdet*(c*j*m - b*g*m - c*e*n + a*g*n + b*e*o - a*f*o);
    inv[15] = invdet*(-c*f*i + b*g*i + c*e*j - a*g*j - b*e*k + a*f*k);
}

/* BlockDiag_precond.c:3-27: z_node = invBD_node * p_node, z[neq] = 0; z may alias p */
void orc_blockdiag_apply(int neq, int nnodes, const int *Id, const double *invBD,
                         const double *p, double *z)
{
    for (int I = 0; I < nnodes; I++) {
        const int *id = Id + 4 * I;
        const double *M = invBD + 16 * I;
        double a = p[id[0]], b = p[id[1]], c = p[id[2]], d = p[id[3]];
        for (int r = 0; r < 4; r++)
            z[id[r]] = M[4 * r] * a + M[4 * r + 1] * b + M[4 * r + 2] * c + M[4 * r + 3] * d;
        z[neq] = 0;
    }
}

/* BlockDiag_precond_EBE_setup.c:6-86 (2D, 3 nodes of 4 dofs): node blocks
 * summed in element order from the same-node 4x4 blocks of the 12x12 element
 * matrices (index table I_diag, :18-19), cofactor inverses, F <- BD^{-1} F.
 * vertex[nel*3]; BD, invBD [nnodes*16] (BD zeroed first). */
void orc_blockdiag_setup_ebe(int neq, int nel, int nnodes, const int *vertex, const double *A,
                             const int *Id, double *BD, double *invBD, double *F)
{
    orc_blockdiag_setup_items(neq, nel, 3, nnodes, vertex, A, Id, BD, invBD, F);
}

/* the node blocks of items of nv = 3 nodes (elements, k = 12) or nv = 2
 * (edges, k = 8: BlockDiag_precond_EDE_setup.c:105-116, I_diag[v][4r+c] =
 * 36 v + 8 r + c); items ascending, as both setups loop */
void orc_blockdiag_setup_items(int neq, int nitems, int nv, int nnodes, const int *nodes,
                               const double *A, const int *Id, double *BD, double *invBD,
                               double *F)
{
    const int k = 4 * nv;
    memset(BD, 0, sizeof(double) * 16 * (size_t)nnodes);
    for (long E = 0; E < nitems; E++)
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++)
                for (int v = 0; v < nv; v++) {
                    const int node = nodes[(long)nv * E + v];
                    /* I_diag[v][4r+c] = 4 (k + 1) v + k r + c */
                    BD[16 * node + 4 * r + c] += A[(long)k * k * E + 4 * (k + 1) * v + k * r + c];
                }
    for (int I = 0; I < nnodes; I++) orc_matrix_inverse4(BD + 16 * I, invBD + 16 * I);
    if (F) orc_blockdiag_apply(neq, nnodes, Id, invBD, F, F);
}

/* ---- the edge structure of the Euler driver ------------------------------
 * 03_Euler_Equations/2D/01_CommonFiles/ede_Initialization.c:23-88: each
 * element's vertices sorted (order[] = local index of the smallest, middle,
 * largest), its edges (I1,I2), (I1,I3), (I2,I3) entered in a per-node list of
 * the larger end, sorted ascending (ede_List_insertA, :205-241), a new entry
 * taking the next edge number.  Here the lists are arrays of (J, number)
 * linked by index. */
typedef struct { int *head, *next, *J, *num, used; } orc_elist;

static int orc_elist_insert(orc_elist *L, int I, int J, int *nedge)
{
    int prev = -1, cur = L->head[I];
    while (cur >= 0 && L->J[cur] < J) { prev = cur; cur = L->next[cur]; }
    if (cur >= 0 && L->J[cur] == J) return L->num[cur];
    const int n = L->used++;
    L->J[n] = J;
    L->num[n] = (*nedge)++;
    L->next[n] = cur;
    if (prev < 0) L->head[I] = n; else L->next[prev] = n;
    return L->num[n];
}

int orc_ede4_init(int nnodes, int nel, int neq, const int *vertex, const int *ids, int *order,
                  int *edge_by_elem, int *lm2)
{
    orc_elist L;
    const long cap = 3L * nel + 1;
    L.head = (int *)malloc(sizeof(int) * (size_t)(nnodes > 0 ? nnodes : 1));
    L.next = (int *)malloc(sizeof(int) * cap);
    L.J = (int *)malloc(sizeof(int) * cap);
    L.num = (int *)malloc(sizeof(int) * cap);
    L.used = 0;
    for (int i = 0; i < nnodes; i++) L.head[i] = -1;
    int nedge = 0;
    for (long K = 0; K < nel; K++) {
        const int *V = vertex + 3 * K;
        int o0, o1, o2;
        if (V[0] < V[1] && V[0] < V[2]) { o0 = 0; o1 = V[1] < V[2] ? 1 : 2; }
        else if (V[1] < V[2]) { o0 = 1; o1 = V[0] < V[2] ? 0 : 2; }
        else { o0 = 2; o1 = V[0] < V[1] ? 0 : 1; }
        o2 = 3 - o0 - o1;
        int *o = order + 3 * K;
        o[0] = o0; o[1] = o1; o[2] = o2;
        const int I1 = V[o0], I2 = V[o1], I3 = V[o2];
        /* EDGE_by_Element (:96-129): the numbers of (I1,I2), (I1,I3), (I2,I3) */
        edge_by_elem[3 * K + 0] = orc_elist_insert(&L, I1, I2, &nedge);
        edge_by_elem[3 * K + 1] = orc_elist_insert(&L, I1, I3, &nedge);
        edge_by_elem[3 * K + 2] = orc_elist_insert(&L, I2, I3, &nedge);
    }
    /* lm2 (:131-183): the smaller node's 4 dofs, then the larger's; id -1 ->
     * neq, written once per element that holds the edge */
    if (lm2)
        for (long K = 0; K < nel; K++) {
            const int *V = vertex + 3 * K, *o = order + 3 * K, *eb = edge_by_elem + 3 * K;
            const int Io[3] = {V[o[0]], V[o[1]], V[o[2]]};
            static const int ends[3][2] = {{0, 1}, {0, 2}, {1, 2}};
            for (int s = 0; s < 3; s++)
                for (int i = 0; i < 4; i++)
                    for (int h = 0; h < 2; h++) {
                        const int id = ids[4 * Io[ends[s][h]] + i];
                        lm2[8L * eb[s] + 4 * h + i] = id == -1 ? neq : id;
                    }
        }
    free(L.head); free(L.next); free(L.J); free(L.num);
    return nedge;
}

/* BlockDiag_precond_EDE_setup.c:82-103: the two nodes of every edge
 * (smaller first) from order and EDGE_by_Element */
void orc_ede4_edge_nodes(int nel, const int *vertex, const int *order, const int *edge_by_elem,
                         int *enodes)
{
    static const int ends[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (long E = 0; E < nel; E++) {
        const int *V = vertex + 3 * E, *o = order + 3 * E;
        const int Io[3] = {V[o[0]], V[o[1]], V[o[2]]};
        for (int s = 0; s < 3; s++) {
            enodes[2L * edge_by_elem[3 * E + s]] = Io[ends[s][0]];
            enodes[2L * edge_by_elem[3 * E + s] + 1] = Io[ends[s][1]];
        }
    }
}

/* ede_assembly.c:3-113 of the Euler driver over zeroed edge matrices
 * (setzeros.c:11-13): per element, edges (I1,I2), (I1,I3), (I2,I3) in turn,
 * entry (r, c) of the 8x8 edge matrix += Me[row][col] of the element, times
 * 0.5 when row and column belong to the same node */
void orc_ede4_assembly(int nel, int nedge, const double *A12, const int *order,
                       const int *edge_by_elem, double *Aede)
{
    static const int ends[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    memset(Aede, 0, sizeof(double) * 64 * (size_t)nedge);
    for (long E = 0; E < nel; E++) {
        const double *Me = A12 + 144 * E;
        const int *o = order + 3 * E;
        for (int s = 0; s < 3; s++) {
            double *M = Aede + 64L * edge_by_elem[3 * E + s];
            const int node[2] = {o[ends[s][0]], o[ends[s][1]]};
            for (int r = 0; r < 8; r++)
                for (int c = 0; c < 8; c++) {
                    const int rn = r / 4, cn = c / 4;
                    const double v = Me[12 * (4 * node[rn] + r % 4) + 4 * node[cn] + c % 4];
                    M[8 * r + c] += rn == cn ? 0.5 * v : v;
                }
        }
    }
}

/* Block_scaling.c:10-59 Block_scaling_EBE (2D, NDOF 4): the block-diagonal
 * setup (F preconditioned), then per element the nine 4x4 blocks A_vw with
 * the prescribed dofs of each node removed (matrix_Bondary_Conditions,
 * :73-92: row and column zeroed, 1 on the diagonal of its own block) and
 * every block row v left-multiplied by invBD of its node (matrix_Scaling,
 * :95-118, sum over j ascending from 0).  ids[nnodes*4] = Node[].id (-1 =
 * prescribed).  A is scaled in place; BD/invBD as orc_blockdiag_setup_ebe. */
void orc_block_scaling_ebe(int neq, int nel, int nnodes, const int *vertex, const int *ids,
                           const int *Id, double *A, double *BD, double *invBD, double *F)
{
    orc_blockdiag_setup_ebe(neq, nel, nnodes, vertex, A, Id, BD, invBD, F);
    for (long E = 0; E < nel; E++) {
        double *Ae = A + 144 * E;
        const int *V = vertex + 3 * E;
        double M[3][3][4][4];   /* M[v][w][i][j] = Ae[12(4v+i) + 4w+j] */
        for (int v = 0; v < 3; v++)
            for (int w = 0; w < 3; w++)
                for (int i = 0; i < 4; i++)
                    for (int j = 0; j < 4; j++) M[v][w][i][j] = Ae[12 * (4 * v + i) + 4 * w + j];
        /* boundary conditions, Block_scaling.c:46-50 call order */
        for (int I = 0; I < 4; I++)
            for (int v = 0; v < 3; v++) {
                if (ids[4 * V[v] + I] >= 0) continue;
                for (int i = 0; i < 4; i++) {
                    M[v][v][I][i] = 0;
                    M[v][v][i][I] = 0;
                    for (int w = 0; w < 3; w++)
                        if (w != v) {
                            M[v][w][I][i] = 0;   /* row I of v's off-diagonal blocks */
                            M[w][v][i][I] = 0;   /* column I of the blocks into v */
                        }
                }
                M[v][v][I][I] = 1;
            }
        /* scaling of block row v by invBD of its node */
        for (int v = 0; v < 3; v++) {
            const double *iD = invBD + 16 * V[v];
            for (int i = 0; i < 4; i++)
                for (int k = 0; k < 4; k++)
                    for (int w = 0; w < 3; w++) {
                        double s = 0;
                        for (int j = 0; j < 4; j++) s += iD[4 * i + j] * M[v][w][j][k];
                        Ae[12 * (4 * v + i) + 4 * w + k] = s;
                    }
        }
    }
}

/* ======================================================================== */
/* transient 2D SUPG (02_Transport_Equation)                                 */
/* ======================================================================== */
/* 02 CAU_ShockCapture.c:3-30: the element residual includes the time
 * derivative (Ru = du_b + beta.grad u + gamma u_b - f_b, :10) */
static double cau_shock_tr(double kx, double ky, double Be_x, double Be_y, double gamma,
                           double ue1, double ue2, double ue3, double ueb, double dueb,
                           double feb, double y23, double y31, double y12, double x32,
                           double x13, double x21, double invArea, double h)
{
    const double third = 1. / 3.;
    double BetaGradu = 0.5 * invArea * (ue1 * (Be_x * y23 + Be_y * x32) +
                                        ue2 * (Be_x * y31 + Be_y * x13) +
                                        ue3 * (Be_x * y12 + Be_y * x21));
    double Ru = dueb + BetaGradu + gamma * ueb - feb;
    double g0 = 0.5 * invArea * (y23 * ue1 + y31 * ue2 + y12 * ue3);
    double g1 = 0.5 * invArea * (x32 * ue1 + x13 * ue2 + x21 * ue3);
    double ng = sqrt(g0 * g0 + g1 * g1);
    double bxp = BetaGradu * g0 / (ng * ng);
    double byp = BetaGradu * g1 / (ng * ng);
    double kt = bxp * bxp * kx + byp * byp * ky;
    double nbp = sqrt(bxp * bxp + byp * byp);
    double Pee = 0.5 * h * nbp * nbp * nbp / kt;
    if (ng > 0) {
        if (third * Pee < 0.7)
            return 0.5 * h * third * Pee * fabs(Ru) / ng;
        return 0.5 * h * 0.7 * fabs(Ru) / ng;
    }
    return 0;
}

/* Build_M_K_R_SUPG.c:6-184 (operand order kept: bitwise at -ffp-contract=off) */
void orc_build_supg2d_transient(int nnodes, int nel, int neq, const double *x, const double *y,
                                const int *vertex, const int *lm, const double *U,
                                const double *dU, const orc_prob2d *pb, double alpha, double dt,
                                double *A, double *R)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0, twelfth = 1.0 / 12.0;
    double kx, ky, gamma;
    (void)nnodes;
    if (pb->kind == ORC_P_VALIDACAO) { kx = 1.0; ky = 1.0; gamma = 1; }
    else { kx = pb->kappa; ky = pb->kappa; gamma = pb->gamma; }
    for (int r = 0; r <= neq; r++) R[r] = 0.0;
    for (long E = 0; E < nel; E++) {
        const int J1 = vertex[3 * E], J2 = vertex[3 * E + 1], J3 = vertex[3 * E + 2];
        const double X0 = x[J1], X1 = x[J2], X2 = x[J3];
        const double Y0 = y[J1], Y1 = y[J2], Y2 = y[J3];
        double y23 = Y1 - Y2, y31 = Y2 - Y0, y12 = Y0 - Y1;
        double x32 = X2 - X1, x13 = X0 - X2, x21 = X1 - X0;
        double TwoA = x21 * y31 - x13 * y12;
        double Area = 0.5 * TwoA;
        double invArea = 1.0 / Area;
        double h = sqrt(fabs(TwoA));
        double XB = third * (X0 + X1 + X2);
        double YB = third * (Y0 + Y1 + Y2);
        double Beta[2];
        orc_prob2d_velocity(pb, XB, YB, Beta);
        double Be_x = Beta[0], Be_y = Beta[1];
        double norm = sqrt(Be_x * Be_x + Be_y * Be_y);
        double tau = norm > 0 ? 0.5 * h / norm : 0;
        double fe1 = orc_prob2d_font(pb, X0, Y0, kx, gamma, Be_x, Be_y);
        double fe2 = orc_prob2d_font(pb, X1, Y1, kx, gamma, Be_x, Be_y);
        double fe3 = orc_prob2d_font(pb, X2, Y2, kx, gamma, Be_x, Be_y);
        double feb = third * (fe1 + fe2 + fe3);
        const double ue[3] = {U[J1], U[J2], U[J3]};
        const double due[3] = {dU[J1], dU[J2], dU[J3]};
        double ueb = third * (ue[0] + ue[1] + ue[2]);
        double dueb = third * (due[0] + due[1] + due[2]);

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
        /* 02 h_shock_2sqrtArea.c:3-6 */
        double hs = fabs(sqrt(2 * Area));
        double Eu = cau_shock_tr(kx, ky, Be_x, Be_y, gamma, ue[0], ue[1], ue[2], ueb, dueb, feb,
                                 y23, y31, y12, x32, x13, x21, invArea, hs);
        double Csc12 = Eu * 0.25 * invArea * (y23 * y31 + x32 * x13);
        double Csc13 = Eu * 0.25 * invArea * (y23 * y12 + x32 * x21);
        double Csc23 = Eu * 0.25 * invArea * (y31 * y12 + x13 * x21);

        double me[3][3], ke[3][3];
        me[0][0] = sixth * Area + tau * C11;
        me[0][1] = twelfth * Area + tau * C11;
        me[0][2] = twelfth * Area + tau * C11;
        me[1][0] = twelfth * Area + tau * C12;
        me[1][1] = sixth * Area + tau * C12;
        me[1][2] = twelfth * Area + tau * C12;
        me[2][0] = twelfth * Area + tau * C13;
        me[2][1] = twelfth * Area + tau * C13;
        me[2][2] = sixth * Area + tau * C13;
        ke[0][0] = -(D12 + D13) + C11 - (Cpg12 + Cpg13) + R11 + Rpg11 - (Csc12 + Csc13);
        ke[0][1] = D12 + C12 + Cpg12 + R12 + Rpg11 + Csc12;
        ke[0][2] = D13 + C13 + Cpg13 + R12 + Rpg11 + Csc13;
        ke[1][0] = D12 + C11 + Cpg12 + R12 + Rpg21 + Csc12;
        ke[1][1] = -(D12 + D23) + C12 - (Cpg12 + Cpg23) + R11 + Rpg21 - (Csc12 + Csc23);
        ke[1][2] = D23 + C13 + Cpg23 + R12 + Rpg21 + Csc23;
        ke[2][0] = D13 + C11 + Cpg13 + R12 + Rpg31 + Csc13;
        ke[2][1] = D23 + C12 + Cpg23 + R12 + Rpg31 + Csc23;
        ke[2][2] = -(D13 + D23) + C13 - (Cpg13 + Cpg23) + R11 + Rpg31 - (Csc13 + Csc23);
        double *a = A + 9 * E;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) a[3 * i + j] = me[i][j] + alpha * dt * ke[i][j];

        const int *le = lm + 3 * E;
        R[le[0]] += Area * twelfth * (2 * fe1 + fe2 + fe3) + sixth * tau * (fe1 + fe2 + fe3) * (Be_x * y23 + Be_y * x32);
        R[le[1]] += Area * twelfth * (fe1 + 2 * fe2 + fe3) + sixth * tau * (fe1 + fe2 + fe3) * (Be_x * y31 + Be_y * x13);
        R[le[2]] += Area * twelfth * (fe1 + fe2 + 2 * fe3) + sixth * tau * (fe1 + fe2 + fe3) * (Be_x * y12 + Be_y * x21);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) R[le[i]] -= me[i][j] * due[j] + ke[i][j] * ue[j];
        R[neq] = 0;
    }
}

/* ======================================================================== */
/* Euler 2D element build (03_Euler_Equations)                               */
/* ======================================================================== */
/* Ax_Ay_calculations.c:3-49 (dimensionless, reference Mach number) and
 * :51-95 (dimensional): the flux Jacobians at the barycentre.  Ay[0][3],
 * which the reference leaves unset and never reads, is 0 here. */
static void eu_jacobians(double gamma, double Mach, int dimless, const double Ub[4],
                         double Ax[4][4], double Ay[4][4])
{
    double v1 = Ub[1] / Ub[0], v2 = Ub[2] / Ub[0], E = Ub[3] / Ub[0];
    memset(Ax, 0, 16 * sizeof(double));
    memset(Ay, 0, 16 * sizeof(double));
    Ax[0][1] = 1.0;
    Ay[0][2] = 1.0;
    if (dimless) {
        double nv = v1 * v1 + v2 * v2, e2 = Mach * Mach;
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
        double rho = Ub[0], nU = (Ub[1] * Ub[1]) + (Ub[2] * Ub[2]);
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

/* x^T A0 x in the grouping of Delta_CAU.c:35-44 (upper triangle, off-diagonal
 * terms doubled) */
static double eu_quad(const double A0[4][4], const double v[4])
{
    return v[0] * (A0[0][0] * v[0] + 2.0 * A0[0][1] * v[1] + 2.0 * A0[0][2] * v[2] + 2.0 * A0[0][3] * v[3]) +
           v[1] * (A0[1][1] * v[1] + 2.0 * A0[1][2] * v[2] + 2.0 * A0[1][3] * v[3]) +
           v[2] * (A0[2][2] * v[2] + 2.0 * A0[2][3] * v[3]) +
           v[3] * A0[3][3] * v[3];
}

/* Delta_CAU.c:4-78: the CAU discontinuity-capturing coefficient.  The 4x4
 * products run over k ascending; the terms the reference leaves out are
 * exact zeros of Ax/Ay (rows 0, Ax[2][3], Ay[1][3]), so the sums agree. */
static double eu_delta_cau(double tol, const double gx[4], const double gy[4],
                           const double Ax[4][4], const double Ay[4][4], const double A0[4][4],
                           const double dUb[4], double y23, double y31, double x32, double x13,
                           double twoArea)
{
    double Agx[4], Agy[4], kx[4], ky[4], L[4];
    for (int i = 0; i < 4; i++) {
        double sx = 0, sy = 0;
        int first = 1;
        for (int k = 0; k < 4; k++) {
            if (first) { sx = Ax[i][k] * gx[k]; sy = Ay[i][k] * gy[k]; first = 0; }
            else { sx = sx + Ax[i][k] * gx[k]; sy = sy + Ay[i][k] * gy[k]; }
        }
        Agx[i] = sx;
        Agy[i] = sy;
    }
    for (int i = 0; i < 4; i++) {
        kx[i] = (y23 * gx[i] + x32 * gy[i]) / twoArea;
        ky[i] = (y31 * gx[i] + x13 * gy[i]) / twoArea;
    }
    double ngk = sqrt(eu_quad(A0, kx) + eu_quad(A0, ky));
    for (int i = 0; i < 4; i++) L[i] = dUb[i] + Agx[i] + Agy[i];
    double nL = sqrt(eu_quad(A0, L));
    double ngu = sqrt(eu_quad(A0, gx) + eu_quad(A0, gy));
    return ngu >= tol ? nL / ngk : 0.0;
}

/* rotation.c:3-45: the slip-wall rotation of node `a`'s block row (every
 * column block) and of its load entries */
static void eu_rotate(int a, double theta, double M[12][12], double F[12])
{
    const double cs = cos(theta), sn = sin(theta);
    const int I = 4 * a;
    for (int K = 0; K < 3; K++) {
        const int J = 4 * K;
        double b = M[I][J + 1], c = M[I][J + 2], e = M[I + 1][J], f = M[I + 1][J + 1];
        double g = M[I + 1][J + 2], h = M[I + 1][J + 3], i = M[I + 2][J], j = M[I + 2][J + 1];
        double k = M[I + 2][J + 2], l = M[I + 2][J + 3], n = M[I + 3][J + 1], o = M[I + 3][J + 2];
        M[I][J + 1] = b * cs - c * sn;
        M[I][J + 2] = b * sn + c * cs;
        M[I + 1][J] = e * cs - i * sn;
        M[I + 1][J + 1] = cs * (f * cs - j * sn) - sn * (g * cs - k * sn);
        M[I + 1][J + 2] = cs * (g * cs - k * sn) + sn * (f * cs - j * sn);
        M[I + 1][J + 3] = h * cs - l * sn;
        M[I + 2][J] = e * sn + i * cs;
        M[I + 2][J + 1] = cs * (f * sn + j * cs) - sn * (g * sn + k * cs);
        M[I + 2][J + 2] = cs * (g * sn + k * cs) + sn * (f * sn + j * cs);
        M[I + 2][J + 3] = h * sn + l * cs;
        M[I + 3][J + 1] = n * cs - o * sn;
        M[I + 3][J + 2] = n * sn + o * cs;
    }
    double b = F[I + 1], c = F[I + 2];
    F[I + 1] = b * cs - c * sn;
    F[I + 2] = b * sn + c * cs;
}

/* Build_M_K_F_SUPG.c:6-1073 with Delta_CAU (see header) */
void orc_build_euler2d(int nnodes, int nel, int neq, const double *x, const double *y,
                       const int *vertex, const int *lm, const double *U, const double *dU,
                       const int *rot, const double *theta, double gamma, double Mach,
                       int dimless, double alpha, double dt, double tol, double *A, double *R)
{
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    (void)nnodes;
    for (int r = 0; r <= neq; r++) R[r] = 0.0;
    for (long e = 0; e < nel; e++) {
        const int Jn[3] = {vertex[3 * e], vertex[3 * e + 1], vertex[3 * e + 2]};
        const double x1 = x[Jn[0]], x2 = x[Jn[1]], x3 = x[Jn[2]];
        const double y1 = y[Jn[0]], y2 = y[Jn[1]], y3 = y[Jn[2]];
        const double y23 = y2 - y3, y31 = y3 - y1, y12 = y1 - y2;
        const double x32 = x3 - x2, x13 = x1 - x3, x21 = x2 - x1;
        const double twoArea = fabs(x21 * y31 - x13 * y12);
        const double Area = 0.5 * twoArea;
        /* node a's geometric factors (y_a, x_a): (y23, x32), (y31, x13), (y12, x21) */
        const double ya[3] = {y23, y31, y12}, xa[3] = {x32, x13, x21};
        double Ue[12], dUe[12], Ub[4], dUb[4], gx[4], gy[4];
        for (int a = 0; a < 3; a++)
            for (int d = 0; d < 4; d++) {
                Ue[4 * a + d] = U[4 * Jn[a] + d];
                dUe[4 * a + d] = dU[4 * Jn[a] + d];
            }
        for (int i = 0; i < 4; i++) {
            Ub[i] = (Ue[i] + Ue[i + 4] + Ue[i + 8]) * third;
            dUb[i] = (dUe[i] + dUe[i + 4] + dUe[i + 8]) * third;
        }
        for (int i = 0; i < 4; i++) {
            gx[i] = (Ue[i] * y23 + Ue[i + 4] * y31 + Ue[i + 8] * y12) / twoArea;
            gy[i] = (Ue[i] * x32 + Ue[i + 4] * x13 + Ue[i + 8] * x21) / twoArea;
        }
        const double v1 = Ub[1] / Ub[0], v2 = Ub[2] / Ub[0], E = Ub[3] / Ub[0];
        const double nU23 = (Ub[1] * Ub[1]) + (Ub[2] * Ub[2]);
        double Ax[4][4], Ay[4][4], A0[4][4];
        eu_jacobians(gamma, Mach, dimless, Ub, Ax, Ay);
        /* :131-159 inverse Riemannian metric A0^{-1} from entropy variables */
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
        /* :161-237 AxAx, AxAy, AyAx, AyAy (k ascending, see eu_delta_cau) */
        double P[4][4][4];   /* P[0]=AxAx P[1]=AxAy P[2]=AyAx P[3]=AyAy */
        const double (*Lm[4])[4] = {Ax, Ax, Ay, Ay}, (*Rm[4])[4] = {Ax, Ay, Ax, Ay};
        for (int q = 0; q < 4; q++)
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    double s = Lm[q][i][0] * Rm[q][0][j];
                    for (int k = 1; k < 4; k++) s = s + Lm[q][i][k] * Rm[q][k][j];
                    P[q][i][j] = s;
                }
        const double delta = eu_delta_cau(tol, gx, gy, Ax, Ay, A0, dUb, y23, y31, x32, x13, twoArea);
        /* :243-269 SUPG tau */
        double betax = 2.0 * (gx[0] * Ub[0] + gx[1] * Ub[1] + gx[2] * Ub[2] + gx[3] * Ub[3]);
        double betay = 2.0 * (gy[0] * Ub[0] + gy[1] * Ub[1] + gy[2] * Ub[2] + gy[3] * Ub[3]);
        const double betaxy = sqrt(betax * betax + betay * betay);
        if (betaxy != 0.0) { betax = betax / betaxy; betay = betay / betaxy; }
        const double h = sqrt(twoArea);
        const double avb = fabs(v1 * betax + v2 * betay);
        const double c = sqrt(gamma * (gamma - 1.0) * (E - 0.5 * (pow(v1, 2) + pow(v2, 2))));
        const double CFL = ((c + avb) * dt) / h;
        const double psi = (2.0 * alpha * CFL) * (1.0 / (1.0 + 2.0 * alpha * CFL));
        const double tau_d = delta / pow((c + avb), 2);
        const double tau_a = h / (2.0 * (c + avb));
        const double tau_t = (2.0 / (3.0 * (1.0 + 2.0 * alpha * CFL))) * tau_a;
        double tau = tau_t + psi * (tau_a - tau_d);
        if (tau < 0.0) tau = 0.0;
        /* :273-332 C_a = (y_a Ax + x_a Ay)/6 (k-sums as above, Ca14 etc. = 0) */
        double Cm[3][4][4];
        for (int a = 0; a < 3; a++)
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++)
                    Cm[a][i][j] = (ya[a] * Ax[i][j] + xa[a] * Ay[i][j]) * sixth;
        /* the reference writes row 0 and the two structural zeros directly */
        for (int a = 0; a < 3; a++) {
            Cm[a][0][0] = 0.0; Cm[a][0][1] = ya[a] * sixth; Cm[a][0][2] = xa[a] * sixth;
            Cm[a][0][3] = 0.0;
            Cm[a][1][3] = ya[a] * Ax[1][3] * sixth;
            Cm[a][2][3] = xa[a] * Ay[2][3] * sixth;
        }
        const double Mg1 = Area * sixth, Mg2 = Area / 12.0;
        /* :618-737 a_a = y_a [AxAx|AxAy] + x_a [AyAx|AyAy]; :739-855 Kpg */
        double ax[3][4][4], ay[3][4][4];
        for (int a = 0; a < 3; a++)
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    ax[a][i][j] = ya[a] * P[0][i][j] + xa[a] * P[2][i][j];
                    ay[a][i][j] = ya[a] * P[1][i][j] + xa[a] * P[3][i][j];
                }
        const double t4a = tau / (4 * Area);
        const double d4a = delta / (4.0 * Area);
        double Kc[3][3];   /* :858-861 */
        Kc[0][1] = Kc[1][0] = d4a * (y23 * y31 + x32 * x13);
        Kc[0][2] = Kc[2][0] = d4a * (y23 * y12 + x32 * x21);
        Kc[1][2] = Kc[2][1] = d4a * (y31 * y12 + x13 * x21);
        double Me[12][12], Ke[12][12];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int i = 0; i < 4; i++)
                    for (int j = 0; j < 4; j++) {
                        const double mpg = Cm[a][i][j] * tau;
                        Me[4 * a + i][4 * b + j] = i == j ? (a == b ? Mg1 : Mg2) + mpg : mpg;
                        if (a != b) {
                            const double kpg = t4a * (ya[b] * ax[a][i][j] + xa[b] * ay[a][i][j]);
                            double v = Cm[b][i][j] + kpg;
                            if (i == j) v = v + Kc[a][b];
                            Ke[4 * a + i][4 * b + j] = v;
                        }
                    }
        for (int a = 0; a < 3; a++) {
            const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    const double k1p = t4a * (ya[b1] * ax[a][i][j] + xa[b1] * ay[a][i][j]);
                    const double k2p = t4a * (ya[b2] * ax[a][i][j] + xa[b2] * ay[a][i][j]);
                    double v = Cm[a][i][j] - (k1p + k2p);
                    if (i == j) v = v - (Kc[a][b1] + Kc[a][b2]);
                    Ke[4 * a + i][4 * a + j] = v;
                }
        }
        /* :1034-1044 Ae, Re */
        double Ae[12][12], Re[12];
        for (int i = 0; i < 12; i++) {
            double md = 0, ku = 0;
            for (int j = 0; j < 12; j++) {
                md += Me[i][j] * dUe[j];
                ku += Ke[i][j] * Ue[j];
                Ae[i][j] = Me[i][j] + alpha * dt * Ke[i][j];
            }
            Re[i] = -md - ku;
        }
        /* :1046-1057 slip walls */
        if (rot)
            for (int a = 0; a < 3; a++)
                if (rot[Jn[a]]) eu_rotate(a, theta[Jn[a]], Ae, Re);
        const int *le = lm + 12 * e;
        for (int i = 0; i < 12; i++) R[le[i]] += Re[i];
        R[neq] = 0;
        memcpy(A + 144 * e, Ae, sizeof Ae);
    }
}

/* ---- time integration: Predictor_Old.c:4-145, Criteria.c:25-67 ---------- */
int orc_predictor_old_tr(int nnodes, int nel, int neq, const double *x, const double *y,
                         const int *type, const int *id, const int *vertex, const int *lm,
                         const orc_prob2d *pb, double ud, const orc_pred_params *pp, int diag,
                         const double *u0, double *u, int *hist, int hist_cap, int *nsteps,
                         double *t_final)
{
    const double alpha = pp->alpha, dt = pp->dt;
    double *a = calloc(neq + 1, sizeof(double)), *Da = calloc(neq + 1, sizeof(double));
    double *u_old = calloc(neq + 1, sizeof(double)), *R = calloc(neq + 1, sizeof(double));
    double *A = calloc((size_t)nel * 9, sizeof(double));
    double *U = calloc(nnodes, sizeof(double)), *dU = calloc(nnodes, sizeof(double));
    double *Dg = calloc(neq + 1, sizeof(double)), *iD = calloc(neq + 1, sizeof(double));
    double *diff = calloc(neq + 1, sizeof(double));
    int rc = 0, passo = 0;
    memcpy(u, u0, sizeof(double) * (neq + 1));   /* InitialSolution */
    double t = 0.0;
    int stop_time;
    do {
        passo++;
        t += dt;
        int i = 0;
        for (int I = 0; I < neq; I++) {   /* predictor (:57-62) */
            u_old[I] = u[I];
            u[I] += (1.0 - alpha) * dt * a[I];
            a[I] = 0.0;
        }
        double norm_a = sqrt(orc_ddot(neq, a, a));
        double norm_Da = norm_a + 1.0;
        int stop;
        do {   /* multicorrection (:68-125) */
            i++;
            for (int I = 0; I < nnodes; I++) {   /* eval_U_dU.c */
                if (type[I] > 0) {
                    U[I] = u[id[I]];
                    dU[I] = a[id[I]];
                } else {
                    U[I] = ud;
                    dU[I] = 0.0;
                }
            }
            orc_build_supg2d_transient(nnodes, nel, neq, x, y, vertex, lm, U, dU, pb, alpha, dt,
                                       A, R);
            if (diag) orc_diag_setup_ebe(neq, nel, 3, A, lm, Dg, iD, R);
            if (orc_dense_solve_ebe(neq, nel, 3, A, lm, diag ? iD : NULL, R, Da) && !rc) rc = 1;
            orc_daxpy(neq, 1, Da, a);
            orc_daxpy(neq, alpha * dt, Da, u);
            norm_a = sqrt(orc_ddot(neq, a, a));
            norm_Da = sqrt(orc_ddot(neq, Da, Da));
            if (pp->stop_by_norm)   /* StopByNorm (Criteria.c:35-41) */
                stop = norm_Da < pp->nl_tol * norm_a || i >= pp->nl_maxiter;
            else   /* StopByIterations (:26-32) */
                stop = i >= pp->nl_maxiter;
        } while (!stop);
        if (passo - 1 < hist_cap) hist[passo - 1] = i;
        if (pp->steady) {   /* StopBySteadyState (:44-63) */
            for (int I = 0; I < neq; I++) diff[I] = u[I] - u_old[I];
            const double norm_diff = sqrt(orc_ddot(neq, diff, diff));
            const double norm_u = sqrt(orc_ddot(neq, u, u));
            stop_time = pp->time_tol * norm_u > norm_diff || fabs(pp->final_time - t) < dt;
        } else {   /* StopByTime (:65-72) */
            stop_time = fabs(pp->final_time - t) < 1e-15;
        }
        if (!stop_time && pp->cap > 0 && passo >= pp->cap) stop_time = 1;
    } while (!stop_time);
    *nsteps = passo;
    *t_final = t;
    free(a); free(Da); free(u_old); free(R); free(A); free(U); free(dU); free(Dg); free(iD);
    free(diff);
    return rc;
}

int orc_predictor_old_eu(int nnodes, int nel, int neq, const double *x, const double *y,
                         const int *vertex, const int *lm, const int *ids, const double *presc,
                         const int *rot, const double *theta, double gamma, double mach,
                         int dimless, double stab_tol, const orc_pred_params *pp,
                         int block_scaling, const double *u0, double *u, int *hist, int hist_cap,
                         int *nsteps, double *t_final)
{
    return orc_predictor_old_eu_scheme(nnodes, nel, neq, x, y, vertex, lm, ids, presc, rot,
                                       theta, gamma, mach, dimless, stab_tol, pp, block_scaling,
                                       0, u0, u, hist, hist_cap, nsteps, t_final);
}

/* the dense matrix of a scheme's product, column j = product of e_j, then
 * the direct solve (X = M^{-1} B, X[neq] = 0) */
static int dense_solve_scheme(int scheme, int neq, int nel, int nedge, int nnodes,
                              const double *A, const int *lm, const double *Aede, const int *lm2,
                              const int *Id, const double *iBD, const double *B, double *X)
{
    double *M = calloc((size_t)neq * neq, sizeof(double));
    double *e = calloc(neq + 1, sizeof(double)), *q = calloc(neq + 1, sizeof(double));
    for (int j = 0; j < neq; j++) {
        memset(e, 0, sizeof(double) * (neq + 1));
        e[j] = 1.0;
        if (scheme == 1) orc_ebe2mv_items(neq, nel, 12, A, lm, nnodes, Id, iBD, NULL, e, q);
        else if (scheme == 2) orc_ebemv(neq, nedge, 8, Aede, lm2, e, q);
        else orc_ebe2mv_items(neq, nedge, 8, Aede, lm2, nnodes, Id, iBD, NULL, e, q);
        for (int i = 0; i < neq; i++) M[(size_t)i * neq + j] = q[i];
    }
    memcpy(X, B, sizeof(double) * neq);
    const int rc = orc_dense_solve(neq, M, X);
    X[neq] = 0.0;
    free(M); free(e); free(q);
    return rc;
}

/* scheme (setMatrixVectorProductType.c / setPreconditioner.c of the Euler
 * driver): 0 EBE (ebemvNDOF4, NO_precond_setup; block_scaling: NO_scaling or
 * Block_scaling_EBE), 1 EBE2 (ebe2mvNDOF4, BlockDiag_precond_EBE_setup on R),
 * 2 EDE (ede_assembly, edemvNDOF4, NO_precond_setup), 3 EDE2 (ede_assembly,
 * ede2mvNDOF4, BlockDiag_precond_EDE_setup on R); the solve is direct */
int orc_predictor_old_eu_scheme(int nnodes, int nel, int neq, const double *x, const double *y,
                                const int *vertex, const int *lm, const int *ids,
                                const double *presc, const int *rot, const double *theta,
                                double gamma, double mach, int dimless, double stab_tol,
                                const orc_pred_params *pp, int block_scaling, int scheme,
                                const double *u0, double *u, int *hist, int hist_cap,
                                int *nsteps, double *t_final)
{
    const double alpha = pp->alpha, dt = pp->dt;
    int nedge = 0, *order = NULL, *ebe = NULL, *lm2 = NULL, *enodes = NULL;
    double *Aede = NULL;
    if (scheme >= 2) {
        order = calloc(3 * (size_t)nel, sizeof(int));
        ebe = calloc(3 * (size_t)nel, sizeof(int));
        nedge = orc_ede4_init(nnodes, nel, neq, vertex, ids, order, ebe, NULL);
        lm2 = calloc(8 * (size_t)nedge + 1, sizeof(int));
        enodes = calloc(2 * (size_t)nedge + 1, sizeof(int));
        Aede = calloc(64 * (size_t)nedge + 1, sizeof(double));
        orc_ede4_init(nnodes, nel, neq, vertex, ids, order, ebe, lm2);
        orc_ede4_edge_nodes(nel, vertex, order, ebe, enodes);
    }
    double *a = calloc(neq + 1, sizeof(double)), *Da = calloc(neq + 1, sizeof(double));
    double *u_old = calloc(neq + 1, sizeof(double)), *R = calloc(neq + 1, sizeof(double));
    double *A = calloc((size_t)nel * 144, sizeof(double));
    double *U = calloc(4 * (size_t)nnodes, sizeof(double));
    double *dU = calloc(4 * (size_t)nnodes, sizeof(double));
    double *BD = calloc(16 * (size_t)nnodes, sizeof(double));
    double *iBD = calloc(16 * (size_t)nnodes, sizeof(double));
    int *Id = calloc(4 * (size_t)nnodes, sizeof(int));
    double *diff = calloc(neq + 1, sizeof(double));
    for (long q = 0; q < 4L * nnodes; q++) Id[q] = ids[q] < 0 ? neq : ids[q];
    int rc = 0, passo = 0, stop_time;
    memcpy(u, u0, sizeof(double) * (neq + 1));
    double t = 0.0;
    do {
        passo++;
        t += dt;
        int i = 0;
        for (int I = 0; I < neq; I++) {
            u_old[I] = u[I];
            u[I] += (1.0 - alpha) * dt * a[I];
            a[I] = 0.0;
        }
        double norm_a = sqrt(orc_ddot(neq, a, a));
        double norm_Da = norm_a + 1.0;
        int stop;
        do {
            i++;
            for (long I = 0; I < nnodes; I++)   /* eval_U_dU.c of the 03 driver */
                for (int k = 0; k < 4; k++) {
                    const int id = ids[4 * I + k];
                    U[4 * I + k] = id >= 0 ? u[id] : presc[k];
                    dU[4 * I + k] = id >= 0 ? a[id] : 0.0;
                }
            orc_build_euler2d(nnodes, nel, neq, x, y, vertex, lm, U, dU, rot, theta, gamma, mach,
                              dimless, alpha, dt, stab_tol, A, R);
            if (scheme >= 2) orc_ede4_assembly(nel, nedge, A, order, ebe, Aede);
            if (scheme == 0 && block_scaling)
                orc_block_scaling_ebe(neq, nel, nnodes, vertex, ids, Id, A, BD, iBD, R);
            if (scheme == 1) orc_blockdiag_setup_ebe(neq, nel, nnodes, vertex, A, Id, BD, iBD, R);
            if (scheme == 3)
                orc_blockdiag_setup_items(neq, nedge, 2, nnodes, enodes, Aede, Id, BD, iBD, R);
            const int src = scheme == 0
                                ? orc_dense_solve_ebe(neq, nel, 12, A, lm, NULL, R, Da)
                                : dense_solve_scheme(scheme, neq, nel, nedge, nnodes, A, lm, Aede,
                                                     lm2, Id, iBD, R, Da);
            if (src && !rc) rc = 1;
            orc_daxpy(neq, 1, Da, a);
            orc_daxpy(neq, alpha * dt, Da, u);
            norm_a = sqrt(orc_ddot(neq, a, a));
            norm_Da = sqrt(orc_ddot(neq, Da, Da));
            if (pp->stop_by_norm)
                stop = norm_Da < pp->nl_tol * norm_a || i >= pp->nl_maxiter;
            else
                stop = i >= pp->nl_maxiter;
        } while (!stop);
        if (passo - 1 < hist_cap) hist[passo - 1] = i;
        if (pp->steady) {
            for (int I = 0; I < neq; I++) diff[I] = u[I] - u_old[I];
            const double norm_diff = sqrt(orc_ddot(neq, diff, diff));
            const double norm_u = sqrt(orc_ddot(neq, u, u));
            stop_time = pp->time_tol * norm_u > norm_diff || fabs(pp->final_time - t) < dt;
        } else {
            stop_time = fabs(pp->final_time - t) < 1e-15;
        }
        if (!stop_time && pp->cap > 0 && passo >= pp->cap) stop_time = 1;
    } while (!stop_time);
    *nsteps = passo;
    *t_final = t;
    free(a); free(Da); free(u_old); free(R); free(A); free(U); free(dU); free(BD); free(iBD);
    free(Id); free(diff);
    free(order); free(ebe); free(lm2); free(enodes); free(Aede);
    return rc;
}

/* ---- scaling.c:10-180: Left / LeftRight diagonal scalings -----------------
 * Diag, invDiag [neq+1] from the Diag setup of the same matrix (which also
 * did F *= invDiag); invDiag[neq] reads as invDiag[neq-1] (the reference
 * writes it so, :19).  Items: A [nitems*k*k] row-major, lm [nitems*k]. */
static double orc_inv_at(const double *iD, int g, int neq) { return iD[g < neq ? g : neq - 1]; }

void orc_scaling_items(int neq, int nitems, int k, double *A, const int *lm, const double *Diag,
                       const double *invDiag, int leftright, double *F)
{
    for (long E = 0; E < nitems; E++)
        for (int I = 0; I < k; I++)
            for (int J = 0; J < k; J++) {
                double *a = &A[(long)k * k * E + k * I + J];
                const double di = orc_inv_at(invDiag, lm[(long)k * E + I], neq);
                if (leftright) *a *= sqrt(di * orc_inv_at(invDiag, lm[(long)k * E + J], neq));
                else *a *= di;
            }
    if (leftright)
        for (int I = 0; I < neq; I++) F[I] *= sqrt(Diag[I]);
}

void orc_scaling_csr(int neq, double *AA, const int *JA, const int *IA, const double *Diag,
                     const double *invDiag, int leftright, double *F)
{
    for (int I = 0; I < neq; I++)
        for (int J = IA[I]; J < IA[I + 1]; J++) {
            if (leftright) AA[J] *= sqrt(invDiag[I] * orc_inv_at(invDiag, JA[J], neq));
            else AA[J] *= invDiag[I];
        }
    if (leftright)
        for (int I = 0; I < neq; I++) F[I] *= sqrt(Diag[I]);
}

/* Left_unscaling (:170-180) */
void orc_left_unscaling(int neq, const double *invDiag, double *x)
{
    for (int I = 0; I < neq; I++) x[I] *= sqrt(invDiag[I]);
}

/* ---- the drivers' file reads, as Preprocess.c does them with fscanf ------
 * Parameter lines: 's' "%s\t:%[^\n]", 'd' "%lf\t:%[^\n]", 'i' "%d\t:%[^\n]"
 * (01 2D Preprocess.c:44-61, 3D :45-83, 03 :45-74), 'p' "\n%[^\t]%[^\n]\n"
 * (03 Preprocess.c:54).  Each value is written back as text ('d' with %.17g)
 * into out, NUL-separated.  Returns the number of fields read. */
int orc_read_params(const char *path, const char *kinds, char *out, int cap)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char sval[4096], label[4096];
    int n = 0, used = 0;
    for (const char *k = kinds; *k; k++) {
        char txt[4200];
        int ok;
        if (*k == 'd') {
            double v;
            ok = fscanf(f, "%lf\t:%[^\n]", &v, label) >= 1;
            snprintf(txt, sizeof txt, "%.17g", v);
        } else if (*k == 'i') {
            int v;
            ok = fscanf(f, "%d\t:%[^\n]", &v, label) >= 1;
            snprintf(txt, sizeof txt, "%d", v);
        } else if (*k == 'p') {
            ok = fscanf(f, "\n%[^\t]%[^\n]\n", sval, label) >= 1;
            snprintf(txt, sizeof txt, "%s", sval);
        } else {
            ok = fscanf(f, "%s\t:%[^\n]", sval, label) >= 1;
            snprintf(txt, sizeof txt, "%s", sval);
        }
        if (!ok) break;
        const int len = (int)strlen(txt);
        if (used + len + 1 > cap) break;
        memcpy(out + used, txt, (size_t)len + 1);
        used += len + 1;
        n++;
    }
    fclose(f);
    return n;
}

/* mesh .dat: "%d" nnodes, per node ndim "%lf" and nint "%d", "%d" nel, per
 * element nvert + ecols "%d" (01 2D Preprocess.c:71-99, 03 :80-100) */
int orc_read_mesh(const char *path, int ndim, int nint, int nvert, int ecols, int *nnodes,
                  int *nel, double *xyz, int *ni, int *vertex, int *ei)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    int nn = 0, ne = 0;
    if (fscanf(f, "%d", &nn) != 1) { fclose(f); return -1; }
    for (long i = 0; i < nn; i++) {
        for (int d = 0; d < ndim; d++) {
            double v = 0;
            if (fscanf(f, "%lf", &v) != 1) { fclose(f); return -1; }
            if (xyz) xyz[i * ndim + d] = v;
        }
        for (int d = 0; d < nint; d++) {
            int v = 0;
            if (fscanf(f, "%d", &v) != 1) { fclose(f); return -1; }
            if (ni) ni[i * nint + d] = v;
        }
    }
    if (fscanf(f, "%d", &ne) != 1) { fclose(f); return -1; }
    for (long e = 0; e < ne; e++) {
        for (int d = 0; d < nvert + ecols; d++) {
            int v = 0;
            if (fscanf(f, "%d", &v) != 1) { fclose(f); return -1; }
            if (d < nvert) { if (vertex) vertex[e * nvert + d] = v; }
            else if (ei) ei[e * ecols + d - nvert] = v;
        }
    }
    fclose(f);
    *nnodes = nn;
    *nel = ne;
    return 0;
}
