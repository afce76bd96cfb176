/*
 * eu2d_adapter.c -- drop-in GPU objects for the 2D compressible Euler driver
 * (03_Euler_Equations/2D).
 *
 * Compiled with the driver's own flags and header (-DEulerEquations2D,
 * 03_Euler_Equations/2D/01_CommonFiles/EulerEquations.h), this file exports
 * the reference symbols the driver binds in setMatrixVectorProductType.c:
 * 10-28, setPreconditioner.c:6-92, setScaling.c:6-16 and setSolver.c:7:
 *
 *   ebemvNDOF4 / ebe2mvNDOF4 (P*, M*, FS*, double *P, double *Q)
 *        MatrixVector_Operations/EBE/ebemvNDOF4.c:5-112 (the 2D, 12x12
 *        branch) and EBE/ebe2mvNDOF4.c:4-79 (the product without each row
 *        node's own block, then BlockDiag2_precond or NOBlockDiag2_precond
 *        as Parameters->Preconditioner selects, :72-75)
 *   edemvNDOF4 / ede2mvNDOF4 (...)
 *        EDE/edemvNDOF4.c:4-72 and EDE/ede2mvNDOF4.c:4-61 (8x8 edge matrices)
 *   csrmv (...)                      CSR/csrmv.c:3-24
 *   BlockDiag_precond_EBE_setup / BlockDiag_precond_EDE_setup
 *   NOBlockDiag2_precond_EBE_setup / NOBlockDiag2_precond_EDE_setup
 *        (P*, M*, FS*, int tag, double *F)
 *        Solvers_and_Preconditioners/BlockDiag_precond_EBE_setup.c:6-86,
 *        BlockDiag_precond_EDE_setup.c:6-128 and the NOBlockDiag2 forms
 *        (the node block diagonal without preconditioning F)
 *   BlockDiag_precond (P*, M*, FS*, double *p, double *z)
 *        BlockDiag_precond.c:3-27
 *   Block_scaling_EBE (P*, M*, FS*), NO_scaling, NO_unscaling
 *        Block_scaling.c:10-59, :120-128 (the whole object)
 *   JacobiDOF4_precond_EBE_setup / JacobiDOF4_precond_EBE
 *        EBE/JacobiDOF4_precond_EBE_setup.c:4-64, EBE/JacobiDOF4_precond_EBE.c:7-83
 *   pgmres (P*, M*, FS*, FF*, double *B, double *X)
 *        the 6-argument solver Predictor_Old.c:111 calls and
 *        EulerEquations.h:168 types (setSolver.c:7); the reference tree has
 *        no implementation of it
 *
 * so that the driver linked with this object (instead of the objects of
 * those reference files) and libfemgpu.so keeps its products, block setups,
 * block scaling, element Jacobi and its GMRES solves on the GPU; its Build
 * (Build_M_K_F_SUPG), Predictor_Old, criteria, ede_Initialization and
 * assembly stay the reference's.
 *
 * Semantics kept:
 *  - products overwrite Q, P[neq] = Q[neq] = 0 (the element and edge forms);
 *  - the block setups sum the same-node 4x4 blocks of the stored element
 *    (or edge) matrices in element (edge) order, invert them with
 *    matrix_Inverse's cofactor expressions (entry 14 as the reference writes
 *    it) and left-precondition F in place (the NOBlockDiag2 forms leave F);
 *    the blocks stay on the device: MatrixData->BlockDiag / invBlockDiag /
 *    Id are not filled, and every routine that reads them (BlockDiag_precond,
 *    ebe2mvNDOF4 / ede2mvNDOF4 through BlockDiag2_precond and
 *    NOBlockDiag2_precond) is exported here;
 *  - Block_scaling_EBE rewrites the element matrices on the device (the host
 *    matrix keeps its unscaled values, see fga_core.h) and F in place;
 *  - JacobiDOF4: invDe = 1/A_jj per element dof, F preconditioned in place;
 *    the apply is correct in place (the reference zeroes z first, so its
 *    in-place call returns 0); MatrixData->invDe is not filled;
 *  - pgmres(B, X): solves A X = B with B already left-preconditioned;
 *    reads KrylovBasisVectorsQuantity, LinearMaxIter (cycle cap) and
 *    SolverTolerance, adds the Arnoldi step count to Parameters->iterations
 *    (Block_scaling.c:20 passes iterations + 1 as the setup tag); left
 *    preconditioners NO_precond, BlockDiag_precond, JacobiDOF4_precond_EBE
 *    and, for CSR, Diag_precond (MatrixData->invDiag uploaded); precondR
 *    must be NO_precond.
 * The edge numbering is computed on the device side from the mesh
 * (fg_ede4_create, bit for bit ede_Initialization.c) and checked against the
 * driver's lm2.  Anything else is reported and returns 1: there is no CPU
 * fallback.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "EulerEquations.h"
#include "fga_core.h"

int NO_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int Diag_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int BlockDiag_precond(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int JacobiDOF4_precond_EBE(ParametersType *, MatrixDataType *, FemStructsType *, double *,
                           double *);
int ebemvNDOF4(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int ebe2mvNDOF4(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int edemvNDOF4(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int ede2mvNDOF4(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);
int csrmv(ParametersType *, MatrixDataType *, FemStructsType *, double *, double *);

enum { SCH_EBE = 0, SCH_EDE = 1, SCH_CSR = 2 };

static struct {
    int scheme, neq, nel, nnodes, nitems;
    const void *key;            /* lm / lm2 / JA host pointer of the cached structure */
    fg_op *op;                  /* k = 12 elements, k = 8 edges, or CSR              */
    fg_ede4 *edges;             /* edge numbering (EDE)                               */
    fg_prec *blocks;            /* node block diagonal of op (EBE / EDE)              */
    fg_op *two;                 /* ebe2 / ede2 operator over op and blocks            */
    int two_form;
    fg_prec *jac;               /* JacobiDOF4 (EBE)                                   */
    fg_prec *diag;              /* Diag of a host setup (CSR)                         */
    int blocks_ready, jac_ready;
} g;

static void drop_op(void)
{
    if (g.two) fg_op_destroy(g.two);
    if (g.blocks) fg_prec_destroy(g.blocks);
    if (g.jac) fg_prec_destroy(g.jac);
    if (g.diag) fg_prec_destroy(g.diag);
    if (g.op) fg_op_destroy(g.op);
    if (g.edges) fg_ede4_destroy(g.edges);
    memset(&g, 0, sizeof g);
    fga_new_operator();
}

static int scheme_of_mv(FemFunctionsType *FF)
{
    if (FF->mv == ebemvNDOF4 || FF->mv == ebe2mvNDOF4) return SCH_EBE;
    if (FF->mv == edemvNDOF4 || FF->mv == ede2mvNDOF4) return SCH_EDE;
    if (FF->mv == csrmv) return SCH_CSR;
    return -1;
}

/* vertex[nel*3] and Node[].id[4] of the mesh */
static void mesh_arrays(ParametersType *Pm, FemStructsType *FS, int **vertex, int **ids)
{
    int nel = Pm->nel, nn = Pm->nnodes;
    *vertex = (int *)malloc(sizeof(int) * 3 * (size_t)(nel ? nel : 1));
    *ids = (int *)malloc(sizeof(int) * 4 * (size_t)(nn ? nn : 1));
    for (int e = 0; e < nel; e++)
        for (int a = 0; a < 3; a++) (*vertex)[3 * e + a] = FS->Element[e].Vertex[a];
    for (int i = 0; i < nn; i++)
        for (int d = 0; d < 4; d++) (*ids)[4 * i + d] = FS->Node[i].id[d] < 0 ? -1 : FS->Node[i].id[d];
}

static int ensure_structure(int scheme, ParametersType *Pm, MatrixDataType *M,
                            FemStructsType *FS)
{
    if (fga_ensure_ctx()) return 1;
    int neq = Pm->neq;
    const void *key = scheme == SCH_EBE ? (const void *)FS->lmaux
                    : scheme == SCH_EDE ? (const void *)FS->lm2aux : (const void *)M->JA;
    int nitems = scheme == SCH_EBE ? Pm->nel : scheme == SCH_EDE ? Pm->nedge : Pm->nnzero;
    if (g.op && g.scheme == scheme && g.neq == neq && g.key == key && g.nitems == nitems)
        return 0;
    drop_op();
    g.scheme = scheme; g.neq = neq; g.key = key; g.nitems = nitems;
    g.nel = Pm->nel; g.nnodes = Pm->nnodes;
    if (scheme == SCH_CSR) {
        if (fg_op_csr_create(fga_ctx, neq, Pm->nnzero, M->IA, M->JA, &g.op) != FG_OK)
            return fga_fail("operator");
        return 0;
    }
    int *vertex = NULL, *ids = NULL;
    mesh_arrays(Pm, FS, &vertex, &ids);
    int rc = FG_OK;
    if (scheme == SCH_EBE) {
        rc = fg_op_ebe_create(fga_ctx, neq, Pm->nel, 12, FS->lmaux, 0, &g.op);
        if (rc == FG_OK) rc = fg_prec_blockdiag_create(g.op, Pm->nnodes, vertex, ids, &g.blocks);
    } else {
        int nedge = 0;
        rc = fg_ede4_create(fga_ctx, Pm->nnodes, Pm->nel, neq, vertex, ids, &g.edges);
        if (rc == FG_OK) rc = fg_ede4_info(g.edges, &nedge);
        if (rc == FG_OK && nedge != Pm->nedge) {
            printf("femgpu: %d edges from the mesh, the driver has %d\n", nedge, Pm->nedge);
            rc = FG_ERR_ARG;
        }
        int *lm2 = NULL, *en = NULL;
        if (rc == FG_OK) {
            lm2 = (int *)malloc(sizeof(int) * 8 * (size_t)(nedge ? nedge : 1));
            en = (int *)malloc(sizeof(int) * 2 * (size_t)(nedge ? nedge : 1));
            rc = fg_ede4_get(g.edges, NULL, NULL, lm2, en);
        }
        if (rc == FG_OK && memcmp(lm2, FS->lm2aux, sizeof(int) * 8 * (size_t)nedge) != 0) {
            printf("femgpu: the edge numbering differs from the driver's lm2\n");
            rc = FG_ERR_ARG;
        }
        if (rc == FG_OK) rc = fg_op_ebe_create(fga_ctx, neq, nedge, 8, lm2, 0, &g.op);
        if (rc == FG_OK) rc = fg_prec_blockdiag_create(g.op, Pm->nnodes, en, ids, &g.blocks);
        free(lm2);
        free(en);
    }
    free(vertex);
    free(ids);
    if (rc != FG_OK) return fga_fail("operator");
    return 0;
}

static int upload_values(MatrixDataType *M)
{
    const double *vals = g.scheme == SCH_CSR ? M->AA : M->Aaux;
    if (fg_op_set_values_host(g.op, vals) != FG_OK) return fga_fail("values");
    g.blocks_ready = 0;
    g.jac_ready = 0;
    fga_uploaded();
    return 0;
}

static int ensure_build(int scheme, ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    if (ensure_structure(scheme, Pm, M, FS)) return 1;
    if (fga_needs_build_upload()) return upload_values(M);
    return 0;
}

/* the ebe2 / ede2 operator in the form Parameters->Preconditioner selects */
static int ensure_two(ParametersType *Pm)
{
    int form = strcasecmp(Pm->Preconditioner, "NOT") == 0 ? FG_EBE2_NOBLOCKDIAG
                                                          : FG_EBE2_BLOCKDIAG;
    if (g.two && g.two_form == form) return 0;
    if (g.two) fg_op_destroy(g.two);
    g.two = NULL;
    if (fg_op_ebe2_create(g.op, g.blocks, form, &g.two) != FG_OK) return fga_fail("ebe2");
    g.two_form = form;
    return 0;
}

static int product(int scheme, int two, ParametersType *Pm, MatrixDataType *M,
                   FemStructsType *FS, double *P, double *Q)
{
    if (ensure_structure(scheme, Pm, M, FS)) return 1;
    if (fga_needs_product_upload() && upload_values(M)) return 1;
    fg_op *op = g.op;
    if (two) {
        if (!g.blocks_ready) {
            printf("femgpu: ebe2mvNDOF4 / ede2mvNDOF4 need the block setup of this build\n");
            return 1;
        }
        if (ensure_two(Pm)) return 1;
        op = g.two;
    }
    if (fga_product(op, Pm->neq, P, Q)) return 1;
    if (scheme != SCH_CSR) P[Pm->neq] = 0.0;
    return 0;
}

int ebemvNDOF4(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_EBE, 0, Pm, M, FS, P, Q);
}

int ebe2mvNDOF4(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_EBE, 1, Pm, M, FS, P, Q);
}

int edemvNDOF4(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_EDE, 0, Pm, M, FS, P, Q);
}

int ede2mvNDOF4(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_EDE, 1, Pm, M, FS, P, Q);
}

int csrmv(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *P, double *Q)
{
    return product(SCH_CSR, 0, Pm, M, FS, P, Q);
}

/* ---- node block diagonal ------------------------------------------------ */
static int block_setup(int scheme, int with_f, ParametersType *Pm, MatrixDataType *M,
                       FemStructsType *FS, double *F)
{
    if (ensure_build(scheme, Pm, M, FS)) return 1;
    if (with_f) {
        if (fga_f_roundtrip(Pm->neq, F)) return 1;
        if (fg_prec_setup(g.blocks, fga_dv[0]) != FG_OK) return fga_fail("block setup");
        if (fga_f_back(Pm->neq, F)) return 1;
    } else if (fg_prec_setup(g.blocks, NULL) != FG_OK) {
        return fga_fail("block setup");
    }
    g.blocks_ready = 1;
    return 0;
}

int BlockDiag_precond_EBE_setup(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS,
                                int tag, double *F)
{
    (void)tag;
    return block_setup(SCH_EBE, 1, Pm, M, FS, F);
}

int BlockDiag_precond_EDE_setup(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS,
                                int tag, double *F)
{
    (void)tag;
    return block_setup(SCH_EDE, 1, Pm, M, FS, F);
}

int NOBlockDiag2_precond_EBE_setup(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS,
                                   int tag, double *F)
{
    (void)tag;
    return block_setup(SCH_EBE, 0, Pm, M, FS, F);
}

int NOBlockDiag2_precond_EDE_setup(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS,
                                   int tag, double *F)
{
    (void)tag;
    return block_setup(SCH_EDE, 0, Pm, M, FS, F);
}

int BlockDiag_precond(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *p,
                      double *z)
{
    (void)M; (void)FS;
    if (!g.blocks || !g.blocks_ready) {
        printf("femgpu BlockDiag_precond: no block setup of this build\n");
        return 1;
    }
    return fga_precond(g.blocks, Pm->neq, p, z);
}

/* Block_scaling.c:10-59: the block setup on F, then every element row block
 * left-multiplied by its node's inverse block, prescribed dofs removed */
int Block_scaling_EBE(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    if (ensure_build(SCH_EBE, Pm, M, FS)) return 1;
    if (fga_f_roundtrip(Pm->neq, FS->F)) return 1;
    if (fg_block_scaling_ebe(g.blocks, fga_dv[0]) != FG_OK) return fga_fail("block scaling");
    if (fga_f_back(Pm->neq, FS->F)) return 1;
    g.blocks_ready = 1;
    return 0;
}

int NO_scaling(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS)
{
    (void)Pm; (void)M; (void)FS;
    return 0;
}

int NO_unscaling(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *x)
{
    (void)Pm; (void)M; (void)FS; (void)x;
    return 0;
}

/* ---- JacobiDOF4 ---------------------------------------------------------- */
int JacobiDOF4_precond_EBE_setup(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS,
                                 int tag, double *F)
{
    (void)tag;
    if (ensure_build(SCH_EBE, Pm, M, FS)) return 1;
    if (!g.jac && fg_prec_create(g.op, FG_PREC_JACOBI, &g.jac) != FG_OK)
        return fga_fail("Jacobi");
    if (fga_f_roundtrip(Pm->neq, F)) return 1;
    if (fg_prec_setup(g.jac, fga_dv[0]) != FG_OK) return fga_fail("Jacobi setup");
    if (fga_f_back(Pm->neq, F)) return 1;
    g.jac_ready = 1;
    return 0;
}

int JacobiDOF4_precond_EBE(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, double *p,
                           double *z)
{
    (void)M; (void)FS;
    if (!g.jac || !g.jac_ready) {
        printf("femgpu JacobiDOF4_precond_EBE: no setup of this build\n");
        return 1;
    }
    return fga_precond(g.jac, Pm->neq, p, z);
}

/* ---- the 6-argument solver ----------------------------------------------- */
int pgmres(ParametersType *Pm, MatrixDataType *M, FemStructsType *FS, FemFunctionsType *FF,
           double *B, double *X)
{
    const int scheme = scheme_of_mv(FF);
    if (scheme < 0) {
        printf("femgpu pgmres: unsupported matrix-vector product\n");
        return 1;
    }
    const int two = FF->mv == ebe2mvNDOF4 || FF->mv == ede2mvNDOF4;
    if (FF->precondR != NO_precond) {
        printf("femgpu pgmres: only NO_precond is supported as right preconditioner\n");
        return 1;
    }
    const int bd = FF->precond == BlockDiag_precond, jac = FF->precond == JacobiDOF4_precond_EBE;
    const int diag = FF->precond == Diag_precond;
    if (FF->precond != NO_precond && !bd && !jac && !diag) {
        printf("femgpu pgmres: unsupported left preconditioner\n");
        return 1;
    }
    if ((bd && (scheme == SCH_CSR || two)) || (jac && scheme != SCH_EBE) ||
        (diag && scheme != SCH_CSR)) {
        printf("femgpu pgmres: preconditioner and product scheme do not combine\n");
        return 1;
    }
    if (ensure_build(scheme, Pm, M, FS)) return 1;
    fg_op *op = g.op;
    fg_prec *pc = NULL;
    if (two) {
        if (!g.blocks_ready) {
            printf("femgpu pgmres: ebe2/ede2 need the block setup of this build\n");
            return 1;
        }
        if (ensure_two(Pm)) return 1;
        op = g.two;
    }
    if (bd) {
        if (!g.blocks_ready) {
            printf("femgpu pgmres: BlockDiag_precond needs its setup in this build\n");
            return 1;
        }
        pc = g.blocks;
    }
    if (jac) {
        if (!g.jac_ready) {
            printf("femgpu pgmres: JacobiDOF4_precond_EBE needs its setup in this build\n");
            return 1;
        }
        pc = g.jac;
    }
    if (diag) {
        if (!g.diag && fg_prec_create(g.op, FG_PREC_DIAG, &g.diag) != FG_OK)
            return fga_fail("Diag");
        if (fg_prec_set_invdiag(g.diag, M->invDiag) != FG_OK) return fga_fail("invDiag");
        pc = g.diag;
    }
    fg_gmres_opts o;
    fga_opts(&o, Pm->KrylovBasisVectorsQuantity, Pm->LinearMaxIter, Pm->SolverTolerance);
    fg_gmres_info info;
    if (fga_solve(op, pc, Pm->neq, &o, B, X, &info)) return 1;
    g.blocks_ready = 0;
    g.jac_ready = 0;
    Pm->iterations += info.steps;
    return 0;
}
