#!/usr/bin/env python3
"""Benchmark of the FEM_CODES solve path on one MI355X.

Workload (the 2D SUPG steady-transport system at its single-GPU size):
  * unit square, 2048 x 2048 cells split into 8,388,608 P1 triangles
    (4,198,401 nodes, 4,190,209 equations), SUPG element matrices built on
    the GPU for a seeded rotating velocity (-y+0.5, x-0.5)*U[0.5,2] and
    diffusivity 10^U[-6,-3];
  * one "step" = one FP64 EBE matrix-vector product Q = A P over the whole
    mesh (ebemv, MatrixVector_Operations/EBE/ebemv.c:4-43), inputs resident
    in HBM; `value` = algorithmic bytes / time in GB/s;
  * the same product in FP32 and with the EDE and CSR storage schemes;
  * time to relative residual 1e-8 of the element-Jacobi-preconditioned
    GMRES(30) (--prec diag: the global diagonal), FP64 vs FP32-IR, on the
    constant-velocity SUPG system (beta=(cos t, sin t), k=1e-4, f=1) at
    2048^2 and at 64^2.  The rotating system above does not converge with
    GMRES(30)+Jacobi in a practical number of steps (the element Jacobi
    stagnates on it already at 16^2), so it is used for the products only.
  * cpu_baseline: the reference ebemv compiled from its own sources
    (oracle/_ref, -Ofast) or, where that library is absent, the oracle
    restatement, timed on one host core over the full mesh.

Prints one JSON line.  --gpus N>1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("matvec HBM GB/s (% of peak) + time-to-1e-8 solve, FP32-IR vs FP64 GMRES")
PEAK_GBS = 8000.0
ENGINES = {"auto": 0, "kernels": 1, "fused": 2, "small": 3}
ORTHS = {"mgs": 0, "cgs": 1}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def setup_system(F, ctx, n, prob, schemes=("EBE", "EDE", "CSR")):
    x, y, t, v = F.structured_mesh2d(n)
    mesh = F.Mesh2D(ctx, x, y, t, v)
    neq, nel = mesh.neq, mesh.nel
    lm = mesh.lm()
    ke = ctx.empty(9 * nel)
    Fv = ctx.empty(neq + 1)
    Fv.zero()
    mesh.build_supg(prob, ke, Fv)
    ops = {}
    if "EBE" in schemes:
        ops["EBE"] = F.Operator.ebe(ctx, neq, lm, 3)
        ops["EBE"].set_values(ke)
    if "CSR" in schemes:
        nnz, IA, JA, cbe = F.csr_pattern(neq, lm)
        ops["CSR"] = F.Operator.csr(ctx, neq, IA, JA)
        F.assemble_csr(ops["CSR"], cbe, ke)
        del IA, JA, cbe
    if "EDE" in schemes:
        nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
        ops["EDE"] = F.Operator.ede(ctx, neq, lm2)
        F.assemble_ede(ops["EDE"], order, ebe, ke)
        del order, ebe, lm2
    return dict(x=x, y=y, type=t, vertex=v, mesh=mesh, neq=neq, nel=nel, lm=lm, ke=ke, F=Fv,
                ops=ops)


def time_solve(F, ctx, sysd, precision, engine, kmax=30, tol=1e-8, lmax=100000, prec="diag",
               orth=0, scheme="EBE"):
    op = sysd["ops"].get("EBE_default", sysd["ops"]["EBE"]) if scheme == "EBE" else \
        sysd["ops"][scheme]
    pc = F.Preconditioner(op, F.PREC_JACOBI if prec == "jacobi" else F.PREC_DIAG)
    B = ctx.empty(sysd["neq"] + 1)
    F.lib.fg_memcpy_d2d(ctx.h, B.ptr, sysd["F"].ptr, B.nbytes)
    pc.setup(B)
    X = ctx.empty(sysd["neq"] + 1)
    ctx.sync()
    t0 = time.perf_counter()
    info, _, _ = F.gmres(ctx, op, pc, B, X, kmax=kmax, lmax=lmax, tol=tol, precision=precision,
                         engine=engine, orth=orth if precision == F.GMRES_FP32_IR else 0)
    ctx.sync()
    wall = time.perf_counter() - t0
    return info, wall, X


def cpu_baseline(sysd, ke_host, budget_s=12.0):
    """reference ebemv (or the oracle port) on one core over the full mesh"""
    import oracle as O
    neq, nel = sysd["neq"], sysd["nel"]
    P = np.random.default_rng(7).standard_normal(neq + 1)
    kind = "port"
    if O.Reference.available():
        r = O.Reference(fast=True)
        r.setup(sysd["x"], sysd["y"], sysd["type"], sysd["vertex"], "EBE")
        r.set_matrix(ke_host)
        run = lambda: r.matvec(P)  # noqa: E731
        kind = "reference"
    else:
        o = O.Oracle(fast=True)
        lm = sysd["lm"]
        run = lambda: o.ebemv(neq, nel, 3, ke_host, lm, P.copy())  # noqa: E731
    run()
    reps, t0 = 0, time.perf_counter()
    while True:
        run()
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s or reps >= 200:
            break
    sec = el / reps
    bytes_ = nel * (9 * 8 + 12) + 2 * neq * 8
    return dict(value=bytes_ / sec / 1e9, unit="GB/s", cores=1, kind=kind,
                sample=f"{reps} EBE FP64 matvecs of the full {nel}-element mesh "
                       f"({sec * 1e3:.1f} ms each), -Ofast single thread",
                ms_per_matvec=sec * 1e3, cpu=platform.processor() or platform.machine(),
                nproc=os.cpu_count())


def cpu_solve_small(n, prob_args):
    """reference pgmres at the 64^2 constant-velocity system (Diag, GMRES(30))"""
    import oracle as O
    o = O.Oracle()
    x, y, t, v = o.mesh2d(n)
    neq, idv = o.number_nodes(t)
    nel = len(v) // 3
    lm = o.fill_lm(neq, nel, v, t, idv)
    ke, Fv = o.build_supg2d(x, y, t, idv, v, neq, np.zeros(neq + 1), O.prob2d(*prob_args))
    if O.Reference.available():
        r = O.Reference(fast=True)
        r.setup(x, y, t, v, "EBE")
        r.set_problem(O.prob2d(*prob_args))
        r.set_matrix(ke)
        Fs = r.precond_setup("Diag", Fv)
        t0 = time.perf_counter()
        X, steps = r.pgmres(30, 100000, 1e-8, Fs)
        sec = time.perf_counter() - t0
        kind = "reference"
    else:
        fo = O.Oracle(fast=True)
        D, iD, Fs = fo.diag_setup_ebe(neq, nel, 3, ke, lm, Fv)
        t0 = time.perf_counter()
        X, steps, _, _, _ = fo.pgmres(neq, 30, 100000, 1e-8, Fs, "EBE", (ke, lm, 3, nel), invDiag=iD)
        sec = time.perf_counter() - t0
        kind = "port"
    return dict(seconds=sec, steps=int(steps), kind=kind), X, ke, Fv


def batched_leg(F, ctx, nprob, seed, prec="jacobi", kmax=30, tol=1e-8, sequential=True):
    """nprob independent constant-velocity SUPG problems, n ~ U{32..512} cells
    per side, angle t ~ U[0, 2 pi), diffusivity 10^U[-5,-2], assembled on the
    GPU, solved (a) together by the batched GMRES (FP64 and FP32-IR, graph
    replay of each cycle) and (b) one after another by the single-system
    GMRES, from the same left-preconditioned right-hand sides."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(32, 513, nprob)
    t0 = time.perf_counter()
    parts, kes, Fs, probs = [], [], [], []
    for n in sizes:
        th = float(rng.uniform(0.0, 2 * np.pi))
        kappa = float(10.0 ** rng.uniform(-5.0, -2.0))
        prob = F.Supg2dProblem(F.P_CONST, kappa, 0.0, 1.0, float(np.cos(th)), float(np.sin(th)),
                               1.0, F.H_2SQRTAREA)
        x, y, t, v = F.structured_mesh2d(int(n))
        mesh = F.Mesh2D(ctx, x, y, t, v)
        ke = ctx.empty(9 * mesh.nel)
        Fv = ctx.empty(mesh.neq + 1)
        Fv.zero()
        mesh.build_supg(prob, ke, Fv)
        parts.append((mesh.neq, mesh.lm()))
        kes.append(ke)
        Fs.append(Fv)
        probs.append((int(n), th, kappa))
        mesh.close()
    lm, seg = F.concat_lm(parts)
    N = int(seg[-1])
    nel = sum(len(p[1]) // 3 for p in parts)
    ke_all = ctx.empty(9 * nel)
    B = ctx.empty(N + 1)
    B.zero()
    off = 0
    for p, (ke, Fv) in enumerate(zip(kes, Fs)):
        F.lib.fg_memcpy_d2d(ctx.h, ke_all.ptr + 8 * off, ke.ptr, ke.nbytes)
        F.lib.fg_memcpy_d2d(ctx.h, B.ptr + 8 * int(seg[p]), Fv.ptr, 8 * parts[p][0])
        off += ke.n
    op = F.Operator.ebe(ctx, N, lm, 3)
    op.set_values(ke_all)
    kind = F.PREC_JACOBI if prec == "jacobi" else F.PREC_DIAG
    pc = F.Preconditioner(op, kind)
    pc.setup(B)
    ctx.sync()
    log(f"batched: {nprob} problems, {nel} elements, {N} equations, setup "
        f"{time.perf_counter() - t0:.1f}s")
    res = {"problems": int(nprob), "elements": int(nel), "equations": N,
           "sizes": {"min": int(sizes.min()), "max": int(sizes.max()),
                     "mean_cells_per_side": float(sizes.mean())},
           "precond": prec, "kmax": kmax, "tol": tol}
    Xs = {}
    for precision, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
        X = ctx.empty(N + 1)
        ctx.sync()
        w0 = time.perf_counter()
        rc, infos = F.gmres_batched(ctx, op, pc, seg, B, X, kmax=kmax, lmax=100000, tol=tol,
                                    precision=precision, graph=True)
        ctx.sync()
        wall = time.perf_counter() - w0
        Xs[tag] = X.get()
        res[f"batched_{tag}"] = dict(seconds=wall, status=int(rc),
                                     converged=int(sum(f.status == F.FG_OK for f in infos)),
                                     max_steps=int(max(f.steps for f in infos)),
                                     total_steps=int(sum(f.steps for f in infos)))
        log(f"batched {tag}", res[f"batched_{tag}"])
    d = Xs["fp32_ir"][:N] - Xs["fp64"][:N]
    errs = [float(np.linalg.norm(d[seg[p]:seg[p + 1]]) /
                  max(np.linalg.norm(Xs["fp64"][seg[p]:seg[p + 1]]), 1e-300))
            for p in range(nprob) if seg[p + 1] > seg[p]]
    res["fp32_ir_max_rel_err_vs_fp64"] = max(errs)
    res["speedup_fp32_ir"] = res["batched_fp64"]["seconds"] / res["batched_fp32_ir"]["seconds"]
    del ke_all, op, pc
    if sequential:
        # the same problems, one fg_gmres call each (operators built first)
        Bh = B.get()
        tot = {"fp64": 0.0, "fp32_ir": 0.0}
        for p, (neq_p, lm_p) in enumerate(parts):
            opp = F.Operator.ebe(ctx, neq_p, lm_p, 3)
            opp.set_values(kes[p])
            pcp = F.Preconditioner(opp, kind)
            Bp = ctx.to_device(np.append(Bh[seg[p]:seg[p + 1]], 0.0))
            pcp.setup(None)
            for precision, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
                Xp = ctx.empty(neq_p + 1)
                ctx.sync()
                w0 = time.perf_counter()
                F.gmres(ctx, opp, pcp, Bp, Xp, kmax=kmax, lmax=100000, tol=tol,
                        precision=precision)
                ctx.sync()
                tot[tag] += time.perf_counter() - w0
            opp.close()
        res["sequential_seconds"] = tot
        res["batched_speedup_vs_sequential"] = {
            k: tot[k] / res[f"batched_{k}"]["seconds"] for k in tot}
        log("sequential", tot)
    return res


def ss3d_leg(F, ctx, n=128, sweeps=4, seed=20260303):
    """3D SUPG steady transport (DIFCONV) on the n^3 cube cut into 6 n^3
    tetrahedra, a seeded random node numbering then reverse Cuthill-McKee;
    FP32 EBE 4x4 products + ILU(0) of the CSR matrix (its triangular solves
    as `sweeps` Jacobi sweeps) inside FP64 refinement (FP32-IR), GMRES(50)
    to 1e-10, against FP64 GMRES with the same operator and preconditioner."""
    t0 = time.perf_counter()
    x, y, z, t, v = F.structured_mesh3d(n)
    nn = len(x)
    perm = np.random.default_rng(seed).permutation(nn)
    inv = np.empty(nn, np.int64)
    inv[perm] = np.arange(nn)
    x, y, z, t, v = x[perm], y[perm], z[perm], t[perm], inv[v].astype(np.int32)
    m = F.Mesh3D(ctx, x, y, z, t, v)
    neq, nel = m.neq, m.nel
    lm = m.lm()
    nnz, IA, JA, cbe = F.csr_pattern_k(neq, lm, 4)
    bw0 = F.csr_bandwidth(neq, IA, JA)
    P = F.rcm(neq, IA, JA)
    lm = F.renumber_lm(neq, lm, P)
    nnz, IA, JA, cbe = F.csr_pattern_k(neq, lm, 4)
    bw1 = F.csr_bandwidth(neq, IA, JA)
    ke = ctx.empty(16 * nel)
    F0 = ctx.empty(neq + 1)
    F0.zero()
    m.build_supg(F.Supg3dProblem(F.P3_DIFCONV, 1.0, 0, 0, 0, 0, 0, 0), ke, F0)
    Fh = F0.get()
    Fn = np.zeros(neq + 1)
    Fn[:neq] = Fh[P]
    csr = F.Operator.csr(ctx, neq, IA, JA)
    F.assemble_csr(csr, cbe, ke, k=4)
    ebe = F.Operator.ebe(ctx, neq, lm, 4)
    ebe.set_values(ke)
    del cbe, IA, JA
    pc = F.Preconditioner(csr, F.PREC_ILU0)
    levels = pc.ilu_levels()
    pc.ilu_sweeps(sweeps)
    B = ctx.to_device(Fn)
    pc.setup(B)
    ctx.sync()
    log(f"ss3d: n={n} nel={nel} neq={neq} nnz={nnz} bandwidth {bw0} -> {bw1} (RCM), ILU levels "
        f"{levels[0]}, setup {time.perf_counter() - t0:.1f}s")
    rng = np.random.default_rng(1)
    res = {"n": n, "elements": nel, "equations": neq, "nnz": int(nnz),
           "bandwidth_random": int(bw0), "bandwidth_rcm": int(bw1), "ilu_levels": levels[0],
           "ilu_sweeps": sweeps, "kmax": 50, "tol": 1e-10}
    mv = {}
    for nm, op in (("EBE", ebe), ("CSR", csr)):
        for dt in (np.float64, np.float32):
            Pd = ctx.to_device(rng.standard_normal(neq + 1).astype(dt))
            Qd = ctx.empty(neq + 1, dt)
            ms = op.time_apply(Pd, Qd, reps=10)
            b = op.algorithmic_bytes(dt == np.float32)
            mv[f"{nm}_{'f32' if dt == np.float32 else 'f64'}"] = dict(ms=ms, GBs=b / ms / 1e6)
    res["matvec"] = mv
    Xs = {}
    for prec, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
        X = ctx.empty(neq + 1)
        ctx.sync()
        w0 = time.perf_counter()
        info, _, _ = F.gmres(ctx, ebe, pc, B, X, kmax=50, lmax=400, tol=1e-10, precision=prec)
        ctx.sync()
        Xs[tag] = X.get()
        res[tag] = dict(seconds=time.perf_counter() - w0, steps=info.steps, cycles=info.cycles,
                        status=info.status, true_rel_res=info.true_res / (info.eps / 1e-10))
        log("ss3d", tag, res[tag])
    res["speedup_fp32_ir"] = res["fp64"]["seconds"] / res["fp32_ir"]["seconds"]
    res["fp32_ir_rel_err_vs_fp64"] = float(np.linalg.norm(Xs["fp32_ir"][:neq] - Xs["fp64"][:neq]) /
                                           np.linalg.norm(Xs["fp64"][:neq]))
    return res


def transient_leg(F, ctx, n=1024, steps=10):
    """2D transient transport, the reference's CONE problem (02_Transport_
    Equation/2D/CONE: kappa 1e-8, no source or reaction, boundary value 0,
    velocity the rotation about (5,5), initial state exp(-r^2/2) about
    (5,7.5)) on the n x n grid of P1 triangles of [0,10]^2; Predictor_Old
    with alpha 1/2, dt 1e-3 (Courant number below 1 near the cone at
    n = 1024), corrections stopped by ||Da|| < 1e-3 ||a|| (at most 10),
    `steps` time steps; each correction is the transient SUPG + CAU build on
    the device and a Diag-preconditioned GMRES(30) solve to 1e-8 (at most 200
    restarts), FP64 and FP32-IR.  u, a, Da, R and the element matrices stay
    in HBM for the whole run."""
    x, y, t, v = F.structured_mesh2d(n)
    x, y = 10.0 * x, 10.0 * y
    mesh = F.Mesh2D(ctx, x, y, t, v)
    neq = mesh.neq
    op = F.Operator.ebe(ctx, neq, mesh.lm(), 3)
    pc = F.Preconditioner(op, F.PREC_DIAG)
    pb = F.Supg2dProblem(F.P_ROTATE_ABOUT, 1e-8, 0.0, 0.0, 5.0, 5.0, 1.0, F.H_2SQRTAREA)
    up = ctx.to_device(np.zeros(len(x)))
    _, idv = F.number_nodes(t)
    u0 = np.zeros(neq + 1)
    free = t > 0
    u0[idv[free]] = np.exp(-0.5 * ((x[free] - 5.0) ** 2 + (y[free] - 7.5) ** 2))
    dt = 1e-3
    res = {"n": n, "equations": neq, "elements": mesh.nel, "steps": steps, "dt": dt,
           "problem": "CONE", "precond": "Diag", "kmax": 30, "tol": 1e-8}
    U = {}
    for prec, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
        u = ctx.to_device(u0)
        ctx.sync()
        t0 = time.perf_counter()
        info, hist = F.predictor_old_transport(
            mesh, pb, op, pc, up, u, alpha=0.5, dt=dt, final_time=steps * dt + 0.5 * dt,
            nl_tol=1e-3, nl_maxiter=10, stop_by_norm=True, steady=False, max_steps=steps,
            kmax=30, lmax=200, tol=1e-8, precision=prec, engine=F.ENGINE_KERNELS,
            orth=F.ORTH_CGS if prec == F.GMRES_FP32_IR else F.ORTH_MGS)
        wall = time.perf_counter() - t0
        U[tag] = u.get()
        res[tag] = dict(seconds=wall, ms_per_step=1e3 * wall / max(1, info.steps),
                        time_steps=info.steps, corrections=info.corrections,
                        corrections_per_step=hist.tolist(),
                        krylov_steps=int(info.krylov_steps), unconverged=info.unconverged)
        log(f"transient {tag}", res[tag])
    res["speedup_fp32_ir"] = res["fp64"]["seconds"] / res["fp32_ir"]["seconds"]
    res["fp32_ir_rel_err_vs_fp64"] = float(np.linalg.norm(U["fp32_ir"][:neq] - U["fp64"][:neq]) /
                                           np.linalg.norm(U["fp64"][:neq]))
    return res


def euler_leg(F, ctx, n=1024, seed=20260305):
    """2D Euler: the n x n grid of P1 triangles, 4 conserved variables per
    node, a uniform Mach-3 inflow (rho 1 with a seeded 1% perturbation, u 1,
    v 0, c 1/3), every variable prescribed on the inflow side x = 0; one
    implicit backward-Euler step (alpha 1, dt 1e-3): the SUPG + CAU element
    build, then (Me + dt Ke) x = R with the node block-diagonal left
    preconditioner and GMRES(30) to 1e-8, FP64 and FP32-IR (12x12 block EBE
    products in FP32)."""
    t0 = time.perf_counter()
    x, y, t, v = F.structured_mesh2d(n)
    nn, nel = len(x), len(v) // 3
    inflow = np.abs(x) <= 1e-15
    ids = np.full(4 * nn, -1, np.int64)
    free = np.repeat(~inflow, 4)
    neq = int(free.sum())
    ids[free] = np.arange(neq)
    lm = ids[(4 * v[:, None].astype(np.int64) + np.arange(4)).reshape(-1)]
    lm = np.where(lm < 0, neq, lm).astype(np.int32)
    rng = np.random.default_rng(seed)
    gam = 1.4
    rho = 1.0 + 0.01 * rng.standard_normal(nn)
    u1 = np.ones(nn)
    u2 = np.zeros(nn)
    p = 1.0 / (9.0 * gam)
    E = p / ((gam - 1.0) * rho) + 0.5 * (u1 ** 2 + u2 ** 2)
    U = np.stack([rho, rho * u1, rho * u2, rho * E], 1).reshape(-1)
    m = F.Euler2D(ctx, x, y, v, lm, neq)
    A = ctx.empty(144 * nel)
    R = ctx.empty(neq + 1)
    Ud, dUd = ctx.to_device(U), ctx.to_device(np.zeros(4 * nn))
    prm = F.EulerParams(gam, 3.0, 0, 1.0, 1e-3, 1e-10)
    m.build(prm, Ud, dUd, A, R)
    ctx.sync()
    bt = []
    for _ in range(3):
        w0 = time.perf_counter()
        m.build(prm, Ud, dUd, A, R)
        ctx.sync()
        bt.append(time.perf_counter() - w0)
    op = F.Operator.ebe(ctx, neq, lm, 12)
    op.set_values(A)
    ids32 = np.where(ids < 0, -1, ids).astype(np.int32)
    pc = F.Preconditioner.blockdiag(op, nn, v, ids32)
    Rh = R.get()
    log(f"euler: n={n} nel={nel} neq={neq} setup {time.perf_counter() - t0:.1f}s")
    res = {"n": n, "elements": nel, "equations": neq, "element_build_ms": min(bt) * 1e3,
           "element_matrix_bytes": 144 * 8 * nel, "precond": "BlockDiag", "kmax": 30,
           "tol": 1e-8}
    mv = {}
    for dt in (np.float64, np.float32):
        Pd = ctx.to_device(rng.standard_normal(neq + 1).astype(dt))
        Qd = ctx.empty(neq + 1, dt)
        ms = op.time_apply(Pd, Qd, reps=20)
        b = op.algorithmic_bytes(dt == np.float32)
        mv["f32" if dt == np.float32 else "f64"] = dict(ms=ms, GBs=b / ms / 1e6,
                                                        pct_peak=100 * b / ms / 1e6 / PEAK_GBS,
                                                        bytes=b)
    res["matvec_ebe12"] = mv
    Xs = {}
    for prec, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
        B = ctx.to_device(Rh)
        pc.setup(B)
        X = ctx.empty(neq + 1)
        ctx.sync()
        w0 = time.perf_counter()
        info, _, _ = F.gmres(ctx, op, pc, B, X, kmax=30, lmax=2000, tol=1e-8, precision=prec)
        ctx.sync()
        Xs[tag] = X.get()
        res[tag] = dict(seconds=time.perf_counter() - w0, steps=info.steps, cycles=info.cycles,
                        status=info.status, true_rel_res=info.true_res / (info.eps / 1e-8))
        log("euler", tag, res[tag])
    res["speedup_fp32_ir"] = res["fp64"]["seconds"] / res["fp32_ir"]["seconds"]
    res["fp32_ir_rel_err_vs_fp64"] = float(np.linalg.norm(Xs["fp32_ir"][:neq] - Xs["fp64"][:neq]) /
                                           max(np.linalg.norm(Xs["fp64"][:neq]), 1e-300))
    # the edge-by-edge scheme (EDE): the same system as 8x8 edge matrices
    t1 = time.perf_counter()
    g = F.Ede4(ctx, nn, v, ids32, neq)
    _, _, lm2, en = g.arrays()
    Aede = ctx.empty(64 * g.nedge)
    op8 = F.Operator.ebe(ctx, neq, lm2, 8)
    pc8 = F.Preconditioner.blockdiag(op8, nn, en, ids32)
    ede = {"edges": g.nedge, "setup_s": time.perf_counter() - t1,
           "edge_matrix_bytes": 64 * 8 * g.nedge}
    g.assemble(A, Aede)
    ctx.sync()
    at = []
    for _ in range(3):
        w0 = time.perf_counter()
        g.assemble(A, Aede)
        ctx.sync()
        at.append(time.perf_counter() - w0)
    ede["edge_assembly_ms"] = min(at) * 1e3
    op8.set_values(Aede)
    mv8 = {}
    for dt in (np.float64, np.float32):
        Pd = ctx.to_device(rng.standard_normal(neq + 1).astype(dt))
        Qd = ctx.empty(neq + 1, dt)
        ms = op8.time_apply(Pd, Qd, reps=20)
        b = op8.algorithmic_bytes(dt == np.float32)
        mv8["f32" if dt == np.float32 else "f64"] = dict(ms=ms, GBs=b / ms / 1e6,
                                                         pct_peak=100 * b / ms / 1e6 / PEAK_GBS,
                                                         bytes=b)
    ede["matvec_ede8"] = mv8
    B = ctx.to_device(Rh)
    pc8.setup(B)
    X = ctx.empty(neq + 1)
    ctx.sync()
    w0 = time.perf_counter()
    info, _, _ = F.gmres(ctx, op8, pc8, B, X, kmax=30, lmax=2000, tol=1e-8)
    ctx.sync()
    ede["fp64"] = dict(seconds=time.perf_counter() - w0, steps=info.steps, status=info.status)
    ede["rel_diff_vs_ebe_fp64"] = float(np.linalg.norm(X.get()[:neq] - Xs["fp64"][:neq]) /
                                        max(np.linalg.norm(Xs["fp64"][:neq]), 1e-300))
    log("euler ede", ede)
    res["ede"] = ede
    g.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=20260202)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--engine", default="kernels", choices=sorted(ENGINES),
                    help="GMRES engine for the time-to-solve runs")
    ap.add_argument("--prec", default="jacobi", choices=("diag", "jacobi"),
                    help="preconditioner of the time-to-solve runs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the products, run a cast kernel of known traffic (64Mi "
                         "doubles read, 64Mi floats written) three times, the scale of "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--orth", default="cgs", choices=sorted(ORTHS),
                    help="orthogonalisation of the FP32-IR inner cycle (FP64 is always MGS)")
    ap.add_argument("--batched", type=int, default=0, metavar="NPROB",
                    help="also run the batched leg: NPROB independent problems of "
                         "32..512 cells per side (256 = the batched workload)")
    ap.add_argument("--ss3d", type=int, default=0, metavar="N",
                    help="also run the 3D leg on the N^3 Kuhn cube (128 = the 3D workload)")
    ap.add_argument("--transient", type=int, default=0, metavar="N",
                    help="also run the transient transport leg (Predictor_Old) on an N^2 grid")
    ap.add_argument("--euler", type=int, default=0, metavar="N",
                    help="also run the Euler leg on the N x N grid (1024 = the Euler workload)")
    ap.add_argument("--matvec-only", action="store_true",
                    help="only the EBE FP64 product (for profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              "(one GPU per process; the sharding over several GPUs is a design only)", file=sys.stderr)
        sys.exit(2)

    import fem_codes_amd as F
    ctx = F.Context(0)
    rng = np.random.default_rng(args.seed)
    scale = float(rng.uniform(0.5, 2.0))
    kappa = float(10.0 ** rng.uniform(-6.0, -3.0))
    theta = float(rng.uniform(0.0, 2 * np.pi))
    n = args.n
    t0 = time.perf_counter()
    rot = F.Supg2dProblem(F.P_ROTATING, kappa, 0.0, 1.0, 0.0, 0.0, scale, F.H_2SQRTAREA)
    schemes = ("EBE",) if args.matvec_only else ("EBE", "EDE", "CSR")
    sysd = setup_system(F, ctx, n, rot, schemes)
    ctx.sync()
    log(f"setup n={n} neq={sysd['neq']} nel={sysd['nel']} in {time.perf_counter() - t0:.1f}s")

    neq = sysd["neq"]
    P64 = ctx.to_device(rng.standard_normal(neq + 1))
    Q64 = ctx.empty(neq + 1)
    P32 = ctx.to_device(rng.standard_normal(neq + 1).astype(np.float32))
    Q32 = ctx.empty(neq + 1, np.float32)
    # elements per block of the headline product: the autotuner's choice
    from fem_codes_amd import autotune
    (best_epb, best_pipe), tune_ms, ebe_tuned = autotune.tune_ebe(ctx, neq, sysd["lm"], 3,
                                                                  sysd["ke"])
    log("autotune EBE f64 ms by (elements/block, workgroups/CU):",
        {k: round(v, 4) for k, v in tune_ms.items()}, "->", (best_epb, best_pipe))
    sysd["ops"]["EBE_default"] = sysd["ops"]["EBE"]
    sysd["ops"]["EBE"] = ebe_tuned
    ebe = ebe_tuned
    for _ in range(args.warmup):
        ebe.apply(P64, Q64)
    ctx.sync()
    w0 = time.perf_counter()
    ms = ebe.time_apply(P64, Q64, reps=args.steps)
    ctx.sync()
    wall = time.perf_counter() - w0
    ms_main = ebe.time_apply(P64, Q64, reps=args.steps, main_only=True)
    bytes64 = ebe.algorithmic_bytes(False)
    value = bytes64 / (ms * 1e-3) / 1e9
    achieved = bytes64 / (ms_main * 1e-3) / 1e9
    log(f"EBE f64: {ms:.4f} ms/matvec (block kernel {ms_main:.4f} ms) -> {value:.1f} GB/s")

    matvec = {}
    for name, op in sysd["ops"].items():
        r = {}
        for fp32, (P, Q) in ((False, (P64, Q64)), (True, (P32, Q32))):
            t = op.time_apply(P, Q, reps=args.steps)
            b = op.algorithmic_bytes(fp32)
            r["f32" if fp32 else "f64"] = dict(ms=t, GBs=b / (t * 1e-3) / 1e9,
                                              pct_peak=100 * b / (t * 1e-3) / 1e9 / PEAK_GBS,
                                              bytes=b)
        matvec[name] = r
        log(name, {k: round(v["GBs"], 1) for k, v in r.items()})

    copy = None
    if args.calibrate_copy:
        m = 64 << 20
        X = ctx.to_device(np.ones(m))
        Y = ctx.empty(m, np.float32)
        ctx.sync()
        c0 = time.perf_counter()
        for _ in range(3):
            F.lib.fg_cast_f64_to_f32(ctx.h, m, X.ptr, Y.ptr)
        ctx.sync()
        cs = (time.perf_counter() - c0) / 3
        copy = dict(bytes=12 * m, ms=cs * 1e3, GBs=12 * m / cs / 1e9)
        log("copy calibration", copy)
        del X, Y

    # HBM bytes per launch from the committed counter runs, when they measured
    # the kernel this run chose
    kernel = (f"k_ebe_apply_pipe<3, {best_epb // 256}, double, 0>" if best_pipe
              else f"k_ebe_apply<3, {best_epb // 256}, double, 0>")
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ebe_f64_2048_traffic.json")
    if os.path.exists(tpath) and n == 2048:
        with open(tpath) as f:
            tj = json.load(f)
        if tj.get("kernel") == kernel:
            traffic = tj.get("hbm_bytes_per_launch")

    solve = {}
    if not args.no_solve and not args.matvec_only:
        const = F.Supg2dProblem(F.P_CONST, 1e-4, 0.0, 1.0, float(np.cos(theta)),
                                float(np.sin(theta)), 1.0, F.H_2SQRTAREA)
        del sysd["ops"]["EDE"], sysd["ops"]["CSR"]
        sysd["mesh"].build_supg(const, sysd["ke"], sysd["F"])
        sysd["ops"]["EBE"].set_values(sysd["ke"])
        del sysd["ops"]["EBE_default"]
        res = {"engine": args.engine, "precond": args.prec, "fp32_ir_orth": args.orth}
        X64 = X32 = None
        for prec, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
            info, wall_s, X = time_solve(F, ctx, sysd, prec, ENGINES[args.engine],
                                         prec=args.prec, orth=ORTHS[args.orth])
            res[tag] = dict(seconds=wall_s, gpu_ms=info.ms, steps=info.steps, cycles=info.cycles,
                            true_rel_res=info.true_res / (info.eps / 1e-8), status=info.status)
            if tag == "fp64":
                X64 = X.get()
            else:
                X32 = X.get()
            log(tag, res[tag])
        res["speedup_fp32_ir"] = res["fp64"]["seconds"] / res["fp32_ir"]["seconds"]
        res["fp32_ir_rel_err_vs_fp64"] = float(np.linalg.norm(X32[:neq] - X64[:neq]) /
                                               np.linalg.norm(X64[:neq]))
        solve[f"const_{n}"] = res
        # 64^2, the reference's own case (3,969 equations): Diag, GMRES(30)
        # to 1e-8, FP64 on each engine (auto = the one-workgroup engine at
        # this size), EBE and CSR; the reference pgmres on one host core
        small = setup_system(F, ctx, 64, const, ("EBE", "CSR"))
        sres = {"precond": "diag"}
        runs = [(sch, F.GMRES_FP64, eng) for sch in ("EBE", "CSR") for eng in ("auto", "kernels")]
        runs.append(("EBE", F.GMRES_FP32_IR, "kernels"))
        for sch, prec, eng in runs:
            tag = f"{sch}_{'fp64' if prec == F.GMRES_FP64 else 'fp32_ir'}_{eng}"
            best = None
            for _ in range(3):   # first call of an engine includes its code-object load
                info, wall_s, X = time_solve(F, ctx, small, prec, ENGINES[eng], prec="diag",
                                             orth=ORTHS[args.orth], scheme=sch)
                best = wall_s if best is None else min(best, wall_s)
            sres[tag] = dict(seconds=best, steps=info.steps, cycles=info.cycles,
                             status=info.status, X=X.get())
        if not args.no_cpu:
            cs, Xc, _, _ = cpu_solve_small(64, (0, 1e-4, 0.0, 1.0, float(np.cos(theta)),
                                                float(np.sin(theta)), 1.0, 0))
            sres["cpu"] = cs
            for k, r in sres.items():
                if isinstance(r, dict) and "X" in r:
                    Xg = r.pop("X")
                    r["rel_err_vs_cpu"] = float(np.linalg.norm(Xg[:-1] - Xc[:-1]) /
                                                np.linalg.norm(Xc[:-1]))
                    r["speedup_vs_cpu"] = cs["seconds"] / r["seconds"]
        else:
            for r in sres.values():
                if isinstance(r, dict):
                    r.pop("X", None)
        solve["const_64"] = sres
        log("64^2", sres)

    batched = None
    if args.batched > 0:
        batched = batched_leg(F, ctx, args.batched, args.seed + 4, prec=args.prec)
    ss3d = ss3d_leg(F, ctx, args.ss3d) if args.ss3d > 0 else None
    euler = euler_leg(F, ctx, args.euler) if args.euler > 0 else None
    transient = transient_leg(F, ctx, args.transient) if args.transient > 0 else None

    cpu = None
    if not args.no_cpu and not args.matvec_only:
        if "ke" in sysd:
            sysd["mesh"].build_supg(rot, sysd["ke"], None)
        cpu = cpu_baseline(sysd, sysd["ke"].get())
        log("cpu", cpu)

    out = {
        "metric": METRIC, "value": value, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2D SUPG transport, {n}x{n} grid of P1 triangles "
                               f"({sysd['nel']} elements, {neq} equations), EBE 3x3 FP64 matvec",
                   "velocity_scale": scale, "diffusivity": kappa, "pct_of_peak": 100 * value / PEAK_GBS,
                   "wall_ms_per_step": wall * 1e3 / args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_GBS, "traffic": traffic,
                     "kernel": kernel,
                     "schedule": {"elems_per_block": best_epb, "wg_per_cu": best_pipe},
                     "algorithmic_bytes": bytes64, "ms": ms_main},
        "cpu_baseline": cpu,
        "matvec": matvec,
        "copy_calibration": copy,
        "ebe_autotune_ms": {f"{k[0]}/{k[1]}": v for k, v in tune_ms.items()},
        "solve": solve,
        "batched": batched,
        "ss3d": ss3d,
        "euler": euler, "transient": transient,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
