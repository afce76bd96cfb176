#!/usr/bin/env python3
"""Benchmark of the FEM_CODES solve path on one MI355X.

Every workload of the solve path runs by default (one leg each; --legs picks
a subset):

  headline  the 2D SUPG steady-transport system at its single-GPU size: unit
            square, 2048 x 2048 cells split into 8,388,608 P1 triangles
            (4,190,209 equations), element matrices built on the GPU for the
            seeded rotating velocity (-y+0.5, x-0.5)*U[0.5,2] and
            diffusivity 10^U[-6,-3].  One "step" = one FP64 EBE product
            Q = A P over the whole mesh (ebemv, MatrixVector_Operations/EBE/
            ebemv.c:4-43), inputs resident in HBM; `value` = algorithmic
            bytes / time in GB/s.  Also: the FP32 product and the EDE and CSR
            schemes; GMRES(30) to 1e-8 on this rotating system (it stagnates:
            the steps and the residual reached within a cap are recorded) and,
            FP64 vs FP32-IR, on the constant-velocity system (beta = (cos t,
            sin t), kappa 1e-4) with each storage scheme;
  small     the reference's own case: 64 x 64, CSR/EBE, Diag, GMRES(30) to
            1e-8, against the reference pgmres on one host core;
  batched   256 independent SUPG problems, 32..512 cells per side, one
            batched FP64 / FP32-IR GMRES(30) with graph-replayed cycles;
  ss3d      3D SUPG (DIFCONV) on the 128^3 Kuhn cube (12.6M tetrahedra), a
            seeded random numbering then RCM, EBE 4x4 products + CSR ILU(0),
            GMRES(50) to 1e-10, FP64 vs FP32-IR;
  euler     2D Euler, 1024^2 grid, 12x12 block element matrices, one
            implicit step with the node block-diagonal preconditioner,
            FP64 vs FP32-IR; and the edge scheme (8x8 edge matrices);
  transient the 02 driver's CONE problem through Predictor_Old at 1024^2.

Each leg reports `roofline` for its dominant kernel (algorithmic bytes of one
launch over that kernel's mean duration, HIP events on the library stream)
and `cpu_baseline`: the reference routine compiled from its own sources with
its own speed flags (oracle/_ref, -Ofast) timed on one host core over a
bounded sample, with the CPU model.  Prints one JSON line.  --gpus N>1 exits:
this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("matvec HBM GB/s (% of peak) + time-to-1e-8 solve, FP32-IR vs FP64 GMRES")
PEAK_GBS = 8000.0
ENGINES = {"auto": 0, "kernels": 1, "fused": 2, "small": 3, "graph": 4}
ORTHS = {"mgs": 0, "cgs": 1}
LEGS = ("headline", "small", "batched", "ss3d", "euler", "transient")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_info():
    """model name and core counts of the host (for the CPU baselines)"""
    model, fields = None, {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break   # first processor only
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k in ("cpu family", "model", "cpu MHz", "cache size"):
                    fields[k.replace(" ", "_")] = v
    except OSError:
        pass
    if fields:
        model = f"{model} (family {fields.get('cpu_family')}, model {fields.get('model')}, " \
                f"{fields.get('cpu_MHz')} MHz, cache {fields.get('cache_size')})"
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu": model, "nproc": os.cpu_count(), "cores_available": avail}


def roofline(ms, nbytes, kernel):
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": PEAK_GBS, "unit": "GB/s",
            "frac": gbs / PEAK_GBS, "traffic": None, "kernel": kernel,
            "algorithmic_bytes": nbytes, "ms": ms}


def leg_traffic(leg, kernel):
    """measured HBM bytes per launch of a leg's roofline kernel, from the
    FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/traffic_<leg>.json
    (null when the file is absent or names another kernel)"""
    path = os.path.join(ROOT, "profiles", f"traffic_{leg}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        tj = json.load(f)
    rec = tj.get("kernels", {}).get(kernel)
    return rec.get("hbm_bytes_per_launch") if rec else None


def ebe_kernel_name(k, fp32, op=None):
    """the dominant kernel of an element product: the sliced rows where the
    operator uses them, else the blocked kernel at the library's default
    block size"""
    t = "float" if fp32 else "double"
    if op is not None and op.product_form(fp32) == 2:   # k = 4 neighbour tables
        return f"k_ebe_nrows<{t}, 4>"
    if op is not None and k in (8, 12) and op.product_sliced(fp32):
        return f"k_ebe_node_rows<{t}, {k}>"
    if op is not None and op.product_sliced(fp32):
        return f"k_ebe_rows<{t}, {k}>"
    ept = 4 if k <= 2 else 2 if k == 3 else 4 if k == 4 else 1   # the library's defaults
    return f"k_ebe_apply<{k}, {ept}, {t}, 0>"


def time_cpu(run, budget_s, max_reps):
    """mean seconds of run() over a bounded sample (one warm-up call)"""
    run()
    reps, t0 = 0, time.perf_counter()
    while True:
        run()
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s or reps >= max_reps:
            return el / reps, reps


def cpu_product_baseline(name, run, nbytes, budget_s, max_reps, what, kind="reference"):
    sec, reps = time_cpu(run, budget_s, max_reps)
    out = dict(value=nbytes / sec / 1e9, unit="GB/s", cores=1, kind=kind,
               sample=f"{reps} FP64 {name} products of the full {what} "
                      f"({sec * 1e3:.1f} ms each), -Ofast, one thread",
               ms_per_matvec=sec * 1e3)
    out.update(cpu_info())
    return out


def matvec_table(op, P64, Q64, P32, Q32, reps):
    r = {}
    for fp32, (P, Q) in ((False, (P64, Q64)), (True, (P32, Q32))):
        t = op.time_apply(P, Q, reps=reps)
        b = op.algorithmic_bytes(fp32)
        r["f32" if fp32 else "f64"] = dict(ms=t, GBs=b / (t * 1e-3) / 1e9,
                                          pct_peak=100 * b / (t * 1e-3) / 1e9 / PEAK_GBS,
                                          bytes=b, sliced=op.product_sliced(fp32))
    return r


# ---------------------------------------------------------------------------
def setup_system(F, ctx, n, prob, schemes=("EBE", "EDE", "CSR")):
    x, y, t, v = F.structured_mesh2d(n)
    mesh = F.Mesh2D(ctx, x, y, t, v)
    neq, nel = mesh.neq, mesh.nel
    lm = mesh.lm()
    ke = ctx.empty(9 * nel)
    Fv = ctx.empty(neq + 1)
    Fv.zero()
    mesh.build_supg(prob, ke, Fv)
    ops, aux = {}, {}
    if "EBE" in schemes:
        ops["EBE"] = F.Operator.ebe(ctx, neq, lm, 3)
        ops["EBE"].set_values(ke)
    if "CSR" in schemes:
        nnz, IA, JA, cbe = F.csr_pattern(neq, lm)
        ops["CSR"] = F.Operator.csr(ctx, neq, IA, JA)
        F.assemble_csr(ops["CSR"], cbe, ke)
        aux["CSR"] = cbe
    if "EDE" in schemes:
        nedge, order, ebe, lm2 = F.ede_pattern(neq, lm)
        ops["EDE"] = F.Operator.ede(ctx, neq, lm2)
        F.assemble_ede(ops["EDE"], order, ebe, ke)
        aux["EDE"] = (order, ebe)
        aux["lm2"] = lm2
    return dict(x=x, y=y, type=t, vertex=v, mesh=mesh, neq=neq, nel=nel, lm=lm, ke=ke, F=Fv,
                ops=ops, aux=aux)


def reassemble(F, sysd, prob):
    """rebuild the element matrices for another problem, into every scheme"""
    sysd["F"].zero()
    sysd["mesh"].build_supg(prob, sysd["ke"], sysd["F"])
    for name, op in sysd["ops"].items():
        if name == "CSR":
            F.assemble_csr(op, sysd["aux"]["CSR"], sysd["ke"])
        elif name == "EDE":
            F.assemble_ede(op, *sysd["aux"]["EDE"], sysd["ke"])
        else:
            op.set_values(sysd["ke"])


def time_solve(F, ctx, sysd, op, prec_kind, precision, engine, kmax=30, tol=1e-8,
               lmax=100000, orth=0):
    pc = F.Preconditioner(op, prec_kind)
    B = ctx.empty(sysd["neq"] + 1)
    F.lib.fg_memcpy_d2d(ctx.h, B.ptr, sysd["F"].ptr, B.nbytes)
    pc.setup(B)
    X = ctx.empty(sysd["neq"] + 1)
    # untimed: one restart cycle of the same solve (first launches load the
    # kernels' code objects)
    F.gmres(ctx, op, pc, B, X, kmax=kmax, lmax=1, tol=tol, precision=precision, engine=engine,
            orth=orth if precision == F.GMRES_FP32_IR else 0)
    ctx.sync()
    t0 = time.perf_counter()
    info, _, _ = F.gmres(ctx, op, pc, B, X, kmax=kmax, lmax=lmax, tol=tol, precision=precision,
                         engine=engine, orth=orth if precision == F.GMRES_FP32_IR else 0)
    ctx.sync()
    wall = time.perf_counter() - t0
    return info, wall, X


def solve_record(info, wall, tol):
    return dict(seconds=wall, gpu_ms=info.ms, steps=info.steps, cycles=info.cycles,
                true_rel_res=info.true_res / (info.eps / tol) if info.eps > 0 else None,
                status=info.status)


def headline_leg(F, ctx, args, rng):
    """config 2: products of the three storage schemes, roofline, the
    reference CPU product, and the time-to-solve runs"""
    scale = float(rng.uniform(0.5, 2.0))
    kappa = float(10.0 ** rng.uniform(-6.0, -3.0))
    theta = float(rng.uniform(0.0, 2 * np.pi))
    n = args.n
    t0 = time.perf_counter()
    rot = F.Supg2dProblem(F.P_ROTATING, kappa, 0.0, 1.0, 0.0, 0.0, scale, F.H_2SQRTAREA)
    schemes = ("EBE",) if args.matvec_only else ("EBE", "EDE", "CSR")
    sysd = setup_system(F, ctx, n, rot, schemes)
    ctx.sync()
    neq = sysd["neq"]
    log(f"setup n={n} neq={neq} nel={sysd['nel']} in {time.perf_counter() - t0:.1f}s")
    P64 = ctx.to_device(rng.standard_normal(neq + 1))
    Q64 = ctx.empty(neq + 1)
    P32 = ctx.to_device(rng.standard_normal(neq + 1).astype(np.float32))
    Q32 = ctx.empty(neq + 1, np.float32)
    from fem_codes_amd import autotune
    default_op = sysd["ops"]["EBE"]          # the library default: 512 per block, BFS
    if args.no_autotune:
        best_epb, best_ord, best_pipe, tune_ms = 512, F.ORDER_BFS, 0, {}
        fp32_ms = None
        ebe_tuned = default_op
    else:
        (best_epb, best_ord, best_pipe), tune_ms, ebe_tuned = autotune.tune_ebe(
            ctx, neq, sysd["lm"], 3, sysd["ke"], orderings=autotune.ORDERINGS)
        log("autotune EBE f64 ms by elements/block / ordering / workgroups per CU:",
            {autotune.key_str(k): round(v, 4) for k, v in tune_ms.items()}, "->",
            autotune.key_str((best_epb, best_ord, best_pipe)),
            "; FP32 products:", getattr(ebe_tuned, "other_precision_ms", None))
        fp32_ms = getattr(ebe_tuned, "other_precision_ms", {}).get("fp32")
        if best_pipe == autotune.ROWS or (best_epb, best_ord, best_pipe) == (512, F.ORDER_BFS, 0):
            # the same kernel choice per precision on the setup operator
            for fp32, prec in ((False, F.PRODUCT_FP64), (True, F.PRODUCT_FP32)):
                default_op.set_product_kernel(F.PRODUCT_AUTO if ebe_tuned.product_sliced(fp32)
                                              else F.PRODUCT_BLOCKED, prec)
            # the tuner chose the default configuration: keep the operator
            # built at setup (allocated before the tuning candidates)
            ebe_tuned.close()
            ebe_tuned = default_op
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
    log(f"EBE f64: {ms:.4f} ms/matvec (block kernel {ms_main:.4f} ms) -> {value:.1f} GB/s")
    if ebe.product_sliced():
        kernel = "k_ebe_rows<double, 3>"
        schedule = {"kernel": "sliced rows"}
    else:
        kernel = (f"k_ebe_apply_pipe<3, {best_epb // 256}, double, 0>" if best_pipe
                  else f"k_ebe_apply<3, {best_epb // 256}, double, 0>")
        schedule = {"kernel": "blocked", "elems_per_block": best_epb, "wg_per_cu": best_pipe,
                    "ordering": autotune.ORDER_NAMES[best_ord]}
    roof = roofline(ms_main, bytes64, kernel)
    roof["schedule"] = schedule
    if n == 2048:
        roof["traffic"] = leg_traffic("headline", kernel)

    # the edge scheme: edges per block x schedule tuned the same way
    ede_tune = None
    if "EDE" in sysd["ops"] and not args.no_autotune:
        from fem_codes_amd import autotune
        vals = sysd["ops"]["EDE"].get_values()
        (epb_e, ord_e, pipe_e), tms_e, ede_tuned = autotune.tune_ebe(
            ctx, neq, sysd["aux"]["lm2"], 2, vals, edges=True, orderings=autotune.ORDERINGS)
        del vals
        log("autotune EDE f64 ms by edges/block / ordering / workgroups per CU:",
            {autotune.key_str(k): round(v, 4) for k, v in tms_e.items()}, "->",
            autotune.key_str((epb_e, ord_e, pipe_e)))
        sysd["ops"]["EDE"].close()
        sysd["ops"]["EDE"] = ede_tuned
        ede_tune = {"kernel": "rows" if pipe_e == autotune.ROWS else "blocked",
                    "edges_per_block": epb_e, "wg_per_cu": pipe_e,
                    "ordering": autotune.ORDER_NAMES[ord_e],
                    "ms": {autotune.key_str(k): v for k, v in tms_e.items()}}
    matvec = {}
    extra = [] if ebe_tuned is default_op else [("EBE_default", default_op)]
    for name, op in list(sysd["ops"].items()) + extra:
        matvec[name] = matvec_table(op, P64, Q64, P32, Q32, args.steps)
        log(name, {k: round(v["GBs"], 1) for k, v in matvec[name].items()})
        if name in ("EBE", "EDE"):
            # the other product kernel of each precision, for comparison
            s64, s32 = op.product_sliced(False), op.product_sliced(True)
            try:
                op.set_product_kernel(F.PRODUCT_BLOCKED if s64 else F.PRODUCT_SLICED,
                                      F.PRODUCT_FP64)
                op.set_product_kernel(F.PRODUCT_BLOCKED if s32 else F.PRODUCT_SLICED,
                                      F.PRODUCT_FP32)
            except F.FemGpuError:
                continue           # no sliced rows on this mesh
            alt = name + "_other"
            matvec[alt] = matvec_table(op, P64, Q64, P32, Q32, args.steps)
            log(alt, {k: (v["sliced"], round(v["GBs"], 1)) for k, v in matvec[alt].items()})
            op.set_product_kernel(F.PRODUCT_AUTO if s64 else F.PRODUCT_BLOCKED, F.PRODUCT_FP64)
            op.set_product_kernel(F.PRODUCT_AUTO if s32 else F.PRODUCT_BLOCKED, F.PRODUCT_FP32)
    if ebe_tuned is not default_op:
        default_op.close()
    res = {"value": value, "ms": ms, "wall_ms_per_step": wall * 1e3 / args.steps,
           "scale": scale, "kappa": kappa, "theta": theta, "roofline": roof,
           "matvec": matvec, "ebe_autotune_ms": {autotune.key_str(k): v for k, v in tune_ms.items()},
           "ede_autotune": ede_tune, "neq": neq, "nel": sysd["nel"],
           "ebe_fp32_product_ms": fp32_ms}

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
        res["copy_calibration"] = dict(bytes=12 * m, ms=cs * 1e3, GBs=12 * m / cs / 1e9)
        del X, Y

    if not args.no_cpu and not args.matvec_only:
        res["cpu_baseline"] = headline_cpu(sysd)
        log("cpu", res["cpu_baseline"])

    if args.no_solve or args.matvec_only:
        return res, sysd
    solve = {}
    eng = ENGINES[args.engine]
    # (a) the workload's own system: rotating flow, kappa 10^U[-6,-3];
    # closed streamlines, GMRES(30) stagnates: capped at args.rot_cycles
    rot_res = {"precond": "Diag", "kmax": 30, "cycle_cap": args.rot_cycles}
    for prec, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
        info, w, X = time_solve(F, ctx, sysd, ebe, F.PREC_DIAG, prec, eng,
                                lmax=args.rot_cycles, orth=ORTHS[args.orth])
        rot_res[tag] = solve_record(info, w, 1e-8)
        log("rotating", tag, rot_res[tag])
    solve[f"rotating_{n}"] = rot_res
    # (b) the constant-velocity system: FP64 vs FP32-IR, each scheme (Diag),
    # and the element Jacobi of Jacobi_precond_EBE.c on the EBE scheme
    const = F.Supg2dProblem(F.P_CONST, 1e-4, 0.0, 1.0, float(np.cos(theta)),
                            float(np.sin(theta)), 1.0, F.H_2SQRTAREA)
    reassemble(F, sysd, const)
    cres = {"kmax": 30, "tol": 1e-8, "engine": args.engine, "fp32_ir_orth": args.orth}
    runs = [("EBE", F.PREC_JACOBI, "EBE_jacobi")] + \
           [(s, F.PREC_DIAG, f"{s}_diag") for s in sysd["ops"]]
    for scheme, pk, tag in runs:
        r, Xs = {}, {}
        for prec, pt in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
            info, w, X = time_solve(F, ctx, sysd, sysd["ops"][scheme], pk, prec, eng,
                                    orth=ORTHS[args.orth])
            r[pt] = solve_record(info, w, 1e-8)
            Xs[pt] = X.get()
        r["speedup_fp32_ir"] = r["fp64"]["seconds"] / r["fp32_ir"]["seconds"]
        r["fp32_ir_rel_err_vs_fp64"] = float(np.linalg.norm(Xs["fp32_ir"][:neq] - Xs["fp64"][:neq])
                                             / np.linalg.norm(Xs["fp64"][:neq]))
        if tag == "EBE_jacobi":
            Xref = Xs["fp64"]
        else:
            r["rel_diff_vs_EBE_jacobi_fp64"] = float(np.linalg.norm(Xs["fp64"][:neq] -
                                                                    Xref[:neq]) /
                                                     np.linalg.norm(Xref[:neq]))
        cres[tag] = r
        log("const", tag, {k: (v if not isinstance(v, dict) else
                               {kk: vv for kk, vv in v.items() if kk in ("seconds", "steps")})
                           for k, v in r.items()})
    solve[f"const_{n}"] = cres
    res["solve"] = solve
    return res, sysd


def headline_cpu(sysd):
    """the reference ebemv on one core over the full mesh"""
    import oracle as O
    neq, nel = sysd["neq"], sysd["nel"]
    P = np.random.default_rng(7).standard_normal(neq + 1)
    ke_host = sysd["ke"].get()
    bytes_ = nel * (9 * 8 + 12) + 2 * neq * 8
    if O.Reference.available():
        r = O.Reference(fast=True)
        r.setup(sysd["x"], sysd["y"], sysd["type"], sysd["vertex"], "EBE")
        r.set_matrix(ke_host)
        return cpu_product_baseline("ebemv", lambda: r.matvec(P), bytes_, 12.0, 200,
                                    f"{nel}-element mesh")
    o = O.Oracle(fast=True)
    lm = sysd["lm"]
    return cpu_product_baseline("ebemv", lambda: o.ebemv(neq, nel, 3, ke_host, lm, P.copy()),
                                bytes_, 12.0, 200, f"{nel}-element mesh", kind="port")


def small_leg(F, ctx, args, theta):
    """config 1: 64^2 (3,969 equations), Diag, GMRES(30) to 1e-8; each engine,
    CSR and EBE; the reference pgmres on one host core"""
    const = F.Supg2dProblem(F.P_CONST, 1e-4, 0.0, 1.0, float(np.cos(theta)),
                            float(np.sin(theta)), 1.0, F.H_2SQRTAREA)
    small = setup_system(F, ctx, 64, const, ("EBE", "CSR"))
    sres = {"precond": "diag", "kmax": 30, "tol": 1e-8}
    runs = [(sch, F.GMRES_FP64, eng) for sch in ("CSR", "EBE") for eng in ("auto", "kernels")]
    runs.append(("EBE", F.GMRES_FP32_IR, "kernels"))
    for sch, prec, eng in runs:
        tag = f"{sch}_{'fp64' if prec == F.GMRES_FP64 else 'fp32_ir'}_{eng}"
        best = None
        for _ in range(3):   # the first call of an engine includes its code-object load
            info, wall_s, X = time_solve(F, ctx, small, small["ops"][sch], F.PREC_DIAG, prec,
                                         ENGINES[eng], orth=ORTHS[args.orth])
            best = wall_s if best is None else min(best, wall_s)
        sres[tag] = dict(seconds=best, steps=info.steps, cycles=info.cycles,
                         status=info.status, X=X.get())
    if not args.no_cpu:
        cs, Xc = cpu_solve_small(64, (0, 1e-4, 0.0, 1.0, float(np.cos(theta)),
                                      float(np.sin(theta)), 1.0, 0))
        sres["cpu_baseline"] = cs
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
    log("64^2", {k: (v["seconds"] if isinstance(v, dict) else v) for k, v in sres.items()})
    return sres


def cpu_solve_small(n, prob_args):
    """the reference pgmres at the 64^2 constant-velocity system (Diag,
    GMRES(30)), CSR storage as the workload states"""
    import oracle as O
    o = O.Oracle()
    x, y, t, v = o.mesh2d(n)
    neq, idv = o.number_nodes(t)
    nel = len(v) // 3
    lm = o.fill_lm(neq, nel, v, t, idv)
    if O.Reference.available():
        r = O.Reference(fast=True)
        r.setup(x, y, t, v, "CSR")
        r.set_problem(O.prob2d(*prob_args))
        A, Fv = r.build(np.zeros(r.neq + 1))
        Fs = r.precond_setup("Diag", Fv)
        best = None
        for _ in range(3):
            t0 = time.perf_counter()
            X, steps = r.pgmres(30, 100000, 1e-8, Fs)
            sec = time.perf_counter() - t0
            best = sec if best is None else min(best, sec)
        kind = "reference"
    else:
        ke, Fv = o.build_supg2d(x, y, t, idv, v, neq, np.zeros(neq + 1), O.prob2d(*prob_args))
        fo = O.Oracle(fast=True)
        D, iD, Fs = fo.diag_setup_ebe(neq, nel, 3, ke, lm, Fv)
        t0 = time.perf_counter()
        X, steps, _, _, _ = fo.pgmres(neq, 30, 100000, 1e-8, Fs, "EBE", (ke, lm, 3, nel),
                                      invDiag=iD)
        best = time.perf_counter() - t0
        kind = "port"
    out = dict(value=best, unit="s", seconds=best, steps=int(steps), cores=1, kind=kind,
               sample="one pgmres solve of the 64^2 system (CSR, Diag, GMRES(30), 1e-8), "
                      "best of 3, -Ofast, one thread")
    out.update(cpu_info())
    return out, X


# ---------------------------------------------------------------------------
def batched_leg(F, ctx, args, nprob, seed, kmax=30, tol=1e-8):
    """nprob independent constant-velocity SUPG problems, n ~ U{32..512} cells
    per side, angle t ~ U[0, 2 pi), diffusivity 10^U[-5,-2], assembled on the
    GPU, solved (a) together by the batched GMRES (FP64 and FP32-IR, graph
    replay of each cycle), (b) one after another by the single-system GMRES,
    from the same Diag-preconditioned right-hand sides; the reference pgmres
    on one host core solves a bounded sample of the same problems."""
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
    pc = F.Preconditioner(op, F.PREC_DIAG)
    pc.setup(B)
    ctx.sync()
    log(f"batched: {nprob} problems, {nel} elements, {N} equations, setup "
        f"{time.perf_counter() - t0:.1f}s")
    res = {"problems": int(nprob), "elements": int(nel), "equations": N,
           "sizes": {"min": int(sizes.min()), "max": int(sizes.max()),
                     "mean_cells_per_side": float(sizes.mean())},
           "precond": "diag", "kmax": kmax, "tol": tol, "graph_replay": not args.no_graph,
           "fp32_ir_orth": args.orth}
    # the batched product over every problem at once: the dominant kernel
    P32 = ctx.to_device(rng.standard_normal(N + 1).astype(np.float32))
    Q32 = ctx.empty(N + 1, np.float32)
    P64 = ctx.to_device(rng.standard_normal(N + 1))
    Q64 = ctx.empty(N + 1)
    res["matvec"] = matvec_table(op, P64, Q64, P32, Q32, 20)
    ms_main = op.time_apply(P32, Q32, reps=20, main_only=True)
    res["roofline"] = roofline(ms_main, op.algorithmic_bytes(True), ebe_kernel_name(3, True, op))
    res["roofline"]["traffic"] = leg_traffic("batched", res["roofline"]["kernel"])
    Xs = {}
    for precision, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
        X = ctx.empty(N + 1)
        # untimed: one restart cycle (loads the kernels' code objects)
        F.gmres_batched(ctx, op, pc, seg, B, X, kmax=kmax, lmax=1, tol=tol, precision=precision,
                        graph=not args.no_graph,
                        orth=ORTHS[args.orth] if precision == F.GMRES_FP32_IR else F.ORTH_MGS)
        ctx.sync()
        w0 = time.perf_counter()
        rc, infos = F.gmres_batched(ctx, op, pc, seg, B, X, kmax=kmax, lmax=100000, tol=tol,
                                    precision=precision, graph=not args.no_graph,
                                    orth=ORTHS[args.orth] if precision == F.GMRES_FP32_IR
                                    else F.ORTH_MGS)
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
    if not args.no_autotune:
        res["tuned"] = batched_tuned(F, ctx, args, N, lm, ke_all, seg, B, Xs["fp64"], tol,
                                     res["batched_fp32_ir"]["seconds"],
                                     res["batched_fp32_ir"]["max_steps"])
    Bh = B.get()
    del ke_all, op, pc
    if not args.no_sequential:
        # the same problems, one fg_gmres call each (operators built first)
        tot = {"fp64": 0.0, "fp32_ir": 0.0}
        for p, (neq_p, lm_p) in enumerate(parts):
            opp = F.Operator.ebe(ctx, neq_p, lm_p, 3)
            opp.set_values(kes[p])
            pcp = F.Preconditioner(opp, F.PREC_DIAG)
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
    if not args.no_cpu:
        res["cpu_baseline"] = batched_cpu(probs, Xs["fp64"], seg, res["batched_fp32_ir"]["seconds"],
                                          nprob)
        log("batched cpu", res["cpu_baseline"])
    return res


def batched_tuned(F, ctx, args, N, lm, ke_all, seg, B, X64, tol, default_s, default_steps):
    """the FP32-IR batched solve with the autotuner's choices: elements per
    block x ordering x schedule of the FP32 product, then the restart length
    over full solves (B already carries the Diag left-preconditioning, and
    the tuned operator's Diag is the same sum of element diagonals).  Each
    trial solve is capped at three times the Arnoldi steps of the default
    solve, so a restart length that stagnates ends as 'not converged'."""
    from fem_codes_amd import autotune
    key, tms, op_t = autotune.tune_ebe(ctx, N, lm, 3, ke_all, fp32=True, reps=20,
                                       orderings=autotune.ORDERINGS)
    log("batched autotune FP32 product ms:", {autotune.key_str(k): round(v, 4)
                                             for k, v in tms.items()}, "->", autotune.key_str(key))
    pc_t = F.Preconditioner(op_t, F.PREC_DIAG)
    pc_t.setup(None)

    def run(km, X=None):
        X = X if X is not None else ctx.empty(N + 1)
        ctx.sync()
        w0 = time.perf_counter()
        lmax = -(-3 * default_steps // km)
        rc, infos = F.gmres_batched(ctx, op_t, pc_t, seg, B, X, kmax=km, lmax=lmax, tol=tol,
                                    precision=F.GMRES_FP32_IR, graph=not args.no_graph,
                                    orth=ORTHS[args.orth])
        ctx.sync()
        if rc != F.FG_OK:
            st = [f.status for f in infos]
            log(f"batched tuned kmax {km}: status {rc}, problem statuses",
                {s: st.count(s) for s in set(st)}, "max steps", max(f.steps for f in infos))
        return time.perf_counter() - w0, rc == F.FG_OK, infos

    try:
        best_km, km_s = autotune.tune_restart(lambda km: run(km)[:2])
    except F.FemGpuError as e:
        log("batched restart tuning:", e)
        op_t.close()
        pc_t.close()
        return {"product": autotune.key_str(key), "error": str(e)}
    log("batched autotune FP32-IR seconds by restart length:", km_s, "->", best_km)
    X = ctx.empty(N + 1)
    sec, ok, infos = run(best_km, X)
    d = X.get()[:N] - X64[:N]
    err = max(float(np.linalg.norm(d[seg[p]:seg[p + 1]]) /
                    max(np.linalg.norm(X64[seg[p]:seg[p + 1]]), 1e-300))
              for p in range(len(seg) - 1) if seg[p + 1] > seg[p])
    out = {"product": autotune.key_str(key),
           "product_ms": {autotune.key_str(k): v for k, v in tms.items()},
           "restart_seconds": {str(k): v for k, v in km_s.items()}, "kmax": best_km,
           "fp32_ir": dict(seconds=sec, converged=int(sum(f.status == F.FG_OK for f in infos)),
                           total_steps=int(sum(f.steps for f in infos)),
                           max_steps=int(max(f.steps for f in infos))),
           "max_rel_err_vs_fp64_kmax30": err, "speedup_vs_default_fp32_ir": default_s / sec}
    log("batched tuned", {k: v for k, v in out.items() if k not in ("product_ms",)})
    pc_t.close()
    op_t.close()
    return out


def batched_cpu(probs, X_gpu, seg, gpu_seconds, nprob, budget_s=15.0):
    """the reference pgmres (Diag, GMRES(30), 1e-8) on one core, problem by
    problem in order, until the budget is spent; problems per second over
    that sample"""
    import oracle as O
    if not O.Reference.available():
        return None
    solved, sec, steps, err = 0, 0.0, 0, 0.0
    for p, (n, th, kappa) in enumerate(probs):
        if sec > budget_s:
            break
        o = O.Oracle()
        x, y, t, v = o.mesh2d(n)
        r = O.Reference(fast=True)
        r.setup(x, y, t, v, "EBE")
        r.set_problem(O.prob2d(0, kappa, 0.0, 1.0, float(np.cos(th)), float(np.sin(th)), 1.0, 0))
        A, Fv = r.build(np.zeros(r.neq + 1))
        Fsp = r.precond_setup("Diag", Fv)
        t0 = time.perf_counter()
        X, it = r.pgmres(30, 100000, 1e-8, Fsp)
        sec += time.perf_counter() - t0
        steps += it
        Xg = X_gpu[seg[p]:seg[p + 1]]
        err = max(err, float(np.linalg.norm(Xg - X[:-1]) / max(np.linalg.norm(X[:-1]), 1e-300)))
        solved += 1
    cells = sum(probs[p][0] ** 2 for p in range(solved))
    out = dict(value=solved / sec, unit="problems/s", cores=1, kind="reference",
               sample=f"reference pgmres, one problem after another, the first {solved} of "
                      f"{nprob} problems ({cells} cells, {steps} Arnoldi steps) in "
                      f"{sec:.1f} s, -Ofast, one thread",
               seconds=sec, problems=solved,
               gpu_fp32_ir_problems_per_s=nprob / gpu_seconds,
               gpu_fp64_max_rel_diff_vs_cpu_on_sample=err)
    out.update(cpu_info())
    return out


# ---------------------------------------------------------------------------
def ss3d_leg(F, ctx, args, n=128, sweeps=4, seed=20260303):
    """3D SUPG steady transport (DIFCONV) on the n^3 cube cut into 6 n^3
    tetrahedra, a seeded random node numbering then reverse Cuthill-McKee;
    FP32 EBE 4x4 products + ILU(0) of the CSR matrix (its triangular solves
    as `sweeps` Jacobi sweeps) inside FP64 refinement (FP32-IR), GMRES(50)
    to 1e-10, against FP64 GMRES with the same operator and preconditioner,
    and FP64 GMRES with the exact (level-scheduled) ILU(0) solves."""
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
    t_struct = time.perf_counter() - t0
    ke = ctx.empty(16 * nel)
    F0 = ctx.empty(neq + 1)
    F0.zero()
    m.build_supg(F.Supg3dProblem(F.P3_DIFCONV, 1.0, 0, 0, 0, 0, 0, 0), ke, F0)
    Fh = F0.get()
    Fn = np.zeros(neq + 1)
    Fn[:neq] = Fh[P]
    csr = F.Operator.csr(ctx, neq, IA, JA)
    F.assemble_csr(csr, cbe, ke, k=4)
    ss3d_tune = None
    if args.no_autotune:
        ebe = F.Operator.ebe(ctx, neq, lm, 4)
        ebe.set_values(ke)
        ebe_kernel = ebe_kernel_name(4, True, ebe)
    else:
        # the FP32 product of the inner cycle: elements per block x ordering
        from fem_codes_amd import autotune
        key, tms, ebe = autotune.tune_ebe(ctx, neq, lm, 4, ke, fp32=True, reps=10,
                                          pipelines=(0,), orderings=autotune.ORDERINGS)
        ss3d_tune = {"product": autotune.key_str(key),
                     "ms": {autotune.key_str(k): v for k, v in tms.items()}}
        log("ss3d autotune FP32 product ms:", {autotune.key_str(k): round(v, 4)
                                               for k, v in tms.items()}, "->", autotune.key_str(key))
        ebe_kernel = (ebe_kernel_name(4, True, ebe) if ebe.product_sliced(True)
                      else f"k_ebe_apply<4, {key[0] // 256}, float, 0>")
    del cbe, IA, JA
    pc = F.Preconditioner(csr, F.PREC_ILU0)
    levels = pc.ilu_levels()
    pc.ilu_sweeps(sweeps)
    B = ctx.to_device(Fn)
    pc.setup(B)
    ctx.sync()
    log(f"ss3d: n={n} nel={nel} neq={neq} nnz={nnz} bandwidth {bw0} -> {bw1} (RCM), ILU levels "
        f"{levels[0]}, setup {time.perf_counter() - t0:.1f}s (structures {t_struct:.1f}s)")
    rng = np.random.default_rng(1)
    res = {"n": n, "elements": nel, "equations": neq, "nnz": int(nnz),
           "bandwidth_random": int(bw0), "bandwidth_rcm": int(bw1), "ilu_levels": levels[0],
           "ilu_sweeps": sweeps, "kmax": 50, "tol": 1e-10, "structure_seconds": t_struct,
           "fp32_ir_orth": args.orth, "engine": args.engine}
    mv = {}
    P64 = ctx.to_device(rng.standard_normal(neq + 1))
    Q64 = ctx.empty(neq + 1)
    P32 = ctx.to_device(rng.standard_normal(neq + 1).astype(np.float32))
    Q32 = ctx.empty(neq + 1, np.float32)
    for nm, op in (("EBE", ebe), ("CSR", csr)):
        mv[nm] = matvec_table(op, P64, Q64, P32, Q32, 10)
    res["matvec"] = mv
    ms_main = ebe.time_apply(P32, Q32, reps=10, main_only=True)
    res["roofline"] = roofline(ms_main, ebe.algorithmic_bytes(True), ebe_kernel)
    res["autotune"] = ss3d_tune
    res["roofline"]["traffic"] = leg_traffic("ss3d", res["roofline"]["kernel"])
    # ILU(0) apply: 4 sweeps (the solve's) and the exact level-scheduled solves
    Z64 = ctx.empty(neq + 1)
    Z32 = ctx.empty(neq + 1, np.float32)
    ilu = {}
    for tag, (Pv, Zv) in (("f64", (P64, Z64)), ("f32", (P32, Z32))):
        pc.apply(Pv, Zv)
        ctx.sync()
        w0 = time.perf_counter()
        for _ in range(5):
            pc.apply(Pv, Zv)
        ctx.sync()
        ilu[f"sweeps{sweeps}_{tag}_ms"] = (time.perf_counter() - w0) / 5 * 1e3
    pcx = F.Preconditioner(csr, F.PREC_ILU0)
    Bx = ctx.to_device(Fn)
    pcx.setup(Bx)
    pcx.apply(P64, Z64)
    ctx.sync()
    w0 = time.perf_counter()
    for _ in range(3):
        pcx.apply(P64, Z64)
    ctx.sync()
    exact_ms = (time.perf_counter() - w0) / 3 * 1e3
    ilu["exact_f64_ms"] = exact_ms
    res["ilu_apply"] = ilu
    log("ss3d ILU(0) apply", ilu)
    if args.no_solve:
        return res
    Xs = {}
    for prec, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
        X = ctx.empty(neq + 1)
        # untimed: one restart cycle (loads the kernels' code objects)
        F.gmres(ctx, ebe, pc, B, X, kmax=50, lmax=1, tol=1e-10, precision=prec,
                orth=ORTHS[args.orth] if prec == F.GMRES_FP32_IR else F.ORTH_MGS,
                engine=ENGINES[args.engine])
        ctx.sync()
        w0 = time.perf_counter()
        info, _, _ = F.gmres(ctx, ebe, pc, B, X, kmax=50, lmax=400, tol=1e-10, precision=prec,
                             orth=ORTHS[args.orth] if prec == F.GMRES_FP32_IR else F.ORTH_MGS,
                             engine=ENGINES[args.engine])
        ctx.sync()
        wall = time.perf_counter() - w0     # the solution stays in HBM; its copy is not timed
        Xs[tag] = X.get()
        res[tag] = solve_record(info, wall, 1e-10)
        log("ss3d", tag, res[tag])
    res["speedup_fp32_ir"] = res["fp64"]["seconds"] / res["fp32_ir"]["seconds"]
    res["fp32_ir_rel_err_vs_fp64"] = float(np.linalg.norm(Xs["fp32_ir"][:neq] - Xs["fp64"][:neq]) /
                                           np.linalg.norm(Xs["fp64"][:neq]))
    # FP64 with the exact ILU(0) solves (the drop-in's preconditioner): a full
    # solve when its estimated time fits the cap, else one restart cycle and
    # the time per step
    est = 2.2 * exact_ms * 1e-3 * res["fp64"]["steps"]
    cap_cycles = 400 if est < args.exact_ilu_cap_s else 1
    X = ctx.empty(neq + 1)
    ctx.sync()
    w0 = time.perf_counter()
    info, _, _ = F.gmres(ctx, ebe, pcx, Bx, X, kmax=50, lmax=cap_cycles, tol=1e-10)
    ctx.sync()
    ex = solve_record(info, time.perf_counter() - w0, 1e-10)
    ex["cycle_cap"] = cap_cycles
    ex["ms_per_step"] = ex["seconds"] * 1e3 / max(1, info.steps)
    if cap_cycles == 1:
        ex["seconds_extrapolated_to_sweep_steps"] = ex["ms_per_step"] * res["fp64"]["steps"] * 1e-3
    else:
        ex["rel_diff_vs_sweeps_fp64"] = float(np.linalg.norm(X.get()[:neq] - Xs["fp64"][:neq]) /
                                              np.linalg.norm(Xs["fp64"][:neq]))
    res["fp64_exact_ilu"] = ex
    log("ss3d exact ILU fp64", ex)
    if not args.no_cpu:
        import oracle as O
        if O.Reference3D.available():
            r = O.Reference3D(fast=True)
            r.ebemv_setup(neq, lm, ke.get())
            Ph = rng.standard_normal(neq + 1)
            b = nel * (16 * 8 + 16) + 2 * neq * 8
            res["cpu_baseline"] = cpu_product_baseline("ebemv3Dndof1", lambda: r.ebemv(Ph), b,
                                                       10.0, 40, f"{nel}-tetrahedron mesh")
            del r
            log("ss3d cpu", res["cpu_baseline"])
    return res


# ---------------------------------------------------------------------------
def transient_leg(F, ctx, args, n=1024, steps=10):
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
    neq, nel = mesh.neq, mesh.nel
    lm = mesh.lm()
    op = F.Operator.ebe(ctx, neq, lm, 3)
    pc = F.Preconditioner(op, F.PREC_DIAG)
    pb = F.Supg2dProblem(F.P_ROTATE_ABOUT, 1e-8, 0.0, 0.0, 5.0, 5.0, 1.0, F.H_2SQRTAREA)
    up = ctx.to_device(np.zeros(len(x)))
    _, idv = F.number_nodes(t)
    u0 = np.zeros(neq + 1)
    free = t > 0
    u0[idv[free]] = np.exp(-0.5 * ((x[free] - 5.0) ** 2 + (y[free] - 7.5) ** 2))
    dt = 1e-3
    res = {"n": n, "equations": neq, "elements": nel, "steps": steps, "dt": dt,
           "problem": "CONE", "precond": "Diag", "kmax": 30, "tol": 1e-8,
           "engine": args.engine}
    U = {}
    for prec, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
        # untimed: one time step of the same loop (loads the kernels' code
        # objects), on a copy of the initial state
        uw = ctx.to_device(u0)
        F.predictor_old_transport(
            mesh, pb, op, pc, up, uw, alpha=0.5, dt=dt, final_time=steps * dt + 0.5 * dt,
            nl_tol=1e-3, nl_maxiter=10, stop_by_norm=True, steady=False, max_steps=1,
            kmax=30, lmax=200, tol=1e-8, precision=prec, engine=ENGINES[args.engine],
            orth=F.ORTH_CGS if prec == F.GMRES_FP32_IR else F.ORTH_MGS)
        del uw
        u = ctx.to_device(u0)
        ctx.sync()
        t0 = time.perf_counter()
        info, hist = F.predictor_old_transport(
            mesh, pb, op, pc, up, u, alpha=0.5, dt=dt, final_time=steps * dt + 0.5 * dt,
            nl_tol=1e-3, nl_maxiter=10, stop_by_norm=True, steady=False, max_steps=steps,
            kmax=30, lmax=200, tol=1e-8, precision=prec, engine=ENGINES[args.engine],
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
    # the product inside every Arnoldi step (the element matrices of the last
    # correction)
    rng = np.random.default_rng(3)
    P64 = ctx.to_device(rng.standard_normal(neq + 1))
    Q64 = ctx.empty(neq + 1)
    P32 = ctx.to_device(rng.standard_normal(neq + 1).astype(np.float32))
    Q32 = ctx.empty(neq + 1, np.float32)
    res["matvec"] = matvec_table(op, P64, Q64, P32, Q32, 20)
    ms_main = op.time_apply(P64, Q64, reps=20, main_only=True)
    res["roofline"] = roofline(ms_main, op.algorithmic_bytes(False), ebe_kernel_name(3, False, op))
    res["roofline"]["traffic"] = leg_traffic("transient", res["roofline"]["kernel"])
    if not args.no_cpu:
        import oracle as O
        if O.Reference.available():
            r = O.Reference(fast=True)
            r.setup(x, y, t, v, "EBE")
            r.set_matrix(op.get_values())
            Ph = rng.standard_normal(neq + 1)
            b = nel * (9 * 8 + 12) + 2 * neq * 8
            res["cpu_baseline"] = cpu_product_baseline("ebemv", lambda: r.matvec(Ph), b, 6.0, 200,
                                                       f"{nel}-element mesh")
            log("transient cpu", res["cpu_baseline"])
    return res


# ---------------------------------------------------------------------------
def euler_leg(F, ctx, args, n=1024, seed=20260305):
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
    if not args.no_autotune:
        # the FP32 12x12 product of the inner cycle: elements per block x
        # ordering x schedule (blocks of 512 do not fit the LDS at k = 12)
        from fem_codes_amd import autotune
        key, tms, op_t = autotune.tune_ebe(ctx, neq, lm, 12, A, fp32=True, reps=10,
                                           candidates=(256, 512), pipelines=(0, 2, 3),
                                           orderings=autotune.ORDERINGS)
        res["autotune"] = {"product": autotune.key_str(key),
                           "ms": {autotune.key_str(k): v for k, v in tms.items()}}
        log("euler autotune FP32 12x12 product ms:", {autotune.key_str(k): round(v, 4)
                                                      for k, v in tms.items()},
            "->", autotune.key_str(key))
        if key == (256, F.ORDER_BFS, 0):
            op_t.close()            # the default: keep the operator built above
        else:
            pc.close()
            op.close()
            op = op_t
            pc = F.Preconditioner.blockdiag(op, nn, v, ids32)
    P64 = ctx.to_device(rng.standard_normal(neq + 1))
    Q64 = ctx.empty(neq + 1)
    P32 = ctx.to_device(rng.standard_normal(neq + 1).astype(np.float32))
    Q32 = ctx.empty(neq + 1, np.float32)
    res["matvec_ebe12"] = matvec_table(op, P64, Q64, P32, Q32, 20)
    # the row sums on the matrix cores against the vector row sums: the
    # block kernel alone, both precisions; the faster form (FP32, the inner
    # cycle's product) is kept for the blocked products (the node rows, where
    # the operator uses them, stay in use)
    rows_in_use = (op.product_sliced(False), op.product_sliced(True))
    op.set_product_kernel(F.PRODUCT_BLOCKED)
    trial, Qr = {}, {}
    for kind, tag in ((F.ROWS_VALU, "valu"), (F.ROWS_MFMA, "mfma")):
        op.set_row_kernel(kind)
        trial[tag] = {"fp64_ms": op.time_apply(P64, Q64, reps=20, main_only=True),
                      "fp32_ms": op.time_apply(P32, Q32, reps=20, main_only=True)}
        op.apply(P64, Q64)
        Qr[tag] = Q64.get()
    trial["rel_diff_fp64"] = float(np.max(np.abs(Qr["mfma"][:neq] - Qr["valu"][:neq])) /
                                   np.max(np.abs(Qr["valu"][:neq])))
    trial["kept"] = "mfma" if trial["mfma"]["fp32_ms"] < trial["valu"]["fp32_ms"] else "valu"
    op.set_row_kernel(F.ROWS_MFMA if trial["kept"] == "mfma" else F.ROWS_VALU)
    for prec, rows in zip((F.PRODUCT_FP64, F.PRODUCT_FP32), rows_in_use):
        op.set_product_kernel(F.PRODUCT_AUTO if rows else F.PRODUCT_BLOCKED, prec)
    res["rows_trial"] = trial
    log("euler rows VALU vs MFMA", trial)
    ms_main = op.time_apply(P32, Q32, reps=20, main_only=True)
    pipe = getattr(op, "pipeline", 0)
    res["roofline"] = roofline(ms_main, op.algorithmic_bytes(True),
                               ebe_kernel_name(12, True, op) if rows_in_use[1]
                               else "k_ebe_apply_mfma12<float, 0>" if trial["kept"] == "mfma"
                               else "k_ebe_apply_pipe<12, 1, float, 0>" if pipe
                               else ebe_kernel_name(12, True))
    res["roofline"]["traffic"] = leg_traffic("euler", res["roofline"]["kernel"])
    res["fp32_ir_orth"] = args.orth
    res["engine"] = args.engine
    Xs = {}
    for prec, tag in ((F.GMRES_FP64, "fp64"), (F.GMRES_FP32_IR, "fp32_ir")):
        B = ctx.to_device(Rh)
        pc.setup(B)
        X = ctx.empty(neq + 1)
        # untimed: one restart cycle (loads the kernels' code objects)
        F.gmres(ctx, op, pc, B, X, kmax=30, lmax=1, tol=1e-8, precision=prec,
                orth=ORTHS[args.orth] if prec == F.GMRES_FP32_IR else F.ORTH_MGS,
                engine=ENGINES[args.engine])
        ctx.sync()
        w0 = time.perf_counter()
        info, _, _ = F.gmres(ctx, op, pc, B, X, kmax=30, lmax=2000, tol=1e-8, precision=prec,
                             orth=ORTHS[args.orth] if prec == F.GMRES_FP32_IR else F.ORTH_MGS,
                             engine=ENGINES[args.engine])
        ctx.sync()
        wall = time.perf_counter() - w0     # the solution stays in HBM; its copy is not timed
        Xs[tag] = X.get()
        res[tag] = solve_record(info, wall, 1e-8)
        log("euler", tag, res[tag])
    res["speedup_fp32_ir"] = res["fp64"]["seconds"] / res["fp32_ir"]["seconds"]
    res["fp32_ir_rel_err_vs_fp64"] = float(np.linalg.norm(Xs["fp32_ir"][:neq] - Xs["fp64"][:neq]) /
                                           max(np.linalg.norm(Xs["fp64"][:neq]), 1e-300))
    if not args.no_cpu:
        import oracle as O
        if O.ReferenceEuler.available():
            r = O.ReferenceEuler(fast=True)
            r.setup(nn, nel, neq, v, ids32, lm, A.get())
            Ph = rng.standard_normal(neq + 1)
            b = nel * (144 * 8 + 48) + 2 * neq * 8
            res["cpu_baseline"] = cpu_product_baseline("ebemvNDOF4", lambda: r.ebemv(Ph), b, 8.0,
                                                       40, f"{nel}-element mesh")
            del r
            log("euler cpu", res["cpu_baseline"])
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
    ede["matvec_ede8"] = matvec_table(op8, P64, Q64, P32, Q32, 20)
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
    log("euler ede", {k: v for k, v in ede.items() if k != "matvec_ede8"})
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
    ap.add_argument("--legs", default="all",
                    help="comma-separated subset of " + ",".join(LEGS) + " (default all)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-solve", action="store_true",
                    help="headline and 3D legs: products (and the 3D ILU(0) applies) only")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batched leg: skip the one-by-one comparison")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched leg: launch every cycle's kernels directly instead of "
                         "replaying a captured hipGraph (kernel-trace profiling runs)")
    ap.add_argument("--engine", default="kernels", choices=sorted(ENGINES),
                    help="GMRES engine for the 2048^2 time-to-solve runs and the single-system "
                         "solves of the 3D, Euler and transient legs")
    ap.add_argument("--orth", default="cgs", choices=sorted(ORTHS),
                    help="orthogonalisation of the FP32-IR inner cycle (FP64 is always MGS)")
    ap.add_argument("--rot-cycles", type=int, default=40,
                    help="restart cycles allowed to the stagnating rotating-flow solve")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the products, run a cast kernel of known traffic (64Mi "
                         "doubles read, 64Mi floats written) three times, the scale of "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--batched", type=int, default=256, metavar="NPROB")
    ap.add_argument("--ss3d", type=int, default=128, metavar="N")
    ap.add_argument("--exact-ilu-cap-s", type=float, default=40.0,
                    help="3D leg: the exact-ILU(0) FP64 solve runs in full when its estimate "
                         "is below this, else one restart cycle")
    ap.add_argument("--transient", type=int, default=1024, metavar="N")
    ap.add_argument("--euler", type=int, default=1024, metavar="N")
    ap.add_argument("--matvec-only", action="store_true",
                    help="only the headline EBE products (for profiling runs)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="headline leg: the default 512-element blocks, no tuning sweep")
    ap.add_argument("--no-headline", action="store_true",
                    help="profiling runs of the other legs: skip the headline leg (the JSON "
                         "line then has no value)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              "(one GPU per process; the sharding over several GPUs is a design only)",
              file=sys.stderr)
        sys.exit(2)
    legs = set(LEGS) if args.legs == "all" else set(args.legs.split(","))
    unknown = legs - set(LEGS)
    if unknown:
        ap.error(f"unknown legs {sorted(unknown)}")
    if args.matvec_only:
        legs = {"headline"}
    if not args.no_headline:
        legs.add("headline")   # `value` is the headline product

    import fem_codes_amd as F
    ctx = F.Context(0)
    rng = np.random.default_rng(args.seed)
    t_start = time.perf_counter()
    if "headline" in legs:
        head, sysd = headline_leg(F, ctx, args, rng)
        # free the 2048^2 system before the other legs
        for op in sysd["ops"].values():
            op.close()
        sysd["mesh"].close()
        del sysd
        ctx.trim()
    else:
        rng.uniform(0.5, 2.0), rng.uniform(-6.0, -3.0)
        head = {"theta": float(rng.uniform(0.0, 2 * np.pi)), "value": None, "ms": None,
                "nel": None, "neq": None, "scale": None, "kappa": None, "roofline": None,
                "matvec": None, "ebe_autotune_ms": None, "wall_ms_per_step": None}
    theta = head["theta"]
    small = small_leg(F, ctx, args, theta) if "small" in legs else None
    batched = batched_leg(F, ctx, args, args.batched, args.seed + 4) if "batched" in legs else None
    ctx.trim()
    ss3d = ss3d_leg(F, ctx, args, args.ss3d) if "ss3d" in legs else None
    ctx.trim()
    euler = euler_leg(F, ctx, args, args.euler) if "euler" in legs else None
    ctx.trim()
    transient = transient_leg(F, ctx, args, args.transient) if "transient" in legs else None
    log(f"total {time.perf_counter() - t_start:.1f}s")

    n = args.n
    out = {
        "metric": METRIC, "value": head["value"], "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D SUPG transport, {n}x{n} grid of P1 triangles "
                               f"({head['nel']} elements, {head['neq']} equations), EBE 3x3 "
                               f"FP64 matvec",
                   "velocity_scale": head["scale"], "diffusivity": head["kappa"],
                   "pct_of_peak": 100 * head["value"] / PEAK_GBS if head["value"] else None,
                   "wall_ms_per_step": head["wall_ms_per_step"]},
        "roofline": head["roofline"],
        "cpu_baseline": head.get("cpu_baseline"),
        "matvec": head["matvec"],
        "copy_calibration": head.get("copy_calibration"),
        "ebe_autotune_ms": head["ebe_autotune_ms"],
        "ebe_fp32_product_ms": head.get("ebe_fp32_product_ms"),
        "ede_autotune": head.get("ede_autotune"),
        "solve": head.get("solve"),
        "small": small,
        "batched": batched,
        "ss3d": ss3d,
        "euler": euler,
        "transient": transient,
        "host": cpu_info(),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
