#!/usr/bin/env python3
"""Benchmark of the SIFEL linear-system path on one MI355X.

Workload (default): 3D steady heat conduction, hex8 mesh 128^3 elements
(2 113 407 unknowns), lognormal conductivity (mean 1, sigma_log 0.5, seed 2),
T=0 on z=0 and T=100 on z=1. The matrix is assembled on the GPU by the batched
hex8 conduction kernel; inputs are resident in HBM before timing starts.

A step is one solve of that system to relative error <= 1e-8 against x_ref
(FP64 GPU PCG run to the limit of FP64, itself checked against the CPU
restatement of the reference on smaller meshes by the tests):
  * value = solves/s of the mixed form (FP32 Jacobi-PCG, FP64 reliable updates)
  * fp64  = solves/s of the plain FP64 GPU PCG (the drop-in recurrence)
The roofline figure is the FP32 SpMV of the mixed solve: the bytes of the
format it streams per launch, divided by its mean duration, sampled inside the
timed solves with the launch's own start/stop events (hipExtLaunchKernelGGL).
The post-solve element pass (temperature gradients and fluxes at the 8 points of
every element, assembled internal fluxes) runs on the mixed solution and is
reported under "post" with its own bytes and kernel time. The nonstationary leg
("transient") repeats TRFEL's one_step_linear on the same cube with device-resident
vectors. Further legs, each with its CPU baseline: reverse Cuthill-McKee of the
cube's mesh ("renumber"), cg_prec with the incomplete LDL^T preconditioner
("ildl"), comprow bicg with nonsymmetric values ("bicg"), the FP64 and mixed
solves of the 64x64x256 elasticity column ("elasticity"), 512 quad4 plates with
8 load cases ("batched") and 24 served requests ("served"); every leg has a
--no-* switch. --only-assembly times the element kernels and the post pass of
the 64x64x256 elasticity column. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12



TOL = 1e-8  # the relative error every timed result must meet


def gated(speedup, *errs):
    """A speedup stands only when every result it compares meets TOL; else None."""
    return speedup if all(e is not None and e <= TOL for e in errs) else None

def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    """CPU model and the threads this process may use (the box's host cores)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "available_to_process": avail}


class OracleSolve(threading.Thread):
    """The oracle (CPU restatement of the reference's symcomprow assembly and
    cg_prec, one thread) on a problem, in the background: its own assembly from
    the element data, then cg_prec from x = 0 to `res`. The ctypes calls release
    the GIL, so it runs beside the GPU legs."""

    def __init__(self, prob, res, ni=400000, rhs=None):
        super().__init__(daemon=True)
        self.prob, self.res, self.ni, self.rhs = prob, res, ni, rhs
        self.err = None

    def run(self):
        try:
            import oracle as O
            t = time.perf_counter()
            self.adr, self.ci, self.a, rhs = O.assemble_problem(self.prob, "scr")
            self.y = rhs[0] if self.rhs is None else self.rhs
            self.t_assemble = time.perf_counter() - t
            t = time.perf_counter()
            self.x, self.ani, self.ares, self.status = O.scr_cg_prec(self.adr, self.ci, self.a,
                                                                     self.y, self.ni, self.res)
            self.t_solve = time.perf_counter() - t
        except Exception as e:  # reported by the caller
            self.err = repr(e)

    def summary(self):
        return {"res": self.res, "status": int(self.status), "ani": int(self.ani),
                "ares": float(self.ares), "assembly_s": self.t_assemble, "solve_s": self.t_solve,
                "threads": 1}


def cpu_parallel(adr, ci, a, y, its, k=10):
    """P = the process's available CPUs (at most 16) concurrent oracle cg_prec runs
    of k iterations each on the same system (independent vectors), as the host
    would solve independent systems side by side; scaled to `its` iterations per
    solve. Returns the cpu_baseline fields of this leg."""
    import oracle as O
    P = max(1, min(16, cpu_info()["available_to_process"]))
    done = []

    def one():
        O.scr_cg_prec(adr, ci, a, y, k, 0.0)
        done.append(1)

    th = [threading.Thread(target=one) for _ in range(P)]
    t = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    dt = time.perf_counter() - t
    return {"value": P / (dt / k * its), "unit": "solves/s", "cores": P,
            "sample": f"{P} concurrent oracle cg_prec runs of {k} iterations each ({dt:.2f} s), "
                      f"scaled to {its} iterations per solve"}


def bench_batched(ctx, args, gpu, mesh, torch):
    """512 independent quad4 plane-stress plates (1e3..6e4 DOF, seed 4), 8 load cases
    each: 4096 solves in one batch. Compared with the drop-in path, which solves
    the 4096 systems one after the other with the FP64 GPU PCG (timed on every
    `seq_stride`-th plate, all 8 load cases, and scaled to the 512 plates).
    Errors are against the batched FP64 recurrence run to 1e-28, itself checked
    against the oracle's FP64 CPU cg_prec (to 1e-30) on a sample of plates. The
    mixed form stops on its true residual, which bounds the error of these
    high-condition plates less tightly than the FP64 recurrence's recursive
    residual does: it runs at --batch-res-mixed (1e-22)."""
    t0 = time.time()
    probs = mesh.plate_batch(args.batch_count, seed=4, nrhs=8)
    mats, rhs = [], []
    for p in probs:
        m, r = gpu.build_problem_matrix(ctx, p)
        mats.append(m)
        rhs.append(r)
    b = gpu.Batch(ctx, mats, 8)
    ndof = int(sum(p.n for p in probs))
    log(f"[batch] {len(probs)} plates, {ndof} DOF, {b.nnz} entries, 8 rhs "
        f"({time.time() - t0:.1f}s to generate and assemble)")
    lanes64 = b.autotune(gpu.FP64)
    lanes32 = b.autotune(gpu.MIXED)
    dev = torch.device("cuda:0")
    y = torch.tensor(b.pack(rhs), dtype=torch.float64, device=dev)
    x = torch.zeros_like(y)
    xref = torch.zeros_like(y)
    torch.cuda.synchronize()
    b.solve_dev(y.data_ptr(), xref.data_ptr(), ni=40000, res=1e-28)
    nq = len(probs) * 8
    row0 = b.row0

    starts = np.asarray(row0[:-1], dtype=np.int64)

    def pair_errors():
        """relative error of every (plate, load case) pair, in pair order s * 8 + r
        (per-plate sums over its contiguous rows on the host: a device index_add
        into 512 x 8 sums contends on a few addresses and takes seconds)"""
        d = (x - xref).view(-1, 8).cpu().numpy()
        rr = xref.view(-1, 8).cpu().numpy()
        num = np.add.reduceat(d * d, starts, axis=0)
        den = np.add.reduceat(rr * rr, starts, axis=0)
        return np.sqrt(num / den).reshape(-1)

    oracle_check = None
    if not args.no_cpu:
        # x_ref against the oracle's FP64 CPU solve on a sample of plates (the
        # largest among them), first load case
        import oracle as O
        xr = b.unpack(xref.cpu().numpy())
        big = int(np.argmax([p.n for p in probs]))
        sample = sorted({0, len(probs) // 3, 2 * len(probs) // 3, big})
        errs = {}
        t = time.perf_counter()
        for s in sample:
            adr, ci = mats[s].pattern()
            xo, ani, _, st = O.scr_cg_prec(adr, ci, mats[s].get_values(), rhs[s][0], 400000, 1e-30)
            errs[str(s)] = {"n": int(probs[s].n), "oracle_status": int(st),
                            "oracle_iterations": int(ani + 1),
                            "rel_err_x_ref": float(np.linalg.norm(xr[s][0] - xo) / np.linalg.norm(xo))}
        oracle_check = {"plates": errs, "seconds": time.perf_counter() - t,
                        "max_rel_err_x_ref": max(v["rel_err_x_ref"] for v in errs.values())}
        log(f"[batch] x_ref vs the oracle: {oracle_check}")
    out = {}
    for name, prec, res in (("fp64", gpu.FP64, args.batch_res),
                            ("mixed", gpu.MIXED, args.batch_res_mixed)):
        for _ in range(args.warmup):
            b.solve_dev(y.data_ptr(), x.data_ptr(), ni=40000, res=res, precision=prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            infos = b.solve_dev(y.data_ptr(), x.data_ptr(), ni=40000, res=res,
                                precision=prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        its = [i.iterations for i in infos]
        err = pair_errors()
        st = np.array([i.status for i in infos])
        by_status = {int(k): {"pairs": int((st == k).sum()), "max_rel_err": float(err[st == k].max())}
                     for k in np.unique(st)}
        out[name] = {"solves_per_s": nq / dt, "ms_per_batch": 1e3 * dt, "res": res,
                     "max_rel_err": float(err.max()), "by_status": by_status,
                     "max_iterations": int(max(its)), "mean_iterations": float(np.mean(its))}
        log(f"[batch] {name}: {out[name]}")

    # the drop-in path: one FP64 solve per (plate, load case), one after the other
    stride = args.seq_stride
    sample = list(range(0, len(probs), stride))
    ys = [torch.tensor(rhs[s], dtype=torch.float64, device=dev) for s in sample]
    xs = [torch.zeros(probs[s].n, dtype=torch.float64, device=dev) for s in sample]
    for j, s in enumerate(sample[:4]):
        mats[s].solve_dev(ys[j][0].data_ptr(), xs[j].data_ptr(), ni=40000, res=args.batch_res)
    torch.cuda.synchronize()
    t = time.perf_counter()
    nseq = 0
    for j, s in enumerate(sample):
        for r in range(8):
            mats[s].solve_dev(ys[j][r].data_ptr(), xs[j].data_ptr(), ni=40000, res=args.batch_res)
            nseq += 1
    torch.cuda.synchronize()
    dt_seq = (time.perf_counter() - t) * (len(probs) / len(sample))
    seq = nq / dt_seq
    log(f"[batch] sequential FP64 drop-in: {nseq} solves timed, {seq:.1f} solves/s (scaled)")
    for m in mats:
        m.close()
    return {
        "workload": f"{len(probs)} quad4 plates (seed 4), {ndof} DOF, 8 load cases each",
        "solves": nq, "res": {"fp64": args.batch_res, "mixed": args.batch_res_mixed},
        "error_reference": "batched FP64 recurrence to 1e-28 (x_ref), checked against the oracle",
        "oracle_check_x_ref": oracle_check,
        "launch": "one hipGraph launch per batched solve (the graph holds the chunks the "
                  "previous solve needed, plus one; persistent kernels over the running systems)",
        "batched_mixed_solves_per_s": out["mixed"]["solves_per_s"],
        "batched_fp64_solves_per_s": out["fp64"]["solves_per_s"],
        "sequential_fp64_solves_per_s": seq,
        "speedup_mixed_vs_sequential_fp64": gated(out["mixed"]["solves_per_s"] / seq,
                                                  out["mixed"]["max_rel_err"]),
        "speedup_mixed_vs_batched_fp64": gated(out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"],
                                               out["mixed"]["max_rel_err"], out["fp64"]["max_rel_err"]),
        "speedup_fp64_batched_vs_sequential_fp64": gated(out["fp64"]["solves_per_s"] / seq,
                                                         out["fp64"]["max_rel_err"]),
        "max_rel_err": {"mixed": out["mixed"]["max_rel_err"], "fp64": out["fp64"]["max_rel_err"]},
        "meets_1e-8": {"mixed": out["mixed"]["max_rel_err"] <= TOL,
                       "fp64": out["fp64"]["max_rel_err"] <= TOL},
        "detail": out, "spmm_format_lanes": {"fp64": lanes64, "fp32": lanes32},
        "sequential_sample": f"every {stride}th plate ({len(sample)} plates x 8 load cases), "
                             f"time scaled by {len(probs) / len(sample):.0f}",
    }


def bench_served(ctx, args, gpu, mesh, torch):
    """Served solves: a Poisson stream of requests (seed 5), half hex8 elasticity
    columns and half heat cubes of 5e3..2e6 DOF (log-uniform), admitted under a
    200 GB HBM budget, for each arrival rate of --serve-rates and each precision
    of --serve-precisions (FP64: the cg_prec recurrence per request; mixed: FP32
    recurrences refreshed in FP64 every chunk). Latency = finished - submitted per
    request (server clock). Compared with the drop-in path, which builds,
    assembles and solves the same requests one after the other with the FP64 GPU
    PCG; errors are against those drop-in solutions, which are checked against the
    oracle's FP64 CPU cg_prec on the --serve-oracle smallest requests."""
    rng = np.random.default_rng(5)
    t0 = time.time()
    probs = [mesh.served_request(rng) for _ in range(args.serve_count)]
    ndof = [p.n for p in probs]
    log(f"[serve] {len(probs)} requests, {sum(ndof)} DOF in total (max {max(ndof)}), "
        f"generated in {time.time() - t0:.1f}s")
    out = {"workload": f"{len(probs)} requests (seed 5): hex8 elasticity / heat, "
                       f"5e3..2e6 DOF, {sum(ndof)} DOF in total",
           "res": args.serve_res, "budget_bytes": 200 << 30, "chunk": args.serve_chunk}
    # the drop-in path, one request after the other: pattern, assembly, FP64 solve
    t = time.perf_counter()
    seq_x = []
    for p in probs:
        m, rhs = gpu.build_problem_matrix(ctx, p)
        x, info = m.solve(rhs[0], ni=100000, res=args.serve_res)
        seq_x.append(x)
        m.close()
    dt = time.perf_counter() - t
    out["sequential_fp64_solves_per_s"] = len(probs) / dt
    log(f"[serve] sequential drop-in {out['sequential_fp64_solves_per_s']:.2f} solves/s")
    if not args.no_cpu and args.serve_oracle > 0:
        import oracle as O
        small = np.argsort(ndof)[:args.serve_oracle]
        chk = {}
        for k in small:
            adr, ci, a, rhs_o = O.assemble_problem(probs[k], "scr")
            xo, ani, _, st = O.scr_cg_prec(adr, ci, a, rhs_o[0], 400000, 1e-30)
            chk[str(int(k))] = {"n": int(ndof[k]), "oracle_iterations": int(ani + 1),
                                "rel_err_drop_in": float(np.linalg.norm(seq_x[k] - xo) /
                                                         np.linalg.norm(xo))}
        out["oracle_check_drop_in"] = chk
        log(f"[serve] drop-in vs the oracle: {chk}")
    for pname in args.serve_precisions:
        prec = gpu.MIXED if pname == "mixed" else gpu.FP64
        rates = {}
        for rate in args.serve_rates:
            arr = np.cumsum(np.random.default_rng(7).exponential(1.0 / rate, len(probs)))
            srv = gpu.Server(budget_bytes=200 << 30, res=args.serve_res, chunk=args.serve_chunk,
                             prep_threads=args.serve_prep, precision=prec)
            try:
                tickets = []
                start = time.perf_counter()
                for p, ta in zip(probs, arr):
                    dt = ta - (time.perf_counter() - start)
                    if dt > 0:
                        time.sleep(dt)
                    tickets.append(srv.submit(p))
                res = [srv.wait(t) for t in tickets]
                st = srv.stats()
            finally:
                srv.close()
            times = np.array([r[2] for r in res])
            lat = times[:, 3] - times[:, 0]
            span = times[:, 3].max() - times[:, 0].min()
            errs = np.array([np.linalg.norm(r[0] - b) / np.linalg.norm(b)
                             for r, b in zip(res, seq_x)])
            stat = np.array([r[1].status for r in res])
            rates[str(rate)] = {
                "offered_req_per_s": rate, "solves_per_s": len(probs) / span,
                "p50_latency_s": float(np.percentile(lat, 50)),
                "p99_latency_s": float(np.percentile(lat, 99)),
                "max_latency_s": float(lat.max()),
                "mean_queue_s": float(np.mean(times[:, 1] - times[:, 0])),
                "mean_assembly_s": float(np.mean(times[:, 2] - times[:, 1])),
                "mean_solve_s": float(np.mean(times[:, 3] - times[:, 2])),
                "status_counts": {str(int(k)): int(v) for k, v in
                                  zip(*np.unique(stat, return_counts=True))},
                "max_rel_err_vs_drop_in": float(errs.max()),
                "peak_bytes": st["peak_bytes"], "steps": st["steps"],
                "prep_thread_s": {"pattern": st["pattern_s"], "setup": st["setup_s"],
                                  "assembly": st["assembly_s"]}}
            # served results are compared with the drop-in, itself checked against the
            # oracle: the error against the oracle is at most the sum of the two
            err = float(errs.max()) + max(
                (v["rel_err_drop_in"] for v in out.get("oracle_check_drop_in", {}).values()),
                default=0.0)
            rates[str(rate)]["meets_1e-8"] = err <= TOL
            rates[str(rate)]["speedup_vs_sequential_fp64"] = gated(
                rates[str(rate)]["solves_per_s"] / out["sequential_fp64_solves_per_s"], err)
            log(f"[serve] {pname} rate {rate}/s: {rates[str(rate)]}")
        out[pname] = rates
    best = {p: max(v["solves_per_s"] for v in out[p].values()) for p in args.serve_precisions}
    out["best_solves_per_s"] = best
    return out


def bench_assembly(ctx, args, gpu, mesh, torch=None):
    """Element kernels + assembly by entry of the 64x64x256 hex8 elasticity column
    (1 048 576 elements, 3 244 800 unknowns): the whole ssf_matrix_assemble call
    (host uploads of the element arrays, element kernel, per-entry gather, value
    expansion) after a warm-up call that builds the assembly plan."""
    prob = mesh.elastic_column(args.asm_nel, args.asm_nel, 4 * args.asm_nel, seed=3)
    m = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, prob.elem_codes())
    m.assemble(prob, lift=False)  # builds the assembly plan
    ts = []
    for _ in range(args.asm_reps):
        m.zero()
        t = time.perf_counter()
        m.assemble(prob, lift=False)
        ts.append(time.perf_counter() - t)
    dt = min(ts)
    flop = 69120.0 * prob.ne
    out = {"workload": f"hex8 elasticity {args.asm_nel}x{args.asm_nel}x{4 * args.asm_nel}, "
                       f"{prob.ne} elements, n={prob.n}",
           "assemble_call_s": dt, "elements_per_s": prob.ne / dt,
           "reference_flop_per_element": 69120,
           "gflops_call": flop / dt / 1e9}
    log(f"[assembly] {out}")
    m.close()
    if args.asm_heat_nel:
        # the conductivity matrix of the heat cube (hex8, 8 DOFs per element)
        hp = mesh.heat_cube(args.asm_heat_nel, seed=2)
        mh = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, hp.n, hp.elem_codes())
        mh.assemble(hp, lift=True)
        th = []
        for _ in range(args.asm_reps):
            mh.zero()
            t = time.perf_counter()
            mh.assemble(hp, lift=True)
            th.append(time.perf_counter() - t)
        out["heat"] = {"workload": f"hex8 heat {args.asm_heat_nel}^3, {hp.ne} elements, n={hp.n}",
                       "assemble_call_s": min(th), "elements_per_s": hp.ne / min(th)}
        log(f"[assembly] heat {out['heat']}")
        mh.close()
    if torch is not None:
        x = torch.tensor(np.random.default_rng(1).standard_normal(prob.n) * 1e-3,
                         dtype=torch.float64, device=torch.device("cuda:0"))
        out["post"] = bench_post(ctx, args, gpu, mesh, torch, prob, x,
                                 f"hex8 elasticity {args.asm_nel}x{args.asm_nel}x{4 * args.asm_nel}")
    return out


def post_bytes(prob, ngp, ncomp, mesh):
    """Algorithmic HBM bytes of one post-solve pass: every input read once (element
    nodes and code numbers as int32, material, node coordinates, the solution and
    the prescribed values) and every output written once (strains, stresses, the
    assembled forces). Returns (element kernel bytes, whole pass bytes)."""
    ne, nne, nd = prob.ne, prob.nne, prob.ndofe
    nmat = 1 + (prob.kind != mesh.KIND_HEX8_HEAT) + (prob.kind == mesh.KIND_QUAD4_PSTRESS)
    rd = ne * (4 * nne + 4 * nd + 8 * nmat) + 24 * prob.coords.shape[0] + 8 * prob.n
    rd += 8 * prob.pv.size
    ip = 2 * ne * ngp * ncomp * 8
    return rd + ip, rd + ip + 8 * prob.n


def bench_post(ctx, args, gpu, mesh, torch, prob, x_dev, label):
    """Post-solve element pass (integration-point strains / stresses or gradients /
    fluxes, assembled internal forces) of the solution x_dev on prob: mean device
    time of the element kernel (its own start / stop events) and of the whole call,
    against the algorithmic bytes; CPU baseline: the oracle's post pass on a
    smaller mesh of the same kind, scaled per element."""
    el = gpu.Elements(ctx, prob)
    dev = torch.device("cuda:0")
    shp = (prob.ne, el.ngp, el.ncomp)
    eps = torch.empty(shp, dtype=torch.float64, device=dev)
    sig = torch.empty(shp, dtype=torch.float64, device=dev)
    ifor = torch.empty(prob.n, dtype=torch.float64, device=dev)
    pv = torch.tensor(prob.pv, dtype=torch.float64, device=dev) if prob.pv.size else None
    torch.cuda.synchronize()
    pvp = pv.data_ptr() if pv is not None else None
    npv = int(prob.pv.size)
    tk, tc = el.bench(x_dev.data_ptr(), pvp, npv, eps.data_ptr(), sig.data_ptr(),
                      ifor.data_ptr(), reps=args.post_reps)
    bk, bc = post_bytes(prob, el.ngp, el.ncomp, mesh)
    out = {"workload": f"{label}, {prob.ne} elements, n={prob.n}",
           "element_kernel_us": tk, "call_us": tc,
           "element_kernel_bytes": bk, "call_bytes": bc,
           "element_kernel_gbs": bk / (tk * 1e-6) / 1e9, "call_gbs": bc / (tc * 1e-6) / 1e9,
           "elements_per_s": prob.ne / (tc * 1e-6)}
    out["element_kernel_frac"] = out["element_kernel_gbs"] * 1e9 / PEAK_HBM
    el.close()
    del eps, sig, ifor, pv
    torch.cuda.empty_cache()
    if not args.no_cpu:
        import oracle as O
        small = (mesh.heat_cube(32, seed=2) if prob.kind == mesh.KIND_HEX8_HEAT else
                 mesh.elastic_column(16, 16, 64, seed=3) if prob.kind == mesh.KIND_HEX8_ELAST
                 else mesh.plate(256, 128, seed=4, nrhs=1))
        xs = np.random.default_rng(0).standard_normal(small.n) * 1e-3
        t = time.perf_counter()
        O.elem_results(small, xs)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": small.ne / dt, "unit": "elements/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle post pass (cr_elem_results_mesh, 1 thread) on "
                                         f"{small.name}, {small.ne} elements"}
        out["speedup_vs_cpu"] = out["elements_per_s"] / out["cpu_baseline"]["value"]
    log(f"[post] {out}")
    return out


def bench_transient(ctx, args, gpu, mesh, torch, prob, K, rhs):
    """Nonstationary heat on the heat cube: one_step_linear (generalized trapezoidal
    rule, alpha 0.5, dt = --np-dt) repeated --np-steps times from T = v = 0 with
    c = rho = 1 and the Dirichlet lift as the right-hand side. Everything stays in
    HBM between steps; the system matrix C + alpha dt K is formed once. CPU
    baseline: the oracle's one_step_linear on the same system, one step."""
    cn = prob.elem_codes()
    t0 = time.time()
    C = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    A = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    C.assemble_capacity(prob, np.ones(prob.ne))
    log(f"[np] capacity matrix and system matrix set up ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda:0")
    f = torch.tensor(rhs, dtype=torch.float64, device=dev)
    out = {"workload": f"heat hex8 {args.nel}^3 transient, n={prob.n}, alpha 0.5, "
                       f"dt {args.np_dt}, c = rho = 1, {args.np_steps} steps from T = 0"}
    for name, prec in (("fp64", gpu.FP64), ("mixed", gpu.MIXED)):
        T = torch.zeros(prob.n, dtype=torch.float64, device=dev)
        v = torch.zeros(prob.n, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        gpu.np_step_linear(K, C, A, 0.5, args.np_dt, f.data_ptr(), T.data_ptr(), v.data_ptr(),
                           res=args.res, precision=prec)  # warm-up step (forms A)
        torch.cuda.synchronize()
        its = []
        t = time.perf_counter()
        for _ in range(args.np_steps):
            info = gpu.np_step_linear(K, C, A, 0.5, args.np_dt, f.data_ptr(), T.data_ptr(),
                                      v.data_ptr(), res=args.res, precision=prec)
            its.append(int(info.iterations))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        out[name] = {"steps_per_s": args.np_steps / dt, "ms_per_step": 1e3 * dt / args.np_steps,
                     "iterations_per_step": float(np.mean(its)), "status": int(info.status),
                     "T_max": float(T.max().item())}
        if name == "fp64":
            Tref = T.clone()
        else:
            out[name]["rel_diff_vs_fp64"] = (torch.linalg.norm(T - Tref) /
                                             torch.linalg.norm(Tref)).item()
        log(f"[np] {name}: {out[name]}")
    if not args.no_cpu:
        import oracle as O
        adr, ci = K.pattern()
        aK, aC = K.get_values(), C.get_values()
        Tc, vc = np.zeros(prob.n), np.zeros(prob.n)
        t = time.perf_counter()
        st, ani, _, _ = O.np_step_linear(adr, ci, aK, aC, 0.5, args.np_dt, rhs, Tc, vc, 200000,
                                         args.res)
        dtc = time.perf_counter() - t
        out["cpu_baseline"] = {"value": 1.0 / dtc, "unit": "steps/s", "cores": 1, "kind": "port",
                               "sample": f"one oracle one_step_linear (1 thread) on the same "
                                         f"system from T = 0 ({ani + 1} iterations)"}
        out["speedup_fp64_vs_cpu"] = out["fp64"]["steps_per_s"] * dtc
    C.close()
    A.close()
    return out


def skew_values(adr, ci, a, gamma, seed):
    """Nonsymmetric values on a comprow pattern: each off-diagonal a_ij scaled by
    1 + gamma t_ij, t_ij uniform in [-1, 1] per entry (a_ij and a_ji independently)."""
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(adr.shape[0] - 1, dtype=np.int64), np.diff(adr))
    t = rng.uniform(-1.0, 1.0, a.shape[0])
    return np.where(ci == rows, a, a * (1.0 + gamma * t))


def bench_elastic(ctx, args, gpu, mesh, torch):
    """Linear elasticity on the hex8 column nel x nel x 4 nel (64 x 64 x 256 by
    default: 3.2M unknowns, 257M stored entries of the full matrix; lognormal E,
    uniform nu, clamped base, nodal body forces; seed 3). Host pattern with row
    patterns, device assembly, then FP64 Jacobi-PCG (the drop-in) and the mixed
    solve to --res from x = 0; errors against an FP64 solve to 1e-30 (which ends on
    the zero-denominator test once the residual is at rounding level). The SpMV
    launches are timed live as in the heat leg. CPU baseline: --cpu-iters
    iterations of the oracle's cg_prec on the same system, scaled to the FP64
    iteration count."""
    nel = args.elast_nel
    # the oracle's FP64 CPU solve of a mid-size column of the same family, beside
    oref = None
    if not args.no_cpu:
        mid = mesh.elastic_column(args.elast_oracle_nel, args.elast_oracle_nel,
                                  4 * args.elast_oracle_nel, seed=3)
        oref = OracleSolve(mid, 1e-30)
        oref.start()
    t0 = time.time()
    prob = mesh.elastic_column(nel, nel, 4 * nel, seed=3)
    m = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, prob.elem_codes())
    t_pat = time.time() - t0
    t0 = time.time()
    lift = m.assemble(prob)
    rhs = prob.rhs[0] + (lift if lift is not None else 0.0)
    t_asm = time.time() - t0
    if args.elast_tiling64:
        m.set_tune(gpu.FP64, args.elast_tiling64)
        tune64 = tuple(args.elast_tiling64)
    else:
        tune64 = m.autotune(gpu.FP64)
    if args.elast_tiling32:
        m.set_tune(gpu.MIXED, args.elast_tiling32)
        tune32 = tuple(args.elast_tiling32)
    else:
        tune32 = m.autotune(gpu.MIXED)
    log(f"[elast] n={prob.n} nnz={m.nnz}: pattern {t_pat:.1f}s, assembly {t_asm:.2f}s, "
        f"tilings {tune64} {tune32}")
    dev = torch.device("cuda:0")
    b = torch.tensor(rhs, dtype=torch.float64, device=dev)
    x = torch.zeros(prob.n, dtype=torch.float64, device=dev)
    xref = torch.zeros(prob.n, dtype=torch.float64, device=dev)
    info_ref = m.solve_dev(b.data_ptr(), xref.data_ptr(), ni=400000, res=1e-30)
    torch.cuda.synchronize()
    xref_norm = torch.linalg.norm(xref).item()
    log(f"[elast] x_ref: status {info_ref.status} ani {info_ref.ani}")
    out = {"workload": f"hex8 elasticity column {nel}x{nel}x{4 * nel}, n={prob.n}, nnz={m.nnz}",
           "res": args.res, "refinement_interval": args.interval,
           "spmv_tiling": {"fp64": list(tune64), "fp32": list(tune32)},
           "pattern_s": t_pat, "assembly_s": t_asm}
    for name, prec, which in (("fp64", gpu.FP64, 0), ("mixed", gpu.MIXED, 1)):
        def solve():
            if prec == gpu.FP64:
                return m.solve_dev(b.data_ptr(), x.data_ptr(), ni=400000, res=args.res)
            return m.solve_dev(b.data_ptr(), x.data_ptr(), ni=400000, res=args.res,
                               precision=gpu.MIXED, inner_ni=args.interval)
        solve()  # warm-up
        torch.cuda.synchronize()
        ctx.set_profiling(True)
        t = time.perf_counter()
        info = solve()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        spmv_us, samples = ctx.kernel_time(which)
        ctx.set_profiling(False)
        can, act = m.spmv_bytes(prec)
        out[name] = {"solves_per_s": 1.0 / dt, "ms_per_solve": 1e3 * dt,
                     "rel_err": torch.linalg.norm(x - xref).item() / xref_norm,
                     "iterations": int(info.iterations), "refinements": int(info.outer),
                     "status": int(info.status), "spmv_us_live": spmv_us, "spmv_samples": samples,
                     "spmv_bytes": act, "spmv_gbs": act / (spmv_us * 1e-6) / 1e9 if spmv_us else None}
        log(f"[elast] {name}: {out[name]}")
    out["speedup_mixed_vs_fp64"] = gated(out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"],
                                         out["mixed"]["rel_err"], out["fp64"]["rel_err"])
    out["meets_1e-8"] = {k: out[k]["rel_err"] <= TOL for k in ("fp64", "mixed")}
    out["error_reference"] = "the GPU FP64 recurrence to 1e-30 on this column"
    if oref is not None:
        # both forms on the mid-size column against the oracle's solution
        mm, rr = gpu.build_problem_matrix(ctx, mid)
        oref.join()
        chk = {"workload": f"hex8 elasticity column {mid.name}, n={mid.n}"}
        if oref.err:
            chk["error"] = oref.err
        else:
            chk["oracle"] = oref.summary()
            no = np.linalg.norm(oref.x)
            for name, prec in (("fp64", gpu.FP64), ("mixed", gpu.MIXED)):
                xm, im = mm.solve(rr[0], ni=400000, res=args.res, precision=prec,
                                  inner_ni=args.interval)
                chk[name] = {"rel_err": float(np.linalg.norm(xm - oref.x) / no),
                             "status": int(im.status), "iterations": int(im.iterations)}
        mm.close()
        out["oracle_check_mid_size"] = chk
        log(f"[elast] mid-size check against the oracle: {chk}")
    if not args.no_cpu:
        import oracle as O
        adr, ci = m.pattern()
        a = m.get_values()
        k = max(1, args.cpu_iters // 10)
        t = time.perf_counter()
        O.scr_cg_prec(adr, ci, a, rhs, k, 0.0)
        per_it = (time.perf_counter() - t) / k
        its = out["fp64"]["iterations"]
        out["cpu_baseline"] = {"value": 1.0 / (per_it * its), "unit": "solves/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{k} iterations of the oracle cg_prec (scr, Jacobi, 1 "
                                         f"thread) on the same system, scaled to the {its} "
                                         f"iterations of the FP64 recurrence"}
        out["speedup_fp64_vs_cpu"] = out["fp64"]["solves_per_s"] / out["cpu_baseline"]["value"]
    m.close()
    return out


def bench_bicg(ctx, args, gpu, mesh, torch, prob):
    """comprow::bicg on the heat cube stored as comprow with nonsymmetric values
    (off-diagonals scaled by 1 + 0.3 t, seeded), from x = 0 to --res, FP64 (the
    reference's bicg has no preconditioner). The dual SELL product A d, A^T dd is
    the dominant kernel; its launches are timed live (second product of every
    iteration chunk). CPU baseline: --cpu-iters iterations of the oracle's bicg
    on the same system, scaled to the GPU's iteration count."""
    cn = prob.elem_codes()
    t0 = time.time()
    m = gpu.Matrix.from_elements(ctx, gpu.STORAGE_CR, prob.n, cn)
    lift = m.assemble(prob)
    rhs = prob.rhs[0] + lift
    adr, ci = m.pattern()
    m.set_values(skew_values(adr, ci, m.get_values(), 0.3, 11))
    nnz = int(adr[-1])
    log(f"[bicg] comprow matrix nnz={nnz} set up ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda:0")
    b = torch.tensor(rhs, dtype=torch.float64, device=dev)
    x = torch.zeros(prob.n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def solve():
        return m.solve_dev(b.data_ptr(), x.data_ptr(), ni=200000, res=args.res,
                           prec=gpu.PREC_NONE, linsolver=gpu.SOLVER_BICG)

    t0 = time.time()
    info = solve()  # builds the union layout of A and A^T, gathers both values
    torch.cuda.synchronize()
    log(f"[bicg] first solve incl. layout {time.time() - t0:.1f}s: status {info.status} "
        f"iterations {info.iterations}")
    ctx.set_profiling(True)
    t = time.perf_counter()
    infos = [solve() for _ in range(args.bicg_steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    us, samples = ctx.kernel_time(0)
    ctx.set_profiling(False)
    its = int(infos[-1].iterations)
    # algorithmic bytes of one dual product: a_ij and a_ji (8 + 8) per entry of
    # the (structurally symmetric) pattern, d and dd read once, p and pp written
    # once; no column bytes (the stencil form of the heat cube's union layout
    # streams none; the 16-bit column form adds 2 per entry)
    alg = 16 * nnz + 32 * prob.n
    gbs = alg / (us * 1e-6) / 1e9 if us else 0.0
    out = {"workload": f"heat hex8 {args.nel}^3 as comprow, nonsymmetric values "
                       f"(off-diagonals x (1 + 0.3 t)), n={prob.n}, nnz={nnz}, res {args.res}",
           "solves_per_s": args.bicg_steps / dt, "ms_per_solve": 1e3 * dt / args.bicg_steps,
           "iterations": its, "status": int(infos[-1].status),
           "rel_residual": float(infos[-1].rel_residual),
           "dual_spmv": {"kernel": "k_dual_spmv (A d and A^T dd, one SELL sweep)",
                         "usec_per_launch": us, "samples": samples,
                         "algorithmic_bytes_per_launch": alg, "achieved_gbs": gbs,
                         "frac_of_hbm_peak": gbs * 1e9 / PEAK_HBM}}
    log(f"[bicg] {out['solves_per_s']:.2f} solves/s, {its} iterations, dual spmv {us:.1f} us "
        f"{gbs:.0f} GB/s")
    if not args.no_cpu:
        import oracle as O
        a = m.get_values()
        k = args.cpu_iters
        t = time.perf_counter()
        O.cr_bicg(adr, ci, a, rhs, k, 0.0)
        per_it = (time.perf_counter() - t) / k
        out["cpu_baseline"] = {"value": 1.0 / (per_it * its), "unit": "solves/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{k} iterations of the oracle's comprow bicg (1 thread) "
                                         f"on the same system, scaled to the GPU's {its}"}
        out["speedup_vs_cpu"] = out["solves_per_s"] / out["cpu_baseline"]["value"]
    m.close()
    return out


def bench_ildl(ctx, args, gpu, mesh, torch, prob, m, rhs):
    """cg_prec with the incomplete LDL^T preconditioner (incomdec) on the heat cube's
    symcomprow matrix, FP64 from x = 0 to --res, against the Jacobi cg_prec on the
    same matrix. The factor and the sweeps run by dependency levels, one launch per
    level. CPU baseline: --cpu-iters iterations of the oracle's incomdec cg_prec on
    the GPU's factor, scaled to the GPU's iteration count; the factorization is
    left out of it (the oracle forms every dropped-fill product, as the reference
    does, which takes minutes at this size)."""
    t0 = time.time()
    nf, nb, us_apply = m.ildl_bench(reps=5)  # plan, factor and the apply graph
    setup_s = time.time() - t0
    _, _, us_direct = m.ildl_bench(reps=5, graph=False)
    dev = torch.device("cuda:0")
    b = torch.tensor(rhs, dtype=torch.float64, device=dev)
    x = torch.zeros(prob.n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def solve():
        return m.solve_dev(b.data_ptr(), x.data_ptr(), ni=200000, res=args.res,
                           prec=gpu.PREC_ILDL)

    solve()
    torch.cuda.synchronize()
    t = time.perf_counter()
    infos = [solve() for _ in range(args.ildl_steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    its = int(infos[-1].iterations)
    out = {"workload": f"heat hex8 {args.nel}^3 symcomprow, n={prob.n}, res {args.res}, "
                       f"incomplete LDL^T (weight 0) from x = 0",
           "solves_per_s": args.ildl_steps / dt, "ms_per_solve": 1e3 * dt / args.ildl_steps,
           "iterations": its, "status": int(infos[-1].status),
           "rel_residual": float(infos[-1].rel_residual),
           "levels": {"forward": nf, "backward": nb},
           "usec_per_application": us_apply,
           "usec_per_application_stream_launches": us_direct,
           "setup_s_levels_and_factor": setup_s}
    log(f"[ildl] {out['solves_per_s']:.3f} solves/s, {its} iterations, levels {nf}/{nb}, "
        f"apply {us_apply:.0f} us (graph) / {us_direct:.0f} us (stream), set-up {setup_s:.1f}s")
    if not args.no_cpu:
        import oracle as O
        adr, ci = m.pattern()
        a = m.get_values()
        inc = m.ildl_factor()
        k = args.cpu_iters
        t = time.perf_counter()
        O.scr_cg_prec_ildl(adr, ci, a, rhs, k, 0.0, incdec=inc)
        per_it = (time.perf_counter() - t) / k
        out["cpu_baseline"] = {"value": 1.0 / (per_it * its), "unit": "solves/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{k} iterations of the oracle's incomdec cg_prec "
                                         f"(1 thread) on the same system and factor, scaled to "
                                         f"the GPU's {its}; factorization not included"}
        out["speedup_vs_cpu"] = out["solves_per_s"] / out["cpu_baseline"]["value"]
    return out


def bench_renumber(ctx, args, gpu, mesh, prob):
    """Reverse Cuthill-McKee ordering of the heat cube's mesh (edge graph of its hex8
    elements) on the device. CPU baseline: the oracle's restatement of the
    reference's routine, whose search for the next parent scans every node per
    number, on a 40^3 cube."""
    nn = prob.coords.shape[0]
    t = time.perf_counter()
    o = gpu.cuthill_mckee(ctx, nn, prob.conn, reverse=True)
    dt = time.perf_counter() - t
    out = {"workload": f"reverse Cuthill-McKee of the heat hex8 {args.nel}^3 mesh, nn={nn}",
           "seconds": dt, "nodes_per_s": nn / dt, "first": int(o[0]), "last": int(o[-1])}
    log(f"[renumber] {nn} nodes in {dt * 1e3:.0f} ms")
    if not args.no_cpu:
        import oracle as O
        small = mesh.heat_cube(40, seed=2)
        ns = small.coords.shape[0]
        t = time.perf_counter()
        O.cuthill_mckee(ns, small.conn, True)
        dts = time.perf_counter() - t
        out["cpu_baseline"] = {"value": ns / dts, "unit": "nodes/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle's cuthill_mckee_renumb (nn^2 parent scan, "
                                         f"1 thread) on the 40^3 cube ({ns} nodes, {dts:.2f} s); "
                                         f"its cost grows as nn^2"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nel", type=int, default=128, help="elements per cube edge")
    ap.add_argument("--res", type=float, default=1e-22, help="squared relative residual test")
    ap.add_argument("--interval", type=int, default=128,
                    help="FP32 iterations between FP64 refinements (mixed)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--batch-res", type=float, default=1e-20)
    ap.add_argument("--batch-res-mixed", type=float, default=1e-22)
    ap.add_argument("--seq-stride", type=int, default=16)
    ap.add_argument("--only-batch", action="store_true")
    ap.add_argument("--tiling64", type=int, nargs="+", default=None,
                    help="pin the FP64 SpMV tiling (format lanes blocks idx16 [map]) instead of "
                         "autotuning")
    ap.add_argument("--tiling32", type=int, nargs="+", default=None,
                    help="pin the FP32 SpMV tiling instead of autotuning")
    ap.add_argument("--elast-tiling64", type=int, nargs="+", default=None,
                    help="pin the elasticity leg's FP64 SpMV tiling instead of autotuning")
    ap.add_argument("--elast-tiling32", type=int, nargs="+", default=None,
                    help="pin the elasticity leg's FP32 SpMV tiling instead of autotuning")
    ap.add_argument("--no-serve", action="store_true", help="skip the served leg")
    ap.add_argument("--only-serve", action="store_true")
    ap.add_argument("--only-assembly", action="store_true")
    ap.add_argument("--asm-nel", type=int, default=64)
    ap.add_argument("--asm-reps", type=int, default=3)
    ap.add_argument("--asm-heat-nel", type=int, default=128)
    ap.add_argument("--post-reps", type=int, default=10)
    ap.add_argument("--no-post", action="store_true", help="skip the post-solve element pass")
    ap.add_argument("--no-np", action="store_true", help="skip the nonstationary heat leg")
    ap.add_argument("--np-steps", type=int, default=10)
    ap.add_argument("--np-dt", type=float, default=1e-3)
    ap.add_argument("--no-bicg", action="store_true", help="skip the comprow bicg leg")
    ap.add_argument("--no-ildl", action="store_true", help="skip the incomplete LDL^T leg")
    ap.add_argument("--no-renumber", action="store_true", help="skip the Cuthill-McKee leg")
    ap.add_argument("--no-elast", action="store_true", help="skip the elasticity column leg")
    ap.add_argument("--elast-nel", type=int, default=64,
                    help="elasticity column nel x nel x 4 nel elements")
    ap.add_argument("--elast-oracle-nel", type=int, default=16,
                    help="column checked against the oracle: nel x nel x 4 nel")
    ap.add_argument("--ildl-steps", type=int, default=1)
    ap.add_argument("--bicg-steps", type=int, default=2)
    ap.add_argument("--serve-count", type=int, default=24)
    ap.add_argument("--serve-res", type=float, default=1e-22)
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[50.0])
    ap.add_argument("--serve-precisions", nargs="+", default=["fp64", "mixed"])
    ap.add_argument("--serve-chunk", type=int, default=32)
    ap.add_argument("--serve-prep", type=int, default=4,
                    help="server prep threads (pattern, layout and assembly of new requests)")
    ap.add_argument("--serve-oracle", type=int, default=3,
                    help="smallest requests whose drop-in solution is checked by the oracle")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; the path shards by independent "
              "systems (see the design notes) but no multi-GPU figure is produced", flush=True)
        sys.exit(2)

    import torch
    from ssf_amd import gpu, mesh

    if args.only_batch:
        ctx = gpu.Context(0)
        print(json.dumps(bench_batched(ctx, args, gpu, mesh, torch)), flush=True)
        return
    if args.only_serve:
        ctx = gpu.Context(0)
        print(json.dumps(bench_served(ctx, args, gpu, mesh, torch)), flush=True)
        return
    if args.only_assembly:
        ctx = gpu.Context(0)
        print(json.dumps(bench_assembly(ctx, args, gpu, mesh, torch)), flush=True)
        return
    t0 = time.time()
    prob = mesh.heat_cube(args.nel, seed=2)
    log(f"[bench] problem {prob.name}: n={prob.n} ne={prob.ne} ({time.time() - t0:.1f}s)")
    # the reference's FP64 CPU solve of the same system (the oracle: its own
    # assembly and cg_prec to 1e-30), beside the GPU legs
    oref = None
    if not args.no_cpu:
        oref = OracleSolve(prob, 1e-30)
        oref.start()
    ctx = gpu.Context(0)
    t0 = time.time()
    m = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, prob.elem_codes())
    log(f"[bench] pattern nnz_full={m.nnz} negm={m.negm} ({time.time() - t0:.1f}s)")
    t0 = time.time()
    lift = m.assemble(prob)
    rhs = prob.rhs[0] + lift
    log(f"[bench] assembled ({time.time() - t0:.1f}s)")
    if args.tiling64:
        m.set_tune(gpu.FP64, args.tiling64)
        tune64 = tuple(args.tiling64)
    else:
        tune64 = m.autotune(gpu.FP64)
    if args.tiling32:
        m.set_tune(gpu.MIXED, args.tiling32)
        tune32 = tuple(args.tiling32)
    else:
        tune32 = m.autotune(gpu.MIXED)
    log(f"[bench] spmv tiling fp64 {tune64} fp32 {tune32}")

    dev = torch.device("cuda:0")
    b = torch.tensor(rhs, dtype=torch.float64, device=dev)
    x = torch.zeros(prob.n, dtype=torch.float64, device=dev)
    xref = torch.zeros(prob.n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    info_ref = m.solve_dev(b.data_ptr(), xref.data_ptr(), ni=200000, res=1e-30)
    torch.cuda.synchronize()
    log(f"[bench] x_ref: status {info_ref.status} ani {info_ref.ani} "
        f"rel_residual {info_ref.rel_residual:.3e}")
    xref_norm = torch.linalg.norm(xref).item()

    def rel_err():
        return torch.linalg.norm(x - xref).item() / xref_norm

    def fp64_solve():
        return m.solve_dev(b.data_ptr(), x.data_ptr(), ni=200000, res=args.res)

    def mixed_solve():
        return m.solve_dev(b.data_ptr(), x.data_ptr(), ni=200000, res=args.res,
                           precision=gpu.MIXED, inner_ni=args.interval)

    results = {}
    for name, fn, which in (("fp64", fp64_solve, 0), ("mixed", mixed_solve, 1)):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ctx.set_profiling(True)
        t = time.perf_counter()
        infos = [fn() for _ in range(args.steps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        spmv_us, samples = ctx.kernel_time(which)
        ctx.set_profiling(False)
        err = rel_err()
        results[name] = {"solves_per_s": args.steps / dt, "ms_per_solve": 1e3 * dt / args.steps,
                         "rel_err": err, "iterations": int(infos[-1].iterations),
                         "refinements": int(infos[-1].outer), "status": int(infos[-1].status),
                         "rel_residual": float(infos[-1].rel_residual),
                         "spmv_us_live": spmv_us, "spmv_samples": samples}
        log(f"[bench] {name}: {results[name]}")

    # bytes per SpMV launch (canonical CSR count and the format actually streamed)
    b32_can, b32_act = m.spmv_bytes(gpu.MIXED)
    b64_can, b64_act = m.spmv_bytes(gpu.FP64)
    if tune32[0] == 2 and tune32[3] == 4:
        # the half form's FP32 product in the mixed solve also reads the residual
        # (the four dots of the fused step): 4 bytes per row more
        b32_act += 4.0 * prob.n
    us32 = results["mixed"]["spmv_us_live"]
    us64 = results["fp64"]["spmv_us_live"]
    gbs32 = b32_act / (us32 * 1e-6) / 1e9 if us32 else 0.0
    gbs64 = b64_act / (us64 * 1e-6) / 1e9 if us64 else 0.0
    copy_bytes = 1 << 30
    copy_us = ctx.copy_bench(copy_bytes, reps=20)
    copy_gbs = 2 * copy_bytes / (copy_us * 1e-6) / 1e9
    log(f"[bench] spmv fp32 {us32:.1f} us {gbs32:.0f} GB/s; fp64 {us64:.1f} us {gbs64:.0f} GB/s; "
        f"copy {copy_gbs:.0f} GB/s")

    cpu = None
    oracle_ref = None
    if oref is not None:
        # the timed solutions against the oracle's; the solutions of the last
        # timed solves are kept
        xs = {}
        for name, fn in (("fp64", fp64_solve), ("mixed", mixed_solve)):
            fn()
            torch.cuda.synchronize()
            xs[name] = x.cpu().numpy()
        t = time.perf_counter()
        oref.join()
        waited = time.perf_counter() - t
        if oref.err:
            log(f"[bench] oracle failed: {oref.err}")
        else:
            xo = oref.x
            no = np.linalg.norm(xo)
            oracle_ref = oref.summary()
            oracle_ref["waited_s"] = waited
            for name in ("fp64", "mixed"):
                results[name]["rel_err_vs_gpu_ref"] = results[name]["rel_err"]
                results[name]["rel_err"] = float(np.linalg.norm(xs[name] - xo) / no)
            oracle_ref["gpu_ref_rel_err"] = float(
                np.linalg.norm(xref.cpu().numpy() - xo) / no)
            its = results["fp64"]["iterations"]
            per_it = oref.t_solve / (oref.ani + 1)
            ci = cpu_info()
            cpu = {"value": 1.0 / (per_it * its), "unit": "solves/s", "cores": 1, "kind": "port",
                   "cpu_model": ci["model"], "cpus_available": ci["available_to_process"],
                   "sample": f"the oracle's full cg_prec solve (scr, Jacobi, 1 thread) of the same "
                             f"{prob.n}-unknown system to 1e-30 ({oref.ani + 1} iterations, "
                             f"{oref.t_solve:.1f} s, run beside the GPU legs), per-iteration time "
                             f"scaled to the {its} iterations of the FP64 recurrence at {args.res}"}
            cpu["parallel"] = cpu_parallel(oref.adr, oref.ci, oref.a, oref.y, its)
            log(f"[bench] oracle x: {oracle_ref}; errors fp64 {results['fp64']['rel_err']:.2e} "
                f"mixed {results['mixed']['rel_err']:.2e}")
            log(f"[bench] cpu: {per_it * 1e3:.1f} ms/iteration -> {cpu['value']:.4f} solves/s; "
                f"{cpu['parallel']['cores']} threads: {cpu['parallel']['value']:.4f} solves/s")

    # HBM traffic per launch measured by rocprofv3 counters (calibrated on the copy
    # kernel) for this workload and tiling, when such a measurement is committed
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "spmv_traffic_heat128.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        t = tr["fp32"]["tiling"]
        if (tr["n"] == prob.n and tr["nnz"] == m.nnz and t[0] == tune32[0] and t[3] == tune32[3]
                and (t[0] == 2 or t[1] == tune32[1])):
            traffic = tr["fp32"]["bytes_per_launch"]

    post = None
    if not args.no_post:
        post = bench_post(ctx, args, gpu, mesh, torch, prob, x, f"heat hex8 {args.nel}^3")
    transient = None
    if not args.no_np:
        transient = bench_transient(ctx, args, gpu, mesh, torch, prob, m, rhs)
    renumber = None
    if not args.no_renumber:
        renumber = bench_renumber(ctx, args, gpu, mesh, prob)
    ildl = None
    if not args.no_ildl:
        ildl = bench_ildl(ctx, args, gpu, mesh, torch, prob, m, rhs)
    bicg = None
    if not args.no_bicg:
        bicg = bench_bicg(ctx, args, gpu, mesh, torch, prob)
    elast = None
    if not args.no_elast:
        elast = bench_elastic(ctx, args, gpu, mesh, torch)

    batched = None
    if not args.no_batch:
        m.close()
        del b, x, xref
        torch.cuda.empty_cache()
        batched = bench_batched(ctx, args, gpu, mesh, torch)
    served = None
    if not args.no_serve:
        served = bench_served(ctx, args, gpu, mesh, torch)

    mixed = results["mixed"]
    fp64 = results["fp64"]
    line = {
        "metric": "FE solves/s (mixed-precision PCG at <=1e-8 rel err) vs FP64 GPU PCG; SpMV GB/s",
        "value": mixed["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp32-inner/fp64-refinement",
        "data": "synthetic (seeded lognormal conductivity field, generated mesh)",
        "config": {"workload": f"heat hex8 {args.nel}^3, n={prob.n}, nnz={m.nnz}",
                   "res": args.res, "refinement_interval": args.interval,
                   "spmv_tiling": {"fp64": tune64, "fp32": tune32}},
        "fp64_solves_per_s": fp64["solves_per_s"],
        "speedup_vs_fp64": gated(mixed["solves_per_s"] / fp64["solves_per_s"],
                                 mixed["rel_err"], fp64["rel_err"]),
        "meets_1e-8": {"mixed": mixed["rel_err"] <= TOL, "fp64": fp64["rel_err"] <= TOL},
        "rel_err": {"mixed": mixed["rel_err"], "fp64": fp64["rel_err"],
                    "reference": ("the oracle's FP64 CPU cg_prec of the same system to 1e-30"
                                  if oracle_ref else "the GPU FP64 recurrence to 1e-30"),
                    "vs_gpu_fp64_ref": {"mixed": mixed.get("rel_err_vs_gpu_ref"),
                                        "fp64": fp64.get("rel_err_vs_gpu_ref")}},
        "oracle_reference_solve": oracle_ref,
        "iterations": {"mixed_fp32": mixed["iterations"], "mixed_refinements": mixed["refinements"],
                       "fp64": fp64["iterations"]},
        "spmv_gbs": {"fp32": gbs32, "fp64": gbs64, "copy_kernel": copy_gbs},
        "roofline": {"bound": "hbm", "achieved": gbs32, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": gbs32 * 1e9 / PEAK_HBM, "traffic": traffic,
                     "kernel": {2: ("k_spmv_half<float, 3, %d> (FP32 SpMV of the mixed solve, "
                                    "symmetric half of the stencil form: each row's lower slots "
                                    "stored once, the upper entries read from the mirrored rows; "
                                    "the residual read for the fused step's dots)"
                                    % (round((b32_act - 12 * prob.n) / (4 * prob.n)) - 1)
                                    if tune32[3] == 4 else
                                    "k_spmv_sell_sten<float, 1, %d> (FP32 SELL-64 SpMV of the mixed "
                                    "solve, stencil form)" % round((b32_act - 8 * prob.n) / 4 / (64 * ((prob.n + 63) // 64)) / 4)
                                    if tune32[3] == 3 else
                                    "k_spmv_sell_pat<float, 1, %d> (FP32 SELL-64 SpMV of the mixed "
                                    "solve, row patterns)" % (8 if tune32[1] >= 8 else 4 if tune32[1] >= 4 else 2)
                                    if tune32[3] == 2 else
                                    "k_spmv_sell<float, %d, 1> (FP32 SELL-64 SpMV of the mixed solve, %s)"
                                    % (tune32[3], "16-bit columns" if tune32[3] else "32-bit columns")),
                                1: "k_spmv_stream<float> (FP32 stream SpMV of the mixed solve)"}.get(
                                    tune32[0], "k_spmv_vec<float> (FP32 SpMV of the mixed solve)"),
                     "bytes_per_launch": b32_act, "bytes_per_launch_canonical_csr": b32_can,
                     "usec_per_launch": us32, "samples": mixed["spmv_samples"]},
        "cpu_baseline": cpu,
        "post": post,
        "transient": transient,
        "renumber": renumber,
        "elasticity": elast,
        "ildl": ildl,
        "bicg": bicg,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
