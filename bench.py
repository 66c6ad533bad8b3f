"""H-matvec benchmark on one MI355X.

A step is one H-matvec y = H x in the original DoF ordering
(BigWhamIO::MatVec, src/io/bigwham_io.cc:919-941) over one synthetic input
vector already resident in HBM.  The timed region is `--steps` FP64 matvecs
bracketed by torch.cuda.synchronize(); `value` is matvecs per second.

Alongside the headline line it reports:
  roofline      the dominant kernel's algorithmic bytes / its average HIP-event
                duration over the timed steps, against 8 TB/s;
  fp32          the same for FP32 factors with FP32 accumulation;
  gmres         time to ||b - H64 x|| / ||b|| <= 1e-8 for the FP64 path and the
                mixed-precision path (on by default; --no-gmres skips it);
  cpu_baseline  the CPU oracle's OpenMP matvec (restatement of Hmat::matvec,
                hmat.cc:472-537) on this host's cores, on the same H.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))  # cpu_baseline leg only

HBM_PEAK_GBS = 8000.0
# FP32 copy: low-rank blocks recompressed to this relative accuracy (the FP64
# copy is the host's blocks unchanged)
FP32_RANK_TOL = 1e-7

CONFIGS = {
    "c1": dict(kernel="2DS0-H", mesh=("line", 2000), leaf=32, eta=3.0, eps=1e-6,
               workload="2DS0-H straight crack, 2000 segments (4000 DoF), leaf 32, eta 3, ACA eps 1e-6"),
    "c2": dict(kernel="3DR0-H", mesh=("grid", 200), leaf=64, eta=3.0, eps=1e-6,
               workload="3DR0-H square crack 200x200 rectangles (120000 DoF), leaf 64, eta 3, ACA eps 1e-6"),
    "c3": dict(kernel="3DT0-H", mesh=("penny", 105000), leaf=64, eta=3.0, eps=1e-6,
               workload="3DT0-H penny crack r=1, ~200k Delaunay triangles (~600k DoF), leaf 64, eta 3, ACA eps 1e-6"),
    "c4": dict(kernel="3DR0-H", mesh=("grid", 224), leaf=64, eta=3.0, eps=1e-6, nrhs=32,
               n_meshes=16, mesh_elems=(1000, 40000),
               workload="3DR0-H square crack 224x224 rectangles (150528 DoF) x 32 right-hand "
                        "sides on FP32 MFMA, plus 16 square cracks of 1k-40k rectangles "
                        "(log-uniform, seed 7) as one operator; leaf 64, eta 3, ACA eps 1e-6"),
    "c5": dict(kernel="3DR0-H-mode1", mesh=("grid", 500), leaf=64, eta=3.0, eps=1e-6,
               workload="3DR0-H-mode1 500x500 rectangles (250000 DoF), leaf 64, eta 3, ACA eps 1e-6"),
}


# HBM bytes per launch of the slab kernels, measured with rocprofv3 PMC counters
# (FETCH_SIZE and WRITE_SIZE, scaled by a copy of known size) in separate runs;
# the counters cannot be read in the timed run itself.
TRAFFIC_FILES = {"c2": "profiles/hbm_traffic_3dr0_120k_dof.md",
                 "c3": "profiles/hbm_traffic_3dt0_600k_dof.md"}


def load_traffic(config):
    """({(precision, pass): HBM bytes per launch}, FP32 rank tolerance of the
    measured copy) from the PMC summary, if any."""
    import re
    path = TRAFFIC_FILES.get(config)
    if not path or not (ROOT / path).exists():
        return {}, None
    m = re.search(r"```json\n(.*?)```", (ROOT / path).read_text(), re.S)
    if not m:
        return {}, None
    doc = json.loads(m.group(1))
    out = {}
    for name, v in doc["traffic_bytes"].items():
        k = re.match(r"k_slab<(\w+), \w+, \d+, (\d)>", name)
        if k:
            out[("fp64" if k.group(1) == "double" else "fp32", int(k.group(2)))] = int(v)
    return out, doc.get("fp32_rank_tolerance", 0.0)


def load_call_traffic(path, tol32, precision="fp32"):
    """HBM bytes per batched call of one precision from a PMC summary (None if
    absent or, for FP32, measured on a copy of another recompression
    tolerance)."""
    import re
    if not (ROOT / path).exists():
        return None
    m = re.search(r"```json\n(.*?)```", (ROOT / path).read_text(), re.S)
    if not m:
        return None
    doc = json.loads(m.group(1))
    if precision == "fp32" and doc.get("fp32_rank_tolerance", 0.0) != tol32:
        return None
    return doc.get("traffic_bytes_per_call", {}).get(precision)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_mesh(kind, n):
    from bigwham_dev_amd import meshes
    if kind == "line":
        return meshes.segment_line(n)
    if kind == "grid":
        return meshes.rect_grid(n)
    if kind == "penny":
        return meshes.penny_delaunay(n, seed=0)
    raise ValueError(kind)


def split_entries(host):
    """Entries of the F blocks, the A factors and the B factors; total rank T."""
    fp, lp = host.fr_pattern(), host.lr_pattern()
    d = host.dof
    e_fr = int(np.sum(d * (fp[3] - fp[1]) * d * (fp[4] - fp[2])))
    e_a = int(np.sum(d * (lp[3] - lp[1]) * lp[5]))
    e_b = int(np.sum(lp[5] * d * (lp[4] - lp[2])))
    return e_fr, e_a, e_b, int(np.sum(lp[5]))


def cpu_baseline(host, x, budget_s=15.0):
    """Time the oracle's OpenMP matvec (1 warm-up + up to 10 timed calls,
    mat_vec.cc:158-168 protocol) within a time budget, built here for this
    host's CPU (gcc -O3 -march=native -fopenmp).  Returns the result of the
    default oracle build (the parity reference) and the timing."""
    import tempfile
    import oracle as O
    import build_oracle
    flat = host.flat()
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
    native = O.native_lib(tempfile.mkdtemp(prefix="bw_oracle_"))
    y = O.matvec_original(flat, x, threads)
    t0 = time.perf_counter()
    y_native = O.matvec_original(flat, x, threads, use=native)
    t_warm = time.perf_counter() - t0
    reps = int(max(1, min(10, budget_s // max(t_warm, 1e-6))))
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.matvec_original(flat, x, threads, use=native)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    agree = float(np.linalg.norm(y_native - y) / np.linalg.norm(y))
    flags = " ".join(build_oracle.NATIVE_FLAGS)

    def host_matvec(v):
        return O.matvec_original(flat, v, threads, use=native)
    host_matvec.flat = flat  # keeps the flat view's buffers alive with the closure
    return y, host_matvec, dict(value=1.0 / med, unit="H-matvec/s", cores=threads, kind="port",
                   sample=f"{reps} timed calls (+1 warm-up) of the full matvec, median "
                          f"{med * 1e3:.2f} ms, OpenMP restatement of Hmat::matvecOriginal "
                          f"built with gcc {flags} on this host ({threads} threads); "
                          f"rel. diff to the reference build {agree:.1e}")


def host_gmres(host_matvec, diag, b, tol, restart):
    """The host path's solve, as a user of the reference runs it: scipy's GMRES
    on a LinearOperator whose matvec is the CPU baseline's OpenMP matvec, with
    the Jacobi preconditioner 1/diag(H) (bigwham4py.py:373-380), restart 30.
    Returns the wall time, the matvec count and the true FP64 residual."""
    from scipy.sparse import diags
    from scipy.sparse.linalg import LinearOperator, gmres as sp_gmres
    n = b.size
    count = [0]

    def mv(v):
        count[0] += 1
        return host_matvec(np.ascontiguousarray(v, dtype=np.float64).ravel())
    A = LinearOperator((n, n), matvec=mv, dtype=np.float64)
    M = diags(1.0 / diag)
    t0 = time.perf_counter()
    x, flag = sp_gmres(A, b, rtol=tol, atol=0.0, restart=restart, maxiter=100, M=M)
    dt = time.perf_counter() - t0
    nmv = count[0]
    rel = float(np.linalg.norm(b - host_matvec(x)) / np.linalg.norm(b))
    log(f"host GMRES: {dt:.1f} s, {nmv} matvecs, scipy flag {flag}, rel. residual {rel:.2e}")
    return {"seconds": round(dt, 2), "matvecs": nmv, "scipy_info": int(flag),
            "rel_residual_fp64": rel, "solver": "scipy.sparse.linalg.gmres, restart 30, "
            "Jacobi 1/diag, OpenMP restatement of Hmat::matvecOriginal"}


def dense_check(host, dev, x, max_dense=12000, blocks=8, m=8):
    """Matvec error against the exact dense operator (toArray2D,
    matrix_generator.h:35-40, evaluated by the host kernels): the whole
    operator when N <= max_dense, else `blocks` random blocks of m
    consecutive points' rows (permuted numbering).  The gap measures the ACA
    approximation, not the arithmetic.  Returns (summary, dense or None)."""
    import torch
    n, dof = host.size, host.dof
    dev0 = "cuda:0"
    if n <= max_dense:
        D = host.dense(original=True)
        y_ex = D @ x
        xd = torch.from_numpy(x).to(dev0)
        got = {p: dev.matvec(xd, original=True, precision=p).cpu().numpy()
               for p in dev.precisions}
        rows = n
    else:
        D = None
        perm = host.permutation()
        idx = (dof * perm[:, None] + np.arange(dof)[None, :]).ravel()
        xp = x[idx]  # permuted numbering: z = x gathered through permutation_1_
        xd = torch.from_numpy(xp).to(dev0)
        yp = {p: dev.matvec(xd, original=False, precision=p).cpu().numpy()
              for p in dev.precisions}
        npts = n // dof
        starts = np.sort(np.random.default_rng(99).choice(npts - m, blocks, replace=False))
        sl = np.concatenate([np.arange(dof * b, dof * (b + m)) for b in starts])
        y_ex = np.concatenate([host.dense_block(int(b), 0, m, npts) @ xp for b in starts])
        got = {p: yp[p][sl] for p in yp}
        rows = int(sl.size)
    nrm = np.linalg.norm(y_ex)
    out = {"rows": rows, "of": n, "sampled": D is None}
    for p in got:
        out[f"{p}_rel_err"] = float(np.linalg.norm(got[p] - y_ex) / nrm)
    return out, D


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default=os.environ.get("BW_BENCH_CONFIG", "c3"))
    ap.add_argument("--gmres", action="store_true", default=True,
                    help="also time GMRES to 1e-8, FP64 and mixed (default)")
    ap.add_argument("--no-gmres", dest="gmres", action="store_false")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--wi-bytes", type=int, default=0)
    ap.add_argument("--fp32-rank-tol", type=float, default=FP32_RANK_TOL,
                    help="recompress the FP32 copy's low-rank blocks to this relative "
                         "accuracy (0: the FP64 factors rounded to FP32)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates HBM counters)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default unroll depth instead of timing 4/8/16 first")
    ap.add_argument("--host-gmres", action="store_true",
                    help="with --gmres: also run the host path's GMRES to 1e-8 (scipy GMRES on "
                         "the CPU baseline's OpenMP matvec, Jacobi; about 200 host matvecs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: bigwham_dev_amd runs on one GPU only; multi-GPU sharding is a design "
              "note (block-row ownership + RCCL all-gather of x), not a measured path.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        n = 1 << 27  # 1 GiB of float64 read and 1 GiB written per copy
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"probe": "copy", "bytes_read": n * 8, "bytes_written": n * 8,
                          "launches": 5}))
        return
    from bigwham_dev_amd import HipHmat, HostHmat, set_tuning
    from bigwham_dev_amd import solvers

    cfg = CONFIGS[args.config]
    if args.config == "c4":
        return bench_batching(args, cfg)
    if args.config == "c5":
        return bench_adjoint(args, cfg)
    coor, conn = make_mesh(*cfg["mesh"])
    t0 = time.perf_counter()
    host = HostHmat.from_mesh(cfg["kernel"], coor, conn, [1.0, 0.25])
    host.build(cfg["leaf"], cfg["eta"], cfg["eps"])
    t_build = time.perf_counter() - t0
    n = host.size
    E = host.nb_entries()
    e_fr, e_a, e_b, T = split_entries(host)
    log(f"config {args.config}: N={n} E={E} (FR {e_fr}, A {e_a}, B {e_b}) build {t_build:.1f}s")

    if args.wi_bytes:
        set_tuning(args.wi_bytes, args.wi_bytes)
    t0 = time.perf_counter()
    dev = HipHmat(host, precisions=("fp64", "fp32"), device=0, fp32_rank_tol=args.fp32_rank_tol)
    torch.cuda.synchronize()
    t_upload = time.perf_counter() - t0
    ent = dev.entries
    log(f"layout + upload {t_upload:.1f}s, device bytes {dev.device_bytes / 1e9:.2f} GB, "
        f"entries fp64 {sum(ent['fp64'][:3])} fp32 {sum(ent['fp32'][:3])}")
    tuning = None
    if not args.no_autotune:
        from bigwham_dev_amd import autotune
        t0 = time.perf_counter()
        tr = autotune.tune_unroll(dev, use_cache=False)
        tuning = {"unroll": tr.unroll, "ms": tr.ms, "seconds": round(time.perf_counter() - t0, 2)}
        log(f"autotune {tuning}")

    rng = np.random.default_rng(1234)
    x_np = rng.uniform(-1.0, 1.0, n)
    x = torch.from_numpy(x_np).to("cuda:0")
    y = torch.empty(n, dtype=torch.float64, device="cuda:0")

    def run(precision, steps, warmup):
        for _ in range(warmup):
            dev.matvec(x, original=True, precision=precision, out=y)
        dev.set_timing(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            dev.matvec(x, original=True, precision=precision, out=y)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        stats = dev.timing_stats()
        dev.set_timing(False)
        return dt, stats

    dt64, st64 = run("fp64", args.steps, args.warmup)
    y64 = y.cpu().numpy()
    dt32, st32 = run("fp32", args.steps, args.warmup)
    y32 = y.cpu().numpy()

    # the same matvecs replayed from one captured HIP graph each (what a
    # solver loop on fixed vectors can use): removes the per-call host and
    # launch overhead, which matters for the small matrices
    graph_ms = {}
    for prec in ("fp64", "fp32"):
        g = dev.graph(x, y, precision=prec)
        for _ in range(args.warmup):
            g.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.launch()
        torch.cuda.synchronize()
        graph_ms[prec] = round((time.perf_counter() - t0) / args.steps * 1e3, 4)
        g.close()
    log(f"fp64 {dt64 / args.steps * 1e3:.3f} ms/matvec, fp32 {dt32 / args.steps * 1e3:.3f} ms/matvec")
    log(f"fp64 kernels {st64}")
    log(f"fp32 kernels {st32}")

    traffic, traffic_tol32 = load_traffic(args.config)

    def roofline(stats, sm, sv, prec):
        # algorithmic bytes per launch of each slab kernel, from the entries
        # this precision's copy streams
        pe_fr, pe_a, pe_b, pT = ent[prec]
        b1 = sm * pe_b + sv * (n + pT)
        b2 = sm * (pe_fr + pe_a) + sv * (n + pT) + 8 * n  # y is written in float64
        k1, k2 = stats["p1_slab"], stats["p2_slab"]
        dom, b, ms = ("k_slab pass 2", b2, k2) if k2 >= k1 else ("k_slab pass 1", b1, k1)
        tr = traffic.get((prec, int(dom[-1])))
        if prec == "fp32" and traffic_tol32 != args.fp32_rank_tol:
            tr = None  # the committed PMC figures are of another FP32 copy
        ach = b / (ms * 1e-3) / 1e9
        pE = pe_fr + pe_a + pe_b
        whole = (sm * pE + sv * 4 * n + 2 * 8 * n / host.dof) / (stats["total"] * 1e-3) / 1e9
        return dict(bound="hbm", kernel=dom, achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=tr,
                    traffic_source=TRAFFIC_FILES.get(args.config) if tr else None,
                    avg_kernel_ms=round(ms, 4), alg_bytes=int(b),
                    whole_matvec_GBs=round(whole, 1),
                    whole_matvec_frac=round(whole / HBM_PEAK_GBS, 4))

    rf64 = roofline(st64, 8, 8, "fp64")
    rf32 = roofline(st32, 4, 4, "fp32")

    out = {
        "metric": "H-matvec/s",
        "value": round(args.steps / dt64, 3),
        "unit": "matvec/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt64 / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "n_dof": n, "nb_entries": E,
                   "compression_ratio": round(E / float(n) / n, 6),
                   "ordering": "original", "inputs": "x ~ U[-1,1] seed 1234, resident in HBM"},
        "roofline": rf64,
        "fp32": {"value": round(args.steps / dt32, 3), "unit": "matvec/s",
                 "ms_per_step": round(dt32 / args.steps * 1e3, 4), "roofline": rf32,
                 "rank_tolerance": args.fp32_rank_tol,
                 "nb_entries": int(sum(ent["fp32"][:3])),
                 "entries_fr_a_b_rank": [int(v) for v in ent["fp32"]],
                 "entries_vs_fp64": round(sum(ent["fp32"][:3]) / E, 4)},
        "graph_replay_ms_per_step": graph_ms,
        "setup_s": {"host_build": round(t_build, 2), "layout_upload": round(t_upload, 2)},
        "tuning": tuning,
    }

    if not args.no_cpu:
        y_ref, host_matvec, cb = cpu_baseline(host, x_np)
        out["cpu_baseline"] = cb
        out["parity"] = {
            "fp64_rel_err_vs_oracle": float(np.linalg.norm(y64 - y_ref) / np.linalg.norm(y_ref)),
            "fp32_rel_err_vs_oracle": float(np.linalg.norm(y32 - y_ref) / np.linalg.norm(y_ref)),
        }
        out["vs_cpu"] = round(out["value"] / cb["value"], 2)

    t0 = time.perf_counter()
    out["vs_dense_operator"], D = dense_check(host, dev, x_np)
    out["vs_dense_operator"]["seconds"] = round(time.perf_counter() - t0, 2)
    log(f"vs dense operator {out['vs_dense_operator']}")

    if args.gmres:
        b = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        b[host.dof - 1::host.dof] = 1.0  # unit normal pressure, local frame (normal last)
        solvers.warm_up(dev, b)
        x64, info64 = solvers.solve_fp64(dev, b, tol=1e-8)
        xm, infom = solvers.solve_mixed(dev, b, tol=1e-8)
        out["gmres"] = {
            "tol": 1e-8, "restart": 30, "precond": "jacobi",
            "fp64": {"seconds": round(info64.seconds, 4), "iterations": info64.iterations,
                     "rel_residual_fp64": info64.rel_residual},
            "mixed": {"seconds": round(infom.seconds, 4), "iterations": infom.iterations,
                      "fp32_matvecs": infom.matvecs_fp32, "fp64_matvecs": infom.matvecs_fp64,
                      "rel_residual_fp64": infom.rel_residual},
            "speedup_mixed_vs_fp64": round(info64.seconds / infom.seconds, 3),
            "matvec_time_ratio_fp64_fp32": round(dt64 / dt32, 4),
            "solution_rel_diff": float(torch.linalg.vector_norm(xm - x64) /
                                       torch.linalg.vector_norm(x64)),
        }
        if D is not None:  # small N: the dense FP64 solve of the exact operator
            xs = np.linalg.solve(D, b.cpu().numpy())
            for key, sol in (("fp64", x64), ("mixed", xm)):
                out["gmres"][key]["rel_err_vs_dense_solve"] = float(
                    np.linalg.norm(sol.cpu().numpy() - xs) / np.linalg.norm(xs))
        if "cpu_baseline" in out:
            # the host path would run the same FP64 GMRES (scipy on BEMatrix) with
            # one OpenMP matvec per iteration: bounded by its matvec count times
            # the measured host matvec time (Krylov work on the host not counted)
            t_host = info64.matvecs_fp64 / out["cpu_baseline"]["value"]
            out["gmres"]["host_reference_estimate"] = {
                "seconds_lower_bound": round(t_host, 2),
                "basis": f"{info64.matvecs_fp64} FP64 matvecs x host matvec time",
                "speedup_mixed_vs_host": round(t_host / infom.seconds, 1)}
            if args.host_gmres:
                out["gmres"]["host_reference"] = host_gmres(
                    host_matvec, dev.diagonal(original=True), b.cpu().numpy(), 1e-8, 30)
                out["gmres"]["host_reference"]["speedup_mixed_vs_host"] = round(
                    out["gmres"]["host_reference"]["seconds"] / infom.seconds, 1)
    print(json.dumps(out), flush=True)


def _build_host(kernel, coor, conn, cfg):
    from bigwham_dev_amd import HostHmat
    h = HostHmat.from_mesh(kernel, coor, conn, [1.0, 0.25])
    h.build(cfg["leaf"], cfg["eta"], cfg["eps"])
    return h


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def dense_gemm_reference(E64, E32, k, reps=10):
    """The vendor BLAS (torch.mm on ROCm, TF32 off) on a dense GEMM with the
    same bytes and flops as the 32-RHS call (E64 FP64 entries, or the E32
    entries of the FP32 copy): A (M x K, M*K = E) times
    B (K x k), for two shapes of A in both storage orders.  A reference point
    for what a matrix-core GEMM of this arithmetic intensity reaches on the
    chip, not a bound."""
    import torch
    torch.backends.cuda.matmul.allow_tf32 = False
    out = {}
    for dt, name, E in ((torch.float32, "fp32", E32), (torch.float64, "fp64", E64)):
        runs = {}
        for M in (32768, 150528):
            K = E // M
            for colmajor in (False, True):
                g = torch.Generator(device="cuda:0")
                g.manual_seed(1)
                A = torch.rand((K, M) if colmajor else (M, K), device="cuda:0", dtype=dt,
                               generator=g) - 0.5
                B = torch.rand((K, k), device="cuda:0", dtype=dt, generator=g) - 0.5
                Am = A.t() if colmajor else A
                for _ in range(3):
                    torch.mm(Am, B)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    torch.mm(Am, B)
                e1.record()
                torch.cuda.synchronize()
                runs[f"M{M}_{'colmajor' if colmajor else 'rowmajor'}"] = round(
                    e0.elapsed_time(e1) / reps, 4)
                del A, B, Am
                torch.cuda.empty_cache()
        best = min(runs.values())
        out[name] = {"ms": runs, "best_ms": best,
                     "best_tflops": round(2.0 * k * E / (best * 1e-3) / 1e12, 1)}
    return out


def bench_batching(args, cfg):
    """Many right-hand sides (FP32 MFMA thin GEMM) and many meshes in one
    operator.  value = right-hand sides per second of the 32-RHS FP32 call."""
    import torch
    from bigwham_dev_amd import HipHmat
    from bigwham_dev_amd import meshes

    coor, conn = make_mesh(*cfg["mesh"])
    t0 = time.perf_counter()
    host = _build_host(cfg["kernel"], coor, conn, cfg)
    t_build = time.perf_counter() - t0
    n, E, k = host.size, host.nb_entries(), cfg["nrhs"]
    log(f"c4 batch matrix: N={n} E={E} build {t_build:.1f}s")
    # the FP32 copy as in the c3 line: low-rank blocks recompressed to FP32 accuracy
    tol32 = args.fp32_rank_tol
    dev = HipHmat(host, precisions=("fp64", "fp32"), device=0, fp32_rank_tol=tol32)
    E32 = int(sum(dev.entries["fp32"][:3]))
    rng = np.random.default_rng(1234)
    X = torch.from_numpy(rng.uniform(-1.0, 1.0, (n, k))).to("cuda:0")
    steps, warmup = args.steps, args.warmup
    tuning = None
    if not args.no_autotune:
        from bigwham_dev_amd import autotune
        tuning = {"fp32": autotune.tune_mfma(dev, k, precision="fp32"),
                  "fp64": autotune.tune_mfma(dev, k, precision="fp64")}
        log(f"mfma schedules (ms per 32-RHS call): {tuning}")
    t_b32 = _timed(lambda: dev.matvec_batched(X, precision="fp32"), steps, warmup)
    x1 = X[:, 0].contiguous()
    y1 = torch.empty(n, dtype=torch.float64, device="cuda:0")
    t_b64 = _timed(lambda: dev.matvec_batched(X, precision="fp64"), steps, warmup)
    t_s32 = _timed(lambda: dev.matvec(x1, precision="fp32", out=y1), steps, warmup)
    t_s64 = _timed(lambda: dev.matvec(x1, precision="fp64", out=y1), steps, warmup)
    Yb = dev.matvec_batched(X, precision="fp32")
    ys = torch.stack([dev.matvec(X[:, j].contiguous(), precision="fp32") for j in (0, k - 1)], 1)
    ys64 = torch.stack([dev.matvec(X[:, j].contiguous()) for j in (0, k - 1)], 1)
    err = float(torch.linalg.matrix_norm(Yb[:, [0, k - 1]] - ys64) / torch.linalg.matrix_norm(ys64))
    Yb64 = dev.matvec_batched(X, precision="fp64")
    err64 = float(torch.linalg.matrix_norm(Yb64[:, [0, k - 1]] - ys64) / torch.linalg.matrix_norm(ys64))
    # the first column's single-RHS products against the exact dense operator
    # (sampled rows): with the batched-vs-FP64 gap above, the batched error
    vs_dense, _ = dense_check(host, dev, X[:, 0].cpu().numpy())
    log(f"vs dense operator {vs_dense}")
    flops = 2.0 * k * E
    flops32 = 2.0 * k * E32
    bytes_b = 4 * E32 + k * 8 * 2 * n    # FP32 factors once, X read and Y written in float64
    bytes_b64 = 8 * E + k * 8 * 2 * n
    rhs_s = k / t_b32
    log(f"32-RHS fp32 batched {t_b32 * 1e3:.3f} ms, single fp32 {t_s32 * 1e3:.3f} ms, "
        f"single fp64 {t_s64 * 1e3:.3f} ms; rel err vs fp64 {err:.2e}")

    # 16 uneven square cracks in one operator
    lo, hi = cfg["mesh_elems"]
    sides = meshes.crack_family(cfg["n_meshes"], lo, hi, seed=7)
    t0 = time.perf_counter()
    hosts = [_build_host(cfg["kernel"], *meshes.rect_grid(m), cfg) for m in sides]
    t_mbuild = time.perf_counter() - t0
    sizes = [h.size for h in hosts]
    log(f"16 meshes: sides {sides}, DoF {sum(sizes)}, build {t_mbuild:.1f}s")
    multi = HipHmat.many(hosts, precisions=("fp64", "fp32"), device=0, fp32_rank_tol=tol32)
    singles = [HipHmat(h, precisions=("fp64", "fp32"), device=0, fp32_rank_tol=tol32)
               for h in hosts]
    ntot = int(multi.row_offsets[-1])
    xm = torch.from_numpy(rng.uniform(-1.0, 1.0, ntot)).to("cuda:0")
    res = {}
    for prec in ("fp64", "fp32"):
        g = multi.graph(xm, precision=prec)
        t_graph = _timed(g.launch, steps, warmup)
        ym = torch.empty(ntot, dtype=torch.float64, device="cuda:0")
        t_multi = _timed(lambda: multi.matvec(xm, precision=prec, out=ym), steps, warmup)
        parts = [xm[multi.col_offsets[i]:multi.col_offsets[i + 1]] for i in range(len(hosts))]
        outs = [torch.empty_like(p) for p in parts]

        def sep():
            for d, p, o in zip(singles, parts, outs):
                d.matvec(p, precision=prec, out=o)
        t_sep = _timed(sep, steps, warmup)
        sep()
        yg = g.launch().clone()
        torch.cuda.synchronize()
        diff = float(torch.linalg.vector_norm(yg - torch.cat(outs)) /
                     torch.linalg.vector_norm(torch.cat(outs)))
        res[prec] = {"graph_ms": round(t_graph * 1e3, 4), "one_operator_ms": round(t_multi * 1e3, 4),
                     "16_separate_ms": round(t_sep * 1e3, 4),
                     "speedup_graph_vs_separate": round(t_sep / t_graph, 2),
                     "rel_diff_vs_separate": diff}
        g.close()
    # 32 right-hand sides over all 16 meshes: one graph replay per precision
    Xm = torch.from_numpy(rng.uniform(-1.0, 1.0, (ntot, k))).to("cuda:0")
    xparts = [Xm[multi.col_offsets[i]:multi.col_offsets[i + 1]] for i in range(len(hosts))]
    e_multi = sum(h.nb_entries() for h in hosts)
    e_multi32 = int(sum(multi.entries["fp32"][:3]))
    for prec, es in (("fp32", 4), ("fp64", 8)):
        e_m = e_multi32 if prec == "fp32" else e_multi
        if not args.no_autotune:
            from bigwham_dev_amd import autotune
            res[prec]["batched_tuning"] = autotune.tune_mfma(multi, k, precision=prec)
        gb = multi.graph_batched(Xm, precision=prec)
        t_gb = _timed(gb.launch, steps, warmup)
        Ym = torch.empty((ntot, k), dtype=torch.float64, device="cuda:0")
        t_eb = _timed(lambda: multi.matvec_batched(Xm, precision=prec, out=Ym), steps, warmup)
        outs_b = [torch.empty_like(p) for p in xparts]

        def sep_b():
            for d, p, o in zip(singles, xparts, outs_b):
                d.matvec_batched(p, precision=prec, out=o)
        t_sb = _timed(sep_b, steps, warmup)
        sep_b()
        Yg = gb.launch().clone()
        torch.cuda.synchronize()
        Ycat = torch.cat(outs_b)
        res[prec]["batched_32rhs"] = {
            "graph_ms": round(t_gb * 1e3, 4), "one_operator_ms": round(t_eb * 1e3, 4),
            "16_separate_ms": round(t_sb * 1e3, 4),
            "speedup_graph_vs_separate": round(t_sb / t_gb, 2),
            "rhs_per_s": round(k / t_gb, 1),
            "hbm_GBs": round((es * e_m + k * 8 * 2 * ntot) / t_gb / 1e9, 1),
            "tflops": round(2.0 * k * e_m / t_gb / 1e12, 2),
            "rel_diff_vs_separate": float(torch.linalg.matrix_norm(Yg - Ycat) /
                                          torch.linalg.matrix_norm(Ycat))}
        gb.close()
    log(f"16 meshes x 32 RHS: {({p: res[p]['batched_32rhs'] for p in res})}")
    dense = dense_gemm_reference(E, E32, k)
    dense["fp32"]["dense_best_over_h_call"] = round(dense["fp32"]["best_ms"] / (t_b32 * 1e3), 3)
    dense["fp64"]["dense_best_over_h_call"] = round(dense["fp64"]["best_ms"] / (t_b64 * 1e3), 3)
    log(f"dense GEMM of the same size through the vendor BLAS: {dense}")

    def roof(t, es, peak_tf, nbytes, flops):
        # attainable roof: the slower of the MFMA peak and HBM at 8 TB/s
        t_mfma = flops / (peak_tf * 1e12)
        t_hbm = nbytes / (HBM_PEAK_GBS * 1e9)
        if t_hbm >= t_mfma:
            return {"bound": "hbm", "achieved": round(nbytes / t / 1e9, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(t_hbm / t, 4), "traffic": None,
                    "attainable_ms": round(t_hbm * 1e3, 4),
                    "tflops": round(flops / t / 1e12, 2),
                    "attainable_tflops": round(flops / t_hbm / 1e12, 2), "mfma_peak_tflops": peak_tf,
                    "alg_bytes": int(nbytes)}
        return {"bound": "mfma", "achieved": round(flops / t / 1e12, 2), "peak": peak_tf,
                "unit": "TFLOP/s", "frac": round(t_mfma / t, 4), "traffic": None,
                "attainable_ms": round(t_mfma * 1e3, 4), "hbm_GBs": round(nbytes / t / 1e9, 1),
                "alg_bytes": int(nbytes)}

    out = {
        "metric": "batched H-matvec RHS/s", "value": round(rhs_s, 2), "unit": "RHS/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(t_b32 * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": cfg["workload"], "n_dof": n, "nb_entries": E, "nrhs": k,
                   "fp32_rank_tolerance": tol32, "fp32_entries": E32,
                   "fp32_entries_vs_fp64": round(E32 / E, 4)},
        "roofline": {**roof(t_b32, 4, 157.3, bytes_b, flops32),
                     "traffic": load_call_traffic("profiles/hbm_traffic_3dr0_150k_dof_32rhs.md",
                                                  tol32),
                     "traffic_source": "profiles/hbm_traffic_3dr0_150k_dof_32rhs.md"},
        "single_rhs": {"fp32_ms": round(t_s32 * 1e3, 4), "fp64_ms": round(t_s64 * 1e3, 4),
                       "batched_speedup_vs_32_fp32_calls": round(k * t_s32 / t_b32, 2)},
        "fp64_batched": {"ms_per_call": round(t_b64 * 1e3, 4), "rhs_per_s": round(k / t_b64, 1),
                         "roofline": {**roof(t_b64, 8, 78.6, bytes_b64, flops),
                                      "traffic": load_call_traffic(
                                          "profiles/hbm_traffic_3dr0_150k_dof_32rhs.md", tol32,
                                          "fp64"),
                                      "traffic_source":
                                          "profiles/hbm_traffic_3dr0_150k_dof_32rhs.md"},
                         "speedup_vs_32_fp64_calls": round(k * t_s64 / t_b64, 2)},
        "parity": {"batched_fp32_rel_err_vs_fp64": err, "batched_fp64_rel_err_vs_fp64": err64},
        "vs_dense_operator": vs_dense,
        "multi_mesh": {"sides": sides, "dof": sizes, "nb_entries": e_multi, **res},
        "dense_gemm_reference": dense,
        "setup_s": {"host_build": round(t_build, 2), "multi_build": round(t_mbuild, 2)},
        "tuning": tuning,
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args, cfg):
    """Gradient of ||H w - t_obs||^2 through the adjoint kernel (autograd).
    value = gradient evaluations (one H and one H^T matvec) per second."""
    import torch
    from bigwham_dev_amd import HipHmat
    from bigwham_dev_amd.autograd import hmatvec

    coor, conn = make_mesh(*cfg["mesh"])
    t0 = time.perf_counter()
    host = _build_host(cfg["kernel"], coor, conn, cfg)
    t_build = time.perf_counter() - t0
    n, E = host.size, host.nb_entries()
    log(f"c5: N={n} E={E} build {t_build:.1f}s")
    dev = HipHmat(host, precisions=("fp64",), adjoint=True, device=0)
    tuning = None
    if not args.no_autotune:
        from bigwham_dev_amd import autotune
        tuning = autotune.tune_unroll(dev, use_cache=False).unroll
        log(f"autotune {tuning}")
    xc = torch.from_numpy(host.collocation_points()).to("cuda:0")
    r2 = (xc[:, 0] ** 2 + xc[:, 1] ** 2)
    w_true = torch.exp(-r2 / (2 * 0.3 ** 2))  # Gaussian-bump opening
    t_clean = dev.matvec(w_true)
    g = torch.Generator(device="cuda:0").manual_seed(11)
    t_obs = t_clean + 0.01 * t_clean.std() * torch.randn(n, generator=g, device="cuda:0",
                                                          dtype=torch.float64)

    def loss_fn(w):
        r = hmatvec(dev, w) - t_obs
        return torch.dot(r, r)

    w = torch.zeros(n, dtype=torch.float64, device="cuda:0", requires_grad=True)

    def step():
        w.grad = None
        loss_fn(w).backward()
    t_grad = _timed(step, args.steps, args.warmup)
    # central differences on 20 random directions at a generic point
    w0 = (w_true + 0.1 * torch.randn(n, generator=g, device="cuda:0", dtype=torch.float64))
    w0.requires_grad_(True)
    L0 = loss_fn(w0)
    L0.backward()
    grad = w0.grad.detach()
    errs = []
    with torch.no_grad():
        for _ in range(20):
            d = torch.randn(n, generator=g, device="cuda:0", dtype=torch.float64)
            eps = 1e-3 * float(torch.linalg.vector_norm(w0) / torch.linalg.vector_norm(d))
            fd = (loss_fn(w0 + eps * d) - loss_fn(w0 - eps * d)) / (2 * eps)
            an = torch.dot(grad, d)
            errs.append(float(abs(fd - an) / abs(an)))
    # the FP64 matvec against the exact dense operator (sampled rows)
    vs_dense, _ = dense_check(host, dev, w0.detach().cpu().numpy())
    log(f"vs dense operator {vs_dense}")
    # adjoint consistency <H x, y> = <x, H^T y>
    x = torch.randn(n, generator=g, device="cuda:0", dtype=torch.float64)
    y = torch.randn(n, generator=g, device="cuda:0", dtype=torch.float64)
    lhs = float(torch.dot(dev.matvec(x), y))
    rhs = float(torch.dot(x, dev.rmatvec(y)))
    bytes_step = 2 * 8 * E
    out = {
        "metric": "loss+gradient evaluations/s", "value": round(1.0 / t_grad, 3),
        "unit": "grad/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_grad * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": cfg["workload"], "n_dof": n, "nb_entries": E,
                   "loss": "||H w - t_obs||^2, t_obs = H w_gauss + 1% noise"},
        "roofline": {"bound": "hbm", "achieved": round(bytes_step / t_grad / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_step / t_grad / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "fd_check": {"directions": 20, "max_rel_err": max(errs), "median_rel_err": float(np.median(errs))},
        "adjoint_identity_rel": abs(lhs - rhs) / abs(lhs),
        "vs_dense_operator": vs_dense,
        "setup_s": {"host_build": round(t_build, 2)},
        "tuning": tuning,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
