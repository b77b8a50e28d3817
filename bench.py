"""Benchmark of the BSE pipeline on one MI355X.

Workload (one "step" = one BSE matrix, build + solve + eps(omega)):
seeded random 2D hexagonal Wannier model, a = 3.2 A, nwann = 22, |R| <= 3a
(37 R vectors), hoppings N(0,1) exp(-|R|/1.2) eV, ~2 eV gap; 36x36x1
Monkhorst-Pack grid; nv = nc = 2 -> BSE dimension 5184; Rytova-Keldysh
potential r0 = 40 A, eps = 1; eps(omega) on 0-6 eV, 1001 points, CSHIFT 0.05.
The timed region runs, with the model already resident in HBM: k grid, H(k)
phase GEMM, per-k eigensolves, dipoles, W table, BSE assembly, Hermitian
eigensolve, activities and eps(omega).  By default the value is the precision
form (COMPLEX*8 matrix, FP32 MFMA assembly and trailing updates, FP64 norms,
tridiagonal solve and spectrum) and the plain FP64 form is timed beside it
with the same steps, together with the precision form's error against it.  The
line also carries the iterative form's measurement on the 60x60 model
(`iterative_form`: lowest 256 excitons + Lanczos eps against the full FP64 solve
of that matrix; skip it with --no-iterative).

    python bench.py --form davidson --config c5   # the iterative form alone
    python bench.py --form batch                  # the batching form: a set of uneven
                                                  # models, batch API vs single plans
    python bench.py --form bands                  # exciton bands (finite Q) along
                                                  # Gamma-K-K'-Gamma, s per q-point

    python bench.py [--steps K] [--warmup W] [--precision mixed|fp64] [--no-cpu]
    python bench.py --temp 3000 --ta VE --st 0.3 --phavg 0.05   # finite temperature
                                                  # (bsesolvertemp) on the same workload
    python bench.py --copy-probe      # one 1 GiB device copy, to scale HBM counters
    python bench.py --no-graph ...    # direct launches (for kernel tracers)

Prints one JSON line.  The CPU baseline is the reference's own code (its
subroutines compiled from source with flang + MKL, oracle/_ref/ref_dump.x) on
the same model, timed on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

CONFIG = dict(name="hex2d_nw22_36x36x1_nv2nc2", seed=2, nw=22, ngrid=(36, 36, 1), nc=2, nv=2,
              coultype="V2DRK", r0=40.0, ediel=(1.0, 1.0, 1.0), lc=1.0, ez=1.0, w=0.0, ktol=0.001,
              cshift=0.05, ebse0=0.0, ebsef=6.0, numbse=1001, dtfull=True, renorm=True,
              shift=(0.0, 0.0, 0.0))
SMALL = dict(CONFIG, name="hex2d_nw8_12x12x1_nv2nc2", seed=1, nw=8, ngrid=(12, 12, 1))
LARGE2D = dict(CONFIG, name="hex2d_nw24_60x60x1_nv2nc2", seed=5, nw=24, ngrid=(60, 60, 1))
# large single solve, 3D bulk: cubic a = 5.4 A, nw = 16, |R| <= 2 cells, 16x16x16 k, nv = nc = 3
BULK3D = dict(CONFIG, name="cubic3d_nw16_16x16x16_nv3nc3", lattice="cubic", seed=3, nw=16, ngrid=(16, 16, 16),
              nc=3, nv=3, coultype="V3DL", ediel=(6.0, 6.0, 6.0))
CONFIGS = {"c2": CONFIG, "c1": SMALL, "c3": BULK3D, "c5": LARGE2D}
REDUCTION = {"auto": None, "one": False, "two": True}
HBM_PEAK = 8000.0       # GB/s
FP64_PEAK = 78.6        # TFLOP/s, dense FP64 MFMA
FP32_PEAK = 157.3       # TFLOP/s, dense FP32 MFMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_model(cfg):
    from wantibexos_amd import tbmodel
    if cfg.get("lattice") == "cubic":
        return tbmodel.cubic3d(cfg["seed"], cfg["nw"], cfg["ngrid"])
    return tbmodel.hex2d(cfg["seed"], cfg["nw"], cfg["ngrid"])


def params(cfg, precision):
    from wantibexos_amd import bse
    return bse.BSEParams(ngrid=cfg["ngrid"], mshift=cfg["shift"], nc=cfg["nc"], nv=cfg["nv"],
                         coultype=cfg["coultype"], ediel=cfg["ediel"], lc=cfg["lc"], ez=cfg["ez"],
                         w=cfg["w"], r0=cfg["r0"], ktol=cfg["ktol"], cshift=cfg["cshift"],
                         ebse0=cfg["ebse0"], ebsef=cfg["ebsef"], numbse=cfg["numbse"],
                         dtfull=cfg["dtfull"], cpol=True, renorm=cfg["renorm"], f154=True,
                         precision=bse.MIXED if precision == "mixed" else bse.FP64,
                         bset=cfg.get("bset", False), temp=cfg.get("temp", 0.0), ta=cfg.get("ta", "FA"),
                         st=cfg.get("st", 0.0), phavg=cfg.get("phavg", 0.0))


def cpu_baseline(cfg, model, max_threads):
    """The reference's own subroutines (oracle/_ref/ref_dump.x) on the same model."""
    from wantibexos_amd import tbmodel
    exe = ROOT / "oracle" / "_ref" / "ref_dump.x"
    if not exe.exists():
        return None
    nthreads = max(1, min(max_threads, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        (td / "out").mkdir()
        tbmodel.write_tb(model, td / "tb.dat")
        (td / "input.dat").write_text(tbmodel.input_dat(cfg, "tb.dat", nthreads=nthreads))
        env = dict(os.environ, OMP_NUM_THREADS=str(nthreads), MKL_NUM_THREADS=str(nthreads))
        t0 = time.perf_counter()
        r = subprocess.run([str(exe)], stdin=open(td / "input.dat"), cwd=td, env=env,
                           capture_output=True, text=True, timeout=900)
        wall = time.perf_counter() - t0
    if r.returncode != 0:
        log("cpu baseline failed:", r.stderr[-800:])
        return None
    m = re.search(r"TIMES.*", r.stdout)
    vals = dict(re.findall(r"(\w+)=\s*([0-9.Ee+-]+)", m.group(0)))
    stages = {k: float(vals[k]) for k in ("ab", "dip", "build", "diag", "osc", "eps")}
    return dict(value=sum(stages.values()), unit="s/matrix", cores=nthreads, kind="reference",
                sample=(f"one full {cfg['name']} matrix (N={int(vals['N'])}) through the reference's "
                        f"own subroutines (flang -O2 + MKL ZHEEV, LWORK as the reference), "
                        f"stage sums excl. file I/O; wall {wall:.1f}s"),
                stages=stages)


def cpu_baseline_1core(cfg, N):
    """The reference's own subroutines on ONE thread, on a smaller matrix of the same model
    family (hexagonal, same nw, 24x24 k grid: N = 2304), scaled to the workload's N stage by
    stage: ZHEEV as N^3, the build and oscillator strengths as N^2, the rest as N."""
    from wantibexos_amd import tbmodel
    g = 24
    small = dict(cfg, name=cfg["name"] + "_1core_sample", ngrid=(g, g, 1))
    m = tbmodel.hex2d(cfg["seed"], cfg["nw"], small["ngrid"])
    r = cpu_baseline(small, m, 1)
    if not r:
        return None
    n_s = g * g * cfg["nc"] * cfg["nv"]
    f = N / n_s
    st = r["stages"]
    scaled = {"ab": st["ab"] * f, "dip": st["dip"] * f, "build": st["build"] * f * f, "diag": st["diag"] * f ** 3,
              "osc": st["osc"] * f * f, "eps": st["eps"] * f}
    return dict(value=round(sum(scaled.values()), 3), unit="s/matrix", cores=1, kind="reference",
                sample=(f"the reference's subroutines on one thread on the same model family at N={n_s} "
                        f"(nw={cfg['nw']}, {g}x{g}x1 k grid): {r['value']:.2f} s measured, scaled to N={N} "
                        f"(ZHEEV N^3, build and oscillators N^2, the rest N)"),
                measured_s=round(r["value"], 3), stages_scaled={k: round(v, 3) for k, v in scaled.items()})


def tridiag_bytes(N, nb=32, elem=16.0):
    """Algorithmic HBM bytes of the one-stage reduction: each Hermitian GEMV reads the
    lower triangle of its trailing matrix once; each panel update reads and writes the
    lower triangle of its trailing matrix once."""
    hemv = sum(elem * (N - j - 1) * (N - j) / 2 for j in range(N - 1))
    her2k = sum(2 * elem * (N - s) * (N - s + 1) / 2 for s in range(nb, N, nb))
    return hemv, her2k


def bench_davidson(args, cfg, model):
    """Iterative form: build + lowest nev excitons (block Davidson, FP32 products with an FP64
    polish) + eps(omega) from Lanczos chains, timed against the full FP64 solve of the plain
    path on the same model, which is also the accuracy reference."""
    import dataclasses
    from wantibexos_amd import bse, learned
    ctx = bse.Context(0)
    p32 = dataclasses.replace(params(cfg, "mixed"), f154=False)
    p64 = dataclasses.replace(params(cfg, "fp64"), f154=False)
    net = None
    if args.learned:
        net = bse.MLP(ctx)
        net.set_params(learned.load())
    plan = bse.BSEPlan(ctx, model, p32)
    N = plan.N
    log(f"iterative form: N={N} nev={args.nev} lanczos={args.lanczos_steps} learned={args.learned}")
    kw = dict(nev=args.nev, guard=args.guard, tol=args.tol, lanczos_steps=args.lanczos_steps, mlp=net,
              max_sub=args.max_sub, coarse=args.coarse)
    for _ in range(args.warmup):
        plan.lowest(**kw)
    times, r = [], None
    for _ in range(args.steps):
        ctx.sync()
        t0 = time.perf_counter()
        r = plan.lowest(**kw)
        ctx.sync()
        times.append(time.perf_counter() - t0)
        log(f"  step {len(times)}: {times[-1] * 1e3:.1f} ms  {r['info']}")
    plan.close()
    dt = float(np.mean(times))
    coarse = args.coarse if args.coarse is not None else min(10 * (args.nev + args.guard), N // 4)
    # the reference point: full FP64 solve (plain path), timed once after a warm-up
    with bse.BSEPlan(ctx, model, p64) as full:
        if args.no_graph:
            full.set_graph(False)
        full.execute()
        ctx.sync()
        t0 = time.perf_counter()
        full.execute()
        ref = full.results()
        t_full = time.perf_counter() - t0
    log(f"full FP64 solve: {t_full * 1e3:.1f} ms")
    dW = np.abs(r["W"] - ref["W"][:args.nev])
    eps_err = {}
    if r["eps"] is not None:
        for c in range(6):
            for q in range(2):
                rr = ref["eps"][c, q]
                if np.abs(rr).max() > 1e-8:
                    eps_err[f"{c}{'ri'[q]}"] = float(np.abs(r["eps"][c, q] - rr).max() / np.abs(rr).max())
    if net is not None:
        net.close()
    ctx.close()
    out = {
        "metric": "BSE build + lowest excitons + eps(omega) s/matrix, iterative form vs full FP64",
        "value": round(dt, 6), "unit": "s/matrix", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 products, c128 Rayleigh-Ritz", "data": "synthetic (seeded random tight-binding model)",
        "config": {"workload": cfg["name"], "bse_dim": N, "nev": args.nev, "lanczos_steps": args.lanczos_steps,
                   "tol_eV": args.tol, "guard": args.guard, "max_sub": args.max_sub or 5 * (args.nev + args.guard),
                   "coarse": coarse,
                   "preconditioner": ("learned" if args.learned else "diagonal")
                   + (f" + coarse block of {coarse} states" if coarse else "")},
        "full_fp64_s": round(t_full, 6), "speedup_vs_full_fp64": round(t_full / dt, 3),
        "info": r["info"], "max_dW_eV": float(dW.max()), "max_resid_eV": float(r["resid"].max()),
        "eps_rel_err": eps_err,
    }
    return out


def bench_batch(args):
    """Batching form: a set of uneven models (different Wannier counts and k grids) through
    the batch API (one captured graph for the set: per-matrix build and post-processing on
    --lanes streams, the one-stage reductions in lockstep, one launch per kernel and column
    step for all matrices), against the same plans run one after the other (results equal
    bit for bit) and against plain FP64 plans of the same models one after the other."""
    import dataclasses
    from concurrent.futures import ProcessPoolExecutor
    from wantibexos_amd import bse, tbmodel
    spec = [(12, 8), (14, 6), (16, 10), (18, 8), (20, 12), (22, 6), (24, 8), (16, 10)]   # (k grid, nw)
    if args.config == "c4":
        # 64 seeded models: nw uniform in 6..40, k grids uniform in 12x12 .. 30x30 (N 576 .. 3600)
        rng = np.random.default_rng(4)
        spec = [(int(rng.integers(12, 31)), int(rng.integers(6, 41))) for _ in range(64)]
    # the models are generated on host worker processes, before this process touches the GPU
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
        models = list(pool.map(tbmodel.hex2d, [200 + i for i in range(len(spec))], [nw for _, nw in spec],
                               [(ng, ng, 1) for ng, _ in spec]))
    log(f"{len(models)} models generated in {time.perf_counter() - t0:.1f}s")
    prms = [params(dict(CONFIG, ngrid=(ng, ng, 1)), args.precision) for ng, _ in spec]
    ctx = bse.Context(0)
    plans = [bse.BSEPlan(ctx, m, p) for m, p in zip(models, prms)]
    dims = [pl.N for pl in plans]
    log(f"batching form: {len(plans)} matrices, N = {dims}, precision {args.precision}, lanes {args.lanes}")

    def run_seq(pls):
        for pl in pls:
            pl.execute()
        ctx.sync()

    def time_seq(pls):
        for _ in range(args.warmup):
            run_seq(pls)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run_seq(pls)
        return (time.perf_counter() - t0) / args.steps

    t_single = time_seq(plans) if not args.no_single else None
    ref = [pl.results() for pl in plans] if not args.no_single else None
    for pl in plans:
        pl.close()
    t_fp64 = t_single
    if args.precision != "fp64" and not args.no_single:
        # the sequential plain FP64 solves of the same models
        plans = [bse.BSEPlan(ctx, m, params(dict(CONFIG, ngrid=(ng, ng, 1)), "fp64")) for m, (ng, _) in zip(models, spec)]
        t_fp64 = time_seq(plans)
        for pl in plans:
            pl.close()
    if t_fp64 is not None:
        log(f"sequential plain FP64 plans {t_fp64 * 1e3:.1f} ms per set")
    with bse.BSEBatch(ctx, models, prms, lanes=args.lanes) as b:
        tiles = b.tune().tolist() if args.tune else None
        for _ in range(args.warmup):
            b.execute()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.execute()
        ctx.sync()
        t_batch = (time.perf_counter() - t0) / args.steps
        same = all(np.array_equal(b.results(i)["W"], ref[i]["W"]) for i in range(len(ref))) if ref and not args.tune else None
        dW = max(float(np.abs(b.results(i)["W"] - ref[i]["W"]).max()) for i in range(len(ref))) if ref else None
    ctx.close()
    log(f"single plans {t_single * 1e3:.1f} ms per set, batch {t_batch * 1e3:.1f} ms per set" if t_single else
        f"batch {t_batch * 1e3:.1f} ms per set")
    rnd = lambda x: round(x, 6) if x is not None else None
    ratio = lambda x: round(x / t_batch, 3) if x is not None else None
    out = {
        "metric": "BSE build+solve of a set of uneven matrices, batch API vs sequential single plans",
        "value": round(t_batch, 6), "unit": "s/set", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_batch * 1e3, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128" if args.precision == "fp64" else "c64/c128", "data": "synthetic (seeded random tight-binding models)",
        "config": {"workload": f"hex2d set of {len(spec)}", "models": [{"kgrid": ng, "nwann": nw, "bse_dim": N} for (ng, nw), N in zip(spec, dims)],
                   "lanes": args.lanes, "precision": args.precision, "tuned_panels": tiles},
        "single_plans_s": rnd(t_single), "speedup_vs_single": ratio(t_single),
        "fp64_plain_sequential_s": rnd(t_fp64), "speedup_vs_fp64_sequential": ratio(t_fp64),
        "matrices_per_s": round(len(dims) / t_batch, 2), "bitwise_equal_to_single": same, "max_dW_eV": dW,
    }
    print(json.dumps(out), flush=True)


HEX_PATH = [(0.0, 0.0, 0.0), (0.6667, -0.3333, 0.0), (0.6667, -0.3333, 0.0), (0.3333, 0.3333, 0.0),
            (0.3333, 0.3333, 0.0), (0.0, 0.0, 0.0)]


def cpu_baseline_bands(cfg, model, max_threads):
    """The unmodified reference program (oracle/_ref/wtb.x, BSE_BND= T) on the same model
    for a two-point path (Gamma, K): wall time per q-point."""
    from wantibexos_amd import tbmodel
    exe = ROOT / "oracle" / "_ref" / "wtb.x"
    if not exe.exists():
        return None
    nthreads = max(1, min(max_threads, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        (td / "out").mkdir()
        tbmodel.write_tb(model, td / "tb.dat")
        (td / "kpath.dat").write_text("2\n2\n0.0 0.0 0.0\n0.6667 -0.3333 0.0\n")
        (td / "input.dat").write_text(tbmodel.input_dat(dict(cfg, kpath_bse="kpath.dat"), "tb.dat", nthreads=nthreads))
        env = dict(os.environ, OMP_NUM_THREADS=str(nthreads), MKL_NUM_THREADS=str(nthreads))
        t0 = time.perf_counter()
        r = subprocess.run([str(exe)], stdin=open(td / "input.dat"), cwd=td, env=env, capture_output=True,
                           text=True, timeout=900)
        wall = time.perf_counter() - t0
        ok = (td / "out" / "bands_bse.dat").exists()
    if not ok:
        log("cpu baseline (bands) failed:", r.stdout[-400:], r.stderr[-400:])
        return None
    return dict(value=wall / 2, unit="s/q-point", cores=nthreads, kind="reference",
                sample=(f"the reference program's bsebnds on {cfg['name']} (N={cfg['N']}) for a two-point path "
                        f"(Gamma, K): wall {wall:.1f}s incl. its k-grid setup and file output"))


def bench_bands(args, cfg, model):
    """Exciton bands along Gamma-K-K'-Gamma (--npts points per segment): s per q-point with
    the plans kept (one finite-Q plan and stream per lane, graphs captured in the warmup)."""
    from wantibexos_amd import bse
    ctx = bse.Context(0)
    qp, _ = bse.kpathbse(model.rlat, np.array(HEX_PATH), args.npts)
    nq = len(qp)
    out = {}
    ref = None
    for prec in (["mixed", "fp64"] if args.precision == "mixed" else ["fp64"]):
        for lanes in sorted({args.lanes, 1}, reverse=True):
            with bse.BSEBands(ctx, model, params(cfg, prec), lanes=lanes) as b:
                if args.no_graph:
                    b.set_graph(False)
                for _ in range(max(1, args.warmup)):
                    exk = b.run(qp)
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    exk = b.run(qp)
                dt = (time.perf_counter() - t0) / args.steps
            out[(prec, lanes)] = (dt, exk)
            log(f"bands {prec} lanes={lanes}: {dt * 1e3:.1f} ms per path of {nq} q-points")
            if prec == "fp64" and lanes == args.lanes:
                ref = exk
    N = out[(args.precision, args.lanes)][1].shape[0]
    cfg = dict(cfg, N=N)
    # dominant kernel: the reduction's GEMV in one q-point's execute (per-launch HIP events)
    with bse.BSEPlan(ctx, model, params(cfg, args.precision), finite_q=True) as plan:
        plan.set_q(qp[1])
        plan.execute()
        plan.energies()
        plan.set_timing(2)
        plan.execute()
        plan.energies()
        k_ms, k_n = plan.kernel_ms()
        k_flops = plan.kernel_flops()
    elem = 8.0 if args.precision == "mixed" else 16.0
    hemv_b, _ = tridiag_bytes(N, elem=elem)
    avg_s = k_ms * 1e-3 / max(k_n, 1)
    if k_flops > 0:
        per_launch = k_flops / max(k_n, 1)
        ach = per_launch / avg_s / 1e12
        peak = FP32_PEAK if args.precision == "mixed" else FP64_PEAK
        roofline = dict(bound="mfma", achieved=round(ach, 2), peak=peak, unit="TFLOP/s", frac=round(ach / peak, 4),
                        traffic=None, kernel="k_her2k_full", launches=k_n, avg_launch_us=round(avg_s * 1e6, 3),
                        flops_per_launch=round(per_launch))
    else:
        per_launch = hemv_b / max(k_n, 1)
        ach = per_launch / avg_s / 1e9
        roofline = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK, unit="GB/s",
                        frac=round(ach / HBM_PEAK, 4), traffic=None, kernel="k_tri_hemv", launches=k_n,
                        avg_launch_us=round(avg_s * 1e6, 3), bytes_per_launch=round(per_launch))
    cpu = None if args.no_cpu else cpu_baseline_bands(cfg, model, args.cpu_threads)
    ctx.close()
    dt, exk = out[(args.precision, args.lanes)]
    checks = {"lanes_bitwise_equal": bool(np.array_equal(exk, out[(args.precision, 1)][1])),
              "sorted": bool(np.all(np.diff(exk, axis=0) >= 0)),
              "path_ends_equal_at_shared_points": float(np.abs(exk[:, args.npts - 1] - exk[:, args.npts]).max())}
    if ref is not None and args.precision == "mixed":
        checks["mixed_vs_fp64_max_dW_eV"] = float(np.abs(exk - ref).max())
    res = {
        "metric": "exciton bands (finite-Q BSE) s per q-point",
        "value": round(dt / nq, 6),
        "unit": "s/q-point",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if args.precision == "fp64" else "c64/c128",
        "data": "synthetic (seeded random tight-binding model)",
        "config": {"workload": f"{cfg['name']}_bands_GKK'G_{nq}q", "bse_dim": N, "q_points": nq,
                   "lanes": args.lanes, "precision": args.precision},
        "lanes_1_s_per_q": round(out[(args.precision, 1)][0] / nq, 6),
        "fp64_s_per_q": round(out[("fp64", args.lanes)][0] / nq, 6) if ("fp64", args.lanes) in out else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "checks": checks,
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed",
                    help="mixed (default): the precision form, timed beside the plain FP64 form")
    ap.add_argument("--config", choices=["c2", "c1", "c3", "c4", "c5"], default="c2",
                    help="c4: the 64-model set of the batching form (--form batch only)")
    ap.add_argument("--form", choices=["full", "davidson", "batch", "bands"], default="full",
                    help="full: the whole eigensolve (plain FP64 or precision form); davidson: the "
                         "iterative form (lowest --nev excitons, Lanczos eps) against the full FP64 solve")
    ap.add_argument("--nev", type=int, default=256)
    ap.add_argument("--lanes", type=int, default=4, help="batching form and bands: streams")
    ap.add_argument("--no-single", action="store_true", help="batching form: time the batch only (profiling)")
    ap.add_argument("--npts", type=int, default=4, help="bands: q-points per path segment (3 segments)")
    ap.add_argument("--tune", action="store_true", help="batching form: run the panel autotuner first")
    ap.add_argument("--refine", type=int, default=256,
                    help="precision form: FP64 Rayleigh-quotient refinement of this many lowest excitons (0: off)")
    ap.add_argument("--no-cpu-1core", action="store_true",
                    help="skip the one-thread timing of the reference's subroutines (a smaller matrix of the "
                         "same model family, scaled to the workload)")
    ap.add_argument("--lanczos-steps", type=int, default=250)
    ap.add_argument("--guard", type=int, default=32, help="iterative form: extra Ritz pairs carried along")
    ap.add_argument("--coarse", type=int, default=None,
                    help="iterative form: coarse block of the preconditioner (states; default 10 (nev + guard), "
                         "at most N/4; 0: diagonal only)")
    ap.add_argument("--max-sub", type=int, default=None,
                    help="iterative form: basis size that triggers a restart (default 5 (nev + guard))")
    ap.add_argument("--tol", type=float, default=1e-6,
                    help="iterative form: residual norm (eV) below which an exciton counts as converged; "
                         "a Ritz value lies within its residual norm of an eigenvalue")
    ap.add_argument("--learned", action="store_true", help="iterative form with the trained preconditioner")
    ap.add_argument("--temp", type=float, default=None,
                    help="finite temperature (K): the bsesolvertemp path (BSET= T) on the same workload")
    ap.add_argument("--ta", choices=["FA", "VE", "BE"], default="VE", help="finite temperature: gap correction")
    ap.add_argument("--st", type=float, default=0.3, help="finite temperature: ST")
    ap.add_argument("--phavg", type=float, default=0.05, help="finite temperature: PHAVG")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-iterative", action="store_true",
                    help="full form: skip the iterative form's measurement on the 60x60 model that the "
                         "line carries beside the metric")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--reduction", choices=["auto", "one", "two"], default="auto",
                    help="tridiagonal reduction: one-stage Householder, two-stage (dense -> band -> "
                         "tridiagonal), or by size (auto: two-stage from N = 8192 on, one-stage where "
                         "exciton vectors are formed)")
    ap.add_argument("--band", type=int, default=32, choices=[16, 32], help="two-stage reduction: band width")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the kernels directly instead of replaying the captured HIP graph")
    ap.add_argument("--pmc-light", action="store_true",
                    help="one execute and nothing else (for hardware-counter passes)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy kernel (counter scale)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 1 GiB three times")
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    from wantibexos_amd import bse
    if args.form == "batch":
        return bench_batch(args)
    if args.config == "c4":
        print("bench.py: --config c4 is the batching form's set; use --form batch", file=sys.stderr)
        sys.exit(2)
    cfg = CONFIGS[args.config]
    if args.temp is not None:
        cfg = dict(cfg, name=f"{cfg['name']}_T{args.temp:g}K_{args.ta}", bset=True, temp=args.temp, ta=args.ta,
                   st=args.st, phavg=args.phavg)
    t0 = time.perf_counter()
    model = make_model(cfg)
    log(f"model {cfg['name']} generated in {time.perf_counter() - t0:.1f}s")
    if args.form == "davidson":
        print(json.dumps(bench_davidson(args, cfg, model)), flush=True)
        return
    if args.form == "bands":
        return bench_bands(args, cfg, model)
    prm = params(cfg, args.precision)
    ctx = bse.Context(0)
    plan = bse.BSEPlan(ctx, model, prm)
    if args.no_graph:
        plan.set_graph(False)
    N = plan.N
    plan.set_reduction(REDUCTION[args.reduction], args.band)
    nref = min(args.refine, N) if args.precision == "mixed" else 0
    if nref:
        plan.set_refine(nref)
    log(f"plan N={N} created" + (f", FP64 refinement of the lowest {nref} excitons" if nref else ""))
    if args.pmc_light:
        plan.execute()
        plan.results()
        log("pmc-light: one execute done")
        return
    for _ in range(args.warmup):
        plan.execute()
    res = plan.results()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute()
    ctx.sync()
    dt = (time.perf_counter() - t0) / args.steps
    res = plan.results()
    log(f"timed {args.steps} steps: {dt * 1e3:.2f} ms/step")
    rq_res, w_unref = plan.refine_results() if nref else (None, None)
    dt_unref = None
    if nref:
        # the same form without the refinement, for the record
        plan.set_refine(0)
        for _ in range(args.warmup):
            plan.execute()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.execute()
        ctx.sync()
        dt_unref = (time.perf_counter() - t0) / args.steps
        plan.set_refine(nref)
        log(f"without the refinement: {dt_unref * 1e3:.2f} ms/step")
    # per-stage device times (HIP events on the context stream), separate run
    plan.set_timing(True)
    plan.execute()
    plan.results()
    stage = plan.stage_ms()
    plan.set_timing(False)
    log("stage ms:", json.dumps({k: round(v, 3) for k, v in stage.items()}))
    # size-independent checks of the result
    eigv = plan.fetch("eigv")
    dip = plan.fetch("dipole")
    fw = plan.fetch("fdeh") if prm.bset else np.ones(N)      # f_v - f_c (1 at T = 0)
    nv = prm.nv
    dE = (eigv[:, nv:][:, :, None] - eigv[:, :nv][:, None, :]).reshape(-1)
    sx, sy = (np.sum(np.abs(dip[i]) ** 2 * fw) for i in (0, 1))
    checks = dict(
        trace_rel=abs(res["W"].sum() - dE.sum()) / abs(dE.sum()),           # sum W = trace H (V(0)=0)
        sumrule_xx_rel=abs(res["act"][0].sum() - sx) / sx,                   # sum act = sum f |h|^2
        sumrule_yy_rel=abs(res["act"][1].sum() - sy) / sy,
        sorted=bool(np.all(np.diff(res["W"]) >= 0)),
        ground_state_eV=float(res["W"][0]),
    )
    plain = None
    if args.precision == "mixed":
        # the same workload in the plain FP64 form, timed the same way: the precision
        # form's reference point and error bar
        with bse.BSEPlan(ctx, model, params(cfg, "fp64")) as p64:
            if args.no_graph:
                p64.set_graph(False)
            p64.set_reduction(REDUCTION[args.reduction], args.band)
            for _ in range(args.warmup):
                p64.execute()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                p64.execute()
            ctx.sync()
            dt64 = (time.perf_counter() - t0) / args.steps
            r64 = p64.results()
        log(f"plain FP64: {dt64 * 1e3:.2f} ms/step")
        plain = {"value": round(dt64, 6), "ms_per_step": round(dt64 * 1e3, 3),
                 "speedup_mixed_vs_plain": round(dt64 / dt, 3)}
        checks["mixed_vs_fp64_max_dW_eV"] = float(np.abs(res["W"] - r64["W"]).max())
        checks["mixed_vs_fp64_eps_rel"] = float(max(
            np.abs(res["eps"][c, q] - r64["eps"][c, q]).max() / max(np.abs(r64["eps"][c, q]).max(), 1e-30)
            for c in range(6) for q in range(2) if np.abs(r64["eps"][c, q]).max() > 1e-8))
        # per component (xx yy zz xy xz yz) and part: the off-diagonal spectra of a cubic model
        # are small, so their error relative to their own maximum is the largest
        names = ["xx", "yy", "zz", "xy", "xz", "yz"]
        checks["mixed_vs_fp64_eps_rel_by_component"] = {
            f"{names[c]}{'ri'[q]}": float(np.abs(res["eps"][c, q] - r64["eps"][c, q]).max()
                                          / max(np.abs(r64["eps"][c, q]).max(), 1e-30))
            for c in range(6) for q in range(2) if np.abs(r64["eps"][c, q]).max() > 1e-8}
        checks["mixed_vs_fp64_eps_rel_of_diagonal_max"] = float(max(
            np.abs(res["eps"][c, q] - r64["eps"][c, q]).max() for c in range(6) for q in range(2))
            / max(np.abs(r64["eps"][c, q]).max() for c in range(3) for q in range(2)))
        log("mixed vs fp64:", json.dumps({k: checks[k] for k in checks if k.startswith("mixed")}))
        if nref:
            # per exciton: the refined energy against the FP64 solve, and the residual bound
            # |H x - theta x| / |x| of the FP64 matrix (an eigenvalue lies within it)
            dref = np.abs(res["W"][:nref] - r64["W"][:nref])
            dun = np.abs(w_unref - r64["W"][:nref])
            refine = {"nref": nref, "max_dW_refined_eV": float(dref.max()), "mean_dW_refined_eV": float(dref.mean()),
                      "max_dW_unrefined_eV": float(dun.max()), "max_resid_eV": float(rq_res.max()),
                      "median_resid_eV": float(np.median(rq_res)),
                      "dW_within_resid": bool(np.all(dref <= rq_res + 1e-12)),
                      "per_exciton_dW_eV_first8": [float(x) for x in dref[:8]],
                      "per_exciton_resid_eV_first8": [float(x) for x in rq_res[:8]],
                      "s_without_refinement": round(dt_unref, 6)}
            checks["refinement"] = refine
            log("refinement:", json.dumps(refine))
    # dominant kernel: the reduction's Hermitian GEMV, one launch per column, each
    # bracketed by HIP events on the context stream (a separate, non-graph execute)
    elem = 8.0 if args.precision == "mixed" else 16.0
    hemv_b, her2k_b = tridiag_bytes(N, elem=elem)
    if nref:
        plan.set_refine(0)          # the reduction the timed form uses for the matrix itself
    plan.set_timing(2)
    plan.execute()
    plan.results()
    k_ms, k_n = plan.kernel_ms()
    k_flops = plan.kernel_flops()
    plan.set_timing(False)
    if nref:
        plan.set_refine(nref)
    avg_s = k_ms * 1e-3 / max(k_n, 1)
    if k_flops > 0:
        # two-stage reduction: the dense -> band stage's trailing update, MFMA-bound
        per_launch = k_flops / max(k_n, 1)
        ach = per_launch / avg_s / 1e12
        peak = FP32_PEAK if args.precision == "mixed" else FP64_PEAK
        log(f"k_her2k_full: {k_n} launches, avg {avg_s * 1e6:.2f} us, {per_launch / 1e9:.3f} GFLOP/launch")
        roofline = dict(bound="mfma", achieved=round(ach, 2), peak=peak, unit="TFLOP/s", frac=round(ach / peak, 4),
                        traffic=None, kernel="k_her2k_full", launches=k_n, avg_launch_us=round(avg_s * 1e6, 3),
                        flops_per_launch=round(per_launch))
    else:
        per_launch = hemv_b / max(k_n, 1)
        ach = per_launch / avg_s / 1e9
        log(f"k_tri_hemv: {k_n} launches, avg {avg_s * 1e6:.2f} us, {per_launch / 1e6:.2f} MB/launch")
        # HBM traffic per launch: the measured traffic/algorithmic ratio of the counter passes
        # (profiles/, same workload and precision) applied to the mean launch
        traffic = None
        tfile = ROOT / "profiles" / f"k_tri_hemv_traffic_hex2d_nw22_k36_N5184_{args.precision}.json"
        if tfile.exists() and args.config == "c2" and not prm.bset:
            tj = json.loads(tfile.read_text())
            traffic = round(tj["traffic_over_algorithmic"] * per_launch)
        roofline = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK, unit="GB/s",
                        frac=round(ach / HBM_PEAK, 4), traffic=traffic,
                        kernel="k_tri_hemv", launches=k_n, avg_launch_us=round(avg_s * 1e6, 3),
                        bytes_per_launch=round(per_launch))
    cpu = None
    cpu1 = None
    if not args.no_cpu:
        log("cpu baseline: running the reference's subroutines on the same model ...")
        cpu = cpu_baseline(cfg, model, args.cpu_threads)
        if cpu:
            log(f"cpu baseline {cpu['value']:.2f} s/matrix on {cpu['cores']} threads: {cpu['stages']}")
        if not args.no_cpu_1core and cfg.get("lattice") != "cubic":
            cpu1 = cpu_baseline_1core(cfg, N)
            if cpu1:
                log(f"cpu baseline, one thread: {cpu1['value']:.1f} s/matrix ({cpu1['sample']})")
    plan.close()
    ctx.close()
    iterative = None
    if not args.no_iterative and args.config == "c2" and args.precision == "mixed" and not prm.bset:
        # the iterative form at the target accuracy (every residual within 1e-6 eV, 250-step
        # chains) on the 60x60 model, against the full FP64 solve of the same matrix:
        # its cost grows like N^2 and the full solve's like N^3, so it pays on larger matrices
        import copy
        ia = copy.copy(args)
        ia.tol, ia.lanczos_steps, ia.steps, ia.warmup, ia.learned = 1e-6, 250, 2, 1, False
        lcfg = CONFIGS["c5"]
        t0 = time.perf_counter()
        lmodel = make_model(lcfg)
        log(f"iterative form on {lcfg['name']} (model generated in {time.perf_counter() - t0:.1f}s)")
        d = bench_davidson(ia, lcfg, lmodel)
        iterative = {k: d[k] for k in ("metric", "value", "unit", "config", "full_fp64_s", "speedup_vs_full_fp64",
                                       "max_dW_eV", "max_resid_eV", "eps_rel_err")}
        iterative["iterations"] = d["info"]["iters"]
    out = {
        "metric": "BSE build+solve s/matrix, mixed vs FP64 plain, at <=1 meV exciton error",
        "value": round(dt, 6),
        "unit": "s/matrix",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if args.precision == "fp64" else "c64/c128",
        "data": "synthetic (seeded random tight-binding model)",
        "config": {"workload": cfg["name"], "bse_dim": N, "nwann": cfg["nw"],
                   "kgrid": list(cfg["ngrid"]), "potential": cfg["coultype"],
                   "precision": args.precision, "n_omega": cfg["numbse"],
                   **({"temperature_K": cfg["temp"], "ta": cfg["ta"], "st": cfg["st"], "phavg": cfg["phavg"]}
                      if prm.bset else {})},
        "fp64_plain": plain,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu1,
        "stage_ms": {k: round(v, 3) for k, v in stage.items()},
        "checks": checks,
        "iterative_form": iterative,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
