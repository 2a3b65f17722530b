#!/usr/bin/env python3
"""Benchmark: cell-updates/s per SSP-RK3 stage of the WENO-4 HLLC residual
path on one MI355X, FP64 and mixed FP32/FP64.

A "step" is one SSP-RK3 time step (three residual evaluations + stage
updates) over the whole mesh; the state, the PRESTORE arrays and the face
data are resident in HBM before the timed region starts.  `value` is the
FP64 path on the default workload (64^3 hexes split into 6 jittered
tetrahedra each, 1,572,864 cells); the mixed path runs on the same mesh and
is reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c3|c1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "cell-updates/s per RK stage (WENO-4 HLLC) FP64 vs mixed; adjoint/primal cost"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def recon_bytes_per_cell(case, precision: int) -> float:
    """Algorithmic bytes one cell costs the reconstruction kernel: stencil
    pseudo-inverses, packed indicator, stencil indices, face Gauss points,
    basis means and the cell's own state (neighbour re-reads and the ULEFT
    round trip are implementation overheads and not counted)."""
    b = 4 if precision == 1 else 8
    adm = case.array("admis").astype(np.float64).mean()
    faces = 6 if case.shape == 1 else 4
    nqp = faces * (case.qp_quad if case.shape == 1 else case.qp_tri)
    by = adm * case.imax * case.ideg * b
    if case.iweno == 1:
        by += case.ideg * (case.ideg + 1) / 2 * b
    by += adm * case.imax * 4 + nqp * 3 * b + case.ideg * b + 5 * 8
    return by


def recon_kernel_name(case, precision: int) -> str:
    """the reconstruction kernel the library launches for this case: the
    row-stream kernels for tetrahedra (5 stencils) and hexahedra (7)"""
    tet = case.shape == 2
    ts, nql = (5, 3) if tet else (7, 7)
    return f"k_recon_{'s32' if precision else 's64'}<19,39,{nql},{ts}>"


def committed_traffic(kernel: str, config: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by the copy calibration) and committed under
    profiles/ as 'traffic <kernel> <config> <bytes>' lines; None if absent"""
    d = os.path.join(REPO, "profiles")
    for fn in sorted(os.listdir(d)) if os.path.isdir(d) else []:
        if not fn.endswith(".txt"):
            continue
        for line in open(os.path.join(d, fn)):
            f = line.split()
            if len(f) == 4 and f[0] == "traffic" and f[1] == kernel and f[2] == config:
                return float(f[3])
    return None


def run_path(case, precision: int, steps: int, warmup: int, cfl: float, u0: np.ndarray):
    import torch
    from ucns3d_amd.solver import Solver

    t0 = time.time()
    s = Solver(case, precision)
    s.set_state(u0)
    log(f"  upload {'mixed' if precision else 'fp64'}: {time.time() - t0:.1f}s")
    if warmup > 0:
        s.rk3_steps(warmup, cfl=cfl)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    t_sim = s.rk3_steps(steps, cfl=cfl)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    u = s.get_state()
    # per-kernel durations with HIP events on the library stream (short extra run)
    s.reset_stats()
    s.rk3_steps(2, cfl=cfl)
    st = s.stats()
    rhs = s.residual()
    s.close()
    cons = float(np.abs(rhs.sum(axis=0)).max() / max(np.abs(rhs).sum(axis=0).max(), 1e-300))
    return {
        "elapsed": elapsed, "t_sim": t_sim, "u": u, "stats": st, "conservation": cons,
    }


def run_adjoint(case, precision: int, steps: int, cfl: float, u0: np.ndarray) -> dict:
    """adjoint / primal cost: ucns_rk3_adjoint_device over `steps` steps (its
    forward sweep with stage checkpoints plus the reverse sweep) against
    ucns_rk3_steps over the same steps, device-resident cotangents.  Timed
    twice: with every stage's forward reconstruction kept for the reverse
    sweep (the library's default, as many as fit in its memory budget) and
    with none kept (a context with adj_keep_stages < 0: every reversed stage
    runs its reconstruction again)."""
    import torch
    from ucns3d_amd.solver import Solver

    s = Solver(case, precision)
    s.set_state(u0)
    dt = s.compute_dt(cfl)
    g = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randn(case.ncell, 5, dtype=torch.float64, device="cuda", generator=g)
    u0b = torch.empty_like(a)
    s.rk3_steps(1, dt_fixed=dt)
    s.set_state(u0)
    s.rk3_adjoint_device(steps, a.data_ptr(), u0b.data_ptr(), dt_fixed=dt)  # allocates checkpoints, kept stages
    s.set_state(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.rk3_steps(steps, dt_fixed=dt)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0

    def adjoint() -> float:
        s.set_state(u0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.rk3_adjoint_device(steps, a.data_ptr(), u0b.data_ptr(), dt_fixed=dt)
        torch.cuda.synchronize()
        return time.perf_counter() - t0
    t_adj = adjoint()
    finite = bool(torch.isfinite(u0b).all().item())
    s.close()
    s = Solver(case, precision, adj_keep_stages=-1)
    s.set_state(u0)
    s.rk3_adjoint_device(steps, a.data_ptr(), u0b.data_ptr(), dt_fixed=dt)  # allocates checkpoints
    t_rec = adjoint()
    s.set_state(u0)
    s.reset_stats()
    s.rk3_adjoint_device(1, a.data_ptr(), u0b.data_ptr(), dt_fixed=dt)
    st = s.stats()
    s.close()
    return {"steps": steps, "primal_ms_per_step": t_fwd / steps * 1e3, "adjoint_ms_per_step": t_adj / steps * 1e3,
            "adjoint_over_primal": t_adj / t_fwd, "reverse_sweep_over_primal": (t_adj - t_fwd) / t_fwd,
            "recompute_adjoint_ms_per_step": t_rec / steps * 1e3, "recompute_adjoint_over_primal": t_rec / t_fwd,
            "recon_vjp_ms": st["ms_recon_vjp"] / max(st["n_recon_vjp"], 1),
            "flux_vjp_ms": st["ms_flux_vjp"] / max(st["n_flux_vjp"], 1),
            "recon_adj_ms": st["ms_recon_adj"] / max(st["n_recon_adj"], 1),
            "recon_ms": st["ms_recon"] / max(st["n_recon"], 1), "finite": finite,
            "note": "adjoint = forward sweep storing every stage state (and keeping each stage's "
                    "reconstruction) + reverse sweep (each stage: flux product, reconstruction product, "
                    "transposed scatter); recompute = the same with the reconstruction run again per "
                    "reversed stage; dt fixed"}


def cpu_baseline(spec_key: str, seconds: float):
    """The CPU oracle (C restatement of the reference algorithm, OpenMP) on a
    reduced mesh of the same construction, timed on this host."""
    from oracle.oracle import Oracle
    from ucns3d_amd.case import Case, reduced

    n = 24 if spec_key == "c2" else 40
    case = Case(reduced(spec_key, n))
    orc = Oracle(case)
    u = case.init_state()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and steps < 50:
        u, _ = orc.rk3(u, 1, cfl=case.spec.cfl)
        steps += 1
    el = time.perf_counter() - t0
    val = case.ncell * 3 * steps / el
    return {"value": val, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"{case.spec.name}: {case.ncell} cells, {steps} SSP-RK3 steps ({3 * steps} stages) in "
                      f"{el:.1f}s, oracle/ucns_oracle.cpp with OpenMP"}


def rel_errors(a: np.ndarray, b: np.ndarray) -> dict:
    """per conserved variable: relative L2 and relative Linf of a against b"""
    d = a - b
    l2 = np.sqrt((d ** 2).sum(0)) / np.sqrt((b ** 2).sum(0))
    linf = np.abs(d).max(0) / np.abs(b).max(0)
    return {"rel_l2": l2.tolist(), "rel_linf": linf.tolist()}


def accuracy(case, u0: np.ndarray, steps: list) -> dict:
    """FP64 and mixed contexts side by side from the same state with the same
    fixed dt (the FP64 CFL step of the initial state), compared at each
    requested step count; the CPU oracle joins at step 1."""
    from ucns3d_amd.native import UCNS_FP64, UCNS_MIXED
    from ucns3d_amd.solver import Solver

    fp = Solver(case, UCNS_FP64)
    mx = Solver(case, UCNS_MIXED)
    fp.set_state(u0)
    mx.set_state(u0)
    dt = fp.compute_dt(case.spec.cfl)
    out = {"dt": dt, "steps": {}}
    done = 0
    for n in sorted(steps):
        fp.rk3_steps(n - done, dt_fixed=dt)
        mx.rk3_steps(n - done, dt_fixed=dt)
        done = n
        a, b = mx.get_state(), fp.get_state()
        rec = {"mixed_vs_fp64": rel_errors(a, b)}
        if n == 1:
            from oracle.oracle import Oracle
            t0 = time.time()
            ref, _ = Oracle(case).rk3(u0, 1, dt_fixed=dt)
            rec["fp64_vs_oracle"] = rel_errors(b, ref)
            rec["mixed_vs_oracle"] = rel_errors(a, ref)
            log(f"  oracle step: {time.time() - t0:.1f}s")
        out["steps"][str(n)] = rec
        log(f"step {n}: {rec}")
    fp.close()
    mx.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3"])
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--adjoint-steps", type=int, default=2,
                    help="steps of the adjoint / primal cost measurement on the FP64 path (0: skip)")
    ap.add_argument("--accuracy", default="",
                    help="comma-separated step counts: report the mixed path's per-variable relative L2 and Linf "
                         "against the FP64 path (and, at step 1, the CPU oracle) at each, then exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (the reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs the whole mesh as one partition on one GPU; "
                                   "--gpus must be 1 (the multi-GPU halo design is documented, not built)"}))
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0

    import torch
    from ucns3d_amd.case import PRESETS, Case
    from ucns3d_amd.native import UCNS_FP64, UCNS_MIXED

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        n = 1 << 28
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)
        return 0
    spec = PRESETS[args.config]
    t0 = time.time()
    case = Case(spec)
    u0 = case.init_state()
    log(f"{spec.name}: {case.ncell} cells, preprocessing {time.time() - t0:.1f}s")

    if args.accuracy:
        rep = accuracy(case, u0, [int(x) for x in args.accuracy.split(",")])
        print(json.dumps({"workload": spec.name, "ncell": case.ncell, "accuracy": rep}), flush=True)
        return 0

    fp = run_path(case, UCNS_FP64, args.steps, args.warmup, spec.cfl, u0)
    ms_step = fp["elapsed"] / args.steps * 1e3
    value = case.ncell * 3 * args.steps / fp["elapsed"]
    st = fp["stats"]
    rk_ms = st["ms_recon"] / max(st["n_recon"], 1)
    bytes_cell = recon_bytes_per_cell(case, UCNS_FP64)
    achieved = bytes_cell * case.ncell / (rk_ms * 1e-3) / 1e9
    log(f"fp64: {ms_step:.2f} ms/step, {value / 1e6:.1f} M cell-updates/s/stage, recon kernel {rk_ms:.3f} ms "
        f"({achieved:.0f} GB/s), flux {st['ms_flux'] / max(st['n_flux'], 1):.3f} ms, "
        f"update {st['ms_update'] / max(st['n_update'], 1):.3f} ms")
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": spec.name, "ncell": case.ncell, "cfl": spec.cfl, "stages_per_step": 3,
                   "wenocentralweight": spec.lwci1, "seed": spec.seed, "jitter": spec.jitter},
        "roofline": {"bound": "hbm", "kernel": recon_kernel_name(case, UCNS_FP64),
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": committed_traffic(recon_kernel_name(case, UCNS_FP64), args.config),
                     "algorithmic_bytes": bytes_cell * case.ncell, "bytes_per_cell": bytes_cell,
                     "kernel_ms": rk_ms},
        "kernels_ms": {"recon": rk_ms, "flux": st["ms_flux"] / max(st["n_flux"], 1),
                       "update": st["ms_update"] / max(st["n_update"], 1)},
        "conservation_fp64": fp["conservation"],
    }
    if not args.no_mixed:
        mx = run_path(case, UCNS_MIXED, args.steps, args.warmup, spec.cfl, u0)
        v_mx = case.ncell * 3 * args.steps / mx["elapsed"]
        stm = mx["stats"]
        rk_mx = stm["ms_recon"] / max(stm["n_recon"], 1)
        bytes_mx = recon_bytes_per_cell(case, UCNS_MIXED)
        ach_mx = bytes_mx * case.ncell / (rk_mx * 1e-3) / 1e9
        a, b = mx["u"], fp["u"]
        err = (np.sqrt(((a - b) ** 2).sum(0)) / np.sqrt((b ** 2).sum(0))).tolist()
        log(f"mixed: {mx['elapsed'] / args.steps * 1e3:.2f} ms/step, {v_mx / 1e6:.1f} M/s, recon {rk_mx:.3f} ms "
            f"({ach_mx:.0f} GB/s), rel L2 vs fp64 after {args.warmup + args.steps} steps {err}")
        out["mixed"] = {"value": v_mx, "ms_per_step": mx["elapsed"] / args.steps * 1e3, "speedup_vs_fp64": v_mx / value,
                        "roofline_frac": ach_mx / HBM_PEAK_GBS, "recon_ms": rk_mx, "bytes_per_cell": bytes_mx,
                        "kernel": recon_kernel_name(case, UCNS_MIXED),
                        "traffic": committed_traffic(recon_kernel_name(case, UCNS_MIXED), args.config),
                        "rel_l2_vs_fp64": err, "steps_compared": args.warmup + args.steps}
    if args.adjoint_steps > 0:
        out["adjoint"] = run_adjoint(case, UCNS_FP64, args.adjoint_steps, spec.cfl, u0)
        log(f"adjoint: {out['adjoint']}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.config, args.cpu_seconds)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
