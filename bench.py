#!/usr/bin/env python3
"""Benchmark: cell-updates/s per SSP-RK3 stage of the WENO-4 HLLC residual
path on one MI355X, FP64 and mixed FP32/FP64, and the adjoint's cost.

A "step" is one SSP-RK3 time step (three residual evaluations + stage
updates) over the whole mesh; the state, the PRESTORE arrays and the face
data are resident in HBM before the timed region starts.  `value` is the
FP64 path on the default workload (64^3 hexes split into 6 jittered
tetrahedra each, 1,572,864 cells); the mixed path runs on the same mesh and
is reported beside it.  The same line carries a block for the 128^3-hex
WENO-4 mesh (FP64, mixed, and the adjoint over a march long enough that the
kept reconstructions run out of their memory budget) and one for the
507,600-cell mixed-element channel (WENO-3, FP64 and mixed, the adjoint over
its 50-step march).  `--config c5` or `--extra c5` measures the WENO-5 CWENO
blast box (FP64 and mixed) at `--c5-n` hexes per axis.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c3|c1|c4|c5] [--extra c3,c4]
                       [--c5-n N]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "cell-updates/s per RK stage (WENO-4 HLLC) FP64 vs mixed; adjoint/primal cost"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def face_points(case) -> np.ndarray:
    """Gauss points per cell (qp_quad on quadrilaterals, qp_tri on triangles)"""
    nv = (case.array("face_nodes").reshape(case.ncell, -1, 4) >= 0).sum(2)
    return np.where(nv == 4, case.qp_quad, np.where(nv == 3, case.qp_tri, 0)).sum(1)


def recon_bytes_per_cell(case, precision: int) -> float:
    """Algorithmic bytes one cell costs the reconstruction kernel, averaged
    over the cells: stencil pseudo-inverses, packed indicator, stencil
    indices, face Gauss points, basis means and the cell's own state
    (neighbour re-reads and the ULEFT round trip are implementation
    overheads and not counted)."""
    b = 4 if precision == 1 else 8
    adm = case.array("admis").astype(np.float64).mean()
    nqp = face_points(case).mean()
    if case.spec.ees == 5:
        # CWENO: the central stencil's pseudo-inverse, the sectors' low-order
        # ones over their first imax2 members, both indicators, mbar and mbar - mbar_c
        cw = case.cweno()
        id2, im2 = cw.idegfree2, cw.inumneighbours2 - 1
        by = case.imax * case.ideg * b + (adm - 1) * im2 * id2 * b
        by += case.ideg * (case.ideg + 1) / 2 * b + id2 * (id2 + 1) / 2 * b
        by += (case.imax + (adm - 1) * im2) * 4 + nqp * 3 * b + (case.ideg + id2) * b + 5 * 8
        return float(by)
    by = adm * case.imax * case.ideg * b
    if case.iweno == 1:
        by += case.ideg * (case.ideg + 1) / 2 * b
    by += adm * case.imax * 4 + nqp * 3 * b + case.ideg * b + 5 * 8
    return float(by)


def recon_kernel_name(case, precision: int) -> str:
    """the reconstruction kernel the library launches: the row-stream
    kernels for 5 or 7 stencils (tetrahedra, hexahedra), the
    register-prefetch ones for other stencil counts"""
    ts = int(case.array("admis").max())
    nv = (case.array("face_nodes").reshape(case.ncell, -1, 4) >= 0).sum(2)
    nf = int((nv > 0).sum(1).max())
    nqm = max(case.qp_quad if (nv == 4).any() else 0, case.qp_tri if (nv == 3).any() else 0)
    nfq = nf * nqm
    if case.ideg == 34:  # WENO-5: the per-cell kernel (CWENO: its compact-layout variant)
        r = "float" if precision else "double"
        return f"k_recon<{r},34,69,7{',true,9,19' if case.spec.ees == 5 else ''}>"
    stream = ((ts == 5 and nfq <= 24) or (ts == 7 and nfq > 24)) if case.ideg == 19 else (ts in (5, 7) and nfq <= 24)
    nql = 3 if nfq <= 24 else 7
    if stream:
        return f"k_recon_{'s32' if precision else 's64'}<{case.ideg},{case.imax},{nql},{ts}>"
    return f"k_recon_{'w32' if precision else 'w64'}<{case.ideg},{case.imax},{nql}>"


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


def make_solver(case, precision: int, **opts):
    from ucns3d_amd.solver import Solver
    s = Solver(case, precision, **opts)
    if case.spec.mach > 0:  # the channel: characteristic inflow/outflow at the free stream
        fs = case.spec.freestream()
        s.set_boundary(1, fs, fs)
    elif any(b != 0 for b in case.spec.bc):  # outflow sides with BOUNDTYPE 0: the ghosts copy the cell
        amb = np.array([1.0, 0.0, 0.0, 0.0, 1.0 / (case.spec.gamma - 1.0)])
        s.set_boundary(0, amb, amb)
    return s


def run_path(case, precision: int, steps: int, warmup: int, cfl: float, u0: np.ndarray):
    import torch

    t0 = time.time()
    s = make_solver(case, precision)
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
    # sum of the residual over all cells relative to its absolute sum: round-off on a closed (periodic)
    # domain; with inflow/outflow faces it is the net boundary flux, so it is not reported there
    cons = None
    if all(b == 0 for b in case.spec.bc):
        cons = float(np.abs(rhs.sum(axis=0)).max() / max(np.abs(rhs).sum(axis=0).max(), 1e-300))
    return {"elapsed": elapsed, "t_sim": t_sim, "u": u, "stats": st, "conservation": cons}


def path_block(case, key: str, precision: int, r: dict, steps: int) -> dict:
    """throughput, kernel times and the reconstruction kernel's roofline"""
    st = r["stats"]
    value = case.ncell * 3 * steps / r["elapsed"]
    rk_ms = st["ms_recon"] / max(st["n_recon"], 1)
    by = recon_bytes_per_cell(case, precision)
    achieved = by * case.ncell / (rk_ms * 1e-3) / 1e9
    kernel = recon_kernel_name(case, precision)
    return {"value": value, "ms_per_step": r["elapsed"] / steps * 1e3,
            "kernels_ms": {"recon": rk_ms, "flux": st["ms_flux"] / max(st["n_flux"], 1),
                           "update": st["ms_update"] / max(st["n_update"], 1)},
            "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": committed_traffic(kernel, key), "algorithmic_bytes": by * case.ncell,
                         "bytes_per_cell": by, "kernel_ms": rk_ms},
            "step_roofline_frac": by * case.ncell * 3 / (r["elapsed"] / steps) / (HBM_PEAK_GBS * 1e9)}


def run_adjoint(case, precision: int, steps: int, cfl: float, u0: np.ndarray) -> dict:
    """adjoint / primal cost: ucns_rk3_adjoint_device over `steps` steps (its
    forward sweep with stage checkpoints plus the reverse sweep) against
    ucns_rk3_steps over the same steps, device-resident cotangents.  Timed
    twice: with the library's default, which keeps the forward
    reconstruction of as many of the last stages as fit in its memory
    budget (64 GiB), and with none kept (a context with adj_keep_stages < 0:
    every reversed stage runs its reconstruction again)."""
    import torch

    s = make_solver(case, precision)
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
    s = make_solver(case, precision, adj_keep_stages=-1)
    s.set_state(u0)
    s.rk3_adjoint_device(steps, a.data_ptr(), u0b.data_ptr(), dt_fixed=dt)  # allocates checkpoints
    t_rec = adjoint()
    s.set_state(u0)
    s.reset_stats()
    s.rk3_adjoint_device(1, a.data_ptr(), u0b.data_ptr(), dt_fixed=dt)
    st = s.stats()
    s.close()
    return {"steps": steps, "precision": "mixed" if precision else "fp64",
            "primal_ms_per_step": t_fwd / steps * 1e3, "adjoint_ms_per_step": t_adj / steps * 1e3,
            "adjoint_over_primal": t_adj / t_fwd,
            "recompute_adjoint_ms_per_step": t_rec / steps * 1e3, "recompute_adjoint_over_primal": t_rec / t_fwd,
            "recon_vjp_ms": st["ms_recon_vjp"] / max(st["n_recon_vjp"], 1),
            "flux_vjp_ms": st["ms_flux_vjp"] / max(st["n_flux_vjp"], 1),
            "recon_adj_ms": st["ms_recon_adj"] / max(st["n_recon_adj"], 1),
            "recon_ms": st["ms_recon"] / max(st["n_recon"], 1), "finite": finite,
            "note": "adjoint = forward sweep storing every stage state (and keeping the last stages' "
                    "reconstructions within the memory budget) + reverse sweep (each stage: flux product, "
                    "reconstruction product, transposed scatter); recompute = every reversed stage runs "
                    "its reconstruction again; dt fixed"}


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def cpu_baseline(case, u0: np.ndarray, n1: int, n2: int):
    """The CPU oracle (C++ restatement of the reference algorithm, OpenMP on
    the host's cores) on the same mesh and state as the GPU run, timed by
    differencing: (t(n2 steps) - t(n1 steps)) / (3 (n2 - n1)) per stage,
    which leaves out set-up.  The steps run in one sequence, so t(n1) is a
    prefix of t(n2)."""
    from oracle.oracle import Oracle

    orc = Oracle(case)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    dt = orc.dt(u0, case.spec.cfl)
    u = u0
    t0 = time.perf_counter()
    t_n1 = None
    for k in range(1, n2 + 1):
        u, _ = orc.rk3(u, 1, dt_fixed=dt)
        if k == n1:
            t_n1 = time.perf_counter() - t0
    t_n2 = time.perf_counter() - t0
    per_stage = (t_n2 - t_n1) / (3 * (n2 - n1))
    return {"value": case.ncell / per_stage, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{case.spec.name}: {case.ncell} cells (the GPU run's mesh and state), steps {n1 + 1}-{n2} "
                      f"of a {n2}-step march timed ({t_n2 - t_n1:.1f} s of {t_n2:.1f} s), oracle/ucns_oracle.cpp "
                      f"with OpenMP"}


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

    fp = make_solver(case, UCNS_FP64)
    mx = make_solver(case, UCNS_MIXED)
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
            from oracle import oracle
            from oracle.oracle import Oracle
            t0 = time.time()
            if case.spec.mach > 0:
                fs = case.spec.freestream()
                oracle.set_boundary(1, fs, fs)
            ref, _ = Oracle(case).rk3(u0, 1, dt_fixed=dt)
            rec["fp64_vs_oracle"] = rel_errors(b, ref)
            rec["mixed_vs_oracle"] = rel_errors(a, ref)
            log(f"  oracle step: {time.time() - t0:.1f}s")
        out["steps"][str(n)] = rec
        log(f"step {n}: {rec}")
    fp.close()
    mx.close()
    return out


def config_block(key: str, steps: int, warmup: int, adjoint_steps: int) -> dict:
    """FP64 and mixed throughput on one configuration, the mixed path's error
    after the march, and the adjoint's cost (FP64)"""
    from ucns3d_amd.case import PRESETS, Case
    from ucns3d_amd.native import UCNS_FP64, UCNS_MIXED

    spec = spec_of(key)
    t0 = time.time()
    case = Case(spec)
    u0 = case.init_state()
    log(f"{spec.name}: {case.ncell} cells, preprocessing {time.time() - t0:.1f}s")
    fp = run_path(case, UCNS_FP64, steps, warmup, spec.cfl, u0)
    mx = run_path(case, UCNS_MIXED, steps, warmup, spec.cfl, u0)
    b64 = path_block(case, key, UCNS_FP64, fp, steps)
    bmx = path_block(case, key, UCNS_MIXED, mx, steps)
    a, b = mx["u"], fp["u"]
    err = (np.sqrt(((a - b) ** 2).sum(0)) / np.sqrt((b ** 2).sum(0))).tolist()
    out = {"workload": spec.name, "ncell": case.ncell, "steps": steps, "warmup": warmup, "fp64": b64, "mixed": bmx,
           "mixed_speedup_vs_fp64": bmx["value"] / b64["value"],
           "mixed_rel_l2_vs_fp64": err, "steps_compared": warmup + steps}
    log(f"{key}: fp64 {b64['value'] / 1e6:.1f} M/s ({b64['roofline']['frac']:.3f} of HBM in the reconstruction), "
        f"mixed {bmx['value'] / 1e6:.1f} M/s ({bmx['roofline']['frac']:.3f}), x{out['mixed_speedup_vs_fp64']:.2f}")
    if adjoint_steps > 0:
        out["adjoint"] = run_adjoint(case, UCNS_FP64, adjoint_steps, spec.cfl, u0)
        log(f"{key} adjoint: {out['adjoint']}")
    return out


C5_N = 121


def spec_of(key: str):
    """the configuration's spec; C5 at --c5-n hexes per axis (the blast box's
    WENO-5 preprocessing runs at about 0.3 ms per cell on 8 cores, so its full
    121^3 takes most of an hour)"""
    from ucns3d_amd.case import PRESETS, reduced
    if key == "c5" and C5_N != PRESETS["c5"].n:
        return reduced("c5", C5_N)
    return PRESETS[key]


def main() -> int:
    global C5_N
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--c5-n", type=int, default=40,
                    help="hexes per axis of the C5 blast box (121: the configuration's 8.1M cells)")
    ap.add_argument("--extra", default="c3,c4",
                    help="comma-separated further configurations reported in blocks of their own ('' for none)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", default="1,3",
                    help="N1,N2: the CPU oracle's per-stage time is (t(N2) - t(N1)) / (3 (N2 - N1)) steps")
    ap.add_argument("--adjoint-steps", type=int, default=10,
                    help="steps of the adjoint / primal cost measurement (0: skip); the channel uses its 50")
    ap.add_argument("--accuracy", default="",
                    help="comma-separated step counts: report the mixed path's per-variable relative L2 and Linf "
                         "against the FP64 path (and, at step 1, the CPU oracle) at each, then exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (the reference point for HBM byte counters)")
    args = ap.parse_args()
    C5_N = args.c5_n
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
    spec = spec_of(args.config)
    t0 = time.time()
    case = Case(spec)
    u0 = case.init_state()
    log(f"{spec.name}: {case.ncell} cells, preprocessing {time.time() - t0:.1f}s")

    if args.accuracy:
        rep = accuracy(case, u0, [int(x) for x in args.accuracy.split(",")])
        print(json.dumps({"workload": spec.name, "ncell": case.ncell, "accuracy": rep}), flush=True)
        return 0

    fp = run_path(case, UCNS_FP64, args.steps, args.warmup, spec.cfl, u0)
    blk = path_block(case, args.config, UCNS_FP64, fp, args.steps)
    log(f"fp64: {blk['ms_per_step']:.2f} ms/step, {blk['value'] / 1e6:.1f} M cell-updates/s/stage, recon kernel "
        f"{blk['kernels_ms']['recon']:.3f} ms ({blk['roofline']['achieved']:.0f} GB/s), "
        f"flux {blk['kernels_ms']['flux']:.3f} ms, update {blk['kernels_ms']['update']:.3f} ms")
    out = {
        "metric": METRIC, "value": blk["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": blk["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": spec.name, "ncell": case.ncell, "cfl": spec.cfl, "stages_per_step": 3,
                   "wenocentralweight": spec.lwci1, "seed": spec.seed, "jitter": spec.jitter},
        "roofline": blk["roofline"],
        "kernels_ms": blk["kernels_ms"],
        "step_roofline_frac": blk["step_roofline_frac"],
        "conservation_fp64": fp["conservation"],
    }
    if not args.no_mixed:
        mx = run_path(case, UCNS_MIXED, args.steps, args.warmup, spec.cfl, u0)
        bmx = path_block(case, args.config, UCNS_MIXED, mx, args.steps)
        a, b = mx["u"], fp["u"]
        err = (np.sqrt(((a - b) ** 2).sum(0)) / np.sqrt((b ** 2).sum(0))).tolist()
        log(f"mixed: {bmx['ms_per_step']:.2f} ms/step, {bmx['value'] / 1e6:.1f} M/s, recon "
            f"{bmx['kernels_ms']['recon']:.3f} ms ({bmx['roofline']['achieved']:.0f} GB/s), rel L2 vs fp64 after "
            f"{args.warmup + args.steps} steps {err}")
        out["mixed"] = {"value": bmx["value"], "ms_per_step": bmx["ms_per_step"],
                        "speedup_vs_fp64": bmx["value"] / blk["value"], "roofline_frac": bmx["roofline"]["frac"],
                        "recon_ms": bmx["kernels_ms"]["recon"], "bytes_per_cell": bmx["roofline"]["bytes_per_cell"],
                        "kernel": bmx["roofline"]["kernel"], "traffic": bmx["roofline"]["traffic"],
                        "step_roofline_frac": bmx["step_roofline_frac"],
                        "rel_l2_vs_fp64": err, "steps_compared": args.warmup + args.steps}
    if args.adjoint_steps > 0 and args.config != "c5":
        out["adjoint"] = run_adjoint(case, UCNS_FP64, args.adjoint_steps, spec.cfl, u0)
        log(f"adjoint: {out['adjoint']}")
    if not args.no_cpu:
        n1, n2 = (int(x) for x in args.cpu_steps.split(","))
        out["cpu_baseline"] = cpu_baseline(case, u0, n1, n2)
        log(f"cpu baseline: {out['cpu_baseline']}")
    del case
    for key in [k for k in args.extra.split(",") if k and k != args.config]:
        # (the adjoint does not cover WENO-5 / CWENO: none for c5)
        adj = 0 if args.adjoint_steps <= 0 or key == "c5" else (PRESETS[key].steps if key == "c4" else args.adjoint_steps)
        out[key] = config_block(key, min(args.steps, 10), 1, adj)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
