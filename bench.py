"""Benchmark: RK-stage DOF-updates/s of the FR residual + SSP-RK3 on one MI355X.

Workload (the headline configuration): nodal FR with the DG correction
function on Gauss-Legendre points, degree p = 3 (4^3 solution points per
cell), Roe flux, SSP-RK3, on a periodic 64^3 hexahedral box [0,1]^3
(262 144 cells, 16.8 M DOFs x 5 variables), FP64 throughout.  Initial state:
rho = 1 + 0.2 sin(2 pi (x+y+z)) + 8 seeded Fourier modes with |k| <= 4 and
amplitude <= 1e-3, u = (1, 0.5, 0.25), p = 1; dt from CFL 0.1.

A "step" is one SSP-RK3 time step = 3 fused residual+stage launches.
value = 3 * n_cells * N / (seconds per step)   [RK-stage DOF-updates/s]

The roofline object prices the stage kernel against HBM with the algorithmic
byte count B_stage = n_cells * (5 * N * 8 * m + 4 * 6): m = 2 for stage 1
(read U0, write U1), 3 for stages 2 and 3 (read U_s, read U0, write U_next),
plus one 4-byte neighbour index per face.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GAMMA = 1.4


def initial_state(x: np.ndarray, seed: int = 2024) -> np.ndarray:
    rng = np.random.default_rng(seed)
    s = x[..., 0] + x[..., 1] + x[..., 2]
    rho = 1 + 0.2 * np.sin(2 * np.pi * s)
    modes = 0
    while modes < 8:
        k = rng.integers(-4, 5, size=3)
        if not (0 < np.linalg.norm(k) <= 4):
            continue
        rho += rng.uniform(-1e-3, 1e-3) * np.sin(2 * np.pi * (x @ k.astype(np.float64)) + rng.uniform(0, 2 * np.pi))
        modes += 1
    vel = (1.0, 0.5, 0.25)
    u = np.empty(rho.shape + (5,))
    u[..., 0] = rho
    for d in range(3):
        u[..., 1 + d] = rho * vel[d]
    u[..., 4] = 1.0 / (GAMMA - 1) + 0.5 * rho * sum(v * v for v in vel)
    return u


def stage_bytes(n_cells: int, n_per_cell: int) -> tuple[float, float]:
    per_cell = lambda m: 5 * n_per_cell * 8 * m + 4 * 6
    return n_cells * per_cell(2), n_cells * per_cell(3)


def read_traffic(name: str):
    """HBM bytes per stage launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(n: int, degree: int) -> dict:
    """The oracle (test infrastructure) timed on the host, on the headline
    workload itself: FR p=degree, Roe, periodic n^3 box, the same initial
    state and time step.

    One SSP-RK3 step with the residual and the RK combination spread over the
    host's threads (at most 16; the threaded oracle step equals the serial one
    bit for bit), and one serial residual for the single-core figure.  The
    oracle's set-up (connectivity, metrics) is not timed.  ctypes releases
    the GIL during the oracle's C++ calls.
    """
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from pvcriemannsolvers_amd import meshgen
    mesh = meshgen.box_hex(n, n, n, periodic=(True, True, True))
    o = ol.FrOracle(mesh, q=degree + 1, riemann=2)
    u = initial_state(o.node_coords()).reshape(-1)
    dt = o.time_step(u, 0.1)
    dofs = n ** 3 * (degree + 1) ** 3
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    o.rk3_threaded(u, dt, n_steps=1, threads=threads)
    secp = time.perf_counter() - t0
    sec1 = o.time_residuals(u, 1)
    del o
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": 3 * dofs / secp, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle FR p={degree} Roe SSP-RK3 on the periodic {n}^3 box itself ({n ** 3} cells, "
                      f"{dofs} DOFs per variable): one step on {threads} threads in {secp:.1f} s; "
                      f"one residual on one thread in {sec1:.1f} s",
            "single_core_value": dofs / sec1, "host_cpus": os.cpu_count(), "cpu_model": model}


def l2_linf(a, b, w):
    """L2 (sqrt of sum of w_q du^2, the mesh/part.hpp:1022-1036 pattern) and Linf of rho and E."""
    out = {}
    for k, name in ((0, "rho"), (4, "E")):
        d = a[..., k] - b[..., k]
        out["l2_" + name] = float(np.sqrt((w * d * d).sum()))
        out["linf_" + name] = float(np.abs(d).max())
    return out


def gl_weights(q: int, h: float) -> np.ndarray:
    x1, w1 = np.polynomial.legendre.leggauss(q)
    return np.einsum("i,j,k->ijk", w1, w1, w1).reshape(-1) * (h / 2) ** 3


def bench_mixed(n: int = 48, degree: int = 4, steps: int = 100) -> dict:
    """Mixed vs FP64 FR on the smooth periodic problem (48^3, p = 4, Roe)."""
    from pvcriemannsolvers_amd import meshgen, spatial
    mesh = meshgen.box_hex(n, n, n, periodic=(True, True, True))
    states, times = {}, {}
    for prec in ("fp64", "mixed"):
        g = spatial.FiniteElement(mesh, degree=degree, riemann="roe", precision=prec)
        x = g.sample_points()
        u0 = initial_state(x)
        del x
        g.set_solution_column(u0.reshape(-1))
        dt = g.get_time_step(0.1)
        for s in range(steps):
            g.rk3_update(s * dt, dt)
        g.synchronize()
        states[prec] = g.get_solution_column().reshape(u0.shape)
        _, kern = g.time_rk3(dt, 10)
        times[prec] = kern / 30
        g.close()
    w = gl_weights(degree + 1, 1.0 / n)[None, :]
    dofs = n ** 3 * (degree + 1) ** 3
    out = {"workload": f"FR p={degree} Roe periodic {n}^3, {steps} SSP-RK3 steps at CFL 0.1",
           "fp64_dof_updates_per_s": dofs / (times["fp64"] * 1e-3),
           "mixed_dof_updates_per_s": dofs / (times["mixed"] * 1e-3),
           "fp64_stage_ms": times["fp64"], "mixed_stage_ms": times["mixed"]}
    out.update({"mixed_vs_fp64_" + k: v for k, v in l2_linf(states["mixed"], states["fp64"], w).items()})
    return out


def tube_cases(n_cases: int = 256, seed: int = 7):
    """Seeded 1-D-in-3-D shock tubes: N x 2 x 2 cubic hexahedra of edge 1/N
    on [0,1]x[0,2/N]^2 (cubic cells keep the reference's length scale
    cbrt(V)/2 a stable CFL measure), N ~ U{100..2000}, p ~ U{1,2,3};
    rho ~ U[0.1,2], u ~ U[-1,1], p ~ U[0.1,10] per side; smart boundaries with
    the side states at the x ends, inviscid walls elsewhere; the jump is
    smoothed over two cells (tanh)."""
    from pvcriemannsolvers_amd import meshgen
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(n_cases):
        n = int(rng.integers(100, 2001))
        p = int(rng.integers(1, 4))
        rho, u, pr = rng.uniform(0.1, 2, 2), rng.uniform(-1, 1, 2), rng.uniform(0.1, 10, 2)
        sides = []
        for i in range(2):
            sides.append(np.array([rho[i], rho[i] * u[i], 0, 0, pr[i] / (GAMMA - 1) + 0.5 * rho[i] * u[i] ** 2]))
        mesh = meshgen.box_hex(n, 2, 2, hi=(1.0, 2.0 / n, 2.0 / n))
        bcs = [("smart", sides[0]), ("smart", sides[1])] + [("inviscid_wall", None)] * 4
        cases.append((mesh, p, bcs, sides, n))
    return cases


def bench_batched(n_cases: int = 256, steps: int = 20) -> dict:
    """256 shock tubes: batched (persistent launch + graph) in FP64 and mixed,
    against one plain FP64 solver per case."""
    from pvcriemannsolvers_amd import spatial
    cases = tube_cases(n_cases)
    spec = [(m, p, bcs) for m, p, bcs, _, _ in cases]
    out = {"workload": f"{n_cases} seeded shock tubes N x 2 x 2 (cubic cells), N in [100, 2000], p in {{1,2,3}}, HLLC, "
                       f"SSP-RK3 {steps} steps, per-case dt at CFL 0.1"}
    finals = {}
    for prec in ("fp64", "mixed"):
        b = spatial.Batch(spec, riemann="hllc", precision=prec)
        dofs = 0
        for i, (m, p, bcs, sides, n) in enumerate(cases):
            x = b.sample_points(i)
            w = 0.5 * (1 + np.tanh((x[..., 0] - 0.5) * n / 2.0))
            b.set_solution_column(i, (1 - w)[..., None] * sides[0] + w[..., None] * sides[1])
            dofs += x.shape[0] * x.shape[1]
        dt = b.get_time_steps(0.1)
        b.set_time_steps(dt)
        case_dt = [float(v) for v in dt]
        tiles, mode = b.autotune(repeats=3)
        b.rk3_steps(steps)
        b.synchronize()
        finals[prec] = [b.get_solution_column(i) for i in range(n_cases)]
        ms = b.time_rk3(10) / 10
        out[prec + "_batched_dof_updates_per_s"] = 3 * dofs / (ms * 1e-3)
        out[prec + "_batched_ms_per_step"] = ms
        out[prec + "_tiles"] = [int(t) for t in tiles if t]
        out[prec + "_launch"] = "persistent" if mode == 1 else "per degree group"
        out["dofs"] = dofs
        out["cells"] = int(sum(m["cells"].shape[0] for m, _, _, _, _ in cases))
        b.close()
    errs = []
    for i, (m, p, bcs, sides, n) in enumerate(cases):
        a = finals["mixed"][i].reshape(-1, (p + 1) ** 3, 5)
        r = finals["fp64"][i].reshape(-1, (p + 1) ** 3, 5)
        w = gl_weights(p + 1, 1.0 / n)[None, :]
        errs.append(l2_linf(a, r, w)["l2_rho"])
    out["mixed_vs_fp64_l2_rho_max"] = float(np.max(errs))
    out["mixed_vs_fp64_l2_rho_median"] = float(np.median(errs))
    out["all_finite"] = bool(all(np.isfinite(f).all() for f in finals["fp64"]))
    # plain FP64: one solver per case (three launches per step and case) on
    # the same 256 cases, initial states, per-case time steps and step count
    plain = []
    for i, (m, p, bcs, sides, n) in enumerate(cases):
        g = spatial.FiniteElement(m, degree=p, riemann="hllc", bcs=bcs)
        x = g.sample_points()
        w = 0.5 * (1 + np.tanh((x[..., 0] - 0.5) * n / 2.0))
        g.set_solution_column(((1 - w)[..., None] * sides[0] + w[..., None] * sides[1]).reshape(-1))
        plain.append(g)
    for g in plain:
        g.synchronize()
    t0 = time.perf_counter()
    for g, dt in zip(plain, case_dt):
        for s in range(steps):
            g.rk3_update(s * dt, dt)
    for g in plain:
        g.synchronize()
    sec = (time.perf_counter() - t0) / steps
    out["fp64_plain_dof_updates_per_s"] = 3 * out["dofs"] / sec
    out["fp64_plain_ms_per_step"] = sec * 1e3
    out["plain_sample"] = f"all {n_cases} cases, one solver each, same initial states, time steps and {steps} steps"
    out["plain_vs_batched_fp64_max_abs_diff"] = float(max(np.abs(g.get_solution_column() - f).max()
                                                          for g, f in zip(plain, finals["fp64"])))
    for g in plain:
        g.close()
    return out


def cons_state(rho, vel, p):
    rho = np.asarray(rho, dtype=np.float64)
    u = np.zeros(rho.shape + (5,))
    u[..., 0] = rho
    for d in range(3):
        u[..., 1 + d] = rho * vel[d]
    u[..., 4] = p / (GAMMA - 1) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return u


def dg_stage_bytes(n_cells: int, n_basis: int, faces_per_cell: int) -> float:
    """Mean algorithmic bytes of one RK stage (stage 1 reads U and writes U1,
    stages 2 and 3 also read U0) plus one int32 neighbour index per face."""
    col = n_cells * n_basis * 5 * 8
    return (2 * col + 2 * 3 * col) / 3 + 4 * faces_per_cell * n_cells


def bench_dg_sod(t_end: float = 0.2) -> dict:
    """Modal DG p = 2 on the shuffled 40 x 4 x 4 Sod tube (exact solver, smart
    ends, inviscid walls), CFL 0.1 time step evaluated every step, to t_end."""
    from pvcriemannsolvers_amd import meshgen, spatial
    mesh = meshgen.box_hex(40, 4, 4, hi=(1.0, 0.1, 0.1), shuffle_seed=2023)
    left, right = cons_state(1.0, (0, 0, 0), 1.0), cons_state(0.125, (0, 0, 0), 0.1)
    bcs = [("smart", left), ("smart", right)] + [("inviscid_wall", None)] * 4
    g = spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="exact", bcs=bcs)
    x = g.sample_points()
    g.approximate(np.where((x[..., 0] < 0.5)[..., None], left, right))
    g.synchronize()
    t, steps = 0.0, 0
    t0 = time.perf_counter()
    while t < t_end:
        dt = g.get_time_step(0.1, t_end - t)
        g.rk3_update(t, dt)
        t += dt
        steps += 1
    g.synchronize()
    wall = time.perf_counter() - t0
    state = g.get_solution_column().reshape(g.n_cells, g.n_per_cell, 5)
    _, kern, st = g.time_rk3(1e-6, 20, per_stage=True)
    dofs = g.n_cells * g.n_per_cell
    out = {"workload": "modal DG p=2 (3^3 Gauss points), exact Riemann, Sod tube 40x4x4 shuffled hexahedra, "
                       f"CFL 0.1 to t={t_end}",
           "steps": steps, "wall_s": wall, "dof_updates_per_s": 3 * steps * dofs / wall,
           "stage_kernels_ms_median": float(np.median(st)),
           "all_finite": bool(np.isfinite(state).all()),
           "rho_mean": float(state[:, 0, 0].mean())}
    g.close()
    return out


def quadrant_states(x: np.ndarray, seed: int = 2024) -> np.ndarray:
    """Four constant quadrants of [0, 1]^2: rho ~ U[0.1, 1.5], p ~ U[0.03, 1.5],
    (u, v) ~ U[-1, 1]^2 per quadrant, seeded."""
    rng = np.random.default_rng(seed)
    rho, p = rng.uniform(0.1, 1.5, 4), rng.uniform(0.03, 1.5, 4)
    vel = rng.uniform(-1, 1, (4, 2))
    q = (x[..., 0] > 0.5).astype(int) + 2 * (x[..., 1] > 0.5).astype(int)
    return cons_state(rho[q], (vel[q, 0], vel[q, 1], 0.0 * q), p[q])


def bench_dg_tets(steps: int = 30, t_end: float = 0.25, seed: int = 22) -> dict:
    """Modal DG p = 2 with the WENO limiter on 96 x 96 x 4 hexahedra split into
    6 tetrahedra each (interior vertices jittered by +-15 % of h, seeded),
    HLLC, seeded four-quadrant Riemann problem.  The x and y sides are smart
    boundaries holding the initial state of the adjacent quadrant at each
    boundary point (a boundary function of the point only, so its values are
    set once), the z sides inviscid walls.  Start-up as the reference's driver on tetrahedra
    (Approximate, then a second limiter pass); a full run to t_end with the
    CFL 0.1 time step of every step, then `steps` steps at the last time step
    timed per stage."""
    from pvcriemannsolvers_amd import meshgen, spatial
    mesh = meshgen.box_tet(96, 96, 4, hi=(1.0, 1.0, 0.04), jitter=0.15, seed=2024)
    bcs = [("smart", [1.0, 0.0, 0.0, 0.0, 2.5])] * 4 + [("inviscid_wall", None)] * 2
    g = spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="hllc", bcs=bcs, limiter="weno")
    for tag in range(4):  # functions of the point only: evaluated once
        g.set_boundary_values(tag, quadrant_states(g.boundary_points(tag), seed))
    g.approximate(quadrant_states(g.sample_points(), seed))
    g.set_solution_column(g.get_solution_column())
    g.synchronize()
    t, n = 0.0, 0
    t0 = time.perf_counter()
    while t < t_end:
        dt = g.get_time_step(0.1, t_end - t)
        g.rk3_update(t, dt)
        t += dt
        n += 1
    g.synchronize()
    wall = time.perf_counter() - t0
    finite = bool(np.isfinite(g.get_solution_column()).all())
    dofs = g.n_cells * g.n_per_cell
    dt = g.get_time_step(0.1)
    total, kern, st = g.time_rk3(dt, steps, per_stage=True)
    flags = g.limiter_flags()
    stage_ms = float(np.median(st))
    achieved = dg_stage_bytes(g.n_cells, g.n_per_cell, 4) / (kern / (3 * steps) * 1e-3) / 1e9
    out = {"workload": "modal DG p=2 + WENO (weno::Eigen detector, weno::Lazy reconstruction), HLLC, "
                       f"{g.n_cells} jittered tetrahedra (96x96x4 hex x 6), four-quadrant Riemann problem "
                       f"(seed {seed}), smart x/y sides holding the quadrant states, walls in z",
           "cells": g.n_cells, "dofs": dofs,
           "t_end": t_end, "steps_to_t_end": n, "wall_s_to_t_end": wall,
           "dof_updates_per_s_to_t_end": 3 * n * dofs / wall, "all_finite": finite,
           "dt": dt, "steps_timed": steps,
           "dof_updates_per_s": 3 * steps * dofs / (total * 1e-3),
           "stage_ms_median": stage_ms, "flagged_fraction": float(flags.mean()),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4)}}
    g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--n", type=int, default=64, help="cells per direction")
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-forms", action="store_true", help="skip the mixed, batched and modal-DG forms")
    ap.add_argument("--dg-only", action="store_true", help="run only the modal-DG forms (for profiling)")
    ap.add_argument("--fr-forms-only", action="store_true",
                    help="run only the mixed-precision and batched FR forms (for profiling)")
    ap.add_argument("--dg-t-end", type=float, default=0.25,
                    help="final time of the tetrahedral DG + WENO run (0.25 is the workload's own)")
    args = ap.parse_args()
    if args.dg_only:
        print(json.dumps({"dg_sod": bench_dg_sod(), "dg_tets_weno": bench_dg_tets(t_end=args.dg_t_end)}), flush=True)
        return
    if args.fr_forms_only:
        print(json.dumps({"mixed": bench_mixed(), "batched": bench_batched()}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from pvcriemannsolvers_amd import meshgen, spatial

    n, p = args.n, args.degree
    mesh = meshgen.box_hex(n, n, n, periodic=(True, True, True))
    g = spatial.FiniteElement(mesh, degree=p, riemann="roe")
    x = g.sample_points()
    u0 = initial_state(x)
    del x
    g.set_solution_column(u0.reshape(-1))
    del u0
    dt = g.get_time_step(0.1)
    print(f"[bench] {n}^3 cells, p={p}, dt={dt:.4e}; warmup {args.warmup} steps", file=sys.stderr, flush=True)
    for s in range(args.warmup):
        g.rk3_update(s * dt, dt)
    g.synchronize()
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    total_ms, kernel_ms, stage_ms = g.time_rk3(dt, args.steps, per_stage=True)
    g.synchronize()
    wall = time.perf_counter() - t0
    state = g.get_solution_column()
    if not np.isfinite(state).all():
        raise RuntimeError("non-finite state after the timed steps")

    n_cells, npc = g.n_cells, g.n_per_cell
    ms_per_step = wall * 1e3 / args.steps
    value = 3 * n_cells * npc / (ms_per_step * 1e-3)
    b1, b23 = stage_bytes(n_cells, npc)
    avg_bytes = (b1 + 2 * b23) / 3
    avg_kernel_s = kernel_ms * 1e-3 / (3 * args.steps)
    achieved = avg_bytes / avg_kernel_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(f"fr_hex_stage_{n}cubed_p{p}_hbm_traffic.json"),
                "kernel": "FrHexStageKernel (Q=%d, Roe, FP64)" % (p + 1),
                "avg_kernel_ms": round(avg_kernel_s * 1e3, 4),
                "stage_ms_p10_median_p90": [round(float(v), 4) for v in np.percentile(stage_ms, [10, 50, 90])],
                "stages_timed": int(stage_ms.size),
                "algorithmic_bytes_per_launch": int(avg_bytes)}
    forms = None
    g.close()
    if not args.no_forms:
        print("[bench] mixed vs FP64 (48^3, p=4)", file=sys.stderr, flush=True)
        forms = {"mixed": bench_mixed()}
        print("[bench] batched shock tubes", file=sys.stderr, flush=True)
        forms["batched"] = bench_batched()
        print("[bench] modal DG Sod tube", file=sys.stderr, flush=True)
        forms["dg_sod"] = bench_dg_sod()
        print("[bench] modal DG + WENO on tetrahedra", file=sys.stderr, flush=True)
        forms["dg_tets_weno"] = bench_dg_tets(t_end=args.dg_t_end)
    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(n, p)
    line = {
        "metric": "RK-stage DOF-updates/s, mixed vs FP64 at stated L2 error; % HBM roofline",
        "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"FR p={p} Gauss-Legendre, Roe, SSP-RK3, periodic {n}^3 hex box",
                   "cells": n_cells, "nodes_per_cell": npc, "dofs_per_variable": n_cells * npc,
                   "precision": "fp64", "rk_stages_per_step": 3, "dt": dt,
                   "event_ms_per_step": total_ms / args.steps},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
