"""Benchmark: particle-steps/s of the dompap per-step hot path on one MI355X.

Default workload: 3-D Lennard-Jones, eps = sigma = 1, truncated and shifted at
r_cut = 2.5, N = 1,048,576 on an FCC 64^3 lattice at density 0.8442, Gaussian
velocities at T = 0.72 (zero total momentum), Langevin G-JF with friction
alpha/m = 0.1 (damping time tau = 10), dt = 0.005, neighbour skin 0.5, mixed
precision.  One step = skin check, rebuild when needed (included in the timed
region), force kernel, integrate kernel; steps run in captured step blocks.

Prints ONE JSON line (see the "metric"/"value"/"roofline"/"cpu_baseline"
keys).  Extra keys report the FP64 GPU mode on the same workload, the
mixed/FP64 ratio, force/energy/virial errors of the mixed mode against the
FP64 GPU mode and the C oracle, and NVE energy drift.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# the step kernels' HIP events are recorded in the first step block of the
# timed region and every TIMING_EVERY-th one after it (every step of those
# blocks): the events' own cost (a few microseconds per recorded step) stays
# off most steps, and the kernel averages still sample the whole region
TIMING_EVERY = 4


def fcc_lattice(cells, density):
    from dompap_amd.positions import generate_positions
    ucc = np.array([(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)], dtype=np.float64)
    a = (4.0 / density) ** (1.0 / 3.0)
    x = generate_positions(ucc, (cells,) * 3, (a, a, a))
    return x, np.full(3, cells * a)


def sc_lattice(cells, density, dim=3):
    from dompap_amd.positions import generate_positions
    a = (1.0 / density) ** (1.0 / dim)
    x = generate_positions(np.zeros((1, dim)), (cells,) * dim, (a,) * dim)
    return x, np.full(dim, cells * a)


def gaussian_velocities(rng, n, dim, T, masses):
    v = rng.normal(0.0, 1.0, size=(n, dim)) * np.sqrt(T / masses.reshape(-1, 1))
    v -= np.sum(v * masses.reshape(-1, 1), axis=0) / np.sum(masses)
    return v


def workload(name, seed):
    """Synthetic inputs of the named workload (sizes as in the docstring)."""
    rng = np.random.default_rng(seed)
    if name == "lj-1m":
        x, box = fcc_lattice(64, 0.8442)
        T, skin = 0.72, 0.5
        desc = "3D LJ FCC 64^3 N=1048576 rho=0.8442 T=0.72 rc=2.5 skin=0.5 dt=0.005 G-JF tau=10"
    elif name == "lj-4k":
        x, box = sc_lattice(16, 0.8)
        T, skin = 0.8, 0.5
        desc = "3D LJ SC 16^3 N=4096 rho=0.8 T=0.8 rc=2.5 skin=0.5 dt=0.005 G-JF tau=10"
    elif name == "lj-binary-262k":
        # A/B mixture: sigma_AA = 1, sigma_BB = 1/1.2, sigma_AB their mean,
        # eps = 1, r_cut = 2.5 sigma_ij (shifted), masses A:B = 2:1, types by a
        # seeded shuffle of 50/50
        x, box = sc_lattice(64, 0.75)
        n = x.shape[0]
        types = np.zeros(n, dtype=np.int32)
        types[rng.permutation(n)[: n // 2]] = 1
        m = np.where(types == 0, 2.0, 1.0)
        sab = (1.0 + 1.0 / 1.2) / 2.0
        sig = [[1.0, sab], [sab, 1.0 / 1.2]]
        v = gaussian_velocities(rng, n, 3, 0.7, m)
        return dict(name=name, desc="3D binary LJ SC 64^3 N=262144 rho=0.75 T=0.7 rc=2.5 sigma_ij skin=0.5 "
                    "dt=0.005 G-JF tau=10 masses 2:1", x=x, v=v, m=m, box=box, pot="4*(r**-12-r**-6)", r_cut=2.5,
                    skin=0.5, dt=0.005, T=0.7, tau=10.0, K=192, seed=seed, types=types, sigma=sig)
    elif name == "harmonic-4d-2m":
        # 4-D (1-r)^2, r_cut = 1, density 1.2 on a 32x32x32x64 hypercubic
        # lattice perturbed by N(0, (0.1 a)^2)
        a = (1.0 / 1.2) ** 0.25
        from dompap_amd.positions import generate_positions
        x = generate_positions(np.zeros((1, 4)), (32, 32, 32, 64), (a,) * 4)
        box = np.array([32 * a, 32 * a, 32 * a, 64 * a])
        x = (x + rng.normal(0.0, 0.1 * a, size=x.shape)) % box
        n = x.shape[0]
        m = np.ones(n)
        v = gaussian_velocities(rng, n, 4, 0.001, m)
        return dict(name=name, desc="4D (1-r)^2 hypercubic 32x32x32x64 N=2097152 rho=1.2 T=0.001 rc=1 skin=0.3 "
                    "dt=0.005 G-JF tau=1", x=x, v=v, m=m, box=box, pot="(1-r)**2", r_cut=1.0, skin=0.3, dt=0.005,
                    T=0.001, tau=1.0, K=64, seed=seed)
    elif name == "ipl-2d-4m":
        # 2-D r^-12, r_cut = 1.5, density 1 on a 2048^2 square lattice
        # perturbed by N(0, 0.05^2)
        x, box = sc_lattice(2048, 1.0, dim=2)
        x = (x + rng.normal(0.0, 0.05, size=x.shape)) % box
        n = x.shape[0]
        m = np.ones(n)
        v = gaussian_velocities(rng, n, 2, 1.0, m)
        return dict(name=name, desc="2D r^-12 square 2048^2 N=4194304 rho=1 T=1 rc=1.5 skin=0.3 dt=0.002 G-JF tau=1",
                    x=x, v=v, m=m, box=box, pot="r**-12", r_cut=1.5, skin=0.3, dt=0.002, T=1.0, tau=1.0, K=32,
                    seed=seed)
    else:
        raise SystemExit(f"unknown workload {name}")
    n = x.shape[0]
    m = np.ones(n)
    v = gaussian_velocities(rng, n, 3, T, m)
    return dict(name=name, desc=desc, x=x, v=v, m=m, box=box, pot="4*(r**-12-r**-6)", r_cut=2.5, skin=skin,
                dt=0.005, T=T, tau=10.0, K=256, seed=seed)


def replica_workload(seed, n_rep=256):
    """256 independent 3-D LJ systems: N log-uniform in [500, 20000], density
    uniform in [0.7, 1.0], T uniform in [0.5, 2.0] (seeded); each starts on a
    seeded subset of an FCC lattice of 4 c^3 >= N sites filling its cubic box,
    displaced by N(0, 0.05^2), with Gaussian velocities at its T (zero
    momentum); r_cut 2.5, skin 0.5, dt 0.005, tau 10."""
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(math.log(500), math.log(20000), n_rep))).astype(np.int64)
    rho = rng.uniform(0.7, 1.0, n_rep)
    T = rng.uniform(0.5, 2.0, n_rep)
    from dompap_amd.positions import generate_positions
    ucc = np.array([(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)], dtype=np.float64)
    xs, vs, boxes = [], [], []
    for r in range(n_rep):
        n = int(ns[r])
        c = int(math.ceil((n / 4.0) ** (1.0 / 3.0) - 1e-9))
        L = (n / rho[r]) ** (1.0 / 3.0)
        a = L / c
        sites = generate_positions(ucc, (c, c, c), (a, a, a))
        x = sites[np.sort(rng.choice(sites.shape[0], size=n, replace=False))]
        x = (x + rng.normal(0.0, 0.05, x.shape)) % L
        xs.append(x)
        vs.append(gaussian_velocities(rng, n, 3, T[r], np.ones(n)))
        boxes.append(np.full(3, L))
    x = np.concatenate(xs)
    return dict(name="replicas-256", desc=f"{n_rep} batched 3D LJ replicas N log-uniform [500,20000] (total "
                f"{int(ns.sum())}) rho U[0.7,1.0] T U[0.5,2.0] FCC-subset starts rc=2.5 skin=0.5 dt=0.005 G-JF "
                "tau=10 100-step graph blocks", x=x, v=np.concatenate(vs), m=np.ones(x.shape[0]),
                box=np.stack(boxes), offsets=np.concatenate([[0], np.cumsum(ns)]).astype(np.int64), T=T,
                ns=ns, pot="4*(r**-12-r**-6)", r_cut=2.5, skin=0.5, dt=0.005, tau=10.0, K=256, seed=seed)


def make_batch_engine(wl, precision, block_steps=100, tuning=None):
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    n = wl["x"].shape[0]
    R = len(wl["ns"])
    eng = Engine(3, n, R, precision)
    eng.set_replicas(wl["offsets"], wl["box"])
    dp = device_potential(wl["pot"], wl["r_cut"])
    eng.load_potential(dp.code, dp.r_cut, dp.u_rc, [[1.0]], [[1.0]])
    eng.set_neighbor_params(np.full(R, wl["r_cut"] + wl["skin"]), np.full(R, wl["skin"]), wl["K"])
    eng.set_integrator(np.full(R, wl["dt"]), wl["T"], np.full(R, wl["tau"]))
    eng.set_seed(wl["seed"], 0)
    eng.set_replica_seeds(np.arange(R, dtype=np.uint64) + np.uint64(wl["seed"]))
    eng.set_graph_block(block_steps)
    if tuning is not None:
        eng.set_tuning(*tuning)
    eng.set_field("positions", wl["x"])
    eng.set_field("velocities", wl["v"])
    eng.set_field("masses", wl["m"])
    eng.build_neighbor_list()
    return eng


def make_engine(wl, precision, pair_precision="fp64", tuning=None, cell_scale=1.0):
    from dompap_amd.engine import Engine
    from dompap_amd.potential import device_potential
    n, dim = wl["x"].shape
    eng = Engine(dim, n, 1, precision)
    eng.set_box(wl["box"])
    dp = device_potential(wl["pot"], wl["r_cut"])
    sig = np.asarray(wl.get("sigma", [[1.0]]), dtype=np.float64)
    eng.load_potential(dp.code, dp.r_cut, dp.u_rc, sig, np.ones_like(sig), pair_precision)
    # list cutoff: largest sigma times r_cut, plus the skin (Simulation.py:353-356)
    eng.set_neighbor_params(float(sig.max()) * wl["r_cut"] + wl["skin"], wl["skin"], wl["K"])
    eng.set_integrator(wl["dt"], wl["T"], wl["tau"])
    eng.set_seed(wl["seed"], 0)
    if tuning is not None:
        eng.set_tuning(*tuning)
    if cell_scale != 1.0:
        eng.set_cell_scale(cell_scale)
    eng.set_field("positions", wl["x"])
    eng.set_field("velocities", wl["v"])
    eng.set_field("masses", wl["m"])
    if "types" in wl:
        eng.set_field("types", wl["types"])
    eng.build_neighbor_list()
    return eng


def sync_all(eng):
    eng.sync()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def tune_cells(eng, args, log):
    """With --cell-scale 0, the cell width chosen by dompap_amd's autotuner
    (timed steps at each width before the warmup); else None."""
    if args.cell_scale:
        return None
    from dompap_amd.tools.autotune import tune_cell_scale
    scale, times = tune_cell_scale(eng, steps=args.tune_steps)
    per = {f"{s:g}": round(t / args.tune_steps * 1e3, 5) for s, t in times.items()}
    log(f"cell width {scale:g} list cutoffs (ms per step by width: {per})")
    return {"chosen": scale, "ms_per_step_by_scale": per, "steps_each": args.tune_steps}


def timed_run(eng, steps, timing=True, breakdown=False):
    """Wall time of `steps` G-JF steps (device work bracketed by stream
    synchronisations), the HIP-event time of the step kernels, launches and
    rebuilds; with breakdown, also (force-kernel ms, integrate-kernel ms)."""
    upd0, st0, _ = eng.stats()
    eng.set_kernel_timing(timing, TIMING_EVERY)
    sync_all(eng)
    t0 = time.perf_counter()
    eng.step(steps, "gjf")
    sync_all(eng)
    t1 = time.perf_counter()
    ms, launches = eng.kernel_time()
    split = eng.kernel_breakdown()
    eng.set_kernel_timing(False)
    upd1, _, _ = eng.stats()
    if breakdown:
        return t1 - t0, ms, launches, int(upd1[0] - upd0[0]), split
    return t1 - t0, ms, launches, int(upd1[0] - upd0[0])


def force_errors(F, Fref):
    dF = np.linalg.norm(F - Fref, axis=1)
    Fn = np.linalg.norm(Fref, axis=1)
    frms = float(np.sqrt(np.mean(Fn ** 2)))
    big = Fn > 0.1 * frms
    return {"max_abs_over_frms": float(dF.max() / frms),
            "max_rel_per_particle": float(np.max(dF[big] / Fn[big])) if big.any() else None,
            "rms_rel": float(np.sqrt(np.mean(dF ** 2)) / frms)}


def nve_drift(eng, steps, every):
    """Slope (per step) of E_tot/N versus step for one friction-free segment;
    E_tot = U(x_n) + (K(v_{n-1/2}) + K(v_{n+1/2}))/2 (leap-frog velocities)."""
    n = eng.n
    m = eng.get_field("masses")
    ts, es = [], []
    for k in range(steps // every):
        eng.step(every - 1, "leapfrog")
        v0 = eng.get_field("velocities")
        U = eng.compute(forces=False, energy=True)["energy"][0]
        eng.step(1, "leapfrog")
        v1 = eng.get_field("velocities")
        K = 0.25 * (np.sum(m[:, None] * v0 * v0) + np.sum(m[:, None] * v1 * v1))
        ts.append((k + 1) * every)
        es.append((U + K) / n)
    t = np.asarray(ts, dtype=float)
    e = np.asarray(es)
    slope = float(np.polyfit(t, e, 1)[0])
    return slope, float(e.std())


def nve_segments(eng, segments, steps, every, dt):
    """Drift per particle per unit time over `segments` consecutive
    friction-free segments of `steps` steps: each segment's own slope, their
    mean and its standard error (segment-to-segment scatter / sqrt(n))."""
    slopes, sds = [], []
    for _ in range(segments):
        sl, sd = nve_drift(eng, steps, every)
        slopes.append(sl / dt)
        sds.append(sd)
    a = np.asarray(slopes)
    se = float(a.std(ddof=1) / np.sqrt(len(a))) if len(a) > 1 else None
    return {"drift_per_particle_per_time": float(a.mean()), "stderr": se, "segments": [float(v) for v in a],
            "energy_std_per_particle": float(np.mean(sds))}


def cpu_baseline(wl, x, rebuild_interval, budget_s=20.0):
    """C oracle (oracle/dp_oracle.c, FP64, OpenMP) on the host cores: one
    neighbour-list build (cell search; substitutes for the reference's O(N^2)
    loop) and a few force + G-JF steps at full N, combined with the GPU run's
    rebuild interval."""
    from oracle import c_oracle as C
    n, dim = x.shape
    box = wl["box"]
    cores = C.threads()
    t0 = time.perf_counter()
    pot = {"4*(r**-12-r**-6)": "lj", "(1-r)**2": "harm2", "r**-12": ("ipl", 12)}[wl["pot"]]
    sig = np.asarray(wl.get("sigma", [[1.0]]), dtype=np.float64)
    types = wl.get("types")
    nl, cnt = C.neighbor_list(x, box, float(sig.max()) * wl["r_cut"] + wl["skin"], wl["K"], "cells")
    t_nl = time.perf_counter() - t0
    v = wl["v"].copy()
    beta = np.zeros_like(v)
    img = np.zeros(x.shape, np.int32)
    steps = 0
    t_steps = 0.0
    xx = x.copy()
    F0 = None
    while steps < 2 or (steps < 10 and t_steps < budget_s / 2):
        t0 = time.perf_counter()
        F = C.pair(xx, box, nl, pot, wl["r_cut"], sigma=sig, eps=np.ones_like(sig), types=types)["forces"]
        if F0 is None:
            F0 = F
        xi = C.normals(n, dim, 1, steps)
        xx, v, beta, img = C.step(xx, v, beta, F, wl["m"], wl["dt"], wl["T"], wl["tau"], xi, box, img)
        t_steps += time.perf_counter() - t0
        steps += 1
    per_step = t_steps / steps + t_nl / max(rebuild_interval, 1.0)
    return {"value": n / per_step, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"C oracle FP64 (OpenMP) on N={n}: 1 cell-list neighbour build ({t_nl:.2f}s) + "
                      f"{steps} force+G-JF steps ({t_steps / steps:.3f}s each), rebuild every "
                      f"{rebuild_interval:.1f} steps as measured on the GPU"}, F0


def copy_calibration():
    """A plain copy of known size: 5 copies of 256 MiB between device buffers.
    FETCH_SIZE / WRITE_SIZE measured on it give the counters' scale factors."""
    import torch
    x = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    x.uniform_()
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_per_copy": 256 * 1024 * 1024, "copies": 5}}), flush=True)


def measured_traffic(workload_desc, plan, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (profiles/traffic.json), if it was measured on this workload and
    tile plan; else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    for e in data.get("entries", []):
        if e.get("workload") == workload_desc and e.get("kernel") == kernel and e.get("plan") == plan:
            return e.get("bytes_per_launch")
    return None


def run_replicas(args, log):
    """The batched-replica workload: one JSON line like main()'s."""
    wl = replica_workload(args.seed)
    n = wl["x"].shape[0]
    R = len(wl["ns"])
    log(f"workload {wl['desc']} precision={args.precision}")
    tuning = None
    if args.tpp or args.tile or args.fused >= 0:
        tuning = (args.tpp, args.tile, args.fused)
    eng = make_batch_engine(wl, args.precision, tuning=tuning)
    cell_tuning = tune_cells(eng, args, log)
    eng.step(args.warmup, "gjf")
    eng.prepare_step_graphs("gjf", timed=True)
    eng.prepare_step_graphs("gjf", timed=False)
    sync_all(eng)
    kbar0 = float(eng.neighbor_counts().mean())
    upd0 = int(eng.stats()[0].sum())
    eng.set_kernel_timing(True, TIMING_EVERY)
    sync_all(eng)
    t0 = time.perf_counter()
    eng.step(args.steps, "gjf")
    sync_all(eng)
    dt = time.perf_counter() - t0
    _, launches = eng.kernel_time()
    f_ms, i_ms = eng.kernel_breakdown()
    eng.set_kernel_timing(False)
    rebuilds = int(eng.stats()[0].sum()) - upd0
    dtg, _, _, _ = timed_run(eng, args.steps, timing=False)
    kbar = 0.5 * (kbar0 + float(eng.neighbor_counts().mean()))
    value = n * args.steps / dt
    nl_ = max(launches, 1)
    force_s = f_ms / 1e3 / nl_
    integ_s = i_ms / 1e3 / nl_
    s_x = 4 if args.precision == "mixed" else 8
    b_force = 4.0 * kbar + 3 * s_x + 24
    achieved = n * b_force / force_s / 1e9
    per_rep_interval = args.steps * R / max(rebuilds, 1)
    log(f"timed {args.steps} steps: {dt:.3f}s -> {value:.4e} particle-steps/s; replica rebuilds={rebuilds} "
        f"(one per {per_rep_interval:.1f} steps and replica); force kernel {force_s * 1e6:.1f} us, "
        f"integrate kernel {integ_s * 1e6:.1f} us per step")
    plan = eng.plan()
    out = {
        "metric": "particle-steps/s on 1 MI355X, mixed vs FP64 GPU, at stated force error",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "i32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded FCC-subset lattices + Gaussian velocities)",
        "config": {"workload": wl["desc"] + f" {args.precision}", "n_particles": n, "replicas": R, "dim": 3,
                   "precision": args.precision, "seed": args.seed},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "dp_tile_force",
                     "bytes_per_particle": b_force, "bytes_definition": "4*k + D*s_x + 8*D",
                     "mean_list_length": kbar, "avg_launch_us": force_s * 1e6,
                     "integrate_kernel": {"kernel": "dp_integrate_gjf", "avg_launch_us": integ_s * 1e6}},
        "plan": plan,
        "cell_tuning": cell_tuning,
        "timing": (f"HIP events recorded inside the step-block graphs of one block in {TIMING_EVERY} "
                   "(100-step blocks, per-replica rebuilds)"),
        "graph_replay_value": n * args.steps / dtg,
        "replica_rebuilds_in_timed_region": rebuilds,
        "rebuild_interval_per_replica": per_rep_interval,
    }
    x_now = eng.get_field("positions")
    if not args.no_fp64 and args.precision == "mixed":
        log("FP64 comparison run")
        e64 = make_batch_engine(wl, "fp64")
        cell_tuning64 = tune_cells(e64, args, log)
        e64.step(args.warmup, "gjf")
        e64.prepare_step_graphs("gjf", timed=False)
        k64 = max(args.steps // 2, 200)
        dt64, _, _, _ = timed_run(e64, k64, timing=False)
        v64 = n * k64 / dt64
        out["fp64_value"] = v64
        out["mixed_over_fp64"] = value / v64
        if cell_tuning64 is not None:
            out["fp64_cell_tuning"] = cell_tuning64
        log(f"fp64: {v64:.4e} particle-steps/s (ratio {value / v64:.2f})")
        e64.set_field("positions", x_now)
        r64 = e64.compute(forces=True, energy=True, virial=True)
        rmx = eng.compute(forces=True, energy=True, virial=True)
        out["mixed_vs_fp64_gpu"] = force_errors(rmx["forces"], r64["forces"])
        out["mixed_vs_fp64_gpu"]["energy_rel_max"] = float(np.max(np.abs(rmx["energy"] - r64["energy"]) /
                                                                  np.abs(r64["energy"])))
        e64.close()
    if not args.no_cpu:
        log("CPU baseline (C oracle)")
        from oracle import c_oracle as C
        cores = C.threads()
        # a bounded sample: the first replicas up to ~100,000 particles, one
        # list build each (amortised over the measured rebuild interval) and
        # three force + G-JF steps each
        t_nl = t_st = 0.0
        n_s = 0
        for r in range(R):
            lo, hi = int(wl["offsets"][r]), int(wl["offsets"][r + 1])
            x = x_now[lo:hi]
            box = wl["box"][r]
            t0 = time.perf_counter()
            nl, _ = C.neighbor_list(x, box, wl["r_cut"] + wl["skin"], wl["K"], "cells")
            t_nl += time.perf_counter() - t0
            v = wl["v"][lo:hi].copy()
            beta = np.zeros_like(v)
            img = np.zeros(x.shape, np.int32)
            t0 = time.perf_counter()
            for k in range(3):
                F = C.pair(x, box, nl, "lj", wl["r_cut"])["forces"]
                xi = C.normals(hi - lo, 3, 1, k)
                x, v, beta, img = C.step(x, v, beta, F, wl["m"][lo:hi], wl["dt"], float(wl["T"][r]), wl["tau"], xi,
                                         box, img)
            t_st += (time.perf_counter() - t0) / 3
            n_s += hi - lo
            if n_s >= 100000:
                break
        per_step = t_st + t_nl / max(per_rep_interval, 1.0)
        out["cpu_baseline"] = {"value": n_s / per_step, "unit": "particle-steps/s", "cores": cores, "kind": "port",
                               "sample": f"C oracle FP64 (OpenMP) on the first {r + 1} replicas ({n_s} particles): "
                                         f"one cell-list build each ({t_nl:.2f}s in all) + 3 force+G-JF steps each "
                                         f"({t_st:.3f}s per step in all), rebuild every {per_rep_interval:.1f} "
                                         "steps per replica as measured on the GPU"}
        log(f"cpu: {out['cpu_baseline']['value']:.3e} particle-steps/s on {cores} threads")
    else:
        out["cpu_baseline"] = None
    eng.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="lj-1m",
                    choices=["lj-1m", "lj-4k", "lj-binary-262k", "harmonic-4d-2m", "ipl-2d-4m", "replicas-256"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--pair-precision", default="fp64", choices=["fp64", "fp32"],
                    help="mixed mode: precision of the in-range pair terms")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tpp", type=int, default=0, help="lanes per particle (0: planner's choice)")
    ap.add_argument("--tile", type=int, default=0, help="home particles per tile (0: planner's choice)")
    ap.add_argument("--cell-scale", type=float, default=1.0,
                    help="cells at least this many list cutoffs wide (1.0: the planner's narrowest, the default); "
                         "0: dompap_amd's autotuner chooses, timing each width on the workload before the warmup "
                         "(in both precision modes)")
    ap.add_argument("--tune-steps", type=int, default=200, help="timed steps per cell width while autotuning")
    ap.add_argument("--fused", type=int, default=-1, choices=[-1, 0, 1],
                    help="1: fused force + integrate kernel, 0: force kernel then integrate kernel, -1: planner")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-nve", action="store_true", help="skip the NVE drift segment")
    ap.add_argument("--nve-steps", type=int, default=2000, help="steps per NVE segment")
    ap.add_argument("--nve-segments", type=int, default=5, help="consecutive NVE segments per precision mode")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy 5 times (HBM counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (dompap_amd has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.workload == "replicas-256":
        run_replicas(args, log)
        return
    wl = workload(args.workload, args.seed)
    n, dim = wl["x"].shape
    log(f"workload {wl['desc']} precision={args.precision}")
    tuning = None
    if args.tpp or args.tile or args.fused >= 0:
        tuning = (args.tpp, args.tile, args.fused)
    eng = make_engine(wl, args.precision, args.pair_precision, tuning, args.cell_scale or 1.0)
    cell_tuning = tune_cells(eng, args, log)
    eng.step(args.warmup, "gjf")
    # every step-block graph the timed region replays (events recorded or
    # not, either buffer parity) is captured and instantiated here, outside
    # the timed region, whatever --warmup is
    eng.prepare_step_graphs("gjf", timed=True)
    eng.prepare_step_graphs("gjf", timed=False)
    sync_all(eng)
    counts = eng.neighbor_counts()
    kbar0 = float(counts.mean())
    log(f"warmup {args.warmup} done, mean list length {kbar0:.1f}")
    plan = eng.plan()
    # the tile plan (what the committed traffic figures were measured on)
    tile_plan = {k: v for k, v in plan.items() if k not in ("steps_per_block", "graph_events")}
    # timed region: step blocks (captured hipGraphs, the rebuild pipeline once
    # per block) with HIP events recorded by the graph around the step kernels
    dt, ms, launches, rebuilds, (f_ms, i_ms) = timed_run(eng, args.steps, timing=True, breakdown=True)
    value = n * args.steps / dt
    kbar1 = float(eng.neighbor_counts().mean())
    kbar = 0.5 * (kbar0 + kbar1)  # mean list length over the timed region (its two ends)
    # the same number of steps again, without events
    dtg, _, _, _ = timed_run(eng, args.steps, timing=False)
    nl_ = max(launches, 1)
    log(f"timed {args.steps} steps: {dt:.3f}s -> {value:.4e} particle-steps/s; rebuilds={rebuilds}; "
        f"force kernel {f_ms / nl_ * 1e3:.1f} us, integrate kernel {i_ms / nl_ * 1e3:.1f} us per step")
    # Algorithmic HBM bytes per particle and force-kernel launch, as the
    # measurement contract defines them: B_force = 4 k (one int32 entry per
    # neighbour, whatever the list's storage) + D s_x (the particle's
    # coordinates once: 4 bytes each in mixed mode, 8 in FP64) + 8 D (its
    # force) [+ 4 with several types].
    typed = "types" in wl
    s_x = 4 if args.precision == "mixed" else 8
    b_force = 4.0 * kbar + dim * s_x + 8 * dim + (4 if typed else 0)
    # The implementation's own bytes: entries of entry_bits, the staged
    # record (mixed: RS ints; FP64: D doubles), the row length, the force and
    # the share of the tile tables.  Integrate kernel: read x, v, beta, F,
    # xref (5 D doubles), cellpack, orig (and mass unless every mass is 1);
    # write x, v, beta (3 D doubles); mixed also writes its record (RS ints,
    # read first when the type slot must carry over).
    ent = plan["entry_bits"] // 8
    rs = 2
    while rs < dim + 1:
        rs *= 2
    home = max(n / max(plan["tiles"], 1), 1.0)
    tab_pp = (8 + 3 * plan["halo_cells_per_tile"] + 2 * plan["home_cells_per_tile"] + 2) * 4 / home
    rec = rs * 4 if args.precision == "mixed" else dim * 8
    b_force_impl = ent * kbar + rec + 4 + 8 * dim + tab_pp
    unit_mass = bool(np.all(wl["m"] == 1.0))
    b_integ = (8 * dim * 5 + 4 + 4 + (0 if unit_mass else 8) + 8 * dim * 3
               + ((rs * 4 * (2 if typed else 1)) if args.precision == "mixed" else 0))
    force_s = f_ms / 1e3 / nl_
    integ_s = i_ms / 1e3 / nl_
    achieved = n * b_force / force_s / 1e9
    out = {
        "metric": "particle-steps/s on 1 MI355X, mixed vs FP64 GPU, at stated force error",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # mixed: int32 fixed-point displacements and screening, pair terms in
        # FP64 (or FP32), FP64 accumulation and integration
        "dtype": (("i32+f64" if args.pair_precision == "fp64" else "i32+f32+f64")
                  if args.precision == "mixed" else "f64"),
        "data": "synthetic (seeded lattice + Gaussian velocities)",
        "config": {"workload": wl["desc"] + f" {args.precision}", "n_particles": n, "dim": dim,
                   "precision": args.precision, "pair_precision": args.pair_precision, "seed": args.seed},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(wl["desc"] + f" {args.precision}", tile_plan,
                                                 "dp_tile_force" if not plan["fused"] else "dp_tile_gjf"),
                     "kernel": "dp_tile_force" if not plan["fused"] else "dp_tile_gjf",
                     "bytes_per_particle": b_force, "mean_list_length": kbar,
                     "bytes_definition": "4*k + D*s_x + 8*D (+4 typed)",
                     "impl_bytes_per_particle": b_force_impl,
                     "impl_achieved_GBs": n * b_force_impl / force_s / 1e9,
                     "avg_launch_us": force_s * 1e6,
                     "integrate_kernel": {"kernel": "dp_integrate_gjf", "avg_launch_us": integ_s * 1e6,
                                          "bytes_per_particle": b_integ,
                                          "achieved_GBs": n * b_integ / integ_s / 1e9 if integ_s > 0 else None}},
        "plan": plan,
        "cell_tuning": cell_tuning,
        "timing": (f"HIP events recorded inside the step-block graphs of one block in {TIMING_EVERY}"
                   if eng.plan()["graph_events"] else "HIP events around single-step launches"),
        "graph_replay_value": n * args.steps / dtg,
        "rebuilds_in_timed_region": rebuilds,
        "rebuild_interval": args.steps / max(rebuilds, 1),
    }
    x_now = eng.get_field("positions")
    if not args.no_fp64 and args.precision == "mixed":
        log("FP64 comparison run")
        eng64 = make_engine(wl, "fp64", cell_scale=args.cell_scale or 1.0)
        cell_tuning64 = tune_cells(eng64, args, log)
        eng64.step(args.warmup, "gjf")
        eng64.prepare_step_graphs("gjf", timed=True)
        eng64.prepare_step_graphs("gjf", timed=False)
        k64 = max(args.steps // 4, 50)
        dt64, ms64, l64, rb64 = timed_run(eng64, k64, timing=True)
        v64 = n * k64 / dt64
        out["fp64_value"] = v64
        out["mixed_over_fp64"] = value / v64
        if cell_tuning64 is not None:
            out["fp64_cell_tuning"] = cell_tuning64
        out["fp64_step_kernel_us"] = ms64 / max(l64, 1) * 1e3
        log(f"fp64: {v64:.4e} particle-steps/s (ratio {value / v64:.2f})")
        # errors of the mixed kernels on the mixed run's configuration
        eng64.set_field("positions", x_now)
        r64 = eng64.compute(forces=True, energy=True, virial=True)
        rmx = eng.compute(forces=True, energy=True, virial=True)
        out["mixed_vs_fp64_gpu"] = force_errors(rmx["forces"], r64["forces"])
        out["mixed_vs_fp64_gpu"]["energy_rel"] = float(abs(rmx["energy"][0] - r64["energy"][0]) / abs(r64["energy"][0]))
        out["mixed_vs_fp64_gpu"]["virial_rel"] = float(abs(rmx["virial"][0] - r64["virial"][0]) / abs(r64["virial"][0]))
        eng64.close()
    if not args.no_nve:
        log("NVE drift segments")
        nve = {}
        for prec in (("mixed", "fp64") if args.precision == "mixed" and not args.no_fp64 else (args.precision,)):
            e2 = make_engine(wl, prec)
            e2.set_field("positions", x_now)
            e2.set_field("velocities", eng.get_field("velocities"))
            e2.build_neighbor_list()
            e2.set_integrator(wl["dt"], wl["T"], np.inf)
            nve[prec] = nve_segments(e2, args.nve_segments, args.nve_steps, 100, wl["dt"])
            e2.close()
            log(f"nve {prec}: drift {nve[prec]['drift_per_particle_per_time']:.3e} +- "
                f"{nve[prec]['stderr'] or 0.0:.1e} per particle per unit time")
        out["nve"] = {"steps_per_segment": args.nve_steps, "segments": args.nve_segments, **nve}
    if not args.no_cpu:
        log("CPU baseline (C oracle)")
        cb, Fcpu = cpu_baseline(wl, x_now, out["rebuild_interval"])
        out["cpu_baseline"] = cb
        Fg = eng.compute(forces=True)["forces"]
        out["mixed_vs_c_oracle" if args.precision == "mixed" else "fp64_vs_c_oracle"] = force_errors(Fg, Fcpu)
        log(f"cpu: {cb['value']:.3e} particle-steps/s on {cb['cores']} threads")
    else:
        out["cpu_baseline"] = None
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
