#!/usr/bin/env python3
"""Benchmark of the MI355X OpenFEA hot path; prints one JSON line.

Default workload (the headline configuration): linear static C3D4, 60x30x30
hex grid split into 324,000 Kuhn tets, interior nodes jittered +/-0.2h
(seed 20261203), E = 210000 MPa, nu = 0.3, FP64 assembly, clamped x = 0,
-5 MPa z traction on x = L.  A "step" is one FP64 assembly of K into the
reference's CSC arrays with mesh, pattern and plans resident in HBM.
value = assembled elements per second over the K timed steps.  The same
line reports time-to-solve to a true relative residual of 1e-8 for plain
FP64 Jacobi PCG and the mixed-precision/learned solvers, the roofline of the
assembly kernel and the oracle's single-core CPU rate.

Other workloads (--workload cfg1|cfg3|cfg4|cfg5) print their own line.
"""
import argparse
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def load_traffic(kernel_tag):
    """HBM bytes per launch measured by rocprofv3 --pmc (committed summary)."""
    path = os.path.join(REPO, "profiles", f"traffic_{kernel_tag}.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("bytes_per_launch")
    return None


def cpu_baseline_fill(m, resort, reps=3):
    import numpy as np
    from oracle import oracle as O

    cp, ri = O.build_csc(m.conn, m.etype, resort, 3 * int(resort.max() + 1))
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        O.fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri)
        times.append(time.perf_counter() - t)
    del np
    return m.ne / statistics.median(times)


def train_learned(workload, args):
    """Trains the learned block-Jacobi correction on a smaller mesh of the same
    family (different jitter seed) from that mesh's own block-Jacobi PCG runs;
    the node-local network then transfers to the benchmark mesh."""
    import numpy as np

    import openfea_amd as ofa
    from openfea_amd import learned as L, meshgen as mg

    t0 = time.perf_counter()
    if workload == "cfg2":
        m, resort, F = mg.cantilever_c3d4(24, 12, 12, seed=7)
    else:
        m, resort, F = mg.cantilever_c3d8(10, 5, 5)
    ctx = ofa.Context(0)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    ctx.assemble(mg.E_STEEL, mg.NU_STEEL, copy=False)
    rng = np.random.default_rng(11)
    rhs = [F] + [ctx.spmv(rng.normal(size=ctx.n)) for _ in range(3)]
    E, R = L.collect_pairs(ctx, rhs, snapshots=(8, 16, 32, 64, 128))
    ln = L.Learned(ctx, seed=0)
    ln.set_samples(E, R)
    hist = ln.train(steps=args.train_steps, lr=2e-3)
    params = ln.get_params()
    ctx.close()
    info = {"train_mesh": f"{m.ne} elements ({'C3D4 24x12x12, jitter seed 7' if workload == 'cfg2' else 'C3D8 10x5x5'})",
            "samples": int(len(E)), "steps": args.train_steps, "loss_first": float(hist[0]),
            "loss_last": float(hist[-1]), "train_s": round(time.perf_counter() - t0, 3), "params": params}
    log(f"learned: samples {len(E)} loss {hist[0]:.4e} -> {hist[-1]:.4e} in {info['train_s']}s")
    return info


def bench_static(args, workload):
    import numpy as np

    import openfea_amd as ofa
    from openfea_amd import _capi as A, learned as L, meshgen as mg

    if workload == "cfg2":
        m, resort, F = mg.cantilever_c3d4()
        wl = "linear static C3D4 60x30x30 Kuhn 6-tet (324,000 tets), jitter +/-0.2h seed 20261203, FP64 assembly + PCG"
        tag = "assemble_tet_cfg2"
    else:
        m, resort, F = mg.cantilever_c3d8()
        wl = "linear static C3D8 20x10x10 (2,000 hexes), FP64 assembly + PCG"
        tag = "assemble_hex_cfg1"
    nn = 4 if workload == "cfg2" else 8
    ctx = ofa.Context(0)
    t0 = time.perf_counter()
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    setup_s = time.perf_counter() - t0
    log(f"{workload}: ne={m.ne} nd={m.nd} n={ctx.n} nnz={ctx.nnz} setup {setup_s:.2f}s")
    E, nu = mg.E_STEEL, mg.NU_STEEL
    for _ in range(args.warmup):
        ctx.assemble(E, nu, copy=False)
    ctx.sync()
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.assemble(E, nu, copy=False)
    ms_total = ctx.timer_stop()
    ms_step = ms_total / args.steps
    value = m.ne / (ms_step * 1e-3)
    log(f"assembly {ms_step * 1e3:.1f} us/step -> {value:.3e} elems/s")
    alg_bytes = 24 * m.nd + 4 * nn * m.ne + 8 * ctx.nnz
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9
    traffic = load_traffic(tag)

    # time-to-solve (RHS resident, x left on device), true FP64 residual 1e-8
    ctx.assemble(E, nu, copy=False)
    ctx.set_rhs(F)
    solves = {}
    all_methods = {"plain_fp64_jacobi": A.PCG_JACOBI, "fp64_block_jacobi": A.PCG_BLOCK_JACOBI,
                   "mixed_fp32val_block_jacobi": A.PCG_MIXED, "mixed_fp32_all_block_jacobi": A.PCG_MIXED32,
                   "learned_fp64": A.PCG_LEARNED, "learned_mixed_fp32val": A.PCG_LEARNED_MIXED}
    chosen = [] if args.no_solve else [m for m in args.solvers.split(",") if m]
    methods = [(m, all_methods[m]) for m in chosen]
    learned_info = None
    if any(m.startswith("learned") for m in chosen):
        learned_info = train_learned(workload, args)
        ln = L.Learned(ctx, seed=0)
        ln.set_params(learned_info.pop("params"))
        ctx.set_rhs(F)
    for name, method in methods:
        ctx.solve(None, method=method, rtol=1e-8, maxit=100000, copy=False)  # warm (graph capture)
        ts = []
        for _ in range(args.solve_reps):
            ctx.sync()
            t = time.perf_counter()
            _, it, rel, ok = ctx.solve(None, method=method, rtol=1e-8, maxit=100000, copy=False)
            ts.append(time.perf_counter() - t)
        solves[name] = {"ms": round(statistics.median(ts) * 1e3, 3), "iters": it, "relres": rel, "converged": ok}
        log(f"solve {name}: {solves[name]}")
    base = solves.get("plain_fp64_jacobi", {}).get("ms")
    for k, v in solves.items():
        v["speedup_vs_plain"] = round(base / v["ms"], 3) if base else None

    cpu = None
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        rate = cpu_baseline_fill(m, resort)
        cpu = {"value": rate, "unit": "elems/s", "cores": 1, "kind": "port",
               "sample": f"oracle fill_csc (reference fill_CSR_sparse_mat algorithm) over the full {m.ne}-element "
                         f"mesh, pattern prebuilt, median of 3 ({time.perf_counter() - t:.1f}s)"}
    ctx.close()
    return {
        "metric": "assembled elems/s + time-to-solve at rel-resid 1e-8 (learned vs plain PCG)",
        "value": value, "unit": "elems/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "elements": m.ne, "nodes": m.nd, "dofs": ctx.n, "nnz": ctx.nnz,
                   "parallelism": "single GPU"},
        "time_to_solve": solves,
        "learned_training": learned_info,
        "setup_s": round(setup_s, 3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_assemble_tet" if nn == 4 else "k_assemble_hex",
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }


def bench_explicit(args):
    import numpy as np

    import openfea_amd as ofa
    from openfea_amd import _capi as A, meshgen as mg
    from oracle import oracle as O

    m = mg.hex_block(100, 50, 50)
    wl = "explicit C3D8 100x50x50 (250,000 hexes), lumped mass, central difference, FP32 state"
    ctx = ofa.Context(0)
    ctx.set_mesh(m.xyz, m.conn, m.etype, mg.free_index(m.nd, m.fix_set))
    mass = ctx.lumped_mass(mg.RHO_STEEL)
    fixed = np.zeros(m.nd, dtype=np.int32)
    fixed[m.fix_set - 1] = 1
    v0 = mg.initial_velocity(m)
    fext = np.zeros((m.nd, 3))
    # dt = 0.9 * 2 / omega_max, omega_max from the element eigenproblem of the unit cube
    Ke = O.c3d8_ke(m.xyz[m.conn[0] - 1], mg.E_STEEL, mg.NU_STEEL)
    me = O.c3d8_mass(m.xyz[m.conn[0] - 1], mg.RHO_STEEL)
    lam = np.linalg.eigvals(Ke / np.repeat(me, 3)[:, None]).real.max()
    dt = 0.9 * 2.0 / np.sqrt(lam)
    nsteps = args.steps if args.steps > 0 else 1000
    ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, min(args.warmup * 50, 100), copy=False)
    ctx.sync()
    ctx.timer_start()
    ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, nsteps, copy=False)
    ms = ctx.timer_stop()
    ms_step = ms / nsteps
    alg_bytes = 32 * m.ne + 64 * m.nd
    ctx.close()
    return {
        "metric": "explicit steps/s", "value": 1e3 / ms_step, "unit": "steps/s", "n_gpus": 1, "steps": nsteps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": wl, "dt": dt, "elements": m.ne},
        "roofline": {"bound": "hbm", "achieved": alg_bytes / (ms_step * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg2", choices=["cfg1", "cfg2", "cfg3"])
    ap.add_argument("--solve-reps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="assembly only (for counter collection)")
    ap.add_argument("--solvers", default="plain_fp64_jacobi,fp64_block_jacobi,mixed_fp32val_block_jacobi,"
                    "mixed_fp32_all_block_jacobi,learned_fp64,learned_mixed_fp32val")
    ap.add_argument("--train-steps", type=int, default=150)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload in ("cfg1", "cfg2"):
        out = bench_static(args, args.workload)
    else:
        out = bench_explicit(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
