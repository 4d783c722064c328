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
assembly kernel and the CPU baseline: the reference's own element code
(oracle/_ref) filling the same K on one host core and on all cores, with the
oracle's single-core rate beside it.  A "configs" block adds compact lines
for configs 1, 3, 4 and 5 (each with its time, roofline and CPU baseline);
--no-configs leaves it out.

Other workloads (--workload cfg1|cfg3|cfg4|cfg5|ca|post|sfem) print their own line.
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
    """HBM bytes per launch measured by rocprofv3 --pmc (committed summary)
    and where the figure comes from."""
    path = os.path.join(REPO, "profiles", f"traffic_{kernel_tag}.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("bytes_per_launch")
    return None


def traffic_source(kernel_tag):
    return (f"profiles/traffic_{kernel_tag}.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of the same kernel "
            f"and workload, scaled by the copy-kernel calibration (profiles/copy_calibration.md); not measured "
            f"in this run")


def cpu_threads():
    n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def cpu_fill_baseline(m, resort, what, E_elem=None, reps=1):
    """K filled on the host by the reference's own element code (oracle/_ref:
    build_ele_stiff_mat of src/Elements/C3D4.cpp / C3D8.cpp and the scatter of
    fill_CSR_sparse_mat) on one core and on all cores, beside the oracle's
    single-core rate (the calibration of the port).  Without oracle/_ref only
    the oracle's rate (kind "port").  Returns the baseline and the K filled."""
    from oracle import oracle as O

    cp, ri = O.build_csc(m.conn, m.etype, resort, 3 * int(resort.max() + 1))

    def best(fn):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t)
        return min(ts), r

    t_port, val = best(lambda: O.fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri, E_elem=E_elem))
    out = {"unit": "elems/s", "nproc": os.cpu_count()}
    if O.ref_lib() is None:
        out.update({"value": m.ne / t_port, "cores": 1, "kind": "port",
                    "sample": f"oracle fill_csc over {what} ({t_port:.2f}s); oracle/_ref not built here"})
        return out, (cp, ri, val)
    t1, (v1, _) = best(lambda: O.ref_fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri, threads=1,
                                              E_elem=E_elem))
    nt = cpu_threads()
    tn, (vn, used) = best(lambda: O.ref_fill_csc(m.conn, m.etype, m.xyz, resort, 210000.0, 0.3, cp, ri, threads=nt,
                                                 E_elem=E_elem))
    out.update({"value": m.ne / t1, "cores": 1, "kind": "reference",
                "sample": f"the reference's element stiffness + CSC scatter over {what}, pattern prebuilt "
                          f"({t1:.2f}s on one core)",
                "all_cores": {"value": m.ne / tn, "cores": used, "s": round(tn, 3)},
                "port": {"value": m.ne / t_port, "cores": 1, "s": round(t_port, 3),
                         "port_time_over_reference_time": round(t_port / t1, 3)}})
    return out, (cp, ri, val)


def cpu_baseline_solve(K, F):
    """Oracle Jacobi PCG (the plain FP64 method, one core) to rel. residual 1e-8."""
    from oracle import oracle as O

    cp, ri, val = K
    t = time.perf_counter()
    _, it, rel = O.pcg_jacobi(cp, ri, val, F, rtol=1e-8, maxit=100000)
    return (time.perf_counter() - t) * 1e3, it, rel


def train_systems(specs, args):
    """Trains the learned correction on the mean loss tr(G K G^T) / n of
    several systems (one context each): specs are (label, mesh, resort, E_elem
    or None).  Returns the parameters and a summary."""
    import openfea_amd as ofa
    from openfea_amd import learned as L, meshgen as mg

    t0 = time.perf_counter()
    ctxs, lns = [], []
    for label, m, resort, Ee in specs:
        c = ofa.Context(0)
        c.set_mesh(m.xyz, m.conn, m.etype, resort)
        c.set_elem_E(Ee)
        c.assemble(0.0 if Ee is not None else mg.E_STEEL, mg.NU_STEEL, copy=False)
        ctxs.append(c)
        lns.append(L.Learned(c, seed=0))
    hist = L.train_multi(lns, steps=args.train_steps, lr=args.train_lr)
    params = lns[0].get_params()
    for c in ctxs:
        c.close()
    info = {"systems": [lbl for lbl, *_ in specs], "loss": "tr(G K G^T)/n (Kaporin, 1 = block Jacobi)",
            "steps": args.train_steps, "lr": args.train_lr, "loss_first": float(hist[0]),
            "loss_last": float(hist[-1]), "train_s": round(time.perf_counter() - t0, 3)}
    log(f"learned: loss {hist[0]:.4f} -> {hist[-1]:.4f} in {info['train_s']}s on {len(specs)} systems")
    return params, info


def train_learned(workload, args):
    """The learned correction for the static workloads, trained on smaller
    meshes that the benchmark does not solve: two C3D4 blocks of other jitter
    seeds and two C3D8 design variants."""
    from openfea_amd import meshgen as mg

    specs = []
    for seed in (7, 8):
        m, resort, _ = mg.cantilever_c3d4(24, 12, 12, seed=seed)
        specs.append((f"C3D4 24x12x12 jitter seed {seed}", m, resort, None))
    mh = mg.hex_block(16, 8, 8)
    rh = mg.free_index(mh.nd, mh.fix_set)
    for v in (0, 1):
        specs.append((f"C3D8 16x8x8 design variant {v}", mh, rh, mg.design_field(mh, v)))
    return train_systems(specs, args)


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
                   "learned_fp64": A.PCG_LEARNED, "learned_mixed_fp32val": A.PCG_LEARNED_MIXED,
                   "two_level_fp64": A.PCG_TWO_LEVEL, "two_level_mixed_fp32val": A.PCG_TWO_LEVEL_MIXED,
                   "two_level_mixed_fp32_all": A.PCG_TWO_LEVEL_MIXED32,
                   "learned_two_level_fp64": A.PCG_LEARNED_TWO_LEVEL,
                   "learned_two_level_mixed_fp32val": A.PCG_LEARNED_TWO_LEVEL_MIXED,
                   "learned_two_level_mixed_fp32val_bf16net": A.PCG_LEARNED_TWO_LEVEL_MIXED,
                   "learned_two_level_mixed_fp32_all": A.PCG_LEARNED_TWO_LEVEL_MIXED32}
    chosen = [] if args.no_solve else [m for m in args.solvers.split(",") if m]
    methods = [(m, all_methods[m]) for m in chosen]
    learned_info, params = {}, None
    if any(m.startswith("learned") for m in chosen):
        params, learned_info = train_learned(workload, args)
    ln = L.Learned(ctx, seed=0) if params is not None else None
    infer = {}
    if ln is not None:  # network inference for the resident K: FP32 vs BF16 on the matrix cores
        ln.set_params(params)
        learned_info["loss_on_benchmark_K"] = float(ln.loss_grad()[0])
        for prec, ptag in ((A.FP32, "fp32"), (A.BF16, "bf16_mfma")):
            ln.precision(prec)
            ts = []
            for _ in range(5):
                ctx.sync()
                t = time.perf_counter()
                ln.refresh()
                ctx.sync()
                ts.append(time.perf_counter() - t)
            infer[ptag] = round(statistics.median(ts) * 1e3, 3)
        log(f"learned inference ms: {infer}")
    ctx.set_rhs(F)
    for name, method in methods:
        if name.startswith("learned"):
            ln.precision(A.BF16 if name.endswith("bf16net") else A.FP32)
        ctx.sync()
        t = time.perf_counter()
        ctx.solve(None, method=method, rtol=1e-8, maxit=100000, copy=False)  # warm (setup, graph capture)
        first_ms = (time.perf_counter() - t) * 1e3
        ts = []
        for _ in range(args.solve_reps):
            ctx.sync()
            t = time.perf_counter()
            _, it, rel, ok = ctx.solve(None, method=method, rtol=1e-8, maxit=100000, copy=False)
            ts.append(time.perf_counter() - t)
        solves[name] = {"ms": round(statistics.median(ts) * 1e3, 3), "iters": it, "relres": rel, "converged": ok,
                        "first_call_ms_incl_setup": round(first_ms, 3)}
        log(f"solve {name}: {solves[name]}")
    base = solves.get("plain_fp64_jacobi", {}).get("ms")
    for k, v in solves.items():
        v["speedup_vs_plain"] = round(base / v["ms"], 3) if base else None

    cpu = None
    if not args.no_cpu_baseline:
        cpu, K = cpu_fill_baseline(m, resort, f"the full {m.ne}-element mesh")
        if workload == "cfg2" and not args.no_solve:
            ms, it, rel = cpu_baseline_solve(K, F)
            cpu["time_to_solve"] = {"ms": round(ms, 1), "iters": it, "relres": rel, "cores": 1, "kind": "port",
                                    "method": "oracle Jacobi PCG, FP64, one core, same system and stopping rule "
                                              "as plain_fp64_jacobi (the reference's own solver is PARDISO)"}
        log(f"cpu baseline: {json.dumps(cpu)}")
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
        "learned_inference_ms": infer,
        "setup_s": round(setup_s, 3),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_source(tag),
                     "kernel": "k_assemble_tet_chunks" if nn == 4 else "k_assemble_hex",
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }


def bench_explicit(args):
    import numpy as np

    import openfea_amd as ofa
    from openfea_amd import _capi as A, meshgen as mg
    from oracle import oracle as O

    m = mg.hex_block(100, 50, 50)
    wl = ("explicit C3D8 100x50x50 (250,000 hexes), lumped mass, central difference dt = 0.9 CFL, "
          "v_z = 1000 sin(pi x/L), x=0 clamped, 1,000 steps, FP32 state")
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
    nsteps = args.explicit_steps
    ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, 100, copy=False)  # warm-up + capture
    ctx.sync()
    ctx.timer_start()
    u32 = ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, nsteps, precision=A.FP32,
                           copy=False)
    ms = ctx.timer_stop()
    ms_step = ms / nsteps
    log(f"explicit fp32 {ms_step * 1e3:.2f} us/step")
    ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, 100, precision=A.FP64, copy=False)
    ctx.sync()
    ctx.timer_start()
    ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, nsteps, precision=A.FP64, copy=False)
    ms64 = ctx.timer_stop() / nsteps
    u32 = ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, nsteps, precision=A.FP32)
    u64 = ctx.explicit_run(mg.E_STEEL, mg.NU_STEEL, mass, fext, v0, fixed, dt, nsteps, precision=A.FP64)
    err = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    log(f"explicit fp64 {ms64 * 1e3:.2f} us/step; fp32 vs fp64 at step {nsteps}: {err:.3e}")
    # both states against the oracle's direct recurrence on the full mesh (K from the
    # oracle's fill over all nodes; the recurrence equals the reference's incremental loop)
    oracle_check = None
    if not args.no_oracle_check:
        t = time.perf_counter()
        allfree = np.arange(m.nd, dtype=np.int32)
        cpa, ria = O.build_csc(m.conn, m.etype, allfree, 3 * m.nd)
        kv = O.fill_csc(m.conn, m.etype, m.xyz, allfree, mg.E_STEEL, mg.NU_STEEL, cpa, ria)
        nt = cpu_threads()
        uo = O.explicit_recurrence(cpa, ria, kv, 1.0 / np.repeat(mass, 3), np.repeat(fixed, 3).astype(np.int32),
                                   v0.ravel(), dt, nsteps, threads=nt)
        del kv, cpa, ria
        no = np.linalg.norm(uo)
        oracle_check = {"fp32_rel_l2_err": float(np.linalg.norm(u32.ravel() - uo) / no),
                        "fp64_rel_l2_err": float(np.linalg.norm(u64.ravel() - uo) / no), "steps": nsteps,
                        "reference": "oracle direct recurrence u += v dt, v -= M^-1 K u dt (K from or_fill_csc), "
                                     f"FP64 on {nt} host threads", "s": round(time.perf_counter() - t, 1)}
        log(f"explicit vs oracle after {nsteps} steps: {oracle_check}")
    alg_bytes = 32 * m.ne + 64 * m.nd
    # FP32 operations of the moment-form hex kernel per element and step, counted in
    # its gfx950 code: 772 FMA (2 each) + 788 add/mul/sub
    flops = 2332 * m.ne
    ctx.close()
    cpu = None
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        k = 3
        O.explicit_run(m.conn, m.etype, m.xyz, mg.E_STEEL, mg.NU_STEEL, O.lumped_mass(m.conn, m.etype, m.xyz,
                       mg.RHO_STEEL), fext, v0, fixed, dt, k)
        el = time.perf_counter() - t
        cpu = {"value": k / el, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"oracle explicit_run (reference cal_in_force recurrence: Ke rebuilt per element per "
                         f"step) on the full mesh, {k} steps, FP64 ({el:.1f}s)"}
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9
    return {
        "metric": "explicit steps/s", "value": 1e3 / ms_step, "unit": "steps/s", "n_gpus": 1, "steps": nsteps,
        "warmup": 100, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": wl, "dt": dt, "elements": m.ne, "nodes": m.nd},
        "fp64_ms_per_step": ms64, "fp32_vs_fp64_rel_err_final": err, "vs_oracle_final": oracle_check,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("explicit_step_cfg3"),
                     "traffic_source": traffic_source("explicit_step_cfg3"),
                     "kernel": "k_tile_forces<float> + k_tile_boundary<float> (one step)",
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "fp32_flops_per_step": flops,
                     "fp32_tflops": flops / (ms_step * 1e-3) / 1e12, "fp32_peak_tflops": FP32_PEAK_TFLOPS,
                     "fp32_unpacked_peak_tflops": FP32_PEAK_TFLOPS / 2},
        "cpu_baseline": cpu,
    }


def bench_batched(args):
    import numpy as np

    import openfea_amd as ofa
    from openfea_amd import _capi as A, meshgen as mg
    from oracle import oracle as O

    nvar, nrhs = 64, 8
    m = mg.hex_block(40, 20, 20)
    resort = mg.free_index(m.nd, m.fix_set)
    wl = ("batched reanalysis C3D8 40x20x20 (16,000 hexes) x 64 design variants (E = 210000 f_v(x), f_v in "
          "[0.5,1.5]) x 8 load cases (1-10 MPa tractions on x=L)")
    ctx = ofa.Context(0)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    Ev = np.stack([mg.design_field(m, v) for v in range(nvar)])
    B = mg.batched_loads(m, resort, nvar, nrhs)
    nu = mg.NU_STEEL
    log(f"cfg4: ne={m.ne} n={ctx.n} nnz={ctx.nnz} variants {nvar} x {nrhs} load cases")
    ctx.assemble_batch(Ev, nu, copy=False)  # moduli resident in HBM from here on
    for _ in range(args.warmup):
        ctx.assemble_batch(None, nu, copy=False, nvar=nvar)
    ctx.sync()
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.assemble_batch(None, nu, copy=False, nvar=nvar)
    ms_step = ctx.timer_stop() / args.steps
    value = m.ne * nvar / (ms_step * 1e-3)
    log(f"batched assembly {ms_step:.3f} ms/launch -> {value:.3e} variant-elems/s")
    # the same 64 variants one ofea_assemble each
    ctx.sync()
    ctx.timer_start()
    for v in range(nvar):
        ctx.set_elem_E(Ev[v])
        ctx.assemble(0.0, nu, copy=False)
    ms_seq = ctx.timer_stop()
    ctx.set_elem_E(None)
    alg_bytes = 24 * m.nd + 4 * 8 * m.ne + 8 * m.ne * nvar + 8 * ctx.nnz * nvar
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9

    batched_solve = None
    if args.no_solve:
        ctx.close()
    else:
        batched_solve = bench_batched_solve(args, ctx, m, Ev, B, nu)
    cpu = None
    if not args.no_cpu_baseline:
        k = 8
        rates = []
        for v in range(k):
            c, _ = cpu_fill_baseline(m, resort, f"design variant {v} (per-element E)", E_elem=Ev[v])
            rates.append(c)
        cpu = dict(rates[0])
        cpu["unit"] = "variant-elems/s"
        cpu["sample"] = (f"{k} of the 64 variants, each filled on the host: " + cpu["sample"].split(" over ")[0] +
                         "; rates are the medians over the variants")
        cpu["value"] = statistics.median(r["value"] for r in rates)
        if "all_cores" in cpu:
            cpu["all_cores"] = {"value": statistics.median(r["all_cores"]["value"] for r in rates),
                                "cores": rates[0]["all_cores"]["cores"]}
            cpu["port"] = {"value": statistics.median(r["port"]["value"] for r in rates), "cores": 1}
    return {
        "metric": "batched assembled variant-elems/s + multi-RHS time-to-solve at rel-resid 1e-8",
        "value": value, "unit": "variant-elems/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "elements": m.ne, "variants": nvar, "load_cases": nrhs, "dofs": int(B.shape[2]),
                   "nnz": ctx.nnz},
        "sequential_assembly_ms_64_variants": ms_seq,
        "batched_solve": batched_solve,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("assemble_hex_batch_cfg4"),
                     "traffic_source": traffic_source("assemble_hex_batch_cfg4"),
                     "kernel": "k_assemble_hex_batch", "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }


def bench_batched_solve(args, ctx, m, Ev, B, nu):
    import numpy as np

    from openfea_amd import _capi as A

    nvar, nrhs = B.shape[0], B.shape[1]
    # multi-RHS solves of all 512 systems (rel. residual 1e-8, FP64 vectors, FP32 block-Jacobi)
    ctx.solve_batch(B[:1], rtol=1e-8, copy=False)  # tile autotune + graph capture
    t = time.perf_counter()
    _, it_all, rr_all, ok_all = ctx.solve_batch(B, rtol=1e-8, copy=False)
    t_all = time.perf_counter() - t
    g, spmv_ms = ctx.batch_tiles(nrhs)
    log(f"batched solve 512 systems {t_all:.3f}s iters {it_all.min()}..{it_all.max()} ok={ok_all} lanes={g}")
    # the same batch with the per-variant coarse space (A_c = Z^T K_v Z built and inverted on the device for
    # every variant inside the timed call, shared by its 8 load cases)
    ctx.solve_batch(B[:1], rtol=1e-8, copy=False, method=A.PCG_TWO_LEVEL)  # per-mesh structure + graph capture
    t = time.perf_counter()
    _, it_tl, rr_tl, ok_tl = ctx.solve_batch(B, rtol=1e-8, copy=False, method=A.PCG_TWO_LEVEL)
    t_tl = time.perf_counter() - t
    log(f"batched two-level solve 512 systems {t_tl:.3f}s iters {it_tl.min()}..{it_tl.max()} ok={ok_tl}")
    # like for like on the first 8 variants: one ofea_solve per system (plain FP64 Jacobi and block Jacobi)
    nv8 = 8
    t = time.perf_counter()
    _, it8, _, ok8 = ctx.solve_batch(B[:nv8], rtol=1e-8, copy=False)
    t_b8 = time.perf_counter() - t
    seq = {}
    seq_all = {"plain_fp64_jacobi": A.PCG_JACOBI, "fp64_block_jacobi": A.PCG_BLOCK_JACOBI,
               "two_level_fp64": A.PCG_TWO_LEVEL}
    for name in [x for x in args.seq_methods.split(",") if x]:
        meth = seq_all[name]
        tot, its = 0.0, []
        ctx.set_elem_E(Ev[0])
        ctx.assemble(0.0, nu, copy=False)
        ctx.solve(B[0, 0], method=meth, rtol=1e-8, copy=False)  # graph capture, per-mesh set-up
        for v in range(nv8):
            # every variant is a new K: the timed solves include its preconditioner set-up
            ctx.set_elem_E(Ev[v])
            ctx.assemble(0.0, nu, copy=False)
            for k in range(nrhs):
                ctx.set_rhs(B[v, k])
                ctx.sync()
                t = time.perf_counter()
                _, it, rel, ok = ctx.solve(None, method=meth, rtol=1e-8, copy=False)
                tot += time.perf_counter() - t
                its.append(it)
        seq[name] = {"s": round(tot, 4), "iters_mean": float(np.mean(its)),
                     "batched_speedup": round(tot / t_b8, 3)}
        log(f"sequential {name} on {nv8}x{nrhs}: {tot:.3f}s (batched {t_b8:.3f}s)")
    # all 512 systems one at a time with the two-level solver: each variant assembled (its own K and
    # coarse set-up, timed) and its 8 load cases solved
    t2 = None
    if "two_level_fp64" in [x for x in args.seq_methods.split(",") if x]:
        its2 = []
        t = time.perf_counter()
        for v in range(nvar):
            ctx.set_elem_E(Ev[v])
            ctx.assemble(0.0, nu, copy=False)
            for k in range(nrhs):
                ctx.set_rhs(B[v, k])
                _, it, rel, ok = ctx.solve(None, method=A.PCG_TWO_LEVEL, rtol=1e-8, copy=False)
                its2.append(it)
        ctx.sync()
        t2 = round(time.perf_counter() - t, 4)
        log(f"two-level one system at a time, all {nvar * nrhs} systems: {t2:.3f}s (batched {t_all:.3f}s), "
            f"iters {min(its2)}..{max(its2)}")
    ctx.set_elem_E(None)
    ctx.close()
    return {"systems": nvar * nrhs, "s": round(t_all, 4), "two_level_sequential_all_s": t2,
            "two_level_batched": {"s": round(t_tl, 4), "iters_min": int(it_tl.min()), "iters_max": int(it_tl.max()),
                                  "relres_max": float(rr_tl.max()), "converged": ok_tl,
                                  "vs_sequential_two_level": (round(t2 / t_tl, 3) if t2 else None)},
            "iters_min": int(it_all.min()),
            "iters_max": int(it_all.max()), "relres_max": float(rr_all.max()), "converged": ok_all,
            "lanes_per_node": g, "multi_rhs_spmv_ms": spmv_ms, "first_8_variants_s": round(t_b8, 4),
            "sequential_first_8_variants": seq}


def bench_nl(args):
    import numpy as np

    import openfea_amd as ofa
    from openfea_amd import _capi as A, learned as L, meshgen as mg
    from oracle import oracle as O

    m = mg.hex_block(80, 40, 40)
    resort = mg.free_index(m.nd, m.fix_set)
    E, nu = mg.E_STEEL, mg.NU_STEEL
    wl = ("geometric nonlinear C3D8 80x40x40 (128,000 hexes), TL St.Venant-Kirchhoff, z traction ramped in 10 "
          "increments to a linear tip deflection of 0.2 L, Newton to |r|/|F| <= 1e-8")
    ctx = ofa.Context(0)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    log(f"cfg5: ne={m.ne} n={ctx.n} nnz={ctx.nnz}")
    F1 = mg.free_rhs(mg.face_traction(m, (0.0, 0.0, 1.0)), resort)
    ctx.assemble(E, nu, copy=False)
    x1, _, _, _ = ctx.solve(F1, method=A.PCG_JACOBI, rtol=1e-10)
    uf = ofa.reset_displacement(resort, x1).reshape(-1, 3)
    L_ = 80.0
    tip = np.abs(m.xyz[:, 0] - L_) < 1e-9
    p_star = 0.2 * L_ / uf[tip, 2].mean()
    F = p_star * F1
    log(f"p* = {p_star:.3f} MPa")
    # TL tangent + internal force throughput at the linear solution
    u0 = 0.5 * p_star * x1
    for _ in range(args.warmup):
        ctx.nl_assemble(E, nu, u0, copy=False)
    ctx.sync()
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.nl_assemble(E, nu, u0, copy=False)
    ms_step = ctx.timer_stop() / args.steps
    value = m.ne / (ms_step * 1e-3)
    log(f"TL assembly {ms_step:.3f} ms -> {value:.3e} elems/s")
    runs, train = {}, None
    if not args.no_solve:
        runs, train = bench_newton(args, ctx, m, resort, E, nu, F, tip, L_)
    ctx.close()
    alg_bytes = 24 * m.nd + 4 * 8 * m.ne + 8 * ctx.nnz + 3 * 8 * ctx.n
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        cp, ri = O.build_csc(m.conn, m.etype, resort, 3 * int(resort.max() + 1))
        us = np.random.default_rng(0).normal(scale=1e-3, size=len(cp) - 1)
        t = time.perf_counter()
        O.nl_fill(m.conn, m.etype, m.xyz, resort, E, nu, us, cp, ri)
        el = time.perf_counter() - t
        cpu = {"value": m.ne / el, "unit": "elems/s", "cores": 1, "kind": "port",
               "sample": f"oracle nl_fill (TL tangent + internal force, reference nl fill algorithm) over the "
                         f"full mesh, one pass ({el:.1f}s)"}
    return {
        "metric": "TL tangent assembled elems/s + Newton time-to-solve (learned vs plain PCG inner solves)",
        "value": value, "unit": "elems/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "elements": m.ne, "dofs": ctx.n, "nnz": ctx.nnz, "p_star_MPa": p_star,
                   "inner_rtol": args.inner_rtol},
        "newton": runs, "learned_training": train,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("nl_assemble_hex_cfg5"),
                     "traffic_source": traffic_source("nl_assemble_hex_cfg5"),
                     "kernel": "k_nl_hex", "algorithmic_bytes_per_launch": alg_bytes,
                     "fp64_flops_per_launch": 185472 * m.ne,
                     "fp64_tflops": 185472 * m.ne / (ms_step * 1e-3) / 1e12, "fp64_peak_tflops": FP64_PEAK_TFLOPS},
        "cpu_baseline": cpu,
    }


def bench_newton(args, ctx, m, resort, E, nu, F, tip, L_):
    import openfea_amd as ofa
    from openfea_amd import _capi as A, learned as L, meshgen as mg

    # learned correction trained on the config-4 systems (40x20x20 C3D8, design variants 0..3)
    m4 = mg.hex_block(40, 20, 20)
    r4 = mg.free_index(m4.nd, m4.fix_set)
    params, train = train_systems([(f"C3D8 40x20x20 design variant {v}", m4, r4, mg.design_field(m4, v))
                                   for v in range(4)], args)
    ln = L.Learned(ctx, seed=0)
    ln.set_params(params)
    runs = {}
    newton_all = {"plain_fp64_jacobi": A.PCG_JACOBI, "fp64_block_jacobi": A.PCG_BLOCK_JACOBI,
                  "learned_fp64": A.PCG_LEARNED, "learned_fp64_bf16net": A.PCG_LEARNED,
                  "two_level_fp64": A.PCG_TWO_LEVEL, "learned_two_level_fp64": A.PCG_LEARNED_TWO_LEVEL,
                  "two_level_fp64_adaptive_inner": A.PCG_TWO_LEVEL,
                  "learned_two_level_fp64_adaptive_inner": A.PCG_LEARNED_TWO_LEVEL}
    for name in [x for x in args.newton_methods.split(",") if x]:
        meth = newton_all[name]
        ln.precision(A.BF16 if name.endswith("bf16net") else A.FP32)
        b0 = L.coarse_builds(ctx)
        ctx.sync()
        t = time.perf_counter()
        # "_adaptive_inner": Eisenstat-Walker forcing terms, no tighter than --inner-rtol
        irt = -args.inner_rtol if name.endswith("_adaptive_inner") else args.inner_rtol
        u, nt, lt, rel, ok = ctx.nl_solve(E, nu, F, n_incr=10, tol=1e-8, max_newton=25, method=meth, inner_rtol=irt)
        el = time.perf_counter() - t
        uz = ofa.reset_displacement(resort, u).reshape(-1, 3)[tip, 2].mean()
        runs[name] = {"s": round(el, 3), "newton_iters": nt, "pcg_iters": lt, "relres": rel, "converged": ok,
                      "tip_uz_over_L": float(uz / L_), "coarse_builds": L.coarse_builds(ctx) - b0}
        log(f"newton {name}: {runs[name]}")
    base = runs.get("plain_fp64_jacobi", {}).get("s")
    for v in runs.values():
        v["speedup_vs_plain"] = round(base / v["s"], 3) if base else None
    tl = [k for k in runs if "two_level" in k]
    if tl:  # the coarse set-up each new tangent costs the two-level inner solves (A_c, Cholesky, inverse)
        ts = []
        for _ in range(3):
            ctx.nl_assemble(E, nu, u, copy=False)  # a new K: the coarse space is rebuilt on its next use
            ctx.sync()
            t = time.perf_counter()
            L.coarse_aggregates(ctx)
            ts.append(time.perf_counter() - t)
        setup_ms = statistics.median(ts) * 1e3
        for k in tl:
            v = runs[k]
            v["coarse_setup_ms_per_build"] = round(setup_ms, 3)
            v["coarse_setup_share"] = round(v["coarse_builds"] * setup_ms / (v["s"] * 1e3), 3)
        log(f"newton: coarse set-up per build {setup_ms:.3f} ms")
    return runs, train


def hole_mesh(m, x0, z0, radius):
    """Elements of m whose centroid lies in the cylinder (x - x0)^2 + (z - z0)^2 < r^2
    (a hole through the y extent), and the mesh without them (nodes renumbered,
    orphans dropped)."""
    import numpy as np

    from openfea_amd import meshgen as mg

    c = mg.element_centroids(m)
    dele = np.nonzero((c[:, 0] - x0) ** 2 + (c[:, 2] - z0) ** 2 < radius ** 2)[0]
    keep = np.setdiff1d(np.arange(m.ne), dele)
    used = np.unique(m.conn[keep, :4] - 1)
    new_id = -np.ones(m.nd, dtype=np.int32)
    new_id[used] = np.arange(len(used), dtype=np.int32)
    conn = np.zeros((len(keep), 8), dtype=np.int32)
    conn[:, :4] = new_id[m.conn[keep, :4] - 1] + 1
    fix = new_id[m.fix_set - 1]
    sub = mg.Mesh(m.xyz[used], conn, m.etype[keep], np.sort(fix[fix >= 0] + 1), m.dims)
    return dele, sub, new_id


def bench_ca(args):
    import numpy as np

    import openfea_amd as ofa
    from openfea_amd import _capi as A, meshgen as mg
    from oracle import oracle as O

    def case(nx, ny, nz):
        m, resort, F = mg.cantilever_c3d4(nx, ny, nz)
        dele, sub, new_id = hole_mesh(m, nx / 2.0, nz / 2.0, nz / 7.5)
        return m, resort, F, dele, sub, new_id

    m, resort, F, dele, sub, new_id = case(60, 30, 30)
    E, nu = mg.E_STEEL, mg.NU_STEEL
    nb = args.ca_basis
    wl = (f"CA reanalysis: C3D4 60x30x30 Kuhn (324,000 tets, jitter seed 20261203), {len(dele)} tets of a "
          f"radius-4h hole through y at x = 30, z = 15 deleted, {nb}-vector CA basis, K0 solves two-level PCG "
          f"to 1e-10, vs a full two-level PCG solve of the modified mesh")
    ctx = ofa.Context(0)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    ctx.assemble(E, nu, copy=False)
    u0, it0, _, ok0 = ctx.solve(F, method=A.PCG_TWO_LEVEL, rtol=1e-10)
    log(f"ca: ne={m.ne} deleted {len(dele)}; K0 solve {it0} it ok={ok0}")
    kind = np.zeros(len(dele), dtype=np.int32)
    cm = (m.conn[dele] - 1).astype(np.int32)
    cm[:, 4:] = 0
    et = np.zeros(len(dele), dtype=np.int32)
    for _ in range(2):  # warm-up (workspace, graphs)
        ctx.ca_delta(E, nu, kind, cm, et, copy=False)
        ctx.ca_solve(nb, u0, F, method=A.PCG_TWO_LEVEL, rtol=1e-10)
    t_d, t_s = [], []
    for _ in range(args.solve_reps):
        ctx.sync()
        t = time.perf_counter()
        ctx.ca_delta(E, nu, kind, cm, et, copy=False)
        t_d.append(time.perf_counter() - t)
        t = time.perf_counter()
        u_ca, k, its, _, ok = ctx.ca_solve(nb, u0, F, method=A.PCG_TWO_LEVEL, rtol=1e-10)
        t_s.append(time.perf_counter() - t)
    ms_delta = statistics.median(t_d) * 1e3
    ms_ca = statistics.median(t_s) * 1e3
    log(f"ca: delta K {ms_delta:.2f} ms, CA solve {ms_ca:.2f} ms ({its} PCG it, basis {k}, ok={ok})")
    n = ctx.n
    nnz = ctx.nnz
    ctx.close()
    # full re-solve of the modified mesh (new pattern, new K, new preconditioner)
    r_sub = mg.free_index(sub.nd, sub.fix_set)
    F_full = ofa.reset_displacement(resort, F).reshape(-1, 3)
    used = new_id >= 0
    F_sub = np.zeros((sub.nd, 3))
    F_sub[new_id[used]] = F_full[used]
    F_sub_free = np.zeros(3 * int(r_sub.max() + 1))
    fr = r_sub >= 0
    F_sub_free.reshape(-1, 3)[r_sub[fr]] = F_sub[fr]
    c2 = ofa.Context(0)
    c2.sync()
    t = time.perf_counter()
    c2.set_mesh(sub.xyz, sub.conn, sub.etype, r_sub)
    c2.assemble(E, nu, copy=False)
    u_full_new, it_full, _, ok_full = c2.solve(F_sub_free, method=A.PCG_TWO_LEVEL, rtol=1e-10)
    ms_full_first = (time.perf_counter() - t) * 1e3
    c2.sync()
    t = time.perf_counter()
    _, it_full, _, _ = c2.solve(F_sub_free, method=A.PCG_TWO_LEVEL, rtol=1e-10, copy=False)
    ms_full_solve = (time.perf_counter() - t) * 1e3
    c2.close()
    # accuracy of CA against the full solution on the kept nodes
    u_ca_nodes = ofa.reset_displacement(resort, u_ca).reshape(-1, 3)[used]
    u_new_nodes = ofa.reset_displacement(r_sub, u_full_new).reshape(-1, 3)[new_id[used]]
    ca_err = float(np.linalg.norm(u_ca_nodes - u_new_nodes) / np.linalg.norm(u_new_nodes))
    u0_err = float(np.linalg.norm(ofa.reset_displacement(resort, u0).reshape(-1, 3)[used] - u_new_nodes) /
                   np.linalg.norm(u_new_nodes))
    log(f"full re-solve {ms_full_first:.1f} ms incl. pattern/assembly/setup, solve {ms_full_solve:.2f} ms; "
        f"CA rel err {ca_err:.3e} (unmodified u0: {u0_err:.3e})")
    # bandwidth of the K0 solves (the dominant kernels): SpMV + vector passes per iteration
    per_it = 12 * nnz + 4 * (n + 1) + 88 * n
    achieved = its * per_it / (ms_ca * 1e-3) / 1e9 if its else 0.0
    cpu = None
    if not args.no_cpu_baseline:
        ms_, rs_, Fs_, ds_, _, _ = case(30, 15, 15)
        t = time.perf_counter()
        ns = 3 * int(rs_.max() + 1)
        cps, ris = O.build_csc(ms_.conn, ms_.etype, rs_, ns)
        K0s = O.fill_csc(ms_.conn, ms_.etype, ms_.xyz, rs_, E, nu, cps, ris)
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        u0s = spla.splu(sp.csc_matrix((K0s, ris, cps), shape=(ns, ns))).solve(Fs_)
        cms = (ms_.conn[ds_] - 1).astype(np.int32)
        cms[:, 4:] = 0
        t1 = time.perf_counter()
        dKs = O.ca_fill(np.zeros(len(ds_), np.int32), cms, np.zeros(len(ds_), np.int32), ms_.xyz, ms_.xyz, rs_, E,
                        nu, cps, ris)
        O.ca_solve(cps, ris, K0s, dKs, u0s, Fs_, nb)
        el = time.perf_counter() - t1
        cpu = {"value": el * 1e3, "unit": "ms", "cores": 1, "kind": "port",
               "sample": f"oracle ca_fill + ca_solve (SuperLU factorisation of K0 inside, reference CA algorithm) "
                         f"on the 30x15x15 Kuhn mesh ({ms_.ne} tets, {len(ds_)} deleted), {nb} vectors; K0 and u0 "
                         f"prepared beforehand ({t1 - t:.1f}s)"}
    return {
        "metric": "CA reanalysis time (delta K + reduced solve) vs full re-solve", "value": ms_delta + ms_ca,
        "unit": "ms", "n_gpus": 1, "steps": args.solve_reps, "warmup": 2, "ms_per_step": ms_delta + ms_ca,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "elements": m.ne, "deleted": int(len(dele)), "dofs": n, "nnz": nnz},
        "delta_K_ms": ms_delta, "ca_solve_ms": ms_ca, "ca_pcg_iters": its, "ca_basis_kept": k,
        "full_resolve_ms_incl_setup": ms_full_first, "full_resolve_solve_only_ms": ms_full_solve,
        "full_resolve_pcg_iters": it_full, "ca_rel_err_vs_full": ca_err, "unmodified_u0_rel_err_vs_full": u0_err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_cg_spmv + k_cg_update of the K0 solves (per iteration 12 nnz + 4 (n+1) + 88 n "
                               "bytes; coarse-space terms not counted)",
                     "algorithmic_bytes_per_launch": per_it},
        "cpu_baseline": cpu,
    }


def bench_post(args):
    import numpy as np

    import openfea_amd as ofa
    from openfea_amd import _capi as A, meshgen as mg
    from oracle import oracle as O

    m, resort, F = mg.cantilever_c3d4()
    E, nu = mg.E_STEEL, mg.NU_STEEL
    wl = ("nodal strain recovery (get_cauchy_stress_3d): C3D4 60x30x30 Kuhn (324,000 tets), displacement = the "
          "resident two-level PCG solution, per-element B u_e + Mises form, nodal averages")
    ctx = ofa.Context(0)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    ctx.assemble(E, nu, copy=False)
    x, it, _, ok = ctx.solve(F, method=A.PCG_TWO_LEVEL, rtol=1e-10)
    for _ in range(args.warmup):
        ctx.post_nodal(E, nu, None, A.POST_STRAIN, copy=False)
    ctx.sync()
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.post_nodal(E, nu, None, A.POST_STRAIN, copy=False)
    ms_step = ctx.timer_stop() / args.steps
    value = m.ne / (ms_step * 1e-3)
    nv, _ = ctx.post_nodal(E, nu, None, A.POST_STRAIN)
    u_full = ofa.reset_displacement(resort, x)
    ref = O.post_node(m.conn, m.etype, m.nd, O.post_elem(m.conn, m.etype, m.xyz, u_full, E, nu, 0))
    err = max(float(np.abs(nv[c] - ref[c]).max() / np.abs(ref[c]).max()) for c in range(7))
    log(f"post: {ms_step * 1e3:.1f} us/call -> {value:.3e} elems/s; max rel err vs oracle {err:.2e}")
    n = ctx.n
    ctx.close()
    alg_bytes = 24 * m.nd + 8 * n + 4 * m.nd + 16 * m.ne + 56 * m.nd
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        O.post_node(m.conn, m.etype, m.nd, O.post_elem(m.conn, m.etype, m.xyz, u_full, E, nu, 0))
        el = time.perf_counter() - t
        cpu = {"value": m.ne / el, "unit": "elems/s", "cores": 1, "kind": "port",
               "sample": f"oracle post_elem + post_node (reference get_cauchy_stress_3d algorithm) over the full "
                         f"mesh, one pass ({el:.2f}s)"}
    return {
        "metric": "nodal strain recovery elems/s", "value": value, "unit": "elems/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "elements": m.ne, "nodes": m.nd},
        "max_rel_err_vs_oracle": err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("post_strain_cfg2"),
                     "traffic_source": traffic_source("post_strain_cfg2"),
                     "kernel": "k_post_chunks (one call)",
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }


def bench_sfem(args):
    import numpy as np

    import openfea_amd as ofa
    from openfea_amd import _capi as A, meshgen as mg
    from oracle import oracle as O

    m, resort, F = mg.cantilever_c3d4()
    E, nu = mg.E_STEEL, mg.NU_STEEL
    wl = ("node-based smoothed FEM (SFEM_C3D4): C3D4 60x30x30 Kuhn (324,000 tets, 58,621 smoothing domains), "
          "smoothed gradients + stabilisation, FP64 assembly into the smoothed CSC pattern + two-level PCG")
    ctx = ofa.Context(0)
    ctx.set_mesh(m.xyz, m.conn, m.etype, resort)
    t = time.perf_counter()
    ctx.sfem_setup()
    setup_s = time.perf_counter() - t
    log(f"sfem: setup {setup_s:.2f}s n={ctx.n} nnz={ctx.nnz}")
    for _ in range(args.warmup):
        ctx.sfem_assemble(E, nu, copy=False)
    ctx.sync()
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.sfem_assemble(E, nu, copy=False)
    ms_step = ctx.timer_stop() / args.steps
    value = m.ne / (ms_step * 1e-3)
    x, it, rel, ok = ctx.solve(F, method=A.PCG_TWO_LEVEL, rtol=1e-8)
    t = time.perf_counter()
    _, it, rel, ok = ctx.solve(F, method=A.PCG_TWO_LEVEL, rtol=1e-8, copy=False)
    solve_ms = (time.perf_counter() - t) * 1e3
    log(f"sfem assembly {ms_step * 1e3:.1f} us -> {value:.3e} elems/s; solve {solve_ms:.1f} ms ({it} it, ok={ok})")
    nnz, n = ctx.nnz, ctx.n
    ctx.close()
    alg_bytes = 24 * m.nd + 16 * m.ne + 8 * nnz
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        ms_, rs_, _ = mg.cantilever_c3d4(30, 15, 15)
        cp, ri = O.sfem_build_csc(ms_.conn, ms_.xyz, rs_)
        t = time.perf_counter()
        O.sfem_fill(ms_.conn, ms_.xyz, rs_, E, nu, cp, ri)
        el = time.perf_counter() - t
        cpu = {"value": ms_.ne / el, "unit": "elems/s", "cores": 1, "kind": "port",
               "sample": f"oracle sfem_fill (reference six-term domain stiffness + linear-search scatter) on the "
                         f"30x15x15 Kuhn mesh ({ms_.ne} tets), one pass ({el:.1f}s)"}
    return {
        "metric": "smoothed-FEM assembled elems/s", "value": value, "unit": "elems/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "elements": m.ne, "nodes": m.nd, "dofs": n, "nnz": nnz},
        "setup_s": round(setup_s, 3), "two_level_solve_ms": solve_ms, "solve_iters": it, "solve_converged": ok,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("sfem_assemble_cfg2"),
                     "traffic_source": traffic_source("sfem_assemble_cfg2"),
                     "kernel": "k_sfem_grad + k_sfem_gather_chunks (one assembly)", "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }


def compact_configs(args):
    """Configs 1, 3, 4 and 5 in short form for the default line: each one's
    timed figure, roofline (with the source of its traffic figure) and CPU
    baseline, from the same functions their own --workload runs use."""
    def sub(**kw):
        a = argparse.Namespace(**vars(args))
        a.steps, a.warmup, a.solve_reps = 10, 2, 3
        for k, v in kw.items():
            setattr(a, k, v)
        return a

    def pick(d, keys):
        return {k: d[k] for k in keys if k in d}

    keys = ["metric", "value", "unit", "ms_per_step", "dtype", "roofline", "cpu_baseline"]
    out = {}
    t = time.perf_counter()
    c1 = bench_static(sub(solvers="plain_fp64_jacobi,fp64_block_jacobi,two_level_fp64"), "cfg1")
    out["cfg1"] = dict(pick(c1, keys), workload=c1["config"]["workload"],
                       time_to_solve={k: pick(v, ["ms", "iters", "converged"]) for k, v in c1["time_to_solve"].items()})
    c3 = bench_explicit(sub(explicit_steps=1000))
    out["cfg3"] = dict(pick(c3, keys + ["fp64_ms_per_step", "fp32_vs_fp64_rel_err_final", "vs_oracle_final"]),
                       workload=c3["config"]["workload"])
    c4 = bench_batched(sub(seq_methods="two_level_fp64"))
    bs = c4.get("batched_solve") or {}
    out["cfg4"] = dict(pick(c4, keys + ["sequential_assembly_ms_64_variants"]), workload=c4["config"]["workload"],
                       batched_solve=pick(bs, ["systems", "s", "two_level_sequential_all_s", "two_level_batched",
                                               "iters_max", "converged"]))
    c5 = bench_nl(sub(newton_methods="fp64_block_jacobi,two_level_fp64,learned_fp64,learned_two_level_fp64,"
                                     "two_level_fp64_adaptive_inner"))
    out["cfg5"] = dict(pick(c5, keys), workload=c5["config"]["workload"],
                       newton={k: pick(v, ["s", "newton_iters", "pcg_iters", "converged", "coarse_builds",
                                           "coarse_setup_ms_per_build", "coarse_setup_share"]) for k, v in
                               (c5.get("newton") or {}).items()},
                       learned_training=pick(c5.get("learned_training") or {}, ["loss_first", "loss_last", "train_s"]))
    out["wall_s"] = round(time.perf_counter() - t, 1)
    log(f"configs block: {out['wall_s']} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg2", choices=["cfg1", "cfg2", "cfg3", "cfg4", "cfg5", "ca", "post", "sfem"])
    ap.add_argument("--explicit-steps", type=int, default=1000)
    ap.add_argument("--inner-rtol", type=float, default=1e-6, help="Newton inner PCG tolerance (cfg5)")
    ap.add_argument("--seq-methods", default="plain_fp64_jacobi,fp64_block_jacobi,two_level_fp64",
                    help="one-system-at-a-time solvers compared with the batched solve (cfg4)")
    ap.add_argument("--newton-methods", default="plain_fp64_jacobi,fp64_block_jacobi,learned_fp64,learned_fp64_bf16net,"
                    "two_level_fp64,learned_two_level_fp64,two_level_fp64_adaptive_inner,"
                    "learned_two_level_fp64_adaptive_inner",
                    help="inner PCG methods of the Newton runs (cfg5)")
    ap.add_argument("--solve-reps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true", help="cfg3: skip the full-size oracle trajectory")
    ap.add_argument("--no-configs", action="store_true", help="default run: leave out the configs 1/3/4/5 block")
    ap.add_argument("--no-solve", action="store_true", help="assembly only (for counter collection)")
    ap.add_argument("--solvers", default="plain_fp64_jacobi,fp64_block_jacobi,mixed_fp32val_block_jacobi,"
                    "mixed_fp32_all_block_jacobi,learned_fp64,learned_mixed_fp32val,two_level_fp64,"
                    "two_level_mixed_fp32val,two_level_mixed_fp32_all,learned_two_level_fp64,learned_two_level_mixed_fp32val,"
                    "learned_two_level_mixed_fp32_all,"
                    "learned_two_level_mixed_fp32val_bf16net")
    ap.add_argument("--train-steps", type=int, default=150)
    ap.add_argument("--train-lr", type=float, default=1e-2)
    ap.add_argument("--ca-basis", type=int, default=4, help="CA basis size (ca workload)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload in ("cfg1", "cfg2"):
        out = bench_static(args, args.workload)
        if args.workload == "cfg2" and not args.no_configs:
            out["configs"] = compact_configs(args)
    elif args.workload == "cfg3":
        out = bench_explicit(args)
    elif args.workload == "cfg4":
        out = bench_batched(args)
    elif args.workload == "ca":
        out = bench_ca(args)
    elif args.workload == "post":
        out = bench_post(args)
    elif args.workload == "sfem":
        out = bench_sfem(args)
    else:
        out = bench_nl(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
