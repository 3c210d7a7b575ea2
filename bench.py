"""Benchmark: FP64 source-timesteps/s of the FETD forward path on one GPU.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--model batched3d|batched3d_refined|plain3d|small]

A "step" is one BDF2 time step of the reference's adaptive loop
(src/FETD_FWD.f90:1281-1443) for every source: right-hand side (SpMV with
Msigma + loop sources), forward/diagonal/backward substitution with the
LDL^T factors, receiver readout, and -- every sp_double steps -- the
doubling check (two fine solves, a trial factorisation, one coarse solve).
The timed region is K steps of the loop from t = 0, including every
factorisation that falls inside it; mesh, assembly and symbolic analysis
are set up before it (inputs resident in HBM).

Default workload (batched3d_refined): 64 transmitters (20 m loops on an
8 x 8 grid, 50 m apart) over the 3-D earth (0.01 S/m host, 1 S/m block,
1e-12 S/m air, 100 km cube), one receiver per transmitter, on a 10 m grid
whose tets are bisected to 1.25 m cells at every wire: 1,690,764 tets /
2,009,893 edges.  batched3d is the same on the unrefined 10 m grid
(1,346,700 tets).  `value` is the batched rate (64 sources per block solve,
one factorisation per dt); the plain single-source path is timed on the
same mesh over the same steps and reported beside it.

Prints one JSON line (with roofline and cpu_baseline objects).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def build_model(name):
    from temf3dt_amd import workloads
    return {"batched3d": workloads.batched_3d_model, "batched3d_refined": workloads.batched_3d_refined_model,
            "plain3d": workloads.plain_3d_model, "small": workloads.small_model,
            "multismall": workloads.multi_source_small_model}[name]()


def setup_gpu(m, loops, leaf=0):
    from temf3dt_amd import fetd
    fe = fetd.Fetd(0)
    fe.set_mesh(m.xyz, m.tet_vex)
    sig, imu = m.sigma_inv_mu()
    fe.set_materials(sig, imu)
    if len(loops) == 1:
        fe.set_source(m.loops[loops[0]], *m.source)
    else:
        fe.set_sources([m.loops[q] for q in loops], *m.source)
    fe.set_receivers(m.receivers, m.rec_charac)
    fe.set_schedule(m.sp_division, m.sp_double, m.sp_dblesize, m.time_maximum)
    fe.assemble()
    t = time.time()
    fe.analyze(leaf)
    return fe, time.time() - t


def timed_run(fe, steps, warmup, profile=False):
    if warmup > 0:
        fe.run(max_steps=warmup)
    if profile:
        fe.set_profiling(True)
    t0 = time.time()
    rows = fe.run(max_steps=steps)
    wall = time.time() - t0
    prof = fe.profile() if profile else None
    if profile:
        fe.set_profiling(False)
    return rows, wall, fe.timers(), prof


def measured_traffic(name, key="traffic_bytes_per_launch"):
    """HBM bytes per launch (or per step) from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes summarised in profiles/<name>, scaled by a 1 GiB device
    copy (the summary gives the commands); None if absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def host_cpu():
    """Host CPU description for the baseline: model name, logical CPUs of the
    machine and of this process's affinity set."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def adjoint_ratio(fe, steps, forward_wall, ckpt):
    """Gradient of a dB/dt misfit over the same bounded run: taped forward
    (checkpoints every `ckpt` steps) + backward adjoint sweep, over the plain
    forward's wall time, same mesh, source and schedule."""
    import numpy as np
    keep = fe.checks()["keep"]
    fe.set_replay(keep)
    fe.set_checkpointing(ckpt)
    t0 = time.time()
    rows = fe.run(max_steps=steps)
    t1 = time.time()
    w = np.zeros_like(rows["fields"])
    w[..., 5] = 1.0  # J = sum of dBz/dt over rows and receivers
    g = fe.adjoint(w)
    t2 = time.time()
    fe.set_checkpointing(0)
    fe.set_replay([])
    return {"steps": int(rows["step"][-1]), "checkpoint_every": ckpt, "taped_forward_s": t1 - t0,
            "backward_s": t2 - t1, "forward_s": forward_wall, "ratio": (t2 - t0) / forward_wall,
            "grad_finite": bool(np.all(np.isfinite(g)))}


def cpu_baseline(m, counts, nsrc):
    """The oracle (C restatement of the reference + MKL PARDISO, the
    reference's solver) on the host cores: one factorisation of transmitter
    0's system at the first dt, timed by phase (analysis 11, numeric 22),
    one solve with 1 and with nsrc right-hand sides on it (phase 33), and one
    right-hand side with its SpMV.  The GPU run's counts (steps, solves,
    factorisations) weight them into source-timesteps/s:
      reference    phase 12 (11 + 22) at every factorisation, one source per
                   run as the reference does (src/FETD_FWD.f90:1307,1361), so
                   nsrc transmitters are nsrc runs at this rate;
      symbolic     the analysis once, phase 22 at every factorisation;
      block        analysis once, and all nsrc sources through one phase-33
                   call with nrhs = nsrc per solve (PARDISO's multi-RHS).
    Threads: this process's CPUs, capped by OMP_NUM_THREADS (the host's
    share per GPU)."""
    import tempfile

    from oracle import pyoracle
    if not pyoracle.pardiso_available():
        return None
    cpu = host_cpu()
    threads = cpu["affinity"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    threads = int(os.environ.get("TF_CPU_THREADS", threads))
    pyoracle.set_threads(threads)
    t0 = time.time()
    with tempfile.TemporaryDirectory() as tmp:
        root = os.path.join(tmp, "cpu")
        m.write(root, loop=0)
        o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
        o.assemble()
        c = o.cpu_costs(m.source[2] / m.sp_division, nrhs_block=nsrc)
        o.close()
    wall = time.time() - t0
    steps, nsolves, nf = counts
    t12 = c["phase11_s"] + c["phase22_s"]
    ref = steps / (nf * t12 + nsolves * (c["phase33_s"] + c["rhs_s"]))
    sym = steps / (c["phase11_s"] + nf * c["phase22_s"] + nsolves * (c["phase33_s"] + c["rhs_s"]))
    blk = nsrc * steps / (c["phase11_s"] + nf * c["phase22_s"] + nsolves * (c["phase33_block_s"] + nsrc * c["rhs_s"]))
    return {"value": ref, "unit": "source-timesteps/s", "cores": threads, "kind": "port",
            "sample": (f"transmitter 0 of the same model on the host with MKL PARDISO (mtype {int(c['mtype'])}): "
                       f"phase 11 {c['phase11_s']:.2f} s, phase 22 {c['phase22_s']:.2f} s, phase 33 "
                       f"{c['phase33_s'] * 1e3:.0f} ms (1 rhs) / {c['phase33_block_s'] * 1e3:.0f} ms ({nsrc} rhs), "
                       f"rhs + SpMV {c['rhs_s'] * 1e3:.0f} ms; weighted by the GPU run's {steps} steps, {nsolves} "
                       f"solves and {nf} factorisations"),
            "variants": {"symbolic_once": sym, "pardiso_nrhs_block": blk},
            "host": cpu, "pardiso_factor_nnz": int(c["factor_nnz"]), "costs": c, "wall_s": wall}


def uneven_bench(steps, n=48):
    """Independent models of uneven size (20k-300k tets, seeded) stepped in
    one context (block-diagonal solve, per-model controllers, replayed step
    graphs) against the same models run one after another, each in its own
    context, over the same number of steps."""
    from temf3dt_amd import fetd, workloads
    ms = workloads.uneven_models(n)
    u = workloads.union(ms)
    fe = fetd.Fetd(0)
    fe.set_mesh(u["xyz"], u["tet_vex"])
    fe.set_models(u["ntet"], u["nrec"])
    fe.set_materials(u["sigma"], u["inv_mu"])
    m0 = ms[0]
    for i, lp in enumerate(u["loops"]):
        fe.set_source(lp, *m0.source, isrc=i)
    fe.set_receivers(u["receivers"], max(m.rec_charac for m in ms))
    fe.set_schedule(m0.sp_division, m0.sp_double, m0.sp_dblesize, m0.time_maximum)
    fe.assemble()
    t = time.time()
    fe.analyze()
    t_an = time.time() - t
    st = fe.factor_stats()
    print(f"[bench] {n} models: {int(u['ntet'].sum())} tets ({int(u['ntet'].min())}-{int(u['ntet'].max())}), "
          f"{fe.nE} edges, nnzL {st['nnzL']:.3e}, analyze {t_an:.2f} s", file=sys.stderr, flush=True)
    out = {}
    fe.run_models(max_steps=3, graphs=True)  # factors every model
    from temf3dt_amd import autotune
    tune = autotune.autotune_single(fe, log=lambda x: print(x, file=sys.stderr, flush=True))
    print(f"[bench] autotune: {tune['best']} ({tune['best_ms']:.3f} ms per solve of all models)", file=sys.stderr,
          flush=True)
    for graphs in (True, False):
        fe.run_models(max_steps=3, graphs=graphs)  # warm-up
        t = time.time()
        fe.run_models(max_steps=steps, graphs=graphs)
        wall = time.time() - t
        done = sum(len(fe.model_rows(i)["t"]) - 1 for i in range(n))
        out["graphs" if graphs else "eager"] = {"wall_s": wall, "model_steps": done, "value": done / wall,
                                                "timers": fe.timers()}
        print(f"[bench] one context, graphs={graphs}: {done} model-steps in {wall:.3f} s", file=sys.stderr,
              flush=True)
    fe.close()
    seq_wall, seq_steps = 0.0, 0
    for m in ms:
        f1, _ = setup_gpu(m, [0])
        for k, v in tune["best"].items():
            f1.set_tuning(k, v)
        f1.run(max_steps=3)
        t = time.time()
        r = f1.run(max_steps=steps)
        seq_wall += time.time() - t
        seq_steps += len(r["t"]) - 1
        f1.close()
    out["sequential"] = {"wall_s": seq_wall, "model_steps": seq_steps, "value": seq_steps / seq_wall}
    print(f"[bench] one context per model, sequential: {seq_steps} model-steps in {seq_wall:.3f} s", file=sys.stderr,
          flush=True)
    res = {"metric": "FP64 source-timesteps/sec, independent models of uneven size in one context",
           "value": out["graphs"]["value"], "unit": "source-timesteps/s", "n_gpus": 1, "steps": steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{n} independent models, {int(u['ntet'].sum())} tets in all", "modes": out,
                      "autotune": tune["best"],
                      "over_sequential": out["graphs"]["value"] / out["sequential"]["value"]}}
    print(json.dumps(res))


# counter-measured HBM traffic (profiles/) per workload: tile launch, plain step, factorisation
TRAFFIC_FILES = {
    "batched3d": ("hbm_traffic_batched3d.txt", "hbm_traffic_plain_step.txt", "hbm_traffic_factor.txt"),
    "batched3d_refined": ("hbm_traffic_batched3d_refined.txt", "hbm_traffic_plain_step_refined.txt",
                          "hbm_traffic_factor_refined.txt"),
}


def single_source_rooflines(prof, nsolves, nfact, st, n, nnzu, lay, model):
    """Roofline entries of a single-source context's profiled run: the step
    (HBM) and the numeric factorisation (FP64 MFMA).
    Step: bytes_1 = 16 nnzL_off (L read in both sweeps) + 8 n (D) + 8 sum of
    struct rows (i32 row indices, both sweeps) + 12 nnzU + 4 (n + 1) (upper
    CSR of Msigma) + 48 n (e2, e3, r, x and two work vectors), over the
    summed duration of its SpMV/right-hand-side and solve kernels per solve.
    Factorisation: F_fact = sum_j (c_j^2 + 2 c_j) flops over the summed
    duration of all factorisation kernels (update, panel, inversion,
    extend-add, scatter) per factorisation.  Traffic: the counter-measured
    bytes of this workload (TRAFFIC_FILES), None for workloads without."""
    files = TRAFFIC_FILES.get(model)
    out = {}
    nnzl_off = st["nnzL"] - n
    bytes1 = 16.0 * nnzl_off + 8.0 * n + 8.0 * lay["struct_rows"] + 12.0 * nnzu + 4.0 * (n + 1) + 48.0 * n
    step_ms = (prof["solve1"][1] + prof["step"][1]) / nsolves
    ach = bytes1 / (step_ms * 1e-3) / 1e9
    out["plain_step"] = {
        "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
        "traffic": measured_traffic(files[1], "traffic_bytes_per_step") if files else None,
        "kernel": (f"single right-hand-side step (k_spmv_rhs + forward/backward solve kernels): bytes_1 = "
                   f"{bytes1:.4e} B per step, {step_ms:.3f} ms per step ({nsolves} solves; HIP events)")}
    fact_ms = (prof["factor_update"][1] + prof["factor_other"][1]) / max(nfact, 1)
    fach = st["flops"] / (fact_ms * 1e-3) / 1e12
    out["factorisation"] = {
        "bound": "mfma", "achieved": fach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": fach / FP64_PEAK_TFLOPS,
        "traffic": measured_traffic(files[2], "traffic_bytes_per_factorisation") if files else None,
        "kernel": (f"multifrontal LDL^T (k_factor_* + inversion + extend-add): F_fact = {st['flops']:.4e} flops "
                   f"per factorisation, {fact_ms:.1f} ms each ({nfact} factorisations; HIP events)")}
    return out


def copy_probe():
    """Calibration for HBM counters: three device-to-device copies of 1 GiB
    (read 1 GiB, write 1 GiB each) through the HIP runtime the library links;
    FETCH_SIZE / WRITE_SIZE of these dispatches against the known bytes give
    the scale applied to the solver kernels' counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    nbytes = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}))


def stream_probe():
    """What HBM sustains on this device for plain streams (context for the
    fraction-of-8-TB/s figures): a read-only FP64 sum and a copy over 8 GiB
    buffers, timed with HIP events over 10 repetitions after a warm-up."""
    import torch
    n = (8 << 30) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    out = {}
    for name, fn, nbytes in (("read_sum", lambda: x.sum(), 8 * n), ("copy", lambda: y.copy_(x), 16 * n)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        out[name] = {"ms": ms, "GB/s": nbytes / ms / 1e6}
    print(json.dumps({"stream_probe_bytes": 8 * n, **out}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--model", default="batched3d_refined")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sources", type=int, default=0,
                    help="use only the first N transmitters (0: all); 1 makes the main run the plain path")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile-kernel autotuner (use defaults)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward measurement")
    ap.add_argument("--ckpt", type=int, default=50, help="adjoint checkpoint interval (steps)")
    ap.add_argument("--leaf", type=int, default=0, help="nested-dissection leaf size (0: the library default)")
    ap.add_argument("--copy-probe", action="store_true", help="1 GiB device copies only (counter calibration)")
    ap.add_argument("--stream-probe", action="store_true", help="sustained HBM read and copy rates only")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="tf_set_tuning key for the batched run (repeatable), applied before the autotuner")
    a = ap.parse_args()
    if a.stream_probe:
        stream_probe()
        return
    if a.copy_probe:
        copy_probe()
        return
    if a.model == "uneven48":
        uneven_bench(a.steps)
        return
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU sharding of the path is a design in "
              "DESIGN.md and is not launched", file=sys.stderr)
        sys.exit(2)
    m = build_model(a.model)
    if a.sources > 0:
        import dataclasses
        m = dataclasses.replace(m, loops=m.loops[:a.sources])
    nsrc = len(m.loops)
    fe, t_an = setup_gpu(m, list(range(nsrc)), a.leaf)
    for kv in a.set:
        key, val = kv.split("=")
        fe.set_tuning(key, int(val))
    st = fe.factor_stats()
    n = fe.nE
    nnzu = fe.nnz
    print(f"[bench] {a.model}: {m.ntet} tets, {n} edges, nnzU {fe.nnz}, nnzL {st['nnzL']:.3e}, "
          f"factor flops {st['flops']:.3e}, {nsrc} sources, analyze {t_an:.2f} s", file=sys.stderr, flush=True)
    tune = None
    if not a.no_tune and nsrc > 1:
        from temf3dt_amd import autotune
        fe.factor(0, m.source[2] / m.sp_division)
        tune = autotune.autotune(fe, slot=0, nrhs=nsrc, log=lambda x: print(x, file=sys.stderr, flush=True))
        fe.release(0)
        print(f"[bench] autotune: {tune['best']} ({tune['best_ms']:.2f} ms per solve, "
              f"{tune['seconds']:.1f} s)", file=sys.stderr, flush=True)
    rows, wall, tm, _ = timed_run(fe, a.steps, a.warmup)
    steps = int(rows["step"][-1])
    nf = int(tm["nfactor"])
    nchk = len(fe.checks()["step"])
    nsolves = steps + 2 * nchk  # a check step costs dblesize fine solves + 1 coarse solve
    value = nsrc * steps / wall
    print(f"[bench] batched: {steps} steps x {nsrc} sources in {wall:.3f} s, {nf} factorisations "
          f"({tm['factor']:.3f} s), t_end {rows['t'][-1]:.3e}", file=sys.stderr, flush=True)
    # the same run again with HIP events around every launch group (which
    # also turns off the step graphs) for the per-kernel breakdown, with the
    # check's factorisation at the check, so that no kernel of the second
    # stream runs inside a timed launch group
    fe.set_tuning("overlap", 0)
    _, _, tm_prof, prof = timed_run(fe, a.steps, 0, profile=True)
    fe.set_tuning("overlap", 1)
    print("[bench] kernel time by category (launches, ms): " +
          ", ".join(f"{k} {v[0]} {v[1]:.1f}" for k, v in prof.items()), file=sys.stderr, flush=True)
    lay = fe.factor_layout()
    fe.close()
    nnzl_off = st["nnzL"] - n
    rooflines = {}
    if nsrc > 1:
        # block-solve tile products (k_mtile, four forms): the factor's
        # off-diagonal entries (L21 and the strict part of L11^-1) each enter
        # one forward and one backward product with K right-hand sides,
        # 4 K nnzL_off flops per solve
        flops_solve = 4.0 * nsrc * nnzl_off
        tile_launches, tile_ms = prof["tile"]
        achieved = flops_solve * nsolves / (tile_ms * 1e-3) / 1e12
        files = TRAFFIC_FILES.get(a.model)
        traffic = measured_traffic(files[0]) if files else None
        rooflines["block_solve_tiles"] = {
            "bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
            "traffic_note": ("HBM bytes per k_mtile launch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on a "
                             f"1 GiB copy), profiles/{files[0]}") if traffic else None,
            "kernel": (f"k_mtile (block-solve tile products, {nsrc} right-hand sides): {nsolves} solves x "
                       f"4*K*nnzL_off = {flops_solve:.3e} flops each, {tile_launches} launches, "
                       f"{tile_ms:.1f} ms (HIP events; breakdown run with the factorisations at the checks)"),
            "step_breakdown_ms": {k: round(v[1], 2) for k, v in prof.items()}}
    if nsrc == 1:
        rooflines.update(single_source_rooflines(prof, nsolves, int(tm_prof["nfactor"]), st, n, nnzu, lay, a.model))
    plain = None
    adjoint = None
    if not a.no_plain:
        fp, _ = setup_gpu(m, [0])
        ptune = None
        if not a.no_tune:
            from temf3dt_amd import autotune
            fp.factor(0, m.source[2] / m.sp_division)
            ptune = autotune.autotune_single(fp, slot=0, log=lambda x: print(x, file=sys.stderr, flush=True))
            fp.release(0)
            print(f"[bench] plain autotune: {ptune['best']} ({ptune['best_ms']:.3f} ms per solve)", file=sys.stderr,
                  flush=True)
        prow, pwall, ptm, _ = timed_run(fp, a.steps, a.warmup)
        psteps = int(prow["step"][-1])
        pchk = len(fp.checks()["step"])
        plain = {"value": psteps / pwall, "steps": psteps, "wall_s": pwall, "factor_s": ptm["factor"],
                 "factorisations": int(ptm["nfactor"]), "autotune": ptune and ptune["best"]}
        print(f"[bench] plain: {psteps} steps in {pwall:.3f} s ({ptm['factor']:.3f} s factorising)",
              file=sys.stderr, flush=True)
        # kernel breakdown of the plain run (factorisations at the checks)
        fp.set_tuning("overlap", 0)
        _, _, pptm, pprof = timed_run(fp, a.steps, 0, profile=True)
        fp.set_tuning("overlap", 1)
        print("[bench] plain kernel time by category (launches, ms): " +
              ", ".join(f"{k} {v[0]} {v[1]:.1f}" for k, v in pprof.items()), file=sys.stderr, flush=True)
        rooflines.update(single_source_rooflines(pprof, psteps + 2 * pchk, int(pptm["nfactor"]), st, n, nnzu,
                                                 lay, a.model))
        if not a.no_adjoint:
            adjoint = adjoint_ratio(fp, a.steps, pwall, a.ckpt)
            print(f"[bench] adjoint: taped forward {adjoint['taped_forward_s']:.3f} s + backward "
                  f"{adjoint['backward_s']:.3f} s = {adjoint['ratio']:.2f} x the plain forward", file=sys.stderr,
                  flush=True)
        fp.close()
    roof = dict(rooflines.get("block_solve_tiles") or rooflines.get("plain_step") or {})
    roof["all"] = {k: {q: v[q] for q in ("bound", "achieved", "peak", "unit", "frac", "traffic")}
                   for k, v in rooflines.items()}
    cpu = None if a.no_cpu else cpu_baseline(m, (steps, nsolves, nf), nsrc)
    out = {
        "metric": "FP64 source-timesteps/sec at 1M tets (batched vs plain); adjoint/forward time",
        "value": value, "unit": "source-timesteps/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.model}: {nsrc}-transmitter BDF2 loop, block solve of {nsrc} sources per "
                               f"factorisation, {m.ntet} tets, {n} edges", "mode": "batched", "sources": nsrc,
                   "nnzL": st["nnzL"], "factorisations": nf,
                   "plain_single_source": plain, "autotune": tune and tune["best"],
                   "batched_over_plain": (value / plain["value"]) if plain else None,
                   "adjoint_over_forward": adjoint},
        "roofline": roof, "rooflines": rooflines, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
