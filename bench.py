"""Benchmark: FP64 source-timesteps/s of the FETD forward path on one GPU.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--model plain3d|small]

A "step" is one BDF2 time step of the reference's adaptive loop
(src/FETD_FWD.f90:1281-1443) for one source: right-hand side (SpMV with
Msigma + loop source), forward/diagonal/backward substitution with the
LDL^T factors, receiver readout, and -- every sp_double steps -- the
doubling check (two fine solves, a trial factorisation, one coarse solve).
The timed region is K steps of the loop from t = 0, including every
factorisation that falls inside it; the mesh, assembly and symbolic analysis
are set up before it (inputs resident in HBM).

Workload (default): the 3-D plain large-loop model, ~0.92M tets / 1.1M
edges (0.01 S/m half-space, 1 S/m block, 800 m loop, 41 receivers, 1e-12 S/m
air, 100 km cube), synthetic tensor-grid mesh.

Prints one JSON line with roofline and cpu_baseline objects.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def build_model(name):
    from temf3dt_amd import workloads
    if name == "plain3d":
        return workloads.plain_3d_model()
    if name == "small":
        return workloads.small_model()
    raise SystemExit(f"unknown model {name}")


def setup_gpu(m):
    from temf3dt_amd import fetd
    fe = fetd.Fetd(0)
    fe.set_mesh(m.xyz, m.tet_vex)
    sig, imu = m.sigma_inv_mu()
    fe.set_materials(sig, imu)
    fe.set_source(m.loops[0], *m.source)
    fe.set_receivers(m.receivers, m.rec_charac)
    fe.set_schedule(m.sp_division, m.sp_double, m.sp_dblesize, m.time_maximum)
    fe.assemble()
    t = time.time()
    fe.analyze()
    return fe, time.time() - t


def cpu_baseline(m, gpu_counts, sample_steps):
    """Oracle (C restatement + MKL PARDISO, the reference's solver) timed on
    the host cores for the first `sample_steps` steps of the same model; the
    per-factorisation and per-solve times are then weighted by the GPU run's
    factorisation and solve counts to express the rate in source-timesteps/s."""
    import tempfile

    from oracle import pyoracle
    if not pyoracle.pardiso_available():
        return None
    threads = int(os.environ.get("TF_CPU_THREADS", os.cpu_count() or 1))
    pyoracle.set_threads(threads)
    with tempfile.TemporaryDirectory() as tmp:
        root = os.path.join(tmp, "cpu")
        m.write(root)
        o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
        o.assemble()
        t0 = time.time()
        r = o.run(max_steps=sample_steps)
        wall = time.time() - t0
    tm = r["timers"]
    nfact = max(1, r["nfact"])
    nsol = max(1, len(r["step"]) - 1)
    t_fact = tm[1] / nfact
    t_step = (wall - tm[1]) / nsol
    steps, nf = gpu_counts
    est = steps / (nf * t_fact + steps * t_step)
    return {"value": est, "unit": "source-timesteps/s", "cores": threads, "kind": "port",
            "sample": (f"first {sample_steps} steps of the same model on the host with MKL PARDISO "
                       f"({nfact} factorisations, {t_fact:.2f} s each; {t_step * 1e3:.1f} ms per step), "
                       f"weighted by the GPU run's {nf} factorisations over {steps} steps"),
            "factor_nnz_pardiso": int(r["factor_nnz"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--model", default="plain3d")
    ap.add_argument("--cpu-sample", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding is described in DESIGN.md "
              "and not launched", file=sys.stderr)
        sys.exit(2)
    m = build_model(a.model)
    fe, t_an = setup_gpu(m)
    st = fe.factor_stats()
    print(f"[bench] model {a.model}: {m.ntet} tets, {fe.nE} edges, nnzU {fe.nnz}, nnzL {st['nnzL']:.3e}, "
          f"factor flops {st['flops']:.3e}, analyze {t_an:.2f} s", file=sys.stderr, flush=True)
    if a.warmup > 0:
        fe.run(max_steps=a.warmup)
    t0 = time.time()
    rows = fe.run(max_steps=a.steps)
    wall = time.time() - t0
    tm = fe.timers()
    steps = int(rows["step"][-1])
    nf = int(tm["nfactor"])
    value = steps / wall
    print(f"[bench] {steps} steps in {wall:.3f} s, {nf} factorisations ({tm['factor']:.3f} s), "
          f"t_end {rows['t'][-1]:.3e}", file=sys.stderr, flush=True)
    # dominant-kernel roofline: the factorisation (FP64 flops) and the solve
    # (bytes of L streamed per step); both reported, factor is the headline
    fact_tflops = st["flops"] / (tm["factor"] / max(nf, 1)) / 1e12 if nf else 0.0
    roof = {"bound": "mfma", "achieved": fact_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": fact_tflops / FP64_PEAK_TFLOPS, "traffic": None,
            "kernel": "multifrontal LDL^T (all factor kernels, event-timed per factorisation)"}
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(m, (steps, nf), a.cpu_sample)
    out = {
        "metric": "FP64 source-timesteps/sec at 1M tets (batched vs plain); adjoint/forward time",
        "value": value, "unit": "source-timesteps/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.model}: plain single-source BDF2 loop, {m.ntet} tets, {fe.nE} edges, "
                               f"{len(m.receivers)} receivers", "mode": "plain", "sources": 1,
                   "nnzL": st["nnzL"], "factorisations": nf},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    fe.close()


if __name__ == "__main__":
    main()
