"""Write full-precision oracle rows for the configuration-size models, with
air at 1e-12 S/m, for the GPU parity tests (tests/test_gpu_configs.py).

The oracle (the CPU restatement with MKL PARDISO, pinned byte for byte to
the reference's shipped output) runs each model as the reference would for
its first steps -- past the first doubling check at step 100 and its
refactorisation -- and the rows, times and check decisions are stored:

  plain3d/oracle_rows.npz    3-D plain large-loop model (0.01 S/m host,
                             1 S/m block, 800 m loop, 41 receivers),
                             924,768 tets, 110 steps
  batched3d/oracle_rows.npz  two transmitters (0 and 45) of the 64-transmitter
                             model (1,346,700 tets), each run on its own on the
                             schedule of transmitter 0 (the GPU block solve
                             replays it), 110 steps
  uneven48/oracle_rows.npz   three of the 48 uneven models (seed 3): the
                             smallest, the median and the largest, 110 steps
  batched3d_refined/oracle_rows.npz
                             transmitters 0 and 45 of the 64-transmitter
                             model on its locally refined mesh (1,690,764
                             tets, 1.25 m cells at the wires), as batched3d

The models come from temf3dt_amd/workloads.py (deterministic), so the GPU
tests rebuild the same meshes.  Sizes and the mesh fingerprint (tet count,
edge count, coordinate sum) are stored to catch a drifting generator.

Usage: python tests/golden/make_config_rows.py [plain3d] [batched3d] [uneven48] [batched3d_refined]
(about 15 minutes on 8 cores; 62 GB of memory is enough).
"""
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import pyoracle  # noqa: E402
from temf3dt_amd import workloads  # noqa: E402

STEPS = 110
BATCH_TX = (0, 45)
UNEVEN_PICK = ("min", "median", "max")


def fingerprint(m):
    return np.array([m.ntet, len(m.xyz), float(np.sum(m.xyz))])


def oracle_rows(m, loop=0, replay=None, steps=STEPS):
    with tempfile.TemporaryDirectory() as tmp:
        root = os.path.join(tmp, "m")
        m.write(root, loop=loop)
        o = pyoracle.Oracle(root + ".mesh", root + ".cntl")
        o.assemble()
        if replay is not None:
            o.set_replay(replay)
        t = time.time()
        r = o.run(max_steps=steps)
        print(f"  {o.nE} edges, {len(r['step'])} rows, checks {list(r['chk_keep'])}, "
              f"{r['nfact']} factorisations, factor nnz {r['factor_nnz']}, {time.time() - t:.1f} s", flush=True)
        nE = o.nE
        o.close()
    return r, nE


def save(sub, **arrays):
    d = os.path.join(HERE, sub)
    os.makedirs(d, exist_ok=True)
    np.savez_compressed(os.path.join(d, "oracle_rows.npz"), **arrays)


def plain3d():
    m = workloads.plain_3d_model()
    print(f"plain3d: {m.ntet} tets", flush=True)
    r, nE = oracle_rows(m)
    save("plain3d", fp=fingerprint(m), nE=nE, step=r["step"], t=r["t"], dt=r["dt"], fields=r["fields"],
         chk_step=r["chk_step"], chk_err=r["chk_err"], chk_keep=r["chk_keep"], factor_nnz=r["factor_nnz"])


def batched3d(name="batched3d", make=workloads.batched_3d_model):
    m = make()
    print(f"{name}: {m.ntet} tets, transmitters {BATCH_TX}", flush=True)
    out = {"fp": fingerprint(m), "tx": np.array(BATCH_TX)}
    keep = None
    for i, tx in enumerate(BATCH_TX):
        r, nE = oracle_rows(m, loop=tx, replay=keep)
        if keep is None:
            keep = list(r["chk_keep"])
            out["chk_keep"] = r["chk_keep"]
            out["chk_step"] = r["chk_step"]
            out["step"], out["t"], out["dt"] = r["step"], r["t"], r["dt"]
        assert np.array_equal(r["t"], out["t"])
        out[f"fields_{i}"] = r["fields"]
        out["nE"] = nE
    save(name, **out)


def batched3d_refined():
    batched3d("batched3d_refined", workloads.batched_3d_refined_model)


def uneven48():
    ms = workloads.uneven_models(48)
    sizes = np.array([m.ntet for m in ms])
    order = np.argsort(sizes, kind="stable")
    pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    print(f"uneven48: models {pick} of sizes {[int(sizes[p]) for p in pick]}", flush=True)
    out = {"pick": np.array(pick), "sizes": sizes}
    for i, k in enumerate(pick):
        r, nE = oracle_rows(ms[k])
        for key in ("step", "t", "dt", "fields", "chk_step", "chk_err", "chk_keep"):
            out[f"{key}_{i}"] = r[key]
        out[f"fp_{i}"] = fingerprint(ms[k])
        out[f"nE_{i}"] = nE
    save("uneven48", **out)


def main():
    pyoracle.set_threads(os.cpu_count() or 8)
    which = sys.argv[1:] or ["plain3d", "batched3d", "uneven48", "batched3d_refined"]
    for w in which:
        {"plain3d": plain3d, "batched3d": batched3d, "uneven48": uneven48, "batched3d_refined": batched3d_refined}[w]()


if __name__ == "__main__":
    main()
