"""Independent models of uneven size in one context (tf_set_models /
tf_run_models): every model, stepped with its own controller inside the
block-diagonal solve, reproduces its own single-model run; replayed step
graphs reproduce eager stepping."""
import dataclasses

import numpy as np
import pytest

from conftest import gpu_from_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def three_models():
    from temf3dt_amd import workloads
    ms = workloads.uneven_models(3, seed=5, tmin=2e4, tmax=6e4)
    # a short schedule with frequent doubling checks, so that the models
    # change dt at different steps; one receiver search radius for all
    rc = max(m.rec_charac for m in ms)
    return [dataclasses.replace(m, sp_double=12, time_maximum=1e-4, rec_charac=rc) for m in ms]


def _union_ctx(models):
    from temf3dt_amd import fetd, workloads
    u = workloads.union(models)
    fe = fetd.Fetd(0)
    fe.set_mesh(u["xyz"], u["tet_vex"])
    fe.set_models(u["ntet"], u["nrec"])
    fe.set_materials(u["sigma"], u["inv_mu"])
    m0 = models[0]
    for i, lp in enumerate(u["loops"]):
        fe.set_source(lp, *m0.source, isrc=i)
    fe.set_receivers(u["receivers"], max(m.rec_charac for m in models))
    fe.set_schedule(m0.sp_division, m0.sp_double, m0.sp_dblesize, m0.time_maximum)
    fe.assemble()
    return fe


@pytest.mark.parametrize("graphs", [True, False])
def test_models_match_single_model_runs(three_models, graphs):
    ms = three_models
    print(f"[gpu] models: {[m.ntet for m in ms]} tets, graphs={graphs}", flush=True)
    fe = _union_ctx(ms)
    fe.run_models(graphs=graphs)
    got = [fe.model_rows(i) for i in range(len(ms))]
    fe.close()
    for i, m in enumerate(ms):
        f1 = gpu_from_model(m)
        r1 = f1.run()
        f1.close()
        g = got[i]
        assert len(g["t"]) == len(r1["t"]), (i, len(g["t"]), len(r1["t"]))
        np.testing.assert_array_equal(g["step"], r1["step"])
        np.testing.assert_array_equal(g["t"], r1["t"])
        np.testing.assert_array_equal(g["dt"], r1["dt"])
        a, b = g["fields"], r1["fields"]
        scale = np.abs(b).max(axis=0) + 1e-300
        rel = (np.abs(a - b) / scale).max(axis=(0, 1))
        print(f"[gpu] model {i}: {len(r1['t'])} rows, dt changes {len(np.unique(r1['dt']))}, "
              f"max diff / scale per component {np.array2string(rel, precision=2)}", flush=True)
        # dB/dt and B: the curl of the field, well determined; E at a receiver in the
        # air also carries the gradient part only the air conductivity pins
        assert rel[3:].max() < 1e-10


def test_model_setup_errors(three_models):
    from temf3dt_amd import fetd, workloads
    ms = three_models
    u = workloads.union(ms)
    fe = fetd.Fetd(0)
    fe.set_mesh(u["xyz"], u["tet_vex"])
    with pytest.raises(fetd.FetdError):
        fe.set_models(u["ntet"][:2], u["nrec"][:2])  # tets do not add up
    fe.set_models(u["ntet"], u["nrec"])
    with pytest.raises(fetd.FetdError):
        fe.run()  # tf_run is for a single model
    fe.close()


def test_models_single_solve_autotuner(three_models):
    """The single-solve autotuner times a context of independent models
    (tf_time_solve over every model's current factor) and its knobs leave the
    models' rows unchanged beyond summation order."""
    from temf3dt_amd import autotune
    fe = _union_ctx(three_models)
    fe.run_models(max_steps=30, graphs=True)
    ref = [fe.model_rows(i)["fields"] for i in range(3)]
    tune = autotune.autotune_single(fe, reps=2, space={"s_units": (64, 8192), "s_chmin": (16, 512),
                                                      "s_tcols": (16, 32)})
    print(f"[gpu] models autotune: {tune['best']} {tune['best_ms']:.3f} ms", flush=True)
    assert tune["best_ms"] > 0 and len(tune["table"]) == 8
    for k, v in (("s_units", 64), ("s_chmin", 16), ("s_tcols", 16)):  # the most split setting
        fe.set_tuning(k, v)
    fe.run_models(max_steps=30, graphs=True)
    for i in range(3):  # dB/dt and B (E in the 1e-12 S/m air carries solver rounding)
        a, b = fe.model_rows(i)["fields"][..., 3:9], ref[i][..., 3:9]
        rel = np.abs(a - b).max() / np.abs(b).max()
        print(f"[gpu] model {i} with the split tiling vs autotuned: {rel:.3e}", flush=True)
        assert rel < 1e-10
    fe.close()
