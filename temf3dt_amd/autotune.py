"""Autotuner for the solve's tile kernels.

The K-source block solve (mtile.hip) has five knobs (tf_set_tuning):
  mt_ks_n / mt_ks_t  k-steps per register stage of the N- and T-form tile
                     kernels (2: two stages in flight, 4: one wider stage)
  mt_units           work units a tile launch aims for (sets how far long
                     inner ranges are split into partial tiles)
  mt_chmin           smallest inner chunk of a split tile
  mt_wide            levels of at most this many fronts use 64-row tiles,
                     the others 32-row tiles
The single right-hand-side solve (one source, or independent models with one
source each) has its own three (autotune_single):
  s_units            work units a tile launch aims for
  s_chmin            smallest inner chunk of a split tile
  s_tcols            output columns per pass of the T-form tiles (16 or 32)
The best setting depends on the front-size profile of the mesh's elimination
tree, so it is measured on the model itself: after a factorisation, every
candidate setting times `reps` device-resident solves (tf_time_solve, HIP
events) and the fastest is kept.  The right-hand-side block width is reported
the same way (source-solves per second at each width), since the sources of
a run share one block of at most 64.
"""
import itertools
import time

SINGLE_SPACE = {
    "s_units": (2048, 4096, 8192, 16384),
    "s_chmin": (256, 512, 1024),
    "s_tcols": (16, 32),
}

DEFAULT_SPACE = {
    "mt_ks_n": (2, 4),
    "mt_ks_t": (2, 4),
    "mt_units": (2048, 4096, 8192),
    "mt_chmin": (256, 512),
    "mt_wide": (16, 64, 256),
}


def autotune(fe, slot=0, nrhs=64, reps=3, space=None, widths=(16, 32, 64), log=None):
    """Tune fe (a factored temf3dt_amd.fetd.Fetd) in place.

    Returns dict(best=settings, best_ms=..., table=[(settings, ms)], widths=[(w, ms, solves_per_s)],
    seconds=tuning wall time)."""
    t0 = time.time()
    space = space or DEFAULT_SPACE
    keys = list(space)
    table = []
    for vals in itertools.product(*(space[k] for k in keys)):
        cfg = dict(zip(keys, vals))
        for k, v in cfg.items():
            fe.set_tuning(k, v)
        ms = fe.time_solve(slot, nrhs, reps)
        table.append((cfg, ms))
        if log:
            log(f"[autotune] {cfg} -> {ms:.3f} ms per {nrhs}-RHS solve")
    best, best_ms = min(table, key=lambda x: x[1])
    for k, v in best.items():
        fe.set_tuning(k, v)
    wtab = []
    for w in widths:
        ms = fe.time_solve(slot, w, reps)
        wtab.append((w, ms, w / ms * 1e3))
        if log:
            log(f"[autotune] width {w}: {ms:.3f} ms, {w / ms * 1e3:.0f} source-solves/s")
    return dict(best=best, best_ms=best_ms, table=table, widths=wtab, seconds=time.time() - t0)


def autotune_single(fe, slot=0, reps=5, space=None, log=None):
    """Tune the single right-hand-side solve of fe in place: every setting of
    `space` times `reps` solves of one right-hand side (tf_time_solve; for a
    context of independent models, one solve of every model's current
    factor).  Returns dict(best, best_ms, table, seconds)."""
    t0 = time.time()
    space = space or SINGLE_SPACE
    keys = list(space)
    table = []
    for vals in itertools.product(*(space[k] for k in keys)):
        cfg = dict(zip(keys, vals))
        for k, v in cfg.items():
            fe.set_tuning(k, v)
        ms = min(fe.time_solve(slot, 1, reps) for _ in range(2))
        table.append((cfg, ms))
        if log:
            log(f"[autotune] {cfg} -> {ms:.3f} ms per single solve")
    best, best_ms = min(table, key=lambda x: x[1])
    for k, v in best.items():
        fe.set_tuning(k, v)
    return dict(best=best, best_ms=best_ms, table=table, seconds=time.time() - t0)
