"""Autotuner of the step's launch shape and linked-cell size.

Times short windows of real steps on the scene itself and keeps the fastest
setting, by coordinate descent: first the cell-size factor
(CellSizeFactor >= 1 scales the cell edge 2 max R_c, Grains.cpp:232-242),
then the block size of the per-particle / per-entry kernels, then the number
of resident blocks per CU of the persistent narrow-phase launch, then one
narrow pass or a screening pass plus a pass over the queued entries, then
the candidate search with or without LDS staging.  Every
window replays a captured graph, the production mode.  The state
is restored from a snapshot before every window, so all candidates run the
same steps.
"""
import time

from . import _abi as A
from .scene import load


def _window(eng, blob, steps):
    # the production mode: steps replayed from a captured graph.  The first
    # call after a layout or launch change allocates, uploads and captures.
    eng.restore(blob)
    eng.step(1, graph=True)
    eng.restore(blob)
    t0 = time.perf_counter()
    eng.step(steps, graph=True)
    return (time.perf_counter() - t0) / steps


def autotune(scene, precision=A.FP32_MIXED, flags=0, gjk_rel_tol=0.0, steps=10,
             cell_factors=(1.0, 1.25, 1.5), blocks=(64, 128, 256), blocks_per_cu=(4, 8, 16),
             narrow_passes=(2, 1), candidate_staging=(1, 2), max_neighbors=64, log=None):
    """Returns dict(cell_size_factor, block_size, narrow_blocks_per_cu,
    narrow_passes, candidate_staging, ms_per_step, trials) for `scene` (all
    its sims share the factor)."""
    from .engine import Engine
    eng = load(Engine(dt=scene.dt, precision=precision, flags=flags, gjk_rel_tol=gjk_rel_tol,
                      max_neighbors=max_neighbors), scene)
    blob = eng.snapshot()
    trials = []
    best = dict(cell_size_factor=1.0, block_size=256, narrow_blocks_per_cu=8, narrow_passes=2, candidate_staging=0)

    def run(cfg):
        for s in scene.sims:
            s["cell_size_factor"] = cfg["cell_size_factor"]
        eng.set_sims(scene)
        eng.set_launch(cfg["block_size"], cfg["narrow_blocks_per_cu"])
        eng.set_narrow_passes(cfg["narrow_passes"])
        eng.set_candidate_staging(cfg["candidate_staging"])
        try:
            t = _window(eng, blob, steps)
        except Exception as ex:  # e.g. candidate capacity exceeded at a large cell
            if log:
                log(f"autotune {cfg}: {ex}")
            return float("inf")
        trials.append(dict(cfg, ms_per_step=t * 1e3))
        if log:
            log(f"autotune {cfg}: {t * 1e3:.3f} ms/step")
        return t

    best_t = float("inf")
    for key, values in (("cell_size_factor", cell_factors), ("block_size", blocks),
                        ("narrow_blocks_per_cu", blocks_per_cu), ("narrow_passes", narrow_passes),
                        ("candidate_staging", candidate_staging)):
        for v in values:
            cfg = dict(best, **{key: v})
            t = run(cfg)
            if t < best_t:
                best_t, best = t, cfg
    for s in scene.sims:
        s["cell_size_factor"] = best["cell_size_factor"]
    eng.close()
    return dict(best, ms_per_step=best_t * 1e3, trials=trials)
