"""Shared helpers for the parity tests: domains on either backend and field comparisons."""
import numpy as np

from cleo_amd import abi
from cleo_amd import populations as P

SEED = 20261202


def make_domain(backend, pop, extra=0):
    d = abi.Domain(backend, pop.n + extra, pop.ndims, *pop.halves, periodic=pop.periodic)
    d.upload(pop.supers)
    d.set_state(pop.state)
    d.set_solute(*P.solute_dimless(pop.solute))
    return d


def cond_params(interval=100, alter=1, maxniters=50, rtol=1e-6, atol=1e-8, minsub=0.001, warm=0, warm_iters=10):
    return abi.CondParams(interval, alter, maxniters, rtol, atol, minsub, warm, warm_iters)


def coll_params(prob=abi.PROB_LONG, param=1.0, enact=abi.ENACT_COALBURE, nfrags=5.0, coaleff=0.85, interval=100,
                nfrags_kind=abi.NFRAGS_CONST, flag_kind=abi.FLAG_CONST, prob_tv=abi.TV_NULL,
                nfrags_tv=abi.TV_NULL, flag_tv=abi.TV_NULL):
    return abi.CollParams(interval, prob, param, enact, nfrags, coaleff, nfrags_kind, flag_kind, prob_tv,
                          nfrags_tv, flag_tv)


def motion_params(interval=100, tv=abi.TV_SIMMEL):
    return abi.MotionParams(interval, tv)


def sdm_params(cond=None, coll=None, motion=None):
    s = abi.SDMParams()
    s.enable_cond = 1 if cond is not None else 0
    s.enable_colls = 1 if coll is not None else 0
    s.enable_motion = 1 if motion is not None else 0
    s.cond = cond if cond is not None else cond_params()
    s.coll = coll if coll is not None else coll_params()
    s.motion = motion if motion is not None else motion_params()
    return s


def rain_population(seed, ndims=(4, 4, 2), mean=128):
    """The coalescence/breakup population with radii log-uniform in 20 um .. 2.5 mm, so that the
    kinetic energies fall on both sides of every flag's boundaries."""
    pop = P.coal_breakup(seed, ndims=ndims, mean=mean)
    rng = np.random.default_rng(seed + 1)
    wet = np.exp(rng.uniform(np.log(20.0), np.log(2500.0), pop.n))
    pop.supers["radius"] = np.maximum(wet, pop.supers["radius"])
    return pop


def injected_stream(refs, rng):
    """Fisher-Yates indices and phis for one collision event over the gridboxes of `refs`."""
    js, phis = [], []
    for g in range(len(refs) - 1):
        n = int(refs[g + 1] - refs[g])
        if n >= 2:
            it = np.arange(n - 1, 0, -1, dtype=np.int64)
            js.append((rng.random(n - 1) * (it + 1)).astype(np.int64).clip(0, it).astype(np.uint32))
        phis.append(rng.random(2 * (n // 2)))
    j = np.concatenate(js) if js else np.zeros(0, np.uint32)
    p = np.concatenate(phis) if phis else np.zeros(0)
    return j.astype(np.uint32), p


def assert_supers_match(a, b, rtol=1e-12, exact_fp=False, what=""):
    """xi, sdgbxindex and sdid bit-exact; radius and solute mass within rtol relative; coordinates
    within rtol of max(|c|, 1e-3 x the largest |c| of the store).  A coordinate near zero (a drop
    that left through z = 0 ends a small step below it) is the difference of its last position and
    step, so its relative error says nothing of the arithmetic; the floor, 1e-3 of the domain's
    extent, keeps the bar at the domain's scale there."""
    assert a["ndomain"] == b["ndomain"], f"{what} ndomain {a['ndomain']} != {b['ndomain']}"
    for f in ("sdgbxindex", "xi", "sdid"):
        bad = np.nonzero(a[f] != b[f])[0]
        assert len(bad) == 0, f"{what} {f} differs at {bad[:10]} ({a[f][bad[:5]]} vs {b[f][bad[:5]]})"
    for f in ("coord3", "coord1", "coord2", "radius", "msol"):
        x, y = a[f], b[f]
        if exact_fp:
            assert np.array_equal(x, y, equal_nan=True), f"{what} {f} not bit-identical"
        else:
            both_nan = np.isnan(x) & np.isnan(y)  # the uninitialised drops a padded store holds
            x, y = np.where(both_nan, 0.0, x), np.where(both_nan, 0.0, y)
            den = np.maximum(np.abs(x), np.abs(y))
            if f.startswith("coord") and len(den):
                den = np.maximum(den, 1e-3 * np.max(np.where(np.isfinite(den), den, 0.0)))
            rel = np.where(den > 0, np.abs(x - y) / np.where(den > 0, den, 1), 0.0)
            assert not np.any(np.isnan(rel)), f"{what} {f} NaN on one side only"
            assert np.all(rel <= rtol), f"{what} {f} max rel diff {rel.max():.3e} at {np.argmax(rel)}"


def assert_state_match(a, b, rtol=1e-12, what=""):
    """temp, qvap, press, volume within rtol relative.  qcond changes by the condensed mass, a sum
    of mass changes that cancel near equilibrium, so it is bounded relative to qvap + qcond
    (the water it is exchanged with), per gridbox."""
    for f in ("temp", "qvap", "press", "volume"):
        x, y = a[f], b[f]
        den = np.maximum(np.abs(x), np.abs(y))
        rel = np.where(den > 0, np.abs(x - y) / np.where(den > 0, den, 1), 0.0)
        assert np.all(rel <= rtol), f"{what} state {f} max rel diff {rel.max():.3e}"
    scale = np.abs(a["qvap"]) + np.abs(a["qcond"])
    err = np.abs(a["qcond"] - b["qcond"])
    assert np.all(err <= rtol * scale), f"{what} state qcond max err/(qv+qc) {(err / scale).max():.3e}"
