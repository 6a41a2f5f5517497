"""Writes the golden fixtures of this directory with the CPU oracle (oracle/liboracle_serial.so).

Each fixture is a small, fully specified case: the input store (SoA drops), gridbox State, grid,
solute and the random stream, the process parameters, and the oracle's outputs.  The reference
holds no golden vectors for this path, so these pin the oracle against unintended change
(tests/test_golden.py) and give the HIP path fixed targets that need no oracle build.

    python tests/golden/make_golden.py          # rewrites coupled_4x4x3.npz, collisions_2x2x2.npz,
                                                # condensation_3x3x2.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from cleo_amd import abi  # noqa: E402
from cleo_amd import populations as P  # noqa: E402
from tests.helpers import SEED, coll_params, cond_params, injected_stream, make_domain, motion_params, sdm_params  # noqa: E402

SUPER_KEYS = ("sdgbxindex", "coord3", "coord1", "coord2", "xi", "radius", "msol", "sdid")
STATE_KEYS = tuple(abi.STATE_FIELDS)

# process parameters of every case (the arguments of tests/helpers.py's builders)
COND = dict(interval=100, alter=1, maxniters=50, rtol=1e-6, atol=1e-8, minsub=0.001)
COLL = dict(prob=abi.PROB_LONG, param=1.0, enact=abi.ENACT_COALBURE, nfrags=5.0, coaleff=0.85, interval=100)
MOTION = dict(interval=100, tv=abi.TV_SIMMEL)


def pack_population(pop):
    out = {"ndims": np.array(pop.ndims, np.int64), "periodic": np.array(pop.periodic, np.int64),
           "solute": np.array([pop.solute["rho"], pop.solute["mr"], pop.solute["ionic"]])}
    for i, h in enumerate(pop.halves):
        out[f"halves{i}"] = np.asarray(h, np.float64)
    for k in SUPER_KEYS:
        out[f"in_{k}"] = np.asarray(pop.supers[k])
    for k in STATE_KEYS:
        out[f"in_state_{k}"] = np.asarray(pop.state[k], np.float64)
    return out


def unpack_population(f):
    """Population of a fixture (the inputs every backend starts from)."""
    sol = f["solute"]
    supers = {k: f[f"in_{k}"].copy() for k in SUPER_KEYS}
    state = {k: f[f"in_state_{k}"].copy() for k in STATE_KEYS}
    halves = tuple(f[f"halves{i}"].copy() for i in range(3))
    return P.Population(tuple(int(v) for v in f["ndims"]), halves, supers, state,
                        dict(rho=float(sol[0]), mr=float(sol[1]), ionic=float(sol[2])),
                        periodic=tuple(int(v) for v in f["periodic"]))


def pack_outputs(d, prefix="out_"):
    s = d.download()
    out = {f"{prefix}{k}": s[k][: d.ntot] for k in SUPER_KEYS}
    out[f"{prefix}ndomain"] = np.array(s["ndomain"], np.int64)
    out[f"{prefix}refs"] = d.refs()
    st = d.state()
    for k in STATE_KEYS:
        out[f"{prefix}state_{k}"] = st[k]
    return out


def coupled_case(backend):
    """3 coupled steps (motion, sort, condensation, Long coalescence/breakup), Philox counter stream."""
    pop = P.coupled(SEED, ndims=(4, 4, 3), mean=24)
    d = make_domain(backend, pop)
    d.set_sdm(sdm_params(cond=cond_params(**COND), coll=coll_params(**COLL), motion=motion_params(**MOTION)))
    d.rng_counter(2024)
    for step in range(3):
        assert d.run_step(step * 100, (step + 1) * 100) == 0
    assert d.error()[0] == 0
    out = pack_population(pop)
    out.update(pack_outputs(d))
    st = d.stats()
    out["stats"] = np.array([st[k] for k in ("cond_drops", "coll_pairs", "coll_events", "nulls", "motion_drops")],
                            np.int64)
    out["rng_seed"] = np.array(2024, np.int64)
    d.close()
    return out


def collisions_case(backend):
    """2 collision events (CoalBuRe, Long) driven by an injected stream of Fisher-Yates indices and phis."""
    pop = P.coal_breakup(SEED, ndims=(2, 2, 2), mean=48)
    d = make_domain(backend, pop)
    out = pack_population(pop)
    rng = np.random.default_rng(7)
    for ev in range(2):
        js, ph = injected_stream(d.refs(), rng)
        out[f"js{ev}"], out[f"phis{ev}"] = js, ph
        d.rng_injected(js, ph)
        assert d.collisions(coll_params(**COLL)) == 0
        assert d.error()[0] == 0
        out.update(pack_outputs(d, prefix=f"ev{ev}_"))
    d.close()
    return out


def condensation_case(backend):
    """One condensation pass with the per-drop Newton iteration counts."""
    pop = P.condensation(SEED, ndims=(3, 3, 2), mean=20)
    d = make_domain(backend, pop)
    d.record_niters(True)
    assert d.condensation(cond_params(**COND)) == 0
    assert d.error()[0] == 0
    out = pack_population(pop)
    out.update(pack_outputs(d))
    out["niters"] = d.niters()
    d.close()
    return out


CASES = {"coupled_4x4x3": coupled_case, "collisions_2x2x2": collisions_case,
         "condensation_3x3x2": condensation_case}


def main():
    from oracle import pyoracle
    orc = pyoracle.load("serial")
    for name, fn in CASES.items():
        out = fn(orc)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out['in_xi'])} drops)")


if __name__ == "__main__":
    main()
