"""Failure paths: each reference Kokkos::abort site forced from the host gives the same error
code in the device error word as in the oracle (every kernel still exits normally)."""
import numpy as np
import pytest

from cleo_amd import abi
from cleo_amd import populations as P
from tests.helpers import SEED, coll_params, cond_params, make_domain, motion_params

pytestmark = pytest.mark.gpu


def _both(gpu, oracle, pop):
    return make_domain(gpu, pop), make_domain(oracle, pop)


def _code(d, rc):
    return d.error()[0]


def test_cfl_violation(gpu, oracle):
    print("case: CFL abort", flush=True)
    pop = P.coupled(SEED, ndims=(4, 4, 2), mean=10)
    pop.state["wvel_lo"][:] = 500.0
    pop.state["wvel_hi"][:] = 500.0
    for d in _both(gpu, oracle, pop):
        rc = d.motion(motion_params())
        assert _code(d, rc) == 8


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_newton_maxiters(gpu, oracle, variant):
    print(f"case: Newton maxniters abort, condensation variant {variant}", flush=True)
    pop = P.condensation(SEED, ndims=(2, 2, 2), mean=20)
    pop.state["qvap"] *= 1.01
    p = cond_params(maxniters=1, rtol=0.0, atol=0.0)
    g, o = _both(gpu, oracle, pop)
    t = abi.Tiles(1, 1, 1, variant, 0, 0, 0, 0, 0)
    assert gpu.lib.cleo_set_tiles(g.h, abi.ctypes_byref(t)) == 0
    for d in (g, o):
        rc = d.condensation(p)
        assert _code(d, rc) == 4


def test_breakup_nfrags(gpu, oracle):
    print("case: breakup nfrags <= 2.5 abort", flush=True)
    pop = P.coal_breakup(SEED, ndims=(2, 2, 1), mean=60)
    # coalbure with coaleff 0: every enacted collision is a breakup
    p = coll_params(prob=abi.PROB_CONST, param=1e30, enact=abi.ENACT_COALBURE, nfrags=2.0, coaleff=0.0)
    for d in _both(gpu, oracle, pop):
        d.rng_counter(1)
        rc = d.collisions(p)
        assert _code(d, rc) == 7


def test_breakup_cke_nfrags_without_velocity(gpu, oracle):
    print("case: kinetic-energy nfrags with no terminal velocity (nfrags = 2.5) abort", flush=True)
    pop = P.coal_breakup(SEED, ndims=(2, 2, 1), mean=60)
    # CKE = 0 without a terminal velocity, so nfrags = 2/3 + 11/6 and breakup.hpp:198-201 aborts;
    # coalbure with coaleff 0 makes every enacted collision a breakup
    p = coll_params(prob=abi.PROB_CONST, param=1e30, enact=abi.ENACT_COALBURE, coaleff=0.0,
                    nfrags_kind=abi.NFRAGS_CKE, nfrags_tv=abi.TV_NULL)
    for d in _both(gpu, oracle, pop):
        d.rng_counter(1)
        rc = d.collisions(p)
        assert _code(d, rc) == 7


def test_timestep_below_minsubstep(gpu, oracle):
    print("case: delt < MINSUBTSTEP abort", flush=True)
    pop = P.condensation(SEED, ndims=(2, 2, 2), mean=5)
    p = cond_params(interval=100, minsub=2.0)
    for d in _both(gpu, oracle, pop):
        rc = d.condensation(p)
        assert rc != 0 and d.error()[0] == 11


def test_bad_injected_index_is_an_error(gpu):
    print("case: injected Fisher-Yates index out of range", flush=True)
    pop = P.coal_breakup(SEED, ndims=(1, 1, 1), mean=6)
    d = make_domain(gpu, pop)
    n = pop.n
    js = np.full(n - 1, 1000, np.uint32)
    d.rng_injected(js, np.full(2 * (n // 2), 0.5))
    d.collisions(coll_params())
    assert d.error()[0] == 12
