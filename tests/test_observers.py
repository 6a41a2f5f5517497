"""Mass-moment observer (observers/massmoments_observer.cpp:23-73): the oracle's restatement pinned
by independent evaluations, and the HIP reduction against it.

Pins of the oracle (CPU):
- mom0 is the integer sum of xi; mom1 and mom2 agree with float64 sums of xi m and xi m^2 where m
  is evaluated independently in SI units (water shell + solute), to float32 accuracy;
- the float sums run in the order of the gridbox's drops: with per-drop terms read back from
  one-drop gridboxes, a big term followed by two small ones rounds differently from the reverse
  order, and each equals numpy's float32 left-to-right sum;
- the raindrop filter keeps radius >= 40e-6 / R0 exactly (39.999, 40, the threshold itself, 40.001);
- a multiplicity of 2^64 - 1 is reported.

GPU (marked gpu): bit-exact mom0, mom1, mom2 against the oracle on the same drops, all drops and
raindrops, after coupled steps (collisions leave a pending order and nulls)."""
import numpy as np
import pytest

from cleo_amd import abi
from cleo_amd import populations as P
from tests.helpers import SEED, coll_params, cond_params, make_domain, motion_params, sdm_params


def _box(ndims, supers, solute=P.AMMONIUM_SULPHATE):
    halves = P.uniform_grid(ndims, 25.0)
    n = len(supers["xi"])
    ngbx = int(np.prod(ndims))
    g = np.asarray(supers["sdgbxindex"], dtype=np.int64)
    nz, nx = ndims[0], ndims[1]
    iz, ix, iy = g % nz, (g // nz) % nx, g // (nz * nx)
    full = dict(sdgbxindex=g.astype(np.uint32),
                coord3=0.5 * (halves[0][iz] + halves[0][iz + 1]),
                coord1=0.5 * (halves[1][ix] + halves[1][ix + 1]),
                coord2=0.5 * (halves[2][iy] + halves[2][iy + 1]),
                xi=np.asarray(supers["xi"], dtype=np.uint64), radius=np.asarray(supers["radius"], dtype=np.float64),
                msol=np.asarray(supers["msol"], dtype=np.float64), sdid=np.arange(n, dtype=np.uint64))
    state = P._state(ngbx, (25.0 / P.COORD0) ** 3, 288.15, 1e5, 1.0)
    return P.Population(ndims, halves, full, state, solute)


def _mass_si(r_dimless, msol_dimless, solute):
    """Drop mass in kg: a water sphere of the drop's radius whose solute part has the solute's
    density instead of water's."""
    r = r_dimless * P.R0
    msol = msol_dimless * P.MASS0
    return 4.0 / 3.0 * np.pi * r ** 3 * P.RHO_L + msol * (1.0 - P.RHO_L / solute["rho"])


def test_oracle_moments_against_si_evaluation(oracle):
    rng = np.random.default_rng(SEED)
    ndims = (3, 2, 2)
    ngbx = 12
    counts = rng.integers(0, 40, ngbx)
    g = np.repeat(np.arange(ngbx), counts)
    n = len(g)
    r = np.exp(rng.uniform(np.log(0.5), np.log(300.0), n))
    msol = (4.0 / 3.0 * np.pi * (0.05 * r * P.R0) ** 3 * P.AMMONIUM_SULPHATE["rho"]) / P.MASS0
    xi = rng.integers(1, 2 ** 40, n).astype(np.uint64)
    pop = _box(ndims, dict(sdgbxindex=g, xi=xi, radius=r, msol=msol))
    d = make_domain(oracle, pop)
    m0, m1, m2 = d.mass_moments()
    m_kg = _mass_si(r, msol, P.AMMONIUM_SULPHATE)
    for gb in range(ngbx):
        sel = g == gb
        assert int(m0[gb]) == int(xi[sel].astype(object).sum())
        e1 = float(np.sum(xi[sel].astype(np.float64) * m_kg[sel])) / P.MASS0
        e2 = float(np.sum(xi[sel].astype(np.float64) * m_kg[sel] ** 2)) / P.MASS0 ** 2
        assert m1[gb] == pytest.approx(e1, rel=1e-5, abs=1e-30)
        assert m2[gb] == pytest.approx(e2, rel=1e-5, abs=1e-30)
    d.close()


def test_oracle_moments_sum_in_drop_order(oracle):
    # gridbox 0: one small drop; 1: one big drop (xi = 2^24 times the small's); 2: big, small, small;
    # 3: small, small, big
    r, ms = 7.3, 0.013
    xs, xb = 1, 2 ** 24
    g = [0, 1, 2, 2, 2, 3, 3, 3]
    xi = [xs, xb, xb, xs, xs, xs, xs, xb]
    pop = _box((4, 1, 1), dict(sdgbxindex=g, xi=xi, radius=[r] * 8, msol=[ms] * 8))
    d = make_domain(oracle, pop)
    m0, m1, m2 = d.mass_moments()
    for mom in (m1, m2):
        ts, tb = np.float32(mom[0]), np.float32(mom[1])
        big_first = (tb + ts) + ts
        small_first = (ts + ts) + tb
        assert big_first != small_first  # the case separates the two orders
        assert mom[2] == big_first and mom[3] == small_first
    assert list(m0) == [xs, xb, xb + 2 * xs, xb + 2 * xs]
    d.close()


def test_oracle_rain_filter(oracle):
    # the threshold is 40e-6 / R0 in double, which rounds to 40.00000000000001: a radius of
    # exactly 40 falls below it, the next double above the threshold does not
    rlim = 40e-6 / P.R0
    g = [0, 0, 0, 0, 1]
    radius = [39.999, 40.0, rlim, 40.001, 12.0]
    xi = [3, 5, 7, 9, 11]
    pop = _box((2, 1, 1), dict(sdgbxindex=g, xi=xi, radius=radius, msol=[0.01] * 5))
    d = make_domain(oracle, pop)
    a0, a1, _ = d.mass_moments()
    r0, r1, _ = d.mass_moments(rain=True)
    assert list(a0) == [24, 11] and list(r0) == [16, 0]
    assert r1[1] == 0.0 and 0.0 < r1[0] < a1[0]
    d.close()


def test_oracle_reports_xi_overflow(oracle):
    pop = _box((1, 1, 1), dict(sdgbxindex=[0, 0], xi=[2 ** 64 - 1, 1], radius=[5.0, 5.0], msol=[0.01] * 2))
    d = make_domain(oracle, pop)
    with pytest.raises(RuntimeError):
        d.mass_moments()
    assert d.error()[0] == abi.ERR_XI_OVERFLOW
    d.close()


@pytest.mark.gpu
@pytest.mark.parametrize("rain", [False, True])
def test_gpu_mass_moments_match_oracle(gpu, oracle, rain):
    print(f"case: mass moments after coupled steps, rain={rain}", flush=True)
    pop = P.coupled(SEED, ndims=(8, 6, 4), mean=64)
    sdm = sdm_params(cond=cond_params(), coll=coll_params(), motion=motion_params())
    doms = []
    for be in (gpu, oracle):
        d = make_domain(be, pop)
        d.set_sdm(sdm)
        d.rng_counter(77)
        for step in range(3):
            assert d.run_step(step * 100, (step + 1) * 100) == 0
        doms.append(d)
    g_mom = doms[0].mass_moments(rain)
    o_mom = doms[1].mass_moments(rain)
    assert np.array_equal(g_mom[0], o_mom[0])
    for a, b in zip(g_mom[1:], o_mom[1:]):
        np.testing.assert_allclose(a, b, rtol=1e-6)
    # the reduction alone, bit for bit: the oracle's drops (in its order) uploaded to the GPU
    o = doms[1].download()
    sup = {k: o[k][: o["ndomain"]] for k, _ in abi.SUPER_FIELDS}
    twin = P.Population(pop.ndims, pop.halves, sup, doms[1].state(), pop.solute, periodic=pop.periodic)
    t = make_domain(gpu, twin)
    t_mom = t.mass_moments(rain)
    for a, b in zip(t_mom, o_mom):
        assert np.array_equal(a, b)
    for d in doms + [t]:
        d.close()
