"""Pins the CPU oracle (oracle/cleo_oracle.cpp) against what the mathematics fixes: published
known-answer vectors, exact integer closed forms, conservation laws, brute force on tiny inputs,
independent root finding and the Golovin analytic solution.  The reference ships no golden
vectors for this path (its tests/ hold two trivial pytest files), so these are the pins."""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

from cleo_amd import abi
from cleo_amd import populations as P
from tests.helpers import SEED, coll_params, cond_params, injected_stream, make_domain, motion_params

# ---------------------------------------------------------------------------------------------
# random numbers


@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(oracle, ctr, key, expect):
    """Random123 kat_vectors for philox4x32_10."""
    out = (C.c_uint32 * 4)()
    oracle.lib.orc_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    assert tuple(out) == expect


# ---------------------------------------------------------------------------------------------
# units


def test_units_and_timesteps(oracle):
    c = (C.c_double * 10)()
    oracle.lib.orc_constants(c)
    rho0 = 1e5 / (1004.64 * 273.15)
    assert c[0] == pytest.approx(rho0, rel=1e-15)
    assert c[1] == pytest.approx(1e-18 * rho0, rel=1e-15)
    assert c[4] == pytest.approx(998.203 / rho0, rel=1e-15)
    assert c[8] == 1.0 and c[9] == 0.001  # 100 model steps = 1 s = 0.001 TIME0


# ---------------------------------------------------------------------------------------------
# enactment: exact integer closed forms and conservation


def _enact(oracle, kind, xi, r, m, prob, phi, phi_out=0.5, nfrags=5.0, coaleff=0.85):
    p = coll_params(enact=kind, nfrags=nfrags, coaleff=coaleff)
    xi = (C.c_uint64 * 2)(*xi)
    r = (C.c_double * 2)(*r)
    m = (C.c_double * 2)(*m)
    err = C.c_int32()
    oracle.lib.orc_enact_pair.restype = C.c_int32
    res = oracle.lib.orc_enact_pair(C.byref(p), xi, r, m, C.c_double(prob), C.c_double(phi),
                                    C.c_double(phi_out), C.byref(err))
    return list(xi), list(r), list(m), res, err.value


def _moment3(xi, r):
    return sum(float(x) * float(rr) ** 3 for x, rr in zip(xi, r))


def test_coalescence_different(oracle):
    xi, r, m, res, err = _enact(oracle, abi.ENACT_COAL, (10, 3), (2.0, 5.0), (0.1, 0.3), 2.3, 0.2)
    # gamma = floor(2.3) + [0.2 < 0.3] = 3 = xi1 // xi2
    assert (xi, res & 1, err) == ([1, 3], 0, 0)
    assert r[1] == pytest.approx((125 + 3 * 8) ** (1 / 3), rel=1e-15)
    assert m[1] == pytest.approx(0.3 + 3 * 0.1, rel=1e-15)
    assert _moment3(xi, r) == pytest.approx(_moment3((10, 3), (2.0, 5.0)), rel=1e-14)


def test_coalescence_gamma_cap_and_no_event(oracle):
    xi, r, m, res, err = _enact(oracle, abi.ENACT_COAL, (10, 4), (2.0, 5.0), (0.1, 0.3), 5.7, 0.9)
    assert xi == [2, 4]  # gamma = min(5, 10 // 4 = 2)
    xi, r, m, res, err = _enact(oracle, abi.ENACT_COAL, (10, 4), (2.0, 5.0), (0.1, 0.3), 0.3, 0.5)
    assert xi == [10, 4] and r == [2.0, 5.0] and res == 0


def test_coalescence_gamma_beyond_uint64(oracle):
    """A probability of 2^64 or more has no defined gamma in the reference (its double -> uint64
    conversion overflows); it is taken as the cap floor(xi1/xi2), which any such gamma reaches."""
    for prob in (2.0 ** 64, 1e30, float("inf")):
        xi, r, m, res, err = _enact(oracle, abi.ENACT_COAL, (9, 4), (2.0, 5.0), (0.1, 0.3), prob, 0.5)
        assert xi == [1, 4] and err == 0  # gamma = 9 // 4 = 2: different case, 9 - 2*4
    xi, r, m, res, err = _enact(oracle, abi.ENACT_COAL, (3, 1), (2.0, 5.0), (0.1, 0.3), 1e30, 0.5)
    assert xi == [0, 1] and (res & 1) == 1  # gamma 3: twin with xi2 = 1, drop 1 becomes null
    # just below 2^64 the reference's conversion applies
    xi, r, m, res, err = _enact(oracle, abi.ENACT_COAL, (9, 4), (2.0, 5.0), (0.1, 0.3), 2.0 ** 63, 0.5)
    assert xi == [1, 4]


def test_coalescence_twin_and_null(oracle):
    xi, r, m, res, err = _enact(oracle, abi.ENACT_COAL, (8, 4), (2.0, 5.0), (0.1, 0.3), 2.5, 0.1)
    # gamma = min(3, 2) = 2 and 8 - 2*4 == 0: twin, xi2/2 each, shared radius and solute
    assert xi == [2, 2] and r[0] == r[1] and m[0] == m[1]
    assert r[0] == pytest.approx((125 + 2 * 8) ** (1 / 3), rel=1e-15)
    assert _moment3(xi, r) == pytest.approx(_moment3((8, 4), (2.0, 5.0)), rel=1e-14)
    xi, r, m, res, err = _enact(oracle, abi.ENACT_COAL, (1, 1), (2.0, 5.0), (0.1, 0.3), 1.0, 0.5)
    assert xi == [0, 1] and (res & 1) == 1 and err == 0


def test_breakup_twin_and_different(oracle):
    xi, r, m, res, err = _enact(oracle, abi.ENACT_BREAKUP, (10, 10), (2.0, 5.0), (0.1, 0.3), 0.7, 0.2)
    assert xi == [25, 25] and err == 0  # round(5*10/2), round(50-25)
    assert _moment3(xi, r) == pytest.approx(_moment3((10, 10), (2.0, 5.0)), rel=1e-14)
    assert sum(x * mm for x, mm in zip(xi, m)) == pytest.approx(10 * 0.1 + 10 * 0.3, rel=1e-14)
    xi, r, m, res, err = _enact(oracle, abi.ENACT_BREAKUP, (10, 3), (2.0, 5.0), (0.1, 0.3), 0.7, 0.2)
    assert xi == [7, 15]  # xi1 - xi2, round(5 * 3)
    assert _moment3(xi, r) == pytest.approx(_moment3((10, 3), (2.0, 5.0)), rel=1e-14)
    xi, r, m, res, err = _enact(oracle, abi.ENACT_BREAKUP, (10, 3), (2.0, 5.0), (0.1, 0.3), 0.7, 0.2,
                                nfrags=2.0)
    assert err == 7  # CLEO_ERR_BREAKUP_NFRAGS


def test_coalbure_flag(oracle):
    a = _enact(oracle, abi.ENACT_COALBURE, (10, 3), (2.0, 5.0), (0.1, 0.3), 2.3, 0.2, phi_out=0.5)
    b = _enact(oracle, abi.ENACT_COAL, (10, 3), (2.0, 5.0), (0.1, 0.3), 2.3, 0.2)
    assert a[:3] == b[:3]  # phi_out < coaleff: coalescence with the same gamma
    c = _enact(oracle, abi.ENACT_COALBURE, (10, 3), (2.0, 5.0), (0.1, 0.3), 2.3, 0.2, phi_out=0.9)
    d = _enact(oracle, abi.ENACT_BREAKUP, (10, 3), (2.0, 5.0), (0.1, 0.3), 0.7, 0.2)
    assert c[:3] == d[:3]  # breakup ignores gamma
    e = _enact(oracle, abi.ENACT_COALBURE, (10, 3), (2.0, 5.0), (0.1, 0.3), 0.1, 0.5, phi_out=0.9)
    assert e[0] == [10, 3] and e[3] == 0


# ---------------------------------------------------------------------------------------------
# probabilities and terminal velocity against SI evaluations


def test_golovin_probability(oracle):
    r1, r2, DELT, VOL = 10.0, 20.0, 1.0, 1e6
    got = oracle.lib.orc_pair_prob(abi.PROB_GOLOVIN, 0.0, r1, 5.0, r2, 3.0, DELT, VOL)
    b = 1.5e3  # s^-1
    v = 4.0 / 3.0 * math.pi * ((r1 * 1e-6) ** 3 + (r2 * 1e-6) ** 3)
    assert got == pytest.approx(b * v * DELT / VOL, rel=1e-13)


def _simmel_si(r_m):
    m_g = 4.0 / 3.0 * math.pi * 998.203 * r_m ** 3 * 1e3
    if r_m >= 1.73892e-3:
        return 9.17
    if r_m >= 7.5582e-4:
        return 17.32 * m_g ** (1 / 6)
    if r_m >= 6.7215e-5:
        return 49.62 * m_g ** (1 / 3)
    return 4579.5 * m_g ** (2 / 3)


@pytest.mark.parametrize("r_um", [0.5, 10.0, 60.0, 67.3, 200.0, 800.0, 1800.0])
def test_simmel_terminal_velocity(oracle, r_um):
    got = oracle.lib.orc_terminalv(abi.TV_SIMMEL, r_um)
    assert got == pytest.approx(_simmel_si(r_um * 1e-6), rel=1e-12)


def test_long_kernel_efficiency(oracle):
    # rbig >= 50 um: efficiency = coaleff
    r1, r2 = 60.0, 20.0
    got = oracle.lib.orc_pair_prob(abi.PROB_LONG, 0.7, r1, 1.0, r2, 1.0, 1.0, 1e6)
    vd = abs(_simmel_si(r1 * 1e-6) - _simmel_si(r2 * 1e-6))
    expect = 0.7 * math.pi * ((r1 + r2) * 1e-6) ** 2 * vd / 1e6
    assert got == pytest.approx(expect, rel=1e-12)
    # below: 4.5e4 cm^-2 R^2 (1 - 3e-4 cm / r), floored at 0.001
    r1, r2 = 30.0, 10.0
    eff = max(4.5e4 * (r1 * 1e-4) ** 2 * (1 - 3e-4 / (r2 * 1e-4)), 0.001)
    got = oracle.lib.orc_pair_prob(abi.PROB_LONG, 1.0, r1, 1.0, r2, 1.0, 1.0, 1e6)
    vd = abs(_simmel_si(r1 * 1e-6) - _simmel_si(r2 * 1e-6))
    assert got == pytest.approx(eff * math.pi * ((r1 + r2) * 1e-6) ** 2 * vd / 1e6, rel=1e-12)
    got = oracle.lib.orc_pair_prob(abi.PROB_LONG, 1.0, 30.0, 1.0, 1.0, 1.0, 1.0, 1e6)
    vd = abs(_simmel_si(30e-6) - _simmel_si(1e-6))
    assert got == pytest.approx(0.001 * math.pi * (31e-6) ** 2 * vd / 1e6, rel=1e-12)


# ---------------------------------------------------------------------------------------------
# Newton / implicit Euler


def _solve(oracle, s, a, b, F, r, delt=1e-3, minsub=1e-6, rtol=1e-6, atol=1e-8, maxit=50, warm=float("nan"),
           warm_limit=10):
    n = C.c_uint32()
    sub = C.c_int32()
    err = C.c_int32()
    newr = oracle.lib.orc_solve_condensation(s, a, b, F, r, delt, minsub, rtol, atol, maxit, warm, warm_limit,
                                             C.byref(n), C.byref(sub), C.byref(err))
    return newr, n.value, sub.value, err.value


def _F(oracle, r):
    temp, press = 290.0 / 273.15, 0.9
    psat = oracle.lib.orc_saturation_pressure(temp)
    return oracle.lib.orc_diffusion_factor(press, temp, psat)


def test_newton_linear_case_closed_form(oracle):
    """a = b = 0: g is linear in 1/z and implicit Euler gives z = r^2 + 2 (s-1) dt / F exactly
    (the NaN uniqueness tests send it through sub-stepping, which sums to the same)."""
    r, s, F = 12.0, 1.004, _F(oracle, 12.0)
    Ff = F / oracle.lib.orc_ventilation_factor(r)
    newr, n, sub, err = _solve(oracle, s, 0.0, 0.0, F, r, minsub=1e-4)
    assert err == 0 and sub == 1
    assert newr == pytest.approx(math.sqrt(r * r + 2 * (s - 1) * 1e-3 / Ff), rel=1e-12)


def _bisect_root(s, a, b, Ff, r, dt):
    g = lambda z: 1 - r * r / z - (s - 1 - a / math.sqrt(z) + b / z ** 1.5) * 2 * dt / (z * Ff)  # noqa: E731
    lo, hi = 1e-8, 1e8
    glo = g(lo)
    for _ in range(400):
        mid = math.sqrt(lo * hi)
        if (g(mid) < 0) == (glo < 0):
            lo, glo = mid, g(mid)
        else:
            hi = mid
    return math.sqrt(math.sqrt(lo * hi))


@pytest.mark.parametrize("s,r,msol", [(1.003, 8.0, 1e-3), (0.97, 0.3, 1e-5), (1.0005, 2.0, 2e-3),
                                       (0.999, 0.05, 1e-6)])
def test_newton_matches_bisection(oracle, s, r, msol):
    """Converged Newton equals the bracketed root of g; in the sub-stepping branch every
    sub-step is one such root from the previous sub-step's radius."""
    temp = 285.0 / 273.15
    a = 3.3e-7 / (273.15 * 1e-6) / temp
    b = 4.3e-6 * P.RHO0 / P.MR_DRY * msol * 2 / (0.05844277 / P.MR_DRY)
    F = _F(oracle, r)
    Ff = F / oracle.lib.orc_ventilation_factor(r)
    minsub = 1e-6
    newr, n, sub, err = _solve(oracle, s, a, b, F, r, rtol=0.0, atol=1e-12, minsub=minsub, maxit=200)
    assert err == 0
    if not sub:
        assert newr == pytest.approx(_bisect_root(s, a, b, Ff, r, 1e-3), rel=1e-9)
    else:
        tcrit = 2.5 * Ff / a * (5 * b / a) ** 1.5
        step, rem, rp = max(tcrit, minsub), 1e-3, r
        while rem > 0:
            dt = min(step, rem)
            rp = _bisect_root(s, a, b, Ff, rp, dt)
            rem -= dt
        assert newr == pytest.approx(rp, rel=1e-9)


def test_newton_iteration_semantics(oracle):
    temp = 285.0 / 273.15
    a = 3.3e-7 / (273.15 * 1e-6) / temp
    r = 0.2
    b = a * r * r  # rcrit^2 = 3b/a = 3 r^2 > r^2: unactivated, unique-root branch
    s = 1 + a / r - b / r ** 3  # equilibrium: the guess r^2 is the root
    F = _F(oracle, r)
    newr, n, sub, err = _solve(oracle, s, a, b, F, r)
    assert (n, sub, err) == (2, 0, 0) and newr == pytest.approx(r, rel=1e-12)
    # not converged after 2: restart from the guess, count everything
    newr, n, sub, err = _solve(oracle, 1.01, a, b, F, 5.0, rtol=0.0, atol=1e-14)
    assert n > 2 and err == 0
    # cannot converge within maxniters: the reference aborts
    newr, n, sub, err = _solve(oracle, 1.01, a, b, F, 5.0, rtol=0.0, atol=0.0, maxit=3)
    assert err == 4  # CLEO_ERR_NEWTON_MAXITERS


def test_newton_warm_start_from_root_takes_at_most_two(oracle):
    temp = 285.0 / 273.15
    a = 3.3e-7 / (273.15 * 1e-6) / temp
    r = 0.2
    b = a * r * r
    s = 1 + 0.5 * math.sqrt(4 * a ** 3 / 27 / b)  # below activation: unique-root branch
    F = _F(oracle, r)
    ref, nref, sref, _ = _solve(oracle, s, a, b, F, r, rtol=0.0, atol=1e-12, maxit=200)
    warm, nwarm, _, _ = _solve(oracle, s, a, b, F, r, rtol=0.0, atol=1e-12, warm=ref * ref, maxit=200)
    assert sref == 0 and nref > 2
    assert 1 <= nwarm <= 2 and warm == pytest.approx(ref, rel=1e-12)
    # a guess off by 1e-4 converges without the reference's restart: at most 3 iterations
    near, nnear, _, _ = _solve(oracle, s, a, b, F, r, warm=ref * ref * (1 + 1e-4))
    plain, nplain, _, _ = _solve(oracle, s, a, b, F, r)
    assert nnear <= 3 and near == pytest.approx(plain, rel=1e-9)
    # a guess that cannot converge within warm_limit falls back to the plain path, counted as a
    # restart, and gives the plain result
    bad, nbad, _, _ = _solve(oracle, s, a, b, F, r, warm=1e6 * ref * ref, warm_limit=1)
    assert bad == pytest.approx(plain, rel=1e-12) and nbad == 1 + nplain


def test_learned_path_starts_from_plain_guess_where_it_is_the_root(oracle):
    """The learned path starts from the learned guess except where |g| at the plain guess is below
    atol.  A drop at its equilibrium radius, whose plain guess r^2 is the root, is then iterated
    from r^2 under a learned guess 1e-3 off: one update passes the test (the reference path makes
    two) and the radius is the reference's to 1e-14.  The same guess on a drop far from its root is
    taken and needs fewer iterations than the plain one."""
    temp = 285.0 / 273.15
    a = 3.3e-7 / (273.15 * 1e-6) / temp
    b = a * 0.2 * 0.2
    s = 1 + 0.5 * math.sqrt(4 * a ** 3 / 27 / b)  # below activation
    lo, hi = (b / a) ** 0.5 * 1.0001, math.sqrt(3 * b / a)  # S - 1 - a/r + b/r^3 changes sign inside
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if s - 1 - a / mid + b / mid ** 3 > 0 else (lo, mid)
    req = 0.5 * (lo + hi)
    F = _F(oracle, req)
    plain, nplain, sub, _ = _solve(oracle, s, a, b, F, req)
    w, nw, _, _ = _solve(oracle, s, a, b, F, req, warm=req * req * (1 + 1e-3))
    assert sub == 0 and nplain == 2 and nw == 1 and w == pytest.approx(plain, rel=1e-14)
    # the learned guess alone (no plain root to start from) needs more than one update there
    g_only, ng_only, _, _ = _solve(oracle, s, a, b, F, req, warm=req * req * (1 + 1e-3), atol=1e-30, rtol=1e-12)
    assert ng_only >= 2 and g_only == pytest.approx(plain, rel=1e-12)
    # from 3 r_eq the plain guess (r^2) is far from the root; a guess 1e-3 off it is taken
    far, nfar, _, _ = _solve(oracle, s, a, b, F, 3 * req)
    wf, nwf, _, _ = _solve(oracle, s, a, b, F, 3 * req, warm=far * far * (1 + 1e-3))
    assert nfar > 3 and nwf < nfar and wf == pytest.approx(far, rel=1e-9)


# ---------------------------------------------------------------------------------------------
# domain-level: shuffle, sort, condensation feedback, motion


def test_fisher_yates_brute_force(oracle):
    """Every index sequence j_{n-1..1} of one gridbox gives the textbook permutation and all n!
    permutations appear exactly once (n = 5)."""
    n = 5
    pop = P.coal_breakup(SEED, ndims=(1, 1, 1), mean=n)
    sup = {k: v[:0] for k, v in pop.supers.items()}
    ids = np.arange(n, dtype=np.uint64)
    base = dict(sdgbxindex=np.zeros(n, np.uint32), coord3=np.full(n, 0.01), coord1=np.full(n, 0.01),
                coord2=np.full(n, 0.01), xi=np.full(n, 10, np.uint64), radius=np.full(n, 1.0),
                msol=np.zeros(n), sdid=ids)
    pop.supers = base
    seen = set()
    for js in itertools.product(*[range(i + 1) for i in range(n - 1, 0, -1)]):
        d = make_domain(oracle, pop)
        d.rng_injected(np.array(js, np.uint32), np.full(2 * (n // 2), 0.999))
        assert d.collisions(coll_params(prob=abi.PROB_CONST, param=0.0, enact=abi.ENACT_COAL)) == 0
        got = list(d.download()["sdid"])
        want = list(range(n))
        for t, it in enumerate(range(n - 1, 0, -1)):
            want[it], want[js[t]] = want[js[t]], want[it]
        assert got == want
        seen.add(tuple(got))
        d.close()
    assert len(seen) == math.factorial(n)
    del sup


def test_sort_is_stable_sort_by_bucket(oracle):
    rng = np.random.default_rng(3)
    pop = P.coal_breakup(SEED, ndims=(3, 2, 2), mean=15)
    keys = rng.integers(0, pop.ngbx + 3, pop.n).astype(np.uint32)
    keys[rng.random(pop.n) < 0.1] = abi.OOB
    pop.supers["sdgbxindex"] = keys
    d = make_domain(oracle, pop)
    out = d.download()
    bucket = np.where(keys < pop.ngbx, keys, pop.ngbx)
    order = np.argsort(bucket, kind="stable")
    assert np.array_equal(out["sdid"], pop.supers["sdid"][order])
    assert out["ndomain"] == int(np.sum(keys < pop.ngbx))
    refs = d.refs()
    assert np.array_equal(refs, np.searchsorted(bucket[order], np.arange(pop.ngbx + 1)))


def test_condensation_conserves_water_and_order_invariance(oracle):
    pop = P.condensation(SEED, ndims=(1, 1, 1), mean=200)
    pop.state["qvap"] *= 1.02  # supersaturate a little so mass moves
    d = make_domain(oracle, pop)
    s0 = d.state()
    assert d.condensation(cond_params()) == 0
    s1 = d.state()
    assert s1["qvap"][0] + s1["qcond"][0] == pytest.approx(s0["qvap"][0] + s0["qcond"][0], rel=1e-15)
    assert s1["qcond"][0] > 0 and s1["temp"][0] > s0["temp"][0]  # condensation warms
    # reversing the drop order changes the sum only by rounding
    pop2 = P.condensation(SEED, ndims=(1, 1, 1), mean=200)
    pop2.state["qvap"] *= 1.02
    pop2.supers = {k: v[::-1].copy() for k, v in pop2.supers.items()}
    d2 = make_domain(oracle, pop2)
    assert d2.condensation(cond_params()) == 0
    assert d2.state()["qcond"][0] == pytest.approx(s1["qcond"][0], rel=1e-12)


def _one_drop_domain(oracle, coords, w=0.0, u=0.0, v=0.0, ndims=(4, 4, 4)):
    halves = P.uniform_grid(ndims, 100.0)
    pop = P.coal_breakup(SEED, ndims=ndims, mean=1)
    nz, nx, ny = ndims
    c3, c1, c2 = coords
    iz, ix, iy = int(c3 / 0.1), int(c1 / 0.1), int(c2 / 0.1)
    pop.supers = dict(sdgbxindex=np.array([iz + nz * (ix + nx * iy)], np.uint32), coord3=np.array([c3]),
                      coord1=np.array([c1]), coord2=np.array([c2]), xi=np.array([5], np.uint64),
                      radius=np.array([10.0]), msol=np.array([0.0]), sdid=np.array([0], np.uint64))
    pop.halves = halves
    for k, val in (("wvel", w), ("uvel", u), ("vvel", v)):
        pop.state[k + "_lo"][:] = val
        pop.state[k + "_hi"][:] = val
    return make_domain(oracle, pop)


def test_motion_uniform_wind_and_periodic_wrap(oracle):
    d = _one_drop_domain(oracle, (0.15, 0.399995, 0.25), w=0.02, u=0.01, v=-0.03)
    assert d.motion(motion_params(tv=abi.TV_NULL)) == 0
    o = d.download()
    delt = 1e-3
    assert o["coord3"][0] == 0.15 + 0.02 * delt and o["coord2"][0] == 0.25 + -0.03 * delt
    # crossed the upper x face (0.4 = domain end): wrapped to x - 0.4 and the gridbox ix = 0
    assert o["coord1"][0] == pytest.approx(0.399995 + 0.01 * delt - 0.4, abs=1e-15)
    assert o["sdgbxindex"][0] == 1 + 4 * (0 + 4 * 2)


def test_motion_leaves_through_finite_bottom(oracle):
    d = _one_drop_domain(oracle, (0.00001, 0.15, 0.15), w=-0.05)
    assert d.motion(motion_params(tv=abi.TV_NULL)) == 0
    o = d.download()
    assert o["ndomain"] == 0 and o["sdgbxindex"][0] == abi.OOB


def test_motion_linear_wind_heun(oracle):
    """Linear w(z) inside one gridbox: the predictor-corrector is Heun's method."""
    d = _one_drop_domain(oracle, (0.13, 0.15, 0.15))
    st = d.state()
    g = 1 + 4 * (1 + 4 * 1)
    st["wvel_lo"][g], st["wvel_hi"][g] = 0.01, 0.04
    d.set_state(st)
    assert d.motion(motion_params(tv=abi.TV_NULL)) == 0
    z0, dt = 0.13, 1e-3
    w = lambda z: 0.01 + (min(max(z, 0.1), 0.2) - 0.1) / 0.1 * 0.03  # noqa: E731
    v0 = w(z0)
    z1 = z0 + 0.5 * (v0 + w(z0 + v0 * dt)) * dt
    assert d.download()["coord3"][0] == pytest.approx(z1, rel=1e-14)


def test_motion_cfl_error(oracle):
    d = _one_drop_domain(oracle, (0.15, 0.15, 0.15), w=200.0)
    assert d.motion(motion_params(tv=abi.TV_NULL)) != 0
    assert d.error()[0] == 8  # CLEO_ERR_CFL


# ---------------------------------------------------------------------------------------------
# Golovin box model against the analytic number decay


def test_golovin_number_decay(oracle):
    """Golovin kernel K = b (v1 + v2) gives N(t) = N0 exp(-b L t) exactly in the mean field
    (L = liquid volume fraction, conserved); 8192 superdroplets follow it within a few per cent."""
    from oracle import pyoracle
    b_omp = pyoracle.load("omp")
    pop = P.golovin_box(SEED)
    d = make_domain(b_omp, pop)
    d.rng_counter(12345)
    p = coll_params(prob=abi.PROB_GOLOVIN, enact=abi.ENACT_COAL)
    xi0 = pop.supers["xi"].astype(np.float64)
    vol = 1e6
    L = np.sum(xi0 * 4 / 3 * np.pi * (pop.supers["radius"] * 1e-6) ** 3) / vol
    n0 = xi0.sum()
    t = 0
    for t_check in (600, 1200):
        while t < t_check:
            assert d.collisions(p) == 0
            t += 1
        o = d.download()
        nd = o["ndomain"]
        n = o["xi"][:nd].astype(np.float64).sum()
        assert n / n0 == pytest.approx(math.exp(-1.5e3 * L * t), rel=0.05)
        # liquid volume conserved to rounding
        lv = np.sum(o["xi"][:nd].astype(np.float64) * 4 / 3 * np.pi * (o["radius"][:nd] * 1e-6) ** 3) / vol
        assert lv == pytest.approx(L, rel=1e-9)


def test_injected_stream_layout(oracle):
    """n-1 indices and 2*floor(n/2) phis per gridbox; a short stream is refused."""
    pop = P.coal_breakup(SEED, ndims=(2, 2, 1), mean=9)
    d = make_domain(oracle, pop)
    js, ph = injected_stream(d.refs(), np.random.default_rng(0))
    d.rng_injected(js[:-1], ph)
    assert d.collisions(coll_params()) != 0
