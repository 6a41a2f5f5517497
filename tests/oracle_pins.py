"""Independent pins of the oracle's thermodynamic equations, terminal velocities, State update and
qcond recompute.  Each pin evaluates the published form of an equation in SI units, with the
coefficients as the source publishes them (unit conversions done here, not taken from the
oracle), and compares the oracle's dimensionless result with it.  Sources:

  Lohmann, Lueoend and Mahrt, "An Introduction to Clouds" (2016): eqs. 6.22/6.24 (Koehler a, b),
    7.23-7.26 (heat and vapour diffusion factors, thermal conductivity, vapour diffusivity);
  Murray (1967) / Tetens form of the saturation pressure with the constants of
    thermodynamic_equations.cpp:44-47 (fixed at the triple point 273.16 K, 611.655 Pa), checked
    against Murphy and Koop (2005, eq. 10) as an independent fit;
  Rogers and Yau, "A Short Course in Cloud Physics" (1989), eqs. 8.5, 8.6, 8.8 in CGS units;
  Rogers et al. (1993) fit to Gunn and Kinzer (1949), diameter in mm;
  the ventilation fit of thermodynamic_equations.cpp:116-145 (Kinzer and Gunn 1951, Pruppacher and
    Rasmussen 1979 data) in its harmonic form and its limits;
  Dalton's law and the mixing-ratio definition for the dry-air density and the saturation ratio;
  the first law at constant pressure for the latent heating of the State update.

Every pin is a function of the loaded oracle (a pyoracle Backend) and raises AssertionError on a
mismatch, so tests/test_oracle_thermo.py runs them and tests/test_oracle_mutations.py runs them
against oracles built with perturbed constants."""
import ctypes as C
import math

import numpy as np

from cleo_amd import abi

# scales of the dimensionless units (libs/cleoconstants.hpp:65-104), pinned by test_units_and_timesteps
R0 = 1e-6
TEMP0 = 273.15
P0 = 1e5
TIME0 = 1000.0
COORD0 = 1000.0
VOL0 = COORD0 ** 3
CP_DRY = 1004.64
RHO0 = P0 / (CP_DRY * TEMP0)
MASS0 = R0 ** 3 * RHO0

# physical constants as libs/cleoconstants.hpp:32-56 states them (SI)
RGAS_UNIV = 8.314462618
MR_WATER = 0.01801528
MR_DRY = 0.028966216
LATENT_V = 2500930.0
CP_V = 1865.01
C_L = 4192.664
RHO_L = 998.203
R_D = RGAS_UNIV / MR_DRY
R_V = RGAS_UNIV / MR_WATER
EPS = MR_WATER / MR_DRY

TOL = 1e-12  # relative; two evaluations of one expression in different arrangements


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def _close(got, want, what, tol=TOL):
    assert _rel(got, want) <= tol, f"{what}: oracle {got!r} vs independent {want!r} (rel {_rel(got, want):.3e})"


def _bind(orc):
    lib = orc.lib
    lib.orc_kohler_factors.argtypes = [C.c_double] * 5 + [C.POINTER(C.c_double)]
    lib.orc_kohler_factors.restype = None
    return lib


# ------------------------------------------------------------------------------------------
# saturation pressure and saturation ratio


def psat_tetens_si(T):
    """e_s [Pa] = 611.655 exp(17.4146 (T - 273.16) / (T - 33.639)), T in K."""
    return 611.655 * math.exp(17.4146 * (T - 273.16) / (T - 33.639))


def psat_murphy_koop_si(T):
    """Murphy and Koop (2005) eq. 10, e_s over liquid water [Pa], 123 K < T < 332 K."""
    lnp = (54.842763 - 6763.22 / T - 4.21 * math.log(T) + 0.000367 * T
           + math.tanh(0.0415 * (T - 218.8)) * (53.878 - 1331.22 / T - 9.44523 * math.log(T) + 0.014025 * T))
    return math.exp(lnp)


def pin_saturation_pressure(orc):
    lib = orc.lib
    # the triple point of water is the formula's reference point: exactly 611.655 Pa
    _close(lib.orc_saturation_pressure(273.16 / TEMP0) * P0, 611.655, "psat(273.16 K)", 1e-14)
    for T in (250.0, 265.3, 280.0, 293.15, 300.0, 310.0):
        got = lib.orc_saturation_pressure(T / TEMP0) * P0
        _close(got, psat_tetens_si(T), f"psat({T} K) vs Tetens form")
        # an independent fit over liquid water agrees to 0.2 % above 260 K (0.5 % at 250 K)
        assert _rel(got, psat_murphy_koop_si(T)) < (5e-3 if T < 260 else 2e-3), T
    # Clausius-Clapeyron: d ln e_s / dT = L / (R_v T^2) within 3 % (L varies with T; the fit knows it)
    for T in (270.0, 290.0):
        h = 1e-3
        dl = (math.log(lib.orc_saturation_pressure((T + h) / TEMP0))
              - math.log(lib.orc_saturation_pressure((T - h) / TEMP0))) / (2 * h)
        assert _rel(dl, LATENT_V / (R_V * T * T)) < 0.03, (T, dl)


def pin_supersaturation_ratio(orc):
    """S = e / e_s, with the vapour mass mixing ratio qv = eps e / (p - e) of vapour pressure e."""
    lib = orc.lib
    for p, e, es in ((1e5, 1500.0, 1700.0), (7.2e4, 2500.0, 2400.0), (9.1e4, 12.5, 300.0)):
        qv = EPS * e / (p - e)
        _close(lib.orc_supersaturation_ratio(p / P0, qv, es / P0), e / es, f"S(p={p}, e={e})", 4e-15)


# ------------------------------------------------------------------------------------------
# diffusion factors


def ffactor_lohmann_si(T, p, es):
    """F_k + F_d [s m^-2] of the radial growth equation r dr/dt = (S - 1) / (F_k + F_d):
    K = 4.1868e-3 (5.69 + 0.017 (T - 273.15)) J m^-1 s^-1 K^-1 (eq. 7.24, from cal cm^-1 s^-1 C^-1),
    D_v = 2.11e-5 (T / 273.15)^1.94 (101325 / p) m^2 s^-1 (eq. 7.26),
    F_k = (L / (R_v T) - 1) L rho_l / (K T) (eq. 7.23), F_d = rho_l R_v T / (D_v e_s) (eq. 7.25)."""
    K = 4.1868e-3 * (5.69 + 0.017 * (T - 273.15))
    Dv = 2.11e-5 * (T / 273.15) ** 1.94 * (101325.0 / p)
    Fk = (LATENT_V / (R_V * T) - 1.0) * LATENT_V * RHO_L / (K * T)
    Fd = RHO_L * R_V * T / (Dv * es)
    return Fk + Fd


def pin_diffusion_factor(orc):
    """The dimensionless factor is F R0^2 / TIME0 (radius in R0, time in TIME0)."""
    lib = orc.lib
    for T, p in ((273.15, 1e5), (290.0, 9e4), (301.5, 7e4), (255.0, 6e4)):
        es = psat_tetens_si(T)
        got = lib.orc_diffusion_factor(p / P0, T / TEMP0, es / P0)
        _close(got, ffactor_lohmann_si(T, p, es) * R0 * R0 / TIME0, f"F({T} K, {p} Pa)", 1e-11)


# ------------------------------------------------------------------------------------------
# ventilation


C1, ALPHA, C2, BETA = 6.954e7, 1.963, 1.069e3, 0.702


def vent_si(R):
    """f_v = 1 + 1 / (1/(c1 R^alpha) + 1/(c2 R^beta)) in its harmonic form, R in m, capped at 20."""
    u, v = C1 * R ** ALPHA, C2 * R ** BETA
    return min(1.0 + u * v / (u + v), 20.0)


def pin_ventilation_factor(orc):
    lib = orc.lib
    for R in (1e-7, 2e-6, 1.5e-5, 1e-4, 4e-4, 1e-3, 2.5e-3):
        _close(lib.orc_ventilation_factor(R / R0), vent_si(R), f"f_v({R} m)")
    # small drops: f_v - 1 -> c1 R^alpha; large drops: f_v - 1 -> c2 R^beta
    R = 1e-7
    assert _rel(lib.orc_ventilation_factor(R / R0) - 1.0, C1 * R ** ALPHA) < 1e-3
    R = 3e-3
    assert _rel(lib.orc_ventilation_factor(R / R0) - 1.0, C2 * R ** BETA) < 0.03
    # the cap: the uncapped fit reaches 20 at R* ~ 3.307 mm; below it the fit, above it exactly 20
    lo, hi = 1e-3, 1e-2
    for _ in range(200):
        m = 0.5 * (lo + hi)
        u, v = C1 * m ** ALPHA, C2 * m ** BETA
        if 1.0 + u * v / (u + v) < 20.0:
            lo = m
        else:
            hi = m
    assert 3.30e-3 < lo < 3.31e-3
    assert lib.orc_ventilation_factor(lo * (1 - 1e-7) / R0) < 20.0
    assert lib.orc_ventilation_factor(lo * (1 + 1e-7) / R0) == 20.0


# ------------------------------------------------------------------------------------------
# Koehler factors


def pin_kohler_factors(orc):
    """a = 3.3e-7 m K / T (eq. 6.24: 2 sigma_w / (rho_l R_v T) with sigma_w(0 C) = 0.0757 N/m)
    and b = 4.3e-6 m^3 mol kg^-1 * i m_s / M_s (eq. 6.22: 3 M_w / (4 pi rho_l) = 4.3e-6), both in
    units of R0 and R0^3."""
    lib = _bind(orc)
    assert _rel(2 * 0.0757 / (RHO_L * R_V), 3.3e-7) < 0.01
    assert _rel(3 * MR_WATER / (4 * math.pi * RHO_L), 4.3e-6) < 0.005
    ab = (C.c_double * 2)()
    for T, m_s, rho_s, M_s, i in ((273.15, 1e-18, 2016.5, 0.05844277, 2.0), (291.3, 3.7e-16, 1770.0, 0.13214, 3.0),
                                  (300.0, 2e-21, 1000.0, 0.1, 1.0)):
        lib.orc_kohler_factors(T / TEMP0, m_s / MASS0, rho_s / RHO0, M_s / MR_DRY, i, ab)
        _close(ab[0], 3.3e-7 / T / R0, f"a({T} K)")
        _close(ab[1], 4.3e-6 * i * m_s / M_s / R0 ** 3, f"b(m_s={m_s})")


# ------------------------------------------------------------------------------------------
# terminal velocities


def tv_rogers_yau_si(R):
    """Rogers and Yau (1989): u = k1 r^2, k1 = 1.19e6 cm^-1 s^-1 (eq. 8.5) below 30 um;
    u = k3 r, k3 = 8e3 s^-1 (eq. 8.8) below 0.6 mm; u = k2 r^(1/2), k2 = 2.01e3 cm^(1/2) s^-1
    (eq. 8.6) below 2 mm; 9 m/s above.  CGS evaluated, returned in m/s."""
    r_cm = R * 100.0
    if R < 3e-5:
        return 1.19e6 * r_cm ** 2 / 100.0
    if R < 6e-4:
        return 8e3 * r_cm / 100.0
    if R < 2e-3:
        return 2.01e3 * math.sqrt(r_cm) / 100.0
    return 9.0


def tv_rogers_gk_si(R):
    """Rogers et al. (1993): v = 4.0 D (1 - exp(-12 D)) for D < 0.745 mm, else
    v = 9.65 - 10.43 exp(-0.6 D), D the diameter in mm, v in m/s."""
    D = 2.0 * R * 1e3
    if D < 0.745:
        return 4.0 * D * (1.0 - math.exp(-12.0 * D))
    return 9.65 - 10.43 * math.exp(-0.6 * D)


def pin_terminal_velocities(orc):
    lib = orc.lib
    for kind, f, edges in ((abi.TV_ROGERS_YAU, tv_rogers_yau_si, (3e-5, 6e-4, 2e-3)),
                           (abi.TV_ROGERS_GK, tv_rogers_gk_si, (3.725e-4,))):
        radii = [1e-6, 1e-5, 2.5e-5, 1e-4, 3e-4, 5e-4, 9e-4, 1.5e-3, 3e-3, 5e-3]
        for e in edges:  # each branch edge from both sides
            radii += [e * (1 - 1e-7), e, e * (1 + 1e-7)]
        for R in radii:
            _close(lib.orc_terminalv(kind, R / R0), f(R), f"tv kind {kind} at R={R!r}")
    # the fits join (nearly) continuously where the publications join them
    for R, tol in ((6e-4, 0.03), (2e-3, 0.002)):
        assert _rel(tv_rogers_yau_si(R * (1 - 1e-9)), tv_rogers_yau_si(R)) < tol
    assert _rel(tv_rogers_gk_si(3.725e-4 * (1 - 1e-9)), tv_rogers_gk_si(3.725e-4)) < 0.005
    assert lib.orc_terminalv(abi.TV_NULL, 10.0) == 0.0


# ------------------------------------------------------------------------------------------
# dry-air density through the qcond recompute, and the State update of condensation


def rho_dry_si(p, T, qv):
    """Dalton's law: the dry air's partial pressure p - e, e = p qv / (eps + qv)."""
    e = p * qv / (EPS + qv)
    return (p - e) / (R_D * T)


def _domain(orc, ndims, spacing_m, sup, state, solute):
    halves = [np.arange(n + 1, dtype=np.float64) * (spacing_m / COORD0) for n in ndims]
    d = abi.Domain(orc, len(sup["xi"]), ndims, *halves)
    d.upload(sup)
    d.set_state(state)
    d.set_solute(solute[0] / RHO0, solute[1] / MR_DRY, solute[2])
    return d


def _drops(gbx, xi, R, m_s, spacing_m, ndims):
    n = len(xi)
    rng = np.random.default_rng(5)
    gbx = np.asarray(gbx, np.uint32)
    nz = ndims[0]
    iz = gbx % nz
    ix = (gbx // nz) % ndims[1]
    iy = gbx // (nz * ndims[1])
    h = spacing_m / COORD0
    return dict(sdgbxindex=gbx, coord3=(iz + rng.uniform(0.1, 0.9, n)) * h, coord1=(ix + rng.uniform(0.1, 0.9, n)) * h,
                coord2=(iy + rng.uniform(0.1, 0.9, n)) * h, xi=np.asarray(xi, np.uint64),
                radius=np.asarray(R, np.float64) / R0, msol=np.asarray(m_s, np.float64) / MASS0,
                sdid=np.arange(n, dtype=np.uint64))


def _state(ngbx, V, p, T, qv, qc):
    z = np.zeros(ngbx)
    return dict(volume=np.full(ngbx, V / VOL0), press=np.asarray(p, np.float64) / P0,
                temp=np.asarray(T, np.float64) / TEMP0, qvap=np.asarray(qv, np.float64),
                qcond=np.asarray(qc, np.float64), wvel_lo=z.copy(), wvel_hi=z.copy(), uvel_lo=z.copy(),
                uvel_hi=z.copy(), vvel_lo=z.copy(), vvel_hi=z.copy())


SOLUTE_NACL = (2016.5, 0.05844277, 2.0)


def condensate_si(xi, R, m_s, rho_s):
    """Liquid water of a drop: rho_l times (wet volume - solute volume), times xi."""
    return xi * RHO_L * (4.0 / 3.0 * math.pi * R ** 3 - m_s / rho_s)


def pin_qcond_recompute(orc):
    """After motion (no wind, no fall speed, so no drop moves) the State's qcond is the gridbox's
    liquid water over its dry-air mass: sum xi rho_l (V_wet - V_sol) / (rho_dry V)."""
    ndims, dx = (2, 1, 1), 20.0
    xi = [10 ** 9, 3 * 10 ** 8, 7 * 10 ** 9, 5, 2 * 10 ** 10]
    R = [1.2e-5, 3e-6, 8e-7, 4e-4, 2e-5]
    m_s = [1e-18, 4e-17, 2e-19, 1e-15, 0.0]
    gbx = [0, 0, 1, 1, 1]
    p, T, qv = [9.3e4, 7.5e4], [288.2, 270.4], [0.011, 0.0031]
    d = _domain(orc, ndims, dx, _drops(gbx, xi, R, m_s, dx, ndims), _state(2, dx ** 3, p, T, qv, [0.5, 0.5]),
                SOLUTE_NACL)
    assert d.motion(abi.MotionParams(100, abi.TV_NULL)) == 0 and d.error()[0] == 0
    got = d.state()["qcond"]
    for g in range(2):
        liq = sum(condensate_si(xi[k], R[k], m_s[k], SOLUTE_NACL[0]) for k in range(5) if gbx[k] == g)
        _close(got[g], liq / (rho_dry_si(p[g], T[g], qv[g]) * dx ** 3), f"qcond gridbox {g}", 1e-11)
    d.close()


def pin_state_update(orc):
    """One condensation pass (condensation.cpp:126-137): the liquid gained, Delta q, follows from each
    drop's radius before and after (the solute volume cancels); qv -= Delta q, qc += Delta q and
    T += L Delta q / c_p with c_p = c_pd + c_pv qv' + c_l qc' of the updated mixing ratios."""
    ndims, dx = (1, 1, 1), 5.0
    xi = [10 ** 9] * 6 + [4 * 10 ** 8] * 3
    R = [1e-5, 1.1e-5, 9e-6, 1.3e-5, 8e-6, 1.05e-5, 3e-5, 5e-5, 2.2e-5]
    m_s = [1e-17] * 9
    p, T = 8.7e4, 283.4
    qv = 1.04 * EPS * psat_tetens_si(T) / (p - 1.04 * psat_tetens_si(T))  # S = 1.04
    qc = 0.012
    sup = _drops([0] * 9, xi, R, m_s, dx, ndims)
    d = _domain(orc, ndims, dx, sup, _state(1, dx ** 3, [p], [T], [qv], [qc]), SOLUTE_NACL)
    assert d.condensation(abi.CondParams(100, 1, 50, 1e-6, 1e-8, 0.001, 0, 10)) == 0 and d.error()[0] == 0
    out = d.download()
    st = d.state()
    order = np.argsort(out["sdid"])
    Rn = out["radius"][order] * R0
    gained = sum(xi[k] * RHO_L * 4.0 / 3.0 * math.pi * (Rn[k] ** 3 - R[k] ** 3) for k in range(9))
    assert gained > 0  # supersaturated: the drops grow
    dq = gained / (rho_dry_si(p, T, qv) * dx ** 3)
    qv1, qc1 = qv - dq, qc + dq
    dT = LATENT_V * dq / (CP_DRY + CP_V * qv1 + C_L * qc1)
    _close(qv - st["qvap"][0], dq, "Delta qv", 1e-9)
    _close(st["qcond"][0] - qc, dq, "Delta qc", 1e-9)
    _close(st["temp"][0] * TEMP0 - T, dT, "Delta T", 1e-9)
    assert dT > 0.01  # large enough that the comparison resolves c_p's 1e-8 terms
    d.close()


def pin_units(orc):
    """Scales and dimensionless constants of libs/cleoconstants.hpp:65-104 from their SI definitions."""
    c = (C.c_double * 10)()
    orc.lib.orc_constants.argtypes = [C.POINTER(C.c_double)]
    orc.lib.orc_constants(c)
    want = [RHO0, MASS0, TIME0 / (RHO0 * R0 * R0), LATENT_V / (TEMP0 * CP_DRY), RHO_L / RHO0, 2016.5 / RHO0,
            0.05844277 / MR_DRY, EPS, 1.0, 1e-3]
    for k, w in enumerate(want):
        _close(c[k], w, f"constant {k}", 1e-15)


PINS = [pin_units, pin_saturation_pressure, pin_supersaturation_ratio, pin_diffusion_factor, pin_ventilation_factor,
        pin_kohler_factors, pin_terminal_velocities, pin_qcond_recompute, pin_state_update]
