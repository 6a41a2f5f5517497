"""Synthetic superdroplet populations and gridbox states for the five workloads.

All quantities are returned in the reference's dimensionless units
(libs/cleoconstants.hpp:65-104): radius / R0, msol / MASS0, coordinates / COORD0,
volume / VOL0, T / TEMP0, p / P0, velocities / W0.

Workloads (sizes are parameters so tests can use small instances):
  golovin_box   1 gridbox of 1e6 m^3, exponential-in-volume radii (Shima et al. 2009,
                examples/boxmodelcollisions/attrgens_shima2009.py:22-62), equal xi, no solute.
  coal_breakup  Poisson counts per gridbox, bimodal lognormal radii, xi ~ U{1e6..1e8},
                log-uniform solute mass, radii raised to the dry radius.
  condensation  lognormal dry radii of ammonium sulphate, wet radius at Koehler equilibrium
                with the gridbox supersaturation, per-gridbox T, p, S.
  coupled       coal_breakup attributes + condensation thermodynamics + the 2-D
                divergence-free flow of cleopy/thermobinary_src/windsgen.py:57-73.
Drops are generated gridbox by gridbox, so the output is already sorted by sdgbxindex.
"""
import numpy as np

# libs/cleoconstants.hpp:32-104
R0 = 1e-6
TIME0 = 1000.0
COORD0 = TIME0 * 1.0
VOL0 = COORD0 ** 3
P0 = 100000.0
TEMP0 = 273.15
CP_DRY = 1004.64
RHO0 = P0 / (CP_DRY * TEMP0)
MASS0 = R0 * R0 * R0 * RHO0
MR_WATER = 0.01801528
MR_DRY = 0.028966216
MR_RATIO = MR_WATER / MR_DRY
RHO_L = 998.203

NACL = dict(rho=2016.5, mr=0.05844277, ionic=2.0)
AMMONIUM_SULPHATE = dict(rho=1770.0, mr=0.13214, ionic=3.0)


def solute_dimless(sol):
    """(rho_sol, mr_sol, ionic) in the units of cleo_solute."""
    return sol["rho"] / RHO0, sol["mr"] / MR_DRY, float(sol["ionic"])


def uniform_grid(ndims, spacing_m):
    """Face coordinates per axis (z, x, y), dimensionless."""
    return [np.arange(n + 1, dtype=np.float64) * (spacing_m / COORD0) for n in ndims]


def psat_dimless(temp):
    """Dimensionless saturation pressure, same formula as thermodynamic_equations.cpp:39-52."""
    T = temp * TEMP0
    return (611.655 * np.exp(17.4146 * (T - 273.16) / (T - 33.639))) / P0


def qvap_from_s(press, temp, s_ratio):
    """Invert s = p qv / ((Mr_ratio + qv) psat) (thermodynamic_equations.hpp:79-81)."""
    psat = psat_dimless(temp)
    return s_ratio * MR_RATIO * psat / (press - s_ratio * psat)


def _dryradius(msol, rho_sol_dimless):
    return np.power(3.0 / (4.0 * np.pi) * msol / rho_sol_dimless, 1.0 / 3.0)


def _koehler_ab(temp, msol, sol):
    """Koehler a, b as kohler_factors (thermodynamic_equations.hpp:96-107), dimensionless."""
    rho_sol, mr_sol, ionic = solute_dimless(sol)
    akoh = 3.3e-7 / (TEMP0 * R0) / temp
    bkoh = 4.3e-6 * RHO0 / MR_DRY * msol * ionic / mr_sol
    return akoh, bkoh


def equilibrium_radius(s_ratio, temp, msol, sol, rdry):
    """Wet radius with 1 + a/r - b/r^3 = S on the stable branch r < r_crit.  Above the activation
    supersaturation no equilibrium exists: such a drop starts on the growing branch at 1.5 r_crit
    (a drop left just below r_crit needs more than 50 Newton iterations in its first solve from the
    reference's 1 mm guess, which the reference aborts on)."""
    akoh, bkoh = _koehler_ab(temp, msol, sol)
    rcrit = np.sqrt(3.0 * bkoh / akoh)
    scrit = 1.0 + np.sqrt(4.0 * akoh ** 3 / 27.0 / bkoh)
    s_eff = np.minimum(s_ratio, 1.0 + (scrit - 1.0) * 0.999)
    lo = rdry * 1.000001
    hi = rcrit.copy()
    f = lambda r: 1.0 + akoh / r - bkoh / r ** 3 - s_eff  # noqa: E731
    # f(lo) < 0 near the dry radius for soluble particles, f(hi) >= 0 when s_eff < scrit
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        neg = f(mid) < 0
        lo = np.where(neg, mid, lo)
        hi = np.where(neg, hi, mid)
    return np.where(s_ratio >= scrit, 1.5 * rcrit, np.maximum(0.5 * (lo + hi), rdry))


class Population:
    """SoA drops + gridbox state + grid description."""

    def __init__(self, ndims, halves, supers, state, solute, periodic=(0, 1, 1)):
        self.ndims = tuple(ndims)
        self.ngbx = int(np.prod(ndims))
        self.halves = halves
        self.supers = supers
        self.state = state
        self.solute = solute
        self.periodic = periodic

    @property
    def n(self):
        return len(self.supers["xi"])


def _counts(rng, ngbx, mean):
    return rng.poisson(mean, ngbx).astype(np.int64)


def _coords(rng, ndims, halves, counts):
    nz, nx, ny = ndims
    g = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    iz, ix, iy = g % nz, (g // nz) % nx, g // (nz * nx)
    zh, xh, yh = halves
    u = rng.random((3, len(g)))
    c3 = zh[iz] + u[0] * (zh[iz + 1] - zh[iz])
    c1 = xh[ix] + u[1] * (xh[ix + 1] - xh[ix])
    c2 = yh[iy] + u[2] * (yh[iy + 1] - yh[iy])
    # keep strictly inside [lower, upper)
    c3 = np.minimum(c3, np.nextafter(zh[iz + 1], -np.inf))
    c1 = np.minimum(c1, np.nextafter(xh[ix + 1], -np.inf))
    c2 = np.minimum(c2, np.nextafter(yh[iy + 1], -np.inf))
    return g.astype(np.uint32), c3, c1, c2


def _state(ngbx, vol_dimless, temp_K, press_Pa, s_ratio):
    temp = np.asarray(temp_K, dtype=np.float64) / TEMP0 * np.ones(ngbx)
    press = np.asarray(press_Pa, dtype=np.float64) / P0 * np.ones(ngbx)
    qv = qvap_from_s(press, temp, np.asarray(s_ratio) * np.ones(ngbx))
    z = np.zeros(ngbx)
    return dict(volume=np.full(ngbx, vol_dimless), press=press, temp=temp, qvap=qv, qcond=z.copy(),
                wvel_lo=z.copy(), wvel_hi=z.copy(), uvel_lo=z.copy(), uvel_hi=z.copy(),
                vvel_lo=z.copy(), vvel_hi=z.copy())


def _bimodal_radii(rng, n):
    pick = rng.random(n) < 0.5
    r1 = rng.lognormal(np.log(0.05e-6), np.log(1.4), n)
    r2 = rng.lognormal(np.log(15e-6), np.log(1.6), n)
    return np.where(pick, r1, r2) / R0


def golovin_box(seed=20261202, nsupers=8192, volume_m3=1e6, r0_m=30.531e-6, numconc=2.0 ** 23):
    rng = np.random.default_rng(seed)
    side = volume_m3 ** (1.0 / 3.0)
    ndims = (1, 1, 1)
    halves = uniform_grid(ndims, side)
    counts = np.array([nsupers])
    g, c3, c1, c2 = _coords(rng, ndims, halves, counts)
    vol0 = 4.0 / 3.0 * np.pi * r0_m ** 3
    vols = rng.exponential(vol0, nsupers)
    radius = np.power(3.0 / (4.0 * np.pi) * vols, 1.0 / 3.0) / R0
    xi = np.full(nsupers, int(round(numconc * volume_m3 / nsupers)), dtype=np.uint64)
    sup = dict(sdgbxindex=g, coord3=c3, coord1=c1, coord2=c2, xi=xi, radius=radius,
               msol=np.zeros(nsupers), sdid=np.arange(nsupers, dtype=np.uint64))
    state = _state(1, volume_m3 / VOL0, 288.15, 1e5, 1.0)
    return Population(ndims, halves, sup, state, NACL)


def coal_breakup(seed=20261202, ndims=(64, 64, 64), spacing_m=50.0, mean=128.0, solute=NACL):
    rng = np.random.default_rng(seed)
    ngbx = int(np.prod(ndims))
    halves = uniform_grid(ndims, spacing_m)
    counts = _counts(rng, ngbx, mean)
    g, c3, c1, c2 = _coords(rng, ndims, halves, counts)
    n = len(g)
    msol = np.exp(rng.uniform(np.log(1e-18), np.log(1e-15), n)) / MASS0
    rho_sol, _, _ = solute_dimless(solute)
    radius = np.maximum(_bimodal_radii(rng, n), _dryradius(msol, rho_sol))
    xi = rng.integers(1_000_000, 100_000_000, n, endpoint=True).astype(np.uint64)
    sup = dict(sdgbxindex=g, coord3=c3, coord1=c1, coord2=c2, xi=xi, radius=radius, msol=msol,
               sdid=np.arange(n, dtype=np.uint64))
    state = _state(ngbx, spacing_m ** 3 / VOL0, 288.15, 1e5, 1.0)
    return Population(ndims, halves, sup, state, solute)


def _thermo(rng, ngbx):
    temp_K = rng.uniform(270.0, 300.0, ngbx)
    press_Pa = rng.uniform(7e4, 1.0e5, ngbx)
    s_ratio = 1.0 + rng.uniform(-0.05, 0.01, ngbx)
    return temp_K, press_Pa, s_ratio


def condensation(seed=20261202, ndims=(64, 64, 64), spacing_m=50.0, mean=122.0,
                 solute=AMMONIUM_SULPHATE):
    rng = np.random.default_rng(seed)
    ngbx = int(np.prod(ndims))
    halves = uniform_grid(ndims, spacing_m)
    counts = _counts(rng, ngbx, mean)
    g, c3, c1, c2 = _coords(rng, ndims, halves, counts)
    n = len(g)
    temp_K, press_Pa, s_ratio = _thermo(rng, ngbx)
    state = _state(ngbx, spacing_m ** 3 / VOL0, temp_K, press_Pa, s_ratio)
    rho_sol, _, _ = solute_dimless(solute)
    rdry = rng.lognormal(np.log(0.04e-6), np.log(1.5), n) / R0
    msol = 4.0 / 3.0 * np.pi * rdry ** 3 * rho_sol
    rdry = _dryradius(msol, rho_sol)
    gi = g.astype(np.int64)
    radius = equilibrium_radius(s_ratio[gi], state["temp"][gi], msol, solute, rdry)
    xi = rng.integers(1_000_000, 100_000_000, n, endpoint=True).astype(np.uint64)
    sup = dict(sdgbxindex=g, coord3=c3, coord1=c1, coord2=c2, xi=xi, radius=radius, msol=msol,
               sdid=np.arange(n, dtype=np.uint64))
    return Population(ndims, halves, sup, state, solute)


def divfree_winds(ndims, halves, wmax_mps=2.0):
    """Face winds of the 2-D (z, x) divergence-free flow (windsgen.py:57-73) with unit density
    and the peak vertical wind equal to wmax; domain length = one wavelength in x, half in z."""
    nz, nx, ny = ndims
    zh, xh, _ = halves
    zlen, xlen = zh[-1] * COORD0, xh[-1] * COORD0
    ztilda, xtilda = zlen / np.pi, xlen / (2.0 * np.pi)
    zc = 0.5 * (zh[:-1] + zh[1:]) * COORD0
    xc = 0.5 * (xh[:-1] + xh[1:]) * COORD0
    zf, xf = zh * COORD0, xh * COORD0
    # gridbox g = iz + nz*(ix + nx*iy)
    g = np.arange(nz * nx * ny)
    iz, ix = g % nz, (g // nz) % nx
    w = lambda z, x: wmax_mps * np.sin(z / ztilda) * np.sin(x / xtilda)  # noqa: E731
    u = lambda z, x: wmax_mps * xtilda / ztilda * np.cos(z / ztilda) * np.cos(x / xtilda)  # noqa: E731
    return dict(wvel_lo=w(zf[iz], xc[ix]), wvel_hi=w(zf[iz + 1], xc[ix]),
                uvel_lo=u(zc[iz], xf[ix]), uvel_hi=u(zc[iz], xf[ix + 1]))


def coupled(seed=20261202, ndims=(128, 128, 128), spacing_m=25.0, mean=64.0, solute=AMMONIUM_SULPHATE,
            wmax_mps=2.0):
    pop = coal_breakup(seed, ndims, spacing_m, mean, solute)
    rng = np.random.default_rng(seed + 1)
    temp_K, press_Pa, s_ratio = _thermo(rng, pop.ngbx)
    st = _state(pop.ngbx, spacing_m ** 3 / VOL0, temp_K, press_Pa, s_ratio)
    st.update(divfree_winds(ndims, pop.halves, wmax_mps))
    pop.state = st
    return pop


def _device_drops(torch, device, gen, ndims, spacing_m, mean):
    """Poisson counts per gridbox, the drops' gridboxes (int32) and coordinates, on the GPU."""
    ngbx = int(np.prod(ndims))
    nz, nx, ny = ndims
    counts = torch.poisson(torch.full((ngbx,), float(mean), device=device, dtype=torch.float32),
                           generator=gen).to(torch.int64)
    gidx = torch.repeat_interleave(torch.arange(ngbx, device=device, dtype=torch.int64), counts)
    n = int(gidx.numel())
    iz, ix, iy = gidx % nz, (gidx // nz) % nx, gidx // (nz * nx)
    dzc = spacing_m / COORD0
    out = {}
    for name, i in (("coord3", iz), ("coord1", ix), ("coord2", iy)):
        u = torch.rand(n, device=device, dtype=torch.float64, generator=gen)
        lo = i.to(torch.float64) * dzc
        c = lo + u * dzc
        out[name] = torch.minimum(c, torch.nextafter(lo + dzc, torch.zeros_like(lo)))
        del u, lo, c
    del iz, ix, iy
    out["sdgbxindex"] = gidx.to(torch.int32)
    return out, n


def _device_coal_breakup_attrs(torch, device, gen, out, n, solute):
    """coal_breakup's attributes on the GPU: log-uniform solute, bimodal lognormal radii raised to
    the dry radius, xi ~ U{1e6..1e8}."""
    rho_sol, _, _ = solute_dimless(solute)
    lmin, lmax = np.log(1e-18), np.log(1e-15)
    msol = torch.exp(lmin + (lmax - lmin) * torch.rand(n, device=device, dtype=torch.float64, generator=gen)) / MASS0
    pick = torch.rand(n, device=device, dtype=torch.float64, generator=gen) < 0.5
    r1 = torch.exp(np.log(0.05e-6) + np.log(1.4) * torch.randn(n, device=device, dtype=torch.float64, generator=gen))
    r2 = torch.exp(np.log(15e-6) + np.log(1.6) * torch.randn(n, device=device, dtype=torch.float64, generator=gen))
    radius = torch.where(pick, r1, r2) / R0
    del r1, r2, pick
    rdry = torch.pow(3.0 / (4.0 * np.pi) * msol / rho_sol, 1.0 / 3.0)
    out["radius"] = torch.maximum(radius, rdry)
    out["msol"] = msol
    del radius, rdry
    out["xi"] = torch.randint(1_000_000, 100_000_001, (n,), device=device, dtype=torch.int64, generator=gen)
    out["sdid"] = torch.arange(n, device=device, dtype=torch.int64)


def coal_breakup_device(torch, device, seed=20261202, ndims=(64, 64, 64), spacing_m=50.0, mean=128.0,
                        solute=NACL):
    """The `coal_breakup` workload drawn on the GPU (same distributions, torch's random stream):
    (arrays dict of device tensors, state dict of numpy arrays, halves), as coupled_device."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    out, n = _device_drops(torch, device, gen, ndims, spacing_m, mean)
    _device_coal_breakup_attrs(torch, device, gen, out, n, solute)
    ngbx = int(np.prod(ndims))
    return out, _state(ngbx, spacing_m ** 3 / VOL0, 288.15, 1e5, 1.0), uniform_grid(ndims, spacing_m)


def condensation_device(torch, device, seed=20261202, ndims=(64, 64, 64), spacing_m=50.0, mean=122.0,
                        solute=AMMONIUM_SULPHATE):
    """The `condensation` workload drawn on the GPU: lognormal dry radii, wet radius at Koehler
    equilibrium with the gridbox's supersaturation (as equilibrium_radius: the same bisection, and
    1.5 r_crit for activated drops)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    out, n = _device_drops(torch, device, gen, ndims, spacing_m, mean)
    ngbx = int(np.prod(ndims))
    rng = np.random.default_rng(seed + 1)
    temp_K, press_Pa, s_ratio = _thermo(rng, ngbx)
    st = _state(ngbx, spacing_m ** 3 / VOL0, temp_K, press_Pa, s_ratio)
    rho_sol, mr_sol, ionic = solute_dimless(solute)
    rdry = torch.exp(np.log(0.04e-6) + np.log(1.5) * torch.randn(n, device=device, dtype=torch.float64,
                                                                  generator=gen)) / R0
    msol = 4.0 / 3.0 * np.pi * rdry ** 3 * rho_sol
    rdry = torch.pow(3.0 / (4.0 * np.pi) * msol / rho_sol, 1.0 / 3.0)
    g = out["sdgbxindex"].to(torch.int64)
    temp = torch.from_numpy(st["temp"]).to(device)[g]
    s = torch.from_numpy(s_ratio).to(device)[g]
    del g
    akoh = 3.3e-7 / (TEMP0 * R0) / temp
    bkoh = 4.3e-6 * RHO0 / MR_DRY * msol * ionic / mr_sol
    rcrit = torch.sqrt(3.0 * bkoh / akoh)
    scrit = 1.0 + torch.sqrt(4.0 * akoh ** 3 / 27.0 / bkoh)
    s_eff = torch.minimum(s, 1.0 + (scrit - 1.0) * 0.999)
    lo, hi = rdry * 1.000001, rcrit
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        neg = (1.0 + akoh / mid - bkoh / mid ** 3 - s_eff) < 0
        lo = torch.where(neg, mid, lo)
        hi = torch.where(neg, hi, mid)
    out["radius"] = torch.where(s >= scrit, 1.5 * rcrit, torch.maximum(0.5 * (lo + hi), rdry))
    out["msol"] = msol
    del lo, hi, mid, neg, akoh, bkoh, rcrit, scrit, s_eff, temp, s
    out["xi"] = torch.randint(1_000_000, 100_000_001, (n,), device=device, dtype=torch.int64, generator=gen)
    out["sdid"] = torch.arange(n, device=device, dtype=torch.int64)
    return out, st, uniform_grid(ndims, spacing_m)


def coupled_device(torch, device, seed=20261202, ndims=(128, 128, 128), spacing_m=25.0, mean=64.0,
                   solute=AMMONIUM_SULPHATE, wmax_mps=2.0):
    """The `coupled` workload drawn on the GPU with torch (same distributions, different
    random stream): returns (arrays dict of device tensors, state dict of numpy arrays, halves).
    sdgbxindex is int32 and xi, sdid int64 tensors; their bits are the u32 / u64 the C ABI reads."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    ngbx = int(np.prod(ndims))
    halves = uniform_grid(ndims, spacing_m)
    out, n = _device_drops(torch, device, gen, ndims, spacing_m, mean)
    _device_coal_breakup_attrs(torch, device, gen, out, n, solute)
    rng = np.random.default_rng(seed + 1)
    temp_K, press_Pa, s_ratio = _thermo(rng, ngbx)
    st = _state(ngbx, spacing_m ** 3 / VOL0, temp_K, press_Pa, s_ratio)
    st.update(divfree_winds(ndims, halves, wmax_mps))
    return out, st, halves
