"""Size-independent properties at the benchmark's full size, where the Serial oracle cannot follow:
the coupled workload of 128 x 128 x 128 gridboxes with Poisson(64) drops (134M superdroplets),
20 steps replayed as a hipGraph, then one more collision event and one more condensation pass.

  - the store after the steps: every drop is in the domain or out of it (in + out = ntot), refs are
    monotone with refs[ngbx] the in-domain count, every in-domain drop's gridbox index is its
    position's bucket and its coordinates lie inside that gridbox's bounds, sdIds are a
    permutation of the uploaded ones;
  - sum over all drops of xi * msol is unchanged by 20 coupled steps (collisions move solute between
    drops, breakup splits it, condensation and motion leave it);
  - one collision event keeps sum xi r^3 and sum xi msol (coalescence and breakup conserve liquid
    volume and solute) to 1e-12 relative;
  - one condensation pass keeps qv + qc of every gridbox (the State update moves the condensed mass
    between them) to 1e-12 relative."""
import ctypes as C

import numpy as np
import pytest

from cleo_amd import abi
from cleo_amd import populations as P
from tests.helpers import coll_params, cond_params, motion_params, sdm_params

pytestmark = pytest.mark.gpu
# before the library: one HIP runtime for both (cleo_amd/lib.py)
torch = pytest.importorskip("torch")


def _ptr(t, ctype):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(ctype))


def _sums(sup):
    xi = sup["xi"].astype(np.float64)
    return float(np.sum(xi * sup["radius"] ** 3)), float(np.sum(xi * sup["msol"]))


def _check_store(sup, refs, halves, ndims, ids0):
    nz, nx, ny = ndims
    ngbx = nz * nx * ny
    nd = int(refs[-1])
    assert nd == sup["ndomain"]
    assert np.all(np.diff(refs.astype(np.int64)) >= 0)
    g = sup["sdgbxindex"]
    assert np.all(g[:nd] < ngbx) and np.all(g[nd:] == abi.OOB)  # in + out = ntot
    expect = np.repeat(np.arange(ngbx, dtype=np.uint32), np.diff(refs.astype(np.int64)))
    assert np.array_equal(g[:nd], expect)  # sorted, and refs are the buckets
    gi = g[:nd].astype(np.int64)
    iz, ix, iy = gi % nz, (gi // nz) % nx, gi // (nz * nx)
    zh, xh, yh = halves
    for c, h, i in ((sup["coord3"][:nd], zh, iz), (sup["coord1"][:nd], xh, ix), (sup["coord2"][:nd], yh, iy)):
        assert np.all(c >= h[i]) and np.all(c < h[i + 1])
    assert np.array_equal(np.sort(sup["sdid"]), ids0)


def test_fullsize_coupled_invariants(gpu):
    L = gpu.lib
    ndims = (128, 128, 128)
    print("case: full-size coupled workload, 134M drops, 20 graph steps", flush=True)
    arrs, st, halves = P.coupled_device(torch, torch.device("cuda:0"), ndims=ndims)
    torch.cuda.synchronize()
    n = int(arrs["xi"].numel())
    xi0 = arrs["xi"].to(torch.float64)
    msol_total0 = float(torch.sum(xi0 * arrs["msol"]).item())
    del xi0
    soa = abi.SupersSoA()
    soa.sdgbxindex = _ptr(arrs["sdgbxindex"], C.c_uint32)
    soa.coord3, soa.coord1, soa.coord2 = (_ptr(arrs[k], C.c_double) for k in ("coord3", "coord1", "coord2"))
    soa.xi = _ptr(arrs["xi"], C.c_uint64)
    soa.radius, soa.msol = _ptr(arrs["radius"], C.c_double), _ptr(arrs["msol"], C.c_double)
    soa.sdid = _ptr(arrs["sdid"], C.c_uint64)
    d = abi.Domain(gpu, n, ndims, *halves)
    d.check(L.cleo_upload_supers_device(d.h, C.byref(soa), n), "upload_supers_device")
    d.ntot = n
    del arrs, soa
    torch.cuda.empty_cache()
    d.set_state(st)
    d.set_solute(*P.solute_dimless(P.AMMONIUM_SULPHATE))
    d.rng_counter(20261202)
    d.set_sdm(sdm_params(cond=cond_params(), coll=coll_params(), motion=motion_params()))
    assert L.cleo_run_steps_graph(d.h, 0, 100, 20) == 0
    assert L.cleo_synchronize(d.h) == 0
    assert d.error()[0] == 0, d.error()
    st1 = d.stats()
    print(f"  20 steps: {st1['coll_events']} collision events, {st1['nulls']} nulls", flush=True)
    sup = d.download()
    _check_store(sup, d.refs(), halves, ndims, np.arange(n, dtype=np.uint64))
    m3_before, ms_before = _sums(sup)
    assert ms_before == pytest.approx(msol_total0, rel=1e-11)
    del sup

    print("  one collision event: sum xi r^3 and sum xi msol", flush=True)
    d.reset_stats()
    assert d.collisions(coll_params()) == 0
    sup = d.download()
    m3_after, ms_after = _sums(sup)
    assert d.stats()["coll_events"] > 0
    assert m3_after == pytest.approx(m3_before, rel=1e-12)
    assert ms_after == pytest.approx(ms_before, rel=1e-12)
    del sup

    print("  one condensation pass: qv + qc per gridbox", flush=True)
    s0 = d.state()
    assert d.condensation(cond_params()) == 0
    s1 = d.state()
    w0, w1 = s0["qvap"] + s0["qcond"], s1["qvap"] + s1["qcond"]
    assert np.all(np.abs(w1 - w0) <= 1e-12 * w0)
    assert np.any(s1["temp"] != s0["temp"])
    assert d.error()[0] == 0
    d.close()
