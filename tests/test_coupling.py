"""One-way coupling to prescribed dynamics (FromFileComms::receive_dynamics,
coupldyn_fromfile/fromfilecomms.cpp:36-62, fromfile_cartesian_dynamics.cpp:22-31,150-215).

The oracle's face indexing is pinned by numpy arrays shaped like the C grid (w on z faces
(ny, nx, nz+1), u on x faces (ny, nx+1, nz), v on y faces (ny+1, nx, nz)) and indexed by
(j, i, k): the lower face of gridbox (k, i, j) and the next face along the axis.  A null wind
array gives zeros; volumes stay.  The HIP entry point takes device arrays and is compared bit for
bit with the oracle (marked gpu).  The thermodynamics file reader (host only) is checked on a file
written in cleopy's layout."""
import ctypes as C

import numpy as np
import pytest

from cleo_amd import abi
from cleo_amd import populations as P
from tests.helpers import SEED, make_domain
from tests.test_binfiles import write_cleo_binary

NDIMS = (5, 4, 3)


def _fields(rng, ndims):
    nz, nx, ny = ndims
    ngbx = nz * nx * ny
    th = {k: rng.uniform(0.5, 1.5, ngbx) for k in ("press", "temp", "qvap", "qcond")}
    w = rng.normal(size=(ny, nx, nz + 1))
    u = rng.normal(size=(ny, nx + 1, nz))
    v = rng.normal(size=(ny + 1, nx, nz))
    return th, w, u, v


def _step(th, w, u, v, keep):
    """DynamicsStep over host arrays; `keep` holds the flattened copies alive."""
    arrs = dict(th)
    arrs.update(wvel_zfaces=w.ravel(), uvel_xfaces=u.ravel(), vvel_yfaces=v.ravel())
    s = abi.DynamicsStep()
    for name, a in arrs.items():
        if a is None:
            continue
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        setattr(s, name, a.ctypes.data)
    return s


def _expected(th, w, u, v, ndims):
    nz, nx, ny = ndims
    g = np.arange(nz * nx * ny)
    k, i, j = g % nz, (g // nz) % nx, g // (nz * nx)
    return dict(press=th["press"], temp=th["temp"], qvap=th["qvap"], qcond=th["qcond"],
                wvel_lo=w[j, i, k], wvel_hi=w[j, i, k + 1], uvel_lo=u[j, i, k], uvel_hi=u[j, i + 1, k],
                vvel_lo=v[j, i, k], vvel_hi=v[j + 1, i, k])


def test_oracle_receive_dynamics_face_indexing(oracle):
    rng = np.random.default_rng(SEED)
    pop = P.coupled(SEED, ndims=NDIMS, mean=4)
    d = make_domain(oracle, pop)
    th, w, u, v = _fields(rng, NDIMS)
    keep = []
    assert oracle.receive_dynamics(d.h, C.byref(_step(th, w, u, v, keep))) == abi.CLEO_OK
    st = d.state()
    for name, e in _expected(th, w, u, v, NDIMS).items():
        assert np.array_equal(st[name], e), name
    assert np.array_equal(st["volume"], pop.state["volume"])
    d.close()


def test_oracle_receive_dynamics_null_winds(oracle):
    rng = np.random.default_rng(SEED + 1)
    pop = P.coupled(SEED, ndims=NDIMS, mean=4)
    d = make_domain(oracle, pop)
    th, w, u, v = _fields(rng, NDIMS)
    keep = []
    s = _step(th, w, u, v, keep)
    s.uvel_xfaces = None
    s.vvel_yfaces = None
    assert oracle.receive_dynamics(d.h, C.byref(s)) == abi.CLEO_OK
    st = d.state()
    for name in ("uvel_lo", "uvel_hi", "vvel_lo", "vvel_hi"):
        assert not st[name].any()
    e = _expected(th, w, u, v, NDIMS)
    assert np.array_equal(st["wvel_lo"], e["wvel_lo"]) and np.array_equal(st["temp"], e["temp"])
    d.close()


def test_thermofile_round_trip(tmp_path):
    from cleo_amd import lib
    so = C.CDLL(lib.LIB_PATH)
    so.cleo_read_thermofile.argtypes = [C.c_char_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]
    rng = np.random.default_rng(SEED)
    nsteps, ngbx = 4, 60
    data = rng.uniform(0.9, 1.1, nsteps * ngbx)
    path = str(tmp_path / "press.dat")
    write_cleo_binary(path, [data], units=[b"P"], scales=[P.P0])
    n = C.c_uint64()
    assert so.cleo_read_thermofile(path.encode(), None, 0, C.byref(n)) == abi.CLEO_OK and n.value == len(data)
    out = np.zeros(n.value)
    assert so.cleo_read_thermofile(path.encode(), out.ctypes.data_as(C.c_void_p), n.value, C.byref(n)) == abi.CLEO_OK
    assert np.array_equal(out, data)
    assert so.cleo_read_thermofile(path.encode(), out.ctypes.data_as(C.c_void_p), n.value - 1,
                                   C.byref(n)) == abi.CLEO_E_ARG
    bad = str(tmp_path / "ints.dat")
    write_cleo_binary(bad, [np.arange(5, dtype=np.uint32)])
    assert so.cleo_read_thermofile(bad.encode(), None, 0, C.byref(n)) == abi.CLEO_E_FILE


@pytest.mark.gpu
@pytest.mark.parametrize("winds", ["all", "w only"])
def test_gpu_receive_dynamics_matches_oracle(gpu, oracle, winds):
    import torch
    print(f"case: receive dynamics from device arrays, winds={winds}", flush=True)
    rng = np.random.default_rng(SEED + 2)
    pop = P.coupled(SEED, ndims=NDIMS, mean=8)
    th, w, u, v = _fields(rng, NDIMS)
    host = {**th, "wvel_zfaces": w.ravel(), "uvel_xfaces": u.ravel(), "vvel_yfaces": v.ravel()}
    if winds == "w only":
        host["uvel_xfaces"] = None
        host["vvel_yfaces"] = None
    dev = {k: (torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") if a is not None else None)
           for k, a in host.items()}
    torch.cuda.synchronize()
    s_dev = abi.DynamicsStep()
    for k, t in dev.items():
        setattr(s_dev, k, t.data_ptr() if t is not None else None)
    g = make_domain(gpu, pop)
    assert gpu.receive_dynamics(g.h, C.byref(s_dev)) == abi.CLEO_OK
    o = make_domain(oracle, pop)
    keep = []
    s_host = abi.DynamicsStep()
    for k, a in host.items():
        if a is not None:
            a = np.ascontiguousarray(a)
            keep.append(a)
            setattr(s_host, k, a.ctypes.data)
    assert oracle.receive_dynamics(o.h, C.byref(s_host)) == abi.CLEO_OK
    gs, os_ = g.state(), o.state()
    for name in abi.STATE_FIELDS:
        assert np.array_equal(gs[name], os_[name]), name
    g.close()
    o.close()
