"""CLEO's binary initialisation files (initialise/readbinary.cpp:22-99, init_all_supers_from_binary.cpp:
57-100, init_supers_from_binary.cpp:44-63, gbx_bounds_from_binary.cpp:29-58) read by
cleo_read_initsupers, cleo_read_initsupers_padded and cleo_read_gridfile.

The fixtures tests/golden/*.dat were written by the reference's own writers (cleopy's
writebinary.writebinary and create_gbxboundaries.write_gridboxboundaries_binary, run by
tests/golden/make_binfiles.py), with what was written kept in tests/golden/binfiles_expected.npz;
they pin the readers bit for bit.  Other cases are written here by `write_cleo_binary`, a
restatement of the layout cleopy/writebinary.py packs (little-endian, unpadded: 4 uint32, the
global metadata string, per variable 3 uint32 + 2 chars + 1 double, then the data), checked against
the fixtures.  The readers are host code, so these tests need no GPU, except the last, which uploads
a padded store read from a fixture to the GPU and to the oracle and compares a step of both."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from cleo_amd import abi
from cleo_amd import populations as P
from tests.helpers import (SEED, assert_state_match, assert_supers_match, coll_params, cond_params, make_domain,
                           sdm_params)

_FMT = {np.dtype(np.uint32): "I", np.dtype(np.uint64): "Q", np.dtype(np.float64): "d"}


def write_cleo_binary(path, arrays, units=None, scales=None, metastr="test file"):
    """Writes `arrays` (a list of 1-D numpy arrays of uint32, uint64 or float64) in the layout of
    cleopy/writebinary.py."""
    nvars = len(arrays)
    mbytes = struct.calcsize("<IIIccd")
    chars = metastr.encode()
    d0 = 16 + len(chars) + nvars * mbytes
    head = struct.pack("<IIII", d0, len(chars), nvars, mbytes) + chars
    metas, data, pos = b"", b"", d0
    for i, a in enumerate(arrays):
        code = _FMT[a.dtype]
        size = struct.calcsize(code)
        unit = (units or [b" "] * nvars)[i]
        sf = float((scales or [1.0] * nvars)[i])
        metas += struct.pack("<IIIccd", pos, size, len(a), code.encode(), unit, sf)
        data += a.astype(a.dtype.newbyteorder("<")).tobytes()
        pos += size * len(a)
    with open(path, "wb") as f:
        f.write(head + metas + data)


def _lib():
    from cleo_amd import lib
    so = C.CDLL(lib.LIB_PATH)
    so.cleo_read_initsupers.argtypes = [C.c_char_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]
    so.cleo_read_gridfile.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p, C.c_void_p]
    return so


def read_initsupers(so, path):
    n = C.c_uint64()
    rc = so.cleo_read_initsupers(path.encode(), None, 0, C.byref(n))
    if rc != abi.CLEO_OK:
        return rc, None
    arrs = abi.empty_supers(n.value)
    s = abi.soa_view(arrs)
    rc = so.cleo_read_initsupers(path.encode(), C.byref(s), n.value, C.byref(n))
    return rc, arrs


def read_gridfile(so, path):
    nd = (C.c_uint32 * 3)()
    rc = so.cleo_read_gridfile(path.encode(), nd, None, None, None)
    if rc != abi.CLEO_OK:
        return rc, None
    halves = [np.zeros(int(n) + 1) for n in nd]
    rc = so.cleo_read_gridfile(path.encode(), nd, *(h.ctypes.data_as(C.c_void_p) for h in halves))
    return rc, (tuple(int(n) for n in nd), halves)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INIT_KEYS = ("sdgbxindex", "xi", "radius", "msol", "coord3", "coord1", "coord2")


def _expected():
    return np.load(os.path.join(GOLDEN, "binfiles_expected.npz"))


def read_initsupers_padded(so, path, maxn):
    arrs = abi.empty_supers(maxn)
    s = abi.soa_view(arrs)
    n = C.c_uint64()
    so.cleo_read_initsupers_padded.argtypes = [C.c_char_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_uint64)]
    rc = so.cleo_read_initsupers_padded(path.encode(), maxn, C.byref(s), C.byref(n))
    return rc, arrs, n.value


@pytest.mark.parametrize("name", ["grid_uneven3d", "grid_delta3d", "grid_box0d"])
def test_reference_written_gridfiles(name):
    """Faces read from the reference writer's file reproduce every gridbox's six bounds and index
    as the reference computed them (dimless_gridboxboundaries), bit for bit."""
    so = _lib()
    e = _expected()
    rc, (nd, halves) = read_gridfile(so, os.path.join(GOLDEN, name + ".dat"))
    assert rc == abi.CLEO_OK
    assert nd == tuple(int(v) for v in e[name + "_ndims"])
    zh, xh, yh = halves
    nz, nx, ny = nd
    got = []
    for g in range(nz * nx * ny):
        k, i, j = g % nz, (g // nz) % nx, g // (nz * nx)
        got += [zh[k], zh[k + 1], xh[i], xh[i + 1], yh[j], yh[j + 1]]
    assert np.array_equal(np.asarray(got), e[name + "_bounds"])
    assert np.array_equal(np.arange(nz * nx * ny, dtype=np.uint32), e[name + "_idxs"])


@pytest.mark.parametrize("name", ["initsupers3d", "initsupers0d"])
def test_reference_written_initsupers(name, tmp_path):
    so = _lib()
    e = _expected()
    path = os.path.join(GOLDEN, name + ".dat")
    rc, got = read_initsupers(so, path)
    assert rc == abi.CLEO_OK
    n = len(e[f"{name}_xi"])
    for k in INIT_KEYS:
        want = e[f"{name}_{k}"]
        if k.startswith("coord") and len(want) == 0:
            assert not got[k].any(), k  # a 0-D file: coordinates read as 0
        else:
            assert np.array_equal(got[k], want), k
    assert np.array_equal(got["sdid"], np.arange(n, dtype=np.uint64))
    # the restated writer produces the reference writer's bytes for the same variables and the same
    # global metadata string
    ref = open(path, "rb").read()
    nchars = struct.unpack_from("<I", ref, 4)[0]
    mine = str(tmp_path / "mine.dat")
    write_cleo_binary(mine, [e[f"{name}_{k}"] for k in INIT_KEYS], units=[b" ", b" ", b"m", b"g", b"m", b"m", b"m"],
                      scales=[1.0, 1.0, P.R0, P.MASS0, 1000.0, 1000.0, 1000.0],
                      metastr=ref[16:16 + nchars].decode())
    assert open(mine, "rb").read() == ref


def test_padded_initsupers():
    """The file's drops, then maxnsupers - n drops out of the domain with NaN attributes and
    coordinates and xi = sdId = 0 (init_supers_from_binary.cpp:44-63); more drops than maxnsupers is
    an error (init_supers_from_binary.hpp:75-79)."""
    so = _lib()
    e = _expected()
    path = os.path.join(GOLDEN, "initsupers3d.dat")
    n = len(e["initsupers3d_xi"])
    rc, got, ninit = read_initsupers_padded(so, path, n + 5)
    assert rc == abi.CLEO_OK and ninit == n
    for k in INIT_KEYS:
        assert np.array_equal(got[k][:n], e[f"initsupers3d_{k}"]), k
    assert np.all(got["sdgbxindex"][n:] == abi.OOB)
    assert np.all(got["xi"][n:] == 0) and np.all(got["sdid"][n:] == 0)
    for k in ("radius", "msol", "coord3", "coord1", "coord2"):
        assert np.all(np.isnan(got[k][n:])), k
    rc, _, _ = read_initsupers_padded(so, path, n)
    assert rc == abi.CLEO_OK
    rc, _, _ = read_initsupers_padded(so, path, n - 1)
    assert rc == abi.CLEO_E_ARG


def _initsupers_arrays(pop):
    s = pop.supers
    return [np.asarray(s["sdgbxindex"], np.uint32), np.asarray(s["xi"], np.uint64),
            np.asarray(s["radius"], np.float64), np.asarray(s["msol"], np.float64),
            np.asarray(s["coord3"], np.float64), np.asarray(s["coord1"], np.float64),
            np.asarray(s["coord2"], np.float64)]


def _gridfile_arrays(ndims, halves):
    zh, xh, yh = halves
    nz, nx, ny = ndims
    bounds = []
    for j in range(ny):
        for i in range(nx):
            for k in range(nz):
                bounds += [zh[k], zh[k + 1], xh[i], xh[i + 1], yh[j], yh[j + 1]]
    return [np.asarray(ndims, np.uint64), np.arange(nz * nx * ny, dtype=np.uint32), np.asarray(bounds, np.float64)]


def test_initsupers_round_trip(tmp_path):
    so = _lib()
    pop = P.coupled(SEED, ndims=(5, 4, 3), mean=20)
    path = str(tmp_path / "initsupers.dat")
    write_cleo_binary(path, _initsupers_arrays(pop), units=[b" ", b" ", b"m", b"g", b"m", b"m", b"m"],
                      scales=[1.0, 1.0, P.R0, P.MASS0 * 1000, P.COORD0, P.COORD0, P.COORD0])
    rc, got = read_initsupers(so, path)
    assert rc == abi.CLEO_OK
    for k in ("sdgbxindex", "xi", "radius", "msol", "coord3", "coord1", "coord2"):
        assert np.array_equal(got[k], pop.supers[k]), k
    assert np.array_equal(got["sdid"], np.arange(pop.n, dtype=np.uint64))


def test_initsupers_without_coordinates_reads_zeros(tmp_path):
    so = _lib()
    g = np.zeros(6, np.uint32)
    xi = np.arange(1, 7, dtype=np.uint64)
    r = np.linspace(1.0, 6.0, 6)
    ms = np.full(6, 0.01)
    empty = np.zeros(0, np.float64)
    path = str(tmp_path / "box.dat")
    write_cleo_binary(path, [g, xi, r, ms, empty, empty, empty])
    rc, got = read_initsupers(so, path)
    assert rc == abi.CLEO_OK
    assert np.array_equal(got["xi"], xi) and np.array_equal(got["radius"], r)
    assert not got["coord3"].any() and not got["coord1"].any() and not got["coord2"].any()


@pytest.mark.parametrize("case", ["truncated", "sizes", "type", "missing", "fewvars"])
def test_initsupers_rejects_bad_files(tmp_path, case):
    so = _lib()
    pop = P.coupled(SEED, ndims=(2, 2, 2), mean=8)
    arrs = _initsupers_arrays(pop)
    path = str(tmp_path / "bad.dat")
    if case == "sizes":
        arrs[2] = arrs[2][:-1]
    elif case == "type":
        arrs[1] = arrs[1].astype(np.uint32)
    elif case == "fewvars":
        arrs = arrs[:4]
    if case != "missing":
        write_cleo_binary(path, arrs)
    if case == "truncated":
        data = open(path, "rb").read()
        open(path, "wb").write(data[:-5])
    rc, _ = read_initsupers(so, path)
    assert rc == abi.CLEO_E_FILE


def test_gridfile_round_trip(tmp_path):
    so = _lib()
    ndims = (7, 3, 2)
    halves = [np.cumsum(np.r_[0.0, np.linspace(0.01, 0.03, n)]) for n in ndims]  # uneven faces
    path = str(tmp_path / "grid.dat")
    write_cleo_binary(path, _gridfile_arrays(ndims, halves))
    rc, (nd, got) = read_gridfile(so, path)
    assert rc == abi.CLEO_OK and nd == ndims
    for a, b in zip(got, halves):
        assert np.array_equal(a, b)


def test_gridfile_rejects_non_cartesian_bounds(tmp_path):
    so = _lib()
    ndims = (3, 2, 2)
    halves = [np.arange(n + 1, dtype=np.float64) * 0.025 for n in ndims]
    arrs = _gridfile_arrays(ndims, halves)
    arrs[2][6 * 5 + 2] += 1e-9  # gridbox 5's lower x face moved
    path = str(tmp_path / "grid.dat")
    write_cleo_binary(path, arrs)
    rc, _ = read_gridfile(so, path)
    assert rc == abi.CLEO_E_FILE


def test_gridfile_rejects_wrapping_gridbox_count(tmp_path):
    """ndims whose product wraps a 64-bit integer (2^22 * 2^22 * 2^20 = 2^64) are not a layout."""
    so = _lib()
    path = str(tmp_path / "grid.dat")
    write_cleo_binary(path, [np.asarray([2 ** 22, 2 ** 22, 2 ** 20], np.uint64), np.zeros(0, np.uint32),
                             np.zeros(0, np.float64)])
    rc, _ = read_gridfile(so, path)
    assert rc == abi.CLEO_E_FILE


@pytest.mark.gpu
def test_gpu_upload_from_binary_files(gpu, oracle):
    """A padded store read from the reference-written fixtures goes to the GPU and to the oracle;
    both hold the same store and refs, and one condensation + collision step of both agrees."""
    print("case: padded store from the reference-written CLEO binary files, GPU against the oracle", flush=True)
    so = _lib()
    rc, (nd, halves) = read_gridfile(so, os.path.join(GOLDEN, "grid_uneven3d.dat"))
    assert rc == abi.CLEO_OK
    n = len(_expected()["initsupers3d_xi"])
    rc, sup, ninit = read_initsupers_padded(so, os.path.join(GOLDEN, "initsupers3d.dat"), n + 40)
    assert rc == abi.CLEO_OK and ninit == n
    ngbx = int(np.prod(nd))
    state = P.condensation(SEED, ndims=nd, mean=2).state
    zh, xh, yh = halves
    gi = np.arange(ngbx)
    k, i, j = gi % nd[0], (gi // nd[0]) % nd[1], gi // (nd[0] * nd[1])
    state["volume"][:] = (zh[k + 1] - zh[k]) * (xh[i + 1] - xh[i]) * (yh[j + 1] - yh[j])
    pop = P.Population(nd, halves, sup, state, P.NACL)
    g, o = make_domain(gpu, pop), make_domain(oracle, pop)
    assert_supers_match(g.download(), o.download(), exact_fp=True, what="padded upload")
    assert np.array_equal(g.refs(), o.refs())
    for d in (g, o):
        d.rng_counter(9)
        d.set_sdm(sdm_params(cond=cond_params(), coll=coll_params()))
        assert d.run_step(0, 100) == 0, d.error()
    assert_supers_match(g.download(), o.download(), what="padded step")
    assert_state_match(g.state(), o.state(), what="padded step")
