"""CLEO's binary initialisation files (initialise/readbinary.cpp:22-99, init_all_supers_from_binary.cpp:
57-100, gbx_bounds_from_binary.cpp:29-58) read by cleo_read_initsupers / cleo_read_gridfile.

The files are written here by `write_cleo_binary`, a restatement of the layout cleopy/writebinary.py
packs (little-endian, unpadded: 4 uint32, the global metadata string, per variable 3 uint32 + 2
chars + 1 double, then the data); the reference holds no binary files of its own.  The readers
are host code, so these tests need no GPU, except the last, which uploads what was read and
checks the store against a direct upload."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from cleo_amd import abi
from cleo_amd import populations as P
from tests.helpers import SEED, make_domain

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


@pytest.mark.gpu
def test_gpu_upload_from_binary_files(gpu, tmp_path):
    print("case: upload from CLEO binary files", flush=True)
    so = _lib()
    pop = P.coupled(SEED, ndims=(6, 5, 4), mean=30)
    ipath, gpath = str(tmp_path / "initsupers.dat"), str(tmp_path / "grid.dat")
    write_cleo_binary(ipath, _initsupers_arrays(pop))
    write_cleo_binary(gpath, _gridfile_arrays(pop.ndims, pop.halves))
    rc, sup = read_initsupers(so, ipath)
    assert rc == abi.CLEO_OK
    rc, (nd, halves) = read_gridfile(so, gpath)
    assert rc == abi.CLEO_OK
    from_files = P.Population(nd, halves, sup, pop.state, pop.solute, periodic=pop.periodic)
    a = make_domain(gpu, from_files).download()
    b = make_domain(gpu, pop).download()
    for k, _ in abi.SUPER_FIELDS:
        assert np.array_equal(a[k], b[k]), k
