"""cleo_sort_supers_noalloc, the entry point that allocates nothing, against the sort that
cleo_upload_supers_device applies (the existing entry point it is a variant of): the same drops in
the same order and the same gridbox refs, bit for bit, on two calls in a row, the second with fewer
drops; sizes beyond its fixed scratch are rejected."""
import ctypes as C

import numpy as np
import pytest

from cleo_amd import abi
from cleo_amd import populations as P
from tests.helpers import SEED

pytestmark = pytest.mark.gpu
# before the library: one HIP runtime for both (cleo_amd/lib.py)
torch = pytest.importorskip("torch")

FIELDS = [("sdgbxindex", np.uint32), ("coord3", np.float64), ("coord1", np.float64), ("coord2", np.float64),
          ("xi", np.uint64), ("radius", np.float64), ("msol", np.float64), ("sdid", np.uint64)]


def _shuffled(seed, ndims, mean):
    """A coupled population in random order with some keys out of the domain."""
    pop = P.coupled(seed, ndims=ndims, mean=mean)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(pop.n)
    sup = {k: np.ascontiguousarray(v[perm]) for k, v in pop.supers.items()}
    sup["sdgbxindex"][rng.random(pop.n) < 0.03] = abi.OOB
    sup["sdgbxindex"][rng.random(pop.n) < 0.02] = pop.ngbx + 5
    pop.supers = sup
    return pop


def _device(torch, sup):
    t = {k: torch.from_numpy(np.ascontiguousarray(sup[k]).view({np.uint32: np.int32, np.uint64: np.int64}.get(dt, dt)))
         .to("cuda") for k, dt in FIELDS}
    return t


def _soa(t):
    s = abi.SupersSoA()
    ptype = dict(abi.SupersSoA._fields_)
    for k, _ in FIELDS:
        setattr(s, k, C.cast(C.c_void_p(t[k].data_ptr()), ptype[k]))
    return s


def _host(t):
    out = {}
    for k, dt in FIELDS:
        a = t[k].cpu().numpy()
        out[k] = a.view(dt) if a.dtype != dt else a
    return out


def _by_context(gpu, torch, pop):
    """The existing entry point: cleo_upload_supers_device sorts device drops into a context."""
    lib = gpu.lib
    d = abi.Domain(gpu, pop.n, pop.ndims, *pop.halves, periodic=pop.periodic)
    t = _device(torch, pop.supers)
    torch.cuda.synchronize()
    s = _soa(t)
    assert lib.cleo_upload_supers_device(d.h, C.byref(s), pop.n) == 0
    d.ntot = pop.n
    out = d.download()
    refs = d.refs()
    d.close()
    return out, refs


def _no_alloc(gpu, torch, pop):
    lib = gpu.lib
    lib.cleo_sort_supers_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]
    lib.cleo_sort_supers_noalloc.restype = C.c_int
    tin = _device(torch, pop.supers)
    tout = {k: torch.empty_like(v) for k, v in tin.items()}
    refs = torch.full((pop.ngbx + 1,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    sin, sout = _soa(tin), _soa(tout)
    rc = lib.cleo_sort_supers_noalloc(C.byref(sin), C.byref(sout), C.c_void_p(refs.data_ptr()), pop.n, pop.ngbx, None)
    torch.cuda.synchronize()
    assert rc == 0
    return _host(tout), refs.cpu().numpy().view(np.uint32)


def test_no_alloc_sort_matches_upload_sort(gpu):
    for seed, ndims, mean in ((SEED, (12, 10, 8), 40), (SEED + 1, (5, 4, 3), 9)):  # the second call is smaller
        print(f"case: no-alloc sort, ndims={ndims} mean={mean}", flush=True)
        pop = _shuffled(seed, ndims, mean)
        want, want_refs = _by_context(gpu, torch, pop)
        got, got_refs = _no_alloc(gpu, torch, pop)
        for k, _ in FIELDS:
            assert np.array_equal(got[k], want[k]), k
        assert np.array_equal(got_refs, want_refs)
        assert int(got_refs[-1]) == want["ndomain"]


def test_no_alloc_sort_rejects_sizes_beyond_its_scratch(gpu):
    lib = gpu.lib
    lib.cleo_sort_supers_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]
    s = abi.SupersSoA()
    refs = (C.c_uint32 * 2)()
    # rejected on the host before any pointer is read or kernel launched
    assert lib.cleo_sort_supers_noalloc(C.byref(s), C.byref(s), refs, (1 << 20) + 1, 1, None) == abi.CLEO_E_ARG
    assert lib.cleo_sort_supers_noalloc(C.byref(s), C.byref(s), refs, 10, 1, None) == abi.CLEO_E_ARG  # null arrays
