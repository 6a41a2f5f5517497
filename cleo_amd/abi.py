"""ctypes mirror of include/cleo_hip.h.

`Backend(lib, prefix)` binds the C entry points of one library: the product
`libcleo_hip.so` (prefix "cleo_") or any library exporting the same signatures
under another prefix.  Struct layouts here must match include/cleo_hip.h.
"""
import ctypes as C

import numpy as np

OOB = 0xFFFFFFFF
ctypes_byref = C.byref

CLEO_OK = 0
CLEO_E_ARG = -1
CLEO_E_HIP = -2
CLEO_E_DEVICE = -3
CLEO_E_STATE = -4
CLEO_E_NOMEM = -5
CLEO_E_FILE = -6

ERR_NAMES = {
    0: "NONE", 1: "XI_ZERO", 2: "RADIUS_BELOW_DRY", 3: "NEG_CONDENSATE", 4: "NEWTON_MAXITERS",
    5: "TEMP_NONPOSITIVE", 6: "COAL_UNDEFINED", 7: "BREAKUP_NFRAGS", 8: "CFL", 9: "BOUNDS",
    10: "SORT_NONLOCAL", 11: "SUBSTEP", 12: "INJECTED_INDEX", 13: "XI_OVERFLOW",
}
ERR_XI_OVERFLOW = 13

PROB_GOLOVIN, PROB_LONG, PROB_CONST, PROB_LOWLIST_COAL, PROB_LOWLIST_BU = 0, 1, 2, 3, 4
NFRAGS_CONST, NFRAGS_CKE = 0, 1
FLAG_CONST, FLAG_SU, FLAG_TS, FLAG_STRAUB = 0, 1, 2, 3
ENACT_COAL, ENACT_BREAKUP, ENACT_COALBURE = 0, 1, 2
TV_NULL, TV_SIMMEL, TV_ROGERS_YAU, TV_ROGERS_GK = 0, 1, 2, 3

MLP_IN = 6
MLP_HID = 64
MLP_NPARAMS = 64 * 6 + 64 + 64 * 64 + 64 + 64 + 1
MLP_NBLOB = MLP_NPARAMS + 6 + 6 + 2

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


class DomainDesc(C.Structure):
    _fields_ = [("ndims", C.c_uint32 * 3), ("zhalf", _dp), ("xhalf", _dp), ("yhalf", _dp),
                ("periodic", C.c_int32 * 3)]


class SupersSoA(C.Structure):
    _fields_ = [("sdgbxindex", _u32p), ("coord3", _dp), ("coord1", _dp), ("coord2", _dp),
                ("xi", _u64p), ("radius", _dp), ("msol", _dp), ("sdid", _u64p)]


class GbxState(C.Structure):
    _fields_ = [(n, _dp) for n in ("volume", "press", "temp", "qvap", "qcond", "wvel_lo", "wvel_hi",
                                   "uvel_lo", "uvel_hi", "vvel_lo", "vvel_hi")]


class Solute(C.Structure):
    _fields_ = [("rho_sol", C.c_double), ("mr_sol", C.c_double), ("ionic", C.c_double)]


class CollParams(C.Structure):
    _fields_ = [("interval", C.c_uint32), ("prob_kind", C.c_int32), ("prob_param", C.c_double),
                ("enact_kind", C.c_int32), ("nfrags", C.c_double), ("coalbure_coaleff", C.c_double),
                ("nfrags_kind", C.c_int32), ("flag_kind", C.c_int32), ("prob_tv", C.c_int32),
                ("nfrags_tv", C.c_int32), ("flag_tv", C.c_int32)]


class CondParams(C.Structure):
    _fields_ = [("interval", C.c_uint32), ("do_alter_thermo", C.c_int32), ("maxniters", C.c_uint32),
                ("rtol", C.c_double), ("atol", C.c_double), ("minsubdelt_s", C.c_double),
                ("warm_start", C.c_int32), ("warm_restart_iters", C.c_uint32)]


class MotionParams(C.Structure):
    _fields_ = [("interval", C.c_uint32), ("terminalv", C.c_int32)]


class SDMParams(C.Structure):
    _fields_ = [("enable_cond", C.c_int32), ("enable_colls", C.c_int32), ("enable_motion", C.c_int32),
                ("cond", CondParams), ("coll", CollParams), ("motion", MotionParams)]


PROCESS_CONDENSATION, PROCESS_COLLISIONS = 0, 1
MAX_PROCESSES = 8


class Process(C.Structure):
    _fields_ = [("kind", C.c_int32), ("cond", CondParams), ("coll", CollParams)]


def cond_process(cp):
    return Process(PROCESS_CONDENSATION, cp, CollParams())


def coll_process(kp):
    return Process(PROCESS_COLLISIONS, CondParams(), kp)


class Error(C.Structure):
    _fields_ = [("code", C.c_int32), ("gbxindex", C.c_uint32), ("drop", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("cond_drops", "newton_iters", "substep_drops", "warm_restarts",
                                          "coll_pairs", "coll_events", "nulls", "sorts", "motion_drops")]


class Tiles(C.Structure):
    _fields_ = [("cond_waves_per_block", C.c_uint32), ("coll_waves_per_block", C.c_uint32),
                ("motion_waves_per_block", C.c_uint32), ("cond_flat", C.c_uint32), ("motion_flat", C.c_uint32),
                ("cond_grid_waves", C.c_uint32), ("coll_grid_waves", C.c_uint32), ("motion_grid_waves", C.c_uint32),
                ("coll_flat", C.c_uint32), ("cond_defer", C.c_uint32)]


class MLPTrainParams(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64)]


SUPER_FIELDS = [("sdgbxindex", np.uint32), ("coord3", np.float64), ("coord1", np.float64),
                ("coord2", np.float64), ("xi", np.uint64), ("radius", np.float64), ("msol", np.float64),
                ("sdid", np.uint64)]
STATE_FIELDS = [f for f, _ in GbxState._fields_]


_CT = {np.uint32: C.c_uint32, np.float64: C.c_double, np.uint64: C.c_uint64}


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def soa_view(arrs):
    """SupersSoA pointing at a dict of contiguous numpy arrays (kept alive by the caller)."""
    s = SupersSoA()
    for name, dt in SUPER_FIELDS:
        a = arrs[name]
        assert a.dtype == dt and a.flags.c_contiguous, name
        setattr(s, name, a.ctypes.data_as(C.POINTER(_CT[dt])))
    return s


def state_view(arrs):
    s = GbxState()
    for name in STATE_FIELDS:
        a = arrs[name]
        assert a.dtype == np.float64 and a.flags.c_contiguous, name
        setattr(s, name, _ptr(a, C.c_double))
    return s


def empty_supers(n):
    return {name: np.zeros(n, dtype=dt) for name, dt in SUPER_FIELDS}


def empty_state(ngbx):
    return {name: np.zeros(ngbx, dtype=np.float64) for name in STATE_FIELDS}


class DynamicsStep(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("press", "temp", "qvap", "qcond", "wvel_zfaces", "uvel_xfaces",
                                           "vvel_yfaces")]


_SIGS = {
    "create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint64, C.POINTER(DomainDesc)]),
    "destroy": (C.c_int, [C.c_void_p]),
    "last_error_message": (C.c_char_p, [C.c_void_p]),
    "upload_supers": (C.c_int, [C.c_void_p, C.POINTER(SupersSoA), C.c_uint64]),
    "download_supers": (C.c_int, [C.c_void_p, C.POINTER(SupersSoA), _u64p, _u64p]),
    "download_refs": (C.c_int, [C.c_void_p, _u32p]),
    "upload_gbx_state": (C.c_int, [C.c_void_p, C.POINTER(GbxState)]),
    "download_gbx_state": (C.c_int, [C.c_void_p, C.POINTER(GbxState)]),
    "set_solute": (C.c_int, [C.c_void_p, C.POINTER(Solute)]),
    "set_rng_counter": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "set_rng_injected": (C.c_int, [C.c_void_p, _u32p, C.c_uint64, _dp, C.c_uint64]),
    "set_sdm": (C.c_int, [C.c_void_p, C.POINTER(SDMParams)]),
    "set_sdm_chain": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "sort": (C.c_int, [C.c_void_p]),
    "motion": (C.c_int, [C.c_void_p, C.POINTER(MotionParams)]),
    "condensation": (C.c_int, [C.c_void_p, C.POINTER(CondParams)]),
    "collisions": (C.c_int, [C.c_void_p, C.POINTER(CollParams)]),
    "run_step": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "get_error": (C.c_int, [C.c_void_p, C.POINTER(Error)]),
    "clear_error": (C.c_int, [C.c_void_p]),
    "get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "reset_stats": (C.c_int, [C.c_void_p]),
    "record_niters": (C.c_int, [C.c_void_p, C.c_int32]),
    "download_niters": (C.c_int, [C.c_void_p, _u32p, C.c_uint64]),
    "set_tiles": (C.c_int, [C.c_void_p, C.c_void_p]),
    "get_tiles": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mass_moments": (C.c_int, [C.c_void_p, C.c_int32, _u64p, _fp, _fp]),
    "receive_dynamics": (C.c_int, [C.c_void_p, C.POINTER(DynamicsStep)]),
}


class Backend:
    """Binds the common entry points of a library under `prefix`."""

    def __init__(self, lib, prefix):
        self.lib = lib
        self.prefix = prefix
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, prefix + name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)


class Domain:
    """Host-side handle for one context of a Backend (product or oracle)."""

    def __init__(self, backend, maxn, ndims, zhalf, xhalf, yhalf, periodic=(0, 1, 1)):
        self.b = backend
        self.ndims = tuple(int(v) for v in ndims)
        self.ngbx = self.ndims[0] * self.ndims[1] * self.ndims[2]
        self._halves = [np.ascontiguousarray(h, dtype=np.float64) for h in (zhalf, xhalf, yhalf)]
        d = DomainDesc()
        d.ndims[:] = self.ndims
        d.zhalf, d.xhalf, d.yhalf = (_ptr(h, C.c_double) for h in self._halves)
        d.periodic[:] = periodic
        self.h = C.c_void_p()
        self.check(self.b.create(C.byref(self.h), int(maxn), C.byref(d)), "create")
        self.ntot = 0

    def check(self, rc, what):
        if rc != CLEO_OK:
            msg = b""
            if self.h and hasattr(self.b, "last_error_message"):
                msg = self.b.last_error_message(self.h) or b""
            err = self.error() if self.h else None
            raise RuntimeError(f"{self.b.prefix}{what} failed rc={rc} {msg.decode()} err={err}")

    def close(self):
        if self.h:
            self.b.destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, sup):
        n = len(sup["xi"])
        arrs = {k: np.ascontiguousarray(sup[k], dtype=dt) for k, dt in SUPER_FIELDS}
        s = soa_view(arrs)
        self.check(self.b.upload_supers(self.h, C.byref(s), n), "upload_supers")
        self.ntot = n

    def download(self):
        arrs = empty_supers(self.ntot)
        s = soa_view(arrs)
        ntot = C.c_uint64()
        ndom = C.c_uint64()
        self.check(self.b.download_supers(self.h, C.byref(s), C.byref(ntot), C.byref(ndom)), "download")
        arrs["ndomain"] = int(ndom.value)
        return arrs

    def refs(self):
        r = np.zeros(self.ngbx + 1, dtype=np.uint32)
        self.check(self.b.download_refs(self.h, _ptr(r, C.c_uint32)), "download_refs")
        return r

    def mass_moments(self, rain=False):
        """(mom0, mom1, mom2) per gridbox: sum of xi, and float sums of xi m and xi m^2."""
        m0 = np.zeros(self.ngbx, dtype=np.uint64)
        m1 = np.zeros(self.ngbx, dtype=np.float32)
        m2 = np.zeros(self.ngbx, dtype=np.float32)
        self.check(self.b.mass_moments(self.h, int(bool(rain)), _ptr(m0, C.c_uint64), _ptr(m1, C.c_float),
                                       _ptr(m2, C.c_float)), "mass_moments")
        return m0, m1, m2

    def set_state(self, st):
        arrs = {k: np.ascontiguousarray(st[k], dtype=np.float64) for k in STATE_FIELDS}
        s = state_view(arrs)
        self.check(self.b.upload_gbx_state(self.h, C.byref(s)), "upload_gbx_state")

    def state(self):
        arrs = empty_state(self.ngbx)
        s = state_view(arrs)
        self.check(self.b.download_gbx_state(self.h, C.byref(s)), "download_gbx_state")
        return arrs

    def set_solute(self, rho_sol, mr_sol, ionic):
        s = Solute(rho_sol, mr_sol, ionic)
        self.check(self.b.set_solute(self.h, C.byref(s)), "set_solute")

    def rng_counter(self, seed, first_event=0):
        self.check(self.b.set_rng_counter(self.h, int(seed), int(first_event)), "set_rng_counter")

    def rng_injected(self, js, phis):
        js = np.ascontiguousarray(js, dtype=np.uint32)
        phis = np.ascontiguousarray(phis, dtype=np.float64)
        self.check(self.b.set_rng_injected(self.h, _ptr(js, C.c_uint32), len(js), _ptr(phis, C.c_double),
                                           len(phis)), "set_rng_injected")

    def set_sdm(self, p):
        self.check(self.b.set_sdm(self.h, C.byref(p)), "set_sdm")

    def set_sdm_chain(self, procs, motion=None):
        """microphysics procs[0] >> procs[1] >> ... (a list of Process), motion a MotionParams or None"""
        arr = (Process * max(1, len(procs)))(*procs)
        m = C.byref(motion) if motion is not None else None
        return self.b.set_sdm_chain(self.h, C.cast(arr, C.c_void_p), len(procs), m)

    def sort(self):
        self.check(self.b.sort(self.h), "sort")

    def motion(self, p):
        return self.b.motion(self.h, C.byref(p))

    def condensation(self, p):
        return self.b.condensation(self.h, C.byref(p))

    def collisions(self, p):
        return self.b.collisions(self.h, C.byref(p))

    def run_step(self, t0, t1):
        return self.b.run_step(self.h, int(t0), int(t1))

    def error(self):
        e = Error()
        self.b.get_error(self.h, C.byref(e))
        return (int(e.code), int(e.gbxindex), int(e.drop))

    def clear_error(self):
        self.b.clear_error(self.h)

    def stats(self):
        s = Stats()
        self.check(self.b.get_stats(self.h, C.byref(s)), "get_stats")
        return {f: int(getattr(s, f)) for f, _ in Stats._fields_}

    def reset_stats(self):
        self.b.reset_stats(self.h)

    def record_niters(self, en=True):
        self.b.record_niters(self.h, 1 if en else 0)

    def niters(self):
        out = np.zeros(self.ntot, dtype=np.uint32)
        self.check(self.b.download_niters(self.h, _ptr(out, C.c_uint32), self.ntot), "download_niters")
        return out
