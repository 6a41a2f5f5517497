"""TEST INFRASTRUCTURE ONLY: loads the CPU oracle (oracle/liboracle_*.so) through the
same ctypes surface as the product library.  Imported only by tests/, the smoke check
and bench.py's cpu_baseline leg."""
import ctypes as C
import os
import subprocess

from cleo_amd.abi import Backend

HERE = os.path.dirname(os.path.abspath(__file__))


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(kind="serial"):
    path = os.path.join(HERE, f"liboracle_{kind}.so")
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    b = Backend(lib, "orc_")
    lib.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.orc_philox4x32_10.restype = None
    lib.orc_terminalv.argtypes = [C.c_int, C.c_double]
    lib.orc_terminalv.restype = C.c_double
    for n in ("orc_saturation_pressure", "orc_ventilation_factor"):
        getattr(lib, n).argtypes = [C.c_double]
        getattr(lib, n).restype = C.c_double
    lib.orc_diffusion_factor.argtypes = [C.c_double] * 3
    lib.orc_diffusion_factor.restype = C.c_double
    lib.orc_supersaturation_ratio.argtypes = [C.c_double] * 3
    lib.orc_supersaturation_ratio.restype = C.c_double
    lib.orc_solve_condensation.argtypes = ([C.c_double] * 9 + [C.c_uint64, C.c_double, C.c_uint64,
                                           C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                                           C.POINTER(C.c_int32)])
    lib.orc_solve_condensation.restype = C.c_double
    lib.orc_pair_prob.argtypes = [C.c_int] + [C.c_double] * 7
    lib.orc_pair_prob.restype = C.c_double
    from cleo_amd.abi import CollParams
    pp = C.POINTER(CollParams)
    lib.orc_cke.argtypes = [C.c_double] * 4
    lib.orc_cke.restype = C.c_double
    lib.orc_surf_energies.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
    lib.orc_surf_energies.restype = None
    lib.orc_lowlist_coaleff.argtypes = [C.c_int, C.c_double, C.c_double]
    lib.orc_lowlist_coaleff.restype = C.c_double
    lib.orc_nfrags.argtypes = [pp, C.c_double, C.c_double]
    lib.orc_nfrags.restype = C.c_double
    lib.orc_coalbure_flag.argtypes = [pp, C.c_double, C.c_double, C.c_double]
    lib.orc_coalbure_flag.restype = C.c_int32
    lib.orc_pair_prob_params.argtypes = [pp, C.c_double, C.c_double, C.c_double, C.c_double]
    lib.orc_pair_prob_params.restype = C.c_double
    lib.orc_constants.argtypes = [C.POINTER(C.c_double)]
    lib.orc_constants.restype = None
    lib.orc_num_threads.restype = C.c_int
    lib.orc_set_num_threads.argtypes = [C.c_int]
    lib.orc_set_num_threads.restype = None
    lib.orc_set_cond_guess.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_uint64]
    return b
