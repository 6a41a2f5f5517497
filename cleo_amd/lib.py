"""Loads the product library libcleo_hip.so (built in-tree by cleo_amd/build.py)."""
import ctypes as C
import os

from .abi import Backend, Domain  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcleo_hip.so")
_backend = None


def load():
    """Backend over libcleo_hip.so; raises if the HIP library has not been built."""
    global _backend
    if _backend is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `python -m cleo_amd.build` (no CPU fallback exists)")
        # torch bundles its own HIP runtime, loaded under another file name than libamdhip64.so.7.
        # Imported first, it provides that soname and the library binds to it; imported after the
        # library, it would start a second runtime in the process, which finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        b = Backend(lib, "cleo_")
        lib.cleo_device_info.argtypes = [C.c_char_p, C.c_size_t]
        lib.cleo_device_info.restype = C.c_int
        lib.cleo_run_steps_graph.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
        lib.cleo_run_steps_graph.restype = C.c_int
        lib.cleo_synchronize.argtypes = [C.c_void_p]
        lib.cleo_synchronize.restype = C.c_int
        lib.cleo_record_guess.argtypes = [C.c_void_p, C.c_int32]
        lib.cleo_download_guess.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_uint64]
        fp = C.POINTER(C.c_float)
        lib.cleo_mlp_set_weights.argtypes = [C.c_void_p, fp]
        lib.cleo_mlp_get_weights.argtypes = [C.c_void_p, fp]
        lib.cleo_mlp_make_samples.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                              C.POINTER(C.c_uint64)]
        lib.cleo_mlp_train.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, fp]
        lib.cleo_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        lib.cleo_mlp_loss_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, fp, fp]
        lib.cleo_upload_supers_device.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        lib.cleo_event_record.argtypes = [C.c_void_p, C.c_int32]
        lib.cleo_event_elapsed.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_float)]
        lib.cleo_ndomain.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        lib.cleo_autotune.argtypes = [C.c_void_p, C.c_void_p]
        lib.cleo_time_kernel.argtypes = [C.c_void_p, C.c_int32, C.c_uint32, C.POINTER(C.c_float)]
        dp = C.POINTER(C.c_double)
        lib.cleo_thermo_eval.argtypes = [C.c_void_p, C.c_uint64, dp, dp, dp, dp, dp, dp]
        lib.cleo_thermo_eval.restype = C.c_int
        _backend = b
    return _backend


def device_info():
    b = load()
    buf = C.create_string_buffer(256)
    rc = b.lib.cleo_device_info(buf, 256)
    return buf.value.decode() if rc == 0 else f"unavailable (rc={rc})"
