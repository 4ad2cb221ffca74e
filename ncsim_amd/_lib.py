"""ctypes binding of libncsv.so (include/ncsv.h).  Loading fails loudly when
the library has not been built: there is no fallback path."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libncsv.so")

c_dp = ctypes.POINTER(ctypes.c_double)
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_u8p = ctypes.POINTER(ctypes.c_uint8)
UNIFORM_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p)


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double),
                ("bytes", ctypes.c_double)]


# (name, argtypes); every function returns an int status unless noted
_SIGS = {
    "ncsv_create": [ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
    "ncsv_destroy": [ctypes.c_void_p],
    "ncsv_clone": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)],
    "ncsv_num_qubits": [ctypes.c_void_p, c_u32p],
    "ncsv_get_precision": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)],
    "ncsv_set_basis": [ctypes.c_void_p, ctypes.c_uint64],
    "ncsv_set_plus": [ctypes.c_void_p],
    "ncsv_set_amplitudes": [ctypes.c_void_p, c_dp, ctypes.c_uint64],
    "ncsv_to_vector": [ctypes.c_void_p, c_dp, ctypes.c_uint64],
    "ncsv_amplitude": [ctypes.c_void_p, ctypes.c_uint64, c_dp],
    "ncsv_apply_gate": [ctypes.c_void_p, ctypes.c_int, c_i64p, c_dp, c_u32p, ctypes.c_uint32],
    "ncsv_apply_matrix": [ctypes.c_void_p, c_dp, c_u32p, ctypes.c_uint32],
    "ncsv_apply_monomial": [ctypes.c_void_p, c_u32p, ctypes.c_uint32, c_u32p, c_dp],
    "ncsv_apply_pauli": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32],
    "ncsv_norm_sq": [ctypes.c_void_p, c_dp],
    "ncsv_expectation": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, c_dp],
    "ncsv_expectation_batch": [ctypes.c_void_p, ctypes.c_uint32, c_u64p, c_u64p, c_u32p, c_dp],
    "ncsv_expectation_sum": [ctypes.c_void_p, ctypes.c_uint32, c_u64p, c_u64p, c_u32p, c_dp, c_dp],
    "ncsv_inner": [ctypes.c_void_p, ctypes.c_void_p, c_dp],
    "ncsv_compare": [ctypes.c_void_p, ctypes.c_void_p, c_dp],
    "ncsv_probability": [ctypes.c_void_p, ctypes.c_uint64, c_dp],
    "ncsv_project": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, c_dp],
    "ncsv_measure": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, UNIFORM_FN,
                     ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)],
    "ncsv_reset": [ctypes.c_void_p, ctypes.c_uint32, UNIFORM_FN, ctypes.c_void_p],
    "ncsv_scale": [ctypes.c_void_p, c_dp],
    "ncsv_normalize": [ctypes.c_void_p],
    "ncsv_sample": [ctypes.c_void_p, ctypes.c_uint64, c_dp, c_u64p],
    "ncsv_sample_noalloc": [ctypes.c_void_p, ctypes.c_uint64, c_dp, c_u64p],
    "ncsv_sample_rng": [ctypes.c_void_p, ctypes.c_uint64, UNIFORM_FN, ctypes.c_void_p, c_u64p],
    "ncsv_tensor": [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)],
    "ncsv_circuit_new": [ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)],
    "ncsv_circuit_parse": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)],
    "ncsv_circuit_destroy": [ctypes.c_void_p],
    "ncsv_circuit_num_qubits": [ctypes.c_void_p, c_u32p],
    "ncsv_circuit_num_measurements": [ctypes.c_void_p, c_u64p],
    "ncsv_circuit_num_gates": [ctypes.c_void_p, c_u64p],
    "ncsv_circuit_push_gate": [ctypes.c_void_p, ctypes.c_int, c_i64p, c_dp, c_u32p, ctypes.c_uint32],
    "ncsv_circuit_push_measure": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32],
    "ncsv_circuit_push_reset": [ctypes.c_void_p, ctypes.c_uint32],
    "ncsv_circuit_push_noise": [ctypes.c_void_p, ctypes.c_int, c_dp, c_u32p, ctypes.c_uint32],
    "ncsv_circuit_push_tick": [ctypes.c_void_p],
    "ncsv_run": [ctypes.c_void_p, ctypes.c_void_p, UNIFORM_FN, ctypes.c_void_p, c_u8p, ctypes.c_uint64, c_u64p],
    "ncsv_set_fusion": [ctypes.c_void_p, ctypes.c_int],
    "ncsv_profile_enable": [ctypes.c_void_p, ctypes.c_int],
    "ncsv_profile_read": [ctypes.c_void_p, ctypes.POINTER(KernelStat), ctypes.c_uint32, c_u32p],
    "ncsv_profile_reset": [ctypes.c_void_p],
    "ncsv_sweep_count": [ctypes.c_void_p, c_u64p],
    "ncsv_synchronize": [ctypes.c_void_p],
    "ncsv_plan_emulate": [ctypes.c_void_p, ctypes.c_int, c_dp, c_u64p],
    "ncsv_jit_check": [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, c_u32p],
    "ncsv_jit_check_expect": [ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32, c_u64p, c_u64p, c_u32p,
                              ctypes.c_char_p, c_u32p],
    "ncsv_batch_add": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, c_u64p, c_u64p, c_u32p, c_u32p],
    "ncsv_batch_draws": [ctypes.c_void_p, ctypes.c_uint32, c_u32p],
    "ncsv_batch_set_tile_bits": [ctypes.c_void_p, ctypes.c_uint32],
    "ncsv_batch_set_blocks": [ctypes.c_void_p, ctypes.c_int32],
    "ncsv_batch_set_shape": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32],
    "ncsv_batch_run": [ctypes.c_void_p, c_dp, c_u64p, c_u8p],
    "ncsv_batch_emulate": [ctypes.c_void_p, c_dp, c_u64p, c_u8p],
    "ncsv_batch_simulate": [ctypes.c_void_p, c_dp, c_u64p, c_u8p],
    "ncsv_batch_result": [ctypes.c_void_p, ctypes.c_uint32, c_dp, c_dp, c_u8p],
    "ncsv_batch_stats": [ctypes.c_void_p, c_dp],
    "ncsv_batch_plan": [ctypes.c_void_p],
    "ncsv_device_arch": [ctypes.c_char_p, ctypes.c_size_t],
    "ncsv_frames_new": [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)],
    "ncsv_frames_destroy": [ctypes.c_void_p],
    "ncsv_frames_shape": [ctypes.c_void_p, c_u64p, c_u64p, c_u64p],
    "ncsv_frames_reference": [ctypes.c_void_p, c_u8p, ctypes.c_uint64],
    "ncsv_frames_sample_packed": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, c_u64p],
    "ncsv_frames_sample_detectors": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, c_u64p, c_u64p],
    "ncsv_frames_emulate_packed": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, c_u64p],
}

_lib = None


def lib():
    """The loaded library; raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m ncsim_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.ncsv_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.ncsv_last_error.restype = ctypes.c_size_t
        L.ncsv_device_count.argtypes = []
        L.ncsv_device_count.restype = ctypes.c_int
        L.ncsv_circuit_to_text.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.ncsv_circuit_to_text.restype = ctypes.c_size_t
        L.ncsv_batch_new.argtypes = [ctypes.c_int]
        L.ncsv_batch_new.restype = ctypes.c_void_p
        L.ncsv_batch_destroy.argtypes = [ctypes.c_void_p]
        L.ncsv_batch_destroy.restype = None
        L.ncsv_batch_size.argtypes = [ctypes.c_void_p]
        L.ncsv_batch_size.restype = ctypes.c_uint32
        _lib = L
    return _lib


def exported_symbols():
    """Names declared by include/ncsv.h (for the export check)."""
    return sorted(list(_SIGS) + ["ncsv_last_error", "ncsv_device_count", "ncsv_circuit_to_text", "ncsv_batch_new",
                                 "ncsv_batch_destroy", "ncsv_batch_size"])
