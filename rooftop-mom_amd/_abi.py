"""ctypes binding of librooftop.so (include/rooftop.h). Loads the in-tree library and fails loudly without it."""

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, 'librooftop.so')

RT_OK, RT_EINVAL, RT_ESIZE, RT_ENOMEM, RT_ESINGULAR, RT_EHIP = 0, -1, -2, -3, -4, -5
RT_PRODUCT_FFT, RT_PRODUCT_SPLIT_FFT, RT_PRODUCT_LEAN_SPLIT_FFT = 0, 1, 2
RT_C128, RT_C64, RT_MIXED, RT_C64_FP64_BASIS = 0, 1, 2, 3
RT_MODIFIER_NONE, RT_MODIFIER_REAL, RT_MODIFIER_IMAG = 0, 1, 2
RT_PRECOND_JACOBI, RT_PRECOND_NEARFIELD = 1, 2
RT_METHOD_CG, RT_METHOD_GMRES, RT_METHOD_BLOCK_CG = 0, 1, 2


class rt_solve_opts(C.Structure):
    _fields_ = [('method', C.c_int32), ('tolerance', C.c_double), ('max_iterations', C.c_int32),
                ('restart', C.c_int32), ('precision', C.c_int32), ('refine', C.c_int32),
                ('inner_tolerance', C.c_double), ('max_outer', C.c_int32)]


class rt_output(C.Structure):
    _fields_ = [('converged', C.c_int32), ('iteration_count', C.c_int32), ('residual_norm', C.c_double),
                ('time_s', C.c_double), ('outer_count', C.c_int32), ('residual_norms_count', C.c_int32),
                ('residual_norms_capacity', C.c_int32), ('residual_norms', C.POINTER(C.c_double))]


class rt_request(C.Structure):
    _fields_ = [('count_x', C.c_int32), ('count_y', C.c_int32), ('size_x', C.c_double), ('size_y', C.c_double),
                ('wavenumber', C.c_double), ('near_order', C.c_int32), ('far_order', C.c_int32),
                ('mask', C.c_void_p), ('rhs', C.c_void_p), ('n', C.c_int64), ('preconditioner', C.c_int32),
                ('opts', rt_solve_opts)]


# rt_apply_fn: apply(user, input, output, n, nrhs) on host arrays (rt_product_create_callback).
APPLY_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64)

# Every exported function: (name, restype, argtypes). The CPU test suite checks this list against include/rooftop.h.
P = C.c_void_p
SIGNATURES = [
    ('rt_ctx_create', C.c_int, [C.c_int32, C.c_uint64, C.POINTER(P)]),
    ('rt_ctx_destroy', None, [P]),
    ('rt_ctx_memory', C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ('rt_ctx_memory_peak', C.c_int, [P, C.c_int32, C.POINTER(C.c_uint64)]),
    ('rt_ctx_synchronize', C.c_int, [P]),
    ('rt_ctx_stream', C.c_int, [P, C.POINTER(P)]),
    ('rt_ctx_set_timing', C.c_int, [P, C.c_int32]),
    ('rt_ctx_kernel_timing', C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ('rt_bandwidth_probe', C.c_int, [P, C.c_uint64, C.c_int32, C.POINTER(C.c_double)]),
    ('rt_problem_create', C.c_int, [P, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32,
                                    C.c_int32, P, C.POINTER(P)]),
    ('rt_problem_destroy', None, [P]),
    ('rt_problem_count', C.c_int, [P, C.POINTER(C.c_int64)]),
    ('rt_problem_generators', C.c_int, [P, P, P, P]),
    ('rt_problem_set_generators', C.c_int, [P, P, P, P]),
    ('rt_product_create', C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(P)]),
    ('rt_precond_create', C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(P)]),
    ('rt_product_create_dense', C.c_int, [P, C.c_int64, P, C.c_int32, C.POINTER(P)]),
    ('rt_product_create_callback', C.c_int, [P, C.c_int64, APPLY_FN, P, C.POINTER(P)]),
    ('rt_product_destroy', None, [P]),
    ('rt_product_apply', C.c_int, [P, P, P, C.c_int64, C.c_int64, C.c_int32]),
    ('rt_product_apply_device', C.c_int, [P, P, P, C.c_int64, C.c_int64, C.c_int32]),
    ('rt_product_apply_device_noalloc', C.c_int, [P, P, P, C.c_int64, C.c_int64, C.c_int32]),
    ('rt_product_memory', C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ('rt_product_spectra', C.c_int, [P, C.POINTER(C.c_int32), P]),
    ('rt_product_estimate_memory', C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint64)]),
    ('rt_product_profile', C.c_int, [P, C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    ('rt_product_tune', C.c_int, [P, C.c_int64, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    ('rt_product_set_shape', C.c_int, [P, C.c_int64, C.c_int32, C.POINTER(C.c_int32)]),
    ('rt_server_create', C.c_int, [C.c_int32, C.c_uint64, C.c_int32, C.POINTER(P)]),
    ('rt_server_destroy', None, [P]),
    ('rt_server_submit', C.c_int, [P, C.POINTER(rt_request), C.POINTER(C.c_int64)]),
    ('rt_server_wait', C.c_int, [P, C.c_int64, P, C.c_int64, C.POINTER(rt_output), C.POINTER(C.c_double)]),
    ('rt_server_stats', C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]),
    ('rt_server_set_batching', C.c_int, [P, C.c_int32]),
    ('rt_server_batch_stats', C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ('rt_request_bytes', C.c_int, [C.POINTER(rt_request), C.POINTER(C.c_uint64)]),
    ('rt_blob_mask', C.c_int, [C.c_int32, C.c_int32, C.c_uint64, C.c_double, P]),
    ('rt_solve', C.c_int, [P, P, P, C.POINTER(rt_solve_opts), P, C.c_int64, C.c_int64, P, P,
                           C.POINTER(rt_output)]),
    ('rt_solve_device', C.c_int, [P, P, P, C.POINTER(rt_solve_opts), P, C.c_int64, C.c_int64, P, P,
                                  C.POINTER(rt_output)]),
    ('rt_device_alloc', C.c_int, [P, C.c_uint64, C.POINTER(P)]),
    ('rt_device_free', C.c_int, [P, P]),
    ('rt_copy_to_device', C.c_int, [P, P, P, C.c_uint64]),
    ('rt_copy_to_host', C.c_int, [P, P, P, C.c_uint64]),
    ('rt_precond_create_circulant', C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                              C.POINTER(P)]),
    ('rt_plane_waves_device', C.c_int, [P, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), P]),
    ('rt_radar_cross_sections_device', C.c_int, [P, C.c_int64, P, P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ('rt_input_impedances_device', C.c_int, [P, C.c_int64, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_double),
                                             P, C.POINTER(C.c_double)]),
    ('rt_last_error_id', C.c_char_p, []),
    ('rt_last_error_message', C.c_char_p, []),
    ('rt_version', C.c_char_p, []),
]

_lib = None


class RooftopError(RuntimeError):
    """An error of librooftop carrying the reference's identifier (e.g. 'rooftopMoM:product:sizeMismatch')."""

    def __init__(self, status, identifier, message):
        super().__init__('%s: %s' % (identifier, message))
        self.status = status
        self.identifier = identifier


def library():
    """The loaded librooftop.so; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError('librooftop.so is not built at %s; run rooftop-mom_amd/build.py '
                               '(there is no CPU fallback)' % LIBRARY)
        lib = C.CDLL(LIBRARY)
        for name, res, args in SIGNATURES:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(status):
    if status != RT_OK:
        lib = library()
        raise RooftopError(status, lib.rt_last_error_id().decode(), lib.rt_last_error_message().decode())
