"""ctypes binding of libfrhip.so.

Follows the reference's LibWrapper convention (``pyfr/ctypesutil.py:8-40``):
per-function argtypes/restype, and an errcheck that turns a nonzero status
into an exception carrying the library's message.  The library is looked up
next to this file, or at ``$PYFR_AMD_LIBRARY_PATH``.  There is no fallback:
if the library is missing or a GPU is absent, using it raises.
"""
import ctypes as ct
import os

_P, _I, _L, _D, _S = ct.c_void_p, ct.c_int, ct.c_long, ct.c_double, ct.c_size_t
_PI32 = ct.c_void_p


class FrHipError(RuntimeError):
    pass


# name -> argtypes (all functions return int status)
_FUNCS = {
    'frhip_init': [_I],
    'frhip_device_info': [ct.c_char_p, _S, ct.POINTER(_S), ct.POINTER(_S)],
    'frhip_malloc': [ct.POINTER(_P), _S],
    'frhip_free': [_P],
    'frhip_host_malloc': [ct.POINTER(_P), _S],
    'frhip_host_free': [_P],
    'frhip_memcpy_htod': [_P, _P, _S, _P],
    'frhip_memcpy_dtoh': [_P, _P, _S, _P],
    'frhip_fill_f64': [_P, _L, _D, _P],
    'frhip_memcpy_dtoh_async': [_P, _P, _S, _P],
    'frhip_memcpy_dtod': [_P, _P, _S, _P],
    'frhip_memset': [_P, _I, _S, _P],
    'frhip_memcpy2d': [_P, _S, _P, _S, _S, _S, _I, _P],
    'frhip_stream_create': [ct.POINTER(_P)],
    'frhip_stream_destroy': [_P],
    'frhip_stream_sync': [_P],
    'frhip_device_sync': [],
    'frhip_event_create': [ct.POINTER(_P)],
    'frhip_event_destroy': [_P],
    'frhip_event_record': [_P, _P],
    'frhip_stream_wait_event': [_P, _P],
    'frhip_event_sync': [_P],
    'frhip_event_query': [_P, ct.POINTER(_I)],
    'frhip_event_elapsed': [ct.POINTER(ct.c_float), _P, _P],
    'frhip_graph_begin': [_P],
    'frhip_graph_end': [_P, ct.POINTER(_P)],
    'frhip_graph_launch': [_P, _P],
    'frhip_graph_destroy': [_P],
    'frhip_mul_plan_create': [ct.POINTER(_P), _I, _I, _P, _I,
                              ct.POINTER(_I)],
    'frhip_mul_plan_destroy': [_P],
    'frhip_mul_plan_tune': [_P, _I, _P, _L, _I, ct.POINTER(_I)],
    'frhip_mul': [_P, _I, _P, _L, _P, _L, _D, _D, _P],
    'frhip_mul2_plan_create': [ct.POINTER(_P), _I, _I, _I, _P, _P, _I,
                               ct.POINTER(_I)],
    'frhip_mul2_plan_destroy': [_P],
    'frhip_mul2_plan_tune': [_P, _I, _P, _L, _P, _L, _I, ct.POINTER(_I)],
    'frhip_mul_plan_set_variant': [_P, _I],
    'frhip_mul2_plan_set_variant': [_P, _I],
    'frhip_mul2': [_P, _I, _P, _L, _P, _L, _P, _L, _I, _P, _L, _P, _D, _D,
                   _I, _P],
    'frhip_tdivf_plan_create': [ct.POINTER(_P), _I, _I, _I, _P, _P, _I],
    'frhip_tdivf_plan_destroy': [_P],
    'frhip_tdivf_plan_tune': [_P, _I, _P, _L, _P, _L, _P, _L, _D, _I,
                              ct.POINTER(_I)],
    'frhip_tdivf_plan_set_variant': [_P, _I],
    'frhip_tdivf_flux_euler': [_P, _I, _P, _L, _P, _L, _P, _L, _P, _L, _I,
                               _P, _L, _P, _D, _D, _I, _D, _P],
    'frhip_tflux_euler': [_I, _I, _I, _I, _P, _L, _P, _L, _P, _L, _P, _D,
                          _P, _P],
    'frhip_tflux_ns': [_I, _I, _I, _I, _P, _L, _P, _L, _P, _L, _P, _L, _P,
                       _P, _L, _P, _P, _P, _P],
    'frhip_gradcoru': [_I, _I, _I, _I, _P, _L, _P, _L, _P, _P, _L, _P],
    'frhip_tflux_ns_grad': [_I, _I, _I, _I, _P, _L, _P, _L, _P, _L, _P, _P,
                            _P, _P],
    'frhip_negdivconf': [_I, _I, _I, _P, _L, _P, _L, _P],
    'frhip_intcflux_euler': [_I, _I, _I, _P, _PI32, _PI32, _P, _L, _D, _D,
                             _P, _P],
    'frhip_intconu': [_I, _I, _P, _PI32, _PI32, _P, _PI32, _PI32, _D, _P],
    'frhip_intcflux_ns': [_I, _I, _I, _P, _PI32, _PI32, _P, _PI32, _PI32,
                          _L, _PI32, _PI32, _P, _L, _P, _D, _P, _P, _P],
    'frhip_bccflux_euler': [_I, _I, _I, _I, _P, _PI32, _P, _L, _P, _L, _P,
                            _D, _P],
    'frhip_bcconu': [_I, _I, _I, _P, _PI32, _P, _PI32, _P, _L, _P, _L, _P,
                     _P],
    'frhip_bccflux_ns': [_I, _I, _I, _I, _P, _PI32, _P, _PI32, _L, _PI32,
                         _P, _L, _P, _L, _P, _D, _P],
    'frhip_bc_program_create': [ct.POINTER(_P), _I, _P, _P, _P, _I],
    'frhip_bc_program_destroy': [_P],
    'frhip_bc_program_set_time': [_P, _D, _P],
    'frhip_bc_program_set_const': [_P, _I, _D, _P],
    'frhip_bc_eval': [_P, _I, _I, _P, _L, _P, _L, _P],
    'frhip_axnpby': [_I, _I, _L, _L, _I, _P, _P, _P],
    'frhip_axnpby_pool': [_I, _I, _L, _L, _I, _P, _P, ct.c_uint, _P, _I, _P],
    'frhip_rkvdh2': [_I, _I, _L, _L, _P, _P, _D, _D, _I, _P],
    'frhip_rkvdh2_mixed': [_I, _L, _L, _P, _P, _D, _D, _I, _P],
    'frhip_convert': [_I, _P, _I, _P, _L, _P],
    'frhip_expr_program_create': [ct.POINTER(_P), _I, _I, _P, _P, _P, _I],
    'frhip_expr_program_destroy': [_P],
    'frhip_reduction': [_P, _I, _I, _I, _I, _I, _I, _P, _P, ct.c_uint, _P,
                        _I, _P, _I, _P, _L, _P, _P],
    'frhip_reduction_workspace': [_P, _I, _I, _I, ct.POINTER(_L)],
    'frhip_reduction_noalloc': [_P, _I, _I, _I, _I, _I, _I, _P, _P,
                                ct.c_uint, _P, _I, _P, _I, _P, _P],
    'frhip_axnpby_scaled': [_I, _I, _L, _L, _I, _I, _P, _P, ct.c_uint, _P,
                            _P, _P],
    'frhip_rk4_mixed': [_I, _L, _L, _I, _P, _P, _P, _P, _D, _P],
    'frhip_rkvdh2_errest': [_I, _I, _L, _L, _P, _P, _P, _P, _D, _D, _D, _D,
                            _I, _I, _P],
}


def lib_path():
    env = os.environ.get('PYFR_AMD_LIBRARY_PATH')
    if env:
        return env
    return os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        'libfrhip.so')


class FrHip:
    """Thin wrapper exposing each C-ABI function as a method."""

    def __init__(self, path=None):
        path = path or lib_path()
        if not os.path.exists(path):
            raise FrHipError(f'libfrhip.so not found at {path}; build it '
                             'with python -m pyfr_amd.build')
        self.path = path
        self.lib = lib = ct.CDLL(path)
        lib.frhip_last_error.restype = ct.c_char_p
        lib.frhip_last_error.argtypes = []
        for name, argtypes in _FUNCS.items():
            fn = getattr(lib, name)
            fn.restype = ct.c_int
            fn.argtypes = argtypes
            fn.errcheck = self._errcheck
            setattr(self, name.removeprefix('frhip_'), fn)

    def _errcheck(self, status, fn, args):
        if status != 0:
            msg = self.lib.frhip_last_error().decode(errors='replace')
            raise FrHipError(f'{fn.__name__}: {msg} (status {status})')
        return status

    def exported_symbols(self):
        return ['frhip_last_error', *_FUNCS]


_lib = None


def get_lib():
    global _lib
    if _lib is None:
        _lib = FrHip()
    return _lib
