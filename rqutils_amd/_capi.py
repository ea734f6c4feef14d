"""ctypes binding of librqx.so (declared in include/rqx.h).

The library is loaded from this package directory.  There is no fallback: if
``librqx.so`` is missing, every entry point raises ``RqxLibraryError`` naming
the build command.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'librqx.so')

RQX_OK = 0
RQX_ERR_ARG = -1
RQX_ERR_HIP = -2
RQX_ERR_DTYPE = -3
RQX_ERR_LIMIT = -4

RQX_F64 = 0
RQX_C128 = 1
RQX_F32 = 2
RQX_C64 = 3
RQX_LOCG_WORK = 168
RQX_LOCG_SCALE = 164
RQX_LOCG_STATIC_PROBLEMS = 4096
RQX_FULL_MAX_QUBITS = 40
RQX_DIGEST_ARRAYS = 16

# every symbol declared in include/rqx.h
EXPORTS = (
    'rqx_version', 'rqx_last_error', 'rqx_plan_create', 'rqx_plan_destroy', 'rqx_plan_get_info',
    'rqx_lookup', 'rqx_diagonal', 'rqx_matvec', 'rqx_count', 'rqx_eigenpair_small',
    'rqx_eigenpair_small_host', 'rqx_sort_unique_u64', 'rqx_grad_coef', 'rqx_expect_terms',
    'rqx_plan_update_coef', 'rqx_plan_create_batch', 'rqx_plan_get_problem', 'rqx_plan_set_launch',
    'rqx_eigenpair_batch', 'rqx_batch_dots', 'rqx_batch_combine', 'rqx_plan_build_cache',
    'rqx_plan_drop_cache', 'rqx_plan_cache_info', 'rqx_plan_use_cache', 'rqx_plan_matvec_path', 'rqx_locg_step',
    'rqx_plan_hits', 'rqx_locg_step_static', 'rqx_full_apply', 'rqx_plan_create_device', 'rqx_pack_to_u64',
    'rqx_plan_digest',
)


class RqxLibraryError(RuntimeError):
    """librqx.so is missing or failed to load."""


class RqxError(RuntimeError):
    """A librqx call returned a HIP / limit error."""

    def __init__(self, code, msg):
        super().__init__(f'librqx error {code}: {msg}')
        self.code = code


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ('n_qubits', ctypes.c_int32),
        ('split_bits', ctypes.c_int32),
        ('n_states', ctypes.c_int64),
        ('n_total', ctypes.c_int64),
        ('n_groups', ctypes.c_int32),
        ('n_classes', ctypes.c_int32),
        ('n_terms', ctypes.c_int64),
        ('n_segments', ctypes.c_int32),
        ('n_tiles', ctypes.c_int32),
        ('max_segment', ctypes.c_int32),
        ('has_diagonal', ctypes.c_int32),
        ('is_real', ctypes.c_int32),
        ('max_group_terms', ctypes.c_int32),
        ('n_problems', ctypes.c_int32),
        ('persistent', ctypes.c_int32),
        ('blocks_per_cu', ctypes.c_int32),
        ('n_cu', ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ProblemInfo(ctypes.Structure):
    _fields_ = [
        ('row_offset', ctypes.c_int64),
        ('n_states', ctypes.c_int64),
        ('term_offset', ctypes.c_int64),
        ('n_terms', ctypes.c_int64),
        ('group_offset', ctypes.c_int32),
        ('n_groups', ctypes.c_int32),
        ('n_qubits', ctypes.c_int32),
        ('split_bits', ctypes.c_int32),
        ('n_segments', ctypes.c_int32),
        ('n_classes', ctypes.c_int32),
        ('n_tiles', ctypes.c_int32),
        ('max_segment', ctypes.c_int32),
        ('has_diagonal', ctypes.c_int32),
        ('is_real', ctypes.c_int32),
        ('cost', ctypes.c_double),
        ('section_bits', ctypes.c_int32),
        ('alpha_rows', ctypes.c_int32),
        ('list_mode', ctypes.c_int32),
        ('reserved', ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lock = threading.Lock()
_lib = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64

_SIGS = {
    'rqx_version': (_i32, []),
    'rqx_last_error': (ctypes.c_char_p, []),
    'rqx_plan_create': (_i32, [ctypes.POINTER(_vp), _i32, _vp, _vp, _i32, _vp, _vp, _i64, _vp, _i64,
                               _i64, _i32, _vp]),
    'rqx_plan_create_device': (_i32, [ctypes.POINTER(_vp), _i32, _vp, _vp, _i32, _vp, _vp, _i64, _vp, _i64,
                                      _i64, _i32, _vp]),
    'rqx_pack_to_u64': (_i32, [_vp, _i64, _i32, _i32, _vp, ctypes.POINTER(_i64), _vp]),
    'rqx_plan_digest': (_i32, [_vp, _vp, _vp]),
    'rqx_plan_destroy': (_i32, [_vp]),
    'rqx_plan_get_info': (_i32, [_vp, ctypes.POINTER(PlanInfo)]),
    'rqx_lookup': (_i32, [_vp, _vp, _vp, _vp]),
    'rqx_diagonal': (_i32, [_vp, _vp, _vp]),
    'rqx_matvec': (_i32, [_vp, _vp, _vp, _i32, _vp]),
    'rqx_count': (_i32, [_vp, _vp, _vp]),
    'rqx_eigenpair_small': (_i32, [_vp, _i32, _vp, _vp]),
    'rqx_eigenpair_small_host': (_i32, [_vp, _i32, _vp]),
    'rqx_grad_coef': (_i32, [_vp, _vp, _vp, _vp, _i32, _vp]),
    'rqx_expect_terms': (_i32, [_vp, _vp, _vp, _i32, _vp]),
    'rqx_plan_update_coef': (_i32, [_vp, _vp, _i32, _vp]),
    'rqx_plan_create_batch': (_i32, [ctypes.POINTER(_vp), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _i32, _vp]),
    'rqx_plan_get_problem': (_i32, [_vp, _i32, ctypes.POINTER(ProblemInfo)]),
    'rqx_plan_set_launch': (_i32, [_vp, _i32, _i32]),
    'rqx_plan_build_cache': (_i32, [_vp, _i32, _i64, _vp]),
    'rqx_plan_drop_cache': (_i32, [_vp, _i32]),
    'rqx_plan_cache_info': (_i32, [_vp, _i32, _vp]),
    'rqx_plan_use_cache': (_i32, [_vp, _i32]),
    'rqx_plan_matvec_path': (_i32, [_vp, _i32, _vp]),
    'rqx_locg_step': (_i32, [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp]),
    'rqx_locg_step_static': (_i32, [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, _vp]),
    'rqx_full_apply': (_i32, [_vp, _vp, _vp, _i64, _i32, _vp]),
    'rqx_plan_hits': (_i32, [_vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    'rqx_eigenpair_batch': (_i32, [_vp, _i32, _i32, _vp, _vp]),
    'rqx_batch_dots': (_i32, [_vp, _i32, _vp, _vp, _vp, _i32, _vp]),
    'rqx_batch_combine': (_i32, [_vp, _i32, _vp, _vp, _vp, _i32, _vp]),
    'rqx_sort_unique_u64': (_i32, [_vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
}


def lib():
    """The loaded library (raises RqxLibraryError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RqxLibraryError(
                    f'{LIB_PATH} is not built; run `python -m rqutils_amd._build` '
                    '(or __graft_entry__.build()) first')
            try:
                handle = ctypes.CDLL(LIB_PATH)
            except OSError as exc:
                raise RqxLibraryError(f'failed to load {LIB_PATH}: {exc}') from exc
            for name, (res, args) in _SIGS.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def last_error():
    msg = lib().rqx_last_error()
    return msg.decode() if msg else ''


def check(code):
    """Map a librqx status to a Python exception (ValueError for argument errors)."""
    if code == RQX_OK:
        return
    msg = last_error()
    if code in (RQX_ERR_ARG, RQX_ERR_DTYPE):
        raise ValueError(msg)
    raise RqxError(code, msg)


def ptr(obj):
    """Raw address of a torch tensor / numpy array / int / None."""
    if obj is None:
        return None
    if isinstance(obj, int):
        return obj
    if hasattr(obj, 'data_ptr'):
        return obj.data_ptr()
    if hasattr(obj, 'ctypes'):
        return obj.ctypes.data
    raise TypeError(f'cannot take the address of {type(obj)}')


def current_stream(device=None):
    import torch
    return torch.cuda.current_stream(device).cuda_stream
