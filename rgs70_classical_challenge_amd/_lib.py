"""ctypes binding of librgs_amd.so (the C ABI declared in include/rgs_amd.h).

The library is loaded from this package directory only; there is no fallback:
if it is missing every call raises RuntimeError naming the build command.
Status codes map to the exception classes the reference raises
(ValueError, ArithmeticError, MemoryError, RuntimeError).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'librgs_amd.so')

RGS_OK, RGS_EINVAL, RGS_EARITH, RGS_ENOMEM, RGS_EHIP = 0, 1, 2, 3, 4
INIT_LOAD, INIT_GRAPH, INIT_COV, INIT_DELTA = 0, 1, 2, 3

_ERRORS = {RGS_EINVAL: ValueError, RGS_EARITH: ArithmeticError, RGS_ENOMEM: MemoryError, RGS_EHIP: RuntimeError}

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_double = ctypes.c_double
c_float = ctypes.c_float
c_char_p = ctypes.c_char_p
P = ctypes.POINTER


class SlabDesc(ctypes.Structure):
    _fields_ = [
        ('state', c_void_p),
        ('nbits', c_int),
        ('init_kind', c_int),
        ('init_scale', c_float),
        ('init_lut', c_void_p),
        ('init_rows', c_void_p),
        ('init_cols', c_void_p),
        ('init_wlut', c_void_p),
        ('do_walsh', c_int),
        ('phase_lut', c_void_p),
        ('phase_rows', c_void_p),
        ('dphase_lut', c_void_p),
        ('rot', c_void_p),
        ('gate_limit', c_int),
        ('mass', c_void_p),
        ('mass_bits', c_int),
        ('low_bits', c_int),
        ('tile_bits', c_int),
        ('lane_bits', c_int),
        ('round_bits', c_int),
        ('phase_round_bits', c_int),
        ('dphase_global', c_int),
        ('init_wbits', c_int),
        ('group_order', c_int),
        ('top_bits', c_int),
        ('group_bits', c_int * 8),
        ('group_lanes', c_int * 8),
        ('group_rounds', c_int * 8),
        ('flip', ctypes.c_uint64),
        ('init_omask', ctypes.c_uint64),
        ('init_wpop', c_int),
        ('defer_leaves', c_int),
        ('fuse_walsh', c_int),
        ('dphase_host', c_void_p),
    ]


class SlabBatchItem(ctypes.Structure):
    _fields_ = [
        ('desc', P(SlabDesc)),
        ('uniforms', c_void_p),
        ('samples', c_int),
        ('out_index', c_void_p),
        ('out_p', c_void_p),
        ('flip', ctypes.c_uint64),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    'rgs_last_error': (c_char_p, []),
    'rgs_abi_version': (c_int, []),
    'rgs_gf2_inverse': (c_int, [c_void_p, c_int, c_void_p]),
    'rgs_gf2_rank': (c_int, [c_void_p, c_int, c_int, P(c_int)]),
    'rgs_graph_phase_lut': (c_int, [c_void_p, c_int, c_void_p]),
    'rgs_cov_luts': (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    'rgs_gate_shear': (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    'rgs_gate_scale': (c_double, [c_void_p, c_int]),
    'rgs_cov_weight_lut': (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_double, c_void_p]),
    'rgs_cov_weight_lut_bits': (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_double, c_int, c_void_p]),
    'rgs_cov_weight_table': (c_int, [c_void_p, c_void_p, c_int, c_double, c_void_p]),
    'rgs_slab_pipeline': (c_int, [P(SlabDesc), c_void_p]),
    'rgs_slab_pass': (c_int, [P(SlabDesc), c_int, c_void_p]),
    'rgs_slab_describe': (c_int, [P(SlabDesc), P(c_int), c_char_p, c_int]),
    'rgs_mass_pyramid': (c_int, [c_void_p, c_int, c_void_p]),
    'rgs_draw': (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, ctypes.c_uint64, c_void_p, c_void_p,
                         c_void_p]),
    'rgs_gather_prob': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    'rgs_slab_leaf_bits': (c_int, [P(SlabDesc), P(c_int)]),
    'rgs_slab_leaf_tiles': (c_int, [P(SlabDesc), c_void_p, c_int, c_void_p, c_void_p]),
    'rgs_slab_finish': (c_int, [P(SlabDesc), c_void_p]),
    'rgs_draw_leaves_workspace': (c_int, [P(SlabDesc), c_int, P(ctypes.c_size_t)]),
    'rgs_draw_leaves': (c_int, [P(SlabDesc), c_void_p, c_void_p, c_int, c_void_p, ctypes.c_size_t, c_void_p,
                                c_void_p, c_void_p]),
    'rgs_slab_batch_create': (c_int, [c_void_p, c_int, P(c_void_p)]),
    'rgs_slab_batch_run': (c_int, [c_void_p, c_void_p]),
    'rgs_slab_batch_info': (c_int, [c_void_p, P(c_int), P(c_int)]),
    'rgs_slab_batch_destroy': (c_int, [c_void_p]),
}

_lib = None


def lib():
    """Load the library once; raise RuntimeError if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'{LIB_PATH} missing: build it with '
                               '"python -m rgs70_classical_challenge_amd.build"')
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def register(name, restype, argtypes):
    """Declare another entry point (used by the scan and MADE modules)."""
    SIGNATURES[name] = (restype, argtypes)
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


def check(rc):
    if rc != RGS_OK:
        msg = lib().rgs_last_error().decode(errors='replace')
        raise _ERRORS.get(rc, RuntimeError)(msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))


def ptr(array):
    """Address of a NumPy array or a torch tensor (device or host)."""
    if array is None:
        return None
    if hasattr(array, 'data_ptr'):
        return array.data_ptr()
    return array.ctypes.data
