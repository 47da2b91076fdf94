"""ctypes binding of libwt.so (the C ABI declared in include/wt.h).

torch is imported before the library is loaded so that libwt binds to the HIP
runtime torch already carries (both are named by the SONAME libamdhip64.so.7).
There is no fallback: if the library is missing and cannot be built, or the
device is absent, every call raises.
"""
import ctypes
import os
import threading

import torch  # noqa: F401  (must precede loading libwt, see above)

HERE = os.path.dirname(os.path.abspath(__file__))
# WT_LIB_PATH selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get('WT_LIB_PATH') or os.path.join(HERE, 'libwt.so')

WT_FP32, WT_FP16, WT_BF16, WT_FP64 = 0, 1, 2, 3
FORMS = {'fp32': WT_FP32, 'fp16': WT_FP16, 'bf16': WT_BF16, 'fp64': WT_FP64}
FORM_NAMES = {v: k for k, v in FORMS.items()}
WT_FLAG_FORCES, WT_FLAG_MASS, WT_FLAG_NO_MAC = 1, 2, 4
# wt_batch_info.tile_order bits
WT_BATCH_BACKWARDS, WT_BATCH_PERSISTENT = 2, 4

c_i32, c_i64, c_u32, c_f32, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p


class Grid(ctypes.Structure):
    """wt_grid"""
    _fields_ = [('nx', c_i32), ('ny', c_i32), ('nz', c_i32), ('form', c_i32), ('cpt', c_i32), ('tile_order', c_i32),
                ('f', c_vp * 2), ('cls', c_vp), ('mac', c_vp), ('shell', c_vp * 2),
                ('partials', c_vp), ('sums', c_vp)]


class Params(ctypes.Structure):
    """wt_params"""
    _fields_ = [('uin', c_f32), ('tau', c_f32), ('smag', c_f32)]


class SeedBox(ctypes.Structure):
    """wt_seed_box"""
    _fields_ = [('x', c_f32), ('cy', c_f32), ('cz', c_f32), ('hy', c_f32), ('hz', c_f32)]


class BatchInfo(ctypes.Structure):
    """wt_batch_info"""
    _fields_ = [('total_tiles', c_i64), ('max_chunks', c_i32), ('form', c_i32), ('ntunnels', c_i32),
                ('tile_order', c_i32), ('max_cpt', c_i32), ('persist_blocks', c_i32)]


P = ctypes.POINTER
# name -> (restype, argtypes); mirrors include/wt.h
SIGNATURES = {
    'wt_last_error': (ctypes.c_char_p, []),
    'wt_version': (c_i32, []),
    'wt_device_info': (c_i32, [ctypes.c_char_p, c_i32, P(c_i64)]),
    'wt_buffer_sizes': (c_i32, [c_i32, c_i32, c_i32, c_i32, P(c_i64)]),
    'wt_state_bytes': (c_i64, [c_i32, c_i32, c_i32, c_i32]),
    'wt_tile_count': (c_i64, [c_i32, c_i32, c_i32]),
    'wt_cache_policy': (c_i32, [c_i32, c_i32, c_i32, c_i32]),
    'wt_init': (c_i32, [P(Grid), c_f32, c_vp]),
    'wt_set_solid': (c_i32, [P(Grid), c_vp, c_vp]),
    'wt_step': (c_i32, [P(Grid), c_i32, c_i32, P(Params), c_i32, c_vp]),
    'wt_step_noalloc': (c_i32, [P(Grid), c_i32, c_i32, P(Params), c_i32, c_vp]),
    'wt_step_advect': (c_i32, [P(Grid), c_i32, P(Params), c_i32, c_vp, c_vp, c_vp, c_vp, c_i64, c_f32, c_f32, c_f32,
                               c_u32, P(SeedBox), c_vp]),
    'wt_forces': (c_i32, [P(Grid), c_i32, c_vp]),
    'wt_read_sums': (c_i32, [P(Grid), P(ctypes.c_double), c_vp]),
    'wt_batch_workspace_bytes': (c_i64, [c_i32]),
    'wt_batch_gate_offset': (c_i64, [c_i32]),
    'wt_batch_prepare': (c_i32, [P(Grid), P(Params), P(c_i32), c_i32, c_vp, c_vp, P(BatchInfo), c_vp]),
    'wt_batch_encode': (c_i32, [P(Grid), P(Params), P(c_i32), c_i32, c_vp, c_vp, P(BatchInfo)]),
    'wt_batch_step': (c_i32, [c_vp, P(BatchInfo), c_i32, c_i32, c_vp]),
    'wt_advect': (c_i32, [P(Grid), c_vp, c_vp, c_i64, c_f32, c_f32, c_f32, c_u32, P(SeedBox), c_vp]),
    'wt_advect_ordered': (c_i32, [P(Grid), c_vp, c_vp, c_vp, c_i64, c_f32, c_f32, c_f32, c_u32, P(SeedBox), c_vp]),
    'wt_particle_order_workspace_bytes': (c_i64, [c_i64]),
    'wt_particle_order': (c_i32, [P(Grid), c_vp, c_i64, c_vp, c_vp, c_i64, c_vp]),
    'wt_pack_particles': (c_i32, [P(Grid), c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    'wt_advect_slots': (c_i32, [P(Grid), c_vp, c_vp, c_vp, c_i64, c_f32, c_f32, c_f32, c_u32, P(SeedBox), c_vp]),
    'wt_particle_gather': (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    'wt_pack_particles_slots': (c_i32, [P(Grid), c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    'wt_slice': (c_i32, [P(Grid), c_vp, c_i32, c_i32, c_i32, c_f32, c_f32, c_f32, c_f32, c_vp]),
    'wt_cp': (c_i32, [P(Grid), c_vp, c_vp, c_vp, c_i64, c_f32, c_f32, c_vp]),
    'wt_link_classes': (c_i32, [P(Grid), c_vp, c_vp]),
    'wt_selftest_bf16': (c_i32, [c_vp, c_vp]),
    'wt_stream_copy': (c_i32, [c_vp, c_vp, c_i64, c_i32, c_vp]),
    'wt_voxelize': (c_i32, [c_vp, c_i64, c_i32, c_i32, c_i32, c_i32, c_vp, P(c_i64)]),
    'wt_ellipsoids': (c_i32, [c_vp, c_i32, c_i32, c_i32, c_i32, c_i32, c_vp, P(c_i64)]),
}

_lock = threading.Lock()
_lib = None


def load(build_if_missing=True):
    """Load libwt.so, building it first only if it does not exist."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            if not build_if_missing:
                raise RuntimeError(f'{LIB_PATH} is not built')
            from . import build as _build
            _build.build()
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error():
    return load().wt_last_error().decode(errors='replace')


def check(rc, what=''):
    """Raise RuntimeError(wt_last_error()) on a negative status."""
    if rc < 0:
        raise RuntimeError(f'{what}: {last_error()}' if what else last_error())
    return rc


def stream_handle(stream=None):
    """hipStream_t of a torch stream (the current stream by default)."""
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def require_gpu():
    """The solver runs on the GPU only; fail loudly without one."""
    if not torch.cuda.is_available():
        raise RuntimeError('wind-tunnel solver needs a HIP GPU (torch.cuda.is_available() is False)')
