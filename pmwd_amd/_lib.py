"""ctypes binding of libpmx.so (the C ABI declared in include/pmx.h).

torch is imported first so that the HIP runtime torch ships is the one the
library binds to (same soname), and device pointers/streams are shared.
There is no fallback: if the library is missing or cannot run, the call
raises.
"""
import atexit
import contextlib
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libpmx.so')

_lib = None
_torch_pool = False  # contexts made by lib() take memory from torch's caching allocator
_lock = threading.Lock()

c_int, c_ll, c_f, c_d, c_p = ctypes.c_int, ctypes.c_longlong, ctypes.c_float, ctypes.c_double, ctypes.c_void_p


class PmxError(RuntimeError):
    """Error returned by the pmx C ABI; ``rc`` is its pmx_status code."""

    def __init__(self, msg, rc=None):
        super().__init__(msg)
        self.rc = rc


ENOMEM = 2  # PMX_ENOMEM


class CicGeom(ctypes.Structure):
    """pmx_cic_geom (include/pmx.h)."""
    _fields_ = [('mesh_shape', c_int * 3), ('cell_size', c_d), ('offset', c_d * 3),
                ('out_cell_size', c_d), ('out_shape', c_int * 3)]


# name -> argtypes (all return int status unless listed in _RESTYPE)
# pmx_alloc_fn / pmx_free_fn (include/pmx.h)
_ALLOC_T = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                            ctypes.c_void_p)
_FREE_T = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                           ctypes.c_void_p)


@_ALLOC_T
def _torch_alloc(nbytes, device, stream, user):
    # a block of torch's caching allocator for the context's stream; NULL
    # (PMX_ENOMEM in the library) when torch has no room
    try:
        import torch
        return torch.cuda.caching_allocator_alloc(int(nbytes), int(device), int(stream or 0))
    except Exception:
        return None


@_FREE_T
def _torch_free(ptr, nbytes, device, stream, user):
    try:
        import torch
        torch.cuda.caching_allocator_delete(int(ptr))
    except Exception:
        pass


_SIGS = {
    'pmx_version': [],
    'pmx_last_error': [],
    'pmx_ctx_create': [c_int, c_int, c_int, c_d, c_p, ctypes.POINTER(c_p)],
    'pmx_ctx_destroy': [c_p],
    'pmx_ctx_set_stream': [c_p, c_p],
    'pmx_ctx_bytes': [c_p],
    'pmx_ctx_trim': [c_p, c_int],
    'pmx_set_allocator': [_ALLOC_T, _FREE_T, c_p],
    'pmx_cell_keys': [c_p, c_p, c_p, c_ll, c_p],
    'pmx_sort_perm': [c_p, c_p, c_ll, c_p],
    'pmx_tile_keys': [c_p, c_p, c_p, c_ll, c_p],
    'pmx_tile_shape': [c_p, ctypes.POINTER(c_int * 3)],
    'pmx_scatter_mass': [c_p, c_p, c_p, c_ll, c_f, c_p, c_int],
    'pmx_scatter_mass_noalloc': [c_p, c_p, c_ll, c_int, c_int, c_int, c_d, c_f, c_p, c_int, c_p],
    'pmx_scatter_cic': [c_p, ctypes.POINTER(CicGeom), c_p, c_p, c_ll, c_p, c_f, c_int, c_p],
    'pmx_gather_cic': [c_p, ctypes.POINTER(CicGeom), c_p, c_p, c_ll, c_p, c_int, c_p, c_f, c_p],
    'pmx_scatter_cic_adj': [c_p, ctypes.POINTER(CicGeom), c_p, c_p, c_ll, c_p, c_p, c_f, c_int,
                            c_p, c_p],
    'pmx_gather_cic_adj': [c_p, ctypes.POINTER(CicGeom), c_p, c_p, c_ll, c_p, c_p, c_int, c_p, c_p],
    'pmx_enmesh': [c_p, ctypes.POINTER(CicGeom), c_int, c_int, c_p, c_p, c_ll, c_p, c_p, c_p],
    'pmx_gather_cic3': [c_p, c_p, c_p, c_ll, c_p, c_p],
    'pmx_poisson_grad_adj': [c_p, c_p, c_f, c_p],
    'pmx_white_noise': [c_p, ctypes.c_ulonglong, c_ll, c_p],
    'pmx_random_bits': [c_p, ctypes.c_ulonglong, c_ll, c_p],
    'pmx_powspec': [c_p, c_p, c_p, c_int, c_p, c_int, c_int, c_p],
    'pmx_rfftn': [c_p, c_p, c_p],
    'pmx_irfftn': [c_p, c_p, c_p],
    'pmx_poisson_grad': [c_p, c_p, c_f, c_p],
    'pmx_gravity': [c_p, c_p, c_p, c_ll, c_f, c_f, c_p],
    'pmx_kick': [c_p, c_ll, c_p, c_p, c_f],
    'pmx_drift': [c_p, c_ll, c_p, c_p, c_f],
    'pmx_nbody_step': [c_p, c_p, c_p, c_p, c_p, c_ll, c_f, c_f, c_f, c_f, c_f],
    'pmx_nbody_steps': [c_p, c_p, c_p, c_p, c_p, c_ll, c_f, c_f, c_int, c_p, c_p, c_p],
    'pmx_nbody_steps_dev': [c_p, c_p, c_p, c_p, c_p, c_ll, c_f, c_int, c_p],
    'pmx_graph_capture': [c_p, c_p, c_p, c_p, c_p, c_ll, c_f, c_int, c_p, c_p],
    'pmx_graph_launch': [c_p],
    'pmx_lpt': [c_p, c_p, c_f, c_int, c_p, c_p, c_p],
    'pmx_linear_modes': [c_p, c_p, c_f, c_p, c_p, c_int, c_p],
    'pmx_pack_create': [c_p, c_int, c_p],
    'pmx_pack_launch': [c_p, c_p],
    'pmx_pack_destroy': [c_p],
    'pmx_stream_create': [c_int, ctypes.POINTER(c_p)],
    'pmx_stream_destroy': [c_p],
    'pmx_gravity_vjp': [c_p, c_p, c_p, c_ll, c_f, c_f, c_p, c_p, c_p, ctypes.POINTER(c_d)],
    'pmx_kick_adj': [c_p, c_ll, c_p, c_p, c_p, c_p, c_p, c_f, ctypes.POINTER(c_d)],
    'pmx_drift_adj': [c_p, c_ll, c_p, c_p, c_p, c_p, c_f, ctypes.POINTER(c_d)],
    'pmx_nbody_step_adj': [c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_ll, c_f, c_f, c_f, c_f, c_f,
                           c_p, ctypes.POINTER(c_d)],
    'pmx_nbody_steps_adj': [c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_ll, c_f, c_f, c_f, c_int,
                            c_p, c_p, c_p, c_p, ctypes.POINTER(c_d)],
    'pmx_dot': [c_p, c_p, c_p, c_ll, ctypes.POINTER(c_d)],
    'pmx_growth_ode': [c_d, c_d, c_d, c_d, c_p, c_int, c_p, c_d, c_d, c_d, c_int, c_p, c_p],
    'pmx_tophat_var': [c_p, c_p, c_int, c_p, c_int, c_p],
    'pmx_timing_enable': [c_p, c_int],
    'pmx_timing_read': [c_p, ctypes.POINTER(c_d), ctypes.POINTER(c_ll)],
}

TIMER_CLASSES = ('bin', 'scan', 'fill', 'deposit', 'fold', 'z_r2c', 'y_fwd', 'poisson_x',
                 'y_inv', 'z_c2r', 'gather', 'vjp_gather', 'grad_gather', 'kd_adj', 'other')


def timing_read(ctx):
    """{class: (ms, launches)} since the last read (synchronises the stream)."""
    n = len(TIMER_CLASSES)
    ms = (c_d * n)()
    cnt = (c_ll * n)()
    check(lib().pmx_timing_read(ctx.h, ms, cnt))
    return {k: (ms[i], cnt[i]) for i, k in enumerate(TIMER_CLASSES)}
_RESTYPE = {'pmx_last_error': ctypes.c_char_p, 'pmx_ctx_bytes': c_ll}


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise PmxError(f'{LIB_PATH} not found: build it with '
                               '`python -m pmwd_amd.csrc.build` (no CPU fallback exists)')
            L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
            for name, args in _SIGS.items():
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = _RESTYPE.get(name, c_int)
            # contexts take their device memory from torch's caching
            # allocator, so that what one gives back serves the other without
            # new device mappings (PMX_TORCH_ALLOCATOR=0: hipMalloc)
            global _torch_pool
            if os.environ.get('PMX_TORCH_ALLOCATOR', '1') != '0':
                check_rc = L.pmx_set_allocator(_torch_alloc, _torch_free, None)
                if check_rc != 0:
                    raise PmxError(f'pmx_set_allocator failed ({check_rc})', check_rc)
                _torch_pool = True
            _lib = L
    return _lib


def torch_pool():
    """Whether contexts (made after the library loaded) draw from torch's
    caching allocator, so that torch's cached free blocks serve them too."""
    lib()
    return _torch_pool


def exported_symbols():
    return list(_SIGS)


def check(rc):
    if rc != 0:
        msg = lib().pmx_last_error().decode(errors='replace')
        raise PmxError(f'pmx error {rc}: {msg}', rc)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def require_device(t, name='tensor'):
    if not t.is_cuda:
        raise PmxError(f'{name} must be on a GPU device (got {t.device}); '
                       'the pmx kernels have no CPU path')
    return t


TRIM_PIPELINE, TRIM_ALL = 1, 2   # pmx.h PMX_TRIM_*

# Graph captures and the calls that wait for the whole device or queue on
# the null stream (context creation and teardown) exclude each other: such a
# call from another thread while a stream is capturing can invalidate the
# capture.
capture_lock = threading.RLock()


_free_streams = {}          # device index -> [(handle, ExternalStream)] not in use
_streams_lock = threading.Lock()


class PrivateStream:
    """A stream of this process's own (pmx_stream_create), wrapped for torch.
    torch.cuda.Stream() hands out streams from a pool of 32 per device that
    other threads may be using at the same time; a stream that is captured
    must not be shared.  Closing hands the stream back for reuse once its
    queued work has ended; streams are never destroyed, since torch's
    caching allocator keeps the blocks used on a stream tied to it for the
    life of the process."""

    def __init__(self, device):
        self.device = torch.device(device)
        self.index = (self.device.index if self.device.index is not None
                      else torch.cuda.current_device())
        with _streams_lock:
            pool = _free_streams.setdefault(self.index, [])
            item = pool.pop() if pool else None
        if item is None:
            h = c_p()
            check(lib().pmx_stream_create(self.index, ctypes.byref(h)))
            item = (h, torch.cuda.ExternalStream(h.value, device=self.device))
        self._item = item
        self.stream = item[1]

    def close(self):
        if self._item is not None:
            self._item[1].synchronize()
            with _streams_lock:
                _free_streams[self.index].append(self._item)
            self._item = None


class Context:
    """Owner of a pmx_ctx for one (mesh shape, cell size, device)."""

    def __init__(self, mesh_shape, cell_size, device):
        self.mesh_shape = tuple(int(s) for s in mesh_shape)
        self.cell_size = float(cell_size)
        self.device = torch.device(device)
        h = c_p()
        with torch.cuda.device(self.device), capture_lock:
            stream = torch.cuda.current_stream(self.device).cuda_stream
            check(lib().pmx_ctx_create(*self.mesh_shape, self.cell_size, c_p(stream),
                                       ctypes.byref(h)))
        self.h = h

    def bind(self):
        """Point the context at the current torch stream; returns the handle."""
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().pmx_ctx_set_stream(self.h, c_p(stream)))
        return self.h

    @property
    def nbytes(self):
        return int(lib().pmx_ctx_bytes(self.h))

    def trim(self, pipeline_only=False):
        """Free particle-sized scratch (reallocated by the next call that
        needs it): only the multi-step particle copy, or all of it."""
        if self.h:
            check(lib().pmx_ctx_trim(self.h, TRIM_PIPELINE if pipeline_only else TRIM_ALL))

    def close(self):
        if self.h:
            # at interpreter exit (after _close_at_exit) the device memory is
            # left to the process end: the allocator callbacks a destroy
            # would call may already be gone
            if _exiting is False:
                with capture_lock:
                    lib().pmx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx_cache = {}
_tls = threading.local()


@contextlib.contextmanager
def context_scope(provider):
    """Within the scope, ``context()`` on this thread returns
    ``provider(key, make)`` instead of the process-wide cache: ``key`` is
    (mesh shape, cell size, device index) and ``make()`` builds a new Context.
    Concurrent users of one mesh shape (one per stream) each get their own
    scratch this way."""
    prev = getattr(_tls, 'provider', None)
    _tls.provider = provider
    try:
        yield
    finally:
        _tls.provider = prev


def context(mesh_shape, cell_size, device):
    """Cached Context for the mesh of a configuration on a device."""
    device = torch.device(device)
    if device.index is None:
        device = torch.device(device.type, torch.cuda.current_device())
    key = (tuple(int(s) for s in mesh_shape), float(cell_size), device.index)
    provider = getattr(_tls, 'provider', None)
    if provider is not None:
        return provider(key, lambda: Context(mesh_shape, cell_size, device))
    ctx = _ctx_cache.get(key)
    if ctx is None:
        ctx = Context(mesh_shape, cell_size, device)
        _ctx_cache[key] = ctx
    return ctx


def context_any_cell(mesh_shape, device):
    """A context of this mesh shape for work that does not depend on the cell
    size (the FFTs): an existing one of any cell size when there is one."""
    device = torch.device(device)
    if device.index is None:
        device = torch.device(device.type, torch.cuda.current_device())
    shape = tuple(int(s) for s in mesh_shape)
    if getattr(_tls, 'provider', None) is None:
        for (s, _cell, idx), ctx in _ctx_cache.items():
            if s == shape and idx == device.index:
                return ctx
    return context(shape, 1.0, device)


def trim_contexts():
    """Free the particle-sized scratch of every cached context."""
    for c in _ctx_cache.values():
        c.trim()


def clear_contexts():
    for c in _ctx_cache.values():
        c.close()
    _ctx_cache.clear()


_exiting = False


@atexit.register
def _close_at_exit():
    """Destroys the cached contexts while torch and the allocator callbacks
    are alive (atexit handlers run before module teardown, and this one,
    registered after torch's, before torch's)."""
    global _exiting
    try:
        if _lib is not None:
            clear_contexts()
    finally:
        _exiting = True
