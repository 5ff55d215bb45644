"""ctypes binding of libbellman (the bellman-cuda C ABI on MI355X).

This is the host-side mirror of the reference's C interface
(src/bellman-cuda.h): the same entry points, structs passed by value and
bc_error return codes.  Everything below the ABI is HIP; this module only
marshals arguments.  It raises if libbellman.so has not been built — there
is no CPU fallback.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# BELLMAN_AMD_LIB points the binding at another build of the library (e.g. a variant under test).
LIB_PATH = os.environ.get("BELLMAN_AMD_LIB", os.path.join(PKG_DIR, "libbellman.so"))

bc_success = 0
bc_error_invalid_value = 1
bc_error_memory_allocation = 2
bc_error_not_ready = 600


class BellmanError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__("%s returned bc_error %d" % (fn, code))
        self.code = code


class bc_stream(ctypes.Structure):
    _fields_ = [("handle", ctypes.c_void_p)]


class bc_event(ctypes.Structure):
    _fields_ = [("handle", ctypes.c_void_p)]


class bc_mem_pool(ctypes.Structure):
    _fields_ = [("handle", ctypes.c_void_p)]


bc_host_fn = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


class ff_grand_product_configuration(ctypes.Structure):
    _fields_ = [("mem_pool", bc_mem_pool), ("stream", bc_stream), ("inputs", ctypes.c_void_p), ("outputs", ctypes.c_void_p), ("count", ctypes.c_uint)]


class ff_multiply_by_powers_configuration(ctypes.Structure):
    _fields_ = [
        ("mem_pool", bc_mem_pool),
        ("stream", bc_stream),
        ("inputs", ctypes.c_void_p),
        ("base", ctypes.c_void_p),
        ("outputs", ctypes.c_void_p),
        ("count", ctypes.c_uint),
    ]


class ff_inverse_configuration(ctypes.Structure):
    _fields_ = [("mem_pool", bc_mem_pool), ("stream", bc_stream), ("inputs", ctypes.c_void_p), ("outputs", ctypes.c_void_p), ("count", ctypes.c_uint)]


class ff_poly_evaluate_configuration(ctypes.Structure):
    _fields_ = [
        ("mem_pool", bc_mem_pool),
        ("stream", bc_stream),
        ("values", ctypes.c_void_p),
        ("point", ctypes.c_void_p),
        ("result", ctypes.c_void_p),
        ("count", ctypes.c_uint),
    ]


class ff_sort_u32_configuration(ctypes.Structure):
    _fields_ = [
        ("mem_pool", bc_mem_pool),
        ("stream", bc_stream),
        ("values", ctypes.c_void_p),
        ("sorted_values", ctypes.c_void_p),
        ("count", ctypes.c_uint),
    ]


class generate_permutation_polynomials_configuration(ctypes.Structure):
    _fields_ = [
        ("mem_pool", bc_mem_pool),
        ("stream", bc_stream),
        ("indexes", ctypes.c_void_p),
        ("scalars", ctypes.c_void_p),
        ("target", ctypes.c_void_p),
        ("columns_count", ctypes.c_uint),
        ("log_rows_count", ctypes.c_uint),
    ]


class msm_configuration(ctypes.Structure):
    _fields_ = [
        ("mem_pool", bc_mem_pool),
        ("stream", bc_stream),
        ("bases", ctypes.c_void_p),
        ("scalars", ctypes.c_void_p),
        ("results", ctypes.c_void_p),
        ("log_scalars_count", ctypes.c_uint),
        ("h2d_copy_finished", bc_event),
        ("h2d_copy_finished_callback", bc_host_fn),
        ("h2d_copy_finished_callback_data", ctypes.c_void_p),
        ("d2h_copy_finished", bc_event),
        ("d2h_copy_finished_callback", bc_host_fn),
        ("d2h_copy_finished_callback_data", ctypes.c_void_p),
    ]


class ntt_configuration(ctypes.Structure):
    _fields_ = [
        ("mem_pool", bc_mem_pool),
        ("stream", bc_stream),
        ("inputs", ctypes.c_void_p),
        ("outputs", ctypes.c_void_p),
        ("log_values_count", ctypes.c_uint),
        ("bit_reversed_inputs", ctypes.c_bool),
        ("inverse", ctypes.c_bool),
        ("can_overwrite_inputs", ctypes.c_bool),
        ("log_extension_degree", ctypes.c_uint),
        ("coset_index", ctypes.c_uint),
        ("h2d_copy_finished", bc_event),
        ("h2d_copy_finished_callback", bc_host_fn),
        ("h2d_copy_finished_callback_data", ctypes.c_void_p),
        ("d2h_copy_finished", bc_event),
        ("d2h_copy_finished_callback", bc_host_fn),
        ("d2h_copy_finished_callback_data", ctypes.c_void_p),
    ]


# ---------------------------------------------------------------- extensions (include/bellman-amd.h)
class bc_amd_msm_job(ctypes.Structure):
    _fields_ = [
        ("bases", ctypes.c_void_p),
        ("scalars", ctypes.c_void_p),
        ("results", ctypes.c_void_p),
        ("log_scalars_count", ctypes.c_uint),
        ("window_bits", ctypes.c_uint),
        ("piece", ctypes.c_uint),
        ("fixed_base_log_count", ctypes.c_uint),
    ]


class bc_amd_ntt_job(ctypes.Structure):
    _fields_ = [
        ("inputs", ctypes.c_void_p),
        ("outputs", ctypes.c_void_p),
        ("log_values_count", ctypes.c_uint),
        ("bit_reversed_inputs", ctypes.c_bool),
        ("inverse", ctypes.c_bool),
        ("log_extension_degree", ctypes.c_uint),
        ("coset_index", ctypes.c_uint),
    ]


class bc_amd_pn_job(ctypes.Structure):
    _fields_ = [
        ("indexes", ctypes.c_void_p),
        ("scalars", ctypes.c_void_p),
        ("target", ctypes.c_void_p),
        ("columns_count", ctypes.c_uint),
        ("log_rows_count", ctypes.c_uint),
    ]


BC_AMD_JOB_MSM, BC_AMD_JOB_NTT, BC_AMD_JOB_PN = 0, 1, 2


class bc_amd_job(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("msm", bc_amd_msm_job), ("ntt", bc_amd_ntt_job), ("pn", bc_amd_pn_job)]


class bc_amd_job_stats(ctypes.Structure):
    _fields_ = [
        ("submit_ns", ctypes.c_ulonglong),
        ("start_ns", ctypes.c_ulonglong),
        ("done_ns", ctypes.c_ulonglong),
        ("batch_jobs", ctypes.c_uint),
        ("status", ctypes.c_int),
        ("enqueued_ns", ctypes.c_ulonglong),
        ("slot", ctypes.c_uint),
    ]


class bc_amd_msm_tuning(ctypes.Structure):
    _fields_ = [("window_bits", ctypes.c_uint), ("piece", ctypes.c_uint)]


class bc_amd_plan(ctypes.Structure):
    _fields_ = [("handle", ctypes.c_void_p)]


class bc_amd_server(ctypes.Structure):
    _fields_ = [("handle", ctypes.c_void_p)]


_lib = None


def lib():
    """Loads libbellman.so (raises if it is missing: the HIP library is the only path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libbellman.so not built at %s (run make -C era-bellman-cuda_amd)" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH)
        for name in HEADER_SYMBOLS + EXTENSION_SYMBOLS:
            getattr(_lib, name).restype = ctypes.c_int
    return _lib


def check(fn, code):
    if code != bc_success:
        raise BellmanError(fn, code)


def call(name, *args):
    code = getattr(lib(), name)(*args)
    check(name, code)


# Every entry point declared by include/bellman-cuda.h (70, as src/bellman-cuda.h:44-549).
HEADER_SYMBOLS = [
    "bc_get_device_count", "bc_get_device", "bc_set_device", "bc_stream_create", "bc_stream_wait_event", "bc_stream_synchronize",
    "bc_stream_query", "bc_stream_destroy", "bc_launch_host_fn", "bc_event_create", "bc_event_record", "bc_event_synchronize",
    "bc_event_query", "bc_event_destroy", "bc_event_elapsed_time", "bc_mem_get_info", "bc_malloc", "bc_malloc_host", "bc_free",
    "bc_free_host", "bc_host_register", "bc_host_unregister", "bc_device_disable_peer_access", "bc_device_enable_peer_access",
    "bc_memcpy", "bc_memcpy_async", "bc_memset", "bc_memset_async", "bc_mem_pool_create", "bc_mem_pool_destroy",
    "bc_mem_pool_disable_peer_access", "bc_mem_pool_enable_peer_access", "bc_malloc_from_pool_async", "bc_free_async",
    "ff_set_up", "ff_set_value", "ff_set_value_zero", "ff_set_value_one", "ff_ax", "ff_a_plus_x", "ff_x_plus_y", "ff_ax_plus_y",
    "ff_x_minus_y", "ff_ax_minus_y", "ff_x_minus_ay", "ff_x_mul_y", "ff_grand_product", "ff_multiply_by_powers", "ff_inverse",
    "ff_poly_evaluate", "ff_get_powers_of_w", "ff_get_powers_of_g", "ff_omega_shift", "ff_bit_reverse", "ff_bit_reverse_multigpu",
    "ff_select", "ff_sort_u32", "ff_tear_down", "pn_set_up", "pn_generate_permutation_polynomials", "pn_set_values_from_packed_bits",
    "pn_distribute_values", "pn_tear_down", "msm_set_up", "msm_execute_async", "msm_tear_down", "ntt_set_up", "ntt_execute_async",
    "ntt_execute_async_multigpu", "ntt_tear_down",
]

# Extension entry points (include/bellman-amd.h) with no counterpart in the reference.
EXTENSION_SYMBOLS = [
    "bc_amd_generator_multiples", "bc_amd_msm_fold", "bc_amd_msm_execute_no_alloc", "bc_amd_msm_fixed_base_table", "bc_amd_msm_batch_async", "bc_amd_ntt_batch_async", "bc_amd_msm_get_tuning", "bc_amd_msm_set_tuning",
    "bc_amd_ntt_get_split", "bc_amd_ntt_set_split", "bc_amd_msm_autotune", "bc_amd_ntt_autotune", "bc_amd_tuning_save",
    "bc_amd_tuning_load", "bc_amd_plan_create", "bc_amd_plan_workspace_bytes", "bc_amd_plan_launch", "bc_amd_plan_destroy",
    "bc_amd_job_footprint", "bc_amd_job_workspace_bytes", "bc_amd_server_create", "bc_amd_server_submit", "bc_amd_server_wait", "bc_amd_server_destroy",
]


# ---------------------------------------------------------------- helpers
class Device:
    """A stream, a pool and small helpers for moving numpy arrays in and out."""

    def __init__(self, device_id=0):
        call("bc_set_device", ctypes.c_int(device_id))
        self.stream = bc_stream()
        call("bc_stream_create", ctypes.byref(self.stream), ctypes.c_bool(False))
        self.pool = bc_mem_pool()
        call("bc_mem_pool_create", ctypes.byref(self.pool), ctypes.c_int(device_id))
        self._bufs = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        call("bc_malloc", ctypes.byref(p), ctypes.c_size_t(max(int(nbytes), 16)))
        self._bufs.append(p.value)
        return p.value

    def free(self, ptr):
        if ptr in self._bufs:
            self._bufs.remove(ptr)
            call("bc_free", ctypes.c_void_p(ptr))

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        ptr = self.malloc(a.nbytes)
        call("bc_memcpy", ctypes.c_void_p(ptr), a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes))
        return ptr

    def download(self, ptr, shape, dtype=np.uint32):
        out = np.zeros(shape, dtype=dtype)
        call("bc_memcpy", out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes))
        return out

    def sync(self):
        call("bc_stream_synchronize", self.stream)

    def close(self):
        self.sync()
        for p in list(self._bufs):
            self.free(p)
        call("bc_mem_pool_destroy", self.pool)
        call("bc_stream_destroy", self.stream)


def ptr(a):
    return ctypes.c_void_p(a) if isinstance(a, int) else a.ctypes.data_as(ctypes.c_void_p)


def msm_job(bases, scalars, results, log_n, window_bits=0, piece=0, fixed_base_log_count=0):
    return bc_amd_msm_job(bases, scalars, results, log_n, window_bits, piece, fixed_base_log_count)


def ntt_job(inputs, outputs, log_n, bitrev=False, inverse=False, log_ext=0, coset=0):
    return bc_amd_ntt_job(inputs, outputs, log_n, bitrev, inverse, log_ext, coset)


def array_of(ctype, items):
    arr = (ctype * max(len(items), 1))()
    for i, it in enumerate(items):
        arr[i] = it
    return arr


class Server:
    """The continuous-batching request loop (bc_amd_server_*)."""

    def __init__(self, hbm_budget_bytes, max_in_flight=2):
        self.handle = bc_amd_server()
        call("bc_amd_server_create", ctypes.byref(self.handle), ctypes.c_size_t(int(hbm_budget_bytes)), ctypes.c_uint(max_in_flight))

    def submit(self, job):
        t = ctypes.c_ulonglong()
        call("bc_amd_server_submit", self.handle, ctypes.byref(job), ctypes.byref(t))
        return t.value

    def wait(self, ticket):
        st = bc_amd_job_stats()
        call("bc_amd_server_wait", self.handle, ctypes.c_ulonglong(ticket), ctypes.byref(st))
        return st

    def close(self):
        if self.handle.handle:
            call("bc_amd_server_destroy", self.handle)
            self.handle.handle = None
