"""Thin wrappers that drive libbellman through its C ABI (ctypes) for tests."""
import ctypes

import numpy as np

import bellman_amd as B


def u(v):
    return ctypes.c_uint(int(v))


def vp(p):
    return ctypes.c_void_p(p)


def ntt(dev, x, log_n, bitrev=False, inverse=False, log_ext=0, coset=0, host_in=False, host_out=False, can_overwrite=False):
    """Runs ntt_execute_async; inputs/outputs on device unless host_* is set."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    n = 1 << log_n
    out_host = np.zeros((n, 8), dtype=np.uint32)
    d_in = None if host_in else dev.upload(x)
    d_out = None if host_out else dev.malloc(n * 32)
    cfg = B.ntt_configuration()
    cfg.mem_pool = dev.pool
    cfg.stream = dev.stream
    cfg.inputs = x.ctypes.data if host_in else d_in
    cfg.outputs = out_host.ctypes.data if host_out else d_out
    cfg.log_values_count = log_n
    cfg.bit_reversed_inputs = bitrev
    cfg.inverse = inverse
    cfg.can_overwrite_inputs = can_overwrite
    cfg.log_extension_degree = log_ext
    cfg.coset_index = coset
    B.call("ntt_execute_async", cfg)
    dev.sync()
    res = out_host if host_out else dev.download(d_out, (n, 8))
    for p in (d_in, d_out):
        if p is not None:
            dev.free(p)
    return res


def msm(dev, d_bases, scalars, log_n, host_scalars=False, host_results=False):
    scalars = np.ascontiguousarray(scalars, dtype=np.uint32)
    res_host = np.zeros((254, 24), dtype=np.uint32)
    d_s = None if host_scalars else dev.upload(scalars)
    d_r = None if host_results else dev.malloc(254 * 96)
    cfg = B.msm_configuration()
    cfg.mem_pool = dev.pool
    cfg.stream = dev.stream
    cfg.bases = d_bases
    cfg.scalars = scalars.ctypes.data if host_scalars else d_s
    cfg.results = res_host.ctypes.data if host_results else d_r
    cfg.log_scalars_count = log_n
    B.call("msm_execute_async", cfg)
    dev.sync()
    res = res_host if host_results else dev.download(d_r, (254, 24))
    for p in (d_s, d_r):
        if p is not None:
            dev.free(p)
    return res


class CopyHooks:
    """The h2d/d2h events and callbacks of msm_configuration and
    ntt_configuration: two events, and one host function that records the
    user data it is called with (0x11 for h2d, 0x22 for d2h)."""

    H2D, D2H = 0x11, 0x22

    def __init__(self):
        self.calls = []
        self.events = []
        for _ in range(2):
            ev = B.bc_event()
            B.call("bc_event_create", ctypes.byref(ev), ctypes.c_bool(False), ctypes.c_bool(True))
            self.events.append(ev)
        self.fn = B.bc_host_fn(lambda data: self.calls.append(data))

    def attach(self, cfg):
        cfg.h2d_copy_finished, cfg.d2h_copy_finished = self.events
        cfg.h2d_copy_finished_callback = cfg.d2h_copy_finished_callback = self.fn
        cfg.h2d_copy_finished_callback_data = self.H2D
        cfg.d2h_copy_finished_callback_data = self.D2H

    def wait_calls(self, count, seconds=5.0):
        """The host functions run on a runtime thread; a stream synchronize
        returns once they have, but allow a bounded grace period."""
        import time

        end = time.time() + seconds
        while len(self.calls) < count and time.time() < end:
            time.sleep(0.01)
        return list(self.calls)

    def events_complete(self):
        return all(B.lib().bc_event_query(ev) == 0 for ev in self.events)

    def close(self):
        for ev in self.events:
            B.call("bc_event_destroy", ev)


def ff_binary(dev, name, a, x, y, a_on_device=False):
    """Element-wise ops; `a` passed by host pointer or device pointer."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    n = x.shape[0]
    dx = dev.upload(x)
    dy = dev.upload(np.ascontiguousarray(y, dtype=np.uint32)) if y is not None else None
    dr = dev.malloc(n * 32)
    args = []
    a_host = None
    if a is not None:
        a_host = np.ascontiguousarray(a, dtype=np.uint32).reshape(1, 8)
        args.append(vp(dev.upload(a_host)) if a_on_device else B.ptr(a_host))
    args.append(vp(dx))
    if dy is not None:
        args.append(vp(dy))
    args += [vp(dr), u(n), dev.stream]
    B.call(name, *args)
    dev.sync()
    out = dev.download(dr, (n, 8))
    for p in (dx, dy, dr):
        if p is not None:
            dev.free(p)
    return out


def pn_cycles(cols, log_rows, seed):
    """Variable indexes forming random cycles of length 1..64 over the cells."""
    import random

    cells = cols << log_rows
    rng = random.Random(seed)
    perm = list(range(cells))
    rng.shuffle(perm)
    idx = np.zeros(cells, dtype=np.uint32)
    pos = var = 0
    while pos < cells:
        L = rng.randint(1, 64)
        idx[perm[pos:pos + L]] = var
        pos += L
        var += 1
    return idx


def pn(dev, idx, ks, cols, log_rows):
    """pn_generate_permutation_polynomials on device copies; returns the target."""
    di, dk = dev.upload(idx), dev.upload(ks)
    dt = dev.malloc((cols << log_rows) * 32)
    cfg = B.generate_permutation_polynomials_configuration(dev.pool, dev.stream, di, dk, dt, cols, log_rows)
    B.call("pn_generate_permutation_polynomials", cfg)
    dev.sync()
    out = dev.download(dt, (cols << log_rows, 8))
    for p in (di, dk, dt):
        dev.free(p)
    return out


def msm_batch(dev, jobs):
    """bc_amd_msm_batch_async over a list of bc_amd_msm_job; waits for the stream."""
    arr = B.array_of(B.bc_amd_msm_job, jobs)
    B.call("bc_amd_msm_batch_async", arr, ctypes.c_uint(len(jobs)), dev.pool, dev.stream)
    dev.sync()


def ntt_batch(dev, jobs):
    arr = B.array_of(B.bc_amd_ntt_job, jobs)
    B.call("bc_amd_ntt_batch_async", arr, ctypes.c_uint(len(jobs)), dev.pool, dev.stream)
    dev.sync()
