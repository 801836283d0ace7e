"""HIP-event timing on the engine's own stream."""
from __future__ import annotations

import ctypes


class HipEvents:
    """hipEventCreate / hipEventRecord / hipEventElapsedTime via ctypes."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(f"HIP event call failed ({rc})")

    def create(self):
        e = ctypes.c_void_p()
        self._ok(self.hip.hipEventCreate(ctypes.byref(e)))
        return e

    def record(self, e, stream):
        self._ok(self.hip.hipEventRecord(e, ctypes.c_void_p(stream)))

    def elapsed_ms(self, a, b):
        self._ok(self.hip.hipEventSynchronize(b))
        ms = ctypes.c_float()
        self._ok(self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b))
        return float(ms.value)

    def destroy(self, e):
        self.hip.hipEventDestroy(e)


def time_region(ev: HipEvents, sweeper, fn, count: int) -> float:
    """Milliseconds for `count` calls of fn, bracketed by events recorded on
    the sweeper's stream (the stream its kernels run on)."""
    start, stop = ev.create(), ev.create()
    stream = sweeper.stream()
    sweeper.synchronize()
    ev.record(start, stream)
    for _ in range(count):
        fn()
    ev.record(stop, stream)
    ms = ev.elapsed_ms(start, stop)
    ev.destroy(start)
    ev.destroy(stop)
    return ms
