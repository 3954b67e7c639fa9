"""The no-allocation gradient entry point, mpm_mlp_loss_grad_noalloc, against
the allocating mpm_mlp_loss_grad on the same batches, through the C ABI.

Two calls in a row, the second on a smaller batch (fewer workgroup partials
than the first left in the shared scratch), one with loss weights and one
without.  Floating point: the two entry points sum the same per-sample terms
over different numbers of workgroup partials, so they agree to f32 round-off:
the loss within 1e-6 relative, every gradient entry within 1e-5 of the
largest.  The test's own device buffers come from the HIP runtime the engine
library uses (the caller allocates inputs and outputs; the entry point
allocates nothing).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Caller-side device memory for the test, from the engine's HIP runtime."""

    def __init__(self, hip, host):
        self.hip = hip
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def read(self, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def hip_runtime(engine):
    engine.lib()  # loads the engine library and with it its HIP runtime
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def test_noalloc_gradient_matches_the_allocating_entry_point(engine):
    import litter_physics_amd.surrogate as S

    hip = hip_runtime(engine)
    net = S.load_network()
    P = net.param_count()
    rng = np.random.default_rng(17)
    for n, weighted in ((5000, True), (500, False)):
        print(f"noalloc case n={n} weighted={weighted}", flush=True)
        x = (0.5 * rng.standard_normal((n, 12))).astype(np.float32)
        t = (0.1 * rng.standard_normal((n, 7))).astype(np.float32)
        w = np.linspace(0.5, 2.0, 7).astype(np.float32) if weighted else None
        loss_ref, g_ref = net.loss_grad(x, t, w)
        g0 = np.full(P, np.nan, np.float32)
        l0 = np.full(1, np.nan, np.float64)
        bufs = [DeviceBuffer(hip, a) for a in (x, t, g0, l0)]
        wb = DeviceBuffer(hip, w) if weighted else None
        try:
            net.loss_grad_device(n, bufs[0].ptr.value, bufs[1].ptr.value, bufs[3].ptr.value, bufs[2].ptr.value,
                                 wb.ptr.value if weighted else None)
            g = bufs[2].read(g0)
            loss = float(bufs[3].read(l0)[0])
        finally:
            for b in bufs + ([wb] if wb else []):
                b.free()
        assert np.all(np.isfinite(g))
        assert abs(loss - loss_ref) <= 1e-6 * abs(loss_ref), (loss, loss_ref)
        assert np.abs(g - g_ref).max() <= 1e-5 * np.abs(g_ref).max()


def test_noalloc_rejects_an_oversized_batch(engine):
    import litter_physics_amd.surrogate as S

    net = S.load_network()
    with pytest.raises(engine.MpmError, match="exceeds the maximum"):
        # Rejected before any pointer is read.
        net.loss_grad_device(2 ** 31, 1, 1)
