"""aurora_hip_mcc_collide_device, the collide that allocates nothing (its
per-wave energy rows live in a fixed device array), against
aurora_hip_mcc_collide on the same velocities: velocities, counts and energy
change identical, on two calls in a row (the second smaller), so nothing the
first call left in the shared scratch leaks into the second.  Device buffers
come from the HIP runtime directly (hipMalloc through ctypes)."""
import ctypes as C
import os

import numpy as np
import pytest

from aurora_pic import workloads

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        self.lib = C.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def put(self, dst, arr):
        assert self.lib.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1) == 0

    def get(self, src, arr):
        assert self.lib.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2) == 0
        return arr

    def free(self, p):
        assert self.lib.hipFree(p) == 0


def test_collide_without_allocation_equals_collide():
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    dt = 1.0 / (400 * 13.56e6)
    rng = np.random.default_rng(9)
    rt = _Hip()
    with hip.Context(0) as c:
        c.mesh([65], [0.067], hip.DIRICHLET, workloads.EPS0)
        sid = c.add_species(-workloads.EV, workloads.ME, 1.0, 16)
        ion = c.add_species(workloads.EV, 6.67e-27, 1.0, 16)
        c.mcc_load(sid, gas, product_species=(sid, ion), seed=4242)
        nch = len(gas["channels"])
        cap = 300_000
        dv = rt.alloc(cap * 3 * 4)
        dcounts = rt.alloc((nch + 3) * 8)
        denergy = rt.alloc(nch * 8)
        dstatus = rt.alloc(4)
        try:
            for call, n in enumerate((300_000, 70_001)):
                print(f"no-alloc collide call {call}: {n} velocities", flush=True)
                # electrons up to ~60 eV: every channel fires
                v = (rng.normal(0.0, 2.5e6, size=(n, 3))).astype(np.float32)
                ref_v, ref_counts, ref_e = c.mcc_collide(v.astype(np.float64), dt, step=call,
                                                         stream_id=3)
                rt.put(dv, v)
                c.mcc_collide_device(n, dv, dt, call, dcounts, denergy, dstatus, stream_id=3)
                c.sync()
                out_v = rt.get(dv, np.zeros((n, 3), dtype=np.float32))
                counts = rt.get(dcounts, np.zeros(nch + 3, dtype=np.uint64))
                energy = rt.get(denergy, np.zeros(nch))
                status = rt.get(dstatus, np.zeros(1, dtype=np.uint32))
                assert status[0] == 0
                assert np.array_equal(out_v.astype(np.float64), ref_v)
                assert np.array_equal(counts, ref_counts)
                assert np.array_equal(energy, ref_e)
                assert counts[2 + 2] > 0            # ionizations happened
        finally:
            for p in (dv, dcounts, denergy, dstatus):
                rt.free(p)
