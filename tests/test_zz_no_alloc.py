"""odba_dtw_all_pairs_noalloc against odba_dtw_all_pairs_resident through the
C ABI: bit-identical distances and flags on two calls in a row (the second
on a smaller set, into buffers pre-filled with garbage), in two modes and
two value types, and a set too long for the fixed scratch refused before
anything runs. The test allocates its device buffers itself (the entry
point allocates nothing)."""
import ctypes as C

import numpy as np
import pytest

from opendba_amd import native as N
from opendba_amd import workloads as WL

pytestmark = pytest.mark.gpu

ODBA_E_INVALID = 101


class DeviceBuffer:
    """Device memory from the HIP runtime the library itself uses."""

    def __init__(self, nbytes):
        N.load()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        assert self.hip.hipMalloc(C.byref(self.ptr), max(nbytes, 1)) == 0

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def noalloc(ctx, sset, os_, oe, dist_buf, flag_buf):
    return ctx.L.odba_dtw_all_pairs_noalloc(ctx.h, sset.h, os_, oe, dist_buf.ptr, flag_buf.ptr)


def _sets():
    big = [s.astype(np.float32) for s in WL.random_walks(40, 200, 2600, seed=41)]
    small = [s.astype(np.float32) for s in WL.random_walks(9, 100, 700, seed=42)]
    ibig = [np.round(s * 10).astype(np.int32) for s in big[:24]]
    ismall = [np.round(s * 10).astype(np.int32) for s in small]
    return [(big, small), (ibig, ismall)]


@pytest.mark.parametrize("mode", [(0, 0), (0, 1)])
@pytest.mark.parametrize("which", [0, 1])
def test_noalloc_equals_resident_two_calls(mode, which):
    os_, oe = mode
    first, second = _sets()[which]
    npairs_max = len(first) * (len(first) - 1) // 2
    dist = DeviceBuffer(npairs_max * 8)
    flags = DeviceBuffer(npairs_max * 4)
    try:
        with N.Context(0) as ctx:
            for seqs in (first, second):
                n = len(seqs)
                npairs = n * (n - 1) // 2
                garbage = np.frombuffer(np.full(npairs_max * 8, 0xA5, np.uint8).tobytes(), np.uint8)
                dist.upload(garbage)
                flags.upload(garbage[: npairs_max * 4])
                ss = ctx.seqset(seqs)
                try:
                    exp, efl, _ = ss.all_pairs(os_, oe, flags=True)
                    print(f"noalloc case: n={n} dtype={seqs[0].dtype} mode={mode}", flush=True)
                    rc = noalloc(ctx, ss, os_, oe, dist, flags)
                    assert rc == 0, ctx.L.odba_last_error(ctx.h).decode()
                    got = dist.download(seqs[0].dtype, npairs)
                    gfl = flags.download(np.int32, npairs)
                finally:
                    ss.close()
                assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))
                assert np.array_equal(gfl, efl)
    finally:
        dist.free()
        flags.free()


def test_noalloc_refuses_sequences_beyond_the_fixed_scratch():
    rng = np.random.default_rng(5)
    seqs = [rng.normal(size=1_100_000), rng.normal(size=300)]  # float64: 8 B per value
    dist = DeviceBuffer(8)
    try:
        with N.Context(0) as ctx:
            ss = ctx.seqset(seqs)
            try:
                rc = ctx.L.odba_dtw_all_pairs_noalloc(ctx.h, ss.h, 0, 0, dist.ptr, None)
                assert rc == ODBA_E_INVALID
                assert "fixed scratch" in ctx.L.odba_last_error(ctx.h).decode()
            finally:
                ss.close()
    finally:
        dist.free()
