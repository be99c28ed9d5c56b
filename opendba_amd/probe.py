"""Roofline probes (opendba_amd/lib/libodba_probe.so), used by bench.py.

`cell_mix_rate` measures the VALU issue rate of the DTW cell update's own
instruction mix (v_sub_f32, v_mul_f32, v_min3_f32, v_add_f32 on independent
chains). `path_mix_rate` does the same for the path cell (10 lane-ops: the
four above, two more adds, two compares and two carry-adds that shift the
step bits into plane words) with the path kernels' dependencies: 16 columns
per lane chained through the left value, the left edge from the previous
lane by a DPP wave shift. `copy_ms` times a float4 copy of a known size; run under
`rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` it is the calibration for
those counters.

    python3 -m opendba_amd.probe mix [blocks_per_cu ...]
    python3 -m opendba_amd.probe pathmix [blocks_per_cu ...]
    python3 -m opendba_amd.probe copy [bytes]
"""
import ctypes as C
import os
import sys

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libodba_probe.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} is not built (make -C opendba_amd/csrc)")
        lib = C.CDLL(_LIB)
        lib.odba_probe_cell_mix.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.odba_probe_cell_mix.restype = C.c_int
        lib.odba_probe_path_mix.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.odba_probe_path_mix.restype = C.c_int
        lib.odba_probe_copy.argtypes = [C.c_int, C.c_uint64, C.POINTER(C.c_double)]
        lib.odba_probe_copy.restype = C.c_int
        _lib = lib
    return _lib


def cell_mix_rate(blocks_per_cu=16, iters=20000, device=0):
    """(lane-ops/s, ms) of one timed cell-mix launch with `blocks_per_cu`
    256-thread workgroups per CU."""
    rate, ms = C.c_double(), C.c_double()
    rc = _load().odba_probe_cell_mix(device, blocks_per_cu, iters, C.byref(rate), C.byref(ms))
    if rc:
        raise RuntimeError(f"odba_probe_cell_mix failed ({rc})")
    return rate.value, ms.value


def path_mix_rate(blocks_per_cu=4, iters=20000, device=0):
    """(lane-ops/s, ms) of one timed path-cell-mix launch with
    `blocks_per_cu` 256-thread workgroups per CU (10 lane-ops per cell)."""
    rate, ms = C.c_double(), C.c_double()
    rc = _load().odba_probe_path_mix(device, blocks_per_cu, iters, C.byref(rate), C.byref(ms))
    if rc:
        raise RuntimeError(f"odba_probe_path_mix failed ({rc})")
    return rate.value, ms.value


def copy_ms(nbytes=1 << 30, device=0):
    ms = C.c_double()
    rc = _load().odba_probe_copy(device, nbytes, C.byref(ms))
    if rc:
        raise RuntimeError(f"odba_probe_copy failed ({rc})")
    return ms.value


def main(argv):
    mode = argv[1] if len(argv) > 1 else "mix"
    if mode == "mix":
        for bpc in [int(a) for a in argv[2:]] or [1, 2, 4, 8, 16]:
            rate, ms = cell_mix_rate(bpc)
            print(f"cell mix, {bpc:2d} workgroups/CU: {ms:8.3f} ms  {rate / 1e12:7.2f} T lane-ops/s", flush=True)
    elif mode == "pathmix":
        for bpc in [int(a) for a in argv[2:]] or [1, 2, 4, 8, 16]:
            rate, ms = path_mix_rate(bpc)
            print(f"path cell mix, {bpc:2d} workgroups/CU: {ms:8.3f} ms  {rate / 1e12:7.2f} T lane-ops/s  "
                  f"{rate / 10 / 1e9:8.1f} Gcells/s  {rate / 78.65e12:.3f} of the FP32 vector peak", flush=True)
    elif mode == "copy":
        nbytes = int(argv[2]) if len(argv) > 2 else 1 << 30
        ms = copy_ms(nbytes)
        print(f"copy {nbytes} B: {ms:.3f} ms  {2.0 * nbytes / (ms * 1e-3) / 1e9:.1f} GB/s read+write", flush=True)
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
