"""Known-size device copies for calibrating rocprofv3's FETCH_SIZE and
WRITE_SIZE counters (run under `rocprofv3 --pmc FETCH_SIZE`, then
`--pmc WRITE_SIZE`): copies a buffer of --mib MiB device to device three times
with hipMemcpy through the HIP runtime (no torch), so every dispatch of the
runtime's copy kernel reads and writes exactly that many bytes.  The ratio
bytes / (counter KiB * 1024) is the scale bench.py applies to those counters.

Measurement infrastructure, not a test (pytest does not collect this file).
"""
import argparse
import ctypes
import json
import sys

HIP_MEMCPY_DEVICE_TO_DEVICE = 3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--copies", type=int, default=3)
    args = ap.parse_args()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    n = args.mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), n) or hip.hipMalloc(ctypes.byref(dst), n):
        print("hipMalloc failed", file=sys.stderr)
        return 1
    try:
        if hip.hipMemset(src, 1, n):
            return 1
        for _ in range(args.copies):
            if hip.hipMemcpy(dst, src, n, HIP_MEMCPY_DEVICE_TO_DEVICE):
                print("hipMemcpy failed", file=sys.stderr)
                return 1
        if hip.hipDeviceSynchronize():
            return 1
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    print(json.dumps({"bytes_read_per_copy": n, "bytes_written_per_copy": n,
                      "copies": args.copies}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
