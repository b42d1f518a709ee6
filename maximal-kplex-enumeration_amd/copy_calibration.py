"""Plain device-to-device copy of a known size on the current HIP device.

Run under `rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE): the copy kernel reads
and writes exactly N bytes, which calibrates what the counters report per
HBM byte for the roofline traffic figure.  Usage: copy_calibration.py [N]
"""
import ctypes
import sys


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 1 << 30
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)):
            raise SystemExit("hipMalloc failed")
    if hip.hipMemset(a, 1, ctypes.c_size_t(n)) or hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3):
        raise SystemExit("hipMemset/hipMemcpy failed")
    if hip.hipDeviceSynchronize():
        raise SystemExit("hipDeviceSynchronize failed")
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"copied {n} bytes device-to-device", flush=True)


if __name__ == "__main__":
    main()
