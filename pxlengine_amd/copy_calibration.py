"""Reference copy for HBM-traffic counters: three device-to-device copies of 256 MiB
through the HIP runtime (hipMemcpy), so that rocprofv3's FETCH_SIZE / WRITE_SIZE for the
copy kernel can be compared with the known 256 MiB read + 256 MiB written per copy."""
import ctypes
import sys

N = 256 * 1024 * 1024


def main():
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(N)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(N)):
        sys.exit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(N))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(N), 3):  # hipMemcpyDeviceToDevice
            sys.exit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"copied {N} bytes x3")


if __name__ == "__main__":
    main()
