"""Minimal device-memory helper (HIP runtime through ctypes) for callers that
keep solutions resident on the GPU without a framework."""
import ctypes as C

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipDeviceGetName.argtypes = [C.c_char_p, C.c_int, C.c_int]
        _hip.hipDeviceGetAttribute.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_int]
    return _hip


def device_name(device=0):
    """Marketing name and compute-unit count of a HIP device, e.g. 'AMD
    Instinct MI355X/256cu' (the autotuner's cache key)."""
    buf = C.create_string_buffer(256)
    rc = hip().hipDeviceGetName(buf, 256, device)
    if rc != 0:
        raise RuntimeError(f"hipDeviceGetName({device}) failed: {rc}")
    cus = C.c_int()
    # hipDeviceAttributeMultiprocessorCount
    HIP_ATTR_CU = 63
    rc = hip().hipDeviceGetAttribute(C.byref(cus), HIP_ATTR_CU, device)
    return f"{buf.value.decode(errors='replace').strip()}/{cus.value if rc == 0 else '?'}cu"


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), nbytes)
        if rc != 0:
            raise MemoryError(f"hipMalloc({nbytes}) failed: {rc}")

    def to_host(self, out):
        """Copies into a contiguous numpy array of the same byte size."""
        assert out.nbytes == self.nbytes
        rc = hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
        return out

    def from_host(self, a):
        assert a.nbytes == self.nbytes
        rc = hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed: {rc}")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")
