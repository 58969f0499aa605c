"""Device buffers for the GPU tests (the package's own helper over the HIP
runtime that libscb.so links)."""
from scbf_amd.devmem import DeviceBuffer, device_free_bytes, synchronize  # noqa: F401
