"""The part of ``numba.cuda`` the reference's drivers use (run_randommode_3d.py:244-290,
tests/verification_3d.py:771-905): ``to_device``, ``device_array``,
``synchronize`` and ``copy_to_host`` on the returned arrays.  Memory comes
from the library's allocator on the current HIP device."""
from .dense import DeviceArray, device_array, synchronize, to_device

DeviceNDArray = DeviceArray

__all__ = ["DeviceArray", "DeviceNDArray", "device_array", "synchronize", "to_device"]
