"""MI355X-native time-stepping core of the AlpineQMS quasi-magnetostatic
A-V solver (reference: DanRich123/AlpineQMS, 3D/QMS_3D.f90).

* ``native``   -- ctypes binding of libqms.so (HIP kernels, C ABI in include/qms.h)
* ``qms_io``   -- inputs.txt / geom.bin / Fortran-record formats
* ``program``  -- the QMS_3D program flow (drop-in executable ``alpineqms_amd/QMS_3D``)
* ``configs``  -- seeded generators of the benchmark designs
"""
__all__ = ["native", "qms_io", "program", "configs"]
