"""MI355X-native groundwater steady/transient solve and adjoint for DPFEHM.

The package directory is ``dpfehm.jl_amd``; because of the dot it is loaded
under the module name ``dpfehm_jl_amd`` (``load_package`` in
``__graft_entry__.py`` and ``tests/conftest.py``).

* ``api``       -- the reference's groundwater interface on the GPU
                   (``groundwater_steadystate``, its rrule, ``groundwater_residuals``,
                   the transient solve, the batched ensemble)
* ``_lib``      -- ctypes binding of the C-ABI library ``libdpfgpu.so``
* ``grid``      -- regulargrid1d/2d/3d with the reference's ordering
* ``fields``    -- seeded Matérn log-conductivity fields
* ``workloads`` -- the synthetic problems used by tests and the benchmark
"""
from . import fields, grid, workloads  # noqa: F401
