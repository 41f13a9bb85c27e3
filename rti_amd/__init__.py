"""MI355X (gfx950) Monte Carlo flux tracer: the hot path of rti behind a C ABI.

* ``rti_amd._lib``      -- ctypes binding of librti_hip.so (include/rti/rti_c.h):
                           ``Context``, ``run_scene``, ``default_params``
* ``rti_amd.geometry``  -- seeded synthetic trench / hole / rough-surface / batch scenes
* ``rti_amd.tune``      -- autotuner of the persistent kernel's launch knobs
* ``rti_amd.autograd``  -- torch.autograd binding of the adjoint (d flux / ds, d flux / dn)
* ``rti_amd.build``     -- in-tree hipcc build of librti_hip.so and the command-line tool
* ``rti_amd.cli``       -- rti_cli, the file-level replacement of the reference's executable
* ``rti_amd.tools``     -- drivers and summaries for rocprofv3 measurements

The C++ drop-in for the reference's ``rti::device`` is include/rti/device.hpp.
"""
from . import geometry  # noqa: F401
from ._lib import Context, RtiError, default_params, run_scene  # noqa: F401
