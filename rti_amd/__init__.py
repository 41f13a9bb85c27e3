"""MI355X (gfx950) Monte Carlo flux tracer: the hot path of rti behind a C ABI.

* ``rti_amd._lib``      -- ctypes binding of librti_hip.so (include/rti/rti_c.h)
* ``rti_amd.device``    -- Python mirror of the reference's rti::device API
* ``rti_amd.geometry``  -- seeded synthetic trench / hole / rough-surface scenes
* ``rti_amd.build``     -- in-tree hipcc build of librti_hip.so
"""
from . import geometry  # noqa: F401
from ._lib import Context, RtiError, default_params, run_scene  # noqa: F401
