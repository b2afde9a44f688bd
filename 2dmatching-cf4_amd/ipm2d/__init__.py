"""ipm2d: MI355X (gfx950) particle hot path of iPM2D (push, absorption,
deposit, null-collision MCC, cell sort) behind the reference's per-step
particle interface.  The compute runs in the HIP library built from
``csrc/`` (see ``_native``); this package is the host side."""
from . import _native, geometry, xsec, workload  # noqa: F401
from .hotpath import HotPath, from_config, from_config_device  # noqa: F401
from ._native import IPMError, CapacityError  # noqa: F401
from . import circuit, recombination, refconfig, rezoning  # noqa: F401
