"""MI355X-native NEP molecular-dynamics step (dmgx).

The compute path is libdmgx.so (HIP kernels for gfx950 behind the C ABI of
include/dmgx.h); this package is the host-side mirror of the reference's NEP
force-evaluator and integrator interfaces.
"""
from .nep import MD, NEP, Box, Server, K_B, PRESSURE_UNIT_CONVERSION, TIME_UNIT_CONVERSION  # noqa: F401
from ._lib import DmgxError, lib  # noqa: F401
