"""MI355X-native PRNU common-source-identification hot path (Rocket's CSI
pipeline): extraction filters, NCC Gram on MFMA and PCE, behind a C ABI
(include/prnu_api.h) implemented by libprnu_hip.so."""
from . import _lib  # noqa: F401
from ._lib import PrnuError, PrnuInvalidArgument  # noqa: F401

__all__ = ["PrnuError", "PrnuInvalidArgument"]
