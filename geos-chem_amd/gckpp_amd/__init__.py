"""gckpp_amd: host side of the MI355X-native KPP chemistry solve.

The compute path is the HIP library csrc/libgckpp_hip.so behind the C ABI in
include/gckpp_hip.h; this package only mirrors the reference's operator
interface (Update_RCONST / Integrate over a batch of cells) and builds inputs.
"""
from . import fields, mechanism  # noqa: F401
