"""MI355X-native (gfx950) ThunderEgg patch-parallel GMG hot path.

Host-side Python view of the C ABI (include/thunderegg_amd.h); the C++ API
mirroring the reference's classes lives in include/ThunderEgg/ (ThunderEgg/amd.h).
"""
from . import mesh  # noqa: F401
