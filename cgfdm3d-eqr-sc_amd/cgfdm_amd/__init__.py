"""MI355X-native CGFDM3D-EQR-SC time-stepping path (HIP kernels behind a C ABI)."""
from .model import Model, HALO  # noqa: F401
