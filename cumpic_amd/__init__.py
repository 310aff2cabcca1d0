"""MI355X-native CUMC3D ideal-MHD time step (PPM/TVD + HLL + upwind CT + SSP-RK3)."""
from .config import Config  # noqa: F401
