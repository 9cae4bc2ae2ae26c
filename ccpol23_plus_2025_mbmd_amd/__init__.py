"""MI355X-native CCpol23+(2025) three-body FS2/FS3 term (drop-in for
``threebody_term`` of source/threebody_potential.f90)."""
from .params import ThreeBodyParams, Data3B, load_variant, read_data3b, read_control_3b  # noqa: F401

__version__ = "0.1.0"
