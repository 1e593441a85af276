"""Host-side Python layer of the HIP batched MuJoCo step (mjmlx C API)."""
from . import scenes  # noqa: F401
