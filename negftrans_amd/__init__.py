"""MI355X-native RGF energy loop of NEGFtrans (HIP kernels behind a C-ABI)."""
from . import negf  # noqa: F401
