"""MI355X-native NEGF charge path (RGF, lead self-energies, charge
integration, Broyden SCF) with a C ABI in include/negf_amd.h."""
from ._lib import load, check, NegfError, LIB_PATH  # noqa: F401

__all__ = ["load", "check", "NegfError", "LIB_PATH"]
