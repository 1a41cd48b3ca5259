"""bmx -- MI355X (gfx950) multi-particle tracker for Bmad's bmad_standard
tracking, behind the libbmx C ABI (include/bmx.h)."""
from . import abi, beam, lattice  # noqa: F401

__all__ = ["abi", "beam", "lattice", "tracker"]
