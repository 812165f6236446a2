"""sphinxsys_amd -- MI355X-native weakly-compressible SPH step behind the
SPHinXsys computing-kernel dynamics API.

`load()` binds the in-tree HIP library (sphinxsys_amd/lib/libsphx.so).  There
is no CPU fallback: if the library is missing or was not built for a GPU that
is present, `load()` raises.
"""
import os

from ._capi import SphxError, SphxLibrary  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "lib", "libsphx.so")
_lib = None


def load():
    """Return the SphxLibrary bound to the HIP build; raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise SphxError(
                f"HIP library not built: {LIBRARY_PATH} is missing "
                "(run __graft_entry__.build() or python sphinxsys_amd/build.py)")
        lib = SphxLibrary(LIBRARY_PATH)
        if lib.backend != "hip-gfx950":
            raise SphxError(f"unexpected backend {lib.backend!r} at {LIBRARY_PATH}")
        _lib = lib
    return _lib
