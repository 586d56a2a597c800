"""Matrix-vector backend preference (src/pref.jl:1-21).

The reference chooses the BLAS used for its three mantle products at
precompile time: "LinearAlgebra" (OpenBLAS/MKL) or "Octavian".  Here every
product runs in the gfx950 RHS kernel; the preference is kept so code that
sets it behaves as before (same values, same error, same message).
"""
import logging

_BACKENDS = ("LinearAlgebra", "Octavian")
_current = "LinearAlgebra"
log = logging.getLogger("oetqf_amd")


def get_matvecmul():
    """get_matvecmul!() (pref.jl:1-3); default "LinearAlgebra"."""
    return _current


def set_matvecmul(backend):
    """set_matvecmul!(backend) (pref.jl:7-13): ValueError (ArgumentError) for
    an unknown backend, otherwise logs the restart notice."""
    global _current
    if backend not in _BACKENDS:
        raise ValueError("Invalid backend: %s" % backend)
    _current = backend
    msg = "New backend %s set; restart your Julia session for this change to take effect!" % backend
    log.info(msg)
    return msg
