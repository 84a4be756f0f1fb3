"""Loads the in-tree HIP library (flowfpmpp_amd/lib/libffpm.so).

There is no CPU fallback: if the library is missing or cannot be loaded, the
import of any solver entry point raises. Build it with
``python -m flowfpmpp_amd.build`` (hipcc, gfx950).
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libffpm.so")

# include/ffpm.h status codes
OK = 0
STATUS_NAMES = {
    1: "FFPM_ERR_ARGUMENT", 2: "FFPM_ERR_ISOLATED", 3: "FFPM_ERR_NO_OUTER_FACE",
    4: "FFPM_ERR_METHOD", 5: "FFPM_ERR_SHAPE", 6: "FFPM_ERR_DECOMPOSITION", 7: "FFPM_ERR_DEVICE",
}

_lib = None


class FFPMError(RuntimeError):
    """A failed library call; ``status`` is the include/ffpm.h code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"HIP library {LIB_PATH} is not built; run `python -m flowfpmpp_amd.build`")
        _lib = C.CDLL(LIB_PATH)
        _lib.ffpm_last_error.restype = C.c_char_p
    return _lib


def check(status: int) -> None:
    if status != OK:
        msg = lib().ffpm_last_error().decode(errors="replace")
        raise FFPMError(status, f"{STATUS_NAMES.get(status, status)}: {msg}")


def symbols() -> list:
    """Names declared in include/ffpm.h (function prototypes)."""
    import re
    hdr = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                       "ffpm.h")
    text = open(hdr).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ffpm_[a-z0-9_]+)\s*\(", text)))
