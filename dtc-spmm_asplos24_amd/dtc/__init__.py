"""Host-side helpers around the ``DTCSpMM`` extension (libdtc, gfx950).

``load_extension()`` imports the in-tree extension and raises if it has not
been built: there is no CPU fallback anywhere in the product path.
"""
from __future__ import annotations

import os
import sys

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_extension():
    """Import and return the DTCSpMM extension module built next to this package."""
    if PKG_DIR not in sys.path:
        sys.path.insert(0, PKG_DIR)
    try:
        import DTCSpMM  # noqa: N813
    except ImportError as e:  # pragma: no cover - exercised only when unbuilt
        raise ImportError(
            "the DTCSpMM HIP extension is not built; run "
            f"`python {os.path.join(PKG_DIR, 'build.py')}` first ({e})") from e
    return DTCSpMM
