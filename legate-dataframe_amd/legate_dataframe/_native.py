"""Loads the pybind11 extension; there is no fallback when it (or libldf_hip.so) is missing."""

from __future__ import annotations

try:
    from legate_dataframe import _ldf  # noqa: F401
except ImportError as e:  # pragma: no cover - reported to the user
    raise ImportError(
        "legate_dataframe native extension is not built (run `python legate-dataframe_amd/build.py`): " + str(e)
    ) from e

lib = _ldf


def sync_stream() -> None:
    """Runs the operators on torch's current stream when torch is in use."""
    import sys

    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available() and torch.cuda.is_initialized():
        lib.set_stream(torch.cuda.current_stream().cuda_stream)
