"""flurrypp_amd: FlurryPP's FR residual and RK update on MI355X (gfx950).

The hot path lives in ``lib/libflurry_hip.so`` (C ABI, include/flurry_hip.h);
host set-up in ``lib/libflurry_host.so`` (include/flurry_host.h).
"""
from .input import InputParams  # noqa: F401

__all__ = ["InputParams"]
