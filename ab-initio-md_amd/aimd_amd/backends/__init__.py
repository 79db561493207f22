"""
Force backends. Importing this package registers the ``hf`` backend: native
RHF/UHF energies and analytic gradients on the MI355X engine.
"""

from __future__ import annotations

from aimd_amd.backends.base import ForceBackend, GradientResult
from aimd_amd.backends.registry import get_backend, list_backends, register_backend
from aimd_amd.backends import hf as _hf  # noqa: F401,E402  (registration side effect)

__all__ = ["ForceBackend", "GradientResult", "get_backend", "list_backends", "register_backend"]
