"""Backend registry: string name -> ForceBackend subclass (aimd/backends/registry.py:21-39)."""

from __future__ import annotations

_REGISTRY: dict[str, type] = {}


def register_backend(cls: type) -> type:
    key = str(cls.name).strip().lower()
    if not key:
        raise ValueError(f"{cls.__name__} must define a non-empty 'name'")
    _REGISTRY[key] = cls
    return cls


def get_backend(name: str) -> type:
    key = str(name).strip().lower()
    if key not in _REGISTRY:
        raise ValueError(f"Unknown backend '{name}'. Available backends: {list_backends()}")
    return _REGISTRY[key]


def list_backends() -> list[str]:
    return sorted(_REGISTRY)
