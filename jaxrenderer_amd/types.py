"""Containers and layout conventions (renderer/types.py).

Buffers are x-major: zbuffer [width, height], canvas [width, height, channel]
(renderer/types.py:63-64).  Arrays are torch tensors on the GPU.
"""

from __future__ import annotations

from typing import Any, Generic, NamedTuple, Tuple, TypeVar

import numpy as np


class LightSource(NamedTuple):
    """renderer/types.py:134-143."""

    direction: Any = np.array((0.0, 0.0, -1.0), dtype=np.float32)
    colour: Any = np.ones(3, dtype=np.float32)


_TargetsT = TypeVar("_TargetsT", bound=Tuple[Any, ...])


class Buffers(NamedTuple):
    """renderer/types.py:148-155: zbuffer [W, H] and target buffers [W, H, ...]."""

    zbuffer: Any
    targets: Tuple[Any, ...]


class DtypeInfo(NamedTuple):
    """renderer/types.py:98-131."""

    min: Any
    max: Any
    bits: int
    dtype: type

    @classmethod
    def create(cls, dtype) -> "DtypeInfo":
        dt = np.dtype(str(dtype).replace("torch.", ""))
        if np.issubdtype(dt, np.floating):
            fi = np.finfo(dt)
            return cls(min=fi.min, max=fi.max, bits=fi.bits, dtype=dt.type)
        if np.issubdtype(dt, np.integer):
            ii = np.iinfo(dt)
            return cls(min=ii.min, max=ii.max, bits=ii.bits, dtype=dt.type)
        raise ValueError(f"Unexpected dtype {dtype}")


__all__ = ["Buffers", "DtypeInfo", "LightSource"]
