"""The shader hook ABI (renderer/shader.py).

A shader is a class of static hooks: vertex, primitive_chooser, interpolate,
fragment and mix.  The seven built-in shaders (jaxrenderer_amd.shaders) run as
fused HIP kernels selected by class identity.  Any other subclass runs its own
vertex / interpolate / fragment / mix hooks as torch code on the GPU around the
HIP visibility kernel (the default primitive_chooser); the hooks receive one
vertex or one fragment at a time and are vectorised with torch.func.vmap,
like jax.vmap in the reference.
"""

from __future__ import annotations

from typing import Any, NamedTuple, Tuple

import torch

from .geometry import Interpolation, interpolate

ID = Any


class PerVertex(NamedTuple):
    """renderer/shader.py:42-51: clip-space position of one vertex."""

    gl_Position: Any


class PerFragment(NamedTuple):
    """renderer/shader.py:54-64."""

    gl_FragDepth: Any = float("inf")
    keeps: Any = True
    use_default_depth: Any = False


class MixerOutput(NamedTuple):
    """renderer/shader.py:80-88."""

    keep: Any
    zbuffer: Any


def _t(x, like: torch.Tensor) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x
    return torch.as_tensor(x, device=like.device)


class Shader:
    """renderer/shader.py:91-396.  Subclass and override the hooks."""

    #: id of the fused HIP kernel implementing this class (built-ins only)
    _jr_shader_id = None

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra) -> Tuple[PerVertex, tuple]:
        raise NotImplementedError("vertex shader not implemented")

    @staticmethod
    def primitive_chooser(gl_FragCoord, gl_FrontFacing, gl_PointCoord, keeps, values,
                          barycentric_screen, barycentric_clip):
        """Default: the kept, front-facing primitive with the smallest window
        depth, first index on ties; index 0 when none qualifies
        (renderer/shader.py:207-251).  Operates on [primitives, ...] tensors."""
        inf = torch.full_like(gl_FragCoord[:, 2], float("inf"))
        depths = torch.where(keeps & gl_FrontFacing, gl_FragCoord[:, 2], inf)
        idx = torch.argmin(depths).reshape(1)
        take = lambda a: torch.index_select(a, 0, idx)  # noqa: E731
        return (
            take(gl_FragCoord),
            take(gl_FrontFacing),
            take(gl_PointCoord),
            take(keeps),
            tuple(take(v) for v in values),
            take(barycentric_screen),
            take(barycentric_clip),
        )

    @staticmethod
    def interpolate(values, barycentric_screen, barycentric_clip):
        """Default: every field SMOOTH (perspective-correct) interpolated."""
        out = tuple(interpolate(v, barycentric_screen, barycentric_clip, Interpolation.SMOOTH)
                    for v in values)
        return type(values)(*out) if hasattr(values, "_fields") else out

    @staticmethod
    def fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra):
        """Default: keep the fragment and use the default depth."""
        return PerFragment(use_default_depth=True), varying

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        """Default: the kept fragment of smallest depth (renderer/shader.py:374-396)."""
        inf = torch.full_like(gl_FragDepth, float("inf"))
        depths = torch.where(keeps, gl_FragDepth, inf)
        idx = torch.argmin(depths).reshape(1)
        pick = lambda a: torch.index_select(a, 0, idx).squeeze(0)  # noqa: E731
        picked = tuple(pick(v) for v in extra)
        return MixerOutput(keep=pick(keeps), zbuffer=pick(depths)), (
            type(extra)(*picked) if hasattr(extra, "_fields") else picked)


__all__ = ["ID", "MixerOutput", "PerFragment", "PerVertex", "Shader"]
