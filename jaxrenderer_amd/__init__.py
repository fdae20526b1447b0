"""MI355X-native rasteriser with jaxrenderer's API (the `renderer` package).

    import jaxrenderer_amd as renderer

The rasterisation pipeline runs as hand-written HIP kernels for gfx950
(libjrhip.so, loaded through ctypes); everything above it -- cameras, model
merging, the Shader hook classes -- mirrors the reference's Python API.
"""

from .geometry import (
    Camera,
    normalise,
    quaternion,
    quaternion_mul,
    rotation_matrix,
    transform_matrix_from_rotation,
)
from .model import Model, ModelObject, batch_models, merge_objects
from .pipeline import render
from .renderer import CameraParameters, LightParameters, Renderer, ShadowParameters
from .shader import MixerOutput, PerFragment, PerVertex, Shader
from .shadow import Shadow
from .shapes.capsule import UpAxis, create_capsule
from .shapes.cube import create_cube
from .types import Buffers, LightSource
from .utils import build_texture_from_PyTinyrenderer, merge_canvases, transpose_for_display

__all__ = [
    "batch_models",
    "Buffers",
    "build_texture_from_PyTinyrenderer",
    "Camera",
    "CameraParameters",
    "create_capsule",
    "create_cube",
    "LightParameters",
    "LightSource",
    "merge_canvases",
    "merge_objects",
    "MixerOutput",
    "Model",
    "ModelObject",
    "normalise",
    "PerFragment",
    "PerVertex",
    "quaternion",
    "quaternion_mul",
    "render",
    "Renderer",
    "rotation_matrix",
    "Shader",
    "Shadow",
    "ShadowParameters",
    "transform_matrix_from_rotation",
    "transpose_for_display",
    "UpAxis",
]
