"""Layout helpers (renderer/utils.py), for torch tensors and numpy arrays."""

from __future__ import annotations

import numpy as np
import torch


def transpose_for_display(matrix, flip_vertical: bool = True):
    """[W, H, ...] -> [H, W, ...], rows flipped by default
    (renderer/utils.py:72-96)."""
    if isinstance(matrix, torch.Tensor):
        mat = matrix.transpose(0, 1)
        return torch.flip(mat, dims=(0,)) if flip_vertical else mat
    mat = np.swapaxes(np.asarray(matrix), 0, 1)
    return mat[::-1, ...] if flip_vertical else mat


def build_texture_from_PyTinyrenderer(texture, width: int, height: int):
    """renderer/utils.py:99-125: [H][W][C] C-order flat -> [W, H, C] with the
    second axis reversed."""
    t = np.reshape(np.asarray(texture), (width, height, -1), order="C")
    return np.swapaxes(t, 0, 1)[:, ::-1, :].astype(np.float32)


def merge_canvases(zbuffers, canvases):
    """renderer/utils.py:36-69: per pixel, the layer with the largest z."""
    z = torch.as_tensor(zbuffers)
    c = torch.as_tensor(canvases)
    idx = torch.argmax(z, dim=0)
    zb = torch.gather(z, 0, idx[None]).squeeze(0)
    cv = torch.gather(c, 0, idx[None, ..., None].expand((1,) + tuple(c.shape[1:]))).squeeze(0)
    return zb, cv


__all__ = ["build_texture_from_PyTinyrenderer", "merge_canvases", "transpose_for_display"]
