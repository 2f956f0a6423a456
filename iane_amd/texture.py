"""Texture images: the reference's ``texture_image`` / ``texture_image_view``
(include/iane/resource/cuda/texture_image.hpp:13-176,
include/iane/view/texture_image_view.hpp:20-98, docs/cuda_texture_images.md).

gfx950 offers no image/texture sampling to HIP, so a texture image owns a
packed device tensor of texels, and a texture view is a read-only view of it
with the reference's lookup rules: an integer lookup returns the texel without
interpolation or normalisation, coordinates are local to the view, and a
subview keeps the storage and moves its origin (a crop is an independent image
for the border rule of ``convolve``). Shapes, corners and sizes are given in
tensor order (last dimension = IANE axis 0), like every image in this package.
"""
from __future__ import annotations

from . import _lib

TEXEL_DTYPES = ("uint8", "uint16", "int32", "float32")


class texture_image_view:
    """Read-only view of texture texels (texture_image_view.hpp:20-98)."""

    def __init__(self, texels):
        self._texels = texels

    @property
    def shape(self):
        return tuple(self._texels.shape)

    @property
    def dtype(self):
        return self._texels.dtype

    def dimensions(self):
        """Logical extent in IANE order (axis 0 first)."""
        return tuple(reversed(self._texels.shape))

    def subview(self, corner, size):
        """In-bounds region whose coordinates start at zero (texture_image_view.hpp:74-80)."""
        corner, size = tuple(corner), tuple(size)
        if len(corner) != self._texels.dim() or len(size) != self._texels.dim():
            raise _lib.invalid_argument("subview requires a corner and size of the view's rank")
        index = []
        for c, s, n in zip(corner, size, self._texels.shape):
            if c < 0 or s < 0 or c + s > n:
                raise _lib.invalid_argument("subview must lie inside the view")
            index.append(slice(c, c + s))
        return texture_image_view(self._texels[tuple(index)])

    def texels(self):
        """The texels of this view as a device tensor view (this backend's storage)."""
        return self._texels


class texture_image:
    """Owner of a 1-, 2- or 3-D texture image (texture_image.hpp:13-98).

    ``shape`` is in tensor order; an empty extent allocates nothing.
    Raises ``invalid_argument`` for negative extents and ``unsupported`` for
    ranks outside 1..3 or element types other than u8, u16, i32 and f32.
    """

    def __init__(self, shape, dtype=None, device="cuda"):
        import torch
        dtype = torch.float32 if dtype is None else dtype
        shape = tuple(int(n) for n in shape)
        if not 1 <= len(shape) <= 3:
            raise _lib.unsupported("texture rank must be 1..3")
        if any(n < 0 for n in shape):
            raise _lib.invalid_argument("Image dimensions must be nonnegative")
        if str(dtype).replace("torch.", "") not in TEXEL_DTYPES:
            raise _lib.unsupported("texture images hold u8, u16, i32 or f32 texels")
        self._shape = shape
        self._texels = torch.empty(shape, dtype=dtype, device=device) if all(shape) else \
            torch.empty(shape, dtype=dtype)

    def dimensions(self):
        return tuple(reversed(self._shape))

    def const_view(self) -> texture_image_view:
        return texture_image_view(self._texels)

    def array(self):
        """Device storage of the texels (the interoperability handle); None when empty."""
        return self._texels if all(self._shape) else None

    def upload(self, source, policy=None):
        """copy(source, texture, policy), texture_image.hpp:100-176: a host or device
        tensor of the same shape and element type with contiguous rows; rows and
        3-D slices may be padded."""
        from . import _stream, hip_execution, view_desc
        if tuple(source.shape) != self._shape:
            raise _lib.invalid_argument("Texture upload requires matching dimensions")
        if source.dtype != self._texels.dtype:
            raise _lib.invalid_argument("Texture upload requires the texel element type")
        if not all(self._shape):
            # nothing to transfer: the handle refers to the policy's stream (the null stream by default)
            return hip_execution(0 if policy is None else _stream(policy))
        stream = _stream(policy)
        esize = source.element_size()
        rank = len(self._shape)
        row_bytes = self._shape[-1] * esize
        if source.stride(-1) != 1:
            raise _lib.invalid_argument("Texture upload requires contiguous rows")
        if rank >= 2:
            pitch = source.stride(-2) * esize
            if pitch < row_bytes:
                raise _lib.invalid_argument("Texture upload requires a valid row pitch")
        if rank == 3:
            slice_bytes = source.stride(-3) * esize
            if slice_bytes < row_bytes or (slice_bytes - row_bytes) // pitch < self._shape[-2] - 1:
                raise _lib.invalid_argument("Texture upload requires a valid slice stride")
        kind = 3 if source.device.type == "cuda" else 1
        _lib.call("iane_hip_copy_view", view_desc(source, require_device=False), view_desc(self._texels), kind,
                  stream)
        return hip_execution(stream)


def copy(source, destination, policy=None):
    """copy(source, texture) uploads; copy(texture view, tensor) materialises the view."""
    from . import _stream, hip_execution, view_desc
    if isinstance(destination, texture_image):
        return destination.upload(source, policy)
    if isinstance(source, texture_image_view):
        stream = _stream(policy)
        texels = source.texels()
        if tuple(texels.shape) != tuple(destination.shape) or texels.dtype != destination.dtype:
            raise _lib.invalid_argument("copy requires views of equal extents and element type")
        if texels.numel():
            _lib.call("iane_hip_copy_view", view_desc(texels), view_desc(destination), 3, stream)
        return hip_execution(stream)
    raise _lib.invalid_argument("copy: expected a texture image destination or a texture view source")
