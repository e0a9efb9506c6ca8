"""Sampled-image descriptors: the stand-in for the reference's CUDA texture objects.

Python/texture.py:27-55 copies an array into a cudaArray and wraps it in a
bindless texture object with an address mode, a filter mode and normalised
coordinates; the kernels then sample it with tex2D.  On gfx950 the kernels
read the array directly and apply the same semantics in code (bilinear with
exact weights, half-sample symmetric mirror edges), so a "texture" here is a
descriptor around a device tensor.

``create_texture_object`` keeps the reference's signature and its copy
semantics (the descriptor holds a snapshot of the array).
"""

import torch

# Address / filter / read mode values of the CUDA runtime enums the reference
# passes (cudaAddressModeWrap=0, Clamp=1, Mirror=2, Border=3; cudaFilterModePoint=0,
# Linear=1; cudaReadModeElementType=0).
cudaAddressModeWrap = 0
cudaAddressModeClamp = 1
cudaAddressModeMirror = 2
cudaAddressModeBorder = 3
cudaFilterModePoint = 0
cudaFilterModeLinear = 1
cudaReadModeElementType = 0


class SampledImage:
    """A device array plus its sampling state.

    Attributes: ``tensor`` (H, W) or (H, W, 2) on the GPU; ``address_mode``,
    ``filter_mode``, ``normalized``.  ``shape`` is the array's (H, W).
    """

    __slots__ = ("tensor", "address_mode", "filter_mode", "read_mode", "normalized")

    def __init__(self, tensor, address_mode=cudaAddressModeMirror, filter_mode=cudaFilterModeLinear,
                 read_mode=cudaReadModeElementType, normalized=1):
        if not isinstance(tensor, torch.Tensor):
            raise TypeError("SampledImage wraps a torch tensor")
        if tensor.dtype not in (torch.uint8, torch.int16, torch.float32):
            raise TypeError(f"unsupported texture element type {tensor.dtype}")
        if tensor.dim() == 3 and tensor.shape[2] != 2:
            raise ValueError("3-D textures must have 2 channels (float2)")
        self.tensor = tensor.contiguous()
        self.address_mode = address_mode
        self.filter_mode = filter_mode
        self.read_mode = read_mode
        self.normalized = normalized

    @property
    def shape(self):
        return tuple(self.tensor.shape[:2])

    def require(self, address_mode=cudaAddressModeMirror, filter_mode=None, normalized=1):
        """Check that the kernels support this descriptor (mirror, normalised coordinates)."""
        if self.address_mode != address_mode:
            raise NotImplementedError(f"address mode {self.address_mode} is not supported by this kernel")
        if filter_mode is not None and self.filter_mode != filter_mode:
            raise NotImplementedError(f"filter mode {self.filter_mode} is not supported by this kernel")
        if self.normalized != normalized:
            raise NotImplementedError("only normalised coordinates are supported by this kernel")
        return self.tensor


def create_texture_object(img_gpu, addressMode=cudaAddressModeBorder, filterMode=cudaFilterModePoint,
                          readMode=cudaReadModeElementType, normalizedCoords=0):
    """Python/texture.py:27-55: snapshot ``img_gpu`` into a SampledImage.

    Returns None for element types the reference does not support, as it does.
    """
    if img_gpu.dtype not in (torch.uint8, torch.int16, torch.float32):
        return None
    if img_gpu.dim() == 3 and img_gpu.shape[2] != 2:
        raise AssertionError("3-D textures must have 2 channels")
    return SampledImage(img_gpu.clone(memory_format=torch.contiguous_format), addressMode, filterMode, readMode,
                        normalizedCoords)


def wrap(tensor, filter_mode=cudaFilterModeLinear):
    """Descriptor over ``tensor`` without a copy (Mirror, normalised), for arrays that are not modified later."""
    return SampledImage(tensor, cudaAddressModeMirror, filter_mode, cudaReadModeElementType, 1)
