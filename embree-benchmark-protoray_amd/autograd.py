"""torch.autograd.Function over the C ABI: per-triangle RGB albedo -> image.

forward  = prt_hip_render_albedo_image (mean image over spp passes)
backward = prt_hip_adjoint_albedo (path replay; the gradient is accumulated
           with atomics into a zeroed tensor)

Tensors stay on the GPU; the library runs on its own HIP stream, so torch's
current stream is synchronised before each call and the library's stream
is synchronised inside each call.  This module imports torch; the core
binding (the package __init__) does not.
"""
from __future__ import annotations

import torch


class AlbedoRender(torch.autograd.Function):
    """image = render(albedo); albedo: (T, 3) float32 on the device."""

    @staticmethod
    def forward(ctx, albedo: torch.Tensor, dev, pass0: int, spp: int, max_depth: int) -> torch.Tensor:
        if not albedo.is_cuda or albedo.dtype != torch.float32 or albedo.shape != (dev.tri_count, 3):
            raise ValueError("albedo must be a (T, 3) float32 tensor on the GPU")
        albedo = albedo.contiguous()
        image = torch.empty((dev.height, dev.width, 3), device=albedo.device, dtype=torch.float32)
        torch.cuda.current_stream().synchronize()
        dev.set_albedo_device(albedo.data_ptr())
        dev.render_albedo_image_device(pass0, spp, max_depth, image.data_ptr())
        ctx.dev, ctx.pass0, ctx.spp, ctx.max_depth = dev, pass0, spp, max_depth
        ctx.save_for_backward(albedo)
        return image

    @staticmethod
    def backward(ctx, grad_image: torch.Tensor):
        (albedo,) = ctx.saved_tensors
        grad_image = grad_image.contiguous().to(torch.float32)
        grad = torch.zeros_like(albedo)
        torch.cuda.current_stream().synchronize()
        ctx.dev.set_albedo_device(albedo.data_ptr())
        ctx.dev.adjoint_albedo_device(ctx.pass0, ctx.spp, ctx.max_depth, grad_image.data_ptr(), grad.data_ptr())
        return grad, None, None, None, None


def render_albedo(albedo: torch.Tensor, dev, pass0: int = 0, spp: int = 1, max_depth: int = 6) -> torch.Tensor:
    return AlbedoRender.apply(albedo, dev, pass0, spp, max_depth)
