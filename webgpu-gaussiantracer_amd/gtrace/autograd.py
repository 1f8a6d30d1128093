"""torch.autograd wrapper of the forward renderer and its hand-written adjoint.

    rgba = render(raw, u, W, H, alpha_min, mode="checkpoint"|"recompute")

raw is the (n, 59) PLY-raw parameter tensor (mean, f_dc, f_rest, opacity
logit, log-scale, raw quaternion); rgba is the (H, W, 4) device image of
gt_render.  The backward kernel returns dLoss/draw in the same layout.  The
scene (proxies + BVH) is built from raw's values at each forward call (on the
GPU when raw is a CUDA tensor, gt_scene_create_device), so the hit lists the
adjoint replays are those of this forward.
"""
import torch

from . import SENTINEL, Scene, backward as _backward

CHECKPOINT_CAP = 256  # initial composited ids stored per ray; grows when a ray composites more


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, raw, u, width, height, alpha_min, mode):
        dev = torch.device("cuda", torch.cuda.current_device())
        stream = torch.cuda.current_stream().cuda_stream
        if raw.is_cuda:  # built on the GPU from the device parameters
            r = raw.detach().to(torch.float32).contiguous()
            scene = Scene(raw_dev=r.data_ptr(), n=r.shape[0], alpha_min=alpha_min, stream_ptr=stream)
        else:
            scene = Scene(raw.detach().to("cpu", torch.float32).numpy(), alpha_min=alpha_min)
        rgba = torch.empty((height, width, 4), dtype=torch.float32, device=dev)
        if mode == "checkpoint":
            cap = CHECKPOINT_CAP
            ncomp = torch.empty(height * width, dtype=torch.int32, device=dev)
            while True:
                cids = torch.empty((height * width, cap), dtype=torch.int32, device=dev)
                scene.render_device(u, width, height, rgba.data_ptr(), stream, comp_ids_ptr=cids.data_ptr(),
                                    n_comp_ptr=ncomp.data_ptr(), comp_cap=cap)
                most = int(ncomp.max().item()) if ncomp.numel() else 0
                if most <= cap:
                    break
                cap = 1 << (most - 1).bit_length()  # a ray composited more: render again with room for it
            ctx.ckpt = (cids, ncomp, cap)
        elif mode == "recompute":
            scene.render_device(u, width, height, rgba.data_ptr(), stream)
            ctx.ckpt = None
        else:
            raise ValueError(f"mode {mode!r}")
        ctx.scene, ctx.u, ctx.size, ctx.n = scene, u, (width, height), raw.shape[0]
        ctx.raw_device, ctx.raw_dtype = raw.device, raw.dtype
        return rgba

    @staticmethod
    def backward(ctx, grad_rgba):
        width, height = ctx.size
        dev = grad_rgba.device
        g = grad_rgba.contiguous().to(torch.float32)
        grads = torch.zeros((ctx.n, 59), dtype=torch.float32, device=dev)
        overflow = torch.zeros(1, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream().cuda_stream
        if ctx.ckpt is not None:
            cids, ncomp, cap = ctx.ckpt
            _backward(ctx.scene, ctx.u, width, height, g.data_ptr(), grads.data_ptr(), cids.data_ptr(),
                      ncomp.data_ptr(), cap, overflow.data_ptr(), 0, stream)
        else:
            _backward(ctx.scene, ctx.u, width, height, g.data_ptr(), grads.data_ptr(), None, None,
                      CHECKPOINT_CAP, overflow.data_ptr(), 0, stream)
        if int(overflow.item()) != 0:
            raise RuntimeError(f"{int(overflow.item())} rays overflowed the backward checkpoint")
        return grads.to(ctx.raw_device, ctx.raw_dtype), None, None, None, None, None


def render(raw, u, width, height, alpha_min, mode="checkpoint"):
    return RenderFunction.apply(raw, u, width, height, alpha_min, mode)


__all__ = ["render", "RenderFunction", "CHECKPOINT_CAP", "SENTINEL"]
