"""A training step over the PLY-raw Gaussian parameters, all on the GPU.

The reference renders only and plans training (README.md:68-69).  This is the
loop the library's pieces make possible, with the parameters resident in HBM:

    for each view:  gt_render with the per-ray checkpoint of composited ids
                    dL/dRGBA = 2 (render - target)      (loss = sum of squared RGB error)
                    gt_backward (accumulates dL/draw, n x 59)
    gt_adam_step    per-field learning rates
    gt_scene_refit_device    records, proxies and the tree's boxes from the
                    updated parameters (topology kept; a GPU rebuild when the
                    set of proxies changed), or gt_scene_update_device, a
                    full GPU rebuild

Which hits a ray finds does not depend on the tree, so both updates give the
same renders and gradients bit for bit.  They differ in speed: the refit is
cheaper than a rebuild and keeps the tree of the first build, by default the
host's binned-SAH tree, which renders up to 2 % faster than the GPU's
tree (a Morton radix tree with clustered subtrees).

Default learning rates are those of the 3D Gaussian Splatting reference
optimiser (Kerbl et al. 2023): means 1.6e-4 x spatial scale, f_dc 2.5e-3,
f_rest 2.5e-3 / 20, opacity logit 5e-2, log-scale 5e-3, quaternion 1e-3,
eps 1e-15.
"""
import numpy as np
import torch

from . import GT_RAW_FLOATS, Scene, adam_step, backward

FIELDS = {"mean": slice(0, 3), "f_dc": slice(3, 6), "f_rest": slice(6, 51), "opacity": slice(51, 52),
          "scale": slice(52, 55), "rotation": slice(55, 59)}


def default_lr(spatial_scale=1.0):
    """Per-field learning rates (59 floats) of the 3DGS optimiser."""
    lr = np.zeros(GT_RAW_FLOATS, np.float32)
    lr[FIELDS["mean"]] = 1.6e-4 * spatial_scale
    lr[FIELDS["f_dc"]] = 2.5e-3
    lr[FIELDS["f_rest"]] = 2.5e-3 / 20.0
    lr[FIELDS["opacity"]] = 5e-2
    lr[FIELDS["scale"]] = 5e-3
    lr[FIELDS["rotation"]] = 1e-3
    return lr


class Trainer:
    """Fits raw (a CUDA float32 tensor (n, 59), updated in place) to target
    images.  views: list of (uniforms, width, height, target) with target a
    CUDA float32 (height, width, 4) tensor (its alpha channel is ignored)."""

    def __init__(self, raw, views, alpha_min=0.01, lr=None, betas=(0.9, 0.999), eps=1e-15, stream=None,
                 build="host", refit=True):
        if not (raw.is_cuda and raw.dtype == torch.float32 and raw.is_contiguous() and raw.dim() == 2
                and raw.shape[1] == GT_RAW_FLOATS):
            raise ValueError("raw must be a contiguous CUDA float32 tensor of shape (n, 59)")
        self.raw = raw
        self.n = raw.shape[0]
        self.views = list(views)
        self.alpha_min = float(alpha_min)
        self.lr = default_lr() if lr is None else np.broadcast_to(np.asarray(lr, np.float32), (GT_RAW_FLOATS,))
        self.betas, self.eps = betas, eps
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        sp = self.stream.cuda_stream
        if build not in ("host", "gpu"):
            raise ValueError("build must be 'host' (binned SAH) or 'gpu' (radix tree with PLOC subtrees)")
        if build == "host" and not refit:
            raise ValueError("a host-built tree is kept by refit; rebuilds (refit=False) need build='gpu'")
        self.refit = bool(refit)
        self.rebuilds = 0
        if build == "host":
            self.stream.synchronize()
            self.scene = Scene(raw.detach().cpu().numpy(), alpha_min=self.alpha_min)
        else:
            self.scene = Scene(raw_dev=raw.data_ptr(), n=self.n, alpha_min=self.alpha_min, stream_ptr=sp)
        self.m = torch.zeros_like(raw)
        self.v = torch.zeros_like(raw)
        self.grads = torch.zeros_like(raw)
        self.step_count = 0
        self._cap = {}
        self._bufs = {}
        self.overflow = torch.zeros(1, dtype=torch.int32, device=raw.device)

    def _buffers(self, W, H):
        key = (W, H)
        if key not in self._bufs:
            dev = self.raw.device
            self._bufs[key] = (torch.empty((H, W, 4), dtype=torch.float32, device=dev),
                               torch.empty((H, W, 4), dtype=torch.float32, device=dev),
                               torch.empty(W * H, dtype=torch.int32, device=dev))
            self._cap[key] = (64, None)
        return self._bufs[key]

    def _render_checkpoint(self, u, W, H, rgba, ncomp):
        """Render with the composited-id checkpoint, growing its capacity when a ray needs more."""
        sp = self.stream.cuda_stream
        key = (W, H)
        cap, cids = self._cap[key]
        while True:
            if cids is None:
                cids = torch.empty((W * H, cap), dtype=torch.int32, device=self.raw.device)
            self.scene.render_device(u, W, H, rgba.data_ptr(), sp, comp_ids_ptr=cids.data_ptr(),
                                     n_comp_ptr=ncomp.data_ptr(), comp_cap=cap)
            most = int(ncomp.max().item())
            if most <= cap:
                break
            cap, cids = 1 << (most - 1).bit_length(), None
        self._cap[key] = (cap, cids)
        return cids, cap

    def step(self):
        """One training step over all views; returns the loss before the update."""
        sp = self.stream.cuda_stream
        with torch.cuda.stream(self.stream):
            self.grads.zero_()
            self.overflow.zero_()
            loss = torch.zeros((), dtype=torch.float64, device=self.raw.device)
            for u, W, H, target in self.views:
                rgba, gimg, ncomp = self._buffers(W, H)
                cids, cap = self._render_checkpoint(u, W, H, rgba, ncomp)
                torch.sub(rgba, target, out=gimg)
                gimg[..., 3] = 0.0
                loss += (gimg.double() ** 2).sum()
                gimg.mul_(2.0)
                backward(self.scene, u, W, H, gimg.data_ptr(), self.grads.data_ptr(), cids.data_ptr(),
                         ncomp.data_ptr(), cap, self.overflow.data_ptr(), 0, sp)
            self.step_count += 1
            adam_step(self.raw.data_ptr(), self.grads.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self.n,
                      self.lr, self.step_count, self.betas[0], self.betas[1], self.eps, sp)
            if self.refit:
                self.rebuilds += int(self.scene.refit_device(self.raw.data_ptr(), sp))
            else:
                self.scene.update_device(self.raw.data_ptr(), sp)
        if int(self.overflow.item()) != 0:
            raise RuntimeError("backward checkpoint overflowed")
        return float(loss.item())

    def close(self):
        self.scene.close()


__all__ = ["Trainer", "default_lr", "FIELDS"]
