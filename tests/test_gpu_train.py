"""GPU tests of the training step: gt_adam_step and the Trainer loop
(render with checkpoint -> dL/dRGBA -> gt_backward -> Adam -> GPU rebuild).

Bar: Adam within FP32 rounding of an FP64 evaluation of the same update
(and of torch.optim.Adam); a short fit lowers the loss and keeps every
parameter finite; the rebuilt scene equals a fresh build from the updated
parameters (hit lists and pixels bit-exact).
"""
import numpy as np
import pytest
import torch

from helpers import uniforms
import gtrace as gt
from gtrace.train import FIELDS, Trainer, default_lr

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _adam64(p, g, m, v, lr, step, b1, b2, eps):
    m = m + (1 - b1) * (g - m)
    v = b2 * v + (1 - b2) * g * g
    bc1, bc2 = 1 - b1 ** step, 1 - b2 ** step
    p = p - (lr / bc1) * m / (np.sqrt(v) / np.sqrt(bc2) + eps)
    return p, m, v


def test_adam_matches_fp64_and_torch():
    _print("case adam vs fp64 and torch.optim.Adam")
    rng = np.random.default_rng(0)
    n = 3001
    dev = torch.device("cuda", 0)
    p0 = rng.standard_normal((n, 59)).astype(np.float32)
    lr = default_lr(2.0)
    p, m, v = (torch.from_numpy(p0).to(dev), torch.zeros((n, 59), device=dev), torch.zeros((n, 59), device=dev))
    P, M, V = p0.astype(np.float64), np.zeros((n, 59)), np.zeros((n, 59))
    grads = [rng.standard_normal((n, 59)).astype(np.float32) * s for s in (1.0, 1e-3, 10.0, 0.5, 1e-6)]
    for step, g in enumerate(grads, 1):
        gd = torch.from_numpy(g).to(dev)
        gt.adam_step(p.data_ptr(), gd.data_ptr(), m.data_ptr(), v.data_ptr(), n, lr, step, 0.9, 0.999, 1e-15)
        P, M, V = _adam64(P, g.astype(np.float64), M, V, lr.astype(np.float64), step, 0.9, 0.999, 1e-15)
    torch.cuda.synchronize()
    got = p.cpu().numpy().astype(np.float64)
    # FP32 rounding of the stored parameter (a few ulp) + 1e-4 of the displacement scale (lr per step)
    ulp = np.spacing(np.abs(p0).astype(np.float32) + np.float32(0.01)).astype(np.float64)
    tol = 4 * ulp + 1e-4 * lr.astype(np.float64)[None, :] * len(grads)
    err = np.abs(got - P)
    assert (err <= tol).all(), f"Adam differs from FP64 by {(err / tol).max()} x the tolerance"
    # torch.optim.Adam with one learning rate for every field
    q = torch.from_numpy(p0).to(dev).requires_grad_(True)
    opt = torch.optim.Adam([q], lr=1e-3, betas=(0.9, 0.999), eps=1e-15)
    p2, m2, v2 = torch.from_numpy(p0).to(dev), torch.zeros((n, 59), device=dev), torch.zeros((n, 59), device=dev)
    for step, g in enumerate(grads, 1):
        gd = torch.from_numpy(g).to(dev)
        q.grad = gd.clone()
        opt.step()
        gt.adam_step(p2.data_ptr(), gd.data_ptr(), m2.data_ptr(), v2.data_ptr(), n, 1e-3, step, 0.9, 0.999, 1e-15)
    torch.cuda.synchronize()
    d = (p2 - q.detach()).abs().cpu().numpy().astype(np.float64)
    tol = 4 * ulp + 1e-4 * 1e-3 * len(grads)
    assert (d <= tol).all(), f"differs from torch.optim.Adam by {(d / tol).max()} x the tolerance"


def test_adam_zero_lr_and_bad_arguments():
    dev = torch.device("cuda", 0)
    n = 100
    p0 = torch.randn((n, 59), device=dev)
    p, m, v = p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)
    g = torch.randn((n, 59), device=dev)
    lr = np.zeros(59, np.float32)
    lr[FIELDS["f_dc"]] = 1e-2
    gt.adam_step(p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), n, lr, 1)
    torch.cuda.synchronize()
    changed = (p != p0).any(dim=0).cpu().numpy()
    assert changed[3:6].all() and not changed[:3].any() and not changed[6:].any()
    for kw in (dict(step=0), dict(beta1=1.0), dict(eps=0.0)):
        args = dict(step=1, beta1=0.9, beta2=0.999, eps=1e-15)
        args.update(kw)
        with pytest.raises(gt.GTError):
            gt.adam_step(p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), n, 1e-3, args["step"],
                         args["beta1"], args["beta2"], args["eps"])
    with pytest.raises(gt.GTError):
        gt.adam_step(p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), n, -1.0, 1)
    with pytest.raises(gt.GTError):
        gt.adam_step(None, g.data_ptr(), m.data_ptr(), v.data_ptr(), n, 1e-3, 1)


def test_trainer_fits_perturbed_scene():
    """Targets from a scene with perturbed means and colours; training from the
    unperturbed parameters lowers the loss, and the trained scene equals a
    fresh build from the trained parameters."""
    _print("case trainer fit")
    dev = torch.device("cuda", 0)
    raw = gt.generate_raw(n=6000, seed=13, config="c4")
    rng = np.random.default_rng(4)
    raw_t = raw.copy()
    raw_t[:, 0:3] += rng.normal(0, 0.03, (raw.shape[0], 3)).astype(np.float32)
    raw_t[:, 3:6] += rng.normal(0, 0.3, (raw.shape[0], 3)).astype(np.float32)
    W, H = 64, 48
    poses = [[0.0, 0.3, 3.0], [2.0, 0.5, 2.0], [-2.5, -0.4, 1.5]]
    us = [uniforms(p, [0, 0, 0], W, H, hfov_deg=50, alpha_min=0.01, k=16, T_min=1e-3) for p in poses]
    ts = gt.Scene(raw_t, alpha_min=0.01)
    views = []
    for u in us:
        img = torch.from_numpy(ts.render(u, W, H)["rgba"]).to(dev)
        views.append((u, W, H, img))
    x = torch.from_numpy(raw).to(dev)
    lr = default_lr()
    lr[FIELDS["mean"]] = 2e-3
    tr = Trainer(x, views, alpha_min=0.01, lr=lr)
    losses = [tr.step() for _ in range(30)]
    _print("  losses", [round(l, 3) for l in losses[:3]], "...", [round(l, 3) for l in losses[-3:]])
    assert np.all(np.isfinite(losses)) and torch.isfinite(x).all()
    assert min(losses[-5:]) < 0.6 * losses[0], losses
    # the scene the trainer holds is the build of the trained parameters
    fresh = gt.Scene(x.cpu().numpy(), alpha_min=0.01)
    a = tr.scene.render(us[1], W, H, max_rounds=64)
    b = fresh.render(us[1], W, H, max_rounds=64)
    np.testing.assert_array_equal(a["hit_ids"], b["hit_ids"])
    np.testing.assert_array_equal(a["rgba"].view(np.uint32), b["rgba"].view(np.uint32))
    tr.close()
