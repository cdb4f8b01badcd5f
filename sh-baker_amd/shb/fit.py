"""Fitting harness: light intensities and material albedos from a target
lightmap, differentiating the GPU bake with the path-replay adjoint.

`BakeSH` is a torch.autograd.Function over the C ABI: forward sets the
per-light / per-material RGB scales and bakes (shb_bake), backward runs
shb_bake_adjoint with the incoming dL/dSH. PyTorch only carries the
parameters and the optimizer (CPU tensors); the bake and its adjoint run in
the HIP library. The library is loaded before torch is imported so both use
the same HIP runtime.
"""
from __future__ import annotations

import numpy as np

from . import get_lib

get_lib()

import torch  # noqa: E402  (after the HIP library, see the module docstring)


class BakeSH(torch.autograd.Function):
    """SH9 lightmap (texels, 27) as a function of light scales (lights, 3)
    and albedo scales (materials, 3); scales must be > 0 for the backward."""

    @staticmethod
    def forward(ctx, light_scale, albedo_scale, baker, samples, bounces, seed):
        baker.set_params(light_scale.detach().cpu().numpy().astype(np.float32),
                         albedo_scale.detach().cpu().numpy().astype(np.float32))
        sh, _, _ = baker.bake(samples, bounces, seed=seed)
        ctx.baker = baker
        ctx.cfg = (samples, bounces, seed)
        ctx.save_for_backward(light_scale, albedo_scale)
        return torch.from_numpy(sh)

    @staticmethod
    def backward(ctx, grad_sh):
        light_scale, albedo_scale = ctx.saved_tensors
        baker = ctx.baker
        samples, bounces, seed = ctx.cfg
        # the replay needs the forward's parameters (another forward may
        # have changed them since)
        baker.set_params(light_scale.detach().cpu().numpy().astype(np.float32),
                         albedo_scale.detach().cpu().numpy().astype(np.float32))
        g = grad_sh.detach().cpu().numpy().astype(np.float32)
        dl, dm, _ = baker.bake_adjoint(samples, bounces, g, seed=seed)
        return (torch.from_numpy(dl).to(light_scale.dtype),
                torch.from_numpy(dm).to(albedo_scale.dtype), None, None, None, None)


def lognormal_factors(n_lights, n_materials, sigma=0.3, seed=0):
    """Seeded per-channel lognormal factors (lights, 3), (materials, 3)."""
    rng = np.random.default_rng(seed)
    return (np.exp(rng.normal(0.0, sigma, (n_lights, 3))).astype(np.float32),
            np.exp(rng.normal(0.0, sigma, (n_materials, 3))).astype(np.float32))


def fit(baker, target_sh, samples, bounces, iters=50, lr=0.05, seed=12345,
        init_lights=None, init_albedos=None, callback=None):
    """Adam on log-scales so every scale stays > 0. L = 0.5 * sum((SH -
    target)^2) / texels. Returns (light_scale, albedo_scale, losses)."""
    nl, nm = len(baker.scene.lights), len(baker.scene.materials)
    ll = torch.zeros((nl, 3), dtype=torch.float32) if init_lights is None else \
        torch.log(torch.as_tensor(init_lights, dtype=torch.float32))
    la = torch.zeros((nm, 3), dtype=torch.float32) if init_albedos is None else \
        torch.log(torch.as_tensor(init_albedos, dtype=torch.float32))
    ll.requires_grad_(True)
    la.requires_grad_(True)
    opt = torch.optim.Adam([ll, la], lr=lr)
    target = torch.as_tensor(np.asarray(target_sh, np.float32))
    n = float(target.shape[0])
    losses = []
    for it in range(iters):
        opt.zero_grad()
        sh = BakeSH.apply(torch.exp(ll), torch.exp(la), baker, samples, bounces, seed)
        loss = 0.5 * ((sh - target) ** 2).sum() / n
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
        if callback:
            callback(it, losses[-1])
    return torch.exp(ll).detach().numpy(), torch.exp(la).detach().numpy(), losses
