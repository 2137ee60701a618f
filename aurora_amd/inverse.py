"""Inverse rendering on the GPU path tracer.

``RenderFunction`` is a ``torch.autograd.Function``: forward renders the mean
image for per-triangle diffuse albedo, per-material roughness and per-material
Kelemen specular weight (``aur_render``); backward runs the path-replay
adjoint (``aur_adjoint``) with the same random numbers, reusing the radiance
records the forward render kept (``aur_keep_path_records``) instead of
replaying each path a first time.  ``make_problem`` and
``optimize`` set up and run the inverse-rendering workload on the box room:
512x512, 32 spp forward and adjoint, depth 6, ground truth albedo
~U[0.1, 0.9]^3 per triangle, roughness ~U[0.05, 0.8] and spec weight ~U[0, 1]
per material, optimised from (0.5, 0.4, 0.5) with Adam on an L2 image loss.
``TextureRenderFunction`` (``make_texture_function``) does the same for an
FP32 albedo texture (``aur_set_albedo_texture``), the gradient target of the
large differentiable scene.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import scenes
from .renderer import Renderer


def _torch():
    import torch
    return torch


class _Ctx:
    renderer: Renderer = None
    passes: int = 1
    first_pass: int = 0


def render_image(r: Renderer, albedo, rough, spec, first_pass: int, passes: int) -> np.ndarray:
    r.set_tri_albedo(np.asarray(albedo, np.float32))
    r.set_material_params(np.asarray(rough, np.float32), np.asarray(spec, np.float32))
    acc, _ = r.render_passes(first_pass, passes)
    return acc[..., :3] / passes


def make_function():
    torch = _torch()

    class RenderFunction(torch.autograd.Function):
        """image = render(albedo (n_tris, 3), roughness (n_attrs,), spec weight (n_attrs,))"""

        @staticmethod
        def forward(ctx, albedo, rough, spec, renderer, first_pass, passes):
            # the render keeps its per-path records for the backward replay
            renderer.keep_path_records(first_pass, passes)
            img = render_image(renderer, albedo.detach().cpu().numpy(), rough.detach().cpu().numpy(),
                               spec.detach().cpu().numpy(), first_pass, passes)
            ctx.renderer, ctx.first_pass, ctx.passes = renderer, first_pass, passes
            ctx.dev = albedo.device
            return torch.from_numpy(img).to(albedo.device)

        @staticmethod
        def backward(ctx, grad_img):
            g = grad_img.detach().cpu().numpy().astype(np.float32)
            ga, gr, gs, _ = ctx.renderer.adjoint(ctx.first_pass, ctx.passes, g, albedo=True)
            dev = ctx.dev
            return (torch.from_numpy(ga).to(dev), torch.from_numpy(gr).to(dev), torch.from_numpy(gs).to(dev),
                    None, None, None)

    return RenderFunction


def make_texture_function():
    """image = render(albedo texture (H, W, 3)) for a renderer, the materials
    the texture drives, and a pass range: forward through
    aur_set_albedo_texture + aur_render, backward through aur_adjoint's
    texture gradient (the same seed and texture)."""
    torch = _torch()

    class TextureRenderFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, tex, renderer, attrs, first_pass, passes):
            renderer.keep_path_records(first_pass, passes)
            renderer.set_albedo_texture(tex.detach().cpu().numpy().astype(np.float32), attrs)
            acc, _ = renderer.render_passes(first_pass, passes)
            ctx.renderer, ctx.first_pass, ctx.passes = renderer, first_pass, passes
            ctx.dev = tex.device
            return torch.from_numpy(acc[..., :3] / passes).to(tex.device)

        @staticmethod
        def backward(ctx, grad_img):
            g = grad_img.detach().cpu().numpy().astype(np.float32)
            *_, gt = ctx.renderer.adjoint(ctx.first_pass, ctx.passes, g, albedo=False, texture=True)
            return torch.from_numpy(gt).to(ctx.dev), None, None, None, None

    return TextureRenderFunction


@dataclass
class Problem:
    renderer: Renderer
    truth: dict
    init: dict
    target: np.ndarray
    kelemen: np.ndarray   # mask of optimisable materials
    spp: int


def make_problem(out_dir: str, res: int = 512, spp: int = 32, depth: int = 6, seed: int = 4, device: int = 0) -> Problem:
    path = scenes.write_config("box_room_kelemen", out_dir, res=res, spp=spp, maxdepth=depth, mindepth=depth)
    r = Renderer(path, os.path.join(out_dir, "config.json"), device=device)
    r.build()
    info = r.info()
    rng = scenes.SplitMix64(seed)
    nt, nm = info["n_tris"], info["n_attrs"]
    albedo = np.array([[rng.uniform(0.1, 0.9) for _ in range(3)] for _ in range(nt)], np.float32)
    rough = np.array([rng.uniform(0.05, 0.8) for _ in range(nm)], np.float32)
    spec = np.array([rng.uniform(0.0, 1.0) for _ in range(nm)], np.float32)
    kel = np.array([m.type == 2 for m in r.export()["materials"]])
    truth = {"albedo": albedo, "rough": rough, "spec": np.maximum(spec, 1e-3)}
    init = {"albedo": np.full((nt, 3), 0.5, np.float32), "rough": np.full(nm, 0.4, np.float32),
            "spec": np.full(nm, 0.5, np.float32)}
    r.set_seed(seed * 7919 + 1)
    target = render_image(r, truth["albedo"], truth["rough"], truth["spec"], 1 << 20, spp)
    return Problem(r, truth, init, target, kel, spp)


def optimize(p: Problem, steps: int = 200, lr: float = 0.02, log_every: int = 0):
    """Adam on the L2 image loss; returns the loss history and final parameters."""
    torch = _torch()
    fn = make_function()
    albedo = torch.tensor(p.init["albedo"], requires_grad=True)
    rough = torch.tensor(p.init["rough"], requires_grad=True)
    spec = torch.tensor(p.init["spec"], requires_grad=True)
    target = torch.from_numpy(p.target)
    opt = torch.optim.Adam([albedo, rough, spec], lr=lr)
    hist = []
    for it in range(steps):
        opt.zero_grad()
        p.renderer.set_seed(1000 + it)
        img = fn.apply(albedo, rough, spec, p.renderer, 0, p.spp)
        loss = torch.mean((img - target) ** 2)
        loss.backward()
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.01, 1.0)
            rough.clamp_(0.02, 0.99)
            spec.clamp_(0.001, 1.0)
        hist.append(float(loss.detach()))
        if log_every and it % log_every == 0:
            print("step %d loss %.6g" % (it, hist[-1]), flush=True)
    return hist, {"albedo": albedo.detach().numpy(), "rough": rough.detach().numpy(), "spec": spec.detach().numpy()}
