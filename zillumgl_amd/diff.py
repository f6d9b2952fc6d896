"""Differentiable rendering through the C ABI: a torch.autograd.Function over the
path-replay adjoint of the naive path tracer.

The reference has no differentiable mode.  The estimator is the naive MIS path tracer
(src/shader/path_integ_naive.glsl:35-174) with detached sampling: the uploaded scene
fixes every sampling decision, and the parameters enter through BSDF values of
principled/Lambertian materials and the environment radiance.  Parameters:

  * material_params (n_materials, 5): baseColor r, g, b, raw roughness (before the
    loader's mix(0.0134, 1, r)), metallic -- record slots 0, 1, 2, 3, 5;
  * env_rgb (env_height, env_width, 3): FP32 environment texels.

The image is the mean over the rendered passes, (H, W, 3).
"""
import numpy as np
import torch

from . import render


def scene_parameters(packed):
    """(material_params, env_rgb) of a packed scene, the values the uploaded scene holds."""
    m = packed.materials
    params = np.stack([m[:, 0], m[:, 1], m[:, 2], m[:, 3], m[:, 5]], 1).astype(np.float32)
    env = packed.env_half.view(np.float16).astype(np.float32)
    return params, env


class Renderer:
    """A context with an uploaded scene and a fixed pass range."""

    def __init__(self, ctx: render.Context, packed, n_passes, max_depth, first_pass=0, free_counter=1):
        self.ctx = ctx
        self.n_materials = packed.materials.shape[0]
        self.env_shape = (packed.env_height, packed.env_width)
        self.n_passes = n_passes
        self.max_depth = max_depth
        self.first_pass = first_pass
        self.free_counter = free_counter

    def image(self, material_params, env_rgb):
        """Mean image of the passes at the given evaluation point (numpy, (H, W, 3))."""
        self.ctx.set_eval_params(material_params, env_rgb)
        self.ctx.clear()
        self.ctx.render_replay(self.first_pass, self.n_passes, self.free_counter, self.max_depth)
        return self.ctx.read_frame()[..., :3] / np.float32(self.n_passes)

    def gradients(self, material_params, env_rgb, d_image):
        """Gradients of sum(d_image . image) w.r.t. (material_params, env_rgb)."""
        self.ctx.set_eval_params(material_params, env_rgb)
        d_frame = np.asarray(d_image, np.float32) / np.float32(self.n_passes)
        return self.ctx.adjoint(d_frame, self.n_materials, self.env_shape, self.first_pass, self.n_passes,
                                self.free_counter, self.max_depth)


class RenderFunction(torch.autograd.Function):
    """image = render(material_params, env_rgb); backward runs the path-replay adjoint."""

    @staticmethod
    def forward(fctx, material_params, env_rgb, renderer):
        mp = material_params.detach().cpu().numpy().astype(np.float32)
        env = env_rgb.detach().cpu().numpy().astype(np.float32)
        img = renderer.image(mp, env)
        fctx.renderer = renderer
        fctx.save_for_backward(material_params, env_rgb)
        return torch.from_numpy(img).to(material_params.device)

    @staticmethod
    def backward(fctx, grad_image):
        material_params, env_rgb = fctx.saved_tensors
        gm, ge = fctx.renderer.gradients(material_params.detach().cpu().numpy().astype(np.float32),
                                         env_rgb.detach().cpu().numpy().astype(np.float32),
                                         grad_image.detach().cpu().numpy())
        return (torch.from_numpy(gm).to(material_params.device), torch.from_numpy(ge).to(env_rgb.device), None)


def render_image(material_params, env_rgb, renderer):
    return RenderFunction.apply(material_params, env_rgb, renderer)
