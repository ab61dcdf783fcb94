"""Plain per-op FP32 path: the reference's op composition of one
advect -> render -> L2 step, run on libgft's per-op HIP kernels. It is
the baseline the fused step is measured against.

* advection: VelocityGrid.warp (data_structures.py:870-905) as separate
  SampleGridLut passes (forward trace, back trace, min, max) plus elementwise
  combine and where passes, with the LuT from the centred velocity
  (data_structures.py:864-868);
* render: the 4-channel light/density grid is materialised
  (renderer.py:573-612 with an ambient light, scaled by the depth step,
  renderer.py:420-437), resampled into every camera frustum with
  SampleGridTransform (TRANSFORM_LINDEPTH) and blended along depth with
  ReduceGridBlend (the non-fused Renderer path, renderer.py:463-526);
* loss: mean squared error of the rgb channels;
* backward: autograd through the ops' own gradient kernels (the reference's
  registered, count-normalised gradients: grad_mode "REFERENCE").
"""
import numpy as np
import torch

from .ops import blending_ops, sampling_ops


class PlainTransportRenderStep(object):
    def __init__(self, grid_dhw, matrix_m, matrix_v, matrix_p, frustum_params, output_shape, depth_step,
                 light=(1.0, 1.0, 1.0), order=2, clamp="MC_SMOOTH", dt=1.0, grad_mode="REFERENCE", device="cuda"):
        self.grid_dhw = tuple(int(s) for s in grid_dhw)
        self.M, self.V, self.P, self.F = matrix_m, matrix_v, matrix_p, frustum_params
        self.out_dhw = tuple(int(s) for s in output_shape)
        self.depth_step = float(depth_step)
        self.light = torch.tensor(light, dtype=torch.float32, device=device)
        self.order, self.clamp, self.dt = int(order), str(clamp).upper(), float(dt)
        self.grad_mode = grad_mode
        self.num_views = int(np.asarray(matrix_v).reshape(-1, 16).shape[0])

    def _lut_sample(self, x, lut, interpolation="LINEAR"):
        return sampling_ops.sample_grid_lut(x, lut, interpolation=interpolation, boundary="CLAMP",
                                            separate_camera_batch=False, relative_coords=True,
                                            grad_mode=self.grad_mode)[:, 0]

    def advect(self, d, vel):
        zero = torch.zeros_like(vel[..., :1])
        lut = torch.cat([-vel * self.dt, zero], dim=-1)
        f = self._lut_sample(d, lut)
        out = f
        if self.order == 2:
            back = torch.cat([vel * self.dt, zero], dim=-1)
            b = self._lut_sample(f, back)
            corr = f + 0.5 * (d - b)
            if self.clamp == "MC_SMOOTH":
                with torch.no_grad():
                    mn = self._lut_sample(d, lut.detach(), "MIN")
                    mx = self._lut_sample(d, lut.detach(), "MAX")
                    mask = (corr < mn) | (corr > mx)
                corr = torch.where(mask, f, corr)
            out = corr
        if self.clamp == "NEGATIVE":
            out = torch.relu(out)
        return out

    def render(self, d1):
        grid = torch.cat([d1 * self.light, d1], dim=-1) * self.depth_step
        frusta = sampling_ops.sample_grid_transform(grid, self.M, self.V, self.P, self.F, self.out_dhw,
                                                    interpolation="LINEAR", boundary="BORDER",
                                                    coordinate_mode="TRANSFORM_LINDEPTH",
                                                    grad_mode=self.grad_mode)  # [1,V,D,H,W,4]
        n = frusta.shape[1]
        img = blending_ops.reduce_grid_blend(frusta.reshape((n,) + tuple(frusta.shape[2:])), "BEER_LAMBERT",
                                             grad_mode=self.grad_mode)  # [V,H,W,4]
        return img.unsqueeze(0)

    def run(self, d0, vel, target):
        with torch.enable_grad():
            d = d0.detach().requires_grad_(True)
            v = vel.detach().requires_grad_(True)
            img = self.render(self.advect(d, v))
            loss = torch.mean((img[..., :3] - target) ** 2)
            loss.backward()
        self.images = img.detach()
        return loss.detach(), d.grad, v.grad

    def plain_bytes(self):
        """Compulsory HBM traffic of the per-op composition at FP32: every op
        reads its inputs once and writes its outputs once (V voxels, N
        pixels, S samples per ray, C = 4 materialised channels)."""
        D, H, W = self.grid_dhw
        V = D * H * W
        N = self.num_views * self.out_dhw[1] * self.out_dhw[2]
        NS = N * self.out_dhw[0]
        adv_f = 2 * 28 * V + 4 * 24 * V + 16 * V + 20 * V if self.order == 2 else 28 * V + 24 * V
        ren_f = 20 * V + (16 * V + 16 * NS) + (16 * NS + 16 * N)
        loss = 12 * N + 12 * N
        ren_b = 12 * N + (16 * NS + 32 * N + 16 * NS) + (16 * NS + 16 * V + 4 * V) + (16 * V + 4 * V)
        adv_b = 2 * (4 * V + 16 * V + 4 * V + 4 * V + 16 * V) + 40 * V if self.order == 2 else \
            (4 * V + 16 * V + 4 * V + 4 * V + 16 * V) + 16 * V
        return adv_f + ren_f + loss + ren_b + adv_b
