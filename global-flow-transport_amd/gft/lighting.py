"""Lights of the density renderer (phitest/render/lighting.py).

``Light`` is the ambient light: its grid lighting is density * colour *
intensity (lighting.py:10-32). ``PointLight`` adds the inverse-square-like
falloff 1 / (1 + (|p - l| s)^2) in world-scaled grid coordinates
(lighting.py:34-75). ``SpotLight`` multiplies that by a shadow transmittance:
the density resampled into the light's frustum, blended along depth with
every prefix kept, shifted by one cell, exponentiated (BEER_LAMBERT) or
clipped (ALPHA), cone-masked and resampled back to the grid
(lighting.py:77-147). All grid work runs on the HIP ops through the renderer;
the frustum part runs as one fused shadow op where the renderer's settings
allow (the op sequence otherwise).
"""
import numpy as np
import torch

from .camera import Camera
from .transform import GridTransform


class Light(object):
    def __init__(self, color=(1., 1., 1.), intensity=1.):
        self.c = np.asarray(color, dtype=np.float64)
        self.i = float(intensity)

    @property
    def monochrome(self):
        return np.isscalar(self.c) or np.asarray(self.c).shape in ((), (1,))

    def color_intensity(self):
        return np.asarray(self.c, dtype=np.float64).reshape(-1) * self.i

    def grid_lighting(self, density_transform, renderer=None):
        c = torch.tensor(self.color_intensity(), dtype=torch.float32, device=density_transform.data.device)
        return density_transform.data * c


class PointLight(Light):
    def __init__(self, transform, color=(1., 1., 1.), intensity=1., range_scale=1, range_limit=0):
        super().__init__(color, intensity)
        self.t = transform
        self.r = range_limit
        self.s = range_scale

    def falloff_factor(self, density_transform):
        """[1,D,H,W,1] float32 falloff 1 / (1 + (|p - l| s)^2) without the colour
        (lighting.py:54-73)."""
        return self._falloff_impl(density_transform, None)

    def _falloff(self, density_transform):
        return self._falloff_impl(density_transform, self.color_intensity())

    def _falloff_impl(self, density_transform, color):
        data = density_transform.data
        D, H, W = data.shape[1:4]
        cell = density_transform.cell_size_world()
        light_ws = np.asarray(self.t.position_global(), dtype=np.float64)
        light_os = (np.linalg.inv(np.asarray(density_transform.get_transform_matrix(), dtype=np.float64)) @ light_ws)[:3]
        lpos = light_os * cell
        dev = data.device
        if color is None:
            # the factor depends on nothing but these (not on the density): a
            # step that renders the same grid again reuses it
            key = (int(D), int(H), int(W), tuple(float(c) for c in cell), tuple(float(c) for c in lpos), float(self.s),
                   str(dev))
            cached = getattr(self, "_falloff_cache", None)
            if cached is not None and cached[0] == key:
                return cached[1]
        x = (torch.arange(W, dtype=torch.float64, device=dev) + 0.5) * cell[0] - lpos[0]
        y = (torch.arange(H, dtype=torch.float64, device=dev) + 0.5) * cell[1] - lpos[1]
        z = (torch.arange(D, dtype=torch.float64, device=dev) + 0.5) * cell[2] - lpos[2]
        dist = torch.sqrt(z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2)
        fall = 1.0 / (1.0 + (dist * self.s) ** 2)
        if color is None:
            out = fall[..., None].to(torch.float32)[None]
            self._falloff_cache = (key, out)
            return out
        c = torch.tensor(color, dtype=torch.float64, device=dev)
        return (fall[..., None] * c).to(torch.float32)[None]

    def grid_lighting(self, density_transform, renderer=None):
        return density_transform.data * self._falloff(density_transform)

    def light_factor(self, density_transform, renderer):
        """Per-cell factor of this light without colour and intensity (the
        falloff), [1,D,H,W,1]."""
        return self.falloff_factor(density_transform)


class SpotLight(PointLight):
    def __init__(self, transform, color=(1., 1., 1.), intensity=1., range_scale=1, range_limit=0, angle_deg=45,
                 cast_shadows=False, shadow_resolution=(64, 64, 64), shadow_clip=(1, 10), cone_mask=True):
        super().__init__(transform, color, intensity, range_scale, range_limit)
        self.a = angle_deg
        self.cast_shadows = cast_shadows
        self.shadow_resolution = list(shadow_resolution)
        self.shadow_cam = Camera(GridTransform.from_transform(transform, self.shadow_resolution), list(shadow_clip),
                                 fov=self.a)
        self.cone_mask = cone_mask

    def _shadow_mask(self, device):
        a = torch.linspace(-1.0, 1.0, self.shadow_resolution[1], device=device)
        b = torch.linspace(-1.0, 1.0, self.shadow_resolution[2], device=device)
        gy, gx = torch.meshgrid(a, b, indexing="ij")
        return (torch.sqrt(gy * gy + gx * gx) <= 1).to(torch.float32)[None, None, :, :, None]

    def _shadow_ops(self, density_transform, renderer):
        """The op sequence of lighting.py:108-135: frustum resample, depth
        prefix blend, one-cell shift, transmittance."""
        shadow = renderer.sample_camera(density_transform.data, density_transform, self.shadow_cam, inverse=False,
                                        use_step_channel=[0])
        shadow = renderer.blending.reduce_grid_blend(shadow, renderer.blend_mode, keep_dims=True,
                                                     grad_mode=renderer.grad_mode)
        if renderer.boundary_mode == "BORDER":
            pad = torch.zeros_like(shadow[:, :1])
        elif renderer.boundary_mode == "CLAMP":
            pad = shadow[:, :1]
        elif renderer.boundary_mode == "WRAP":
            pad = shadow[:, -1:]
        else:
            raise ValueError("Unknow boundary_mode %s" % renderer.boundary_mode)
        shadow = torch.cat([pad, shadow[:, :-1]], dim=1)
        if renderer.blend_mode == "BEER_LAMBERT":
            return torch.exp(-shadow)
        if renderer.blend_mode == "ALPHA":
            return 1 - torch.clamp(shadow, 0, 1)
        raise ValueError("Unknown blend_mode '%s'" % renderer.blend_mode)

    def render_shadow_map(self, density_transform, renderer, fused=None):
        """Shadow transmittance on the density grid (lighting.py:106-141).
        With shadows cast, the fused shadow op computes the light-frustum
        transmittance in one pass where the renderer's settings allow it
        (``fused`` None: the renderer decides; False: always the op
        sequence); the reverse resample to the grid follows either way."""
        dev = density_transform.data.device
        mask = self._shadow_mask(dev) if self.cone_mask else None
        if self.cast_shadows:
            transmission = None
            if fused is not False:
                transmission = renderer.shadow_transmission(density_transform, self.shadow_cam, mask)
            if transmission is None:
                transmission = self._shadow_ops(density_transform, renderer)
                if mask is not None:
                    transmission = transmission * mask
        else:
            transmission = torch.ones([1] + self.shadow_resolution + [1], dtype=torch.float32, device=dev)
            if mask is not None:
                transmission = transmission * mask
        return renderer.sample_camera(transmission.contiguous(), density_transform, self.shadow_cam, inverse=True,
                                      use_step_channel=None)

    def grid_lighting(self, density_transform, renderer=None):
        light = self.render_shadow_map(density_transform, renderer)
        return light * super().grid_lighting(density_transform, renderer)

    def light_factor(self, density_transform, renderer):
        """Per-cell factor of this light without colour and intensity: the
        shadow transmittance on the grid times the falloff, [1,D,H,W,1]."""
        t = self.render_shadow_map(density_transform, renderer)
        return t.reshape(density_transform.data.shape) * self.falloff_factor(density_transform)
