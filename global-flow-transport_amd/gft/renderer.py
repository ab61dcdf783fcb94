"""Renderer subset of phitest/render/renderer.py on the HIP ops.

Same method names, arguments and defaults as the reference's Renderer for
the parts the reconstruction loop drives: camera parameter packing
(``_get_camera_params_batch`` :108-112), frustum resampling with the
renderer's custom gather gradient (``_sample_transform`` :188-227,
``_sample_camera_transform`` :229-281, ``sample_camera`` :335-398), the fused
raymarch (``_raymarch_camera_transform`` :320-333, ``raymarch_camera``
:400-441), LuT sampling (``_sample_LuT`` :445-459,
``resample_grid3D_offset`` :695-709), depth blending (``_blend_grid``
:463-526), lighting and the image composition of ``render_density``
(:555-651): rgb + optical depth, optional background * exp(-depth), alpha
cut, tonemapping.

Differences: static-camera LuT caches (``_sample_camera_LuT``) and custom
op hooks are not part of this path; an ambient-only BEER_LAMBERT render with
``fused=True`` runs the fused density kernel without materialising the
4-channel light/density grid.
"""
import numpy as np
import torch

from . import _lib
from .camera import camera_params_batch
from .lighting import Light
from .ops import blending_ops, lighting_ops, raymarching_ops, sampling_ops
from .pipeline import render_density as fused_render_density
from .pipeline import render_density_lit as fused_render_density_lit


def _as_list(x):
    return list(x) if isinstance(x, (list, tuple)) else [x]


class Renderer(object):
    def __init__(self, profiler=None, filter_mode='LINEAR', boundary_mode='BORDER', mipmapping='LINEAR', num_mips=2,
                 mip_bias=0.0, sample_gradients=False, fast_gradient_mip_bias_add=0.0, blend_mode='BEER_LAMBERT',
                 name='Renderer', grad_mode="REFERENCE", **kwargs):
        self.blending = blending_ops
        self.filter_mode = filter_mode
        self.boundary_mode = boundary_mode
        self.mip_mode = mipmapping
        self.num_mips = num_mips
        self.mip_bias = mip_bias
        self.gradient_mip_bias_add = fast_gradient_mip_bias_add
        self.sample_gradients = sample_gradients
        self.blend_mode = blend_mode
        self.grad_mode = grad_mode
        self.luma = kwargs.get("luma", [0.2126, 0.7152, 0.0722])
        self.allow_fused = kwargs.get("fused", False)
        self.fused_shadow = kwargs.get("fused_shadow", True)
        self.fused_lit = kwargs.get("fused_lit", True)
        self.scattering_ratio = kwargs.get("scattering_ratio", 1.0)

    @property
    def can_render_fused(self):
        return self.allow_fused and (self.blend_mode in ["BEER_LAMBERT", "ADDITIVE", "ALPHA_ADDITIVE", "ALPHA"])

    @staticmethod
    def _get_camera_params_batch(cameras):
        return camera_params_batch(_as_list(cameras))

    def _sort_cameras(self, cameras):
        sizes = []
        for cam in cameras:
            if list(cam.transform.grid_size) not in sizes:
                sizes.append(list(cam.transform.grid_size))
        return [(size, [cam for cam in cameras if list(cam.transform.grid_size) == size]) for size in sizes]

    # ------------------------------------------------------------ sampling
    def _sample_grad_mode(self):
        return "GATHER" if self.sample_gradients else self.grad_mode

    def _sample_transform(self, data, from_transformations, to_transformations, inverse=False, fix_scale_center=False):
        """Resample between grids (TRANSFORM / TRANSFORM_REVERSE, identity
        projection, frustum [-1,1,-1,1,1,-1]); renderer.py:188-227."""
        cm = "TRANSFORM_REVERSE" if inverse else "TRANSFORM"
        M, V, P, F = self._sample_transform_params(from_transformations, to_transformations, fix_scale_center)
        out_shape = from_transformations[0].grid_size if inverse else to_transformations[0].grid_size
        return sampling_ops.sample_grid_transform(data, M, V, P, F, out_shape, interpolation=self.filter_mode,
                                                  boundary=self.boundary_mode, mipmapping=self.mip_mode,
                                                  num_mipmaps=self.num_mips, mip_bias=self.mip_bias,
                                                  coordinate_mode=cm, grad_mode=self._sample_grad_mode(),
                                                  gradient_mip_bias_add=self.gradient_mip_bias_add)

    @staticmethod
    def _sample_transform_params(from_transformations, to_transformations, fix_scale_center=False):
        """M, V, P, F of a grid-to-grid resample (renderer.py:191-198)."""
        from .transform import GridTransform
        M = np.stack([np.asarray(t.get_transform_matrix(), dtype=np.float32).T for t in from_transformations])
        if fix_scale_center:
            V = [(GridTransform(t.grid_size, scale=[2, 2, 2], center=True, normalize='ALL').get_transform_matrix()
                  @ t.get_inverse_transform()).T for t in to_transformations]
        else:
            V = [np.asarray(t.get_inverse_transform()).T for t in to_transformations]
        V = np.stack(V).astype(np.float32)
        P = np.stack([np.eye(4, dtype=np.float32) for _ in to_transformations])
        F = np.array([[-1, 1, -1, 1, 1, -1] for _ in to_transformations], dtype=np.float32)
        return M, V, P, F

    def _sample_camera_transform(self, data, transformations, cameras, inverse=False):
        cm = "TRANSFORM_LINDEPTH_REVERSE" if inverse else "TRANSFORM_LINDEPTH"
        M = np.stack([np.asarray(t.get_transform_matrix(), dtype=np.float32).T for t in transformations])
        V, P, F = self._get_camera_params_batch(cameras)
        out_shape = transformations[0].grid_size if inverse else cameras[0].transform.grid_size
        return sampling_ops.sample_grid_transform(data, M, V, P, F, out_shape, interpolation=self.filter_mode,
                                                  boundary=self.boundary_mode, mipmapping=self.mip_mode,
                                                  num_mipmaps=self.num_mips, mip_bias=self.mip_bias,
                                                  coordinate_mode=cm, grad_mode=self._sample_grad_mode(),
                                                  gradient_mip_bias_add=self.gradient_mip_bias_add)

    def _step_scale(self, data, cameras, use_step_channel, fail_on_mixed=False):
        """Premultiply the channels in use_step_channel by the depth step
        (renderer.py:365-385); returns (data, per-camera scale or None)."""
        C = data.shape[-1]
        if use_step_channel is None or (not np.isscalar(use_step_channel) and len(use_step_channel) == 0):
            return data, None
        if np.isscalar(use_step_channel):
            return data * float(use_step_channel), None
        ch = sorted({c % C for c in use_step_channel if -C <= c < C})
        steps = [cam.depth_step for cam in cameras]
        if all(s == steps[0] for s in steps):
            scale = torch.tensor([steps[0] if c in ch else 1.0 for c in range(C)], dtype=torch.float32,
                                 device=data.device)
            return data * scale, None
        if fail_on_mixed:
            raise ValueError("All cameras must have the same step size for batched rendering.")
        post = torch.tensor([[s if c in ch else 1.0 for c in range(C)] for s in steps], dtype=torch.float32,
                            device=data.device)
        return data, post

    def _check_batch(self, data, transformations):
        if data.dim() != 5:
            raise ValueError('data must be 5D (NDHWC)')
        if not isinstance(transformations, (list, tuple)):
            if data.shape[0] != 1:
                raise ValueError('transformation and data batch size mismatch.')
            return [transformations], True
        if len(transformations) != data.shape[0]:
            raise ValueError('transformation and data batch size mismatch.')
        return list(transformations), False

    def _check_cameras(self, cameras):
        cameras = _as_list(cameras)
        if len(cameras) > 1 and not all(list(c.transform.grid_size) == list(cameras[0].transform.grid_size)
                                        for c in cameras[1:]):
            raise ValueError('all cameras must have the same resolution (DHW). (use Renderer._sort_cameras() for '
                             'batching.)')
        return cameras

    def sample_camera(self, data, transformations, cameras, inverse=False, use_step_channel=None):
        transformations, no_batch = self._check_batch(data, transformations)
        cameras = self._check_cameras(cameras)
        data, post = self._step_scale(data, cameras, use_step_channel)
        sampled = self._sample_camera_transform(data.contiguous(), transformations, cameras, inverse)  # NVDHWC
        if post is not None:
            sampled = sampled * post[None, :, None, None, None, :]
        return sampled[0] if no_batch else sampled

    def _raymarch_camera_transform(self, data, transformations, cameras, **kwargs):
        M = np.stack([np.asarray(t.get_transform_matrix(), dtype=np.float32).T for t in transformations])
        V, P, F = self._get_camera_params_batch(cameras)
        return raymarching_ops.raymarch_grid_transform(data, M, V, P, F, cameras[0].transform.grid_size,
                                                       interpolation=kwargs.get("filter_mode", self.filter_mode),
                                                       boundary=kwargs.get("boundary_mode", self.boundary_mode),
                                                       blending_mode=kwargs.get("blend_mode", self.blend_mode),
                                                       grad_mode=self.grad_mode)

    def shadow_transmission(self, density_transform, camera, cone_mask=None):
        """Light-frustum transmittance of a spot light's shadow map
        (lighting.py:108-135 and the cone mask) as one fused op, or None
        where the renderer's settings fall outside it (fused_shadow off,
        batch or channels other than 1, MIN/MAX filters, WRAP, blend modes
        other than BEER_LAMBERT and ALPHA)."""
        data = density_transform.data
        if not self.fused_shadow or data.dim() != 5 or data.shape[0] != 1 or data.shape[-1] != 1:
            return None
        if str(self.filter_mode).upper() not in ("LINEAR", "NEAREST"):
            return None
        if str(self.boundary_mode).upper() not in ("BORDER", "CLAMP"):
            return None
        if str(self.blend_mode).upper() not in ("BEER_LAMBERT", "ALPHA"):
            return None
        M = np.asarray(density_transform.get_transform_matrix(), dtype=np.float32).T[None]
        V, P, F = self._get_camera_params_batch([camera])
        return lighting_ops.shadow_transmission(data, M, V, P, F, camera.transform.grid_size, camera.depth_step,
                                                interpolation=self.filter_mode, boundary=self.boundary_mode,
                                                mipmapping=self.mip_mode, num_mipmaps=self.num_mips,
                                                mip_bias=self.mip_bias, blending_mode=self.blend_mode,
                                                cone_mask=cone_mask, grad_mode=self._sample_grad_mode(),
                                                gradient_mip_bias_add=self.gradient_mip_bias_add)

    def raymarch_camera(self, data, transformations, cameras, use_step_channel=None, **raymarch_kwargs):
        transformations, no_batch = self._check_batch(data, transformations)
        cameras = self._check_cameras(cameras)
        data, _ = self._step_scale(data, cameras, use_step_channel, fail_on_mixed=True)
        sampled = self._raymarch_camera_transform(data.contiguous(), transformations, cameras, **raymarch_kwargs)
        return sampled[0] if no_batch else sampled

    def _sample_LuT(self, data, luts, combined_batch=False, relative=False, normalized=False, cell_center_offset=0.0):
        if luts.shape[-1] == 3:
            luts = torch.cat([luts, torch.zeros_like(luts[..., :1])], dim=-1)
        if luts.shape[-1] != 4:
            raise ValueError("lookup table must have 3 or 4 channels")
        out = sampling_ops.sample_grid_lut(data, luts.contiguous(), interpolation=self.filter_mode,
                                           boundary=self.boundary_mode, mipmapping=self.mip_mode,
                                           num_mipmaps=self.num_mips, mip_bias=self.mip_bias,
                                           cell_center_offset=cell_center_offset,
                                           separate_camera_batch=not combined_batch, relative_coords=relative,
                                           normalized_coords=normalized, grad_mode=self.grad_mode)
        return out[:, 0] if combined_batch else out

    def resample_grid3D_aligned(self, data, target_shape, align_x='BORDER', align_y='BORDER', align_z='BORDER'):
        """Resample NDHWC data to target_shape (D,H,W) with per-axis alignment
        (renderer.py:653-692): BORDER aligns the outer cell borders, CENTER
        the outer cell centres, STAGGER_INPUT input centres to output
        borders, STAGGER_OUTPUT input borders to output centres. The input
        is scaled into the centred [-1,1] cube that the TRANSFORM resample
        maps the output grid onto; the sample position of output cell q is
        then (q+0.5) in/out - 0.5 (BORDER), q (in-1)/(out-1) (CENTER),
        (q+0.5) (in-1)/out (STAGGER_INPUT), (q+0.5) (in+1)/out - 1
        (STAGGER_OUTPUT)."""
        if data.dim() != 5:
            raise ValueError("input shape must be NDHWC, is : %s" % list(data.shape))
        if len(target_shape) != 3:
            raise ValueError("target_shape must be (D,H,W)")
        in_transform, target_transform = self.aligned_transforms(data.shape[1:4], target_shape, align_x, align_y,
                                                                 align_z)
        return self._sample_transform(data.contiguous(), [in_transform] * int(data.shape[0]),
                                      [target_transform])[:, 0]

    @staticmethod
    def aligned_transforms(in_shape, target_shape, align_x='BORDER', align_y='BORDER', align_z='BORDER'):
        """Input and target grid transforms of resample_grid3D_aligned."""
        from .transform import GridTransform
        in_dhw = [int(s) for s in in_shape]
        out_dhw = [int(s) for s in target_shape]

        def get_scale(align, n_in, n_out):
            a = str(align).upper()
            if a == 'BORDER':
                return 2. / n_in
            if a == 'CENTER':
                return 2. / n_out * (n_out - 1.) / (n_in - 1.)
            if a == 'STAGGER_INPUT':
                return 2. / (n_in - 1.)
            if a == 'STAGGER_OUTPUT':
                return 2. / (n_in + 1.)
            raise ValueError("unknown alignment {}".format(align))

        scale = [get_scale(align_x, in_dhw[2], out_dhw[2]), get_scale(align_y, in_dhw[1], out_dhw[1]),
                 get_scale(align_z, in_dhw[0], out_dhw[0])]
        return GridTransform(in_dhw, scale=scale, center=True), GridTransform(out_dhw)

    def resample_grid3D_offset(self, data, offsets, target_shape):
        offsets = torch.as_tensor(np.asarray(offsets, dtype=np.float32), device=data.device)
        if offsets.dim() != 2 or offsets.shape[1] != 3:
            raise ValueError("Shape of offsets must be (N,3), is %s" % list(offsets.shape))
        if len(target_shape) != 3:
            raise ValueError("target_shape must be (3,), is %s" % (target_shape,))
        n = offsets.shape[0]
        lut = torch.cat([offsets, torch.zeros_like(offsets[:, :1])], dim=1).reshape(n, 1, 1, 1, 4)
        lut = lut.expand([n] + list(target_shape) + [4]).contiguous()
        return self._sample_LuT(data, lut, relative=True)

    # ------------------------------------------------------------ blending
    def _blend_grid(self, data, blend_mode=None, keep_dims=False):
        """Blend along depth: MAX/MEAN/MIN reductions, otherwise the blend
        kernel (BEER_LAMBERT with self-attenuation: rgb = sum light *
        exp(-cumsum density), alpha = cumsum density)."""
        blend_mode = (blend_mode or self.blend_mode).upper()
        if blend_mode == 'MAX':
            return data.amax(dim=-4, keepdim=keep_dims)
        if blend_mode == 'MEAN':
            return data.mean(dim=-4, keepdim=keep_dims)
        if blend_mode == 'MIN':
            return data.amin(dim=-4, keepdim=keep_dims)
        shape = data.shape
        flat = data.reshape((-1,) + tuple(shape[-4:])).contiguous()
        out = blending_ops.reduce_grid_blend(flat, blend_mode, keep_dims, grad_mode=self.grad_mode)
        return out.reshape(tuple(shape[:-4]) + tuple(out.shape[1:]))

    def _tonemap(self, image, mode='NONE'):
        mode = mode.upper()
        if mode == 'NONE':
            return image
        if mode == 'CLIP_NEGATIVE':
            return torch.clamp_min(image, 0)
        if mode == 'SATURATE':
            return torch.clamp(image, 0, 1)
        if mode == 'NORMALIZE':
            mn = image.min()
            if mn < 0:
                image = image - mn
            mx = image.max()
            return image / mx if mx > 0 else image
        raise ValueError("Unknown tonemapping '%s'" % mode)

    # ------------------------------------------------------------ rendering
    def _build_light_grid(self, density_transform, light_list, monochrome=False):
        if len(light_list) == 0:
            raise ValueError('light list is empty')
        total = None
        for i, light in enumerate(light_list):
            if isinstance(light, Light):
                g = light.grid_lighting(density_transform, self)
                if monochrome and not light.monochrome:
                    g = (g * torch.tensor(self.luma, dtype=torch.float32, device=g.device)).sum(-1, keepdim=True)
                if not monochrome and light.monochrome:
                    g = g.expand(tuple(g.shape[:-1]) + (3,))
            elif isinstance(light, torch.Tensor):
                g = light
                if g.shape[-1] == 3 and monochrome:
                    g = (g * torch.tensor(self.luma, dtype=torch.float32, device=g.device)).sum(-1, keepdim=True)
            else:
                raise ValueError("Type of light {} '{}' is not supported.".format(i, type(light).__name__))
            total = g if total is None else total + g
        shape = tuple(density_transform.data.shape[:-1]) + (1 if monochrome else 3,)
        return total.expand(shape)

    def _render_cameras(self, light_density, grid_transform, cameras):
        if self.can_render_fused:
            return self.raymarch_camera(light_density, grid_transform, cameras,
                                        use_step_channel=[0, 1, 2, 3] if self.blend_mode not in ['MAX', 'MEAN', 'MIN']
                                        else None)
        frustum = self.sample_camera(light_density, grid_transform, cameras, False,
                                     use_step_channel=[0, 1, 2, 3] if self.blend_mode not in ['MAX', 'MEAN', 'MIN']
                                     else None)
        return self._blend_grid(frustum, self.blend_mode)

    def _ambient_only(self, light_list, monochrome):
        return (not monochrome and len(light_list) > 0
                and all(type(l) is Light and not l.monochrome for l in light_list))

    def _lit_split(self, light_list, monochrome):
        """For the fused lit render: (color, lights with a per-cell factor and
        their scales, ambient scale) when every light is a Light whose
        colour * intensity is a multiple of one colour and at least one is a
        point or spot light; None otherwise."""
        if monochrome or not light_list or not all(isinstance(l, Light) for l in light_list):
            return None
        vecs = [np.asarray(l.color_intensity(), dtype=np.float64).reshape(-1) for l in light_list]
        if any(v.shape != (3,) for v in vecs):
            return None
        ref = max(vecs, key=lambda v: float(np.abs(v).max()))
        if not np.abs(ref).max() > 0:
            return None
        color = ref / np.abs(ref).max()
        scales = []
        for v in vecs:
            k = float(v @ color / (color @ color))
            if np.abs(v - k * color).max() > 1e-6 * max(np.abs(v).max(), 1e-30):
                return None
            scales.append(k)
        factored = [(l, k) for l, k in zip(light_list, scales) if type(l) is not Light]
        if not factored:
            return None
        ambient = float(sum(k for l, k in zip(light_list, scales) if type(l) is Light))
        return tuple(float(c) for c in color), factored, ambient

    def render_density(self, density_transform, light_list, camera_list, cut_alpha=True, background=None,
                       monochrome=False, split_cameras=False, tonemapping="NONE"):
        """Images of the density under the lights, one [1,H,W,C] tensor per
        camera in camera order; C = 3 (+1 optical depth unless cut_alpha).

        With ``fused`` and BEER_LAMBERT, an ambient-only light list runs the
        fused density kernel, and a list of point/spot lights plus ambient
        lights with proportional colours (the production [SpotLight, Light])
        runs the fused lit kernel on BORDER: the light grid d * (c lambda +
        c a) is never materialised; lambda (shadow transmittance times falloff,
        with the intensities) is built with the differentiable ops as the
        reference's _build_light_grid does (renderer.py:573-601)."""
        density = density_transform.data
        single = density.shape[0] == 1 and density.shape[-1] == 1
        fast = (self.allow_fused and self.blend_mode == "BEER_LAMBERT" and self._ambient_only(light_list, monochrome)
                and single)
        lit = None
        if (self.allow_fused and self.fused_lit and self.blend_mode == "BEER_LAMBERT" and not fast and single
                and self.boundary_mode == "BORDER"):
            lit = self._lit_split(light_list, monochrome)
        lam = None
        if lit is not None:
            color, factored, ambient = lit
            for l, k in factored:
                f = l.light_factor(density_transform, self) * k
                lam = f if lam is None else lam + f
            lam = lam.contiguous()
        data = None
        if not fast and lit is None:
            light_data = self._build_light_grid(density_transform, light_list, monochrome)
            data = torch.cat([light_data, density], dim=-1)
        cam_images = [None] * len(camera_list)
        for cam_size, cams in self._sort_cameras(camera_list):
            groups = [[c] for c in cams] if split_cameras else [cams]
            for group in groups:
                steps = [c.depth_step for c in group]
                if fast and all(s == steps[0] for s in steps):
                    M = np.asarray(density_transform.get_transform_matrix(), dtype=np.float32).T[None]
                    V, P, F = self._get_camera_params_batch(group)
                    light = np.sum([l.color_intensity() for l in light_list], axis=0)
                    images = fused_render_density(density.contiguous(), M, V, P, F, cam_size, steps[0],
                                                  tuple(float(v) for v in light), self.boundary_mode)[0]
                elif lit is not None and all(s == steps[0] for s in steps):
                    M = np.asarray(density_transform.get_transform_matrix(), dtype=np.float32).T[None]
                    V, P, F = self._get_camera_params_batch(group)
                    images = fused_render_density_lit(density.contiguous(), lam, M, V, P, F, cam_size, steps[0],
                                                      color=lit[0], ambient=lit[2])[0]
                else:
                    if data is None:
                        data = torch.cat([self._build_light_grid(density_transform, light_list, monochrome), density],
                                         dim=-1)
                    images = self._render_cameras(data, density_transform, group)
                t = torch.exp(-images[..., -1:])
                if cut_alpha:
                    images = images[..., :-1]
                if background is not None:
                    bk = torch.stack([torch.as_tensor(background[camera_list.index(c)], dtype=torch.float32,
                                                      device=images.device).expand(tuple(images.shape[1:]))
                                      for c in group])
                    images = images + bk * t
                images = self._tonemap(images, mode=tonemapping)
                for k, c in enumerate(group):
                    cam_images[camera_list.index(c)] = images[k:k + 1]
        return cam_images
