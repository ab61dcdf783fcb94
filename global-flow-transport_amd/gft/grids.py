"""Density and staggered-velocity grids with the reference's advection
semantics, on the HIP kernels.

Restates the transport part of phitest/render/data_structures.py:

* DensityGrid (:97-330): the density [1,D,H,W,1], optional inflow added and
  clamped at 0 (``with_inflow`` :167-171), ``warped`` (:263-266);
* VelocityGrid (:380-936): staggered MAC components x [1,D,H,W+1,1],
  y [1,D,H+1,W,1], z [1,D+1,H,W,1] (``component_shapes`` :382-389), the
  centred velocity as the mean of the two faces of a cell (``centered``
  :819-833), the LuT -v*dt (``centered_lut_grid`` :864-868), scalar warps
  (``warp`` :870-905), self-advection of each component with the velocity
  interpolated to its faces (``_sampled_to_component_shape`` :835-862,
  ``_warp_vel_component`` :594-619, ``warped`` :621-651) and ``divergence``
  (:929-936);
* State / Sequence (:1074-1325): ``density_advected``, ``velocity_advected``,
  ``densities_advect_fwd`` and ``velocities_advect_fwd``.

Warps run on the fused MacCormack kernel (one pass forward, exact adjoint
w.r.t. the data and the velocity) when ``grad_mode`` is "EXACT"; with
"REFERENCE" they are composed of SampleGridLut passes exactly as the reference
composes them, so their gradients are the reference's count-normalised ones.
"""
import os

import numpy as np
import torch

from . import _lib
from .ops import sampling_ops
from .pipeline import advect as fused_advect

WARP_CLAMPS = ("NONE", "MC", "MC_SMOOTH", "NEGATIVE")


def _to_numpy(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def _check_clamp(clamp):
    clamp = "NONE" if clamp is None else str(clamp).upper()
    if clamp not in WARP_CLAMPS:
        raise _lib.InvalidArgumentError("invalid warp clamp '%s'" % clamp)
    if clamp == "MC":
        # data_structures.py:890-893
        raise NotImplementedError("MIM and MAX warp sampling have wrong gradients.")
    return clamp


def _lut_sample(data, lut, interpolation="LINEAR", grad_mode="REFERENCE", boundary="CLAMP"):
    """Renderer._sample_LuT(data, lut, combined_batch=True, relative=True)
    (renderer.py:445-459): LoD channel padded with 0, batch squeezed."""
    if lut.shape[-1] == 3:
        lut = torch.cat([lut, torch.zeros_like(lut[..., :1])], dim=-1)
    return sampling_ops.sample_grid_lut(data, lut.contiguous(), interpolation=interpolation, boundary=boundary,
                                        separate_camera_batch=False, relative_coords=True,
                                        grad_mode=grad_mode)[:, 0]


def warp_composed(data, lut, order=1, clamp="NONE", grad_mode="REFERENCE", boundary="CLAMP", negative_clamp=True):
    """VelocityGrid.warp / _warp_vel_component as the reference composes them
    from SampleGridLut passes: forward trace, MacCormack back trace and
    correction, MIN/MAX revert clamp, optional clamp of negatives."""
    clamp = _check_clamp(clamp)
    if order not in (1, 2):
        raise _lib.InvalidArgumentError("Unsupported warp order '%s'" % order)
    f = _lut_sample(data, lut, grad_mode=grad_mode, boundary=boundary)
    out = f
    if order == 2:
        b = _lut_sample(f, -lut, grad_mode=grad_mode, boundary=boundary)
        corr = f + 0.5 * (data - b)
        if clamp == "MC_SMOOTH":
            with torch.no_grad():
                mn = _lut_sample(data, lut.detach(), "MIN", boundary=boundary)
                mx = _lut_sample(data, lut.detach(), "MAX", boundary=boundary)
                oob = (corr < mn) | (corr > mx)
            corr = torch.where(oob, f, corr)
        out = corr
    if clamp == "NEGATIVE" and negative_clamp:
        out = torch.clamp_min(out, 0.0)
    return out


def _fused_warp(data, vel, order, dt, clamp):
    """The fused kernel warps a scalar grid by a velocity given at its cells."""
    return fused_advect(data.contiguous(), vel.contiguous(), dt, order, clamp)


def _scale_renderer(renderer, boundary):
    """The renderer that rescales grids at grow steps; by default the
    reference's scale renderer without mips (LINEAR filter, the velocity
    boundary, reconstruct_sequence.py:528-535)."""
    if renderer is not None:
        return renderer
    from .renderer import Renderer
    return Renderer(filter_mode="LINEAR", boundary_mode=str(boundary).upper(), mipmapping="NONE",
                    blend_mode="ADDITIVE")


class DensityGrid(object):
    def __init__(self, shape, constant=0.1, d=None, inflow=None, inflow_offset=None, inflow_mask=None, device="cuda",
                 grad_mode="EXACT", hull=None, scale_renderer=None):
        self.shape = [int(s) for s in shape]
        self.hull = hull
        self.scale_renderer = scale_renderer
        self.inflow_offset = None if inflow_offset is None else [int(o) for o in inflow_offset]
        if d is None:
            d = torch.full([1] + self.shape + [1], float(constant), dtype=torch.float32, device=device)
        if tuple(d.shape) != tuple([1] + self.shape + [1]):
            raise ValueError("Shape mismatch: density %s for shape %s" % (list(d.shape), self.shape))
        self._d = d
        self._inflow = inflow
        self.inflow_mask = inflow_mask
        self.grad_mode = grad_mode
        if inflow is not None:
            # pad the inflow block to the grid at inflow_offset (z, y, x)
            off = [0, 0, 0] if inflow_offset is None else [int(o) for o in inflow_offset]
            s = list(inflow.shape[1:4])
            self._inflow_padding = [(0, 0)] + [(off[i], self.shape[i] - off[i] - s[i]) for i in range(3)] + [(0, 0)]

    @property
    def d(self):
        return self._d

    @property
    def inflow(self):
        if self._inflow is None:
            return torch.zeros_like(self._d)
        inf = self._inflow * self.inflow_mask if self.inflow_mask is not None else self._inflow
        pad = []
        for lo, hi in reversed(self._inflow_padding):
            pad += [lo, hi]
        return torch.nn.functional.pad(inf, pad)

    def with_inflow(self):
        density = self.d
        if self._inflow is not None:
            density = torch.clamp_min(density + self.inflow, 0.0)
        return density

    def warped(self, vel_grid, order=1, dt=1.0, clamp="NONE"):
        if not isinstance(vel_grid, VelocityGrid):
            raise ValueError("Invalid velocity grid")
        return vel_grid.warp(self.with_inflow(), order=order, dt=dt, clamp=clamp)

    def copy_warped(self, vel_grid, order=1, dt=1.0, clamp="NONE"):
        return DensityGrid(self.shape, d=self.warped(vel_grid, order, dt, clamp), grad_mode=self.grad_mode,
                           scale_renderer=self.scale_renderer)

    # ------------------------------------------------------------ grid growth
    def scaled(self, new_shape, with_inflow=False):
        """The density resampled to new_shape with border alignment
        (data_structures.py:236-247); the same shape returns a copy."""
        if not (isinstance(new_shape, (list, tuple)) and len(new_shape) == 3):
            raise ValueError("Invalid shape")
        new_shape = [int(s) for s in new_shape]
        density = self.d if not with_inflow else self.with_inflow()
        if new_shape != self.shape:
            return _scale_renderer(self.scale_renderer, "CLAMP").resample_grid3D_aligned(density, new_shape)
        return density.clone()

    def copy_scaled(self, new_shape):
        """A new grid of new_shape holding scaled(); inflow and hull are not
        carried over (data_structures.py:249-261)."""
        d = self.scaled(new_shape)
        return DensityGrid(new_shape, d=d, device=d.device, grad_mode=self.grad_mode,
                           scale_renderer=self.scale_renderer)

    # ------------------------------------------------------------ state I/O
    def save(self, path):
        """DensityGrid.save (data_structures.py:332-353): compressed npz with the
        density as the first positional array (``arr_0``) and the optional
        ``hull``, ``inflow``, ``inflow_mask``, ``inflow_offset``."""
        save = {}
        if self.hull is not None:
            save["hull"] = _to_numpy(self.hull)
        if self._inflow is not None:
            save["inflow"] = _to_numpy(self._inflow)
            if self.inflow_mask is not None:
                save["inflow_mask"] = _to_numpy(self.inflow_mask)
            save["inflow_offset"] = np.asarray(self.inflow_offset if self.inflow_offset is not None else [0, 0, 0])
        np.savez_compressed(path, _to_numpy(self._d), **save)

    @classmethod
    def from_file(cls, path, hull=None, inflow=None, inflow_offset=None, inflow_mask=None, device="cuda",
                  grad_mode="EXACT"):
        """DensityGrid.from_file (data_structures.py:174-193); the stored hull
        and inflow fields are used where the caller passes none. Returns None
        (as the reference does) when the file cannot be read."""
        try:
            with np.load(path) as f:
                d = f["arr_0"]
                if "hull" in f and hull is None:
                    hull = torch.from_numpy(np.ascontiguousarray(f["hull"])).to(device)
                if "inflow" in f and inflow is None:
                    inflow = torch.from_numpy(np.ascontiguousarray(f["inflow"], dtype=np.float32)).to(device)
                    if "inflow_mask" in f and inflow_mask is None:
                        inflow_mask = torch.from_numpy(np.ascontiguousarray(f["inflow_mask"],
                                                                            dtype=np.float32)).to(device)
                    if "inflow_offset" in f and inflow_offset is None:
                        inflow_offset = f["inflow_offset"].tolist()
        except (OSError, KeyError, ValueError):
            return None
        d = torch.from_numpy(np.ascontiguousarray(d, dtype=np.float32)).to(device)
        return cls(list(d.shape[1:4]), d=d, inflow=inflow, inflow_offset=inflow_offset, inflow_mask=inflow_mask,
                   device=device, grad_mode=grad_mode, hull=hull)


class VelocityGrid(object):
    @staticmethod
    def component_shapes(centered_shape):
        x_shape = list(centered_shape)
        x_shape[2] += 1
        y_shape = list(centered_shape)
        y_shape[1] += 1
        z_shape = list(centered_shape)
        z_shape[0] += 1
        return x_shape, y_shape, z_shape

    def __init__(self, centered_shape, std=0.1, x=None, y=None, z=None, boundary="CLAMP", device="cuda",
                 grad_mode="EXACT", seed=None, scale_renderer=None):
        self.centered_shape = [int(s) for s in centered_shape]
        self.scale_renderer = scale_renderer
        self.x_shape, self.y_shape, self.z_shape = VelocityGrid.component_shapes(self.centered_shape)
        self.boundary = boundary
        self.grad_mode = grad_mode
        gen = torch.Generator(device="cpu")
        if seed is not None:
            gen.manual_seed(int(seed))
        comps = []
        for c, shape in zip((x, y, z), (self.x_shape, self.y_shape, self.z_shape)):
            if c is None:
                c = (torch.randn([1] + shape + [1], generator=gen) * std).to(device)
            if tuple(c.shape) != tuple([1] + shape + [1]):
                raise ValueError("Shape mismatch: component %s for %s" % (list(c.shape), [1] + shape + [1]))
            comps.append(c)
        self._x, self._y, self._z = comps

    @classmethod
    def from_centered(cls, vel_c, **kw):
        """Staggered grid whose interior faces average the two adjacent centred
        values (edge faces copy the edge cell); ``centered()`` of the result
        is the face average, not the input."""
        v = vel_c
        x = torch.cat([v[:, :, :, :1, 0:1], 0.5 * (v[:, :, :, 1:, 0:1] + v[:, :, :, :-1, 0:1]), v[:, :, :, -1:, 0:1]],
                      dim=3)
        y = torch.cat([v[:, :, :1, :, 1:2], 0.5 * (v[:, :, 1:, :, 1:2] + v[:, :, :-1, :, 1:2]), v[:, :, -1:, :, 1:2]],
                      dim=2)
        z = torch.cat([v[:, :1, :, :, 2:3], 0.5 * (v[:, 1:, :, :, 2:3] + v[:, :-1, :, :, 2:3]), v[:, -1:, :, :, 2:3]],
                      dim=1)
        return cls(list(v.shape[1:4]), x=x.contiguous(), y=y.contiguous(), z=z.contiguous(), **kw)

    @property
    def x(self):
        return self._x

    @property
    def y(self):
        return self._y

    @property
    def z(self):
        return self._z

    def variables(self):
        return [self._x, self._y, self._z]

    def centered(self, concat=True):
        h = 0.5
        vc = [(self.x[:, :, :, 1:] + self.x[:, :, :, :-1]) * h,
              (self.y[:, :, 1:] + self.y[:, :, :-1]) * h,
              (self.z[:, 1:] + self.z[:, :-1]) * h]
        return torch.cat(vc, dim=-1) if concat else vc

    def centered_lut_grid(self, dt=1.0):
        return self.centered() * (-dt)

    def warp(self, data, order=1, dt=1.0, clamp="NONE"):
        """Warp a scalar grid of the centred shape (VelocityGrid.warp)."""
        clamp = _check_clamp(clamp)
        if list(data.shape[1:4]) != self.centered_shape:
            raise ValueError("Shape mismatch")
        if order not in (1, 2):
            raise ValueError("Unsupported warp order '%s'" % order)
        if self.grad_mode == "EXACT":
            return _fused_warp(data, self.centered(), order, dt, clamp)
        return warp_composed(data, self.centered_lut_grid(dt), order, clamp, grad_mode="REFERENCE",
                             boundary=self.boundary)

    def _offset_lut(self, shape, offset, device):
        lut = torch.zeros([1] + list(shape) + [4], dtype=torch.float32, device=device)
        lut[..., 0], lut[..., 1], lut[..., 2] = offset[0], offset[1], offset[2]
        return lut

    def _sampled_to_component_shape(self, component, concat=True):
        """The three velocity components at the faces of ``component`` (x, y
        or z): the own component as is, the others resampled with a constant
        half-cell offset (renderer.resample_grid3D_offset, renderer.py:695-709)."""
        component = component.upper()
        frm, to = 0.5, -0.5
        dev = self._x.device
        gm = "REFERENCE" if self.grad_mode == "REFERENCE" else "EXACT"
        target = {"X": self.x_shape, "Y": self.y_shape, "Z": self.z_shape}[component]

        def res(data, offset):
            return _lut_sample(data, self._offset_lut(target, offset, dev), grad_mode=gm, boundary=self.boundary)
        vx = self.x if component == "X" else res(self.x, [frm, to, 0.0] if component == "Y" else [frm, 0.0, to])
        vy = self.y if component == "Y" else res(self.y, [to, frm, 0.0] if component == "X" else [0.0, frm, to])
        vz = self.z if component == "Z" else res(self.z, [to, 0.0, frm] if component == "X" else [0.0, to, frm])
        v = [vx, vy, vz]
        return torch.cat(v, dim=-1) if concat else v

    def _warp_vel_component(self, data, vel, order=1, dt=1.0, clamp="NONE"):
        clamp = _check_clamp(clamp)
        # the component warp knows only the MacCormack clamps (data_structures.py:594-619)
        kclamp = clamp if clamp == "MC_SMOOTH" else "NONE"
        if self.grad_mode == "EXACT":
            return _fused_warp(data, vel, order, dt, kclamp)
        return warp_composed(data, -vel * dt, order, kclamp, grad_mode="REFERENCE", boundary=self.boundary)

    def warped(self, vel_grid=None, order=1, dt=1.0, clamp="NONE"):
        """Self-advection of the three components (VelocityGrid.warped with
        vel_grid None, the production path)."""
        if vel_grid is not None and vel_grid is not self:
            raise NotImplementedError("warping by a different velocity grid is not part of this path")
        return tuple(self._warp_vel_component(comp, self._sampled_to_component_shape(name), order, dt, clamp)
                     for name, comp in (("X", self.x), ("Y", self.y), ("Z", self.z)))

    def copy_warped(self, vel_grid=None, order=1, dt=1.0, clamp="NONE"):
        x, y, z = self.warped(vel_grid, order, dt, clamp)
        return VelocityGrid(self.centered_shape, x=x, y=y, z=z, boundary=self.boundary, device=self._x.device,
                            grad_mode=self.grad_mode, scale_renderer=self.scale_renderer)

    # ------------------------------------------------------------ grid growth
    def scaled(self, centered_shape, scale_magnitude=True):
        """Components resampled to the staggered shapes of centered_shape,
        each aligned at the outer cell centres along its own axis and at the
        borders along the others, and with scale_magnitude multiplied by
        new/old size per axis (data_structures.py:551-573)."""
        if not (isinstance(centered_shape, (list, tuple)) and len(centered_shape) == 3):
            raise ValueError("Invalid shape")
        centered_shape = [int(s) for s in centered_shape]
        if centered_shape == self.centered_shape:
            return self.x.clone(), self.y.clone(), self.z.clone()
        r = _scale_renderer(self.scale_renderer, self.boundary)
        x_shape, y_shape, z_shape = VelocityGrid.component_shapes(centered_shape)
        x = r.resample_grid3D_aligned(self.x, x_shape, align_x='center')
        y = r.resample_grid3D_aligned(self.y, y_shape, align_y='center')
        z = r.resample_grid3D_aligned(self.z, z_shape, align_z='center')
        if scale_magnitude:
            s = [o / i for i, o in zip(self.centered_shape, centered_shape)]  # z, y, x
            x, y, z = x * s[2], y * s[1], z * s[0]
        return x, y, z

    def copy_scaled(self, centered_shape, scale_magnitude=True):
        x, y, z = self.scaled(centered_shape, scale_magnitude)
        return VelocityGrid(centered_shape, x=x, y=y, z=z, boundary=self.boundary, device=x.device,
                            grad_mode=self.grad_mode, scale_renderer=self.scale_renderer)

    def divergence(self, world_scale=(1, 1, 1)):
        x_div = self.x[:, :, :, 1:, :] - self.x[:, :, :, :-1, :]
        y_div = self.y[:, :, 1:, :, :] - self.y[:, :, :-1, :, :]
        z_div = self.z[:, 1:, :, :, :] - self.z[:, :-1, :, :, :]
        return x_div * world_scale[0] + y_div * world_scale[1] + z_div * world_scale[2]

    def magnitude(self, world_scale=(1, 1, 1)):
        v = self.centered() * torch.tensor(list(world_scale), dtype=torch.float32, device=self._x.device)
        return torch.linalg.norm(v, dim=-1, keepdim=True)


    # ------------------------------------------------------------ state I/O
    def save(self, path):
        """VelocityGrid.save (data_structures.py:680-681): compressed npz with
        ``centered_shape`` and the staggered components ``vel_x``, ``vel_y``,
        ``vel_z`` (NDHWC, N = C = 1)."""
        np.savez_compressed(path, centered_shape=np.asarray(self.centered_shape), vel_x=_to_numpy(self.x),
                            vel_y=_to_numpy(self.y), vel_z=_to_numpy(self.z))

    @classmethod
    def from_file(cls, path, boundary="CLAMP", device="cuda", grad_mode="EXACT"):
        """VelocityGrid.from_file (data_structures.py:487-503), including the
        legacy files without ``centered_shape`` (shape from vel_x, one fewer
        cell in x). Returns None when the file cannot be read."""
        try:
            with np.load(path) as f:
                if "centered_shape" not in f:
                    shape = list(f["vel_x"].shape)
                    shape[-2] -= 1
                    shape = shape[1:-1]
                else:
                    shape = f["centered_shape"].tolist()
                comps = [torch.from_numpy(np.ascontiguousarray(f[k], dtype=np.float32)).to(device)
                         for k in ("vel_x", "vel_y", "vel_z")]
        except (OSError, KeyError, ValueError):
            return None
        return cls(shape, x=comps[0], y=comps[1], z=comps[2], boundary=boundary, device=device, grad_mode=grad_mode)


class State(object):
    def __init__(self, density, velocity, frame, prev=None, next=None, targets=None):
        self.density = density
        self.velocity = velocity
        self.frame = frame
        self.prev = prev
        self.next = next
        self.targets = targets

    def density_advected(self, dt=1.0, order=1, clamp="NONE"):
        return self.density.warped(self.velocity, order=order, dt=dt, clamp=clamp)

    def velocity_advected(self, dt=1.0, order=1, clamp="NONE"):
        return self.velocity.copy_warped(order=order, dt=dt, clamp=clamp)

    def rescale_density(self, shape):
        """Grow step (data_structures.py:1178-1179)."""
        self.density = self.density.copy_scaled(shape)

    def rescale_velocity(self, shape, scale_magnitude=True):
        self.velocity = self.velocity.copy_scaled(shape, scale_magnitude=scale_magnitude)

    def rescale(self, dens_shape, vel_shape):
        self.rescale_density(dens_shape)
        self.rescale_velocity(vel_shape)

    def save(self, path, suffix=None):
        """State.save (data_structures.py:1213-1215): density.npz and
        velocity.npz (or density_<suffix>.npz, velocity_<suffix>.npz) in path."""
        self.density.save(os.path.join(path, "density.npz" if suffix is None else "density_" + suffix + ".npz"))
        self.velocity.save(os.path.join(path, "velocity.npz" if suffix is None else "velocity_" + suffix + ".npz"))

    @classmethod
    def from_file(cls, path, frame, boundary="CLAMP", device="cuda", density_filename="density.npz",
                  velocity_filename="velocity.npz"):
        """State.from_file (data_structures.py:1123-1128)."""
        density = DensityGrid.from_file(os.path.join(path, density_filename), device=device)
        velocity = VelocityGrid.from_file(os.path.join(path, velocity_filename), boundary=boundary, device=device)
        return cls(density, velocity, frame)


class Sequence(list):
    @classmethod
    def from_file(cls, load_path, frames, boundary="CLAMP", device="cuda", density_filename="density.npz",
                  velocity_filename="velocity.npz"):
        """Sequence.from_file (data_structures.py:1241-1258): one State per
        frame from load_path/frame_{:06d}, linked prev/next."""
        seq = cls()
        prev = None
        for frame in frames:
            st = State.from_file(os.path.join(load_path, "frame_{:06d}".format(frame)), frame, boundary=boundary,
                                 device=device, density_filename=density_filename,
                                 velocity_filename=velocity_filename)
            st.prev = prev
            if prev is not None:
                prev.next = st
            prev = st
            seq.append(st)
        return seq

    def save(self, path, suffix=None):
        """Sequence.save (data_structures.py:1302-1307): each state into
        path/frame_{:06d}."""
        for st in self:
            d = os.path.join(path, "frame_{:06d}".format(st.frame))
            os.makedirs(d, exist_ok=True)
            st.save(d, suffix)

    def densities_advect_fwd(self, dt=1.0, order=1, clamp="NONE"):
        """data_structures.py:1309-1320; LOCAL clamps negatives after each
        warp, GLOBAL after the whole chain."""
        mode = "NONE" if clamp is None else str(clamp).upper()
        with torch.no_grad():
            for i in range(1, len(self)):
                if mode == "LOCAL":
                    d = torch.clamp_min(self[i - 1].density_advected(order=order, dt=dt), 0.0)
                elif mode == "GLOBAL":
                    d = self[i - 1].density_advected(order=order, dt=dt)
                else:
                    d = self[i - 1].density_advected(order=order, dt=dt, clamp=clamp)
                self[i].density._d = d.detach()
            if mode == "GLOBAL":
                for i in range(1, len(self)):
                    self[i].density._d = torch.clamp_min(self[i].density._d, 0.0)

    def velocities_advect_fwd(self, dt=1.0, order=1, clamp="NONE"):
        with torch.no_grad():
            for i in range(1, len(self)):
                x, y, z = self[i - 1].velocity.warped(order=order, dt=dt, clamp=clamp)
                v = self[i].velocity
                v._x, v._y, v._z = x.detach(), y.detach(), z.detach()


def staggered_shapes(centered_shape):
    return VelocityGrid.component_shapes(centered_shape)


def numpy_centered(x, y, z):
    """Face average of numpy staggered components (for host-side checks)."""
    return np.concatenate([0.5 * (x[:, :, :, 1:] + x[:, :, :, :-1]), 0.5 * (y[:, :, 1:] + y[:, :, :-1]),
                           0.5 * (z[:, 1:] + z[:, :-1])], axis=-1)
