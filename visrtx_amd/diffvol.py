"""Differentiable emission-absorption volume rendering (torch.autograd).

``VolumeIntegral`` wraps the HIP forward/adjoint pair of ``include/vrx_hip.h``
(``vrx_diffvol_forward`` / ``vrx_diffvol_backward``): the forward is the
volume integral of the reference's ray march (gpu/volumeIntegration.h:64-103
in the VisRTX tree, jitter off), the backward is a hand-written adjoint.  By
default it replays the march in forward order and takes the suffix sums it
needs from the forward's FP64 totals; ``adjoint="replay"`` replays the march
backward from transmittance checkpoints stored every 64 steps.
Gradients flow to the density grid and to the 256-entry RGBA transfer
function LUT.  Torch provides device memory and the stream only.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N


@dataclass
class GridSpec:
    dims: tuple                 # (nx, ny, nz)
    origin: tuple = (0.0, 0.0, 0.0)
    spacing: tuple = (1.0, 1.0, 1.0)
    value_range: tuple = (0.0, 1.0)
    unit_distance: float = 1.0
    step: float = 0.5           # world units
    early_stop: bool = True     # stop at opacity 0.99, as the reference does
    checkpoint: int = 64
    # rays are images of image[0] x image[1] pixels (vrx_make_rays order);
    # lets the adjoint tile its workgroups.  None = unstructured rays.
    image: tuple | None = None
    # backward schedule: "suffix" (forward-order replay, FP64 suffix sums) or
    # "replay" (backward replay from the transmittance checkpoints)
    adjoint: str = "suffix"

    @property
    def max_steps(self) -> int:
        ext = [(d - 1) * s for d, s in zip(self.dims, self.spacing)]
        diag = math.sqrt(sum(e * e for e in ext))
        return int(math.ceil(diag / self.step)) + 2

    def struct(self, density: torch.Tensor, lut: torch.Tensor) -> N.vrx_diffvol:
        v = N.vrx_diffvol()
        v.d_density = density.data_ptr()
        v.dims[:] = list(self.dims)
        v.origin[:] = list(self.origin)
        v.spacing[:] = list(self.spacing)
        v.d_lut = lut.data_ptr()
        v.value_range[:] = list(self.value_range)
        v.unit_distance = self.unit_distance
        v.step = self.step
        v.early_stop = 1 if self.early_stop else 0
        v.checkpoint = self.checkpoint
        v.max_steps = self.max_steps
        if self.image is not None:
            v.image_w, v.image_h = int(self.image[0]), int(self.image[1])
        if self.adjoint not in ("suffix", "replay"):
            raise ValueError(f"adjoint must be 'suffix' or 'replay', not {self.adjoint!r}")
        v.adjoint = N.VRX_ADJOINT_SUFFIX if self.adjoint == "suffix" else N.VRX_ADJOINT_REPLAY
        return v


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _check_tensor(t: torch.Tensor, name: str):
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous float32 CUDA tensor")


class VolumeIntegral(torch.autograd.Function):
    """rgba[i] = (C.rgb, O) of ray i through density (nz, ny, nx) classified by lut (256, 4)."""

    @staticmethod
    def forward(ctx, density, lut, org, dirs, spec: GridSpec):
        for t, n in [(density, "density"), (lut, "lut"), (org, "org"), (dirs, "dirs")]:
            _check_tensor(t, n)
        L = N.hip()
        n = org.shape[0]
        v = spec.struct(density, lut)
        rgba = torch.empty((n, 4), device=org.device, dtype=torch.float32)
        ck = torch.empty(L.vrx_diffvol_ckpt_floats(C.byref(v), n), device=org.device,
                         dtype=torch.float32)
        N.check(L.vrx_diffvol_forward(C.byref(v), n, org.data_ptr(), dirs.data_ptr(),
                                      rgba.data_ptr(), ck.data_ptr(), _stream()),
                "vrx_diffvol_forward")
        ctx.save_for_backward(density, lut, org, dirs, rgba, ck)
        ctx.spec = spec
        return rgba

    @staticmethod
    def backward(ctx, grad_rgba):
        density, lut, org, dirs, rgba, ck = ctx.saved_tensors
        spec = ctx.spec
        g = grad_rgba.contiguous().float()
        gd = torch.zeros_like(density)
        gl = torch.zeros_like(lut)
        L = N.hip()
        v = spec.struct(density, lut)
        N.check(L.vrx_diffvol_backward(C.byref(v), org.shape[0], org.data_ptr(), dirs.data_ptr(),
                                       rgba.data_ptr(), g.data_ptr(), ck.data_ptr(),
                                       gd.data_ptr(), gl.data_ptr(), _stream()),
                "vrx_diffvol_backward")
        return gd, gl, None, None, None


def render(density, lut, org, dirs, spec: GridSpec):
    return VolumeIntegral.apply(density, lut, org, dirs, spec)


def camera(pos, target, up=(0.0, 1.0, 0.0), fovy=math.radians(40.0), aspect=1.0) -> N.vrx_camera:
    c = N.vrx_camera()
    d = (C.c_float * 3)(*[float(t - p) for t, p in zip(target, pos)])
    P = (C.c_float * 3)(*[float(x) for x in pos])
    U = (C.c_float * 3)(*[float(x) for x in up])
    N.check(N.hip().vrx_camera_perspective(P, d, U, fovy, aspect, None, C.byref(c)))
    return c


def orbit_cameras(n: int, center=(0.5, 0.5, 0.5), radius=2.0, seed: int = 5, fovy=40.0):
    """n seeded views on a sphere around ``center``."""
    rng = np.random.default_rng(seed)
    cams = []
    for i in range(n):
        phi = 2.0 * math.pi * (i + rng.uniform(0, 0.5)) / n
        z = rng.uniform(-0.6, 0.6)
        r = math.sqrt(1 - z * z)
        pos = (center[0] + radius * r * math.cos(phi), center[1] + radius * z,
               center[2] + radius * r * math.sin(phi))
        cams.append(camera(pos, center, fovy=math.radians(fovy)))
    return cams


def make_rays(cams, width: int, height: int, device="cuda"):
    n = len(cams) * width * height
    arr = (N.vrx_camera * len(cams))(*cams)
    org = torch.empty((n, 3), device=device, dtype=torch.float32)
    dirs = torch.empty((n, 3), device=device, dtype=torch.float32)
    N.check(N.hip().vrx_make_rays(arr, len(cams), width, height, org.data_ptr(), dirs.data_ptr(),
                                  _stream()), "vrx_make_rays")
    return org, dirs


def linear_lut(alpha_max=0.5) -> np.ndarray:
    """256 RGBA entries, every channel linear in the index (no interior kinks)."""
    t = np.arange(256, dtype=np.float64) / 255.0
    lut = np.stack([0.2 + 0.6 * t, 0.9 - 0.7 * t, 0.3 + 0.4 * t, alpha_max * t], 1)
    return lut.astype(np.float32)


def colormap_lut(colors: np.ndarray, opacities: np.ndarray) -> np.ndarray:
    """256-entry RGBA LUT from colour and opacity control points spread evenly
    over the value range (the transferFunction1D discretisation,
    scene/volume/TransferFunction1D.cpp:101-150)."""
    p = np.arange(256, dtype=np.float64) / 255.0
    cx = np.linspace(0.0, 1.0, len(colors))
    ox = np.linspace(0.0, 1.0, len(opacities))
    lut = np.empty((256, 4), np.float64)
    for k in range(3):
        lut[:, k] = np.interp(p, cx, colors[:, k])
    lut[:, 3] = np.interp(p, ox, opacities)
    return lut.astype(np.float32)
