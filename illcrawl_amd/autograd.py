"""torch.autograd binding of the projective smoothing path.

``projected_image(pos, w, h, camera)`` renders the exact-mode quartic image of
particles at world positions ``pos`` (n, 3) with weights ``w`` and smoothing
lengths ``h`` through the C ABI (project_particles, grid build, gather,
retrieve), and its backward calls ilc_psg_backward, then carries the
projected-frame position gradient back to world coordinates through the
projection's FP32 matrix (projective_smoothing_reconstruction.cpp:175-205).
Particles that the depth mask removes get dL/dw = 0.  The reference renders
only; this binding exists so that gradients can be checked against finite
differences.
"""
from __future__ import annotations

import torch

from .backend import (ProjectiveSmoothingGrid, camera_pixel_points, camera_projection,
                      project_particles)


class ProjectedImage(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, w, h, camera, kernel, scale, target_particles_per_cell):
        dev = pos.device
        p4 = torch.cat([pos.detach().float(), w.detach().float().reshape(-1, 1)], 1).contiguous()
        proj = project_particles(camera, p4, dev.index)
        # depth mask: project unit weights and see which become 0
        ones = torch.cat([pos.detach().float(), torch.ones_like(p4[:, :1])], 1).contiguous()
        kept = project_particles(camera, ones, dev.index)[:, 3] != 0
        psg = ProjectiveSmoothingGrid(dev.index, target_particles_per_cell, mode="exact",
                                      kernel=kernel)
        psg.init_backend(proj.shape[0])
        psg.setup_projected_particles(proj, h.detach().float())
        pts = camera_pixel_points(camera, dev.index)
        psg.setup_evaluation_points(pts, layout_nx=int(camera["nx"]))
        psg.run()
        img = psg.retrieve_results(float(scale))
        ctx.psg, ctx.scale, ctx.kept = psg, float(scale), kept
        ctx.keep_pts = pts
        minv, _ = camera_projection(camera)
        ctx.jac = torch.from_numpy(minv[:2].copy()).to(dev)   # d(x', y') / d(x, y, z)
        ctx.dtypes = (pos.dtype, w.dtype, h.dtype)
        return img.reshape(int(camera.get("ny", camera["nx"])), int(camera["nx"]))

    @staticmethod
    def backward(ctx, grad_img):
        gw, gh, gx, gy = ctx.psg.backward(grad_img.contiguous().reshape(-1), ctx.scale)
        gw = torch.where(ctx.kept, gw, torch.zeros_like(gw))
        gpos = gx[:, None] * ctx.jac[0][None, :] + gy[:, None] * ctx.jac[1][None, :]
        tp, tw, th = ctx.dtypes
        return gpos.to(tp), gw.to(tw), gh.to(th), None, None, None, None


def projected_image(pos, w, h, camera, kernel="quartic", scale=1.0,
                    target_particles_per_cell=1000):
    """Image (ny, nx) = scale * sum_j w_j W(|x_pixel - x'_j|, h_j), differentiable
    in pos, w and h."""
    return ProjectedImage.apply(pos, w, h, camera, kernel, scale, target_particles_per_cell)
