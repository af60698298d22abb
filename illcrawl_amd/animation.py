"""Animations of projected views: the reference's camera steppers and its
animation target, with every frame rendered in one batched launch.

The reference renders frame i of an animation by resetting the camera to its
initial state, rotating it about a centre (camera_movement::rotation_around_point
and dual_axis_rotation_around_point, animation.cpp:98-178) and running one
projection per frame (frame_rendering::integrated_projection,
animation.cpp:181-210), one frame after another, or frames distributed over MPI
ranks (animation.cpp:80-96).  Here the frames are the views of one
BatchRenderer batch: one persistent gather launch renders all of them.

camera::rotate (camera.cpp:107-122) rotates the position about the centre, the
look-at vector and the two screen basis vectors with the same matrix.  A view
here is (position, look_at, roll, width, nx, ny, depth), whose basis the
library derives from look_at and roll as camera::update_basis_vectors does.
So a rotated frame gets the roll that makes that derivation give the rotated
basis.  The two agree to rounding (about 1e-16), not bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

from . import backend as B


def rotation_matrix(axis, alpha):
    """Rotation by alpha (radians) about the unit vector axis, right-handed
    (the Rodrigues form of math::matrix_create_rotation_matrix, math.hpp:109-140)."""
    a = np.asarray(axis, np.float64)
    c, s = math.cos(alpha), math.sin(alpha)
    k = 1.0 - c
    x, y, z = a
    return np.array([[x * x * k + c, x * y * k - z * s, x * z * k + y * s],
                     [x * y * k + z * s, y * y * k + c, y * z * k - x * s],
                     [x * z * k - y * s, y * z * k + x * s, z * z * k + c]])


def _unrolled_axis(look_at):
    # camera::update_basis_vectors: v1 = z x look_at, or y x look_at when
    # looking along +z exactly
    la = np.asarray(look_at, np.float64)
    v = np.array([0.0, 0.0, 1.0])
    if la[0] == 0.0 and la[1] == 0.0 and la[2] == 1.0:
        v = np.array([0.0, 1.0, 0.0])
    u = np.cross(v, la)
    return u / np.linalg.norm(u)


def rotate_view(view, R, center):
    """camera::rotate: the view rotated by the matrix R about center."""
    b0, _, _, _ = B.camera_basis(view)
    c = np.asarray(center, np.float64)
    pos = R @ (np.asarray(view["position"], np.float64) - c) + c
    la = R @ np.asarray(view["look_at"], np.float64)
    t = R @ b0
    u = _unrolled_axis(la)
    roll = math.atan2(float(np.dot(t, np.cross(la, u))), float(np.dot(t, u)))
    out = dict(view)
    out.update(position=pos, look_at=la, roll=roll)
    return out


class RotationAroundPoint:
    """camera_movement::rotation_around_point (animation.cpp:98-140): frame i
    of n is the initial view rotated by i * range / n about the axis through
    center; range in degrees.  ``matrix_creator(alpha)`` replaces the axis
    rotation when given."""

    def __init__(self, center, axis, initial_view, rotation_range_degree,
                 matrix_creator=None):
        self.center = np.asarray(center, np.float64)
        self.range = (2.0 * math.pi / 360.0) * float(rotation_range_degree)
        self.initial = dict(initial_view)
        ax = np.asarray(axis, np.float64) if axis is not None else None
        self.matrix_creator = matrix_creator or (lambda alpha: rotation_matrix(ax, alpha))

    def __call__(self, frame_id, num_frames):
        angle_per_frame = self.range / float(num_frames)
        R = self.matrix_creator(frame_id * angle_per_frame)
        return rotate_view(self.initial, R, self.center)


class DualAxisRotationAroundPoint:
    """camera_movement::dual_axis_rotation_around_point (animation.cpp:142-178):
    the rotation M_phi(alpha) M_theta(alpha * theta_range / phi_range), ranges
    in degrees."""

    def __init__(self, center, phi_axis, theta_axis, initial_view, rotation_range_phi_degree,
                 rotation_range_theta_degree):
        ratio = float(rotation_range_theta_degree) / float(rotation_range_phi_degree)
        pa = np.asarray(phi_axis, np.float64)
        ta = np.asarray(theta_axis, np.float64)

        def creator(alpha_phi):
            return rotation_matrix(pa, alpha_phi) @ rotation_matrix(ta, alpha_phi * ratio)

        self._rotation = RotationAroundPoint(center, None, initial_view,
                                             rotation_range_phi_degree, creator)

    def __call__(self, frame_id, num_frames):
        return self._rotation(frame_id, num_frames)


def illcrawl_render_animation_stepper(center, initial_view, phi_speed):
    """The animation target of illcrawl_render (illcrawl_render.cpp:189-228):
    phi about y at phi_speed turns over the animation, theta about x at one
    turn."""
    if phi_speed <= 0:
        raise ValueError("animation.phi_speed must be > 0.")
    return DualAxisRotationAroundPoint(center, (0.0, 1.0, 0.0), (1.0, 0.0, 0.0), initial_view,
                                       phi_speed * 360.0, 360.0)


def animation_views(stepper, num_frames, first_frame=0, end_frame=None):
    """Views of frames first_frame .. end_frame - 1 of a num_frames animation
    (animation.cpp:41-78 with total_frame_range = num_frames)."""
    end_frame = num_frames if end_frame is None else end_frame
    if end_frame < first_frame:
        raise ValueError("end_frame < first_frame")
    return [stepper(i, num_frames) for i in range(first_frame, end_frame)]


def render_animation(renderer, views, f64=False):
    """All frames in one batched render: ``renderer`` is a BatchRenderer whose
    particles (and Gaunt table, for X-ray views) are set.  Returns the frames
    as (frames, ny, nx) float32 (float64 accumulators with f64=True), the
    reference's multi_array{nx, ny, frames} layout."""
    renderer.set_views(views)
    renderer.render()
    get = renderer.image_f64 if f64 else renderer.image
    return np.stack([get(k).cpu().numpy() for k in range(len(views))], 0)
