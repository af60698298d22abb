"""Camera steppers of animations (host math, no GPU): the rotated view's
basis, as the library derives it from look_at and roll, equals the
reference's rotated basis (camera::rotate, camera.cpp:107-122)."""
import math

import numpy as np
import pytest

A = pytest.importorskip("illcrawl_amd.animation")
B = pytest.importorskip("illcrawl_amd.backend")


def view0():
    return dict(position=np.array([10.0, 20.0, -30.0]), look_at=np.array([0.0, 0.6, 0.8]),
                roll=0.4, width=50.0, nx=64, ny=48, depth=100.0)


def test_rotation_matrix_closed_forms():
    R = A.rotation_matrix((0.0, 0.0, 1.0), math.pi / 2)
    np.testing.assert_allclose(R @ [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], atol=1e-15)
    ax = np.array([1.0, 2.0, 2.0]) / 3.0
    R = A.rotation_matrix(ax, 0.7)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-15)
    np.testing.assert_allclose(R @ ax, ax, atol=1e-15)
    assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)


@pytest.mark.parametrize("frame", [0, 1, 5, 11])
def test_rotated_view_basis_equals_rotated_basis(frame):
    v = view0()
    center = np.array([1.0, 2.0, 3.0])
    st = A.DualAxisRotationAroundPoint(center, (0, 1, 0), (1, 0, 0), v, 3 * 360.0, 360.0)
    w = st(frame, 12)
    alpha = frame * (3 * 2 * math.pi) / 12
    R = A.rotation_matrix((0, 1, 0), alpha) @ A.rotation_matrix((1, 0, 0), alpha / 3)
    b0, b1, mn, ps = B.camera_basis(v)
    c0, c1, cmn, cps = B.camera_basis(w)
    np.testing.assert_allclose(c0, R @ b0, atol=1e-13)
    np.testing.assert_allclose(c1, R @ b1, atol=1e-13)
    np.testing.assert_allclose(w["position"], R @ (v["position"] - center) + center, atol=1e-12)
    # the screen corner follows: min = position - nx/2 ps b0 - ny/2 ps b1
    np.testing.assert_allclose(cmn, R @ (mn - center) + center, atol=1e-11)
    assert cps == ps


def test_animation_frames_and_ranges():
    v = view0()
    st = A.RotationAroundPoint((0, 0, 0), (0, 0, 1), v, 360.0)
    frames = A.animation_views(st, 8)
    assert len(frames) == 8
    np.testing.assert_allclose(frames[0]["position"], v["position"], atol=1e-12)
    part = A.animation_views(st, 8, 3, 6)
    for k, f in enumerate(part):
        np.testing.assert_allclose(f["position"], frames[3 + k]["position"], atol=0)
    # a quarter turn about z maps (x, y) to (-y, x)
    np.testing.assert_allclose(frames[2]["position"], [-20.0, 10.0, -30.0], atol=1e-12)
    with pytest.raises(ValueError):
        A.illcrawl_render_animation_stepper((0, 0, 0), v, 0.0)
