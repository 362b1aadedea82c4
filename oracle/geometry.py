"""Geometry primitives (oracle).  Restates cmm_s2/core/utils.py.

Pinned by tests/test_oracle_geometry.py through closed forms: q = sum b_i v_i
for bary_coords, unit length of pi_proj / sphere_tan_proj outputs, tangency,
sphere2cart/cart2sphere round trips.
"""
import numpy as np


def sphere2cart(phi, theta):
    """utils.py:9-13 -- colatitude theta, longitude phi."""
    x = np.sin(theta) * np.cos(phi)
    y = np.sin(theta) * np.sin(phi)
    z = np.cos(theta)
    return [x, y, z]


def cart2sphere(XYZ):
    """utils.py:18-23 -- phi in [0, 2pi), theta = atan2(rho, z)."""
    x, y, z = XYZ[0], XYZ[1], XYZ[2]
    phi = (np.arctan2(y, x) + 2 * np.pi) % (2 * np.pi)
    theta = np.arctan2(np.sqrt(y ** 2 + x ** 2), z)
    return [phi, theta]


def dot(v1, v2):
    """utils.py:25 -- contraction over axis 0."""
    return v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]


def tan_proj(X, A):
    """utils.py:29-34 -- (I - X X^T) A."""
    outx = (1 - X[0] ** 2) * A[0] + (-X[0] * X[1]) * A[1] + (-X[0] * X[2]) * A[2]
    outy = (-X[0] * X[1]) * A[0] + (1 - X[1] ** 2) * A[1] + (-X[1] * X[2]) * A[2]
    outz = (-X[0] * X[2]) * A[0] + (-X[1] * X[2]) * A[1] + (1 - X[2] ** 2) * A[2]
    return [outx, outy, outz]


def div_norm(x):
    """utils.py:36-38 -- x is (N,3); returns (3,N) normalised columns."""
    Norm = np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2 + x[:, 2] ** 2)
    return np.array([x[:, 0] / Norm, x[:, 1] / Norm, x[:, 2] / Norm])


def det_vec(A):
    """utils.py:49-56 -- determinant of [a b c] (columns), rows are points (N,3)."""
    v_1, v_2, v_3 = A[0], A[1], A[2]
    return (v_1[:, 0] * v_2[:, 1] * v_3[:, 2] + v_2[:, 0] * v_3[:, 1] * v_1[:, 2]
            + v_1[:, 1] * v_2[:, 2] * v_3[:, 0]
            - (v_3[:, 0] * v_2[:, 1] * v_1[:, 2] + v_3[:, 1] * v_2[:, 2] * v_1[:, 0]
               + v_2[:, 0] * v_1[:, 1] * v_3[:, 2]))


def bary_coords(v_1, v_2, v_3, v):
    """utils.py:58-67 -- homogeneous (Cramer) barycentrics; q = sum b_i v_i."""
    denom = det_vec([v_1, v_2, v_3])
    return np.stack([det_vec([v, v_2, v_3]) / denom,
                     det_vec([v_1, v, v_3]) / denom,
                     det_vec([v_1, v_2, v]) / denom], axis=1)


def cross(a, b):
    """utils.py:103-104."""
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def grad_x(xyz):
    """utils.py:87-88 -- tangential gradient of x."""
    return np.array(tan_proj(xyz, np.array([0 * xyz[0] + 1, 0 * xyz[1], 0 * xyz[2]])))


def grad_y(xyz):
    """utils.py:90-91."""
    return np.array(tan_proj(xyz, np.array([0 * xyz[0], 0 * xyz[1] + 1, 0 * xyz[2]])))


def grad_z(xyz):
    """utils.py:93-94."""
    return np.array(tan_proj(xyz, np.array([0 * xyz[0], 0 * xyz[1], 0 * xyz[2] + 1])))


def pi_proj(y, n):
    """utils.py:136-139 -- lift tangent-plane point y at n back to the sphere.
    y, n are (N,3); returns (3,N)."""
    tt = -dot(y.T, n.T) + np.sqrt(dot(y.T, n.T) ** 2 - dot(y.T, y.T) + 1)
    return (y + tt[:, None] * n).T


def sphere_tan_proj(a, b):
    """mesh_functions.py:320-329 -- unit tangential part of a at base point b.
    a, b are (N,3); returns (3,N)."""
    c = dot(b.T, a.T)
    d = np.sqrt(dot(a.T, a.T) - 2 * c ** 2 + dot(b.T, b.T) * c ** 2)
    return np.array([(a[:, 0] - c * b[:, 0]) / d, (a[:, 1] - c * b[:, 1]) / d,
                     (a[:, 2] - c * b[:, 2]) / d])


def rotation_matrix(axis, angle):
    """Rodrigues rotation (test support: solid-body-rotation known answers)."""
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K
