"""Geometry helpers with the signatures of cmm_s2/core/utils.py (utils.py:9-139).

These small elementwise helpers accept NumPy arrays or torch tensors (host
set-up code and callers use them); the hot path never calls them -- its
arithmetic lives in the HIP kernels.
"""
import numpy as np
import torch


def _xp(a):
    return torch if isinstance(a, torch.Tensor) else np


def sphere2cart(phi, theta):
    """utils.py:9-13."""
    xp = _xp(phi)
    return [xp.sin(theta) * xp.cos(phi), xp.sin(theta) * xp.sin(phi), xp.cos(theta)]


def norm(x):
    xp = _xp(x[0])
    return xp.sqrt(x[0] ** 2 + x[1] ** 2 + x[2] ** 2)


def cart2sphere(XYZ):
    """utils.py:18-23: phi in [0, 2 pi)."""
    x, y, z = XYZ[0], XYZ[1], XYZ[2]
    xp = _xp(x)
    if xp is torch:
        phi = torch.remainder(torch.atan2(y, x) + 2 * np.pi, 2 * np.pi)
        theta = torch.atan2(torch.sqrt(y ** 2 + x ** 2), z)
    else:
        phi = (np.arctan2(y, x) + 2 * np.pi) % (2 * np.pi)
        theta = np.arctan2(np.sqrt(y ** 2 + x ** 2), z)
    return [phi, theta]


def dot(v1, v2):
    return v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]


def tan_proj(X, A):
    outx = (1 - X[0] ** 2) * A[0] + (-X[0] * X[1]) * A[1] + (-X[0] * X[2]) * A[2]
    outy = (-X[0] * X[1]) * A[0] + (1 - X[1] ** 2) * A[1] + (-X[1] * X[2]) * A[2]
    outz = (-X[0] * X[2]) * A[0] + (-X[1] * X[2]) * A[1] + (1 - X[2] ** 2) * A[2]
    return [outx, outy, outz]


def div_norm(x):
    """utils.py:36-38: (N,3) -> (3,N) normalised."""
    xp = _xp(x)
    n = xp.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2 + x[:, 2] ** 2)
    return xp.stack([x[:, 0] / n, x[:, 1] / n, x[:, 2] / n])


def det_vec(A):
    v_1, v_2, v_3 = A[0], A[1], A[2]
    return (v_1[:, 0] * v_2[:, 1] * v_3[:, 2] + v_2[:, 0] * v_3[:, 1] * v_1[:, 2]
            + v_1[:, 1] * v_2[:, 2] * v_3[:, 0]
            - (v_3[:, 0] * v_2[:, 1] * v_1[:, 2] + v_3[:, 1] * v_2[:, 2] * v_1[:, 0]
               + v_2[:, 0] * v_1[:, 1] * v_3[:, 2]))


def bary_coords(v_1, v_2, v_3, v):
    xp = _xp(v)
    d = det_vec([v_1, v_2, v_3])
    return xp.stack([det_vec([v, v_2, v_3]) / d, det_vec([v_1, v, v_3]) / d,
                     det_vec([v_1, v_2, v]) / d], 1)


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def identity(xyz):
    return [xyz[0], xyz[1], xyz[2]]


def grad_x(xyz):
    xp = _xp(xyz)
    return xp.stack(tan_proj(xyz, [0 * xyz[0] + 1, 0 * xyz[1], 0 * xyz[2]]))


def grad_y(xyz):
    xp = _xp(xyz)
    return xp.stack(tan_proj(xyz, [0 * xyz[0], 0 * xyz[1] + 1, 0 * xyz[2]]))


def grad_z(xyz):
    xp = _xp(xyz)
    return xp.stack(tan_proj(xyz, [0 * xyz[0], 0 * xyz[1], 0 * xyz[2] + 1]))


def d_Proj(X, A):
    return tan_proj(X, A)


def Rot_x(alpha):
    return np.array([[1, 0, 0], [0, np.cos(alpha), -np.sin(alpha)], [0, np.sin(alpha), np.cos(alpha)]])


def Rot_y(alpha):
    return np.array([[np.cos(alpha), 0, -np.sin(alpha)], [0, 1, 0], [np.sin(alpha), 0, np.cos(alpha)]])


def Rot_z(alpha):
    return np.array([[np.cos(alpha), -np.sin(alpha), 0], [np.sin(alpha), np.cos(alpha), 0], [0, 0, 1]])


def rotate(L, XYZ):
    return [L[0, 0] * XYZ[0] + L[0, 1] * XYZ[1] + L[0, 2] * XYZ[2],
            L[1, 0] * XYZ[0] + L[1, 1] * XYZ[1] + L[1, 2] * XYZ[2],
            L[2, 0] * XYZ[0] + L[2, 1] * XYZ[1] + L[2, 2] * XYZ[2]]


def Mmul(A, B):
    return np.matmul(A, B)


def pi_proj(y, n):
    """utils.py:136-139: lift tangent-plane points y (N,3) at n (N,3) to the
    sphere; returns (3,N)."""
    xp = _xp(y)
    yn = dot(y.T, n.T)
    tt = -yn + xp.sqrt(yn ** 2 - dot(y.T, y.T) + 1)
    return (y + tt[:, None] * n).T
