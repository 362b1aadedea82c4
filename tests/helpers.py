"""Shared test helpers (oracle <-> product index mapping, tolerances)."""
import numpy as np


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def tri_map(prod_simplices, oracle_simplices):
    """oracle row -> product row, by sorted vertex triples."""
    key = {tuple(sorted(t)): i for i, t in enumerate(np.asarray(prod_simplices).tolist())}
    return np.array([key[tuple(sorted(t))] for t in np.asarray(oracle_simplices).tolist()])


def random_unit(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1)[:, None]


def perturbed_identity_jets(vertices, seed, amp=0.02):
    """Smooth map jets: identity plus a low-degree polynomial perturbation,
    vals[c,0] = f_c(v), vals[d,1+c] = d_d f_c (tangential gradient)."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((3, 3, 3)) * amp
    A = A + A.transpose(0, 2, 1)
    V = vertices
    nv = len(V)
    jets = np.zeros((3, 4, nv))
    for c in range(3):
        f = V[:, c] + np.einsum("ni,ij,nj->n", V, A[c], V)
        g = np.eye(3)[c][None, :] + 2 * V @ A[c]
        g = g - np.einsum("ni,ni->n", g, V)[:, None] * V
        jets[c, 0] = f
        jets[:, 1 + c] = g.T
    return jets
