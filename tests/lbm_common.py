"""Shared test helpers: lattice constants, symmetry group, comparison."""
import itertools

import numpy as np

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1])
C = np.stack([CX, CY, CZ], axis=1)
OPP = np.array([0] + [q + 1 if q % 2 == 1 else q - 1 for q in range(1, 27)])
W = np.array([np.prod([(2 / 3 if c == 0 else 1 / 6) for c in C[q]]) for q in range(27)])


def qindex(c):
    return int(np.where((C == np.asarray(c)).all(axis=1))[0][0])


def cube_symmetries():
    """The 48 signed permutation matrices of the cube group."""
    mats = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product([1, -1], repeat=3):
            M = np.zeros((3, 3), dtype=int)
            for i in range(3):
                M[i, perm[i]] = signs[i]
            mats.append(M)
    return mats


def perm_of(M):
    """q -> index of M c_q."""
    return np.array([qindex(M @ C[q]) for q in range(27)])


def rel_max(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def stream_periodic(f):
    """Pure streaming S on a periodic lattice: f_q(r + c_q) <- f_q(r); f is (27, X, Y, Z)."""
    out = np.empty_like(f)
    for q in range(27):
        out[q] = np.roll(f[q], shift=tuple(C[q]), axis=(0, 1, 2))
    return out
