"""Loading helpers for the converted reference fixtures (tests/golden/*.npz)
and small shared constructions used by the tests."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def cvec(a):
    """(..., 2) float array [re, im] -> complex, flattened."""
    a = np.asarray(a, dtype=np.float64)
    return (a[..., 0] + 1j * a[..., 1]).reshape(-1)


def gate(d, prefix):
    """matchgate stored as <prefix>center / <prefix>corner, each (2,2,2) -> complex (8,)"""
    return np.concatenate([cvec(d[prefix + "center"]), cvec(d[prefix + "corner"])])


def gates(d, prefix, n):
    """n matchgates stored as <prefix>center<i> / <prefix>corner<i> -> complex (n, 8)"""
    return np.stack([np.concatenate([cvec(d[f"{prefix}center{i}"]), cvec(d[f"{prefix}corner{i}"])]) for i in range(n)])


def perms(d, n, key="perm"):
    return np.stack([np.asarray(d[f"{key}{i}"], dtype=np.int32) for i in range(n)])


def scramble_target(L):
    """Dense matrix of the index-scrambling linear map used as 'ufunc' by the
    reference's target tests (test/test_matchgate_target.c:21-36)."""
    n = 1 << L
    U = np.zeros((n, n), np.complex128)
    for i in range(n):
        U[i, ((i + 3) * 113) % n] += -1.1 + 0.8j
        U[i, ((i + 9) * 173) % n] += 0.4 - 0.7j
        U[i, i] += 0.5 + 0.1j
        U[i, ((i + 4) * 199) % n] += -0.3 + 0.2j
    return U


def haar_u2(rng):
    """Haar-random 2x2 unitary (complex Gaussian QR with phase fix, Mezzadri 2007)."""
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return (q * (d / np.abs(d))).reshape(-1)


def haar_matchgates(rng, n):
    return np.stack([np.concatenate([haar_u2(rng), haar_u2(rng)]) for _ in range(n)])


def random_perms(rng, L, nlayers):
    return np.stack([rng.permutation(L).astype(np.int32) for _ in range(nlayers)])
