"""Brute-force dense references (test infrastructure): the full 2^L x 2^L matrix
of a gate program built by tensor contraction, so that its index conventions are
derived independently of the statevector kernels' bit arithmetic."""
import numpy as np


def matchgate_matrix(g):
    """4x4 matrix in the basis |q_i q_j> = 00, 01, 10, 11 (reference
    test/rqcopt_matfree/util.py:100-113: center on {01,10}, corner on {00,11})."""
    g = np.asarray(g)
    M = np.zeros((4, 4), np.complex128)
    M[1, 1], M[1, 2], M[2, 1], M[2, 2] = g[0], g[1], g[2], g[3]
    M[0, 0], M[0, 3], M[3, 0], M[3, 3] = g[4], g[5], g[6], g[7]
    return M


def apply_dense_gate(M, i, j, L, psi):
    """Apply the 4x4 M to wires (i, j) of psi (qubit 0 = most significant)."""
    t = np.asarray(psi, np.complex128).reshape((2,) * L)
    t = np.moveaxis(t, (i, j), (0, 1)).reshape(4, -1)
    t = (M @ t).reshape((2, 2) + (2,) * (L - 2))
    return np.moveaxis(t, (0, 1), (i, j)).reshape(-1)


def program_matrix(gates, param, wires, L):
    S = 1 << L
    W = np.eye(S, dtype=np.complex128)
    for g in range(len(param)):
        M = matchgate_matrix(gates[param[g]])
        W = np.stack([apply_dense_gate(M, wires[2 * g], wires[2 * g + 1], L, W[:, b]) for b in range(S)], axis=1)
    return W


def target_f(U, gates, param, wires, L):
    W = program_matrix(gates, param, wires, L)
    return -np.sum(np.conj(U) * W)


def fd_gradient(U, gates, param, wires, L, h=1e-5):
    """Holomorphic gradient df/dV by central differences."""
    grad = np.zeros_like(gates)
    for p in range(gates.shape[0]):
        for e in range(8):
            gp = gates.copy(); gp[p, e] += h
            gm = gates.copy(); gm[p, e] -= h
            grad[p, e] = (target_f(U, gp, param, wires, L) - target_f(U, gm, param, wires, L)) / (2 * h)
    return grad
