"""Synthetic workloads of the matchgate-brickwall optimisation: seeded Haar
matchgates, brickwall layouts and the 1D Fermi-Hubbard target exp(-iHt).

Randomness comes from numpy's PCG64 generator (``np.random.default_rng(seed)``),
so every workload is reproducible from its seed.
"""
import numpy as np


def haar_u2(rng):
    """Haar-random 2x2 unitary: QR of a complex Gaussian matrix with the phases of
    R's diagonal moved into Q (Mezzadri, Notices AMS 54, 592 (2007))."""
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return (q * (d / np.abs(d))).reshape(-1)


def haar_matchgates(rng, n):
    """n matchgates with independent Haar U(2) center (odd parity) and corner (even parity) blocks."""
    return np.stack([np.concatenate([haar_u2(rng), haar_u2(rng)]) for _ in range(n)])


def open_brickwall(L, depth):
    """Open-boundary brickwall with one independent gate per instance: layer k acts on
    (0,1),(2,3),... for even k and on (1,2),(3,4),... for odd k.
    Returns (param, wires, ngates) with param[g] = g."""
    wires = []
    for k in range(depth):
        for a in range(k % 2, L - 1, 2):
            wires += [a, a + 1]
    G = len(wires) // 2
    return np.arange(G, dtype=np.int32), np.asarray(wires, dtype=np.int32), G


def fermi_hubbard_1d_unitary(nsites, J=1.0, U=4.0, t=0.25):
    """exp(-i H t), dense complex128, for the open-chain Fermi-Hubbard model
        H = -J sum_{j,s} (c+_{j,s} c_{j+1,s} + h.c.) + U sum_j n_{j,up} n_{j,down}
    with spin-up of site j on qubit 2j and spin-down on qubit 2j+1 (qubit 0 = most
    significant bit), Jordan-Wigner in qubit order: the same-spin hop between qubits
    2j+s and 2j+2+s carries a Z on qubit 2j+1+s.  H conserves (N_up, N_down), so it is
    exponentiated block by block with a Hermitian eigendecomposition."""
    L = 2 * nsites
    S = 1 << L
    idx = np.arange(S)

    def bit(s, q):
        return (s >> (L - 1 - q)) & 1

    nup = sum(bit(idx, 2 * j) for j in range(nsites))
    ndn = sum(bit(idx, 2 * j + 1) for j in range(nsites))
    diag = U * sum(bit(idx, 2 * j) * bit(idx, 2 * j + 1) for j in range(nsites)).astype(np.float64)
    W = np.zeros((S, S), np.complex128)
    for nu in range(nsites + 1):
        for nd in range(nsites + 1):
            states = idx[(nup == nu) & (ndn == nd)]
            pos = {int(s): k for k, s in enumerate(states)}
            n = len(states)
            H = np.diag(diag[states])
            for k, s in enumerate(states):
                s = int(s)
                for j in range(nsites - 1):
                    for sp in range(2):
                        a, b = 2 * j + sp, 2 * j + 2 + sp
                        if bit(s, a) != bit(s, b):
                            s2 = s ^ (1 << (L - 1 - a)) ^ (1 << (L - 1 - b))
                            sign = -1.0 if bit(s, 2 * j + 1 + sp) else 1.0
                            H[pos[s2], k] += -J * sign
            w, v = np.linalg.eigh(H)
            blk = (v * np.exp(-1j * t * w)) @ v.conj().T
            W[np.ix_(states, states)] = blk
    return W


def sweep_flops(L, ngates, ntarget_gates=0):
    """Algorithmic FP64/FP32 flops of one f + gradient + HVP evaluation over all
    2^L basis columns: 18 complex MACs per amplitude per gate (forward 2, gradient 4,
    tangent forward/adjoint and the two outer products 12), 1 for the trace and 2 per
    matrix-free target gate; 8 real flops per complex MAC."""
    S = 1 << L
    return 8.0 * S * S * (18 * ngates + 2 * ntarget_gates + 1)


def brickwall_job(rng, L, depth, niter, target_factor=3):
    """One optimisation job: an open brickwall of `depth` layers with seeded Haar
    gates, and a target that is a seeded random open brickwall of target_factor x depth
    layers (a circuit target, applied matrix-free)."""
    param, wires, G = open_brickwall(L, depth)
    gates = haar_matchgates(rng, G)
    _, twires, Gt = open_brickwall(L, target_factor * depth)
    tgates = haar_matchgates(rng, Gt)
    return {"L": L, "depth": depth, "param": param, "wires": wires, "G": G, "gates": gates,
            "target_gates": tgates, "target_wires": twires, "Gt": Gt, "niter": int(niter)}


def batched_instances(n=64, seed=0, Ls=(8, 10, 12, 14), depths=(3, 5, 7, 9, 11), niter=20):
    """The batched workload: n independent instances with L uniform in Ls, odd depth
    uniform in depths, targets random brickwalls of 3x the depth."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = int(rng.choice(Ls))
        d = int(rng.choice(depths))
        out.append(brickwall_job(rng, L, d, niter))
    return out


def served_stream(n, seed=1, rate=200.0, Ls=(8, 10, 12, 14, 16), depths=tuple(range(3, 16, 2)),
                  budget=(5, 50)):
    """The served stream: Poisson arrivals (mean `rate` jobs/s) of jobs with L uniform
    in Ls, odd depth uniform in depths, iteration budget uniform in [budget[0],
    budget[1]], targets random brickwalls of 3x the depth."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        L = int(rng.choice(Ls))
        d = int(rng.choice(depths))
        it = int(rng.integers(budget[0], budget[1] + 1))
        job = brickwall_job(rng, L, d, it)
        job["arrival"] = t
        out.append(job)
    return out


def matchgate_from_matrix(M):
    """The 8 matchgate entries of a parity-conserving 4x4 matrix in the basis |q_i q_j> =
    00, 01, 10, 11: center block on {01, 10} (entries 0-3), corner block on {00, 11} (4-7)."""
    return np.array([M[1, 1], M[1, 2], M[2, 1], M[2, 2], M[0, 0], M[0, 3], M[3, 0], M[3, 3]], np.complex128)


def xxz_bond_matrix(tau, hi=0.0, hj=0.0):
    """exp(-i tau h) for the bond term h = (X X + Y Y)/2 + 2 Z Z + hi Z_i + hj Z_j
    (4x4, basis 00, 01, 10, 11 of the bond's qubits i < j)."""
    h = np.zeros((4, 4))
    h[1, 2] = h[2, 1] = 1.0                       # (XX + YY)/2 flips 01 <-> 10
    h += 2.0 * np.diag([1.0, -1.0, -1.0, 1.0])    # 2 ZZ
    h += hi * np.diag([1.0, 1.0, -1.0, -1.0]) + hj * np.diag([1.0, -1.0, 1.0, -1.0])
    w, v = np.linalg.eigh(h)
    return (v * np.exp(-1j * tau * w)) @ v.conj().T


def xxz_fields(L, seed=0):
    """Random fields h_j ~ U(-1, 1) of the L=16 workload, seeded."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, L)


def xxz_trotter_target(L, steps=50, t=1.0, seed=0):
    """Second-order Trotter circuit of H = sum_j [(X_j X_{j+1} + Y_j Y_{j+1})/2 + 2 Z_j Z_{j+1}]
    + sum_j h_j Z_j (open chain, h_j ~ U(-1, 1) seeded) for time t in `steps` steps, as
    parity-conserving nearest-neighbour matchgates.  H = H_even + H_odd by bond parity;
    each site's field is split equally between its bonds (a boundary site gives all of it
    to its one bond).  The steps' half layers are merged:
    E(dt/2) [O(dt) E(dt)]^(steps-1) O(dt) E(dt/2), i.e. 2 steps + 1 layers.
    Returns (gates (Gt, 8), wires (2 Gt,), fields)."""
    h = xxz_fields(L, seed)
    dt = t / steps
    nb = np.array([2.0] * L)
    nb[0] = nb[-1] = 1.0                            # bonds per site
    layers = [(0, dt / 2)] + [(1, dt), (0, dt)] * (steps - 1) + [(1, dt), (0, dt / 2)]
    gates, wires = [], []
    for start, tau in layers:
        for a in range(start, L - 1, 2):
            M = xxz_bond_matrix(tau, h[a] / nb[a], h[a + 1] / nb[a + 1])
            gates.append(matchgate_from_matrix(M))
            wires += [a, a + 1]
    return np.stack(gates), np.asarray(wires, np.int32), h


def xxz_hamiltonian(L, h):
    """Dense H of xxz_trotter_target (qubit 0 = most significant bit), for small L."""
    S = 1 << L
    idx = np.arange(S)
    bits = [(idx >> (L - 1 - q)) & 1 for q in range(L)]
    z = [1.0 - 2.0 * b for b in bits]
    H = np.zeros((S, S), np.complex128)
    diag = sum(2.0 * z[j] * z[j + 1] for j in range(L - 1)) + sum(h[j] * z[j] for j in range(L))
    H[idx, idx] = diag
    for j in range(L - 1):
        flip = bits[j] != bits[j + 1]
        s2 = idx ^ (1 << (L - 1 - j)) ^ (1 << (L - 2 - j))
        H[s2[flip], idx[flip]] += 1.0
    return H


def large_instance(L=16, depth=11, steps=50, t=1.0, seed=0):
    """The large single instance: an open brickwall of `depth` layers with seeded Haar
    gates, and the XXZ Trotter circuit as its matrix-free target."""
    rng = np.random.default_rng(seed)
    param, wires, G = open_brickwall(L, depth)
    gates = haar_matchgates(rng, G)
    tg, tw, h = xxz_trotter_target(L, steps, t, seed)
    return {"L": L, "param": param, "wires": wires, "G": G, "gates": gates, "target_gates": tg,
            "target_wires": tw, "Gt": len(tg), "fields": h}
