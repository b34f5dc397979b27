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
