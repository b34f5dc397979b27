"""Column-sharded evaluation across ranks (one process per GPU).

The 2^L basis columns of the trace are independent: rank r of `world` evaluates the contiguous
range [r 2^L / world, (r+1) 2^L / world) (`column_range`, the split of the reference's
`asign_state_vectors`, benchmark_hpc/benchmark_mpi.c:26-45), one all-reduce adds the partial
f (1 complex), complex gradient (8P complex) and D^2 f[Z] (8P complex) sums, and every rank then
applies the Riemannian projection (linear in those sums) to the totals.  This replaces the
reference's MPI reduction (src/matchgate_target.c:979-1355) with one torch.distributed
all-reduce: RCCL ("nccl") on device tensors between GPUs, or gloo on host tensors.

The per-rank partial is a callable `partial(b0, b1) -> (f, dgrad (P, 8), dhz (P, 8))`; on a GPU
rank it is `rqcopt_mg.evaluate_columns(target, program, gates, Z, what, precision, b0, b1)`.
"""
import numpy as np


def column_range(L, rank, world):
    """Basis columns [b0, b1) of `rank` among `world` ranks: contiguous ranges, the first
    2^L mod world ranks one column longer (the reference's rule)."""
    if world < 1 or not 0 <= rank < world:
        raise ValueError("invalid rank / world size")
    q, extra = divmod(1 << L, world)
    b0 = rank * q + min(rank, extra)
    return b0, b0 + q + (1 if rank < extra else 0)


def pack(f, dgrad, dhz):
    """[Re f, Im f, dgrad (re, im interleaved), dhz (re, im interleaved)] as float64 (2 + 32 P)"""
    dg = np.ascontiguousarray(dgrad, np.complex128).reshape(-1)
    dh = np.ascontiguousarray(dhz, np.complex128).reshape(-1) if dhz is not None else np.zeros_like(dg)
    return np.concatenate([[np.real(f), np.imag(f)], dg.view(np.float64), dh.view(np.float64)])


def unpack(buf, P):
    buf = np.ascontiguousarray(buf, np.float64)
    f = complex(buf[0], buf[1])
    dg = buf[2:2 + 16 * P].view(np.complex128).reshape(P, 8)
    dh = buf[2 + 16 * P:2 + 32 * P].view(np.complex128).reshape(P, 8)
    return f, dg, dh


def sharded_riemannian_gradient_hvp(partial, gates, v, L, rank, world, group=None, device=None):
    """f, Riemannian gradient (8P) and, with v, H v (8P) of the full trace, each rank evaluating its
    column range with `partial` and adding the partial sums with one all-reduce over
    torch.distributed (`device`: a torch device for the reduction buffer, e.g. "cuda" with the
    nccl backend; None keeps it on the host, as gloo needs)."""
    import torch
    import torch.distributed as dist

    from . import riemannian_project

    gates = np.ascontiguousarray(gates, np.complex128)
    P = gates.shape[0]
    b0, b1 = column_range(L, rank, world)
    f, dg, dh = partial(b0, b1)
    buf = torch.from_numpy(pack(f, dg, dh if v is not None else None))
    if device is not None:
        buf = buf.to(device)
    if world > 1:
        dist.all_reduce(buf, group=group)
    f, dg, dh = unpack(buf.cpu().numpy(), P)
    g, hv = riemannian_project(gates, dg, v, dh if v is not None else None)
    return f, g, hv
