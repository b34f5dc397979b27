"""One rank of a column-sharded evaluation (tests/test_sharded.py starts two).

Usage: shard_worker.py RANK WORLD PORT MODE OUT.npz
MODE "oracle": the rank's partial sums come from the CPU oracle (no GPU);
MODE "gpu":    from rqcopt_mg.evaluate_columns on the GPU.
The partial sums are added with torch.distributed (gloo, host tensors, 127.0.0.1) and
projected with rqcopt_mg.sharded.sharded_riemannian_gradient_hvp."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "rqcopt_hpc-matchgate_amd"))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402


def problem(L=6, depth=3, seed=5):
    from dense_ref import program_matrix
    from rqcopt_mg.workloads import brickwall_job

    rng = np.random.default_rng(seed)
    job = brickwall_job(rng, L, depth, 1)
    W = program_matrix(job["target_gates"], np.arange(job["Gt"]), job["target_wires"], L)
    v = rng.standard_normal(8 * job["G"])
    return job, W, v


def main():
    rank, world, port, mode, out = int(sys.argv[1]), int(sys.argv[2]), int(sys.argv[3]), sys.argv[4], sys.argv[5]
    import torch.distributed as dist

    import rqcopt_mg as mg
    from rqcopt_mg.sharded import sharded_riemannian_gradient_hvp

    job, W, v = problem()
    L, V = job["L"], job["gates"]
    if mode == "oracle":
        import oracle as orc

        Z = orc.real_to_directions(V, v)

        def partial(b0, b1):
            return orc.target_gradient_hvp(W, V, Z, job["param"], job["wires"], L, b0=b0, b1=b1)
    else:
        prog = mg.Program.create(L, job["G"], job["param"], job["wires"])
        tgt = mg.Target.dense(W)
        Z = mg.real_to_directions(V, v)

        def partial(b0, b1):
            return mg.evaluate_columns(tgt, prog, V, Z, mg.EVAL_F_GRAD_HVP, mg.MG_FP64, b0, b1)

    dist.init_process_group("gloo", init_method=f"tcp://127.0.0.1:{port}", rank=rank, world_size=world)
    try:
        f, g, hv = sharded_riemannian_gradient_hvp(partial, V, v, L, rank, world)
    finally:
        dist.destroy_process_group()
    np.savez(out, f=np.array([f]), g=g, hv=hv)
    print(f"rank {rank}: done", flush=True)


if __name__ == "__main__":
    main()
