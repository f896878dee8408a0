"""Workload driver for hardware-counter profiles of the step kernels.

    python3 profile_probe.py [--grid 12800] [--steps 6] [--copy-mib 1024]
                                       [--sort MODE]

Runs, in one process:
  1. one device-to-device copy of --copy-mib MiB through the HIP runtime
     (a plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE give the
     scale that turns those counters into bytes);
  2. init + --steps iterations of the dense v2-style workload on a
     --grid x --grid torus (capacity 4x the initial agents), with the agent
     sort of --sort (0 radix every iteration, 1 automatic, 8..11 row
     buckets; default automatic).
It prints the exact event sums of the run, so the per-dispatch counters can
be divided by the algorithmic bytes of each kernel.
"""
import argparse
import ctypes
import json
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def device_copy(mib):
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    hipMemcpyDeviceToDevice = 3
    if hip.hipMemcpy(b, a, ctypes.c_size_t(n), hipMemcpyDeviceToDevice) != 0:
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=12800)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--sort", type=int, default=1)
    args = ap.parse_args()
    from pphpc_amd import Params, Simulation
    copied = device_copy(args.copy_mib)
    L = args.grid
    cells = L * L
    d = dict(init_sheep=cells * 1600 // 10000, sheep_gain_from_food=4,
             sheep_reproduce_threshold=2, sheep_reproduce_prob=10,
             init_wolves=cells * 800 // 10000, wolves_gain_from_food=20,
             wolves_reproduce_threshold=2, wolves_reproduce_prob=5,
             grass_restart=5, grid_x=L, grid_y=L, iters=args.steps)
    n0 = d["init_sheep"] + d["init_wolves"]
    sim = Simulation(Params.from_values(**d), seed=0, capacity=4 * n0,
                     device=0)
    sim.set_sort(args.sort)
    sim.init()
    sim.run(args.steps, collect=False)
    out = {"copy_bytes": copied, "grid": L, "steps": args.steps,
           "info": sim.info(), "counters": sim.counters()}
    sim.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
