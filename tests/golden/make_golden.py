"""Write the stencil golden fixtures from the reference's own kernels.

Runs in the build container, where the reference tree is mounted: the two C
kernels (src/modeling/fd4t10s-damp-zjh.c, fd4t10s-zjh.c) are compiled
unchanged into oracle/_ref/libref.so by oracle/Makefile.  This script drives
them on seeded inputs and stores inputs and outputs as data.

    make -C oracle && python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle  # noqa: E402

# (nx, nz, nb, seed): padded dims and the damping width handed to the kernel
CASES = [(48, 40, 14, 11), (20, 16, 8, 12), (37, 29, 20, 13)]
# multi-step case on the 201x301 (nb=20) padded grid: point source, damped steps
C0 = dict(nx=353, nz=233, nb=26, steps=40, seed=14)


def main():
    ref = pyoracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref.so missing: run `make -C oracle` with the reference tree present")
    out = {}
    for k, (nx, nz, nb, seed) in enumerate(CASES):
        rng = np.random.default_rng(seed)
        cur = rng.standard_normal((nx, nz)).astype(np.float32)
        prev = rng.standard_normal((nx, nz)).astype(np.float32)
        vel = rng.uniform(4.9, 44.0, (nx, nz)).astype(np.float32)
        pd, ud = prev.copy(), np.zeros_like(cur)
        ref.fd4t10s_damp_zjh_2d_vtrans(pd, cur, vel, ud, nx, nz, nb)
        pu, uu = prev.copy(), np.zeros_like(cur)
        ref.fd4t10s_zjh_2d_vtrans(pu, cur, vel, uu, nx, nz)
        out.update({f"c{k}_cur": cur, f"c{k}_prev": prev, f"c{k}_vel": vel, f"c{k}_nb": np.int32(nb),
                    f"c{k}_damp_next": pd, f"c{k}_damp_u2": ud, f"c{k}_undamp_next": pu, f"c{k}_undamp_u2": uu})
    # 201x301 (nb=20) grid: a unit spike stepped C0["steps"] times with the damped kernel
    nx, nz, nb = C0["nx"], C0["nz"], C0["nb"]
    rng = np.random.default_rng(C0["seed"])
    vel = rng.uniform(4.9, 44.0, (nx, nz)).astype(np.float32)
    p0 = np.zeros((nx, nz), np.float32)
    p1 = np.zeros((nx, nz), np.float32)
    u2 = np.zeros((nx, nz), np.float32)
    p1[nx // 2, 7] = 1.0
    for _ in range(C0["steps"]):
        ref.fd4t10s_damp_zjh_2d_vtrans(p0, p1, vel, u2, nx, nz, nb)
        p0, p1 = p1, p0
    out["c0grid_vel_seed"] = np.int32(C0["seed"])
    out["c0grid_steps"] = np.int32(C0["steps"])
    out["c0grid_sha256"] = np.frombuffer(hashlib.sha256(p1.tobytes()).digest(), np.uint8)
    out["c0grid_column"] = p1[nx // 2].copy()
    out["c0grid_row7"] = p1[:, 7].copy()
    np.savez_compressed(os.path.join(HERE, "fd4t10s_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "fd4t10s_golden.npz"))


if __name__ == "__main__":
    main()
