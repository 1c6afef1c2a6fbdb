"""Write the golden vectors under tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

The oracle is itself pinned by the reference's known-answer tests (Philox
Random123 vectors, advection/Hanna/Langevin/deposition closed forms,
gridding/compaction/Morton vectors); these fixtures freeze its output on the
point-release workload shape so that later changes to the oracle or to the
HIP path are checked against the same bits.  Inputs that are large (met
fields) are regenerated from their seeds by flexpart-gpu_amd/workloads.py;
the fixtures hold particle states, parameters and expected outputs only.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from conftest import load_package, met_dict, pbl_dict  # noqa: E402

import oracle  # noqa: E402


def point_release_case():
    """64 particles released at the grid centre of the point-release met,
    spread by 10 steps of the production step (900 s, dry deposition)."""
    load_package()
    import flexpart_gpu_amd.workloads as W
    w = W.point_release()
    p = w.particles[:64].copy()
    p["flags"] = 1
    p["turb_u"] = np.linspace(-0.5, 0.5, 64, dtype=np.float32)
    p["pos_z"] = np.linspace(5.0, 1500.0, 64, dtype=np.float32)
    p["flags"][::9] = 0
    init = p.copy()
    params = dict(dt_seconds=w.dt_seconds, alpha=0.0, x_scale=w.x_scale, y_scale=w.y_scale,
                  key=w.key, counter=(0, 0, 0, 0), dry_mode=1, dry_vd=w.dry_vd, dry_href=15.0)
    states, counters = [], []
    for k in range(10):
        params["alpha"] = (k * w.dt_seconds) / w.met.seconds_between_levels
        nxt = oracle.step(p, met_dict(w.met), pbl_dict(w.pbl), params)
        params["counter"] = tuple(int(v) for v in nxt)
        states.append(p.copy())
        counters.append(nxt.copy())
    counts, mass = oracle.grid_accumulate(p, w.out_shape, w.outheights, w.refine, 0, 1.0e6)
    return dict(init=init.view(np.uint32), step1=states[0].view(np.uint32),
                step10=states[9].view(np.uint32), counters=np.array(counters, np.uint32),
                grid_counts=counts, grid_mass=mass)


def compaction_sort_case():
    load_package()
    import flexpart_gpu_amd.workloads as W
    w = W.small_random(512, shape=(16, 12, 8), seed=3)
    p = w.particles
    a, out, smap = oracle.compact(p)
    sorted10, n2o10 = oracle.sort_morton(p, 10)
    sorted3, n2o3 = oracle.sort_morton(p, 3, bounds=(0.0, 16.0, 0.0, 12.0, 0.0, 2500.0))
    return dict(particles=p.view(np.uint32), active=np.array([a], np.uint64),
                compacted=out.view(np.uint32), survivor_map=smap,
                sorted10=sorted10.view(np.uint32), n2o10=n2o10,
                sorted3=sorted3.view(np.uint32), n2o3=n2o3)


def philox_case():
    key = np.array([0xDECAFBAD, 0x12345678], np.uint32)
    base = np.array([0xFFFFFFF0, 0xFFFFFFFF, 0, 1], np.uint32)
    return dict(key=key, base=base, blocks=oracle.philox_stream(base, key, 64))


def main():
    np.savez_compressed(os.path.join(HERE, "point_release_64.npz"), **point_release_case())
    np.savez_compressed(os.path.join(HERE, "compaction_sort_512.npz"), **compaction_sort_case())
    np.savez_compressed(os.path.join(HERE, "philox_64.npz"), **philox_case())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
