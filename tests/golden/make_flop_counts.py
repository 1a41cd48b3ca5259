"""Algorithmic FP64 FLOPs per particle-element-turn of the benchmark lattices,
counted by running the oracle's op-counting build (oracle/liboracle_count.so:
add/sub/mul/div/sqrt = 1, each transcendental call = 1, every per-element
constant recomputed per particle exactly as the reference does -- i.e. the
reference's own operation count, not the HIP kernel's).

Writes tests/golden/flops_per_pet.json.  Run after building the oracle.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "bmad-ecosystem_amd")]

import oracle  # noqa: E402
from bmx import abi, beam, lattice  # noqa: E402


def count(eles, coords, n_turns=1):
    oracle.flops_read_reset()
    r = oracle.track_bunch(eles, abi.ELECTRON, n_turns, coords, counting=True)
    flops = oracle.flops_read_reset()
    alive = int(np.sum(r["state"] == abi.ALIVE))
    assert alive == coords.shape[1], "counting sample must not lose particles"
    n = coords.shape[1]
    per_turn = flops / (n * n_turns)
    return dict(flops_per_particle_turn=per_turn, n_ele=len(eles), flops_per_pet=per_turn / len(eles),
                sample_particles=n, turns=n_turns)


def per_type(eles, coords):
    out = {}
    for key in sorted(set(e.key for e in eles)):
        sub = [lattice.EleRec_copy(e) for e in eles if e.key == key][:1]
        c = count(sub, coords)
        out[abi.KEY_NAMES[key]] = c["flops_per_pet"]
    return out


if __name__ == "__main__":
    res = {}
    c0 = lattice.fodo_ring_c0()
    coords, _ = beam.c0_bunch(c0[:7], n=256, seed=1)
    res["C0"] = count(c0, coords)
    res["C0"]["per_element_type"] = per_type(c0, coords)
    c1 = lattice.ring_c1()
    tw = beam.periodic_twiss(lattice.ring_c1(n_rf=0)[:8])
    coords = beam.matched_gaussian(256, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed=1)
    res["C1"] = count(c1, coords)
    res["C1"]["per_element_type"] = per_type(c1, coords)
    res["convention"] = ("FP64 add/sub/mul/div/sqrt = 1, sin/cos/tan/sinh/cosh/asin/atan2 = 1; counted on the "
                         "oracle's restatement of the reference routines, per live particle-element-turn")
    with open(os.path.join(HERE, "flops_per_pet.json"), "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print(json.dumps(res, indent=1))
