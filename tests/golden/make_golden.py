"""Writes tests/golden/rng_vectors.json: known-answer vectors of the project
generator, computed by an independent pure-Python restatement (integer
arithmetic only, no numpy, no oracle) of

  * SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
    generators", OOPSLA 2014), used to expand (seed << 32 | id) into a state;
  * Marsaglia's xorwow ("Xorshift RNGs", J. Stat. Softw. 8(14), 2003, p. 5):
    five 32-bit xorshift words plus a Weyl counter d += 362437, output d + v;
  * the uniform map ((r >> 8) + 1) * 2^-24, exact in FP32 and FP64;
  * the block-linear seed id of a 32x16 launch grid
    (src/array2d_psng.cu:8-9 of the reference).

The oracle and the GPU engine must both reproduce these numbers bit for bit.
Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os

M32 = (1 << 32) - 1
M64 = (1 << 64) - 1


def splitmix64_stream(state: int, n: int):
    out = []
    for _ in range(n):
        state = (state + 0x9E3779B97F4A7C15) & M64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out.append(z ^ (z >> 31))
    return out


def seed_state(id_: int, seed: int):
    a, b, c = splitmix64_stream(((seed & M32) << 32) | (id_ & M32), 3)
    st = [a & M32, a >> 32, b & M32, b >> 32, c & M32, c >> 32]
    if (st[0] | st[1] | st[2] | st[3] | st[4]) == 0:
        st[0] = 1
    return st


def xorwow(st):
    x, y, z, w, v, d = st
    t = (x ^ (x >> 2)) & M32
    x, y, z, w = y, z, w, v
    v = ((v ^ ((v << 4) & M32)) ^ (t ^ ((t << 1) & M32))) & M32
    d = (d + 362437) & M32
    st[:] = [x, y, z, w, v, d]
    return (d + v) & M32


def uniform(r: int) -> float:
    return ((r >> 8) + 1) / 16777216.0


def seed_id(x, y, width, bx=32, by=16):
    gx = (width - 1) // bx + 1
    return ((y // by) * gx + (x // bx)) * bx * by + (y % by) * bx + (x % bx)


def build():
    cases = []
    for id_, seed in [(0, 0), (1, 0), (0, 7), (2593, 7), (4294967295, 7), (12345, 4294967295)]:
        st = seed_state(id_, seed)
        init = list(st)
        outs = [xorwow(st) for _ in range(16)]
        cases.append(dict(id=id_, seed=seed, state=init, outputs=outs,
                          uniforms=[uniform(r) for r in outs], final_state=list(st)))
    ids = []
    for (x, y, w) in [(0, 0, 64), (31, 0, 64), (32, 0, 64), (0, 1, 64), (0, 16, 64),
                      (33, 17, 100), (99, 35, 100), (47, 35, 48), (2047, 1535, 2048)]:
        ids.append(dict(x=x, y=y, width=w, id=seed_id(x, y, w)))
    return dict(
        splitmix64_seed0=splitmix64_stream(0, 3),
        rng=cases,
        seed_ids=ids,
    )


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_vectors.json")
    with open(path, "w") as f:
        json.dump(build(), f, indent=1)
    print("wrote", path)
