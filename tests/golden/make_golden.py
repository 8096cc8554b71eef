"""Writes tests/golden/keys_lebesgue_dyadic.npz: Lebesgue keys of points whose coordinates
are dyadic in the box [-1, 1]^3, from the closed form of the reference's key bisection
(tree.cu:157-240): with u = i / 2^21, level l's digit is (1 - bit_l(ix)) + 2 (1 - bit_l(iy))
+ 4 (1 - bit_l(iz)) (the lower half sets the bit; a point on a midpoint goes up).  The
points include both box corners so any bounding box that contains them and 0 is [-1, 1]^3.
Independent of the oracle and of the HIP code.  Run: python tests/golden/make_golden.py"""
from pathlib import Path

import numpy as np

LEVELS = 21


def closed_form_lebesgue(ix, iy, iz):
    key = np.zeros(len(ix), dtype=np.uint64)
    for lv in range(LEVELS):
        sh = LEVELS - 1 - lv
        d = (1 - ((ix >> sh) & 1)) + 2 * (1 - ((iy >> sh) & 1)) + 4 * (1 - ((iz >> sh) & 1))
        key |= d.astype(np.uint64) << np.uint64(3 * (LEVELS - 1 - lv))
    return key


def main():
    rng = np.random.default_rng(20261202)
    n = 4093
    ii = rng.integers(0, 2 ** LEVELS, size=(n, 3))
    hand = np.array([[0, 0, 0], [2 ** LEVELS, 2 ** LEVELS, 2 ** LEVELS],   # corners -1 and +1
                     [2 ** 20, 0, 0], [2 ** 20, 2 ** 20, 2 ** 20],          # on the first midpoints
                     [2 ** 20 - 1, 2 ** 20, 2 ** 20 + 1], [1, 2, 3], [2 ** 21 - 1] * 3])
    ii = np.concatenate([hand, ii]).astype(np.int64)
    # the upper corner i = 2^21 is u = 1: on every level it is not below the midpoint
    keys = closed_form_lebesgue(np.minimum(ii[:, 0], 2 ** LEVELS - 1) | (ii[:, 0] >> LEVELS) * (2 ** LEVELS - 1),
                                np.minimum(ii[:, 1], 2 ** LEVELS - 1) | (ii[:, 1] >> LEVELS) * (2 ** LEVELS - 1),
                                np.minimum(ii[:, 2], 2 ** LEVELS - 1) | (ii[:, 2] >> LEVELS) * (2 ** LEVELS - 1))
    x = 2.0 * ii / 2 ** LEVELS - 1.0
    out = Path(__file__).with_name("keys_lebesgue_dyadic.npz")
    np.savez(out, x=x[:, 0], y=x[:, 1], z=x[:, 2], keys=keys)
    print(f"wrote {out} ({len(keys)} points)")


if __name__ == "__main__":
    main()
