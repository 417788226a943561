"""Writes tests/golden/cuszp_golden.npz: small seeded inputs for every
(dim, mode, dtype) with the oracle's compressed stream and decompressed array.

The fixtures freeze the format: tests/test_golden.py checks the oracle (CPU)
and the HIP library (GPU) against them, so a change to either that alters a
byte shows up even when both sides change together.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.workloads import mixed_field  # noqa: E402

SHAPES = {1: [(0, 0, 0, 1000), (0, 0, 0, 33)], 2: [(12, 10, 3, 360), (8, 8, 1, 64)], 3: [(8, 9, 5, 360), (4, 4, 4, 64)]}


def cases():
    for dtype in (np.float32, np.float64):
        for mode in (0, 1, 2):
            for dim, shapes in SHAPES.items():
                for k, (x, y, z, n) in enumerate(shapes):
                    yield dtype, mode, dim, (x, y, z), n, k


def main():
    out = {}
    for dtype, mode, dim, dims, n, k in cases():
        tag = f"{np.dtype(dtype).name}_m{mode}_d{dim}_{k}"
        data = mixed_field(n, seed=1000 + n + 17 * mode + dim, dtype=dtype)
        eb = 1e-2
        c = O.compress(data, eb, mode, dim, dims)
        d = O.decompress(c, n, dtype, eb, mode, dim, dims)
        out[f"{tag}__x"] = data
        out[f"{tag}__cmp"] = np.frombuffer(c, np.uint8)
        out[f"{tag}__dec"] = d
        out[f"{tag}__meta"] = np.array([mode, dim, dims[0], dims[1], dims[2], n], np.int64)
        out[f"{tag}__eb"] = np.array([eb], np.float64)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cuszp_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes,", len(out) // 5, "cases")


if __name__ == "__main__":
    main()
