"""Frozen fixtures (tests/golden/cuszp_golden.npz, written by make_golden.py):
the oracle must reproduce them on the CPU, the HIP library on the GPU."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "cuszp_golden.npz"), allow_pickle=False)
TAGS = sorted({k.split("__")[0] for k in G.files})


def _case(tag):
    mode, dim, x, y, z, n = (int(v) for v in G[f"{tag}__meta"])
    return G[f"{tag}__x"], float(G[f"{tag}__eb"][0]), mode, dim, (x, y, z), n


@pytest.mark.parametrize("tag", TAGS)
def test_oracle_matches_golden(oracle, tag):
    x, eb, mode, dim, dims, n = _case(tag)
    assert oracle.compress(x, eb, mode, dim, dims) == G[f"{tag}__cmp"].tobytes()
    d = oracle.decompress(G[f"{tag}__cmp"].tobytes(), n, x.dtype, eb, mode, dim, dims)
    assert np.array_equal(d.view(np.uint8), G[f"{tag}__dec"].view(np.uint8))


@pytest.mark.gpu
@pytest.mark.parametrize("tag", TAGS)
def test_gpu_matches_golden(device, tag):
    from tests.helpers import gpu_compress, gpu_decompress

    x, eb, mode, dim, dims, n = _case(tag)
    assert gpu_compress(x, eb, mode, dim, dims) == G[f"{tag}__cmp"].tobytes()
    d = gpu_decompress(G[f"{tag}__cmp"].tobytes(), n, x.dtype, eb, mode, dim, dims)
    assert np.array_equal(d.view(np.uint8), G[f"{tag}__dec"].view(np.uint8))
