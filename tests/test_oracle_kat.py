"""Known-answer streams derived by hand from the stream format, checked against
the C++ oracle and the numpy restatement.

Every expected byte string below was worked out on paper from the format
(header = one rate byte per block, payload = [outlier prefix][sign field]
[planes, LSB first], element e at byte e/8, bit 7-e%8) and the reference's
kernels (src/cuSZp_kernels_1D_f32.cu, _2D_f32.cu, _3D_f32.cu); with eb = 0.5
the reciprocal is exactly 1, so integer inputs quantise to themselves.
"""
import numpy as np
import pytest

from oracle import np_oracle as NPO

H1 = bytes(1024)  # 1D header of a single 32768-element chunk, before the rate bytes


def hexb(s: str) -> bytes:
    return bytes.fromhex(s.replace(" ", ""))


def check(oracle, x, eb, mode, dim, dims, expected):
    got = oracle.compress(x, eb, mode, dim, dims)
    assert got == expected, (got.hex(), expected.hex())
    assert NPO.compress(x, eb, mode, dim, dims) == expected
    y = oracle.decompress(got, x.size, x.dtype, eb, mode, dim, dims)
    y2 = NPO.decompress(got, x.size, x.dtype, eb, mode, dim, dims)
    assert np.array_equal(y.view(np.uint8), y2.view(np.uint8))
    return y


def header_1d(rates, n):
    h = bytearray(-(-n // 32768) * 1024)
    for i, r in enumerate(rates):
        h[i] = r
    return bytes(h)


def test_1d_plain_ramp(oracle):
    # q = 0..31, residuals [0,1,1,...] -> rate 1; sign word 0; plane 0 = 0x7FFFFFFF (element 0 clear)
    x = np.arange(32, dtype=np.float32)
    exp = header_1d([1], 32) + hexb("00000000 7fffffff")
    y = check(oracle, x, 0.5, 1, 1, (0, 0, 0), exp)
    assert np.array_equal(y, x)
    # reference bug B1: a single-threadblock grid reports cmpSize = 1024 (header only);
    # the true stream length is 1032
    assert len(exp) == 1032


def test_1d_fixed_signs_and_planes(oracle):
    # q = [-3, 0, 5, 0...]: sign of element 0; |q| = 3, 0, 5 -> rate 3
    x = np.zeros(32, np.float32)
    x[0], x[2] = -3, 5
    exp = header_1d([3], 32) + hexb("80000000 a0000000 80000000 20000000")
    y = check(oracle, x, 0.5, 0, 1, (0, 0, 0), exp)
    assert np.array_equal(y, x)


def test_1d_outlier_selected(oracle):
    # d = [1000, 0, 1, -1, 0...]: plain needs 10 planes (44 bytes); outlier form
    # is 2 prefix bytes + sign + 1 plane = 10 bytes -> byte 0x80 | (2-1)<<5 | 1
    x = np.full(32, 1000, np.float32)
    x[2] = 1001
    exp = header_1d([0xA1], 32) + hexb("e803 10000000 30000000")
    y = check(oracle, x, 0.5, 2, 1, (0, 0, 0), exp)
    assert np.array_equal(y, x)


def test_1d_outlier_tie_prefers_plain(oracle):
    # d = [2^24, 0, ..., 0, 2^23]: plain = 4 + 4*25 = 104 bytes; outlier = 4 prefix
    # bytes + 4 + 4*24 = 104 -> equal sizes keep the plain encoding (size1 <= size2)
    x = np.full(32, 2.0**24, np.float32)
    x[31] = 2.0**24 + 2.0**23
    got = oracle.compress(x, 0.5, 2, 1)
    assert got[0] == 25 and len(got) == 1024 + 104
    assert NPO.compress(x, 0.5, 2, 1) == got
    # one smaller plane on the outlier side wins: d = [2^24, 0, ..., 2^22]
    x[31] = 2.0**24 + 2.0**22
    got = oracle.compress(x, 0.5, 2, 1)
    assert got[0] == (0x80 | (3 << 5) | 23) and len(got) == 1024 + 4 + 4 + 4 * 23


def test_1d_outlier_rate_zero_remainder(oracle):
    # constant 200: d = [200, 0, ...] -> plain 4 + 4*8 = 36 bytes vs outlier 1 + 4 = 5
    x = np.full(32, 200, np.float32)
    exp = header_1d([0x80], 32) + hexb("c8 00000000")
    y = check(oracle, x, 0.5, 2, 1, (0, 0, 0), exp)
    assert np.array_equal(y, x)


def test_1d_rate_zero_and_header_padding(oracle):
    x = np.zeros(40000, np.float32)
    exp = bytes(2048)   # ceil(40000/32768) * 1024 header bytes, no payload
    y = check(oracle, x, 1e-3, 1, 1, (0, 0, 0), exp)
    assert not np.signbit(y).any() and not y.any()


def test_1d_f64_fixed(oracle):
    # eb = 0.25: r = 2, 2.5 -> q = 5 = 0b101, rate 3
    x = np.array([2.5], np.float64)
    exp = header_1d([3], 1) + hexb("00000000 80000000 00000000 80000000")
    y = check(oracle, x, 0.25, 0, 1, (0, 0, 0), exp)
    assert y[0] == 2.5


def test_2d_plain_tile(oracle):
    # x(r,c) = 10r + c: d = 1 (c>=1), 10 (c=0, r>=1), 0 at (0,0) -> rate 4
    x = (10 * np.arange(8)[:, None] + np.arange(8)[None, :]).astype(np.float32).ravel()
    planes = ("7f" * 8) + "0080808080808080" + ("00" * 8) + "0080808080808080"
    exp = hexb("04000000") + bytes(8) + hexb(planes)
    y = check(oracle, x, 0.5, 1, 2, (8, 8, 1), exp)
    assert np.array_equal(y, x)


def test_2d_header_pad_bytes_zero(oracle):
    x = np.zeros(8 * 8 * 3, np.float32)
    assert oracle.compress(x, 0.5, 1, 2, (8, 8, 3)) == bytes(4)


def test_3d_plain_block(oracle):
    # x(z,y,x) = 100z + 10y + x: d = 1 (x>=1), 10 (x=0,y>=1), 100 (x=y=0,z>=1) -> rate 7
    z, y, xx = np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij")
    x = (100 * z + 10 * y + xx).astype(np.float32).ravel()
    p0 = "77" * 8
    p1 = "0888088808880888"
    p2 = "0000800080008000"
    planes = p0 + p1 + p2 + p1 + ("00" * 8) + p2 + p2
    exp = hexb("07000000") + bytes(8) + hexb(planes)
    out = check(oracle, x, 0.5, 1, 3, (4, 4, 4), exp)
    assert np.array_equal(out, x)


def test_2d_partial_tile_edges(oracle):
    # 4 x 3 slice: columns 4..7 and rows 3..7 of the 8x8 tile are padding (|d| = 0)
    x = np.arange(12, dtype=np.float32)   # rows: [0..3], [4..7], [8..11]
    # d: row 0 = [0,1,1,1]; rows 1,2 col 0 = 4; others 1 -> rate 3 (|4| = 0b100)
    # plane 0 rows: 0b0111_0000, 0b0111_0000, 0b0111_0000; plane 2 rows 1,2: 0b1000_0000
    planes = "707070" + "0000000000" + ("00" * 8) + "008080" + "0000000000"
    exp = hexb("03000000") + bytes(8) + hexb(planes)
    y = check(oracle, x, 0.5, 1, 2, (4, 3, 1), exp)
    assert np.array_equal(y, x)
