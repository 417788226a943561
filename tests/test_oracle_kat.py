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


def test_2d_fixed_tile(oracle):
    # fixed mode: d = q. q(0,0) = -3, q(1,2) = 5, q(7,7) = 2 -> rate 3 (|5| = 0b101).
    # Byte k of the sign field and of each plane is row k; column c is bit 7-c.
    x = np.zeros((8, 8), np.float32)
    x[0, 0], x[1, 2], x[7, 7] = -3, 5, 2
    sign = "8000000000000000"     # (0,0) negative
    p0 = "8020000000000000"       # 3 (0,0) and 5 (1,2) have bit 0
    p1 = "8000000000000001"       # 3 (0,0) and 2 (7,7) have bit 1
    p2 = "0020000000000000"       # 5 (1,2) has bit 2
    exp = hexb("03000000") + hexb(sign + p0 + p1 + p2)
    y = check(oracle, x.ravel(), 0.5, 0, 2, (8, 8, 1), exp)
    assert np.array_equal(y, x.ravel())


def test_2d_outlier_selected_two_prefix_bytes(oracle):
    # outlier mode, one 8x8 tile of 1000 with x(0,1) = 1001:
    # d(0,0) = 1000, d(0,1) = 1, d(0,2) = -1, every other residual 0.
    # R1 = bitlen(1000) = 10 -> 8 + 8*10 = 88 bytes; R2 = 1, nb = ceil(10/8) = 2
    # -> 2 + 8 + 8 = 18 bytes: selected, byte 0x80 | (2-1)<<5 | 1 = 0xA1.
    # Payload: |d(0,0)| little-endian (e8 03), sign (e=2: byte 0 bit 5),
    # plane 0 with element 0 cleared (e=1, e=2: bits 6 and 5 of byte 0).
    x = np.full((8, 8), 1000, np.float32)
    x[0, 1] = 1001
    exp = hexb("a1000000") + hexb("e803" + "2000000000000000" + "6000000000000000")
    y = check(oracle, x.ravel(), 0.5, 2, 2, (8, 8, 1), exp)
    assert np.array_equal(y, x.ravel())


def test_2d_plain_second_tile_only(oracle):
    # 16 x 8 slice = two tiles along x (b = by*Bx + bx); tile 0 is all zero (rate 0,
    # no payload); tile 1 has x(0,8) = 1, its predictor restarts at the tile:
    # d(0,0) = 1, d(0,1) = -1, d(1,0) = -1 -> rate 1.
    x = np.zeros((8, 16), np.float32)
    x[0, 8] = 1
    exp = hexb("00010000") + hexb("4080000000000000" + "c080000000000000")
    y = check(oracle, x.ravel(), 0.5, 1, 2, (16, 8, 1), exp)
    assert np.array_equal(y, x.ravel())


def test_3d_fixed_block(oracle):
    # fixed mode, 4x4x4 block, e = 16z + 4y + x: q(0,0,0) = 6, q(1,2,3) = -1 (e = 27),
    # q(3,3,3) = 4 (e = 63) -> rate 3. e = 27 is byte 3, bit 4; e = 63 byte 7, bit 0.
    x = np.zeros((4, 4, 4), np.float32)
    x[0, 0, 0], x[1, 2, 3], x[3, 3, 3] = 6, -1, 4
    sign = "0000001000000000"
    p0 = "0000001000000000"       # 1 (e=27)
    p1 = "8000000000000000"       # 6 (e=0)
    p2 = "8000000000000001"       # 6 (e=0), 4 (e=63)
    exp = hexb("03000000") + hexb(sign + p0 + p1 + p2)
    y = check(oracle, x.ravel(), 0.5, 0, 3, (4, 4, 4), exp)
    assert np.array_equal(y, x.ravel())


def test_3d_outlier_selected(oracle):
    # outlier mode, a 4x4x4 block of 300 with x(0,0,1) = 302:
    # d(0,0,0) = 300, d(0,0,1) = 2, d(0,0,2) = -2, others 0.
    # R1 = bitlen(300) = 9 -> 8 + 72 = 80 bytes; R2 = 2, nb = 2 -> 2 + 8 + 16 = 26:
    # selected, byte 0x80 | 1<<5 | 2 = 0xA2. Prefix 300 = 2c 01; sign e=2 (byte 0
    # bit 5); plane 0 empty; plane 1: e=1, e=2 (0x60).
    x = np.full((4, 4, 4), 300, np.float32)
    x[0, 0, 1] = 302
    exp = hexb("a2000000") + hexb("2c01" + "2000000000000000" + "0000000000000000" + "6000000000000000")
    y = check(oracle, x.ravel(), 0.5, 2, 3, (4, 4, 4), exp)
    assert np.array_equal(y, x.ravel())


def test_3d_outlier_not_selected_keeps_plain_bytes(oracle):
    # outlier mode where the plain encoding is smaller: d(0,0,0) = 3, one residual
    # 2 -> R1 = 2: 8 + 16 = 24 bytes vs R2 = 2, nb = 1: 1 + 8 + 16 = 25 -> plain,
    # byte 2, and element 0 stays in the planes.
    x = np.full((4, 4, 4), 3, np.float32)
    x[3, 3, 3] = 5   # d(3,3,3) = 5 - 3 = 2 (e = 63: byte 7, bit 0)
    exp = hexb("02000000") + hexb("0000000000000000" + "8000000000000000" + "8000000000000001")
    y = check(oracle, x.ravel(), 0.5, 2, 3, (4, 4, 4), exp)
    assert np.array_equal(y, x.ravel())
