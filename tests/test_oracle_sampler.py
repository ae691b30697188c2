"""Pins the oracle's z-Sobol sampler (sobol.h:17-248, hash.h:140-252) without running the reference.

Checks: generator matrices against the reference table text (fixture) and their closed forms,
MurmurHash64A against the SMHasher verification value, hash/scrambler constants against the fixture,
and the (0,m,2)-net stratification every pixel's sample set must have.
"""

import numpy as np
import pytest

M64 = (1 << 64) - 1


def test_matrices_match_reference_table(ora, ref_tables):
    assert ora.sobol_matrices().tolist() == ref_tables["sobol_matrices_dim01"]


def test_matrices_closed_forms(ref_tables):
    w = ref_tables["sobol_matrices_dim01"]
    dim0, dim1 = w[:52], w[52:]
    assert dim0[:32] == [0x80000000 >> i for i in range(32)]
    assert dim0[32:] == [0] * 20
    v = [0x80000000]
    for _ in range(31):
        v.append(v[-1] ^ (v[-1] >> 1))
    assert dim1[:32] == v
    assert dim1[32:] == v[:20]


def _murmur_py(key: bytes, seed: int, m: int, r: int) -> int:
    h = (seed ^ (len(key) * m)) & M64
    n8 = len(key) // 8
    for i in range(n8):
        k = int.from_bytes(key[8 * i : 8 * i + 8], "little")
        k = (k * m) & M64
        k ^= k >> r
        k = (k * m) & M64
        h ^= k
        h = (h * m) & M64
    tail = key[8 * n8 :]
    if tail:
        for i in reversed(range(len(tail))):
            h ^= tail[i] << (8 * i)
        h = (h * m) & M64
    h ^= h >> r
    h = (h * m) & M64
    h ^= h >> r
    return h


def test_murmur64a_smhasher_verification(ora):
    # SMHasher VerificationTest: hash keys {0..i-1} with seed 256-i, then hash the 256 hashes with
    # seed 0; the low 32 bits of that are MurmurHash64A's published verification value 0x1F0D3804.
    hashes = b"".join(ora.murmur64a(bytes(range(i)), 256 - i).to_bytes(8, "little") for i in range(256))
    assert ora.murmur64a(hashes, 0) & 0xFFFFFFFF == 0x1F0D3804


def test_murmur64a_constants(ora, ref_tables):
    rng = np.random.default_rng(3)
    for n in (0, 1, 7, 8, 9, 12, 16, 23):
        key = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        seed = int(rng.integers(0, 2**63))
        assert ora.murmur64a(key, seed) == _murmur_py(key, seed, ref_tables["murmur64a_m"], ref_tables["murmur64a_r"])


def test_mix_bits_constants_and_bijection(ora, ref_tables):
    c1, c2 = ref_tables["mix_bits_constants"]
    s1, s2, s3 = ref_tables["mix_bits_shifts"]

    def mix(v):
        v ^= v >> s1
        v = (v * c1) & M64
        v ^= v >> s2
        v = (v * c2) & M64
        v ^= v >> s3
        return v

    rng = np.random.default_rng(4)
    vals = [int(x) for x in rng.integers(0, 2**63, 200)] + [0, 1, M64]
    outs = [ora.mix_bits(v) for v in vals]
    assert outs == [mix(v) for v in vals]
    assert len(set(outs)) == len(set(vals))  # xorshift + odd multiply: invertible


def test_owen_scrambler_constants(ora, ref_tables):
    a, b, c = ref_tables["owen_scrambler_constants"]
    M32 = 0xFFFFFFFF

    def rev(x):
        return int(f"{x:032b}"[::-1], 2)

    def scr(v, seed):
        v = rev(v)
        v ^= (v * a) & M32
        v = (v + seed) & M32
        v = (v * ((seed >> 16) | 1)) & M32
        v ^= (v * b) & M32
        v ^= (v * c) & M32
        return rev(v)

    rng = np.random.default_rng(5)
    for _ in range(200):
        v, s = (int(x) for x in rng.integers(0, 2**32, 2))
        assert ora.owen_scramble(v, s) == scr(v, s)


def _is_net(points_u32: np.ndarray, m: int) -> bool:
    """(0, m, 2)-net check on 2^m points given as uint32 pairs: every elementary interval of volume 2^-m
    (2^a x 2^(m-a) grid cells, a = 0..m) holds exactly one point."""
    x = points_u32[:, 0].astype(np.uint64)
    y = points_u32[:, 1].astype(np.uint64)
    for a in range(m + 1):
        cx = x >> np.uint64(32 - a) if a else np.zeros_like(x)
        cy = y >> np.uint64(32 - (m - a)) if m - a else np.zeros_like(y)
        cells = cx * np.uint64(1 << (m - a)) + cy
        if len(np.unique(cells)) != len(cells):
            return False
    return True


@pytest.mark.parametrize("spp", [1, 2, 4, 8, 16, 32, 64])
def test_pixel_samples_form_nets(ora, spp):
    # The spp samples of one pixel are one aligned block of the (0,2)-sequence, re-ordered by the
    # per-digit permutations and Owen-scrambled; every 2D draw of the pixel must form a (0,m,2)-net
    # and every 1D draw a (0,m,1)-net.
    m = spp.bit_length() - 1
    kinds = [2, 1, 2, 1, 2, 1]
    rng = np.random.default_rng(spp)
    W, H = 256, 144
    for _ in range(6):
        px, py = int(rng.integers(0, W)), int(rng.integers(0, H))
        vals = np.stack([ora.sobol(W, H, spp, 7, px, py, s, kinds)[0] for s in range(spp)])
        col = 0
        for k in kinds:
            if k == 2:
                assert _is_net(vals[:, col : col + 2], m), (px, py, col)
            else:
                cells = vals[:, col].astype(np.uint64) >> np.uint64(32 - m) if m else np.zeros(spp, np.uint64)
                assert len(np.unique(cells)) == spp
            col += k


def test_distinct_pixels_decorrelated(ora):
    # neighbouring pixels draw different points (the Morton index and permutations differ)
    a = ora.sobol(64, 64, 16, 0, 10, 10, 3, [2])[0]
    b = ora.sobol(64, 64, 16, 0, 11, 10, 3, [2])[0]
    c = ora.sobol(64, 64, 16, 1, 10, 10, 3, [2])[0]
    assert not np.array_equal(a, b) and not np.array_equal(a, c)


def test_float_map(ora):
    u, f = ora.sobol(32, 32, 4, 0, 1, 2, 0, [2, 1, 2])
    expect = np.minimum(u.astype(np.float32) * np.float32(2.0**-32), np.float32(np.nextafter(np.float32(1), 0)))
    assert np.array_equal(f, expect)
    assert (f < 1.0).all() and (f >= 0.0).all()
