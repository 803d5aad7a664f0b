"""The key layout and hash (gpu.py:231-252) pinned by closed-form values and by two independent restatements
(the oracle's C uint64 code and a numpy int64 one), including negative indices and the +-2^20 field wrap."""

import numpy as np

# Worked by hand from the closed form: x = (ix + 2^20) & (2^21 - 1), key = x | y << 21 | z << 42,
# h = key * 0x9E3779B97F4A7C15 (mod 2^64), h ^= h >> 29, slot_hash = h & 0x7fffffff.
KNOWN = [
    ((0, 0, 0), 0x4000020000100000, 1171322174),
    ((-1, -1, -1), 0x3FFFFDFFFFEFFFFF, 916897566),
    ((1, 2, 3), 0x40000E0000500001, 303645719),
]


def test_known_keys(orc):
    for ijk, key, h in KNOWN:
        assert orc.voxel_key_ijk(*ijk) == key
        assert int(orc.voxel_key_ijk_np(*ijk)) == key
        assert orc.slot_hash(key) == h
        assert int(orc.slot_hash_np(key)) == h
    assert orc.slot_hash(0x4000020000100000) & ((1 << 21) - 1) == 1111358


def test_field_wrap(orc):
    # 1048575 + 2^20 = 2^21 - 1 fills the x field; -2^20 + 2^20 = 0 leaves y empty
    assert orc.voxel_key_ijk(1048575, -1048576, 0) == 0x40000000001FFFFF
    # one past the range wraps into the same field (mask), as the reference's int64 & does
    assert orc.voxel_key_ijk(1048576, 0, 0) == orc.voxel_key_ijk(-1048576, 0, 0)


def test_two_restatements_agree(orc):
    rng = np.random.default_rng(1)
    ijk = np.concatenate([rng.integers(-2000, 2000, (4000, 3)), rng.integers(-(1 << 21), 1 << 21, (4000, 3))])
    keys_np = orc.voxel_key_ijk_np(ijk[:, 0], ijk[:, 1], ijk[:, 2])
    h_np = orc.slot_hash_np(keys_np)
    for (a, b, c), k, h in zip(ijk[::37], keys_np[::37], h_np[::37]):
        assert orc.voxel_key_ijk(int(a), int(b), int(c)) == int(k)
        assert orc.slot_hash(int(k)) == int(h)
    assert (h_np >= 0).all()


def test_voxel_key_floor(orc):
    # floor, not truncation: -0.01 m at 0.25 m voxels is index -1
    w = np.array([[-0.01, 0.0, 0.2499], [0.25, -0.25, 1e-9]], np.float32)
    k = orc.voxel_key_np(w, 4.0)
    assert int(k[0]) == orc.voxel_key_ijk(-1, 0, 0)
    assert int(k[1]) == orc.voxel_key_ijk(1, -1, 0)
