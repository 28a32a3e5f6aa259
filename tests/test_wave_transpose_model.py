"""CPU model of the batched BFS's wave-wide 64 x 64 bit transpose
(hornet_amd/csrc/msbfs.hip, wave_transpose64): the same six block-swap steps
over 64 simulated lanes must leave in lane b the bits of column b, so that
its popcount is the number of lanes whose word has bit b set."""
import numpy as np

M = [0x00000000FFFFFFFF, 0x0000FFFF0000FFFF, 0x00FF00FF00FF00FF,
     0x0F0F0F0F0F0F0F0F, 0x3333333333333333, 0x5555555555555555]
FULL = (1 << 64) - 1


def transpose_model(words):
    x = [int(w) for w in words]
    for k in range(6):
        j = 32 >> k
        m = M[k]
        p = [x[l ^ j] for l in range(64)]   # __shfl_xor(x, j)
        x = [((x[l] & ~m & FULL) | ((p[l] & ~m & FULL) >> j)) if (l & j)
             else ((x[l] & m) | (((p[l] & m) << j) & FULL)) for l in range(64)]
    return x


def bits(words):
    return np.array([[(int(w) >> b) & 1 for b in range(64)] for w in words], dtype=np.uint8)


def test_transpose_matches_numpy():
    rng = np.random.default_rng(3)
    for density in (0.02, 0.5, 0.97):
        B = (rng.random((64, 64)) < density).astype(np.uint8)
        words = [int(sum(int(B[l, b]) << b for b in range(64))) for l in range(64)]
        T = bits(transpose_model(words))
        assert np.array_equal(T, B.T)
        # the per-source counts the collect adds
        assert [bin(t).count("1") for t in transpose_model(words)] == list(B.sum(axis=0))


def test_transpose_special_cases():
    assert transpose_model([0] * 64) == [0] * 64
    assert transpose_model([FULL] * 64) == [FULL] * 64
    # identity is its own transpose; one set bit moves to the mirrored place
    eye = [1 << l for l in range(64)]
    assert transpose_model(eye) == eye
    w = [0] * 64
    w[5] = 1 << 40
    t = transpose_model(w)
    assert t[40] == 1 << 5 and sum(1 for v in t if v) == 1
