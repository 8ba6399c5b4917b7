"""Pins the oracle's random numbers (include/util/random.h) against published answers."""
import numpy as np

MASK64 = (1 << 64) - 1


def pcg32_paper(initstate, initseq, n):
    """pcg32_srandom_r / pcg32_random_r as printed in O'Neill, "PCG: A Family of Simple Fast
    Space-Efficient Statistically Good Algorithms for Random Number Generation" (2014)."""
    state = 0
    inc = ((initseq << 1) | 1) & MASK64
    out = []

    def step():
        nonlocal state
        old = state
        state = (old * 6364136223846793005 + inc) & MASK64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    step()
    state = (state + initstate) & MASK64
    step()
    for _ in range(n):
        out.append(step())
    return out


def test_pcg32_reference_demo(orc):
    # pcg32-demo of the PCG C reference implementation, seeded with (42, 54).
    expected = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    assert list(orc.pcg32_stream(42, 54, 6)) == expected
    assert pcg32_paper(42, 54, 6) == expected


def test_pcg32_matches_paper_restatement(orc):
    rng = np.random.default_rng(0)
    for _ in range(20):
        s, q = (int(x) for x in rng.integers(0, 2**63, 2))
        assert list(orc.pcg32_stream(s, q, 200)) == pcg32_paper(s, q, 200)


def test_next_float_and_u64_conventions(orc):
    u = orc.pcg32_stream(7, 0xDA3E39CB94B95BDB, 1000, kind=0)
    f = orc.pcg32_stream(7, 0xDA3E39CB94B95BDB, 1000, kind=1)
    # NextFloat = (u >> 9) * 2^-23 exactly (random.h:76-81)
    assert np.array_equal(f, ((u >> 9).astype(np.float64) / 2**23).astype(np.float32))
    assert f.min() >= 0.0 and f.max() < 1.0
    # NextUint64: low word from the first draw
    w = orc.pcg32_stream(7, 0xDA3E39CB94B95BDB, 500, kind=2)
    assert np.array_equal(w, u[0::2].astype(np.uint64) | (u[1::2].astype(np.uint64) << np.uint64(32)))


def test_tea_published_vector(orc):
    # TEA (Wheeler & Needham 1994), 32 rounds, all-zero key and block.
    assert orc.tea_keyed(0, 0, [0, 0, 0, 0], 32) == (0x41EA3A0A, 0x94BAA940)


def test_sample_tea_is_keyed_tea(orc):
    key = [0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E]
    rng = np.random.default_rng(1)
    for a, b in rng.integers(0, 2**32, (50, 2)):
        v0, v1 = orc.tea_keyed(int(a), int(b), key, 4)
        assert orc.tea64(int(a), int(b), 4) == (v0 | (v1 << 32))
        assert orc.tea32(int(a), int(b), 4) == v0
