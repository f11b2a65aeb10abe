"""The 16-byte key word's mixing function, restated for the tests that build
keys sharing a key word: the SplitMix64 output function (Steele, Lea, Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014) of
c + seed * 0x9E3779B97F4A7C15. Pinned by the generator's published first
outputs from state 0 (the state advances by the golden ratio before each
output, so output i is wide_mix(0, i + 1))."""

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def wide_mix(c, seed):
    z = (c + seed * GOLDEN) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_splitmix64_known_outputs():
    # SplitMix64 seeded with 0: its first three outputs
    assert [wide_mix(0, i) for i in (1, 2, 3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]

