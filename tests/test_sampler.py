"""The seeded soil sampler: Philox4x32-10 against the published known-answer
vectors (Random123's kat_vectors for philox4x32_10), and gf_soil_sample
against an independent pure-Python restatement of the same draws."""

import math

import pytest

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox_py(ctr, key):
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    for _ in range(10):
        p0, p1 = M0 * c0, M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ c3 ^ k1) & MASK, p0 & MASK
        k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
    return [c0, c1, c2, c3]


KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([MASK] * 4, [MASK, MASK], [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(fdtd, ctr, key, want):
    assert fdtd.Philox4x32(ctr, key) == want
    assert philox_py(ctr, key) == want


def sample_py(seed, index):
    key = [seed & MASK, (seed >> 32) & MASK]

    def u(block):
        return [(x + 0.5) / 4294967296.0 for x in philox_py([index, block, 0, 0], key)]

    above = 1 + min(2, int(u(0)[0] * 3))
    layers, vs_min = [], 1e30
    for l in range(above + 1):
        a = u(1 + l)
        vs = 120 + 280 * a[0]
        layer = (0.5 + 4.5 * u(5 + l)[0] if l < above else 0.0, vs * (1.7 + 0.8 * a[1]), vs, 1600 + 500 * a[2],
                 10 + 70 * a[3], 2 * (10 + 70 * a[3]))
        layers.append(layer)
        vs_min = min(vs_min, vs)
    h = 0.9 * vs_min / 1000.0

    def rnd(x):  # C lround: half away from zero
        return int(math.floor(x + 0.5))

    nr = max(600, min(2048, rnd(218.0 / h) + 21))
    nz = max(300, min(1024, rnd(108.0 / h) + 21))
    return layers, h, (nr - 21) * h, (nz - 21) * h


@pytest.mark.parametrize("seed", [0, 1, 2 ** 40 + 3])
def test_sampler_matches_an_independent_restatement(fdtd, seed):
    for index in [0, 1, 7, 255, 4095, 70000]:
        s = fdtd.SampleSoil(seed, index)
        layers, h, rmax, depth = sample_py(seed, index)
        got = [(l.Thickness, l.Vp, l.Vs, l.Density, l.Qs, l.Qp) for l in s.Stack]
        assert got == layers
        assert (s.Spacing, s.MaxRange, s.Depth, s.PMLCells) == (h, rmax, depth, 20)


def test_sampled_models_are_valid_and_span_the_stated_ranges(fdtd):
    nrs, nzs = [], []
    for i in range(256):
        s = fdtd.SampleSoil(0, i)
        assert 2 <= len(s.Stack) <= 4
        for l in s.Stack:
            assert 120 <= l.Vs <= 400 and 1.7 <= l.Vp / l.Vs <= 2.5 and 1600 <= l.Density <= 2100
            assert 10 <= l.Qs <= 80 and l.Qp == 2 * l.Qs
        assert all(0.5 <= l.Thickness <= 5 for l in s.Stack[:-1]) and s.Stack[-1].Thickness == 0
        m = fdtd.Model(Stack=s.Stack, MaxRange=s.MaxRange, Depth=s.Depth, Spacing=s.Spacing, PMLCells=20,
                       GslsMechanisms=3)
        dt, nr, nz = m.Grid()
        assert 600 <= nr <= 2048 and 300 <= nz <= 1024
        nrs.append(nr)
        nzs.append(nz)
        assert s.MaxRange > 50  # every receiver of the bank grid is inside
    assert max(nrs) - min(nrs) > 800 and max(nzs) - min(nzs) > 400
    assert fdtd.SampleSoil(0, 17) == fdtd.SampleSoil(0, 17)
