"""Pins the oracle's Philox4x32-10 to the Random123 known-answer vectors."""

import numpy as np

from oracle import philox


def test_philox_known_answers():
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        got = philox.philox4x32_10(np.array(ctr, dtype=np.uint32), key)
        assert tuple(int(v) for v in got) == want


def test_host_key_split_matches_oracle_philox():
    from flowmc_amd import random as fr

    key = fr.PRNGKey(1234)
    kids = fr.split(key, 3)
    for i in range(3):
        b = philox.philox4x32_10(np.array([i, 0, 0, 0x53504C54], dtype=np.uint32), key)
        assert (kids[i] == b[:2]).all()
    assert (fr.split(key, 3) == kids).all()  # determinism


def test_uniform_open_interval_and_normal_moments():
    bits = np.array([0, 1, 511, 512, 0xFFFFFFFF], dtype=np.uint32)
    u = philox.u01(bits)
    assert u.min() > 0 and u.max() < 1
    assert u[0] == 2.0 ** -24 and u[-1] == 1 - 2.0 ** -24
    z = philox.normals((7, 9), 512, 8, 4, 0, philox.RNG_MALA_NOISE)
    assert z.shape == (4, 512, 8)
    assert abs(z.mean()) < 0.03 and abs(z.var() - 1) < 0.03
