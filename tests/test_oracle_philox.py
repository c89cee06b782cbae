"""Pins the oracle's Philox4x32-10 and uniform conversion.

Known answers: the Random123 published philox4x32_10 vectors, and 66 vectors
of rocRAND's independent host engine (tests/golden/philox_rocrand.txt,
written by tests/golden/make_philox_rocrand.sh)."""
import os

import numpy as np

from conftest import GOLDEN

RANDOM123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def load_rocrand_vectors():
    rows = []
    with open(os.path.join(GOLDEN, "philox_rocrand.txt")) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([int(t, 16) for t in line.split()])
    return np.array(rows, dtype=np.uint64).astype(np.uint32)


def test_random123_kat(O):
    for ctr, key, exp in RANDOM123_KAT:
        assert O.philox(ctr, key)[0].tolist() == list(exp)


def test_rocrand_engine_vectors(O):
    v = load_rocrand_vectors()
    assert len(v) == 65
    got = O.philox(v[:, 0:4], v[:, 4:6])
    np.testing.assert_array_equal(got, v[:, 6:10])


def test_counter_sensitivity(O):
    # every counter word and key word changes every output word
    base = O.philox((1, 2, 3, 4), (5, 6))[0]
    for i in range(4):
        c = [1, 2, 3, 4]
        c[i] ^= 1
        assert np.all(O.philox(c, (5, 6))[0] != base)
    assert np.all(O.philox((1, 2, 3, 4), (5, 7))[0] != base)


def test_uniform_conversion(O):
    L = O.lib()
    L.ppo_u01.restype = __import__("ctypes").c_float
    L.ppo_u01.argtypes = [__import__("ctypes").c_uint32]
    L.ppo_u01_open0.restype = __import__("ctypes").c_float
    L.ppo_u01_open0.argtypes = [__import__("ctypes").c_uint32]
    assert L.ppo_u01(0) == 0.0
    assert L.ppo_u01(0xFFFFFFFF) == 1.0 - 2.0**-24
    assert L.ppo_u01_open0(0) == 2.0**-24
    assert L.ppo_u01_open0(0xFFFFFFFF) == 1.0
    assert L.ppo_u01(0x80000000) == 0.5
