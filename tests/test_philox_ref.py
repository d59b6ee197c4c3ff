"""The reference Philox4x32-10 used to check the device noise stream, pinned
to the published known-answer vectors of Random123 (kat_vectors)."""
import numpy as np

from philox_ref import philox4x32_10

KAT = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
       ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
       ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]


def test_philox_known_answers():
    for ctr, key, expect in KAT:
        got = philox4x32_10([np.uint64(c) for c in ctr], [np.uint64(k) for k in key])
        assert tuple(int(x) for x in got) == expect
