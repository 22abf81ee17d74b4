"""Writes the known-answer fixtures used by the oracle tests.

Sources (none of them is the reference code):
  * Philox4x32-10: the known-answer vectors published with Random123
    (Salmon et al., SC'11; kat_vectors, "philox4x32 10" lines).
  * MT19937 init_by_array({0x123, 0x234, 0x345, 0x456}): CPython's `random`
    module, an independent MT19937 implementation whose integer seeding is
    init_by_array over the 32-bit words of the seed; its first outputs are the
    published mt19937ar.out values.
Run:  python tests/golden/make_golden.py
"""
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))

PHILOX_KAT = [
    # counter (4 x u32), key (2 x u32), expected (4 x u32)
    ([0x00000000, 0x00000000, 0x00000000, 0x00000000], [0x00000000, 0x00000000],
     [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff], [0xffffffff, 0xffffffff],
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def main():
    with open(os.path.join(HERE, "philox4x32_10_kat.json"), "w") as f:
        json.dump([{"ctr": c, "key": k, "out": o} for c, k, o in PHILOX_KAT], f, indent=1)
    key = [0x123, 0x234, 0x345, 0x456]
    seed = sum(v << (32 * i) for i, v in enumerate(key))
    r = random.Random(seed)
    outs = [r.getrandbits(32) for _ in range(2000)]
    with open(os.path.join(HERE, "mt19937_init_by_array_123_234_345_456.json"), "w") as f:
        json.dump({"key": key, "outputs": outs}, f)


if __name__ == "__main__":
    main()
