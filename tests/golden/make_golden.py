"""Writes the committed fixtures of tests/golden/.

philox_kat.json: the Philox4x32-10 known-answer vectors published with
Random123 (Salmon et al., SC'11; file kat_vectors of the Random123
distribution, lines "philox4x32 10 ..."): counter/key all zero, all ones, and
the pi-digit counter with key a4093822 299f31d0.  rocRAND's independent
engine reproduces all three (tests/test_rng.py runs it).

Run: python tests/golden/make_golden.py
"""
import json
import pathlib

HERE = pathlib.Path(__file__).resolve().parent


def h(s):
    return [int(x, 16) for x in s.split()]


KATS = [
    {"ctr": h("00000000 00000000 00000000 00000000"), "key": h("00000000 00000000"),
     "out": h("6627e8d5 e169c58d bc57ac4c 9b00dbd8")},
    {"ctr": h("ffffffff ffffffff ffffffff ffffffff"), "key": h("ffffffff ffffffff"),
     "out": h("408f276d 41c83b0e a20bc7c6 6d5451fd")},
    {"ctr": h("243f6a88 85a308d3 13198a2e 03707344"), "key": h("a4093822 299f31d0"),
     "out": h("d16cfe09 94fdcceb 5001e420 24126ea1")},
]

if __name__ == "__main__":
    (HERE / "philox_kat.json").write_text(json.dumps({"random123": KATS}, indent=1) + "\n")
    print("wrote", HERE / "philox_kat.json")
