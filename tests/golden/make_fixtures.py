#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from data the reference holds.

Run in the build container (it reads /root/reference as data, never runs it):

  fuzz_msm.json  every input of test/fuzz/corpus/blst_fuzz_msm/ decoded the
                 way test/fuzz/targets/msm.cpp and
                 test/fuzz/internal/adapter_msm.cpp:19-50, 251-258 do
                 (hex: seeds, n = first byte % 16 + 1, points (i+1)G, one-byte
                 scalars input[i % len] + i + 1, nbits 8), with the expected
                 MSM and the bit-0, window-8 tile of the 7-bit scalars, both
                 computed by the textbook affine group law of pyref.py
  h2c_g1.json    the G1 outputs P of the hash-to-curve vector files
                 (bindings/vectors/hash_to_curve/BLS12381G1_*.json): known
                 subgroup points with their compressed encodings
  h2c_g2.json    the G2 outputs P of the G2 hash-to-curve vector files
                 (bindings/vectors/hash_to_curve/BLS12381G2_*.json), with
                 compressed encodings by the textbook model of pyref2.py
  kat.json       the compressed generators of test/Conformance_tests.cpp:15-21
                 (draft-irtf-cfrg-pairing-friendly-curves-13, B.1)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import pyref  # noqa: E402
import pyref2  # noqa: E402

REF = "/root/reference"


def decode_hex_seed(data, capacity=64):
    """test/fuzz/support/seed_encoding.h:56-93"""
    if not data.startswith(b"hex:"):
        return None
    body = data[4:]
    while body and body[-1:] in (b"\n", b"\r"):
        body = body[:-1]
    if not body or len(body) % 2 or len(body) // 2 > capacity:
        return None
    if any(ch not in b"0123456789abcdefABCDEF" for ch in body):
        return None
    return bytes(int(body[i:i + 2], 16) for i in range(0, len(body), 2))


def fuzz_cases():
    d = os.path.join(REF, "test/fuzz/corpus/blst_fuzz_msm")
    cases = []
    for name in sorted(os.listdir(d)):
        raw = open(os.path.join(d, name), "rb").read()
        dec = decode_hex_seed(raw)
        inp = dec if dec is not None else raw
        n = 1 if len(inp) == 0 else inp[0] % 16 + 1
        scalars = [(i + 1) if not inp else (inp[i % len(inp)] + i + 1) % 256 for i in range(n)]
        points = [pyref.mul(i + 1, pyref.G) for i in range(n)]
        expected = pyref.msm(points, scalars)
        tile_scalars = [s & 0x7F for s in scalars]
        tile = pyref.msm(points, tile_scalars)
        cases.append({
            "file": name,
            "n": n,
            "nbits": 8,
            "point_multipliers": [i + 1 for i in range(n)],
            "scalars": scalars,
            "expected": pyref.compress(expected).hex(),
            "tile_scalars": tile_scalars,
            "tile_bit0": 0,
            "tile_window": 8,
            "tile_expected": pyref.compress(tile).hex(),
        })
    return cases


def h2c_points():
    out = []
    d = os.path.join(REF, "bindings/vectors/hash_to_curve")
    for fn in sorted(os.listdir(d)):
        if not fn.startswith("BLS12381G1_") or not fn.endswith(".json"):
            continue
        js = json.load(open(os.path.join(d, fn)))
        for v in js["vectors"]:
            x = int(v["P"]["x"], 16)
            y = int(v["P"]["y"], 16)
            assert pyref.on_curve((x, y))
            out.append({"suite": fn, "msg": v["msg"], "x": "%096x" % x, "y": "%096x" % y,
                        "compressed": pyref.compress((x, y)).hex()})
    return out


def h2c_points_g2():
    out = []
    d = os.path.join(REF, "bindings/vectors/hash_to_curve")
    for fn in sorted(os.listdir(d)):
        if not fn.startswith("BLS12381G2_") or not fn.endswith(".json"):
            continue
        js = json.load(open(os.path.join(d, fn)))
        for v in js["vectors"]:
            x = tuple(int(c, 16) for c in v["P"]["x"].split(","))
            y = tuple(int(c, 16) for c in v["P"]["y"].split(","))
            assert pyref2.on_curve((x, y))
            out.append({"suite": fn, "msg": v["msg"], "x": ["%096x" % c for c in x], "y": ["%096x" % c for c in y],
                        "compressed": pyref2.compress((x, y)).hex()})
    return out


def main():
    fz = fuzz_cases()
    with open(os.path.join(HERE, "fuzz_msm.json"), "w") as f:
        json.dump(fz, f, indent=1)
    hc = h2c_points()
    with open(os.path.join(HERE, "h2c_g1.json"), "w") as f:
        json.dump(hc, f, indent=1)
    hc2 = h2c_points_g2()
    with open(os.path.join(HERE, "h2c_g2.json"), "w") as f:
        json.dump(hc2, f, indent=1)
    kat = {
        "g1_generator_compressed": "97f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac5"
                                   "86c55e83ff97a1aeffb3af00adb22c6bb",
        "g1_infinity_compressed": "c0" + "00" * 47,
        "g2_generator_compressed": "93e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f504"
                                   "9334cf11213945d57e5ac7d055d042b7e024aa2b2f08f0a91260805272dc510"
                                   "51c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121b"
                                   "db8",
        "g2_infinity_compressed": "c0" + "00" * 95,
        "source": "test/Conformance_tests.cpp:15-21 (draft-irtf-cfrg-pairing-friendly-curves-13, "
                  "Appendix B.1); infinity: blst/scalar/curve/point.cpp:895-898",
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    print("fuzz cases %d, h2c points %d + %d" % (len(fz), len(hc), len(hc2)))


if __name__ == "__main__":
    main()
