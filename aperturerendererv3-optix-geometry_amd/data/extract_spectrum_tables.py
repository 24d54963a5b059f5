"""Extracts the numeric tables of the reference's spectral upsampling into
binary data files of this package (run once, in a container that has the
reference mounted; the outputs are committed).

Sources (tables only; no code is taken):
  core/spectrum/spectrum_basis.cuh:24-2259  RGB/LAM MLP weights (W32 D8 K4)
  core/spectrum/ciexyz.cuh                  D65 SPD (421) and CIE 1931 2-degree CMF (421 x XYZ)

Outputs (little-endian float32, concatenated in the order listed):
  spectrum_mlp.bin : RGB_IN_W[1024] RGB_IN_B[32] {RGB_BLKk_W[1024] RGB_BLKk_B[32]} k=0..6 RGB_OUT_W[128]
                     {LAM_BLKk_W[1024] LAM_BLKk_B[32]} k=0..6 LAM_OUT_W[128]
  cie_tables.bin   : D65_SPD[421] CIEXYZ_CMF[421*3]
"""
import os
import re
import sys

import numpy as np

REF = "/root/reference/core/spectrum"
HERE = os.path.dirname(os.path.abspath(__file__))

MLP_ORDER = (["RGB_IN_W", "RGB_IN_B"] + [f"RGB_BLK{k}_{p}" for k in range(7) for p in "WB"] + ["RGB_OUT_W"]
             + [f"LAM_BLK{k}_{p}" for k in range(7) for p in "WB"] + ["LAM_OUT_W"])

NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?f?")


def array_body(text, name):
    """Numbers of the host (#else) initialiser of `name`."""
    m = re.search(r"inline constexpr float " + re.escape(name) + r"\[[^\]]*\]\s*=", text)
    i = text.index("{", m.end()) + 1
    depth, j = 1, i
    while depth:
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
        j += 1
    body = text[i:j - 1]
    return np.array([float(x.rstrip("f")) for x in NUM.findall(body)], dtype=np.float32)


def main():
    basis = open(os.path.join(REF, "spectrum_basis.cuh")).read()
    parts = []
    for name in MLP_ORDER:
        a = array_body(basis, name)
        want = 128 if name.endswith("OUT_W") else (32 if name.endswith("_B") else 1024)
        assert a.size == want, (name, a.size)
        parts.append(a)
    np.concatenate(parts).astype("<f4").tofile(os.path.join(HERE, "spectrum_mlp.bin"))

    cie = open(os.path.join(REF, "ciexyz.cuh")).read()
    d65 = array_body(cie, "D65_SPD")
    m = re.search(r"inline constexpr float3 CIEXYZ_CMF\[[^\]]*\]\s*=", cie)
    i = cie.index("{", m.start())
    j = cie.index("};", i)
    cmf = np.array([float(x.rstrip("f")) for x in NUM.findall(cie[i:j].replace("float3", ""))], dtype=np.float32)
    assert d65.size == 421 and cmf.size == 421 * 3, (d65.size, cmf.size)
    np.concatenate([d65, cmf]).astype("<f4").tofile(os.path.join(HERE, "cie_tables.bin"))
    print("wrote", sum(p.size for p in parts), "MLP floats and", d65.size + cmf.size, "CIE floats")


if __name__ == "__main__":
    sys.exit(main())
