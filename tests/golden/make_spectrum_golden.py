"""Writes spectrum_golden.json: the reference's own known-answer vectors for
the RGB->spectrum MLP, inputTest[4][100] and gtTest[100]
(core/spectrum/spectrum_basis.cuh:2472-2489), produced by the reference's
eval-mode PyTorch model (pytorch/spectrum/train_spectrum.py:417-443). Rows of
inputTest are R, G, B and wavelength (nm); gtTest is the clamped basis dot
product. Run in a container with the reference mounted; the JSON is committed.
"""
import json
import os
import re

REF = "/root/reference/core/spectrum/spectrum_basis.cuh"
NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?f?")


def body(text, name):
    m = re.search(r"inline constexpr float " + re.escape(name) + r"\[[^=]*=", text)
    i = text.index("{", m.end())
    depth, j = 1, i + 1
    while depth:
        depth += {"{": 1, "}": -1}.get(text[j], 0)
        j += 1
    return [float(x.rstrip("f")) for x in NUM.findall(text[i + 1:j - 1])]


def main():
    text = open(REF).read()
    inp = body(text, "inputTest")
    gt = body(text, "gtTest")
    assert len(inp) == 400 and len(gt) == 100
    out = {"source": "core/spectrum/spectrum_basis.cuh:2472-2489",
           "rgb": [[inp[i], inp[100 + i], inp[200 + i]] for i in range(100)],
           "lambda_nm": inp[300:400], "gt": gt}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "spectrum_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
