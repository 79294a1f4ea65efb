"""Runs the CPU oracle's whole path (FP32 and FP64 builds) over every CFA
layout, the three level filters, local tone mapping, the high-noise branch
and a size whose pyramid is not an exact 2x chain.  Started by
tests/test_oracle_sanitize.py with the sanitizer build of the oracle."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from oracle import oracle as O  # noqa: E402

for dbl in (False, True):
    o = O.Oracle(double=dbl)
    for w, h in ((64, 48), (70, 50)):
        for bayer in range(4):
            for alg in range(3):
                raw = O.generate_bayer(w, h, bayer, 1000 + bayer, 2000 + alg, 0.004, 2e-5, 64.0, 4095.0)
                p = O.default_params(bayer, 64.0, 4095.0, levels=4, algorithm=alg)
                p.localToneMapping = int(alg == 0)
                p.noiseLevel = 0.7 if alg == 1 else 0.0
                p.dp[0].gradientBoost = 2.0
                p.dp[0].sharpening = 1.5
                o.raw_to_rgb(raw, p)
print("oracle sanitizer run: ok")
