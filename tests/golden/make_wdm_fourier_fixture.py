"""Writes wdm_fourier_coefficients.json: known answers for the WDM band-edge
Fourier coefficients, read from the reference's coefficient table
cwb-core/FourierCoefficients.icc (Cos[beta][j], first 16 per beta, and the
table sizes).  Run in the build container, where /root/reference exists."""

import json
import os
import re

SRC = "/root/reference/cwb-core/FourierCoefficients.icc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wdm_fourier_coefficients.json")

vals, sizes = {}, {}
with open(SRC) as f:
    for line in f:
        m = re.match(r"\s*Cos\[(\d)\]\[(\d+)\]\s*=\s*([-+0-9.eE]+);", line)
        if m and int(m.group(2)) < 16:
            vals.setdefault(m.group(1), {})[int(m.group(2))] = float(m.group(3))
        m = re.match(r"\s*CosSize\[(\d)\]\s*=\s*(\d+);", line)
        if m:
            sizes[m.group(1)] = int(m.group(2))
out = {"source": "cwb-core/FourierCoefficients.icc Cos[beta][0..15], CosSize[beta]",
       "cos": {b: [v[j] for j in sorted(v)] for b, v in sorted(vals.items())},
       "size": sizes}
with open(OUT, "w") as f:
    json.dump(out, f, indent=1)
print(OUT)
