"""Writes wdm_fourier_coefficients.json: known answers for the WDM band-edge
Fourier coefficients, read from the reference's coefficient tables
cwb-core/FourierCoefficients.icc (Cos, Cos2 and SinCos [beta][j], first 16 per
beta, and the table sizes).  Run in the build container, where
/root/reference exists."""

import json
import os
import re

SRC = "/root/reference/cwb-core/FourierCoefficients.icc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wdm_fourier_coefficients.json")

tables = {"Cos": ({}, {}), "Cos2": ({}, {}), "SinCos": ({}, {})}
with open(SRC) as f:
    for line in f:
        for name, (vals, sizes) in tables.items():
            m = re.match(r"\s*%s\[(\d)\]\[(\d+)\]\s*=\s*([-+0-9.eE]+);" % name, line)
            if m and int(m.group(2)) < 16:
                vals.setdefault(m.group(1), {})[int(m.group(2))] = float(m.group(3))
            m = re.match(r"\s*%sSize\[(\d)\]\s*=\s*(\d+);" % name, line)
            if m:
                sizes[m.group(1)] = int(m.group(2))


def dense(v):
    """Coefficients 0..15; entries the table leaves at its zero fill are 0."""
    return [v.get(j, 0.0) for j in range(16)]


out = {"source": "cwb-core/FourierCoefficients.icc Cos/Cos2/SinCos[beta][0..15] and sizes",
       "cos": {b: dense(v) for b, v in sorted(tables["Cos"][0].items())},
       "size": tables["Cos"][1],
       "cos2": {b: dense(v) for b, v in sorted(tables["Cos2"][0].items())},
       "cos2_size": tables["Cos2"][1],
       "sincos": {b: dense(v) for b, v in sorted(tables["SinCos"][0].items())},
       "sincos_size": tables["SinCos"][1]}
with open(OUT, "w") as f:
    json.dump(out, f, indent=1)
print(OUT)
