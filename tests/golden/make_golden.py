"""Writes tests/golden/kernel_values.json: the cubic spline kernel
(SPH_kernels.h:42-94) at 11 q points, evaluated exactly with rationals
(pi kept symbolic until the final division).  Run: python tests/golden/make_golden.py"""
import json
import math
import os
from fractions import Fraction as Fr

HK = Fr(1, 10)


def W(q):
    # k = 8 / (pi h^3); q <= 1/2: k (6q^3 - 6q^2 + 1); q <= 1: 2k (1 - q)^3
    if q <= Fr(1, 2):
        poly = 6 * q ** 3 - 6 * q ** 2 + 1
    elif q <= 1:
        poly = 2 * (1 - q) ** 3
    else:
        poly = Fr(0)
    return float(8 * poly / HK ** 3) / math.pi


qs = [Fr(k, 10) for k in range(11)]
out = {"hk": float(HK), "q": [float(q) for q in qs], "W": [W(q) for q in qs]}
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernel_values.json")
with open(path, "w") as f:
    json.dump(out, f, indent=1)
print("wrote", path)
