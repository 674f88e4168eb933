"""Explicit-height octahedron model, written independently of any slope
encoding, used to pin the oracle's update rules and MCS compositions.

A site is a local minimum when all four neighbours are one unit higher and a
local maximum when all four are one unit lower (Odor et al. 2009, the model
behind kpz/schedulerCPU.inl:21-33); a p-flip raises a minimum by 2, a q-flip
lowers a maximum by 2.  Heights are compared modulo a global constant.
"""
import math

import numpy as np


def thresh23(p):
    """number of 23-bit mantissas r with r * 2^-23 < p"""
    if not p > 0.0:
        return 0
    if p >= 1.0:
        return 1 << 23
    return int(math.ceil(p * 8388608.0))


class HeightLattice:
    def __init__(self, h):
        self.h = np.array(h, dtype=np.int64)
        self.Y, self.X = self.h.shape

    def kind(self, x, y):
        h = self.h
        X, Y = self.X, self.Y
        c = h[y, x]
        nb = (h[y, (x - 1) % X], h[y, (x + 1) % X], h[(y - 1) % Y, x], h[(y + 1) % Y, x])
        if all(v == c + 1 for v in nb):
            return 1
        if all(v == c - 1 for v in nb):
            return -1
        return 0

    def update(self, x, y, accept_min=True, accept_max=False):
        k = self.kind(x, y)
        if k == 1 and accept_min:
            self.h[y, x] += 2
        elif k == -1 and accept_max:
            self.h[y, x] -= 2

    def normalized(self):
        return self.h - self.h[0, 0]


def normalized(h):
    h = np.asarray(h, dtype=np.int64)
    return h - h[0, 0]
