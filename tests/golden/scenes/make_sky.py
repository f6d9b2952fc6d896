"""Writes sky.pfm, the 8 x 4 environment of room.xml: a vertical gradient (brighter at
the zenith row) with one bright texel."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", ".."))
from zillumgl_amd import image  # noqa: E402

h, w = 4, 8
env = np.zeros((h, w, 3), np.float32)
for i in range(h):
    env[i, :] = np.array([0.2, 0.3, 0.5], np.float32) * np.float32(1.0 + (h - 1 - i))
env[0, 3] = (20.0, 18.0, 15.0)
image.write_pfm(os.path.join(os.path.dirname(os.path.abspath(__file__)), "sky.pfm"), env)
