"""Binary monitor slices, the format the reference drivers write
(examples/pearson-3d-main.cpp:116-157) and its plotting script reads
(scripts/plot-pearson-bin.py:48-60): six int32 -- global position of the
lower corner (x, y, z) and the extents (x, y, z) -- then FP64 planes: U then V
of one slice (x = mid for monitorX, y = mid for monitorY), rows of z."""
from __future__ import annotations

from typing import Dict

import numpy as np


def read_monitor(path: str) -> Dict[str, np.ndarray]:
    with open(path, "rb") as f:
        head = np.frombuffer(f.read(24), dtype=np.int32)
        body = np.frombuffer(f.read(), dtype=np.float64)
    gx, gy, gz, sx, sy, sz = (int(v) for v in head)
    rows = sy if "monitorX" in path else sx
    planes = body.reshape(2, rows, sz)
    return {"position": (gx, gy, gz), "extent": (sx, sy, sz), "U": planes[0], "V": planes[1]}
