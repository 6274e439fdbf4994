"""Cuts a 3-frame fixture out of the reference's data file data/synthetic_liver_l1_single_sweep.npz (read with
numpy, allow_pickle=False: arrays only). Run in a container where the reference is mounted:
    python tests/golden/make_liver_fixture.py /root/reference/data/synthetic_liver_l1_single_sweep.npz
Frames 0, 120 and 239 keep every per-frame array; scalar metadata is copied as is."""

import sys
from pathlib import Path

import numpy as np

src = Path(sys.argv[1])
out = Path(__file__).resolve().parent / "liver_3frames.npz"
raw = np.load(src, allow_pickle=False)
n = raw["images"].shape[0]
pick = np.array([0, n // 2, n - 1])
data = {}
for k in raw.files:
    a = raw[k]
    data[k] = a[pick] if (a.ndim >= 1 and a.shape[0] == n) else a
np.savez_compressed(out, **data)
print(out, {k: v.shape for k, v in data.items()})
