"""Writes tests/golden/ct_stack.npz from the reference's sample CT stack
(data/input/ct_stack: three 744 x 789 uint16 little-endian RAW slices) and
checks the known answer its documentation states for it: background 0 and
768151 foreground voxels (docs/en-us/examples/crop.md:111).

Run here, where the reference tree is present:
    python tests/golden/make_ct_fixture.py /root/reference
"""
import sys
from pathlib import Path

import numpy as np


def main(ref: str) -> None:
    files = sorted((Path(ref) / "data/input/ct_stack").iterdir())
    vol = np.stack([np.fromfile(f, dtype="<u2").reshape(789, 744) for f in files])
    assert vol.shape == (3, 789, 744)
    assert int(np.count_nonzero(vol != 0)) == 768151
    out = Path(__file__).resolve().parent / "ct_stack.npz"
    np.savez_compressed(out, samples=vol, names=np.array([f.name for f in files]))
    print(out, vol.shape, vol.dtype)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
