"""Copies the reference's sample particle mesh (data/input/particles.stl: a binary
STL of 5,600 faces, 15 closed particles) to tests/golden/particles.stl, after
checking its header's face count against its length.

Run here, where the reference tree is present:
    python tests/golden/copy_particles_stl.py /root/reference
"""
import shutil
import sys
from pathlib import Path


def main(ref: str) -> None:
    src = Path(ref) / "data/input/particles.stl"
    data = src.read_bytes()
    n = int.from_bytes(data[80:84], "little")
    assert n == 5600 and len(data) == 84 + 50 * n
    out = Path(__file__).resolve().parent / "particles.stl"
    shutil.copyfile(src, out)
    print(out, n, "faces")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
