"""Copy the reference's regression mesh into the test fixtures, gzipped (data only).

The reference's tests hold ``tests/data/SquareCylinder-Coarse.msh`` (an ASCII Fluent mesh
written by GAMBIT: 5 824 quads, square cylinder D = 4 in a channel; see that folder's
README). The fixture is the same bytes, gzip-compressed; the reader takes .gz directly.
Run from the repository root: ``python tests/golden/make_fluent_fixture.py``.
"""
import gzip
import os

SRC = "/root/reference/tests/data/SquareCylinder-Coarse.msh"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "SquareCylinder-Coarse.msh.gz")

if __name__ == "__main__":
    with open(SRC, "rb") as fh:
        data = fh.read()
    with open(DST, "wb") as fh:
        fh.write(gzip.compress(data, 9, mtime=0))
    print(f"wrote {DST}: {len(data)} bytes -> {os.path.getsize(DST)} gzipped")
