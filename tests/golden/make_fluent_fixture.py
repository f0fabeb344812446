"""Copy the reference's Fluent mesh files into the test fixtures, gzipped (data only).

* ``tests/data/SquareCylinder-Coarse.msh``: the reference's regression mesh, an ASCII Fluent
  mesh written by GAMBIT (5 824 quads, square cylinder D = 4 in a channel; see that folder's
  README).
* ``examples/heated_duct_tri.msh`` (ASCII, 2 400 skewed triangles, zones inlet, outlet, floor,
  ceiling), ``examples/cavity_quad.msh`` (binary, double precision, a 40 x 40 cavity with zone
  ``lid``) and ``examples/elbow-mix.msh.txt`` (ASCII from GAMBIT, a 2-D mixing elbow with two
  inlets): the reference's example meshes (``examples/README.md``).

Each fixture is the same bytes, gzip-compressed; the reader takes .gz directly.
Run from the repository root: ``python tests/golden/make_fluent_fixture.py``.
"""
import gzip
import os

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "tests/data/SquareCylinder-Coarse.msh": "SquareCylinder-Coarse.msh.gz",
    "examples/heated_duct_tri.msh": "example_heated_duct_tri.msh.gz",
    "examples/cavity_quad.msh": "example_cavity_quad.msh.gz",
    "examples/elbow-mix.msh.txt": "example_elbow_mix.msh.gz",
}

if __name__ == "__main__":
    for src, dst in FILES.items():
        with open(os.path.join(REF, src), "rb") as fh:
            data = fh.read()
        out = os.path.join(HERE, dst)
        with open(out, "wb") as fh:
            fh.write(gzip.compress(data, 9, mtime=0))
        print(f"wrote {out}: {len(data)} bytes -> {os.path.getsize(out)} gzipped")
