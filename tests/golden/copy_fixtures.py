"""Copies the mesh fixtures of the reference's own tests into tests/golden/.

Data files only (ALBERT meshes used by test/*.c and examples/):
    test/cube.dat   6 Kuhn tetrahedra of the unit cube
    test/tetra.dat  the reference tetrahedron
    test/cube4.dat  cube.dat refined to 384 tetrahedra
Run once in a tree where the reference is present:
    python tests/golden/copy_fixtures.py [/path/to/reference]
"""
import os
import shutil
import sys

FILES = ("test/cube.dat", "test/tetra.dat", "test/cube4.dat")


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    for f in FILES:
        shutil.copyfile(os.path.join(ref, f), os.path.join(here, os.path.basename(f)))
        print("copied", f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
