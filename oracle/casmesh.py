"""Fluent ``.cas`` writer used to run the compiled reference for comparisons (test
infrastructure): the writer itself is the library's, mpdugks_amd/mesh.py."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from mpdugks_amd.mesh import write_cas  # noqa: E402,F401

if __name__ == "__main__":
    # usage: casmesh.py OUT NX NY [walls]
    write_cas(sys.argv[1], int(sys.argv[2]), int(sys.argv[3]),
              walls_y=(len(sys.argv) > 4 and sys.argv[4] == "walls"))
