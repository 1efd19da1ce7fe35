"""Import path of the engine's host mirror (the package directory name is not
a Python identifier, so it is put on sys.path)."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, 'ngspice-sf-mirror_amd')
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import ngmi  # noqa: E402,F401
import workloads  # noqa: E402,F401
