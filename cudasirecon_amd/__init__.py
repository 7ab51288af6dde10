"""MI355X-native 3D-SIM reconstruction engine with cudasirecon's interfaces.

Layout:
  csrc/        HIP kernels (gfx950), C++ host orchestration, extern "C" ABI
  engine.py    ctypes binding of the stage-level ABI (include/simr.h)
  _abi.py      ctypes mirrors of the ABI structs
  mrc.py       MRC/DV codec
"""
from ._abi import SimrFit, SimrParams, norders_of, output_shape, set_params  # noqa: F401

__version__ = "1.2.0"


def build(jobs=8):
    """Compile the HIP library and tools in-tree (hipcc, --offload-arch=gfx950)."""
    import os
    import subprocess

    here = os.path.dirname(os.path.abspath(__file__))
    subprocess.run(["make", "-s", "-j%d" % jobs, "-C", os.path.join(here, "csrc")], check=True)
