"""Build recipe of the HBM streaming probe (stream_probe.hip -> measure/stream_probe, gfx950).

    python -m tcsem_fdtd_amd.measure.build

The probe is measurement tooling, built on request and not by the library build.
"""
import os
import sys

from tcsem_fdtd_amd import build as B

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE = os.path.join(HERE, "stream_probe")


def build():
    B._run([B.hipcc()] + B.HIP_FLAGS + ["-x", "hip", os.path.join(HERE, "stream_probe.hip"), "-o", PROBE])
    return PROBE


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    build()
