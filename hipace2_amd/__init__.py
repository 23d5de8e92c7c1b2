"""hipace2_amd: MI355X-native per-slice quasi-static loop of HiPACE++.

The compute path is libhpq.so (HIP kernels for gfx950 + C ABI, include/hpq.h);
`Simulation` drives it from Python with device buffers allocated through torch.
"""
from .inputs import BeamArrays, BeamSpec, SliceConfig, make_beam, merge_beams, parse_deck  # noqa: F401


def __getattr__(name):
    if name == "Simulation":
        from .simulation import Simulation
        return Simulation
    raise AttributeError(name)
