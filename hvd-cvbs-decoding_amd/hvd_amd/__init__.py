"""MI355X (gfx950) engine for hvd-core's Y/C-separation hot path.

`hvd_amd.gpu` binds the C ABI of libhvd_gpu.so (include/hvd_gpu.h);
`hvd_amd.synth` generates the seeded NTSC / PAL test fields.
"""
from . import synth  # noqa: F401

__all__ = ["synth", "gpu"]
