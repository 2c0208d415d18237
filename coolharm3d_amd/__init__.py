"""MI355X-native per-substep GRMHD advance for coolharm3d.

The product is ``libhgpu.so`` (HIP kernels for gfx950 behind the C ABI of
``include/hgpu.h``); :class:`Harm3D` is the Python mirror of the reference's
``step_ch()`` interface over it.
"""
from .problems import GridSpec, balsara1, balsara1_grid, torus, torus_grid, bbh_grid, bbh_disk  # noqa: F401
from .harm import Harm3D, config_for, risco_prograde  # noqa: F401
