"""Cell-type constants and small helpers (utils.py:1-62 of the reference).

``sample`` (utils.py:12-32) is used only inside operator-assembly kernels;
here it lives in the HIP strategies (csrc/mgpcg_strategies.hip).  ``splat``
(utils.py:34-57) belongs to the APIC transfer, which is not part of the
solver path.
"""
FLUID = 0
AIR = 1
SOLID = 2


def copy(src, dst):
    """utils.py:59-62: dst[I] = src[I] (torch tensors of the same shape)."""
    dst.copy_(src)
