"""qcu_mrhs_amd -- MI355X (gfx950) multi-RHS SU(N) Wilson dslash and BiCGStab.

The product is the C-ABI library ``libqcu.so`` (hand-written HIP, built by
``qcu_mrhs_amd.build.build()``); :mod:`qcu_mrhs_amd.api` mirrors its entry
points in Python and :mod:`qcu_mrhs_amd.fields` generates seeded inputs.
"""
from . import fields  # noqa: F401

__all__ = ["fields", "api", "build"]
