"""tigergrav_amd -- MI355X (gfx950) gravity-interaction kernels for the tigergrav tree code.

The eight reference kernels (PP/PC/CC/CP, direct and Ewald) behind a C ABI
(include/tg_gravity.h), Python wrappers under the reference's names (tigergrav_amd.gravity), the
host-side tree / interaction-list builder and the batched device-resident force solve.
"""
from ._lib import FORCE_DTYPE, MULTI_DTYPE, TGError, lib  # noqa: F401

__all__ = ["FORCE_DTYPE", "MULTI_DTYPE", "TGError", "lib"]
