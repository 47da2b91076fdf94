"""MI355X-native D3Q19 lattice-Boltzmann wind tunnel.

Host code over libwt.so (hand-written HIP kernels for gfx950, C ABI in
include/wt.h). The drop-in for the reference solver module is `lbm` beside
this package.
"""
from ._lib import FORMS, FORM_NAMES, WT_FLAG_FORCES, WT_FLAG_MASS, load  # noqa: F401

__all__ = ['FORMS', 'FORM_NAMES', 'WT_FLAG_FORCES', 'WT_FLAG_MASS', 'load']
