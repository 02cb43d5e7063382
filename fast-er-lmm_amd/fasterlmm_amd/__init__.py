"""MI355X-native LMM association scan with the drop-in API of FaST-ER-LMM (fasterlmm.core / perms / io).

    import fasterlmm_amd.core as core      # loco_scan_compat, snp_wald_scan_compat, ...
    from fasterlmm_amd.perms import perm_threshold

The compute path is libfasterlmm_amd.so (hand-written gfx950 kernels, C ABI in include/fasterlmm_amd.h).
"""
__version__ = "0.1.0"
REFERENCE_VERSION = "1.6.2"
