"""dachshund_amd: gfx950 (MI355X) Wiener-filter reconstruction.

The product is libdachshund_gpu.so (HIP kernels behind the C ABI of
include/dachshund_gpu.h) and the C++ drop-in of the reference's library API
built on it; this package is the Python mirror of the same interface.
"""
from .wiener import (DS_F32_MFMA, DS_F32_MFMA_CLAMPED, DS_F32_VALU,
                     DS_PREC_F32, DS_PREC_F64,
                     DS_PRECOND_BLOCK, DS_PRECOND_JACOBI,
                     BatchWF, Context, PCGParams,
                     PCGResult,
                     SignalCovarParams, WFVectorSN, WFVectorW,
                     default_context, init_uniform_map_coords, map_covar_diag,
                     npixel_to_pixel, pixel_to_npixel, pixel_to_wpixel,
                     plan_tiles, points_from_pixels, read_pixel_data,
                     sigma_m_noise_dom, smp_product, smp_product_grid,
                     wpixel_to_pixel)

__all__ = [
    "DS_F32_MFMA", "DS_F32_MFMA_CLAMPED", "DS_F32_VALU", "DS_PRECOND_BLOCK", "DS_PRECOND_JACOBI", "DS_PREC_F32", "DS_PREC_F64", "BatchWF", "Context", "PCGParams", "PCGResult",
    "SignalCovarParams", "WFVectorSN", "WFVectorW", "default_context",
    "init_uniform_map_coords", "map_covar_diag", "npixel_to_pixel", "pixel_to_npixel",
    "pixel_to_wpixel", "plan_tiles", "points_from_pixels", "read_pixel_data",
    "sigma_m_noise_dom", "smp_product", "smp_product_grid", "wpixel_to_pixel",
]
