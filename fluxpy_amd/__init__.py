"""fluxpy_amd -- MI355X-native form-factor / visibility / compressed-solve path
of fluxpy, behind fluxpy's own module names:

    fluxpy_amd.shape                    (flux.shape)
    fluxpy_amd.form_factors             (flux.form_factors)
    fluxpy_amd.compressed_form_factors  (flux.compressed_form_factors)
    fluxpy_amd.quadtree / octree        (flux.quadtree / flux.octree)
    fluxpy_amd.linalg / util            (flux.linalg / flux.util)
    fluxpy_amd.solve / model            (flux.solve / flux.model)
    fluxpy_amd.thermal                  (flux.thermal)
    fluxpy_amd.config                   (flux.config)

Compute runs in libfluxhip.so (hand-written HIP for gfx950) through a C ABI
(include/fluxhip.h); PyTorch-ROCm provides device memory and streams.
"""
__version__ = "0.1.0"
