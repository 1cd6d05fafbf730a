"""varrcwa_amd: MI355X (gfx950) scattering-matrix solver for the varRCWA hot path.

The compute path lives in ``libvarrcwa.so`` (hand-written HIP kernels behind the
C ABI in ``include/varrcwa.h``); this package is its Python binding.
"""
from .api import (Context, DifferentialIntegrator, GDSIISampler, RCWASolver, gds_polygons, find_guided_modes, modal_power, Layer, RCWAIntegrator, RedhefferIntegrator, StackLayerSampler,
                  VarLayerSampler, VrcError, assemble_pq, assemble_toeplitz, batch_stack_rcwa, eig_pq,
                  layer_smatrix, project_left, redheffer_segment, stack_efficiencies, star_product,
                  wave_eqn_derivatives)

__all__ = ["Context", "GDSIISampler", "RCWASolver", "Layer", "VarLayerSampler", "StackLayerSampler", "RCWAIntegrator", "RedhefferIntegrator",
           "DifferentialIntegrator", "batch_stack_rcwa", "assemble_toeplitz", "assemble_pq", "eig_pq",
           "layer_smatrix", "star_product", "project_left", "redheffer_segment", "stack_efficiencies",
           "modal_power", "find_guided_modes", "wave_eqn_derivatives", "gds_polygons",
           "VrcError"]
