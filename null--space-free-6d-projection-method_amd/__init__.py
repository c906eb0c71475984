"""MI355X-native hot path of the null-space-free 6D Fourier projection eigensolver.

The compute path is libfame6d.so (hand-written HIP kernels for gfx950 behind the C ABI
of include/fame6d.h); this package is its host-side mirror of the reference's
operator interface (the `mass` struct of build_mass_backend_large_scalar, the
q-system, K_r^{-1}, Lanczos, recovery, Taylor reconstruction and the Yee/LWRQ closure).  There is no
CPU fallback: every operation raises when the library or the GPU is missing.
"""
from ._lib import FP32, FP64, MIXED, Fame6dError, load  # noqa: F401
from .mass import (MassBackend, P_DEFAULT, build_mass_backend_large_scalar, cosine_material,  # noqa: F401
                   smooth_scalar_material, start_vector)
from .eigs import (QSystem, build_experiment1_q_system_C_compact, compute_fourier_diagnostics,  # noqa: F401
                   krinv_nested, make_explicit_krinv_compact, method_a_top, recover_original_eigenvectors,
                   run_method_A, run_method_B, run_method_C, run_method_C_batch,
                   solve_tracked_modes, solve_tracking_scan)
from .taylor import (auto_taylor_blocks, block_centers, block_partition, build_multi_index_list_3d,  # noqa: F401
                     compute_field_errors, direct_reconstruct, projected_q_modes, reconstruct_yee_field, taylor_evaluate,
                     taylor_moments, yee_block_centers, yee_points)
from .yee import (apply_cropped_A, compute_lwrq, compute_yee_lwrq_metrics, crop_plus_field,  # noqa: F401
                  sample_scalar_yee_material)
from . import results  # noqa: F401,E402
