"""FoilCFD MI355X solver core.

Hand-written HIP kernels for gfx950 behind a C ABI (include/foilcfd.h); this
package is the Python-side binding used by the tests and the benchmark.
"""
from ._abi import LIB_PATH, load  # noqa: F401
from .solver import (  # noqa: F401
    CX, CY, CZ, FLUID, INLET, INTERFACE, OUTLET, Q, SLIP_BACK, SLIP_BOTTOM, SLIP_FRONT, SLIP_TOP,
    SOLID, W, LBMSolver, build_boundary_flags, build_clean_foil_flags, build_fine_patch_flags,
    close_trailing_edge_gaps, compute_scaling, derive_patch_box, flow_through_steps, force_to_coefficient,
    generate_naca4, make_fine_layout, make_layout, ramped_tau, ramped_u, refined_scaling,
    scaling_for_tau, stamp_interface_shell, voxelize_airfoil)
