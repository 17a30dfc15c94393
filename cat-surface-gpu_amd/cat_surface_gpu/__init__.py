"""MI355X-native CAT_SurfWarp registration hot path (HIP, gfx950).

Drop-in for the registration part of ``cat_surface_gpu`` (ayakacxy/cat-surface-gpu,
src/cat_surface_gpu/__init__.py:5-127) minus the Surf2Sphere names, which
belong to a different program.  Every device computation runs in
libcsg_hip.so; nothing falls back to the CPU.
"""

from .dartel_grid import (
    DeviceUnavailable,
    GridSpec,
    apply_membrane_system,
    composition,
    composition_jacobian,
    dartel_step,
    expdef,
    fmg2_membrane,
    from_c_layout,
    initialise_objfun,
    jac_div_smalldef,
    jacobian_of_displacement,
    make_identity_map,
    regularization_operator,
    relax_membrane,
    resize_field,
    resolve_device,
    sample_field,
    smalldef_jac,
    squaring_update,
    to_c_layout,
)
from .surface_warp import (
    DartelSolveResult,
    clear_dartel_cuda_graph_cache,
    default_dartel_parameters,
    prepare_dartel_cuda_graph,
    solve_dartel_maps,
)

__version__ = "0.1.0"
