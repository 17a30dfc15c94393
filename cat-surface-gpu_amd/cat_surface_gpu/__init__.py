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
    set_fmg_schedule,
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
from .surface_warp import (
    apply_flow_to_sphere,
    apply_flow_to_stenciled_sphere,
    solve_dartel_from_surfaces,
    solve_dartel_from_surfaces_jobs,
    solve_dartel_maps_jobs,
)
from .surface_stencil import SurfaceStencil, SurfaceStencilDevice
from .surface_rotation import (
    RotationCostInputs,
    RotationGridIndex,
    RotationGridIndexDevice,
    RotationSearchResult,
    official_seed_grid_angles,
)
from .rotation_pipeline import (
    RotationPipeline,
    read_rotation_points,
    read_rotation_values,
    run_rotation_pipeline,
    run_rotation_pipeline_from_files,
)
from .gpu_pipeline import (
    CatSurfaceGpuResult,
    GiftiMesh,
    PipelineJob,
    read_gifti_mesh,
    run_cat_surface_gpu_pipeline,
    run_cat_surface_gpu_pipelines,
    run_cat_surface_gpu_stream,
    write_gifti_mesh,
)
from .dartel_adjoint import ExpWarpSSD, dartel_rhs, exp_warp_ssd
