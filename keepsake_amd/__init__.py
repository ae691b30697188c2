"""keepsake_amd: keepsake's small_pt path tracer rebuilt for AMD MI355X (gfx950).

The compute path is libkeepsake_amd.so (hand-written HIP kernels behind a C ABI, declared in
include/keepsake_amd.h); this package holds its ctypes bindings and the build script.
"""

from .api import (  # noqa: F401
    KS_GUIDE_OFF,
    KS_GUIDE_TRAIN_AND_SAMPLE,
    KS_MIS_CONSISTENT,
    KS_MIS_REFERENCE,
    Device,
    Film,
    Guide,
    HostScene,
    KsError,
    Scene,
    camera_desc,
    device_count,
    guide_config,
    load_library,
    procedural_scene,
    small_pt_params,
    small_pt_render,
    small_pt_render_batch,
    small_pt_render_guided,
    vmm_pdf,
    vmm_sample,
)
