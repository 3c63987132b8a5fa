"""MI355X-native DFSPH time step (gfx950 HIP kernels behind a C ABI).

Python here is plumbing: the solver is libdfsph_hip.so (include/dfsph_hip.h)
and the C++ host classes of libdfsph_host.so (include/dfsph_host.h).
"""
from . import scene  # noqa: F401
from ._native import (  # noqa: F401
    DFSPHError, DFSPHParams, HostSolver, Solver, default_params, load_hip, load_host,
    FIELD_POS, FIELD_VEL, FIELD_MASS, FIELD_DENSITY, FIELD_FACTOR, FIELD_DENSITY_ADV, FIELD_KAPPA,
    FIELD_KAPPAV, FIELD_ACC, FIELD_ID, FIELD_NBR_FLUID, FIELD_NBR_BOUNDARY, FIELD_BOUNDARY_POS,
    FIELD_BOUNDARY_PSI, WARMSTART_LEARNED, WARMSTART_NONE, WARMSTART_REFERENCE,
    LISTS_AUTO, LISTS_NO_LDS, LISTS_GLOBAL, LEARNED_ALWAYS, LEARNED_ABOVE_FLOORS, FIELD_TAG, FIELD_KAPPA_TOTAL, FIELD_KAPPAV_TOTAL, DFSPHSurface, surface,
    TAG_ACTIVE, TAG_ACTIVE_NEIGHBORS, TAG_UNTAGGED, TAG_REMOVAL, TAG_AIR,
    Scene, read_particles_text, write_particles_text,
    DFSPHRigidBody, rigid_body, BODY_DYNAMIC, BODY_KINEMATIC, BODY_SINE,
)
