"""MI355X-native MPM snow-simulation step (libmpm_hip.so) and its host-side mirror of the
reference's simulation managers (luoluobuli/WebGPU-MPM-based-Snow-Simulation,
src/lib/gpu/mpm/, src/lib/gpu/uniforms/, src/lib/gpu/collider/)."""
from ._lib import (load_library, default_config, MpmError, MpmConfig, EXPORTED, LIB_PATH,
                   METHOD_EXPLICIT, METHOD_PBMPM, UNIFORMS_BYTES, PARTICLE_BYTES)
from .particles import PARTICLE_DTYPE, make_particles, state_columns, matrix
from .managers import (GpuUniformsBufferManager, GpuMpmBufferManager, GpuColliderBufferManager, GpuMeshBufferManager, triangle_soup, load_glb,
                       GpuMpmPipelineManager, MpmSimulation, MpmEnsemble, build_bvh, mlp_grad)
from . import scenes
