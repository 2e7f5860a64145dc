"""MI355X (gfx950) support-tomography sweep: the per-orientation hot path of cfms_tomo's
TomoSh_INT3 engine as HIP kernels behind a C ABI (libtomo_hip.so).

Modules: _lib (ctypes binding), cfms_tomo (drop-in API), mesh (trimesh-free front end),
synth (seeded benchmark workloads), build (hipcc build of the library), cvss (continuous
support volume with a HIP backward, torch.autograd wrapper), refine (coarse sweep + gradient
refinement + voxel judging against the plain dense grid).
"""
