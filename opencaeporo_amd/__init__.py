"""MI355X-native EoS flash + FIM Jacobian assembly for OpenCAEPoro.

The compute path is the HIP library libocpgpu.so (C-ABI in include/ocpgpu.h);
``opencaeporo_amd.lib`` loads it and raises if it is not built.
"""
