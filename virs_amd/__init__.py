"""MI355X-native FDTD time-stepping path of VIRS (room acoustics).

virs_amd.grid   host grid pipeline (OBJ -> flags -> shell -> normals -> boundary list)
virs_amd.fdtd   the GPU time loop over libvirs_hip.so (C ABI in include/virs_fdtd.h)
virs_amd.rooms  synthetic benchmark rooms
"""
__all__ = ["grid", "fdtd", "rooms"]
