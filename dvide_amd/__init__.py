"""dvide_amd: MI355X-native (gfx950) contact time step of DVIDE.

Collision detection, contact Jacobian assembly, the APGD cone-complementarity
solve and the velocity/position update run as hand-written HIP kernels in
libdvide_hip.so, reached through the C ABI declared in include/dvide.h.
"""
