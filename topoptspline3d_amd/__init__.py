"""MI355X-native (gfx950) topology-optimisation hot path: matrix-free geometric
multigrid for SIMP elasticity on voxel grids, spline densities, sensitivity filter and
OC/MMA updates, behind the C ABI of include/tos_api.h."""
