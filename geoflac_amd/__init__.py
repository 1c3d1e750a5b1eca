"""MI355X-native GeoFLAC explicit time step (HIP kernels behind a C-ABI)."""
