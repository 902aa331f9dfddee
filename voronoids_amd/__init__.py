"""voronoids_amd: MI355X-native parallel Delaunay construction (2D/3D)."""
