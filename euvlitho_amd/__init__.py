"""MI355X-native EUV lithography mask -> EM -> M3D -> image path (gfx950 HIP via libeuvl.so)."""
