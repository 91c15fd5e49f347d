"""MI355X-native backend for PolySpec's hot path (HEALPix SHTs, gNL/tauNL-loc
radial-node kernels, CG S^-1 filter).  Kernels live in libpolyspec_hip.so and
are called through the C ABI declared in include/polyspec_hip.h."""
