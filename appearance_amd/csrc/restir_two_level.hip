// The ReSTIR kernels of restir.hip instantiated for two-level scenes
// (instanced models): every traversal in them maps rays into the instances'
// object space.  A translation unit of its own, so it compiles in parallel
// with restir.hip; it defines restir_sample_two_level only.
#define APT_RESTIR_TWO_LEVEL 1
#include "restir.hip"
