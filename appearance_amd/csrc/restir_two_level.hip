// The ReSTIR kernels of restir.hip instantiated for two-level scenes
// (instanced models): every traversal in them maps rays into the instances'
// object space.  A translation unit of its own, so it compiles in parallel
// with restir.hip; it defines restir_sample_variant<true, false> only.
#define APT_RESTIR_INST 1
#define APT_RESTIR_TEX 0
#include "restir.hip"
