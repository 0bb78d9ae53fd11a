// The ReSTIR kernels of restir.hip instantiated for textured two-level scenes
// (spawned models with textures).  A translation unit of its own, so it
// compiles in parallel with restir.hip; it defines
// restir_sample_variant<true, true> only.
#define APT_RESTIR_INST 1
#define APT_RESTIR_TEX 1
#include "restir.hip"
