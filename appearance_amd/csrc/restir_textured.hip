// The ReSTIR kernels of restir.hip instantiated for textured scenes (uvs,
// tangent frames and material textures).  A translation unit of its own, so
// it compiles in parallel with restir.hip; it defines
// restir_sample_variant<false, true> only.
#define APT_RESTIR_INST 0
#define APT_RESTIR_TEX 1
#include "restir.hip"
