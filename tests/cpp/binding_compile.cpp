// Compile-only check of the reference-side binding (nuray_amd/binding/rendererptnee_mi355x.h)
// against the reference's own headers and Qt 5: nothing here is linked or run.
#include "rendererptnee_mi355x.h"

Renderer* make_ptnee_gpu() { return new RendererPTNEE_MI355X(0); }
Renderer* make_pt_gpu() { return new RendererPTNEE_MI355X(1); }
