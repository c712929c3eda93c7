// surface.h -- free-surface mirror (Propagator::apply_symmetry).
#pragma once
#include "klrtm_internal.h"

namespace klrtm {
// arguments the mirror accepts (tprlen 0 when untapered)
int apply_symmetry_check(int nx, int ny, int nz, int iz0, int tprlen);
int apply_symmetry(float *wf, int nx, int ny, int nz, int iz0, int sym, int tprlen, int taper);
}  // namespace klrtm
