// Density-matrix storage layout of an emitter group (host side, shared by
// the device layer and the manager's byte accounting).
#pragma once

#include "../../include/chiml/chiml_dev.h"

// True when the group keeps rho Hermitian (every member's H0 and the three
// dipole matrices Hermitian, and a Lindblad operator that maps Hermitian to
// Hermitian) and CHIML_ML_UNPACKED is unset: the density kernel then stores
// the upper triangle only, N (N + 1) / 2 complex words per matrix.
bool chiml_qe_packed(const chiml_qe_t& qi);
