// Host part of the eigen-step (tridiag.cpp); plain C++, no HIP.
#pragma once

namespace spamd {

// R largest eigenpairs of the symmetric tridiagonal T with diagonal d[0..P)
// and off-diagonal e[0..P-1): w descending, Z holds R unit vectors of length
// P, vector j at Z + j*P.  Throws std::invalid_argument unless 0 < R <= P.
void tridiag_top(const double *d, const double *e, int P, int R, double *w, double *Z);

}  // namespace spamd
