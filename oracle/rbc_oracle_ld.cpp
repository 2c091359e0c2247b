// The oracle (rbc_oracle.cpp) evaluated in x87 extended precision (long double, 64-bit
// significand) instead of double: the same algorithm, operation for operation, with every
// intermediate carried to ~1e-19 relative.
//
// TEST INFRASTRUCTURE ONLY, like rbc_oracle.cpp.  It serves as the accuracy yardstick of
// the FP64 comparisons: the distance of the FP64 oracle and of the GPU step from this
// evaluation tells rounding (which no two FP64 evaluations share) apart from a deviation
// of the algorithm.  Inputs and outputs cross the C interface below as double and are
// converted exactly (double -> long double) on the way in.
//
// The eigen-solver's deflation threshold is the double epsilon of the source, so confined
// tensor set-ups converge to FP64 accuracy here; the periodic set-up has no eigen-solve.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#define double long double
#undef M_PI
#define M_PI 3.141592653589793238462643383279502884L
#include "rbc_oracle.cpp"
#undef double

namespace {
using ldbl = long double;
template <class T> void out_dbl(const Arr<T>& a, double* dst) {
  const ldbl* s = reinterpret_cast<const ldbl*>(a.d.data());
  const size_t n = a.d.size() * (sizeof(T) / sizeof(ldbl));
  for (size_t i = 0; i < n; ++i) dst[i] = (double)s[i];
}
template <class T> void in_dbl(Arr<T>& a, const double* src) {
  ldbl* d = reinterpret_cast<ldbl*>(a.d.data());
  const size_t n = a.d.size() * (sizeof(T) / sizeof(ldbl));
  for (size_t i = 0; i < n; ++i) d[i] = (ldbl)src[i];
}
}  // namespace

extern "C" {
const char* orcld_last_error() { return g_err.c_str(); }
void* orcld_navier_create(int nx, int ny, double ra, double pr, double dt, double aspect, int adiabatic,
                          int periodic) {
  return orc_navier_create(nx, ny, (ldbl)ra, (ldbl)pr, (ldbl)dt, (ldbl)aspect, adiabatic, periodic);
}
void orcld_navier_destroy(void* p) { orc_navier_destroy(p); }
// which: field index as orc_navier_get; space 0 physical, 1 spectral
int orcld_navier_get(void* p, int which, int space, double* out) {
  Handle* h = (Handle*)p;
  GUARD({
    if (h->periodic) {
      Field2<cd>* f = pick(*h->c, which);
      if (space) out_dbl(f->vhat, out); else out_dbl(f->v, out);
    } else {
      Field2<ldbl>* f = pick(*h->r, which);
      if (space) out_dbl(f->vhat, out); else out_dbl(f->v, out);
    }
  })
}
int orcld_navier_set(void* p, int which, int space, const double* in) {
  Handle* h = (Handle*)p;
  GUARD({
    if (h->periodic) {
      Field2<cd>* f = pick(*h->c, which);
      if (space) in_dbl(f->vhat, in); else in_dbl(f->v, in);
    } else {
      Field2<ldbl>* f = pick(*h->r, which);
      if (space) in_dbl(f->vhat, in); else in_dbl(f->v, in);
    }
  })
}
int orcld_navier_forward(void* p, int which) { return orc_navier_forward(p, which); }
int orcld_navier_set_velocity(void* p, double amp, double m, double n) {
  return orc_navier_set_velocity(p, (ldbl)amp, (ldbl)m, (ldbl)n);
}
int orcld_navier_update(void* p, int nsteps) { return orc_navier_update(p, nsteps); }
}
