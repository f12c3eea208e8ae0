// modes_stub.hip -- entry points not yet ported to the GPU in this build.
#include <cstdio>

#include "solver.h"

namespace mf {
int solve_nufft_tv(Context&, const SolveArgs&, RESULT*, mfista_opts*) {
  fprintf(stderr, "[mfista_amd] L1+TV (FGP) is not available in this build\n");
  return MFISTA_E_ARG;
}
}  // namespace mf

extern "C" {
void mfista_imaging_core_fft(int*, int*, double*, double*, double*, int, int, int, int, double,
                             double, double, double, double, double*, double*, int, unsigned int,
                             int, float*, struct RESULT*) {
  fprintf(stderr, "[mfista_amd] FFT mode is not available in this build\n");
}
void mfista_imaging_core_dft(double*, double*, int*, int*, int, int, int, double, double, double,
                             double, double, double*, double*, int, int, int, float*,
                             struct RESULT*) {
  fprintf(stderr, "[mfista_amd] DFT mode is not available in this build\n");
}
}
