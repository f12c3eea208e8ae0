// cli_common.h -- option handling shared by the three imaging tools.
//
// Positional arguments and the options -nonneg, -maxiter, -eps, -cl_box,
// -log, -fftw_measure (FFT), -t/-rec/-looe (DFT) mean what they mean for the
// reference tools (c/mfista_imaging_{nufft,fft,dft}.c); any other argument
// is taken as the initial-image file, as there.  Extra options select what
// the reference fixes at build time: -preset c24|cxx12 (NUFFT gridding
// kernel; the C tool's is c24, the C++ tool's cxx12), -fgp N (FGP iterations
// of the TV prox, FGPITER = 100 by default) and -rel_tol r (relative stopping
// rule instead of the absolute eps rule).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mfista.h"
#include "../../include/mfista_io.h"

namespace cli {

struct Common {
  int nonneg_flag = 0, box_flag = 0, log_flag = 0, init_flag = 0, maxiter = MAXITER;
  double eps = EPS;
  std::string box_fname, log_fname, init_fname;
  mfista_opts opts{};
};

// Consumes argv[i] (and its value) if it is a common option; returns false
// for an argument that is not one.  Exits with status 1 on a missing value.
inline bool common_option(int argc, char** argv, int& i, Common& c) {
  auto value = [&](const char* opt) -> const char* {
    if (i + 1 >= argc) {
      std::printf("option %s needs a value.\n", opt);
      std::exit(1);
    }
    return argv[++i];
  };
  const char* a = argv[i];
  if (!std::strcmp(a, "-log")) {
    c.log_flag = 1;
    c.log_fname = value(a);
  } else if (!std::strcmp(a, "-maxiter")) {
    c.maxiter = std::atoi(value(a));
  } else if (!std::strcmp(a, "-eps")) {
    c.eps = std::atof(value(a));
  } else if (!std::strcmp(a, "-nonneg")) {
    c.nonneg_flag = 1;
  } else if (!std::strcmp(a, "-cl_box")) {
    c.box_flag = 1;
    c.box_fname = value(a);
  } else if (!std::strcmp(a, "-preset")) {
    const char* p = value(a);
    if (!std::strcmp(p, "c24"))
      c.opts.preset = MFISTA_PRESET_C24;
    else if (!std::strcmp(p, "cxx12"))
      c.opts.preset = MFISTA_PRESET_CXX12;
    else {
      std::printf("unknown preset %s (c24 or cxx12).\n", p);
      std::exit(1);
    }
  } else if (!std::strcmp(a, "-fgp")) {
    c.opts.fgp_iter = std::atoi(value(a));
  } else if (!std::strcmp(a, "-rel_tol")) {
    c.opts.rel_tol = std::atof(value(a));
  } else {
    return false;
  }
  return true;
}

inline void print_lambdas(double l1, double ltv, double ltsv, double c, const Common& o) {
  std::printf("lambda_l1 = %g\nlambda_tv = %g\nlambda_tsv = %g\nc = %g\n", l1, ltv, ltsv, c);
  if (o.nonneg_flag == 1) std::printf("x is nonnegative.\n");
  if (o.log_flag == 1) std::printf("Log will be saved to %s.\n", o.log_fname.c_str());
  std::printf("\n");
}

// Initial image (zeros without one) and CLEAN box, read as the reference
// tools read them; a short file is reported and used as far as it goes.
inline bool read_init_and_box(const Common& o, int NN, std::vector<double>& xinit, std::vector<float>& box) {
  xinit.assign((size_t)NN, 0.0);
  box.assign((size_t)NN, 0.f);
  if (o.init_flag) {
    std::printf("Initializing x with %s.\n", o.init_fname.c_str());
    const int n = read_V_vector((char*)o.init_fname.c_str(), NN, xinit.data());
    if (n < 0) return false;
    if (n != NN) std::printf("Number of read data is shorter than expected.\n");
  }
  if (o.box_flag) {
    std::printf("Restrict x with CLEAN box defined in %s.\n", o.box_fname.c_str());
    const int n = read_f_vector((char*)o.box_fname.c_str(), NN, box.data());
    if (n < 0) return false;
    if (n != NN) std::printf("Number of read data is shorter than expected.\n");
  }
  return true;
}

// Writes x, prints the report to stdout and, with -log, to the log file.
inline int finish(const char* prog, const char* xout_fname, std::vector<double>& x, IO_FNAMES& io,
                  RESULT& res, const Common& o) {
  if (write_X_vector((char*)xout_fname, (int)x.size(), x.data()) < 0) return 1;
  io.in_fname = o.init_flag ? (char*)o.init_fname.c_str() : nullptr;
  io.out_fname = (char*)xout_fname;
  show_io_fnames(stdout, (char*)prog, &io);
  show_result(stdout, (char*)prog, &res);
  if (o.log_flag) {
    FILE* fp = std::fopen(o.log_fname.c_str(), "w");
    if (!fp) {
      std::fprintf(stderr, " --- Can't fopen(w) %s\n", o.log_fname.c_str());
      return 1;
    }
    show_io_fnames(fp, (char*)prog, &io);
    show_result(fp, (char*)prog, &res);
    std::fclose(fp);
  }
  return 0;
}

}  // namespace cli
