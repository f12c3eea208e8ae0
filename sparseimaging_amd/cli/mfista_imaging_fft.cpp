// mfista_imaging_fft -- command-line FFT-mode imaging on the GPU solver.
// Same arguments, data file and report as c/mfista_imaging_fft.c:65-249
// (and c++/mfista_imaging_fft.cpp); -fftw_measure is accepted and has no
// effect (there is no planner to tune).
#include "cli_common.h"

static void usage(const char* s) {
  std::printf(
      "%s <fft_data fname> <double lambda_l1> <double lambda_tv> <double lambda_tsv> <double c> <X outfile> "
      "{X initfile} {-nonneg} {-cl_box box_fname} {-maxiter N} {-eps epsilon} {-fftw_measure} {-log log_fname} "
      "{-fgp N} {-rel_tol r}\n\n",
      s);
  std::printf(
      "  fft_data file: header \"M  = ..\", \"NX = ..\", \"NY = ..\", a title line, then rows\n"
      "                 u, v, y_r, y_i, noise_std_dev   (u, v: integer cells of the NX x NY FFT)\n"
      "  X files are NX*NY doubles (x[i*NY+j]); the CLEAN box is NX*NY floats.\n\n"
      "  Minimises |W(y - F x)|^2/2 + lambda_l1 |x|_1 (+ lambda_tv TV(x) or lambda_tsv TSV(x))\n"
      "  with MFISTA; -nonneg restricts x to x >= 0.  c is the initial step-size constant.\n\n");
  std::exit(1);
}

int main(int argc, char** argv) {
  if (argc < 7) usage(argv[0]);
  cli::Common o;
  mfista_default_opts(&o.opts);
  const double l1 = std::atof(argv[2]), ltv = std::atof(argv[3]), ltsv = std::atof(argv[4]),
               cinit = std::atof(argv[5]);
  for (int i = 7; i < argc; i++) {
    if (!std::strcmp(argv[i], "-fftw_measure")) continue;
    if (!cli::common_option(argc, argv, i, o)) {
      o.init_flag = 1;
      o.init_fname = argv[i];
    }
  }
  cli::print_lambdas(l1, ltv, ltsv, cinit, o);

  int M = 0, NX = 0, NY = 0;
  int *iu, *iv;
  double *yr, *yi, *sd;
  const int rc = mfista_read_fft_data(argv[1], &M, &NX, &NY, &iu, &iv, &yr, &yi, &sd);
  if (rc == MFISTA_IO_E_OPEN) return 1;
  if (rc == MFISTA_IO_E_ROW) std::printf("cannot read data.\n");
  if (rc != MFISTA_IO_OK) return 0;
  std::printf("number of u-v points:  %d\nX-dim of image:        %d\nY-dim of image:        %d\n", M, NX, NY);

  const int NN = NX * NY;
  std::vector<double> xinit, x((size_t)NN, 0.0);
  std::vector<float> box;
  if (!cli::read_init_and_box(o, NN, xinit, box)) return 1;

  RESULT res{};
  const int st = mfista_imaging_core_fft_ex(iu, iv, yr, yi, sd, M, NX, NY, o.maxiter, o.eps, l1, ltv, ltsv, cinit,
                                            xinit.data(), x.data(), o.nonneg_flag, o.box_flag, box.data(), &res,
                                            &o.opts);
  mfista_amd_finalize();  // release the GPU state while the HIP runtime is up
  mfista_io_free(iu);
  mfista_io_free(iv);
  mfista_io_free(yr);
  mfista_io_free(yi);
  mfista_io_free(sd);
  if (st != MFISTA_OK && st != MFISTA_E_NONNEG) {
    std::fprintf(stderr, "%s: solve failed with status %d\n", argv[0], st);
    return 1;
  }
  IO_FNAMES io{};
  io.fft = 1;
  io.fft_fname = argv[1];
  return cli::finish(argv[0], argv[6], x, io, res, o);
}
