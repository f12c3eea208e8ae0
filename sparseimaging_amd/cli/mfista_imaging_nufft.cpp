// mfista_imaging_nufft -- command-line NUFFT imaging on the GPU solver.
// Same arguments, data file and report as c/mfista_imaging_nufft.c:65-242
// (and c++/mfista_imaging_nufft.cpp), plus -preset / -fgp / -rel_tol.
#include "cli_common.h"

static void usage(const char* s) {
  std::printf(
      "%s <nufft_data fname> <double lambda_l1> <double lambda_tv> <double lambda_tsv> <double c> <X outfile> "
      "{X initfile} {-nonneg} {-cl_box box_fname} {-maxiter N} {-eps epsilon} {-log log_fname} "
      "{-preset c24|cxx12} {-fgp N} {-rel_tol r}\n\n",
      s);
  std::printf(
      "  nufft_data file: header \"M  = ..\", \"NX = ..\", \"NY = ..\", a title line, then rows\n"
      "                   u_rad, v_rad, vis_r, vis_i, noise_std_dev\n"
      "  X files are NX*NY doubles (x[i*NY+j]); the CLEAN box is NX*NY floats.\n\n"
      "  Minimises |W(v - A x)|^2/2 + lambda_l1 |x|_1 (+ lambda_tv TV(x) or lambda_tsv TSV(x))\n"
      "  with MFISTA; -nonneg restricts x to x >= 0.  c is the initial step-size constant.\n"
      "  -preset: gridding kernel, c24 (24 taps, the C library; default) or cxx12 (12 taps, the C++ library)\n"
      "  -fgp: FGP iterations of the TV prox (default 100); -rel_tol: stop when the cost falls by less than\n"
      "        r * cost over 50 iterations (default: the absolute rule with eps).\n\n");
  std::exit(1);
}

int main(int argc, char** argv) {
  if (argc < 7) usage(argv[0]);
  cli::Common o;
  mfista_default_opts(&o.opts);
  const double l1 = std::atof(argv[2]), ltv = std::atof(argv[3]), ltsv = std::atof(argv[4]),
               cinit = std::atof(argv[5]);
  for (int i = 7; i < argc; i++) {
    if (!cli::common_option(argc, argv, i, o)) {
      o.init_flag = 1;
      o.init_fname = argv[i];
    }
  }
  cli::print_lambdas(l1, ltv, ltsv, cinit, o);

  int M = 0, NX = 0, NY = 0;
  double *u, *v, *vr, *vi, *sd;
  const int rc = mfista_read_nufft_data(argv[1], &M, &NX, &NY, &u, &v, &vr, &vi, &sd);
  if (rc == MFISTA_IO_E_OPEN) return 1;
  if (rc == MFISTA_IO_E_ROW) std::printf("cannot read data.\n");
  if (rc != MFISTA_IO_OK) return 0;  // the reference tool exits with 0 on a malformed file
  std::printf("number of u-v points:  %d\nX-dim of image:        %d\nY-dim of image:        %d\n", M, NX, NY);

  const int NN = NX * NY;
  std::vector<double> xinit, x((size_t)NN, 0.0);
  std::vector<float> box;
  if (!cli::read_init_and_box(o, NN, xinit, box)) return 1;

  RESULT res{};
  const int st = mfista_imaging_core_nufft_ex(u, v, vr, vi, sd, M, NX, NY, o.maxiter, o.eps, l1, ltv, ltsv, cinit,
                                              xinit.data(), x.data(), o.nonneg_flag, o.box_flag, box.data(), &res,
                                              &o.opts);
  mfista_amd_finalize();  // release the GPU state while the HIP runtime is up
  mfista_io_free(u);
  mfista_io_free(v);
  mfista_io_free(vr);
  mfista_io_free(vi);
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
