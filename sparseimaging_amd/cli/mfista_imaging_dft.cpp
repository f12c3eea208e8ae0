// mfista_imaging_dft -- command-line imaging with an explicit matrix A on the
// GPU solver.  Same arguments and report as c/mfista_imaging_dft.c:73-256:
// y is M doubles, A is M x N doubles stored column-major (row-major with -t),
// the image is sqrt(N) square or NX wide with -rec NX.  -looe is accepted and
// has no effect, as in the reference, whose LOOE call is disabled.
#include <cmath>

#include "cli_common.h"

static void usage(const char* s) {
  std::printf(
      "%s <int m> <int n> <V fname> <A fname> <double lambda_l1> <double lambda_tv> <double lambda_tsv> "
      "<double c> <X outfile> {X initfile} {-t} {-rec NX} {-nonneg} {-maxiter N} {-eps epsilon} "
      "{-cl_box box_fname} {-log log_fname} {-fgp N} {-rel_tol r}\n\n",
      s);
  std::printf(
      "  m, n: rows and columns of A; V: m doubles; A: m x n doubles, column-major (-t: row-major).\n"
      "  -rec NX: the image is NX x (n / NX) instead of sqrt(n) x sqrt(n).\n"
      "  Minimises |v - A x|^2/2 + lambda_l1 |x|_1 (+ lambda_tv TV(x) or lambda_tsv TSV(x)) with MFISTA.\n\n");
  std::exit(1);
}

int main(int argc, char** argv) {
  if (argc < 10) usage(argv[0]);
  cli::Common o;
  mfista_default_opts(&o.opts);
  int trans_flag = 0, rec_flag = 0, NX = 0;
  for (int i = 10; i < argc; i++) {
    if (!std::strcmp(argv[i], "-t")) {
      trans_flag = 1;
    } else if (!std::strcmp(argv[i], "-rec")) {
      if (i + 1 >= argc) usage(argv[0]);
      rec_flag = 1;
      NX = std::atoi(argv[++i]);
    } else if (!std::strcmp(argv[i], "-looe")) {
      continue;
    } else if (!cli::common_option(argc, argv, i, o)) {
      o.init_flag = 1;
      o.init_fname = argv[i];
    }
  }
  const int M = std::atoi(argv[1]), N = std::atoi(argv[2]);
  std::printf("M is %d\nN is %d\n", M, N);
  if (!rec_flag) NX = (int)std::sqrt((double)N);
  if (M <= 0 || N <= 0 || NX <= 0 || N % NX != 0) {
    std::printf("N = %d is not NX x NY with NX = %d.\n", N, NX);
    return 1;
  }
  const int NY = N / NX;
  std::printf("NX is %d\n", NX);

  std::vector<double> y((size_t)M), A((size_t)M * (size_t)N), xinit, x((size_t)N, 0.0);
  std::vector<float> box;
  if (!cli::read_init_and_box(o, N, xinit, box)) return 1;
  const int ny = read_V_vector(argv[3], M, y.data());
  if (ny < 0) return 1;
  if (ny != M) std::printf("Number of read data is shorter than expected in %s.\n", argv[3]);
  const unsigned long na = read_A_matrix(argv[4], M, N, A.data());
  if (na != (unsigned long)M * (unsigned long)N) {
    std::printf("Number of read data is shorter than expected in %s.\n", argv[4]);
    std::printf("expected num is %lu but it only had %lu.\n", (unsigned long)M * (unsigned long)N, na);
  }
  const double l1 = std::atof(argv[5]), ltv = std::atof(argv[6]), ltsv = std::atof(argv[7]),
               cinit = std::atof(argv[8]);
  cli::print_lambdas(l1, ltv, ltsv, cinit, o);
  if (trans_flag) transpose_matrix(A.data(), M, N);  // row-major M x N -> column-major

  RESULT res{};
  const int st = mfista_imaging_core_dft_ex(y.data(), A.data(), M, N, NX, NY, o.maxiter, o.eps, l1, ltv, ltsv,
                                            cinit, xinit.data(), x.data(), o.nonneg_flag, o.box_flag, box.data(),
                                            &res, &o.opts);
  mfista_amd_finalize();  // release the GPU state while the HIP runtime is up
  if (st != MFISTA_OK && st != MFISTA_E_NONNEG) {
    std::fprintf(stderr, "%s: solve failed with status %d\n", argv[0], st);
    return 1;
  }
  IO_FNAMES io{};
  io.fft = 0;
  io.v_fname = argv[3];
  io.A_fname = argv[4];
  const int r = cli::finish(argv[0], argv[9], x, io, res, o);
  std::free(res.residual);
  return r;
}
