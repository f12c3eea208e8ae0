// mfista_io.cpp -- host-side file formats and the text report of the imaging
// command-line tools (declared in include/mfista_io.h).  No GPU code.
//
// Binary vectors are raw native-endian arrays (fread/fwrite of int32,
// float32 or float64), as c/mfista_io.c reads and writes them.  The text data
// files have a three-line header, a column-title line and one comma-separated
// row per visibility (c/mfista_imaging_nufft.c:140-170,
// c/mfista_imaging_fft.c:153-183).  The report labels are those of
// c/mfista_io.c:141-215, kept verbatim so that scripts that scrape the
// reference's logs read these too.
#include "../../include/mfista_io.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

FILE* open_or_report(const char* fn, const char* mode) {
  FILE* fp = fn ? std::fopen(fn, mode) : nullptr;
  if (!fp) std::fprintf(stderr, " --- Can't fopen(%s) %s\n", mode[0] == 'r' ? "r" : "w", fn ? fn : "(null)");
  return fp;
}

template <typename T>
int read_raw(const char* fname, int length, T* v) {
  if (length < 0 || (length > 0 && !v)) return -1;
  FILE* fp = open_or_report(fname, "rb");
  if (!fp) return -1;
  const size_t n = std::fread(v, sizeof(T), (size_t)length, fp);
  std::fclose(fp);
  return (int)n;
}

// Header "M  = %d", "NX = %d", "NY = %d" (whitespace-insensitive, as fscanf's
// directives are), then an optional column-title line.
bool read_header(FILE* fp, int* M, int* NX, int* NY) {
  if (std::fscanf(fp, " M = %d", M) != 1) return false;
  if (std::fscanf(fp, " NX = %d", NX) != 1) return false;
  if (std::fscanf(fp, " NY = %d", NY) != 1) return false;
  int c;
  do {
    c = std::fgetc(fp);
  } while (c != EOF && std::isspace(c));
  if (c == EOF) return *M == 0;
  if (std::isalpha(c)) {  // title line: skip to its end
    while (c != EOF && c != '\n') c = std::fgetc(fp);
  } else {
    std::ungetc(c, fp);
  }
  return *M >= 0 && *NX > 0 && *NY > 0;
}

}  // namespace

extern "C" {

int read_int_vector(char* fname, int length, int* vector) { return read_raw(fname, length, vector); }

int read_f_vector(char* fname, int length, float* vector) { return read_raw(fname, length, vector); }

int read_V_vector(char* fname, int length, double* vector) { return read_raw(fname, length, vector); }

unsigned long read_A_matrix(char* fname, int height, int width, double* matrix) {
  if (height < 0 || width < 0 || !matrix) return 0;
  const size_t n = (size_t)height * (size_t)width;
  std::printf("Reading %d x %d = %zu entries of matrix A.\n", height, width, n);
  FILE* fp = open_or_report(fname, "rb");
  if (!fp) return 0;
  const size_t got = std::fread(matrix, sizeof(double), n, fp);
  std::fclose(fp);
  return (unsigned long)got;
}

int write_X_vector(char* fname, int length, double* vector) {
  if (length < 0 || (length > 0 && !vector)) return -1;
  FILE* fp = open_or_report(fname, "wb");
  if (!fp) return -1;
  const size_t n = std::fwrite(vector, sizeof(double), (size_t)length, fp);
  std::fclose(fp);
  return (int)n;
}

void transpose_matrix(double* matrix, int origheight, int origwidth) {
  if (!matrix || origheight <= 0 || origwidth <= 0) return;
  const size_t H = (size_t)origheight, W = (size_t)origwidth;
  std::vector<double> t(H * W);
  for (size_t i = 0; i < H; i++)
    for (size_t j = 0; j < W; j++) t[j * H + i] = matrix[i * W + j];
  std::memcpy(matrix, t.data(), H * W * sizeof(double));
}

void show_io_fnames(FILE* fid, char* fname, struct IO_FNAMES* io) {
  if (!fid || !io) return;
  std::fprintf(fid, "\n\nIO files of %s.\n\n\n", fname ? fname : "");
  if (io->fft == 0) {
    std::fprintf(fid, " input vector file:      %s\n", io->v_fname ? io->v_fname : "(null)");
    std::fprintf(fid, " input matrix file:      %s\n", io->A_fname ? io->A_fname : "(null)");
  } else {
    std::fprintf(fid, " FFTW file:              %s\n", io->fft_fname ? io->fft_fname : "(null)");
  }
  if (io->in_fname) std::fprintf(fid, " x was initialized with: %s\n", io->in_fname);
  if (io->out_fname) std::fprintf(fid, " x is saved to:          %s\n", io->out_fname);
  std::fprintf(fid, "\n");
}

void show_result(FILE* fid, char* fname, struct RESULT* r) {
  if (!fid || !r) return;
  std::fprintf(fid, "\n\nOutput of %s.\n\n\n", fname ? fname : "");
  std::fprintf(fid, " Size of the problem:\n\n");
  std::fprintf(fid, " size of input vector:   %d\n", r->M);
  std::fprintf(fid, " size of output vector:  %d\n", r->N);
  if (r->NX != 0) std::fprintf(fid, " size of image:          %d x %d\n", r->NX, r->NY);
  std::fprintf(fid, "\n\n Problem Setting:\n\n");
  if (r->nonneg == 1)
    std::fprintf(fid, " x is a nonnegative vector.\n\n");
  else if (r->nonneg == 0)
    std::fprintf(fid, " x is a real vector (takes 0, positive, and negative value).\n\n");
  if (r->lambda_l1 != 0) std::fprintf(fid, " Lambda_1:               %e\n", r->lambda_l1);
  if (r->lambda_tsv != 0) std::fprintf(fid, " Lambda_TSV:             %e\n", r->lambda_tsv);
  if (r->lambda_tv != 0) std::fprintf(fid, " Lambda_TV:              %e\n", r->lambda_tv);
  std::fprintf(fid, " MAXITER:                %d\n", r->maxiter);
  std::fprintf(fid, "\n\n Results:\n\n");
  std::fprintf(fid, " # of iterations:        %d\n", r->ITER);
  std::fprintf(fid, " cost:                   %e\n", r->finalcost);
  std::fprintf(fid, " computaion time[sec]:   %e\n", r->comp_time);
  std::fprintf(fid, " Est. Lipschitzs const:  %e\n\n", r->Lip_const);
  std::fprintf(fid, " # of nonzero pixels:    %d\n", r->N_active);
  std::fprintf(fid, " Squared Error (SE):     %e\n", r->sq_error);
  std::fprintf(fid, " Mean SE:                %e\n", r->mean_sq_error);
  if (r->lambda_l1 != 0) std::fprintf(fid, " L1 cost:                %e\n", r->l1cost);
  if (r->lambda_tsv != 0) std::fprintf(fid, " TSV cost:               %e\n", r->tsvcost);
  if (r->lambda_tv != 0) std::fprintf(fid, " TV cost:                %e\n", r->tvcost);
  std::fprintf(fid, "\n");
}

int mfista_read_nufft_data(const char* fname, int* M, int* NX, int* NY, double** u_dx, double** v_dy,
                           double** vis_r, double** vis_i, double** vis_std) {
  if (!M || !NX || !NY || !u_dx || !v_dy || !vis_r || !vis_i || !vis_std) return MFISTA_IO_E_ARG;
  FILE* fp = open_or_report(fname, "r");
  if (!fp) return MFISTA_IO_E_OPEN;
  int m = 0, nx = 0, ny = 0;
  if (!read_header(fp, &m, &nx, &ny)) {
    std::fclose(fp);
    return MFISTA_IO_E_HEADER;
  }
  const size_t n = (size_t)(m > 0 ? m : 1);
  double* a[5];
  for (int k = 0; k < 5; k++) a[k] = (double*)std::malloc(n * sizeof(double));
  if (!a[0] || !a[1] || !a[2] || !a[3] || !a[4]) {
    for (int k = 0; k < 5; k++) std::free(a[k]);
    std::fclose(fp);
    return MFISTA_IO_E_NOMEM;
  }
  for (int i = 0; i < m; i++) {
    if (std::fscanf(fp, " %lf , %lf , %lf , %lf , %lf", a[0] + i, a[1] + i, a[2] + i, a[3] + i, a[4] + i) != 5) {
      for (int k = 0; k < 5; k++) std::free(a[k]);
      std::fclose(fp);
      return MFISTA_IO_E_ROW;
    }
  }
  std::fclose(fp);
  *M = m;
  *NX = nx;
  *NY = ny;
  *u_dx = a[0];
  *v_dy = a[1];
  *vis_r = a[2];
  *vis_i = a[3];
  *vis_std = a[4];
  return MFISTA_IO_OK;
}

int mfista_read_fft_data(const char* fname, int* M, int* NX, int* NY, int** u_idx, int** v_idx, double** y_r,
                         double** y_i, double** noise_stdev) {
  if (!M || !NX || !NY || !u_idx || !v_idx || !y_r || !y_i || !noise_stdev) return MFISTA_IO_E_ARG;
  FILE* fp = open_or_report(fname, "r");
  if (!fp) return MFISTA_IO_E_OPEN;
  int m = 0, nx = 0, ny = 0;
  if (!read_header(fp, &m, &nx, &ny)) {
    std::fclose(fp);
    return MFISTA_IO_E_HEADER;
  }
  const size_t n = (size_t)(m > 0 ? m : 1);
  int* iu = (int*)std::malloc(n * sizeof(int));
  int* iv = (int*)std::malloc(n * sizeof(int));
  double* d[3];
  for (int k = 0; k < 3; k++) d[k] = (double*)std::malloc(n * sizeof(double));
  auto release = [&]() {
    std::free(iu);
    std::free(iv);
    for (int k = 0; k < 3; k++) std::free(d[k]);
  };
  if (!iu || !iv || !d[0] || !d[1] || !d[2]) {
    release();
    std::fclose(fp);
    return MFISTA_IO_E_NOMEM;
  }
  for (int i = 0; i < m; i++) {
    if (std::fscanf(fp, " %d , %d , %lf , %lf , %lf", iu + i, iv + i, d[0] + i, d[1] + i, d[2] + i) != 5) {
      release();
      std::fclose(fp);
      return MFISTA_IO_E_ROW;
    }
  }
  std::fclose(fp);
  *M = m;
  *NX = nx;
  *NY = ny;
  *u_idx = iu;
  *v_idx = iv;
  *y_r = d[0];
  *y_i = d[1];
  *noise_stdev = d[2];
  return MFISTA_IO_OK;
}

int mfista_write_nufft_data(const char* fname, int M, int NX, int NY, const double* u_dx, const double* v_dy,
                            const double* vis_r, const double* vis_i, const double* vis_std, int digits) {
  if (M < 0 || NX <= 0 || NY <= 0 || digits < 1 || digits > 17) return MFISTA_IO_E_ARG;
  if (M > 0 && (!u_dx || !v_dy || !vis_r || !vis_i || !vis_std)) return MFISTA_IO_E_ARG;
  FILE* fp = open_or_report(fname, "w");
  if (!fp) return MFISTA_IO_E_OPEN;
  std::fprintf(fp, "M  = %d\nNX = %d\nNY = %d\n\nu_rad, v_rad, vis_r, vis_i, noise_std_dev\n\n", M, NX, NY);
  for (int i = 0; i < M; i++)
    std::fprintf(fp, "%.*e, %.*e, %.*e, %.*e, %.*e\n", digits, u_dx[i], digits, v_dy[i], digits, vis_r[i], digits,
                 vis_i[i], digits, vis_std[i]);
  return std::fclose(fp) == 0 ? MFISTA_IO_OK : MFISTA_IO_E_OPEN;
}

int mfista_write_fft_data(const char* fname, int M, int NX, int NY, const int* u_idx, const int* v_idx,
                          const double* y_r, const double* y_i, const double* noise_stdev, int digits) {
  if (M < 0 || NX <= 0 || NY <= 0 || digits < 1 || digits > 17) return MFISTA_IO_E_ARG;
  if (M > 0 && (!u_idx || !v_idx || !y_r || !y_i || !noise_stdev)) return MFISTA_IO_E_ARG;
  FILE* fp = open_or_report(fname, "w");
  if (!fp) return MFISTA_IO_E_OPEN;
  std::fprintf(fp, "M  = %d\nNX = %d\nNY = %d\n\nu, v, y_r, y_i, noise_std_dev\n\n", M, NX, NY);
  for (int i = 0; i < M; i++)
    std::fprintf(fp, "%d, %d, %.*e, %.*e, %.*e\n", u_idx[i], v_idx[i], digits, y_r[i], digits, y_i[i], digits,
                 noise_stdev[i]);
  return std::fclose(fp) == 0 ? MFISTA_IO_OK : MFISTA_IO_E_OPEN;
}

void mfista_io_free(void* p) { std::free(p); }

}  // extern "C"
