/*
 * mfista_io.h -- file formats and report of the imaging command-line tools,
 * exported by libmfista_amd.so (host code only; no GPU call).
 *
 * The first block has the names, arguments and return values of the
 * reference's I/O helpers (c/mfista.h:68-95, c/mfista_io.c), so that a
 * program written against them links unchanged.  One difference: where the
 * reference calls exit(1) because a file cannot be opened (fopenr/fopenw,
 * c/mfista_io.c:9-24), these functions print the same message to stderr and
 * return -1 (0 for read_A_matrix), leaving the decision to the caller.
 *
 * The second block reads and writes the two text data files the command-line
 * tools take (header "M  = ..", "NX = ..", "NY = .." then one CSV row per
 * visibility), parsed as c/mfista_imaging_nufft.c:140-170 and
 * c/mfista_imaging_fft.c:153-183 parse them.
 */
#ifndef MFISTA_AMD_IO_H
#define MFISTA_AMD_IO_H

#include <stdio.h>

#include "mfista.h"

#ifdef __cplusplus
extern "C" {
#endif

/* c/mfista.h:68-75 */
struct IO_FNAMES {
  unsigned int fft;
  char *fft_fname;
  char *v_fname;
  char *A_fname;
  char *in_fname;
  char *out_fname;
};

/* c/mfista_io.c:27-38: binary int32 vector; returns the number read */
int read_int_vector(char *fname, int length, int *vector);
/* c/mfista_io.c:40-51: binary float32 vector (CLEAN box) */
int read_f_vector(char *fname, int length, float *vector);
/* c/mfista_io.c:53-64: binary float64 vector (x init, DFT y) */
int read_V_vector(char *fname, int length, double *vector);
/* c/mfista_io.c:66-101: binary float64 height x width matrix, as stored */
unsigned long read_A_matrix(char *fname, int height, int width, double *matrix);
/* c/mfista_io.c:103-114: binary float64 vector (x out); returns the number written */
int write_X_vector(char *fname, int length, double *vector);
/* c/mfista_io.c:118-137: in place, row-major origheight x origwidth ->
 * row-major origwidth x origheight */
void transpose_matrix(double *matrix, int origheight, int origwidth);
/* c/mfista_io.c:141-165: the "IO files of ..." block of the report */
void show_io_fnames(FILE *fid, char *fname, struct IO_FNAMES *mfista_io);
/* c/mfista_io.c:167-215: the "Output of ..." block of the report */
void show_result(FILE *fid, char *fname, struct RESULT *mfista_result);

/* ---- extensions (no reference counterpart) ----------------------------- */

#define MFISTA_IO_OK 0
#define MFISTA_IO_E_OPEN -1   /* file cannot be opened */
#define MFISTA_IO_E_HEADER -2 /* header lines do not parse */
#define MFISTA_IO_E_ROW -3    /* a data row does not parse ("cannot read data.") */
#define MFISTA_IO_E_ARG -4    /* invalid argument */
#define MFISTA_IO_E_NOMEM -5

/* NUFFT data file: "u_rad, v_rad, vis_r, vis_i, noise_std_dev" rows.  The
 * five arrays are malloc'ed by the library (release with mfista_io_free).
 * On failure nothing is allocated. */
int mfista_read_nufft_data(const char *fname, int *M, int *NX, int *NY, double **u_dx,
                           double **v_dy, double **vis_r, double **vis_i, double **vis_std);
/* FFT data file: "u, v, y_r, y_i, noise_std_dev" rows with integer cells */
int mfista_read_fft_data(const char *fname, int *M, int *NX, int *NY, int **u_idx, int **v_idx,
                         double **y_r, double **y_i, double **noise_stdev);
/* writers of the same formats; digits = significant digits after the point
 * (6 writes what the reference's data files hold, 16 round-trips a double) */
int mfista_write_nufft_data(const char *fname, int M, int NX, int NY, const double *u_dx,
                            const double *v_dy, const double *vis_r, const double *vis_i,
                            const double *vis_std, int digits);
int mfista_write_fft_data(const char *fname, int M, int NX, int NY, const int *u_idx,
                          const int *v_idx, const double *y_r, const double *y_i,
                          const double *noise_stdev, int digits);
void mfista_io_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
