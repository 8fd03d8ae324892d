/* C restatement of rfx's FDTD step — TEST INFRASTRUCTURE ONLY.
 *
 * Same algorithm and sub-step order as oracle/fdtd_np.py (which cites the
 * reference file:line for every piece); plain loops with OpenMP so the
 * oracle finishes larger cases in seconds and so the CPU baseline of
 * bench.py times a compiled restatement rather than numpy.  Built by
 * oracle/Makefile into oracle/liboracle_cstep.so; loaded by oracle/cstep.py.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_EPS0 8.8541878128e-12
#define ORACLE_MU0 1.25663706212e-6

typedef struct {
  long nx, ny, nz;
  int graded;               /* 0: uniform (divide by dx), 1: graded inv arrays */
  double dx;
  const float *inv_e[3];    /* graded: E-update inverse spacings per axis */
  const float *inv_h[3];    /* graded: H-update inverse spacings per axis */
  double dt;
  const float *eps_r, *sigma, *mu_r;
  long cpml_n;              /* layers per face, 0 = none */
  int cpml_axes;            /* bit a set = axis a has CPML */
  const float *prof_b[3][2], *prof_c[3][2], *prof_k[3][2];
  double face_d[3][2];
  int pec_axes;             /* bit a set = outer-face PEC on axis a */
  const unsigned char *pec_bits; /* per cell: 1 ex, 2 ey, 4 ez masked; NULL = none */
  long n_steps;
  int record_order;         /* 0 uniform, 1 graded */
  int n_src;
  const long *src_cell;
  const int *src_comp;
  const float *src_wave;    /* [n_steps][n_src] */
  int n_probe;
  const long *probe_cell;
  const int *probe_comp;
  int n_port;
  const long *port_cell;
  const int *port_comp;
  const double *port_dpar;
  const double *port_loop;  /* 3 per port */
  int n_plane;
  const int *plane_axis;
  const long *plane_index;
  const int *plane_comp;
  const int *plane_nf;
  const float *const *plane_freqs;
  double *const *plane_acc; /* complex128 accumulators [nf][n1][n2] */
  const double *const *init_fields; /* optional 6 initial fields */
  const float *aniso_eps[3];  /* per-component eps_r (subpixel smoothing), NULL = eps_r */
} oracle_desc;

typedef struct {
  double *ts;       /* [n_steps][n_probe] */
  double *port_v;   /* [n_port][n_steps] */
  double *port_i;
  double *const *fields; /* optional 6 output fields */
} oracle_out;

#define REAL double
#define SUFFIX _f64
#include "cstep_impl.h"
#undef REAL
#undef SUFFIX

#define REAL float
#define SUFFIX _f32
#include "cstep_impl.h"
#undef REAL
#undef SUFFIX

int oracle_run(const oracle_desc *d, oracle_out *out, int use_double) {
  return use_double ? run_f64(d, out) : run_f32(d, out);
}
