/* rfxhip — C ABI of the MI355X (gfx950) FDTD time-stepping backend.
 *
 * The reference (bk-squared/rfx) has no FFI: its hot path is the Python
 * operator boundary
 *   rfx.simulation.run(grid, materials, n_steps, *, boundary, cpml_axes,
 *                      pec_axes, sources, probes, dft_planes,
 *                      lumped_port_sparams, pec_mask, ...) -> SimResult
 *   (rfx/simulation.py:1686-1725, result :2125-2138) and
 *   rfx.nonuniform.run_nonuniform(...) (rfx/nonuniform.py:1589-1618).
 * This header is the seam that replaces the body of those two functions
 * (the jax.lax.scan at rfx/simulation.py:1977 and rfx/nonuniform.py:1796):
 * the host sets the problem up (grid, materials, CPML profiles, source
 * tables) and hands device buffers to a plan that advances the fields.
 *
 * Conventions: every call returns 0 on success and a negative code on
 * failure; rfxh_last_error() gives a thread-local message.  No C++
 * exception crosses this boundary.  Device pointers ("d_" prefix) are owned
 * by the caller and must stay alive for the plan's lifetime; host arrays are
 * copied at rfxh_plan_create.  Field arrays are float32, [nx][ny][nz]
 * row-major (z fastest).  A plan is used from one host thread at a time;
 * distinct plans are independent.  Stream arguments are hipStream_t passed
 * as void* (NULL = the default stream); all work is stream-ordered and
 * asynchronous.
 */
#ifndef RFXHIP_H
#define RFXHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RFXH_OK 0
#define RFXH_EINVAL -1
#define RFXH_EHIP -2
#define RFXH_ENOMEM -3

/* Material flag bits (per material id). */
#define RFXH_ZERO_EX 1u   /* tangential-edge PEC mask on Ex (rfx/boundaries/pec.py:179-208) */
#define RFXH_ZERO_EY 2u
#define RFXH_ZERO_EZ 4u
#define RFXH_POINT 8u     /* cell carries point operations (source/probe/port) */
#define RFXH_ANISO 16u    /* subpixel smoothing: per-component eps at the cell; ca, cb of
                             mat_coef hold the Ex coefficients, aniso_coef the Ey, Ez ones
                             (rfx/core/yee.py:824-901, graded update_e_nu_aniso :668-719) */
#define RFXH_POLES 0xFFFF0000u /* bit 16 + q: dispersion pole q acts at the cell
                                   (per-pole masks, rfx/materials/debye.py:76-179,
                                   rfx/materials/lorentz.py:103-198) */

/* Point operation kinds. */
#define RFXH_OP_PORT 0    /* record V = -E_c * d_par and the Ampere-loop current */
#define RFXH_OP_INJECT 1  /* E_c += wave[n][slot]   (rfx/simulation.py:1520-1523) */
#define RFXH_OP_PROBE 2   /* out[n][slot] = F_c     (rfx/simulation.py:1565-1567) */

typedef struct {
  int64_t cell;    /* linear index (i*ny + j)*nz + k */
  int32_t kind;    /* RFXH_OP_* */
  int32_t comp;    /* 0..5 = ex, ey, ez, hx, hy, hz */
  int32_t slot;    /* column in the source table / probe output / port output */
  int32_t order;   /* execution order within one cell (ascending) */
  float d_par;     /* port: length along the port component */
  float loop[3];   /* port: per-axis length multiplying each H difference */
} rfxh_point_op;

typedef struct {
  int32_t axis;        /* plane normal 0/1/2 */
  int32_t index;       /* plane index along the normal */
  int32_t comp;        /* 0..5 */
  int32_t n_freq;
  const double* freqs; /* host, n_freq Hz */
  double* d_acc;       /* device complex128 [n_freq][n1][n2] (re, im interleaved) */
} rfxh_plane;

typedef struct {
  int32_t axis;        /* plane normal 0/1/2 */
  int32_t index;       /* plane index along the normal */
  int32_t lo1, n1;     /* region along the plane's first array dimension */
  int32_t lo2, n2;     /* region along the second */
  int32_t n_freq;
  const double* freqs; /* host, n_freq Hz */
  const float* weight; /* host [max_steps + 1]: DFT window weight per step value; NULL = 1 */
  double* d_acc;       /* device complex128 [4][n_freq][n1][n2]: tangential E1, E2, H1, H2 */
} rfxh_flux;

/* Near-to-far-field (Huygens) box: DFTs of the tangential fields on its six
 * faces (rfx/farfield.py:32-80 NTFFBox, :208-313 accumulate_ntff; in-scan
 * rfx/simulation.py:1570-1572, rfx/nonuniform.py:1524-1527).  Faces lie on
 * the index planes lo[a] and hi[a] of each axis and cover the half-open
 * ranges [lo, hi) of the other two axes.  Per face, in this order,
 *   x faces (Ey, Ez, Hy, Hz) over [lo1, hi1) x [lo2, hi2) = j x k,
 *   y faces (Ex, Ez, Hx, Hz) over i x k,
 *   z faces (Ex, Ey, Hx, Hy) over i x j,
 * E at t = n dt and H at t = (n + 1/2) dt for step n (scan index), the
 * fields of the plane itself (no averaging), float64 sums (the reference
 * uses complex64 with Kahan compensation). */
typedef struct {
  int32_t lo[3];       /* i_lo, j_lo, k_lo */
  int32_t hi[3];       /* i_hi, j_hi, k_hi (hi > lo, hi < n_axis) */
  int32_t n_freq;
  const double* freqs; /* host, n_freq Hz */
  double* d_acc[6];    /* device complex128 [4][n_freq][n1][n2] per face x_lo, x_hi, y_lo, y_hi,
                          z_lo, z_hi */
} rfxh_ntff;

typedef struct {
  int32_t nx, ny, nz;
  int32_t graded;          /* 0: uniform spacing; 1: graded (per-index inverse spacings) */
  const float* inv_e[3];   /* host, length n_axis: E-update inverse spacings
                              (uniform: 1/dx; graded: rfx/nonuniform.py:212-224 inv_d_e) */
  const float* inv_h[3];   /* host: H-update inverse spacings (graded: inv_d_h, last = 0) */
  double dt;
  /* materials: id per cell + per-id coefficient table */
  const uint16_t* d_mat;   /* device [nx*ny*nz] */
  int32_t n_mat;
  const float* mat_coef;   /* host [n_mat][4] = ca, cb, ce (= dt/(eps_r eps0)), ch (= dt/(mu_r mu0)) */
  const uint32_t* mat_flags; /* host [n_mat] RFXH_ZERO_* | RFXH_POINT | pole bits (RFXH_POLES) */
  /* CPML, faces ordered x_lo, x_hi, y_lo, y_hi, z_lo, z_hi */
  int32_t cpml_n[6];       /* slab depth per face (0 = no CPML on that face) */
  const float* cpml_b[6];  /* host [cpml_n[f]] in slab-local order (index 0 = slab's first cell) */
  const float* cpml_c[6];
  const float* cpml_km1[6];/* 1/kappa - 1 */
  float cpml_inv_d[6];     /* 1 / face cell size */
  int32_t pec_axes;        /* bit a: zero tangential E on both faces of axis a (rfx/boundaries/pec.py:13-48) */
  /* state (device, caller-owned) */
  float* d_f[6];           /* ex, ey, ez, hx, hy, hz */
  float* d_psi;            /* CPML memory, rfxh_psi_floats() floats, zeroed by the caller */
  float* d_f2[6];          /* optional second field set; when given (with d_psi2 if there is
                              CPML) each step reads one set and writes the other, which lets
                              one fused kernel do the whole step.  NULL: in-place two-pass step */
  float* d_psi2;           /* second CPML memory set (same size as d_psi) */
  /* point operations */
  int32_t n_ops;
  const rfxh_point_op* ops;/* host, any order (sorted internally by cell, order) */
  int32_t n_src;
  const float* d_src_wave; /* device [max_steps][n_src] */
  int32_t n_probe;
  float* d_probe_out;      /* device [max_steps][n_probe] */
  int32_t n_port;
  float* d_port_out;       /* device [max_steps][n_port][2] = (V, I) */
  int64_t max_steps;       /* rows of the tables above */
  /* DFT planes (running sums in float64) */
  int32_t n_plane;
  const rfxh_plane* planes;
  int32_t plane_t_offset;  /* t = (n + offset) dt: 1 uniform lane (rfx/simulation.py:1575),
                              0 graded lane (rfx/nonuniform.py:1458-1476) */
  /* flux monitors (rfx/probes/probes.py:476-613, in-scan rfx/simulation.py:1590-1644,
     rfx/nonuniform.py:1484-1522): tangential (E1, E2, H1, H2) with
     (E1, E2, H1, H2) = (Ey, Ez, Hy, Hz) | (Ez, Ex, Hz, Hx) | (Ex, Ey, Hx, Hy) for
     normal x | y | z; E of the plane at t = (n + offset) dt, H averaged over
     the planes index-1 and index at t - dt/2; float64 sums */
  int32_t n_flux;
  const rfxh_flux* flux;
  /* near-to-far-field boxes */
  int32_t n_ntff;
  const rfxh_ntff* ntff;
  /* dispersion (E update rfx/simulation.py:361-500): at the cells whose
     material flags carry pole bits, with E the plain update ca E^n + cb curl H,
       Lorentz poles (rfx/materials/lorentz.py):  P_l^{n+1} = a P_l^n + b P_l^{n-1} + c E^n,
       E^{n+1} = E + sum_d cc_d P_d^n - cc_l sum_l (P_l^{n+1} - P_l^n),
       Debye poles (rfx/materials/debye.py):  P_d^{n+1} = alpha P_d^n + beta (E^{n+1} + E^n);
     ca, cb of every material are then the dispersive ones */
  int32_t n_pole;          /* 0..16: poles [0, n_debye) Debye, [n_debye, n_pole) Lorentz */
  int32_t n_debye;
  const float* pole_coef;  /* host [n_mat][n_pole][4]: Debye (alpha, beta, cc, 0), Lorentz
                              (a, b, c, cc_l); zero where the material lacks the pole */
  int32_t pol_lo[3];       /* box holding every cell with poles */
  int32_t pol_n[3];
  float* d_pol;            /* device, zeroed by the caller: per Debye pole P[3][box cells],
                              then per Lorentz pole P[3][box cells], P^{n-1}[3][box cells] */
  /* subpixel smoothing (rfx/runners/uniform.py:271-278, rfx/geometry/smoothing.py):
     per id flagged RFXH_ANISO, (ca, cb) of Ey and of Ez from that component's eps_r
     and the id's sigma, the ce of the CPML correction staying the isotropic one;
     such plans run the two-pass kernels (0, 4, 5, 6) and have no reverse mode */
  const float* aniso_coef; /* host [n_mat][4] = ca_y, cb_y, ca_z, cb_z, or NULL */
} rfxh_desc;

typedef struct rfxh_plan rfxh_plan;

/* Number of floats of CPML memory the descriptor needs. */
int64_t rfxh_psi_floats(const rfxh_desc* desc);

/* Validate the descriptor, copy host tables to device, build the kernels'
 * parameter block.  Replaces the set-up half of rfx/simulation.py:626-1086. */
int rfxh_plan_create(const rfxh_desc* desc, rfxh_plan** out);
int rfxh_plan_destroy(rfxh_plan* plan);

/* Advance the fields by n_steps time steps starting at absolute step step0
 * (rows step0 .. step0+n_steps-1 of the source/probe/port tables).  The loop
 * body is rfx/simulation.py:1256-1681 (uniform) or rfx/nonuniform.py:
 * 1256-1565 (graded).  use_graph != 0 replays chunks of the step sequence
 * from a captured hipGraph. */
int rfxh_run(rfxh_plan* plan, int64_t step0, int64_t n_steps, void* stream, int use_graph);

/* With two field sets: which set (0 = d_f/d_psi, 1 = d_f2/d_psi2) holds the
 * current state; rfxh_set_state_set tells the plan after the caller wrote a
 * state (initial conditions, a restored checkpoint) into a set. */
int rfxh_state_set(rfxh_plan* plan);
int rfxh_set_state_set(rfxh_plan* plan, int which);

/* Average duration in milliseconds of the kernels of the sampled steps of
 * runs made with timing enabled (rfxh_set_timing(plan, 1), which also resets
 * the sums), measured with HIP events on the stream the kernels run on.
 * which: 0 = single-launch step kernel (two-pass: H pass), 1 = E pass
 * (two-pass only), 2 = DFT-plane kernels. */
int rfxh_set_timing(rfxh_plan* plan, int enable);
int rfxh_kernel_time_ms(rfxh_plan* plan, int which, double* avg_ms, int64_t* count);

/* Kernel variant of a plan.  kernel 0: two-pass step (param = workgroup
 * order, 1 flat, 0 XCD bands); 1: fused single-launch step (same param);
 * 2 / 3: marching step with 6- / 14-row tiles (param = number of x
 * segments, 0 = automatic); 4 / 6 / 5: two-pass step with columns of
 * 2 / 3 / 4 cells along x per thread (param = workgroup order); 7: wave-march
 * step, each wavefront walking a pair of rows along x with no barriers
 * (param = x segments, 0 = automatic; at most 1024 material ids); 8:
 * temporal-blocked march, two time steps per sweep (temporal-block depth 2;
 * param = x segments, 0 = automatic; at most 1024 material ids and 4 DFT
 * planes, no flux monitors, NTFF boxes or dispersion; an odd last step of
 * a run takes the two-pass kernels; rfxh_kernel_time_ms which = 0 is then
 * the two-step sweep).  Kernels 1-3, 7 and 8 need the second field set.
 * Variants compute the same step; they differ in speed only.  The
 * project's autotuner (rfx_amd/autotune.py) times them on the actual
 * problem.  No counterpart in the reference (its XLA lowering has no
 * kernel choice). */
int rfxh_set_variant(rfxh_plan* plan, int kernel, int param);
int rfxh_get_variant(rfxh_plan* plan, int* kernel, int* param);

/* ---- reverse mode (gradient of a probe-series objective w.r.t. eps_r) ----
 * Replaces jax.grad through the scan (rfx/api/_execute.py:2447-2858 with
 * rfx/simulation.py:1975, 2011-2055 rematerialisation): the forward run
 * records, per step and per cell of a box region, E^n, curl H^{n+1/2} and
 * the CPML-E term of the E update (9 floats); the reverse sweep runs the
 * transposed step backwards and accumulates dJ/d eps_r on the region. */
typedef struct {
  int32_t i0, j0, k0;      /* region origin (cells) */
  int32_t ni, nj, nk;      /* region extent */
} rfxh_region;

/* Forward steps step0 .. step0+n_steps-1 (two-pass sub-steps on the current
 * field set) recording 9 x region cells floats per step into
 * d_rec[n][9][region] (row n = step step0 + n). */
int rfxh_run_record(rfxh_plan* plan, int64_t step0, int64_t n_steps, const rfxh_region* region,
                    float* d_rec, void* stream);
/* The same, also recording into d_rec_h[n][3][region] (device; NULL: not
 * recorded) the H update's term R^n = -curl_fwd E^n + CPML-H term, with
 * H^{n+1/2} = H^{n-1/2} + ch R^n, which dJ/d mu_r needs (the gradient of
 * mu_r_override, rfx/api/_execute.py:2800-2805). */
int rfxh_run_record_h(rfxh_plan* plan, int64_t step0, int64_t n_steps, const rfxh_region* region,
                      float* d_rec, float* d_rec_h, void* stream);
/* rfxh_run_record_h plus, for plans with Debye / Lorentz poles (required
 * there, ignored elsewhere), d_rec_k[n][3][region]: the dispersive part of
 * the E update, K^n = sum_d cc_d P_d^n - cc_L sum_l (P_l^{n+1} - P_l^n)
 * (zero at cells without poles), which the eps_r and sigma gradients of
 * dispersive cells need (the reference differentiates its ADE,
 * rfx/materials/debye.py:182-250 and lorentz.py:200-275, with jax.grad). */
int rfxh_run_record_ex(rfxh_plan* plan, int64_t step0, int64_t n_steps, const rfxh_region* region,
                       float* d_rec, float* d_rec_h, float* d_rec_k, void* stream);

typedef struct {
  float* d_adj[6];           /* adjoint fields E-bar x,y,z, H-bar x,y,z; zeroed before the first call */
  float* d_adj_psi;          /* adjoint CPML memory, rfxh_psi_floats() floats, zeroed */
  float* d_gbar;             /* scratch, rfxh_psi_floats() floats */
  const float* d_probe_adj;  /* [max_steps][n_probe] dJ/d(probe sample) */
  float* d_src_adj;          /* [max_steps][n_src] out: dJ/d(source table entry); may be NULL */
  const float* d_rec;        /* forward record rows rec_step0 .. rec_step0+rec_rows-1 */
  int64_t rec_step0, rec_rows;
  rfxh_region region;
  float* d_grad;             /* [region cells] dJ/d eps_r, accumulated (+=); NULL: no gradient */
  const float* dcoef;        /* host [n_mat][3]: d ca, d cb, d ce / d eps_r per material id */
  /* dJ/d sigma on the same region (the gradient of sigma_override,
   * rfx/api/_execute.py:2800-2805), from the same record: ca and cb depend
   * on sigma, ce does not.  NULL: not computed (needs d_grad). */
  float* d_grad_sigma;       /* [region cells], accumulated (+=) */
  const float* dcoef_sigma;  /* host [n_mat][2]: d ca, d cb / d sigma per material id */
  /* dJ/d mu_r on the same region from the H record of rfxh_run_record_h
   * (same rows as d_rec).  NULL: not computed (needs d_grad). */
  const float* d_rec_h;      /* [rec_rows][3][region cells] */
  float* d_grad_mu;          /* [region cells], accumulated (+=) */
  const float* dcoef_mu;     /* host [n_mat]: d ch / d mu_r per material id */
  /* dJ/d(port samples) [max_steps][n_port][2] = dJ/dV, dJ/dI (device;
   * NULL: none): objectives on lumped-port S-parameters (rfx/probes/
   * probes.py:1240-1274 differentiated through the scan). */
  const float* d_port_adj;
  /* dJ/d(DFT plane accumulators): host array of the plan's n_plane device
   * pointers (complex128 [n_freq][n1][n2] each, torch's gradient of the
   * accumulator: dJ/dRe + j dJ/dIm; a NULL entry, or a NULL array: none)
   * (rfx/probes/probes.py:379-473 differentiated through the scan). */
  const double* const* d_plane_adj;
  /* dJ/d(flux accumulators): host array of one device pointer per flux
   * monitor, in the plan's order, then six per NTFF box (faces x-lo, x-hi,
   * y-lo, y-hi, z-lo, z-hi), each complex128 [4][n_freq][n1][n2] like the
   * accumulator (dJ/dRe + j dJ/dIm); a NULL entry, or a NULL array: none
   * (rfx/farfield.py:208-313 and the flux DFT of rfx/api/__init__.py:
   * 2577-2645, differentiated through the scan). */
  const double* const* d_flux_adj;
  /* Plans with Debye / Lorentz poles: the adjoint polarisation, device
   * float32 laid out like the plan's polarisation state (d_pol), zeroed by
   * the caller before the first (latest) segment and carried between
   * segments; with d_grad also the dispersive record of rfxh_run_record_ex
   * (rows as d_rec) and per material id the host pair (-(d gamma/d eps_r)/
   * gamma, -(d gamma/d sigma)/gamma) of the ADE normaliser gamma = eps_inf +
   * sum_d beta_d + sigma dt/2. */
  float* d_pol_adj;
  const float* d_rec_k;
  const float* dcoef_disp;
} rfxh_adjoint_desc;

/* Reverse sweep over steps hi-1 down to lo.  The adjoint state carries
 * across calls, so a run can be swept segment by segment (latest first),
 * each segment with its own forward record. */
int rfxh_adjoint(rfxh_plan* plan, const rfxh_adjoint_desc* desc, int64_t lo, int64_t hi,
                 void* stream);

/* ---- batched sweeps ----
 * Many independent problems (uneven grid sizes and step counts) advanced
 * together: every time step is one H launch and one E launch covering all
 * instances that still have steps left, instead of two launches per
 * instance.  Replaces the per-instance loop of rfx/batch.py:530-559
 * (run_batch) and the vmap program of rfx/vmap_sweep.py:434-643, whose
 * uneven-shape case runs instances one after another.  Each instance keeps
 * its own plan (fields, tables, outputs); a batch only groups them. */
typedef struct rfxh_batch rfxh_batch;
int rfxh_batch_create(rfxh_plan* const* plans, int n_plans, rfxh_batch** out);
int rfxh_batch_destroy(rfxh_batch* batch);
/* Instance q advances n_steps[q] steps from absolute step step0[q]
 * (two-pass sub-steps on its current field set, one cell per thread, each
 * instance in its plan's workgroup order).  use_graph != 0 replays captured
 * chunks of steps wherever the set of active instances is constant. */
int rfxh_batch_run(rfxh_batch* batch, const int64_t* step0, const int64_t* n_steps, void* stream,
                   int use_graph);

/* Batched reverse mode: the gradients of a sweep's instances computed
 * together, where the reference takes jax.grad of each instance's run
 * (rfx/vmap_sweep.py:1017-1022 maps the gradient over the sweep;
 * rfx/api/_execute.py:2447-2858 per instance).  Same semantics as
 * rfxh_run_record / rfxh_adjoint per instance; every time step is one
 * launch per sub-step covering all instances still active.
 *
 * Record: instance q advances n_steps[q] steps from step0[q] (0 = sits
 * out) recording its region regions[q] into d_rec[q] (device, as
 * rfxh_run_record); accumulate != 0 also takes the probe / DFT outputs of
 * those steps, as rfxh_run_record does (0: a recomputation whose outputs
 * are already known). */
int rfxh_batch_run_record(rfxh_batch* batch, const int64_t* step0, const int64_t* n_steps,
                          const rfxh_region* regions, float* const* d_rec, int accumulate,
                          void* stream);
/* Reverse sweep of instance q over steps hi[q]-1 down to lo[q] with
 * descs[q] (NULL = sits out), as rfxh_adjoint; instances sweep their own
 * step ranges side by side. */
int rfxh_batch_adjoint(rfxh_batch* batch, const rfxh_adjoint_desc* const* descs,
                       const int64_t* lo, const int64_t* hi, void* stream);

/* ---- near-to-far-field transform ----
 * Far-field radiation of an NTFF box from its accumulated face DFTs
 * (rfx/farfield.py:393-556 compute_far_field, post-processing in NumPy in
 * the reference).  Per face with outward normal n (sign s = -1 on lo faces,
 * +1 on hi faces): J = n x H, M = -n x E; per frequency k = 2 pi f / c0 and
 * direction r(theta, phi):
 *   N = sum J e^{j k r.p} dS,  L = sum M e^{j k r.p} dS,
 *   E_theta = -j k / (4 pi) (L_phi + eta0 N_theta),
 *   E_phi   =  j k / (4 pi) (L_theta - eta0 N_phi)
 * with p the position of the face cell (pos[a][index]) and dS the product
 * of the widths (width[a][index]) of its two tangential indices. */
typedef struct {
  int32_t lo[3], hi[3];      /* the box, as in rfxh_ntff */
  int32_t n_freq;
  const double* freqs;       /* host, n_freq Hz */
  const double* d_acc[6];    /* device, the rfxh_ntff accumulators (read only) */
  int32_t n_axis[3];         /* lengths of pos[a] / width[a] (the grid shape) */
  const double* pos[3];      /* host: position (m) of each index along each axis */
  const double* width[3];    /* host: cell width (m) of each index along each axis */
  double c0;                 /* speed of light (m/s) */
  double eta0;               /* free-space impedance (ohm) */
} rfxh_farfield_desc;

/* e_theta, e_phi: device complex128 [n_freq][n_theta][n_phi] (re, im
 * interleaved), caller-owned.  theta, phi: host arrays (radians).  Copies
 * the small host tables to the device, runs one kernel on `stream` and waits
 * for it (the call returns with the outputs written). */
int rfxh_ntff_far_field(const rfxh_farfield_desc* desc, int32_t n_theta, const double* theta,
                        int32_t n_phi, const double* phi, double* e_theta, double* e_phi,
                        void* stream);

/* The same transform with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Arguments and results as rfxh_ntff_far_field (the direction
 * vectors are computed on the device, so results agree with it to float64
 * rounding, not bit for bit).  Its tables live in one file-scope device array
 * per GPU of 1,048,576 doubles (8 MiB, held until the process ends), which
 * must hold 9 per direction + 2 per frequency + 2 per grid index of each axis
 * + the theta and phi values; larger requests return RFXH_EINVAL.  Every call
 * on a GPU shares that GPU's array: calls that overlap on one GPU (other host
 * threads, other streams) race.  Waits for its stream before returning. */
/* Reverse mode of rfxh_ntff_far_field (the reference differentiates its
 * far field with jax.grad through rfx/farfield.py:393-556): from g_theta,
 * g_phi (device complex128 [n_freq][n_theta][n_phi], torch's gradient
 * dJ/dRe + j dJ/dIm of E_theta and E_phi) writes dJ/d(face accumulator)
 * into d_gacc[face] (device complex128 [4][n_freq][n1][n2] per face,
 * caller-owned, overwritten).  desc->d_acc is not read.  The transform is
 * linear: the result is the conjugate transpose of the forward map applied
 * to (g_theta, g_phi).  Allocates its tables and scratch, waits for its
 * stream before returning. */
int rfxh_ntff_far_field_adjoint(const rfxh_farfield_desc* desc, int32_t n_theta,
                                const double* theta, int32_t n_phi, const double* phi,
                                const double* g_theta, const double* g_phi,
                                double* const* d_gacc, void* stream);

int rfxh_ntff_far_field_noalloc(const rfxh_farfield_desc* desc, int32_t n_theta,
                                const double* theta, int32_t n_phi, const double* phi,
                                double* e_theta, double* e_phi, void* stream);

/* ---- Device-side material assembly ----------------------------------------
 * The per-cell uint16 material ids, built on the GPU in the reference's
 * order: box fill in insertion order, later boxes winning
 * (rfx/api/_compile.py:206-224 with Box.mask_on_coords, rfx/geometry/
 * csg.py:235-307: each box's node set is a product of three index ranges);
 * CPML pad extension (rfx/geometry/rasterize_grid.py:237-414); tangential
 * PEC zero bits (rfx/boundaries/pec.py:90-176), folded into the ids through
 * a table the caller builds from the (id, bits) pairs that occur.  Replaces
 * the host loops of _assemble_materials (rfx/api/_compile.py:121-367). */
typedef struct {
  int32_t lo[3], hi[3];  /* half-open node-index range per axis (lo >= hi: empty) */
  int32_t id;            /* material id the box paints */
  int32_t pec;           /* 1: the box marks perfect-conductor cells (sigma >= 1e6)
                            instead of painting (rfx/api/_compile.py:118, 216-220) */
} rfxh_asm_box;

/* d_ids (device uint16 [nx][ny][nz]) <- 0 (vacuum) painted by the non-PEC
 * boxes in order; d_pec (device uint8, same shape) <- 1 where a PEC box
 * covers the cell.  pads (host [6], or NULL: none) = CPML pad widths lo_x,
 * hi_x, lo_y, hi_y, lo_z, hi_z: along x, then y, then z, the lo pad copies
 * the first interior slice and the hi pad the last one, except where that
 * slice is vacuum and the one inside it is not (then the inner slice, also
 * written onto the last interior slice).  vacuum: host uint8 [n_mat], 1 for
 * ids whose (eps_r, sigma, mu_r) is (1, 0, 1).  boxes: host [n_box]. */
int rfxh_assemble_fill(int32_t nx, int32_t ny, int32_t nz, const rfxh_asm_box* boxes,
                       int32_t n_box, const int32_t* pads, const uint8_t* vacuum, int32_t n_mat,
                       uint16_t* d_ids, uint8_t* d_pec, void* stream);

/* The tangential-zero bits of the PEC cells (bit a: E along axis a is zeroed
 * because a neighbour along a is PEC too, none past the domain): bit b of
 * d_used[id] (device uint32 [n_mat], caller-zeroed) is set when a PEC cell
 * with that id carries bits b != 0. */
int rfxh_assemble_pec_pairs(int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_ids,
                            const uint8_t* d_pec, uint32_t* d_used, int32_t n_mat, void* stream);

/* Each PEC cell with bits b != 0 takes id map[id * 8 + b] (host uint16
 * [n_mat * 8]; the caller fills the pairs d_used reported). */
int rfxh_assemble_pec_apply(int32_t nx, int32_t ny, int32_t nz, uint16_t* d_ids,
                            const uint8_t* d_pec, const uint16_t* map, int32_t n_mat,
                            void* stream);

const char* rfxh_last_error(void);
int rfxh_version(void);

#ifdef __cplusplus
}
#endif
#endif
