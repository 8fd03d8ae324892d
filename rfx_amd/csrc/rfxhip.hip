// rfxhip.hip — MI355X (gfx950) FDTD time-stepping kernels and their C ABI.
//
// One time step follows rfx/simulation.py:1256-1681 (uniform) or
// rfx/nonuniform.py:1256-1565 (graded):
//   H update (rfx/core/yee.py:225-291 / :411-451) + CPML-H (rfx/boundaries/
//   cpml.py:804-1065) -> E update (yee.py:346-404 / :456-492) + CPML-E
//   (cpml.py:522-801) -> outer-face PEC (rfx/boundaries/pec.py:13-48) ->
//   tangential PEC mask (pec.py:179-208) -> port V/I sampling, source
//   injection, probes -> DFT planes.
//
// Two kernel designs (no MFMA: ~33 flops per cell-step):
//   two-pass (default): an H launch and an E launch per step, in place; 12
//     components read + 6 written per cell-step.
//   fused (RFXH_KERNEL=fused, needs two field sets): one launch per step.  Fields and CPML memory are
//     double-buffered (read set A, write set B, swap), so a thread can
//     compute H^{n+1/2} at its own cell and, redundantly, at its three
//     backward neighbours, then E^{n+1} at its cell, without waiting for
//     other workgroups.  HBM traffic ~ 6 components read + 6 written +
//     a 2-byte material id per cell-step (the other loads hit L1/L2), but
//     four H evaluations per cell make it VALU-bound on MI355X.
// Fields are float32 [nx][ny][nz], z fastest.  Each workgroup covers 256
// consecutive cells of one (y, z) plane (flat chunks: neighbouring
// workgroups share only the cache line at their boundary; chunk index
// fastest, XCD-contiguous).  RFXH_ORDER=bands cuts the plane into 8 bands,
// one per XCD, swept plane after plane (lowest L2 miss traffic, measured
// slightly slower).
// Material coefficients come through a uint16 id per cell and a small
// cache-resident table.  The step index lives in device memory and is
// advanced by a one-thread kernel at the end of each step, so one captured
// hipGraph of a chunk of steps is replayed for every chunk.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <mutex>
#include <vector>

#include "rfxhip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return fail(RFXH_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));       \
  } while (0)

constexpr int kBX = 64;             // lanes (one wavefront)
// 4 wavefronts per workgroup: against 8, measured 1.9 % off the 2-cell
// column step, 12 % off the one-cell step and 10 % more batched throughput
// at 533^3 (more, smaller workgroups in flight per CU)
constexpr int kTY = 4;
constexpr int kWG = kBX * kTY;      // 256 cells per workgroup
constexpr int kTimingSamples = 64;  // sampled steps per timed run

struct KParams {
  int nx, ny, nz;
  int graded;
  long long plane;               // ny * nz
  float inv_e0, inv_h0;          // uniform spacing
  const float* inv_e[3];         // graded spacings (device)
  const float* inv_h[3];
  const uint16_t* mat;
  const float4* coef;            // ca, cb, ce, ch per material id
  const uint32_t* flags;
  const float4* ecoef;           // ca, cb, ce, flag bits (as float bits): the E update's row
  float* f[6];                   // two-pass kernels: the field set updated in place
  int cn[6];                     // CPML slab depth per face
  float* psi_h[6][2];            // CPML memory of set 0 (set 1 at + psi_set)
  float* psi_e[6][2];
  const float* pb[6];
  const float* pc[6];
  const float* pk[6];            // 1/kappa - 1
  float cinv[6];                 // 1 / face cell size
  int pec_axes;
  const rfxh_point_op* ops;
  int n_ops;
  const float* src_wave;
  int n_src;
  float* probe_out;
  int n_probe;
  float* port_out;
  int n_port;
  long long* step;               // current absolute step (device)
  int order;                     // workgroup order: 0 bands, 1 flat chunks
  float inv_nz;                  // 1/nz for the flat-index split
  int slo[3], shi[3];            // CPML slab extents: cells i < slo or i >= shi per axis
  float ch_all;                  // > 0: every material has this ch (one mu_r), and the
                                 // column H pass reads no material id
  const float4* acoef;           // RFXH_ANISO ids: ca, cb of Ey and of Ez
};

struct FusedIO {
  const float* eo[3];
  const float* ho[3];
  float* en[3];
  float* hn[3];
};

// CPML memory pointers of one read set and one write set, per face and term.
struct PsiIO {
  const float* hr[6][2];
  float* hw[6][2];
  const float* er[6][2];
  float* ew[6][2];
  // dispersive polarisation (in place, shared by both field sets) over the
  // box pol_lo + [0, pol_n): poles [0, n_debye) are Debye (P[3] each),
  // the rest Lorentz (P[3] then P^{n-1}[3]); per material and pole a row of
  // 4 floats, Debye (alpha, beta, cc, 0), Lorentz (a, b, c, cc)
  float* pol;
  const float* pcoef;
  int n_pole, n_debye;
  int pol_lo[3], pol_n[3];
};

constexpr int kDftK = 16;   // steps staged per fold of the DFT-plane accumulators
// ring slots of a DFT plane: two groups of kDftK, so a two-step sweep can
// stage the first step of the next group before the fold of the current one
constexpr int kDftRing = 2 * kDftK;
constexpr int kFoldF = 4;    // accumulators a fold keeps in flight

struct PlaneParams {
  const float* fld;
  double2* acc;
  const double2* tw;  // [max_steps][nf] twiddles exp(-j 2 pi f t) dt
  long long* step;
  float* snap;        // [kDftRing][n1*n2] ring of plane values not yet folded in
  const long long* acc_from;  // first step not yet folded into acc (device)
  long long max_steps;
  int nf;
  int axis, index;
  int n1, n2;
  int ny, nz;
};

struct FluxParams {
  const float* f[4];     // E1, E2, H1, H2 field arrays (current set)
  double2* acc;          // [4][nf][n1][n2]
  const double2* twe;    // [max_steps][nf] E twiddles (dt and window folded in)
  const double2* twh;    // [max_steps][nf] H twiddles (half a step off E)
  float* snap;           // [kDftK][4][n1*n2] ring of values not yet folded in
  const long long* acc_from;  // first step not yet folded into acc (device)
  long long* step;
  long long max_steps;
  int nf, axis, index, index_m1;
  int lo1, n1, lo2, n2;
  int ny, nz;
};

// Loads/stores at a 32-bit byte offset from a (wave-uniform) base pointer: the
// base stays in SGPRs and the address is one VGPR offset, no 64-bit math.
// Every array indexed this way is smaller than 4 GiB (checked at plan
// creation).
// The pointers are cast to the global address space: where a base pointer
// comes from device memory (the batched step's instance table) rather than
// a kernel argument, the compiler cannot infer it and would emit flat
// accesses, which also count against the scalar-load counter and stall the
// scalar parameter loads behind them.
#define RFX_G __attribute__((address_space(1)))
__device__ __forceinline__ float ldf(const float* base, unsigned cell) {
  return *(const RFX_G float*)((const RFX_G char*)base + (cell << 2));
}
__device__ __forceinline__ void stf(float* base, unsigned cell, float v) {
  *(RFX_G float*)((RFX_G char*)base + (cell << 2)) = v;
}
__device__ __forceinline__ unsigned ldu16(const uint16_t* base, unsigned cell) {
  return *(const RFX_G uint16_t*)((const RFX_G char*)base + (cell << 1));
}
// Element idx of a small global table (coefficients, profiles, spacings).
template <class T>
__device__ __forceinline__ T gld(const T* base, int idx) {
  return ((const RFX_G T*)base)[idx];
}
// ch of material m alone (one dword instead of the 16-byte row: the H
// update needs nothing else; measured 1.5 % off the H pass)
__device__ __forceinline__ float ld_ch(const float4* coef, unsigned m) {
  return gld(reinterpret_cast<const float*>(coef), 4 * (int)m + 3);
}
__device__ __forceinline__ float4 gld(const float4* base, int idx) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = ((const RFX_G v4*)base)[idx];
  return make_float4(v.x, v.y, v.z, v.w);
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so each XCD
// receives a contiguous range of logical workgroup ids.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const unsigned ub = (unsigned)b, un = (unsigned)nb;
  const unsigned per = un >> 3, rem = un & 7u;
  const unsigned x = ub & 7u, s = ub >> 3;
  return (int)(x < rem ? x * (per + 1) + s : rem * (per + 1) + (x - rem) * per + s);
}

// q = n / d, r = n % d for wave-uniform 0 <= n < 2^31, d >= 1: the product
// with the hardware double reciprocal (relative error near 2^-50, so the
// estimate is within one of the quotient) corrected once in integers --
// a few instructions instead of the generic integer-division sequence,
// which every workgroup of the step kernels runs before its first load.
__device__ __forceinline__ int div_fast(int n, int d, int& r) {
  int q = (int)((double)n * __builtin_amdgcn_rcp((double)d));
  r = n - q * d;
  if (r < 0) {
    --q;
    r += d;
  } else if (r >= d) {
    ++q;
    r -= d;
  }
  return q;
}

// Cell of this thread in logical workgroup L of the problem p.  All index
// math is 32-bit: a y-z plane holds fewer than 2^24 cells (checked at plan
// creation), so the in-plane index f is exact in float and the float
// quotient f/nz is within one of the integer one.
__device__ __forceinline__ bool cell_at(const KParams& p, int L, int& i, int& j, int& k,
                                        int nxb = -1) {
  // nxb: number of x positions the grid walks (x blocks of several planes
  // for the column kernels; default the planes themselves)
  if (nxb < 0) nxb = p.nx;
  const int pl = (int)p.plane;
  const int t = (int)(threadIdx.y * kBX + threadIdx.x);
  int f;
  if (p.order == 0) {
    // bands: the (y, z) plane is cut into 8 flat bands, one per XCD; the XCD
    // sweeps its band plane by plane with kWG-cell chunks fastest
    const int bw = (int)(((unsigned)pl + 7u) >> 3);
    const int ncb = (int)(((unsigned)bw + kWG - 1) / kWG);
    int r, c;
    const int b = div_fast(L, nxb * ncb, r);
    i = div_fast(r, ncb, c);
    f = b * bw + c * kWG + t;
    if (f >= (b + 1) * bw) return false;
  } else {
    // flat chunks of kWG cells, chunk index fastest, then plane
    const int nch = (int)(((unsigned)pl + kWG - 1) / kWG);
    int c;
    i = div_fast(L, nch, c);
    f = c * kWG + t;
  }
  if (f >= pl) return false;
  int jj = (int)((float)f * p.inv_nz);
  int kk = f - jj * p.nz;
  if (kk < 0) {
    --jj;
    kk += p.nz;
  } else if (kk >= p.nz) {
    ++jj;
    kk -= p.nz;
  }
  j = jj;
  k = kk;
  return true;
}

__device__ __forceinline__ bool cell_of(const KParams& p, int& i, int& j, int& k) {
  return cell_at(p, xcd_remap((int)blockIdx.x, (int)gridDim.x), i, j, k);
}

__device__ __forceinline__ bool in_slab(const KParams& p, int i, int j, int k) {
  return (i < p.slo[0]) | (i >= p.shi[0]) | (j < p.slo[1]) | (j >= p.shi[1]) |
         (k < p.slo[2]) | (k >= p.shi[2]);
}

// x-face slab index for psi: [l][j][k]; y-face: [i][l][k]; z-face: [i][j][l]
__device__ __forceinline__ long long psi_idx(int face, int l, int i, int j, int k, int n,
                                             int ny, int nz) {
  const int ax = face >> 1;
  if (ax == 0) return ((long long)l * ny + j) * nz + k;
  if (ax == 1) return ((long long)i * n + l) * nz + k;
  return ((long long)i * ny + j) * n + l;
}

__device__ __forceinline__ float pick3(int c, float a, float b, float d) {
  return c == 0 ? a : (c == 1 ? b : d);
}

// CPML slab lookup: local layer index of (i, j, k) in face's slab, or -1.
__device__ __forceinline__ int slab_layer(const KParams& p, int face, int i, int j, int k) {
  const int n = p.cn[face];
  if (n == 0) return -1;
  const int ax = face >> 1;
  const int pos = ax == 0 ? i : ax == 1 ? j : k;
  const int N = ax == 0 ? p.nx : ax == 1 ? p.ny : p.nz;
  const int l = pos - ((face & 1) ? N - n : 0);
  return (l < 0 || l >= n) ? -1 : l;
}

// Port V/I sampling, source injection and probes at one cell, in the order
// the host gave (rfx/simulation.py:1449-1567; rfx/nonuniform.py:1353-1542).
// n_step < 0: the step of the device counter.  own = false (a halo thread
// of the temporal-blocked march repeating another thread's cell): inject
// only, record nothing.
__device__ __forceinline__ void point_ops(const KParams& p, unsigned cell, float& e0, float& e1,
                                          float& e2, float hx, float hy, float hz, float hx_j,
                                          float hx_k, float hy_i, float hy_k, float hz_i,
                                          float hz_j, long long n_step = -1, bool own = true) {
  int lo = 0, hi = p.n_ops;  // first op of this cell (ops sorted by cell, order)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (p.ops[mid].cell < (long long)cell) lo = mid + 1;
    else hi = mid;
  }
  const long long n = n_step >= 0 ? n_step : *p.step;
  for (int q = lo; q < p.n_ops && p.ops[q].cell == (long long)cell; ++q) {
    const rfxh_point_op op = p.ops[q];
    if (op.kind == RFXH_OP_PORT) {
      if (!own) continue;
      const float v = -pick3(op.comp, e0, e1, e2) * op.d_par;
      float cur;
      if (op.comp == 2)
        cur = (hy - hy_i) * op.loop[1] - (hx - hx_j) * op.loop[0];
      else if (op.comp == 0)
        cur = (hz - hz_j) * op.loop[2] - (hy - hy_k) * op.loop[1];
      else
        cur = (hx - hx_k) * op.loop[0] - (hz - hz_i) * op.loop[2];
      float* o = p.port_out + (n * p.n_port + op.slot) * 2;
      o[0] = v;
      o[1] = cur;
    } else if (op.kind == RFXH_OP_INJECT) {
      const float w = p.src_wave[n * p.n_src + op.slot];
      if (op.comp == 0) e0 += w;
      else if (op.comp == 1) e1 += w;
      else e2 += w;
    } else if (own) {
      p.probe_out[n * p.n_probe + op.slot] =
          op.comp < 3 ? pick3(op.comp, e0, e1, e2) : pick3(op.comp - 3, hx, hy, hz);
    }
  }
}

// ---------------------------------------------------------------------------
// H at one cell from values: H^{n+1/2} = H^{n-1/2} - ch curl_fwd E^n
// (yee.py:272-289, graded :429-451) plus the CPML-H correction
// (cpml.py:875-1043).  E at the cell and at its forward neighbours (zero
// past the last cell, yee.py:121-128) come in as values; CPML memory is
// read from ps.hr and, when WRITE, the new memory is stored into ps.hw.
// ---------------------------------------------------------------------------
template <bool GRADED, bool WRITE>
__device__ __forceinline__ float3 h_core(const KParams& p, const PsiIO& ps, int i, int j, int k,
                                         float ch, float ex, float ey, float ez, float ex_j,
                                         float ez_j, float ex_k, float ey_k, float ey_i, float ez_i,
                                         float hx, float hy, float hz, bool wr = true) {
  float ix, iy, iz;
  if (GRADED) {
    ix = gld(p.inv_h[0], i);
    iy = gld(p.inv_h[1], j);
    iz = gld(p.inv_h[2], k);
  } else {
    ix = iy = iz = p.inv_h0;
  }
  hx = hx - ch * ((ez_j - ez) * iy - (ey_k - ey) * iz);
  hy = hy - ch * ((ex_k - ex) * iz - (ez_i - ez) * ix);
  hz = hz - ch * ((ey_i - ey) * ix - (ex_j - ex) * iy);
  if (in_slab(p, i, j, k)) {
    // keep the CPML pointer and profile loads inside the slab branch: the
    // compiler would otherwise hoist them into every wavefront (and spill
    // them to VGPR lanes), slab cells or not
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int l = slab_layer(p, face, i, j, k);
      if (l < 0) continue;
      const int ax = face >> 1;
      const unsigned q = (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
      const float bb = gld(p.pb[face], l), cc = gld(p.pc[face], l), km1 = gld(p.pk[face], l);
      const float inv = p.cinv[face];
      float g0, g1;
      if (ax == 0) {
        g0 = (ez_i - ez) * inv;  // Hy += ...
        g1 = (ey_i - ey) * inv;  // Hz -= ...
      } else if (ax == 1) {
        g0 = (ez_j - ez) * inv;  // Hx -= ...
        g1 = (ex_j - ex) * inv;  // Hz += ...
      } else {
        g0 = (ey_k - ey) * inv;  // Hx += ...
        g1 = (ex_k - ex) * inv;  // Hy -= ...
      }
      const float s0 = bb * ldf(ps.hr[face][0], q) + cc * g0;
      const float s1 = bb * ldf(ps.hr[face][1], q) + cc * g1;
      if (WRITE && wr) {
        stf(ps.hw[face][0], q, s0);
        stf(ps.hw[face][1], q, s1);
      }
      const float d0 = ch * s0 + ch * km1 * g0;
      const float d1 = ch * s1 + ch * km1 * g1;
      if (ax == 0) {
        hy += d0;
        hz -= d1;
      } else if (ax == 1) {
        hx -= d0;
        hz += d1;
      } else {
        hx += d0;
        hy -= d1;
      }
    }
  }
  return make_float3(hx, hy, hz);
}

// H at one cell reading E^n and H^{n-1/2} from Eo / Ho, storing into Hn.
template <bool GRADED, bool WRITE, bool ONE_MU = false>
__device__ __forceinline__ float3 h_cell(const KParams& p, const float* const Eo[3],
                                         const float* const Ho[3], float* const Hn[3],
                                         const PsiIO& ps, int i, int j, int k) {
  const unsigned P = (unsigned)p.plane;
  const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const unsigned m = ONE_MU ? 0u : ldu16(p.mat, c);
  // neighbour loads are issued unconditionally from an in-range index and
  // zeroed afterwards, so no load sits behind a branch
  const bool hj = j + 1 < p.ny, hk = k + 1 < p.nz, hi = i + 1 < p.nx;
  const unsigned cj = hj ? c + p.nz : c, ck = hk ? c + 1 : c, ci = hi ? c + P : c;
  const float ex = ldf(Eo[0], c), ey = ldf(Eo[1], c), ez = ldf(Eo[2], c);
  float ex_j = ldf(Eo[0], cj), ez_j = ldf(Eo[2], cj);
  float ex_k = ldf(Eo[0], ck), ey_k = ldf(Eo[1], ck);
  float ey_i = ldf(Eo[1], ci), ez_i = ldf(Eo[2], ci);
  const float hx = ldf(Ho[0], c), hy = ldf(Ho[1], c), hz = ldf(Ho[2], c);
  const float ch = ONE_MU ? p.ch_all : ld_ch(p.coef, m);
  ex_j = hj ? ex_j : 0.f;
  ez_j = hj ? ez_j : 0.f;
  ex_k = hk ? ex_k : 0.f;
  ey_k = hk ? ey_k : 0.f;
  ey_i = hi ? ey_i : 0.f;
  ez_i = hi ? ez_i : 0.f;
  const float3 h = h_core<GRADED, WRITE>(p, ps, i, j, k, ch, ex, ey, ez, ex_j, ez_j, ex_k, ey_k,
                                         ey_i, ez_i, hx, hy, hz);
  if (WRITE) {
    stf(Hn[0], c, h.x);
    stf(Hn[1], c, h.y);
    stf(Hn[2], c, h.z);
  }
  return h;
}

// E at one cell from values: E^{n+1} = ca E^n + cb curl_bwd H (yee.py:379-404),
// CPML-E (cpml.py:608-777), PEC faces and mask, point operations.  H^{n+1/2}
// at the cell and its backward neighbours (zero before the first cell) and
// E^n at the cell come in as values; returns E^{n+1}.
// Dispersive ADE at one cell, with E the plain update ca E^n + cb curl H
// (ca, cb of the material table are the dispersive ones):
//   Lorentz poles first, explicit in E^n (rfx/materials/lorentz.py:200-275):
//     P_l^{n+1} = a P_l^n + b P_l^{n-1} + c E^n,  dP = sum_l (P_l^{n+1} - P_l^n)
//   E^{n+1} = E + sum_d cc_d P_d^n - cc_l dP    (rfx/simulation.py:423-500)
//   Debye poles (rfx/materials/debye.py:182-250):
//     P_d^{n+1} = alpha P_d^n + beta (E^{n+1} + E^n)
// before the CPML correction, PEC and point operations.  A cell outside the
// state box is left alone (the host puts every cell with poles inside it).
__device__ __noinline__ void dispersion_cell(const KParams& p, const PsiIO& ps, int i, int j,
                                             int k, unsigned c, float o0, float o1, float o2,
                                             float& e0, float& e1, float& e2) {
  const int li = i - ps.pol_lo[0], lj = j - ps.pol_lo[1], lk = k - ps.pol_lo[2];
  if (!ps.pol || li < 0 || lj < 0 || lk < 0 || li >= ps.pol_n[0] || lj >= ps.pol_n[1] ||
      lk >= ps.pol_n[2])
    return;
  const unsigned R = (unsigned)ps.pol_n[0] * (unsigned)ps.pol_n[1] * (unsigned)ps.pol_n[2];
  const unsigned r = ((unsigned)li * (unsigned)ps.pol_n[1] + (unsigned)lj) * (unsigned)ps.pol_n[2] +
                     (unsigned)lk;
  const unsigned m = ldu16(p.mat, c);
  const float* pc = ps.pcoef + (size_t)m * ps.n_pole * 4;
  const int nd = ps.n_debye;
  float d0 = 0.f, d1 = 0.f, d2 = 0.f, ccl = 0.f;
  for (int q = nd; q < ps.n_pole; ++q) {
    const float a = pc[4 * q], b = pc[4 * q + 1], cl = pc[4 * q + 2];
    ccl = pc[4 * q + 3];
    float* P = ps.pol + (size_t)nd * 3 * R + (size_t)(q - nd) * 6 * R;
    const float p0 = P[r], p1 = P[R + r], p2 = P[2 * R + r];
    const float n0 = a * p0 + b * P[3 * R + r] + cl * o0;
    const float n1 = a * p1 + b * P[4 * R + r] + cl * o1;
    const float n2 = a * p2 + b * P[5 * R + r] + cl * o2;
    d0 += n0 - p0;
    d1 += n1 - p1;
    d2 += n2 - p2;
    P[3 * R + r] = p0;
    P[4 * R + r] = p1;
    P[5 * R + r] = p2;
    P[r] = n0;
    P[R + r] = n1;
    P[2 * R + r] = n2;
  }
  if (nd > 0) {
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int q = 0; q < nd; ++q) {
      const float cc = pc[4 * q + 2];
      const float* P = ps.pol + (size_t)q * 3 * R;
      s0 += cc * P[r];
      s1 += cc * P[R + r];
      s2 += cc * P[2 * R + r];
    }
    e0 += s0;
    e1 += s1;
    e2 += s2;
  }
  if (nd < ps.n_pole) {
    e0 -= ccl * d0;
    e1 -= ccl * d1;
    e2 -= ccl * d2;
  }
  for (int q = 0; q < nd; ++q) {
    const float a = pc[4 * q], b = pc[4 * q + 1];
    float* P = ps.pol + (size_t)q * 3 * R;
    P[r] = a * P[r] + b * (e0 + o0);
    P[R + r] = a * P[R + r] + b * (e1 + o1);
    P[2 * R + r] = a * P[2 * R + r] + b * (e2 + o2);
  }
}

template <bool GRADED, bool DISP = false>
__device__ __forceinline__ float3 e_core_co(const KParams& p, const PsiIO& ps, int i, int j, int k,
                                            unsigned c, float4 co, uint32_t fl, float e0, float e1,
                                            float e2, float hx, float hy, float hz, float hy_i,
                                            float hz_i, float hx_j, float hz_j, float hx_k,
                                            float hy_k) {
  float ix, iy, iz;
  if (GRADED) {
    ix = gld(p.inv_e[0], i);
    iy = gld(p.inv_e[1], j);
    iz = gld(p.inv_e[2], k);
  } else {
    ix = iy = iz = p.inv_e0;
  }
  const float ca = co.x, cb = co.y, ce = co.z;
  const float o0 = e0, o1 = e1, o2 = e2;
  e0 = ca * e0 + cb * ((hz - hz_j) * iy - (hy - hy_k) * iz);
  if (DISP && (fl & RFXH_ANISO)) {
    // per-component eps (subpixel smoothing): the row holds Ex's ca, cb
    const float4 a = gld(p.acoef, (int)ldu16(p.mat, c));
    e1 = a.x * e1 + a.y * ((hx - hx_k) * iz - (hz - hz_i) * ix);
    e2 = a.z * e2 + a.w * ((hy - hy_i) * ix - (hx - hx_j) * iy);
  } else {
    e1 = ca * e1 + cb * ((hx - hx_k) * iz - (hz - hz_i) * ix);
    e2 = ca * e2 + cb * ((hy - hy_i) * ix - (hx - hx_j) * iy);
  }
  if (DISP && (fl & RFXH_POLES)) dispersion_cell(p, ps, i, j, k, c, o0, o1, o2, e0, e1, e2);
  if (in_slab(p, i, j, k)) {
    // keep the CPML pointer and profile loads inside the slab branch: the
    // compiler would otherwise hoist them into every wavefront (and spill
    // them to VGPR lanes), slab cells or not
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int l = slab_layer(p, face, i, j, k);
      if (l < 0) continue;
      const int ax = face >> 1;
      const unsigned q = (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
      const float bb = gld(p.pb[face], l), cc = gld(p.pc[face], l), km1 = gld(p.pk[face], l);
      const float inv = p.cinv[face];
      float g0, g1;
      if (ax == 0) {
        g0 = (hz - hz_i) * inv;  // Ey -= ...
        g1 = (hy - hy_i) * inv;  // Ez += ...
      } else if (ax == 1) {
        g0 = (hz - hz_j) * inv;  // Ex += ...
        g1 = (hx - hx_j) * inv;  // Ez -= ...
      } else {
        g0 = (hy - hy_k) * inv;  // Ex -= ...
        g1 = (hx - hx_k) * inv;  // Ey += ...
      }
      const float s0 = bb * ldf(ps.er[face][0], q) + cc * g0;
      const float s1 = bb * ldf(ps.er[face][1], q) + cc * g1;
      stf(ps.ew[face][0], q, s0);
      stf(ps.ew[face][1], q, s1);
      const float d0 = ce * s0 + ce * km1 * g0;
      const float d1 = ce * s1 + ce * km1 * g1;
      if (ax == 0) {
        e1 -= d0;
        e2 += d1;
      } else if (ax == 1) {
        e0 += d0;
        e2 -= d1;
      } else {
        e0 -= d0;
        e1 += d1;
      }
    }
  }
  if (p.pec_axes & 1) {
    if (i == 0 || i == p.nx - 1) e1 = e2 = 0.f;
  }
  if (p.pec_axes & 2) {
    if (j == 0 || j == p.ny - 1) e0 = e2 = 0.f;
  }
  if (p.pec_axes & 4) {
    if (k == 0 || k == p.nz - 1) e0 = e1 = 0.f;
    if (k == p.nz - 1) e2 = 0.f;
  }
  if (fl) {
    if (fl & RFXH_ZERO_EX) e0 = 0.f;
    if (fl & RFXH_ZERO_EY) e1 = 0.f;
    if (fl & RFXH_ZERO_EZ) e2 = 0.f;
    if (fl & RFXH_POINT)
      point_ops(p, c, e0, e1, e2, hx, hy, hz, hx_j, hx_k, hy_i, hy_k, hz_i, hz_j);
  }
  return make_float3(e0, e1, e2);
}

template <bool GRADED, bool DISP = false>
__device__ __forceinline__ float3 e_core(const KParams& p, const PsiIO& ps, int i, int j, int k,
                                         unsigned c, unsigned m, float e0, float e1, float e2,
                                         float hx, float hy, float hz, float hy_i, float hz_i,
                                         float hx_j, float hz_j, float hx_k, float hy_k) {
  // one 16-byte row carries everything the E update needs
  const float4 ec = gld(p.ecoef, (int)m);
  return e_core_co<GRADED, DISP>(p, ps, i, j, k, c, ec, __float_as_uint(ec.w), e0, e1, e2, hx, hy, hz,
                           hy_i, hz_i, hx_j, hz_j, hx_k, hy_k);
}

// E at one cell reading E^n from Eo, storing E^{n+1} into En.
template <bool GRADED, bool DISP = false>
__device__ __forceinline__ void e_cell(const KParams& p, const float* const Eo[3],
                                       float* const En[3], const PsiIO& ps, int i, int j, int k,
                                       float hx, float hy, float hz, float hy_i, float hz_i,
                                       float hx_j, float hz_j, float hx_k, float hy_k) {
  const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const unsigned m = ldu16(p.mat, c);
  const float e0 = ldf(Eo[0], c), e1 = ldf(Eo[1], c), e2 = ldf(Eo[2], c);
  const float3 e = e_core<GRADED, DISP>(p, ps, i, j, k, c, m, e0, e1, e2, hx, hy, hz, hy_i, hz_i,
                                        hx_j, hz_j, hx_k, hy_k);
  stf(En[0], c, e.x);
  stf(En[1], c, e.y);
  stf(En[2], c, e.z);
}

// ---------------------------------------------------------------------------
// Fused step: read set A (E^n, H^{n-1/2}, psi), write set B.
// ---------------------------------------------------------------------------
template <bool GRADED>
__global__ __launch_bounds__(kWG) void k_step_fused(KParams p, FusedIO io, PsiIO ps) {
  int i, j, k;
  if (!cell_of(p, i, j, k)) return;
  const float* const Eo[3] = {io.eo[0], io.eo[1], io.eo[2]};
  const float* const Ho[3] = {io.ho[0], io.ho[1], io.ho[2]};
  float* const Hn[3] = {io.hn[0], io.hn[1], io.hn[2]};
  float* const En[3] = {io.en[0], io.en[1], io.en[2]};
  const float3 h = h_cell<GRADED, true>(p, Eo, Ho, Hn, ps, i, j, k);
  float hy_i = 0.f, hz_i = 0.f, hx_j = 0.f, hz_j = 0.f, hx_k = 0.f, hy_k = 0.f;
  if (i > 0) {
    const float3 t = h_cell<GRADED, false>(p, Eo, Ho, Hn, ps, i - 1, j, k);
    hy_i = t.y;
    hz_i = t.z;
  }
  if (j > 0) {
    const float3 t = h_cell<GRADED, false>(p, Eo, Ho, Hn, ps, i, j - 1, k);
    hx_j = t.x;
    hz_j = t.z;
  }
  if (k > 0) {
    const float3 t = h_cell<GRADED, false>(p, Eo, Ho, Hn, ps, i, j, k - 1);
    hx_k = t.x;
    hy_k = t.y;
  }
  e_cell<GRADED>(p, Eo, En, ps, i, j, k, h.x, h.y, h.z, hy_i, hz_i, hx_j, hz_j, hx_k, hy_k);
}

// ---------------------------------------------------------------------------
// Marching fused step: read set A (E^n, H^{n-1/2}, psi), write set B.
// A workgroup owns a tile of MJ rows (y) x 64 columns (z) of the y-z plane
// and walks along x through its segment of planes.  Waves 1..MJ own the
// tile's rows; wave 0 computes H on the row j0-1 and lanes 0..MJ-1 of the
// last wave compute H on the column k0-1 (halo values, never stored), so the
// E update finds every backward H neighbour in LDS (double-buffered, one
// barrier per plane).  Each thread keeps its own column's E of the next
// plane and H of the previous plane in registers, so x-neighbours are never
// re-read, and the inputs of plane i+1 are loaded while plane i is being
// computed.  Tiles are 64 cells wide in z, so every row of every field is
// read and written in whole cache lines.
// ---------------------------------------------------------------------------
// Tile heights (interior rows): 6 (8 waves, two workgroups per CU at the
// kernel's register use) or 14 (16 waves, one 1024-thread workgroup).
constexpr int kMarchRows[2] = {6, 14};

struct MarchIO {
  const float* eo[3];
  const float* ho[3];
  float* en[3];
  float* hn[3];
  int ntk, ntj, seg_len;
};

// Inputs of one plane for one thread's cell.
struct MarchIn {
  float ex_j, ez_j, ex_k, ey_k;  // E at (j+1) and (k+1), zero past the grid
  float hx, hy, hz;              // H^{n-1/2} at the cell
  unsigned m;                    // material id
};

__device__ __forceinline__ MarchIn march_load(const KParams& p, const MarchIO& io, unsigned c,
                                              unsigned dj, unsigned dk, bool hj, bool hk) {
  MarchIn r;
  r.m = ldu16(p.mat, c);
  r.ex_j = ldf(io.eo[0], c + dj);
  r.ez_j = ldf(io.eo[2], c + dj);
  r.ex_k = ldf(io.eo[0], c + dk);
  r.ey_k = ldf(io.eo[1], c + dk);
  r.hx = ldf(io.ho[0], c);
  r.hy = ldf(io.ho[1], c);
  r.hz = ldf(io.ho[2], c);
  r.ex_j = hj ? r.ex_j : 0.f;
  r.ez_j = hj ? r.ez_j : 0.f;
  r.ex_k = hk ? r.ex_k : 0.f;
  r.ey_k = hk ? r.ey_k : 0.f;
  return r;
}

template <bool GRADED, int MJ>
__global__ __launch_bounds__(kBX * (MJ + 2)) void k_step_march(KParams p, MarchIO io, PsiIO ps) {
  constexpr int kMJ = MJ;
  __shared__ float lhx[2][kMJ + 1][kBX + 1];
  __shared__ float lhy[2][kMJ + 1][kBX + 1];
  __shared__ float lhz[2][kMJ + 1][kBX + 1];
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int tk = L % io.ntk;
  const int r = L / io.ntk;
  const int tj = r % io.ntj;
  const int seg = r / io.ntj;
  const int w = (int)threadIdx.y, l = (int)threadIdx.x;
  // role: rows (w <= kMJ) or the halo column (last wave, lanes < kMJ)
  const bool rows = w <= kMJ;
  const bool active = rows || l < kMJ;
  const int j = rows ? tj * kMJ + w - 1 : tj * kMJ + l;
  const int k = rows ? tk * kBX + l : tk * kBX - 1;
  const int srow = rows ? w : l + 1, scol = rows ? l + 1 : 0;  // LDS slot
  const int i0 = seg * io.seg_len;
  const int i1 = min(p.nx, i0 + io.seg_len);
  const bool valid = active && j >= 0 && j < p.ny && k >= 0 && k < p.nz;
  const bool inner = valid && rows && w > 0;
  const unsigned P = (unsigned)p.plane;
  const bool hj = j + 1 < p.ny, hk = k + 1 < p.nz;
  const unsigned c0 = valid ? (unsigned)j * (unsigned)p.nz + (unsigned)k : 0u;
  const unsigned dj = hj ? (unsigned)p.nz : 0u, dk = hk ? 1u : 0u;
  float e0 = 0.f, e1 = 0.f, e2 = 0.f;  // E^n of plane i, own cell
  float n0 = 0.f, n1 = 0.f, n2 = 0.f;  // E^n of plane i+1
  float hyp = 0.f, hzp = 0.f;          // H^{n+1/2} of plane i-1
  MarchIn cur;
  cur.ex_j = cur.ez_j = cur.ex_k = cur.ey_k = cur.hx = cur.hy = cur.hz = 0.f;
  cur.m = 0;
  if (valid && i0 < i1) {
    const unsigned c = (unsigned)i0 * P + c0;
    e0 = ldf(io.eo[0], c);
    e1 = ldf(io.eo[1], c);
    e2 = ldf(io.eo[2], c);
    if (i0 + 1 < p.nx) {
      n0 = ldf(io.eo[0], c + P);
      n1 = ldf(io.eo[1], c + P);
      n2 = ldf(io.eo[2], c + P);
    }
    cur = march_load(p, io, c, dj, dk, hj, hk);
    if (i0 > 0) {
      // H of plane i0-1 (owned by the previous segment): recomputed, not stored
      const unsigned cm = c - P;
      const MarchIn pm = march_load(p, io, cm, dj, dk, hj, hk);
      const float3 h = h_core<GRADED, false>(
          p, ps, i0 - 1, j, k, p.coef[pm.m].w, ldf(io.eo[0], cm), ldf(io.eo[1], cm),
          ldf(io.eo[2], cm), pm.ex_j, pm.ez_j, pm.ex_k, pm.ey_k, e1, e2, pm.hx, pm.hy, pm.hz);
      hyp = h.y;
      hzp = h.z;
    }
  }
  for (int i = i0; i < i1; ++i) {
    const int b = i & 1;
    const unsigned c = (unsigned)i * P + c0;
    // inputs of plane i+1, loaded while plane i is computed
    MarchIn nxt = cur;
    float q0 = 0.f, q1 = 0.f, q2 = 0.f;  // E^n of plane i+2
    if (valid && i + 1 < i1) {
      nxt = march_load(p, io, c + P, dj, dk, hj, hk);
      if (i + 2 < p.nx) {
        q0 = ldf(io.eo[0], c + 2 * P);
        q1 = ldf(io.eo[1], c + 2 * P);
        q2 = ldf(io.eo[2], c + 2 * P);
      }
    }
    float3 h = make_float3(0.f, 0.f, 0.f);
    if (valid)
      h = h_core<GRADED, true>(p, ps, i, j, k, p.coef[cur.m].w, e0, e1, e2, cur.ex_j, cur.ez_j,
                               cur.ex_k, cur.ey_k, n1, n2, cur.hx, cur.hy, cur.hz, inner);
    if (active) {
      lhx[b][srow][scol] = h.x;
      lhy[b][srow][scol] = h.y;
      lhz[b][srow][scol] = h.z;
    }
    __syncthreads();
    if (inner) {
      const float hx_j = lhx[b][w - 1][l + 1], hz_j = lhz[b][w - 1][l + 1];
      const float hx_k = lhx[b][w][l], hy_k = lhy[b][w][l];
      const float3 e = e_core<GRADED>(p, ps, i, j, k, c, cur.m, e0, e1, e2, h.x, h.y, h.z, hyp,
                                      hzp, hx_j, hz_j, hx_k, hy_k);
      stf(io.hn[0], c, h.x);
      stf(io.hn[1], c, h.y);
      stf(io.hn[2], c, h.z);
      stf(io.en[0], c, e.x);
      stf(io.en[1], c, e.y);
      stf(io.en[2], c, e.z);
    }
    hyp = h.y;
    hzp = h.z;
    e0 = n0;
    e1 = n1;
    e2 = n2;
    n0 = q0;
    n1 = q1;
    n2 = q2;
    cur = nxt;
  }
}

// ---------------------------------------------------------------------------
// Wave-march step: read set A, write set B, no barriers in the march.  Each
// wavefront owns two rows (y = j, j+1) of a run of 63 cells (z) plus a halo
// lane at k0-1 and walks along x through its segment of planes; the
// wavefronts of a workgroup take consecutive z runs of one row pair.  Per
// plane a wavefront computes H on its two rows and, again, on the row below
// (j-1, the E update's backward-y neighbour of row j; row j+1 has row j),
// takes the backward-z neighbour H from the next lower lane (cross-lane
// shuffle) and the backward-x neighbour from its registers, then updates E
// on both rows.  Waves never wait for each other.
//
// Software pipeline: every global load a plane needs is issued one plane
// ahead (E two planes ahead, since the H update of plane i reads E of plane
// i+1), and the material coefficients come from a table staged in LDS, so
// no load is consumed in the iteration that issues it.  Vector-memory
// loads complete in order, so a load issued and waited for inside the
// compute phase would also wait for every prefetch issued before it.
// CPML memory and point operations still load inline; only the wavefronts
// whose cells have them pay that wait.
// ---------------------------------------------------------------------------
constexpr int kWMW = 4;              // independent wavefronts per workgroup
constexpr int kWMK = kBX - 1;        // interior columns per wave
constexpr int kWMMat = 1024;         // material table entries staged in LDS (the variant's limit)

struct WMarchIO {
  const float* eo[3];
  const float* ho[3];
  float* en[3];
  float* hn[3];
  int ntk, ntj, seg_len;   // workgroups along z (kWMW runs each), row pairs, planes per segment
  int n_mat;
};

// The values of one plane a wavefront needs: E of the two own rows (e) and
// of the row below (eb), old H of the own rows (h) and of the row below (x,
// z: hb), material ids, forward-y E of row j+2 (ey2: x, z) and lane 63's
// forward-z E (x, y) of rows j, j+1 and j-1 (ek).
struct WMPlane {
  float e[2][3], eb[3], h[2][3], hb[2], ey2[2], ek[3][2];
  unsigned m[2], mb;
};

template <bool GRADED>
__global__ __launch_bounds__(kBX * kWMW) void k_step_wmarch(KParams p, WMarchIO io,
                                                            const PsiIO* __restrict__ psp) {
  // CPML pointers through device memory: as kernel arguments they are
  // hoisted out of the march loop and spilled to VGPR lanes
  const PsiIO& ps = *psp;
  __shared__ float4 s_coef[kWMMat];
  __shared__ uint32_t s_flag[kWMMat];
  // the plan rejects this variant for more than kWMMat materials
  for (int t = (int)(threadIdx.y * kBX + threadIdx.x); t < io.n_mat; t += kBX * kWMW) {
    s_coef[t] = gld(p.coef, t);
    s_flag[t] = gld(p.flags, t);
  }
  __syncthreads();
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int tk = (L % io.ntk) * kWMW + (int)threadIdx.y;
  const int r = L / io.ntk;
  const int j = 2 * (r % io.ntj);
  const int seg = r / io.ntj;
  const int l = (int)threadIdx.x;
  const int k = tk * kWMK + l - 1;
  const int i0 = seg * io.seg_len;
  const int i1 = min(p.nx, i0 + io.seg_len);
  const bool kin = k >= 0 && k < p.nz;
  const bool v0 = kin && j < p.ny, v1 = kin && j + 1 < p.ny;   // own rows exist
  const bool vb = v0 && j > 0;                                 // the row below exists
  const bool h2 = j + 2 < p.ny;                                // row j+2 exists
  const bool inner = l > 0;
  const bool last = l == kBX - 1;               // loads its forward-z E itself
  const unsigned P = (unsigned)p.plane;
  const bool hk = k + 1 < p.nz;
  const unsigned Z = (unsigned)p.nz;
  // in-plane offsets (clamped to a valid cell where a lane has none; such
  // values are never used)
  const unsigned c0 = v0 ? (unsigned)j * Z + (unsigned)k : 0u;
  const unsigned c1 = v1 ? c0 + Z : c0;
  const unsigned cb = vb ? c0 - Z : c0;
  const unsigned c2 = (v1 && h2) ? c0 + 2 * Z : c0;
  const unsigned dk = hk ? 1u : 0u;
  if (i0 >= i1 || tk * kWMK - 1 >= p.nz || j >= p.ny) return;
  auto plane_off = [&](int i) { return (unsigned)min(i, p.nx - 1) * P; };
  // all of a plane's loads except E (E comes two planes ahead, see below)
  auto load_rest = [&](WMPlane& w, unsigned o) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      w.h[0][c] = ldf(io.ho[c], o + c0);
      w.h[1][c] = ldf(io.ho[c], o + c1);
    }
    w.hb[0] = ldf(io.ho[0], o + cb);
    w.hb[1] = ldf(io.ho[2], o + cb);
    w.m[0] = ldu16(p.mat, o + c0);
    w.m[1] = ldu16(p.mat, o + c1);
    w.mb = ldu16(p.mat, o + cb);
    w.ey2[0] = ldf(io.eo[0], o + c2);
    w.ey2[1] = ldf(io.eo[2], o + c2);
    if (last) {
      w.ek[0][0] = ldf(io.eo[0], o + c0 + dk);
      w.ek[0][1] = ldf(io.eo[1], o + c0 + dk);
      w.ek[1][0] = ldf(io.eo[0], o + c1 + dk);
      w.ek[1][1] = ldf(io.eo[1], o + c1 + dk);
      w.ek[2][0] = ldf(io.eo[0], o + cb + dk);
      w.ek[2][1] = ldf(io.eo[1], o + cb + dk);
    }
  };
  auto load_e = [&](float (&e)[2][3], float (&eb)[3], unsigned o) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      e[0][c] = ldf(io.eo[c], o + c0);
      e[1][c] = ldf(io.eo[c], o + c1);
      eb[c] = ldf(io.eo[c], o + cb);
    }
  };
  WMPlane w;            // plane i
  float en[2][3], enb[3];   // E of plane i+1
#pragma unroll
  for (int q = 0; q < 3; ++q) w.ek[q][0] = w.ek[q][1] = 0.f;
  load_e(w.e, w.eb, plane_off(i0));
  load_e(en, enb, plane_off(i0 + 1));
  load_rest(w, plane_off(i0));
  // own-row H of plane i-1 (y, z), for the backward-x neighbours
  float hyp[2] = {0.f, 0.f}, hzp[2] = {0.f, 0.f};
  if (i0 > 0) {
    // owned by the previous segment: recomputed
    const unsigned o = (unsigned)(i0 - 1) * P;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      // a lane whose cell lies outside the grid (k = -1, k >= nz, the row
      // past ny) has no H to recompute: its CPML index would leave the slab
      if (!(q ? v1 : v0)) continue;
      const unsigned cq = q ? c1 : c0;
      const unsigned cy = q ? c2 : c1;     // forward-y neighbour row
      const bool hy = q ? h2 : j + 1 < p.ny;
      const unsigned mm = ldu16(p.mat, o + cq);
      float ex_j = ldf(io.eo[0], o + cy), ez_j = ldf(io.eo[2], o + cy);
      float ex_k = ldf(io.eo[0], o + cq + dk), ey_k = ldf(io.eo[1], o + cq + dk);
      const float3 h = h_core<GRADED, false>(
          p, ps, i0 - 1, j + q, k, s_coef[mm].w, ldf(io.eo[0], o + cq), ldf(io.eo[1], o + cq),
          ldf(io.eo[2], o + cq), hy ? ex_j : 0.f, hy ? ez_j : 0.f, hk ? ex_k : 0.f,
          hk ? ey_k : 0.f, w.e[q][1], w.e[q][2], ldf(io.ho[0], o + cq), ldf(io.ho[1], o + cq),
          ldf(io.ho[2], o + cq));
      hyp[q] = h.y;
      hzp[q] = h.z;
    }
  }
  for (int i = i0; i < i1; ++i) {
    // ---- prefetch: E of plane i+2, everything else of plane i+1
    WMPlane nw;
    float nn[2][3], nnb[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) nw.ek[q][0] = nw.ek[q][1] = 0.f;
    load_e(nn, nnb, plane_off(i + 2));
    load_rest(nw, plane_off(i + 1));
    // ---- compute plane i from the values loaded one iteration earlier
    const bool hi = i + 1 < p.nx;
    // forward-z E from lane l+1 (lane 63 has loaded its own)
    float fk[3][2];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float* src = q < 2 ? w.e[q] : w.eb;
      fk[q][0] = last ? w.ek[q][0] : __shfl_down(src[0], 1);
      fk[q][1] = last ? w.ek[q][1] : __shfl_down(src[1], 1);
      fk[q][0] = hk ? fk[q][0] : 0.f;
      fk[q][1] = hk ? fk[q][1] : 0.f;
    }
    float3 hm = make_float3(0.f, 0.f, 0.f);
    if (vb) {
      // row below: forward-y neighbour is row j; its old Hy is not needed
      hm = h_core<GRADED, false>(p, ps, i, j - 1, k, s_coef[w.mb].w, w.eb[0], w.eb[1], w.eb[2],
                                 w.e[0][0], w.e[0][2], fk[2][0], fk[2][1], hi ? enb[1] : 0.f,
                                 hi ? enb[2] : 0.f, w.hb[0], 0.f, w.hb[1]);
    }
    float3 h[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const bool vq = q ? v1 : v0;
      const bool hy = q ? h2 : v1;                       // forward-y row exists
      const float eyx = q ? w.ey2[0] : w.e[1][0], eyz = q ? w.ey2[1] : w.e[1][2];
      h[q] = make_float3(0.f, 0.f, 0.f);
      if (vq)
        h[q] = h_core<GRADED, true>(p, ps, i, j + q, k, s_coef[w.m[q]].w, w.e[q][0], w.e[q][1],
                                    w.e[q][2], hy ? eyx : 0.f, hy ? eyz : 0.f, fk[q][0], fk[q][1],
                                    hi ? en[q][1] : 0.f, hi ? en[q][2] : 0.f, w.h[q][0],
                                    w.h[q][1], w.h[q][2], inner);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const float hx_k = __shfl_up(h[q].x, 1), hy_k = __shfl_up(h[q].y, 1);
      const bool vq = q ? v1 : v0;
      if (vq && inner) {
        const unsigned c = (unsigned)i * P + (q ? c1 : c0);
        const float3 hj = q ? h[0] : hm;                 // backward-y neighbour
        const float3 e = e_core_co<GRADED>(p, ps, i, j + q, k, c, s_coef[w.m[q]], s_flag[w.m[q]],
                                           w.e[q][0], w.e[q][1], w.e[q][2], h[q].x, h[q].y,
                                           h[q].z, hyp[q], hzp[q], hj.x, hj.z, hx_k, hy_k);
        stf(io.hn[0], c, h[q].x);
        stf(io.hn[1], c, h[q].y);
        stf(io.hn[2], c, h[q].z);
        stf(io.en[0], c, e.x);
        stf(io.en[1], c, e.y);
        stf(io.en[2], c, e.z);
      }
      hyp[q] = h[q].y;
      hzp[q] = h[q].z;
    }
    // ---- rotate the pipeline
    w = nw;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      w.e[0][c] = en[0][c];
      w.e[1][c] = en[1][c];
      w.eb[c] = enb[c];
      en[0][c] = nn[0][c];
      en[1][c] = nn[1][c];
      enb[c] = nnb[c];
    }
  }
}

// ---------------------------------------------------------------------------
// Temporal-blocked march: two time steps per sweep.  Reads set A (E^n,
// H^{n-1/2}, CPML memory), writes set B (E^{n+2}, H^{n+3/2}, memory).
//
// A workgroup of kTBW wavefronts covers a tile of kTBW rows (y) x 64
// columns (z) of the y-z plane, one (y, z) column per thread, and walks
// along x through its segment of planes.  The kTBHalo = 2 outer rows and
// columns on each side are halo: their threads repeat the work of the
// neighbouring tiles' threads, so that the kTBJ x kTBK interior can take
// both steps without waiting for another workgroup (only the interior
// stores).  Along x the four half-steps form a pipeline: the iteration
// that loads E^n of plane p updates
//   D: E^{n+2}   on plane p-4   (H^{n+3/2} of p-4, p-5, E^{n+1} of p-4)
//   B: E^{n+1}   on plane p-2   (H^{n+1/2} of p-2, p-3, E^n of p-2)
//   C: H^{n+3/2} on plane p-3   (E^{n+1} of p-3, p-2, H^{n+1/2} of p-3)
//   A: H^{n+1/2} on plane p-1   (E^n of p-1, p, H^{n-1/2} of p-1)
// in that order.  Neighbours along x are the thread's own registers; along
// y and z they come from the other threads through LDS, each stage reading
// what the previous iteration wrote (two buffers, one barrier per plane).
// The CPML memory of a cell stays in registers between the two steps (a
// cell lies in at most one slab per axis, two terms each).
//
// Per two cell-steps a cell's E, H and material id are read once and E, H
// written once: 50 B, against 2 x (74-80) B for two steps of the two-pass
// passes.  Halo cells are read by two or four tiles (through the L2 when
// the tiles run together).  Point operations of the halo copies inject but
// record nothing; DFT planes are staged into their ring by the interior
// thread of the plane cell (k_dft_fold2 folds them).
// ---------------------------------------------------------------------------
constexpr int kTBW = 16;                    // wavefronts (rows) per workgroup
constexpr int kTBHalo = 2;                  // halo width: the steps per sweep
constexpr int kTBJ = kTBW - 2 * kTBHalo;    // interior rows of a tile
constexpr int kTBK = kBX - 2 * kTBHalo;     // interior columns of a tile
constexpr int kTBMat = 1024;                // material table entries staged in LDS
constexpr int kTBPlanes = 4;                // DFT planes the sweep stages itself
constexpr int kTBProf = 256;                // CPML profile rows staged in LDS (all faces)
constexpr int kTBOps = 256;                 // point operations one tile column may hold
// z-edge tiles: 64 rows x 16 columns (four rows per wavefront), their inner
// 12 columns over a z face's CPML slab, so that the 16 x 64 tiles between
// them need no z-face CPML code
constexpr int kTBEdgeC = 16;
constexpr int kTBEdgeR = kTBW * kBX / kTBEdgeC;

// A point operation of the workgroup's region, staged in LDS at the start
// of the sweep with the source values of both steps, so that the sweep
// issues no global loads besides its plane prefetch and CPML memory (a
// load in a rarely taken branch would make the compiler wait for every
// load in flight where the branches join).
struct TBOp {
  unsigned cell;
  int kind, comp, slot;
  float d_par;
  float loop[3];
  float w[2];      // inject: source values of steps n and n + 1
};

struct TBPlane {
  float* snap;     // ring [kDftRing][npts]
  int axis, index, comp, n2, npts;
};

// CPML memory and profiles of the sweep, addressed from one base pointer
// per field set (a few scalar registers instead of one pointer per face
// and term): face f's four terms (H0, H1, E0, E1) lie at psi + foff[f] +
// t * fs[f], in the slab order of psi_idx; profile row l of face f is
// prof[poff[f] + l] = (b, c, 1/kappa - 1, 0).  The memory after the first
// step goes through `mid` rather than registers: the H terms are written by
// every copy of a cell (the same value), the E terms by the cell's owner.
struct TBCpml {
  const float* psr;      // read set
  float* psw;            // write set
  float* mid;            // memory after the first of the two steps (same layout)
  const float4* prof;
  unsigned foff[6];
  unsigned fs[6];
  int poff[6];
};

// x segments: [0, xa) holds the x-lo CPML slab (with the two planes past
// it that a segment's first updates reach back to), [xb, nx) the x-hi slab;
// [xa, xb) is cut into n_mid segments of seg_len planes.  A workgroup whose
// planes and tile rows and columns (with their halos) miss every slab runs
// the sweep without CPML code.
struct TBIO {
  const float* eo[3];
  const float* ho[3];
  float* en[3];
  float* hn[3];
  int ntk, ntj, n_mat;
  int kz0, kz1;          // the 16 x 64 tiles own columns [kz0, kz1)
  int ntj_e;             // rows of z-edge tiles per z face (0: none)
  int xa, xb, n_mid, seg_len;
  int n_prof;            // CPML profile rows (all faces)
  int n_plane;
  TBPlane pl[kTBPlanes];
  TBCpml cm;
};

struct TBShared {
  float4 ecoef[kTBMat];   // ca, cb, ce, flag bits (as float bits): the E update's row
  float ch[kTBMat];
  // CPML profile rows: a global load inside an update would wait for every
  // plane prefetch issued before it (vector loads complete in order)
  float4 prof[kTBProf];
  // [buffer][stage: E^n, H^{n+1/2}, E^{n+1}, H^{n+3/2}][component][row][lane]
  float x[2][4][3][kTBW * kBX];   // [row * tile columns + column]
  TBOp ops[kTBOps];   // the region's point operations, sorted by (cell, order)
  int n_ops;
  int wcnt[kTBW];
};

// Stable compaction of the point operations whose cells lie in planes
// [qa, qb), rows [ja, jb) and columns [ka, kb) into sh.ops (the plan checks
// that no tile column holds more than kTBOps).  All threads of the
// workgroup call it.
__device__ __forceinline__ void tb_stage_ops(const KParams& p, TBShared& sh, int w, int l,
                                             long long n, int qa, int qb, int ja, int jb,
                                             int ka, int kb) {
  const int tid = w * kBX + l;
  if (tid == 0) sh.n_ops = 0;
  __syncthreads();
  for (int base = 0; base < p.n_ops; base += kBX * kTBW) {
    const int t = base + tid;
    bool in = false;
    rfxh_point_op op;
    if (t < p.n_ops) {
      op = p.ops[t];
      const long long c = op.cell;
      const int i = (int)(c / p.plane);
      const int r = (int)(c - (long long)i * p.plane);
      const int j = r / p.nz, k = r - (r / p.nz) * p.nz;
      in = i >= qa && i < qb && j >= ja && j < jb && k >= ka && k < kb;
    }
    const unsigned long long bal = __ballot(in);
    const int rank = __popcll(bal & ((1ull << l) - 1ull));
    if (l == 0) sh.wcnt[w] = __popcll(bal);
    __syncthreads();
    int off = sh.n_ops;
    for (int u = 0; u < w; ++u) off += sh.wcnt[u];
    if (in && off + rank < kTBOps) {
      TBOp& d = sh.ops[off + rank];
      d.cell = (unsigned)op.cell;
      d.kind = op.kind;
      d.comp = op.comp;
      d.slot = op.slot;
      d.d_par = op.d_par;
      d.loop[0] = op.loop[0];
      d.loop[1] = op.loop[1];
      d.loop[2] = op.loop[2];
      d.w[0] = d.w[1] = 0.f;
      if (op.kind == RFXH_OP_INJECT) {
        d.w[0] = p.src_wave[n * p.n_src + op.slot];
        d.w[1] = p.src_wave[(n + 1) * p.n_src + op.slot];
      }
    }
    __syncthreads();
    if (tid == 0) {
      int tot = sh.n_ops;
      for (int u = 0; u < kTBW; ++u) tot += sh.wcnt[u];
      sh.n_ops = min(tot, kTBOps);
    }
    __syncthreads();
  }
}

// point_ops from the staged list: step n + s (s = 0, 1).
__device__ __forceinline__ void tb_point_ops(const KParams& p, const TBShared& sh, unsigned cell,
                                             float& e0, float& e1, float& e2, float hx, float hy,
                                             float hz, float hx_j, float hx_k, float hy_i,
                                             float hy_k, float hz_i, float hz_j, long long n,
                                             int s, bool own) {
  int lo = 0, hi = sh.n_ops;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sh.ops[mid].cell < cell) lo = mid + 1;
    else hi = mid;
  }
  const long long ns = n + s;
  for (int q = lo; q < sh.n_ops && sh.ops[q].cell == cell; ++q) {
    const TBOp& op = sh.ops[q];
    if (op.kind == RFXH_OP_PORT) {
      if (!own) continue;
      const float v = -pick3(op.comp, e0, e1, e2) * op.d_par;
      float cur;
      if (op.comp == 2)
        cur = (hy - hy_i) * op.loop[1] - (hx - hx_j) * op.loop[0];
      else if (op.comp == 0)
        cur = (hz - hz_j) * op.loop[2] - (hy - hy_k) * op.loop[1];
      else
        cur = (hx - hx_k) * op.loop[0] - (hz - hz_i) * op.loop[2];
      float* o = p.port_out + (ns * p.n_port + op.slot) * 2;
      o[0] = v;
      o[1] = cur;
    } else if (op.kind == RFXH_OP_INJECT) {
      const float w = s ? op.w[1] : op.w[0];
      if (op.comp == 0) e0 += w;
      else if (op.comp == 1) e1 += w;
      else e2 += w;
    } else if (own) {
      p.probe_out[ns * p.n_probe + op.slot] =
          op.comp < 3 ? pick3(op.comp, e0, e1, e2) : pick3(op.comp - 3, hx, hy, hz);
    }
  }
}

// H at one cell with the CPML memory in registers: h_core's update reading
// the memory from pin and leaving the new memory in pout (and, when store,
// in cm.psw).
template <bool GRADED, bool CPML = true>
__device__ __forceinline__ float3 tb_h(const KParams& p, const TBCpml& cm, const float4* prof,
                                       int i, int j, int k, float gy, float gz,
                                       float ch, float ex, float ey, float ez, float ex_j,
                                       float ez_j, float ex_k, float ey_k, float ey_i, float ez_i,
                                       float hx, float hy, float hz, const float (&pin)[3][2],
                                       float (&pout)[3][2], bool store) {
  // graded: gy, gz are the thread's own inverse spacings along y and z
  // (loaded once), the x one is wave-uniform; the pipeline also visits
  // planes outside the grid (results discarded): clamp the index
  float ix, iy, iz;
  if (GRADED) {
    ix = gld(p.inv_h[0], min(max(i, 0), p.nx - 1));
    iy = gy;
    iz = gz;
  } else {
    ix = iy = iz = p.inv_h0;
  }
  hx = hx - ch * ((ez_j - ez) * iy - (ey_k - ey) * iz);
  hy = hy - ch * ((ex_k - ex) * iz - (ez_i - ez) * ix);
  hz = hz - ch * ((ey_i - ey) * ix - (ex_j - ex) * iy);
  if (CPML && in_slab(p, i, j, k)) {
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int l = slab_layer(p, face, i, j, k);
      if (l < 0) continue;
      const int ax = face >> 1;
      const unsigned q = cm.foff[face] + (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
      const unsigned ts = cm.fs[face];
      const float4 pr = prof[cm.poff[face] + l];
      const float inv = p.cinv[face];
      const float bb = pr.x, cc = pr.y, km1 = pr.z;
      float g0, g1;
      if (ax == 0) {
        g0 = (ez_i - ez) * inv;
        g1 = (ey_i - ey) * inv;
      } else if (ax == 1) {
        g0 = (ez_j - ez) * inv;
        g1 = (ex_j - ex) * inv;
      } else {
        g0 = (ey_k - ey) * inv;
        g1 = (ex_k - ex) * inv;
      }
      const float s0 = bb * pin[ax][0] + cc * g0;
      const float s1 = bb * pin[ax][1] + cc * g1;
      pout[ax][0] = s0;
      pout[ax][1] = s1;
      if (store) {
        stf(cm.psw, q, s0);
        stf(cm.psw, q + ts, s1);
      }
      const float d0 = ch * s0 + ch * km1 * g0;
      const float d1 = ch * s1 + ch * km1 * g1;
      if (ax == 0) {
        hy += d0;
        hz -= d1;
      } else if (ax == 1) {
        hx -= d0;
        hz += d1;
      } else {
        hx += d0;
        hy -= d1;
      }
    }
  }
  return make_float3(hx, hy, hz);
}

// E at one cell with the CPML memory in registers (e_core_co's update),
// point operations at step n (recorded only when own).
template <bool GRADED, bool CPML = true>
__device__ __forceinline__ float3 tb_e(const KParams& p, const TBCpml& cm, const float4* prof,
                                       int i, int j, int k, float gy, float gz, unsigned c,
                                       float4 co, uint32_t fl, float e0, float e1,
                                       float e2, float hx, float hy, float hz, float hy_i,
                                       float hz_i, float hx_j, float hz_j, float hx_k, float hy_k,
                                       const float (&pin)[3][2], float (&pout)[3][2], bool store,
                                       const TBShared& sh, long long n, int s, bool own) {
  float ix, iy, iz;
  if (GRADED) {
    ix = gld(p.inv_e[0], min(max(i, 0), p.nx - 1));
    iy = gy;
    iz = gz;
  } else {
    ix = iy = iz = p.inv_e0;
  }
  const float ca = co.x, cb = co.y, ce = co.z;
  e0 = ca * e0 + cb * ((hz - hz_j) * iy - (hy - hy_k) * iz);
  e1 = ca * e1 + cb * ((hx - hx_k) * iz - (hz - hz_i) * ix);
  e2 = ca * e2 + cb * ((hy - hy_i) * ix - (hx - hx_j) * iy);
  if (CPML && in_slab(p, i, j, k)) {
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int l = slab_layer(p, face, i, j, k);
      if (l < 0) continue;
      const int ax = face >> 1;
      const unsigned q = cm.foff[face] + (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
      const unsigned ts = cm.fs[face];
      const float4 pr = prof[cm.poff[face] + l];
      const float inv = p.cinv[face];
      const float bb = pr.x, cc = pr.y, km1 = pr.z;
      float g0, g1;
      if (ax == 0) {
        g0 = (hz - hz_i) * inv;
        g1 = (hy - hy_i) * inv;
      } else if (ax == 1) {
        g0 = (hz - hz_j) * inv;
        g1 = (hx - hx_j) * inv;
      } else {
        g0 = (hy - hy_k) * inv;
        g1 = (hx - hx_k) * inv;
      }
      const float s0 = bb * pin[ax][0] + cc * g0;
      const float s1 = bb * pin[ax][1] + cc * g1;
      pout[ax][0] = s0;
      pout[ax][1] = s1;
      if (store) {
        stf(cm.psw, q + 2 * ts, s0);
        stf(cm.psw, q + 3 * ts, s1);
      }
      const float d0 = ce * s0 + ce * km1 * g0;
      const float d1 = ce * s1 + ce * km1 * g1;
      if (ax == 0) {
        e1 -= d0;
        e2 += d1;
      } else if (ax == 1) {
        e0 += d0;
        e2 -= d1;
      } else {
        e0 -= d0;
        e1 += d1;
      }
    }
  }
  if (p.pec_axes & 1) {
    if (i == 0 || i == p.nx - 1) e1 = e2 = 0.f;
  }
  if (p.pec_axes & 2) {
    if (j == 0 || j == p.ny - 1) e0 = e2 = 0.f;
  }
  if (p.pec_axes & 4) {
    if (k == 0 || k == p.nz - 1) e0 = e1 = 0.f;
    if (k == p.nz - 1) e2 = 0.f;
  }
  if (fl) {
    if (fl & RFXH_ZERO_EX) e0 = 0.f;
    if (fl & RFXH_ZERO_EY) e1 = 0.f;
    if (fl & RFXH_ZERO_EZ) e2 = 0.f;
    if (fl & RFXH_POINT)
      tb_point_ops(p, sh, c, e0, e1, e2, hx, hy, hz, hx_j, hx_k, hy_i, hy_k, hz_i, hz_j, n, s,
                   own);
  }
  return make_float3(e0, e1, e2);
}

// The sweep of one workgroup over planes [i0, i1) of its tile; CPML =
// false leaves out the memory terms and slab tests (the caller has checked
// that the workgroup's needed cells miss every slab).
template <bool GRADED, bool CPML, bool EDGE>
__device__ __forceinline__ void tb2_sweep(const KParams& p, const TBIO& io, TBShared& sh,
                                          int row, int col, int j, int k, int i0, int i1,
                                          int kown1) {
  // tile: TR rows x TC columns (16 x 64, one row per wavefront; z-edge
  // tiles 64 x 16, four rows of 16 columns per wavefront); (row, col) is this
  // thread's place in it, the outer kTBHalo rows and columns are halo
  constexpr int TR = EDGE ? kTBW * kBX / 16 : kTBW;
  constexpr int TC = EDGE ? 16 : kBX;
  const TBCpml& cm = io.cm;
  const float4* const s_ecoef = sh.ecoef;
  const float* const s_ch = sh.ch;
  auto& s_x = sh.x;
  const bool valid = j >= 0 && j < p.ny && k >= 0 && k < p.nz;
  const bool interior = valid && row >= kTBHalo && row < TR - kTBHalo && col >= kTBHalo &&
                        col < TC - kTBHalo && k < kown1;
  const unsigned P = (unsigned)p.plane;
  const unsigned c0 = valid ? (unsigned)j * (unsigned)p.nz + (unsigned)k : 0u;
  const long long n = *p.step;
  const float chc = p.ch_all;
  // graded lane: this thread's inverse spacings along y and z
  float ghy = 0.f, ghz = 0.f, gey = 0.f, gez = 0.f;
  if (GRADED && valid) {
    ghy = gld(p.inv_h[1], j);
    ghz = gld(p.inv_h[2], k);
    gey = gld(p.inv_e[1], j);
    gez = gld(p.inv_e[2], k);
  }
  auto pin = [&](int q) { return valid && q >= 0 && q < p.nx; };
  auto own = [&](int q) { return interior && q >= i0 && q < i1; };
  // neighbour value of stage st from LDS buffer b; zero outside the tile
  auto nb = [&](int b, int st, int cmp, int rr, int cc) {
    const bool ok = rr >= 0 && rr < TR && cc >= 0 && cc < TC;
    const float v = s_x[b][st][cmp][(ok ? rr : 0) * TC + (ok ? cc : 0)];
    return ok ? v : 0.f;
  };
  const int me = row * TC + col;   // this thread's LDS slot
  // memory terms (term0: 0 for H, 2 for E) of the cell at plane q from
  // `base` (zero outside slabs), and their store
  auto ld_psi = [&](float (&dst)[3][2], const float* base, unsigned term0, int q, bool use) {
#pragma unroll
    for (int a = 0; a < 3; ++a) dst[a][0] = dst[a][1] = 0.f;
    if (CPML && use && pin(q) && in_slab(p, q, j, k)) {
#pragma unroll
      for (int face = 0; face < 6; ++face) {
        const int ly = slab_layer(p, face, q, j, k);
        if (ly < 0) continue;
        const unsigned qq = cm.foff[face] + term0 * cm.fs[face] +
                            (unsigned)psi_idx(face, ly, q, j, k, p.cn[face], p.ny, p.nz);
        dst[face >> 1][0] = ldf(base, qq);
        dst[face >> 1][1] = ldf(base, qq + cm.fs[face]);
      }
    }
  };
  auto st_psi = [&](float* base, unsigned term0, int q, const float (&src)[3][2], bool use) {
    if (CPML && use && pin(q) && in_slab(p, q, j, k)) {
#pragma unroll
      for (int face = 0; face < 6; ++face) {
        const int ly = slab_layer(p, face, q, j, k);
        if (ly < 0) continue;
        const unsigned qq = cm.foff[face] + term0 * cm.fs[face] +
                            (unsigned)psi_idx(face, ly, q, j, k, p.cn[face], p.ny, p.nz);
        stf(base, qq, src[face >> 1][0]);
        stf(base, qq + cm.fs[face], src[face >> 1][1]);
      }
    }
  };
  // DFT planes this thread's (j, k) can lie on (x-normal planes: all), so
  // that the waves without plane cells skip the staging
  unsigned snap_bits = 0u;
  for (int t = 0; t < io.n_plane; ++t) {
    const TBPlane& d = io.pl[t];
    const bool hit = d.axis == 0 || (d.axis == 1 ? j == d.index : k == d.index);
    if (valid && hit) snap_bits |= 1u << t;
  }
  auto stage_snap = [&](int q, int cmp0, const float3 v, long long st) {
    if (!snap_bits) return;
    for (int t = 0; t < io.n_plane; ++t) {
      const TBPlane& d = io.pl[t];
      if (!((snap_bits >> t) & 1u) || d.comp < cmp0 || d.comp >= cmp0 + 3) continue;
      const int pos = d.axis == 0 ? q : d.axis == 1 ? j : k;
      if (pos != d.index) continue;
      // plane cell (u, v): the two other indices in axis order
      const int u = d.axis == 0 ? j : q, vv = d.axis == 2 ? j : k;
      const float x = d.comp - cmp0 == 0 ? v.x : d.comp - cmp0 == 1 ? v.y : v.z;
      d.snap[(st % kDftRing) * d.npts + (long long)u * d.n2 + vv] = x;
    }
  };

  // pipeline state (see above); at iteration p:
  float e0c[3], e0b[3], e0n[3];   // E^n of planes p-1, p-2, p
  float h0n[3];                   // H^{n-1/2} of plane p-1
  float h1a[3], h1b[3];           // H^{n+1/2} of planes p-2, p-3
  float e1a[3], e1b[3];           // E^{n+1} of planes p-3, p-4
  float h2a[3], h2by, h2bz;       // H^{n+3/2} of plane p-4, (y, z) of p-5
  unsigned m1, m2, m3, m4;        // material ids of planes p-1 .. p-4
#pragma unroll
  for (int c = 0; c < 3; ++c) e0c[c] = e0b[c] = h1a[c] = h1b[c] = e1a[c] = e1b[c] = h2a[c] = 0.f;
  h2by = h2bz = 0.f;
  m2 = m3 = m4 = 0u;
  // prefetch(pp): E^n of plane pp+1, H^{n-1/2} and id of plane pp, raw
  // (clamped addresses; zeroed outside the grid where consumed).  Two
  // slots: iteration pp fills one with prefetch(pp + 1) and consumes the
  // other, filled one iteration earlier, so each load has two iterations
  // to arrive and waiting for it never waits for the newer one.
  float fe[2][3], fh[2][3];
  unsigned fm[2];
  auto prefetch = [&](int pp, float (&e)[3], float (&h)[3], unsigned& m) {
    const int q = pp + 1;
    const unsigned oq = (unsigned)min(max(q, 0), p.nx - 1) * P + c0;
    const unsigned op = (unsigned)min(max(pp, 0), p.nx - 1) * P + c0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      e[c] = ldf(io.eo[c], oq);
      h[c] = ldf(io.ho[c], op);
    }
    m = ldu16(p.mat, op);
  };
  prefetch(i0 - 3, fe[0], fh[0], fm[0]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    e0n[c] = pin(i0 - 2) ? fe[0][c] : 0.f;
    h0n[c] = pin(i0 - 3) ? fh[0][c] : 0.f;
  }
  m1 = fm[0];
  // Priming in place of a first iteration at pp = i0-2 (whose updates no
  // output needs): publish E^n of plane i0-2 for the first H update's y/z
  // neighbours and shift the planes by one.  The other stages of buffer 0
  // hold the zeros written at the kernel's start.
  s_x[0][0][0][me] = e0n[0];
  s_x[0][0][1][me] = e0n[1];
  s_x[0][0][2][me] = e0n[2];
  prefetch(i0 - 2, fe[1], fh[1], fm[1]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    e0c[c] = e0n[c];
    e0n[c] = pin(i0 - 1) ? fe[1][c] : 0.f;
    h0n[c] = pin(i0 - 2) ? fh[1][c] : 0.f;
  }
  m2 = m1;
  m1 = fm[1];
  // The CPML body loads its memory terms under branches, and the compiler
  // then waits for every load in flight before using them: a second slot
  // would only cost registers there, so it prefetches one plane ahead.
  constexpr bool kDeep = !CPML;
  if (kDeep) prefetch(i0 - 1, fe[0], fh[0], fm[0]);
  __syncthreads();

  // one iteration: planes pp-4 .. pp-1; LDS buffer wb written, 1 - wb read;
  // (ce, chh, cmm) the slot holding prefetch(pp) at the iteration's end,
  // (ne, nh, nm) the slot it fills with prefetch(fill) -- pp + 1 into the
  // other slot, or pp into the same one
  auto iteration = [&](int pp, int wb, const float (&ce)[3], const float (&chh)[3],
                       const unsigned& cmm, float (&ne)[3], float (&nh)[3], unsigned& nm,
                       int fill) {
    const int rb = wb ^ 1;
    // CPML memory of this iteration's four updates, issued before the
    // prefetch so that waiting for it leaves the prefetch in flight (vector
    // loads complete in order)
    float pD[3][2], pB[3][2], pC[3][2], pA[3][2];
    ld_psi(pD, cm.mid, 2u, pp - 4, own(pp - 4));
    ld_psi(pB, cm.psr, 2u, pp - 2, true);
    ld_psi(pC, cm.mid, 0u, pp - 3, true);
    ld_psi(pA, cm.psr, 0u, pp - 1, true);
    prefetch(fill, ne, nh, nm);
    // ---- D: E^{n+2} on plane pp-4
    float3 e2 = make_float3(0.f, 0.f, 0.f);
    {
      // every lane computes (no branch); lanes outside the grid or the
      // planes keep zero, the boundary value their neighbours need
      const int q = pp - 4;
      const float hx_j = nb(rb, 3, 0, row - 1, col), hz_j = nb(rb, 3, 2, row - 1, col);
      const float hx_k = nb(rb, 3, 0, row, col - 1), hy_k = nb(rb, 3, 1, row, col - 1);
      const bool o = own(q);
      const unsigned c = (unsigned)q * P + c0;
      float dummy[3][2];
      const float4 row = s_ecoef[m4];
      e2 = tb_e<GRADED, CPML>(p, cm, sh.prof, q, j, k, gey, gez, c, row, __float_as_uint(row.w), e1b[0], e1b[1],
                              e1b[2], h2a[0], h2a[1], h2a[2], h2by, h2bz, hx_j, hz_j, hx_k, hy_k,
                              pD, dummy, o, sh, n, 1, o);
      if (!pin(q)) e2 = make_float3(0.f, 0.f, 0.f);
      if (o) {
        stf(io.en[0], c, e2.x);
        stf(io.en[1], c, e2.y);
        stf(io.en[2], c, e2.z);
        stage_snap(q, 0, e2, n + 1);
      }
    }
    // ---- B: E^{n+1} on plane pp-2
    float3 e1 = make_float3(0.f, 0.f, 0.f);
    float pe_new[3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) pe_new[a][0] = pe_new[a][1] = 0.f;
    {
      const int q = pp - 2;
      const float hx_j = nb(rb, 1, 0, row - 1, col), hz_j = nb(rb, 1, 2, row - 1, col);
      const float hx_k = nb(rb, 1, 0, row, col - 1), hy_k = nb(rb, 1, 1, row, col - 1);
      const bool o = own(q);
      const unsigned c = (unsigned)q * P + c0;
      const float4 row = s_ecoef[m2];
      e1 = tb_e<GRADED, CPML>(p, cm, sh.prof, q, j, k, gey, gez, c, row, __float_as_uint(row.w), e0b[0], e0b[1],
                              e0b[2], h1a[0], h1a[1], h1a[2], h1b[1], h1b[2], hx_j, hz_j, hx_k,
                              hy_k, pB, pe_new, false, sh, n, 0, o);
      if (!pin(q)) e1 = make_float3(0.f, 0.f, 0.f);
      st_psi(cm.mid, 2u, q, pe_new, o);    // only the owner takes the second E update
      if (o) stage_snap(q, 0, e1, n);
    }
    // ---- C: H^{n+3/2} on plane pp-3
    float3 h2 = make_float3(0.f, 0.f, 0.f);
    {
      const int q = pp - 3;
      const float ex_j = nb(rb, 2, 0, row + 1, col), ez_j = nb(rb, 2, 2, row + 1, col);
      const float ex_k = nb(rb, 2, 0, row, col + 1), ey_k = nb(rb, 2, 1, row, col + 1);
      const bool o = own(q);
      const unsigned c = (unsigned)q * P + c0;
      float dummy[3][2];
      h2 = tb_h<GRADED, CPML>(p, cm, sh.prof, q, j, k, ghy, ghz, chc > 0.f ? chc : s_ch[m3], e1a[0], e1a[1], e1a[2],
                              ex_j, ez_j, ex_k, ey_k, e1.y, e1.z, h1b[0], h1b[1], h1b[2], pC,
                              dummy, o);
      if (!pin(q)) h2 = make_float3(0.f, 0.f, 0.f);
      if (o) {
        stf(io.hn[0], c, h2.x);
        stf(io.hn[1], c, h2.y);
        stf(io.hn[2], c, h2.z);
        stage_snap(q, 3, h2, n + 1);
      }
    }
    // ---- A: H^{n+1/2} on plane pp-1
    float3 h1 = make_float3(0.f, 0.f, 0.f);
    float ph_new[3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) ph_new[a][0] = ph_new[a][1] = 0.f;
    {
      const int q = pp - 1;
      const float ex_j = nb(rb, 0, 0, row + 1, col), ez_j = nb(rb, 0, 2, row + 1, col);
      const float ex_k = nb(rb, 0, 0, row, col + 1), ey_k = nb(rb, 0, 1, row, col + 1);
      h1 = tb_h<GRADED, CPML>(p, cm, sh.prof, q, j, k, ghy, ghz, chc > 0.f ? chc : s_ch[m1], e0c[0], e0c[1], e0c[2],
                              ex_j, ez_j, ex_k, ey_k, e0n[1], e0n[2], h0n[0], h0n[1], h0n[2], pA,
                              ph_new, false);
      if (!pin(q)) h1 = make_float3(0.f, 0.f, 0.f);
      // every copy whose H is right stores it (halo rows take the second H
      // update too): not the tile's last row and lane (their outer
      // neighbours read as zero), not the plane before the pipeline fills
      st_psi(cm.mid, 0u, q, ph_new, q >= i0 - 2 && row < TR - 1 && col < TC - 1);
      if (own(q)) stage_snap(q, 3, h1, n);
    }
    // ---- publish this iteration's planes for the neighbours
    s_x[wb][0][0][me] = e0n[0];
    s_x[wb][0][1][me] = e0n[1];
    s_x[wb][0][2][me] = e0n[2];
    s_x[wb][1][0][me] = h1.x;
    s_x[wb][1][1][me] = h1.y;
    s_x[wb][1][2][me] = h1.z;
    s_x[wb][2][0][me] = e1.x;
    s_x[wb][2][1][me] = e1.y;
    s_x[wb][2][2][me] = e1.z;
    s_x[wb][3][0][me] = h2.x;
    s_x[wb][3][1][me] = h2.y;
    s_x[wb][3][2][me] = h2.z;
    __syncthreads();
    // ---- rotate
    h2by = h2a[1];
    h2bz = h2a[2];
    h2a[0] = h2.x;
    h2a[1] = h2.y;
    h2a[2] = h2.z;
    const bool vq = pin(pp + 1), vp = pin(pp);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      e1b[c] = e1a[c];
      h1b[c] = h1a[c];
      e0b[c] = e0c[c];
      e0c[c] = e0n[c];
      e0n[c] = vq ? ce[c] : 0.f;
      h0n[c] = vp ? chh[c] : 0.f;
    }
    e1a[0] = e1.x;
    e1a[1] = e1.y;
    e1a[2] = e1.z;
    h1a[0] = h1.x;
    h1a[1] = h1.y;
    h1a[2] = h1.z;
    m4 = m3;
    m3 = m2;
    m2 = m1;
    m1 = cmm;
  };
  // an even number of iterations (one past the last plane's when the count
  // is odd: it stores nothing the workgroup owns, and the H memory it may
  // store is the value every other copy of that cell stores)
  for (int pp = i0 - 1; pp < i1 + 4; pp += 2) {
    if (kDeep) {
      iteration(pp, 1, fe[0], fh[0], fm[0], fe[1], fh[1], fm[1], pp + 1);
      iteration(pp + 1, 0, fe[1], fh[1], fm[1], fe[0], fh[0], fm[0], pp + 2);
    } else {
      iteration(pp, 1, fe[0], fh[0], fm[0], fe[0], fh[0], fm[0], pp);
      iteration(pp + 1, 0, fe[0], fh[0], fm[0], fe[0], fh[0], fm[0], pp + 1);
    }
  }
}


// ANY_CPML = false: a plan without CPML.
template <bool GRADED, bool ANY_CPML>
__global__ __launch_bounds__(kBX * kTBW) void k_step_tb2(KParams p, TBIO io) {
  __shared__ TBShared sh;
  // one wavefront per row: the row index is wave-uniform (scalar registers)
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y), l = (int)threadIdx.x;
  const int tid = w * kBX + l;
  // the plan rejects this variant for more than kTBMat materials
  for (int t = tid; t < io.n_mat; t += kBX * kTBW) {
    sh.ecoef[t] = gld(p.ecoef, t);
    sh.ch[t] = ld_ch(p.coef, (unsigned)t);
  }
  // the plan rejects this variant for more than kTBProf CPML layers in all
  for (int t = tid; t < io.n_prof; t += kBX * kTBW) sh.prof[t] = gld(io.cm.prof, t);
  // the first iteration reads buffer 0 (no plane before the segment start
  // but the primed E^n)
  for (int t = tid; t < 4 * 3 * kTBW * kBX; t += kBX * kTBW) (&sh.x[0][0][0][0])[t] = 0.f;
  __syncthreads();
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  // the tiles of one x segment: first the z-edge tiles (slower: CPML on
  // most of their cells), then the 16 x 64 tiles
  const int n_edge = 2 * io.ntj_e;
  const int tiles = n_edge + io.ntj * io.ntk;
  const int t = L % tiles;
  // dispatch order: the two x-slab segments (the slowest workgroups, CPML
  // on every cell) first, so that they do not form the launch's tail;
  // then the middle ones.  seg: 0 x-lo slab, 1..n_mid middle, n_mid + 1 x-hi
  const int so = L / tiles;
  const int seg = so == 0 ? 0 : so == 1 ? io.n_mid + 1 : so - 1;
  int i0, i1;
  if (seg == 0) {
    i0 = 0;
    i1 = io.xa;
  } else if (seg > io.n_mid) {
    i0 = io.xb;
    i1 = p.nx;
  } else {
    i0 = io.xa + (seg - 1) * io.seg_len;
    i1 = min(io.xb, i0 + io.seg_len);
  }
  if (i0 >= i1) return;   // whole workgroup
  const bool edge = t < n_edge;
  int j0, k0, row, col, kown1, tr, tc;
  if (edge) {
    // 64 x 16 tile over a z face's slab: its own columns are [0, kz0) or
    // [kz1, nz)
    const int side = t / io.ntj_e;
    j0 = (t - side * io.ntj_e) * (kTBEdgeR - 2 * kTBHalo);
    k0 = side == 0 ? 0 : io.kz1;
    row = w * (kBX / kTBEdgeC) + l / kTBEdgeC;
    col = l % kTBEdgeC;
    kown1 = side == 0 ? io.kz0 : p.nz;
    tr = kTBEdgeR;
    tc = kTBEdgeC;
  } else {
    const int u = t - n_edge;
    const int tk = u % io.ntk;
    j0 = (u / io.ntk) * kTBJ;
    k0 = io.kz0 + tk * kTBK;
    row = w;
    col = l;
    kown1 = io.kz1;
    tr = kTBW;
    tc = kBX;
  }
  const int j = j0 - kTBHalo + row;
  const int k = k0 - kTBHalo + col;
  // point operations of every cell the sweep's stages visit
  // (without any, no cell carries the point flag and the list is unused)
  if (p.n_ops > 0)
    tb_stage_ops(p, sh, w, l, *p.step, i0 - 6, i1 + 6, j0 - kTBHalo, j0 - kTBHalo + tr,
                 k0 - kTBHalo, k0 - kTBHalo + tc);
  // cells the interior's two steps need: planes [i0-2, i1], rows
  // [j0-2, j0 + inner rows], columns [k0-2, min(k0 + inner columns, own end)]
  const int kneed = min(k0 + tc - 2 * kTBHalo, kown1);
  const bool slab =
      ANY_CPML && ((p.cn[0] > 0 && i0 - 2 < p.slo[0]) || (p.cn[1] > 0 && i1 >= p.shi[0]) ||
                   (p.cn[2] > 0 && j0 - 2 < p.slo[1]) ||
                   (p.cn[3] > 0 && j0 + tr - 2 * kTBHalo >= p.shi[1]) ||
                   (p.cn[4] > 0 && k0 - 2 < p.slo[2]) || (p.cn[5] > 0 && kneed >= p.shi[2]));
  if (edge) tb2_sweep<GRADED, ANY_CPML, true>(p, io, sh, row, col, j, k, i0, i1, kown1);
  else if (slab) tb2_sweep<GRADED, ANY_CPML, false>(p, io, sh, row, col, j, k, i0, i1, kown1);
  else tb2_sweep<GRADED, false, false>(p, io, sh, row, col, j, k, i0, i1, kown1);
}

// ---------------------------------------------------------------------------
// Two-pass step, in place on p.f.
// ---------------------------------------------------------------------------
template <bool GRADED, bool ONE_MU = false>
__device__ __forceinline__ void two_pass_h(const KParams& p, const PsiIO& ps, int i, int j, int k) {
  const float* const Eo[3] = {p.f[0], p.f[1], p.f[2]};
  const float* const Ho[3] = {p.f[3], p.f[4], p.f[5]};
  float* const Hn[3] = {p.f[3], p.f[4], p.f[5]};
  h_cell<GRADED, true, ONE_MU>(p, Eo, Ho, Hn, ps, i, j, k);
}

template <bool GRADED, bool DISP = false>
__device__ __forceinline__ void two_pass_e(const KParams& p, const PsiIO& ps, int i, int j, int k) {
  const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const float* Hx = p.f[3];
  const float* Hy = p.f[4];
  const float* Hz = p.f[5];
  const bool hj = j > 0, hk = k > 0, hi = i > 0;
  const unsigned cj = hj ? c - p.nz : c, ck = hk ? c - 1 : c, ci = hi ? c - (unsigned)p.plane : c;
  const float hx = ldf(Hx, c), hy = ldf(Hy, c), hz = ldf(Hz, c);
  float hx_j = ldf(Hx, cj), hz_j = ldf(Hz, cj);
  float hx_k = ldf(Hx, ck), hy_k = ldf(Hy, ck);
  float hy_i = ldf(Hy, ci), hz_i = ldf(Hz, ci);
  hx_j = hj ? hx_j : 0.f;
  hz_j = hj ? hz_j : 0.f;
  hx_k = hk ? hx_k : 0.f;
  hy_k = hk ? hy_k : 0.f;
  hy_i = hi ? hy_i : 0.f;
  hz_i = hi ? hz_i : 0.f;
  const float* const Eo[3] = {p.f[0], p.f[1], p.f[2]};
  float* const En[3] = {p.f[0], p.f[1], p.f[2]};
  e_cell<GRADED, DISP>(p, Eo, En, ps, i, j, k, hx, hy, hz, hy_i, hz_i, hx_j, hz_j, hx_k, hy_k);
}

template <bool GRADED, bool ONE_MU>
__global__ __launch_bounds__(kWG) void k_update_h(KParams p, PsiIO ps) {
  int i, j, k;
  if (!cell_of(p, i, j, k)) return;
  two_pass_h<GRADED, ONE_MU>(p, ps, i, j, k);
}

// DISP: the plan has dispersive poles (a separate instance: the ADE branch would
// cost the plain kernels registers)
template <bool GRADED, bool DISP>
__global__ __launch_bounds__(kWG) void k_update_e(KParams p, PsiIO ps) {
  int i, j, k;
  if (!cell_of(p, i, j, k)) return;
  two_pass_e<GRADED, DISP>(p, ps, i, j, k);
}

// ---------------------------------------------------------------------------
// Column variants of the two passes: each thread updates NP consecutive
// cells along x.  The x-neighbour one cell needs is the next (H pass) or
// previous (E pass) cell's own value, already in registers, so x-neighbours
// are re-read once per column instead of once per cell, and each thread
// has NP cells' loads in flight at once.
// ---------------------------------------------------------------------------
// Cells i0 .. i0+NP-1 of the column, updating only those below iend
// (iend <= nx; the others are read as neighbours only).
template <bool GRADED, int NP, bool ONE_MU>
__device__ __forceinline__ void col_h(const KParams& p, const PsiIO& ps, int i0, int j, int k,
                                      int iend) {
  const unsigned P = (unsigned)p.plane;
  const bool hj = j + 1 < p.ny, hk = k + 1 < p.nz;
  const unsigned c0 = ((unsigned)i0 * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const unsigned dj = hj ? (unsigned)p.nz : 0u, dk = hk ? 1u : 0u;
  float ex[NP + 1], ey[NP + 1], ez[NP + 1], exj[NP], ezj[NP], exk[NP], eyk[NP];
  float hx[NP], hy[NP], hz[NP];
  unsigned m[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const bool in = i0 + q < p.nx;
    const unsigned c = in ? c0 + (unsigned)q * P : c0;
    m[q] = ONE_MU ? 0u : ldu16(p.mat, c);
    ex[q] = ldf(p.f[0], c);
    ey[q] = ldf(p.f[1], c);
    ez[q] = ldf(p.f[2], c);
    exj[q] = ldf(p.f[0], c + dj);
    ezj[q] = ldf(p.f[2], c + dj);
    exk[q] = ldf(p.f[0], c + dk);
    eyk[q] = ldf(p.f[1], c + dk);
    hx[q] = ldf(p.f[3], c);
    hy[q] = ldf(p.f[4], c);
    hz[q] = ldf(p.f[5], c);
  }
  {
    // E of the cell past the column (zero past the last plane)
    const bool in = i0 + NP < p.nx;
    const unsigned c = in ? c0 + (unsigned)NP * P : c0;
    ey[NP] = ldf(p.f[1], c);
    ez[NP] = ldf(p.f[2], c);
    ey[NP] = in ? ey[NP] : 0.f;
    ez[NP] = in ? ez[NP] : 0.f;
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    if (i0 + q >= iend) break;
    const float exj_ = hj ? exj[q] : 0.f, ezj_ = hj ? ezj[q] : 0.f;
    const float exk_ = hk ? exk[q] : 0.f, eyk_ = hk ? eyk[q] : 0.f;
    const float eyi = (q + 1 < NP) ? ((i0 + q + 1 < p.nx) ? ey[q + 1] : 0.f) : ey[NP];
    const float ezi = (q + 1 < NP) ? ((i0 + q + 1 < p.nx) ? ez[q + 1] : 0.f) : ez[NP];
    const float3 h = h_core<GRADED, true>(p, ps, i0 + q, j, k,
                                          ONE_MU ? p.ch_all : ld_ch(p.coef, m[q]), ex[q], ey[q], ez[q],
                                          exj_, ezj_, exk_, eyk_, eyi, ezi, hx[q], hy[q], hz[q]);
    const unsigned c = c0 + (unsigned)q * P;
    stf(p.f[3], c, h.x);
    stf(p.f[4], c, h.y);
    stf(p.f[5], c, h.z);
  }
}

template <bool GRADED, int NP, bool DISP>
__device__ __forceinline__ void col_e(const KParams& p, const PsiIO& ps, int i0, int j, int k,
                                      int iend) {
  const unsigned P = (unsigned)p.plane;
  const bool bj = j > 0, bk = k > 0;
  const unsigned c0 = ((unsigned)i0 * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const unsigned dj = bj ? (unsigned)p.nz : 0u, dk = bk ? 1u : 0u;
  // index 0 holds the cell before the column (i0-1), 1..NP the column
  float hx[NP + 1], hy[NP + 1], hz[NP + 1], hxj[NP], hzj[NP], hxk[NP], hyk[NP];
  float e0[NP], e1[NP], e2[NP];
  unsigned m[NP];
  {
    const bool in = i0 > 0;
    const unsigned c = in ? c0 - P : c0;
    hy[0] = ldf(p.f[4], c);
    hz[0] = ldf(p.f[5], c);
    hy[0] = in ? hy[0] : 0.f;
    hz[0] = in ? hz[0] : 0.f;
    hx[0] = 0.f;
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const bool in = i0 + q < p.nx;
    const unsigned c = in ? c0 + (unsigned)q * P : c0;
    m[q] = ldu16(p.mat, c);
    hx[q + 1] = ldf(p.f[3], c);
    hy[q + 1] = ldf(p.f[4], c);
    hz[q + 1] = ldf(p.f[5], c);
    hxj[q] = ldf(p.f[3], c - dj);
    hzj[q] = ldf(p.f[5], c - dj);
    hxk[q] = ldf(p.f[3], c - dk);
    hyk[q] = ldf(p.f[4], c - dk);
    e0[q] = ldf(p.f[0], c);
    e1[q] = ldf(p.f[1], c);
    e2[q] = ldf(p.f[2], c);
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    if (i0 + q >= iend) break;
    const unsigned c = c0 + (unsigned)q * P;
    const float hxj_ = bj ? hxj[q] : 0.f, hzj_ = bj ? hzj[q] : 0.f;
    const float hxk_ = bk ? hxk[q] : 0.f, hyk_ = bk ? hyk[q] : 0.f;
    const float3 e = e_core<GRADED, DISP>(p, ps, i0 + q, j, k, c, m[q], e0[q], e1[q], e2[q],
                                          hx[q + 1], hy[q + 1], hz[q + 1], hy[q], hz[q], hxj_, hzj_,
                                          hxk_, hyk_);
    stf(p.f[0], c, e.x);
    stf(p.f[1], c, e.y);
    stf(p.f[2], c, e.z);
  }
}

// The CPML pointers come through device memory (psp) rather than as
// kernel arguments: as arguments the compiler loads all 48 of them up front
// and spills them to VGPR lanes in every wavefront, although only the
// wavefronts with slab cells use them.
// ONE_MU: every material has the same ch (p.ch_all), so the H pass reads
// no material id.
template <bool GRADED, int NP, bool ONE_MU>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(NP == 2 ? 8 : 1)))
void k_update_h_col(KParams p, const PsiIO* __restrict__ psp) {
  const PsiIO& ps = *psp;
  int ib, j, k;
  if (!cell_at(p, xcd_remap((int)blockIdx.x, (int)gridDim.x), ib, j, k, (p.nx + NP - 1) / NP))
    return;
  col_h<GRADED, NP, ONE_MU>(p, ps, ib * NP, j, k, p.nx);
}

// 2-cell columns: an occupancy target of 7 wavefronts per SIMD.  With
// 512-thread workgroups, holding the kernel to 64 VGPRs (8 wavefronts) was
// 3.7 % faster than the compiler's own choice; with 256-thread workgroups
// a target of 7 is 0.7-1 % faster than 8 and the unconstrained build.
template <bool GRADED, int NP, bool DISP>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(NP == 2 ? 7 : 1)))
void k_update_e_col(KParams p, const PsiIO* __restrict__ psp) {
  const PsiIO& ps = *psp;
  int ib, j, k;
  if (!cell_at(p, xcd_remap((int)blockIdx.x, (int)gridDim.x), ib, j, k, (p.nx + NP - 1) / NP))
    return;
  col_e<GRADED, NP, DISP>(p, ps, ib * NP, j, k, p.nx);
}

// ---------------------------------------------------------------------------
// Batched two-pass step over many independent problems: one launch per
// half-step covers every active instance.  Instance q owns logical
// workgroups [start[q], start[q+1]); the instance table lives in device
// memory, so every parameter read below is wave-uniform (scalar loads).
// The instance of a workgroup is a binary search over the few range starts
// (they stay in the scalar cache); a per-workgroup owner table was measured
// slower, its one load per workgroup missing that cache.
// ---------------------------------------------------------------------------
struct BatchEntry {
  KParams p;   // p.f = the instance's current field set
  PsiIO ps;
};

__device__ __forceinline__ int batch_instance(const int* __restrict__ start, int n, int L) {
  int lo = 0, hi = n - 1;  // largest q with start[q] <= L
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (gld(start, mid) <= L) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One cell per thread (the two-pass kernels' bodies).  Measured on a
// 48-instance sweep: 46 Gcell-steps/s, against 40 with the 2-cell column
// bodies and 35 with the column bodies at one cell; the instance
// parameters stay behind a reference into the table (a by-value copy
// measured as slow as the column bodies).
// ONE_MU: every instance has one mu_r (its ch_all), so the H pass reads no
// material id.
template <int PASS, bool ONE_MU, bool DISP>
__global__ __launch_bounds__(kWG) void k_batch_pass(const BatchEntry* __restrict__ ents,
                                                   const int* __restrict__ start, int n_act) {
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int q = batch_instance(start, n_act, L);
  const KParams& p = ents[q].p;
  int i, j, k;
  if (!cell_at(p, L - gld(start, q), i, j, k)) return;
  if (PASS == 0) {
    if (p.graded) two_pass_h<true, ONE_MU>(p, ents[q].ps, i, j, k);
    else two_pass_h<false, ONE_MU>(p, ents[q].ps, i, j, k);
  } else {
    if (p.graded) two_pass_e<true, DISP>(p, ents[q].ps, i, j, k);
    else two_pass_e<false, DISP>(p, ents[q].ps, i, j, k);
  }
}

__global__ void k_batch_advance(const BatchEntry* __restrict__ ents, int n_act) {
  const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q < n_act) {
    long long* st = ents[q].p.step;
    *st = *st + 1;
  }
}

__global__ void k_batch_set_steps(const BatchEntry* __restrict__ ents, const long long* __restrict__ v,
                                  int n) {
  const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q < n) *ents[q].p.step = v[q];
}

// ---------------------------------------------------------------------------
// DFT plane: acc[f][u][v] += tw[n][f] * F(plane)  (float64 sums)
// ---------------------------------------------------------------------------
// The plane values of kDftK consecutive steps are staged in a ring and
// folded into the float64 accumulators together, so the accumulators (16 B
// per frequency and plane cell) are read and written once per kDftK steps
// instead of every step.  The additions happen in the same order as one
// step at a time (acc, then step lo, lo+1, ...), so the sums are unchanged.
__device__ __forceinline__ void dft_fold(const PlaneParams& q, long long pidx, long long npts,
                                         long long lo, long long hi, bool hi_in_reg, float xhi) {
  float xs[kDftK];
  const int cnt = (int)(hi - lo + 1);
#pragma unroll
  for (int m = 0; m < kDftK; ++m) {
    if (m < cnt) {
      const long long st = lo + m;
      xs[m] = (hi_in_reg && st == hi) ? xhi : q.snap[(st % kDftRing) * npts + pidx];
    }
  }
  for (int f0 = 0; f0 < q.nf; f0 += kFoldF) {
    double2 cur[kFoldF];
#pragma unroll
    for (int u = 0; u < kFoldF; ++u)
      if (f0 + u < q.nf) cur[u] = q.acc[(long long)(f0 + u) * npts + pidx];
#pragma unroll
    for (int u = 0; u < kFoldF; ++u) {
      if (f0 + u >= q.nf) break;
#pragma unroll
      for (int m = 0; m < kDftK; ++m) {
        if (m < cnt) {
          const double2 w = q.tw[(lo + m) * q.nf + f0 + u];
          const double x = (double)xs[m];
          cur[u].x += w.x * x;
          cur[u].y += w.y * x;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kFoldF; ++u)
      if (f0 + u < q.nf) q.acc[(long long)(f0 + u) * npts + pidx] = cur[u];
  }
}

__global__ __launch_bounds__(256) void k_dft_plane(PlaneParams q) {
  const long long npts = (long long)q.n1 * q.n2;
  const long long pidx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= npts) return;
  const long long n = *q.step;
  if (n < 0 || n >= q.max_steps) return;
  const int u = (int)(pidx / q.n2), v = (int)(pidx % q.n2);
  int ii[3];
  const int a1 = q.axis == 0 ? 1 : 0, a2 = q.axis == 2 ? 1 : 2;
  ii[q.axis] = q.index;
  ii[a1] = u;
  ii[a2] = v;
  const float x = q.fld[((long long)ii[0] * q.ny + ii[1]) * q.nz + ii[2]];
  if (n % kDftK == kDftK - 1) {
    const long long lo = max(*q.acc_from, n - (kDftK - 1));
    dft_fold(q, pidx, npts, lo, n, true, x);
  } else {
    q.snap[(n % kDftRing) * npts + pidx] = x;
  }
}

// After a two-step sweep of the temporal-blocked march, which staged the
// plane values of steps n and n+1 into the ring itself: fold a group that
// one of them completes.
__global__ __launch_bounds__(256) void k_dft_fold2(PlaneParams q) {
  const long long npts = (long long)q.n1 * q.n2;
  const long long pidx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= npts) return;
  const long long n = *q.step;
  for (long long st = n; st <= n + 1; ++st) {
    if (st < 0 || st >= q.max_steps || st % kDftK != kDftK - 1) continue;
    const long long lo = max(*q.acc_from, st - (kDftK - 1));
    dft_fold(q, pidx, npts, lo, st, false, 0.f);
  }
}

// Folds the staged steps of the last, incomplete group (end of a run).
__global__ __launch_bounds__(256) void k_dft_flush(PlaneParams q) {
  const long long npts = (long long)q.n1 * q.n2;
  const long long pidx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= npts) return;
  const long long last = *q.step - 1;
  if (last < 0) return;
  const int slot = (int)(last % kDftK);
  if (slot == kDftK - 1) return;   // folded by k_dft_plane / k_dft_fold2
  const long long lo = max(*q.acc_from, last - slot);
  if (lo > last) return;
  dft_fold(q, pidx, npts, lo, last, false, 0.f);
}

__global__ void k_set_acc_from(long long* acc_from, const long long* step) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *acc_from = *step;
}

// ---------------------------------------------------------------------------
// Flux monitor: tangential E at the plane, tangential H averaged over the
// planes index-1 and index, accumulated with their own twiddles (float64).
// ---------------------------------------------------------------------------
// Folds steps lo..hi of component t's ring (step hi from a register when
// hi_in_reg) into its accumulators, in step order per frequency: the
// additions a per-step accumulation makes, in the same order.  One thread
// per (point, component) (blockIdx.y = component): four times the threads
// of one per point, so the accumulator round trips overlap more.
__device__ __forceinline__ void flux_fold(const FluxParams& q, int t, long long pidx,
                                          long long npts, long long lo, long long hi,
                                          bool hi_in_reg, float xhi) {
  const int cnt = (int)(hi - lo + 1);
  float xs[kDftK];
#pragma unroll
  for (int m = 0; m < kDftK; ++m) {
    if (m < cnt) {
      const long long st = lo + m;
      xs[m] = (hi_in_reg && st == hi) ? xhi : q.snap[((st % kDftK) * 4 + t) * npts + pidx];
    }
  }
  const double2* tw = t < 2 ? q.twe : q.twh;
  // kFoldF frequencies at a time: their accumulator loads are in flight
  // together instead of one round trip per frequency
  for (int f0 = 0; f0 < q.nf; f0 += kFoldF) {
    double2 cur[kFoldF];
#pragma unroll
    for (int u = 0; u < kFoldF; ++u)
      if (f0 + u < q.nf) cur[u] = q.acc[((long long)t * q.nf + f0 + u) * npts + pidx];
#pragma unroll
    for (int u = 0; u < kFoldF; ++u) {
      if (f0 + u >= q.nf) break;
#pragma unroll
      for (int m = 0; m < kDftK; ++m) {
        if (m < cnt) {
          const double2 w = tw[(lo + m) * q.nf + f0 + u];
          const double x = (double)xs[m];
          cur[u].x += w.x * x;
          cur[u].y += w.y * x;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kFoldF; ++u)
      if (f0 + u < q.nf) q.acc[((long long)t * q.nf + f0 + u) * npts + pidx] = cur[u];
  }
}

__global__ __launch_bounds__(256) void k_flux(FluxParams q) {
  const long long npts = (long long)q.n1 * q.n2;
  const long long pidx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = (int)blockIdx.y;   // component: E1, E2, H1, H2
  if (pidx >= npts) return;
  const long long n = *q.step;
  if (n < 0 || n >= q.max_steps) return;
  const int u = (int)(pidx / q.n2) + q.lo1, v = (int)(pidx % q.n2) + q.lo2;
  int ii[3], im[3];
  const int a1 = q.axis == 0 ? 1 : 0, a2 = q.axis == 2 ? 1 : 2;
  ii[q.axis] = q.index;
  im[q.axis] = q.index_m1;
  ii[a1] = im[a1] = u;
  ii[a2] = im[a2] = v;
  const long long c = ((long long)ii[0] * q.ny + ii[1]) * q.nz + ii[2];
  const long long cm = ((long long)im[0] * q.ny + im[1]) * q.nz + im[2];
  const float* f = q.f[t];
  const float x = t < 2 ? f[c] : (f[cm] + f[c]) * 0.5f;
  const int slot = (int)(n % kDftK);
  if (slot == kDftK - 1) {
    const long long lo = max(*q.acc_from, n - (kDftK - 1));
    flux_fold(q, t, pidx, npts, lo, n, true, x);
  } else {
    q.snap[((long long)slot * 4 + t) * npts + pidx] = x;
  }
}

// Folds the staged steps of the last, incomplete group (end of a run).
__global__ __launch_bounds__(256) void k_flux_flush(FluxParams q) {
  const long long npts = (long long)q.n1 * q.n2;
  const long long pidx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= npts) return;
  const long long last = *q.step - 1;
  if (last < 0) return;
  const int slot = (int)(last % kDftK);
  if (slot == kDftK - 1) return;   // folded by k_flux
  const long long lo = max(*q.acc_from, last - slot);
  if (lo > last) return;
  flux_fold(q, (int)blockIdx.y, pidx, npts, lo, last, false, 0.f);
}

// ---------------------------------------------------------------------------
// Near-to-far-field transform: one thread per (direction, frequency).  The
// face cells stream through LDS in chunks of kFFBlock, each staged once per
// workgroup: its position and its four tangential DFT values at the
// workgroup's frequency times s dS (s = -1 on lo faces, +1 on hi faces).
// Per axis a the four sums S_t = sum s dS f_t e^{j k r.p} over both faces
// then fold into N = sum J dS e^{..} and L = sum M dS e^{..} through the
// face's cross products (J = n x H, M = -n x E).  Float64 throughout: the
// phase k r.p reaches a few hundred radians on large boxes.
// ---------------------------------------------------------------------------
constexpr int kFFBlock = 128;

struct FarFieldParams {
  const double2* acc[6];   // [4][nf][n1][n2] per face
  const double* pos[3];    // per-axis positions (device copies)
  const double* width[3];  // per-axis cell widths
  const double* dirs;      // [n_dir][9]: r-hat, theta-hat, phi-hat
  const double* kwave;     // [nf] 2 pi f / c0
  double2* e_theta;        // [nf][n_dir]
  double2* e_phi;
  int lo[3], hi[3];
  int nf, n_dir;
  double eta0;
};

__global__ __launch_bounds__(kFFBlock) void k_far_field(FarFieldParams q) {
  __shared__ double s_p[3][kFFBlock];
  __shared__ double2 s_v[4][kFFBlock];
  const int tid = threadIdx.x;
  const int f = blockIdx.y;
  const int d = blockIdx.x * kFFBlock + tid;
  const bool live = d < q.n_dir;
  const double k = q.kwave[f];
  double r[3] = {0.0, 0.0, 0.0};
  if (live)
    for (int c = 0; c < 3; ++c) r[c] = q.dirs[9 * (long long)d + c];
  double2 Nv[3], Lv[3];
  for (int c = 0; c < 3; ++c) Nv[c] = Lv[c] = make_double2(0.0, 0.0);
  for (int ax = 0; ax < 3; ++ax) {
    const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
    const int n2 = q.hi[t2] - q.lo[t2];
    const long long npts = (long long)(q.hi[t1] - q.lo[t1]) * n2;
    double2 S[4];
    for (int t = 0; t < 4; ++t) S[t] = make_double2(0.0, 0.0);
    for (int side = 0; side < 2; ++side) {
      const double2* acc = q.acc[2 * ax + side];
      const double sgn = side ? 1.0 : -1.0;
      const double pn = q.pos[ax][side ? q.hi[ax] : q.lo[ax]];
      for (long long base = 0; base < npts; base += kFFBlock) {
        __syncthreads();   // the previous chunk has been consumed
        const long long c = base + tid;
        if (c < npts) {
          const int i1 = q.lo[t1] + (int)(c / n2), i2 = q.lo[t2] + (int)(c % n2);
          s_p[ax][tid] = pn;
          s_p[t1][tid] = q.pos[t1][i1];
          s_p[t2][tid] = q.pos[t2][i2];
          const double ds = sgn * (q.width[t1][i1] * q.width[t2][i2]);
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const double2 a = acc[((long long)t * q.nf + f) * npts + c];
            s_v[t][tid] = make_double2(a.x * ds, a.y * ds);
          }
        }
        __syncthreads();
        const int m = (int)min((long long)kFFBlock, npts - base);
        for (int e = 0; e < m; ++e) {
          const double ph = k * (r[0] * s_p[0][e] + r[1] * s_p[1][e] + r[2] * s_p[2][e]);
          double sn, cs;
          sincos(ph, &sn, &cs);
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const double2 v = s_v[t][e];
            S[t].x += cs * v.x - sn * v.y;
            S[t].y += cs * v.y + sn * v.x;
          }
        }
      }
    }
    // (f0, f1, f2, f3) = x: (Ey, Ez, Hy, Hz); y: (Ex, Ez, Hx, Hz); z: (Ex, Ey, Hx, Hy)
    auto add = [](double2& a, const double2& b, double s) {
      a.x += s * b.x;
      a.y += s * b.y;
    };
    if (ax == 0) {          // J = (0, -f3, f2), M = (0, f1, -f0)
      add(Nv[1], S[3], -1.0); add(Nv[2], S[2], 1.0);
      add(Lv[1], S[1], 1.0);  add(Lv[2], S[0], -1.0);
    } else if (ax == 1) {   // J = (f3, 0, -f2), M = (-f1, 0, f0)
      add(Nv[0], S[3], 1.0);  add(Nv[2], S[2], -1.0);
      add(Lv[0], S[1], -1.0); add(Lv[2], S[0], 1.0);
    } else {                // J = (-f3, f2, 0), M = (f1, -f0, 0)
      add(Nv[0], S[3], -1.0); add(Nv[1], S[2], 1.0);
      add(Lv[0], S[1], 1.0);  add(Lv[1], S[0], -1.0);
    }
  }
  if (!live) return;
  const double* th = q.dirs + 9 * (long long)d + 3;
  const double* ph = th + 3;
  double2 nth = make_double2(0.0, 0.0), nph = nth, lth = nth, lph = nth;
  for (int c = 0; c < 3; ++c) {
    nth.x += Nv[c].x * th[c]; nth.y += Nv[c].y * th[c];
    nph.x += Nv[c].x * ph[c]; nph.y += Nv[c].y * ph[c];
    lth.x += Lv[c].x * th[c]; lth.y += Lv[c].y * th[c];
    lph.x += Lv[c].x * ph[c]; lph.y += Lv[c].y * ph[c];
  }
  // E_theta = -j k/(4 pi) (L_phi + eta0 N_theta); E_phi = j k/(4 pi) (L_theta - eta0 N_phi)
  const double g = k / (4.0 * M_PI);
  const double2 a = make_double2(lph.x + q.eta0 * nth.x, lph.y + q.eta0 * nth.y);
  const double2 b = make_double2(lth.x - q.eta0 * nph.x, lth.y - q.eta0 * nph.y);
  const long long o = (long long)f * q.n_dir + d;
  q.e_theta[o] = make_double2(g * a.y, -g * a.x);   // -j g a
  q.e_phi[o] = make_double2(-g * b.y, g * b.x);     //  j g b
}

// Reverse mode of k_far_field, first half: per (direction, frequency) the
// gradients of the twelve per-axis sums S[ax][t] from the gradients G_theta,
// G_phi of the outputs (torch's complex convention: for y = a x, G_x =
// conj(a) G_y).  E_theta = -j g A, E_phi = j g B give G_A = j g G_theta,
// G_B = -j g G_phi; A = L_phi + eta0 N_theta, B = L_theta - eta0 N_phi give
// G_L = phi-hat G_A + theta-hat G_B, G_N = eta0 (theta-hat G_A - phi-hat G_B);
// each S[ax][t] enters one component of N or L with sign +-1.
__global__ __launch_bounds__(256) void k_far_field_adj_s(FarFieldParams q,
                                                         const double2* __restrict__ g_th,
                                                         const double2* __restrict__ g_ph,
                                                         double2* __restrict__ gs) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (d >= q.n_dir) return;
  const double g = q.kwave[f] / (4.0 * M_PI);
  const long long o = (long long)f * q.n_dir + d;
  const double2 gt = g_th[o], gp = g_ph[o];
  const double2 GA = make_double2(-g * gt.y, g * gt.x);    //  j g G_theta
  const double2 GB = make_double2(g * gp.y, -g * gp.x);    // -j g G_phi
  const double* th = q.dirs + 9 * (long long)d + 3;
  const double* ph = th + 3;
  double2 GL[3], GN[3];
  for (int c = 0; c < 3; ++c) {
    GL[c] = make_double2(ph[c] * GA.x + th[c] * GB.x, ph[c] * GA.y + th[c] * GB.y);
    GN[c] = make_double2(q.eta0 * (th[c] * GA.x - ph[c] * GB.x),
                         q.eta0 * (th[c] * GA.y - ph[c] * GB.y));
  }
  auto neg = [](double2 v) { return make_double2(-v.x, -v.y); };
  double2* out = gs + 12 * o;
  // x: N1 -= S3, N2 += S2, L1 += S1, L2 -= S0
  out[0] = neg(GL[2]); out[1] = GL[1]; out[2] = GN[2]; out[3] = neg(GN[1]);
  // y: N0 += S3, N2 -= S2, L0 -= S1, L2 += S0
  out[4] = GL[2]; out[5] = neg(GL[0]); out[6] = neg(GN[2]); out[7] = GN[0];
  // z: N0 -= S3, N1 += S2, L0 += S1, L1 -= S0
  out[8] = neg(GL[1]); out[9] = GL[0]; out[10] = GN[1]; out[11] = neg(GN[0]);
}

// Second half, one thread per (face cell, frequency): dJ/d acc[t][f][c] =
// s dS(c) sum_d e^{-j k r_d.p_c} G_S[ax][t](f, d), the directions streaming
// through LDS in chunks of kFFBlock (r-hat and the face axis's four G_S).
__global__ __launch_bounds__(kFFBlock) void k_far_field_adj(FarFieldParams q,
                                                            const double2* __restrict__ gs,
                                                            int face, double2* __restrict__ gacc) {
  __shared__ double s_r[3][kFFBlock];
  __shared__ double2 s_g[4][kFFBlock];
  const int ax = face / 2, side = face & 1;
  const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
  const int n2 = q.hi[t2] - q.lo[t2];
  const long long npts = (long long)(q.hi[t1] - q.lo[t1]) * n2;
  const int tid = threadIdx.x;
  const int f = blockIdx.y;
  const long long c = (long long)blockIdx.x * kFFBlock + tid;
  const bool live = c < npts;
  const double k = q.kwave[f];
  double p[3] = {0.0, 0.0, 0.0};
  double ds = 0.0;
  if (live) {
    const int i1 = q.lo[t1] + (int)(c / n2), i2 = q.lo[t2] + (int)(c % n2);
    p[ax] = q.pos[ax][side ? q.hi[ax] : q.lo[ax]];
    p[t1] = q.pos[t1][i1];
    p[t2] = q.pos[t2][i2];
    ds = (side ? 1.0 : -1.0) * (q.width[t1][i1] * q.width[t2][i2]);
  }
  double2 acc[4];
  for (int t = 0; t < 4; ++t) acc[t] = make_double2(0.0, 0.0);
  for (int base = 0; base < q.n_dir; base += kFFBlock) {
    __syncthreads();   // the previous chunk has been consumed
    const int d = base + tid;
    if (d < q.n_dir) {
      for (int a = 0; a < 3; ++a) s_r[a][tid] = q.dirs[9 * (long long)d + a];
      const double2* g = gs + 12 * ((long long)f * q.n_dir + d) + 4 * ax;
#pragma unroll
      for (int t = 0; t < 4; ++t) s_g[t][tid] = g[t];
    }
    __syncthreads();
    const int m = min(kFFBlock, q.n_dir - base);
    for (int e = 0; e < m; ++e) {
      const double ph = k * (s_r[0][e] * p[0] + s_r[1][e] * p[1] + s_r[2][e] * p[2]);
      double sn, cs;
      sincos(ph, &sn, &cs);
#pragma unroll
      for (int t = 0; t < 4; ++t) {   // e^{-j ph} G
        const double2 v = s_g[t][e];
        acc[t].x += cs * v.x + sn * v.y;
        acc[t].y += cs * v.y - sn * v.x;
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int t = 0; t < 4; ++t)
    gacc[((long long)t * q.nf + f) * npts + c] = make_double2(acc[t].x * ds, acc[t].y * ds);
}

// Direction vectors and wave numbers of a far-field transform computed on
// the device (the allocation-free entry point has no host tables to
// upload): dirs[9 t] = r-hat, theta-hat, phi-hat of direction t = it n_phi +
// ip; kw[f] = 2 pi f / c0.
__global__ __launch_bounds__(256) void k_far_field_dirs(const double* __restrict__ theta,
                                                        int n_theta,
                                                        const double* __restrict__ phi, int n_phi,
                                                        const double* __restrict__ freqs, int nf,
                                                        double c0, double* __restrict__ dirs,
                                                        double* __restrict__ kw) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n_dir = (long long)n_theta * n_phi;
  if (t < n_dir) {
    const int it = (int)(t / n_phi), ip = (int)(t % n_phi);
    const double st = sin(theta[it]), ct = cos(theta[it]);
    const double sp = sin(phi[ip]), cp = cos(phi[ip]);
    double* o = dirs + 9 * t;
    o[0] = st * cp;
    o[1] = st * sp;
    o[2] = ct;
    o[3] = ct * cp;
    o[4] = ct * sp;
    o[5] = -st;
    o[6] = -sp;
    o[7] = cp;
    o[8] = 0.0;
  }
  if (t < nf) kw[t] = 2.0 * M_PI * freqs[t] / c0;
}

// ---------------------------------------------------------------------------
// Reverse mode.  The forward step is linear in the state (E, H, CPML memory)
// for fixed materials, so its adjoint is the transpose of each sub-step run
// in reverse order.  A reverse step is two gather kernels (k_adj_h, k_adj_e;
// k_batch_adj<false / true> over the instances of a batch):
//   H-bar: H-bar += (E update)^T E-bar + (CPML-E)^T g-bar_E, then the H-probe
//            adjoint and the CPML-H memory adjoint (writes g-bar_H);
//   E-bar:   E-bar <- ca E-bar + (H update)^T H-bar + (CPML-H)^T g-bar_H, then
//            the prologue of the previous step: E-probe adjoint, source
//            adjoint record, PEC projection, CPML-E memory adjoint (writes
//            g-bar_E) and the eps_r gradient from the forward record.
// k_record stores, for the gradient region, E^n, curl_bwd H^{n+1/2} and the
// CPML-E term of the forward E update (9 floats per cell per step), so that
// dE'/d(ca, cb, ce) is known without keeping whole forward states.
// ---------------------------------------------------------------------------
struct RegionParams {
  int i0, j0, k0;   // box origin
  int ni, nj, nk;   // box extent
  long long n;      // ni * nj * nk
};

struct AdjParams {
  float* a[6];            // adjoint fields: E-bar x, y, z, H-bar x, y, z
  float* ph[6][2];        // adjoint CPML-H memory (psi layout)
  float* pe[6][2];        // adjoint CPML-E memory
  float* gh[6][2];        // g-bar of the CPML-H terms (scratch, psi layout)
  float* ge[6][2];        // g-bar of the CPML-E terms
  const float* probe_adj; // [max_steps][n_probe] dJ/dy
  float* src_adj;         // [max_steps][n_src]   dJ/dw
  const float* rec;       // [rec_rows][9][region n]
  long long rec_step0, rec_rows;
  RegionParams reg;
  float* grad;            // [region n] dJ/d eps_r
  const float4* dcoef;    // per material id: d ca, d cb, d ce / d eps_r
  float* grad_s;          // [region n] dJ/d sigma (null: not computed)
  const float2* dcoef_s;  // per material id: d ca, d cb / d sigma
  const float* rec_h;     // [rec_rows][3][region n]: the H update's R^n
  float* grad_mu;         // [region n] dJ/d mu_r (null: not computed)
  const float* dmu;       // per material id: d ch / d mu_r
  const float* port_adj;  // [max_steps][n_port][2] dJ/dV, dJ/dI (null: none)
  // dispersive plans: the adjoint polarisation (the forward state's layout),
  // the plan's pole rows and box, the dispersive record [rec_rows][3][region
  // n] and per material (-gamma'/gamma w.r.t. eps_r, w.r.t. sigma)
  float* pol_adj;
  const float* pcoef;
  const float* rec_k;
  const float2* dkdis;
  int n_pole, n_debye;
  int pol_lo[3], pol_n[3];
  // single-instance sweep: per-cell cb and ch (the neighbours' coefficients
  // without an id load and a table gather each) and whether any probe
  // samples an H component
  const float* cbv;
  const float* chv;
  int h_probe;
};

__device__ __forceinline__ long long region_index(const RegionParams& r, int i, int j, int k) {
  const int a = i - r.i0, b = j - r.j0, c = k - r.k0;
  if (a < 0 || b < 0 || c < 0 || a >= r.ni || b >= r.nj || c >= r.nk) return -1;
  return ((long long)a * r.nj + b) * r.nk + c;
}

// the record of region cell q (row-major in the region box) into rec
// R = -curl_fwd E^n + CPML-H term at a cell, with H^{n+1/2} = H^{n-1/2} +
// ch R (h_core's update; the CPML memory already holds its new values);
// stored at rec_h[v * r.n + q].
template <bool GRADED>
__device__ __forceinline__ void record_h_cell(const KParams& p, const PsiIO& ps,
                                              const RegionParams& r, long long q, int i, int j,
                                              int k, float* rec_h) {
  const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const bool hj = j + 1 < p.ny, hk = k + 1 < p.nz, hi = i + 1 < p.nx;
  const float ex = ldf(p.f[0], c), ey = ldf(p.f[1], c), ez = ldf(p.f[2], c);
  const float ex_j = hj ? ldf(p.f[0], c + p.nz) : 0.f, ez_j = hj ? ldf(p.f[2], c + p.nz) : 0.f;
  const float ex_k = hk ? ldf(p.f[0], c + 1) : 0.f, ey_k = hk ? ldf(p.f[1], c + 1) : 0.f;
  const float ey_i = hi ? ldf(p.f[1], c + (unsigned)p.plane) : 0.f;
  const float ez_i = hi ? ldf(p.f[2], c + (unsigned)p.plane) : 0.f;
  float ix, iy, iz;
  if (GRADED) {
    ix = p.inv_h[0][i];
    iy = p.inv_h[1][j];
    iz = p.inv_h[2][k];
  } else {
    ix = iy = iz = p.inv_h0;
  }
  float rx = -((ez_j - ez) * iy - (ey_k - ey) * iz);
  float ry = -((ex_k - ex) * iz - (ez_i - ez) * ix);
  float rz = -((ey_i - ey) * ix - (ex_j - ex) * iy);
  if (in_slab(p, i, j, k)) {
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int l = slab_layer(p, face, i, j, k);
      if (l < 0) continue;
      const int ax = face >> 1;
      const unsigned pq = (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
      const float km1 = p.pk[face][l], inv = p.cinv[face];
      float g0, g1;
      if (ax == 0) {
        g0 = (ez_i - ez) * inv;
        g1 = (ey_i - ey) * inv;
      } else if (ax == 1) {
        g0 = (ez_j - ez) * inv;
        g1 = (ex_j - ex) * inv;
      } else {
        g0 = (ey_k - ey) * inv;
        g1 = (ex_k - ex) * inv;
      }
      const float d0 = ldf(ps.hr[face][0], pq) + km1 * g0;
      const float d1 = ldf(ps.hr[face][1], pq) + km1 * g1;
      if (ax == 0) {
        ry += d0;
        rz -= d1;
      } else if (ax == 1) {
        rx -= d0;
        rz += d1;
      } else {
        rx += d0;
        ry -= d1;
      }
    }
  }
  rec_h[q] = rx;
  rec_h[r.n + q] = ry;
  rec_h[2 * r.n + q] = rz;
}

template <bool GRADED>
__device__ __forceinline__ void record_cell(const KParams& p, const PsiIO& ps,
                                            const RegionParams& r, long long q, float* rec) {
  const int c3 = (int)(q % r.nk);
  const long long t = q / r.nk;
  const int i = r.i0 + (int)(t / r.nj), j = r.j0 + (int)(t % r.nj), k = r.k0 + c3;
  const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const float* Hx = p.f[3];
  const float* Hy = p.f[4];
  const float* Hz = p.f[5];
  const float hx = ldf(Hx, c), hy = ldf(Hy, c), hz = ldf(Hz, c);
  const float hx_j = j > 0 ? ldf(Hx, c - p.nz) : 0.f, hz_j = j > 0 ? ldf(Hz, c - p.nz) : 0.f;
  const float hx_k = k > 0 ? ldf(Hx, c - 1) : 0.f, hy_k = k > 0 ? ldf(Hy, c - 1) : 0.f;
  const float hy_i = i > 0 ? ldf(Hy, c - (unsigned)p.plane) : 0.f;
  const float hz_i = i > 0 ? ldf(Hz, c - (unsigned)p.plane) : 0.f;
  float ix, iy, iz;
  if (GRADED) {
    ix = p.inv_e[0][i];
    iy = p.inv_e[1][j];
    iz = p.inv_e[2][k];
  } else {
    ix = iy = iz = p.inv_e0;
  }
  float cp[3] = {0.f, 0.f, 0.f};
  if (in_slab(p, i, j, k)) {
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int l = slab_layer(p, face, i, j, k);
      if (l < 0) continue;
      const int ax = face >> 1;
      const unsigned pq = (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
      const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
      const float inv = p.cinv[face];
      float g0, g1;
      if (ax == 0) {
        g0 = (hz - hz_i) * inv;
        g1 = (hy - hy_i) * inv;
      } else if (ax == 1) {
        g0 = (hz - hz_j) * inv;
        g1 = (hx - hx_j) * inv;
      } else {
        g0 = (hy - hy_k) * inv;
        g1 = (hx - hx_k) * inv;
      }
      const float t0 = bb * ldf(ps.er[face][0], pq) + cc * g0 + km1 * g0;
      const float t1 = bb * ldf(ps.er[face][1], pq) + cc * g1 + km1 * g1;
      if (ax == 0) {
        cp[1] -= t0;
        cp[2] += t1;
      } else if (ax == 1) {
        cp[0] += t0;
        cp[2] -= t1;
      } else {
        cp[0] -= t0;
        cp[1] += t1;
      }
    }
  }
  const float vals[9] = {ldf(p.f[0], c), ldf(p.f[1], c), ldf(p.f[2], c),
                         (hz - hz_j) * iy - (hy - hy_k) * iz,
                         (hx - hx_k) * iz - (hz - hz_i) * ix,
                         (hy - hy_i) * ix - (hx - hx_j) * iy,
                         cp[0], cp[1], cp[2]};
  // the CPML-E rows only for cells in a slab (zero elsewhere; the reverse
  // sweep reads them only there)
  const int nv = in_slab(p, i, j, k) ? 9 : 6;
#pragma unroll
  for (int v = 0; v < 9; ++v)
    if (v < nv) rec[(long long)v * r.n + q] = vals[v];
}

// The dispersive part of the E update at one cell and step n, from the state
// before the E pass (E^n, P_d^n, P_l^n, P_l^{n-1}): K = sum_d cc_d P_d^n -
// cc_L sum_l (P_l^{n+1} - P_l^n), P_l^{n+1} = a P_l^n + b P_l^{n-1} + c E^n
// (dispersion_cell's terms); zero for a cell without poles.  Every cc is
// 1/gamma times a constant, so d K / d(eps_r or sigma) = -(gamma'/gamma) K.
__device__ void disp_k_cell(const KParams& p, const PsiIO& ps, int i, int j, int k, unsigned c,
                            float* out, long long stride) {
  float k0 = 0.f, k1 = 0.f, k2 = 0.f;
  const int li = i - ps.pol_lo[0], lj = j - ps.pol_lo[1], lk = k - ps.pol_lo[2];
  const unsigned m = ldu16(p.mat, c);
  if ((p.flags[m] & RFXH_POLES) && ps.pol && li >= 0 && lj >= 0 && lk >= 0 &&
      li < ps.pol_n[0] && lj < ps.pol_n[1] && lk < ps.pol_n[2]) {
    const unsigned R = (unsigned)ps.pol_n[0] * (unsigned)ps.pol_n[1] * (unsigned)ps.pol_n[2];
    const unsigned r = ((unsigned)li * (unsigned)ps.pol_n[1] + (unsigned)lj) *
                       (unsigned)ps.pol_n[2] + (unsigned)lk;
    const float* pc = ps.pcoef + (size_t)m * ps.n_pole * 4;
    const int nd = ps.n_debye;
    const float e0 = ldf(p.f[0], c), e1 = ldf(p.f[1], c), e2 = ldf(p.f[2], c);
    for (int q = 0; q < nd; ++q) {
      const float cc = pc[4 * q + 2];
      const float* P = ps.pol + (size_t)q * 3 * R;
      k0 += cc * P[r];
      k1 += cc * P[R + r];
      k2 += cc * P[2 * R + r];
    }
    float d0 = 0.f, d1 = 0.f, d2 = 0.f, ccl = 0.f;
    for (int q = nd; q < ps.n_pole; ++q) {
      const float a = pc[4 * q], b = pc[4 * q + 1], cl = pc[4 * q + 2];
      ccl = pc[4 * q + 3];
      const float* P = ps.pol + (size_t)nd * 3 * R + (size_t)(q - nd) * 6 * R;
      d0 += a * P[r] + b * P[3 * R + r] + cl * e0 - P[r];
      d1 += a * P[R + r] + b * P[4 * R + r] + cl * e1 - P[R + r];
      d2 += a * P[2 * R + r] + b * P[5 * R + r] + cl * e2 - P[2 * R + r];
    }
    k0 -= ccl * d0;
    k1 -= ccl * d1;
    k2 -= ccl * d2;
  }
  out[0] = k0;
  out[stride] = k1;
  out[2 * stride] = k2;
}

template <bool GRADED>
__global__ __launch_bounds__(256) void k_record(KParams p, PsiIO ps, RegionParams r, float* rec,
                                                float* rec_h, float* rec_k) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= r.n) return;
  record_cell<GRADED>(p, ps, r, q, rec);
  const int c3 = (int)(q % r.nk);
  const long long t = q / r.nk;
  const int i = r.i0 + (int)(t / r.nj), j = r.j0 + (int)(t % r.nj), k = r.k0 + c3;
  if (rec_h) record_h_cell<GRADED>(p, ps, r, q, i, j, k, rec_h);
  if (rec_k) {
    const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
    disp_k_cell(p, ps, i, j, k, c, rec_k + q, r.n);
  }
}

// ch or cb of the cell's material (w = ch, y = cb)
__device__ __forceinline__ float4 coef_at(const KParams& p, unsigned c) {
  return p.coef[ldu16(p.mat, c)];
}

template <bool GRADED>
__device__ __forceinline__ void adj_h_cell(const KParams& p, const AdjParams& a, int i, int j,
                                           int k, long long n) {
  const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const unsigned P = (unsigned)p.plane;
  const bool hj = j + 1 < p.ny, hk = k + 1 < p.nz, hi = i + 1 < p.nx;
  float ix, iy, iz, ixn, iyn, izn;  // E-update spacings at the cell and at its forward neighbour
  if (GRADED) {
    ix = p.inv_e[0][i];
    iy = p.inv_e[1][j];
    iz = p.inv_e[2][k];
    ixn = hi ? p.inv_e[0][i + 1] : 0.f;
    iyn = hj ? p.inv_e[1][j + 1] : 0.f;
    izn = hk ? p.inv_e[2][k + 1] : 0.f;
  } else {
    ix = iy = iz = p.inv_e0;
    ixn = hi ? p.inv_e0 : 0.f;
    iyn = hj ? p.inv_e0 : 0.f;
    izn = hk ? p.inv_e0 : 0.f;
  }
  // v = cb * E-bar at the cell and its forward neighbours
  const float cb0 = ldf(a.cbv, c);
  const float vx = cb0 * ldf(a.a[0], c), vy = cb0 * ldf(a.a[1], c), vz = cb0 * ldf(a.a[2], c);
  const unsigned cj = hj ? c + p.nz : c, ck = hk ? c + 1 : c, ci = hi ? c + P : c;
  const float cbj = ldf(a.cbv, cj), cbk = ldf(a.cbv, ck), cbi = ldf(a.cbv, ci);
  const float vx_j = cbj * ldf(a.a[0], cj), vz_j = cbj * ldf(a.a[2], cj);
  const float vx_k = cbk * ldf(a.a[0], ck), vy_k = cbk * ldf(a.a[1], ck);
  const float vy_i = cbi * ldf(a.a[1], ci), vz_i = cbi * ldf(a.a[2], ci);
  float hx = ldf(a.a[3], c), hy = ldf(a.a[4], c), hz = ldf(a.a[5], c);
  // transpose of E += cb curl_bwd H (yee.py:379-404)
  hz += vx * iy - vx_j * iyn - vy * ix + vy_i * ixn;
  hy += -vx * iz + vx_k * izn + vz * ix - vz_i * ixn;
  hx += vy * iz - vy_k * izn - vz * iy + vz_j * iyn;
  // transpose of the CPML-E differences g = (H(x) - H(x - e)) / d
  if (in_slab(p, i, j, k) || in_slab(p, hi ? i + 1 : i, hj ? j + 1 : j, hk ? k + 1 : k)) {
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      if (p.cn[face] == 0) continue;
      const int ax = face >> 1;
      const float inv = p.cinv[face];
      const int l0 = slab_layer(p, face, i, j, k);
      const bool nb = ax == 0 ? hi : ax == 1 ? hj : hk;
      const int ii = i + (ax == 0), jj = j + (ax == 1), kk = k + (ax == 2);
      const int l1 = nb ? slab_layer(p, face, ii, jj, kk) : -1;
      float d0 = 0.f, d1 = 0.f;
      if (l0 >= 0) {
        const unsigned q = (unsigned)psi_idx(face, l0, i, j, k, p.cn[face], p.ny, p.nz);
        d0 += ldf(a.ge[face][0], q);
        d1 += ldf(a.ge[face][1], q);
      }
      if (l1 >= 0) {
        const unsigned q = (unsigned)psi_idx(face, l1, ii, jj, kk, p.cn[face], p.ny, p.nz);
        d0 -= ldf(a.ge[face][0], q);
        d1 -= ldf(a.ge[face][1], q);
      }
      d0 *= inv;
      d1 *= inv;
      if (ax == 0) {
        hz += d0;
        hy += d1;
      } else if (ax == 1) {
        hz += d0;
        hx += d1;
      } else {
        hy += d0;
        hx += d1;
      }
    }
  }
  // H-probe adjoints of step n
  if (a.h_probe && (p.flags[ldu16(p.mat, c)] & RFXH_POINT) && a.probe_adj) {
    int lo = 0, hi2 = p.n_ops;
    while (lo < hi2) {
      const int mid = (lo + hi2) >> 1;
      if (p.ops[mid].cell < (long long)c) lo = mid + 1;
      else hi2 = mid;
    }
    for (int q = lo; q < p.n_ops && p.ops[q].cell == (long long)c; ++q) {
      const rfxh_point_op op = p.ops[q];
      if (op.kind == RFXH_OP_PROBE && op.comp >= 3) {
        const float g = a.probe_adj[n * p.n_probe + op.slot];
        if (op.comp == 3) hx += g;
        else if (op.comp == 4) hy += g;
        else hz += g;
      }
    }
  }
  // CPML-H memory adjoint (cpml.py:875-1043 transposed)
  if (in_slab(p, i, j, k)) {
    const float ch = ldf(a.chv, c);
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int l = slab_layer(p, face, i, j, k);
      if (l < 0) continue;
      const int ax = face >> 1;
      const unsigned q = (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
      const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
      float s0, s1, h0, h1;  // signs and adjoints of the two H components the face corrects
      if (ax == 0) {
        s0 = 1.f; h0 = hy; s1 = -1.f; h1 = hz;
      } else if (ax == 1) {
        s0 = -1.f; h0 = hx; s1 = 1.f; h1 = hz;
      } else {
        s0 = 1.f; h0 = hx; s1 = -1.f; h1 = hy;
      }
      const float t0 = ldf(a.ph[face][0], q) + s0 * ch * h0;
      const float t1 = ldf(a.ph[face][1], q) + s1 * ch * h1;
      stf(a.gh[face][0], q, cc * t0 + s0 * ch * km1 * h0);
      stf(a.gh[face][1], q, cc * t1 + s1 * ch * km1 * h1);
      stf(a.ph[face][0], q, bb * t0);
      stf(a.ph[face][1], q, bb * t1);
    }
  }
  // mu_r gradient: dJ/dH^{n+1/2} . R^n times d ch / d mu_r
  if (a.grad_mu) {
    const long long r = region_index(a.reg, i, j, k);
    const long long row = n - a.rec_step0;
    if (r >= 0 && row >= 0 && row < a.rec_rows) {
      const long long N = a.reg.n;
      const float* rr = a.rec_h + row * 3 * N + r;
      a.grad_mu[r] += a.dmu[ldu16(p.mat, c)] * (hx * rr[0] + hy * rr[N] + hz * rr[2 * N]);
    }
  }
  stf(a.a[3], c, hx);
  stf(a.a[4], c, hy);
  stf(a.a[5], c, hz);
}

template <bool GRADED>
__device__ __forceinline__ void adj_e_cell(const KParams& p, const AdjParams& a, int i, int j,
                                           int k, int gather, long long m) {
  const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const unsigned P = (unsigned)p.plane;
  const unsigned mid0 = ldu16(p.mat, c);
  const float4 co = p.coef[mid0];
  float ex = ldf(a.a[0], c), ey = ldf(a.a[1], c), ez = ldf(a.a[2], c);
  if (gather) {
    const bool hj = j > 0, hk = k > 0, hi = i > 0;
    float ix, iy, iz, ixp, iyp, izp;  // H-update spacings at the cell and its backward neighbour
    if (GRADED) {
      ix = p.inv_h[0][i];
      iy = p.inv_h[1][j];
      iz = p.inv_h[2][k];
      ixp = hi ? p.inv_h[0][i - 1] : 0.f;
      iyp = hj ? p.inv_h[1][j - 1] : 0.f;
      izp = hk ? p.inv_h[2][k - 1] : 0.f;
    } else {
      ix = iy = iz = p.inv_h0;
      ixp = hi ? p.inv_h0 : 0.f;
      iyp = hj ? p.inv_h0 : 0.f;
      izp = hk ? p.inv_h0 : 0.f;
    }
    const float ch0 = co.w;
    const float ux = ch0 * ldf(a.a[3], c), uy = ch0 * ldf(a.a[4], c), uz = ch0 * ldf(a.a[5], c);
    const unsigned cj = hj ? c - p.nz : c, ck = hk ? c - 1 : c, ci = hi ? c - P : c;
    const float chj = ldf(a.chv, cj), chk = ldf(a.chv, ck), chi = ldf(a.chv, ci);
    const float ux_j = chj * ldf(a.a[3], cj), uz_j = chj * ldf(a.a[5], cj);
    const float ux_k = chk * ldf(a.a[3], ck), uy_k = chk * ldf(a.a[4], ck);
    const float uy_i = chi * ldf(a.a[4], ci), uz_i = chi * ldf(a.a[5], ci);
    // transpose of H -= ch curl_fwd E (yee.py:272-289), plus ca E-bar
    ex = co.x * ex + uy * iz - uy_k * izp - uz * iy + uz_j * iyp;
    ey = co.x * ey - ux * iz + ux_k * izp + uz * ix - uz_i * ixp;
    ez = co.x * ez + ux * iy - ux_j * iyp - uy * ix + uy_i * ixp;
    // transpose of the CPML-H differences g = (E(x + e) - E(x)) / d
    if (in_slab(p, i, j, k) || in_slab(p, hi ? i - 1 : i, hj ? j - 1 : j, hk ? k - 1 : k)) {
#pragma unroll
      for (int face = 0; face < 6; ++face) {
        if (p.cn[face] == 0) continue;
        const int ax = face >> 1;
        const float inv = p.cinv[face];
        const int l0 = slab_layer(p, face, i, j, k);
        const bool nb = ax == 0 ? hi : ax == 1 ? hj : hk;
        const int ii = i - (ax == 0), jj = j - (ax == 1), kk = k - (ax == 2);
        const int l1 = nb ? slab_layer(p, face, ii, jj, kk) : -1;
        float d0 = 0.f, d1 = 0.f;
        if (l1 >= 0) {
          const unsigned q = (unsigned)psi_idx(face, l1, ii, jj, kk, p.cn[face], p.ny, p.nz);
          d0 += ldf(a.gh[face][0], q);
          d1 += ldf(a.gh[face][1], q);
        }
        if (l0 >= 0) {
          const unsigned q = (unsigned)psi_idx(face, l0, i, j, k, p.cn[face], p.ny, p.nz);
          d0 -= ldf(a.gh[face][0], q);
          d1 -= ldf(a.gh[face][1], q);
        }
        d0 *= inv;
        d1 *= inv;
        if (ax == 0) {
          ez += d0;
          ey += d1;
        } else if (ax == 1) {
          ez += d0;
          ex += d1;
        } else {
          ey += d0;
          ex += d1;
        }
      }
    }
  }
  if (m >= 0) {
    // prologue of step m: probes read E after injection, injection follows
    // the PEC projection (rfx/simulation.py:1392-1567)
    const uint32_t fl = p.flags[mid0];
    if (fl & RFXH_POINT) {
      int lo = 0, hi2 = p.n_ops;
      while (lo < hi2) {
        const int md = (lo + hi2) >> 1;
        if (p.ops[md].cell < (long long)c) lo = md + 1;
        else hi2 = md;
      }
      for (int q = lo; q < p.n_ops && p.ops[q].cell == (long long)c; ++q) {
        const rfxh_point_op op = p.ops[q];
        if (op.kind == RFXH_OP_PROBE && op.comp < 3 && a.probe_adj) {
          const float g = a.probe_adj[m * p.n_probe + op.slot];
          if (op.comp == 0) ex += g;
          else if (op.comp == 1) ey += g;
          else ez += g;
        }
      }
      for (int q = lo; q < p.n_ops && p.ops[q].cell == (long long)c; ++q) {
        const rfxh_point_op op = p.ops[q];
        if (op.kind == RFXH_OP_INJECT && a.src_adj)
          a.src_adj[m * p.n_src + op.slot] = pick3(op.comp, ex, ey, ez);
      }
      // port voltage V = -E_c d_par, sampled before the injection: its
      // adjoint is not part of dJ/d(source table)
      if (a.port_adj)
        for (int q = lo; q < p.n_ops && p.ops[q].cell == (long long)c; ++q) {
          const rfxh_point_op op = p.ops[q];
          if (op.kind != RFXH_OP_PORT) continue;
          const float g = -op.d_par * a.port_adj[(m * p.n_port + op.slot) * 2];
          if (op.comp == 0) ex += g;
          else if (op.comp == 1) ey += g;
          else ez += g;
        }
    }
    if (p.pec_axes & 1) {
      if (i == 0 || i == p.nx - 1) ey = ez = 0.f;
    }
    if (p.pec_axes & 2) {
      if (j == 0 || j == p.ny - 1) ex = ez = 0.f;
    }
    if (p.pec_axes & 4) {
      if (k == 0 || k == p.nz - 1) ex = ey = 0.f;
      if (k == p.nz - 1) ez = 0.f;
    }
    if (fl & RFXH_ZERO_EX) ex = 0.f;
    if (fl & RFXH_ZERO_EY) ey = 0.f;
    if (fl & RFXH_ZERO_EZ) ez = 0.f;
    // CPML-E memory adjoint (cpml.py:608-777 transposed)
    if (in_slab(p, i, j, k)) {
      const float ce = co.z;
#pragma unroll
      for (int face = 0; face < 6; ++face) {
        const int l = slab_layer(p, face, i, j, k);
        if (l < 0) continue;
        const int ax = face >> 1;
        const unsigned q = (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
        const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
        float s0, s1, e0v, e1v;
        if (ax == 0) {
          s0 = -1.f; e0v = ey; s1 = 1.f; e1v = ez;
        } else if (ax == 1) {
          s0 = 1.f; e0v = ex; s1 = -1.f; e1v = ez;
        } else {
          s0 = -1.f; e0v = ex; s1 = 1.f; e1v = ey;
        }
        const float t0 = ldf(a.pe[face][0], q) + s0 * ce * e0v;
        const float t1 = ldf(a.pe[face][1], q) + s1 * ce * e1v;
        stf(a.ge[face][0], q, cc * t0 + s0 * ce * km1 * e0v);
        stf(a.ge[face][1], q, cc * t1 + s1 * ce * km1 * e1v);
        stf(a.pe[face][0], q, bb * t0);
        stf(a.pe[face][1], q, bb * t1);
      }
    }
    // eps_r gradient from the forward record of step m
    if (a.grad) {
      const long long r = region_index(a.reg, i, j, k);
      const long long row = m - a.rec_step0;
      if (r >= 0 && row >= 0 && row < a.rec_rows) {
        const float* rr = a.rec + row * 9 * a.reg.n + r;
        const long long N = a.reg.n;
        const float A = ex * rr[0] + ey * rr[N] + ez * rr[2 * N];
        const float B = ex * rr[3 * N] + ey * rr[4 * N] + ez * rr[5 * N];
        // the record holds the CPML-E rows only for slab cells
        const float C = in_slab(p, i, j, k) ? ex * rr[6 * N] + ey * rr[7 * N] + ez * rr[8 * N]
                                            : 0.f;
        const float4 dc = a.dcoef[mid0];
        a.grad[r] += dc.x * A + dc.y * B + dc.z * C;
        if (a.grad_s) {
          const float2 ds = a.dcoef_s[mid0];
          a.grad_s[r] += ds.x * A + ds.y * B;
        }
      }
    }
  }
  stf(a.a[0], c, ex);
  stf(a.a[1], c, ey);
  stf(a.a[2], c, ez);
}

// Port current adjoint of step n into H-bar^{n+1/2}: the Ampere-loop
// current I = sum of +-loop * H over the cell and its backward neighbours
// (point_ops), transposed; one thread per point operation (few ports),
// before the H-bar pass of step n, so that pass sees the whole
// dJ/dH^{n+1/2}.
__global__ void k_port_adj_h(KParams p, AdjParams a, long long n) {
  const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= p.n_ops) return;
  const rfxh_point_op op = p.ops[q];
  if (op.kind != RFXH_OP_PORT) return;
  const float g = a.port_adj[(n * p.n_port + op.slot) * 2 + 1];
  const long long c = op.cell;
  const int k = (int)(c % p.nz), j = (int)((c / p.nz) % p.ny), i = (int)(c / p.plane);
  const long long P = p.plane;
  float* const hx = a.a[3];
  float* const hy = a.a[4];
  float* const hz = a.a[5];
  if (op.comp == 2) {         // I = (hy - hy_i) l1 - (hx - hx_j) l0
    atomicAdd(hy + c, g * op.loop[1]);
    if (i > 0) atomicAdd(hy + c - P, -g * op.loop[1]);
    atomicAdd(hx + c, -g * op.loop[0]);
    if (j > 0) atomicAdd(hx + c - p.nz, g * op.loop[0]);
  } else if (op.comp == 0) {  // I = (hz - hz_j) l2 - (hy - hy_k) l1
    atomicAdd(hz + c, g * op.loop[2]);
    if (j > 0) atomicAdd(hz + c - p.nz, -g * op.loop[2]);
    atomicAdd(hy + c, -g * op.loop[1]);
    if (k > 0) atomicAdd(hy + c - 1, g * op.loop[1]);
  } else {                    // I = (hx - hx_k) l0 - (hz - hz_i) l2
    atomicAdd(hx + c, g * op.loop[0]);
    if (k > 0) atomicAdd(hx + c - 1, -g * op.loop[0]);
    atomicAdd(hz + c, -g * op.loop[2]);
    if (i > 0) atomicAdd(hz + c - P, g * op.loop[2]);
  }
}

// DFT-plane adjoint of step n into the adjoint field `tgt` (the plane's
// component): acc[f] += tw[n][f] x over the steps, so dJ/dx[n] = sum_f
// Re(conj(G[f]) tw[n][f]) with G = dJ/dRe(acc) + j dJ/dIm(acc).
__global__ __launch_bounds__(256) void k_plane_adj(PlaneParams q, const double2* __restrict__ G,
                                                   long long n, float* tgt) {
  const long long npts = (long long)q.n1 * q.n2;
  const long long pidx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= npts || n < 0 || n >= q.max_steps) return;
  const int u = (int)(pidx / q.n2), v = (int)(pidx % q.n2);
  int ii[3];
  const int a1 = q.axis == 0 ? 1 : 0, a2 = q.axis == 2 ? 1 : 2;
  ii[q.axis] = q.index;
  ii[a1] = u;
  ii[a2] = v;
  double g = 0.0;
  for (int f = 0; f < q.nf; ++f) {
    const double2 gf = G[(long long)f * npts + pidx];
    const double2 tw = q.tw[n * q.nf + f];
    g += gf.x * tw.x + gf.y * tw.y;
  }
  const long long c = ((long long)ii[0] * q.ny + ii[1]) * q.nz + ii[2];
  tgt[c] += (float)g;
}

// Flux-monitor / NTFF-face adjoint of step n: component t of the monitor
// (E1, E2 when e_part, else H1, H2; blockIdx.y picks one of the two) takes
// g = sum_f Re(conj(G[t][f]) tw[n][f]) into its adjoint field tgt[y]; the
// H components were sampled as the mean of planes index_m1 and index, so
// each gets half (all of it when the two planes coincide).
__global__ __launch_bounds__(256) void k_flux_adj(FluxParams q, const double2* G, long long n,
                                                  int e_part, float* tgt0, float* tgt1) {
  const long long npts = (long long)q.n1 * q.n2;
  const long long pidx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= npts || n < 0 || n >= q.max_steps) return;
  const int y = (int)blockIdx.y;
  const int t = (e_part ? 0 : 2) + y;
  float* tgt = y ? tgt1 : tgt0;
  const int u = (int)(pidx / q.n2) + q.lo1, v = (int)(pidx % q.n2) + q.lo2;
  int ii[3], im[3];
  const int a1 = q.axis == 0 ? 1 : 0, a2 = q.axis == 2 ? 1 : 2;
  ii[q.axis] = q.index;
  im[q.axis] = q.index_m1;
  ii[a1] = im[a1] = u;
  ii[a2] = im[a2] = v;
  const double2* tw = t < 2 ? q.twe : q.twh;
  double g = 0.0;
  for (int f = 0; f < q.nf; ++f) {
    const double2 gf = G[((long long)t * q.nf + f) * npts + pidx];
    const double2 w = tw[n * q.nf + f];
    g += gf.x * w.x + gf.y * w.y;
  }
  const long long c = ((long long)ii[0] * q.ny + ii[1]) * q.nz + ii[2];
  const long long cm = ((long long)im[0] * q.ny + im[1]) * q.nz + im[2];
  if (t < 2 || cm == c) {
    tgt[c] += (float)g;
  } else {
    tgt[cm] += (float)(0.5 * g);
    tgt[c] += (float)(0.5 * g);
  }
}

// The single-instance reverse kernels: the bodies of adj_h_cell and
// adj_e_cell written out in the kernels (the same arithmetic; through a
// device function taking the by-value kernel arguments by reference the
// compiler copies them into registers whole and spills, 4 % slower on a
// 277 x 277 x 149 grid).  The batched kernels below call the functions.
template <bool GRADED>
__global__ __launch_bounds__(kWG) void k_adj_h(KParams p, AdjParams a, long long n) {
  int i, j, k;
  if (!cell_of(p, i, j, k)) return;
  const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const unsigned P = (unsigned)p.plane;
  const bool hj = j + 1 < p.ny, hk = k + 1 < p.nz, hi = i + 1 < p.nx;
  float ix, iy, iz, ixn, iyn, izn;  // E-update spacings at the cell and at its forward neighbour
  if (GRADED) {
    ix = p.inv_e[0][i];
    iy = p.inv_e[1][j];
    iz = p.inv_e[2][k];
    ixn = hi ? p.inv_e[0][i + 1] : 0.f;
    iyn = hj ? p.inv_e[1][j + 1] : 0.f;
    izn = hk ? p.inv_e[2][k + 1] : 0.f;
  } else {
    ix = iy = iz = p.inv_e0;
    ixn = hi ? p.inv_e0 : 0.f;
    iyn = hj ? p.inv_e0 : 0.f;
    izn = hk ? p.inv_e0 : 0.f;
  }
  // v = cb * E-bar at the cell and its forward neighbours
  const float cb0 = ldf(a.cbv, c);
  const float vx = cb0 * ldf(a.a[0], c), vy = cb0 * ldf(a.a[1], c), vz = cb0 * ldf(a.a[2], c);
  const unsigned cj = hj ? c + p.nz : c, ck = hk ? c + 1 : c, ci = hi ? c + P : c;
  const float cbj = ldf(a.cbv, cj), cbk = ldf(a.cbv, ck), cbi = ldf(a.cbv, ci);
  const float vx_j = cbj * ldf(a.a[0], cj), vz_j = cbj * ldf(a.a[2], cj);
  const float vx_k = cbk * ldf(a.a[0], ck), vy_k = cbk * ldf(a.a[1], ck);
  const float vy_i = cbi * ldf(a.a[1], ci), vz_i = cbi * ldf(a.a[2], ci);
  float hx = ldf(a.a[3], c), hy = ldf(a.a[4], c), hz = ldf(a.a[5], c);
  // transpose of E += cb curl_bwd H (yee.py:379-404)
  hz += vx * iy - vx_j * iyn - vy * ix + vy_i * ixn;
  hy += -vx * iz + vx_k * izn + vz * ix - vz_i * ixn;
  hx += vy * iz - vy_k * izn - vz * iy + vz_j * iyn;
  // transpose of the CPML-E differences g = (H(x) - H(x - e)) / d
  if (in_slab(p, i, j, k) || in_slab(p, hi ? i + 1 : i, hj ? j + 1 : j, hk ? k + 1 : k)) {
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      if (p.cn[face] == 0) continue;
      const int ax = face >> 1;
      const float inv = p.cinv[face];
      const int l0 = slab_layer(p, face, i, j, k);
      const bool nb = ax == 0 ? hi : ax == 1 ? hj : hk;
      const int ii = i + (ax == 0), jj = j + (ax == 1), kk = k + (ax == 2);
      const int l1 = nb ? slab_layer(p, face, ii, jj, kk) : -1;
      float d0 = 0.f, d1 = 0.f;
      if (l0 >= 0) {
        const unsigned q = (unsigned)psi_idx(face, l0, i, j, k, p.cn[face], p.ny, p.nz);
        d0 += ldf(a.ge[face][0], q);
        d1 += ldf(a.ge[face][1], q);
      }
      if (l1 >= 0) {
        const unsigned q = (unsigned)psi_idx(face, l1, ii, jj, kk, p.cn[face], p.ny, p.nz);
        d0 -= ldf(a.ge[face][0], q);
        d1 -= ldf(a.ge[face][1], q);
      }
      d0 *= inv;
      d1 *= inv;
      if (ax == 0) {
        hz += d0;
        hy += d1;
      } else if (ax == 1) {
        hz += d0;
        hx += d1;
      } else {
        hy += d0;
        hx += d1;
      }
    }
  }
  // H-probe adjoints of step n
  if (a.h_probe && (p.flags[ldu16(p.mat, c)] & RFXH_POINT) && a.probe_adj) {
    int lo = 0, hi2 = p.n_ops;
    while (lo < hi2) {
      const int mid = (lo + hi2) >> 1;
      if (p.ops[mid].cell < (long long)c) lo = mid + 1;
      else hi2 = mid;
    }
    for (int q = lo; q < p.n_ops && p.ops[q].cell == (long long)c; ++q) {
      const rfxh_point_op op = p.ops[q];
      if (op.kind == RFXH_OP_PROBE && op.comp >= 3) {
        const float g = a.probe_adj[n * p.n_probe + op.slot];
        if (op.comp == 3) hx += g;
        else if (op.comp == 4) hy += g;
        else hz += g;
      }
    }
  }
  // CPML-H memory adjoint (cpml.py:875-1043 transposed)
  if (in_slab(p, i, j, k)) {
    const float ch = ldf(a.chv, c);
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int l = slab_layer(p, face, i, j, k);
      if (l < 0) continue;
      const int ax = face >> 1;
      const unsigned q = (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
      const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
      float s0, s1, h0, h1;  // signs and adjoints of the two H components the face corrects
      if (ax == 0) {
        s0 = 1.f; h0 = hy; s1 = -1.f; h1 = hz;
      } else if (ax == 1) {
        s0 = -1.f; h0 = hx; s1 = 1.f; h1 = hz;
      } else {
        s0 = 1.f; h0 = hx; s1 = -1.f; h1 = hy;
      }
      const float t0 = ldf(a.ph[face][0], q) + s0 * ch * h0;
      const float t1 = ldf(a.ph[face][1], q) + s1 * ch * h1;
      stf(a.gh[face][0], q, cc * t0 + s0 * ch * km1 * h0);
      stf(a.gh[face][1], q, cc * t1 + s1 * ch * km1 * h1);
      stf(a.ph[face][0], q, bb * t0);
      stf(a.ph[face][1], q, bb * t1);
    }
  }
  // mu_r gradient: dJ/dH^{n+1/2} . R^n times d ch / d mu_r
  if (a.grad_mu) {
    const long long r = region_index(a.reg, i, j, k);
    const long long row = n - a.rec_step0;
    if (r >= 0 && row >= 0 && row < a.rec_rows) {
      const long long N = a.reg.n;
      const float* rr = a.rec_h + row * 3 * N + r;
      a.grad_mu[r] += a.dmu[ldu16(p.mat, c)] * (hx * rr[0] + hy * rr[N] + hz * rr[2 * N]);
    }
  }
  stf(a.a[3], c, hx);
  stf(a.a[4], c, hy);
  stf(a.a[5], c, hz);
}

// Per-cell cb and ch of the single-instance reverse sweep.
__global__ __launch_bounds__(256) void k_fill_cbch(KParams p, float* cbv, float* chv, long long n) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const float4 co = p.coef[ldu16(p.mat, (unsigned)c)];
  cbv[c] = co.y;
  chv[c] = co.w;
}

// Row of a cell with poles in the adjoint polarisation box (-1: none); R =
// the box's cell count.
__device__ __forceinline__ long long adj_pol_row(const AdjParams& a, int i, int j, int k,
                                                 uint32_t fl, unsigned& R) {
  const int li = i - a.pol_lo[0], lj = j - a.pol_lo[1], lk = k - a.pol_lo[2];
  R = (unsigned)a.pol_n[0] * (unsigned)a.pol_n[1] * (unsigned)a.pol_n[2];
  if (!(fl & RFXH_POLES) || !a.pol_adj || li < 0 || lj < 0 || lk < 0 || li >= a.pol_n[0] ||
      lj >= a.pol_n[1] || lk >= a.pol_n[2])
    return -1;
  return ((long long)li * a.pol_n[1] + lj) * a.pol_n[2] + lk;
}

template <bool GRADED, bool DISP = false>
__global__ __launch_bounds__(kWG) void k_adj_e(KParams p, AdjParams a, int gather, long long m) {
  int i, j, k;
  if (!cell_of(p, i, j, k)) return;
  const unsigned c = ((unsigned)i * (unsigned)p.ny + (unsigned)j) * (unsigned)p.nz + (unsigned)k;
  const unsigned P = (unsigned)p.plane;
  const unsigned mid0 = ldu16(p.mat, c);
  const float4 co = p.coef[mid0];
  float ex = ldf(a.a[0], c), ey = ldf(a.a[1], c), ez = ldf(a.a[2], c);
  const float bx = ex, by = ey, bz = ez;   // adjoint of E^{m+2} before the E update's transpose
  if (gather) {
    const bool hj = j > 0, hk = k > 0, hi = i > 0;
    float ix, iy, iz, ixp, iyp, izp;  // H-update spacings at the cell and its backward neighbour
    if (GRADED) {
      ix = p.inv_h[0][i];
      iy = p.inv_h[1][j];
      iz = p.inv_h[2][k];
      ixp = hi ? p.inv_h[0][i - 1] : 0.f;
      iyp = hj ? p.inv_h[1][j - 1] : 0.f;
      izp = hk ? p.inv_h[2][k - 1] : 0.f;
    } else {
      ix = iy = iz = p.inv_h0;
      ixp = hi ? p.inv_h0 : 0.f;
      iyp = hj ? p.inv_h0 : 0.f;
      izp = hk ? p.inv_h0 : 0.f;
    }
    const float ch0 = co.w;
    const float ux = ch0 * ldf(a.a[3], c), uy = ch0 * ldf(a.a[4], c), uz = ch0 * ldf(a.a[5], c);
    const unsigned cj = hj ? c - p.nz : c, ck = hk ? c - 1 : c, ci = hi ? c - P : c;
    const float chj = ldf(a.chv, cj), chk = ldf(a.chv, ck), chi = ldf(a.chv, ci);
    const float ux_j = chj * ldf(a.a[3], cj), uz_j = chj * ldf(a.a[5], cj);
    const float ux_k = chk * ldf(a.a[3], ck), uy_k = chk * ldf(a.a[4], ck);
    const float uy_i = chi * ldf(a.a[4], ci), uz_i = chi * ldf(a.a[5], ci);
    // transpose of H -= ch curl_fwd E (yee.py:272-289), plus ca E-bar
    ex = co.x * ex + uy * iz - uy_k * izp - uz * iy + uz_j * iyp;
    ey = co.x * ey - ux * iz + ux_k * izp + uz * ix - uz_i * ixp;
    ez = co.x * ez + ux * iy - ux_j * iyp - uy * ix + uy_i * ixp;
    // transpose of the CPML-H differences g = (E(x + e) - E(x)) / d
    if (in_slab(p, i, j, k) || in_slab(p, hi ? i - 1 : i, hj ? j - 1 : j, hk ? k - 1 : k)) {
#pragma unroll
      for (int face = 0; face < 6; ++face) {
        if (p.cn[face] == 0) continue;
        const int ax = face >> 1;
        const float inv = p.cinv[face];
        const int l0 = slab_layer(p, face, i, j, k);
        const bool nb = ax == 0 ? hi : ax == 1 ? hj : hk;
        const int ii = i - (ax == 0), jj = j - (ax == 1), kk = k - (ax == 2);
        const int l1 = nb ? slab_layer(p, face, ii, jj, kk) : -1;
        float d0 = 0.f, d1 = 0.f;
        if (l1 >= 0) {
          const unsigned q = (unsigned)psi_idx(face, l1, ii, jj, kk, p.cn[face], p.ny, p.nz);
          d0 += ldf(a.gh[face][0], q);
          d1 += ldf(a.gh[face][1], q);
        }
        if (l0 >= 0) {
          const unsigned q = (unsigned)psi_idx(face, l0, i, j, k, p.cn[face], p.ny, p.nz);
          d0 -= ldf(a.gh[face][0], q);
          d1 -= ldf(a.gh[face][1], q);
        }
        d0 *= inv;
        d1 *= inv;
        if (ax == 0) {
          ez += d0;
          ey += d1;
        } else if (ax == 1) {
          ez += d0;
          ex += d1;
        } else {
          ey += d0;
          ex += d1;
        }
      }
    }
    if (DISP) {
      // transpose of the ADE (dispersion_cell) at this cell: with e-bar the
      // adjoint of the new E before the CPML term, P-bar / Q-bar those of
      // the new polarisation: Debye P' = alpha P + beta (E' + E) and E' +=
      // cc P give E-bar += beta P-bar', P-bar = alpha P-bar' + cc e-bar;
      // Lorentz P' = a P + b Q + c E, Q' = P and E' -= ccl (P' - P) give
      // E-bar += c P-bar' - ccl c e-bar, P-bar = a P-bar' + Q-bar' - ccl
      // (a - 1) e-bar, Q-bar = b P-bar' - ccl b e-bar
      unsigned R;
      const long long r = adj_pol_row(a, i, j, k, p.flags[mid0], R);
      if (r >= 0) {
        const float* pc = a.pcoef + (size_t)mid0 * a.n_pole * 4;
        const int nd = a.n_debye;
        const float eb[3] = {bx, by, bz};
        float ed[3] = {0.f, 0.f, 0.f};
        for (int q = 0; q < nd; ++q) {
          const float al = pc[4 * q], be = pc[4 * q + 1], cc = pc[4 * q + 2];
          float* Pb = a.pol_adj + (size_t)q * 3 * R;
#pragma unroll
          for (int v = 0; v < 3; ++v) {
            const float pb = Pb[(size_t)v * R + r];
            ed[v] += be * pb;
            Pb[(size_t)v * R + r] = al * pb + cc * eb[v];
          }
        }
        for (int q = nd; q < a.n_pole; ++q) {
          const float A = pc[4 * q], B = pc[4 * q + 1], cl = pc[4 * q + 2], ccl = pc[4 * q + 3];
          float* Pb = a.pol_adj + (size_t)nd * 3 * R + (size_t)(q - nd) * 6 * R;
#pragma unroll
          for (int v = 0; v < 3; ++v) {
            const float pb = Pb[(size_t)v * R + r], qb = Pb[(size_t)(3 + v) * R + r];
            ed[v] += cl * pb - ccl * cl * eb[v];
            Pb[(size_t)v * R + r] = A * pb + qb - ccl * (A - 1.f) * eb[v];
            Pb[(size_t)(3 + v) * R + r] = B * pb - ccl * B * eb[v];
          }
        }
        ex += ed[0];
        ey += ed[1];
        ez += ed[2];
      }
    }
  }
  if (m >= 0) {
    // prologue of step m: probes read E after injection, injection follows
    // the PEC projection (rfx/simulation.py:1392-1567)
    const uint32_t fl = p.flags[mid0];
    if (fl & RFXH_POINT) {
      int lo = 0, hi2 = p.n_ops;
      while (lo < hi2) {
        const int md = (lo + hi2) >> 1;
        if (p.ops[md].cell < (long long)c) lo = md + 1;
        else hi2 = md;
      }
      for (int q = lo; q < p.n_ops && p.ops[q].cell == (long long)c; ++q) {
        const rfxh_point_op op = p.ops[q];
        if (op.kind == RFXH_OP_PROBE && op.comp < 3 && a.probe_adj) {
          const float g = a.probe_adj[m * p.n_probe + op.slot];
          if (op.comp == 0) ex += g;
          else if (op.comp == 1) ey += g;
          else ez += g;
        }
      }
      for (int q = lo; q < p.n_ops && p.ops[q].cell == (long long)c; ++q) {
        const rfxh_point_op op = p.ops[q];
        if (op.kind == RFXH_OP_INJECT && a.src_adj)
          a.src_adj[m * p.n_src + op.slot] = pick3(op.comp, ex, ey, ez);
      }
      // port voltage V = -E_c d_par, sampled before the injection: its
      // adjoint is not part of dJ/d(source table)
      if (a.port_adj)
        for (int q = lo; q < p.n_ops && p.ops[q].cell == (long long)c; ++q) {
          const rfxh_point_op op = p.ops[q];
          if (op.kind != RFXH_OP_PORT) continue;
          const float g = -op.d_par * a.port_adj[(m * p.n_port + op.slot) * 2];
          if (op.comp == 0) ex += g;
          else if (op.comp == 1) ey += g;
          else ez += g;
        }
    }
    if (p.pec_axes & 1) {
      if (i == 0 || i == p.nx - 1) ey = ez = 0.f;
    }
    if (p.pec_axes & 2) {
      if (j == 0 || j == p.ny - 1) ex = ez = 0.f;
    }
    if (p.pec_axes & 4) {
      if (k == 0 || k == p.nz - 1) ex = ey = 0.f;
      if (k == p.nz - 1) ez = 0.f;
    }
    if (fl & RFXH_ZERO_EX) ex = 0.f;
    if (fl & RFXH_ZERO_EY) ey = 0.f;
    if (fl & RFXH_ZERO_EZ) ez = 0.f;
    // CPML-E memory adjoint (cpml.py:608-777 transposed)
    if (in_slab(p, i, j, k)) {
      const float ce = co.z;
#pragma unroll
      for (int face = 0; face < 6; ++face) {
        const int l = slab_layer(p, face, i, j, k);
        if (l < 0) continue;
        const int ax = face >> 1;
        const unsigned q = (unsigned)psi_idx(face, l, i, j, k, p.cn[face], p.ny, p.nz);
        const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
        float s0, s1, e0v, e1v;
        if (ax == 0) {
          s0 = -1.f; e0v = ey; s1 = 1.f; e1v = ez;
        } else if (ax == 1) {
          s0 = 1.f; e0v = ex; s1 = -1.f; e1v = ez;
        } else {
          s0 = -1.f; e0v = ex; s1 = 1.f; e1v = ey;
        }
        const float t0 = ldf(a.pe[face][0], q) + s0 * ce * e0v;
        const float t1 = ldf(a.pe[face][1], q) + s1 * ce * e1v;
        stf(a.ge[face][0], q, cc * t0 + s0 * ce * km1 * e0v);
        stf(a.ge[face][1], q, cc * t1 + s1 * ce * km1 * e1v);
        stf(a.pe[face][0], q, bb * t0);
        stf(a.pe[face][1], q, bb * t1);
      }
    }
    // the CPML term's adjoint is E-bar as it stands; the Debye poles read
    // the new E before that term: E-bar of it adds sum_d beta_d P-bar_d
    const float fx = ex, fy = ey, fz = ez;
    if (DISP) {
      unsigned R;
      const long long r = adj_pol_row(a, i, j, k, fl, R);
      if (r >= 0) {
        const float* pc = a.pcoef + (size_t)mid0 * a.n_pole * 4;
        for (int q = 0; q < a.n_debye; ++q) {
          const float be = pc[4 * q + 1];
          const float* Pb = a.pol_adj + (size_t)q * 3 * R;
          ex += be * Pb[r];
          ey += be * Pb[R + r];
          ez += be * Pb[2 * (size_t)R + r];
        }
      }
    }
    // eps_r gradient from the forward record of step m
    if (a.grad) {
      const long long r = region_index(a.reg, i, j, k);
      const long long row = m - a.rec_step0;
      if (r >= 0 && row >= 0 && row < a.rec_rows) {
        const float* rr = a.rec + row * 9 * a.reg.n + r;
        const long long N = a.reg.n;
        const float A = ex * rr[0] + ey * rr[N] + ez * rr[2 * N];
        const float B = ex * rr[3 * N] + ey * rr[4 * N] + ez * rr[5 * N];
        // the record holds the CPML-E rows only for slab cells
        const float C = in_slab(p, i, j, k) ? fx * rr[6 * N] + fy * rr[7 * N] + fz * rr[8 * N]
                                            : 0.f;
        const float4 dc = a.dcoef[mid0];
        float K = 0.f;
        if (DISP && a.rec_k) {
          const float* rk = a.rec_k + row * 3 * N + r;
          K = ex * rk[0] + ey * rk[N] + ez * rk[2 * N];
        }
        a.grad[r] += dc.x * A + dc.y * B + dc.z * C + (DISP ? a.dkdis[mid0].x * K : 0.f);
        if (a.grad_s) {
          const float2 ds = a.dcoef_s[mid0];
          a.grad_s[r] += ds.x * A + ds.y * B + (DISP ? a.dkdis[mid0].y * K : 0.f);
        }
      }
    }
  }
  stf(a.a[0], c, ex);
  stf(a.a[1], c, ey);
  stf(a.a[2], c, ez);
}

// ---------------------------------------------------------------------------
// Batched reverse mode: the forward record and the reverse sweep of many
// instances per launch, over a batch's instance table as k_batch_pass does
// for the forward (instance q owns logical workgroups [start[q],
// start[q+1]), all instance parameters read through the table).
// ---------------------------------------------------------------------------
struct BatchAdjEntry {
  AdjParams a;
  long long lo, hi;      // this instance's reverse sweep: steps hi-1 down to lo
};

struct BatchRecEntry {
  RegionParams r;
  float* rec;            // [rows][9][r.n]
  long long rec_step0;   // step of row 0
};

// E pass of the active instances, each cell of an instance's region first
// recording its E^n, curl H^{n+1/2} and CPML-E term at the instance's
// current step (the update then finds the same inputs in cache).  Plain
// isotropic plans only (the record path refuses the others).
__global__ __launch_bounds__(kWG) void k_batch_e_rec(const BatchEntry* __restrict__ ents,
                                                     const BatchRecEntry* __restrict__ recs,
                                                     const int* __restrict__ start, int n_act) {
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int q = batch_instance(start, n_act, L);
  const KParams& p = ents[q].p;
  int i, j, k;
  if (!cell_at(p, L - gld(start, q), i, j, k)) return;
  const BatchRecEntry& e = recs[q];
  const long long t = region_index(e.r, i, j, k);
  if (t >= 0) {
    float* rec = e.rec + (*p.step - e.rec_step0) * 9 * e.r.n;
    if (p.graded) record_cell<true>(p, ents[q].ps, e.r, t, rec);
    else record_cell<false>(p, ents[q].ps, e.r, t, rec);
  }
  if (p.graded) two_pass_e<true, false>(p, ents[q].ps, i, j, k);
  else two_pass_e<false, false>(p, ents[q].ps, i, j, k);
}

// Reverse step r (counted from each instance's hi) of the active instances:
// E = false the H-bar gather of step n = hi-1-r; E = true the E-bar gather
// and the prologue of step n-1 (gather = 0: only the prologue of step
// hi-1, before the first reverse step).  GM: 0 every instance uniform, 1
// every instance graded, 2 mixed (one body per lane kind, chosen per
// instance).
template <bool E, int GM>
__global__ __launch_bounds__(kWG) void k_batch_adj(const BatchEntry* __restrict__ ents,
                                                   const BatchAdjEntry* __restrict__ adj,
                                                   const int* __restrict__ start, int n_act,
                                                   int gather, long long r) {
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int q = batch_instance(start, n_act, L);
  const KParams& p = ents[q].p;
  int i, j, k;
  if (!cell_at(p, L - gld(start, q), i, j, k)) return;
  const BatchAdjEntry& e = adj[q];
  const long long n = e.hi - 1 - r;
  const bool graded = GM == 2 ? p.graded != 0 : GM == 1;
  if (E) {
    const long long m = !gather ? n : n - 1 >= e.lo ? n - 1 : -1;
    if (GM != 0 && graded) adj_e_cell<true>(p, e.a, i, j, k, gather, m);
    else if (GM != 1) adj_e_cell<false>(p, e.a, i, j, k, gather, m);
  } else {
    if (GM != 0 && graded) adj_h_cell<true>(p, e.a, i, j, k, n);
    else if (GM != 1) adj_h_cell<false>(p, e.a, i, j, k, n);
  }
}

__global__ void k_set_step(long long* step, long long value) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *step = value;
}

__global__ void k_advance(long long* step) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *step = *step + 1;
}

__global__ void k_advance2(long long* step) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *step = *step + 2;
}

}  // namespace

struct rfxh_plan {
  KParams kp;
  // Debye / Lorentz polarisation state (caller-owned) and per-material pole coefficients
  float* pol = nullptr;
  const float* pcoef = nullptr;
  int n_pole = 0, n_debye = 0;
  int pol_lo[3] = {0, 0, 0}, pol_n[3] = {0, 0, 0};
  dim3 grid, block;
  dim3 grid_col;                // column kernels (ncol cells per thread)
  bool fused = false;           // ping-pong step over two field sets (fused or marching kernel)
  bool march = false;           // the ping-pong step uses the marching kernel
  int march_seg = 0;            // x segments of the marching kernel (0: automatic)
  int march_tile = 0;           // index into kMarchRows
  int ncol = 1;                 // two-pass step: cells per thread along x (1, 2, 3, 4)
  bool wmarch = false;          // the ping-pong step uses the wave-march kernel
  int wmarch_seg = 0;           // its x segments (0: automatic)
  bool aniso = false;           // ids with per-component eps (subpixel smoothing)
  bool tb = false;              // two steps per sweep (temporal-blocked march), ping-pong
  int tb_seg = 0;               // its x segments (0: automatic)
  int tb_ops_max = 0;           // most point operations in one of its tile columns
  float* psi_base = nullptr;    // CPML memory of set 0 (set 1 at + psi_set)
  float* psi_mid = nullptr;     // the sweep's memory after its first step (owned, same layout)
  long long psi_floats = 0;
  float4* cprof = nullptr;      // CPML profile rows (b, c, 1/kappa - 1, 0) of every face
  int cprof_off[6] = {0, 0, 0, 0, 0, 0};
  unsigned psi_foff[6] = {0, 0, 0, 0, 0, 0};  // float offset of each face's terms
  unsigned psi_fs[6] = {0, 0, 0, 0, 0, 0};    // floats per term of each face
  float* set[2][6];             // field sets (set[1] null in two-pass mode)
  long long psi_set = 0;        // float offset of CPML set 1 from set 0
  long long* acc_from = nullptr;  // first step not yet folded into the DFT planes (device)
  PsiIO* d_psio = nullptr;      // in-place CPML read/write pointers of set 0 and set 1 (device)
  int cur = 0;                  // which set holds the current state
  std::vector<PlaneParams> planes;
  std::vector<int> plane_comp;
  std::vector<FluxParams> fluxes;
  std::vector<int> flux_comp;     // 4 field components per monitor
  std::vector<void*> owned;     // device allocations owned by the plan
  long long max_steps = 0;
  hipStream_t cap_stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  int graph_chunk = 16;         // even: a chunk returns the state to its starting set
  int timing = 0;
  double t_sum[3] = {0, 0, 0};
  long long t_cnt[3] = {0, 0, 0};
  std::vector<hipEvent_t> tev;
  float4* dcoef = nullptr;      // per material d(ca, cb, ce)/d eps_r (reverse mode)
  float2* dcoef_s = nullptr;    // per material d(ca, cb)/d sigma (reverse mode)
  float2* dkdis = nullptr;      // per material -gamma'/gamma of the ADE (reverse mode)
  float* cbv = nullptr;         // per-cell cb, ch of the reverse sweep (built on first use)
  float* chv = nullptr;
  int h_probe = 0;              // some probe samples an H component
  float* dmu = nullptr;         // per material d ch / d mu_r (reverse mode)
  int n_mat = 0;
};

// One reverse-mode launch (E-bar pass when e) over the first n_act
// instances of a table; gm as k_batch_adj's GM.
static void launch_batch_adj(hipStream_t s, bool e, int gm, unsigned nb, const BatchEntry* ents,
                             const BatchAdjEntry* adj, const int* start, int n_act, int gather,
                             long long r) {
  const dim3 g(nb), blk(kBX, kTY);
  if (e) {
    if (gm == 0) hipLaunchKernelGGL((k_batch_adj<true, 0>), g, blk, 0, s, ents, adj, start, n_act, gather, r);
    else if (gm == 1) hipLaunchKernelGGL((k_batch_adj<true, 1>), g, blk, 0, s, ents, adj, start, n_act, gather, r);
    else hipLaunchKernelGGL((k_batch_adj<true, 2>), g, blk, 0, s, ents, adj, start, n_act, gather, r);
  } else {
    if (gm == 0) hipLaunchKernelGGL((k_batch_adj<false, 0>), g, blk, 0, s, ents, adj, start, n_act, gather, r);
    else if (gm == 1) hipLaunchKernelGGL((k_batch_adj<false, 1>), g, blk, 0, s, ents, adj, start, n_act, gather, r);
    else hipLaunchKernelGGL((k_batch_adj<false, 2>), g, blk, 0, s, ents, adj, start, n_act, gather, r);
  }
}

// Logical workgroups of one instance in a batched launch (its plan's
// workgroup order, one cell per thread).
static long long batch_blocks(const rfxh_plan* pl) {
  const KParams& kq = pl->kp;
  const long long flat = kq.plane, nxb = kq.nx;
  return kq.order == 0 ? 8LL * nxb * (((flat + 7) / 8 + kWG - 1) / kWG)
                       : nxb * ((flat + kWG - 1) / kWG);
}


namespace {

template <typename T>
int dev_copy(rfxh_plan* pl, const T* host, size_t n, T** out) {
  if (n == 0) {
    *out = nullptr;
    return 0;
  }
  void* d = nullptr;
  HIPCHK(hipMalloc(&d, n * sizeof(T)));
  pl->owned.push_back(d);
  HIPCHK(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<T*>(d);
  return 0;
}

long long face_slab_floats(const rfxh_desc* d, int face) {
  const int n = d->cpml_n[face];
  const int ax = face >> 1;
  long long s = n;
  if (ax != 0) s *= d->nx;
  if (ax != 1) s *= d->ny;
  if (ax != 2) s *= d->nz;
  return s;
}

// Tiles and x segments of the marching kernel.  Automatic segments: enough
// workgroups for about eight per CU, at least four planes per segment.
long long march_blocks(const KParams& kp, int rows, int want_seg, MarchIO* io, int* nseg_out) {
  const int ntj = (kp.ny + rows - 1) / rows, ntk = (kp.nz + kBX - 1) / kBX;
  const long long tiles = (long long)ntj * ntk;
  int nseg = want_seg;
  if (nseg <= 0) {
    nseg = (int)std::max<long long>(1, (2048 + tiles - 1) / tiles);
    nseg = std::min(nseg, std::max(1, kp.nx / 4));
  }
  nseg = std::min(nseg, kp.nx);
  const int seg_len = (kp.nx + nseg - 1) / nseg;
  nseg = (kp.nx + seg_len - 1) / seg_len;
  io->ntj = ntj;
  io->ntk = ntk;
  io->seg_len = seg_len;
  *nseg_out = nseg;
  return tiles * nseg;
}

// DFT-plane and flux-monitor kernels on field set `set` (after the step).
void launch_accumulators(rfxh_plan* pl, hipStream_t s, int set) {
  for (size_t q = 0; q < pl->planes.size(); ++q) {
    PlaneParams pp = pl->planes[q];
    pp.fld = pl->set[set][pl->plane_comp[q]];
    const long long npts = (long long)pp.n1 * pp.n2;
    hipLaunchKernelGGL(k_dft_plane, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, pp);
  }
  for (size_t q = 0; q < pl->fluxes.size(); ++q) {
    FluxParams fp = pl->fluxes[q];
    for (int t = 0; t < 4; ++t) fp.f[t] = pl->set[set][pl->flux_comp[4 * q + t]];
    const long long npts = (long long)fp.n1 * fp.n2;
    hipLaunchKernelGGL(k_flux, dim3((unsigned)((npts + 255) / 256), 4), dim3(256), 0, s, fp);
  }
}

// Start of a run at step0: steps before it are already folded into the
// DFT-plane accumulators.
void dft_begin(rfxh_plan* pl, hipStream_t s, long long step0) {
  if (pl->acc_from) hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, pl->acc_from, step0);
}

// End of a run: fold the staged steps of the last incomplete group.
void dft_flush(rfxh_plan* pl, hipStream_t s) {
  if (pl->planes.empty() && pl->fluxes.empty()) return;
  for (size_t q = 0; q < pl->planes.size(); ++q) {
    const PlaneParams& pp = pl->planes[q];
    const long long npts = (long long)pp.n1 * pp.n2;
    hipLaunchKernelGGL(k_dft_flush, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, pp);
  }
  for (size_t q = 0; q < pl->fluxes.size(); ++q) {
    const FluxParams& fp = pl->fluxes[q];   // the flush reads only the ring
    const long long npts = (long long)fp.n1 * fp.n2;
    hipLaunchKernelGGL(k_flux_flush, dim3((unsigned)((npts + 255) / 256), 4), dim3(256), 0, s,
                       fp);
  }
  hipLaunchKernelGGL(k_set_acc_from, dim3(1), dim3(64), 0, s, pl->acc_from, pl->kp.step);
}

template <int PASS, bool GRADED, int NP>
void launch_pass_n(const rfxh_plan* pl, hipStream_t s, const KParams& kp, const PsiIO& ps,
                   int set) {
  const dim3 g = NP == 1 ? pl->grid : pl->grid_col, blk = pl->block;
  if (NP == 1) {
    if (PASS == 0) {
      if (kp.ch_all > 0.f) hipLaunchKernelGGL((k_update_h<GRADED, true>), g, blk, 0, s, kp, ps);
      else hipLaunchKernelGGL((k_update_h<GRADED, false>), g, blk, 0, s, kp, ps);
    }
    else if (pl->n_pole > 0 || pl->aniso)
      hipLaunchKernelGGL((k_update_e<GRADED, true>), g, blk, 0, s, kp, ps);
    else hipLaunchKernelGGL((k_update_e<GRADED, false>), g, blk, 0, s, kp, ps);
  } else {
    const PsiIO* psp = pl->d_psio + set;
    if (PASS == 0) {
      if (kp.ch_all > 0.f)
        hipLaunchKernelGGL((k_update_h_col<GRADED, NP, true>), g, blk, 0, s, kp, psp);
      else
        hipLaunchKernelGGL((k_update_h_col<GRADED, NP, false>), g, blk, 0, s, kp, psp);
    } else {
      if (pl->n_pole > 0 || pl->aniso)
        hipLaunchKernelGGL((k_update_e_col<GRADED, NP, true>), g, blk, 0, s, kp, psp);
      else
        hipLaunchKernelGGL((k_update_e_col<GRADED, NP, false>), g, blk, 0, s, kp, psp);
    }
  }
}

template <int PASS, bool GRADED>
void launch_pass_g(const rfxh_plan* pl, hipStream_t s, const KParams& kp, const PsiIO& ps,
                   int set) {
  switch (pl->ncol) {
    case 2: launch_pass_n<PASS, GRADED, 2>(pl, s, kp, ps, set); break;
    case 3: launch_pass_n<PASS, GRADED, 3>(pl, s, kp, ps, set); break;
    case 4: launch_pass_n<PASS, GRADED, 4>(pl, s, kp, ps, set); break;
    default: launch_pass_n<PASS, GRADED, 1>(pl, s, kp, ps, set); break;
  }
}

// One two-pass half step (pass 0: H, pass 1: E) in place on field set `set`
// (kp.f and ps point at it), with the plan's column width.
void launch_pass(const rfxh_plan* pl, hipStream_t s, const KParams& kp, const PsiIO& ps, int pass,
                 int set) {
  if (pass == 0) {
    if (kp.graded) launch_pass_g<0, true>(pl, s, kp, ps, set);
    else launch_pass_g<0, false>(pl, s, kp, ps, set);
  } else {
    if (kp.graded) launch_pass_g<1, true>(pl, s, kp, ps, set);
    else launch_pass_g<1, false>(pl, s, kp, ps, set);
  }
}

// One time step starting from field set `from` (fused: writes set 1-from).
int launch_step(rfxh_plan* pl, hipStream_t s, hipEvent_t* ev, int from) {
  // ev (optional): four events bracketing the main kernel(s) and the DFT
  // kernels of this step, recorded on the stream the kernels run on.
  const int to = pl->fused ? 1 - from : from;
  PsiIO ps;
  ps.pol = pl->pol;
  ps.pcoef = pl->pcoef;
  ps.n_pole = pl->n_pole;
  ps.n_debye = pl->n_debye;
  for (int a = 0; a < 3; ++a) {
    ps.pol_lo[a] = pl->pol_lo[a];
    ps.pol_n[a] = pl->pol_n[a];
  }
  for (int f = 0; f < 6; ++f)
    for (int t = 0; t < 2; ++t) {
      const long long ro = from * pl->psi_set, wo = to * pl->psi_set;
      float* h0 = pl->kp.psi_h[f][t];
      float* e0 = pl->kp.psi_e[f][t];
      ps.hr[f][t] = h0 ? h0 + ro : nullptr;
      ps.hw[f][t] = h0 ? h0 + wo : nullptr;
      ps.er[f][t] = e0 ? e0 + ro : nullptr;
      ps.ew[f][t] = e0 ? e0 + wo : nullptr;
    }
  if (ev) HIPCHK(hipEventRecord(ev[0], s));
  if (pl->wmarch) {
    WMarchIO io;
    for (int a = 0; a < 3; ++a) {
      io.eo[a] = pl->set[from][a];
      io.ho[a] = pl->set[from][3 + a];
      io.en[a] = pl->set[to][a];
      io.hn[a] = pl->set[to][3 + a];
    }
    const KParams& kp = pl->kp;
    io.n_mat = pl->n_mat;
    io.ntj = (kp.ny + 1) / 2;                                // one row pair per workgroup
    io.ntk = ((kp.nz + kWMK - 1) / kWMK + kWMW - 1) / kWMW;  // groups of kWMW z tiles
    const long long tiles = (long long)io.ntj * io.ntk;
    int nseg = pl->wmarch_seg;
    if (nseg <= 0) {
      // about 32 K wavefronts, at least four planes per segment
      nseg = (int)std::max<long long>(1, (8192 + tiles - 1) / tiles);
      nseg = std::min(nseg, std::max(1, kp.nx / 4));
    }
    nseg = std::min(nseg, kp.nx);
    io.seg_len = (kp.nx + nseg - 1) / nseg;
    nseg = (kp.nx + io.seg_len - 1) / io.seg_len;
    const dim3 g((unsigned)(tiles * nseg)), blk(kBX, kWMW);
    const PsiIO* psp = pl->d_psio + 2 + from;   // read set `from`, write set `to`
    if (kp.graded) hipLaunchKernelGGL(k_step_wmarch<true>, g, blk, 0, s, kp, io, psp);
    else hipLaunchKernelGGL(k_step_wmarch<false>, g, blk, 0, s, kp, io, psp);
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
  } else if (pl->march) {
    MarchIO io;
    for (int a = 0; a < 3; ++a) {
      io.eo[a] = pl->set[from][a];
      io.ho[a] = pl->set[from][3 + a];
      io.en[a] = pl->set[to][a];
      io.hn[a] = pl->set[to][3 + a];
    }
    int nseg = 1;
    const int rows = kMarchRows[pl->march_tile];
    const long long nblk = march_blocks(pl->kp, rows, pl->march_seg, &io, &nseg);
    const dim3 g((unsigned)nblk), blk(kBX, rows + 2);
    if (pl->march_tile == 0) {
      if (pl->kp.graded) hipLaunchKernelGGL((k_step_march<true, 6>), g, blk, 0, s, pl->kp, io, ps);
      else hipLaunchKernelGGL((k_step_march<false, 6>), g, blk, 0, s, pl->kp, io, ps);
    } else {
      if (pl->kp.graded) hipLaunchKernelGGL((k_step_march<true, 14>), g, blk, 0, s, pl->kp, io, ps);
      else hipLaunchKernelGGL((k_step_march<false, 14>), g, blk, 0, s, pl->kp, io, ps);
    }
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
  } else if (pl->fused) {
    FusedIO io;
    for (int a = 0; a < 3; ++a) {
      io.eo[a] = pl->set[from][a];
      io.ho[a] = pl->set[from][3 + a];
      io.en[a] = pl->set[to][a];
      io.hn[a] = pl->set[to][3 + a];
    }
    if (pl->kp.graded)
      hipLaunchKernelGGL(k_step_fused<true>, pl->grid, pl->block, 0, s, pl->kp, io, ps);
    else
      hipLaunchKernelGGL(k_step_fused<false>, pl->grid, pl->block, 0, s, pl->kp, io, ps);
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
  } else {
    KParams kp = pl->kp;
    for (int c = 0; c < 6; ++c) kp.f[c] = pl->set[from][c];
    launch_pass(pl, s, kp, ps, 0, from);
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
    launch_pass(pl, s, kp, ps, 1, from);
  }
  if (ev) HIPCHK(hipEventRecord(ev[2], s));
  launch_accumulators(pl, s, to);
  if (ev) HIPCHK(hipEventRecord(ev[3], s));
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, pl->kp.step);
  HIPCHK(hipGetLastError());
  return 0;
}

// Tiles and x segments of the temporal-blocked march.  Automatic segments:
// at least about four workgroups per CU (one fits a CU), at least eight
// planes per segment.
// The x-slab segments and the middle ones (want_seg of them, 0: about
// 6144 workgroups, at least 24 planes each; measured at 533^3, 16 middle
// segments (7290 workgroups) beat 4 by 22 % and 8 by 7 %: the workgroups
// over CPML slabs take longer, and many small workgroups balance them).
void tb_blocks(const KParams& kp, int want_seg, TBIO* io, int* nseg_out) {
  // z-edge tiles when both z faces have CPML slabs no deeper than an edge
  // tile's inner columns less the halo
  const int ei = kTBEdgeC - 2 * kTBHalo;
  const bool edges = kp.cn[4] > 0 && kp.cn[5] > 0 && std::max(kp.cn[4], kp.cn[5]) + kTBHalo <= ei &&
                     kp.nz > 2 * ei;
  io->kz0 = edges ? ei : 0;
  io->kz1 = edges ? kp.nz - ei : kp.nz;
  io->ntj = (kp.ny + kTBJ - 1) / kTBJ;
  io->ntk = (io->kz1 - io->kz0 + kTBK - 1) / kTBK;
  io->ntj_e = edges ? (kp.ny + kTBEdgeR - 2 * kTBHalo - 1) / (kTBEdgeR - 2 * kTBHalo) : 0;
  const long long tiles = (long long)io->ntj * io->ntk + 2LL * io->ntj_e;
  io->xa = kp.cn[0] > 0 ? std::min(kp.nx, kp.cn[0] + 2) : 0;
  io->xb = kp.cn[1] > 0 ? std::max(io->xa, kp.nx - kp.cn[1] - 1) : kp.nx;
  const int mid = io->xb - io->xa;
  int nmid = want_seg;
  if (nmid <= 0) {
    nmid = (int)std::max<long long>(1, (6144 + tiles - 1) / tiles);
    nmid = std::min(nmid, std::max(1, mid / 24));
  }
  nmid = std::max(1, std::min(nmid, std::max(1, mid)));
  io->seg_len = std::max(1, (mid + nmid - 1) / nmid);
  io->n_mid = mid > 0 ? (mid + io->seg_len - 1) / io->seg_len : 1;
  *nseg_out = io->n_mid + 2;
}

// Two time steps starting from field set `from` (writes set 1-from): the
// temporal-blocked sweep, the folds of the DFT groups it completes, and the
// step counter.  ev (optional): four events as in launch_step; the sweep
// kernel's interval covers both steps.
int launch_tb(rfxh_plan* pl, hipStream_t s, hipEvent_t* ev, int from) {
  const int to = 1 - from;
  const KParams& kp = pl->kp;
  TBIO io;
  std::memset(&io, 0, sizeof(io));
  for (int a = 0; a < 3; ++a) {
    io.eo[a] = pl->set[from][a];
    io.ho[a] = pl->set[from][3 + a];
    io.en[a] = pl->set[to][a];
    io.hn[a] = pl->set[to][3 + a];
  }
  io.n_mat = pl->n_mat;
  int nseg = 1;
  tb_blocks(kp, pl->tb_seg, &io, &nseg);
  io.n_plane = (int)pl->planes.size();
  for (int q = 0; q < io.n_plane; ++q) {
    const PlaneParams& pp = pl->planes[q];
    io.pl[q].snap = pp.snap;
    io.pl[q].axis = pp.axis;
    io.pl[q].index = pp.index;
    io.pl[q].comp = pl->plane_comp[q];
    io.pl[q].n2 = pp.n2;
    io.pl[q].npts = pp.n1 * pp.n2;
  }
  if (pl->psi_base) {
    io.cm.psr = pl->psi_base + from * pl->psi_set;   // read set `from`, write set `to`
    io.cm.psw = pl->psi_base + to * pl->psi_set;
  }
  io.cm.prof = pl->cprof;
  io.cm.mid = pl->psi_mid;
  io.n_prof = kp.cn[0] + kp.cn[1] + kp.cn[2] + kp.cn[3] + kp.cn[4] + kp.cn[5];
  for (int f = 0; f < 6; ++f) {
    io.cm.foff[f] = pl->psi_foff[f];
    io.cm.fs[f] = pl->psi_fs[f];
    io.cm.poff[f] = pl->cprof_off[f];
  }
  const dim3 g((unsigned)(((long long)io.ntj * io.ntk + 2LL * io.ntj_e) * nseg)), blk(kBX, kTBW);
  if (ev) HIPCHK(hipEventRecord(ev[0], s));
  const bool cpml = pl->psi_base && pl->cprof;   // the plan has CPML faces
  if (kp.graded) {
    if (cpml) hipLaunchKernelGGL((k_step_tb2<true, true>), g, blk, 0, s, kp, io);
    else hipLaunchKernelGGL((k_step_tb2<true, false>), g, blk, 0, s, kp, io);
  } else {
    if (cpml) hipLaunchKernelGGL((k_step_tb2<false, true>), g, blk, 0, s, kp, io);
    else hipLaunchKernelGGL((k_step_tb2<false, false>), g, blk, 0, s, kp, io);
  }
  if (ev) {
    HIPCHK(hipEventRecord(ev[1], s));
    HIPCHK(hipEventRecord(ev[2], s));
  }
  for (size_t q = 0; q < pl->planes.size(); ++q) {
    const PlaneParams& pp = pl->planes[q];
    const long long npts = (long long)pp.n1 * pp.n2;
    hipLaunchKernelGGL(k_dft_fold2, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, pp);
  }
  if (ev) HIPCHK(hipEventRecord(ev[3], s));
  hipLaunchKernelGGL(k_advance2, dim3(1), dim3(64), 0, s, kp.step);
  HIPCHK(hipGetLastError());
  return 0;
}

// Launch grid of the plan's workgroup order.
// pl->grid: one cell per thread (every kernel that maps with cell_of);
// pl->grid_col: the column kernels' grid (ncol cells per thread along x).
int plan_grid(rfxh_plan* pl) {
  const KParams& kp = pl->kp;
  const long long flat = kp.plane;
  auto blocks = [&](long long nxb) {
    return kp.order == 0 ? 8LL * nxb * (((flat + 7) / 8 + kWG - 1) / kWG)
                         : nxb * ((flat + kWG - 1) / kWG);
  };
  const long long n1 = blocks(kp.nx), nc = blocks((kp.nx + pl->ncol - 1) / pl->ncol);
  if (n1 > 0x7fffffffLL) return fail(RFXH_EINVAL, "grid too large");
  pl->grid = dim3((unsigned)n1);
  pl->grid_col = dim3((unsigned)nc);
  pl->block = dim3(kBX, kTY);
  return 0;
}

// Sum the sampled kernel durations once the run has been queued.
int collect_timing(rfxh_plan* pl, int n_samples) {
  for (int q = 0; q < n_samples; ++q) {
    hipEvent_t* ev = pl->tev.data() + 4 * q;
    HIPCHK(hipEventSynchronize(ev[3]));
    float ms;
    for (int w = 0; w < 3; ++w) {
      HIPCHK(hipEventElapsedTime(&ms, ev[w], ev[w + 1]));
      pl->t_sum[w] += ms;
      pl->t_cnt[w] += 1;
    }
  }
  return 0;
}

// Snapshot ring of a flux monitor / NTFF face and the plan's fold marker.
int flux_ring(rfxh_plan* pl, FluxParams& fp) {
  int rc;
  if (!pl->acc_from) {
    const long long zero = 0;
    if ((rc = dev_copy(pl, &zero, 1, &pl->acc_from))) return rc;
  }
  fp.acc_from = pl->acc_from;
  const std::vector<float> ring((size_t)kDftK * 4 * fp.n1 * fp.n2, 0.f);
  return dev_copy(pl, ring.data(), ring.size(), &fp.snap);
}

}  // namespace

extern "C" {

int64_t rfxh_psi_floats(const rfxh_desc* d) {
  if (!d) return -1;
  long long tot = 0;
  for (int f = 0; f < 6; ++f) tot += 4 * face_slab_floats(d, f);
  return tot;
}

static PsiIO psi_io(const rfxh_plan* pl, int set);

int rfxh_plan_create(const rfxh_desc* d, rfxh_plan** out) {
  if (!d || !out) return fail(RFXH_EINVAL, "null descriptor or output");
  *out = nullptr;
  if (d->nx < 2 || d->ny < 2 || d->nz < 2) return fail(RFXH_EINVAL, "grid must be at least 2 cells per axis");
  if (d->n_mat < 1 || d->n_mat > 65536 || !d->mat_coef || !d->mat_flags || !d->d_mat)
    return fail(RFXH_EINVAL, "material table missing or larger than 65536 entries");
  for (int c = 0; c < 6; ++c)
    if (!d->d_f[c]) return fail(RFXH_EINVAL, "field pointer missing");
  bool has2 = true;
  for (int c = 0; c < 6; ++c) has2 = has2 && d->d_f2[c];
  const int N[3] = {d->nx, d->ny, d->nz};
  bool any_cpml = false;
  for (int f = 0; f < 6; ++f) {
    if (d->cpml_n[f] < 0 || d->cpml_n[f] > N[f >> 1]) return fail(RFXH_EINVAL, "CPML depth outside the axis");
    if (d->cpml_n[f] > 0 && (!d->cpml_b[f] || !d->cpml_c[f] || !d->cpml_km1[f] || !d->d_psi))
      return fail(RFXH_EINVAL, "CPML profile or psi buffer missing");
    any_cpml = any_cpml || d->cpml_n[f] > 0;
  }
  if (has2 && any_cpml && !d->d_psi2) return fail(RFXH_EINVAL, "second CPML set missing");
  if (d->max_steps < 0) return fail(RFXH_EINVAL, "max_steps < 0");
  if ((d->n_src > 0 && !d->d_src_wave) || (d->n_probe > 0 && !d->d_probe_out) ||
      (d->n_port > 0 && !d->d_port_out))
    return fail(RFXH_EINVAL, "point-operation table missing");
  const long long ncell = (long long)d->nx * d->ny * d->nz;
  if (ncell >= (1LL << 30))
    return fail(RFXH_EINVAL, "grid larger than 2^30 cells (field arrays must stay below 4 GiB)");
  if ((long long)d->ny * d->nz >= (1LL << 24))
    return fail(RFXH_EINVAL, "y-z plane of 2^24 cells or more (the in-plane index must be exact in float32)");
  std::vector<rfxh_point_op> ops(d->ops, d->ops + d->n_ops);
  for (const rfxh_point_op& op : ops) {
    if (op.cell < 0 || op.cell >= ncell) return fail(RFXH_EINVAL, "point operation outside the grid");
    if (op.comp < 0 || op.comp > 5) return fail(RFXH_EINVAL, "bad point-operation component");
    const int lim = op.kind == RFXH_OP_PORT ? d->n_port : op.kind == RFXH_OP_INJECT ? d->n_src : d->n_probe;
    if (op.kind < 0 || op.kind > 2 || op.slot < 0 || op.slot >= lim)
      return fail(RFXH_EINVAL, "bad point-operation slot/kind");
    if (op.kind != RFXH_OP_PROBE && op.comp > 2) return fail(RFXH_EINVAL, "ports/sources drive E only");
  }
  std::stable_sort(ops.begin(), ops.end(), [](const rfxh_point_op& a, const rfxh_point_op& b) {
    return a.cell != b.cell ? a.cell < b.cell : a.order < b.order;
  });

  rfxh_plan* pl = new rfxh_plan();
  for (const rfxh_point_op& op : ops)
    if (op.kind == RFXH_OP_PROBE && op.comp >= 3) pl->h_probe = 1;
  {
    // point operations per tile column of the temporal-blocked march, over
    // all x: 16 x 64 tiles (rows [j0 - halo, + kTBW), columns [k0 - halo,
    // + kBX)) and, with z-edge tiles, 64 x 16 ones
    KParams tk_p;
    std::memset(&tk_p, 0, sizeof(tk_p));
    tk_p.nx = d->nx;
    tk_p.ny = d->ny;
    tk_p.nz = d->nz;
    for (int f = 0; f < 6; ++f) tk_p.cn[f] = d->cpml_n[f];
    TBIO tio;
    std::memset(&tio, 0, sizeof(tio));
    int nseg = 1;
    tb_blocks(tk_p, 0, &tio, &nseg);
    std::map<long long, int> cnt;
    const long long P = (long long)d->ny * d->nz;
    auto count = [&](long long key) { pl->tb_ops_max = std::max(pl->tb_ops_max, ++cnt[key]); };
    for (const rfxh_point_op& op : ops) {
      const int r = (int)(op.cell % P), j = r / d->nz, k = r % d->nz;
      for (int tj = std::max(0, (j - kTBHalo) / kTBJ - 1); tj < tio.ntj; ++tj) {
        if (tj * kTBJ - kTBHalo > j) break;
        if (j >= tj * kTBJ - kTBHalo + kTBW) continue;
        for (int tk = 0; tk < tio.ntk; ++tk) {
          const int k0 = tio.kz0 + tk * kTBK;
          if (k >= k0 - kTBHalo && k < k0 - kTBHalo + kBX) count((long long)tj * tio.ntk + tk);
        }
      }
      for (int side = 0; side < 2 && tio.ntj_e > 0; ++side) {
        const int k0 = side == 0 ? 0 : tio.kz1;
        if (k < k0 - kTBHalo || k >= k0 - kTBHalo + kTBEdgeC) continue;
        for (int te = 0; te < tio.ntj_e; ++te) {
          const int j0 = te * (kTBEdgeR - 2 * kTBHalo);
          if (j >= j0 - kTBHalo && j < j0 - kTBHalo + kTBEdgeR)
            count(-1 - (long long)(side * tio.ntj_e + te));
        }
      }
    }
  }
  KParams& kp = pl->kp;
  std::memset(&kp, 0, sizeof(kp));
  kp.nx = d->nx;
  kp.ny = d->ny;
  kp.nz = d->nz;
  kp.graded = d->graded ? 1 : 0;
  kp.plane = (long long)d->ny * d->nz;
  kp.inv_nz = 1.0f / (float)d->nz;
  for (int a = 0; a < 3; ++a) {
    kp.slo[a] = d->cpml_n[2 * a];
    kp.shi[a] = N[a] - d->cpml_n[2 * a + 1];
  }
  int rc = 0;
  auto bail = [&](int r) {
    rfxh_plan_destroy(pl);
    return r;
  };
  for (int a = 0; a < 3; ++a)
    if (!d->inv_e[a] || !d->inv_h[a]) return bail(fail(RFXH_EINVAL, "inverse spacing arrays missing"));
  kp.inv_e0 = d->inv_e[0][0];
  kp.inv_h0 = d->inv_h[0][0];
  for (int a = 0; a < 3; ++a) {
    float* pe;
    float* ph;
    if ((rc = dev_copy(pl, d->inv_e[a], (size_t)N[a], &pe))) return bail(rc);
    if ((rc = dev_copy(pl, d->inv_h[a], (size_t)N[a], &ph))) return bail(rc);
    kp.inv_e[a] = pe;
    kp.inv_h[a] = ph;
  }
  kp.mat = d->d_mat;
  {
    float4* co;
    std::vector<float4> tmp(d->n_mat);
    for (int m = 0; m < d->n_mat; ++m)
      tmp[m] = make_float4(d->mat_coef[4 * m], d->mat_coef[4 * m + 1], d->mat_coef[4 * m + 2],
                           d->mat_coef[4 * m + 3]);
    if ((rc = dev_copy(pl, tmp.data(), tmp.size(), &co))) return bail(rc);
    kp.coef = co;
    // one mu_r everywhere (the usual case): ch is a constant and the H
    // update need not read the material id (2 B per cell of its traffic)
    kp.ch_all = tmp[0].w > 0.f ? tmp[0].w : 0.f;
    for (int m = 1; m < d->n_mat; ++m)
      if (tmp[m].w != tmp[0].w) kp.ch_all = 0.f;
    const char* chm = std::getenv("RFXH_CH_PER_CELL");
    if (chm && std::string(chm) == "1") kp.ch_all = 0.f;
    uint32_t* fl;
    if ((rc = dev_copy(pl, d->mat_flags, (size_t)d->n_mat, &fl))) return bail(rc);
    kp.flags = fl;
    for (int m = 0; m < d->n_mat; ++m) {
      float bits;
      std::memcpy(&bits, &d->mat_flags[m], sizeof(bits));
      tmp[m].w = bits;
    }
    float4* ec;
    if ((rc = dev_copy(pl, tmp.data(), tmp.size(), &ec))) return bail(rc);
    kp.ecoef = ec;
  }
  // default: the two-pass step with 2-cell columns along x, the fastest
  // variant measured on MI355X at 533^3 and on the graded workloads;
  // RFXH_KERNEL = onecell | fused | march selects another (fused and march
  // need the second field set); rfxh_set_variant changes it later
  const char* kern = std::getenv("RFXH_KERNEL");
  const std::string kname = kern ? kern : "";
  pl->fused = has2 && d->n_pole == 0 && !d->aniso_coef && (kname == "fused" || kname == "march");
  pl->march = has2 && kname == "march";
  pl->ncol = (pl->fused || kname == "onecell") ? 1 : 2;
  for (int c = 0; c < 6; ++c) {
    pl->set[0][c] = d->d_f[c];
    pl->set[1][c] = has2 ? d->d_f2[c] : nullptr;
    kp.f[c] = d->d_f[c];
  }
  long long off = 0;
  pl->psi_base = d->d_psi;
  std::vector<float4> prof_rows;
  for (int f = 0; f < 6; ++f) {
    kp.cn[f] = d->cpml_n[f];
    kp.cinv[f] = d->cpml_inv_d[f];
    const long long s = face_slab_floats(d, f);
    pl->psi_foff[f] = (unsigned)off;
    pl->psi_fs[f] = (unsigned)s;
    pl->cprof_off[f] = (int)prof_rows.size();
    for (int l = 0; l < d->cpml_n[f]; ++l)
      prof_rows.push_back(make_float4(d->cpml_b[f][l], d->cpml_c[f][l], d->cpml_km1[f][l], 0.f));
    if (d->cpml_n[f] > 0) {
      kp.psi_h[f][0] = d->d_psi + off;
      kp.psi_h[f][1] = d->d_psi + off + s;
      kp.psi_e[f][0] = d->d_psi + off + 2 * s;
      kp.psi_e[f][1] = d->d_psi + off + 3 * s;
      float *b, *c, *k;
      if ((rc = dev_copy(pl, d->cpml_b[f], (size_t)d->cpml_n[f], &b))) return bail(rc);
      if ((rc = dev_copy(pl, d->cpml_c[f], (size_t)d->cpml_n[f], &c))) return bail(rc);
      if ((rc = dev_copy(pl, d->cpml_km1[f], (size_t)d->cpml_n[f], &k))) return bail(rc);
      kp.pb[f] = b;
      kp.pc[f] = c;
      kp.pk[f] = k;
    }
    off += 4 * s;
  }
  if (!prof_rows.empty() && (rc = dev_copy(pl, prof_rows.data(), prof_rows.size(), &pl->cprof)))
    return bail(rc);
  pl->psi_floats = off;
  // Dispersive poles: cells whose material flags carry pole bits update their
  // polarisation in the box pol_lo + [0, pol_n)
  if (d->n_pole < 0 || d->n_pole > 16) return bail(fail(RFXH_EINVAL, "n_pole must be 0..16"));
  if (d->n_debye < 0 || d->n_debye > d->n_pole)
    return bail(fail(RFXH_EINVAL, "n_debye must be 0..n_pole"));
  if (d->n_pole > 0) {
    if (!d->pole_coef || !d->d_pol) return bail(fail(RFXH_EINVAL, "dispersion tables missing"));
    for (int a = 0; a < 3; ++a)
      if (d->pol_lo[a] < 0 || d->pol_n[a] < 1 || d->pol_lo[a] + d->pol_n[a] > N[a])
        return bail(fail(RFXH_EINVAL, "polarisation state box outside the grid"));
    for (int m = 0; m < d->n_mat; ++m)
      if ((d->mat_flags[m] & RFXH_POLES) >> 16 >= (1u << d->n_pole))
        return bail(fail(RFXH_EINVAL, "material flags name a pole past n_pole"));
    if ((long long)d->pol_n[0] * d->pol_n[1] * d->pol_n[2] * (6LL * d->n_pole - 3LL * d->n_debye) >=
        (1LL << 32))
      return bail(fail(RFXH_EINVAL, "dispersion state larger than 2^32 floats"));
    float* pc;
    if ((rc = dev_copy(pl, d->pole_coef, (size_t)d->n_mat * d->n_pole * 4, &pc))) return bail(rc);
    pl->pcoef = pc;
    pl->pol = d->d_pol;
    pl->n_pole = d->n_pole;
    pl->n_debye = d->n_debye;
    for (int a = 0; a < 3; ++a) {
      pl->pol_lo[a] = d->pol_lo[a];
      pl->pol_n[a] = d->pol_n[a];
    }
  } else {
    for (int m = 0; m < d->n_mat; ++m)
      if (d->mat_flags[m] & RFXH_POLES)
        return bail(fail(RFXH_EINVAL, "material flags name poles but n_pole is 0"));
  }
  // subpixel smoothing: rows (ca, cb of Ey, of Ez) for the flagged ids
  {
    bool flagged = false;
    for (int m = 0; m < d->n_mat; ++m) flagged = flagged || (d->mat_flags[m] & RFXH_ANISO);
    if (flagged && !d->aniso_coef)
      return bail(fail(RFXH_EINVAL, "material flags name per-component eps but aniso_coef is NULL"));
    if (d->aniso_coef) {
      float4* ac;
      if ((rc = dev_copy(pl, reinterpret_cast<const float4*>(d->aniso_coef), (size_t)d->n_mat,
                         &ac)))
        return bail(rc);
      kp.acoef = ac;
      pl->aniso = true;
    }
  }
  // CPML set 1 is addressed as set-0 pointer + psi_set floats
  pl->psi_set = (has2 && any_cpml) ? (long long)(d->d_psi2 - d->d_psi) : 0;
  {
    // [0], [1]: in place on set 0 / 1; [2], [3]: read set 0 / 1, write the
    // other (the ping-pong steps)
    const PsiIO s0 = psi_io(pl, 0), s1 = psi_io(pl, has2 ? 1 : 0);
    PsiIO tab[4] = {s0, s1, s0, s1};
    for (int f = 0; f < 6; ++f)
      for (int t = 0; t < 2; ++t) {
        tab[2].hw[f][t] = s1.hw[f][t];
        tab[2].ew[f][t] = s1.ew[f][t];
        tab[3].hw[f][t] = s0.hw[f][t];
        tab[3].ew[f][t] = s0.ew[f][t];
      }
    if ((rc = dev_copy(pl, tab, 4, &pl->d_psio))) return bail(rc);
  }
  kp.pec_axes = d->pec_axes;
  {
    rfxh_point_op* dops;
    if ((rc = dev_copy(pl, ops.data(), ops.size(), &dops))) return bail(rc);
    kp.ops = dops;
    kp.n_ops = (int)ops.size();
  }
  kp.src_wave = d->d_src_wave;
  kp.n_src = d->n_src;
  kp.probe_out = d->d_probe_out;
  kp.n_probe = d->n_probe;
  kp.port_out = d->d_port_out;
  kp.n_port = d->n_port;
  pl->max_steps = d->max_steps;
  pl->n_mat = d->n_mat;
  {
    const long long init = -1;
    long long* st;
    if ((rc = dev_copy(pl, &init, 1, &st))) return bail(rc);
    kp.step = st;
  }
  for (int q = 0; q < d->n_plane; ++q) {
    const rfxh_plane& pd = d->planes[q];
    if (pd.axis < 0 || pd.axis > 2 || pd.index < 0 || pd.index >= N[pd.axis] || pd.comp < 0 ||
        pd.comp > 5 || pd.n_freq < 1 || !pd.d_acc || !pd.freqs)
      return bail(fail(RFXH_EINVAL, "bad DFT plane"));
    PlaneParams pp;
    pp.fld = d->d_f[pd.comp];
    pp.acc = reinterpret_cast<double2*>(pd.d_acc);
    pp.step = kp.step;
    if (!pl->acc_from) {
      const long long zero = 0;
      if ((rc = dev_copy(pl, &zero, 1, &pl->acc_from))) return bail(rc);
    }
    pp.acc_from = pl->acc_from;
    pp.max_steps = d->max_steps;
    pp.nf = pd.n_freq;
    pp.axis = pd.axis;
    pp.index = pd.index;
    pp.n1 = N[pd.axis == 0 ? 1 : 0];
    pp.n2 = N[pd.axis == 2 ? 1 : 2];
    pp.ny = d->ny;
    pp.nz = d->nz;
    {
      const std::vector<float> ring((size_t)kDftRing * pp.n1 * pp.n2, 0.f);
      if ((rc = dev_copy(pl, ring.data(), ring.size(), &pp.snap))) return bail(rc);
    }
    std::vector<double2> tw((size_t)d->max_steps * pd.n_freq);
    for (long long n = 0; n < d->max_steps; ++n) {
      const double t = (double)(n + d->plane_t_offset) * d->dt;
      for (int f = 0; f < pd.n_freq; ++f) {
        const double w = -2.0 * M_PI * pd.freqs[f] * t;
        tw[n * pd.n_freq + f] = make_double2(std::cos(w) * d->dt, std::sin(w) * d->dt);
      }
    }
    double2* dtw;
    if ((rc = dev_copy(pl, tw.data(), tw.size(), &dtw))) return bail(rc);
    pp.tw = dtw;
    pl->planes.push_back(pp);
    pl->plane_comp.push_back(pd.comp);
  }
  static const int kFluxComp[3][4] = {{1, 2, 4, 5}, {2, 0, 5, 3}, {0, 1, 3, 4}};
  for (int q = 0; q < d->n_flux; ++q) {
    const rfxh_flux& fd = d->flux[q];
    if (fd.axis < 0 || fd.axis > 2 || fd.index < 0 || fd.index >= N[fd.axis] || fd.n_freq < 1 ||
        !fd.d_acc || !fd.freqs || fd.n1 < 1 || fd.n2 < 1 || fd.lo1 < 0 || fd.lo2 < 0)
      return bail(fail(RFXH_EINVAL, "bad flux monitor"));
    const int t1 = fd.axis == 0 ? 1 : 0, t2 = fd.axis == 2 ? 1 : 2;
    if (fd.lo1 + fd.n1 > N[t1] || fd.lo2 + fd.n2 > N[t2])
      return bail(fail(RFXH_EINVAL, "flux region outside the plane"));
    FluxParams fp;
    std::memset(&fp, 0, sizeof(fp));
    fp.acc = reinterpret_cast<double2*>(fd.d_acc);
    fp.step = kp.step;
    fp.max_steps = d->max_steps;
    fp.nf = fd.n_freq;
    fp.axis = fd.axis;
    fp.index = fd.index;
    fp.index_m1 = fd.index > 0 ? fd.index - 1 : 0;
    fp.lo1 = fd.lo1;
    fp.n1 = fd.n1;
    fp.lo2 = fd.lo2;
    fp.n2 = fd.n2;
    fp.ny = d->ny;
    fp.nz = d->nz;
    std::vector<double2> te((size_t)d->max_steps * fd.n_freq), th(te.size());
    for (long long n = 0; n < d->max_steps; ++n) {
      const long long stp = n + d->plane_t_offset;
      const double w = fd.weight ? (double)fd.weight[stp] : 1.0;
      const double t = (double)stp * d->dt;
      for (int f = 0; f < fd.n_freq; ++f) {
        const double ae = -2.0 * M_PI * fd.freqs[f] * t;
        const double ah = -2.0 * M_PI * fd.freqs[f] * (t - d->dt * 0.5);
        te[n * fd.n_freq + f] = make_double2(std::cos(ae) * d->dt * w, std::sin(ae) * d->dt * w);
        th[n * fd.n_freq + f] = make_double2(std::cos(ah) * d->dt * w, std::sin(ah) * d->dt * w);
      }
    }
    double2 *dte, *dth;
    if ((rc = dev_copy(pl, te.data(), te.size(), &dte))) return bail(rc);
    if ((rc = dev_copy(pl, th.data(), th.size(), &dth))) return bail(rc);
    fp.twe = dte;
    fp.twh = dth;
    if ((rc = flux_ring(pl, fp))) return bail(rc);
    pl->fluxes.push_back(fp);
    for (int t = 0; t < 4; ++t) pl->flux_comp.push_back(kFluxComp[fd.axis][t]);
  }
  // NTFF boxes: six faces each, run by the flux kernel with the H planes
  // not averaged (index_m1 = index: (h + h) * 0.5f is h exactly) and the
  // box's own time base, t = n dt for E and (n + 1/2) dt for H on both
  // lanes (rfx/farfield.py:234-245, called with the scan index).
  static const int kNtffComp[3][4] = {{1, 2, 4, 5}, {0, 2, 3, 5}, {0, 1, 3, 4}};
  for (int q = 0; q < d->n_ntff; ++q) {
    const rfxh_ntff& nb = d->ntff[q];
    if (nb.n_freq < 1 || !nb.freqs)
      return bail(fail(RFXH_EINVAL, "bad NTFF box: no frequencies"));
    for (int a = 0; a < 3; ++a)
      if (nb.lo[a] < 0 || nb.hi[a] <= nb.lo[a] || nb.hi[a] >= N[a])
        return bail(fail(RFXH_EINVAL, "bad NTFF box: need 0 <= lo < hi < n on every axis"));
    std::vector<double2> te((size_t)d->max_steps * nb.n_freq), th(te.size());
    for (long long n = 0; n < d->max_steps; ++n) {
      const double t = (double)n * d->dt;
      for (int f = 0; f < nb.n_freq; ++f) {
        const double ae = -2.0 * M_PI * nb.freqs[f] * t;
        const double ah = -2.0 * M_PI * nb.freqs[f] * (t + d->dt * 0.5);
        te[n * nb.n_freq + f] = make_double2(std::cos(ae) * d->dt, std::sin(ae) * d->dt);
        th[n * nb.n_freq + f] = make_double2(std::cos(ah) * d->dt, std::sin(ah) * d->dt);
      }
    }
    double2 *dte, *dth;
    if ((rc = dev_copy(pl, te.data(), te.size(), &dte))) return bail(rc);
    if ((rc = dev_copy(pl, th.data(), th.size(), &dth))) return bail(rc);
    for (int face = 0; face < 6; ++face) {
      const int ax = face / 2;
      if (!nb.d_acc[face]) return bail(fail(RFXH_EINVAL, "NTFF face accumulator is NULL"));
      const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
      FluxParams fp;
      std::memset(&fp, 0, sizeof(fp));
      fp.acc = reinterpret_cast<double2*>(nb.d_acc[face]);
      fp.step = kp.step;
      fp.max_steps = d->max_steps;
      fp.nf = nb.n_freq;
      fp.axis = ax;
      fp.index = fp.index_m1 = (face & 1) ? nb.hi[ax] : nb.lo[ax];
      fp.lo1 = nb.lo[t1];
      fp.n1 = nb.hi[t1] - nb.lo[t1];
      fp.lo2 = nb.lo[t2];
      fp.n2 = nb.hi[t2] - nb.lo[t2];
      fp.ny = d->ny;
      fp.nz = d->nz;
      fp.twe = dte;
      fp.twh = dth;
      if ((rc = flux_ring(pl, fp))) return bail(rc);
      pl->fluxes.push_back(fp);
      for (int t = 0; t < 4; ++t) pl->flux_comp.push_back(kNtffComp[ax][t]);
    }
  }
  // workgroup order (RFXH_ORDER = flat (default) | bands, for comparisons;
  // rfxh_set_variant changes it later)
  const char* ord = std::getenv("RFXH_ORDER");
  kp.order = (ord && std::string(ord) == "bands") ? 0 : 1;
  if ((rc = plan_grid(pl))) return bail(rc);
  pl->tev.resize(4 * kTimingSamples);
  for (auto& e : pl->tev) {
    if (hipEventCreate(&e) != hipSuccess) return bail(fail(RFXH_EHIP, "hipEventCreate"));
  }
  *out = pl;
  return RFXH_OK;
}

int rfxh_plan_destroy(rfxh_plan* pl) {
  if (!pl) return RFXH_OK;
  for (int g = 0; g < 2; ++g)
    if (pl->graph_exec[g]) (void)hipGraphExecDestroy(pl->graph_exec[g]);
  if (pl->cap_stream) (void)hipStreamDestroy(pl->cap_stream);
  if (pl->ev_in) (void)hipEventDestroy(pl->ev_in);
  if (pl->ev_out) (void)hipEventDestroy(pl->ev_out);
  for (auto e : pl->tev)
    if (e) (void)hipEventDestroy(e);
  for (void* p : pl->owned) (void)hipFree(p);
  delete pl;
  return RFXH_OK;
}

int rfxh_state_set(rfxh_plan* pl) {
  if (!pl) return fail(RFXH_EINVAL, "null plan");
  return pl->cur;
}

int rfxh_set_state_set(rfxh_plan* pl, int which) {
  if (!pl || which < 0 || which > 1 || (which == 1 && !pl->set[1][0]))
    return fail(RFXH_EINVAL, "bad state set");
  pl->cur = which;
  return RFXH_OK;
}

int rfxh_set_timing(rfxh_plan* pl, int enable) {
  if (!pl) return fail(RFXH_EINVAL, "null plan");
  pl->timing = enable;
  for (int w = 0; w < 3; ++w) {
    pl->t_sum[w] = 0;
    pl->t_cnt[w] = 0;
  }
  return RFXH_OK;
}

int rfxh_kernel_time_ms(rfxh_plan* pl, int which, double* avg_ms, int64_t* count) {
  if (!pl || which < 0 || which > 2 || !avg_ms) return fail(RFXH_EINVAL, "bad arguments");
  *avg_ms = pl->t_cnt[which] ? pl->t_sum[which] / pl->t_cnt[which] : 0.0;
  if (count) *count = pl->t_cnt[which];
  return RFXH_OK;
}

// rfxh_run with the temporal-blocked march: pairs of steps per sweep (each
// flips the field set), an odd last step on the two-pass kernels in place.
static int run_tb(rfxh_plan* pl, hipStream_t s, int64_t n_steps, int use_graph) {
  const int64_t pairs = n_steps / 2;
  if (!use_graph || pl->timing || n_steps < pl->graph_chunk) {
    const int64_t every = std::max<int64_t>(1, pairs / kTimingSamples);
    int n_samples = 0;
    for (int64_t q = 0; q < pairs; ++q) {
      hipEvent_t* ev = nullptr;
      if (pl->timing && q % every == every - 1 && n_samples < kTimingSamples)
        ev = pl->tev.data() + 4 * (n_samples++);
      const int rc = launch_tb(pl, s, ev, pl->cur);
      if (rc) return rc;
      pl->cur ^= 1;
    }
    if (n_steps & 1) {
      const int rc = launch_step(pl, s, nullptr, pl->cur);
      if (rc) return rc;
    }
    dft_flush(pl, s);
    HIPCHK(hipGetLastError());
    if (n_samples) return collect_timing(pl, n_samples);
    return RFXH_OK;
  }
  if (!pl->cap_stream) {
    HIPCHK(hipStreamCreateWithFlags(&pl->cap_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&pl->ev_in, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&pl->ev_out, hipEventDisableTiming));
  }
  const int g = pl->cur;
  const int sweeps = pl->graph_chunk / 2;   // even: a chunk returns the state to set g
  if (!pl->graph_exec[g]) {
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(pl->cap_stream, hipStreamCaptureModeThreadLocal));
    int from = g;
    for (int q = 0; q < sweeps; ++q) {
      const int rc = launch_tb(pl, pl->cap_stream, nullptr, from);
      if (rc) {
        (void)hipStreamEndCapture(pl->cap_stream, &graph);
        return rc;
      }
      from ^= 1;
    }
    HIPCHK(hipStreamEndCapture(pl->cap_stream, &graph));
    HIPCHK(hipGraphInstantiate(&pl->graph_exec[g], graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(graph));
  }
  HIPCHK(hipEventRecord(pl->ev_in, s));
  HIPCHK(hipStreamWaitEvent(pl->cap_stream, pl->ev_in, 0));
  int64_t done = 0;
  while (n_steps - done >= 2 * sweeps) {
    HIPCHK(hipGraphLaunch(pl->graph_exec[g], pl->cap_stream));
    done += 2 * sweeps;
  }
  for (; n_steps - done >= 2; done += 2) {
    const int rc = launch_tb(pl, pl->cap_stream, nullptr, pl->cur);
    if (rc) return rc;
    pl->cur ^= 1;
  }
  if (done < n_steps) {
    const int rc = launch_step(pl, pl->cap_stream, nullptr, pl->cur);
    if (rc) return rc;
  }
  dft_flush(pl, pl->cap_stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(pl->ev_out, pl->cap_stream));
  HIPCHK(hipStreamWaitEvent(s, pl->ev_out, 0));
  return RFXH_OK;
}

int rfxh_run(rfxh_plan* pl, int64_t step0, int64_t n_steps, void* stream, int use_graph) {
  if (!pl) return fail(RFXH_EINVAL, "null plan");
  if (step0 < 0 || n_steps < 0 || step0 + n_steps > pl->max_steps)
    return fail(RFXH_EINVAL, "steps outside the source/probe tables (max_steps)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, pl->kp.step, (long long)step0);
  dft_begin(pl, s, step0);
  HIPCHK(hipGetLastError());
  if (n_steps == 0) return RFXH_OK;
  if (pl->tb) return run_tb(pl, s, n_steps, use_graph);
  const int flip = pl->fused ? 1 : 0;
  if (!use_graph || pl->timing || n_steps < pl->graph_chunk) {
    // timing: sample up to kTimingSamples evenly spaced steps; the host does
    // not wait inside the loop, so the queue stays full and each event pair
    // brackets only its kernels
    const int64_t every = std::max<int64_t>(1, n_steps / kTimingSamples);
    int n_samples = 0;
    for (int64_t n = 0; n < n_steps; ++n) {
      hipEvent_t* ev = nullptr;
      if (pl->timing && n % every == every - 1 && n_samples < kTimingSamples)
        ev = pl->tev.data() + 4 * (n_samples++);
      const int rc = launch_step(pl, s, ev, pl->cur);
      if (rc) return rc;
      pl->cur ^= flip;
    }
    dft_flush(pl, s);
    HIPCHK(hipGetLastError());
    if (n_samples) return collect_timing(pl, n_samples);
    return RFXH_OK;
  }
  if (!pl->cap_stream) {
    HIPCHK(hipStreamCreateWithFlags(&pl->cap_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&pl->ev_in, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&pl->ev_out, hipEventDisableTiming));
  }
  const int g = pl->cur;
  if (!pl->graph_exec[g]) {
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(pl->cap_stream, hipStreamCaptureModeThreadLocal));
    int from = g;
    for (int n = 0; n < pl->graph_chunk; ++n) {
      const int rc = launch_step(pl, pl->cap_stream, nullptr, from);
      if (rc) {
        (void)hipStreamEndCapture(pl->cap_stream, &graph);
        return rc;
      }
      from ^= flip;
    }
    HIPCHK(hipStreamEndCapture(pl->cap_stream, &graph));
    HIPCHK(hipGraphInstantiate(&pl->graph_exec[g], graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(graph));
  }
  HIPCHK(hipEventRecord(pl->ev_in, s));
  HIPCHK(hipStreamWaitEvent(pl->cap_stream, pl->ev_in, 0));
  int64_t done = 0;
  while (n_steps - done >= pl->graph_chunk) {
    HIPCHK(hipGraphLaunch(pl->graph_exec[g], pl->cap_stream));
    done += pl->graph_chunk;  // even chunk: the state is back in set g
  }
  for (; done < n_steps; ++done) {
    const int rc = launch_step(pl, pl->cap_stream, nullptr, pl->cur);
    if (rc) return rc;
    pl->cur ^= flip;
  }
  dft_flush(pl, pl->cap_stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(pl->ev_out, pl->cap_stream));
  HIPCHK(hipStreamWaitEvent(s, pl->ev_out, 0));
  return RFXH_OK;
}

static int check_region(const rfxh_plan* pl, const rfxh_region* r) {
  if (!r || r->ni < 1 || r->nj < 1 || r->nk < 1 || r->i0 < 0 || r->j0 < 0 || r->k0 < 0 ||
      r->i0 + r->ni > pl->kp.nx || r->j0 + r->nj > pl->kp.ny || r->k0 + r->nk > pl->kp.nz)
    return fail(RFXH_EINVAL, "gradient region outside the grid");
  return RFXH_OK;
}

static RegionParams region_params(const rfxh_region* r) {
  RegionParams rp;
  rp.i0 = r->i0;
  rp.j0 = r->j0;
  rp.k0 = r->k0;
  rp.ni = r->ni;
  rp.nj = r->nj;
  rp.nk = r->nk;
  rp.n = (long long)r->ni * r->nj * r->nk;
  return rp;
}

static PsiIO psi_io(const rfxh_plan* pl, int set) {
  PsiIO ps;
  ps.pol = pl->pol;
  ps.pcoef = pl->pcoef;
  ps.n_pole = pl->n_pole;
  ps.n_debye = pl->n_debye;
  for (int a = 0; a < 3; ++a) {
    ps.pol_lo[a] = pl->pol_lo[a];
    ps.pol_n[a] = pl->pol_n[a];
  }
  for (int f = 0; f < 6; ++f)
    for (int t = 0; t < 2; ++t) {
      const long long o = set * pl->psi_set;
      float* h0 = pl->kp.psi_h[f][t];
      float* e0 = pl->kp.psi_e[f][t];
      ps.hr[f][t] = h0 ? h0 + o : nullptr;
      ps.hw[f][t] = h0 ? h0 + o : nullptr;
      ps.er[f][t] = e0 ? e0 + o : nullptr;
      ps.ew[f][t] = e0 ? e0 + o : nullptr;
    }
  return ps;
}

int rfxh_set_variant(rfxh_plan* pl, int kernel, int param) {
  if (!pl || kernel < 0 || kernel > 8) return fail(RFXH_EINVAL, "bad variant");
  if (kernel == 8) {
    if (param < 0) return fail(RFXH_EINVAL, "bad segment count");
    if (!pl->set[1][0]) return fail(RFXH_EINVAL, "the temporal-blocked march needs a second field set");
    if (pl->n_pole > 0 || pl->aniso)
      return fail(RFXH_EINVAL, "dispersive (Debye / Lorentz) and subpixel-smoothed materials run on "
                               "the two-pass kernels (0, 4, 5, 6)");
    if (pl->n_mat > kTBMat)
      return fail(RFXH_EINVAL, "the temporal-blocked march takes at most 1024 materials");
    if ((int)pl->planes.size() > kTBPlanes)
      return fail(RFXH_EINVAL, "the temporal-blocked march stages at most 4 DFT planes");
    if (pl->tb_ops_max > kTBOps)
      return fail(RFXH_EINVAL, "the temporal-blocked march stages at most 256 point operations "
                               "(sources, probes, ports) per 16 x 64 tile column");
    if (!pl->fluxes.empty())
      return fail(RFXH_EINVAL, "flux monitors and NTFF boxes run on the single-step kernels");
    const int N[3] = {pl->kp.nx, pl->kp.ny, pl->kp.nz};
    for (int a = 0; a < 3; ++a)
      if (pl->kp.cn[2 * a] + pl->kp.cn[2 * a + 1] > N[a])
        return fail(RFXH_EINVAL, "the temporal-blocked march needs CPML slabs that do not overlap");
    if (pl->kp.cn[0] + pl->kp.cn[1] + pl->kp.cn[2] + pl->kp.cn[3] + pl->kp.cn[4] + pl->kp.cn[5] >
        kTBProf)
      return fail(RFXH_EINVAL, "the temporal-blocked march takes at most 256 CPML layers in all");
    if ((long long)pl->psi_foff[5] + 4LL * pl->psi_fs[5] >= (1LL << 30))
      return fail(RFXH_EINVAL, "the temporal-blocked march addresses CPML memory below 2^30 floats");
    if (pl->cprof && !pl->psi_mid) {
      void* m = nullptr;
      HIPCHK(hipMalloc(&m, (size_t)pl->psi_floats * sizeof(float)));
      pl->owned.push_back(m);
      pl->psi_mid = static_cast<float*>(m);
    }
    pl->tb = true;
    pl->tb_seg = param;
    pl->fused = pl->march = pl->wmarch = false;
    for (int g = 0; g < 2; ++g)
      if (pl->graph_exec[g]) {
        (void)hipGraphExecDestroy(pl->graph_exec[g]);
        pl->graph_exec[g] = nullptr;
      }
    return RFXH_OK;
  }
  const bool pingpong = kernel == 1 || kernel == 2 || kernel == 3 || kernel == 7;
  if (pingpong && !pl->set[1][0])
    return fail(RFXH_EINVAL, "the fused and marching steps need a second field set");
  if (pingpong && (pl->n_pole > 0 || pl->aniso))
    return fail(RFXH_EINVAL, "dispersive (Debye / Lorentz) and subpixel-smoothed materials run on "
                             "the two-pass kernels (0, 4, 5, 6)");
  if (kernel == 2 || kernel == 3 || kernel == 7) {
    if (param < 0) return fail(RFXH_EINVAL, "bad segment count");
    if (kernel == 7) {
      if (pl->n_mat > kWMMat)
        return fail(RFXH_EINVAL, "the wave-march step takes at most 1024 materials");
      pl->wmarch_seg = param;
    } else {
      pl->march_seg = param;
      pl->march_tile = kernel - 2;
    }
  } else {
    if (param < 0 || param > 1) return fail(RFXH_EINVAL, "bad workgroup order");
    const int old_order = pl->kp.order, old_ncol = pl->ncol;
    pl->kp.order = param;
    pl->ncol = kernel == 4 ? 2 : kernel == 5 ? 4 : kernel == 6 ? 3 : 1;
    int rc = plan_grid(pl);
    if (rc) {
      pl->kp.order = old_order;
      pl->ncol = old_ncol;
      (void)plan_grid(pl);
      return rc;
    }
  }
  pl->fused = pingpong;
  pl->march = kernel == 2 || kernel == 3;
  pl->wmarch = kernel == 7;
  pl->tb = false;
  for (int g = 0; g < 2; ++g)
    if (pl->graph_exec[g]) {
      (void)hipGraphExecDestroy(pl->graph_exec[g]);
      pl->graph_exec[g] = nullptr;
    }
  return RFXH_OK;
}

int rfxh_get_variant(rfxh_plan* pl, int* kernel, int* param) {
  if (!pl || !kernel || !param) return fail(RFXH_EINVAL, "bad arguments");
  *kernel = pl->tb ? 8
          : pl->wmarch ? 7
          : pl->march ? 2 + pl->march_tile
          : pl->fused ? 1
          : pl->ncol == 2 ? 4 : pl->ncol == 4 ? 5 : pl->ncol == 3 ? 6 : 0;
  *param = pl->tb ? pl->tb_seg
         : pl->wmarch ? pl->wmarch_seg : pl->march ? pl->march_seg : pl->kp.order;
  return RFXH_OK;
}

int rfxh_run_record(rfxh_plan* pl, int64_t step0, int64_t n_steps, const rfxh_region* region,
                    float* d_rec, void* stream) {
  return rfxh_run_record_h(pl, step0, n_steps, region, d_rec, nullptr, stream);
}

int rfxh_run_record_h(rfxh_plan* pl, int64_t step0, int64_t n_steps, const rfxh_region* region,
                      float* d_rec, float* d_rec_h, void* stream) {
  return rfxh_run_record_ex(pl, step0, n_steps, region, d_rec, d_rec_h, nullptr, stream);
}

int rfxh_run_record_ex(rfxh_plan* pl, int64_t step0, int64_t n_steps, const rfxh_region* region,
                       float* d_rec, float* d_rec_h, float* d_rec_k, void* stream) {
  if (!pl) return fail(RFXH_EINVAL, "null plan");
  if (pl->aniso)
    return fail(RFXH_EINVAL, "subpixel-smoothed materials: the reverse mode covers isotropic "
                             "plans only");
  if (pl->n_pole > 0 && !d_rec_k)
    return fail(RFXH_EINVAL, "dispersive plan: the dispersive record (d_rec_k) is missing");
  if (step0 < 0 || n_steps < 0 || step0 + n_steps > pl->max_steps)
    return fail(RFXH_EINVAL, "steps outside the source/probe tables (max_steps)");
  int rc = check_region(pl, region);
  if (rc) return rc;
  if (!d_rec) return fail(RFXH_EINVAL, "record buffer missing");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const RegionParams rp = region_params(region);
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, pl->kp.step, (long long)step0);
  dft_begin(pl, s, step0);
  // two-pass sub-steps in place on the current set, with the record taken
  // between the H and the E update (E^n and H^{n+1/2} both in memory)
  KParams kp = pl->kp;
  for (int c = 0; c < 6; ++c) kp.f[c] = pl->set[pl->cur][c];
  const PsiIO ps = psi_io(pl, pl->cur);
  const unsigned rb = (unsigned)((rp.n + 255) / 256);
  for (int64_t n = 0; n < n_steps; ++n) {
    launch_pass(pl, s, kp, ps, 0, pl->cur);
    float* rh = d_rec_h ? d_rec_h + n * 3 * rp.n : nullptr;
    float* rk = (d_rec_k && pl->n_pole > 0) ? d_rec_k + n * 3 * rp.n : nullptr;
    if (kp.graded)
      hipLaunchKernelGGL(k_record<true>, dim3(rb), dim3(256), 0, s, kp, ps, rp, d_rec + n * 9 * rp.n,
                         rh, rk);
    else
      hipLaunchKernelGGL(k_record<false>, dim3(rb), dim3(256), 0, s, kp, ps, rp, d_rec + n * 9 * rp.n,
                         rh, rk);
    launch_pass(pl, s, kp, ps, 1, pl->cur);
    launch_accumulators(pl, s, pl->cur);
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, pl->kp.step);
  }
  dft_flush(pl, s);
  HIPCHK(hipGetLastError());
  return RFXH_OK;
}

// Checks of one instance's reverse sweep over steps [lo, hi) and its
// kernel parameters.
static int adjoint_params(rfxh_plan* pl, const rfxh_adjoint_desc* ad, int64_t lo, int64_t hi,
                          AdjParams* out) {
  if (!pl || !ad) return fail(RFXH_EINVAL, "null plan or adjoint descriptor");
  if (pl->aniso)
    return fail(RFXH_EINVAL, "subpixel-smoothed materials: the reverse mode covers isotropic "
                             "plans only");
  if (pl->n_pole > 0 && !ad->d_pol_adj)
    return fail(RFXH_EINVAL, "dispersive plan: the adjoint polarisation (d_pol_adj) is missing");
  if (pl->n_pole > 0 && ad->d_grad && (!ad->d_rec_k || !ad->dcoef_disp))
    return fail(RFXH_EINVAL, "dispersive plan: the dispersive record or its table is missing");
  if (lo < 0 || hi < lo || hi > pl->max_steps) return fail(RFXH_EINVAL, "bad step range");
  for (int c = 0; c < 6; ++c)
    if (!ad->d_adj[c]) return fail(RFXH_EINVAL, "adjoint field pointer missing");
  bool any_cpml = false;
  for (int f = 0; f < 6; ++f) any_cpml = any_cpml || pl->kp.cn[f] > 0;
  if (any_cpml && (!ad->d_adj_psi || !ad->d_gbar)) return fail(RFXH_EINVAL, "adjoint CPML buffers missing");
  if (pl->kp.n_probe > 0 && !ad->d_probe_adj) return fail(RFXH_EINVAL, "probe adjoint table missing");
  if (ad->d_grad) {
    int rc = check_region(pl, &ad->region);
    if (rc) return rc;
    if (!ad->d_rec || !ad->dcoef || ad->rec_rows < 0) return fail(RFXH_EINVAL, "forward record missing");
  }
  AdjParams& a = *out;
  std::memset(&a, 0, sizeof(a));
  for (int c = 0; c < 6; ++c) a.a[c] = ad->d_adj[c];
  // adjoint CPML memory and g-bar share the psi layout
  {
    // offsets of each face/term inside one psi-sized buffer (as in the forward)
    long long off = 0;
    for (int f = 0; f < 6; ++f) {
      const int n = pl->kp.cn[f];
      const int ax = f >> 1;
      long long sz = n;
      if (ax != 0) sz *= pl->kp.nx;
      if (ax != 1) sz *= pl->kp.ny;
      if (ax != 2) sz *= pl->kp.nz;
      if (n > 0) {
        a.ph[f][0] = ad->d_adj_psi + off;
        a.ph[f][1] = ad->d_adj_psi + off + sz;
        a.pe[f][0] = ad->d_adj_psi + off + 2 * sz;
        a.pe[f][1] = ad->d_adj_psi + off + 3 * sz;
        a.gh[f][0] = ad->d_gbar + off;
        a.gh[f][1] = ad->d_gbar + off + sz;
        a.ge[f][0] = ad->d_gbar + off + 2 * sz;
        a.ge[f][1] = ad->d_gbar + off + 3 * sz;
      }
      off += 4 * sz;
    }
  }
  a.probe_adj = ad->d_probe_adj;
  a.src_adj = ad->d_src_adj;
  a.port_adj = ad->d_port_adj;
  if (pl->n_pole > 0) {
    a.pol_adj = ad->d_pol_adj;
    a.pcoef = pl->pcoef;
    a.n_pole = pl->n_pole;
    a.n_debye = pl->n_debye;
    for (int c = 0; c < 3; ++c) {
      a.pol_lo[c] = pl->pol_lo[c];
      a.pol_n[c] = pl->pol_n[c];
    }
    if (ad->d_grad) {
      if (!pl->dkdis) {
        std::vector<float2> tmp(pl->n_mat);
        for (int m = 0; m < pl->n_mat; ++m)
          tmp[m] = make_float2(ad->dcoef_disp[2 * m], ad->dcoef_disp[2 * m + 1]);
        int rc = dev_copy(pl, tmp.data(), tmp.size(), &pl->dkdis);
        if (rc) return rc;
      }
      a.rec_k = ad->d_rec_k;
      a.dkdis = pl->dkdis;
    }
  }
  if (ad->d_grad) {
    if (!pl->dcoef) {
      std::vector<float4> tmp(pl->n_mat);
      for (int m = 0; m < pl->n_mat; ++m)
        tmp[m] = make_float4(ad->dcoef[3 * m], ad->dcoef[3 * m + 1], ad->dcoef[3 * m + 2], 0.f);
      int rc = dev_copy(pl, tmp.data(), tmp.size(), &pl->dcoef);
      if (rc) return rc;
    }
    a.rec = ad->d_rec;
    a.rec_step0 = ad->rec_step0;
    a.rec_rows = ad->rec_rows;
    a.reg = region_params(&ad->region);
    a.grad = ad->d_grad;
    a.dcoef = pl->dcoef;
    if (ad->d_grad_sigma) {
      if (!ad->dcoef_sigma) return fail(RFXH_EINVAL, "sigma derivative table missing");
      if (!pl->dcoef_s) {
        std::vector<float2> tmp(pl->n_mat);
        for (int m = 0; m < pl->n_mat; ++m)
          tmp[m] = make_float2(ad->dcoef_sigma[2 * m], ad->dcoef_sigma[2 * m + 1]);
        int rc = dev_copy(pl, tmp.data(), tmp.size(), &pl->dcoef_s);
        if (rc) return rc;
      }
      a.grad_s = ad->d_grad_sigma;
      a.dcoef_s = pl->dcoef_s;
    }
    if (ad->d_grad_mu) {
      if (!ad->dcoef_mu || !ad->d_rec_h) return fail(RFXH_EINVAL, "mu_r gradient: H record or table missing");
      if (!pl->dmu) {
        int rc = dev_copy(pl, ad->dcoef_mu, (size_t)pl->n_mat, &pl->dmu);
        if (rc) return rc;
      }
      a.rec_h = ad->d_rec_h;
      a.grad_mu = ad->d_grad_mu;
      a.dmu = pl->dmu;
    }
  }
  return RFXH_OK;
}

}  // extern "C"

// Per-cell cb and ch of a plan's reverse sweep, built once (the material ids
// are fixed), into a.
static int use_cbch(rfxh_plan* pl, hipStream_t s, AdjParams& a) {
  if (!pl->cbv) {
    const long long ncell = (long long)pl->kp.nx * pl->kp.ny * pl->kp.nz;
    void* d = nullptr;
    HIPCHK(hipMalloc(&d, 2 * (size_t)ncell * sizeof(float)));
    pl->owned.push_back(d);
    pl->cbv = static_cast<float*>(d);
    pl->chv = pl->cbv + ncell;
    hipLaunchKernelGGL(k_fill_cbch, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, pl->kp,
                       pl->cbv, pl->chv, ncell);
  }
  a.cbv = pl->cbv;
  a.chv = pl->chv;
  a.h_probe = pl->h_probe;
  return RFXH_OK;
}

extern "C" {

int rfxh_adjoint(rfxh_plan* pl, const rfxh_adjoint_desc* ad, int64_t lo, int64_t hi, void* stream) {
  AdjParams a;
  const int rc = adjoint_params(pl, ad, lo, hi, &a);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (hi == lo) return RFXH_OK;
  {
    const int rc2 = use_cbch(pl, s, a);
    if (rc2) return rc2;
  }
  // DFT-plane adjoints: planes of an E component take theirs just before
  // the prologue of their step (the sample is E^{n+1} after the injection,
  // like a probe's), planes of an H component before the H-bar pass of
  // their step
  // (flux monitors and NTFF faces the same, with E and H parts both)
  bool e_planes = false, h_planes = false;
  if (ad->d_plane_adj)
    for (size_t q = 0; q < pl->planes.size(); ++q)
      if (ad->d_plane_adj[q]) (pl->plane_comp[q] < 3 ? e_planes : h_planes) = true;
  if (ad->d_flux_adj)
    for (size_t q = 0; q < pl->fluxes.size(); ++q)
      if (ad->d_flux_adj[q]) e_planes = h_planes = true;
  auto plane_adj = [&](bool e, long long step) {
    if (ad->d_plane_adj)
      for (size_t q = 0; q < pl->planes.size(); ++q) {
        const int comp = pl->plane_comp[q];
        if (!ad->d_plane_adj[q] || (comp < 3) != e) continue;
        const PlaneParams& pp = pl->planes[q];
        const long long npts = (long long)pp.n1 * pp.n2;
        hipLaunchKernelGGL(k_plane_adj, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, pp,
                           reinterpret_cast<const double2*>(ad->d_plane_adj[q]), step, a.a[comp]);
      }
    if (ad->d_flux_adj)
      for (size_t q = 0; q < pl->fluxes.size(); ++q) {
        if (!ad->d_flux_adj[q]) continue;
        const FluxParams& fp = pl->fluxes[q];
        const long long npts = (long long)fp.n1 * fp.n2;
        const int* comp = &pl->flux_comp[4 * q + (e ? 0 : 2)];
        hipLaunchKernelGGL(k_flux_adj, dim3((unsigned)((npts + 255) / 256), 2), dim3(256), 0, s,
                           fp, reinterpret_cast<const double2*>(ad->d_flux_adj[q]), step,
                           e ? 1 : 0, a.a[comp[0]], a.a[comp[1]]);
      }
  };
  const bool disp = pl->n_pole > 0;
  auto adj_e = [&](int gather, long long m) {
    if (pl->kp.graded) {
      if (disp)
        hipLaunchKernelGGL((k_adj_e<true, true>), pl->grid, pl->block, 0, s, pl->kp, a, gather, m);
      else
        hipLaunchKernelGGL(k_adj_e<true>, pl->grid, pl->block, 0, s, pl->kp, a, gather, m);
    } else {
      if (disp)
        hipLaunchKernelGGL((k_adj_e<false, true>), pl->grid, pl->block, 0, s, pl->kp, a, gather, m);
      else
        hipLaunchKernelGGL(k_adj_e<false>, pl->grid, pl->block, 0, s, pl->kp, a, gather, m);
    }
  };
  // prologue of step hi-1 (E-bar holds the adjoint of E^{hi})
  if (e_planes) plane_adj(true, (long long)(hi - 1));
  adj_e(0, (long long)(hi - 1));
  const bool ports = a.port_adj && pl->kp.n_port > 0;
  for (int64_t n = hi - 1; n >= lo; --n) {
    const long long m = n - 1 >= lo ? (long long)(n - 1) : -1;
    if (ports)
      hipLaunchKernelGGL(k_port_adj_h, dim3((unsigned)((pl->kp.n_ops + 63) / 64)), dim3(64), 0, s,
                         pl->kp, a, (long long)n);
    if (h_planes) plane_adj(false, (long long)n);
    if (pl->kp.graded)
      hipLaunchKernelGGL(k_adj_h<true>, pl->grid, pl->block, 0, s, pl->kp, a, (long long)n);
    else
      hipLaunchKernelGGL(k_adj_h<false>, pl->grid, pl->block, 0, s, pl->kp, a, (long long)n);
    if (e_planes && m >= 0) {
      // the gather, then the E-plane adjoints of step m, then its prologue
      adj_e(1, -1);
      plane_adj(true, m);
      adj_e(0, m);
    } else {
      adj_e(1, m);
    }
  }
  HIPCHK(hipGetLastError());
  return RFXH_OK;
}

// ---- batched sweeps --------------------------------------------------------
}  // extern "C"

struct rfxh_batch {
  std::vector<rfxh_plan*> plans;
  BatchEntry* d_ents = nullptr;   // instance table in run order
  int* d_start = nullptr;         // first logical workgroup per instance (+ total)
  long long* d_steps = nullptr;   // step0 per instance, run order
  std::vector<int> order;         // run order -> plan index
  std::vector<int> start;         // first logical workgroup per instance (+ total)
  hipStream_t cap_stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::map<int, hipGraphExec_t> graphs;  // keyed by active count
  int graph_chunk = 16;
  // tables of the record and reverse-mode calls (the run's captured graphs
  // read d_ents, so these calls keep their own)
  BatchEntry* d_xents = nullptr;
  int* d_xstart = nullptr;
  long long* d_xsteps = nullptr;
  BatchRecEntry* d_rents = nullptr;
  BatchAdjEntry* d_aents = nullptr;
};

namespace {

void batch_drop_graphs(rfxh_batch* b) {
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
}

// Instance table rows (current field sets) and first logical workgroups of
// the plans in `order`.
int batch_table(const rfxh_batch* b, const std::vector<int>& order, std::vector<BatchEntry>& ents,
                std::vector<int>& start) {
  const int n = (int)order.size();
  ents.assign(n, BatchEntry());
  start.assign(n + 1, 0);
  for (int r = 0; r < n; ++r) {
    rfxh_plan* pl = b->plans[order[r]];
    ents[r].p = pl->kp;
    for (int c = 0; c < 6; ++c) ents[r].p.f[c] = pl->set[pl->cur][c];
    ents[r].ps = psi_io(pl, pl->cur);
    const long long nb = (long long)start[r] + batch_blocks(pl);
    if (nb > 0x7fffffffLL) return fail(RFXH_EINVAL, "batch too large for one launch");
    start[r + 1] = (int)nb;
  }
  return RFXH_OK;
}

// One pass (0: H, 1: E) of the first n_act instances of a table whose rows
// are the plans in `order`; nb = their logical workgroups.
void batch_pass(const rfxh_batch* b, hipStream_t s, const BatchEntry* d_ents, const int* d_start,
                const std::vector<int>& order, unsigned nb, int n_act, int pass) {
  if (pass == 0) {
    bool one_mu = true;
    for (int q = 0; q < n_act; ++q) one_mu = one_mu && b->plans[order[q]]->kp.ch_all > 0.f;
    if (one_mu)
      hipLaunchKernelGGL((k_batch_pass<0, true, false>), dim3(nb), dim3(kBX, kTY), 0, s, d_ents,
                         d_start, n_act);
    else
      hipLaunchKernelGGL((k_batch_pass<0, false, false>), dim3(nb), dim3(kBX, kTY), 0, s, d_ents,
                         d_start, n_act);
  } else {
    bool disp = false;
    for (int q = 0; q < n_act; ++q)
      disp = disp || b->plans[order[q]]->n_pole > 0 || b->plans[order[q]]->aniso;
    if (disp)
      hipLaunchKernelGGL((k_batch_pass<1, false, true>), dim3(nb), dim3(kBX, kTY), 0, s, d_ents,
                         d_start, n_act);
    else
      hipLaunchKernelGGL((k_batch_pass<1, false, false>), dim3(nb), dim3(kBX, kTY), 0, s, d_ents,
                         d_start, n_act);
  }
}

// One step of the first n_act instances (run order).
int batch_step(rfxh_batch* b, hipStream_t s, int n_act) {
  const unsigned nb = (unsigned)b->start[n_act];
  batch_pass(b, s, b->d_ents, b->d_start, b->order, nb, n_act, 0);
  batch_pass(b, s, b->d_ents, b->d_start, b->order, nb, n_act, 1);
  for (int q = 0; q < n_act; ++q) {
    rfxh_plan* pl = b->plans[b->order[q]];
    launch_accumulators(pl, s, pl->cur);
  }
  hipLaunchKernelGGL(k_batch_advance, dim3((unsigned)((n_act + 63) / 64)), dim3(64), 0, s, b->d_ents,
                     n_act);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int rfxh_batch_create(rfxh_plan* const* plans, int n_plans, rfxh_batch** out) {
  if (!plans || n_plans < 1 || !out) return fail(RFXH_EINVAL, "no plans");
  *out = nullptr;
  for (int q = 0; q < n_plans; ++q)
    if (!plans[q]) return fail(RFXH_EINVAL, "null plan in batch");
  rfxh_batch* b = new rfxh_batch();
  b->plans.assign(plans, plans + n_plans);
  auto bail = [&](int r) {
    rfxh_batch_destroy(b);
    return r;
  };
  if (hipMalloc(&b->d_ents, sizeof(BatchEntry) * n_plans) != hipSuccess ||
      hipMalloc(&b->d_start, sizeof(int) * (n_plans + 1)) != hipSuccess ||
      hipMalloc(&b->d_steps, sizeof(long long) * n_plans) != hipSuccess ||
      hipMalloc(&b->d_xents, sizeof(BatchEntry) * n_plans) != hipSuccess ||
      hipMalloc(&b->d_xstart, sizeof(int) * (n_plans + 1)) != hipSuccess ||
      hipMalloc(&b->d_xsteps, sizeof(long long) * n_plans) != hipSuccess ||
      hipMalloc(&b->d_rents, sizeof(BatchRecEntry) * n_plans) != hipSuccess ||
      hipMalloc(&b->d_aents, sizeof(BatchAdjEntry) * n_plans) != hipSuccess)
    return bail(fail(RFXH_ENOMEM, "batch tables"));
  *out = b;
  return RFXH_OK;
}

int rfxh_batch_destroy(rfxh_batch* b) {
  if (!b) return RFXH_OK;
  batch_drop_graphs(b);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->d_ents) (void)hipFree(b->d_ents);
  if (b->d_start) (void)hipFree(b->d_start);
  if (b->d_steps) (void)hipFree(b->d_steps);
  for (void* d : {(void*)b->d_xents, (void*)b->d_xstart, (void*)b->d_xsteps, (void*)b->d_rents,
                  (void*)b->d_aents})
    if (d) (void)hipFree(d);
  delete b;
  return RFXH_OK;
}

int rfxh_batch_run(rfxh_batch* b, const int64_t* step0, const int64_t* n_steps, void* stream,
                   int use_graph) {
  if (!b || !step0 || !n_steps) return fail(RFXH_EINVAL, "bad batch arguments");
  const int n = (int)b->plans.size();
  for (int q = 0; q < n; ++q) {
    const rfxh_plan* pl = b->plans[q];
    if (step0[q] < 0 || n_steps[q] < 0 || step0[q] + n_steps[q] > pl->max_steps)
      return fail(RFXH_EINVAL, "batch instance steps outside its tables (max_steps)");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  // run order: most steps first, so the active instances are always a prefix
  std::vector<int> order(n);
  for (int q = 0; q < n; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return n_steps[a] > n_steps[c]; });
  std::vector<BatchEntry> ents;
  std::vector<int> start;
  const int trc = batch_table(b, order, ents, start);
  if (trc) return trc;
  std::vector<long long> st(n);
  for (int r = 0; r < n; ++r) st[r] = step0[order[r]];
  // the captured graphs launch per-instance plane kernels: keep them only
  // while the instance table is unchanged
  bool same = order == b->order && start == b->start;
  if (same) {
    std::vector<BatchEntry> old(n);
    HIPCHK(hipMemcpy(old.data(), b->d_ents, sizeof(BatchEntry) * n, hipMemcpyDeviceToHost));
    same = std::memcmp(old.data(), ents.data(), sizeof(BatchEntry) * n) == 0;
  }
  if (!same) {
    batch_drop_graphs(b);
    HIPCHK(hipMemcpyAsync(b->d_ents, ents.data(), sizeof(BatchEntry) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->d_start, start.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    b->order = order;
    b->start = start;
  }
  HIPCHK(hipMemcpyAsync(b->d_steps, st.data(), sizeof(long long) * n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_batch_set_steps, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, b->d_ents,
                     b->d_steps, n);
  for (int q = 0; q < n; ++q) dft_begin(b->plans[q], s, step0[q]);
  HIPCHK(hipGetLastError());
  if (use_graph && !b->cap_stream) {
    HIPCHK(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
  }
  hipStream_t w = s;
  if (use_graph) {
    HIPCHK(hipEventRecord(b->ev_in, s));
    HIPCHK(hipStreamWaitEvent(b->cap_stream, b->ev_in, 0));
    w = b->cap_stream;
  }
  int64_t done = 0;
  for (int act = n; act > 0; --act) {
    const int64_t until = n_steps[order[act - 1]];  // the shortest active instance ends here
    int64_t len = until - done;
    if (len <= 0) continue;
    if (use_graph && len >= b->graph_chunk) {
      auto it = b->graphs.find(act);
      if (it == b->graphs.end()) {
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        for (int c = 0; c < b->graph_chunk; ++c) {
          const int rc = batch_step(b, b->cap_stream, act);
          if (rc) {
            (void)hipStreamEndCapture(b->cap_stream, &graph);
            return rc;
          }
        }
        HIPCHK(hipStreamEndCapture(b->cap_stream, &graph));
        hipGraphExec_t ge;
        HIPCHK(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(graph));
        it = b->graphs.emplace(act, ge).first;
      }
      while (len >= b->graph_chunk) {
        HIPCHK(hipGraphLaunch(it->second, w));
        len -= b->graph_chunk;
        done += b->graph_chunk;
      }
    }
    for (; len > 0; --len, ++done) {
      const int rc = batch_step(b, w, act);
      if (rc) return rc;
    }
  }
  for (int q = 0; q < n; ++q) dft_flush(b->plans[q], w);
  HIPCHK(hipGetLastError());
  if (use_graph) {
    HIPCHK(hipEventRecord(b->ev_out, b->cap_stream));
    HIPCHK(hipStreamWaitEvent(s, b->ev_out, 0));
  }
  return RFXH_OK;
}

int rfxh_batch_run_record(rfxh_batch* b, const int64_t* step0, const int64_t* n_steps,
                          const rfxh_region* regions, float* const* d_rec, int accumulate,
                          void* stream) {
  if (!b || !step0 || !n_steps || !regions || !d_rec)
    return fail(RFXH_EINVAL, "bad batch record arguments");
  const int n = (int)b->plans.size();
  std::vector<int> order;
  for (int q = 0; q < n; ++q) {
    rfxh_plan* pl = b->plans[q];
    if (step0[q] < 0 || n_steps[q] < 0 || step0[q] + n_steps[q] > pl->max_steps)
      return fail(RFXH_EINVAL, "batch instance steps outside its tables (max_steps)");
    if (n_steps[q] == 0) continue;
    if (pl->n_pole > 0 || pl->aniso)
      return fail(RFXH_EINVAL, "dispersive or subpixel-smoothed materials: the reverse mode "
                               "covers plain isotropic plans only");
    const int rc = check_region(pl, &regions[q]);
    if (rc) return rc;
    if (!d_rec[q]) return fail(RFXH_EINVAL, "record buffer missing");
    order.push_back(q);
  }
  if (order.empty()) return RFXH_OK;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return n_steps[a] > n_steps[c]; });
  const int na = (int)order.size();
  std::vector<BatchEntry> ents;
  std::vector<int> start;
  int rc = batch_table(b, order, ents, start);
  if (rc) return rc;
  std::vector<BatchRecEntry> recs(na);
  std::vector<long long> st(na);
  for (int r = 0; r < na; ++r) {
    const int q = order[r];
    recs[r].r = region_params(&regions[q]);
    recs[r].rec = d_rec[q];
    recs[r].rec_step0 = step0[q];
    st[r] = step0[q];
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIPCHK(hipMemcpyAsync(b->d_xents, ents.data(), sizeof(BatchEntry) * na, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(b->d_xstart, start.data(), sizeof(int) * (na + 1), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(b->d_xsteps, st.data(), sizeof(long long) * na, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(b->d_rents, recs.data(), sizeof(BatchRecEntry) * na, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_batch_set_steps, dim3((unsigned)((na + 63) / 64)), dim3(64), 0, s, b->d_xents,
                     b->d_xsteps, na);
  if (accumulate)
    for (int q : order) dft_begin(b->plans[q], s, step0[q]);
  int act = na;
  for (int64_t t = 0; t < n_steps[order[0]]; ++t) {
    while (act > 0 && n_steps[order[act - 1]] <= t) --act;
    const unsigned nb = (unsigned)start[act];
    batch_pass(b, s, b->d_xents, b->d_xstart, order, nb, act, 0);
    hipLaunchKernelGGL(k_batch_e_rec, dim3(nb), dim3(kBX, kTY), 0, s, b->d_xents, b->d_rents,
                       b->d_xstart, act);
    if (accumulate)
      for (int r = 0; r < act; ++r) launch_accumulators(b->plans[order[r]], s, b->plans[order[r]]->cur);
    hipLaunchKernelGGL(k_batch_advance, dim3((unsigned)((act + 63) / 64)), dim3(64), 0, s,
                       b->d_xents, act);
  }
  if (accumulate)
    for (int q : order) dft_flush(b->plans[q], s);
  HIPCHK(hipGetLastError());
  return RFXH_OK;
}

int rfxh_batch_adjoint(rfxh_batch* b, const rfxh_adjoint_desc* const* descs, const int64_t* lo,
                       const int64_t* hi, void* stream) {
  if (!b || !descs || !lo || !hi) return fail(RFXH_EINVAL, "bad batch adjoint arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int n = (int)b->plans.size();
  std::vector<int> order;
  std::vector<BatchAdjEntry> all(n);
  for (int q = 0; q < n; ++q) {
    if (!descs[q]) continue;
    if (descs[q]->d_grad_mu || descs[q]->d_port_adj || descs[q]->d_plane_adj ||
        descs[q]->d_flux_adj || b->plans[q]->n_pole > 0)
      return fail(RFXH_EINVAL, "the batched reverse mode computes eps_r and sigma gradients of "
                               "probe objectives only");
    int rc = adjoint_params(b->plans[q], descs[q], lo[q], hi[q], &all[q].a);
    if (rc) return rc;
    if ((rc = use_cbch(b->plans[q], s, all[q].a))) return rc;
    all[q].lo = lo[q];
    all[q].hi = hi[q];
    if (hi[q] > lo[q]) order.push_back(q);
  }
  if (order.empty()) return RFXH_OK;
  // longest sweep first: the active instances are always a prefix
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return hi[a] - lo[a] > hi[c] - lo[c]; });
  const int na = (int)order.size();
  std::vector<BatchEntry> ents;
  std::vector<int> start;
  int rc = batch_table(b, order, ents, start);
  if (rc) return rc;
  std::vector<BatchAdjEntry> adj(na);
  for (int r = 0; r < na; ++r) adj[r] = all[order[r]];
  HIPCHK(hipMemcpyAsync(b->d_xents, ents.data(), sizeof(BatchEntry) * na, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(b->d_xstart, start.data(), sizeof(int) * (na + 1), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(b->d_aents, adj.data(), sizeof(BatchAdjEntry) * na, hipMemcpyHostToDevice, s));
  int n_graded = 0;
  for (int q : order) n_graded += b->plans[q]->kp.graded ? 1 : 0;
  const int gm = n_graded == 0 ? 0 : n_graded == na ? 1 : 2;
  // prologue of step hi-1 of every instance
  launch_batch_adj(s, true, gm, (unsigned)start[na], b->d_xents, b->d_aents, b->d_xstart, na, 0,
                   0LL);
  int act = na;
  const long long len0 = hi[order[0]] - lo[order[0]];
  for (long long r = 0; r < len0; ++r) {
    while (act > 0 && hi[order[act - 1]] - lo[order[act - 1]] <= r) --act;
    const unsigned nb = (unsigned)start[act];
    launch_batch_adj(s, false, gm, nb, b->d_xents, b->d_aents, b->d_xstart, act, 1, r);
    launch_batch_adj(s, true, gm, nb, b->d_xents, b->d_aents, b->d_xstart, act, 1, r);
  }
  HIPCHK(hipGetLastError());
  return RFXH_OK;
}

}  // extern "C"

// Checks of a far-field request and its device tables (positions and widths
// per axis, directions, wave numbers) in one allocation `*blk` (extra:
// further doubles reserved after them, at *extra_at), uploaded on stream s;
// fills q except the accumulators and outputs.
static int far_field_setup(const rfxh_farfield_desc* dsc, int32_t n_theta, const double* theta,
                           int32_t n_phi, const double* phi, hipStream_t s, size_t extra,
                           FarFieldParams& q, double** blk, double** extra_at) {
  *blk = nullptr;
  if (dsc->n_freq < 1 || dsc->n_freq > 65535 || !dsc->freqs)
    return fail(RFXH_EINVAL, "far field: need 1..65535 frequencies");
  if ((long long)n_theta * n_phi > (1LL << 30))
    return fail(RFXH_EINVAL, "far field: too many directions");
  for (int a = 0; a < 3; ++a) {
    if (!dsc->pos[a] || !dsc->width[a] || dsc->n_axis[a] < 1)
      return fail(RFXH_EINVAL, "far field: missing position / width table");
    if (dsc->lo[a] < 0 || dsc->hi[a] <= dsc->lo[a] || dsc->hi[a] >= dsc->n_axis[a])
      return fail(RFXH_EINVAL, "far field: need 0 <= lo < hi < n on every axis");
  }
  const int n_dir = n_theta * n_phi;
  std::vector<double> dirs((size_t)n_dir * 9);
  for (int it = 0; it < n_theta; ++it) {
    const double st = std::sin(theta[it]), ct = std::cos(theta[it]);
    for (int ip = 0; ip < n_phi; ++ip) {
      const double sp = std::sin(phi[ip]), cp = std::cos(phi[ip]);
      double* o = dirs.data() + 9 * ((size_t)it * n_phi + ip);
      o[0] = st * cp; o[1] = st * sp; o[2] = ct;     // r-hat
      o[3] = ct * cp; o[4] = ct * sp; o[5] = -st;    // theta-hat
      o[6] = -sp;     o[7] = cp;      o[8] = 0.0;    // phi-hat
    }
  }
  std::vector<double> kw(dsc->n_freq);
  for (int f = 0; f < dsc->n_freq; ++f) kw[f] = 2.0 * M_PI * dsc->freqs[f] / dsc->c0;
  // one device block for every table: positions and widths per axis, the
  // directions and the wave numbers
  size_t total = dirs.size() + kw.size();
  for (int a = 0; a < 3; ++a) total += 2 * (size_t)dsc->n_axis[a];
  total = (total + 1) & ~(size_t)1;   // the extra space holds complex128: 16-byte aligned
  if (hipMalloc(blk, (total + extra) * sizeof(double)) != hipSuccess) {
    *blk = nullptr;
    return fail(RFXH_ENOMEM, "far field: device tables");
  }
  std::memset(&q, 0, sizeof(q));
  size_t off = 0;
  auto put = [&](const double* src, size_t n) -> const double* {
    double* dst = *blk + off;
    off += n;
    return hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess
               ? dst : nullptr;
  };
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    ok = ok && (q.pos[a] = put(dsc->pos[a], dsc->n_axis[a])) != nullptr;
    ok = ok && (q.width[a] = put(dsc->width[a], dsc->n_axis[a])) != nullptr;
  }
  ok = ok && (q.dirs = put(dirs.data(), dirs.size())) != nullptr;
  ok = ok && (q.kwave = put(kw.data(), kw.size())) != nullptr;
  if (extra_at) *extra_at = *blk + total;
  for (int a = 0; a < 3; ++a) {
    q.lo[a] = dsc->lo[a];
    q.hi[a] = dsc->hi[a];
  }
  q.nf = dsc->n_freq;
  q.n_dir = n_dir;
  q.eta0 = dsc->eta0;
  // the caller waits for the stream before freeing *blk
  return ok ? RFXH_OK : fail(RFXH_EHIP, "far field: table upload");
}

extern "C" {

int rfxh_ntff_far_field(const rfxh_farfield_desc* dsc, int32_t n_theta, const double* theta,
                        int32_t n_phi, const double* phi, double* e_theta, double* e_phi,
                        void* stream) {
  if (!dsc || !theta || !phi || !e_theta || !e_phi || n_theta < 1 || n_phi < 1)
    return fail(RFXH_EINVAL, "far field: missing arguments");
  for (int face = 0; face < 6; ++face)
    if (!dsc->d_acc[face]) return fail(RFXH_EINVAL, "far field: face accumulator is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FarFieldParams q;
  double* blk = nullptr;
  int rc = far_field_setup(dsc, n_theta, theta, n_phi, phi, s, 0, q, &blk, nullptr);
  if (rc == RFXH_OK) {
    for (int face = 0; face < 6; ++face) q.acc[face] = reinterpret_cast<const double2*>(dsc->d_acc[face]);
    q.e_theta = reinterpret_cast<double2*>(e_theta);
    q.e_phi = reinterpret_cast<double2*>(e_phi);
    hipLaunchKernelGGL(k_far_field, dim3((unsigned)((q.n_dir + kFFBlock - 1) / kFFBlock),
                                         (unsigned)dsc->n_freq), dim3(kFFBlock), 0, s, q);
    if (hipGetLastError() != hipSuccess) rc = fail(RFXH_EHIP, "far field: launch");
  }
  if (!blk) return rc;
  // the uploads and the kernel must finish before the tables are freed
  if (hipStreamSynchronize(s) != hipSuccess && rc == RFXH_OK)
    rc = fail(RFXH_EHIP, "far field: kernel");
  (void)hipFree(blk);
  return rc;
}

int rfxh_ntff_far_field_adjoint(const rfxh_farfield_desc* dsc, int32_t n_theta,
                                const double* theta, int32_t n_phi, const double* phi,
                                const double* g_theta, const double* g_phi,
                                double* const* d_gacc, void* stream) {
  if (!dsc || !theta || !phi || !g_theta || !g_phi || !d_gacc || n_theta < 1 || n_phi < 1)
    return fail(RFXH_EINVAL, "far field adjoint: missing arguments");
  for (int face = 0; face < 6; ++face)
    if (!d_gacc[face]) return fail(RFXH_EINVAL, "far field adjoint: face gradient is NULL");
  if ((long long)n_theta * n_phi > (1LL << 30))
    return fail(RFXH_EINVAL, "far field: too many directions");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FarFieldParams q;
  double* blk = nullptr;
  double* gs = nullptr;
  // scratch: twelve complex sums per (frequency, direction)
  const size_t extra = (size_t)24 * (size_t)std::max(dsc->n_freq, 0) * (size_t)n_theta * n_phi;
  int rc = far_field_setup(dsc, n_theta, theta, n_phi, phi, s, extra, q, &blk, &gs);
  if (rc == RFXH_OK) {
    hipLaunchKernelGGL(k_far_field_adj_s, dim3((unsigned)((q.n_dir + 255) / 256), (unsigned)q.nf),
                       dim3(256), 0, s, q, reinterpret_cast<const double2*>(g_theta),
                       reinterpret_cast<const double2*>(g_phi), reinterpret_cast<double2*>(gs));
    for (int face = 0; face < 6; ++face) {
      const int ax = face / 2;
      const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
      const long long npts = (long long)(q.hi[t1] - q.lo[t1]) * (q.hi[t2] - q.lo[t2]);
      hipLaunchKernelGGL(k_far_field_adj, dim3((unsigned)((npts + kFFBlock - 1) / kFFBlock),
                                               (unsigned)q.nf), dim3(kFFBlock), 0, s, q,
                         reinterpret_cast<const double2*>(gs), face,
                         reinterpret_cast<double2*>(d_gacc[face]));
    }
    if (hipGetLastError() != hipSuccess) rc = fail(RFXH_EHIP, "far field adjoint: launch");
  }
  if (!blk) return rc;
  if (hipStreamSynchronize(s) != hipSuccess && rc == RFXH_OK)
    rc = fail(RFXH_EHIP, "far field adjoint: kernel");
  (void)hipFree(blk);
  return rc;
}

}  // extern "C"

// ---- far-field transform without allocation ---------------------------------
// The tables of rfxh_ntff_far_field_noalloc live in one file-scope device
// array per GPU (8 MiB of float64, held until the process ends).  Host code
// never passes the array's name to a kernel: each device's address comes
// from hipGetSymbolAddress, taken once per device with that device current.
constexpr long long kFFScratch = 1LL << 20;   // doubles
static __device__ __align__(256) double scratch[kFFScratch];

namespace {
constexpr int kScratchDevices = 64;
double* g_scratch[kScratchDevices] = {};
std::mutex g_scratch_mu;

// Address of this device's scratch array (the current device's copy).
int scratch_address(double** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices)
    return fail(RFXH_EINVAL, "far field: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    double* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(RFXH_EHIP, "far field: no address for the scratch array");
    g_scratch[dev] = ptr;
  }
  *out = g_scratch[dev];
  return 0;
}
}  // namespace

extern "C" {

int rfxh_ntff_far_field_noalloc(const rfxh_farfield_desc* dsc, int32_t n_theta,
                                const double* theta, int32_t n_phi, const double* phi,
                                double* e_theta, double* e_phi, void* stream) {
  if (!dsc || !theta || !phi || !e_theta || !e_phi || n_theta < 1 || n_phi < 1)
    return fail(RFXH_EINVAL, "far field: missing arguments");
  if (dsc->n_freq < 1 || dsc->n_freq > 65535 || !dsc->freqs)
    return fail(RFXH_EINVAL, "far field: need 1..65535 frequencies");
  for (int a = 0; a < 3; ++a) {
    if (!dsc->pos[a] || !dsc->width[a] || dsc->n_axis[a] < 1)
      return fail(RFXH_EINVAL, "far field: missing position / width table");
    if (dsc->lo[a] < 0 || dsc->hi[a] <= dsc->lo[a] || dsc->hi[a] >= dsc->n_axis[a])
      return fail(RFXH_EINVAL, "far field: need 0 <= lo < hi < n on every axis");
  }
  for (int face = 0; face < 6; ++face)
    if (!dsc->d_acc[face]) return fail(RFXH_EINVAL, "far field: face accumulator is NULL");
  const long long n_dir = (long long)n_theta * n_phi;
  // layout: pos / width per axis, directions, wave numbers, theta, phi, freqs
  long long total = 9 * n_dir + 2LL * dsc->n_freq + n_theta + n_phi;
  for (int a = 0; a < 3; ++a) total += 2LL * dsc->n_axis[a];
  if (total > kFFScratch)
    return fail(RFXH_EINVAL,
                "far field (no allocation): tables need " + std::to_string(total) +
                    " doubles, more than the 1048576 of the scratch array (9 per direction, "
                    "2 per frequency, 2 per grid index); use rfxh_ntff_far_field");
  double* base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  FarFieldParams q;
  std::memset(&q, 0, sizeof(q));
  long long off = 0;
  bool ok = true;
  auto put = [&](const double* src, long long n) -> double* {
    double* dst = base + off;
    off += n;
    if (src)
      ok = ok && hipMemcpyAsync(dst, src, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s) ==
                     hipSuccess;
    return dst;
  };
  for (int a = 0; a < 3; ++a) {
    q.pos[a] = put(dsc->pos[a], dsc->n_axis[a]);
    q.width[a] = put(dsc->width[a], dsc->n_axis[a]);
  }
  double* dirs = put(nullptr, 9 * n_dir);
  double* kw = put(nullptr, dsc->n_freq);
  const double* d_theta = put(theta, n_theta);
  const double* d_phi = put(phi, n_phi);
  const double* d_freq = put(dsc->freqs, dsc->n_freq);
  if (!ok) return fail(RFXH_EHIP, "far field: table upload");
  const long long nt = std::max<long long>(n_dir, dsc->n_freq);
  hipLaunchKernelGGL(k_far_field_dirs, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s,
                     d_theta, n_theta, d_phi, n_phi, d_freq, dsc->n_freq, dsc->c0, dirs, kw);
  q.dirs = dirs;
  q.kwave = kw;
  for (int face = 0; face < 6; ++face) q.acc[face] = reinterpret_cast<const double2*>(dsc->d_acc[face]);
  for (int a = 0; a < 3; ++a) {
    q.lo[a] = dsc->lo[a];
    q.hi[a] = dsc->hi[a];
  }
  q.nf = dsc->n_freq;
  q.n_dir = (int)n_dir;
  q.eta0 = dsc->eta0;
  q.e_theta = reinterpret_cast<double2*>(e_theta);
  q.e_phi = reinterpret_cast<double2*>(e_phi);
  hipLaunchKernelGGL(k_far_field, dim3((unsigned)((n_dir + kFFBlock - 1) / kFFBlock),
                                       (unsigned)dsc->n_freq), dim3(kFFBlock), 0, s, q);
  if (hipGetLastError() != hipSuccess) return fail(RFXH_EHIP, "far field: launch");
  // the host tables are read by the copies: wait for them (and the kernel),
  // as rfxh_ntff_far_field does
  HIPCHK(hipStreamSynchronize(s));
  return RFXH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device-side material assembly: box fill, CPML pad extension, tangential
// PEC bits on the uint16 id array (one thread per cell or per column).
// ---------------------------------------------------------------------------
namespace {

// Box fill: the last non-PEC box covering the cell paints it (insertion
// order, later wins); any PEC box covering it marks it.
__global__ __launch_bounds__(256) void k_asm_fill(const rfxh_asm_box* __restrict__ boxes,
                                                  int n_box, int nx, int ny, int nz,
                                                  uint16_t* __restrict__ ids,
                                                  uint8_t* __restrict__ pec) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long P = (long long)ny * nz;
  if (c >= (long long)nx * P) return;
  const int i = (int)(c / P);
  const int r = (int)(c - (long long)i * P);
  const int j = r / nz, k = r - (r / nz) * nz;
  int id = 0;
  bool painted = false, p = false;
  for (int b = n_box - 1; b >= 0; --b) {
    const rfxh_asm_box& bx = boxes[b];
    const bool in = i >= bx.lo[0] && i < bx.hi[0] && j >= bx.lo[1] && j < bx.hi[1] &&
                    k >= bx.lo[2] && k < bx.hi[2];
    if (!in) continue;
    if (bx.pec) {
      p = true;
    } else if (!painted) {
      id = bx.id;
      painted = true;
    }
  }
  ids[c] = (uint16_t)id;
  pec[c] = p ? 1 : 0;
}

// Pad extension along axis a: one thread per line along a (the other two
// indices), lo pad first, then the hi pad with its vacuum rule.
__global__ __launch_bounds__(256) void k_asm_pads(int a, int plo, int phi, int nx, int ny, int nz,
                                                  const uint8_t* __restrict__ vacuum, int n_mat,
                                                  uint16_t* __restrict__ ids) {
  const int N[3] = {nx, ny, nz};
  const long long stride[3] = {(long long)ny * nz, (long long)nz, 1};
  const int a1 = a == 0 ? 1 : 0, a2 = a == 2 ? 1 : 2;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N[a1] * N[a2]) return;
  const int u = (int)(t / N[a2]), v = (int)(t - (t / N[a2]) * N[a2]);
  const long long base = u * stride[a1] + v * stride[a2];
  const long long sa = stride[a];
  const int n = N[a];
  if (plo > 0) {
    const uint16_t first = ids[base + plo * sa];
    for (int q = 0; q < plo; ++q) ids[base + q * sa] = first;
  }
  if (phi > 0) {
    const uint16_t outer = ids[base + (long long)(n - phi - 1) * sa];
    uint16_t src = outer;
    if (n - plo - phi >= 2) {
      const uint16_t inner = ids[base + (long long)(n - phi - 2) * sa];
      // (ids come from the fill, below n_mat)
      if (outer < n_mat && inner < n_mat && vacuum[outer] && !vacuum[inner]) src = inner;
      ids[base + (long long)(n - phi - 1) * sa] = src;
    }
    for (int q = n - phi; q < n; ++q) ids[base + q * sa] = src;
  }
}

// Tangential-zero bits of a PEC cell: bit a when a neighbour along a is
// PEC (none past the domain).
__device__ __forceinline__ unsigned asm_pec_bits(const uint8_t* __restrict__ pec, long long c,
                                                 int i, int j, int k, int nx, int ny, int nz) {
  const long long P = (long long)ny * nz;
  unsigned bits = 0;
  if ((i > 0 && pec[c - P]) || (i < nx - 1 && pec[c + P])) bits |= 1u;
  if ((j > 0 && pec[c - nz]) || (j < ny - 1 && pec[c + nz])) bits |= 2u;
  if ((k > 0 && pec[c - 1]) || (k < nz - 1 && pec[c + 1])) bits |= 4u;
  return bits;
}

template <bool APPLY>
__global__ __launch_bounds__(256) void k_asm_pec(int nx, int ny, int nz, uint16_t* __restrict__ ids,
                                                 const uint8_t* __restrict__ pec,
                                                 uint32_t* __restrict__ used,
                                                 const uint16_t* __restrict__ map, int n_mat) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long P = (long long)ny * nz;
  if (c >= (long long)nx * P || !pec[c]) return;
  const int i = (int)(c / P);
  const int r = (int)(c - (long long)i * P);
  const int j = r / nz, k = r - (r / nz) * nz;
  const unsigned bits = asm_pec_bits(pec, c, i, j, k, nx, ny, nz);
  if (!bits) return;
  const unsigned id = ids[c];
  if (id >= (unsigned)n_mat) return;   // not an id of the table given
  if (APPLY) ids[c] = map[id * 8u + bits];
  else atomicOr(&used[id], 1u << bits);
}

int asm_check(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1 || (long long)nx * ny * nz >= (1LL << 30))
    return fail(RFXH_EINVAL, "assembly: grid must have 1 .. 2^30 cells");
  return 0;
}

// host table -> device (synchronous: set-up path); *out freed by the caller
template <typename T>
int asm_upload(const T* host, size_t n, T** out) {
  *out = nullptr;
  if (n == 0) return 0;
  if (hipMalloc((void**)out, n * sizeof(T)) != hipSuccess)
    return fail(RFXH_ENOMEM, "assembly: device table");
  if (hipMemcpy(*out, host, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(*out);
    *out = nullptr;
    return fail(RFXH_EHIP, "assembly: table upload");
  }
  return 0;
}

}  // namespace

extern "C" {

int rfxh_assemble_fill(int32_t nx, int32_t ny, int32_t nz, const rfxh_asm_box* boxes,
                       int32_t n_box, const int32_t* pads, const uint8_t* vacuum, int32_t n_mat,
                       uint16_t* d_ids, uint8_t* d_pec, void* stream) {
  int rc = asm_check(nx, ny, nz);
  if (rc) return rc;
  if (!d_ids || !d_pec || n_box < 0 || (n_box > 0 && !boxes) || n_mat < 1 || n_mat > 65536 ||
      !vacuum)
    return fail(RFXH_EINVAL, "assembly: bad arguments");
  for (int b = 0; b < n_box; ++b)
    if (!boxes[b].pec && (boxes[b].id < 0 || boxes[b].id >= n_mat))
      return fail(RFXH_EINVAL, "assembly: box material id outside the table");
  const int N[3] = {nx, ny, nz};
  if (pads)
    for (int a = 0; a < 3; ++a)
      if (pads[2 * a] < 0 || pads[2 * a + 1] < 0 || pads[2 * a] + pads[2 * a + 1] >= N[a])
        return fail(RFXH_EINVAL, "assembly: pads leave no interior");
  hipStream_t s = static_cast<hipStream_t>(stream);
  rfxh_asm_box* dbox = nullptr;
  uint8_t* dvac = nullptr;
  if ((rc = asm_upload(boxes, (size_t)n_box, &dbox))) return rc;
  if ((rc = asm_upload(vacuum, (size_t)n_mat, &dvac))) {
    (void)hipFree(dbox);
    return rc;
  }
  const long long cells = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_asm_fill, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, dbox,
                     n_box, nx, ny, nz, d_ids, d_pec);
  if (pads) {
    for (int a = 0; a < 3; ++a) {
      if (pads[2 * a] == 0 && pads[2 * a + 1] == 0) continue;
      const long long lines = cells / N[a];
      hipLaunchKernelGGL(k_asm_pads, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, s, a,
                         pads[2 * a], pads[2 * a + 1], nx, ny, nz, dvac, n_mat, d_ids);
    }
  }
  rc = hipGetLastError() == hipSuccess ? 0 : fail(RFXH_EHIP, "assembly: launch");
  // the tables are freed once the kernels are done with them
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(RFXH_EHIP, "assembly: fill");
  (void)hipFree(dbox);
  (void)hipFree(dvac);
  return rc;
}

int rfxh_assemble_pec_pairs(int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_ids,
                            const uint8_t* d_pec, uint32_t* d_used, int32_t n_mat, void* stream) {
  int rc = asm_check(nx, ny, nz);
  if (rc) return rc;
  if (!d_ids || !d_pec || !d_used || n_mat < 1 || n_mat > 65536)
    return fail(RFXH_EINVAL, "assembly: bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long cells = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_asm_pec<false>, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, nx,
                     ny, nz, const_cast<uint16_t*>(d_ids), d_pec, d_used, nullptr, n_mat);
  HIPCHK(hipGetLastError());
  return RFXH_OK;
}

int rfxh_assemble_pec_apply(int32_t nx, int32_t ny, int32_t nz, uint16_t* d_ids,
                            const uint8_t* d_pec, const uint16_t* map, int32_t n_mat,
                            void* stream) {
  int rc = asm_check(nx, ny, nz);
  if (rc) return rc;
  if (!d_ids || !d_pec || !map || n_mat < 1 || n_mat > 65536)
    return fail(RFXH_EINVAL, "assembly: bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint16_t* dmap = nullptr;
  if ((rc = asm_upload(map, (size_t)n_mat * 8, &dmap))) return rc;
  const long long cells = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_asm_pec<true>, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, nx,
                     ny, nz, d_ids, d_pec, nullptr, dmap, n_mat);
  rc = hipGetLastError() == hipSuccess ? 0 : fail(RFXH_EHIP, "assembly: launch");
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(RFXH_EHIP, "assembly: pec apply");
  (void)hipFree(dmap);
  return rc;
}

const char* rfxh_last_error(void) { return g_err.c_str(); }

int rfxh_version(void) { return 2; }

}  // extern "C"
