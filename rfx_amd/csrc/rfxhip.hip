// rfxhip.hip — MI355X (gfx950) FDTD time-stepping kernels and their C ABI.
//
// One time step (rfx/simulation.py:1256-1681 uniform, rfx/nonuniform.py:
// 1256-1565 graded) is three kinds of launch on one stream:
//   k_update_h  H update (rfx/core/yee.py:225-291 / :411-451) with the CPML
//               H correction (rfx/boundaries/cpml.py:804-1065) fused in;
//   k_update_e  E update (yee.py:346-404 / :456-492) with the CPML E
//               correction (cpml.py:522-801), outer-face PEC (rfx/boundaries/
//               pec.py:13-48), the tangential PEC mask (pec.py:179-208) and
//               the per-cell point operations (port V/I sampling, source
//               injection, probes) fused in;
//   k_dft_plane one per DFT plane: float64 running sums (rfx/simulation.py:
//               1574-1588, rfx/nonuniform.py:1458-1476).
// Fields are float32 structure-of-arrays [nx][ny][nz] with z fastest; a
// 64-lane wavefront covers 64 consecutive z cells of one (i, j) row, so every
// field access of a wavefront is one contiguous 256-byte segment.  Material
// coefficients are read through a uint16 id per cell and a small table that
// stays cache-resident (2 bytes per cell per pass instead of 12-16).  The
// step index lives in device memory (advanced by k_update_h) so that one
// captured hipGraph of a chunk of steps can be replayed for every chunk.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
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

constexpr int kBX = 64;  // lanes along z (one wavefront)
constexpr int kTY = 8;   // rows along y per workgroup (one wavefront per row)
constexpr int kTimingSamples = 64;  // sampled steps per timed run

struct KParams {
  int nx, ny, nz;
  int graded;
  long long plane;  // ny * nz
  float inv_e0, inv_h0;          // uniform spacing
  const float* inv_e[3];         // graded spacings (device)
  const float* inv_h[3];
  const uint16_t* mat;
  const float4* coef;            // ca, cb, ce, ch
  const uint32_t* flags;
  float* f[6];
  int cn[6];                     // CPML slab depth per face
  float* psi_h[6][2];
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
  int ichunk;                    // planes per workgroup march
  int order;                     // cell kernels: 0 flat chunks, 1 column tiles
};

struct PlaneParams {
  const float* fld;
  double2* acc;
  const double2* tw;  // [max_steps][nf] twiddles exp(-j 2 pi f t) dt
  long long* step;
  long long max_steps;
  int nf;
  int axis, index;
  int n1, n2;
  int ny, nz;
};

__device__ __forceinline__ float ld(const float* __restrict__ a, long long c) { return a[c]; }

// x-face slab index for psi: [l][j][k]; y-face: [i][l][k]; z-face: [i][j][l]
__device__ __forceinline__ long long psi_idx(int face, int l, int i, int j, int k, int n,
                                             int ny, int nz) {
  int ax = face >> 1;
  if (ax == 0) return ((long long)l * ny + j) * nz + k;
  if (ax == 1) return ((long long)i * n + l) * nz + k;
  return ((long long)i * ny + j) * n + l;
}

// ---------------------------------------------------------------------------
// Tile order.  Each workgroup owns a 64 (z) x kTY (y) tile of columns and
// marches along x over a chunk [i0, i1) of planes, so the x neighbour of every
// field comes from registers (loaded once), the y neighbour from the tile row
// above/below through LDS, and the z neighbour from the adjacent lane.
// Workgroups are dispatched round-robin over the 8 XCDs; xcd_remap hands each
// XCD a contiguous range of logical tiles so the halo rows two neighbouring
// tiles share are fetched into the same L2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int per = nb / 8, rem = nb % 8;
  const int x = b % 8, s = b / 8;
  return x < rem ? x * (per + 1) + s : rem * (per + 1) + (x - rem) * per + s;
}

struct TileIdx {
  int i0, i1, j0, k0;
};

__device__ __forceinline__ TileIdx tile_of(const KParams& p) {
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int nkb = (p.nz + kBX - 1) / kBX, njb = (p.ny + kTY - 1) / kTY;
  const int kb = L % nkb;
  const int t = L / nkb;
  const int jb = t % njb;
  const int ic = t / njb;
  TileIdx r;
  r.i0 = ic * p.ichunk;
  r.i1 = min(p.nx, r.i0 + p.ichunk);
  r.j0 = jb * kTY;
  r.k0 = kb * kBX;
  return r;
}

__device__ __forceinline__ float pick3(int c, float a, float b, float d) {
  return c == 0 ? a : (c == 1 ? b : d);
}

// ---------------------------------------------------------------------------
// H pass: H(i) = H(i) - ch curl_fwd E (yee.py:272-289) + CPML-H (cpml.py:875-1043)
// ---------------------------------------------------------------------------
template <bool GRADED>
__global__ __launch_bounds__(kBX* kTY) void k_h_march(KParams p) {
  __shared__ float sEx[2][kTY + 1][kBX];
  __shared__ float sEz[2][kTY + 1][kBX];
  const TileIdx T = tile_of(p);
  if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    // advance the step counter; the E pass and DFT kernels of this step read it
    *p.step = *p.step + 1;
  }
  const int lane = threadIdx.x, ty = threadIdx.y;
  const int k = T.k0 + lane, j = T.j0 + ty;
  const bool in = (k < p.nz) && (j < p.ny);
  const long long P = p.plane;
  const float* __restrict__ Ex = p.f[0];
  const float* __restrict__ Ey = p.f[1];
  const float* __restrict__ Ez = p.f[2];
  float* __restrict__ Hx = p.f[3];
  float* __restrict__ Hy = p.f[4];
  float* __restrict__ Hz = p.f[5];
  long long c = ((long long)T.i0 * p.ny + (in ? j : 0)) * p.nz + (in ? k : 0);
  float ex0 = 0.f, ey0 = 0.f, ez0 = 0.f;
  if (in) {
    ex0 = Ex[c];
    ey0 = Ey[c];
    ez0 = Ez[c];
  }
  float iy = p.inv_h0, iz = p.inv_h0;
  if (GRADED && in) {
    iy = p.inv_h[1][j];
    iz = p.inv_h[2][k];
  }
  const bool j_edge = (ty == kTY - 1) && in && (j + 1 < p.ny);
  const bool k_edge = (lane == kBX - 1) && in && (k + 1 < p.nz);
  for (int i = T.i0; i < T.i1; ++i, c += P) {
    const int buf = (i - T.i0) & 1;
    float ex1 = 0.f, ey1 = 0.f, ez1 = 0.f;
    if (in && i + 1 < p.nx) {
      ex1 = Ex[c + P];
      ey1 = Ey[c + P];
      ez1 = Ez[c + P];
    }
    sEx[buf][ty][lane] = ex0;
    sEz[buf][ty][lane] = ez0;
    if (ty == kTY - 1) {
      float a = 0.f, b = 0.f;
      if (j_edge) {
        a = Ex[c + p.nz];
        b = Ez[c + p.nz];
      }
      sEx[buf][kTY][lane] = a;
      sEz[buf][kTY][lane] = b;
    }
    float ex_k = __shfl_down(ex0, 1), ey_k = __shfl_down(ey0, 1);
    if (lane == kBX - 1) {
      ex_k = k_edge ? Ex[c + 1] : 0.f;
      ey_k = k_edge ? Ey[c + 1] : 0.f;
    } else if (k + 1 >= p.nz) {
      ex_k = ey_k = 0.f;
    }
    __syncthreads();
    if (in) {
      const float ex_j = sEx[buf][ty + 1][lane];
      const float ez_j = sEz[buf][ty + 1][lane];
      const float ix = GRADED ? p.inv_h[0][i] : p.inv_h0;
      const float ch = p.coef[p.mat[c]].w;
      float hx = Hx[c], hy = Hy[c], hz = Hz[c];
      const float cx = (ez_j - ez0) * iy - (ey_k - ey0) * iz;
      const float cy = (ex_k - ex0) * iz - (ez1 - ez0) * ix;
      const float cz = (ey1 - ey0) * ix - (ex_j - ex0) * iy;
      hx = hx - ch * cx;
      hy = hy - ch * cy;
      hz = hz - ch * cz;
#pragma unroll
      for (int face = 0; face < 6; ++face) {
        const int n = p.cn[face];
        if (n == 0) continue;
        const int ax = face >> 1;
        const int pos = ax == 0 ? i : ax == 1 ? j : k;
        const int N = ax == 0 ? p.nx : ax == 1 ? p.ny : p.nz;
        const int l = pos - ((face & 1) ? N - n : 0);
        if (l < 0 || l >= n) continue;
        const long long q = psi_idx(face, l, i, j, k, n, p.ny, p.nz);
        const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
        const float inv = p.cinv[face];
        float g0, g1;
        if (ax == 0) {
          g0 = (ez1 - ez0) * inv;   // Hy += ...
          g1 = (ey1 - ey0) * inv;   // Hz -= ...
        } else if (ax == 1) {
          g0 = (ez_j - ez0) * inv;  // Hx -= ...
          g1 = (ex_j - ex0) * inv;  // Hz += ...
        } else {
          g0 = (ey_k - ey0) * inv;  // Hx += ...
          g1 = (ex_k - ex0) * inv;  // Hy -= ...
        }
        float* ps0 = p.psi_h[face][0];
        float* ps1 = p.psi_h[face][1];
        const float s0 = bb * ps0[q] + cc * g0;
        const float s1 = bb * ps1[q] + cc * g1;
        ps0[q] = s0;
        ps1[q] = s1;
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
      Hx[c] = hx;
      Hy[c] = hy;
      Hz[c] = hz;
    }
    ex0 = ex1;
    ey0 = ey1;
    ez0 = ez1;
  }
}

// ---------------------------------------------------------------------------
// E pass: E = ca E + cb curl_bwd H (yee.py:379-404), CPML-E (cpml.py:608-777),
// outer-face PEC (pec.py:24-46), PEC mask (pec.py:179-208), point operations
// ---------------------------------------------------------------------------
__device__ __forceinline__ void point_ops(const KParams& p, long long cell, float& e0, float& e1,
                                          float& e2, float hx, float hy, float hz, float hx_j,
                                          float hx_k, float hy_i, float hy_k, float hz_i,
                                          float hz_j) {
  // binary search the first op of this cell (ops sorted by cell, order)
  int lo = 0, hi = p.n_ops;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (p.ops[mid].cell < cell) lo = mid + 1;
    else hi = mid;
  }
  const long long n = *p.step;
  for (int q = lo; q < p.n_ops && p.ops[q].cell == cell; ++q) {
    const rfxh_point_op op = p.ops[q];
    if (op.kind == RFXH_OP_PORT) {
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
    } else {
      p.probe_out[n * p.n_probe + op.slot] =
          op.comp < 3 ? pick3(op.comp, e0, e1, e2) : pick3(op.comp - 3, hx, hy, hz);
    }
  }
}

// Cell-per-thread tile order: logical workgroup L covers plane i of column tile
// (jb, kb) with i fastest, and each XCD receives a contiguous run of logical
// ids, so the workgroups of planes i and i+1 of one column tile run back to
// back on the same XCD and the x neighbours one loads are still in its L2.
__device__ __forceinline__ bool cell_of(const KParams& p, int& i, int& j, int& k) {
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const long long pl = (long long)p.ny * p.nz;
  if (p.order == 2) {
    // bands: the (y, z) plane is split into 8 flat bands, one per XCD; the
    // XCD sweeps its band plane by plane (512-cell chunks fastest), so the
    // band's x neighbours stay in that XCD's L2 between consecutive planes
    const long long bw = (pl + 7) / 8;
    const int ncb = (int)((bw + kBX * kTY - 1) / (kBX * kTY));
    const int b = L / (p.nx * ncb);
    const int r = L % (p.nx * ncb);
    i = r / ncb;
    const long long f = (long long)b * bw + (long long)(r % ncb) * (kBX * kTY) +
                        threadIdx.y * kBX + threadIdx.x;
    if (f >= pl || f >= (long long)(b + 1) * bw) return false;
    j = (int)(f / p.nz);
    k = (int)(f % p.nz);
    return true;
  }
  if (p.order == 0) {
    // flat chunks: workgroup L covers 512 consecutive cells of the (y, z)
    // plane of plane i (chunk index fastest), so neighbouring workgroups
    // share only the cache line at their common boundary
    const int nch = (int)((pl + kBX * kTY - 1) / (kBX * kTY));
    const int ch = L % nch;
    i = L / nch;
    const long long f = (long long)ch * (kBX * kTY) + threadIdx.y * kBX + threadIdx.x;
    if (f >= pl) return false;
    j = (int)(f / p.nz);
    k = (int)(f % p.nz);
    return true;
  }
  const int nkb = (p.nz + kBX - 1) / kBX;
  i = L % p.nx;
  const int t = L / p.nx;
  const int kb = t % nkb;
  const int jb = t / nkb;
  k = kb * kBX + threadIdx.x;
  j = jb * kTY + threadIdx.y;
  return k < p.nz && j < p.ny;
}

// ---------------------------------------------------------------------------
// Cell-per-thread passes (one plane per workgroup)
// ---------------------------------------------------------------------------
template <bool GRADED>
__global__ __launch_bounds__(kBX* kTY) void k_update_h(KParams p) {
  int i, j, k;
  const bool ok = cell_of(p, i, j, k);
  if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    // advance the step counter; the E pass and DFT kernels of this step read it
    *p.step = *p.step + 1;
  }
  if (!ok) return;
  const long long P = p.plane;
  const long long c = ((long long)i * p.ny + j) * p.nz + k;
  const float* __restrict__ Ex = p.f[0];
  const float* __restrict__ Ey = p.f[1];
  const float* __restrict__ Ez = p.f[2];
  const float ex = ld(Ex, c), ey = ld(Ey, c), ez = ld(Ez, c);
  const bool hj = j + 1 < p.ny, hk = k + 1 < p.nz, hi = i + 1 < p.nx;
  const float ex_j = hj ? ld(Ex, c + p.nz) : 0.f;
  const float ez_j = hj ? ld(Ez, c + p.nz) : 0.f;
  const float ex_k = hk ? ld(Ex, c + 1) : 0.f;
  const float ey_k = hk ? ld(Ey, c + 1) : 0.f;
  const float ey_i = hi ? ld(Ey, c + P) : 0.f;
  const float ez_i = hi ? ld(Ez, c + P) : 0.f;
  float ix, iy, iz;
  if (GRADED) {
    ix = p.inv_h[0][i];
    iy = p.inv_h[1][j];
    iz = p.inv_h[2][k];
  } else {
    ix = iy = iz = p.inv_h0;
  }
  const float4 co = p.coef[p.mat[c]];
  const float ch = co.w;
  float hx = p.f[3][c], hy = p.f[4][c], hz = p.f[5][c];
  // yee.py:272-289: H -= (dt/mu) * curl E (forward differences)
  const float cx = (ez_j - ez) * iy - (ey_k - ey) * iz;
  const float cy = (ex_k - ex) * iz - (ez_i - ez) * ix;
  const float cz = (ey_i - ey) * ix - (ex_j - ex) * iy;
  hx = hx - ch * cx;
  hy = hy - ch * cy;
  hz = hz - ch * cz;
  // CPML-H (cpml.py:875-1043); ch equals dt/(mu_r mu0) of cpml.py:817
#pragma unroll
  for (int face = 0; face < 6; ++face) {
    const int n = p.cn[face];
    if (n == 0) continue;
    const int ax = face >> 1;
    const int pos = ax == 0 ? i : ax == 1 ? j : k;
    const int N = ax == 0 ? p.nx : ax == 1 ? p.ny : p.nz;
    const int lo = (face & 1) ? N - n : 0;
    const int l = pos - lo;
    if (l < 0 || l >= n) continue;
    const long long q = psi_idx(face, l, i, j, k, n, p.ny, p.nz);
    const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
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
    float* ps0 = p.psi_h[face][0];
    float* ps1 = p.psi_h[face][1];
    const float s0 = bb * ps0[q] + cc * g0;
    const float s1 = bb * ps1[q] + cc * g1;
    ps0[q] = s0;
    ps1[q] = s1;
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
  p.f[3][c] = hx;
  p.f[4][c] = hy;
  p.f[5][c] = hz;
}

// ---------------------------------------------------------------------------
// E pass
// ---------------------------------------------------------------------------
template <bool GRADED>
__global__ __launch_bounds__(kBX* kTY) void k_update_e(KParams p) {
  int i, j, k;
  if (!cell_of(p, i, j, k)) return;
  const long long P = p.plane;
  const long long c = ((long long)i * p.ny + j) * p.nz + k;
  const float* __restrict__ Hx = p.f[3];
  const float* __restrict__ Hy = p.f[4];
  const float* __restrict__ Hz = p.f[5];
  const float hx = ld(Hx, c), hy = ld(Hy, c), hz = ld(Hz, c);
  const float hx_j = j > 0 ? ld(Hx, c - p.nz) : 0.f;
  const float hz_j = j > 0 ? ld(Hz, c - p.nz) : 0.f;
  const float hx_k = k > 0 ? ld(Hx, c - 1) : 0.f;
  const float hy_k = k > 0 ? ld(Hy, c - 1) : 0.f;
  const float hy_i = i > 0 ? ld(Hy, c - P) : 0.f;
  const float hz_i = i > 0 ? ld(Hz, c - P) : 0.f;
  float ix, iy, iz;
  if (GRADED) {
    ix = p.inv_e[0][i];
    iy = p.inv_e[1][j];
    iz = p.inv_e[2][k];
  } else {
    ix = iy = iz = p.inv_e0;
  }
  const unsigned m = p.mat[c];
  const float4 co = p.coef[m];
  const float ca = co.x, cb = co.y, ce = co.z;
  float e0 = p.f[0][c], e1 = p.f[1][c], e2 = p.f[2][c];
  // yee.py:379-404: E = ca E + cb curl H (backward differences)
  const float cx = (hz - hz_j) * iy - (hy - hy_k) * iz;
  const float cy = (hx - hx_k) * iz - (hz - hz_i) * ix;
  const float cz = (hy - hy_i) * ix - (hx - hx_j) * iy;
  e0 = ca * e0 + cb * cx;
  e1 = ca * e1 + cb * cy;
  e2 = ca * e2 + cb * cz;
  // CPML-E (cpml.py:608-777) with ce = dt/(eps_r eps0) (cpml.py:538-546)
#pragma unroll
  for (int face = 0; face < 6; ++face) {
    const int n = p.cn[face];
    if (n == 0) continue;
    const int ax = face >> 1;
    const int pos = ax == 0 ? i : ax == 1 ? j : k;
    const int N = ax == 0 ? p.nx : ax == 1 ? p.ny : p.nz;
    const int lo = (face & 1) ? N - n : 0;
    const int l = pos - lo;
    if (l < 0 || l >= n) continue;
    const long long q = psi_idx(face, l, i, j, k, n, p.ny, p.nz);
    const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
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
    float* ps0 = p.psi_e[face][0];
    float* ps1 = p.psi_e[face][1];
    const float s0 = bb * ps0[q] + cc * g0;
    const float s1 = bb * ps1[q] + cc * g1;
    ps0[q] = s0;
    ps1[q] = s1;
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
  // outer-face PEC (pec.py:24-46)
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
  const uint32_t fl = p.flags[m];
  if (fl & RFXH_ZERO_EX) e0 = 0.f;
  if (fl & RFXH_ZERO_EY) e1 = 0.f;
  if (fl & RFXH_ZERO_EZ) e2 = 0.f;
  if (fl & RFXH_POINT) point_ops(p, c, e0, e1, e2, hx, hy, hz, hx_j, hx_k, hy_i, hy_k, hz_i, hz_j);
  p.f[0][c] = e0;
  p.f[1][c] = e1;
  p.f[2][c] = e2;
}

// ---------------------------------------------------------------------------
// Flat-chunk marching passes (default): workgroup = 512 consecutive cells of
// the (y, z) plane, marching over a chunk of x planes.  The x neighbour is
// carried in registers from the previous plane; y and z neighbours are plain
// loads of cells the neighbouring lanes / workgroups fetch in the same
// iteration (L1/L2 hits).  No LDS, no barriers: waves run independently.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool fm_cell(const KParams& p, int& i0, int& i1, int& j, int& k) {
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const long long pl = (long long)p.ny * p.nz;
  const int nch = (int)((pl + kBX * kTY - 1) / (kBX * kTY));
  const int ch = L % nch;
  const int ic = L / nch;
  i0 = ic * p.ichunk;
  i1 = min(p.nx, i0 + p.ichunk);
  const long long f = (long long)ch * (kBX * kTY) + threadIdx.y * kBX + threadIdx.x;
  if (f >= pl) return false;
  j = (int)(f / p.nz);
  k = (int)(f % p.nz);
  return true;
}

template <bool GRADED>
__global__ __launch_bounds__(kBX* kTY) void k_h_fm(KParams p) {
  int i0, i1, j, k;
  const bool ok = fm_cell(p, i0, i1, j, k);
  if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    // advance the step counter; the E pass and DFT kernels of this step read it
    *p.step = *p.step + 1;
  }
  if (!ok) return;
  const long long P = p.plane;
  const float* __restrict__ Ex = p.f[0];
  const float* __restrict__ Ey = p.f[1];
  const float* __restrict__ Ez = p.f[2];
  float* __restrict__ Hx = p.f[3];
  float* __restrict__ Hy = p.f[4];
  float* __restrict__ Hz = p.f[5];
  const bool hj = j + 1 < p.ny, hk = k + 1 < p.nz;
  float iy = p.inv_h0, iz = p.inv_h0;
  if (GRADED) {
    iy = p.inv_h[1][j];
    iz = p.inv_h[2][k];
  }
  long long c = ((long long)i0 * p.ny + j) * p.nz + k;
  float ex = Ex[c], ey = Ey[c], ez = Ez[c];
  for (int i = i0; i < i1; ++i, c += P) {
    const bool hi = i + 1 < p.nx;
    const float ex_j = hj ? Ex[c + p.nz] : 0.f;
    const float ez_j = hj ? Ez[c + p.nz] : 0.f;
    const float ex_k = hk ? Ex[c + 1] : 0.f;
    const float ey_k = hk ? Ey[c + 1] : 0.f;
    const float ex_i = hi ? Ex[c + P] : 0.f;
    const float ey_i = hi ? Ey[c + P] : 0.f;
    const float ez_i = hi ? Ez[c + P] : 0.f;
    const float ix = GRADED ? p.inv_h[0][i] : p.inv_h0;
    const float ch = p.coef[p.mat[c]].w;
    float hx = Hx[c], hy = Hy[c], hz = Hz[c];
    const float cx = (ez_j - ez) * iy - (ey_k - ey) * iz;
    const float cy = (ex_k - ex) * iz - (ez_i - ez) * ix;
    const float cz = (ey_i - ey) * ix - (ex_j - ex) * iy;
    hx = hx - ch * cx;
    hy = hy - ch * cy;
    hz = hz - ch * cz;
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int n = p.cn[face];
      if (n == 0) continue;
      const int ax = face >> 1;
      const int pos = ax == 0 ? i : ax == 1 ? j : k;
      const int N = ax == 0 ? p.nx : ax == 1 ? p.ny : p.nz;
      const int l = pos - ((face & 1) ? N - n : 0);
      if (l < 0 || l >= n) continue;
      const long long q = psi_idx(face, l, i, j, k, n, p.ny, p.nz);
      const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
      const float inv = p.cinv[face];
      float g0, g1;
      if (ax == 0) {
        g0 = (ez_i - ez) * inv;   // Hy += ...
        g1 = (ey_i - ey) * inv;   // Hz -= ...
      } else if (ax == 1) {
        g0 = (ez_j - ez) * inv;   // Hx -= ...
        g1 = (ex_j - ex) * inv;   // Hz += ...
      } else {
        g0 = (ey_k - ey) * inv;   // Hx += ...
        g1 = (ex_k - ex) * inv;   // Hy -= ...
      }
      float* ps0 = p.psi_h[face][0];
      float* ps1 = p.psi_h[face][1];
      const float s0 = bb * ps0[q] + cc * g0;
      const float s1 = bb * ps1[q] + cc * g1;
      ps0[q] = s0;
      ps1[q] = s1;
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
    Hx[c] = hx;
    Hy[c] = hy;
    Hz[c] = hz;
    ex = ex_i;
    ey = ey_i;
    ez = ez_i;
  }
}

template <bool GRADED>
__global__ __launch_bounds__(kBX* kTY) void k_e_fm(KParams p) {
  int i0, i1, j, k;
  if (!fm_cell(p, i0, i1, j, k)) return;
  const long long P = p.plane;
  float* __restrict__ Ex = p.f[0];
  float* __restrict__ Ey = p.f[1];
  float* __restrict__ Ez = p.f[2];
  const float* __restrict__ Hx = p.f[3];
  const float* __restrict__ Hy = p.f[4];
  const float* __restrict__ Hz = p.f[5];
  const bool hj = j > 0, hk = k > 0;
  float iy = p.inv_e0, iz = p.inv_e0;
  if (GRADED) {
    iy = p.inv_e[1][j];
    iz = p.inv_e[2][k];
  }
  long long c = ((long long)i0 * p.ny + j) * p.nz + k;
  float hy_i = 0.f, hz_i = 0.f;  // H at plane i-1
  if (i0 > 0) {
    hy_i = Hy[c - P];
    hz_i = Hz[c - P];
  }
  for (int i = i0; i < i1; ++i, c += P) {
    const float hx = Hx[c], hy = Hy[c], hz = Hz[c];
    const float hx_j = hj ? Hx[c - p.nz] : 0.f;
    const float hz_j = hj ? Hz[c - p.nz] : 0.f;
    const float hx_k = hk ? Hx[c - 1] : 0.f;
    const float hy_k = hk ? Hy[c - 1] : 0.f;
    const float ix = GRADED ? p.inv_e[0][i] : p.inv_e0;
    const unsigned m = p.mat[c];
    const float4 co = p.coef[m];
    const float ca = co.x, cb = co.y, ce = co.z;
    float e0 = Ex[c], e1 = Ey[c], e2 = Ez[c];
    const float cx = (hz - hz_j) * iy - (hy - hy_k) * iz;
    const float cy = (hx - hx_k) * iz - (hz - hz_i) * ix;
    const float cz = (hy - hy_i) * ix - (hx - hx_j) * iy;
    e0 = ca * e0 + cb * cx;
    e1 = ca * e1 + cb * cy;
    e2 = ca * e2 + cb * cz;
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      const int n = p.cn[face];
      if (n == 0) continue;
      const int ax = face >> 1;
      const int pos = ax == 0 ? i : ax == 1 ? j : k;
      const int N = ax == 0 ? p.nx : ax == 1 ? p.ny : p.nz;
      const int l = pos - ((face & 1) ? N - n : 0);
      if (l < 0 || l >= n) continue;
      const long long q = psi_idx(face, l, i, j, k, n, p.ny, p.nz);
      const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
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
      float* ps0 = p.psi_e[face][0];
      float* ps1 = p.psi_e[face][1];
      const float s0 = bb * ps0[q] + cc * g0;
      const float s1 = bb * ps1[q] + cc * g1;
      ps0[q] = s0;
      ps1[q] = s1;
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
    const uint32_t fl = p.flags[m];
    if (fl & RFXH_ZERO_EX) e0 = 0.f;
    if (fl & RFXH_ZERO_EY) e1 = 0.f;
    if (fl & RFXH_ZERO_EZ) e2 = 0.f;
    if (fl & RFXH_POINT) point_ops(p, c, e0, e1, e2, hx, hy, hz, hx_j, hx_k, hy_i, hy_k, hz_i, hz_j);
    Ex[c] = e0;
    Ey[c] = e1;
    Ez[c] = e2;
    hy_i = hy;
    hz_i = hz;
  }
}

template <bool GRADED>
__global__ __launch_bounds__(kBX* kTY) void k_e_march(KParams p) {
  __shared__ float sHx[2][kTY + 1][kBX];
  __shared__ float sHz[2][kTY + 1][kBX];
  const TileIdx T = tile_of(p);
  const int lane = threadIdx.x, ty = threadIdx.y;
  const int k = T.k0 + lane, j = T.j0 + ty;
  const bool in = (k < p.nz) && (j < p.ny);
  const long long P = p.plane;
  float* __restrict__ Ex = p.f[0];
  float* __restrict__ Ey = p.f[1];
  float* __restrict__ Ez = p.f[2];
  const float* __restrict__ Hx = p.f[3];
  const float* __restrict__ Hy = p.f[4];
  const float* __restrict__ Hz = p.f[5];
  long long c = ((long long)T.i0 * p.ny + (in ? j : 0)) * p.nz + (in ? k : 0);
  float hy_p = 0.f, hz_p = 0.f;  // H at plane i-1
  if (in && T.i0 > 0) {
    hy_p = Hy[c - P];
    hz_p = Hz[c - P];
  }
  float iy = p.inv_e0, iz = p.inv_e0;
  if (GRADED && in) {
    iy = p.inv_e[1][j];
    iz = p.inv_e[2][k];
  }
  const bool j_edge = (ty == 0) && in && (j > 0);
  const bool k_edge = (lane == 0) && in && (k > 0);
  for (int i = T.i0; i < T.i1; ++i, c += P) {
    const int buf = (i - T.i0) & 1;
    float hx = 0.f, hy = 0.f, hz = 0.f;
    if (in) {
      hx = Hx[c];
      hy = Hy[c];
      hz = Hz[c];
    }
    sHx[buf][ty + 1][lane] = hx;
    sHz[buf][ty + 1][lane] = hz;
    if (ty == 0) {
      float a = 0.f, b = 0.f;
      if (j_edge) {
        a = Hx[c - p.nz];
        b = Hz[c - p.nz];
      }
      sHx[buf][0][lane] = a;
      sHz[buf][0][lane] = b;
    }
    float hx_k = __shfl_up(hx, 1), hy_k = __shfl_up(hy, 1);
    if (lane == 0) {
      hx_k = k_edge ? Hx[c - 1] : 0.f;
      hy_k = k_edge ? Hy[c - 1] : 0.f;
    }
    __syncthreads();
    if (in) {
      const float hx_j = sHx[buf][ty][lane];
      const float hz_j = sHz[buf][ty][lane];
      const float ix = GRADED ? p.inv_e[0][i] : p.inv_e0;
      const unsigned m = p.mat[c];
      const float4 co = p.coef[m];
      const float ca = co.x, cb = co.y, ce = co.z;
      float e0 = Ex[c], e1 = Ey[c], e2 = Ez[c];
      const float cx = (hz - hz_j) * iy - (hy - hy_k) * iz;
      const float cy = (hx - hx_k) * iz - (hz - hz_p) * ix;
      const float cz = (hy - hy_p) * ix - (hx - hx_j) * iy;
      e0 = ca * e0 + cb * cx;
      e1 = ca * e1 + cb * cy;
      e2 = ca * e2 + cb * cz;
#pragma unroll
      for (int face = 0; face < 6; ++face) {
        const int n = p.cn[face];
        if (n == 0) continue;
        const int ax = face >> 1;
        const int pos = ax == 0 ? i : ax == 1 ? j : k;
        const int N = ax == 0 ? p.nx : ax == 1 ? p.ny : p.nz;
        const int l = pos - ((face & 1) ? N - n : 0);
        if (l < 0 || l >= n) continue;
        const long long q = psi_idx(face, l, i, j, k, n, p.ny, p.nz);
        const float bb = p.pb[face][l], cc = p.pc[face][l], km1 = p.pk[face][l];
        const float inv = p.cinv[face];
        float g0, g1;
        if (ax == 0) {
          g0 = (hz - hz_p) * inv;  // Ey -= ...
          g1 = (hy - hy_p) * inv;  // Ez += ...
        } else if (ax == 1) {
          g0 = (hz - hz_j) * inv;  // Ex += ...
          g1 = (hx - hx_j) * inv;  // Ez -= ...
        } else {
          g0 = (hy - hy_k) * inv;  // Ex -= ...
          g1 = (hx - hx_k) * inv;  // Ey += ...
        }
        float* ps0 = p.psi_e[face][0];
        float* ps1 = p.psi_e[face][1];
        const float s0 = bb * ps0[q] + cc * g0;
        const float s1 = bb * ps1[q] + cc * g1;
        ps0[q] = s0;
        ps1[q] = s1;
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
      const uint32_t fl = p.flags[m];
      if (fl & RFXH_ZERO_EX) e0 = 0.f;
      if (fl & RFXH_ZERO_EY) e1 = 0.f;
      if (fl & RFXH_ZERO_EZ) e2 = 0.f;
      if (fl & RFXH_POINT) point_ops(p, c, e0, e1, e2, hx, hy, hz, hx_j, hx_k, hy_p, hy_k, hz_p, hz_j);
      Ex[c] = e0;
      Ey[c] = e1;
      Ez[c] = e2;
    }
    hy_p = hy;
    hz_p = hz;
  }
}

// ---------------------------------------------------------------------------
// DFT plane: acc[f][u][v] += tw[n][f] * F(plane)  (float64 sums)
// ---------------------------------------------------------------------------
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
  const double x = (double)q.fld[((long long)ii[0] * q.ny + ii[1]) * q.nz + ii[2]];
  const double2* tw = q.tw + n * q.nf;
  for (int f = 0; f < q.nf; ++f) {
    const double2 w = tw[f];
    double2* a = q.acc + (long long)f * npts + pidx;
    double2 cur = *a;
    cur.x += w.x * x;
    cur.y += w.y * x;
    *a = cur;
  }
}

__global__ void k_set_step(long long* step, long long value) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *step = value;
}

}  // namespace

struct rfxh_plan {
  KParams kp;
  dim3 grid, block;
  int kind = 0;  // 0 flat-chunk march (default), 1 LDS tile march, 2 one plane per workgroup
  std::vector<PlaneParams> planes;
  std::vector<void*> owned;  // device allocations owned by the plan
  long long max_steps = 0;
  hipStream_t cap_stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_chunk = 0;
  int timing = 0;
  double t_sum[3] = {0, 0, 0};
  long long t_cnt[3] = {0, 0, 0};
  std::vector<hipEvent_t> tev;
};

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

int launch_step(rfxh_plan* pl, hipStream_t s, hipEvent_t* ev) {
  // ev (optional): four events bracketing the H pass, the E pass and the DFT
  // kernels of this step, recorded on the stream the kernels run on.
  if (ev) HIPCHK(hipEventRecord(ev[0], s));
  if (pl->kind == 0) {
    if (pl->kp.graded)
      hipLaunchKernelGGL(k_h_fm<true>, pl->grid, pl->block, 0, s, pl->kp);
    else
      hipLaunchKernelGGL(k_h_fm<false>, pl->grid, pl->block, 0, s, pl->kp);
  } else if (pl->kind == 1) {
    if (pl->kp.graded)
      hipLaunchKernelGGL(k_h_march<true>, pl->grid, pl->block, 0, s, pl->kp);
    else
      hipLaunchKernelGGL(k_h_march<false>, pl->grid, pl->block, 0, s, pl->kp);
  } else {
    if (pl->kp.graded)
      hipLaunchKernelGGL(k_update_h<true>, pl->grid, pl->block, 0, s, pl->kp);
    else
      hipLaunchKernelGGL(k_update_h<false>, pl->grid, pl->block, 0, s, pl->kp);
  }
  if (ev) HIPCHK(hipEventRecord(ev[1], s));
  if (pl->kind == 0) {
    if (pl->kp.graded)
      hipLaunchKernelGGL(k_e_fm<true>, pl->grid, pl->block, 0, s, pl->kp);
    else
      hipLaunchKernelGGL(k_e_fm<false>, pl->grid, pl->block, 0, s, pl->kp);
  } else if (pl->kind == 1) {
    if (pl->kp.graded)
      hipLaunchKernelGGL(k_e_march<true>, pl->grid, pl->block, 0, s, pl->kp);
    else
      hipLaunchKernelGGL(k_e_march<false>, pl->grid, pl->block, 0, s, pl->kp);
  } else {
    if (pl->kp.graded)
      hipLaunchKernelGGL(k_update_e<true>, pl->grid, pl->block, 0, s, pl->kp);
    else
      hipLaunchKernelGGL(k_update_e<false>, pl->grid, pl->block, 0, s, pl->kp);
  }
  if (ev) HIPCHK(hipEventRecord(ev[2], s));
  for (const PlaneParams& q : pl->planes) {
    const long long npts = (long long)q.n1 * q.n2;
    const unsigned nb = (unsigned)((npts + 255) / 256);
    hipLaunchKernelGGL(k_dft_plane, dim3(nb), dim3(256), 0, s, q);
  }
  if (ev) HIPCHK(hipEventRecord(ev[3], s));
  HIPCHK(hipGetLastError());
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

}  // namespace

extern "C" {

int64_t rfxh_psi_floats(const rfxh_desc* d) {
  if (!d) return -1;
  long long tot = 0;
  for (int f = 0; f < 6; ++f) tot += 4 * face_slab_floats(d, f);
  return tot;
}

int rfxh_plan_create(const rfxh_desc* d, rfxh_plan** out) {
  if (!d || !out) return fail(RFXH_EINVAL, "null descriptor or output");
  *out = nullptr;
  if (d->nx < 2 || d->ny < 2 || d->nz < 2) return fail(RFXH_EINVAL, "grid must be at least 2 cells per axis");
  if (d->n_mat < 1 || d->n_mat > 65536 || !d->mat_coef || !d->mat_flags || !d->d_mat)
    return fail(RFXH_EINVAL, "material table missing or larger than 65536 entries");
  for (int c = 0; c < 6; ++c)
    if (!d->d_f[c]) return fail(RFXH_EINVAL, "field pointer missing");
  const int N[3] = {d->nx, d->ny, d->nz};
  for (int f = 0; f < 6; ++f) {
    if (d->cpml_n[f] < 0 || d->cpml_n[f] > N[f >> 1]) return fail(RFXH_EINVAL, "CPML depth outside the axis");
    if (d->cpml_n[f] > 0 && (!d->cpml_b[f] || !d->cpml_c[f] || !d->cpml_km1[f] || !d->d_psi))
      return fail(RFXH_EINVAL, "CPML profile or psi buffer missing");
  }
  if (d->max_steps < 0) return fail(RFXH_EINVAL, "max_steps < 0");
  if ((d->n_src > 0 && !d->d_src_wave) || (d->n_probe > 0 && !d->d_probe_out) ||
      (d->n_port > 0 && !d->d_port_out))
    return fail(RFXH_EINVAL, "point-operation table missing");
  const long long ncell = (long long)d->nx * d->ny * d->nz;
  std::vector<rfxh_point_op> ops(d->ops, d->ops + d->n_ops);
  for (const rfxh_point_op& op : ops) {
    if (op.cell < 0 || op.cell >= ncell) return fail(RFXH_EINVAL, "point operation outside the grid");
    if (op.comp < 0 || op.comp > 5) return fail(RFXH_EINVAL, "bad point-operation component");
    int lim = op.kind == RFXH_OP_PORT ? d->n_port : op.kind == RFXH_OP_INJECT ? d->n_src : d->n_probe;
    if (op.kind < 0 || op.kind > 2 || op.slot < 0 || op.slot >= lim)
      return fail(RFXH_EINVAL, "bad point-operation slot/kind");
    if (op.kind != RFXH_OP_PROBE && op.comp > 2) return fail(RFXH_EINVAL, "ports/sources drive E only");
  }
  std::stable_sort(ops.begin(), ops.end(), [](const rfxh_point_op& a, const rfxh_point_op& b) {
    return a.cell != b.cell ? a.cell < b.cell : a.order < b.order;
  });

  rfxh_plan* pl = new rfxh_plan();
  KParams& kp = pl->kp;
  std::memset(&kp, 0, sizeof(kp));
  kp.nx = d->nx;
  kp.ny = d->ny;
  kp.nz = d->nz;
  kp.graded = d->graded ? 1 : 0;
  kp.plane = (long long)d->ny * d->nz;
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
    uint32_t* fl;
    if ((rc = dev_copy(pl, d->mat_flags, (size_t)d->n_mat, &fl))) return bail(rc);
    kp.flags = fl;
  }
  for (int c = 0; c < 6; ++c) kp.f[c] = d->d_f[c];
  long long off = 0;
  for (int f = 0; f < 6; ++f) {
    kp.cn[f] = d->cpml_n[f];
    kp.cinv[f] = d->cpml_inv_d[f];
    const long long s = face_slab_floats(d, f);
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
  {
    long long init = -1;
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
    pp.max_steps = d->max_steps;
    pp.nf = pd.n_freq;
    pp.axis = pd.axis;
    pp.index = pd.index;
    pp.n1 = N[pd.axis == 0 ? 1 : 0];
    pp.n2 = N[pd.axis == 2 ? 1 : 2];
    pp.ny = d->ny;
    pp.nz = d->nz;
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
  }
  // kernel design (RFXH_KERNEL = fm | march | cell; fm is the default) and
  // planes per march (RFXH_ICHUNK, for tuning sweeps)
  const char* kern = std::getenv("RFXH_KERNEL");
  pl->kind = 2;
  if (kern && std::string(kern) == "march") pl->kind = 1;
  if (kern && std::string(kern) == "fm") pl->kind = 0;
  const char* ord = std::getenv("RFXH_ORDER");
  kp.order = 2;
  if (ord && std::string(ord) == "flat") kp.order = 0;
  if (ord && std::string(ord) == "tile") kp.order = 1;
  const long long flat = (long long)d->ny * d->nz;
  const long long nch = (flat + kBX * kTY - 1) / (kBX * kTY);
  const long long nkb = (d->nz + kBX - 1) / kBX, njb = (d->ny + kTY - 1) / kTY;
  const long long tiles = pl->kind == 1 ? nkb * njb : nch;
  // default march length: enough workgroups to fill 256 CUs several times
  long long ichunk = pl->kind == 2 ? 1 : std::max<long long>(1, d->nx * tiles / 8192);
  if (const char* env = std::getenv("RFXH_ICHUNK")) {
    const long long v = std::atoll(env);
    if (v >= 1) ichunk = v;
  }
  ichunk = std::min<long long>(ichunk, d->nx);
  const long long nic = (d->nx + ichunk - 1) / ichunk;
  long long nblocks = tiles * nic;
  if (pl->kind == 2) {
    const long long ncb = ((flat + 7) / 8 + kBX * kTY - 1) / (kBX * kTY);
    nblocks = kp.order == 0   ? (long long)d->nx * nch
              : kp.order == 1 ? (long long)d->nx * nkb * njb
                              : 8LL * d->nx * ncb;
  }
  if (nblocks > 0x7fffffffLL) return bail(fail(RFXH_EINVAL, "grid too large"));
  kp.ichunk = (int)ichunk;
  pl->grid = dim3((unsigned)nblocks);
  pl->block = dim3(kBX, kTY);
  pl->tev.resize(4 * kTimingSamples);
  for (auto& e : pl->tev) {
    if (hipEventCreate(&e) != hipSuccess) return bail(fail(RFXH_EHIP, "hipEventCreate"));
  }
  *out = pl;
  return RFXH_OK;
}

int rfxh_plan_destroy(rfxh_plan* pl) {
  if (!pl) return RFXH_OK;
  if (pl->graph_exec) (void)hipGraphExecDestroy(pl->graph_exec);
  if (pl->cap_stream) (void)hipStreamDestroy(pl->cap_stream);
  if (pl->ev_in) (void)hipEventDestroy(pl->ev_in);
  if (pl->ev_out) (void)hipEventDestroy(pl->ev_out);
  for (auto e : pl->tev)
    if (e) (void)hipEventDestroy(e);
  for (void* p : pl->owned) (void)hipFree(p);
  delete pl;
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

int rfxh_run(rfxh_plan* pl, int64_t step0, int64_t n_steps, void* stream, int use_graph) {
  if (!pl) return fail(RFXH_EINVAL, "null plan");
  if (step0 < 0 || n_steps < 0 || step0 + n_steps > pl->max_steps)
    return fail(RFXH_EINVAL, "steps outside the source/probe tables (max_steps)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, pl->kp.step, (long long)(step0 - 1));
  HIPCHK(hipGetLastError());
  if (n_steps == 0) return RFXH_OK;
  const int chunk = 16;
  if (!use_graph || pl->timing || n_steps < chunk) {
    // timing: sample up to kTimingSamples evenly spaced steps; the host does
    // not wait inside the loop, so the queue stays full and each event pair
    // brackets only its kernel
    const int64_t every = std::max<int64_t>(1, n_steps / kTimingSamples);
    int n_samples = 0;
    for (int64_t n = 0; n < n_steps; ++n) {
      hipEvent_t* ev = nullptr;
      if (pl->timing && n % every == every - 1 && n_samples < kTimingSamples)
        ev = pl->tev.data() + 4 * (n_samples++);
      int rc = launch_step(pl, s, ev);
      if (rc) return rc;
    }
    if (n_samples) return collect_timing(pl, n_samples);
    return RFXH_OK;
  }
  if (!pl->cap_stream) {
    HIPCHK(hipStreamCreateWithFlags(&pl->cap_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&pl->ev_in, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&pl->ev_out, hipEventDisableTiming));
  }
  if (!pl->graph_exec) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(pl->cap_stream, hipStreamCaptureModeThreadLocal));
    for (int n = 0; n < chunk; ++n) {
      int rc = launch_step(pl, pl->cap_stream, nullptr);
      if (rc) {
        hipStreamEndCapture(pl->cap_stream, &g);
        return rc;
      }
    }
    HIPCHK(hipStreamEndCapture(pl->cap_stream, &g));
    HIPCHK(hipGraphInstantiate(&pl->graph_exec, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    pl->graph_chunk = chunk;
  }
  HIPCHK(hipEventRecord(pl->ev_in, s));
  HIPCHK(hipStreamWaitEvent(pl->cap_stream, pl->ev_in, 0));
  int64_t done = 0;
  while (n_steps - done >= pl->graph_chunk) {
    HIPCHK(hipGraphLaunch(pl->graph_exec, pl->cap_stream));
    done += pl->graph_chunk;
  }
  for (; done < n_steps; ++done) {
    int rc = launch_step(pl, pl->cap_stream, nullptr);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(pl->ev_out, pl->cap_stream));
  HIPCHK(hipStreamWaitEvent(s, pl->ev_out, 0));
  return RFXH_OK;
}

const char* rfxh_last_error(void) { return g_err.c_str(); }

int rfxh_version(void) { return 1; }

}  // extern "C"
