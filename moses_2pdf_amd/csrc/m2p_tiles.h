// m2p_tiles.h -- per-step device bodies of the adNOC_2Pb engine, shared by the
// single-scenario kernels (m2p_kernels.hip) and the batched-ensemble kernels
// (m2p_batch.hip).  Each body works on one scenario's buffers and one unit of work (a
// ghost line, a block of CFL faces, a stage tile); the kernels around them decide which
// scenario and which unit a block takes.  Templated on the state's scalar type (double:
// the plain path; float: the FP32-bulk variant).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "m2p_kernels.h"
#include "m2p_numerics.h"

namespace m2p {

// Whether a value is anything but +0, bit for bit (-0 and NaN count).
__device__ __forceinline__ bool nonzero_bits(double v) { return __double_as_longlong(v) != 0; }
__device__ __forceinline__ bool nonzero_bits(float v) { return __float_as_uint(v) != 0u; }

// Inflow1Kernel + Inflow2Kernel (I/K:818-887).  In the reference every thread of the
// grid repeats the whole loop; one thread does it here, in the same order (FP64
// arithmetic, stored in the state's type).
template <typename T>
__device__ __forceinline__ void inflow_body(const Grid& g, T* __restrict__ U, const double* __restrict__ M,
                                            const double* __restrict__ table, const double* __restrict__ loc,
                                            const int* __restrict__ dire, int nloc, int nsamp,
                                            double phiS0, int ic, unsigned char* __restrict__ LIM,
                                            const DryMap& dm) {
  const long long P = g.P;
  for (int i = 0; i < nloc; i++) {
    const int direX = dire[i * 3 + 1], direY = dire[i * 3 + 2];
    const int xloc = (int)loc[i * 3 + 1], yloc = (int)loc[i * 3 + 2];
    const double Hin = table[0 * (long long)nsamp * nloc + (long long)ic * nloc + i];
    for (int mm = 0; mm < direX + MD + 3; mm++)
      for (int nn = 0; nn < direY + MD + 3; nn++) {
        const long long c = idx(g, yloc + nn, xloc + mm);
        U[0 * P + c] = (T)(Hin * (1 - phiS0) / phiS0);
        U[3 * P + c] = (T)(Hin * (1.0 - phiS0) * M[MC_C * P + c]);
      }
  }
  for (int i = 0; i < nloc; i++) {
    const int direX = dire[i * 3 + 1], direY = dire[i * 3 + 2];
    const int xloc = (int)loc[i * 3 + 1], yloc = (int)loc[i * 3 + 2];
    const double Uin = table[1 * (long long)nsamp * nloc + (long long)ic * nloc + i];
    const double Vin = table[2 * (long long)nsamp * nloc + (long long)ic * nloc + i];
    for (int mm = 0; mm < direX + MD + 3; mm++)
      for (int nn = 0; nn < direY + MD + 3; nn++) {
        const long long c = idx(g, yloc + nn, xloc + mm);
        const double u0 = (double)U[0 * P + c], u3 = (double)U[3 * P + c];
        U[1 * P + c] = (T)(Uin * (u0 + H0));
        U[2 * P + c] = (T)(Vin * (u0 + H0));
        U[4 * P + c] = (T)(Uin * (u3 + H0));
        U[5 * P + c] = (T)(Vin * (u3 + H0));
        U[6 * P + c] = (T)(u0 / (u0 + u3 + H0));
        if (LIM) LIM[c] = 0;
        const int r = yloc + nn, cc = xloc + mm;
        if (dm.nseg > 0 && r >= MD && r < g.ny + MD && cc >= MD && cc < g.nx + MD)
          dm.z0[(long long)(r - MD) * dm.nseg + (cc - MD) / 63] = 1;
      }
  }
}

// Boundary5..8Kernel (L/K:853-956; inflow I/K:936-997) as one launch.  Thread (k, line):
// lines [0, nyd) fill the x-ghosts of a row, lines [nyd, nyd+nxd) the y-ghosts of a
// column.  Momenta (k = 1,2,4,5) of every ghost cell written are 0 (Boundary7/8).
//   landslide: x-ghosts on every row; on the ghost rows the sources (columns 3,4 and
//     nx+1,nx+2) are the previous stage's ghost-row values, held in `prev` (Boundary6
//     skips the corner columns, so the reference's corners lag one stage).  y-ghosts on
//     columns [3, nx+2] only.
//   inflow: x-ghosts on rows [3, ny+2]; y-ghosts on every column, the corner columns from
//     the x-ghost values of rows 3,4 / ny+1,ny+2 (Boundary6 covers every column).
template <typename T>
__device__ __forceinline__ void extrap3(T a, T b, T& g2, T& g1, T& g0) {
  g2 = T(0.5) * (a + b);   // u[2] = (u[3]+u[4])/2
  g1 = T(0.5) * (g2 + a);  // u[1] = (u[2]+u[3])/2
  g0 = T(0.5) * (g1 + g2); // u[0] = (u[1]+u[2])/2
}

// Returns whether it wrote a value other than +0 (bit for bit).
template <typename T>
__device__ __forceinline__ bool ghost_line(const Grid& g, int mode, T* __restrict__ Tg,
                                           const T* __restrict__ prev, int line, int k) {
  const long long P = g.P;
  T* A = Tg + k * P;
  const bool mom = (k == 1 || k == 2 || k == 4 || k == 5);
  const int nx = g.nx, ny = g.ny;
  if (line < g.nyd) {
    const int r = line;
    const bool ghost_row = (r < MD || r >= ny + MD);
    if (ghost_row && mode == 1) return true;
    const T* src = (ghost_row ? prev + k * P : A) + (long long)r * g.nxd;
    T* row = A + (long long)r * g.nxd;
    if (mom) {
      row[0] = row[1] = row[2] = T(0);
      row[nx + MD] = row[nx + MD + 1] = row[nx + MD + 2] = T(0);
      return false;
    }
    T g2, g1, g0, h3, h4, h5;
    extrap3(src[MD], src[MD + 1], g2, g1, g0);
    extrap3(src[nx + MD - 1], src[nx + MD - 2], h3, h4, h5);
    row[2] = g2;
    row[1] = g1;
    row[0] = g0;
    row[nx + MD] = h3;
    row[nx + MD + 1] = h4;
    row[nx + MD + 2] = h5;
    return nonzero_bits(g0) || nonzero_bits(g1) || nonzero_bits(g2) || nonzero_bits(h3) ||
           nonzero_bits(h4) || nonzero_bits(h5);
  } else {
    const int c = line - g.nyd;
    const bool ghost_col = (c < MD || c >= nx + MD);
    if (ghost_col && mode == 0) return false;
    const long long n = g.nxd;
    T* col = A + c;
    if (mom) {
      col[0] = col[n] = col[2 * n] = T(0);
      col[(ny + MD) * n] = col[(ny + MD + 1) * n] = col[(ny + MD + 2) * n] = T(0);
      return false;
    }
    T a3, a4, b2, b1;
    if (!ghost_col) {
      a3 = col[MD * n];
      a4 = col[(MD + 1) * n];
      b2 = col[(ny + MD - 1) * n];
      b1 = col[(ny + MD - 2) * n];
    } else {
      // x-ghost value of column c on rows 3, 4, ny+2, ny+1 (what Boundary5 just wrote)
      T v[4];
      const int rows[4] = {MD, MD + 1, ny + MD - 1, ny + MD - 2};
      for (int q = 0; q < 4; q++) {
        const T* rr = A + (long long)rows[q] * n;
        T e2, e1, e0;
        if (c < MD) {
          extrap3(rr[MD], rr[MD + 1], e2, e1, e0);
          v[q] = (c == 2) ? e2 : (c == 1) ? e1 : e0;
        } else {
          extrap3(rr[nx + MD - 1], rr[nx + MD - 2], e2, e1, e0);
          v[q] = (c == nx + MD) ? e2 : (c == nx + MD + 1) ? e1 : e0;
        }
      }
      a3 = v[0];
      a4 = v[1];
      b2 = v[2];
      b1 = v[3];
    }
    T g2, g1, g0, h3, h4, h5;
    extrap3(a3, a4, g2, g1, g0);
    extrap3(b2, b1, h3, h4, h5);
    col[2 * n] = g2;
    col[1 * n] = g1;
    col[0] = g0;
    col[(ny + MD) * n] = h3;
    col[(ny + MD + 1) * n] = h4;
    col[(ny + MD + 2) * n] = h5;
    return nonzero_bits(g0) || nonzero_bits(g1) || nonzero_bits(g2) || nonzero_bits(h3) ||
           nonzero_bits(h4) || nonzero_bits(h5);
  }
}

// Whether a landslide's ghost line reads only cells of dry segments (z: the flags of the
// buffer it fills, valid for nseg > 0): an interior row reads columns 3, 4, nx+1 and nx+2
// of its own row, an interior column rows 3, 4, ny+1 and ny+2 of its own column.  The
// ghost rows (they read the other buffer's ghost rows) and columns (which a landslide's
// fill leaves alone) count as not dry.
__device__ __forceinline__ bool ghost_sources_dry(const Grid& g, const unsigned char* __restrict__ z,
                                                  int nseg, int line) {
  if (line < g.nyd) {
    const int r = line;
    if (r < MD || r >= g.ny + MD) return false;
    const unsigned char* f = z + (long long)(r - MD) * nseg;
    return (f[0] | f[(g.nx - 2) / 63] | f[(g.nx - 1) / 63]) == 0;
  }
  const int c = line - g.nyd;
  if (c < MD || c >= g.nx + MD) return false;
  const int sg = (c - MD) / 63;
  return (z[sg] | z[nseg + sg] | z[(long long)(g.ny - 2) * nseg + sg] |
          z[(long long)(g.ny - 1) * nseg + sg]) == 0;
}

// ghost_line with the dry skip: a landslide line (plane k) whose sources are dry and whose
// ghosts the last fill left at +0 is not filled (the fill would write the same +0 bits);
// otherwise it is filled and its flag says whether it wrote anything but +0.
template <typename T>
__device__ __forceinline__ void ghost_line_skip(const Grid& g, int mode, T* __restrict__ Tg,
                                                const T* __restrict__ prev, int line, int k,
                                                const unsigned char* __restrict__ z,
                                                unsigned char* __restrict__ gf, int nseg) {
  if (mode != 0 || nseg <= 0 || !gf) {
    ghost_line(g, mode, Tg, prev, line, k);
    return;
  }
  unsigned char* f = gf + (long long)k * (g.nyd + g.nxd) + line;
  const unsigned char was = *f;
  if (was == 0 && ghost_sources_dry(g, z, nseg, line)) return;
  const unsigned char now = ghost_line(g, mode, Tg, prev, line, k) ? 1 : 0;
  if (now != was) *f = now;
}

// Max of one wave (64 lanes) with fmax semantics (NaN ignored, as fmaxf in reduceKernel).
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Flux1..6Kernel + CFL1/2Kernel + reduceKernel (L/K:1426-1848, 1938-2004) for the dt of
// the predictor: the max over faces [1, N-3]^2 of em/dx, em/dy (solid and fluid).  Faces
// outside that square have a zero face-averaged metric in the reference, so their em is
// the 1e-10 floor, which never exceeds these.  Each block's maximum goes to its own slot,
// which the dt kernel reduces.  k_cfl covers the faces of the interior box [3, ny+2] x [3, nx+2]
// (InterfacesKernel states); its ring blocks the rest of [1, N-3]^2, where Interfaces2Kernel
// overrides the states.  Speeds are formed in T; the quotient by dx and the max in FP64.
//
// The reference divides each face's speed by the cell size before the max (CFL1Kernel,
// L/K:1939-1957).  The cells are square (dx = dy; one cellsize is read) and a correctly
// rounded division by a positive constant is monotonic, so the max of the quotients is the
// quotient of the max: the faces here give em, and the dt kernel divides the max once.
//
// CFL values carry a non-finite mark in their sign.  A face value is at least the floor
// 1e-10 > 0; it is negated when one of the face's four one-sided speeds (apE, apW, apFE,
// apFW) is not finite.  The reference's speed floor and maxima drop a NaN there without a
// trace (fmax, fmaxf), so this is the only place it can still be seen.  cfl_join keeps
// the larger magnitude, negative if either value is, so the magnitude of a reduction is
// the same fmax the dt control has always used, and its sign says whether any face met a
// non-finite state.  Slots with no faces hold +0.
template <typename T, bool UNI = false>
__device__ __forceinline__ double face_cfl(const T* __restrict__ U, const T* __restrict__ M,
                                           const Grid& g, const Phys& ph, int dir, int r, int c,
                                           bool edge) {
  T L[6], R[6];
  constexpr int SRC = UNI ? FM_UNIFORM : FM_STORED;
  FaceMetIn<T, SRC> fm;
  if (edge) {
    final_states_edge(U, g, dir, r, c, L, R);
    face_metric_load(M, g, dir, r, c, fm);
  } else {   // the face's metric and state loads issued together
    Stencil<T> sn;
    face_metric_load(M, g, dir, r, c, fm);
    load_stencil(U, g, dir, r, c, sn);
    __builtin_amdgcn_sched_barrier(0);
    stencil_states(sn, L, R);
  }
  FaceMet<T> f;
  face_metric_form(fm, dir, f);
  FaceSpeed<T> s;
  face_speed(dir, L, R, f, ph, s);
  const double v = fmax((double)s.emS, (double)s.emF);
  return s.finite ? v : -v;
}

__device__ __forceinline__ double cfl_join(double a, double b) {
  const double m = fmax(fabs(a), fabs(b));
  return (a < 0.0 || b < 0.0) ? -m : m;
}

__device__ __forceinline__ double wave_cfl(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = cfl_join(v, __shfl_xor(v, off, 64));
  return v;
}

// Max over the block of a per-thread value (fmax: NaN ignored unless all are NaN), written
// by thread 0 to *out.  Every thread of the block must call it; it may be called again
// right away (it ends with a barrier).  Each block writes its own slot and the dt kernel
// reduces the slots, so no two blocks contend for one address.  JOIN: combine with
// cfl_join (CFL values) instead of fmax.
template <bool JOIN = false>
__device__ __forceinline__ void block_max_store(double m, double* out) {
  m = JOIN ? wave_cfl(m) : wave_max(m);
  __shared__ double part[16];
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int nw = (blockDim.x * blockDim.y + 63) / 64;
  if ((tid & 63) == 0) part[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    double b = part[0];
    for (int w = 1; w < nw; w++) b = JOIN ? cfl_join(b, part[w]) : fmax(b, part[w]);
    *out = b;
  }
  __syncthreads();
}

__device__ __forceinline__ void block_cfl_store(double m, double* out) {
  block_max_store<true>(m, out);
}

// fmax of p[0..n) by all threads of a 1-D block; every thread gets the result.  The max
// of a set does not depend on the order it is formed in.  JOIN: cfl_join instead.
template <bool JOIN = false>
__device__ __forceinline__ double block_max_of(const double* __restrict__ p, int n) {
  auto op = [](double a, double b) { return JOIN ? cfl_join(a, b) : fmax(a, b); };
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;   // four independent chains of loads
  const int s = blockDim.x;
  int i = threadIdx.x;
  for (; i + 3 * s < n; i += 4 * s) {
    m0 = op(m0, p[i]);
    m1 = op(m1, p[i + s]);
    m2 = op(m2, p[i + 2 * s]);
    m3 = op(m3, p[i + 3 * s]);
  }
  for (; i < n; i += s) m0 = op(m0, p[i]);
  double m = op(op(m0, m1), op(m2, m3));
  m = JOIN ? wave_cfl(m) : wave_max(m);
  __shared__ double part[16];
  __syncthreads();
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  double b = part[0];
  for (int w = 1; w < (int)((blockDim.x + 63) / 64); w++) b = op(b, part[w]);
  __syncthreads();
  return b;
}

__device__ __forceinline__ double block_cfl_of(const double* __restrict__ p, int n) {
  return block_max_of<true>(p, n);
}

// The dry test of a region [ra, rb] x [ca, cb] of U (clipped to the padded grid) as a list
// of items, each one load and test: first the row-segment flags of its interior cells,
// then the corner ghost cells it holds (components 0..5).  The other ghost cells need no
// test: the ghost fill that ran before the caller made them 0.5-weighted sums of the
// interior cells within two cells of the boundary in the same row or column (momenta 0),
// and for any region of a stage tile or CFL box (at least 7 rows and 67 columns) those
// cells lie in the region's interior part, so their flags are among the items.  The corner
// blocks (rows and columns both ghost) may lag a stage behind, so they are read.
struct DryRegion {
  int ia, ib, sa, ns, nf;        // interior rows [ia, ib], segments [sa, sa+ns), nf flag items
  int rt, nrt, rbo, ncg, nrg;    // corner rows: [rt, rt+nrt) and [rbo, ...), nrg in all
  int cl, ncl, cr;               // corner columns: [cl, cl+ncl) and [cr, ...), ncg in all
  int n;                         // items
};

__device__ __forceinline__ DryRegion dry_region_of(const Grid& g, int ra, int rb, int ca, int cb) {
  DryRegion d;
  ra = max(ra, 0);
  rb = min(rb, g.nyd - 1);
  ca = max(ca, 0);
  cb = min(cb, g.nxd - 1);
  const int xmax = g.nx + MD - 1, ymax = g.ny + MD - 1;
  d.ia = max(ra, MD);
  d.ib = min(rb, ymax);
  const int ib = d.ib;
  d.sa = (max(ca, MD) - MD) / 63;
  d.ns = (min(cb, xmax) - MD) / 63 - d.sa + 1;
  d.nf = (ib - d.ia + 1) * d.ns;
  d.rt = ra;
  d.nrt = max(0, min(rb, MD - 1) - ra + 1);
  d.rbo = max(ra, ymax + 1);
  d.nrg = d.nrt + max(0, rb - d.rbo + 1);
  d.cl = ca;
  d.ncl = max(0, min(cb, MD - 1) - ca + 1);
  d.cr = max(ca, xmax + 1);
  d.ncg = d.ncl + max(0, cb - d.cr + 1);
  d.n = d.nf + d.nrg * d.ncg;
  return d;
}

// 1 if item e of the region is wet (NaN counts as wet)
template <typename T>
__device__ __forceinline__ int dry_item_wet(const DryRegion& d, const T* __restrict__ U,
                                            const Grid& g, const unsigned char* __restrict__ z,
                                            int nseg, int e) {
  if (e < d.nf) return z[(long long)(d.ia + e / d.ns - MD) * nseg + d.sa + e % d.ns] != 0;
  e -= d.nf;
  const int ir = e / d.ncg, ic = e % d.ncg;
  const int r = ir < d.nrt ? d.rt + ir : d.rbo + (ir - d.nrt);
  const int c = ic < d.ncl ? d.cl + ic : d.cr + (ic - d.ncl);
  const long long i = idx(g, r, c);
  int wet = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) wet |= (U[k * g.P + i] != T(0)) ? 1 : 0;
  return wet;
}

// The OR of the interior flag items of a region at most NR rows tall whose columns span at
// most three segments (every stage tile's and CFL box's region: 67 columns), without the
// per-item divisions of dry_item_wet: lane q of L takes rows d.ia + q, d.ia + q + L, ...
// The region's corner items (e from d.nf) are left to the caller.
template <int NR, int L>
__device__ __forceinline__ int dry_flags_wet3(const DryRegion& d, const unsigned char* __restrict__ z,
                                              int nseg, int q) {
  int wet = 0;
#pragma unroll
  for (int k = 0; k < (NR + L - 1) / L; k++) {
    const int r = d.ia + q + k * L;
    if (r <= d.ib) {
      const unsigned char* zr = z + (long long)(r - MD) * nseg + d.sa;
#pragma unroll
      for (int j = 0; j < 3; j++)
        if (j < d.ns) wet |= zr[j];
    }
  }
  return wet;
}

// True when the region holds only zeros.  Every thread of the block must call it; tid in
// [0, nthr).  A region of more than max_items items is not tested (false).
template <typename T>
__device__ __forceinline__ bool dry_region(const T* __restrict__ U, const Grid& g,
                                           const unsigned char* __restrict__ z, int nseg, int ra,
                                           int rb, int ca, int cb, int tid, int nthr,
                                           int max_items = INT_MAX) {
  const DryRegion d = dry_region_of(g, ra, rb, ca, cb);
  if (d.n > max_items) return false;
  int wet = 0;
  for (int e = tid; e < d.n; e += nthr) wet |= dry_item_wet(d, U, g, z, nseg, e);
  return !__syncthreads_or(wet);
}

// The same test by one wave on its own (every lane of the wave calls it; lane in [0, 64)).
template <typename T>
__device__ __forceinline__ bool dry_region_wave(const T* __restrict__ U, const Grid& g,
                                                const unsigned char* __restrict__ z, int nseg,
                                                int ra, int rb, int ca, int cb, int lane,
                                                int max_items) {
  const DryRegion d = dry_region_of(g, ra, rb, ca, cb);
  if (d.n > max_items) return false;
  int wet = 0;
  for (int e = lane; e < d.n; e += 64) wet |= dry_item_wet(d, U, g, z, nseg, e);
  return __ballot(wet) == 0ull;
}

// CFL boxes: 64x4 blocks of interior faces (both directions), box (bx, by) starting at
// column MD + 64 bx, row MD + 4 by.  A box's faces read columns c-1..c+2 (x-faces) and rows
// r-1..r+2 (y-faces): the stencil [rlo-1, rlo+5] x [clo-1, clo+65].  A box whose stencil
// is all zero gives the wave-speed floor 1e-10 over dx and dy on every face.
template <typename T>
__device__ __forceinline__ double cfl_floor(const Grid& g) {
  (void)g;
  return (double)T(1.0e-10);
}

__device__ __forceinline__ DryRegion box_region(const Grid& g, int bx, int by) {
  const int clo = MD + 64 * bx, rlo = MD + 4 * by;
  return dry_region_of(g, rlo - 1, rlo + 5, clo - 1, clo + 65);
}

template <typename T, bool UNI = false>
__device__ __forceinline__ double cfl_box(const T* __restrict__ U, const T* __restrict__ M,
                                          const Grid& g, const Phys& ph, int bx, int by) {
  const int c = bx * 64 + threadIdx.x + MD;
  const int r = by * 4 + threadIdx.y + MD;
  double m = 0.0;
  if (r <= g.ny + MD - 1 && c <= g.nx + MD - 1) {
    m = face_cfl<T, UNI>(U, M, g, ph, 0, r, c, false);
    m = cfl_join(m, face_cfl<T, UNI>(U, M, g, ph, 1, r, c, false));
  }
  return m;
}

// The ring [1, N-3]^2 minus the interior box, one thread per (face index, direction).
// Face indices are enumerated as 4 strips: rows 1-2 and ny+3..ny+4 (columns 1..N-3), then
// columns 1-2 and nx+3..nx+4 on rows 3..ny+2.
__host__ __device__ __forceinline__ int ring_threads(const Grid& g) {
  return 2 * (4 * (g.nxd - 3) + 4 * g.ny);
}

// Position of ring thread t's face; false past the ring.
__device__ __forceinline__ bool ring_face_pos(const Grid& g, int t, int& r, int& c, int& dir) {
  const int f = t >> 1;
  dir = t & 1;
  const int w = g.nxd - 3;        // columns 1..N-3
  const int nrow = 4 * w;         // 4 ghost face rows
  const int ncol = 4 * g.ny;      // 4 ghost face columns on interior rows
  if (f >= nrow + ncol) return false;
  if (f < nrow) {
    const int k = f / w;
    r = (k < 2) ? 1 + k : g.ny + MD + (k - 2);
    c = 1 + f % w;
  } else {
    const int q = f - nrow, k = q / g.ny;
    r = MD + q % g.ny;
    c = (k < 2) ? 1 + k : g.nx + MD + (k - 2);
  }
  return true;
}

// The CFL value of ring thread t (0 past the ring).  Called by every lane of a wave, whose
// threads are consecutive.  With dry-tile skipping the wave first tests the cells its faces
// read: a ring face
// on a ghost row or column takes its states from rows 3,4 (ny+2,ny+1) and columns 3,4
// (nx+2,nx+1) instead (Interfaces2Kernel), and base_states reads a row r' or column c'
// and the cells from r'-1 to r'+2 (c'-1 to c'+2); so the wave's faces read within the
// bounding box of those effective rows and columns, widened by 1 before and 2 after.  When
// it is dry, each face gives the floor 1e-10 over its spacing, as computing it would.  A
// wave whose faces lie on two strips far apart (its box would span the grid) is not tested.
template <typename T, bool UNI = false>
__device__ __forceinline__ double cfl_ring_face(const T* __restrict__ U, const T* __restrict__ M,
                                                const Grid& g, const Phys& ph, int t,
                                                const DryMap& dm) {
  int r = 0, c = 0, dir = 0;
  const bool valid = ring_face_pos(g, t, r, c, dir);
  if (dm.nseg > 0) {
    int ra = INT_MAX, rb = INT_MIN, ca = INT_MAX, cb = INT_MIN;
    if (valid) {
      const bool gr = r <= 2 || r >= g.ny + MD, gc = c <= 2 || c >= g.nx + MD;
      ra = gr ? (r <= 2 ? MD : g.ny + MD - 2) : r;
      rb = gr ? (r <= 2 ? MD + 1 : g.ny + MD - 1) : r;
      ca = gc ? (c <= 2 ? MD : g.nx + MD - 2) : c;
      cb = gc ? (c <= 2 ? MD + 1 : g.nx + MD - 1) : c;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      ra = min(ra, __shfl_xor(ra, off, 64));
      rb = max(rb, __shfl_xor(rb, off, 64));
      ca = min(ca, __shfl_xor(ca, off, 64));
      cb = max(cb, __shfl_xor(cb, off, 64));
    }
    if (rb >= ra && dry_region_wave(U, g, dm.z0, dm.nseg, ra - 1, rb + 2, ca - 1, cb + 2,
                                    (int)(threadIdx.x & 63), 4 * 64))
      return valid ? (double)T(1.0e-10) : 0.0;
  }
  return valid ? face_cfl<T, UNI>(U, M, g, ph, dir, r, c, true) : 0.0;
}

// CFL3Kernel + the host's dt control + CFL4Kernel (L/K:2006-2036, L/R:978-991; inflow
// sample advance I/R:1078-1081), on the device.  mx is the max wave speed over the CFL
// faces, negative when some face met a non-finite state (face_cfl); its magnitude over dx
// is rounded to float as the reference's fmaxf reduction does (L/K:1988, 1996).
// Also the halting rules, which have no reference counterpart: a step budget (halt_at),
// an end time (t_end: the step that reaches it is shortened to end on it) and a stop speed
// (vstop2: once the squared max flow speed v2 of a corrector has reached it and a later one
// falls below it, the next step is the last); and the same budget and end time for the
// duration of one m2p_run_to call (call_halt, call_t_end).
__device__ __forceinline__ bool dt_halted(DevState* __restrict__ st) {
  if (st->skip) return true;
  if (st->nt >= st->halt_at || st->nt >= st->call_halt) {
    st->skip = 1;
    return true;
  }
  return false;
}

__device__ __forceinline__ void dt_body(const DevParams& dp, DevState* __restrict__ st,
                                        const double* __restrict__ tf,
                                        const double* __restrict__ itime, double mx) {
  if (mx < 0.0) {
    if (!st->nonfinite) st->nonfinite_nt = st->nt;
    st->nonfinite = 1;
    mx = -mx;
  }
  mx = mx / dp.dx;   // the CFL faces give em; dx = dy
  if (st->vstop2 > 0.0 && st->v2_nt == st->nt) {
    const double v2 = st->v2;
    if (v2 >= st->vstop2)
      st->moved = 1;
    else if (st->moved && st->halt_at > st->nt + 1)
      st->halt_at = st->nt + 1;
  }
  const double dmax = (double)(float)mx;
  double dtval = dp.cfl / dmax;
  dtval = fmin(dp.dt_max, dtval);
  dtval = fmax(dp.dt_min, dtval);
  const double T = st->time;
  st->step_start = T;
  st->dmax = dmax;
  if (st->iter < dp.ntf && (T + dtval) >= tf[st->iter]) {
    dtval = tf[st->iter] - T;
    st->schreiben = 1;
    st->iter = st->iter + 1;
  }
  if (dp.nsamp > 0 && T >= itime[st->inflowCount] && st->inflowCount < (dp.nsamp - 1))
    st->inflowCount = st->inflowCount + 1;
  // the nearer of the two end times shortens the step; each rule whose end is reached
  // halts the scenario after it
  const bool own = st->t_end > 0.0, call = st->call_t_end > 0.0;
  if (own || call) {
    const double te = (own && call) ? fmin(st->t_end, st->call_t_end) : own ? st->t_end : st->call_t_end;
    if ((T + dtval) >= te) {
      dtval = te - T;
      if (own && st->t_end == te && st->halt_at > st->nt + 1) st->halt_at = st->nt + 1;
      if (call && st->call_t_end == te && st->call_halt > st->nt + 1) st->call_halt = st->nt + 1;
    }
  }
  st->dtval = dtval;
  st->dt_pred = dtval / (2.0 - 0);
  st->dt_corr = dtval / (2.0 - 1);
  st->dtdx_pred = st->dt_pred / dp.dx;
  st->dtdy_pred = st->dt_pred / dp.dy;
  st->dtdx_corr = st->dt_corr / dp.dx;
  st->dtdy_corr = st->dt_corr / dp.dy;
  st->time = st->dt_corr + T;
  st->nt = st->nt + 1;
}


// The viscous term's quotients of cell i (PDSide): u4, u5 and their limited slopes along
// direction `tdir` (0: x, 1: y), each over u3 + H0, in the reference's operations.
template <typename T>
__device__ __forceinline__ PDSide<T> pd_side(const T* __restrict__ U, const Grid& g, long long i,
                                             int tdir) {
  const long long P = g.P, st = tdir == 0 ? 1 : g.nxd;
  const T u3 = U[3 * P + i], u4 = U[4 * P + i], u5 = U[5 * P + i];
  const T t4 = tvd_slope(U[4 * P + i - st], u4, U[4 * P + i + st]);
  const T t5 = tvd_slope(U[5 * P + i - st], u5, U[5 * P + i + st]);
  const T d = u3 + T(H0);
  PDSide<T> p;
  p.q4 = qdiv(u4, d);
  p.q5 = qdiv(u5, d);
  p.s4 = qdiv(t4, d);
  p.s5 = qdiv(t5, d);
  return p;
}

// One stage (predictor IO = 0, corrector IO = 1): Flux8..21Kernel (L/K:2038-2726) and the
// crossing limiter, Flux22..24 / Flux25..28 (L/K:2730-2937).
//   IO = 0: Uin = U0 = u^n (also uzero).  Writes u^{n+1/2} into Uout = U1 and the limited
//           fluid momenta uone[4], uone[5] into ONE.
//   IO = 1: Uin = U1, uzero = U0, old velocities from (U1[0..3], ONE).  Writes u^{n+1}
//           into U0 in place (each thread reads and writes only its own cell of U0).
// A block of 8 waves owns a tile of TX = 63 columns x TY = 7 rows of cells.  Each face is
// computed once: the 64 x-faces of a tile row by one wave, then (through LDS) the 64
// y-faces of a face row by one wave; a cell reads its two faces of each direction from
// LDS.  The update keeps the reference's left-to-right sum, so the x terms are formed
// first and held in registers while the y faces are computed.
constexpr int ST_TX = 63, ST_NQ = 12;

// Phase A: InterfacesKernel states of one face into the tile (L: q 0-5, R: q 6-11).
template <typename T, int TY, bool LB>
__device__ __forceinline__ void stage_states(T (*Tl)[TY + 1][64], int w, int lane,
                                             const T* __restrict__ U, const Grid& g, int dir, int r,
                                             int c) {
  T L[6], R[6];
  base_states<T, LB>(U, g, dir, r, c, L, R);
#pragma unroll
  for (int k = 0; k < 6; k++) {
    Tl[k][w][lane] = L[k];
    Tl[6 + k][w][lane] = R[k];
  }
}

// Phase B: the face's final states.  The only Interfaces2Kernel overrides a stage needs are
// the x-face at column 2 (from the faces at columns 3, 4: lanes +1, +2) and the y-face at
// row 2 (from rows 3, 4: tile rows +1, +2).
template <typename T, int TY>
__device__ __forceinline__ void load_states(T (*Tl)[TY + 1][64], int w, int lane, bool ovr,
                                            int dir, T L[6], T R[6]) {
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (!ovr) {
      L[k] = Tl[k][w][lane];
      R[k] = Tl[6 + k][w][lane];
    } else if (dir == 0) {
      L[k] = T(0.5) * (Tl[k][w][lane + 1] + Tl[k][w][lane + 2]);
      R[k] = T(0.5) * (Tl[6 + k][w][lane + 1] + Tl[6 + k][w][lane + 2]);
    } else {
      L[k] = T(0.5) * (Tl[k][w + 1][lane] + Tl[k][w + 2][lane]);
      R[k] = T(0.5) * (Tl[6 + k][w + 1][lane] + Tl[6 + k][w + 2][lane]);
    }
  }
}

// Where a stage takes its face metric from (face_metric).
template <typename T, bool UNI>
constexpr int stage_fm_src() {
  return UNI ? FM_UNIFORM : std::is_same<T, double>::value ? FM_STORED : FM_CELLS;
}

// Phase C: fluxes of one face from its final states and its loaded metric, into the tile.
// pl, pr: the viscous quotients of the face's left/lower and right/upper cells.
template <typename T, int TY, int SRC>
__device__ __forceinline__ void flux_store(T (*Tl)[TY + 1][64], int w, int lane,
                                           const FaceMetIn<T, SRC>& fm, const Grid& g,
                                           const Phys& ph, int dir, T L[6], T R[6],
                                           const PDSide<T>& pl, const PDSide<T>& pr) {
  FaceMet<T> f;
  face_metric_form(fm, dir, f);
  FaceSpeed<T> sp;
  face_speed(dir, L, R, f, ph, sp);
  FaceFlux<T> o;
  face_flux(dir, L, R, f, sp, pl, pr, g, ph, o);
#pragma unroll
  for (int k = 0; k < 6; k++) Tl[k][w][lane] = o.H[k];
  Tl[6][w][lane] = o.P1;
  Tl[7][w][lane] = o.P2;
  Tl[8][w][lane] = o.P4;
  Tl[9][w][lane] = o.P5;
  Tl[10][w][lane] = o.D4;
  Tl[11][w][lane] = o.D5;
}

// Returns, for IO = 1 and want_v2 (a stop speed is set; uniform over the launch), the larger
// squared flow speed (solid, fluid) of the cell's new state if its flow depth exceeds
// WET_SPEED (0 otherwise, and for IO = 0).
constexpr double WET_SPEED = 1.0e-2;  // scaled depth (1 mm)

// FLAGS: write the dry-tile flags of the output rows (only the planned stage kernels run
// with dry-tile skipping on; the others are compiled without the flag code).  UNI: the
// grid's metric passed the uniform-metric check, so the inverse Jacobian is formed from
// svec and cvalue instead of loaded (FM_UNIFORM): in FP64 the same values bit for bit, in
// FP32 formed in float from the rounded svec and cvalue.  LB: issue each face's loads, and
// the update's, together ahead of their arithmetic (base_states); off for the FP32 kernels
// held to 64 VGPRs, which spill with it.  The results are the same either way.
template <int IO, typename T, int TY, bool FLAGS, bool UNI = false, bool LB = true>
__device__ __forceinline__ double stage_tile(T (*F)[TY + 1][64], const Grid& g, const Phys& ph,
                                           const T* __restrict__ Uin, const T* Uzero, T* Uout,
                                           T* __restrict__ ONE, const T* __restrict__ M,
                                           unsigned char* __restrict__ LIM, double dt_d, int tx,
                                           int ty, const DryMap& dm, double dtdx_d = 0.0,
                                           double dtdy_d = 0.0, bool want_v2 = true) {
  const int lane = threadIdx.x;
  const int w = threadIdx.y;
  const int c0 = tx * ST_TX + MD;
  const int r0 = ty * TY + MD;
  const int xmax = g.nx + MD - 1, ymax = g.ny + MD - 1;
  const long long P = g.P;

  // The viscous quotients (y-slopes) of the cells the x-faces touch, formed once per cell
  // and shared through LDS: wave w < TY forms row r0+w, columns c0-1..c0+62 (slot lane), and
  // the last wave, idle in the x phase, forms column c0+63 of every row (slot 64); face
  // c0-1+lane takes slots lane and lane+1.  The y phase reuses the array.
  __shared__ PDSide<T> Q[(TY + 2) * 64];
  if (w < TY) {
    const int r = r0 + w, c = c0 - 1 + lane;
    if (r <= ymax && c <= xmax + 1) Q[w * 65 + lane] = pd_side(Uin, g, idx(g, r, c), 1);
  } else if (lane < TY) {
    const int r = r0 + lane, c = c0 + 63;
    if (r <= ymax && c <= xmax + 1) Q[lane * 65 + 64] = pd_side(Uin, g, idx(g, r, c), 1);
  }
  __syncthreads();

  // this thread's cell; its phi is loaded now, with the x-phase's loads
  const int r = r0 + w, c = c0 + lane;
  const bool active = (w < TY) && (lane < ST_TX) && r <= ymax && c <= xmax;
  const long long i = idx(g, r, c);
  T phi = T(0);
  if (active) phi = Uin[6 * P + i];

  // x faces (r0+w, c0-1+lane), faces (r, c)-(r, c+1).  Only tiles in the first tile column
  // hold the overridden face at column 2; the others form their face states in registers
  // and skip the LDS round trip (the branch is uniform over the block).  A face's metric
  // and state loads are issued together, ahead of its arithmetic.
  constexpr int SRC = stage_fm_src<T, UNI>();
  {
    const int r = r0 + w, c = c0 - 1 + lane;
    const bool has = (w < TY && r <= ymax && c <= xmax);
    T L[6], R[6];
    FaceMetIn<T, SRC> fm;
    if (tx == 0) {
      if (has) stage_states<T, TY, LB>(F, w, lane, Uin, g, 0, r, c);
      __syncthreads();
      if (has) {
        face_metric_load(M, g, 0, r, c, fm);
        load_states<T, TY>(F, w, lane, c == 2, 0, L, R);
      }
      __syncthreads();
    } else if (has) {
      Stencil<T> sn;
      face_metric_load(M, g, 0, r, c, fm);
      load_stencil(Uin, g, 0, r, c, sn);
      if constexpr (LB) __builtin_amdgcn_sched_barrier(0);
      stencil_states(sn, L, R);
    }
    if (has) flux_store<T, TY, SRC>(F, w, lane, fm, g, ph, 0, L, R, Q[w * 65 + lane], Q[w * 65 + lane + 1]);
  }
  __syncthreads();

  const T dt = T(dt_d);
  // FP64: dt/dx and dt/dy as the dt kernel formed them (the same IEEE quotients)
  constexpr bool F64 = std::is_same<T, double>::value;
  const T dtdx = F64 ? T(dtdx_d) : qdiv(dt, T(g.dx));
  const T dtdy = F64 ? T(dtdy_d) : qdiv(dt, T(g.dy));
  const T epsilon0 = T(1.0);
  T fP = T(0), fD = T(0);
  T tH[6], tP[4], tD[2];
  if (active) {
    fP = epsilon0 * phi;
    fD = qdiv(epsilon0 * (T(1.0) - phi), T(ph.N_R));
#pragma unroll
    for (int k = 0; k < 6; k++) tH[k] = (dtdx) * (F[k][w][lane + 1] - F[k][w][lane]);
#pragma unroll
    for (int q = 0; q < 4; q++) tP[q] = (dtdx) * (F[6 + q][w][lane + 1] - F[6 + q][w][lane]) * (fP);
#pragma unroll
    for (int q = 0; q < 2; q++) tD[q] = (dtdx) * (F[10 + q][w][lane + 1] - F[10 + q][w][lane]) * (fD);
  }
  // The viscous quotients of cell rows r0-1 .. r0+TY (x-slopes), for the y-faces: face row
  // r0-1+w takes row w as its lower cells and row w+1 as its upper ones, so each cell's
  // quotients are formed once (wave 0 also forms the extra top row).  The x phase's reads
  // of Q all precede the barrier after the x-face fluxes, so these stores may reuse it.
  {
    const int cq = c0 + lane;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int row = w + k * (TY + 1);   // k = 1: the extra row, wave 0 only
      if (row > TY + 1) break;
      const int rq = r0 - 1 + row;
      if (lane < ST_TX && rq <= ymax + 1 && cq <= xmax) Q[row * 64 + lane] = pd_side(Uin, g, idx(g, rq, cq), 0);
    }
  }
  __syncthreads();

  // y faces (r0-1+w, c0+lane), faces (r, c)-(r+1, c); the overridden face row 2 lies in
  // the first tile row only
  {
    const int rf = r0 - 1 + w, cf = c0 + lane;
    const bool has = (lane < ST_TX && rf <= ymax && cf <= xmax);
    T L[6], R[6];
    FaceMetIn<T, SRC> fm;
    if (ty == 0) {
      if (has) stage_states<T, TY, LB>(F, w, lane, Uin, g, 1, rf, cf);
      __syncthreads();
      if (has) {
        face_metric_load(M, g, 1, rf, cf, fm);
        load_states<T, TY>(F, w, lane, rf == 2, 1, L, R);
      }
      __syncthreads();
    } else if (has) {
      Stencil<T> sn;
      face_metric_load(M, g, 1, rf, cf, fm);
      load_stencil(Uin, g, 1, rf, cf, sn);
      if constexpr (LB) __builtin_amdgcn_sched_barrier(0);
      stencil_states(sn, L, R);
    }
    if (has) flux_store<T, TY, SRC>(F, w, lane, fm, g, ph, 1, L, R, Q[w * 64 + lane], Q[(w + 1) * 64 + lane]);
  }
  __syncthreads();
  if (!active) return 0.0;

  // the update's loads, issued together: the cell's state, metric, uzero (the same buffer as
  // Uin in the predictor, so taken from u there) and the old fluid momenta of the corrector
  T u[7];
#pragma unroll
  for (int k = 0; k < 7; k++) u[k] = Uin[k * P + i];
  CellMet<T> cm;
  cell_metric_load<T, UNI>(M, P, i, cm);
  T uz[6];
#pragma unroll
  for (int k = 0; k < 6; k++) uz[k] = IO == 0 ? u[k] : Uzero[k * P + i];
  T one4 = T(0), one5 = T(0);
  if (IO == 1) {
    one4 = ONE[i];
    one5 = ONE[P + i];
  }
  if constexpr (LB) __builtin_amdgcn_sched_barrier(0);
  cell_metric_form<T, UNI>(cm);
  T s[6];
  cell_source(u, cm, ph, s);

  // Flux18Kernel L/K:2641-2652 (Pp and PD are stored zeros for the components they skip)
  T vnew[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int q = k == 1 ? 0 : k == 2 ? 1 : k == 4 ? 2 : 3;  // P1, P2, P4, P5
    const bool hasP = (k == 1 || k == 2 || k == 4 || k == 5);
    const bool hasD = (k == 4 || k == 5);
    const T zero = T(0.0);
    const T yH = (dtdy) * (F[k][w + 1][lane] - F[k][w][lane]);
    const T xP = hasP ? tP[q] : (dtdx) * (zero - zero) * (fP);
    const T yP = hasP ? (dtdy) * (F[6 + q][w + 1][lane] - F[6 + q][w][lane]) * (fP)
                      : (dtdy) * (zero - zero) * (fP);
    const T xD = hasD ? tD[k - 4] : (dtdx) * (zero - zero) * (fD);
    const T yD = hasD ? (dtdy) * (F[10 + (k - 4)][w + 1][lane] - F[10 + (k - 4)][w][lane]) * (fD)
                      : (dtdy) * (zero - zero) * (fD);
    vnew[k] = uz[k] * cm.det - tH[k] - yH + xP + yP + xD + yD + (dt)*s[k];
  }
  const T h0 = T(H0);
  T un[7];
#pragma unroll
  for (int k = 0; k < 6; k++) un[k] = qdiv(vnew[k], cm.det);  // Flux19
  if (un[0] <= T(H00)) {                                     // Flux20
    un[0] = T(0);
    un[1] = T(0);
    un[2] = T(0);
  }
  if (un[3] <= T(H00)) {
    un[3] = T(0);
    un[4] = T(0);
    un[5] = T(0);
  }
  un[6] = qdiv(un[0], un[0] + un[3] + h0);  // Flux21

  // velocity-crossing limiter (Flux23/24, Flux26/27)
  T o0, o1, o2, o3, o4, o5;
  if (IO == 0) {
    o0 = uz[0]; o1 = uz[1]; o2 = uz[2]; o3 = uz[3]; o4 = uz[4]; o5 = uz[5];
  } else {
    o0 = u[0]; o1 = u[1]; o2 = u[2]; o3 = u[3]; o4 = one4; o5 = one5;
  }
  const T usxnew = qdiv(un[1], un[0] + h0);
  const T ufxnew = qdiv(un[4], un[3] + h0);
  const T usxold = qdiv(o1, o0 + h0);
  const T ufxold = qdiv(o4, o3 + h0);
  const T usynew = qdiv(un[2], un[0] + h0);
  const T ufynew = qdiv(un[5], un[3] + h0);
  const T usyold = qdiv(o2, o0 + h0);
  const T ufyold = qdiv(o5, o3 + h0);
  T lim4, lim5;
  if constexpr (std::is_same<T, double>::value) {
    lim4 = ((usxnew - ufxnew) * (usxold - ufxold) >= T(0.0)) ? un[4] : usxnew * un[3];
    lim5 = ((usynew - ufynew) * (usyold - ufyold) >= T(0.0)) ? un[5] : usynew * un[3];
  } else {
    // Limiter memory (FP32 only).  When a stage limits, the stored fluid momentum is
    // us*hf, and in FP64 the next stage's old relative velocity is us*H0/(hf+H0): about
    // 1e-8 of us, with the sign of us.  Float storage rounds that to noise, so the flag
    // of the stored state says to take the sign of us instead.  Bits 0-1: flags of the
    // stage-1 output (U0, read by stage 0); bits 2-3: flags of ONE (read by stage 1).
    const unsigned f = IO == 0 ? (LIM[i] & 3u) : ((LIM[i] >> 2) & 3u);
    const T dxo = (f & 1u) ? usxold : (usxold - ufxold);
    const T dyo = (f & 2u) ? usyold : (usyold - ufyold);
    const bool kx = (usxnew - ufxnew) * dxo >= T(0.0);
    const bool ky = (usynew - ufynew) * dyo >= T(0.0);
    lim4 = kx ? un[4] : usxnew * un[3];
    lim5 = ky ? un[5] : usynew * un[3];
    const unsigned nb = (kx ? 0u : 1u) | (ky ? 0u : 2u);
    LIM[i] = (unsigned char)(IO == 0 ? ((f & 3u) | (nb << 2)) : nb);
  }
  if (IO == 1) {
    un[4] = lim4;
    un[5] = lim5;
  }
  if (FLAGS && dm.nseg > 0) {
    // this row segment's flag of the output state (lane 0 holds column c0, always active)
    bool nz = false;
#pragma unroll
    for (int k = 0; k < 7; k++) nz = nz || (un[k] != T(0));
    const unsigned long long wet = __ballot(nz);
    if (lane == 0) (IO == 0 ? dm.z1 : dm.z0)[(long long)(r - MD) * dm.nseg + tx] = wet != 0ull;
  }
  if (IO == 0) {
#pragma unroll
    for (int k = 0; k < 7; k++) Uout[k * P + i] = un[k];
    ONE[i] = lim4;
    ONE[P + i] = lim5;
    return 0.0;
  } else {
#pragma unroll
    for (int k = 0; k < 7; k++) Uout[k * P + i] = un[k];
    if (!want_v2 || (double)un[0] + (double)un[3] <= WET_SPEED) return 0.0;
    const double ufx = (double)qdiv(lim4, un[3] + h0), ufy = (double)qdiv(lim5, un[3] + h0);
    const double sx = (double)usxnew, sy = (double)usynew;
    return fmax(sx * sx + sy * sy, ufx * ufx + ufy * ufy);
  }
}

// Dry-tile skipping.  A stage tile reads cells [r0-2, r0+TY+1] x [c0-2, c0+64] of its
// input (face stencils, the boundary overrides of edge tiles and the viscous term's
// tangential slopes) and its own cells of uzero (and of ONE, which is zero wherever U1
// is).  When all of them are zero, every output is +0 and the flow speed 0.
// plan_tile is called by the PLAN_LANES lanes of one lane group (aligned within a wave,
// lane index q in the group) and returns, in every lane of the group, the tile's plan
// entry: t to compute it, t | PLAN_ZERO (predictor) to clear its U1 and ONE because its U1
// flags say they may hold values of an earlier step, or -1 for nothing to do.  A group
// whose tile does not exist passes t < 0 (it still takes part in the wave's ballots).
constexpr int PLAN_LANES = 2;

template <int IO, typename T, int TY>
__device__ __forceinline__ int plan_tile(const Grid& g, const T* __restrict__ Uin,
                                         const DryMap& dm, int tx, int ty, int t, int q) {
  int wet = 0, own = 0;
  if (t >= 0 && dm.nseg > 0) {
    const int c0 = tx * ST_TX + MD, r0 = ty * TY + MD, ymax = g.ny + MD - 1;
    const DryRegion d = dry_region_of(g, r0 - 2, r0 + TY + 1, c0 - 2, c0 + 64);
    const unsigned char* zin = IO == 0 ? dm.z0 : dm.z1;
    // the interior flags row by row (single-scenario dry step 3-4% faster than the item
    // list), then the corner ghost cells, which only corner tiles have
    wet = dry_flags_wet3<TY + 4, PLAN_LANES>(d, zin, dm.nseg, q);
    for (int e = d.nf + q; e < d.n; e += PLAN_LANES) wet |= dry_item_wet(d, Uin, g, zin, dm.nseg, e);
    const unsigned char* zown = IO == 0 ? dm.z1 : dm.z0;   // flags of the tile's output rows
    const int rown = min(TY, ymax - r0 + 1);
    for (int k = q; k < rown; k += PLAN_LANES)
      own |= zown[(long long)(r0 + k - MD) * dm.nseg + tx];
  }
  const unsigned long long grp = ((1ull << PLAN_LANES) - 1) << (threadIdx.x & ~(PLAN_LANES - 1) & 63);
  const bool anywet = (__ballot(wet) & grp) != 0ull, anyown = (__ballot(own != 0) & grp) != 0ull;
  if (t < 0) return -1;
  if (dm.nseg == 0) return t;
  if (anywet || (IO == 1 && anyown)) return t;   // the corrector also needs uzero zero
  if (IO == 0 && anyown) return t | PLAN_ZERO;
  return -1;
}

// A dry predictor tile whose U1 and ONE may hold values of an earlier step: writes the +0
// that computing it would, and its U1 flags.
template <typename T, int TY>
__device__ __forceinline__ void stage_zero(const Grid& g, T* Uout, T* __restrict__ ONE,
                                           const DryMap& dm, int tx, int ty) {
  const int lane = threadIdx.x, w = threadIdx.y;
  const int r = ty * TY + MD + w, c = tx * ST_TX + MD + lane;
  if (w >= TY || r > g.ny + MD - 1) return;
  if (lane < ST_TX && c <= g.nx + MD - 1) {
    const long long i = idx(g, r, c), P = g.P;
#pragma unroll
    for (int k = 0; k < 7; k++) Uout[k * P + i] = T(0);
    ONE[i] = T(0);
    ONE[P + i] = T(0);
  }
  if (lane == 0) dm.z1[(long long)(r - MD) * dm.nseg + tx] = 0;
}

// The plan entry of a planned stage block: entry blockIdx.x / 8 of list blockIdx.x % 8, or
// -1 past the list's end.
__device__ __forceinline__ int plan_entry(const TilePlan& pl) {
  const int x = blockIdx.x % 8, j = blockIdx.x / 8;
  return j < pl.cnt[x] ? pl.list[x * pl.per + j] : -1;
}

// Appends the plan entries of the lane groups of one wave (e >= 0 in a group's leader lane,
// list x) to the lists with one atomic per list present in the wave (a wave's 32 tiles
// are consecutive, so two or three).  Every lane of the wave calls it.
__device__ __forceinline__ void plan_append(const TilePlan& pl, bool leader, int e, int x) {
  const int lane = threadIdx.x & 63;
  bool pend = leader && e >= 0;
  while (true) {
    const unsigned long long any = __ballot(pend);
    if (!any) return;
    const int first = __ffsll((long long)any) - 1;
    const int xf = __shfl(x, first, 64);
    const bool mine = pend && x == xf;
    const unsigned long long m = __ballot(mine);
    int base = 0;
    if (lane == first) base = atomicAdd(&pl.cnt[xf], __popcll(m));
    base = __shfl(base, first, 64);
    if (mine) {
      pl.list[xf * pl.per + base + __popcll(m & ((1ull << lane) - 1))] = e;
      pend = false;
    }
  }
}

// Tile height and the waves-per-SIMD cap of the single-scenario stage kernel: FP64 is held
// to 128 VGPRs (4 waves); the FP32 body needs ~70-80 VGPRs and runs at 6.  On the
// uniform-metric form the dense FP32 predictor fits 64 VGPRs without spilling and runs at
// 8 (W_PRED), and the dense corrector at 7 (W_CORR, with the compiler's load order, LB off).
// The corrector spills at 8, the stored-metric kernels at 7 or 8 (16 VGPRs: their
// predictor at 8 took 1.9x as long), so those and the planned kernels stay at W.
template <typename T> struct PlainStage;
template <> struct PlainStage<double> { static constexpr int TY = 7, W = 4, W_PRED = 4, W_CORR = 4; };
template <> struct PlainStage<float> { static constexpr int TY = 7, W = 6, W_PRED = 8, W_CORR = 7; };

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).  Block b runs on
// XCD b % 8; it is given tile number (b % 8) * per + b / 8 in row-major tile order, so
// each XCD sweeps a contiguous band of tile rows and vertically adjacent tiles, which
// share halo rows, meet in the same L2.
constexpr int N_XCD = 8;
__host__ __device__ __forceinline__ int stage_tiles(const Grid& g, int TY) {
  return ((g.nx + ST_TX - 1) / ST_TX) * ((g.ny + TY - 1) / TY);
}
__device__ __forceinline__ bool tile_of(const Grid& g, int TY, int t, int& tx, int& ty) {
  const int ntx = (g.nx + ST_TX - 1) / ST_TX;
  if (t >= stage_tiles(g, TY)) return false;
  tx = t % ntx;
  ty = t / ntx;
  return true;
}

// The single-scenario stage's tile height for a grid.  A grid with fewer than
// PLAIN_SMALL_TILES 7-row tiles does not fill the GPU (two 8-wave blocks on each of 256
// CUs); it runs 3-row tiles, twice as many blocks with half the dependent chain each
// (128x128 FP64: 62 -> 43 us per step; 256x256 level; from 512x512 on the 7-row tile, which
// recomputes fewer y-faces, is faster).
constexpr int PLAIN_TY_SMALL = 3, PLAIN_SMALL_TILES = 256;
__host__ __device__ __forceinline__ int plain_ty(const Grid& g, int ty_big) {
  return stage_tiles(g, ty_big) < PLAIN_SMALL_TILES ? PLAIN_TY_SMALL : ty_big;
}

}  // namespace m2p
