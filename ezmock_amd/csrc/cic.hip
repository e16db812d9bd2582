// cic.hip -- cloud-in-cell painting on gfx950 without global atomics.
//
// Replaces the reference's particle loop with eight `omp atomic` adds per
// particle (dens_field.c:233-310, and the momentum painting of
// pop_tracer.c:593-675).  Device-scope floating-point atomics are slow on
// this part, so the painting is organised by Eulerian tile instead:
//
//   1. k_cic_bin<0>   each particle's CIC stencil (base cell = floor of its
//                     displaced grid coordinate, and the +1 neighbours)
//                     touches 1, 2, 4 or 8 tiles; the per-tile list sizes
//                     are counted in LDS over a window of tiles around the
//                     workgroup's particles, with one device atomic per
//                     workgroup and tile;
//   2. scan           list offsets;
//   3. k_cic_bin<1>   each particle index is written into the list of every
//                     tile it touches;
//   4. k_cic_tile     one workgroup per tile accumulates its list in LDS,
//                     keeping the corners inside the tile, and writes the
//                     tile once with plain stores (density and, optionally,
//                     the three momentum fields psi_a * w in the same pass).
// Particle coordinates and weights are formed with the reference's IEEE
// operations (file compiled with -ffp-contract=off); only the summation
// order differs, as it does between the reference's OpenMP threads.
#include "common.h"
#include "kernels.h"

#define CIC_TPB 256
#define CIC_PPT 8
#define CIC_PPB (CIC_TPB * CIC_PPT)
#ifndef CIC_TILE_TPB
#define CIC_TILE_TPB 256
#endif
#ifndef CIC_TILES_PER_WG
#define CIC_TILES_PER_WG 1
#endif

struct ParticleCell {
  real x, y, z;                 // displaced grid coordinates
  int x0, y0, z0;               // base cell (floor), before wrapping
};

// dens_field.c:263-289: grid coordinate of particle idx (Lagrangian cell
// (i, j, k), idx = (i Ng + j) Ng + k) displaced by q / cell size.
__device__ __forceinline__ ParticleCell cell_from(real q0, real q1, real q2,
    uint32_t idx, const CicGeom &g) {
  ParticleCell c;
  uint32_t k, j;
  const uint32_t ij = fastdiv_qr(idx, g.fNg, k);
  const uint32_t i = fastdiv_qr(ij, g.fNg, j);
  c.x = q0 * g.igs + (real) (int) i;
  c.y = q1 * g.igs + (real) (int) j;
  c.z = q2 * g.igs + (real) (int) k;
  c.x0 = (int) floor((double) c.x);
  c.y0 = (int) floor((double) c.y);
  c.z0 = (int) floor((double) c.z);
  return c;
}

__device__ __forceinline__ int wrap_once(int a, int Ng) {
  if (a < 0) a += Ng;
  else if (a >= Ng) a -= Ng;
  // displacements beyond one box length (or non-finite ones) are undefined
  // behaviour in the reference; keep every index inside the grid
  if ((unsigned) a >= (unsigned) Ng) a = ((a % Ng) + Ng) % Ng;
  return a;
}

// Tiles touched by the stencil of a particle: base tile (tx, ty, tz) of the
// wrapped base cell, and the next tile along each axis where the base cell
// is the last of its tile (e = 2).  The n = ex ey ez tiles are enumerated by
// v = (a ey + b) ez + d.
struct TileSpan {
  int tx, ty, tz;
  int ex, ey, ez;
};

__device__ __forceinline__ TileSpan tile_span(const ParticleCell &c, const CicGeom &g) {
  TileSpan s;
  const int x0 = wrap_once(c.x0, g.Ng), y0 = wrap_once(c.y0, g.Ng),
      z0 = wrap_once(c.z0, g.Ng);
  s.tx = x0 >> g.lx; s.ty = y0 >> g.ly; s.tz = z0 >> g.lz;
  s.ex = ((x0 & (g.TX - 1)) == g.TX - 1 && g.ntx > 1) ? 2 : 1;
  s.ey = ((y0 & (g.TY - 1)) == g.TY - 1 && g.nty > 1) ? 2 : 1;
  s.ez = ((z0 & (g.TZ - 1)) == g.TZ - 1 && g.ntz > 1) ? 2 : 1;
  return s;
}

// Tile coordinates of the v-th tile of a span (v < ex ey ez).
__device__ __forceinline__ void span_tile(const TileSpan &s, const CicGeom &g,
    int v, int &ux, int &uy, int &uz) {
  int d = 0, b = 0;
  if (s.ez == 2) { d = v & 1; v >>= 1; }
  if (s.ey == 2) { b = v & 1; v >>= 1; }
  ux = s.tx + v; uy = s.ty + b; uz = s.tz + d;
  if (ux == g.ntx) ux = 0;
  if (uy == g.nty) uy = 0;
  if (uz == g.ntz) uz = 0;
}

__device__ __forceinline__ uint32_t tile_key(const CicGeom &g, int ux, int uy, int uz) {
  return (uint32_t) ((ux * g.nty + uy) * g.ntz + uz);
}

// Local tile window of a workgroup of CIC_PPB consecutive particles: the
// tiles within one tile (in x and y) of the tiles holding the particles'
// Lagrangian positions, times all tiles along z.  Particles displaced out of
// the window (several tile widths) take the global-atomic path instead.
#define WMAX 2048

struct Window {
  int direct;                   // whole grid fits: slot = key
  int txlo, tylo, wx, wy;       // window origin and extent in tiles
};

__device__ __forceinline__ Window make_window(const CicGeom &g, uint32_t b0) {
  Window w;
  const size_t ntile = (size_t) g.ntx * g.nty * g.ntz;
  w.direct = ntile <= WMAX;
  const uint32_t pl = (uint32_t) (((size_t) b0 + CIC_PPB < g.N) ? (size_t) b0 + CIC_PPB
                                                                 : g.N) - 1;
  uint32_t r;
  const uint32_t r0 = fastdiv_qr(b0, g.fNg, r), r1 = fastdiv_qr(pl, g.fNg, r);
  uint32_t j0u, j1u;
  const int i0 = (int) fastdiv_qr(r0, g.fNg, j0u), i1 = (int) fastdiv_qr(r1, g.fNg, j1u);
  int j0 = 0, j1 = g.Ng - 1;
  if (i0 == i1) { j0 = (int) j0u; j1 = (int) j1u; }
  w.wx = (i1 >> g.lx) - (i0 >> g.lx) + 3;
  w.wy = (j1 >> g.ly) - (j0 >> g.ly) + 3;
  if (w.wx > g.ntx) w.wx = g.ntx;
  if (w.wy > g.nty) w.wy = g.nty;
  w.txlo = (i0 >> g.lx) - 1; if (w.txlo < 0) w.txlo += g.ntx;
  w.tylo = (j0 >> g.ly) - 1; if (w.tylo < 0) w.tylo += g.nty;
  if (!w.direct && (size_t) w.wx * w.wy * g.ntz > WMAX) {
    w.wx = 1; w.wy = 1;         // degenerate shapes: mostly global path
  }
  return w;
}

// Window slot of tile (ux, uy, uz), or -1 outside the window.
__device__ __forceinline__ int window_slot(const Window &w, const CicGeom &g,
    int ux, int uy, int uz) {
  if (w.direct) return (int) tile_key(g, ux, uy, uz);
  int dx = ux - w.txlo, dy = uy - w.tylo;
  if (dx < 0) dx += g.ntx;
  if (dy < 0) dy += g.nty;
  if (dx >= w.wx || dy >= w.wy) return -1;
  return (dx * w.wy + dy) * g.ntz + uz;
}

__device__ __forceinline__ uint32_t slot_key(const Window &w, const CicGeom &g,
    int slot) {
  if (w.direct) return (uint32_t) slot;
  uint32_t uz;
  const uint32_t q = fastdiv_qr((uint32_t) slot, g.fntz, uz);
  const int dy = (int) (q % (uint32_t) w.wy), dx = (int) (q / (uint32_t) w.wy);
  int ux = w.txlo + dx, uy = w.tylo + dy;
  if (ux >= g.ntx) ux -= g.ntx;
  if (uy >= g.nty) uy -= g.nty;
  return tile_key(g, ux, uy, (int) uz);
}

// Add 1 per active lane to wcnt[slot] with one LDS atomic for the lanes that
// share the first active lane's slot (the common case: a wave's particles
// fall in the same tile) and one per lane otherwise.  RET: return each
// lane's previous count (its rank).
template <bool RET>
__device__ __forceinline__ uint32_t wave_count_add(uint32_t *wcnt, bool active,
    int slot) {
  const int lane = threadIdx.x & 63;
  const uint64_t act = __ballot(active);
  if (!act) return 0;
  const int leader = __ffsll((unsigned long long) act) - 1;
  const int lslot = __shfl(slot, leader, 64);
  const uint64_t peers = __ballot(active && slot == lslot);
  uint32_t base = 0;
  if (lane == leader) {
    if (RET) base = atomicAdd(&wcnt[lslot], (uint32_t) __popcll(peers));
    else atomicAdd(&wcnt[lslot], (uint32_t) __popcll(peers));
  }
  if ((peers >> lane) & 1) {
    if (!RET) return 0;
    base = __shfl(base, leader, 64);
    return base + (uint32_t) __popcll(peers & ((1ull << lane) - 1));
  }
  if (!active) return 0;
  if (RET) return atomicAdd(&wcnt[slot], 1u);
  atomicAdd(&wcnt[slot], 1u);
  return 0;
}

// The particle's tile span (n = 0 for indices past N).
__device__ __forceinline__ int particle_span(const real *q0, const real *q1,
    const real *q2, int u, uint32_t p, const CicGeom &g, TileSpan &s) {
  if ((size_t) p >= g.N) return 0;
  const ParticleCell c = cell_from(q0[u], q1[u], q2[u], p, g);
  s = tile_span(c, g);
  return s.ex * s.ey * s.ez;
}

template <int SCATTER>
__global__ void __launch_bounds__(CIC_TPB)
k_cic_bin(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, CicGeom g, uint32_t *__restrict__ counts,
    unsigned long long *__restrict__ cursor, uint32_t *__restrict__ list) {
  __shared__ uint32_t wcnt[WMAX];
  __shared__ unsigned long long wbase[SCATTER ? WMAX : 1];
  const uint32_t b0 = blockIdx.x * CIC_PPB;
  const Window w = make_window(g, b0);
  const int nslot = w.direct ? (int) ((size_t) g.ntx * g.nty * g.ntz)
                             : w.wx * w.wy * g.ntz;
  for (int s = threadIdx.x; s < nslot; s += CIC_TPB) wcnt[s] = 0;
  __syncthreads();
  // all loads of the block's particles first (independent, in flight
  // together), then the per-particle work
  real q[3][CIC_PPT];
#pragma unroll
  for (int u = 0; u < CIC_PPT; u++) {
    const size_t p = (size_t) b0 + (size_t) u * CIC_TPB + threadIdx.x;
    const bool ok = p < g.N;
    q[0][u] = ok ? p0[p] : (real) 0;
    q[1][u] = ok ? p1[p] : (real) 0;
    q[2][u] = ok ? p2[p] : (real) 0;
  }
  // pass 1: count the window's tiles
#pragma unroll
  for (int u = 0; u < CIC_PPT; u++) {
    const uint32_t p = b0 + (uint32_t) (u * CIC_TPB) + threadIdx.x;
    TileSpan sp;
    const int n = particle_span(q[0], q[1], q[2], u, p, g, sp);
    for (int v = 0; __ballot(v < n); v++) {
      int s = -1;
      if (v < n) {
        int ux, uy, uz;
        span_tile(sp, g, v, ux, uy, uz);
        s = window_slot(w, g, ux, uy, uz);
        if (!SCATTER && s < 0) atomicAdd(&counts[tile_key(g, ux, uy, uz)], 1u);
      }
      (void) wave_count_add<false>(wcnt, s >= 0, s);
    }
  }
  __syncthreads();
  if (!SCATTER) {
    for (int s = threadIdx.x; s < nslot; s += CIC_TPB)
      if (wcnt[s]) atomicAdd(&counts[slot_key(w, g, s)], wcnt[s]);
    return;
  }
  // scatter: reserve each window tile's range, then place the indices
  for (int s = threadIdx.x; s < nslot; s += CIC_TPB) {
    if (wcnt[s])
      wbase[s] = atomicAdd(&cursor[slot_key(w, g, s)], (unsigned long long) wcnt[s]);
    wcnt[s] = 0;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < CIC_PPT; u++) {
    const uint32_t p = b0 + (uint32_t) (u * CIC_TPB) + threadIdx.x;
    TileSpan sp;
    const int n = particle_span(q[0], q[1], q[2], u, p, g, sp);
    for (int v = 0; __ballot(v < n); v++) {
      int s = -1;
      int ux = 0, uy = 0, uz = 0;
      if (v < n) {
        span_tile(sp, g, v, ux, uy, uz);
        s = window_slot(w, g, ux, uy, uz);
      }
      const uint32_t r = wave_count_add<true>(wcnt, s >= 0, s);
      if (v < n) {
        unsigned long long pos;
        if (s >= 0) pos = wbase[s] + r;
        else pos = atomicAdd(&cursor[tile_key(g, ux, uy, uz)], 1ull);
        list[pos] = p;
      }
    }
  }
}

// One workgroup per tile.  MOM: also the three momentum fields, stored
// normalised as the velocity gather uses them (pop_tracer.c:676-682): per
// cell {rho_c, u0, u1, u2} with u_a = m_a / rho_c where rho_c > REAL_TOL
// (the same IEEE division the gather would do per corner).
template <int MOM, int TXX, int TYY, int TZZ>
__global__ void __launch_bounds__(CIC_TILE_TPB)
k_cic_tile(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, CicGeom g, const uint64_t *__restrict__ off,
    const uint32_t *__restrict__ list, real *__restrict__ rho,
    const real *__restrict__ rho_norm, real4 *__restrict__ vel, int tpw) {
  constexpr int TC = TXX * TYY * TZZ;
  constexpr int NF = MOM ? 4 : 1;
  // double accumulators: LDS double atomics are several times faster than
  // float ones on gfx950, and the float result is the correctly rounded
  // sum of the float contributions
  __shared__ double acc[NF][TC];
  // tpw consecutive tiles (z-neighbours, whose particle rows share cache
  // lines) per workgroup, the workgroups of one XCD on consecutive groups
  const int ntile = g.ntx * g.nty * g.ntz;
  const int t0 = (int) xcd_item(blockIdx.x, gridDim.x) * tpw;
  const int t1 = (t0 + tpw < ntile) ? t0 + tpw : ntile;
  for (int t = t0; t < t1; t++) {
  for (int c = threadIdx.x; c < NF * TC; c += CIC_TILE_TPB) (&acc[0][0])[c] = 0;
  __syncthreads();
  const int tz = t % g.ntz, ty = (t / g.ntz) % g.nty, tx = t / (g.ntz * g.nty);
  const int Ng = g.Ng;
  const int ox = tx * TXX, oy = ty * TYY, oz = tz * TZZ;
  const uint64_t s0 = off[t], e0 = off[t + 1];
  constexpr int U = 4;
  for (uint64_t qb = s0; qb < e0; qb += (uint64_t) U * CIC_TILE_TPB) {
  size_t pv[U];
  real qq[U][3];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const uint64_t q = qb + (uint64_t) u * CIC_TILE_TPB + threadIdx.x;
    pv[u] = (q < e0) ? (size_t) list[q] : (size_t) -1;
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const bool ok = pv[u] != (size_t) -1;
    qq[u][0] = ok ? p0[pv[u]] : (real) 0;
    qq[u][1] = ok ? p1[pv[u]] : (real) 0;
    qq[u][2] = ok ? p2[pv[u]] : (real) 0;
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    if (pv[u] == (size_t) -1) continue;
    const size_t p = pv[u];
    const ParticleCell c = cell_from(qq[u][0], qq[u][1], qq[u][2], (uint32_t) p, g);
    const real wx1 = c.x - (real) c.x0, wx0 = 1 - wx1;
    const real wy1 = c.y - (real) c.y0, wy0 = 1 - wy1;
    const real wz1 = c.z - (real) c.z0, wz0 = 1 - wz1;
    const int x0 = wrap_once(c.x0, Ng), y0 = wrap_once(c.y0, Ng),
        z0 = wrap_once(c.z0, Ng);
    const int x1 = (x0 == Ng - 1) ? 0 : x0 + 1;
    const int y1 = (y0 == Ng - 1) ? 0 : y0 + 1;
    const int z1 = (z0 == Ng - 1) ? 0 : z0 + 1;
    // local coordinates inside this tile (outside -> -1)
    const int lx[2] = {(x0 - ox >= 0 && x0 - ox < TXX) ? x0 - ox : -1,
                       (x1 - ox >= 0 && x1 - ox < TXX) ? x1 - ox : -1};
    const int ly[2] = {(y0 - oy >= 0 && y0 - oy < TYY) ? y0 - oy : -1,
                       (y1 - oy >= 0 && y1 - oy < TYY) ? y1 - oy : -1};
    const int lz[2] = {(z0 - oz >= 0 && z0 - oz < TZZ) ? z0 - oz : -1,
                       (z1 - oz >= 0 && z1 - oz < TZZ) ? z1 - oz : -1};
    const real wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1}, wz[2] = {wz0, wz1};
    real qv[3] = {0, 0, 0};
    if (MOM) { qv[0] = qq[u][0]; qv[1] = qq[u][1]; qv[2] = qq[u][2]; }
    // the eight corners, in an order rotated by lane: lanes whose particles
    // share a cell (dense regions) then add to different cells in the same
    // LDS instruction instead of serialising on one address
    const int rot = (int) (threadIdx.x & 7);
#pragma unroll
    for (int ci = 0; ci < 8; ci++) {
      const int cc = (ci + rot) & 7;
      const int a = cc >> 2, b = (cc >> 1) & 1, d = cc & 1;
      const int la = a ? lx[1] : lx[0], lb = b ? ly[1] : ly[0], ld = d ? lz[1] : lz[0];
      if (la < 0 || lb < 0 || ld < 0) continue;
      const real xa = a ? wx[1] : wx[0], yb = b ? wy[1] : wy[0], zd = d ? wz[1] : wz[0];
      const int li = (la * TYY + lb) * TZZ + ld;
      atomicAdd(&acc[0][li], (double) (xa * yb * zd));
      if (MOM) {
#pragma unroll
        for (int f = 0; f < 3; f++) {
          const real u = xa * qv[f];                // pop_tracer.c:641-642
          atomicAdd(&acc[1 + f][li], (double) (u * yb * zd));
        }
      }
    }
  }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < TC; c += CIC_TILE_TPB) {
    const int lz = c % TZZ, ly = (c / TZZ) % TYY, lx = c / (TZZ * TYY);
    const size_t gi = ((size_t) (ox + lx) * Ng + (oy + ly)) * Ng + (oz + lz);
    const real r = (real) acc[0][c];
    if (rho) rho[gi] = r;
    if (MOM) {
      const real rc = rho ? r : rho_norm[gi];
      real u[3];
#pragma unroll
      for (int f = 0; f < 3; f++) {
        const real mv = (real) acc[1 + f][c];
        u[f] = ((double) rc > EZ_REAL_TOL) ? mv / rc : mv;
      }
      vel[gi] = real4{rc, u[0], u[1], u[2]};
    }
  }
  __syncthreads();                      // before the next tile's zeroing
  }
}

__global__ void k_cursor_init(uint64_t *__restrict__ off,
    uint64_t *__restrict__ cursor, size_t nk, const uint64_t *__restrict__ total) {
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nk) cursor[i] = off[i];
  else if (i == nk) off[nk] = *total;
}

// Tile shape: 8 x 8 x 16 cells when the grid allows it (32 KB of double
// accumulators for density + momentum: five workgroups per CU).
static int tile_dim(int Ng, int pref) {
  int d = pref;
  while (d > 1 && Ng % d) d >>= 1;
  return d;
}

CicGeom cic_geometry(int Ng, real igs) {
  CicGeom g;
  g.Ng = Ng;
  g.igs = igs;
  g.TX = tile_dim(Ng, 8);
  g.TY = tile_dim(Ng, 8);
  g.TZ = tile_dim(Ng, 16);
  g.lx = __builtin_ctz(g.TX); g.ly = __builtin_ctz(g.TY); g.lz = __builtin_ctz(g.TZ);
  g.ntx = Ng / g.TX; g.nty = Ng / g.TY; g.ntz = Ng / g.TZ;
  g.N = (size_t) Ng * Ng * Ng;
  g.fNg = fastdiv_make((uint32_t) Ng);
  g.fntz = fastdiv_make((uint32_t) g.ntz);
  return g;
}

size_t cic_nkeys(const CicGeom &g) { return (size_t) g.ntx * g.nty * g.ntz; }

template <int MOM, int A, int B, int C>
static void launch_tile(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    const real *rho_norm, real4 *vel, hipStream_t st) {
  const unsigned nt = (unsigned) (g.ntx * g.nty * g.ntz);
  const int tpw = CIC_TILES_PER_WG;
  hipLaunchKernelGGL((k_cic_tile<MOM, A, B, C>), dim3((nt + tpw - 1) / tpw),
      dim3(CIC_TILE_TPB), 0, st, p0, p1, p2, g, off, list, rho, rho_norm, vel, tpw);
}

template <int MOM>
static hipError_t dispatch_tile(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    const real *rho_norm, real4 *vel, hipStream_t st) {
  if (g.TX == 8 && g.TY == 8 && g.TZ == 32)
    launch_tile<MOM, 8, 8, 32>(p0, p1, p2, g, off, list, rho, rho_norm, vel, st);
  else if (g.TX == 8 && g.TY == 8 && g.TZ == 16)
    launch_tile<MOM, 8, 8, 16>(p0, p1, p2, g, off, list, rho, rho_norm, vel, st);
  else if (g.TX == 8 && g.TY == 8 && g.TZ == 8)
    launch_tile<MOM, 8, 8, 8>(p0, p1, p2, g, off, list, rho, rho_norm, vel, st);
  else if (g.TX == 4 && g.TY == 4 && g.TZ == 4)
    launch_tile<MOM, 4, 4, 4>(p0, p1, p2, g, off, list, rho, rho_norm, vel, st);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// Binning: count, scan and cursor set-up (counts: nkeys u32; off: nkeys + 1
// u64, off[nkeys] = total list length; cursor: nkeys u64; scan_tmp:
// scan_tmp_words(nkeys) + 8 u64).
hipError_t launch_cic_count(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, uint32_t *counts, uint64_t *off, uint64_t *cursor,
    uint64_t *scan_tmp, hipStream_t st) {
  if (g.N >= 0xffffffffull) return hipErrorInvalidValue;
  const size_t nk = cic_nkeys(g);
  hipError_t e;
  if ((e = launch_zero(counts, nk * sizeof(uint32_t), st)) != hipSuccess) return e;
  const unsigned nb = (unsigned) ((g.N + CIC_PPB - 1) / CIC_PPB);
  hipLaunchKernelGGL(k_cic_bin<0>, dim3(nb), dim3(CIC_TPB), 0, st, p0, p1, p2, g,
      counts, (unsigned long long *) nullptr, (uint32_t *) nullptr);
  if ((e = launch_scan_u32(counts, off, nk, scan_tmp, st)) != hipSuccess) return e;
  // off[nk] = total list length (end of the last list); cursor = off
  const size_t nch = (nk + EZ_CELLS_PER_BLOCK - 1) / EZ_CELLS_PER_BLOCK;
  hipLaunchKernelGGL(k_cursor_init, dim3(ez_blocks(nk + 1, 256)), dim3(256), 0, st,
      off, cursor, nk, (const uint64_t *) (scan_tmp + nch));
  return hipGetLastError();
}

// Particle indices into the lists of the tiles they touch (list: off[nkeys]
// u32, between N and 8 N; about 1.35 N for 8 x 8 x 16 tiles).
hipError_t launch_cic_scatter(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, uint64_t *cursor, uint32_t *list, hipStream_t st) {
  const unsigned nb = (unsigned) ((g.N + CIC_PPB - 1) / CIC_PPB);
  hipLaunchKernelGGL(k_cic_bin<1>, dim3(nb), dim3(CIC_TPB), 0, st, p0, p1, p2, g,
      (uint32_t *) nullptr, (unsigned long long *) cursor, list);
  return hipGetLastError();
}

// Tile accumulation: density into rho (unless null) and, when vel is not
// null, the normalised momentum fields.
hipError_t launch_cic_tiles(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    const real *rho_norm, real4 *vel, hipStream_t st) {
  if (vel && !rho && !rho_norm) return hipErrorInvalidValue;
  if (vel) return dispatch_tile<1>(p0, p1, p2, g, off, list, rho, rho_norm, vel, st);
  return dispatch_tile<0>(p0, p1, p2, g, off, list, rho, rho_norm, vel, st);
}
