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
__device__ __forceinline__ ParticleCell cell_at(real q0, real q1, real q2,
    uint32_t i, uint32_t j, uint32_t k, const CicGeom &g) {
  ParticleCell c;
  c.x = q0 * g.igs + (real) (int) i;
  c.y = q1 * g.igs + (real) (int) j;
  c.z = q2 * g.igs + (real) (int) k;
#ifdef SINGLE_PREC
  c.x0 = (int) floorf(c.x);               // = floor of the value in double
  c.y0 = (int) floorf(c.y);
  c.z0 = (int) floorf(c.z);
#else
  c.x0 = (int) floor(c.x);
  c.y0 = (int) floor(c.y);
  c.z0 = (int) floor(c.z);
#endif
  return c;
}

__device__ __forceinline__ ParticleCell cell_from(real q0, real q1, real q2,
    uint32_t idx, const CicGeom &g) {
  uint32_t k, j;
  const uint32_t ij = fastdiv_qr(idx, g.fNg, k);
  const uint32_t i = fastdiv_qr(ij, g.fNg, j);
  return cell_at(q0, q1, q2, i, j, k, g);
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

// (tile counts and window extents are below 2^24: full-rate 24-bit products)
__device__ __forceinline__ uint32_t tile_key(const CicGeom &g, int ux, int uy, int uz) {
  return __umul24(__umul24((uint32_t) ux, (uint32_t) g.nty) + (uint32_t) uy,
                  (uint32_t) g.ntz) + (uint32_t) uz;
}

// Local tile window of a workgroup of CIC_PPB consecutive particles: the
// tiles within one tile (in x and y) of the tiles holding the particles'
// Lagrangian positions, times all tiles along z.  Particles displaced out of
// the window (several tile widths) take the global-atomic path instead.
// The window holds up to WM tiles: 1024 for Ng <= 1536 (3 x 3 x 96 tiles at
// most), 2048 above; the smaller LDS arrays leave room for one more
// workgroup per CU (512^3 scatter: 1.74 -> 1.61 ms).

struct Window {
  int direct;                   // whole grid fits: slot = key
  int txlo, tylo, wx, wy;       // window origin and extent in tiles
};

template <int WM>
__device__ __forceinline__ Window make_window(const CicGeom &g, uint32_t b0) {
  Window w;
  const size_t ntile = (size_t) g.ntx * g.nty * g.ntz;
  w.direct = ntile <= WM;
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
  if (!w.direct && (size_t) w.wx * w.wy * g.ntz > WM) {
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
  return (int) (__umul24(__umul24((uint32_t) dx, (uint32_t) w.wy) + (uint32_t) dy,
                         (uint32_t) g.ntz) + (uint32_t) uz);
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

// The tile span of particle p (n = 0 for indices past N).  Its lattice
// coordinates are divided out again in each pass rather than kept: 24 fewer
// registers per thread, one more workgroup per CU (scatter 1.64 -> 1.58 ms
// at 512^3).
__device__ __forceinline__ int particle_span(const real *q0, const real *q1,
    const real *q2, int u, uint32_t p, const CicGeom &g, TileSpan &s) {
  if ((size_t) p >= g.N) return 0;
  const ParticleCell c = cell_from(q0[u], q1[u], q2[u], p, g);
  s = tile_span(c, g);
  return s.ex * s.ey * s.ez;
}

// SCATTER with cap > 0: tile t owns list[t cap, (t + 1) cap) (cursor
// starts at t cap); entries past the end are dropped and *ovf set.  In that
// mode wbase holds (tile key << 36) | position.
template <int SCATTER, int WM>
__global__ void __launch_bounds__(CIC_TPB)
k_cic_bin(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, CicGeom g, uint32_t *__restrict__ counts,
    unsigned long long *__restrict__ cursor, uint32_t *__restrict__ list, uint32_t cap,
    uint32_t *__restrict__ ovf) {
  __shared__ uint32_t wcnt[WM];
  __shared__ unsigned long long wbase[SCATTER ? WM : 1];
  const uint32_t b0 = blockIdx.x * CIC_PPB;
  const Window w = make_window<WM>(g, b0);
  const int nslot = w.direct ? (int) ((size_t) g.ntx * g.nty * g.ntz)
                             : w.wx * w.wy * g.ntz;
  for (int s = threadIdx.x; s < nslot; s += CIC_TPB) wcnt[s] = 0;
  __syncthreads();
  // all loads of the block's particles first (independent, in flight
  // together), then the per-particle work
  // (unconditional loads at clamped indices: one wait for all of them)
  real q[3][CIC_PPT];
#pragma unroll
  for (int u = 0; u < CIC_PPT; u++) {
    size_t p = (size_t) b0 + (size_t) u * CIC_TPB + threadIdx.x;
    p = p < g.N ? p : g.N - 1;
    q[0][u] = p0[p];
    q[1][u] = p1[p];
    q[2][u] = p2[p];
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
    if (wcnt[s]) {
      const uint32_t key = slot_key(w, g, s);
      const unsigned long long b = atomicAdd(&cursor[key], (unsigned long long) wcnt[s]);
      wbase[s] = cap ? (((unsigned long long) key << 36) | b) : b;
    }
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
        unsigned long long pos, key;
        if (s >= 0) {
          const unsigned long long b = wbase[s];
          key = b >> 36;
          pos = (cap ? (b & ((1ull << 36) - 1)) : b) + r;
        }
        else {
          key = tile_key(g, ux, uy, uz);
          pos = atomicAdd(&cursor[key], 1ull);
        }
        if (!cap || pos < (key + 1) * cap) list[pos] = p;
        else *ovf = 1u;
      }
    }
  }
}

// Whether the cloud of particle c reaches the tile at (ox, oy, oz).
template <int TXX, int TYY, int TZZ>
__device__ __forceinline__ bool cic_touches(const ParticleCell &c, int ox, int oy,
    int oz, int Ng) {
  const int x0 = wrap_once(c.x0, Ng), y0 = wrap_once(c.y0, Ng),
      z0 = wrap_once(c.z0, Ng);
  const int x1 = (x0 == Ng - 1) ? 0 : x0 + 1;
  const int y1 = (y0 == Ng - 1) ? 0 : y0 + 1;
  const int z1 = (z0 == Ng - 1) ? 0 : z0 + 1;
  const bool inx = (unsigned) (x0 - ox) < (unsigned) TXX || (unsigned) (x1 - ox) < (unsigned) TXX;
  const bool iny = (unsigned) (y0 - oy) < (unsigned) TYY || (unsigned) (y1 - oy) < (unsigned) TYY;
  const bool inz = (unsigned) (z0 - oz) < (unsigned) TZZ || (unsigned) (z1 - oz) < (unsigned) TZZ;
  return inx && iny && inz;
}

// Adds the corners of particle c (displacement q) that fall inside the
// tile at (ox, oy, oz) to the LDS accumulators acc[f * TC + cell]
// (pop_tracer.c:616-650, dens_field.c:290-308): weights and momenta are
// formed in `real` in the reference's order, then accumulated in double.
template <int MOM, int TXX, int TYY, int TZZ>
__device__ __forceinline__ void cic_deposit(double *acc, const ParticleCell &c,
    real q0, real q1, real q2, int ox, int oy, int oz, int Ng) {
  constexpr int TC = TXX * TYY * TZZ;
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
  if ((lx[0] < 0 && lx[1] < 0) || (ly[0] < 0 && ly[1] < 0) || (lz[0] < 0 && lz[1] < 0))
    return;
  const real wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1}, wz[2] = {wz0, wz1};
  const real qv[3] = {q0, q1, q2};
  // the eight corners in an order permuted by lane: the two low lane bits
  // flip the corner's x and y sides, so that lanes whose particles share a
  // cell (dense regions) add to different cells in the same LDS instruction
  // instead of serialising on one address, while every lane of an
  // instruction takes the same z side (z-neighbouring cells are
  // bank-neighbours, and a lane-dependent z side puts pairs of lanes of
  // consecutive particles on one bank): tiles 3.28 -> 3.19 ms at 512^3
  // against rotating all three sides by lane, 3.47-3.55 ms with z flipped
  // by lane or no permutation
  const int rot = (int) ((threadIdx.x & 3) << 1);
#pragma unroll
  for (int ci = 0; ci < 8; ci++) {
    const int cc = ci ^ rot;
    const int a = cc >> 2, b = (cc >> 1) & 1, d = cc & 1;
    const int la = a ? lx[1] : lx[0], lb = b ? ly[1] : ly[0], ld = d ? lz[1] : lz[0];
    if (la < 0 || lb < 0 || ld < 0) continue;
    const real xa = a ? wx[1] : wx[0], yb = b ? wy[1] : wy[0], zd = d ? wz[1] : wz[0];
    const int li = (la * TYY + lb) * TZZ + ld;
    atomicAdd(&acc[li], (double) (xa * yb * zd));
    if (MOM) {
#pragma unroll
      for (int f = 0; f < 3; f++) {
        const real u = xa * qv[f];                // pop_tracer.c:641-642
        atomicAdd(&acc[(1 + f) * TC + li], (double) (u * yb * zd));
      }
    }
  }
}

// Writes an accumulated tile.  MOM: also the three momentum fields, stored
// normalised as the velocity gather uses them (pop_tracer.c:676-682): per
// cell {rho_c, u0, u1, u2} with u_a = m_a / rho_c where rho_c > REAL_TOL
// (the same IEEE division the gather would do per corner).
template <int MOM, int TXX, int TYY, int TZZ, int NTH>
__device__ __forceinline__ void cic_store(const double *acc, int ox, int oy, int oz,
    int Ng, real *__restrict__ rho, const real *__restrict__ rho_norm,
    real4 *__restrict__ vel) {
  constexpr int TC = TXX * TYY * TZZ;
  // field offsets from the tile origin fit 32 bits (lx < 8, Ng < 1626);
  // the momentum field is in 2 x 2 x 2 bricks (vel_brick_index): a wave's
  // 64 cells (one lx, four rows of 16 along z in 8 x 8 x 16 tiles) fill 16
  // half-bricks of 64 contiguous bytes (FP32), the other halves coming from
  // the wave of the next lx.  Accumulators
  // are read in cell order (consecutive lanes, consecutive LDS words):
  // reading them in brick order to write whole bricks doubled the LDS
  // conflicts and cost the tile pass 3% at 512^3.
  const size_t o0 = ((size_t) ox * Ng + oy) * Ng + oz;
  const uint32_t ng = (uint32_t) Ng, ng2 = ng * ng, h = ng >> 1, hh = h * h;
  real4 *__restrict__ vb = MOM ? vel + vel_brick_index(ox, oy, oz, Ng) : nullptr;
  for (int c = threadIdx.x; c < TC; c += NTH) {
    const uint32_t lz = c % TZZ, ly = (c / TZZ) % TYY, lx = c / (TZZ * TYY);
    const uint32_t go = lx * ng2 + ly * ng + lz;
    const real r = (real) acc[c];
    if (rho) rho[o0 + go] = r;
    if (MOM) {
      const real rc = rho ? r : rho_norm[o0 + go];
      real u[3];
#pragma unroll
      for (int f = 0; f < 3; f++) {
        const real mv = (real) acc[(1 + f) * TC + c];
        u[f] = ((double) rc > EZ_REAL_TOL) ? mv / rc : mv;
      }
      const uint32_t slot = ((((lx >> 1) * hh + (ly >> 1) * h + (lz >> 1)) << 3)
                             | ((lx & 1) << 2) | ((ly & 1) << 1) | (lz & 1));
      vb[slot] = real4{rc, u[0], u[1], u[2]};
    }
  }
}

// One workgroup per tile: the tile's list of particle indices.
template <int MOM, int TXX, int TYY, int TZZ>
__global__ void __launch_bounds__(CIC_TILE_TPB)
k_cic_tile(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, CicGeom g, const uint64_t *__restrict__ off,
    const uint32_t *__restrict__ list, real *__restrict__ rho,
    const real *__restrict__ rho_norm, real4 *__restrict__ vel, int tpw, uint32_t cap) {
  constexpr int TC = TXX * TYY * TZZ;
  constexpr int NF = MOM ? 4 : 1;
  // double accumulators: LDS double atomics are several times faster than
  // float ones on gfx950, and the float result is the correctly rounded
  // sum of the float contributions
  __shared__ double acc[NF * TC];
  // tpw consecutive tiles (z-neighbours, whose particle rows share cache
  // lines) per workgroup, the workgroups of one XCD on consecutive groups
  const int ntile = g.ntx * g.nty * g.ntz;
  const int t0 = (int) xcd_item(blockIdx.x, gridDim.x) * tpw;
  const int t1 = (t0 + tpw < ntile) ? t0 + tpw : ntile;
  for (int t = t0; t < t1; t++) {
  for (int c = threadIdx.x; c < NF * TC; c += CIC_TILE_TPB) acc[c] = 0;
  __syncthreads();
  const int tz = t % g.ntz, ty = (t / g.ntz) % g.nty, tx = t / (g.ntz * g.nty);
  const int ox = tx * TXX, oy = ty * TYY, oz = tz * TZZ;
  // list range: [off[t], off[t + 1]), or with fixed ranges [t cap,
  // min(cursor[t], (t + 1) cap)) where off is the cursor array
  uint64_t s0, e0;
  if (cap) {
    s0 = (uint64_t) t * cap;
    e0 = off[t] < s0 + cap ? off[t] : s0 + cap;
  }
  else {
    s0 = off[t];
    e0 = off[t + 1];
  }
  constexpr int U = 4;
  for (uint64_t qb = s0; qb < e0; qb += (uint64_t) U * CIC_TILE_TPB) {
    size_t pv[U];
    real qq[U][3];
    // unconditional loads at clamped indices, so that all of them are in
    // flight together (a conditional load waits at its join)
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t q = qb + (uint64_t) u * CIC_TILE_TPB + threadIdx.x;
      const size_t v = (size_t) list[q < e0 ? q : e0 - 1];
      pv[u] = (q < e0) ? v : (size_t) -1;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t p = pv[u] != (size_t) -1 ? pv[u] : 0;
      qq[u][0] = p0[p];
      qq[u][1] = p1[p];
      qq[u][2] = p2[p];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (pv[u] == (size_t) -1) continue;
      const ParticleCell c = cell_from(qq[u][0], qq[u][1], qq[u][2], (uint32_t) pv[u], g);
      cic_deposit<MOM, TXX, TYY, TZZ>(acc, c, qq[u][0], qq[u][1], qq[u][2], ox, oy, oz,
                                       g.Ng);
    }
  }
  __syncthreads();
  cic_store<MOM, TXX, TYY, TZZ, CIC_TILE_TPB>(acc, ox, oy, oz, g.Ng, rho, rho_norm, vel);
  __syncthreads();                      // before the next tile's zeroing
  }
}

// ------------------------------------------------------------ tile scan
// Binning-free variant of steps 1-4 for displacement fields of bounded
// range.  A first pass records, per bounds block of BX x BY x BZ Lagrangian
// lattice points, the extrema of d_a = psi_a * igs (the displacement in
// cells, formed as the particle coordinate is).  The workgroup of tile T
// then visits, in every bounds block near T, only the lattice points i with
//   T.lo - 1 - floor(max d + s) <= i <= T.hi - floor(min d - s)
// along each axis, which contains every particle whose cloud reaches T
// (x = fl(d + i) differs from d + i by less than s = 2^-8 for |x| < 2^16),
// tests each exactly and accumulates those that reach T.  Candidate ranges
// are unwrapped coordinates within one box length, so no particle is
// visited twice; displacement ranges of a box length or more, and non-finite
// displacements, take the binning path.  On the 512^3 mock (2.45 candidates
// per particle with 4 x 4 x 8 bounds blocks) the scan takes 6.3 ms against
// 5.4 ms for count + scatter + tiles: the candidate tests cost as much as
// the binning passes, so binning is the default and the scan an option that
// needs no particle lists (and no host synchronisation to size them).
#define CIC_SCAN_TPB 256
#define CIC_SCAN_CH 256               // candidate blocks per round (<= TPB)
#define CIC_SCAN_U 4
#define CIC_SCAN_SLACK (1.0 / 256)

// Order-preserving unsigned keys of reals (extrema by integer atomics).
__device__ __forceinline__ rkey real_okey(real v) {
#ifdef SINGLE_PREC
  const uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
#else
  const uint64_t b = (uint64_t) __double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
#endif
}

__device__ __forceinline__ real real_from_okey(rkey k) {
#ifdef SINGLE_PREC
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
#else
  return __longlong_as_double((long long) ((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
#endif
}

template <int BZZ>
__device__ __forceinline__ void load_run(const real *__restrict__ p, real *v) {
#ifdef SINGLE_PREC
  if (BZZ % 4 == 0) {
#pragma unroll
    for (int k = 0; k < BZZ; k += 4) {
      const float4 t = *(const float4 *) (p + k);
      v[k] = t.x; v[k + 1] = t.y; v[k + 2] = t.z; v[k + 3] = t.w;
    }
    return;
  }
#else
  if (BZZ % 2 == 0) {
#pragma unroll
    for (int k = 0; k < BZZ; k += 2) {
      const double2 t = *(const double2 *) (p + k);
      v[k] = t.x; v[k + 1] = t.y;
    }
    return;
  }
#endif
#pragma unroll
  for (int k = 0; k < BZZ; k++) v[k] = p[k];
}

__global__ void k_cic_bounds_init(rkey *gkey) {
  const int i = (int) threadIdx.x;
  if (i < CIC_GKEY_WORDS) gkey[i] = (i < 3) ? ~(rkey) 0 : (rkey) 0;
}

// One thread per bounds block (consecutive threads: consecutive blocks
// along z, so each load instruction reads one contiguous run).
template <int BZZ>
__global__ void __launch_bounds__(256)
k_cic_bounds(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, CicGeom g, real4 *__restrict__ bnd,
    rkey *__restrict__ gkey) {
  __shared__ real wlo[3][4], whi[3][4];
  __shared__ int wbad[4];
  const uint32_t nblk = (uint32_t) g.nbx * (uint32_t) g.nby * (uint32_t) g.nbz;
  const uint32_t b = blockIdx.x * 256 + threadIdx.x;
  real lo[3], hi[3];
  bool bad = false;
#pragma unroll
  for (int c = 0; c < 3; c++) { lo[c] = (real) INFINITY; hi[c] = -(real) INFINITY; }
  if (b < nblk) {
    const uint32_t bz = b % (uint32_t) g.nbz, bxy = b / (uint32_t) g.nbz;
    const uint32_t by = bxy % (uint32_t) g.nby, bx = bxy / (uint32_t) g.nby;
    const size_t base = ((size_t) (bx << g.lbx) * g.Ng + (by << g.lby)) * g.Ng
                        + (bz << g.lbz);
    const real *pc[3] = {p0, p1, p2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      for (int a = 0; a < g.BX; a++)
        for (int e = 0; e < g.BY; e++) {
          real v[BZZ];
          load_run<BZZ>(pc[c] + base + ((size_t) a * g.Ng + e) * g.Ng, v);
#pragma unroll
          for (int k = 0; k < BZZ; k++) {
            const real d = v[k] * g.igs;              // as in cell_at
            if (!isfinite(d)) bad = true;
            else { lo[c] = fmin(lo[c], d); hi[c] = fmax(hi[c], d); }
          }
        }
    }
    bnd[2 * (size_t) b] = real4{lo[0], lo[1], lo[2], (real) 0};
    bnd[2 * (size_t) b + 1] = real4{hi[0], hi[1], hi[2], (real) 0};
  }
  // extrema of the workgroup, then one atomic per extremum
#pragma unroll
  for (int o = 32; o; o >>= 1)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      lo[c] = fmin(lo[c], __shfl_xor(lo[c], o, 64));
      hi[c] = fmax(hi[c], __shfl_xor(hi[c], o, 64));
    }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool wb = __ballot(bad) != 0;
  if (lane == 0) {
    for (int c = 0; c < 3; c++) { wlo[c][w] = lo[c]; whi[c][w] = hi[c]; }
    wbad[w] = wb;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int c = threadIdx.x;
    real l = wlo[c][0], h = whi[c][0];
    for (int i = 1; i < 4; i++) { l = fmin(l, wlo[c][i]); h = fmax(h, whi[c][i]); }
    if (l <= h) {
#ifdef SINGLE_PREC
      atomicMin(&gkey[c], real_okey(l));
      atomicMax(&gkey[3 + c], real_okey(h));
#else
      atomicMin((unsigned long long *) &gkey[c], (unsigned long long) real_okey(l));
      atomicMax((unsigned long long *) &gkey[3 + c], (unsigned long long) real_okey(h));
#endif
    }
  }
  else if (threadIdx.x == 3 && (wbad[0] | wbad[1] | wbad[2] | wbad[3])) {
    gkey[6] = 1;
  }
}

// Candidate lattice range of a tile along each axis, relative to its origin.
struct ScanPlan { int lo[3], hi[3]; };

__device__ __forceinline__ bool scan_plan(const rkey *gkey, const CicGeom &g,
    ScanPlan &p) {
  if (gkey[6]) return false;                          // non-finite input
  const int T[3] = {g.TX, g.TY, g.TZ};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double mn = (double) real_from_okey(gkey[a]);
    const double mx = (double) real_from_okey(gkey[3 + a]);
    // keeps every candidate lattice coordinate inside (-Ng, 2 Ng)
    if (!(mn <= mx) || !(mn >= 3.0 - g.Ng) || !(mx <= g.Ng - 3.0)) return false;
    p.lo[a] = -1 - (int) floor(mx + CIC_SCAN_SLACK);
    p.hi[a] = T[a] - 1 - (int) floor(mn - CIC_SCAN_SLACK);
    const int w = p.hi[a] - p.lo[a] + 1;
    if (w >= g.Ng) return false;
  }
  return true;
}

// Lattice points [start, start + count) of one bounds block (unwrapped
// start ub) that can reach the tile at o along one axis; count <= 0: none.
__device__ __forceinline__ int scan_range(int o, int T, int ub, int B, real mn,
    real mx, int &start) {
  int l = o - 1 - (int) floor((double) mx + CIC_SCAN_SLACK);
  int h = o + T - 1 - (int) floor((double) mn - CIC_SCAN_SLACK);
  l = l > ub ? l : ub;
  h = h < ub + B - 1 ? h : ub + B - 1;
  start = l;
  return h - l + 1;
}

__device__ __forceinline__ int wrap_lattice(int a, int n) {
  return a < 0 ? a + n : (a >= n ? a - n : a);
}

// Base cell of a displaced coordinate, wrapped once (the scan admits
// displacements below Ng - 3 cells only), relative to the tile origin o - 1,
// modulo Ng: the cloud reaches [o, o + T) along this axis iff <= T.
__device__ __forceinline__ int scan_rel(real x, int o, int Ng) {
#ifdef SINGLE_PREC
  int a = (int) floorf(x);
#else
  int a = (int) floor(x);
#endif
  a += (a < 0) ? Ng : 0;
  a -= (a >= Ng) ? Ng : 0;
  int t = a - o + 1;
  t += (t < 0) ? Ng : 0;
  t -= (t >= Ng) ? Ng : 0;
  return t;
}

// Tile scan workgroup: the candidate rows (z runs of one bounds block) of
// every bounds block near the tile, one row per thread, tested exactly; the
// particles that reach the tile are compacted per wave and deposited.
template <int MOM, int TXX, int TYY, int TZZ, int BZZ>
__global__ void __launch_bounds__(CIC_SCAN_TPB)
k_cic_scan(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, CicGeom g, const real4 *__restrict__ bnd,
    rkey *__restrict__ gkey, real *__restrict__ rho,
    const real *__restrict__ rho_norm, real4 *__restrict__ vel) {
  constexpr int TC = TXX * TYY * TZZ;
  constexpr int NF = MOM ? 4 : 1;
  constexpr int NW = CIC_SCAN_TPB / 64;
  constexpr int RPW = 64 / BZZ;                 // rows per wave and step
  constexpr int U = CIC_SCAN_U;
  __shared__ double acc[NF * TC];
  __shared__ uint32_t sStart[CIC_SCAN_CH];  // first row of each box
  // box of lattice points: start i, j, k (int16) and extents cy, cz (u8)
  __shared__ uint2 sDesc[CIC_SCAN_CH];
  __shared__ unsigned long long sWave[NW], sTotal;
  // accepted candidates of a wave, deposited 64 at a time
  __shared__ uint32_t sQ[NW][128];
  ScanPlan pl;
  if (!scan_plan(gkey, g, pl)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) gkey[CIC_GKEY_STATUS] = 1;
    return;
  }
  const int Ng = g.Ng;
  const int t = (int) xcd_item(blockIdx.x, gridDim.x);
  const int tz = t % g.ntz, ty = (t / g.ntz) % g.nty, tx = t / (g.ntz * g.nty);
  const int ox = tx * TXX, oy = ty * TYY, oz = tz * TZZ;
  for (int c = threadIdx.x; c < NF * TC; c += CIC_SCAN_TPB) acc[c] = 0;
  // candidate bounds blocks (unwrapped block coordinates)
  const int bx0 = (ox + pl.lo[0]) >> g.lbx, nbxr = ((ox + pl.hi[0]) >> g.lbx) - bx0 + 1;
  const int by0 = (oy + pl.lo[1]) >> g.lby, nbyr = ((oy + pl.hi[1]) >> g.lby) - by0 + 1;
  const int bz0 = (oz + pl.lo[2]) >> g.lbz, nbzr = ((oz + pl.hi[2]) >> g.lbz) - bz0 + 1;
  const int nbk = nbxr * nbyr * nbzr;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t nq = 0;                          // queued candidates of this wave
  for (int c0 = 0; c0 < nbk; c0 += CIC_SCAN_CH) {
    // each thread: the lattice box of one candidate block that can reach
    // the tile, by that block's bounds
    uint32_t rows = 0;
    uint2 desc = {0u, 0u};
    const int cb = c0 + (int) threadIdx.x;
    if (threadIdx.x < CIC_SCAN_CH && cb < nbk) {
      const int cz = cb % nbzr, cxy = cb / nbzr, cy = cxy % nbyr, cx = cxy / nbyr;
      const int ubx = bx0 + cx, uby = by0 + cy, ubz = bz0 + cz;
      const size_t blk = ((size_t) wrap_lattice(ubx, g.nbx) * g.nby
                          + wrap_lattice(uby, g.nby)) * g.nbz + wrap_lattice(ubz, g.nbz);
      const real4 lo = bnd[2 * blk], hi = bnd[2 * blk + 1];
      int i0, j0, k0;
      const int nx = scan_range(ox, TXX, ubx * g.BX, g.BX, lo.x, hi.x, i0);
      const int ny = scan_range(oy, TYY, uby * g.BY, g.BY, lo.y, hi.y, j0);
      const int nz = scan_range(oz, TZZ, ubz * BZZ, BZZ, lo.z, hi.z, k0);
      if (nx > 0 && ny > 0 && nz > 0) {
        rows = (uint32_t) (nx * ny);
        desc = uint2{((uint32_t) i0 & 0xffffu) | ((uint32_t) j0 << 16),
                     ((uint32_t) k0 & 0xffffu) | ((uint32_t) ny << 16) | ((uint32_t) nz << 24)};
      }
    }
    // compaction of the non-empty boxes and their first rows: one scan of
    // (non-empty << 32 | rows)
    const unsigned long long key = rows ? ((1ull << 32) | rows) : 0ull;
    unsigned long long v = key;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned long long u = __shfl_up(v, o, 64);
      if (lane >= o) v += u;
    }
    if (lane == 63) sWave[wv] = v;
    __syncthreads();
    unsigned long long pre = 0;
    for (int w = 0; w < wv; w++) pre += sWave[w];
    const unsigned long long incl = v + pre, excl = incl - key;
    if (rows) {
      const uint32_t slot = (uint32_t) (excl >> 32);
      sStart[slot] = (uint32_t) excl;
      sDesc[slot] = desc;
    }
    if (threadIdx.x == CIC_SCAN_TPB - 1) sTotal = incl;
    __syncthreads();
    const int nd = (int) (sTotal >> 32);
    const uint32_t R = (uint32_t) sTotal;
    int top = 1;
    while (2 * top <= nd) top *= 2;
    // BZZ lanes per row (one lattice point each), RPW rows per wave: the
    // loads of a row are contiguous and the queue keeps z order; U such
    // slabs of rows per lane are in flight together
    const int kz = lane % BZZ;
    for (uint32_t r0 = 0; r0 < R; r0 += U * NW * RPW) {
      uint32_t pk[U];
      int iw[U], jw[U], kw[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t rr = r0 + (uint32_t) (u * NW * RPW + wv * RPW + lane / BZZ);
        pk[u] = 0xffffffffu;
        iw[u] = 0; jw[u] = 0; kw[u] = 0;
        if (rr < R) {
          // last box with first row <= rr
          int e = 0;
          for (int st = top; st; st >>= 1)
            if (e + st < nd && sStart[e + st] <= rr) e += st;
          const uint2 dd = sDesc[e];
          const int cy = (int) ((dd.y >> 16) & 0xffu), nz = (int) (dd.y >> 24);
          if (kz < nz) {
            const int r = (int) (rr - sStart[e]);
            // r = a cy + b, small operands: quotient by a rounded reciprocal
            const int a = (int) (((float) r + 0.5f) * __builtin_amdgcn_rcpf((float) cy));
            iw[u] = wrap_lattice((int) (int16_t) (dd.x & 0xffffu) + a, Ng);
            jw[u] = wrap_lattice(((int) dd.x >> 16) + r - a * cy, Ng);
            kw[u] = wrap_lattice((int) (int16_t) (dd.y & 0xffffu) + kz, Ng);
            pk[u] = ((uint32_t) iw[u] * (uint32_t) Ng + (uint32_t) jw[u]) * (uint32_t) Ng
                    + (uint32_t) kw[u];
          }
        }
      }
      real qq[U][3];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t p = pk[u] != 0xffffffffu ? pk[u] : 0u;
        qq[u][0] = p0[p];
        qq[u][1] = p1[p];
        qq[u][2] = p2[p];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        bool hit = false;
        if (pk[u] != 0xffffffffu) {
          // coordinates as cell_at forms them
          const int sx = scan_rel(qq[u][0] * g.igs + (real) iw[u], ox, Ng);
          const int sy = scan_rel(qq[u][1] * g.igs + (real) jw[u], oy, Ng);
          const int sz = scan_rel(qq[u][2] * g.igs + (real) kw[u], oz, Ng);
          hit = sx <= TXX && sy <= TYY && sz <= TZZ;
        }
        const uint64_t m = __ballot(hit);
        if (hit) sQ[wv][nq + (uint32_t) __popcll(m & ((1ull << lane) - 1))] = pk[u];
        nq += (uint32_t) __popcll(m);
        if (nq < 64) continue;                      // wave-uniform
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        {                                           // one full wave
          const uint32_t p = sQ[wv][nq - 64 + lane];
          const real q0 = p0[p], q1 = p1[p], q2 = p2[p];
          const ParticleCell c = cell_from(q0, q1, q2, p, g);
          cic_deposit<MOM, TXX, TYY, TZZ>(acc, c, q0, q1, q2, ox, oy, oz, Ng);
          nq -= 64;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    __syncthreads();                    // boxes of the next round
  }
  // the wave's remaining queue
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if ((uint32_t) lane < nq) {
    const uint32_t p = sQ[wv][lane];
    const real q0 = p0[p], q1 = p1[p], q2 = p2[p];
    const ParticleCell c = cell_from(q0, q1, q2, p, g);
    cic_deposit<MOM, TXX, TYY, TZZ>(acc, c, q0, q1, q2, ox, oy, oz, Ng);
  }
  __syncthreads();
  cic_store<MOM, TXX, TYY, TZZ, CIC_SCAN_TPB>(acc, ox, oy, oz, Ng, rho, rho_norm, vel);
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
  g.BX = tile_dim(Ng, 4);
  g.BY = tile_dim(Ng, 4);
  g.BZ = tile_dim(Ng, 8);
  g.lbx = __builtin_ctz(g.BX); g.lby = __builtin_ctz(g.BY); g.lbz = __builtin_ctz(g.BZ);
  g.nbx = Ng / g.BX; g.nby = Ng / g.BY; g.nbz = Ng / g.BZ;
  return g;
}

size_t cic_nkeys(const CicGeom &g) { return (size_t) g.ntx * g.nty * g.ntz; }
size_t cic_nblocks(const CicGeom &g) { return (size_t) g.nbx * g.nby * g.nbz; }

template <int MOM, int A, int B, int C>
static void launch_tile(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    const real *rho_norm, real4 *vel, uint32_t cap, hipStream_t st) {
  const unsigned nt = (unsigned) (g.ntx * g.nty * g.ntz);
  const int tpw = CIC_TILES_PER_WG;
  hipLaunchKernelGGL((k_cic_tile<MOM, A, B, C>), dim3((nt + tpw - 1) / tpw),
      dim3(CIC_TILE_TPB), 0, st, p0, p1, p2, g, off, list, rho, rho_norm, vel, tpw, cap);
}

template <int MOM>
static hipError_t dispatch_tile(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    const real *rho_norm, real4 *vel, uint32_t cap, hipStream_t st) {
  if (g.TX == 8 && g.TY == 8 && g.TZ == 32)
    launch_tile<MOM, 8, 8, 32>(p0, p1, p2, g, off, list, rho, rho_norm, vel, cap, st);
  else if (g.TX == 8 && g.TY == 8 && g.TZ == 16)
    launch_tile<MOM, 8, 8, 16>(p0, p1, p2, g, off, list, rho, rho_norm, vel, cap, st);
  else if (g.TX == 8 && g.TY == 8 && g.TZ == 8)
    launch_tile<MOM, 8, 8, 8>(p0, p1, p2, g, off, list, rho, rho_norm, vel, cap, st);
  else if (g.TX == 4 && g.TY == 4 && g.TZ == 4)
    launch_tile<MOM, 4, 4, 4>(p0, p1, p2, g, off, list, rho, rho_norm, vel, cap, st);
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
  if (g.Ng <= 1536)
    hipLaunchKernelGGL((k_cic_bin<0, 1024>), dim3(nb), dim3(CIC_TPB), 0, st, p0, p1, p2, g,
        counts, (unsigned long long *) nullptr, (uint32_t *) nullptr, 0u, (uint32_t *) nullptr);
  else
    hipLaunchKernelGGL((k_cic_bin<0, 2048>), dim3(nb), dim3(CIC_TPB), 0, st, p0, p1, p2, g,
        counts, (unsigned long long *) nullptr, (uint32_t *) nullptr, 0u, (uint32_t *) nullptr);
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
    const CicGeom &g, uint64_t *cursor, uint32_t *list, uint32_t cap, uint32_t *ovf,
    hipStream_t st) {
  if (cap && (!ovf || cic_nkeys(g) >= (1ull << 27))) return hipErrorInvalidValue;
  const unsigned nb = (unsigned) ((g.N + CIC_PPB - 1) / CIC_PPB);
  if (g.Ng <= 1536)
    hipLaunchKernelGGL((k_cic_bin<1, 1024>), dim3(nb), dim3(CIC_TPB), 0, st, p0, p1, p2, g,
        (uint32_t *) nullptr, (unsigned long long *) cursor, list, cap, ovf);
  else
    hipLaunchKernelGGL((k_cic_bin<1, 2048>), dim3(nb), dim3(CIC_TPB), 0, st, p0, p1, p2, g,
        (uint32_t *) nullptr, (unsigned long long *) cursor, list, cap, ovf);
  return hipGetLastError();
}

__global__ void k_cic_fixed_init(uint64_t *__restrict__ cursor, size_t nk,
    uint32_t cap, uint32_t *__restrict__ ovf) {
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nk) cursor[i] = (uint64_t) i * cap;
  if (i == 0) *ovf = 0u;
}

hipError_t launch_cic_fixed_init(uint64_t *cursor, const CicGeom &g, uint32_t cap,
    uint32_t *ovf, hipStream_t st) {
  const size_t nk = cic_nkeys(g);
  hipLaunchKernelGGL(k_cic_fixed_init, dim3(ez_blocks(nk, 256)), dim3(256), 0, st,
      cursor, nk, cap, ovf);
  return hipGetLastError();
}

// Tile accumulation: density into rho (unless null) and, when vel is not
// null, the normalised momentum fields.
hipError_t launch_cic_tiles(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    const real *rho_norm, real4 *vel, uint32_t cap, hipStream_t st) {
  if (vel && !rho && !rho_norm) return hipErrorInvalidValue;
  if (vel) return dispatch_tile<1>(p0, p1, p2, g, off, list, rho, rho_norm, vel, cap, st);
  return dispatch_tile<0>(p0, p1, p2, g, off, list, rho, rho_norm, vel, cap, st);
}

hipError_t launch_cic_bounds(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, real4 *bnd, rkey *gkey, hipStream_t st) {
  if (g.N >= 0xffffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_cic_bounds_init, dim3(1), dim3(64), 0, st, gkey);
  const unsigned nb = ez_blocks(cic_nblocks(g), 256);
  if (g.BZ == 8)
    hipLaunchKernelGGL(k_cic_bounds<8>, dim3(nb), dim3(256), 0, st, p0, p1, p2, g, bnd, gkey);
  else if (g.BZ == 4)
    hipLaunchKernelGGL(k_cic_bounds<4>, dim3(nb), dim3(256), 0, st, p0, p1, p2, g, bnd, gkey);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int MOM, int A, int B, int C>
static void launch_scan_tiles(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const real4 *bnd, rkey *gkey, real *rho,
    const real *rho_norm, real4 *vel, hipStream_t st) {
  const unsigned nt = (unsigned) (g.ntx * g.nty * g.ntz);
  if (g.BZ == 8)
    hipLaunchKernelGGL((k_cic_scan<MOM, A, B, C, 8>), dim3(nt), dim3(CIC_SCAN_TPB), 0, st,
        p0, p1, p2, g, bnd, gkey, rho, rho_norm, vel);
  else
    hipLaunchKernelGGL((k_cic_scan<MOM, A, B, C, 4>), dim3(nt), dim3(CIC_SCAN_TPB), 0, st,
        p0, p1, p2, g, bnd, gkey, rho, rho_norm, vel);
}

template <int MOM>
static hipError_t dispatch_scan(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const real4 *bnd, rkey *gkey, real *rho,
    const real *rho_norm, real4 *vel, hipStream_t st) {
  if (g.TX == 8 && g.TY == 8 && g.TZ == 16)
    launch_scan_tiles<MOM, 8, 8, 16>(p0, p1, p2, g, bnd, gkey, rho, rho_norm, vel, st);
  else if (g.TX == 8 && g.TY == 8 && g.TZ == 8)
    launch_scan_tiles<MOM, 8, 8, 8>(p0, p1, p2, g, bnd, gkey, rho, rho_norm, vel, st);
  else if (g.TX == 4 && g.TY == 4 && g.TZ == 4)
    launch_scan_tiles<MOM, 4, 4, 4>(p0, p1, p2, g, bnd, gkey, rho, rho_norm, vel, st);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_cic_scan(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const real4 *bnd, rkey *gkey, real *rho,
    const real *rho_norm, real4 *vel, hipStream_t st) {
  if (vel && !rho && !rho_norm) return hipErrorInvalidValue;
  if (g.BZ != 8 && g.BZ != 4) return hipErrorInvalidValue;
  if (vel) return dispatch_scan<1>(p0, p1, p2, g, bnd, gkey, rho, rho_norm, vel, st);
  return dispatch_scan<0>(p0, p1, p2, g, bnd, gkey, rho, rho_norm, vel, st);
}
