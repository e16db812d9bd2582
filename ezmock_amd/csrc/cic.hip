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

struct ParticleCell {
  real x, y, z;                 // displaced grid coordinates
  int x0, y0, z0;               // base cell (floor), before wrapping
};

// dens_field.c:263-289
__device__ __forceinline__ ParticleCell particle_cell(const real *__restrict__ p0,
    const real *__restrict__ p1, const real *__restrict__ p2, size_t idx,
    int Ng, real igs) {
  ParticleCell c;
  const int k = (int) (idx % Ng);
  const size_t ij = idx / Ng;
  const int j = (int) (ij % Ng);
  const int i = (int) (ij / Ng);
  c.x = p0[idx] * igs + (real) i;
  c.y = p1[idx] * igs + (real) j;
  c.z = p2[idx] * igs + (real) k;
  int x0 = (int) floor((double) c.x), y0 = (int) floor((double) c.y),
      z0 = (int) floor((double) c.z);
  c.x0 = x0; c.y0 = y0; c.z0 = z0;      // unwrapped: weights use these
  return c;
}

__device__ __forceinline__ ParticleCell cell_from(real q0, real q1, real q2,
    size_t idx, int Ng, real igs) {
  ParticleCell c;
  const int k = (int) (idx % Ng);
  const size_t ij = idx / Ng;
  const int j = (int) (ij % Ng);
  const int i = (int) (ij / Ng);
  c.x = q0 * igs + (real) i;
  c.y = q1 * igs + (real) j;
  c.z = q2 * igs + (real) k;
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

// Tiles touched by the stencil: tx0 and (if the base cell is the last of
// its tile) tx0 + 1, and likewise in y and z.  Returns the count; keys are
// tile ids (tx * nty + ty) * ntz + tz.
__device__ __forceinline__ int tiles_of(const ParticleCell &c, const CicGeom &g,
    uint32_t *keys) {
  const int x0 = wrap_once(c.x0, g.Ng), y0 = wrap_once(c.y0, g.Ng),
      z0 = wrap_once(c.z0, g.Ng);
  const int tx = x0 / g.TX, ty = y0 / g.TY, tz = z0 / g.TZ;
  const int ex = (x0 - tx * g.TX == g.TX - 1 && g.ntx > 1) ? 2 : 1;
  const int ey = (y0 - ty * g.TY == g.TY - 1 && g.nty > 1) ? 2 : 1;
  const int ez = (z0 - tz * g.TZ == g.TZ - 1 && g.ntz > 1) ? 2 : 1;
  int n = 0;
  for (int a = 0; a < ex; a++) {
    const int ux = (tx + a == g.ntx) ? 0 : tx + a;
    for (int b = 0; b < ey; b++) {
      const int uy = (ty + b == g.nty) ? 0 : ty + b;
      for (int d = 0; d < ez; d++) {
        const int uz = (tz + d == g.ntz) ? 0 : tz + d;
        keys[n++] = (uint32_t) ((ux * g.nty + uy) * g.ntz + uz);
      }
    }
  }
  return n;
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

__device__ __forceinline__ Window make_window(const CicGeom &g, size_t b0) {
  Window w;
  const size_t ntile = (size_t) g.ntx * g.nty * g.ntz;
  w.direct = ntile <= WMAX;
  const size_t plane = (size_t) g.Ng * g.Ng;
  const size_t pl = ((b0 + CIC_PPB < g.N) ? b0 + CIC_PPB : g.N) - 1;
  const int i0 = (int) (b0 / plane), i1 = (int) (pl / plane);
  int j0 = 0, j1 = g.Ng - 1;
  if (i0 == i1) { j0 = (int) ((b0 / g.Ng) % g.Ng); j1 = (int) ((pl / g.Ng) % g.Ng); }
  w.wx = i1 / g.TX - i0 / g.TX + 3;
  w.wy = j1 / g.TY - j0 / g.TY + 3;
  if (w.wx > g.ntx) w.wx = g.ntx;
  if (w.wy > g.nty) w.wy = g.nty;
  w.txlo = (i0 / g.TX - 1 + g.ntx) % g.ntx;
  w.tylo = (j0 / g.TY - 1 + g.nty) % g.nty;
  if (!w.direct && (size_t) w.wx * w.wy * g.ntz > WMAX) {
    w.wx = 1; w.wy = 1;         // degenerate shapes: mostly global path
  }
  return w;
}

// Window slot of a tile key, or -1 outside the window.
__device__ __forceinline__ int window_slot(const Window &w, const CicGeom &g,
    uint32_t key) {
  if (w.direct) return (int) key;
  const int uz = (int) (key % g.ntz);
  const int uy = (int) ((key / g.ntz) % g.nty);
  const int ux = (int) (key / ((uint32_t) g.ntz * g.nty));
  const int dx = (ux - w.txlo + g.ntx) % g.ntx, dy = (uy - w.tylo + g.nty) % g.nty;
  if (dx >= w.wx || dy >= w.wy) return -1;
  return (dx * w.wy + dy) * g.ntz + uz;
}

__device__ __forceinline__ uint32_t slot_key(const Window &w, const CicGeom &g,
    int slot) {
  if (w.direct) return (uint32_t) slot;
  const int uz = slot % g.ntz;
  const int dy = (slot / g.ntz) % w.wy, dx = slot / (g.ntz * w.wy);
  const int ux = (w.txlo + dx) % g.ntx, uy = (w.tylo + dy) % g.nty;
  return (uint32_t) ((ux * g.nty + uy) * g.ntz + uz);
}

template <int SCATTER>
__global__ void __launch_bounds__(CIC_TPB)
k_cic_bin(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, CicGeom g, uint32_t *__restrict__ counts,
    unsigned long long *__restrict__ cursor, uint32_t *__restrict__ list) {
  __shared__ uint32_t wcnt[WMAX];
  __shared__ unsigned long long wbase[SCATTER ? WMAX : 1];
  const size_t b0 = (size_t) blockIdx.x * CIC_PPB;
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
    const size_t p = b0 + (size_t) u * CIC_TPB + threadIdx.x;
    const bool ok = p < g.N;
    q[0][u] = ok ? p0[p] : (real) 0;
    q[1][u] = ok ? p1[p] : (real) 0;
    q[2][u] = ok ? p2[p] : (real) 0;
  }
  uint32_t keys[8];
  if (!SCATTER) {
#pragma unroll
    for (int u = 0; u < CIC_PPT; u++) {
      const size_t p = b0 + (size_t) u * CIC_TPB + threadIdx.x;
      if (p >= g.N) continue;
      const ParticleCell c = cell_from(q[0][u], q[1][u], q[2][u], p, g.Ng, g.igs);
      const int n = tiles_of(c, g, keys);
      for (int v = 0; v < n; v++) {
        const int s = window_slot(w, g, keys[v]);
        if (s >= 0) atomicAdd(&wcnt[s], 1u);
        else atomicAdd(&counts[keys[v]], 1u);
      }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nslot; s += CIC_TPB)
      if (wcnt[s]) atomicAdd(&counts[slot_key(w, g, s)], wcnt[s]);
    return;
  }
  // scatter: count the window, reserve each tile's range, place the indices
#pragma unroll
  for (int u = 0; u < CIC_PPT; u++) {
    const size_t p = b0 + (size_t) u * CIC_TPB + threadIdx.x;
    if (p >= g.N) continue;
    const ParticleCell c = cell_from(q[0][u], q[1][u], q[2][u], p, g.Ng, g.igs);
    const int n = tiles_of(c, g, keys);
    for (int v = 0; v < n; v++) {
      const int s = window_slot(w, g, keys[v]);
      if (s >= 0) atomicAdd(&wcnt[s], 1u);
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nslot; s += CIC_TPB) {
    if (wcnt[s])
      wbase[s] = atomicAdd(&cursor[slot_key(w, g, s)], (unsigned long long) wcnt[s]);
    wcnt[s] = 0;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < CIC_PPT; u++) {
    const size_t p = b0 + (size_t) u * CIC_TPB + threadIdx.x;
    if (p >= g.N) continue;
    const ParticleCell c = cell_from(q[0][u], q[1][u], q[2][u], p, g.Ng, g.igs);
    const int n = tiles_of(c, g, keys);
    for (int v = 0; v < n; v++) {
      const int s = window_slot(w, g, keys[v]);
      unsigned long long pos;
      if (s >= 0) pos = wbase[s] + atomicAdd(&wcnt[s], 1u);
      else pos = atomicAdd(&cursor[keys[v]], 1ull);
      list[pos] = (uint32_t) p;
    }
  }
}

// One workgroup per tile.  MOM: also the three momentum fields.
template <int MOM, int TXX, int TYY, int TZZ>
__global__ void __launch_bounds__(CIC_TPB)
k_cic_tile(const real *__restrict__ p0, const real *__restrict__ p1,
    const real *__restrict__ p2, CicGeom g, const uint64_t *__restrict__ off,
    const uint32_t *__restrict__ list, real *__restrict__ rho,
    real *__restrict__ m0, real *__restrict__ m1, real *__restrict__ m2) {
  constexpr int TC = TXX * TYY * TZZ;
  constexpr int NF = MOM ? 4 : 1;
  // double accumulators: LDS double atomics are several times faster than
  // float ones on gfx950, and the float result is the correctly rounded
  // sum of the float contributions
  __shared__ double acc[NF][TC];
  for (int c = threadIdx.x; c < NF * TC; c += CIC_TPB) (&acc[0][0])[c] = 0;
  __syncthreads();
  const int t = blockIdx.x;
  const int tz = t % g.ntz, ty = (t / g.ntz) % g.nty, tx = t / (g.ntz * g.nty);
  const int Ng = g.Ng;
  const int ox = tx * TXX, oy = ty * TYY, oz = tz * TZZ;
  const uint64_t s0 = off[t], e0 = off[t + 1];
  constexpr int U = 4;
  for (uint64_t qb = s0; qb < e0; qb += (uint64_t) U * CIC_TPB) {
  size_t pv[U];
  real qq[U][3];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const uint64_t q = qb + (uint64_t) u * CIC_TPB + threadIdx.x;
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
    const ParticleCell c = cell_from(qq[u][0], qq[u][1], qq[u][2], p, Ng, g.igs);
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
#pragma unroll
    for (int a = 0; a < 2; a++) {
      if (lx[a] < 0) continue;
#pragma unroll
      for (int b = 0; b < 2; b++) {
        if (ly[b] < 0) continue;
#pragma unroll
        for (int d = 0; d < 2; d++) {
          if (lz[d] < 0) continue;
          const int li = (lx[a] * TYY + ly[b]) * TZZ + lz[d];
          atomicAdd(&acc[0][li], (double) (wx[a] * wy[b] * wz[d]));
          if (MOM) {
#pragma unroll
            for (int f = 0; f < 3; f++) {
              const real u = wx[a] * qv[f];          // pop_tracer.c:641-642
              atomicAdd(&acc[1 + f][li], (double) (u * wy[b] * wz[d]));
            }
          }
        }
      }
    }
  }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < TC; c += CIC_TPB) {
    const int lz = c % TZZ, ly = (c / TZZ) % TYY, lx = c / (TZZ * TYY);
    const size_t gi = ((size_t) (ox + lx) * Ng + (oy + ly)) * Ng + (oz + lz);
    if (rho) rho[gi] = (real) acc[0][c];
    if (MOM) {
      m0[gi] = (real) acc[1][c]; m1[gi] = (real) acc[2][c]; m2[gi] = (real) acc[3][c];
    }
  }
}

__global__ void k_cursor_init(uint64_t *__restrict__ off,
    uint64_t *__restrict__ cursor, size_t nk, const uint64_t *__restrict__ total) {
  const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nk) cursor[i] = off[i];
  else if (i == nk) off[nk] = *total;
}

// Tile shape: 8 x 8 x 32 cells when the grid allows it.
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
  g.TZ = tile_dim(Ng, 32);
  g.ntx = Ng / g.TX; g.nty = Ng / g.TY; g.ntz = Ng / g.TZ;
  g.N = (size_t) Ng * Ng * Ng;
  return g;
}

size_t cic_nkeys(const CicGeom &g) { return (size_t) g.ntx * g.nty * g.ntz; }

template <int MOM, int A, int B, int C>
static void launch_tile(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    real *m0, real *m1, real *m2, hipStream_t st) {
  const unsigned nt = (unsigned) (g.ntx * g.nty * g.ntz);
  hipLaunchKernelGGL((k_cic_tile<MOM, A, B, C>), dim3(nt), dim3(CIC_TPB), 0, st,
      p0, p1, p2, g, off, list, rho, m0, m1, m2);
}

template <int MOM>
static hipError_t dispatch_tile(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    real *m0, real *m1, real *m2, hipStream_t st) {
  if (g.TX == 8 && g.TY == 8 && g.TZ == 32)
    launch_tile<MOM, 8, 8, 32>(p0, p1, p2, g, off, list, rho, m0, m1, m2, st);
  else if (g.TX == 8 && g.TY == 8 && g.TZ == 16)
    launch_tile<MOM, 8, 8, 16>(p0, p1, p2, g, off, list, rho, m0, m1, m2, st);
  else if (g.TX == 8 && g.TY == 8 && g.TZ == 8)
    launch_tile<MOM, 8, 8, 8>(p0, p1, p2, g, off, list, rho, m0, m1, m2, st);
  else if (g.TX == 4 && g.TY == 4 && g.TZ == 4)
    launch_tile<MOM, 4, 4, 4>(p0, p1, p2, g, off, list, rho, m0, m1, m2, st);
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
// u32, between N and 8 N; about 1.3 N for 8 x 8 x 32 tiles).
hipError_t launch_cic_scatter(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, uint64_t *cursor, uint32_t *list, hipStream_t st) {
  const unsigned nb = (unsigned) ((g.N + CIC_PPB - 1) / CIC_PPB);
  hipLaunchKernelGGL(k_cic_bin<1>, dim3(nb), dim3(CIC_TPB), 0, st, p0, p1, p2, g,
      (uint32_t *) nullptr, (unsigned long long *) cursor, list);
  return hipGetLastError();
}

// Tile accumulation: density into rho (unless null) and, when m0 is not
// null, the three momentum fields.
hipError_t launch_cic_tiles(const real *p0, const real *p1, const real *p2,
    const CicGeom &g, const uint64_t *off, const uint32_t *list, real *rho,
    real *m0, real *m1, real *m2, hipStream_t st) {
  if (m0) return dispatch_tile<1>(p0, p1, p2, g, off, list, rho, m0, m1, m2, st);
  return dispatch_tile<0>(p0, p1, p2, g, off, list, rho, m0, m1, m2, st);
}
