// ops.hip -- fused operator kernels (see ops.hpp for the staging / epilogue
// contract).  Grids carry the right-hand side q = 6 s + lc; per-sample data
// (moduli, stencils, inverse diagonals, Chebyshev coefficients) is selected
// by s = q / 6.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>
#include <cmath>
#include <stdexcept>

#include "device_common.hpp"
#include "ops.hpp"

namespace femh {

// ---------------------------------------------------------------------------
// fine level: tiled cube-mode voxel kernel
//
// K0 = H^T Lambda H per voxel (tables.hpp).  A workgroup owns a TJ x TK tile
// of node columns (j,k) (16 x 16, 8 x 32 or 16 x 32, one thread per column)
// and marches over a chunk of TI planes in i.  Per voxel slab (planes s, s+1):
//   1. node plane s+1 of the tile plus a one-node halo on each j/k side
//      ((TJ+2) x (TK+2) nodes, periodic wrap) is staged into LDS (a
//      two-buffer ring, see below);
//   2. each of the (TJ+1) x (TK+1) voxels touching the tile transforms its
//      top face ((j,k) butterflies; the bottom face is the top face of the
//      slab below, carried in registers), forms the 8 cube modes per
//      component with the i stage, applies the mode coupling scaled by its
//      modulus and splits the result back into bottom- and top-face modes;
//   3. the bottom-face modes plus the top-face modes of the voxel below (the
//      same node plane s) go through one inverse face transform into LDS,
//      and the 4 tile nodes sharing each corner gather them: plane s is
//      complete and passes through the epilogue.
// The j/k halo voxels are recomputed by neighbouring tiles ((TJ+1)(TK+1) /
// (TJ TK) of the voxel work: 1.13 for 16 x 16, 1.10 for 16 x 32) and the i
// halo slab once per chunk ((TI+1)/TI).
constexpr int kTile = 16;
constexpr int kTN = kTile + 2;   // staged nodes per row
constexpr int kTE = kTile + 1;   // voxels per row
constexpr int kEOH = 13;         // LDS stride of one voxel's 12 plane forces (odd: no bank conflicts)

// periodic index for i in [-n, 2n) without a division
__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// epilogue inputs of one owned node (loaded ahead of the operator work so
// that their latency overlaps it):
//   EP_RES: e0 = b ; EP_CHEB0: e0 = b, e1 = 1/D ; EP_STEP: e0 = rprev, e1 = 1/D,
//   e2 = xprev (0 without one)
template <typename T>
struct EpiIn {
  T e0[3], e1[3], e2[3], e3[3];  // e3: EP_STEP with a dot: b (for b . x)
};

template <typename T, int EP>
__device__ __forceinline__ void epi_load(const OpArgs<T> &a, const T *__restrict__ Ds, size_t off,
                                         int V, int n, EpiIn<T> &p) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t q = off + (size_t)c * V + n;
    if (EP == EP_RES || EP == EP_CHEB0) p.e0[c] = a.b[q];
    if (EP == EP_STEP) p.e0[c] = a.rprev[q];
    if (EP == EP_CHEB0 || EP == EP_STEP || (EP == EP_AX && a.ydinv)) p.e1[c] = Ds[(size_t)c * V + n];
    if (EP == EP_STEP) p.e2[c] = a.xprev ? a.xprev[q] : (T)0;
    if (EP == EP_STEP) p.e3[c] = a.dot ? a.b[q] : (T)0;
  }
}

// FP64 partial sums of the epilogue: EP_AX u.Au ; EP_STEP b.x_new ;
// EP_RES r.r, b.x, x.r (x = the staged vector)
struct DotAcc {
  double a = 0, b = 0, c = 0;
};

// output store of an epilogue; NT: a streaming (non-temporal) store, for
// output vectors far larger than the caches (read back only by a later
// launch, which would find them evicted anyway)
template <bool NT, typename T>
__device__ __forceinline__ void ep_store(T *p, T v) {
  if (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// epilogue at one owned node (all three components)
template <typename T, int EP, bool NT = false>
__device__ __forceinline__ void epilogue(const OpArgs<T> &a, const EpiIn<T> &p, T ti, T c1, T c2,
                                         size_t off, int V, int n, const T Au[3], const T uo[3],
                                         DotAcc &dv) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t q = off + (size_t)c * V + n;
    if (EP == EP_AX) {
      ep_store<NT>(&a.y[q], a.ydinv ? Au[c] * p.e1[c] : Au[c]);
      if (a.pout) ep_store<NT>(&a.pout[q], uo[c]);
      if (a.dot) dv.a += (double)uo[c] * (double)Au[c];
    } else if (EP == EP_RES) {
      const T rv = p.e0[c] - Au[c];
      ep_store<NT>(&a.y[q], rv);
      if (a.dot) {
        dv.a += (double)rv * (double)rv;
        dv.b += (double)p.e0[c] * (double)uo[c];
        dv.c += (double)uo[c] * (double)rv;
      }
    } else if (EP == EP_CHEB0) {
      const T rv = p.e0[c] - Au[c];
      if (a.fin1) {
        const T xv = uo[c] + rv * ti * p.e1[c];
        ep_store<NT>(&a.x[q], xv);
        if (a.dot) dv.a += (double)p.e0[c] * (double)xv;
      } else {
        ep_store<NT>(&a.r[q], rv);
        if (!a.last) ep_store<NT>(&a.dout[q], rv * ti * p.e1[c]);  // `last`: no smoothing step follows
        ep_store<NT>(&a.x[q], uo[c]);
      }
    } else {  // EP_STEP
      const T rv = p.e0[c] - Au[c];
      const T dn = c1 * uo[c] + c2 * rv * p.e1[c];
      T xv = p.e2[c] + uo[c];
      if (a.last) {
        xv += dn;  // the last step's r and d are not read again
      } else {
        ep_store<NT>(&a.r[q], rv);
        ep_store<NT>(&a.dout[q], dn);
      }
      ep_store<NT>(&a.x[q], xv);
      if (a.dot) dv.a += (double)p.e3[c] * (double)xv;
    }
  }
}

#ifndef FEMH_WAVES64
#define FEMH_WAVES64 1
#endif
constexpr int kWaves64 = FEMH_WAVES64;  // FP64 fine operator: waves per SIMD the registers allow

// Tile shapes (TJ x TK owned node columns per workgroup, TJ*TK threads).
// Launch shape encoding: bits 0-7 node planes per workgroup (0 = built-in),
// bits 8-15 tile variant (0 = built-in, tiled_default_tile).
struct TileShape {
  int tj, tk;
};
constexpr TileShape kTileShapes[kNumTileShapes] = {{16, 16}, {16, 16}, {8, 32}, {16, 32}};

// The built-in tile on grids whose z extent is a multiple of 32 (and y of
// 16) and at least 64:
//  * batched launches (24 or more right-hand sides): 16 x 32 -- a tile row is
//    a whole 128-byte line in FP32 and the j/k halo recompute drops from 1.13
//    to 1.10 (64^3, 16 cells: 15% less time than 16 x 16 over the four
//    variants of a PCG iteration);
//  * one cell (6 right-hand sides, too few workgroups for the larger tile):
//    8 x 32 -- as many tiles as 16 x 16, so as many workgroups, with
//    whole-line rows (64^3: 53.3 -> 50.8 us per FP64 launch, 31.2 -> 28.8 us
//    per FP32 launch).
// Smaller or odd grids (the 24^3 and 48^3 cells of an uneven batch, which
// waste less on partial tiles) keep 16 x 16.
int tiled_default_tile(const Dims &g, int R) {
  if (g.nz >= 64 && g.nz % 32 == 0 && g.ny % 16 == 0) return R >= 24 ? 3 : (R <= kRHS ? 2 : 1);
  return 1;
}

static inline TileShape tile_of(const Dims &g, int R, int shape) {
  int v = (shape >> 8) & 0xff;
  if (v <= 0 || v >= kNumTileShapes) v = tiled_default_tile(g, R);
  return kTileShapes[v];
}

static inline int tiled_tiles(const Dims &g, const TileShape &ts) {
  return ((g.ny + ts.tj - 1) / ts.tj) * ((g.nz + ts.tk - 1) / ts.tk);
}

static int tiled_resident(int sz, const TileShape &ts);

// Chunk length for a launch whose active workgroups (Ract active right-hand
// sides) fill the device about once or less: the rounds of resident
// workgroups set the time, so take the length that minimises rounds x slabs
// per workgroup (at most 255 planes, the width of the launch-shape field).
static int round_chunk(const Dims &g, const TileShape &ts, int tiles, int Ract, int sz) {
  const long res = tiled_resident(sz, ts);
  long best = LONG_MAX;
  int TI = g.nx;
  for (int c = (g.nx + 254) / 255; c <= g.nx; c++) {
    const int ti = (g.nx + c - 1) / c, ch = (g.nx + ti - 1) / ti;
    if ((long)tiles * ch > kMaxPartials) continue;
    const long rounds = ((long)tiles * ch * Ract + res - 1) / res, cost = rounds * (ti + 1);
    if (cost < best) {
      best = cost;
      TI = ti;
    }
  }
  return TI;
}

static inline int tiled_chunk(const Dims &g, int R, int shape, int sz) {
  const TileShape ts = tile_of(g, R, shape);
  const int tiles = tiled_tiles(g, ts);
  const int want = shape & 0xff;
  // planes per chunk: `want` when given (an autotuned shape).  One cell (six
  // right-hand sides) gives too few workgroups to fill the device more than
  // about once, so the rounds of resident workgroups set the time: take the
  // chunk length that minimises rounds x slabs per workgroup (64^3 in FP64:
  // 13 planes, 480 workgroups in one round, against 8 planes and 768
  // workgroups in two).  Batches: about 32 planes (halo slab overhead 1/32)
  // but enough chunks to put ~2 workgroups of 256 threads per CU across the
  // right-hand sides.  At most kMaxPartials workgroups per right-hand side
  // either way.
  int TI = 32;
  if (want > 0) {
    TI = want;
  } else if (R <= kRHS) {
    TI = round_chunk(g, ts, tiles, R, sz);
  } else {
    while (TI > 2 && (long)tiles * ((g.nx + TI - 1) / TI) * R * ts.tj * ts.tk < 512L * 256) TI /= 2;
  }
  if (TI > g.nx) TI = g.nx;
  while ((long)tiles * ((g.nx + TI - 1) / TI) > kMaxPartials) TI++;
  return TI;
}

int tiled_slabs(const Dims &g, int R, int shape, int sz) { return tiled_chunk(g, R, shape, sz); }

int tiled_sparse_shape(const Dims &g, int R, int Ract, int shape, int sz) {
  if (shape & 0xff) return shape;  // planes pinned (autotuned or set)
  int v = (shape >> 8) & 0xff;
  if (v <= 0 || v >= kNumTileShapes) v = tiled_default_tile(g, R);
  const TileShape ts = kTileShapes[v];
  return tiled_shape_code(v, round_chunk(g, ts, tiled_tiles(g, ts), Ract, sz));
}

int tiled_parts(const Dims &g, int R, int shape, int sz) {
  const int TI = tiled_chunk(g, R, shape, sz);
  return tiled_tiles(g, tile_of(g, R, shape)) * ((g.nx + TI - 1) / TI);
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b to XCD
// b mod 8), each with its own L2.  xcd_order maps the dispatch index to a
// logical index so that every XCD runs one contiguous range of logical
// workgroups: consecutive logical workgroups -- which share halo rows and
// per-cell data -- then meet in the same L2.
__device__ __forceinline__ int xcd_order(int b, int n) {
  const int x = b & 7, slot = b >> 3, per = n >> 3, rem = n & 7;
  return x * per + (x < rem ? x : rem) + slot;
}

// raw buffer access: base pointer and byte range in a resource, one 32-bit
// lane offset plus a scalar offset per access (no 64-bit address arithmetic)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float *) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ double bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, double *) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// grid: one dimension of tiles * chunks * R workgroups; logical order (cell,
// chunk, tile, load case), load case fastest, so that the six load cases of a
// tile (same moduli and diagonal) and then the neighbouring tiles follow each
// other on one XCD.
//
// LDS: staged node planes in a two-buffer ring, node-interleaved
// (up[buf][3 * slot + c]), so that the 12 values of a voxel face and the
// owned node's value are immediate offsets from one address; eo holds each
// voxel's 12 bottom-face forces (odd stride: no bank conflicts).  A buffer is
// overwritten two slabs after it was filled, behind the barrier that closes
// the slab in between (every read of it -- the FP64 bottom face, the owned
// node's staged value -- happens before that barrier).
// Waves per SIMD the FP32 variants are compiled for: 4 (at most 128 VGPRs),
// except the D^-1-scaled Chebyshev step (two smoothing steps in one launch),
// whose staged diagonal and four epilogue inputs do not fit: held to 128 it
// spilled 16 VGPRs to scratch.
template <typename T, int ST, int EP>
constexpr int tiled_waves() {
  return sizeof(T) == 4 ? ((ST == ST_DSCALE && EP == EP_STEP) ? 3 : 4) : kWaves64;
}

template <typename T, int ST, int EP, int TJ, int TK>
__global__ __launch_bounds__(TJ * TK, (tiled_waves<T, ST, EP>())) void k_tiled(Dims g, const T *__restrict__ E,
                                                                           ModeC<T> mc, OpArgs<T> a, int TI,
                                                                           int tiles_k, int tiles, int chunks) {
  constexpr int NT = TJ * TK;                           // threads = owned nodes per plane
  constexpr int NJ = TJ + 2, NK = TK + 2, NS = NJ * NK;  // staged nodes (one-node halo)
  constexpr int EJ = TJ + 1, EK = TK + 1, NE = EJ * EK;  // voxels touching the tile
  constexpr int QS = (NS + NT - 1) / NT, QE = (NE + NT - 1) / NT;
  static_assert(NT % 64 == 0 && QS <= 2 && QE <= 2, "tile shape");
  constexpr unsigned SZ = sizeof(T);
  const int L = xcd_order(blockIdx.x, gridDim.x);
  const int lc = L % kRHS, tile = (L / kRHS) % tiles, chunk = (L / (kRHS * tiles)) % chunks;
  const int smp = L / (kRHS * tiles * chunks);
  const int r = smp * kRHS + lc;
  if (!a.act[r]) return;
  const int V = g.V;
  const unsigned VB = (unsigned)V * SZ;  // bytes per component plane
  const __amdgpu_buffer_rsrc_t rE = brsrc(E + (size_t)smp * V, VB);
  const T *__restrict__ Ds = a.Dinv ? a.Dinv + (size_t)smp * 3 * V : nullptr;
  const size_t off = (size_t)r * 3 * V;
  const __amdgpu_buffer_rsrc_t rU = brsrc(a.u + off, 3 * VB);
  const __amdgpu_buffer_rsrc_t rU2 = brsrc(ST == ST_PUPD ? a.u2 + off : a.u + off, 3 * VB);
  const __amdgpu_buffer_rsrc_t rD = brsrc(Ds ? Ds : a.u + off, 3 * VB);
  const T ti = a.cheb ? (T)a.cheb[smp] : (T)0, c1 = a.cheb ? (T)a.cheb[kMaxB + smp] : (T)0,
          c2 = a.cheb ? (T)a.cheb[2 * kMaxB + smp] : (T)0;
  const T bt = ST == ST_PUPD ? (T)a.beta[r] : (T)0;
  __shared__ T up[2][3 * NS];
  __shared__ T eo[NE * kEOH];
  const int tid = threadIdx.x;
  // staging slots and voxels go to threads in an order rotated by a wave per
  // workgroup, so that the second round of the extra ones does not fall on
  // the same wave -- hence the same SIMD -- in every workgroup; owned nodes
  // (the epilogue) keep the plain thread order
  const int vt = (tid + 64 * (L & (NT / 64 - 1))) & (NT - 1);
  const int j0 = (tile / tiles_k) * TJ, k0 = (tile % tiles_k) * TK;
  const int i0 = chunk * TI;
  const int iend = min(i0 + TI, g.nx);
  const int plane = g.ny * g.nz;

  // per-thread staging slots (fixed across planes): node offsets in a plane
  int soff[QS], sjc[QS][2], skc[QS][2];
  float sjw[QS], skw[QS];
#pragma unroll
  for (int q = 0; q < QS; q++) {
    const int t = vt + q * NT;
    soff[q] = -1;
    if (t < NS) {
      const int nj = t / NK, nk = t - nj * NK;
      // slots of a tile wider than the grid wrap more than once
      const int jj = ((j0 - 1 + nj) % g.ny + g.ny) % g.ny, kk = ((k0 - 1 + nk) % g.nz + g.nz) % g.nz;
      soff[q] = jj * g.nz + kk;
      if (ST == ST_PROLONG) {
        sjc[q][0] = jj >> 1;
        sjc[q][1] = ((jj + 1) >> 1) % a.gc.ny;
        sjw[q] = (jj & 1) ? 0.5f : 1.0f;
        skc[q][0] = kk >> 1;
        skc[q][1] = ((kk + 1) >> 1) % a.gc.nz;
        skw[q] = (kk & 1) ? 0.5f : 1.0f;
      }
    }
  }
  // per-thread voxels (fixed across slabs): modulus offsets in a plane
  int eoff[QE];
#pragma unroll
  for (int q = 0; q < QE; q++) {
    const int e = vt + q * NT;
    eoff[q] = -1;
    if (e < NE) {
      const int ej = e / EK, ek = e - ej * EK;
      const int jj = ((j0 - 1 + ej) % g.ny + g.ny) % g.ny, kk = ((k0 - 1 + ek) % g.nz + g.nz) % g.nz;
      eoff[q] = jj * g.nz + kk;
    }
  }

  // staging of node plane ip into LDS buffer buf.  PLAIN / DSCALE / PUPD are
  // software-pipelined: fetch() issues the loads of a plane into registers
  // one slab ahead, put() scales and stores them.  PROLONG (one sample only)
  // stages directly.
  T pf[QS][3], pd[QS][3];
  auto fetch = [&](int ip) {
    const int iw = wrap1(ip, g.nx);
#pragma unroll
    for (int q = 0; q < QS; q++) {
      if (soff[q] < 0) continue;
      const unsigned vo = (unsigned)(iw * plane + soff[q]) * SZ;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        pf[q][c] = bload(rU, vo, c * VB, (T *)nullptr);
        if (ST == ST_DSCALE) pd[q][c] = bload(rD, vo, c * VB, (T *)nullptr);
        if (ST == ST_PUPD) pd[q][c] = bload(rU2, vo, c * VB, (T *)nullptr);
      }
    }
  };
  auto put = [&](int buf) {
#pragma unroll
    for (int q = 0; q < QS; q++) {
      if (soff[q] < 0) continue;
      T *dst = &up[buf][3 * (vt + q * NT)];
#pragma unroll
      for (int c = 0; c < 3; c++)
        dst[c] = (ST == ST_DSCALE) ? pf[q][c] * (ti * pd[q][c])
                 : (ST == ST_PUPD) ? pf[q][c] + bt * pd[q][c]
                                   : pf[q][c];
    }
  };
  auto stage = [&](int buf, int ip) {
    if (ST != ST_PROLONG) {
      fetch(ip);
      put(buf);
      return;
    }
    const int iw = wrap1(ip, g.nx);
#pragma unroll
    for (int q = 0; q < QS; q++) {
      if (soff[q] < 0) continue;
      const int t = vt + q * NT;
      const int n = iw * plane + soff[q];
      T v[3];
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = a.u[off + (size_t)c * V + n];
      const int ci0 = iw >> 1, ci1 = ((iw + 1) >> 1) % a.gc.nx;
      const float wi = (iw & 1) ? 0.5f : 1.0f;
      const int ni = (iw & 1) ? 2 : 1, nj = (sjw[q] < 1.0f) ? 2 : 1, nk = (skw[q] < 1.0f) ? 2 : 1;
      const T w = (T)(wi * sjw[q] * skw[q]);
      const size_t offc = (size_t)r * 3 * a.gc.V;
      T sum[3] = {0, 0, 0};
      for (int ai = 0; ai < ni; ai++)
        for (int aj = 0; aj < nj; aj++)
          for (int ak = 0; ak < nk; ak++) {
            const int N = ((ai ? ci1 : ci0) * a.gc.ny + sjc[q][aj]) * a.gc.nz + skc[q][ak];
#pragma unroll
            for (int c = 0; c < 3; c++) sum[c] += a.xc[offc + (size_t)c * a.gc.V + N];
          }
#pragma unroll
      for (int c = 0; c < 3; c++) up[buf][3 * t + c] = v[c] + w * sum[c];
    }
  };

  const int nj = tid / TK, nk = tid % TK;  // owned node, output tile coords
  const int gj = j0 + nj, gk = k0 + nk;
  const bool owns = (gj < g.ny) && (gk < g.nz);
  const int ot = 3 * ((nj + 1) * NK + nk + 1);  // owned node in the staged plane
  DotAcc dv;

  // Per voxel (three components, corners v = 4 bi + 2 bj + bk) the cube-mode
  // transform is separable: the (j,k) transform of each face, then the i
  // stage.  A voxel's top face is the bottom face of the voxel above, so
  // its transformed values are carried to the next slab (FP32; FP64 would
  // spend too many registers and transforms the bottom face again).  On the
  // way back, the i stage splits the modal forces into a bottom and a top
  // face; the top face of slab s-1 and the bottom face of slab s lie on the
  // same node plane s, so they are summed before one inverse face transform
  // -- plane s is then complete after a single LDS exchange.
  constexpr bool kReuse = sizeof(T) == 4;
  T ftop[QE][3][4], tmod[QE][3][4];
  auto face_load = [&](int buf, int ej, int ek, T (&f)[3][4]) {
    const T *src = &up[buf][3 * (ej * NK + ek)];
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
      for (int c = 0; c < 3; c++) f[c][v] = src[3 * ((v >> 1) * NK + (v & 1)) + c];
  };
  // forward (j,k) transform per axis: (a, b) -> (a + b, b - a)
  auto face_fwd = [](T (&f)[3][4]) {
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int mb = 2; mb >= 1; mb >>= 1)
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (!(v & mb)) {
            const T a0 = f[c][v], b0 = f[c][v | mb];
            f[c][v] = a0 + b0;
            f[c][v | mb] = b0 - a0;
          }
  };
  // inverse (j,k) transform per axis: (a, b) -> (a - b, a + b)
  auto face_inv = [](T (&f)[3][4]) {
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int mb = 2; mb >= 1; mb >>= 1)
#pragma unroll
        for (int v = 0; v < 4; v++)
          if (!(v & mb)) {
            const T a0 = f[c][v], b0 = f[c][v | mb];
            f[c][v] = a0 - b0;
            f[c][v | mb] = a0 + b0;
          }
  };

  stage(0, i0 - 1);
  if (ST != ST_PROLONG) fetch(i0);
  for (int sl = i0 - 1, p = 0; sl < iend; sl++, p++) {
    const int cur = p & 1, nxt = cur ^ 1;
    const int sw = wrap1(sl, g.nx);
    // loads in the order they are consumed: moduli of this slab, epilogue
    // inputs of plane sl, then plane sl + 2 for the next slab
    T Ev[QE];
#pragma unroll
    for (int q = 0; q < QE; q++)
      Ev[q] = eoff[q] >= 0 ? bload(rE, (unsigned)(sw * plane + eoff[q]) * SZ, 0u, (T *)nullptr) : (T)0;
    const size_t nown = (size_t)sl * plane + gj * g.nz + gk;
    EpiIn<T> pin;
    if ((EP != EP_AX || a.ydinv) && sl >= i0 && owns) epi_load<T, EP>(a, Ds, off, V, (int)nown, pin);
    if (ST != ST_PROLONG) {
      put(nxt);
      if (sl + 2 <= iend) fetch(sl + 2);
    } else {
      stage(nxt, sl + 1);
    }
    // (this buffer was last read before the barrier that closed the previous
    // slab; the previous slab's gather of eo is behind that barrier too)
    __syncthreads();
    T uo[3];
    if (p > 0 && sl >= i0 && owns) {
#pragma unroll
      for (int c = 0; c < 3; c++) uo[c] = up[cur][ot + c];
    }
#pragma unroll
    for (int q = 0; q < QE; q++) {
      if (eoff[q] < 0) continue;
      const int e = vt + q * NT;
      const int ej = e / EK, ek = e - ej * EK;
      T fb[3][4], ft[3][4];
      face_load(nxt, ej, ek, ft);
      face_fwd(ft);
      if (kReuse && p > 0) {
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int v = 0; v < 4; v++) fb[c][v] = ftop[q][c][v];
      } else {
        face_load(cur, ej, ek, fb);
        face_fwd(fb);
      }
      if (kReuse) {
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int v = 0; v < 4; v++) ftop[q][c][v] = ft[c][v];
      }
      // i stage: modes sm = 4 si + s_jk
      T w[3][8];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          w[c][v] = fb[c][v] + ft[c][v];
          w[c][4 + v] = ft[c][v] - fb[c][v];
        }
      // coupling and inverse i stage: bottom-face (plane sl) and top-face
      // (plane sl+1) modal forces.  FP64 folds the two per component (fewer
      // live registers); FP32 keeps them apart, which schedules without
      // spills at 4 waves per SIMD
      T bm[3][4];
      if constexpr (sizeof(T) == 8) {
        couple_split<T>(w, Ev[q], mc, bm, tmod[q], p > 0);
      } else {
        T z[3][8];
        mode_couple<T>(w, Ev[q], mc, z);
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            bm[c][v] = z[c][v] - z[c][4 + v];
            if (p > 0) bm[c][v] += tmod[q][c][v];
            tmod[q][c][v] = z[c][v] + z[c][4 + v];
          }
      }
      if (p > 0) {
        face_inv(bm);
        T *dst = eo + e * kEOH;
#pragma unroll
        for (int v = 0; v < 4; v++)
#pragma unroll
          for (int c = 0; c < 3; c++) dst[3 * v + c] = bm[c][v];
      }
    }
    // plane i0 - 1 is not owned (uniform branch); its buffer is read above
    // and refilled by the next slab's put, so the slab still closes with a
    // barrier
    __syncthreads();
    if (p == 0) continue;
    // plane sl: voxel columns (nj+dj, nk+dk) hold this node as face corner
    // (1-dj, 1-dk)
    if (sl >= i0 && owns) {
      T Au[3] = {0, 0, 0};
#pragma unroll
      for (int dj = 0; dj < 2; dj++)
#pragma unroll
        for (int dk = 0; dk < 2; dk++) {
          const T *src = eo + ((nj + dj) * EK + nk + dk) * kEOH + 3 * (2 * (1 - dj) + (1 - dk));
#pragma unroll
          for (int c = 0; c < 3; c++) Au[c] += src[c];
        }
      if (a.nt)  // (one uniform branch per node, not per store)
        epilogue<T, EP, true>(a, pin, ti, c1, c2, off, V, (int)nown, Au, uo, dv);
      else
        epilogue<T, EP, false>(a, pin, ti, c1, c2, off, V, (int)nown, Au, uo, dv);
    }
  }
  if ((EP == EP_AX || EP == EP_STEP || EP == EP_RES || EP == EP_CHEB0) && a.dot) {
    const int pidx = chunk * tiles + tile;
    block_partial(a.red, a.slot, r, pidx, block_sum_n<NT>(dv.a));
    if (EP == EP_RES) {
      block_partial(a.red, a.slot2, r, pidx, block_sum_n<NT>(dv.b));
      block_partial(a.red, a.slot3, r, pidx, block_sum_n<NT>(dv.c));
    }
  }
}

// Workgroups of the fine operator resident on the device at once for an
// element size and tile: the occupancy of a representative variant (a
// Chebyshev step) times the CU count, looked up once per process.
template <typename T, int TJ, int TK>
static int tiled_resident_of() {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tiled<T, ST_PLAIN, EP_STEP, TJ, TK>, TJ * TK, 0) !=
          hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu < 1 ||
      cus < 1)
    throw std::runtime_error("occupancy query of the fine operator failed");
  return per_cu * cus;
}

static int tiled_resident(int sz, const TileShape &ts) {
  static std::atomic<int> cache[2][kNumTileShapes] = {};
  const int v = (ts.tj == 8 && ts.tk == 32) ? 2 : (ts.tj == 16 && ts.tk == 32) ? 3 : 1, d = sz == 8 ? 1 : 0;
  int r = cache[d][v].load(std::memory_order_relaxed);
  if (r > 0) return r;
  if (d == 0)
    r = v == 2 ? tiled_resident_of<float, 8, 32>() : v == 3 ? tiled_resident_of<float, 16, 32>()
                                                           : tiled_resident_of<float, 16, 16>();
  else
    r = v == 2 ? tiled_resident_of<double, 8, 32>() : v == 3 ? tiled_resident_of<double, 16, 32>()
                                                            : tiled_resident_of<double, 16, 16>();
  cache[d][v].store(r, std::memory_order_relaxed);
  return r;
}

template <typename T, int ST, int EP, int TJ, int TK>
static void tiled_go(const Dims &g, int R, const T *E, const ModeC<T> &mc, const OpArgs<T> &a,
                     hipStream_t s, int shape, hipEvent_t ev0, hipEvent_t ev1) {
  const int tiles_k = (g.nz + TK - 1) / TK, tiles_j = (g.ny + TJ - 1) / TJ;
  const int TI = tiled_chunk(g, R, shape, (int)sizeof(T));
  const int tiles = tiles_j * tiles_k, chunks = (g.nx + TI - 1) / TI;
  // non-temporal output stores once one output vector (all right-hand
  // sides) passes 128 MiB: 64^3 x 16 cells (302 MB in FP32) gains 1-2% from
  // them; 48^3 x 16 cells (127 MB) loses, since the next launch finds such a
  // vector still in the 256 MB last-level cache (one 64^3 cell in FP64, 38
  // MB, solves as fast either way)
  OpArgs<T> an = a;
  an.nt = (size_t)R * 3 * (size_t)g.V * sizeof(T) >= ((size_t)128 << 20) ? 1 : 0;
  if (R % kRHS) throw std::runtime_error("right-hand sides must come in groups of six");
  // buffer resources address a right-hand side's vector with 32-bit offsets
  if ((size_t)3 * g.V * sizeof(T) >= ((size_t)1 << 31)) throw std::runtime_error("grid too large for the tiled operator");
  if (ev0) {
    // timed launch: the runtime records the events at the kernel's own start
    // and end, so no queue gap is counted
    hipExtLaunchKernelGGL((k_tiled<T, ST, EP, TJ, TK>), dim3(tiles * chunks * R), dim3(TJ * TK), 0, s, ev0, ev1,
                          0, g, E, mc, an, TI, tiles_k, tiles, chunks);
    return;
  }
  hipLaunchKernelGGL((k_tiled<T, ST, EP, TJ, TK>), dim3(tiles * chunks * R), dim3(TJ * TK), 0, s, g, E, mc, an,
                     TI, tiles_k, tiles, chunks);
}

template <typename T, int ST, int EP>
static void tiled_shape(const Dims &g, int R, const T *E, const ModeC<T> &mc, const OpArgs<T> &a, hipStream_t s,
                        int shape, hipEvent_t ev0, hipEvent_t ev1) {
  const TileShape ts = tile_of(g, R, shape);
  if (ts.tj == 8 && ts.tk == 32)
    tiled_go<T, ST, EP, 8, 32>(g, R, E, mc, a, s, shape, ev0, ev1);
  else if (ts.tj == 16 && ts.tk == 32)
    tiled_go<T, ST, EP, 16, 32>(g, R, E, mc, a, s, shape, ev0, ev1);
  else
    tiled_go<T, ST, EP, 16, 16>(g, R, E, mc, a, s, shape, ev0, ev1);
}

template <typename T>
void launch_tiled(int st, int ep, const Dims &g, int R, const T *E, const ModeCoef &m,
                  const OpArgs<T> &a, hipStream_t s, int shape, hipEvent_t ev0, hipEvent_t ev1) {
  // the 9 distinct coefficients, checked against every entry of the table
  ModeC<T> mc;
  double dv[5] = {0, 0, 0, 0, 0}, xv[4] = {0, 0, 0, 0};
  for (int c = 0; c < 3; c++)
    for (int sm = 1; sm < 8; sm++) {
      dv[mode_diag_index(c, sm)] = m.diag[c][sm];
      for (int d = 0; d < 3; d++)
        if (d != c && mode_bit(sm, c) != mode_bit(sm, d)) xv[mode_cross_index(c, d, sm)] = m.cross[c][d][sm];
    }
  for (int c = 0; c < 3; c++)
    for (int sm = 1; sm < 8; sm++) {
      double err = std::fabs(dv[mode_diag_index(c, sm)] - m.diag[c][sm]);
      for (int d = 0; d < 3; d++)
        if (d != c && mode_bit(sm, c) != mode_bit(sm, d))
          err = std::fmax(err, std::fabs(xv[mode_cross_index(c, d, sm)] - m.cross[c][d][sm]));
        else if (d != c && m.cross[c][d][sm] != 0.0)
          err = 1.0;
      if (err > 1e-15 * std::fabs(dv[2])) throw std::runtime_error("mode coupling is not cube-symmetric");
    }
  for (int q = 0; q < 5; q++) mc.d[q] = (T)dv[q];
  for (int q = 0; q < 4; q++) mc.x[q] = (T)xv[q];
#define GO(S_, E_)                                     \
  if (st == S_ && ep == E_) {                          \
    tiled_shape<T, S_, E_>(g, R, E, mc, a, s, shape, ev0, ev1);  \
    return;                                            \
  }
  GO(ST_PLAIN, EP_AX)
  GO(ST_PUPD, EP_AX)
  GO(ST_PLAIN, EP_RES)
  GO(ST_DSCALE, EP_STEP)
  GO(ST_PLAIN, EP_STEP)
  GO(ST_PROLONG, EP_CHEB0)
  GO(ST_PLAIN, EP_CHEB0)
  GO(ST_DSCALE, EP_CHEB0)
#undef GO
  throw std::runtime_error("no operator kernel for this stage/epilogue pair");
}

// ---------------------------------------------------------------------------
// coarse levels: 27-point block stencils, every active RHS per thread (the
// nine coefficients of a block are loaded once for all load cases)
template <typename T, int ST>
__device__ __forceinline__ T stage_at(const OpArgs<T> &a, const T *__restrict__ Ds, T ti, size_t off,
                                      int V, int c, int n) {
  T v = a.u[off + (size_t)c * V + n];
  if (ST == ST_DSCALE) v *= ti * Ds[(size_t)c * V + n];
  return v;
}

// Stencil block `off` of node n (stencil_index layout); `tr`: as the
// transposed block seen from the neighbour, S(m, o) = S(m + o, 26 - o)^T
// (the Galerkin operator is symmetric)
template <typename T>
__device__ __forceinline__ void load_block(const T *__restrict__ Ss, int off, int V, int n, bool tr, T (&sv)[9]) {
  T w[9];
#pragma unroll
  for (int q = 0; q < 9; q++) w[q] = Ss[stencil_index(off, q, V, n)];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int d = 0; d < 3; d++) sv[3 * c + d] = tr ? w[3 * d + c] : w[3 * c + d];
}

// Coarse-level node update for NR right-hand sides of one sample (load cases
// lc0 .. lc0+NR-1, those set in `mask`): the 27 3x3 blocks of the node's
// stencil row are read once and applied to every right-hand side.  With
// `sym` the 13 lower-offset blocks are taken as the transposed upper blocks
// of the neighbours, S(n, o)[c][d] = S(n + o, -o)[d][c] (the Galerkin
// operator is symmetric), so only 14 of the 27 stored blocks are read; a
// level with a dimension of 2 (where n + o = n - o) reads all 27.
template <typename T, int ST, int EP, int NR>
__device__ __forceinline__ void stencil_node(const Dims &g, const T *__restrict__ S, const OpArgs<T> &a,
                                             int n, int smp, int lc0, unsigned mask, bool sym) {
  int i, j, k;
  node_ijk(g, n, i, j, k);
  const int I[3] = {i == 0 ? g.nx - 1 : i - 1, i, i == g.nx - 1 ? 0 : i + 1};
  const int J[3] = {j == 0 ? g.ny - 1 : j - 1, j, j == g.ny - 1 ? 0 : j + 1};
  const int K[3] = {k == 0 ? g.nz - 1 : k - 1, k, k == g.nz - 1 ? 0 : k + 1};
  const int V = g.V;
  const T *__restrict__ Ss = S + (size_t)smp * 243 * V;
  const T *__restrict__ Ds = a.Dinv ? a.Dinv + (size_t)smp * 3 * V : nullptr;
  const T ti = a.cheb ? (T)a.cheb[smp] : (T)0;
  const T c1 = a.cheb ? (T)a.cheb[kMaxB + smp] : (T)0, c2 = a.cheb ? (T)a.cheb[2 * kMaxB + smp] : (T)0;
  const size_t off0 = (size_t)(smp * kRHS + lc0) * 3 * V;
  T acc[NR][3];
#pragma unroll
  for (int t = 0; t < NR; t++) acc[t][0] = acc[t][1] = acc[t][2] = 0;
#pragma unroll 3
  for (int o = 0; o < 27; o++) {
    const int nb = (I[o / 9] * g.ny + J[(o / 3) % 3]) * g.nz + K[o % 3];
    T sv[9];
    if (sym && o < 13)
      load_block(Ss, 26 - o, V, nb, true, sv);
    else
      load_block(Ss, o, V, n, false, sv);
#pragma unroll
    for (int t = 0; t < NR; t++) {
      if (!((mask >> t) & 1u)) continue;
      const size_t off = off0 + (size_t)t * 3 * V;
      const T u0 = stage_at<T, ST>(a, Ds, ti, off, V, 0, nb), u1 = stage_at<T, ST>(a, Ds, ti, off, V, 1, nb),
              u2 = stage_at<T, ST>(a, Ds, ti, off, V, 2, nb);
      acc[t][0] += sv[0] * u0 + sv[1] * u1 + sv[2] * u2;
      acc[t][1] += sv[3] * u0 + sv[4] * u1 + sv[5] * u2;
      acc[t][2] += sv[6] * u0 + sv[7] * u1 + sv[8] * u2;
    }
  }
#pragma unroll
  for (int t = 0; t < NR; t++) {
    if (!((mask >> t) & 1u)) continue;
    const size_t off = off0 + (size_t)t * 3 * V;
    T uo[3];
#pragma unroll
    for (int c = 0; c < 3; c++) uo[c] = stage_at<T, ST>(a, Ds, ti, off, V, c, n);
    DotAcc dv;
    EpiIn<T> pin;
    epi_load<T, EP>(a, Ds, off, V, n, pin);
    epilogue<T, EP>(a, pin, ti, c1, c2, off, V, n, acc[t], uo, dv);
  }
}

// All load cases of a sample per thread (blockIdx.y = sample), reading the
// stencil once; the small levels run k_stencil_split below.
template <typename T, int ST, int EP>
__global__ __launch_bounds__(kBlock) void k_stencil(Dims g, const T *__restrict__ S, OpArgs<T> a,
                                                    int sym) {
  const int smp = blockIdx.y;
  unsigned mask = 0;
#pragma unroll
  for (int t = 0; t < kRHS; t++) mask |= a.act[smp * kRHS + t] ? (1u << t) : 0u;
  if (!mask) return;
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < g.V; n += gridDim.x * kBlock)
    stencil_node<T, ST, EP, kRHS>(g, S, a, n, smp, 0, mask, sym != 0);
}

// Small levels (a few thousand nodes: too few threads to cover the latency of
// a node's 27 dependent block loads): one right-hand side per workgroup row
// (blockIdx.y) and 64 consecutive nodes per workgroup, whose stencil rows are
// split over nine waves -- wave w = 3 di + dj applies the three blocks
// (di, dj, 0..2) -- so a thread's load chain is three blocks long.  The nine
// partial sums meet in LDS and wave 0 finishes the node.  Same block order
// and symmetric reads as stencil_node; only the summation order differs.
constexpr int kSplitNodes = 64, kSplitWaves = 9;

template <typename T, int ST, int EP>
__global__ __launch_bounds__(kSplitNodes *kSplitWaves) void k_stencil_split(Dims g, const T *__restrict__ S,
                                                                             OpArgs<T> a, int sym) {
  const int r = blockIdx.y;
  if (!a.act[r]) return;
  const int smp = r / kRHS, lane = threadIdx.x % kSplitNodes, w = threadIdx.x / kSplitNodes;
  const int n = blockIdx.x * kSplitNodes + lane;
  const bool live = n < g.V;
  const int V = g.V;
  const T *__restrict__ Ss = S + (size_t)smp * 243 * V;
  const T *__restrict__ Ds = a.Dinv ? a.Dinv + (size_t)smp * 3 * V : nullptr;
  const T ti = a.cheb ? (T)a.cheb[smp] : (T)0;
  const size_t off = (size_t)r * 3 * V;
  __shared__ T part[kSplitWaves][3][kSplitNodes];
  T acc[3] = {0, 0, 0};
  if (live) {
    int i, j, k;
    node_ijk(g, n, i, j, k);
    const int di = w / 3, dj = w % 3;
    const int row = (wrap1(i - 1 + di, g.nx) * g.ny + wrap1(j - 1 + dj, g.ny)) * g.nz;
    const int K[3] = {wrap1(k - 1, g.nz), k, wrap1(k + 1, g.nz)};
#pragma unroll
    for (int dk = 0; dk < 3; dk++) {
      const int o = 3 * w + dk, nb = row + K[dk];
      T sv[9];
      if (sym && o < 13)
        load_block(Ss, 26 - o, V, nb, true, sv);
      else
        load_block(Ss, o, V, n, false, sv);
      const T u0 = stage_at<T, ST>(a, Ds, ti, off, V, 0, nb), u1 = stage_at<T, ST>(a, Ds, ti, off, V, 1, nb),
              u2 = stage_at<T, ST>(a, Ds, ti, off, V, 2, nb);
      acc[0] += sv[0] * u0 + sv[1] * u1 + sv[2] * u2;
      acc[1] += sv[3] * u0 + sv[4] * u1 + sv[5] * u2;
      acc[2] += sv[6] * u0 + sv[7] * u1 + sv[8] * u2;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) part[w][c][lane] = acc[c];
  __syncthreads();
  if (w != 0 || !live) return;
  T Au[3], uo[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    T s = part[0][c][lane];
#pragma unroll
    for (int q = 1; q < kSplitWaves; q++) s += part[q][c][lane];
    Au[c] = s;
    uo[c] = stage_at<T, ST>(a, Ds, ti, off, V, c, n);
  }
  const T c1 = a.cheb ? (T)a.cheb[kMaxB + smp] : (T)0, c2 = a.cheb ? (T)a.cheb[2 * kMaxB + smp] : (T)0;
  DotAcc dv;
  EpiIn<T> pin;
  epi_load<T, EP>(a, Ds, off, V, n, pin);
  epilogue<T, EP>(a, pin, ti, c1, c2, off, V, n, Au, uo, dv);
}

// One-plane form of the tiled stencil kernel, for levels with planes of at
// least one 16x16 tile.  A workgroup owns one node plane i of a 16x16 (j,k)
// tile and NG load cases of one sample.  The staged neighbour planes i-1, i,
// i+1 (with the one-node j/k halo, as stage_at stages them) pass through a
// two-buffer LDS ring in three phases -- plane i+1 is loaded into registers
// while plane i-1 is used, and written over it after a barrier -- and each
// phase applies the nine 3x3 blocks of the node's stencil row that couple
// to that plane.  Against a three-plane ring this needs 2/3 of the LDS, so
// more workgroups share a CU and hide the coefficient loads' latency.

// Stencil coefficients through a raw buffer resource (bload): a scalar block
// offset plus one 32-bit lane offset per load (no vector address
// arithmetic), a node's block as two 16-byte loads and one scalar load
// (stencil_index layout), and a load past the sample's stencil returns 0
// instead of faulting.

constexpr int kStBatch = 3;  // stencil blocks whose coefficient loads are issued together

template <typename T, int ST, int EP, int NG, int SYM>
__global__ __launch_bounds__(kBlock) void k_stencil_plane(Dims g, const T *__restrict__ S, OpArgs<T> a,
                                                          int tiles_k, int tiles) {
  constexpr int G = kRHS / NG;  // load-case groups per sample
  // logical order (cell, plane, tile, group), XCD-contiguous (see xcd_order)
  const int L = xcd_order(blockIdx.x, gridDim.x);
  const int grp = L % G, tile = (L / G) % tiles, i = (L / (G * tiles)) % g.nx;
  const int smp = L / (G * tiles * g.nx), lc0 = grp * NG;
  unsigned mask = 0;
#pragma unroll
  for (int t = 0; t < NG; t++) mask |= a.act[smp * kRHS + lc0 + t] ? (1u << t) : 0u;
  if (!mask) return;
  __shared__ T us[2][NG][3][kTN * kTN];
  const int tid = threadIdx.x;
  const int j0 = (tile / tiles_k) * kTile, k0 = (tile % tiles_k) * kTile;
  const int V = g.V, plane = g.ny * g.nz;
  constexpr unsigned SZ = sizeof(T);
  const unsigned vb = (unsigned)V * SZ;  // bytes per V stencil values
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(S + (size_t)smp * 243 * V), (short)0, (int)(243u * vb), 0x00020000);
  const T *__restrict__ Ds = a.Dinv ? a.Dinv + (size_t)smp * 3 * V : nullptr;
  const T ti = a.cheb ? (T)a.cheb[smp] : (T)0;
  const T c1 = a.cheb ? (T)a.cheb[kMaxB + smp] : (T)0, c2 = a.cheb ? (T)a.cheb[2 * kMaxB + smp] : (T)0;
  const size_t off0 = (size_t)(smp * kRHS + lc0) * 3 * V;
  int soff[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int t = tid + q * kBlock;
    soff[q] = -1;
    if (t < kTN * kTN) {
      const int nj = t / kTN, nk = t - nj * kTN;
      soff[q] = wrap1(j0 - 1 + nj, g.ny) * g.nz + wrap1(k0 - 1 + nk, g.nz);
    }
  }
  T stg[2][NG][3];
  auto fetch = [&](int ip) {
    const int iw = wrap1(ip, g.nx);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if (soff[q] < 0) continue;
      const int n = iw * plane + soff[q];
      // the sample's inverse diagonal is shared by its load cases
      T sc[3] = {1, 1, 1};
      if (ST == ST_DSCALE) {
#pragma unroll
        for (int c = 0; c < 3; c++) sc[c] = ti * Ds[(size_t)c * V + n];
      }
#pragma unroll
      for (int r = 0; r < NG; r++)
        // inactive load cases stage zeros: the phases then run branch-free
#pragma unroll
        for (int c = 0; c < 3; c++)
          stg[q][r][c] = ((mask >> r) & 1u) ? a.u[off0 + (size_t)r * 3 * V + (size_t)c * V + n] * sc[c] : (T)0;
    }
  };
  auto put = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if (soff[q] < 0) continue;
      const int t = tid + q * kBlock;
#pragma unroll
      for (int r = 0; r < NG; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) us[buf][r][c][t] = stg[q][r][c];
    }
  };
  const int nj = tid / kTile, nk = tid % kTile;
  const int gj = j0 + nj, gk = k0 + nk;
  const bool owns = gj < g.ny && gk < g.nz;
  const int Jn[3] = {wrap1(gj - 1, g.ny), wrap1(gj, g.ny), wrap1(gj + 1, g.ny)};
  const int Kn[3] = {wrap1(gk - 1, g.nz), wrap1(gk, g.nz), wrap1(gk + 1, g.nz)};
  const int n = i * plane + gj * g.nz + gk;
  T acc[NG][3];
#pragma unroll
  for (int t = 0; t < NG; t++) acc[t][0] = acc[t][1] = acc[t][2] = 0;
  // the nine blocks of offset plane di (0: i-1, 1: i, 2: i+1) against the
  // staged plane in buffer `buf`
  auto phase = [&](int di, int buf) {
    const int pw = wrap1(i - 1 + di, g.nx);
#pragma unroll
    for (int pos = 0; pos < 9; pos++) {
      const int dj = pos / 3, dk = pos % 3;
      const int o = di * 9 + pos;
      const int slot = (nj + dj) * kTN + nk + dk;
      T sv[9];
      {
        // S(n, o) = S(n + o, 26 - o)^T for the lower offsets
        const bool tr = SYM && o < 13;
        const int ob = tr ? 26 - o : o;
        const unsigned node = tr ? (unsigned)(pw * plane + Jn[dj] * g.nz + Kn[dk]) : (unsigned)n;
        const unsigned base = (unsigned)(ob * 9) * vb;  // block ob (stencil_index layout)
        // (the four consecutive loads of each [V][4] group are issued by
        // the compiler as one 16-byte load)
        T w[9];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          w[q] = bload(rs, (node * 4u + q) * SZ, base, (T *)nullptr);
          w[4 + q] = bload(rs, (node * 4u + q) * SZ, base + 4u * vb, (T *)nullptr);
        }
        w[8] = bload(rs, node * SZ, base + 8u * vb, (T *)nullptr);
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int d = 0; d < 3; d++) sv[3 * c + d] = tr ? w[3 * d + c] : w[3 * c + d];
      }
#pragma unroll
      for (int t = 0; t < NG; t++) {
        const T u0 = us[buf][t][0][slot], u1 = us[buf][t][1][slot], u2 = us[buf][t][2][slot];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          T v = acc[t][c];
          v = fma(sv[3 * c], u0, v);
          v = fma(sv[3 * c + 1], u1, v);
          acc[t][c] = fma(sv[3 * c + 2], u2, v);
        }
      }
      // coefficient loads are issued kStBatch blocks at a time: the sums are
      // pinned here, so the compiler neither sinks the FMAs into the
      // epilogue nor hoists all 81 loads of the phase ahead of them (which
      // costs occupancy)
      if (pos % kStBatch == kStBatch - 1) {
#pragma unroll
        for (int t = 0; t < NG; t++)
#pragma unroll
          for (int c = 0; c < 3; c++) asm volatile("" : "+v"(acc[t][c]));
      }
    }
  };
  fetch(i - 1);
  put(0);
  fetch(i);
  put(1);
  fetch(i + 1);  // in flight during the first phase
  __syncthreads();
  if (owns) phase(0, 0);
  __syncthreads();
  put(0);
  T uo[NG][3];
  if (owns) {
    phase(1, 1);
#pragma unroll
    for (int t = 0; t < NG; t++)
#pragma unroll
      for (int c = 0; c < 3; c++) uo[t][c] = us[1][t][c][(nj + 1) * kTN + nk + 1];
  }
  __syncthreads();
  if (owns) {
    phase(2, 0);
#pragma unroll
    for (int t = 0; t < NG; t++) {
      if (!((mask >> t) & 1u)) continue;
      const size_t off = off0 + (size_t)t * 3 * V;
      DotAcc dv;
      EpiIn<T> pin;
      epi_load<T, EP>(a, Ds, off, V, n, pin);
      epilogue<T, EP>(a, pin, ti, c1, c2, off, V, n, acc[t], uo[t], dv);
    }
  }
}

template <typename T>
void launch_stencil(int st, int ep, const Dims &g, int R, const T *S, const OpArgs<T> &a,
                    hipStream_t s) {
  const int blocks = (g.V + kBlock - 1) / kBlock;
  const int nsmp = (R + kRHS - 1) / kRHS;
  const int sym = (g.nx > 2 && g.ny > 2 && g.nz > 2) ? 1 : 0;
  // planes of at least one full 16x16 tile: the one-plane LDS kernel (all six
  // load cases of a sample per workgroup in FP32, 47 KB of LDS: the stencil
  // row is read once per node; three in FP64), once that gives a workgroup
  // per CU
  constexpr int NG = sizeof(T) == 4 ? kRHS : kRHS / 2;
  const int tiles_k = (g.nz + kTile - 1) / kTile, tiles_j = (g.ny + kTile - 1) / kTile;
  const long pl_blocks = (long)tiles_j * tiles_k * g.nx * nsmp * (kRHS / NG);
  // (and a sample's stencil addressable by a buffer resource: < 2 GiB)
  const bool plane = g.ny >= kTile && g.nz >= kTile && g.nx >= 3 && R % kRHS == 0 && pl_blocks >= 256 &&
                     (size_t)243 * g.V * sizeof(T) < ((size_t)1 << 31);
  // all load cases per thread once that still leaves >= 2 workgroups per CU
  const bool multi = (long)blocks * nsmp >= 512 && R % kRHS == 0;
  const int split_blocks = (g.V + kSplitNodes - 1) / kSplitNodes;
#define GO(S_, E_)                                                                               \
  if (st == S_ && ep == E_) {                                                                    \
    if (plane && sym)                                                                            \
      hipLaunchKernelGGL((k_stencil_plane<T, S_, E_, NG, 1>), dim3((unsigned)pl_blocks), dim3(kBlock), \
                         0, s, g, S, a, tiles_k, tiles_j * tiles_k);                             \
    else if (plane)                                                                              \
      hipLaunchKernelGGL((k_stencil_plane<T, S_, E_, NG, 0>), dim3((unsigned)pl_blocks), dim3(kBlock), \
                         0, s, g, S, a, tiles_k, tiles_j * tiles_k);                             \
    else if (multi)                                                                              \
      hipLaunchKernelGGL((k_stencil<T, S_, E_>), dim3(blocks, nsmp), dim3(kBlock), 0, s, g, S,       \
                         a, sym);                                                                \
    else                                                                                         \
      hipLaunchKernelGGL((k_stencil_split<T, S_, E_>), dim3(split_blocks, R),                       \
                         dim3(kSplitNodes * kSplitWaves), 0, s, g, S, a, sym);                   \
    return;                                                                                      \
  }
  GO(ST_PLAIN, EP_AX)
  GO(ST_PLAIN, EP_RES)
  GO(ST_DSCALE, EP_STEP)
  GO(ST_PLAIN, EP_STEP)
  GO(ST_PLAIN, EP_CHEB0)
  GO(ST_DSCALE, EP_CHEB0)
#undef GO
  throw std::runtime_error("no operator kernel for this stage/epilogue pair");
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_add2(int V, const T *__restrict__ xn,
                                                 const T *__restrict__ d, T *__restrict__ x,
                                                 const int *__restrict__ act) {
  const int r = blockIdx.y;
  if (!act[r]) return;
  const size_t off = (size_t)r * 3 * V;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < 3 * V; t += gridDim.x * kBlock)
    x[off + t] = xn[off + t] + d[off + t];
}

template <typename T>
void launch_add2(int V, int R, const T *xn, const T *d, T *x, const int *act, hipStream_t s) {
  hipLaunchKernelGGL(k_add2<T>, dim3(grid_for(3 * V), R), dim3(kBlock), 0, s, V, xn, d, x, act);
}

#define FEMH_OPS_INST(T)                                                                        \
  template void launch_tiled<T>(int, int, const Dims &, int, const T *, const ModeCoef &,       \
                                const OpArgs<T> &, hipStream_t, int, hipEvent_t, hipEvent_t);   \
  template void launch_stencil<T>(int, int, const Dims &, int, const T *, const OpArgs<T> &,    \
                                  hipStream_t);                                                 \
  template void launch_add2<T>(int, int, const T *, const T *, T *, const int *, hipStream_t);

FEMH_OPS_INST(double)
FEMH_OPS_INST(float)

}  // namespace femh
