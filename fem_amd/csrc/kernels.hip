// kernels.hip -- gfx950 setup, transfer, vector and reduction kernels of the
// voxel homogenization engine (the fused operator kernels are in ops.hip).
//
// Per-RHS kernels run a grid-stride loop over one right-hand side q
// (blockIdx.y) and return at once when act[q] == 0.  Block size is 256 (four
// wave64s).  Reductions: see device_common.hpp.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "device_common.hpp"
#include "kernels.hpp"

namespace femh {

// ------------------------------------------------------------------------
// fine level, dense node gather over the 8 incident voxels (diagnostic)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fine_apply(FineOp<T> op, const T *__restrict__ u,
                                                       T *__restrict__ y, const T *__restrict__ b,
                                                       int mode, const int *__restrict__ act,
                                                       Reduce red, int slot, int dot) {
  const int r = blockIdx.y;
  if (!act[r]) return;
  const Dims g = op.g;
  const size_t off = (size_t)r * 3 * g.V;
  const T *__restrict__ ur = u + off;
  const T *__restrict__ Es = op.E + (size_t)(r / kRHS) * g.V;
  const T *__restrict__ Ke = op.Ke;
  double dv = 0;
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < g.V; n += gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(g, n, i, j, k);
    const int I[3] = {i == 0 ? g.nx - 1 : i - 1, i, i == g.nx - 1 ? 0 : i + 1};
    const int J[3] = {j == 0 ? g.ny - 1 : j - 1, j, j == g.ny - 1 ? 0 : j + 1};
    const int K[3] = {k == 0 ? g.nz - 1 : k - 1, k, k == g.nz - 1 ? 0 : k + 1};
    T acc0 = 0, acc1 = 0, acc2 = 0;
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int ax = (a >> 2) & 1, ay = (a >> 1) & 1, az = a & 1;
      const T Ee = Es[(I[1 - ax] * g.ny + J[1 - ay]) * g.nz + K[1 - az]];
      T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
      for (int bb = 0; bb < 8; bb++) {
        const int bx = (bb >> 2) & 1, by = (bb >> 1) & 1, bz = bb & 1;
        const int nb = (I[1 + bx - ax] * g.ny + J[1 + by - ay]) * g.nz + K[1 + bz - az];
        const T u0 = ur[nb], u1 = ur[g.V + nb], u2 = ur[2 * g.V + nb];
        const T *row = Ke + (3 * a) * 24 + 3 * bb;
        s0 += row[0] * u0 + row[1] * u1 + row[2] * u2;
        s1 += row[24] * u0 + row[25] * u1 + row[26] * u2;
        s2 += row[48] * u0 + row[49] * u1 + row[50] * u2;
      }
      acc0 += Ee * s0;
      acc1 += Ee * s1;
      acc2 += Ee * s2;
    }
    T *yr = y + off;
    if (mode == 1) {
      const T *br = b + off;
      acc0 = br[n] - acc0;
      acc1 = br[g.V + n] - acc1;
      acc2 = br[2 * g.V + n] - acc2;
    }
    yr[n] = acc0;
    yr[g.V + n] = acc1;
    yr[2 * g.V + n] = acc2;
    if (dot)
      dv += (double)ur[n] * (double)acc0 + (double)ur[g.V + n] * (double)acc1 +
            (double)ur[2 * g.V + n] * (double)acc2;
  }
  if (dot) {
    double s = block_sum(dv);
    block_partial(red, slot, r, blockIdx.x, s);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fine_diag(FineOp<T> op, T *__restrict__ d) {
  const Dims g = op.g;
  const int s = blockIdx.y;
  const T *Es = op.E + (size_t)s * g.V;
  T *ds = d + (size_t)s * 3 * g.V;
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < g.V; n += gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(g, n, i, j, k);
    double acc[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int ei = (i - ((a >> 2) & 1) + g.nx) % g.nx, ej = (j - ((a >> 1) & 1) + g.ny) % g.ny,
                ek = (k - (a & 1) + g.nz) % g.nz;
      const double Ee = (double)Es[(ei * g.ny + ej) * g.nz + ek];
#pragma unroll
      for (int c = 0; c < 3; c++) acc[c] += Ee * (double)op.Ke[(3 * a + c) * 25];
    }
    // a node enclosed by zero-modulus voxels carries no stiffness: unit
    // diagonal keeps Jacobi/Chebyshev finite there (its force is always 0)
#pragma unroll
    for (int c = 0; c < 3; c++)
      if (!(acc[c] > 0.0)) acc[c] = 1.0;
    ds[n] = (T)acc[0];
    ds[g.V + n] = (T)acc[1];
    ds[2 * g.V + n] = (T)acc[2];
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fine_rhs(FineOp<T> op, const double *__restrict__ gtab,
                                                     T *__restrict__ b, float *__restrict__ b32,
                                                     Reduce red, int slot) {
  const Dims g = op.g;
  const int q = blockIdx.y, lc = q % kRHS;
  const T *Es = op.E + (size_t)(q / kRHS) * g.V;
  T *br = b + (size_t)q * 3 * g.V;
  double ss = 0;
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < g.V; n += gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(g, n, i, j, k);
    double acc[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int ei = (i - ((a >> 2) & 1) + g.nx) % g.nx, ej = (j - ((a >> 1) & 1) + g.ny) % g.ny,
                ek = (k - (a & 1) + g.nz) % g.nz;
      const double Ee = (double)Es[(ei * g.ny + ej) * g.nz + ek];
#pragma unroll
      for (int c = 0; c < 3; c++) acc[c] -= Ee * gtab[lc * 24 + 3 * a + c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T v = (T)acc[c];
      br[(size_t)c * g.V + n] = v;
      if (b32) b32[(size_t)q * 3 * g.V + (size_t)c * g.V + n] = (float)v;
      ss += (double)v * (double)v;
    }
  }
  block_partial(red, slot, q, blockIdx.x, block_sum(ss));
}

// ------------------------------------------------------------------------
// coarse levels
// Level-1 Galerkin stencil straight from the fine moduli: the coarse element
// E_c (2x2x2 voxels) has matrix sum_f E_f M_f; thread per (coarse node, offset).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_build_level1(Dims fg, Dims cg, const T *__restrict__ E,
                                                         const double *__restrict__ M,
                                                         T *__restrict__ S) {
  const int o = blockIdx.y, s = blockIdx.z;
  const int di = o / 9 - 1, dj = (o / 3) % 3 - 1, dk = o % 3 - 1;
  const T *Es = E + (size_t)s * fg.V;
  T *Ss = S + (size_t)s * 243 * cg.V;
  for (int N = blockIdx.x * kBlock + threadIdx.x; N < cg.V; N += gridDim.x * kBlock) {
    int I, J, K;
    node_ijk(cg, N, I, J, K);
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int A = 0; A < 8; A++) {
      const int ax = (A >> 2) & 1, ay = (A >> 1) & 1, az = A & 1;
      const int bx = ax + di, by = ay + dj, bz = az + dk;
      if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
      const int B = 4 * bx + 2 * by + bz;
      const int ei = (I - ax + cg.nx) % cg.nx, ej = (J - ay + cg.ny) % cg.ny,
                ek = (K - az + cg.nz) % cg.nz;
      for (int f = 0; f < 8; f++) {
        const int fi = 2 * ei + ((f >> 2) & 1), fj = 2 * ej + ((f >> 1) & 1), fk = 2 * ek + (f & 1);
        const double Ef = (double)Es[(fi * fg.ny + fj) * fg.nz + fk];
        const double *m = M + f * 576 + (3 * A) * 24 + 3 * B;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int d = 0; d < 3; d++) blk[3 * c + d] += Ef * m[c * 24 + d];
      }
    }
#pragma unroll
    for (int q = 0; q < 9; q++) Ss[stencil_index(o, q, cg.V, N)] = (T)blk[q];
  }
}

// Galerkin triple product on stencils: A_c[N, N+D] = sum P[n,N] A_f[n,m] P[m,N+D]
// with fine n = 2N + o and m = n + d.  Per axis the (o, d) pairs that reach
// coarse offset D with a non-zero weight are enumerated first (at most 7 of
// the 9), so the inner loop visits only contributing (n, m) pairs.
// Weight of the pair (o, d) along one axis for coarse offset D: a coarse
// node I touches fine nodes 2I+o (P weight 1 for o = 0, 1/2 for o = +-1); the
// fine column 2I+o+d belongs to coarse node I+D with P weight 1 (2I+o+d even)
// or 1/2 (odd, shared by two coarse nodes).
__device__ __forceinline__ double pair_weight(int o, int d, int D) {
  const int t = o + d;
  double w;
  if ((t & 1) == 0)
    w = (t / 2 == D) ? 1.0 : 0.0;
  else
    w = (D == (t - 1) / 2 || D == (t + 1) / 2) ? 0.5 : 0.0;
  return o ? 0.5 * w : w;
}

// One axis of the Galerkin product P^T A P with P = Px (x) Py (x) Pz: the
// stencil `Sin` on grid gi is coarsened along AXIS into `Sout` on grid go (gi
// with that dimension halved); the other two axes are untouched.  Three passes
// (x, y, z) give the coarse operator with at most 7 terms per entry and pass.
template <typename T, int AXIS>
__global__ __launch_bounds__(kBlock) void k_galerkin_axis(Dims gi, Dims go, const T *__restrict__ Sin,
                                                          T *__restrict__ Sout) {
  const int oc = blockIdx.y, s = blockIdx.z;
  const int D = AXIS == 0 ? oc / 9 - 1 : (AXIS == 1 ? (oc / 3) % 3 - 1 : oc % 3 - 1);
  constexpr int st = AXIS == 0 ? 9 : (AXIS == 1 ? 3 : 1);
  const T *Sis = Sin + (size_t)s * 243 * gi.V;
  T *Sos = Sout + (size_t)s * 243 * go.V;
  for (int N = blockIdx.x * kBlock + threadIdx.x; N < go.V; N += gridDim.x * kBlock) {
    int i, j, k;
    node_ijk(go, N, i, j, k);
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int o = -1; o <= 1; o++) {
      int fi = i, fj = j, fk = k;
      if (AXIS == 0) fi = (2 * i + o + gi.nx) % gi.nx;
      if (AXIS == 1) fj = (2 * j + o + gi.ny) % gi.ny;
      if (AXIS == 2) fk = (2 * k + o + gi.nz) % gi.nz;
      const int n = (fi * gi.ny + fj) * gi.nz + fk;
#pragma unroll
      for (int d = -1; d <= 1; d++) {
        const double w = pair_weight(o, d, D);
        if (w == 0.0) continue;
        const int os = oc + (d - D) * st;
#pragma unroll
        for (int q = 0; q < 9; q++) blk[q] += w * (double)Sis[stencil_index(os, q, gi.V, n)];
      }
    }
#pragma unroll
    for (int q = 0; q < 9; q++) Sos[stencil_index(oc, q, go.V, N)] = (T)blk[q];
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_invert(size_t n, const T *__restrict__ d, T *__restrict__ di) {
  for (size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock)
    di[t] = (T)1 / d[t];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_stencil_diag(StencilOp<T> op, T *__restrict__ d) {
  const Dims g = op.g;
  const int s = blockIdx.y;
  const T *S = op.S + (size_t)s * 243 * g.V;
  T *ds = d + (size_t)s * 3 * g.V;
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < g.V; n += gridDim.x * kBlock) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T v = S[stencil_index(13, 4 * c, g.V, n)];
      ds[(size_t)c * g.V + n] = v > (T)0 ? v : (T)1;
    }
  }
}

// Grid transfers between a level and the next coarser one (dims exactly
// halved, periodic).  A thread owns one (j,k) column of its output level and
// a chunk of CI coarse planes: the (j,k)-weighted sum of an input plane is
// formed once and shared by the two output planes it contributes to.
// CI: up to 4 planes, fewer when that is needed for ~64K threads.
// (want > 0: that many planes, an autotuned shape; the result is the same
// for every CI, only the work split changes)
static inline int xfer_chunk(const Dims &cg, int cols, int R, int want) {
  if (want > 0) return want < cg.nx ? want : cg.nx;
  int ci = 4;
  while (ci > 1 && (long)cols * R * ((cg.nx + ci - 1) / ci) < 65536) ci /= 2;
  return ci;
}

// rc = P^T rf, full weighting (1/2, 1, 1/2) per axis
template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(Dims fg, Dims cg, const T *__restrict__ rf,
                                                     T *__restrict__ rc, const int *__restrict__ act,
                                                     int CI) {
  const int r = blockIdx.y;
  if (!act[r]) return;
  const int cols = cg.ny * cg.nz;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int chunk = t / cols, col = t - chunk * cols;
  const int I0 = chunk * CI;
  if (I0 >= cg.nx) return;
  const int I1 = min(I0 + CI, cg.nx);
  const int J = col / cg.nz, K = col - J * cg.nz;
  const int fj[3] = {J == 0 ? fg.ny - 1 : 2 * J - 1, 2 * J, 2 * J + 1};
  const int fk[3] = {K == 0 ? fg.nz - 1 : 2 * K - 1, 2 * K, 2 * K + 1};
  const T *f = rf + (size_t)r * 3 * fg.V;
  T *c = rc + (size_t)r * 3 * cg.V;
  auto plane2d = [&](int i, double (&sum)[3]) {
#pragma unroll
    for (int q = 0; q < 3; q++) {
      double acc = 0;
#pragma unroll
      for (int oj = 0; oj < 3; oj++) {
        const T *row = f + (size_t)q * fg.V + ((size_t)i * fg.ny + fj[oj]) * fg.nz;
        const double rs = 0.5 * (double)row[fk[0]] + (double)row[fk[1]] + 0.5 * (double)row[fk[2]];
        acc += (oj == 1 ? 1.0 : 0.5) * rs;
      }
      sum[q] = acc;
    }
  };
  double prev[3];
  plane2d(I0 == 0 ? fg.nx - 1 : 2 * I0 - 1, prev);
  for (int I = I0; I < I1; I++) {
    double cur[3], nxt[3];
    plane2d(2 * I, cur);
    plane2d(2 * I + 1, nxt);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      c[(size_t)q * cg.V + ((size_t)I * cg.ny + J) * cg.nz + K] = (T)(0.5 * prev[q] + cur[q] + 0.5 * nxt[q]);
      prev[q] = nxt[q];
    }
  }
}

// ef += P ec (trilinear interpolation).  Along each axis a fine node 2I takes
// coarse node I with weight 1, a fine node 2I+1 takes I and I+1 with 1/2
// each.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_prolong_add(Dims fg, Dims cg, const T *__restrict__ ec,
                                                        T *ef, const int *__restrict__ act, int CI,
                                                        T *out) {
  const int r = blockIdx.y;
  if (!act[r]) return;
  const int cols = fg.ny * fg.nz;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int chunk = t / cols, col = t - chunk * cols;
  const int I0 = chunk * CI;
  if (I0 >= cg.nx) return;
  const int I1 = min(I0 + CI, cg.nx);
  const int j = col / fg.nz, k = col - j * fg.nz;
  // every lane reads two coarse rows / columns with weights (1, 0) or
  // (1/2, 1/2): no divergence on parity
  const int cj[2] = {j >> 1, (j + 1) >> 1 == cg.ny ? 0 : (j + 1) >> 1};
  const int ck[2] = {k >> 1, (k + 1) >> 1 == cg.nz ? 0 : (k + 1) >> 1};
  const T wj[2] = {(j & 1) ? (T)0.5 : (T)1, (j & 1) ? (T)0.5 : (T)0};
  const T wk[2] = {(k & 1) ? (T)0.5 : (T)1, (k & 1) ? (T)0.5 : (T)0};
  const T *c = ec + (size_t)r * 3 * cg.V;
  const T *f = ef + (size_t)r * 3 * fg.V;
  T *fo = (out ? out : ef) + (size_t)r * 3 * fg.V;
  auto plane2d = [&](int I, T (&sum)[3]) {
#pragma unroll
    for (int q = 0; q < 3; q++) {
      T acc = 0;
#pragma unroll
      for (int b2 = 0; b2 < 2; b2++) {
        const T *row = c + (size_t)q * cg.V + ((size_t)I * cg.ny + cj[b2]) * cg.nz;
        acc += wj[b2] * (wk[0] * row[ck[0]] + wk[1] * row[ck[1]]);
      }
      sum[q] = acc;
    }
  };
  T a[3];
  plane2d(I0, a);
  for (int I = I0; I < I1; I++) {
    T b[3];
    plane2d(I + 1 == cg.nx ? 0 : I + 1, b);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const size_t e = (size_t)q * fg.V + (size_t)(2 * I) * fg.ny * fg.nz + col, e1 = e + (size_t)fg.ny * fg.nz;
      fo[e] = f[e] + a[q];
      fo[e1] = f[e1] + (T)0.5 * (a[q] + b[q]);
      a[q] = b[q];
    }
  }
}

// x[r] = Ainv[s] b[r] for the active right-hand sides r of sample s
// (blockIdx.y): one wave per row i, the row read across the wave's lanes
// (coalesced) and applied to the sample's six right-hand sides at once, FP64
// sums reduced over the wave.  Inactive right-hand sides keep their x.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dense_solve(int n, const T *__restrict__ Ainv,
                                                        const T *__restrict__ b, T *__restrict__ x,
                                                        const int *__restrict__ act) {
  const int s = blockIdx.y, lane = threadIdx.x & 63;
  const int i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  unsigned mask = 0;
#pragma unroll
  for (int t = 0; t < kRHS; t++) mask |= act[s * kRHS + t] ? (1u << t) : 0u;
  if (!mask || i >= n) return;  // (per wave: the kernel has no barrier)
  const T *__restrict__ row = Ainv + (size_t)s * n * n + (size_t)i * n;
  const T *__restrict__ bs = b + (size_t)s * kRHS * n;
  double acc[kRHS] = {0, 0, 0, 0, 0, 0};
  for (int j = lane; j < n; j += 64) {
    const double a = (double)row[j];
#pragma unroll
    for (int t = 0; t < kRHS; t++)
      if ((mask >> t) & 1u) acc[t] += a * (double)bs[(size_t)t * n + j];
  }
#pragma unroll
  for (int t = 0; t < kRHS; t++) acc[t] = wave_sum(acc[t]);
  if (lane == 0)
#pragma unroll
    for (int t = 0; t < kRHS; t++)
      if ((mask >> t) & 1u) x[(size_t)(s * kRHS + t) * n + i] = (T)acc[t];
}

// ------------------------------------------------------------------------
// vector kernels
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cheb_first(int V, const T *__restrict__ r,
                                                       const T *__restrict__ D, T *__restrict__ d,
                                                       T *__restrict__ x, double inv_theta,
                                                       int x_zero, const int *__restrict__ act) {
  const int rr = blockIdx.y;
  if (!act[rr]) return;
  const size_t off = (size_t)rr * 3 * V;
  const T *Ds = D + (size_t)(rr / kRHS) * 3 * V;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < 3 * V; t += gridDim.x * kBlock) {
    const T dv = r[off + t] * (T)inv_theta * Ds[t];
    d[off + t] = dv;
    x[off + t] = x_zero ? dv : x[off + t] + dv;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int V, T *__restrict__ x, T *__restrict__ r,
                                                       const T *__restrict__ p,
                                                       const T *__restrict__ q,
                                                       const T *__restrict__ D, T *__restrict__ z,
                                                       int rzslot, int rz_out,
                                                       const int *__restrict__ act, Reduce red,
                                                       int np_pq, int np_rz,
                                                       double *__restrict__ alpha_out,
                                                       const int *__restrict__ cnt) {
  const int rr = blockIdx.y;
  if (!act[rr]) return;
  const double pq = sum_parts(red, S_PQ, rr, np_pq);
  const double rz = sum_parts(red, rzslot, rr, np_rz);
  const double alpha = pq != 0.0 ? rz / pq : 0.0;
  // Lanczos record: row = iterations this right-hand side has completed
  if (alpha_out && blockIdx.x == 0 && threadIdx.x == 0) alpha_out[(size_t)cnt[rr] * kMaxR + rr] = alpha;
  const T a = (T)alpha;
  const size_t off = (size_t)rr * 3 * V;
  const T *Ds = D ? D + (size_t)(rr / kRHS) * 3 * V : nullptr;
  double srr = 0, srz = 0;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < 3 * V; t += gridDim.x * kBlock) {
    x[off + t] += a * p[off + t];
    const T rv = r[off + t] - a * q[off + t];
    r[off + t] = rv;
    srr += (double)rv * (double)rv;
    if (Ds) {
      const T zv = rv * Ds[t];
      z[off + t] = zv;
      srz += (double)rv * (double)zv;
    }
  }
  double s = block_sum(srr);
  block_partial(red, S_RR, rr, blockIdx.x, s);
  if (Ds) {
    s = block_sum(srz);
    block_partial(red, rz_out, rr, blockIdx.x, s);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_p_update(int V, T *__restrict__ p,
                                                     const T *__restrict__ z, int rz_new,
                                                     int rz_old, int first,
                                                     const int *__restrict__ act, Reduce red,
                                                     int np, double *__restrict__ beta_out,
                                                     const int *__restrict__ cnt) {
  const int rr = blockIdx.y;
  if (!act[rr]) return;
  double beta = 0;
  if (!first) {
    const double den = sum_parts(red, rz_old, rr, np);
    const double num = sum_parts(red, rz_new, rr, np);
    beta = den != 0.0 ? num / den : 0.0;
  }
  // (after the convergence check counted this iteration: row cnt - 1)
  if (beta_out && !first && blockIdx.x == 0 && threadIdx.x == 0) beta_out[(size_t)(cnt[rr] - 1) * kMaxR + rr] = beta;
  const T bt = (T)beta;
  const size_t off = (size_t)rr * 3 * V;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < 3 * V; t += gridDim.x * kBlock)
    p[off + t] = first ? z[off + t] : z[off + t] + bt * p[off + t];
}

// u_dst(i,j,k) = S u_src(sigma(i,j,k)), sigma and S the same axis swap, for
// the samples whose bit is set in mask
__global__ __launch_bounds__(kBlock) void k_permute_case(int n, double *__restrict__ x, int src, int dst,
                                                         int swap, unsigned mask) {
  const int smp = blockIdx.y;
  if (!((mask >> smp) & 1u)) return;
  const size_t V = (size_t)n * n * n;
  const double *us = x + ((size_t)smp * kRHS + src) * 3 * V;
  double *ud = x + ((size_t)smp * kRHS + dst) * 3 * V;
  for (size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x; t < V; t += (size_t)gridDim.x * kBlock) {
    const int k = (int)(t % n), j = (int)((t / n) % n), i = (int)(t / ((size_t)n * n));
    const size_t m = swap == 0 ? ((size_t)j * n + i) * n + k : ((size_t)k * n + j) * n + i;
    const int cx = swap == 0 ? 1 : 2, cy = swap == 0 ? 0 : 1, cz = swap == 0 ? 2 : 0;
    ud[t] = us[(size_t)cx * V + m];
    ud[V + t] = us[(size_t)cy * V + m];
    ud[2 * V + t] = us[(size_t)cz * V + m];
  }
}

__global__ __launch_bounds__(kBlock) void k_beta(int rz_new, int rz_old, int first,
                                                 const int *__restrict__ act, Reduce red, int np,
                                                 double *__restrict__ beta,
                                                 double *__restrict__ beta_out,
                                                 const int *__restrict__ cnt) {
  const int rr = blockIdx.x;
  if (!act[rr]) return;
  double b = 0;
  if (!first) {
    const double den = sum_parts(red, rz_old, rr, np);
    const double num = sum_parts(red, rz_new, rr, np);
    b = den != 0.0 ? num / den : 0.0;
  }
  if (threadIdx.x == 0) {
    beta[rr] = b;
    if (beta_out && !first) beta_out[(size_t)(cnt[rr] - 1) * kMaxR + rr] = b;
  }
}

// PCG convergence test on the device, one block per right-hand side: the
// residual norm from the S_RR partials gives relres = sqrt(r.r / b.b); a
// right-hand side that reaches tol (state 1) or turns non-finite (state 2)
// leaves the activity array, so the rest of the iteration and later
// iterations skip it.  cnt counts each right-hand side's iterations.
__global__ __launch_bounds__(kBlock) void k_pcg_check(int *__restrict__ act, Reduce red, int np,
                                                      const double *__restrict__ bb, double tol,
                                                      int *__restrict__ cnt, double *__restrict__ relres,
                                                      int *__restrict__ state) {
  const int rr = blockIdx.x;
  if (!act[rr]) return;
  const double s = sum_parts(red, S_RR, rr, np);
  if (threadIdx.x == 0) {
    cnt[rr] += 1;
    if (!(s <= 1.7976931348623157e308) || s != s) {
      state[rr] = 2;
      act[rr] = 0;
      return;
    }
    const double rel = sqrt(s / bb[rr]);
    relres[rr] = rel;
    if (rel <= tol) {
      state[rr] = 1;
      act[rr] = 0;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(int V, const T *__restrict__ a,
                                                const T *__restrict__ b, int slot,
                                                const int *__restrict__ act, Reduce red) {
  const int rr = blockIdx.y;
  if (!act[rr]) return;
  const size_t off = (size_t)rr * 3 * V;
  double s = 0;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < 3 * V; t += gridDim.x * kBlock)
    s += (double)a[off + t] * (double)b[off + t];
  s = block_sum(s);
  block_partial(red, slot, rr, blockIdx.x, s);
}

template <typename Ts, typename Td>
__global__ __launch_bounds__(kBlock) void k_convert(size_t n, const Ts *__restrict__ src,
                                                    Td *__restrict__ dst) {
  for (size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock)
    dst[t] = (Td)src[t];
}

__global__ __launch_bounds__(kBlock) void k_add_f32_to_f64(int V, const float *__restrict__ d,
                                                           double *__restrict__ x,
                                                           const int *__restrict__ act) {
  const int rr = blockIdx.y;
  if (!act[rr]) return;
  const size_t off = (size_t)rr * 3 * V;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < 3 * V; t += gridDim.x * kBlock)
    x[off + t] += (double)d[off + t];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mean_sum(int V, const T *__restrict__ x,
                                                     const int *__restrict__ act, Reduce red) {
  const int rr = blockIdx.y / 3, c = blockIdx.y % 3;
  if (!act[rr]) return;
  const T *xc = x + (size_t)rr * 3 * V + (size_t)c * V;
  double s = 0;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < V; t += gridDim.x * kBlock) s += (double)xc[t];
  s = block_sum(s);
  block_partial(red, S_AUX, blockIdx.y, blockIdx.x, s);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mean_sub(int V, T *__restrict__ x,
                                                     const int *__restrict__ act, Reduce red,
                                                     int np) {
  const int rr = blockIdx.y / 3, c = blockIdx.y % 3;
  if (!act[rr]) return;
  const T m = (T)(sum_parts(red, S_AUX, blockIdx.y, np) / (double)V);
  T *xc = x + (size_t)rr * 3 * V + (size_t)c * V;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < V; t += gridDim.x * kBlock) xc[t] -= m;
}

// zero the fields of the right-hand sides flagged in act
__global__ __launch_bounds__(kBlock) void k_zero_rhs(int V, double *__restrict__ x,
                                                     const int *__restrict__ act) {
  const int r = blockIdx.y;
  if (!act[r]) return;
  const size_t off = (size_t)r * 3 * V;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < 3 * V; t += gridDim.x * kBlock) x[off + t] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_hash(size_t n, size_t period, T *__restrict__ x,
                                                      unsigned long long seed) {
  for (size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (t % period + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[t] = (T)((double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(size_t n, T *__restrict__ x, T v) {
  for (size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock) x[t] = v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale_dinv_norm(int V, T *__restrict__ y,
                                                            const T *__restrict__ Dinv,
                                                            const int *__restrict__ act,
                                                            Reduce red) {
  const int rr = blockIdx.y;
  if (!act[rr]) return;
  const size_t off = (size_t)rr * 3 * V;
  const T *Ds = Dinv + (size_t)(rr / kRHS) * 3 * V;
  double s = 0;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < 3 * V; t += gridDim.x * kBlock) {
    const T v = y[off + t] * Ds[t];
    y[off + t] = v;
    s += (double)v * (double)v;
  }
  s = block_sum(s);
  block_partial(red, S_AUX, rr, blockIdx.x, s);
}

// ------------------------------------------------------------------------
// material, C_H, adjoint
__device__ __forceinline__ double interp_E(double rho, int interp, double E0, double Emin, double p) {
  if (interp == 0) return Emin + (E0 - Emin) * pow(rho, p);
  if (interp == 1) return E0 * rho / (3.0 - 2.0 * rho);
  return rho;
}

__device__ __forceinline__ double interp_dE(double rho, int interp, double E0, double Emin,
                                            double p) {
  if (interp == 0) return (E0 - Emin) * p * pow(rho, p - 1.0);
  if (interp == 1) {
    const double den = 3.0 - 2.0 * rho;
    return E0 * 3.0 / (den * den);
  }
  return 1.0;
}

// partials into S_AUX: idx 3s sum E, 3s+1 non-finite count, 3s+2 negative count
// cube_n > 0 (a cube grid of that edge): also counts the voxels whose modulus
// differs from that of their image under the x <-> y or the x <-> z swap by
// more than 1e-12 relative, into S_AUX[3 B + smp]
__global__ __launch_bounds__(kBlock) void k_modulus_stats(int V, const double *__restrict__ E,
                                                          int cube_n, Reduce red) {
  const int smp = blockIdx.y;
  const double *Es = E + (size_t)smp * V;
  double s = 0, nf = 0, ng = 0, na = 0;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < V; t += gridDim.x * kBlock) {
    const double e = Es[t];
    if (!isfinite(e)) nf += 1.0;
    else {
      s += e;
      if (e < 0) ng += 1.0;
    }
    if (cube_n > 0) {
      const int n = cube_n, k = t % n, j = (t / n) % n, i = t / (n * n);
      const double exy = Es[(j * n + i) * n + k], exz = Es[(k * n + j) * n + i];
      const double tol = 1e-12 * fmax(fabs(e), fmax(fabs(exy), fabs(exz)));
      if (!(fabs(e - exy) <= tol && fabs(e - exz) <= tol)) na += 1.0;
    }
  }
  double b = block_sum(s);
  block_partial(red, S_AUX, 3 * smp, blockIdx.x, b);
  b = block_sum(nf);
  block_partial(red, S_AUX, 3 * smp + 1, blockIdx.x, b);
  b = block_sum(ng);
  block_partial(red, S_AUX, 3 * smp + 2, blockIdx.x, b);
  if (cube_n > 0) {
    b = block_sum(na);
    block_partial(red, S_AUX, 3 * gridDim.y + smp, blockIdx.x, b);
  }
}

__global__ __launch_bounds__(kBlock) void k_material(size_t n, const double *__restrict__ rho,
                                                     double *__restrict__ E, int interp, double E0,
                                                     double Emin, double p) {
  for (size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock)
    E[t] = interp_E(rho[t], interp, E0, Emin, p);
}

// W_p(e) = chi_a^T K0 chi_b for the 21 pairs a <= b (row order, p = 0..20; not
// divided by V), handed to sink(p, W_p) one at a time (p is a compile-time
// constant after unrolling).  In cube modes, with w = H chi and z = Lambda w
// (K0 = H^T Lambda H, Lambda symmetric), W_ab = z_a . w_b.  The load cases go
// in three passes of two: a pass keeps z of its two cases and streams the
// later cases' w past them, so the 24 corner values of a case are gathered
// 12 times in all (not once per pair) and each case is coupled once.
__device__ __forceinline__ void chi_modes(const Dims &g, const int (&n)[8], const double *__restrict__ us,
                                          const double *__restrict__ uaff, int lc, double (&w)[3][8]) {
  const double *ul = us + (size_t)lc * 3 * g.V;
  // the corner indices pass through an empty asm: this gather's loads are
  // not hoisted above the previous case's work (which would hold all 288
  // loads of the element in registers at once)
  int m[8];
#pragma unroll
  for (int v = 0; v < 8; v++) {
    m[v] = n[v];
    asm volatile("" : "+v"(m[v]));
  }
#pragma unroll
  for (int v = 0; v < 8; v++)
#pragma unroll
    for (int c = 0; c < 3; c++) w[c][v] = uaff[lc * 24 + 3 * v + c] + ul[(size_t)c * g.V + m[v]];
  // forward: per axis (a, b) -> (a + b, b - a)
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int mb = 4; mb >= 1; mb >>= 1)
#pragma unroll
      for (int v = 0; v < 8; v++)
        if (!(v & mb)) {
          const double x0 = w[c][v], x1 = w[c][v | mb];
          w[c][v] = x0 + x1;
          w[c][v | mb] = x1 - x0;
        }
}

__device__ __forceinline__ double mode_dot(const double (&z)[3][8], const double (&w)[3][8]) {
  double s = 0;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int sm = 1; sm < 8; sm++) s = fma(z[c][sm], w[c][sm], s);  // z[c][0] = 0
  asm volatile("" : "+v"(s));  // ordered before the next gather
  return s;
}

template <typename F>
__device__ __forceinline__ void element_energy(const Dims &g, int e, const ModeT<double> &mc,
                                               const double *__restrict__ uaff,
                                               const double *__restrict__ us, F &&sink) {
  int i, j, k;
  node_ijk(g, e, i, j, k);
  const int I[2] = {i, i == g.nx - 1 ? 0 : i + 1};
  const int J[2] = {j, j == g.ny - 1 ? 0 : j + 1};
  const int K[2] = {k, k == g.nz - 1 ? 0 : k + 1};
  int n[8];
#pragma unroll
  for (int v = 0; v < 8; v++) n[v] = (I[(v >> 2) & 1] * g.ny + J[(v >> 1) & 1]) * g.nz + K[v & 1];
  auto pair = [](int a, int b) { return a * 6 - a * (a - 1) / 2 + (b - a); };
#pragma unroll
  for (int g0 = 0; g0 < 6; g0 += 2) {
    double z0[3][8], z1[3][8];
    {
      double w[3][8];
      chi_modes(g, n, us, uaff, g0, w);
      mode_couple<double>(w, 1.0, mc, z0);
      sink(pair(g0, g0), mode_dot(z0, w));
      chi_modes(g, n, us, uaff, g0 + 1, w);
      sink(pair(g0, g0 + 1), mode_dot(z0, w));
      mode_couple<double>(w, 1.0, mc, z1);
      sink(pair(g0 + 1, g0 + 1), mode_dot(z1, w));
    }
#pragma unroll
    for (int b = g0 + 2; b < 6; b++) {
      double w[3][8];
      chi_modes(g, n, us, uaff, b, w);
      sink(pair(g0, b), mode_dot(z0, w));
      sink(pair(g0 + 1, b), mode_dot(z1, w));
    }
  }
}

__global__ __launch_bounds__(kBlock, 3) void k_ch_energy(Dims g, const double *__restrict__ E,
                                                      ModeT<double> mc,
                                                      const double *__restrict__ uaff,
                                                      const double *__restrict__ u,
                                                      double *__restrict__ partials) {
  const int smp = blockIdx.y;
  const double *Es = E + (size_t)smp * g.V;
  const double *us = u + (size_t)smp * kRHS * 3 * g.V;
  __shared__ double acc[21][kBlock];  // per-thread accumulators
  const int tid = threadIdx.x;
  for (int p = 0; p < 21; p++) acc[p][tid] = 0;
  for (int e = blockIdx.x * kBlock + tid; e < g.V; e += gridDim.x * kBlock) {
    const double Ee = Es[e];
    element_energy(g, e, mc, uaff, us, [&](int p, double w) { acc[p][tid] += Ee * w; });
  }
  for (int p = 0; p < 21; p++) {
    double s = block_sum(acc[p][tid]);
    if (tid == 0) partials[((size_t)smp * 21 + p) * gridDim.x + blockIdx.x] = s;
  }
}

// Thermal stress sums: S_q = sum_e alpha_e E_e (element mean of div chi_q),
// chi_q the total field of load case q.  In the Hadamard modes of the corner
// values the x-, y- and z-difference modes (4, 2, 1) sum the corner values
// signed by their x, y and z side, and the element mean of the divergence is
// (w_x[4] + w_y[2] + w_z[1]) / 4.  Per block partials [B][6][nblk].
__global__ __launch_bounds__(kBlock) void k_thermal(Dims g, const double *__restrict__ E,
                                                    const double *__restrict__ alpha,
                                                    const double *__restrict__ uaff,
                                                    const double *__restrict__ u,
                                                    double *__restrict__ partials) {
  const int smp = blockIdx.y;
  const double *Es = E + (size_t)smp * g.V;
  const double *as = alpha + (size_t)smp * g.V;
  const double *us = u + (size_t)smp * kRHS * 3 * g.V;
  double acc[kRHS] = {0, 0, 0, 0, 0, 0};
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < g.V; e += gridDim.x * kBlock) {
    const double we = Es[e] * as[e];
    if (we == 0.0) continue;
    int i, j, k;
    node_ijk(g, e, i, j, k);
    const int I[2] = {i, i == g.nx - 1 ? 0 : i + 1};
    const int J[2] = {j, j == g.ny - 1 ? 0 : j + 1};
    const int K[2] = {k, k == g.nz - 1 ? 0 : k + 1};
    int n[8];
#pragma unroll
    for (int v = 0; v < 8; v++) n[v] = (I[(v >> 2) & 1] * g.ny + J[(v >> 1) & 1]) * g.nz + K[v & 1];
#pragma unroll
    for (int lc = 0; lc < kRHS; lc++) {
      double w[3][8];
      chi_modes(g, n, us, uaff, lc, w);
      acc[lc] += we * 0.25 * (w[0][4] + w[1][2] + w[2][1]);
    }
  }
  for (int q = 0; q < kRHS; q++) {
    const double s = block_sum(acc[q]);
    if (threadIdx.x == 0) partials[((size_t)smp * kRHS + q) * gridDim.x + blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_sum_partials(const double *__restrict__ partials,
                                                         int nblk, double *__restrict__ out) {
  const int p = blockIdx.x;  // (sample, pair)
  double s = 0;
  for (int t = threadIdx.x; t < nblk; t += kBlock) s += partials[(size_t)p * nblk + t];
  s = block_sum(s);
  if (threadIdx.x == 0) out[p] = s;
}

__global__ __launch_bounds__(kBlock, 2) void k_gradient(Dims g, const double *__restrict__ rho,
                                                     ModeT<double> mc,
                                                     const double *__restrict__ uaff,
                                                     const double *__restrict__ u,
                                                     const double *__restrict__ Gs, int interp,
                                                     double E0, double Emin, double pexp,
                                                     int energy_only, double *__restrict__ out) {
  const int smp = blockIdx.y;
  const double invV = 1.0 / (double)g.V;
  const double *us = u + (size_t)smp * kRHS * 3 * g.V;
  const double *rs = rho + (size_t)smp * g.V;
  const double *G = Gs + (size_t)smp * 21;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < g.V; e += gridDim.x * kBlock) {
    if (energy_only) {
      double *o = out + ((size_t)smp * g.V + e) * 21;
      element_energy(g, e, mc, uaff, us, [&](int p, double w) { o[p] = w * invV; });
    } else {
      double s = 0;
      element_energy(g, e, mc, uaff, us, [&](int p, double w) { s += G[p] * w; });
      out[(size_t)smp * g.V + e] = s * invV * interp_dE(rs[e], interp, E0, Emin, pexp);
    }
  }
}

// ------------------------------------------------------------------------
// projected-gradient line search with the densities on the device
// (fast_fem/run3d.cpp:238-265: x' = clamp(x - h g, lb, ub) per cell)

// per cell s (blockIdx.y): h[s] > 0: out = clamp(x - h[s] g, lb, ub); h[s] < 0:
// out = x; h[s] == 0: out unchanged.  partials[s][b] = block b's sum of the
// resulting out (the cell's density sum).  The product and the difference
// are rounded separately, as the host loop computes them, so both give the
// same trial point: contraction into a fused multiply-add is switched off
// for the kernel (the _rn intrinsics alone do not prevent it).
__global__ __launch_bounds__(kBlock) void k_descent_trial(long long V, const double *__restrict__ x,
                                                          const double *__restrict__ g,
                                                          const double *__restrict__ h, double lb, double ub,
                                                          double *__restrict__ out, double *__restrict__ partials) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  const double hs = h[s];
  const size_t o = (size_t)s * (size_t)V;
  double acc = 0;
  for (long long e = blockIdx.x * (long long)kBlock + threadIdx.x; e < V; e += (long long)gridDim.x * kBlock) {
    double v;
    if (hs > 0) {
      v = x[o + e] - hs * g[o + e];
      v = v < lb ? lb : (v > ub ? ub : v);
      out[o + e] = v;
    } else if (hs < 0) {
      v = x[o + e];
      out[o + e] = v;
    } else {
      v = out[o + e];
    }
    acc += v;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[(size_t)s * gridDim.x + blockIdx.x] = acc;
}

// max that keeps a NaN (as a host max over an array with a NaN does)
__device__ __forceinline__ double max_nan(double a, double b) { return a != a ? a : (b != b ? b : fmax(a, b)); }

// g[s] += shift[s] (when shift is given) and partials[s][b] = block b's
// max |g[s]|
__global__ __launch_bounds__(kBlock) void k_shift_absmax(long long V, double *__restrict__ g,
                                                         const double *__restrict__ shift,
                                                         double *__restrict__ partials) {
  const int s = blockIdx.y;
  const size_t o = (size_t)s * (size_t)V;
  const double sh = shift ? shift[s] : 0.0;
  double m = 0;
  for (long long e = blockIdx.x * (long long)kBlock + threadIdx.x; e < V; e += (long long)gridDim.x * kBlock) {
    double v = g[o + e];
    if (shift) {
      v += sh;
      g[o + e] = v;
    }
    m = max_nan(m, fabs(v));
  }
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) m = max_nan(m, __shfl_xor(m, w, 64));
  __shared__ double sm[kBlock / 64];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; w++) m = max_nan(m, sm[w]);
    partials[(size_t)s * gridDim.x + blockIdx.x] = m;
  }
}

// scalars[slot][idx] = sum of the np partials, for idx < nidx (host reads)
__global__ __launch_bounds__(kBlock) void k_finalize(Reduce red, int slot, int np) {
  const double s = sum_parts(red, slot, blockIdx.x, np);
  if (threadIdx.x == 0) red.scalars[slot * kSlotWidth + blockIdx.x] = s;
}

__global__ void k_reduce_init(unsigned *ctr, double *scal, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    ctr[t] = 0u;
    scal[t] = 0.0;
  }
}

// ------------------------------------------------------------------------
// launch wrappers
static inline dim3 grid2(int V, int y) { return dim3(grid_for(V), y); }

template <typename T>
void launch_dense_apply(const FineOp<T> &op, int R, const T *u, T *y, const T *b, int mode,
                        const int *act, const Reduce *red, int slot, hipStream_t s) {
  Reduce rr = red ? *red : Reduce{nullptr, nullptr, nullptr};
  hipLaunchKernelGGL(k_fine_apply<T>, grid2(op.g.V, R), dim3(kBlock), 0, s, op, u, y, b, mode, act,
                     rr, slot, red ? 1 : 0);
}

int dense_parts(const Dims &g) { return grid_for(g.V); }

template <typename T>
void launch_fine_diag(const FineOp<T> &op, int B, T *d, hipStream_t s) {
  hipLaunchKernelGGL(k_fine_diag<T>, grid2(op.g.V, B), dim3(kBlock), 0, s, op, d);
}
template <typename T>
void launch_fine_rhs(const FineOp<T> &op, int R, const double *g, T *b, float *b32, const Reduce &red,
                     int slot, hipStream_t s) {
  // vec_parts(V) blocks, the partial count the b.b reader expects (blocks
  // past the last node contribute 0)
  hipLaunchKernelGGL(k_fine_rhs<T>, dim3(vec_parts(op.g.V), R), dim3(kBlock), 0, s, op, g, b, b32, red,
                     slot);
}
template <typename T>
void launch_build_level1(Dims fine, Dims coarse, int B, const T *E, const double *M, T *S,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_build_level1<T>, dim3(grid_for(coarse.V), 27, B), dim3(kBlock), 0, s, fine,
                     coarse, E, M, S);
}
template <typename T>
void launch_galerkin(Dims fine, Dims coarse, int B, const T *Sf, T *Sc, T *t1, T *t2, hipStream_t s) {
  Dims a{coarse.nx, fine.ny, fine.nz, coarse.nx * fine.ny * fine.nz};
  Dims b{coarse.nx, coarse.ny, fine.nz, coarse.nx * coarse.ny * fine.nz};
  hipLaunchKernelGGL((k_galerkin_axis<T, 0>), dim3(grid_for(a.V), 27, B), dim3(kBlock), 0, s, fine,
                     a, Sf, t1);
  hipLaunchKernelGGL((k_galerkin_axis<T, 1>), dim3(grid_for(b.V), 27, B), dim3(kBlock), 0, s, a, b,
                     t1, t2);
  hipLaunchKernelGGL((k_galerkin_axis<T, 2>), dim3(grid_for(coarse.V), 27, B), dim3(kBlock), 0, s, b,
                     coarse, t2, Sc);
}
template <typename T>
void launch_stencil_diag(const StencilOp<T> &op, int B, T *d, hipStream_t s) {
  hipLaunchKernelGGL(k_stencil_diag<T>, grid2(op.g.V, B), dim3(kBlock), 0, s, op, d);
}
template <typename T>
void launch_restrict(Dims fine, Dims coarse, int R, const T *rf, T *rc, const int *act,
                     hipStream_t s, int planes) {
  const int cols = coarse.ny * coarse.nz, ci = xfer_chunk(coarse, cols, R, planes);
  const long n = (long)cols * ((coarse.nx + ci - 1) / ci);
  hipLaunchKernelGGL(k_restrict<T>, dim3((unsigned)((n + kBlock - 1) / kBlock), R), dim3(kBlock), 0, s,
                     fine, coarse, rf, rc, act, ci);
}
template <typename T>
void launch_prolong_add(Dims fine, Dims coarse, int R, const T *ec, T *ef, const int *act,
                        hipStream_t s, T *out, int planes) {
  const int cols = fine.ny * fine.nz, ci = xfer_chunk(coarse, cols, R, planes);
  const long n = (long)cols * ((coarse.nx + ci - 1) / ci);
  hipLaunchKernelGGL(k_prolong_add<T>, dim3((unsigned)((n + kBlock - 1) / kBlock), R), dim3(kBlock), 0, s,
                     fine, coarse, ec, ef, act, ci, out);
}
template <typename T>
void launch_dense_solve(int n, int R, const T *Ainv, const T *b, T *x, const int *act,
                        hipStream_t s) {
  if (R % kRHS) throw std::runtime_error("right-hand sides must come in groups of six");
  hipLaunchKernelGGL(k_dense_solve<T>, dim3((n + kBlock / 64 - 1) / (kBlock / 64), R / kRHS), dim3(kBlock), 0, s,
                     n, Ainv, b, x, act);
}
template <typename T>
void launch_cheb_first(int V, int R, const T *r, const T *D, T *d, T *x, double inv_theta,
                       int x_zero, const int *act, hipStream_t s) {
  hipLaunchKernelGGL(k_cheb_first<T>, grid2(3 * V, R), dim3(kBlock), 0, s, V, r, D, d, x,
                     inv_theta, x_zero, act);
}
template <typename T>
void launch_pcg_update(int V, int R, T *x, T *r, const T *p, const T *q, const T *D, T *z,
                       int rzslot, int rz_out, const int *act, const Reduce &red, int np_pq,
                       int np_rz, hipStream_t s, double *alpha_out, const int *cnt) {
  hipLaunchKernelGGL(k_pcg_update<T>, grid2(3 * V, R), dim3(kBlock), 0, s, V, x, r, p, q, D, z,
                     rzslot, rz_out, act, red, np_pq, np_rz, alpha_out, cnt);
}
template <typename T>
void launch_p_update(int V, int R, T *p, const T *z, int rz_new, int rz_old, int first,
                     const int *act, const Reduce &red, int np_rz, hipStream_t s, double *beta_out,
                     const int *cnt) {
  hipLaunchKernelGGL(k_p_update<T>, grid2(3 * V, R), dim3(kBlock), 0, s, V, p, z, rz_new, rz_old,
                     first, act, red, np_rz, beta_out, cnt);
}
void launch_pcg_check(int R, int *act, const Reduce &red, int np, const double *bb, double tol, int *cnt,
                      double *relres, int *state, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_check, dim3(R), dim3(kBlock), 0, s, act, red, np, bb, tol, cnt, relres, state);
}
void launch_permute_case(int n, int nb, double *x, int src, int dst, int swap, unsigned mask, hipStream_t s) {
  hipLaunchKernelGGL(k_permute_case, dim3(grid_for(n * n * n), nb), dim3(kBlock), 0, s, n, x, src, dst, swap,
                     mask);
}
void launch_beta(int R, int rz_new, int rz_old, int first, const int *act, const Reduce &red, int np_rz,
                 double *beta, hipStream_t s, double *beta_out, const int *cnt) {
  hipLaunchKernelGGL(k_beta, dim3(R), dim3(kBlock), 0, s, rz_new, rz_old, first, act, red, np_rz, beta,
                     beta_out, cnt);
}
template <typename T>
void launch_dot(int V, int R, const T *a, const T *b, int slot, const int *act, const Reduce &red,
                hipStream_t s) {
  hipLaunchKernelGGL(k_dot<T>, grid2(3 * V, R), dim3(kBlock), 0, s, V, a, b, slot, act, red);
}
template <typename Ts, typename Td>
void launch_convert(size_t n, const Ts *src, Td *dst, hipStream_t s) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  hipLaunchKernelGGL((k_convert<Ts, Td>), dim3((unsigned)b), dim3(kBlock), 0, s, n, src, dst);
}
template <typename T>
void launch_invert(size_t n, const T *d, T *di, hipStream_t s) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  hipLaunchKernelGGL(k_invert<T>, dim3((unsigned)b), dim3(kBlock), 0, s, n, d, di);
}
void launch_add_f32_to_f64(int V, int R, const float *d, double *x, const int *act, hipStream_t s) {
  hipLaunchKernelGGL(k_add_f32_to_f64, grid2(3 * V, R), dim3(kBlock), 0, s, V, d, x, act);
}
template <typename T>
void launch_remove_mean(int V, int R, T *x, const int *act, const Reduce &red, hipStream_t s) {
  hipLaunchKernelGGL(k_mean_sum<T>, grid2(V, 3 * R), dim3(kBlock), 0, s, V, x, act, red);
  hipLaunchKernelGGL(k_mean_sub<T>, grid2(V, 3 * R), dim3(kBlock), 0, s, V, x, act, red,
                     grid_for(V));
}
template <typename T>
void launch_fill_hash(size_t n, size_t period, T *x, unsigned long long seed, hipStream_t s) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  hipLaunchKernelGGL(k_fill_hash<T>, dim3((unsigned)b), dim3(kBlock), 0, s, n, period, x, seed);
}
template <typename T>
void launch_scale_by_dinv_norm(int V, int R, T *y, const T *Dinv, const int *act,
                               const Reduce &red, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_dinv_norm<T>, grid2(3 * V, R), dim3(kBlock), 0, s, V, y, Dinv, act,
                     red);
}
template <typename T>
void launch_fill(size_t n, T *x, T v, hipStream_t s) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)b), dim3(kBlock), 0, s, n, x, v);
}
void launch_zero_rhs(int V, int R, double *x, const int *act, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_rhs, grid2(3 * V, R), dim3(kBlock), 0, s, V, x, act);
}
void launch_material(size_t n, const double *rho, double *E, int interp, double E0, double Emin,
                     double p, hipStream_t s) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  hipLaunchKernelGGL(k_material, dim3((unsigned)b), dim3(kBlock), 0, s, n, rho, E, interp, E0, Emin,
                     p);
}
int launch_modulus_stats(int V, int B, const double *E, const Reduce &red, hipStream_t s, int cube_n) {
  hipLaunchKernelGGL(k_modulus_stats, grid2(V, B), dim3(kBlock), 0, s, V, E, cube_n, red);
  return grid_for(V);
}
static ModeT<double> mode_table(const ModeCoef &m) {
  ModeT<double> t;
  for (int c = 0; c < 3; c++)
    for (int v = 0; v < 8; v++) {
      t.diag[c][v] = m.diag[c][v];
      for (int d = 0; d < 3; d++) t.cross[c][d][v] = m.cross[c][d][v];
    }
  return t;
}
void launch_ch_energy(Dims g, int B, const double *E, const ModeCoef &m, const double *uaff,
                      const double *u, double *partials, int nblk, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_ch_energy, dim3(nblk, B), dim3(kBlock), 0, s, g, E, mode_table(m), uaff, u,
                     partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(21 * B), dim3(kBlock), 0, s, partials, nblk, out);
}
void launch_thermal(Dims g, int B, const double *E, const double *alpha, const double *uaff, const double *u,
                    double *partials, int nblk, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_thermal, dim3(nblk, B), dim3(kBlock), 0, s, g, E, alpha, uaff, u, partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(kRHS * B), dim3(kBlock), 0, s, partials, nblk, out);
}
void launch_gradient(Dims g, int B, const double *rho, const ModeCoef &m, const double *uaff,
                     const double *u, const double *Gs, int interp, double E0, double Emin,
                     double p, int energy_only, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_gradient, grid2(g.V, B), dim3(kBlock), 0, s, g, rho, mode_table(m), uaff, u,
                     Gs, interp, E0, Emin, p, energy_only, out);
}
void launch_descent_trial(int n, long long V, const double *x, const double *g, const double *h, double lb,
                          double ub, double *out, double *partials, int nblk, hipStream_t s) {
  hipLaunchKernelGGL(k_descent_trial, dim3(nblk, n), dim3(kBlock), 0, s, V, x, g, h, lb, ub, out, partials);
}
void launch_shift_absmax(int n, long long V, double *g, const double *shift, double *partials, int nblk,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_shift_absmax, dim3(nblk, n), dim3(kBlock), 0, s, V, g, shift, partials);
}
int vec_parts(int V) { return grid_for(3 * V); }

void launch_finalize(const Reduce &red, int slot, int nidx, int np, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(nidx), dim3(kBlock), 0, s, red, slot, np);
}

void reduce_init(Reduce &red, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_init, dim3(1), dim3(kBlock), 0, s, red.counters, red.scalars,
                     kNumSlots * kSlotWidth);
}

#define FEMH_INST(T)                                                                              \
  template void launch_dense_apply<T>(const FineOp<T> &, int, const T *, T *, const T *, int,     \
                                      const int *, const Reduce *, int, hipStream_t);             \
  template void launch_fine_diag<T>(const FineOp<T> &, int, T *, hipStream_t);                    \
  template void launch_fine_rhs<T>(const FineOp<T> &, int, const double *, T *, float *,          \
                                   const Reduce &, int, hipStream_t);                             \
  template void launch_build_level1<T>(Dims, Dims, int, const T *, const double *, T *,           \
                                       hipStream_t);                                              \
  template void launch_galerkin<T>(Dims, Dims, int, const T *, T *, T *, T *, hipStream_t);       \
  template void launch_stencil_diag<T>(const StencilOp<T> &, int, T *, hipStream_t);              \
  template void launch_restrict<T>(Dims, Dims, int, const T *, T *, const int *, hipStream_t, int); \
  template void launch_prolong_add<T>(Dims, Dims, int, const T *, T *, const int *, hipStream_t, T *, \
                                      int);                                                       \
  template void launch_dense_solve<T>(int, int, const T *, const T *, T *, const int *,           \
                                      hipStream_t);                                               \
  template void launch_cheb_first<T>(int, int, const T *, const T *, T *, T *, double, int,       \
                                     const int *, hipStream_t);                                   \
  template void launch_pcg_update<T>(int, int, T *, T *, const T *, const T *, const T *, T *,    \
                                     int, int, const int *, const Reduce &, int, int,             \
                                     hipStream_t, double *, const int *);                         \
  template void launch_p_update<T>(int, int, T *, const T *, int, int, int, const int *,          \
                                   const Reduce &, int, hipStream_t, double *, const int *);      \
  template void launch_dot<T>(int, int, const T *, const T *, int, const int *, const Reduce &,   \
                              hipStream_t);                                                       \
  template void launch_remove_mean<T>(int, int, T *, const int *, const Reduce &, hipStream_t);   \
  template void launch_fill_hash<T>(size_t, size_t, T *, unsigned long long, hipStream_t);                \
  template void launch_invert<T>(size_t, const T *, T *, hipStream_t);                            \
  template void launch_fill<T>(size_t, T *, T, hipStream_t);                                      \
  template void launch_scale_by_dinv_norm<T>(int, int, T *, const T *, const int *,               \
                                             const Reduce &, hipStream_t);

FEMH_INST(double)
FEMH_INST(float)
template void launch_convert<double, float>(size_t, const double *, float *, hipStream_t);
template void launch_convert<float, double>(size_t, const float *, double *, hipStream_t);
template void launch_convert<double, double>(size_t, const double *, double *, hipStream_t);
template void launch_fill<int>(size_t, int *, int, hipStream_t);

}  // namespace femh
