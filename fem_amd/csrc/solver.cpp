// solver.cpp -- host orchestration of the device solver.
//
// Method (restating desaic/fem's strain-driven periodic homogenization,
// SamplesGenerator/NumericalCoarseningUsingMultiGrid.cpp:221-313):
//   for each unit macro strain eps_lc (Voigt xx,yy,zz,yz,xz,xy):
//     solve K u_lc = -sum_e E_e K0 u_aff(eps_lc)   on periodic fluctuations
//   C_ij = 1/V sum_e E_e (u_aff,i + u_i)_e^T K0 (u_aff,j + u_j)_e
// Solver: PCG (pcgsolver/pcg_solver.h:243-285 loop) with a Jacobi or a
// Galerkin geometric-multigrid V(nu1,nu2) preconditioner (Chebyshev smoothing,
// dense coarsest solve); all load cases of all samples of a call advance
// together, each right-hand side with its own PCG scalars and stopping test.
// Mixed precision: FP32 multigrid-PCG corrections inside an FP64 iterative
// refinement loop (FP64 residual, FP64 dot products, FP64 fields and C_H).
#include "solver.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

namespace femh {

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    int st = (e == hipErrorOutOfMemory) ? FEMH_OOM : FEMH_DEVICE_ERROR;
    throw Error(st, std::string(what) + ": " + hipGetErrorString(e));
  }
}

namespace {
thread_local Arena *t_arena = nullptr;
}

Arena *current_arena() { return t_arena; }
ArenaScope::ArenaScope(Arena *a) : prev(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev; }

void *Arena::take(size_t n) {
  const size_t at = (used + 255) & ~(size_t)255;
  if (at > cap || n > cap - at)
    throw Error(FEMH_BAD_ARG, "the call needs more device scratch than the fixed region holds (" +
                                  std::to_string(at + n) + " > " + std::to_string(cap) + " bytes)");
  used = at + n;
  return base + at;
}

void DBuf::alloc(size_t n) {
  if (n <= bytes && p) return;
  release();
  if (n == 0) return;
  if (Arena *a = current_arena()) {
    p = a->take(n);
    borrowed = true;
  } else {
    FEMH_HIP(hipMalloc(&p, n));
  }
  bytes = n;
}

void DBuf::release() {
  if (p && !borrowed) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  borrowed = false;
}

// bytes of an [R][3][V] field for capacity B
static size_t field_bytes(const Dims &g, int B, size_t s) { return (size_t)kRHS * B * 3 * g.V * s; }

Context::Context(const femh_opts &opts) : o(opts) {
  if (o.nx < 2 || o.ny < 2 || o.nz < 2) throw Error(FEMH_BAD_ARG, "grid dimensions must be >= 2");
  if ((long long)o.nx * o.ny * o.nz > (1ll << 30)) throw Error(FEMH_BAD_ARG, "grid too large");
  if (!(o.nu > -1.0 && o.nu < 0.5)) throw Error(FEMH_BAD_ARG, "Poisson ratio must lie in (-1, 0.5)");
  if (!(o.E0 > 0)) throw Error(FEMH_BAD_ARG, "E0 must be positive");
  if (o.interp < 0 || o.interp > 2) throw Error(FEMH_BAD_ARG, "unknown interpolation");
  if (o.precond != FEMH_PRECOND_JACOBI && o.precond != FEMH_PRECOND_MG)
    throw Error(FEMH_BAD_ARG, "unknown preconditioner");
  if (o.precision != FEMH_FP64 && o.precision != FEMH_MIXED)
    throw Error(FEMH_BAD_ARG, "unknown precision");
  if (o.nu1 < 1 || o.nu2 < 1 || o.nu1 > kMaxCheb || o.nu2 > kMaxCheb || o.coarse_nu < 0 ||
      o.coarse_nu > kMaxCheb)
    throw Error(FEMH_BAD_ARG, "smoothing degrees must lie in [1,16]");
  if (!(o.cheby_ratio > 1.0)) throw Error(FEMH_BAD_ARG, "cheby_ratio must exceed 1");
  if (!(o.tol > 0) || o.max_iter < 1) throw Error(FEMH_BAD_ARG, "bad tolerance / iteration cap");
  if (o.apply_kernel < 0 || o.apply_kernel > 1) throw Error(FEMH_BAD_ARG, "unknown apply kernel");
  if (o.structure != FEMH_STRUCT_GENERAL && o.structure != FEMH_STRUCT_CUBIC)
    throw Error(FEMH_BAD_ARG, "unknown structure");
  if (o.structure == FEMH_STRUCT_CUBIC && !(o.nx == o.ny && o.ny == o.nz))
    throw Error(FEMH_BAD_ARG, "a cubic structure needs a cube grid");
  if (o.batch < 0 || o.batch > kMaxB)
    throw Error(FEMH_BAD_ARG, "batch must lie in [1," + std::to_string(kMaxB) + "]");
  if (o.check_every < 1) o.check_every = 1;
  B = o.batch > 0 ? o.batch : 1;
  g = Dims{o.nx, o.ny, o.nz, o.nx * o.ny * o.nz};
  FEMH_HIP(hipSetDevice(o.device));
  Arena *arena = current_arena();
  if (arena) {
    stream = arena->stream;
  } else {
    FEMH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    own_stream = true;
  }
  tab = build_tables(o.nu);
  std::memcpy(modecoef.diag, tab.L.diag, sizeof(modecoef.diag));
  std::memcpy(modecoef.cross, tab.L.cross, sizeof(modecoef.cross));

  rho.alloc(sizeof(double) * B * g.V);
  E64.alloc(sizeof(double) * B * g.V);
  Ke64.alloc(sizeof(double) * 576);
  gtab.alloc(sizeof(double) * 6 * 24);
  uaff.alloc(sizeof(double) * 6 * 24);
  Mchild.alloc(sizeof(double) * 8 * 576);
  Gs.alloc(sizeof(double) * 21 * B);
  FEMH_HIP(hipMemcpyAsync(Ke64.p, tab.Ke.data(), sizeof(double) * 576, hipMemcpyHostToDevice, stream));
  FEMH_HIP(hipMemcpyAsync(gtab.p, tab.g.data(), sizeof(double) * 144, hipMemcpyHostToDevice, stream));
  FEMH_HIP(hipMemcpyAsync(uaff.p, tab.u_aff.data(), sizeof(double) * 144, hipMemcpyHostToDevice, stream));
  FEMH_HIP(hipMemcpyAsync(Mchild.p, tab.Mchild.data(), sizeof(double) * 8 * 576, hipMemcpyHostToDevice,
                          stream));
  for (DBuf *bf : {&b, &x, &r, &z, &p, &q}) bf->alloc(field_bytes(g, B, sizeof(double)));
  if (o.precision == FEMH_MIXED)
    for (DBuf *bf : {&r32, &x32, &z32, &p32, &q32}) bf->alloc(field_bytes(g, B, sizeof(float)));
  if (o.warm_start && o.precision != FEMH_MIXED) xw.alloc(field_bytes(g, B, sizeof(double)));
  red_part.alloc(sizeof(double) * kNumSlots * kSlotWidth * kMaxPartials);
  red_ctr.alloc(sizeof(unsigned) * kNumSlots * kSlotWidth);
  red_scal.alloc(sizeof(double) * kNumSlots * kSlotWidth);
  red = Reduce{red_part.as<double>(), red_ctr.as<unsigned>(), red_scal.as<double>()};
  reduce_init(red, stream);
  ch_part.alloc(sizeof(double) * 21 * 1024 * B);
  ch_out.alloc(sizeof(double) * 21 * B);
  act.alloc(sizeof(int) * kActSlots * kMaxR);
  // timing events of every homogenize call and of the autotuner, created
  // once: no event is created or destroyed per call (none on an arena: the
  // set-up and solve times are then not measured)
  if (!arena) {
    for (auto &e : ev_call) FEMH_HIP(hipEventCreate(&e));
    for (auto &e : ev_tune) FEMH_HIP(hipEventCreate(&e));
  }
  FEMH_HIP(hipGetLastError());
  FEMH_HIP(hipStreamSynchronize(stream));
}

// pointers of every device buffer a recorded PCG iteration may touch
template <typename T>
long long Context::hier_hash(const Hierarchy<T> &h) const {
  long long hsh = 1469598103934665603ll;
  auto mix = [&](const void *p) { hsh = (hsh ^ (long long)(uintptr_t)p) * 1099511628211ll; };
  for (const Level<T> &L : h.lv)
    for (const DBuf *b : {&L.S, &L.D, &L.Dinv, &L.x, &L.b, &L.r, &L.d, &L.d2, &L.xn, &L.Ad}) mix(b->p);
  for (const DBuf *b : {&h.E, &h.Ke, &h.Ainv, &h.cheb, &palt, &palt32, &beta_d, &lanczos, &pcg_state, &act, &red_part})
    mix(b->p);
  return hsh;
}

Context::~Context() {
  for (auto &g : graphs_)
    if (g.second) (void)hipGraphExecDestroy(g.second);
  for (auto &pe : pending) {
    (void)hipEventDestroy(pe.a);
    (void)hipEventDestroy(pe.b);
  }
  for (auto e : ev_pool) (void)hipEventDestroy(e);
  for (auto e : ev_call)
    if (e) (void)hipEventDestroy(e);
  for (auto e : ev_tune)
    if (e) (void)hipEventDestroy(e);
  if (stream && own_stream) (void)hipStreamDestroy(stream);
}

void Context::set_act(int slot, const std::vector<int> &v) {
  int h[kMaxR] = {0};
  for (size_t t = 0; t < v.size() && t < (size_t)kMaxR; t++) h[t] = v[t] ? 1 : 0;
  // pageable source: the copy completes before the call returns, so the
  // stack buffer may go away afterwards
  FEMH_HIP(hipMemcpyAsync(act.as<int>() + slot * kMaxR, h, sizeof(h), hipMemcpyHostToDevice, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
}

void Context::read_scalars(int slot, double *out, int n, int np) {
  launch_finalize(red, slot, n, np, stream);
  FEMH_HIP(hipGetLastError());
  FEMH_HIP(hipMemcpyAsync(out, red.scalars + slot * kSlotWidth, sizeof(double) * n,
                          hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
}

void Context::load(const double *src, int n, bool device_src, bool modulus) {
  if (n < 1 || n > B) throw Error(FEMH_BAD_ARG, "sample count exceeds the context's batch");
  nb = n;
  R = kRHS * n;
  const size_t bytes = sizeof(double) * (size_t)n * g.V;
  const hipMemcpyKind kind = device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  FEMH_HIP(hipMemcpyAsync(rho.p, src, bytes, kind, stream));
  if (modulus)
    FEMH_HIP(hipMemcpyAsync(E64.p, src, bytes, kind, stream));
  else
    launch_material((size_t)n * g.V, rho.as<double>(), E64.as<double>(), o.interp, o.E0, o.Emin,
                    o.penal, stream);
  FEMH_HIP(hipGetLastError());
  modulus_input = modulus;
  h64.ready = h32.ready = false;
  have_fields = false;
}

// ---------------------------------------------------------------------------
// Algorithmic HBM bytes of one fine-operator launch: every distinct per-RHS
// vector the launch reads or writes, once per active right-hand side (3
// components each), plus per active sample the moduli and, where the launch
// uses it, the inverse diagonal (3 components).  A vector that is staged and
// also read by the epilogue (a.u == a.b or a.u == a.rprev) counts once;
// outputs count only where the variant writes them (ops.hpp).
template <typename T>
double tiled_alg_bytes(int st, int ep, const OpArgs<T> &a, double V, double sz, double Rn, double ns) {
  const void *in[5] = {a.u, nullptr, nullptr, nullptr, nullptr};
  int nin = 1;
  auto add_in = [&](const void *p) {
    if (!p) return;
    for (int i = 0; i < nin; i++)
      if (in[i] == p) return;
    in[nin++] = p;
  };
  if (st == ST_PUPD) add_in(a.u2);
  if (ep == EP_RES || ep == EP_CHEB0) add_in(a.b);
  if (ep == EP_STEP) {
    add_in(a.rprev);
    add_in(a.xprev);
    if (a.dot) add_in(a.b);
  }
  double vec_out = 0;
  if (ep == EP_AX) vec_out = 1 + (a.pout ? 1 : 0);
  if (ep == EP_RES) vec_out = 1;
  if (ep == EP_CHEB0) vec_out = a.fin1 ? 1 : (a.last ? 2 : 3);
  if (ep == EP_STEP) vec_out = a.last ? 1 : 3;
  double vec = nin + vec_out;
  if (st == ST_PROLONG) vec += 0.125;  // coarse correction, 1/8 of the nodes
  const bool usesD = st == ST_DSCALE || ep == EP_CHEB0 || ep == EP_STEP || (ep == EP_AX && a.ydinv);
  return V * (3.0 * sz * Rn * vec + ns * sz * (usesD ? 4.0 : 1.0));
}

// operator dispatch (with optional kernel timing of the fine level)
template <typename T>
void Context::op(Hierarchy<T> &h, int l, int st, int ep, OpArgs<T> a) {
  const Level<T> &L = h.lv[l];
  if (!a.Dinv) a.Dinv = L.Dinv.template as<T>();
  if (l == 0) {
    const Dims &g0 = L.g;
    const bool dense = o.apply_kernel == 1 && st == ST_PLAIN && (ep == EP_AX || ep == EP_RES);
    auto launch = [&](hipEvent_t e0, hipEvent_t e1) {
      if (dense) {
        FineOp<T> f{g0, h.E.template as<T>(), h.Ke.template as<T>()};
        launch_dense_apply<T>(f, R, a.u, a.y, a.b, ep == EP_RES ? 1 : 0, a.act,
                              a.dot ? &a.red : nullptr, a.slot, stream);
        if (a.ydinv) launch_scale_by_dinv_norm<T>(g0.V, R, a.y, a.Dinv, a.act, red, stream);
      } else {
        launch_tiled<T>(st, ep, g0, R, h.E.template as<T>(), modecoef, a, stream, shape_fine, e0, e1);
      }
    };
    if (prof) {
      Pending pe;
      pe.a = take_event();
      pe.b = take_event();
      pe.which = sizeof(T) == 8 ? 0 : 1;
      int na = 0, ns = 0;
      {
        int hact[kMaxR];
        FEMH_HIP(hipMemcpyAsync(hact, a.act, sizeof(int) * R, hipMemcpyDeviceToHost, stream));
        FEMH_HIP(hipStreamSynchronize(stream));
        for (int t = 0; t < R; t++) na += hact[t] != 0;
        for (int s0 = 0; s0 < R; s0 += kRHS) {
          bool any = false;
          for (int t = s0; t < s0 + kRHS && t < R; t++) any = any || hact[t] != 0;
          ns += any;
        }
      }
      const double Rn = na, V = (double)g0.V, sz = sizeof(T);
      pe.bytes = tiled_alg_bytes(st, ep, a, V, sz, Rn, ns);
      pe.flops = V * Rn * 1152.0;
      if (dense) {
        FEMH_HIP(hipEventRecord(pe.a, stream));
        launch(nullptr, nullptr);
        FEMH_HIP(hipEventRecord(pe.b, stream));
      } else {
        launch(pe.a, pe.b);  // events at the kernel's own start and end
      }
      pending.push_back(pe);
    } else {
      launch(nullptr, nullptr);
    }
  } else {
    launch_stencil<T>(st, ep, L.g, R, L.S.template as<T>(), a, stream);
  }
}

// r = b - K x in FP64 with r.r, b.x and x.r of every flagged RHS into S_RR,
// S_BX and S_XR (one fused launch on the tiled path); returns their partial
// count
int Context::residual_dots(const int *act_) {
  if (o.apply_kernel == 0) {
    OpArgs<double> a;
    a.u = x.as<double>();
    a.y = r.as<double>();
    a.b = b.as<double>();
    a.act = act_;
    a.red = red;
    a.slot = S_RR;
    a.slot2 = S_BX;
    a.slot3 = S_XR;
    a.dot = 1;
    op<double>(h64, 0, ST_PLAIN, EP_RES, a);
    return tiled_parts(g, R, shape_fine, (int)sizeof(double));
  }
  apply<double>(h64, 0, x.as<double>(), r.as<double>(), b.as<double>(), 1, act_, nullptr, 0);
  launch_dot<double>(g.V, R, r.as<double>(), r.as<double>(), S_RR, act_, red, stream);
  launch_dot<double>(g.V, R, b.as<double>(), x.as<double>(), S_BX, act_, red, stream);
  launch_dot<double>(g.V, R, x.as<double>(), r.as<double>(), S_XR, act_, red, stream);
  return vec_parts(g.V);
}

int Context::apply_parts(int sz) const {
  return o.apply_kernel == 1 ? dense_parts(g) : tiled_parts(g, R, shape_fine, sz);
}

template <typename T>
void Context::apply(Hierarchy<T> &h, int l, const T *u, T *y, const T *bb, int mode,
                    const int *act_, const Reduce *rd, int slot) {
  OpArgs<T> a;
  a.u = u;
  a.y = y;
  a.b = bb;
  a.act = act_;
  if (rd) {
    a.red = *rd;
    a.slot = slot;
    a.dot = 1;
  }
  op<T>(h, l, ST_PLAIN, mode == 1 ? EP_RES : EP_AX, a);
}

// ---------------------------------------------------------------------------
// hierarchy
template <typename T>
void Context::build_levels(Hierarchy<T> &h) {
  // level dimensions: halve while every dimension is even and >= 4
  std::vector<Dims> dims{g};
  int maxL = (o.precond == FEMH_PRECOND_MG) ? (o.levels > 0 ? o.levels : 32) : 1;
  while ((int)dims.size() < maxL) {
    Dims d = dims.back();
    if ((d.nx & 1) || (d.ny & 1) || (d.nz & 1) || d.nx < 4 || d.ny < 4 || d.nz < 4) break;
    Dims c{d.nx / 2, d.ny / 2, d.nz / 2, 0};
    c.V = c.nx * c.ny * c.nz;
    dims.push_back(c);
  }
  h.lv.resize(dims.size());
  for (size_t l = 0; l < dims.size(); l++) {
    Level<T> &L = h.lv[l];
    L.g = dims[l];
    L.D.alloc(sizeof(T) * 3 * B * L.g.V);
    L.Dinv.alloc(sizeof(T) * 3 * B * L.g.V);
    L.lmax.assign(B, 0.0);
    // V-cycle work vectors (a one-level hierarchy -- Jacobi -- needs none;
    // the autotuner and apply_level take theirs on first use)
    if (dims.size() > 1)
      for (DBuf *bf : {&L.x, &L.b, &L.r, &L.d, &L.d2, &L.xn, &L.Ad})
        bf->alloc(field_bytes(L.g, B, sizeof(T)));
    if (l > 0) L.S.alloc(sizeof(T) * 243 * B * (size_t)L.g.V);
  }
  if (dims.size() > 2) {
    h.gt1.alloc(sizeof(T) * 243 * B * (size_t)dims[1].V / 2);
    h.gt2.alloc(sizeof(T) * 243 * B * (size_t)dims[1].V / 4);
  }
  h.E.alloc(sizeof(T) * B * g.V);
  h.Ke.alloc(sizeof(T) * 576);
}

constexpr double kLmaxScale = 1.3;  // safety factor on the power-iteration lambda_max

template <typename T>
void Context::estimate_lmax(Hierarchy<T> &h) {
  // power iteration on D^-1 A, one vector per sample (RHS 6 s)
  std::vector<int> pw(R, 0);
  for (int s = 0; s < nb; s++) pw[kRHS * s] = 1;
  set_act(A_POW, pw);
  const int *ap = actp(A_POW);
  for (size_t l = 0; l < h.lv.size(); l++) {
    Level<T> &L = h.lv[l];
    const int V = L.g.V;
    T *xv = L.x.template as<T>(), *yv = L.Ad.template as<T>();
    // every sample starts from the same vector (batched == one-sample results)
    launch_fill_hash<T>((size_t)R * 3 * V, (size_t)kRHS * 3 * V, xv, 0x51ED27ull + l, stream);
    // y_k = (D^-1 A)^k x_0 without normalising (ten steps grow it by at most
    // lambda^10 <= 1e10 or so), one fused launch per step; the estimate
    // ||y_10|| / ||y_9|| is that of the normalised iteration.  It approaches
    // lambda_max from below, and a Chebyshev residual polynomial on
    // [lmax/ratio, lmax] exceeds 1 in magnitude beyond lmax (1 + 1/ratio): an
    // eigenvalue past that point makes the smoother, and so the V-cycle,
    // indefinite (PCG then meets r.z < 0).  Ten steps have underestimated
    // lambda_max by more than 15% on binary cells with a 1e-9 void phase and
    // on SIMP cells with the degree-2 smoother; the estimate is scaled by 1.3.
    // one right-hand side per cell is active: the fine-level launches take
    // a chunk length sized for those workgroups alone (no dot is fused in
    // them, so no partial count depends on it)
    struct Restore {
      int &ref, val;
      ~Restore() { ref = val; }
    } restore{shape_fine, shape_fine};
    if (l == 0 && o.apply_kernel == 0) shape_fine = tiled_sparse_shape(g, R, nb, shape_fine, (int)sizeof(T));
    for (int it = 0; it < 10; it++) {
      OpArgs<T> a;
      a.u = xv;
      a.y = yv;
      a.act = ap;
      a.ydinv = 1;
      op<T>(h, (int)l, ST_PLAIN, EP_AX, a);
      std::swap(xv, yv);
    }
    launch_dot<T>(V, R, yv, yv, S_AUX, ap, red, stream);  // y_9
    launch_dot<T>(V, R, xv, xv, S_RR, ap, red, stream);   // y_10
    FEMH_HIP(hipGetLastError());
    std::vector<double> n9(R), n10(R);
    read_scalars(S_AUX, n9.data(), R, vec_parts(V));
    read_scalars(S_RR, n10.data(), R, vec_parts(V));
    for (int s = 0; s < nb; s++) {
      const double a9 = n9[kRHS * s], a10 = n10[kRHS * s];
      L.lmax[s] = kLmaxScale * std::sqrt(a9 > 0 ? std::max(a10, 0.0) / a9 : 0.0);
      if (!(L.lmax[s] > 0) || !std::isfinite(L.lmax[s]))
        throw Error(FEMH_NONFINITE, "operator is not finite");
    }
  }
}

// Dense inverse of each sample's regularised coarsest operator A + s Z Z^T
// (Z = the three constant displacement fields): for right-hand sides
// orthogonal to Z it returns the zero-mean solution of A x = b.
void Context::dense_coarse(Hierarchy<double> &h) {
  Level<double> &L = h.lv.back();
  const int V = L.g.V, n = 3 * V;
  h.ndense = 0;
  if (n > 648 || h.lv.size() < 2) return;
  std::vector<double> Sh((size_t)243 * V * nb);
  FEMH_HIP(hipMemcpyAsync(Sh.data(), L.S.p, sizeof(double) * Sh.size(), hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
  std::vector<double> inv((size_t)n * n * nb);
  std::vector<double> A((size_t)n * n), col(n);
  const Dims &d = L.g;
  for (int smp = 0; smp < nb; smp++) {
    const double *S = Sh.data() + (size_t)smp * 243 * V;
    std::fill(A.begin(), A.end(), 0.0);
    for (int N = 0; N < V; N++) {
      const int k = N % d.nz, j = (N / d.nz) % d.ny, i = N / (d.nz * d.ny);
      for (int off = 0; off < 27; off++) {
        const int ni = (i + off / 9 - 1 + d.nx) % d.nx, nj = (j + (off / 3) % 3 - 1 + d.ny) % d.ny,
                  nk = (k + off % 3 - 1 + d.nz) % d.nz;
        const int M = (ni * d.ny + nj) * d.nz + nk;
        for (int c = 0; c < 3; c++)
          for (int dd = 0; dd < 3; dd++)
            A[(size_t)(c * V + N) * n + dd * V + M] += S[stencil_index(off, 3 * c + dd, V, N)];
      }
    }
    double dmean = 0;
    for (int t = 0; t < n; t++) dmean += A[(size_t)t * n + t];
    dmean /= n;
    for (int c = 0; c < 3; c++)
      for (int a = 0; a < V; a++)
        for (int bb = 0; bb < V; bb++) A[(size_t)(c * V + a) * n + c * V + bb] += dmean / V;
    // Cholesky A = L L^T (lower, in place)
    for (int j = 0; j < n; j++) {
      double s = A[(size_t)j * n + j];
      for (int k = 0; k < j; k++) s -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
      if (!(s > 0)) throw Error(FEMH_NONFINITE, "coarsest operator is not positive definite");
      s = std::sqrt(s);
      A[(size_t)j * n + j] = s;
      for (int i = j + 1; i < n; i++) {
        double t = A[(size_t)i * n + j];
        for (int k = 0; k < j; k++) t -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
        A[(size_t)i * n + j] = t / s;
      }
    }
    double *Ai = inv.data() + (size_t)smp * n * n;
    for (int c = 0; c < n; c++) {
      std::fill(col.begin(), col.end(), 0.0);
      col[c] = 1.0;
      for (int i = 0; i < n; i++) {
        double s = col[i];
        for (int k = 0; k < i; k++) s -= A[(size_t)i * n + k] * col[k];
        col[i] = s / A[(size_t)i * n + i];
      }
      for (int i = n - 1; i >= 0; i--) {
        double s = col[i];
        for (int k = i + 1; k < n; k++) s -= A[(size_t)k * n + i] * col[k];
        col[i] = s / A[(size_t)i * n + i];
      }
      for (int i = 0; i < n; i++) Ai[(size_t)i * n + c] = col[i];
    }
  }
  h.Ainv.alloc(sizeof(double) * (size_t)n * n * B);
  FEMH_HIP(hipMemcpyAsync(h.Ainv.p, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice,
                          stream));
  FEMH_HIP(hipStreamSynchronize(stream));
  h.ndense = n;
}

// level operators of h from its fine moduli h.E (built in h's precision; the
// Galerkin kernels accumulate in FP64): fine diagonal, level-1 stencil from
// the moduli, Galerkin stencils below, diagonals and their inverses
template <typename T>
void Context::build_ops(Hierarchy<T> &h) {
  FineOp<T> f0{g, h.E.template as<T>(), h.Ke.template as<T>()};
  launch_fine_diag<T>(f0, nb, h.lv[0].D.template as<T>(), stream);
  for (int l = 1; l < nlevels; l++) {
    Level<T> &L = h.lv[l];
    if (l == 1)
      launch_build_level1<T>(g, L.g, nb, h.E.template as<T>(), Mchild.as<double>(), L.S.template as<T>(),
                             stream);
    else
      launch_galerkin<T>(h.lv[l - 1].g, L.g, nb, h.lv[l - 1].S.template as<T>(), L.S.template as<T>(),
                         h.gt1.template as<T>(), h.gt2.template as<T>(), stream);
    StencilOp<T> so{L.g, L.S.template as<T>()};
    launch_stencil_diag<T>(so, nb, L.D.template as<T>(), stream);
  }
  for (int l = 0; l < nlevels; l++)
    launch_invert<T>((size_t)3 * nb * h.lv[l].g.V, h.lv[l].D.template as<T>(), h.lv[l].Dinv.template as<T>(),
                     stream);
  FEMH_HIP(hipGetLastError());
}

// Launch-shape autotuner.  For each (grid, right-hand sides, precision,
// levels) it times the fine operator (y = A u on every right-hand side) for a
// few node-plane depths per workgroup, and the grid transfers (restriction
// plus prolongation) of every level pair for a few planes per thread, on the
// context's own buffers and stream, and keeps the fastest.  A candidate
// within 2% of the built-in choice leaves the built-in choice in place.
// Results are cached for the process, so later contexts of the same shape
// take them without timing.
namespace {
std::mutex g_tune_mu;
std::map<std::array<int, 7>, std::vector<int>> g_tuned;
}  // namespace

template <typename T>
void Context::autotune(Hierarchy<T> &h) {
  const int L = (int)h.lv.size();
  const std::array<int, 7> key{g.nx, g.ny, g.nz, R, (int)sizeof(T), L, o.apply_kernel};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      shape_fine = it->second[0];
      shape_xfer.assign(it->second.begin() + 1, it->second.end());
      return;
    }
  }
  const bool was_prof = prof;
  prof = false;
  std::vector<int> all(R, 1);
  set_act(A_POW, all);
  hipEvent_t e0 = ev_tune[0], e1 = ev_tune[1];
  auto time_ms = [&](auto &&body, int reps) {
    body();  // warm
    FEMH_HIP(hipEventRecord(e0, stream));
    for (int k = 0; k < reps; k++) body();
    FEMH_HIP(hipEventRecord(e1, stream));
    FEMH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    FEMH_HIP(hipEventElapsedTime(&ms, e0, e1));
    return (double)ms / reps;
  };
  auto pick = [](const std::vector<int> &cand, const std::vector<double> &t) {
    // cand[0] is the built-in choice (0)
    size_t best = 0;
    for (size_t i = 1; i < cand.size(); i++)
      if (t[i] < t[best]) best = i;
    return t[best] < 0.98 * t[0] ? cand[best] : 0;
  };
  for (int l = 0; l < L; l++) {
    Level<T> &Lv = h.lv[l];
    Lv.d.alloc(field_bytes(Lv.g, B, sizeof(T)));
    Lv.Ad.alloc(field_bytes(Lv.g, B, sizeof(T)));
    launch_fill_hash<T>((size_t)R * 3 * Lv.g.V, (size_t)kRHS * 3 * Lv.g.V, Lv.d.template as<T>(), 0x7A11ull + l,
                        stream);
  }
  FEMH_HIP(hipGetLastError());
  shape_fine = 0;
  shape_xfer.assign(L > 1 ? L - 1 : 0, 0);
  if (o.apply_kernel == 0) {
    // tile variants x node planes per workgroup (a shape equivalent to an
    // earlier candidate -- same tile, same planes -- is not timed twice)
    std::vector<int> cand{0};
    const int sz = (int)sizeof(T);
    std::vector<std::pair<int, int>> seen{{tiled_default_tile(g, R), tiled_slabs(g, R, 0, sz)}};
    for (int tile = 1; tile < kNumTileShapes; tile++)
      for (int c : {0, 2, 4, 8, 16, 32}) {
        const int code = tiled_shape_code(tile, c);
        if (c > g.nx || (c > 0 && tiled_slabs(g, R, code, sz) != c)) continue;
        const std::pair<int, int> key{tile, tiled_slabs(g, R, code, sz)};
        if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
        seen.push_back(key);
        cand.push_back(code);
      }
    std::vector<double> t;
    for (int c : cand) {
      shape_fine = c;
      OpArgs<T> a;
      a.u = h.lv[0].d.template as<T>();
      a.y = h.lv[0].Ad.template as<T>();
      a.act = actp(A_POW);
      t.push_back(time_ms([&] { op<T>(h, 0, ST_PLAIN, EP_AX, a); }, 3));
    }
    shape_fine = pick(cand, t);
  }
  for (int l = 0; l + 1 < L; l++) {
    Level<T> &F = h.lv[l], &C = h.lv[l + 1];
    std::vector<int> cand{0, 1, 2, 4, 8};
    std::vector<double> t;
    for (int c : cand)
      t.push_back(time_ms(
          [&] {
            launch_restrict<T>(F.g, C.g, R, F.d.template as<T>(), C.Ad.template as<T>(), actp(A_POW), stream, c);
            launch_prolong_add<T>(F.g, C.g, R, C.d.template as<T>(), F.Ad.template as<T>(), actp(A_POW), stream,
                                  nullptr, c);
          },
          3));
    shape_xfer[l] = pick(cand, t);
  }
  FEMH_HIP(hipGetLastError());
  prof = was_prof;
  std::vector<int> val{shape_fine};
  val.insert(val.end(), shape_xfer.begin(), shape_xfer.end());
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned.emplace(key, val);
}

void Context::build_hierarchy() {
  build_levels<double>(h64);
  nlevels = (int)h64.lv.size();
  FEMH_HIP(hipMemcpyAsync(h64.E.p, E64.p, sizeof(double) * nb * g.V, hipMemcpyDeviceToDevice, stream));
  FEMH_HIP(hipMemcpyAsync(h64.Ke.p, Ke64.p, sizeof(double) * 576, hipMemcpyDeviceToDevice, stream));
  h64_ops = false;
  if (o.precision != FEMH_MIXED) {
    if (o.autotune) autotune<double>(h64);
    build_ops<double>(h64);
    h64_ops = true;
    if (o.precond == FEMH_PRECOND_MG) {
      estimate_lmax(h64);
      upload_cheb(h64);
      // the coarsest level is inverted densely when it is small enough, else
      // solved approximately by a long Chebyshev polynomial (cheb_run)
      if (nlevels > 1) dense_coarse(h64);
    }
  } else {
    // mixed precision: the smoothing hierarchy is built in FP32 directly;
    // the FP64 fine operator (residuals) needs only h64.E / h64.Ke
    build_levels<float>(h32);
    launch_convert<double, float>((size_t)nb * g.V, E64.as<double>(), h32.E.as<float>(), stream);
    launch_convert<double, float>(576, Ke64.as<double>(), h32.Ke.as<float>(), stream);
    if (o.autotune) autotune<float>(h32);
    build_ops<float>(h32);
    if (o.precond == FEMH_PRECOND_MG) {
      if (nlevels > 1) {
        // dense coarsest inverse of the FP32 coarsest operator (host Cholesky
        // in FP64 on its widened copy)
        const Level<float> &Lc = h32.lv.back();
        launch_convert<float, double>((size_t)243 * nb * Lc.g.V, Lc.S.as<float>(), h64.lv.back().S.as<double>(),
                                      stream);
        dense_coarse(h64);
        if (h64.ndense) {
          h32.Ainv.alloc(sizeof(float) * (size_t)h64.ndense * h64.ndense * B);
          launch_convert<double, float>((size_t)h64.ndense * h64.ndense * nb, h64.Ainv.as<double>(),
                                        h32.Ainv.as<float>(), stream);
          h32.ndense = h64.ndense;
        }
      }
      estimate_lmax(h32);
      upload_cheb(h32);
      for (int l = 0; l < nlevels; l++) h64.lv[l].lmax = h32.lv[l].lmax;
    }
    FEMH_HIP(hipGetLastError());
    h32.ready = true;
  }
  h64.ready = true;
}

void Context::ensure_h64_ops() {
  if (h64_ops) return;
  build_ops<double>(h64);
  // (spectrum bounds stay those of the FP32 operators)
  if (o.precond == FEMH_PRECOND_MG && nlevels > 1) dense_coarse(h64);
  if (o.precond == FEMH_PRECOND_MG) upload_cheb(h64);
  h64_ops = true;
}

// ---------------------------------------------------------------------------
// multigrid V-cycle with Chebyshev smoothing (fused kernels, see ops.hpp)
// Chebyshev iteration for D^-1 A on [lmax/ratio, lmax]: 1/theta and the
// coefficients (c1, c2) of step `step` (1-based) of the three-term recurrence
// d_s = c1 d_{s-1} + c2 D^-1 r_s.
static void cheb_step(double lmax, double ratio, int step, double *theta_inv, double *c1, double *c2) {
  const double hi = lmax, lo = lmax / ratio;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  *theta_inv = 1.0 / theta;
  double rho = 1.0 / sigma;
  *c1 = *c2 = 0.0;
  for (int s = 1; s <= step; s++) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    *c1 = rn * rho;
    *c2 = 2.0 * rn / delta;
    rho = rn;
  }
}

// the device table of Chebyshev coefficients of h (after its spectrum
// estimates): per level, step and interval, 1/theta and (from step 1 on) c1
// and c2 of every sample
template <typename T>
void Context::upload_cheb(Hierarchy<T> &h) {
  const int L = (int)h.lv.size();
  const int steps = std::max(kMaxCheb, std::max(2, o.coarse_iters)) + 1;
  const size_t per = (size_t)3 * kMaxB;
  std::vector<double> tab((size_t)L * steps * 2 * per, 0.0);
  for (int l = 0; l < L; l++)
    for (int st = 0; st < steps; st++)
      for (int kind = 0; kind < 2; kind++) {
        double *t = tab.data() + (((size_t)l * steps + st) * 2 + kind) * per;
        for (int s = 0; s < nb; s++) {
          double ti, c1, c2;
          cheb_step(h.lv[l].lmax[s], kind ? kCoarseChebRatio : o.cheby_ratio, st, &ti, &c1, &c2);
          t[s] = ti;
          t[kMaxB + s] = st >= 1 ? c1 : 0.0;
          t[2 * kMaxB + s] = st >= 1 ? c2 : 0.0;
        }
      }
  h.cheb.alloc(sizeof(double) * tab.size());
  FEMH_HIP(hipMemcpyAsync(h.cheb.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
  h.cheb_steps = steps;
}

// per-sample 1/theta and the coefficients of Chebyshev step `step` (1-based):
// the launch reads them from the hierarchy's device table
template <typename T>
void Context::set_cheb(const Hierarchy<T> &h, int l, OpArgs<T> &a, int step, double ratio) {
  const int kind = ratio == o.cheby_ratio ? 0 : 1;
  if (kind == 1 && ratio != kCoarseChebRatio) throw Error(FEMH_DEVICE_ERROR, "no Chebyshev table for this interval");
  if (!h.cheb.p || step < 0 || step >= h.cheb_steps || l >= (int)h.lv.size())
    throw Error(FEMH_DEVICE_ERROR, "Chebyshev coefficients not prepared");
  a.cheb = h.cheb.template as<double>() + (((size_t)l * h.cheb_steps + step) * 2 + kind) * 3 * kMaxB;
}

// k Chebyshev steps on level l from x = 0 (pre-smoothing, and the coarsest
// solve when that level is too large for the dense inverse)
template <typename T>
void Context::cheb_run(Hierarchy<T> &h, int l, const T *bb, T *xx, const int *act_, int k, double ratio) {
  Level<T> &L = h.lv[l];
  if (k == 1) {
    // x = b / (theta D) as one fused DSCALE pass with c1 = c2 = 0
    OpArgs<T> a;
    a.act = act_;
    a.u = bb;
    a.rprev = bb;
    a.r = L.r.template as<T>();
    a.x = xx;
    a.dout = L.d.template as<T>();
    a.last = 1;  // c1 = c2 = 0: x = staged value
    set_cheb(h, l, a, 0, ratio);
    op<T>(h, l, ST_DSCALE, EP_STEP, a);
    return;
  }
  T *dcur = L.d.template as<T>(), *dnxt = L.d2.template as<T>();
  OpArgs<T> a;
  a.act = act_;
  a.u = bb;
  a.rprev = bb;
  a.xprev = nullptr;
  a.r = L.r.template as<T>();
  a.x = xx;
  a.dout = dcur;
  a.last = (k == 2);
  set_cheb(h, l, a, 1, ratio);
  op<T>(h, l, ST_DSCALE, EP_STEP, a);
  for (int s = 2; s < k; s++) {
    a.u = dcur;
    a.rprev = L.r.template as<T>();
    a.xprev = xx;
    a.dout = dnxt;
    a.last = (s == k - 1);
    set_cheb(h, l, a, s, ratio);
    op<T>(h, l, ST_PLAIN, EP_STEP, a);
    std::swap(dcur, dnxt);
  }
}

template <typename T>
void Context::pre_smooth(Hierarchy<T> &h, int l, const T *bb, T *xx, const int *act_) {
  cheb_run<T>(h, l, bb, xx, act_, nu_pre(l), o.cheby_ratio);
}

// Returns true when the last launch also left the partials of b . x in
// dot_slot (fine level, degree >= 2).
template <typename T>
bool Context::post_smooth(Hierarchy<T> &h, int l, const T *bb, T *xx, const int *act_, int dot_slot) {
  Level<T> &L = h.lv[l];
  Level<T> &C = h.lv[l + 1];
  const int k = nu_post(l);
  OpArgs<T> a;
  a.act = act_;
  a.u = xx;
  a.b = bb;
  a.r = L.r.template as<T>();
  a.dout = L.d2.template as<T>();
  a.x = L.xn.template as<T>();
  set_cheb(h, l, a, 0, o.cheby_ratio);
  if (k == 1 && l == 0 && R > kRHS) {
    // degree 1 at the fine level: xn = x + P xc out of place, then one launch
    // leaves the smoothed x (and, fused, the partials of b . x = r . z)
    launch_prolong_add<T>(L.g, C.g, R, C.x.template as<T>(), xx, act_, stream, L.xn.template as<T>(),
                          xfer_shape(l));
    a.u = L.xn.template as<T>();
    a.x = xx;
    a.fin1 = 1;
    if (dot_slot >= 0) {
      a.red = red;
      a.slot = dot_slot;
      a.dot = 1;
    }
    op<T>(h, l, ST_PLAIN, EP_CHEB0, a);
    return dot_slot >= 0;
  }
  if (l == 0 && R <= kRHS) {
    // one sample: staging x + P xc inside the operator saves a launch; with
    // more right-hand sides the recomputed interpolation costs more than the
    // separate pass
    a.xc = C.x.template as<T>();
    a.gc = C.g;
    op<T>(h, l, ST_PROLONG, EP_CHEB0, a);  // x3 = x + P xc staged, kept in xn
  } else {
    launch_prolong_add<T>(L.g, C.g, R, C.x.template as<T>(), xx, act_, stream, nullptr, xfer_shape(l));
    op<T>(h, l, ST_PLAIN, EP_CHEB0, a);
  }
  if (k == 1) {
    launch_add2<T>(L.g.V, R, L.xn.template as<T>(), L.d2.template as<T>(), xx, act_, stream);
    return false;
  }
  T *dcur = L.d2.template as<T>(), *dnxt = L.d.template as<T>();
  OpArgs<T> s1;
  s1.act = act_;
  s1.r = L.r.template as<T>();
  s1.rprev = L.r.template as<T>();
  s1.x = xx;
  const bool fuse = l == 0 && dot_slot >= 0;
  for (int s = 1; s < k; s++) {
    s1.u = dcur;
    s1.xprev = (s == 1) ? L.xn.template as<T>() : xx;
    s1.dout = dnxt;
    s1.last = (s == k - 1);
    if (s1.last && fuse) {
      s1.b = bb;
      s1.red = red;
      s1.slot = dot_slot;
      s1.dot = 1;
    }
    set_cheb(h, l, s1, s, o.cheby_ratio);
    op<T>(h, l, ST_PLAIN, EP_STEP, s1);
    std::swap(dcur, dnxt);
  }
  return fuse;
}

template <typename T>
bool Context::vcycle(Hierarchy<T> &h, int l, const T *bb, T *xx, const int *act_, int dot_slot) {
  const int last = (int)h.lv.size() - 1;
  if (l == last) {
    if (h.ndense)
      launch_dense_solve<T>(h.ndense, R, h.Ainv.template as<T>(), bb, xx, act_, stream);
    else
      cheb_run<T>(h, l, bb, xx, act_, std::max(2, o.coarse_iters), kCoarseChebRatio);
    return false;
  }
  Level<T> &L = h.lv[l];
  Level<T> &C = h.lv[l + 1];
  if (nu_pre(l) == 1) {
    // degree-1 pre-smoothing x = b / (theta D) needs no operator apply: one
    // launch stages it and leaves the residual b - A x for the restriction
    OpArgs<T> a;
    a.act = act_;
    a.u = bb;
    a.b = bb;
    a.r = L.r.template as<T>();
    a.dout = L.d.template as<T>();
    a.x = xx;
    a.last = 1;
    set_cheb(h, l, a, 0, o.cheby_ratio);
    op<T>(h, l, ST_DSCALE, EP_CHEB0, a);
  } else {
    pre_smooth<T>(h, l, bb, xx, act_);
    apply<T>(h, l, xx, L.r.template as<T>(), bb, 1, act_, nullptr, 0);
  }
  launch_restrict<T>(L.g, C.g, R, L.r.template as<T>(), C.b.template as<T>(), act_, stream, xfer_shape(l));
  vcycle<T>(h, l + 1, C.b.template as<T>(), C.x.template as<T>(), act_, -1);
  return post_smooth<T>(h, l, bb, xx, act_, dot_slot);
}

// z = M^-1 r, and r . z into rz_slot
template <typename T>
void Context::precond(Hierarchy<T> &h, const T *rv, T *zv, const int *act_, int rz_slot) {
  bool fused = false;
  if (o.precond == FEMH_PRECOND_MG && h.lv.size() > 1) {
    fused = vcycle<T>(h, 0, rv, zv, act_, rz_slot);
  } else {
    // Jacobi: inside the PCG update kernel; this path serves the first step
    launch_cheb_first<T>(h.lv[0].g.V, R, rv, h.lv[0].Dinv.template as<T>(), zv, zv, 1.0, 1, act_,
                         stream);
  }
  if (fused != fuses_rz(h)) throw Error(FEMH_DEVICE_ERROR, "inconsistent r.z reduction path");
  if (!fused) launch_dot<T>(h.lv[0].g.V, R, rv, zv, rz_slot, act_, red, stream);
}

// the V-cycle's last fine-level smoothing launch also reduces r . z
// (degree >= 2, or degree 1 on the multi-sample path)
template <typename T>
bool Context::fuses_rz(const Hierarchy<T> &h) const {
  return o.precond == FEMH_PRECOND_MG && h.lv.size() > 1 && (o.nu2 >= 2 || (o.nu2 == 1 && R > kRHS));
}

template <typename T>
int Context::rz_parts(const Hierarchy<T> &h) const {
  return fuses_rz(h) ? tiled_parts(g, R, shape_fine, (int)sizeof(T)) : vec_parts(g.V);
}

// PCG over all right-hand sides flagged in `active` (updated in place as
// they converge).
template <typename T>
void Context::pcg(Hierarchy<T> &h, const T *bvec, T *xv, T *rv, T *zv, T *pv, T *qv,
                  std::vector<int> &active, double tol, int max_iter, const double *bnorm2,
                  int *iters, double *relres, bool z_ready) {
  const int V = g.V;
  const bool jac = !(o.precond == FEMH_PRECOND_MG && h.lv.size() > 1);
  const T *D0 = h.lv[0].Dinv.template as<T>();
  launch_fill<T>(field_bytes(g, nb, sizeof(T)) / sizeof(T), xv, (T)0, stream);
  // z_ready: the caller also left r = b in rv
  if (!z_ready)
    FEMH_HIP(hipMemcpyAsync(rv, bvec, field_bytes(g, nb, sizeof(T)), hipMemcpyDeviceToDevice, stream));
  for (int t = 0; t < R; t++) {
    iters[t] = 0;
    relres[t] = 0.0;
  }
  int nact = 0;
  for (int t = 0; t < R; t++) nact += active[t] != 0;
  if (!nact) return;
  set_act(A_PCG, active);
  const int *ap = actp(A_PCG);
  const int np_rz = jac ? vec_parts(V) : rz_parts(h);
  // z_ready: the caller already left z = M^-1 b and r.z in S_RZ0
  if (!z_ready) precond<T>(h, rv, zv, ap, S_RZ0);
  // Fine-level tiled operator: the direction update p = z + beta p runs
  // inside the A p launch (staged on the fly, written back by the owner),
  // ping-ponging p between pv and a second buffer since neighbours still
  // read the old direction.  beta comes from a one-block-per-RHS kernel.
  const bool fuse_p = !jac && o.apply_kernel != 1;
  T *pcur = pv, *pnxt = nullptr;
  if (fuse_p) {
    DBuf &alt = sizeof(T) == 8 ? palt : palt32;
    if (!alt.p) alt.alloc(field_bytes(g, B, sizeof(T)));
    if (!beta_d.p) beta_d.alloc(sizeof(double) * kMaxR);
    pnxt = alt.template as<T>();
    launch_beta(R, S_RZ0, S_RZ0, 1, ap, red, np_rz, beta_d.as<double>(), stream);
  } else {
    launch_p_update<T>(V, R, pv, zv, S_RZ0, S_RZ0, 1, ap, red, np_rz, stream);
  }
  // Per-RHS iteration state on the device: iterations done, relative
  // residual, state (1 converged, 2 non-finite), and |b|^2.  The convergence
  // test runs on the device after every update, so an iteration is a fixed
  // launch sequence that needs no host decision.
  pcg_state.alloc(sizeof(double) * 4 * kMaxR);
  int *cnt = pcg_state.as<int>(), *pst = cnt + kMaxR;
  double *relres_d = pcg_state.as<double>() + kMaxR, *bb_d = relres_d + kMaxR;
  launch_fill<int>(2 * kMaxR, cnt, 0, stream);
  FEMH_HIP(hipMemcpyAsync(bb_d, bnorm2, sizeof(double) * R, hipMemcpyHostToDevice, stream));
  int *apw = act.as<int>() + A_PCG * kMaxR;
  const int np_rr = vec_parts(V);
  // one PCG iteration; `cur` selects the r.z parity, `first` the plain p = z
  auto iteration = [&](int cur, bool first) {
    const int rzc = cur ? S_RZ1 : S_RZ0, rzn = cur ? S_RZ0 : S_RZ1;
    if (fuse_p) {
      OpArgs<T> a;
      a.u = zv;
      a.u2 = first ? zv : pcur;  // beta = 0 on the first step: p = z
      a.beta = beta_d.as<double>();
      a.pout = pnxt;
      a.y = qv;
      a.act = ap;
      a.red = red;
      a.slot = S_PQ;
      a.dot = 1;
      op<T>(h, 0, ST_PUPD, EP_AX, a);
      std::swap(pcur, pnxt);
    } else {
      apply<T>(h, 0, pv, qv, nullptr, 0, ap, &red, S_PQ);
    }
    launch_pcg_update<T>(V, R, xv, rv, pcur, qv, jac ? D0 : nullptr, zv, rzc, rzn, ap, red, apply_parts((int)sizeof(T)), np_rz,
                         stream, lz_alpha, cnt);
    launch_pcg_check(R, apw, red, np_rr, bb_d, tol, cnt, relres_d, pst, stream);
    if (!jac) precond<T>(h, rv, zv, ap, rzn);
    if (fuse_p)
      launch_beta(R, rzn, rzc, 0, ap, red, np_rz, beta_d.as<double>(), stream, lz_beta, cnt);
    else
      launch_p_update<T>(V, R, pv, zv, rzn, rzc, 0, ap, red, np_rz, stream, lz_beta, cnt);
  };
  std::vector<int> hact(kMaxR), hst(kMaxR);
  // host look at the device state: false when no right-hand side is left.
  // Only k_pcg_check clears a right-hand side's activity, and it sets the
  // state (1 or 2) when it does, so the state alone tells which are left.
  auto poll = [&]() {
    FEMH_HIP(hipMemcpyAsync(hst.data(), pst, sizeof(int) * R, hipMemcpyDeviceToHost, stream));
    FEMH_HIP(hipStreamSynchronize(stream));
    bool any = false;
    for (int t = 0; t < R; t++) {
      if (hst[t] == 2) throw Error(FEMH_NONFINITE, "residual became non-finite");
      any = any || (active[t] != 0 && hst[t] == 0);
    }
    return any;
  };
  // Iterations 2k and 2k+1 use the same buffers and slots (ping-pong of the
  // direction and of the r.z parity), so with graphs on, the pair is
  // recorded once into a hipGraph and replayed; the host looks at the device
  // state after every pair (or every check_every iterations without graphs).
  FEMH_HIP(hipGetLastError());
  int it = 1;
  iteration(0, true);
  bool live = it < max_iter ? poll() : false;
  const bool graphs = o.graph && !prof;
  hipGraphExec_t exec = nullptr;
  if (graphs && live && max_iter - it >= 2) {
    GraphKey key{};
    key.v = {(long long)sizeof(T), (long long)R, (long long)(uintptr_t)xv, (long long)(uintptr_t)rv,
             (long long)(uintptr_t)zv, (long long)(uintptr_t)pv, (long long)(uintptr_t)qv,
             (long long)(uintptr_t)lz_alpha, (long long)(uintptr_t)h.cheb.p, (long long)h.lv.size(),
             (long long)h.ndense, (long long)shape_fine, (long long)xfer_hash(), (long long)z_ready,
             hier_hash(h)};
    std::memcpy(&key.tol, &tol, sizeof(double));
    exec = find_graph(key);
    if (!exec) {
      hipGraph_t graph = nullptr;
      FEMH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        iteration(1, false);
        iteration(0, false);
      } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(stream, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        throw;
      }
      FEMH_HIP(hipStreamEndCapture(stream, &graph));
      const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      FEMH_HIP(ie);
      graphs_.push_back({key, exec});
    }
  }
  while (live && it < max_iter) {
    if (exec && max_iter - it >= 2) {
      FEMH_HIP(hipGraphLaunch(exec, stream));
      it += 2;
      live = poll();
      continue;
    }
    iteration(it & 1, false);
    it++;
    if (it % o.check_every == 0 || it == max_iter) live = poll();
  }
  if (live) poll();
  // results of the run: iterations and residual per right-hand side, and the
  // activity that remains (not converged within max_iter)
  std::vector<int> hcnt(R);
  std::vector<double> hrel(R);
  FEMH_HIP(hipMemcpyAsync(hcnt.data(), cnt, sizeof(int) * R, hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipMemcpyAsync(hrel.data(), relres_d, sizeof(double) * R, hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipMemcpyAsync(hact.data(), apw, sizeof(int) * R, hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
  for (int t = 0; t < R; t++) {
    if (!active[t]) continue;
    iters[t] = hcnt[t];
    relres[t] = hrel[t];
    active[t] = hact[t] != 0;
  }
  FEMH_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
hipEvent_t Context::take_event() {
  if (!ev_pool.empty()) {
    hipEvent_t e = ev_pool.back();
    ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  FEMH_HIP(hipEventCreate(&e));
  return e;
}

void Context::set_profiling(bool on) {
  flush_profile();
  prof = on;
  kst[0] = KStat{};
  kst[1] = KStat{};
}

void Context::flush_profile() {
  if (pending.empty()) return;
  FEMH_HIP(hipStreamSynchronize(stream));
  for (Pending &pe : pending) {
    float ms = 0;
    FEMH_HIP(hipEventElapsedTime(&ms, pe.a, pe.b));
    KStat &k = kst[pe.which];
    k.launches++;
    k.ms += ms;
    k.bytes += pe.bytes;
    k.flops += pe.flops;
    ev_pool.push_back(pe.a);
    ev_pool.push_back(pe.b);
  }
  pending.clear();
}

void Context::compute_ch(const double *fields, double *C) {
  const int nblk = 1024;
  launch_ch_energy(g, nb, E64.as<double>(), modecoef, uaff.as<double>(), fields,
                   ch_part.as<double>(), nblk, ch_out.as<double>(), stream);
  FEMH_HIP(hipGetLastError());
  std::vector<double> s(21 * nb);
  FEMH_HIP(hipMemcpyAsync(s.data(), ch_out.p, sizeof(double) * s.size(), hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
  for (int smp = 0; smp < nb; smp++) {
    int p = 0;
    double *Cs = C + 36 * smp;
    for (int a = 0; a < 6; a++)
      for (int bb = a; bb < 6; bb++) {
        Cs[6 * a + bb] = Cs[6 * bb + a] = s[21 * smp + p] / (double)g.V;
        p++;
      }
  }
}

// Smallest eigenvalue of the k x k Lanczos matrix of a PCG run from its
// coefficients a_j = alpha_j (j = 0..k-1) and b_j = beta_j (j = 0..k-2),
// stored with stride `stride`:
//   T(j,j) = 1/a_j + b_{j-1}/a_{j-1},  T(j,j+1) = sqrt(b_j)/a_j
// (the Lanczos-CG relation, e.g. Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed., section 6.7.3).  Its eigenvalues (Ritz values) lie in the
// spectrum of the preconditioned operator M^-1 K, and the smallest one
// approaches lambda_min(M^-1 K) from above as k grows.  Bisection on the
// Sturm count of T - x I.  Returns NaN when the coefficients are not those of
// a positive definite run (breakdown, rounding).
static double lanczos_min(const double *a, const double *b, size_t stride, int k) {
  if (k < 1) return NAN;
  std::vector<double> d(k), e2(k > 1 ? k - 1 : 1, 0.0);
  double hi = 0;
  for (int j = 0; j < k; j++) {
    const double aj = a[(size_t)j * stride];
    if (!(aj > 0) || !std::isfinite(aj)) return NAN;
    d[j] = 1.0 / aj;
    if (j > 0) d[j] += b[(size_t)(j - 1) * stride] / a[(size_t)(j - 1) * stride];
    if (j + 1 < k) {
      const double bj = b[(size_t)j * stride];
      if (!(bj >= 0) || !std::isfinite(bj)) return NAN;
      e2[j] = bj / (aj * aj);
    }
  }
  for (int j = 0; j < k; j++) {
    double r = d[j];
    if (j > 0) r += std::sqrt(e2[j - 1]);
    if (j + 1 < k) r += std::sqrt(e2[j]);
    hi = std::max(hi, r);
  }
  // eigenvalues of T - x I below 0: negative pivots of its LDL^T
  auto below = [&](double x) {
    int cnt = 0;
    double piv = 1.0;
    for (int j = 0; j < k; j++) {
      piv = (d[j] - x) - (j > 0 ? e2[j - 1] / piv : 0.0);
      if (piv == 0.0) piv = -1e-300;
      cnt += piv < 0;
    }
    return cnt;
  };
  double lo = 0.0;
  if (below(lo) > 0) return NAN;
  for (int it = 0; it < 200 && hi - lo > 1e-12 * hi; it++) {
    const double mid = 0.5 * (lo + hi);
    if (below(mid) >= 1)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

void Context::homogenize(double *C, femh_stats *st) {
  if (nb < 1) throw Error(FEMH_BAD_ARG, "no input loaded");
  // warm start: each slot begins from the fields of the previous call with
  // the same sample count (valid fields are kept only after a finite result)
  const bool warm = o.warm_start && prev_nb == nb;
  prev_nb = 0;
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t *ev = ev_call;
  const bool timed = ev[0] != nullptr;
  if (timed) FEMH_HIP(hipEventRecord(ev[0], stream));
  have_fields = false;
  // validate the moduli before any device work depends on them
  // cubic structure: a cell takes the two-solve path only if its moduli are
  // invariant under the axis swaps (counted on the device); other cells of
  // the batch are solved with all six load cases
  const bool cubic = o.structure == FEMH_STRUCT_CUBIC;
  std::vector<double> mst(4 * nb);
  std::vector<char> cubic_ok(nb, 0);
  {
    const int np = launch_modulus_stats(g.V, nb, E64.as<double>(), red, stream, cubic ? g.nx : 0);
    read_scalars(S_AUX, mst.data(), (cubic ? 4 : 3) * nb, np);
    if (cubic)
      for (int s = 0; s < nb; s++) cubic_ok[s] = mst[3 * nb + s] == 0.0;
  }
  for (int s = 0; s < nb; s++) {
    if (mst[3 * s + 1] > 0) throw Error(FEMH_NONFINITE, "non-finite modulus");
    if (mst[3 * s + 2] > 0) throw Error(FEMH_BAD_ARG, "negative modulus");
    if (!(mst[3 * s] > 0)) throw Error(FEMH_BAD_ARG, "all-void cell (every modulus is zero)");
  }
  std::vector<int> all(R, 1);
  set_act(A_ALL, all);
  build_hierarchy();
  FineOp<double> f0{g, E64.as<double>(), Ke64.as<double>()};
  // a cold mixed-precision solve starts its first FP32 sweep from r = b:
  // the right-hand side kernel also leaves b in FP32 there
  const bool cold_mixed = o.precision == FEMH_MIXED && !warm;
  launch_fine_rhs<double>(f0, R, gtab.as<double>(), b.as<double>(), cold_mixed ? r32.as<float>() : nullptr,
                          red, S_BB, stream);
  FEMH_HIP(hipGetLastError());
  std::vector<double> bb(R);
  read_scalars(S_BB, bb.data(), R, vec_parts(g.V));
  // right-hand sides that vanish (the affine field already balances, e.g. a
  // homogeneous cell) are solved by u = 0
  std::vector<int> active(R, 0);
  for (int t = 0; t < R; t++) {
    const double sumE = mst[3 * (t / kRHS)];
    const double zero_thresh = 1e-13 * (sumE / g.V) * std::sqrt((double)g.V);
    active[t] = std::sqrt(bb[t]) > zero_thresh;
    // cubic structure: only the xx and yz load cases are solved; the others
    // are formed from them after the solve
    if (cubic_ok[t / kRHS] && t % kRHS != 0 && t % kRHS != 3) active[t] = 0;
  }
  if (timed) FEMH_HIP(hipEventRecord(ev[1], stream));

  std::vector<int> iters(R, 0), outer(R, 0);
  std::vector<double> relres(R, 0.0), ch_est(nb, 0.0), ritz(nb, 0.0);
  fallback_.assign(nb, 0);
  int status = FEMH_OK;
  if (warm) {
    // vanishing right-hand sides restart from zero
    std::vector<int> zero(R);
    bool any = false;
    for (int t = 0; t < R; t++) any |= (zero[t] = !active[t]) != 0;
    if (any) {
      set_act(A_POW, zero);
      launch_zero_rhs(g.V, R, x.as<double>(), actp(A_POW), stream);
    }
  }
  if (o.precision == FEMH_FP64) {
    std::vector<int> act_pcg = active;
    if (warm) {
      // solve K d = b - K x_prev and return x_prev + d
      FEMH_HIP(hipMemcpyAsync(xw.p, x.p, field_bytes(g, nb, sizeof(double)), hipMemcpyDeviceToDevice,
                              stream));
      apply<double>(h64, 0, xw.as<double>(), b.as<double>(), b.as<double>(), 1, actp(A_ALL), nullptr, 0);
      launch_dot<double>(g.V, R, b.as<double>(), b.as<double>(), S_RR, actp(A_ALL), red, stream);
      std::vector<double> r0(R);
      read_scalars(S_RR, r0.data(), R, vec_parts(g.V));
      for (int t = 0; t < R; t++)
        if (act_pcg[t]) {
          relres[t] = std::sqrt(r0[t] / bb[t]);
          if (relres[t] <= o.tol) act_pcg[t] = 0;
        }
    }
    std::vector<int> act_run = act_pcg;
    const std::vector<double> rel0 = relres;
    pcg<double>(h64, b.as<double>(), x.as<double>(), r.as<double>(), z.as<double>(), p.as<double>(),
                q.as<double>(), act_run, o.tol, o.max_iter, bb.data(), iters.data(), relres.data());
    if (warm) {
      for (int t = 0; t < R; t++)
        if (active[t] && !act_pcg[t]) relres[t] = rel0[t];  // converged before any iteration
      launch_add2<double>(g.V, R, x.as<double>(), xw.as<double>(), x.as<double>(), actp(A_ALL), stream);
    }
    for (int t = 0; t < R; t++) {
      outer[t] = act_pcg[t];
      if (act_run[t]) status = FEMH_NOT_CONVERGED;
    }
  } else {
    // FP64 iterative refinement around FP32 multigrid-PCG corrections.
    // Before each sweep the preconditioned residual r.M^-1 r of every load
    // case estimates its energy-norm error ||e||_K^2, and the energy-form C_H
    // error is the Gram matrix of those errors: dC = E^T K E / V (PSD), so
    // ||dC||_F <= trace(dC) = sum_i ||e_i||_K^2 / V.  With C_ii estimated as
    // (a_ii sum E - b.x - x.r)/V, est = sqrt(6) sum_i rz_i / (V trace C) bounds
    // the relative Frobenius error of a sample when M^-1 ~ K^-1.
    std::vector<double> rr = bb;
    std::vector<double> bx(R, 0.0), xr(R, 0.0), rz(R), rn(R), bxn(R), xrn(R);
    if (warm) {
      const int npd = residual_dots(actp(A_ALL));
      read_scalars(S_RR, rr.data(), R, npd);
      read_scalars(S_BX, bx.data(), R, npd);
      read_scalars(S_XR, xr.data(), R, npd);
      for (int t = 0; t < R; t++)
        if (active[t]) {
          relres[t] = std::sqrt(rr[t] / bb[t]);
          if (relres[t] <= o.tol) active[t] = 0;
        }
    }
    // cold start: x = 0 and r = b are implicit in the first sweep (r32 = b
    // came from the right-hand side kernel; x is written as x32 after it)
    bool x_set = warm;
    double aii[kRHS];
    for (int k = 0; k < kRHS; k++) {
      double acc = 0;
      for (int t = 0; t < 24; t++) acc += tab.u_aff[k * 24 + t] * tab.g[k * 24 + t];
      aii[k] = acc;
    }
    std::vector<int> it_in(R);
    std::vector<double> rel_in(R);
    const int max_outer = o.max_outer > 0 ? o.max_outer : 12;
    // Energy-norm bound: r.K^-1 r <= r.M^-1 r / lambda_min(M^-1 K).  The
    // correction sweeps' PCG coefficients give Lanczos matrices whose smallest
    // Ritz value estimates lambda_min, and so does the Rayleigh quotient of
    // each preconditioned residual at a stop test; the smallest of them over
    // all sweeps and load cases of a cell divides that cell's estimate.  All
    // of them lie at or above lambda_min, so the estimate is a bound once
    // they have come down to it (they converge as the sweeps iterate).  A
    // warm start begins with the value of the slot's previous call.
    const bool bound = o.ch_tol > 0;
    std::vector<double> theta(nb, INFINITY);
    if (warm && (int)theta_prev.size() >= nb)
      for (int s_ = 0; s_ < nb; s_++) theta[s_] = theta_prev[s_];
    if (bound) lanczos.alloc(sizeof(double) * 2 * (size_t)kInnerMaxIter * kMaxR);
    std::vector<double> lzh;
    int nact = 0;
    for (int t = 0; t < R; t++) nact += active[t];
    for (int sweep = 0; sweep < max_outer && nact; sweep++) {
      set_act(A_OUT, active);
      const int *ao = actp(A_OUT);
      // the FP32 sweep solves K d = r: its right-hand side is the residual
      // itself; z = M^-1 r and r.z serve the estimate and the first PCG step
      if (x_set) launch_convert<double, float>((size_t)R * 3 * g.V, r.as<double>(), r32.as<float>(), stream);
      precond<float>(h32, r32.as<float>(), z32.as<float>(), ao, S_RZ0);
      FEMH_HIP(hipGetLastError());
      if (o.ch_tol > 0 && (sweep > 0 || warm)) {
        // z.Kz of the preconditioned residual: its Rayleigh quotient
        // mu = z.Kz / r.z (M z = r) lies in the spectrum of M^-1 K, so it is
        // another upper estimate of lambda_min.  A sweep leaves its residual
        // in the modes it has not resolved, so mu catches a small isolated
        // eigenvalue that the sweep's short Lanczos matrix missed (a stiff
        // island in a soft phase: Ritz 0.26 after the first sweep, 0.0054
        // after the second).  One FP32 fine-operator launch.
        apply<float>(h32, 0, z32.as<float>(), q32.as<float>(), nullptr, 0, ao, &red, S_PQ);
        read_scalars(S_RZ0, rz.data(), R, rz_parts(h32));
        std::vector<double> zkz(R);
        read_scalars(S_PQ, zkz.data(), R, apply_parts((int)sizeof(float)));
        for (int s = 0; s < nb; s++) {
          double trC = 0, srz = 0;
          bool any = false, definite = true;
          for (int k = 0; k < kRHS; k++) {
            const int t = kRHS * s + k;
            // (cubic: trace C = 3 (C_xx,xx + C_yz,yz) from the two solved cases)
            if (cubic_ok[s])
              trC += (k == 0 || k == 3) ? 3.0 * (aii[k] * mst[3 * s] - bx[t] - xr[t]) / g.V : 0.0;
            else
              trC += (aii[k] * mst[3 * s] - bx[t] - xr[t]) / g.V;
            if (active[t]) {
              any = true;
              if (!(rz[t] > 0) || !(zkz[t] > 0)) {
                definite = false;  // the preconditioner is not positive definite here
              } else {
                srz += rz[t];
                theta[s] = std::min(theta[s], zkz[t] / rz[t]);
              }
            }
          }
          if (!any) continue;
          // no bound without a positive definite measurement (or, on a warm
          // start's first sweep, without any): the sweep runs
          ch_est[s] = (definite && trC > 0 && std::isfinite(theta[s]) && theta[s] > 0)
                          ? std::sqrt(6.0) * srz / (g.V * trC) / theta[s]
                          : INFINITY;
          if (ch_est[s] <= o.ch_tol)
            for (int k = 0; k < kRHS; k++)
              if (active[kRHS * s + k]) {
                active[kRHS * s + k] = 0;
                nact--;
              }
        }
        if (!nact) break;
      }
      std::vector<int> act_in = active;
      // Per-cell sweep tolerance: with a finite estimate above ch_tol, the
      // estimate scales with r.M^-1 r, i.e. with the square of the residual,
      // so a residual reduction of rho = 0.5 sqrt(ch_tol / est) leaves it a
      // factor 4 below the target; the sweep stops there (never looser than
      // 0.1, never tighter than inner_tol).  The device test compares
      // |r|^2 / bnorm2 with inner^2, so scaling bnorm2 by (rho / inner)^2
      // gives that right-hand side its own tolerance.
      const double inner = o.inner_tol > 0 ? o.inner_tol : 1e-4;
      std::vector<double> rr_in = rr;
      if (o.ch_tol > 0 && o.adapt_inner)
        for (int t = 0; t < R; t++) {
          const double est = ch_est[t / kRHS];
          if (!active[t] || !(est > o.ch_tol) || !std::isfinite(est)) continue;
          const double rho = std::min(0.1, std::max(inner, 0.5 * std::sqrt(o.ch_tol / est)));
          rr_in[t] = rr[t] * (rho / inner) * (rho / inner);
        }
      {
        // the record covers this FP32 sweep only (at most kInnerMaxIter
        // iterations), never the FP64 fallback
        struct Record {
          Context &c;
          Record(Context &c_, bool on) : c(c_) {
            if (on) {
              c.lz_alpha = c.lanczos.as<double>();
              c.lz_beta = c.lz_alpha + (size_t)kInnerMaxIter * kMaxR;
            }
          }
          ~Record() { c.lz_alpha = c.lz_beta = nullptr; }
        } rec(*this, bound);
        pcg<float>(h32, r32.as<float>(), x32.as<float>(), r32.as<float>(), z32.as<float>(),
                   p32.as<float>(), q32.as<float>(), act_in, inner, std::min(o.max_iter, kInnerMaxIter),
                   rr_in.data(), it_in.data(), rel_in.data(), true);
      }
      if (bound) {
        int kmax = 0;
        for (int t = 0; t < R; t++)
          if (active[t]) kmax = std::max(kmax, it_in[t]);
        if (kmax > 0) {
          // rows 0..kmax-1 of alpha and beta (beta rows stop one short)
          lzh.resize((size_t)2 * kInnerMaxIter * kMaxR);
          const double *la = lanczos.as<double>(), *lb = la + (size_t)kInnerMaxIter * kMaxR;
          FEMH_HIP(hipMemcpyAsync(lzh.data(), la, sizeof(double) * (size_t)kmax * kMaxR,
                                  hipMemcpyDeviceToHost, stream));
          FEMH_HIP(hipMemcpyAsync(lzh.data() + (size_t)kInnerMaxIter * kMaxR, lb,
                                  sizeof(double) * (size_t)kmax * kMaxR, hipMemcpyDeviceToHost, stream));
          FEMH_HIP(hipStreamSynchronize(stream));
          for (int t = 0; t < R; t++) {
            if (!active[t] || it_in[t] < 1) continue;
            const double th = lanczos_min(lzh.data() + t, lzh.data() + (size_t)kInnerMaxIter * kMaxR + t,
                                          kMaxR, it_in[t]);
            if (std::isfinite(th) && th > 0) theta[t / kRHS] = std::min(theta[t / kRHS], th);
          }
        }
      }
      set_act(A_OUT, active);
      if (x_set) {
        launch_add_f32_to_f64(g.V, R, x32.as<float>(), x.as<double>(), ao, stream);
      } else {
        // the sweep zeroed x32 for every RHS first: inactive ones stay 0
        launch_convert<float, double>((size_t)R * 3 * g.V, x32.as<float>(), x.as<double>(), stream);
        x_set = true;
      }
      const int npd = residual_dots(ao);
      FEMH_HIP(hipGetLastError());
      read_scalars(S_RR, rn.data(), R, npd);
      read_scalars(S_BX, bxn.data(), R, npd);
      read_scalars(S_XR, xrn.data(), R, npd);
      // A sweep that did not reach its tolerance, or that failed to halve the
      // FP64 residual, means the FP32 hierarchy does not precondition this
      // cell (FP32 rounding of operators with a modulus contrast near its
      // precision can make the coarse levels indefinite); the a-posteriori
      // estimate, built on the same preconditioner, is then not trusted
      // either.  Such right-hand sides are finished by FP64 MG-PCG from the
      // current iterate.
      std::vector<int> fail(R, 0);
      bool any_fail = false;
      for (int t = 0; t < R; t++) {
        if (!active[t]) continue;
        if (!std::isfinite(rn[t])) throw Error(FEMH_NONFINITE, "residual became non-finite");
        iters[t] += it_in[t];
        outer[t]++;
        if (act_in[t] || rn[t] > 0.25 * rr[t]) {
          fail[t] = 1;
          any_fail = true;
        }
        rr[t] = rn[t];
        bx[t] = bxn[t];
        xr[t] = xrn[t];
        relres[t] = std::sqrt(rn[t] / bb[t]);
        if (relres[t] <= o.tol) {
          active[t] = 0;
          fail[t] = 0;
          nact--;
        }
      }
      if (any_fail) {
        if (!dx.p) dx.alloc(field_bytes(g, B, sizeof(double)));
        std::vector<int> it64(R, 0), act64 = fail;
        std::vector<double> rel64(R, 0.0);
        // K d = r (relative to the original right-hand side), x += d
        if (!dq.p) dq.alloc(field_bytes(g, B, sizeof(double)));
        ensure_h64_ops();
        pcg<double>(h64, r.as<double>(), dx.as<double>(), z.as<double>(), p.as<double>(),
                    q.as<double>(), dq.as<double>(), act64, o.tol, o.max_iter, bb.data(), it64.data(),
                    rel64.data());
        set_act(A_POW, fail);
        launch_add2<double>(g.V, R, x.as<double>(), dx.as<double>(), x.as<double>(), actp(A_POW), stream);
        for (int t = 0; t < R; t++) {
          if (!fail[t]) continue;
          iters[t] += it64[t];
          outer[t]++;
          fallback_[t / kRHS] = 1;
          relres[t] = rel64[t];
          active[t] = 0;
          nact--;
          if (act64[t]) status = FEMH_NOT_CONVERGED;
        }
      }
    }
    if (nact) status = FEMH_NOT_CONVERGED;
    theta_prev = theta;
    for (int s_ = 0; s_ < nb; s_++) ritz[s_] = std::isfinite(theta[s_]) ? theta[s_] : 0.0;
    // every right-hand side vanished before a sweep: u = 0
    if (!x_set) launch_fill<double>(field_bytes(g, nb, sizeof(double)) / sizeof(double), x.as<double>(), 0.0, stream);
  }
  // C_H and the adjoint do not see constant shifts of the fields (K0 maps
  // them to zero); the mean is taken out when the fields are read out
  fields_centered = false;
  if (timed) FEMH_HIP(hipEventRecord(ev[2], stream));
  unsigned cmask = 0;
  for (int smp = 0; smp < nb; smp++) cmask |= cubic_ok[smp] ? 1u << smp : 0u;
  if (cmask) {
    // the cell is invariant under axis swaps, so the unsolved load cases are
    // the solved ones with axes swapped (the reference's swapDispXY,
    // NumericalCoarseningUsingMultiGrid.cpp:44-70): yy and xz from xx and yz
    // by x <-> y, zz and xy by x <-> z.  C_H then comes from the energy form
    // of all six fields (second order in the solve error).
    const int n = g.nx;
    launch_permute_case(n, nb, x.as<double>(), 0, 1, 0, cmask, stream);
    launch_permute_case(n, nb, x.as<double>(), 0, 2, 1, cmask, stream);
    launch_permute_case(n, nb, x.as<double>(), 3, 4, 0, cmask, stream);
    launch_permute_case(n, nb, x.as<double>(), 3, 5, 1, cmask, stream);
    FEMH_HIP(hipGetLastError());
  }
  compute_ch(x.as<double>(), C);
  // the exact C_H of such a cell is invariant under the axis permutations,
  // which permute Voigt indices: average it over them, which removes the
  // part of the solve error that breaks that symmetry
  static const int vperm[6][6] = {{0, 1, 2, 3, 4, 5}, {1, 0, 2, 4, 3, 5}, {2, 1, 0, 5, 4, 3},
                                  {0, 2, 1, 3, 5, 4}, {1, 2, 0, 4, 5, 3}, {2, 0, 1, 5, 3, 4}};
  for (int smp = 0; smp < nb; smp++) {
    if (!cubic_ok[smp]) continue;
    double *Cs = C + 36 * smp, avg[36] = {0};
    for (const auto &pm : vperm)
      for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) avg[6 * pm[i] + pm[j]] += Cs[6 * i + j] / 6.0;
    for (int q = 0; q < 36; q++) Cs[q] = avg[q];
  }
  C_last.assign(C, C + 36 * (size_t)nb);
  flush_profile();
  have_fields = true;
  float ms1 = 0, ms2 = 0;
  if (timed) {
    FEMH_HIP(hipEventElapsedTime(&ms1, ev[0], ev[1]));
    FEMH_HIP(hipEventElapsedTime(&ms2, ev[1], ev[2]));
  }
  const double ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (st) {
    for (int s = 0; s < nb; s++) {
      femh_stats &o_ = st[s];
      std::memset(&o_, 0, sizeof(o_));
      for (int k = 0; k < kRHS; k++) {
        o_.iters[k] = iters[kRHS * s + k];
        o_.outer[k] = outer[kRHS * s + k];
        o_.relres[k] = relres[kRHS * s + k];
      }
      o_.ch_err_est = fallback_[s] ? 0.0 : ch_est[s];
      o_.fp64_fallback = fallback_[s];
      o_.ritz_min = ritz[s];
      o_.ms_setup = ms1;
      o_.ms_solve = ms2;
      o_.ms_total = ms_total;
      o_.levels = nlevels;
    }
  }
  for (int k = 0; k < 36 * nb; k++)
    if (!std::isfinite(C[k])) throw Error(FEMH_NONFINITE, "C_H is not finite");
  prev_nb = nb;
  if (status != FEMH_OK) throw Error(status, "PCG did not reach the tolerance within max_iter");
}

// dJ/drho (or, energy_only, the 21 energy terms per voxel) of the last call's
// cells into device memory, queued on the context's stream
static void gradient_launch(Context &c, const double *dJdC, double *d_out, bool energy_only) {
  if (!c.have_fields) throw Error(FEMH_NO_FIELDS, "no solved fields: call femh_homogenize first");
  std::vector<double> Gsym(21 * c.nb);
  for (int s = 0; s < c.nb; s++) {
    const double *G = dJdC + 36 * s;
    int p = 0;
    for (int a = 0; a < 6; a++)
      for (int bb = a; bb < 6; bb++)
        Gsym[21 * s + p++] = energy_only ? 0.0 : (a == bb ? G[6 * a + a] : G[6 * a + bb] + G[6 * bb + a]);
  }
  // (pageable source: the copy is complete when the call returns)
  FEMH_HIP(hipMemcpyAsync(c.Gs.p, Gsym.data(), sizeof(double) * Gsym.size(), hipMemcpyHostToDevice, c.stream));
  launch_gradient(c.g, c.nb, c.rho.as<double>(), c.modecoef, c.uaff.as<double>(), c.x.as<double>(),
                  c.Gs.as<double>(), c.modulus_input ? FEMH_INTERP_MODULUS : c.o.interp, c.o.E0, c.o.Emin,
                  c.o.penal, energy_only ? 1 : 0, d_out, c.stream);
  FEMH_HIP(hipGetLastError());
}

void Context::gradient(const double *dJdC, double *out_host, bool energy_only) {
  const size_t n = (energy_only ? (size_t)21 : 1) * g.V * nb;
  DBuf out;
  out.alloc(sizeof(double) * n);
  gradient_launch(*this, dJdC, out.as<double>(), energy_only);
  FEMH_HIP(hipMemcpyAsync(out_host, out.p, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
}

void Context::gradient_device(const double *dJdC, const double *shift, double *d_out, double *gmax) {
  gradient_launch(*this, dJdC, d_out, false);
  if (!shift && !gmax) {
    FEMH_HIP(hipStreamSynchronize(stream));
    return;
  }
  const int nblk = kRedBlocks;
  ls_part.alloc(sizeof(double) * (size_t)nb * nblk);
  if (shift) {
    ls_h.alloc(sizeof(double) * (size_t)nb);
    FEMH_HIP(hipMemcpyAsync(ls_h.p, shift, sizeof(double) * nb, hipMemcpyHostToDevice, stream));
  }
  launch_shift_absmax(nb, g.V, d_out, shift ? ls_h.as<double>() : nullptr, ls_part.as<double>(), nblk, stream);
  FEMH_HIP(hipGetLastError());
  std::vector<double> part((size_t)nb * nblk);
  FEMH_HIP(hipMemcpyAsync(part.data(), ls_part.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
  if (gmax)
    for (int s = 0; s < nb; s++) {
      double m = 0;
      for (int b = 0; b < nblk; b++) {
        const double v = part[(size_t)s * nblk + b];
        m = (m != m) ? m : (v != v ? v : std::max(m, v));
      }
      gmax[s] = m;
    }
}

void Context::descent_trial(int n, const double *xd, const double *gd, const double *h, double lb, double ub,
                            double *out, double *sums) {
  const int nblk = kRedBlocks;
  ls_part.alloc(sizeof(double) * (size_t)n * nblk);
  ls_h.alloc(sizeof(double) * (size_t)n);
  FEMH_HIP(hipMemcpyAsync(ls_h.p, h, sizeof(double) * n, hipMemcpyHostToDevice, stream));
  launch_descent_trial(n, g.V, xd, gd, ls_h.as<double>(), lb, ub, out, ls_part.as<double>(), nblk, stream);
  FEMH_HIP(hipGetLastError());
  std::vector<double> part((size_t)n * nblk);
  FEMH_HIP(hipMemcpyAsync(part.data(), ls_part.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
  if (sums)
    for (int s = 0; s < n; s++) {
      double t = 0;
      for (int b = 0; b < nblk; b++) t += part[(size_t)s * nblk + b];
      sums[s] = t;
    }
}

void Context::thermal(const double *alpha, double *beta, double *alpha_eff) {
  if (!have_fields) throw Error(FEMH_NO_FIELDS, "no solved fields: call femh_homogenize first");
  alpha_d.alloc(sizeof(double) * (size_t)nb * g.V);
  FEMH_HIP(hipMemcpyAsync(alpha_d.p, alpha, sizeof(double) * (size_t)nb * g.V, hipMemcpyHostToDevice, stream));
  const int nblk = 1024;  // ch_part holds 21 * 1024 partials per sample
  launch_thermal(g, nb, E64.as<double>(), alpha_d.as<double>(), uaff.as<double>(), x.as<double>(),
                 ch_part.as<double>(), nblk, ch_out.as<double>(), stream);
  FEMH_HIP(hipGetLastError());
  std::vector<double> sums(kRHS * (size_t)nb);
  FEMH_HIP(hipMemcpyAsync(sums.data(), ch_out.p, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
  // I : C0 : eps = 3 K tr(eps) with 3 K = E / (1 - 2 nu) for the unit modulus
  const double scale = 1.0 / ((1.0 - 2.0 * o.nu) * (double)g.V);
  for (int smp = 0; smp < nb; smp++) {
    double bq[kRHS];
    for (int q = 0; q < kRHS; q++) {
      bq[q] = sums[kRHS * smp + q] * scale;
      if (!std::isfinite(bq[q])) throw Error(FEMH_NONFINITE, "non-finite thermal stress");
    }
    if (beta)
      for (int q = 0; q < kRHS; q++) beta[kRHS * smp + q] = bq[q];
    if (!alpha_eff) continue;
    // C_H is symmetric positive definite: Cholesky, then two triangular solves
    double L[6][6] = {{0}};
    const double *C = C_last.data() + 36 * smp;
    for (int i = 0; i < 6; i++)
      for (int j = 0; j <= i; j++) {
        double v = C[6 * i + j];
        for (int k = 0; k < j; k++) v -= L[i][k] * L[j][k];
        if (i == j) {
          if (!(v > 0)) throw Error(FEMH_NONFINITE, "C_H is not positive definite");
          L[i][i] = std::sqrt(v);
        } else {
          L[i][j] = v / L[j][j];
        }
      }
    double y[6];
    for (int i = 0; i < 6; i++) {
      double v = bq[i];
      for (int k = 0; k < i; k++) v -= L[i][k] * y[k];
      y[i] = v / L[i][i];
    }
    for (int i = 5; i >= 0; i--) {
      double v = y[i];
      for (int k = i + 1; k < 6; k++) v -= L[k][i] * alpha_eff[kRHS * smp + k];
      alpha_eff[kRHS * smp + i] = v / L[i][i];
    }
  }
}

// The reference's force-driven path (SamplesGenerator/NumericalCoarsening.cpp:
// 1148-1306 with femlib/src/MeshUtilities.cpp:960-1060) applies six face
// tractions under periodic constraints, then reduces C from the tensor shear
// strains at every Gauss point and at the cell-sized hex,
//   C = (1/nelem) G^-T (sum_e sum_q w_q Gv^T C_e Gv) G^-1,
// with C_e the engineering-shear stiffness.  On a cube cell the loads carry no
// net moment, so (without the reference's 1e-5 diagonal regularisation) each
// constrained solution is a macro strain plus the strain-driven fluctuation
// of that strain, G cancels, and C = T^-1 M T^-1 with
// M = (1/V) sum_e E_e chi_a^T KT chi_b over the strain-driven total fields
// chi and KT the element matrix of T C0 T, T = diag(1,1,1,1/2,1/2,1/2).
void Context::coarse_strain(double *C) {
  if (!have_fields) throw Error(FEMH_NO_FIELDS, "no solved fields: call femh_homogenize first");
  if (!(g.nx == g.ny && g.ny == g.nz))
    throw Error(FEMH_BAD_ARG, "the coarse-strain tensor needs a cube cell (on other cells the reference's "
                              "shear loads carry a net moment)");
  ModeCoef mt;
  std::memcpy(mt.diag, tab.LT.diag, sizeof(mt.diag));
  std::memcpy(mt.cross, tab.LT.cross, sizeof(mt.cross));
  const int nblk = 1024;
  launch_ch_energy(g, nb, E64.as<double>(), mt, uaff.as<double>(), x.as<double>(), ch_part.as<double>(), nblk,
                   ch_out.as<double>(), stream);
  FEMH_HIP(hipGetLastError());
  std::vector<double> s(21 * (size_t)nb);
  FEMH_HIP(hipMemcpyAsync(s.data(), ch_out.p, sizeof(double) * s.size(), hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
  static const double Tv[6] = {1, 1, 1, 0.5, 0.5, 0.5};
  for (int smp = 0; smp < nb; smp++) {
    int p = 0;
    double *Cs = C + 36 * smp;
    for (int a = 0; a < 6; a++)
      for (int b2 = a; b2 < 6; b2++) {
        Cs[6 * a + b2] = Cs[6 * b2 + a] = s[21 * smp + p] / ((double)g.V * Tv[a] * Tv[b2]);
        p++;
      }
    for (int q = 0; q < 36; q++)
      if (!std::isfinite(Cs[q])) throw Error(FEMH_NONFINITE, "coarse-strain tensor is not finite");
  }
}

void Context::get_fields(double *u) {
  if (!have_fields) throw Error(FEMH_NO_FIELDS, "no solved fields: call femh_homogenize first");
  if (!fields_centered) {
    std::vector<int> all(R, 1);
    set_act(A_POW, all);
    launch_remove_mean<double>(g.V, R, x.as<double>(), actp(A_POW), red, stream);
    fields_centered = true;
  }
  FEMH_HIP(hipMemcpyAsync(u, x.p, field_bytes(g, nb, sizeof(double)), hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
}

void Context::apply_level(int level, const double *u, double *y) {
  if (!h64.ready) throw Error(FEMH_NO_FIELDS, "no hierarchy: call femh_prepare first");
  ensure_h64_ops();
  if (level < 0 || level >= nlevels) throw Error(FEMH_BAD_ARG, "level out of range");
  Level<double> &L = h64.lv[level];
  const size_t n = (size_t)3 * L.g.V;
  L.x.alloc(field_bytes(L.g, B, sizeof(double)));
  L.Ad.alloc(field_bytes(L.g, B, sizeof(double)));
  std::vector<int> one(R, 0);
  one[0] = 1;
  set_act(A_POW, one);
  FEMH_HIP(hipMemcpyAsync(L.x.p, u, sizeof(double) * n, hipMemcpyHostToDevice, stream));
  apply<double>(h64, level, L.x.as<double>(), L.Ad.as<double>(), nullptr, 0, actp(A_POW), nullptr, 0);
  FEMH_HIP(hipGetLastError());
  FEMH_HIP(hipMemcpyAsync(y, L.Ad.p, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  FEMH_HIP(hipStreamSynchronize(stream));
}

template void Context::pcg<double>(Hierarchy<double> &, const double *, double *, double *, double *,
                                   double *, double *, std::vector<int> &, double, int,
                                   const double *, int *, double *, bool);
template void Context::pcg<float>(Hierarchy<float> &, const float *, float *, float *, float *,
                                  float *, float *, std::vector<int> &, double, int, const double *,
                                  int *, double *, bool);

}  // namespace femh
