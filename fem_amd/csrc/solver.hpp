// solver.hpp -- device-resident homogenization context behind the C ABI.
//
// A context owns device storage for up to B samples of one grid size and
// solves the 6 load cases of every sample together (R = 6 nb right-hand
// sides for a call with nb <= B samples).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/femh.h"
#include "kernels.hpp"
#include "ops.hpp"
#include "tables.hpp"

namespace femh {

struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};

void hip_check(hipError_t e, const char *what);
#define FEMH_HIP(x) ::femh::hip_check((x), #x)

// Device memory arena: a caller-provided device region from which DBuf::alloc
// carves buffers (bump allocation, 256-byte aligned) while an ArenaScope is
// active on the calling thread.  Such buffers are never freed: the region
// belongs to the caller.  A context built inside a scope also works on the
// arena's stream and creates no stream or event of its own.
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  hipStream_t stream = nullptr;
  void *take(size_t n);
};
struct ArenaScope {
  explicit ArenaScope(Arena *a);
  ~ArenaScope();
  ArenaScope(const ArenaScope &) = delete;
  ArenaScope &operator=(const ArenaScope &) = delete;
  Arena *prev;
};
Arena *current_arena();

// device allocation (owning, unless carved from an arena)
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool borrowed = false;  // from an arena: not freed here
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), bytes(o.bytes), borrowed(o.borrowed) {
    o.p = nullptr;
    o.bytes = 0;
    o.borrowed = false;
  }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      borrowed = o.borrowed;
      o.p = nullptr;
      o.bytes = 0;
      o.borrowed = false;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t n);
  void release();
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

template <typename T>
struct Level {
  Dims g{};
  DBuf S;                    // stencil [B][243][V] (levels >= 1)
  DBuf D, Dinv;              // point diagonal and its inverse [B][3][V]
  std::vector<double> lmax;  // per sample: lambda_max(D^-1 A) estimate (with margin)
  DBuf x, b, r, d, d2, xn, Ad;  // V-cycle work, [R][3][V]
};

template <typename T>
struct Hierarchy {
  std::vector<Level<T>> lv;
  DBuf E;     // fine moduli [B][V] in T
  DBuf Ke;    // unit stiffness in T
  DBuf Ainv;  // per sample: dense (regularised) inverse of the coarsest operator
  DBuf gt1, gt2;  // Galerkin intermediates (one- and two-axis coarsened stencils)
  // Chebyshev coefficients of every level, step (0..cheb_steps-1) and
  // interval (0: smoothing ratio, 1: the coarsest solve's) for every sample:
  // [level][step][interval][3][kMaxB] doubles (1/theta, c1, c2)
  DBuf cheb;
  int cheb_steps = 0;
  int ndense = 0;
  bool ready = false;
};

// Chebyshev interval [lmax / ratio, lmax] of the iterative coarsest solve
constexpr double kCoarseChebRatio = 1000.0;
// iteration cap of one FP32 correction sweep in mixed precision
constexpr int kInnerMaxIter = 1000;

// activity-array slots (device int[kMaxR] each)
enum ActSlot : int { A_ALL = 0, A_PCG = 1, A_OUT = 2, A_POW = 3, kActSlots = 4 };

struct Context {
  femh_opts o{};
  Dims g{};
  int B = 1;   // capacity (samples)
  int nb = 0;  // samples of the current call
  int R = 0;   // 6 nb
  ElementTables tab;
  ModeCoef modecoef{};
  hipStream_t stream = nullptr;
  bool own_stream = false;

  // material / tables on device
  DBuf rho, E64, Ke64, gtab, uaff, Mchild, Gs;
  // FP64 solver state, each [R][3][V]
  DBuf b, x, r, z, p, q;
  // FP32 inner solver state (mixed)
  DBuf r32, x32, z32, p32, q32;
  DBuf red_part, red_ctr, red_scal;
  Reduce red{};
  DBuf ch_part, ch_out;
  DBuf alpha_d;                  // per-voxel expansion coefficients (thermal)
  std::vector<double> C_last;    // C_H of the last homogenize call, 36 per sample
  DBuf act;
  Hierarchy<double> h64;
  Hierarchy<float> h32;
  bool have_fields = false;
  bool fields_centered = false;  // x has zero mean per RHS and component
  int prev_nb = 0;  // samples of the last successful call (warm start source)
  DBuf xw;          // warm start: previous fields [R][3][V] (FP64)
  DBuf dx, dq;      // mixed precision: FP64 fallback correction and PCG scratch
  DBuf palt, palt32;  // PCG: second direction buffer of the fused direction update
  DBuf beta_d;        // PCG: per-RHS beta [kMaxR] (double)
  std::vector<int> fallback_;  // per sample: finished in FP64 this call
  // Lanczos record of the mixed path's correction sweeps: alpha and beta of
  // every PCG iteration and right-hand side ([kInnerMaxIter][kMaxR] each);
  // pcg writes them while lz_alpha / lz_beta are set
  DBuf lanczos;
  double *lz_alpha = nullptr, *lz_beta = nullptr;
  std::vector<double> theta_prev;  // per slot: smallest Ritz value of the last call
  // PCG iteration state on the device: iterations, state (int [kMaxR] each),
  // relative residual and |b|^2 (double [kMaxR] each)
  DBuf pcg_state;
  DBuf ls_part, ls_h;  // device line search: per-cell block partials [n][kRedBlocks], steps [n]
  // recorded PCG iteration pairs (hipGraph), keyed by everything their
  // launches bake in: precision, right-hand sides, vector and hierarchy
  // buffers, launch shapes, tolerance
  struct GraphKey {
    std::array<long long, 15> v{};
    long long tol = 0;
    bool operator==(const GraphKey &o) const { return v == o.v && tol == o.tol; }
  };
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs_;
  hipGraphExec_t find_graph(const GraphKey &k) const {
    for (const auto &g : graphs_)
      if (g.first == k) return g.second;
    return nullptr;
  }
  long long xfer_hash() const {
    long long hsh = 1469598103934665603ll;
    for (int v : shape_xfer) hsh = (hsh ^ v) * 1099511628211ll;
    return hsh;
  }
  template <typename T>
  long long hier_hash(const Hierarchy<T> &h) const;
  bool modulus_input = false;  // last load gave moduli (dE/drho = 1)
  int nlevels = 1;
  // launch shapes (0: the built-in choice): node planes per workgroup of the
  // fine operator, coarse planes per thread of the transfers between level l
  // and l + 1; set by the autotuner
  int shape_fine = 0;
  std::vector<int> shape_xfer;
  int xfer_shape(int l) const { return l < (int)shape_xfer.size() ? shape_xfer[l] : 0; }
  template <typename T>
  void autotune(Hierarchy<T> &h);

  // kernel timing
  struct KStat {
    long long launches = 0;
    double ms = 0, bytes = 0, flops = 0;
  };
  struct Pending {
    hipEvent_t a, b;
    int which;
    double bytes, flops;
  };
  bool prof = false;
  KStat kst[2];
  std::vector<Pending> pending;
  std::vector<hipEvent_t> ev_pool;
  hipEvent_t ev_call[3] = {};  // homogenize: start, end of set-up, end of solve
  hipEvent_t ev_tune[2] = {};  // autotuner timing
  void set_profiling(bool on);
  void flush_profile();
  hipEvent_t take_event();

  explicit Context(const femh_opts &opts);
  ~Context();

  // inputs: n samples (<= B) of density (interp applied) or modulus
  void load(const double *src, int n, bool device_src, bool modulus);
  void build_hierarchy();
  // FP64 level operators of a mixed-precision context, built on first use
  // (the FP64 fallback, apply_level); mixed solves smooth with FP32 ones
  void ensure_h64_ops();
  bool h64_ops = false;
  void homogenize(double *C, femh_stats *st);  // C[36 nb], st[nb] (may be null)
  void gradient(const double *dJdC, double *out_host, bool energy_only);
  // dJ/drho of the last call's nb cells into device memory d_out[nb V], plus
  // shift[s] (host, may be null) per cell; gmax[s] (host, may be null)
  // receives max |d_out[s]|
  void gradient_device(const double *dJdC, const double *shift, double *d_out, double *gmax);
  // projected-gradient trial point of n cells of this grid (device x, g,
  // out [n V]; host h[n]): h > 0: out = clamp(x - h g, lb, ub), h < 0:
  // out = x, h == 0: out unchanged; sums[s] (host, may be null) = sum of out[s]
  void descent_trial(int n, const double *x, const double *g, const double *h, double lb, double ub,
                     double *out, double *sums);
  void get_fields(double *u);
  // effective thermal stress beta[6 nb] and expansion C_H^-1 beta (either
  // may be null) for per-voxel expansion coefficients alpha[nb V] (host)
  void thermal(const double *alpha, double *beta, double *alpha_eff);
  // coarse-strain (force-driven) form of C_H of the cells of the last call,
  // C[36 nb]: C = T^-1 [(1/V) sum_e E_e chi^T KT chi] T^-1 with KT the
  // element matrix of T C0 T (cube cells only)
  void coarse_strain(double *C);
  void apply_level(int level, const double *u, double *y);

 private:
  const int *actp(int slot) const { return act.as<int>() + slot * kMaxR; }
  void set_act(int slot, const std::vector<int> &v);
  template <typename T>
  void build_levels(Hierarchy<T> &h);
  template <typename T>
  void build_ops(Hierarchy<T> &h);
  template <typename T>
  void op(Hierarchy<T> &h, int l, int st, int ep, OpArgs<T> a);
  template <typename T>
  void apply(Hierarchy<T> &h, int l, const T *u, T *y, const T *b, int mode, const int *act_,
             const Reduce *red, int slot);
  template <typename T>
  void pre_smooth(Hierarchy<T> &h, int l, const T *b, T *x, const int *act_);
  template <typename T>
  bool post_smooth(Hierarchy<T> &h, int l, const T *b, T *x, const int *act_, int dot_slot);
  template <typename T>
  bool vcycle(Hierarchy<T> &h, int l, const T *b, T *x, const int *act_, int dot_slot);
  template <typename T>
  void precond(Hierarchy<T> &h, const T *r, T *z, const int *act_, int rz_slot);
  template <typename T>
  bool fuses_rz(const Hierarchy<T> &h) const;
  template <typename T>
  int rz_parts(const Hierarchy<T> &h) const;
  template <typename T>
  void pcg(Hierarchy<T> &h, const T *bvec, T *xv, T *rv, T *zv, T *pv, T *qv,
           std::vector<int> &active, double tol, int max_iter, const double *bnorm2, int *iters,
           double *relres, bool z_ready = false);
  template <typename T>
  void set_cheb(const Hierarchy<T> &h, int l, OpArgs<T> &a, int step, double ratio);
  template <typename T>
  void upload_cheb(Hierarchy<T> &h);
  template <typename T>
  void cheb_run(Hierarchy<T> &h, int l, const T *b, T *x, const int *act_, int k, double ratio);
  int apply_parts(int sz) const;  // partials of a fine-operator dot, element size sz
  // Chebyshev degrees of pre- and post-smoothing on level l
  int nu_pre(int l) const { return (l > 0 && o.coarse_nu > 0) ? o.coarse_nu : o.nu1; }
  int nu_post(int l) const { return (l > 0 && o.coarse_nu > 0) ? o.coarse_nu : o.nu2; }
  int residual_dots(const int *act_);
  template <typename T>
  void estimate_lmax(Hierarchy<T> &h);
  void dense_coarse(Hierarchy<double> &h);
  void read_scalars(int slot, double *out, int n, int np);
  void compute_ch(const double *fields, double *C);
};

}  // namespace femh
