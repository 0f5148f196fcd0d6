// capi.cpp -- extern "C" entry points declared in include/femh.h.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/femh.h"
#include "solver.hpp"

struct femh_ctx {
  std::unique_ptr<femh::Context> c;
  femh_opts opts{};
  std::string err;
  int last_status = FEMH_OK;
  // per-grid sub-contexts used by the uneven batch entry point
  // (several per size: batches of one size can be in flight at once)
  std::map<std::tuple<int, int, int>, std::vector<std::unique_ptr<femh::Context>>> sub;
  // peer contexts (opts.streams - 1 of them) for concurrent batches of a
  // many-call
  std::vector<std::unique_ptr<femh::Context>> peers;
};

namespace {

template <typename F>
int guarded(femh_ctx *ctx, F &&f) {
  if (!ctx) return FEMH_BAD_ARG;
  try {
    f();
    return FEMH_OK;
  } catch (const femh::Error &e) {
    ctx->err = e.what();
    ctx->last_status = e.status;
    return e.status;
  } catch (const std::bad_alloc &) {
    ctx->err = "host allocation failed";
    ctx->last_status = FEMH_OOM;
    return FEMH_OOM;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    ctx->last_status = FEMH_DEVICE_ERROR;
    return FEMH_DEVICE_ERROR;
  }
}

int need_ctx(femh_ctx *ctx) {
  if (!ctx) return FEMH_BAD_ARG;
  if (!ctx->c) {
    ctx->last_status = FEMH_BAD_ARG;
    if (ctx->err.empty()) ctx->err = "context was not created successfully";
    return FEMH_BAD_ARG;
  }
  return FEMH_OK;
}

// n samples through a context in chunks of its batch capacity
// n cells in batches of B.  With opts.streams > 1 (and profiling off) the
// batches are spread over that many contexts, each on its own stream and
// host thread, so that one batch's host syncs and small coarse-level
// launches overlap another's work.  Batch b goes to context
// (nbatch - 1 - b) mod streams: the last batch runs on the primary context,
// whose fields the gradient, field and thermal calls then read.
void run_many(femh_ctx *ctx, int n, const double *src, bool device_src, bool modulus, double *C,
              femh_stats *stats) {
  femh::Context &c0 = *ctx->c;
  const size_t V = (size_t)c0.g.V;
  const int B = c0.B, nbatch = (n + B - 1) / B;
  const int S = c0.prof ? 1 : std::min(std::max(ctx->opts.streams, 1), nbatch);
  while ((int)ctx->peers.size() < S - 1) ctx->peers.emplace_back(new femh::Context(ctx->opts));
  // A batch that fails (a cell that does not converge, a non-finite
  // modulus) does not stop the other batches: each batch's error is kept
  // and the first one in batch order is raised after all batches ran, so the
  // cells that are solved -- and which they are -- do not depend on
  // `streams`.  (A batch that raises NOT_CONVERGED has still written its C
  // and stats.)
  std::vector<std::exception_ptr> errs(nbatch);
  auto work = [&](int w) {
    femh::Context *cp = nullptr;
    try {
      cp = w == 0 ? &c0 : ctx->peers[w - 1].get();
      if (w > 0) FEMH_HIP(hipSetDevice(cp->o.device));
    } catch (...) {
      for (int b = 0; b < nbatch; b++)
        if ((nbatch - 1 - b) % S == w) errs[b] = std::current_exception();
      return;
    }
    for (int b = 0; b < nbatch; b++) {
      if ((nbatch - 1 - b) % S != w) continue;
      try {
        const int s0 = b * B, k = std::min(B, n - s0);
        cp->load(src + (size_t)s0 * V, k, device_src, modulus);
        cp->homogenize(C + 36 * (size_t)s0, stats ? stats + s0 : nullptr);
      } catch (...) {
        errs[b] = std::current_exception();
      }
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < S; w++) pool.emplace_back(work, w);
  work(0);
  for (auto &t : pool) t.join();
  for (auto &e : errs)
    if (e) std::rethrow_exception(e);
}

}  // namespace

extern "C" {

int femh_abi_version(void) { return FEMH_ABI_VERSION; }

int femh_element_stiffness(double E, double nu, double *Ke) {
  if (!Ke || !(nu > -1.0 && nu < 0.5) || !(E >= 0)) return FEMH_BAD_ARG;
  femh::ElementTables t = femh::build_tables(nu);
  for (int q = 0; q < 576; q++) Ke[q] = E * t.Ke[q];
  return FEMH_OK;
}

void femh_default_opts(femh_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->nx = o->ny = o->nz = 16;
  o->E0 = 1.0;
  o->Emin = 0.0;
  o->nu = 0.3;
  o->penal = 3.0;
  o->interp = FEMH_INTERP_SIMP;
  o->precond = FEMH_PRECOND_MG;
  o->levels = 0;
  o->nu1 = 2;
  o->nu2 = 2;
  o->cheby_ratio = 20.0;
  o->coarse_iters = 32;
  o->tol = 1e-10;
  o->max_iter = 5000;
  o->precision = FEMH_FP64;
  o->ch_tol = 0.0;
  o->check_every = 1;
  o->device = 0;
  o->inner_tol = 1e-4;
  o->max_outer = 12;
  o->apply_kernel = 0;
  o->batch = 1;
  o->warm_start = 0;
  o->structure = FEMH_STRUCT_GENERAL;
  o->autotune = 0;
  o->streams = 1;
  o->coarse_nu = 3;
  o->batch_threads = 0;
  o->graph = 1;
  o->adapt_inner = 1;
}

int femh_create(const femh_opts *o, femh_ctx **out) {
  if (!o || !out) return FEMH_BAD_ARG;
  *out = nullptr;
  femh_ctx *ctx = new (std::nothrow) femh_ctx();
  if (!ctx) return FEMH_OOM;
  ctx->opts = *o;
  const int st = guarded(ctx, [&] { ctx->c.reset(new femh::Context(*o)); });
  // on failure the context is still returned so the caller can read the message
  *out = ctx;
  return st;
}

void femh_destroy(femh_ctx *ctx) { delete ctx; }

int femh_homogenize(femh_ctx *ctx, const double *rho, double *C, femh_stats *stats) {
  if (int s = need_ctx(ctx)) return s;
  if (!rho || !C) return FEMH_BAD_ARG;
  return guarded(ctx, [&] { run_many(ctx, 1, rho, false, false, C, stats); });
}

int femh_homogenize_modulus(femh_ctx *ctx, const double *E, double *C, femh_stats *stats) {
  if (int s = need_ctx(ctx)) return s;
  if (!E || !C) return FEMH_BAD_ARG;
  return guarded(ctx, [&] { run_many(ctx, 1, E, false, true, C, stats); });
}

int femh_homogenize_many(femh_ctx *ctx, int n, const double *rho, double *C, femh_stats *stats) {
  if (int s = need_ctx(ctx)) return s;
  if (n < 1 || !rho || !C) return FEMH_BAD_ARG;
  return guarded(ctx, [&] { run_many(ctx, n, rho, false, false, C, stats); });
}

int femh_homogenize_many_modulus(femh_ctx *ctx, int n, const double *E, double *C,
                                 femh_stats *stats) {
  if (int s = need_ctx(ctx)) return s;
  if (n < 1 || !E || !C) return FEMH_BAD_ARG;
  return guarded(ctx, [&] { run_many(ctx, n, E, false, true, C, stats); });
}

static int many_device(femh_ctx *ctx, int n, const double *d_rho, double *d_C, void *stream,
                       femh_stats *stats) {
  if (int s = need_ctx(ctx)) return s;
  if (n < 1 || !d_rho || !d_C) return FEMH_BAD_ARG;
  return guarded(ctx, [&] {
    FEMH_HIP(hipStreamSynchronize((hipStream_t)stream));
    std::vector<double> C(36 * (size_t)n);
    run_many(ctx, n, d_rho, true, false, C.data(), stats);
    FEMH_HIP(hipMemcpyAsync(d_C, C.data(), sizeof(double) * C.size(), hipMemcpyHostToDevice,
                            ctx->c->stream));
    FEMH_HIP(hipStreamSynchronize(ctx->c->stream));
  });
}

int femh_homogenize_device(femh_ctx *ctx, const double *d_rho, double *d_C, void *stream,
                           femh_stats *stats) {
  return many_device(ctx, 1, d_rho, d_C, stream, stats);
}

int femh_homogenize_many_device(femh_ctx *ctx, int n, const double *d_rho, double *d_C,
                                void *stream, femh_stats *stats) {
  return many_device(ctx, n, d_rho, d_C, stream, stats);
}

int femh_gradient(femh_ctx *ctx, const double *dJdC, double *dJdrho) {
  if (int s = need_ctx(ctx)) return s;
  if (!dJdC || !dJdrho) return FEMH_BAD_ARG;
  return guarded(ctx, [&] { ctx->c->gradient(dJdC, dJdrho, false); });
}

int femh_gradient_device(femh_ctx *ctx, const double *dJdC, const double *shift, double *d_dJdrho,
                         double *gmax, void *stream) {
  if (int s = need_ctx(ctx)) return s;
  if (!dJdC || !d_dJdrho) return FEMH_BAD_ARG;
  return guarded(ctx, [&] {
    FEMH_HIP(hipStreamSynchronize((hipStream_t)stream));
    ctx->c->gradient_device(dJdC, shift, d_dJdrho, gmax);
  });
}

int femh_descent_trial(femh_ctx *ctx, int n, const double *d_x, const double *d_g, const double *h,
                       double lb, double ub, double *d_out, double *sums, void *stream) {
  if (int s = need_ctx(ctx)) return s;
  if (n < 1 || n > 65535 || !d_x || !d_g || !h || !d_out || !(lb <= ub)) return FEMH_BAD_ARG;
  return guarded(ctx, [&] {
    FEMH_HIP(hipStreamSynchronize((hipStream_t)stream));
    ctx->c->descent_trial(n, d_x, d_g, h, lb, ub, d_out, sums);
  });
}

int femh_energy_terms(femh_ctx *ctx, double *W) {
  if (int s = need_ctx(ctx)) return s;
  if (!W) return FEMH_BAD_ARG;
  return guarded(ctx, [&] {
    std::vector<double> zero(36 * (size_t)std::max(ctx->c->nb, 1), 0.0);
    ctx->c->gradient(zero.data(), W, true);
  });
}

int femh_coarse_strain_tensor(femh_ctx *ctx, double *C) {
  if (int s = need_ctx(ctx)) return s;
  if (!C) return FEMH_BAD_ARG;
  return guarded(ctx, [&] { ctx->c->coarse_strain(C); });
}

int femh_thermal_expansion(femh_ctx *ctx, const double *alpha, double *beta, double *alpha_eff) {
  if (int s = need_ctx(ctx)) return s;
  if (!alpha) return FEMH_BAD_ARG;
  return guarded(ctx, [&] { ctx->c->thermal(alpha, beta, alpha_eff); });
}

int femh_launch_shapes(femh_ctx *ctx, int *shapes, int n) {
  if (int s = need_ctx(ctx)) return -s;
  if (n < 0 || (n > 0 && !shapes)) return -FEMH_BAD_ARG;
  const femh::Context &c = *ctx->c;
  const int L = c.nlevels;
  for (int i = 0; i < n && i < L; i++) shapes[i] = i == 0 ? c.shape_fine : c.xfer_shape(i - 1);
  return L;
}

int femh_set_launch_shapes(femh_ctx *ctx, const int *shapes, int n) {
  if (int s = need_ctx(ctx)) return s;
  if (n < 0 || (n > 0 && !shapes)) return FEMH_BAD_ARG;
  femh::Context &c = *ctx->c;
  if (n > 0) {
    const int tile = (shapes[0] >> 8) & 0xff, planes = shapes[0] & 0xff;
    if (shapes[0] < 0 || (shapes[0] >> 16) != 0 || tile >= femh::kNumTileShapes || planes > 64) {
      ctx->err = "fine-operator shape: tile variant 0..3 in bits 8-15, at most 64 planes in bits 0-7";
      ctx->last_status = FEMH_BAD_ARG;
      return FEMH_BAD_ARG;
    }
  }
  for (int i = 1; i < n; i++)
    if (shapes[i] < 0 || shapes[i] > 8) {
      ctx->err = "transfer shape: 0 (built-in) or 1..8 coarse planes per thread";
      ctx->last_status = FEMH_BAD_ARG;
      return FEMH_BAD_ARG;
    }
  if (n > 0) c.shape_fine = shapes[0];
  if (n > 1) {
    if ((int)c.shape_xfer.size() < n - 1) c.shape_xfer.resize(n - 1, 0);
    for (int i = 1; i < n; i++) c.shape_xfer[i - 1] = shapes[i];
  }
  return FEMH_OK;
}

int femh_get_fields(femh_ctx *ctx, double *u) {
  if (int s = need_ctx(ctx)) return s;
  if (!u) return FEMH_BAD_ARG;
  return guarded(ctx, [&] { ctx->c->get_fields(u); });
}

int femh_homogenize_batch(femh_ctx *ctx, int n, const int *dims, const double *const *E, double *C,
                          femh_stats *stats) {
  if (!ctx) return FEMH_BAD_ARG;
  if (n < 0 || (n > 0 && (!dims || !E || !C))) return FEMH_BAD_ARG;
  return guarded(ctx, [&] {
    // group the samples by grid size; each group runs through a batched
    // sub-context, `batch` samples per launch sequence
    const auto t0 = std::chrono::steady_clock::now();
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;
    for (int s = 0; s < n; s++)
      groups[std::make_tuple(dims[3 * s], dims[3 * s + 1], dims[3 * s + 2])].push_back(s);
    // Work items are batches of up to `batch` cells of one size, largest
    // cells first.  A few host threads take them in turn, each on a free
    // sub-context of that size (created on first need, each with its own
    // stream), so batches of one size and of different sizes are in flight
    // together: one batch of small cells leaves most of the GPU idle, and a
    // slow cell (an FP64 fallback) no longer holds the others back.  Every
    // batch writes its own rows of C; no sub-context warm-starts, so a
    // cell's result does not depend on which sub-context solved it.
    struct Item {
      std::tuple<int, int, int> dims;
      const std::vector<int> *ids;
      size_t s0;
      int k;
    };
    std::vector<Item> items;
    int B = ctx->opts.batch < 1 ? 1 : ctx->opts.batch;
    for (auto it = groups.rbegin(); it != groups.rend(); ++it)
      for (size_t s0 = 0; s0 < it->second.size(); s0 += B)
        items.push_back({it->first, &it->second, s0, (int)std::min<size_t>(B, it->second.size() - s0)});
    std::mutex mu;
    std::map<std::tuple<int, int, int>, std::vector<bool>> busy;
    auto acquire = [&](const std::tuple<int, int, int> &d) -> std::pair<femh::Context *, size_t> {
      std::lock_guard<std::mutex> lk(mu);
      auto &pool = ctx->sub[d];
      auto &b = busy[d];
      b.resize(pool.size(), false);
      for (size_t i = 0; i < pool.size(); i++)
        if (!b[i]) {
          b[i] = true;
          return {pool[i].get(), i};
        }
      femh_opts o = ctx->opts;
      o.nx = std::get<0>(d);
      o.ny = std::get<1>(d);
      o.nz = std::get<2>(d);
      o.interp = FEMH_INTERP_MODULUS;
      o.batch = B;
      o.warm_start = 0;
      pool.emplace_back(new femh::Context(o));
      b.push_back(true);
      return {pool.back().get(), pool.size() - 1};
    };
    auto release = [&](const std::tuple<int, int, int> &d, size_t i) {
      std::lock_guard<std::mutex> lk(mu);
      busy[d][i] = false;
    };
    std::vector<int> maxit(items.size(), 0);
    std::vector<std::exception_ptr> errs(items.size());
    std::atomic<size_t> next{0};
    auto worker = [&] {
      for (size_t w = next++; w < items.size(); w = next++) {
        const Item &it = items[w];
        std::pair<femh::Context *, size_t> slot{nullptr, 0};
        try {
          FEMH_HIP(hipSetDevice(ctx->opts.device));
          slot = acquire(it.dims);
          femh::Context &c = *slot.first;
          const size_t V = (size_t)c.g.V;
          std::vector<double> buf((size_t)it.k * V), Cg(36 * (size_t)c.B);
          std::vector<femh_stats> st(c.B);
          for (int j = 0; j < it.k; j++)
            std::memcpy(buf.data() + j * V, E[(*it.ids)[it.s0 + j]], sizeof(double) * V);
          c.load(buf.data(), it.k, false, true);
          c.homogenize(Cg.data(), st.data());
          for (int j = 0; j < it.k; j++) {
            std::memcpy(C + 36 * (size_t)(*it.ids)[it.s0 + j], Cg.data() + 36 * j, sizeof(double) * 36);
            for (int q = 0; q < 6; q++) maxit[w] = std::max(maxit[w], st[j].iters[q]);
          }
        } catch (...) {
          errs[w] = std::current_exception();
        }
        if (slot.first) release(it.dims, slot.second);
      }
    };
    const int want = ctx->opts.batch_threads > 0 ? ctx->opts.batch_threads : 8;
    const size_t nthreads = std::min<size_t>(items.size(), (size_t)want);
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nthreads; t++) pool.emplace_back(worker);
    if (nthreads) worker();
    for (auto &t : pool) t.join();
    for (auto &e : errs)
      if (e) std::rethrow_exception(e);  // the first failed batch in work order
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      for (int m : maxit) stats->iters[0] = std::max(stats->iters[0], m);
      stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
  });
}

int femh_prepare(femh_ctx *ctx, const double *E) {
  if (int s = need_ctx(ctx)) return s;
  if (!E) return FEMH_BAD_ARG;
  return guarded(ctx, [&] {
    ctx->c->load(E, 1, false, true);
    ctx->c->build_hierarchy();
  });
}

int femh_apply_level(femh_ctx *ctx, int level, const double *u, double *y) {
  if (int s = need_ctx(ctx)) return s;
  if (!u || !y) return FEMH_BAD_ARG;
  return guarded(ctx, [&] { ctx->c->apply_level(level, u, y); });
}

int femh_level_dims(femh_ctx *ctx, int level, int *dims) {
  if (int s = need_ctx(ctx)) return s;
  if (!dims) return FEMH_BAD_ARG;
  femh::Context &c = *ctx->c;
  if (!c.h64.ready || level < 0 || level >= c.nlevels) return FEMH_BAD_ARG;
  dims[0] = c.h64.lv[level].g.nx;
  dims[1] = c.h64.lv[level].g.ny;
  dims[2] = c.h64.lv[level].g.nz;
  return FEMH_OK;
}

int femh_set_profiling(femh_ctx *ctx, int enable) {
  if (int s = need_ctx(ctx)) return s;
  return guarded(ctx, [&] { ctx->c->set_profiling(enable != 0); });
}

int femh_kernel_stats(femh_ctx *ctx, int which, long long *launches, double *ms, double *bytes,
                      double *flops) {
  if (int s = need_ctx(ctx)) return s;
  if (which < 0 || which > 1) return FEMH_BAD_ARG;
  return guarded(ctx, [&] {
    ctx->c->flush_profile();
    const auto &k = ctx->c->kst[which];
    if (launches) *launches = k.launches;
    if (ms) *ms = k.ms;
    if (bytes) *bytes = k.bytes;
    if (flops) *flops = k.flops;
  });
}

int femh_last_error(femh_ctx *ctx, char *buf, size_t len) {
  if (!ctx) return FEMH_BAD_ARG;
  if (buf && len) {
    std::strncpy(buf, ctx->err.c_str(), len - 1);
    buf[len - 1] = '\0';
  }
  return ctx->last_status;
}

}  // extern "C"
