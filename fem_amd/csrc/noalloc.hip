// noalloc.hip -- femh_homogenize_device_noalloc: the Jacobi-PCG FP64
// homogenization of femh_homogenize_device with every device buffer carved
// from one fixed region compiled into the library (`scratch`, below), for
// callers that cannot allocate device memory.
//
// The region is a file-scope device array of kScratchBytes.  Each GPU that
// loads the library holds its own copy until the process ends; its address on
// a device comes from hipGetSymbolAddress, taken once per device with that
// device current (the array's name is a host-side stand-in, never a device
// address).  A call builds a solver context whose buffers are bump-allocated
// from the region (femh::Arena) on the caller's stream: the kernels, their
// order and therefore the results are those of femh_homogenize_device with
// the same options.  Calls on one device share the region, so they are
// serialised by a per-device mutex here (a call returns after its work is
// complete); nothing a call leaves in the region is read by the next one
// before it is written (the context clears what must start at zero with its
// own kernels).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/femh.h"
#include "solver.hpp"

namespace {

// 64 MiB: the Jacobi FP64 solve of one cell needs ~19 MB of reduction
// partials plus ~936 bytes per voxel, so every grid of up to 32^3 = 32768
// voxels fits (FEMH_NOALLOC_MAX_VOXELS)
constexpr size_t kScratchBytes = (size_t)64 << 20;
static_assert(FEMH_NOALLOC_MAX_VOXELS == 32768, "femh.h states the voxel limit");

constexpr int kMaxDevices = 64;
std::mutex g_mu[kMaxDevices];
void *g_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// the region's address on `device` (which is current), looked up on the
// device's first call
void *scratch_on(int device) {
  if (!g_ptr[device]) {
    void *p = nullptr;
    FEMH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw femh::Error(FEMH_DEVICE_ERROR, "no device address for the scratch region");
    g_ptr[device] = p;
  }
  return g_ptr[device];
}

void set_err(char *err, size_t len, const char *msg) {
  if (err && len) {
    std::strncpy(err, msg, len - 1);
    err[len - 1] = '\0';
  }
}

}  // namespace

extern "C" int femh_homogenize_device_noalloc(const femh_opts *o, const double *d_rho, double *d_C,
                                              void *stream, femh_stats *stats, char *err,
                                              size_t err_len) {
  set_err(err, err_len, "");
  if (!o || !d_rho || !d_C) {
    set_err(err, err_len, "null argument");
    return FEMH_BAD_ARG;
  }
  if (o->device < 0 || o->device >= kMaxDevices) {
    set_err(err, err_len, "device ordinal out of range");
    return FEMH_BAD_ARG;
  }
  const long long V = (long long)o->nx * o->ny * o->nz;
  if (o->nx < 2 || o->ny < 2 || o->nz < 2 || V > FEMH_NOALLOC_MAX_VOXELS) {
    const std::string m = "grid " + std::to_string(o->nx) + "x" + std::to_string(o->ny) + "x" +
                          std::to_string(o->nz) + " is outside the no-allocation limit (each edge >= 2, at most " +
                          std::to_string(FEMH_NOALLOC_MAX_VOXELS) + " voxels)";
    set_err(err, err_len, m.c_str());
    return FEMH_BAD_ARG;
  }
  if (o->precond != FEMH_PRECOND_JACOBI || o->precision != FEMH_FP64 || o->warm_start || o->autotune) {
    set_err(err, err_len,
            "the no-allocation entry point runs Jacobi-PCG in FP64 without warm start or autotuning");
    return FEMH_BAD_ARG;
  }
  try {
    std::lock_guard<std::mutex> lk(g_mu[o->device]);
    FEMH_HIP(hipSetDevice(o->device));
    femh::Arena arena;
    arena.base = static_cast<char *>(scratch_on(o->device));
    arena.cap = kScratchBytes;
    arena.stream = static_cast<hipStream_t>(stream);
    femh::ArenaScope scope(&arena);
    femh_opts oo = *o;
    oo.batch = 1;
    oo.streams = 1;
    oo.graph = 0;  // recording a launch sequence allocates
    femh::Context c(oo);
    c.load(d_rho, 1, true, false);
    double C[36];
    c.homogenize(C, stats);
    FEMH_HIP(hipMemcpyAsync(d_C, C, sizeof(C), hipMemcpyHostToDevice, c.stream));
    FEMH_HIP(hipStreamSynchronize(c.stream));
    return FEMH_OK;
  } catch (const femh::Error &e) {
    set_err(err, err_len, e.what());
    return e.status;
  } catch (const std::exception &e) {
    set_err(err, err_len, e.what());
    return FEMH_DEVICE_ERROR;
  }
}
