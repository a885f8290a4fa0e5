// gpue_noalloc.hip -- the allocation-free variant of the vortex detector.
//
// gpue_vortex_marker_2d allocates its scratch (plaquette mask, candidates,
// marker) on its first call; callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc)
// call gpue_vortex_marker_2d_noalloc instead.  Its scratch is one file-scope
// device array of fixed size (enough for every supported square grid up to
// 1024^2), which each GPU the library runs on holds until the process ends.
// Every call on a GPU shares that GPU's copy: calls that overlap on one GPU
// race (each call is synchronous, so calls from one host thread never overlap).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "ctx_internal.h"

using namespace gpue;

namespace {

constexpr int NOALLOC_MAX_XDIM = 1024;
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NP_MAX = (size_t)(NOALLOC_MAX_XDIM - 1) * (NOALLOC_MAX_XDIM - 1);
constexpr size_t OFF_CAND = align256(NP_MAX);                                   // mask: 1 byte per plaquette
constexpr size_t OFF_MARKER = OFF_CAND + align256(NP_MAX * sizeof(int));        // candidates
constexpr size_t OFF_FOUND = OFF_MARKER + align256((size_t)NOALLOC_MAX_XDIM * NOALLOC_MAX_XDIM * sizeof(int));
constexpr size_t SCRATCH_BYTES = OFF_FOUND + 256;

}  // namespace

// The scratch array.  In host code this name is not a device address: host code
// uses the per-device addresses hipGetSymbolAddress returns (scratch_ptr).
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};

// The device address of this GPU's scratch, resolved once per device with that
// device current.
hipError_t scratch_of(int device, unsigned char** out) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_ptr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    scratch_ptr[device] = p;
  }
  *out = (unsigned char*)scratch_ptr[device];
  return hipSuccess;
}

// The plaquette mask of findVortex on the device: sqrt(x_i^2 + x_j^2) < radius
// (src/tracker.cu:111-113), the same predicate the host mask evaluates.
__global__ void k_vortex_mask(const double* x, int xDim, double radius, unsigned char* mask) {
  const int P = xDim - 1;
  const long np = (long)P * P;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < np; idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx / P), j = (int)(idx % P);
    mask[idx] = sqrt(x[i] * x[i] + x[j] * x[j]) < radius ? 1 : 0;
  }
}

}  // namespace

extern "C" {

gpue_status gpue_vortex_marker_2d_noalloc(gpue_ctx* c, double radius, int* marker_out, int* found) {
  if (!c || !marker_out) return GPUE_EINVAL;
  if (c->p.dimnum != 2 || c->dims[0] != c->dims[1])
    return fail(c, GPUE_EUNSUPPORTED, "findVortex needs a square 2D grid (src/tracker.cu:111-113)");
  const int xDim = c->dims[0];
  if (xDim > NOALLOC_MAX_XDIM)
    return fail(c, GPUE_EINVAL, "the allocation-free vortex detector takes grids up to 1024^2");
  unsigned char* base = nullptr;
  hipError_t e = scratch_of(c->p.device, &base);
  if (e != hipSuccess) return fail(c, GPUE_EHIP, std::string("scratch address: ") + hipGetErrorString(e));
  gpue_status st = apply_pending(c);
  if (st != GPUE_OK) return st;
  unsigned char* mask = base;
  int* cand = (int*)(base + OFF_CAND);
  int* marker = (int*)(base + OFF_MARKER);
  int* dfound = (int*)(base + OFF_FOUND);
  const long np = (long)(xDim - 1) * (xDim - 1);
  const int blocks = (int)std::min<long>((np + 255) / 256, 4096);
  hipLaunchKernelGGL(k_vortex_mask, dim3(blocks), dim3(256), 0, c->stream, c->dgrid[0], xDim, radius, mask);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = c->prec == 64 ? launch_vortex<double>(c->psi, mask, xDim, cand, marker, dfound, c->stream)
                      : launch_vortex<float>(c->psi, mask, xDim, cand, marker, dfound, c->stream);
  if (e != hipSuccess) return fail(c, GPUE_EHIP, hipGetErrorString(e));
  int f = 0;
  e = hipMemcpyAsync(marker_out, marker, (size_t)xDim * xDim * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&f, dfound, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, GPUE_EHIP, hipGetErrorString(e));
  if (found) *found = f;
  return GPUE_OK;
}

}  // extern "C"
