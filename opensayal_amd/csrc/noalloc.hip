// noalloc.hip - osy_fluid_traces_noalloc: the path lines of osy_fluid_traces
// without any device allocation.
//
// osy_fluid_traces stages the two trace arrays in a per-handle device buffer
// that it allocates on first use and grows with hipMalloc when a call asks
// for more entries.  Callers that may not allocate (a serving loop that
// forbids it, a harness around captured work) use this variant instead: the
// staging lives in `scratch`, a file-scope device array of fixed size that
// each GPU holds for the life of the process.  The kernels are the same
// (enqueue_traces, kernels_view.hip), so the results are the same bits.
//
// `scratch` names a host-side stand-in: host code uses only the device
// address hipGetSymbolAddress returns for it, looked up once per device
// (with that device current) and kept in g_scratch.  Every call on one GPU
// shares that GPU's array; calls that overlap on one GPU (two threads, or
// two handles' streams) race on it.  The trace kernel writes every entry it
// returns, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "fluid_internal.hpp"

namespace {
// Entries of each of the two staging arrays (traces_x, traces_y): 2^18, so
// the array is 2 MiB per GPU.  The 1920x1080 viewer default (distance 20,
// length 20) needs 103,680.
constexpr size_t kTraceScratchEntries = size_t(1) << 18;
}  // namespace

static __device__ __align__(256) int32_t scratch[2 * kTraceScratchEntries];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up on first use.
int scratch_on(int device, int32_t** out) {
  if (device < 0 || device >= kMaxDevices)
    return osy::set_error(OSY_E_INVALID, "device ordinal out of range for the trace scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    int cur = 0;
    OSY_HIP_CHECK(hipGetDevice(&cur));
    if (cur != device) OSY_HIP_CHECK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) OSY_HIP_CHECK(hipSetDevice(cur));
    if (e != hipSuccess)
      return osy::set_error(OSY_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                           hipGetErrorString(e));
    if (!p) return osy::set_error(OSY_E_HIP, "hipGetSymbolAddress(scratch) gave no address");
    g_scratch[device] = (int32_t*)p;
  }
  *out = g_scratch[device];
  return OSY_OK;
}

}  // namespace

extern "C" int osy_fluid_traces_noalloc(osy_fluid* fl, int32_t distance, int32_t length,
                                        float d_t, int32_t* traces_x, int32_t* traces_y,
                                        size_t bytes) {
  using namespace osy;
  if (!fl || !traces_x || !traces_y) return set_error(OSY_E_INVALID, "null argument");
  if (distance < 1) return set_error(OSY_E_INVALID, "distance must be at least 1");
  if (length < 1 || length > 65536)
    return set_error(OSY_E_INVALID, "length must be in [1, 65536]");
  if (!std::isfinite(d_t)) return set_error(OSY_E_INVALID, "d_t must be finite");
  const size_t seeds = (size_t)(fl->W / distance) * (size_t)(fl->H / distance);
  const size_t entries = seeds * (size_t)length;
  if (bytes != entries * 4)
    return set_error(OSY_E_INVALID,
                     "each buffer must hold (W/distance)*(H/distance)*length int32 entries");
  if (entries > kTraceScratchEntries)
    return set_error(OSY_E_INVALID,
                     "the allocation-free path lines hold at most " +
                         std::to_string(kTraceScratchEntries) + " entries per buffer; this call "
                         "needs " + std::to_string(entries) + " (use osy_fluid_traces)");
  if (entries == 0) return OSY_OK;
  if (int st = batch_wait(fl)) return st;
  int32_t* base = nullptr;
  if (int st = scratch_on(fl->device, &base)) return st;
  int32_t* dx = base;
  int32_t* dy = base + kTraceScratchEntries;
  const StepParams sp = fl->params(d_t, nullptr);
  enqueue_traces(fl->storage, sp, fl->f, fl->uv, distance, length, d_t, dx, dy, fl->stream);
  OSY_HIP_CHECK(hipGetLastError());
  OSY_HIP_CHECK(hipMemcpyAsync(traces_x, dx, bytes, hipMemcpyDeviceToHost, fl->stream));
  OSY_HIP_CHECK(hipMemcpyAsync(traces_y, dy, bytes, hipMemcpyDeviceToHost, fl->stream));
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  return OSY_OK;
}
