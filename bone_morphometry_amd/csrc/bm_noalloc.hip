// bm_noalloc.hip -- the thickness/spacing part of analyze for callers that
// cannot allocate device memory (code that must not call hipMalloc or
// hipHostMalloc: tight serving loops, harnesses that forbid allocation).
//
// Same kernels and results as bm_analyze_thickness_spacing_device; every
// device buffer of the call (both phases' EDT passes, ridge lists, masks,
// maps, quantile state, the plan's descriptors and results) is laid out in
// `scratch`, a device global of kNoAllocScratchBytes.  Each GPU holds its own
// copy from the first call on it until the process ends; its address there is
// looked up once per device with hipGetSymbolAddress, and only that address
// is handed to kernels and copies.  The kernels clear what must start at zero
// (k_prepare) and fully write everything else, so nothing of an earlier
// call's contents is read.  Calls on one device share the array: the library
// serialises them and each returns only when its results are complete.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/bm.h"
#include "bm_analyze.h"
#include "bm_engine.h"

namespace bm {

constexpr size_t kNoAllocScratchBytes = BM_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct NoAllocDevice {
  void* ptr = nullptr;  // device address of `scratch` on this device
  Plan plan;            // laid out entirely in ptr (no allocation)
  std::mutex mu;
};

std::mutex g_mu;
std::map<int, NoAllocDevice*>& devices() {
  static std::map<int, NoAllocDevice*>* m = new std::map<int, NoAllocDevice*>();
  return *m;
}

// With `device` current: its scratch address, looked up on the first call.
NoAllocDevice& device_state(int device) {
  std::lock_guard<std::mutex> g(g_mu);
  auto& m = devices();
  auto it = m.find(device);
  if (it != m.end()) return *it->second;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    throw DeviceError(std::string("hipGetSymbolAddress(scratch) failed: ") +
                      hipGetErrorString(e));
  NoAllocDevice* st = new NoAllocDevice();
  st->ptr = ptr;
  m.emplace(device, st);
  return *st;
}

}  // namespace

// Device bytes the no-allocation form needs for a volume of these dimensions
// in this mode (the plan of its two items plus their descriptors and results).
size_t noalloc_bytes(const Dims& dims, int mode) {
  std::vector<ItemSpec> specs;
  static const uint8_t kNotRead = 0;  // layout only: no pointer is dereferenced
  const double unit[3] = {1.0, 1.0, 1.0};
  analyze_items(&kNotRead, &kNotRead, dims, unit, mode, specs);
  return Plan::bytes_for(specs) + Plan::meta_bytes_for(specs.size());
}

void analyze_noalloc(const uint8_t* d_bone, const uint8_t* d_roi, const Dims& dims,
                     const double spacing[3], int mode, bm_thickness_spacing* out,
                     hipStream_t stream) {
  check_spacing(spacing);
  check_mode(mode);
  const size_t need = noalloc_bytes(dims, mode);
  if (need > kNoAllocScratchBytes)
    throw std::range_error("volume " + std::to_string(dims.w) + "x" + std::to_string(dims.h) +
                           "x" + std::to_string(dims.d) + " needs " + std::to_string(need) +
                           " bytes of device scratch; the no-allocation entry point holds " +
                           std::to_string(kNoAllocScratchBytes));
  if (!d_bone || !d_roi) throw Error("null volume");
  const int device = current_device();
  NoAllocDevice& st = device_state(device);
  hipStream_t s = stream ? stream : engine_for_device(device).stream();
  std::lock_guard<std::mutex> g(st.mu);
  std::vector<ItemSpec> specs;
  analyze_items(d_bone, d_roi, dims, spacing, mode, specs);
  st.plan.build(specs, s, st.ptr, kNoAllocScratchBytes, /*ext_meta=*/true);
  st.plan.launch(s);
  st.plan.fetch(s);
  analyze_finish(st.plan.result(0), st.plan.result(1), dims.n(), spacing, mode, out);
}

}  // namespace bm

namespace {
struct NoAllocCall {
  const uint8_t *bone, *roi;
  uint64_t w, h, d;
  const double* spacing;
  int mode;
  bm_thickness_spacing* out;
  void* stream;
};
void noalloc_call(void* p) {
  NoAllocCall* c = static_cast<NoAllocCall*>(p);
  if (!c->spacing || !c->out) throw bm::Error("null argument");
  const bm::Dims dims = bm::device_dims(c->w, c->h, c->d);
  bm::analyze_noalloc(c->bone, c->roi, dims, c->spacing, c->mode, c->out,
                      static_cast<hipStream_t>(c->stream));
}
}  // namespace

extern "C" int bm_analyze_thickness_spacing_device_noalloc(
    const uint8_t* d_bone, const uint8_t* d_roi, uint64_t width, uint64_t height, uint64_t depth,
    const double voxel_spacing[3], int feature_mode, bm_thickness_spacing* out, void* stream) {
  NoAllocCall c{d_bone, d_roi, width, height, depth, voxel_spacing, feature_mode, out, stream};
  return bm::capture_errors(noalloc_call, &c, nullptr);
}
