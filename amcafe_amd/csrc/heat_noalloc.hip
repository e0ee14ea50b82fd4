// heat_noalloc.hip -- the differentiable temperature field's loss and adjoint gradient
// (the batch form of heat_adjoint.hip, FP32) for callers that cannot allocate device memory:
// every device buffer of a call lives in one static device array, `scratch`.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "amcafe_dev.h"
#include "heat_kernels.h"

using namespace amcafe_heat;

// Every device buffer of a call lives in `scratch`: one array per device, kept until the
// process ends; calls that overlap on one device race on it.  Nothing in it needs clearing:
// the first segment overwrites every block slot.  Host code never passes the array's name to
// a kernel or a copy (on the host it is only a stand-in): it uses the device address that
// hipGetSymbolAddress returns for the current device, fetched once per device.
constexpr size_t kScratchBytes = 8u << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// scratch bytes a call needs: sources, block slots, sums, counter
size_t noalloc_bytes(const amcafe_heat_desc* d, int nb) {
  return align256(sizeof(Src<float>) * 2 * (size_t)nb) + align256(sizeof(float) * 5 * (size_t)n_blocks(d) * nb) +
         align256(sizeof(float) * 5 * (size_t)nb) + align256(sizeof(unsigned long long));
}

int run_noalloc(const amcafe_heat_desc* d, int nb, const double* theta, const double* target, int step0, int nsteps,
                double* loss, double* grad, char* err, int errlen) {
  const size_t need = noalloc_bytes(d, nb);
  if (need > kScratchBytes) {
    snprintf(err, errlen, "the allocation-free heat call needs %zu bytes of scratch for %d instance(s) of %d voxels; "
             "its maximum is %zu", need, nb, d->nX[0] * d->nX[1] * d->nX[2], kScratchBytes);
    return -2;
  }
  int dev = 0;
  HCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    snprintf(err, errlen, "device index %d beyond %d", dev, kMaxDevices);
    return -1;
  }
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      HCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      g_scratch_ptr[dev] = p;
    }
    base = (unsigned char*)g_scratch_ptr[dev];
  }
  std::vector<Src<float>> hs;
  sources<float>(nb, theta, target, hs);
  Bufs<float> B;
  size_t off = 0;
  B.src = (Src<float>*)(base + off);
  off += align256(sizeof(Src<float>) * 2 * (size_t)nb);
  B.part = (float*)(base + off);
  off += align256(sizeof(float) * 5 * (size_t)n_blocks(d) * nb);
  B.out = (float*)(base + off);
  off += align256(sizeof(float) * 5 * (size_t)nb);
  B.evals = nullptr;
  HCHECK(hipMemcpy(B.src, hs.data(), sizeof(Src<float>) * 2 * nb, hipMemcpyHostToDevice));
  launch_all<float, true, true>(d, nb, B, step0, nsteps, 0);
  HCHECK(hipGetLastError());
  std::vector<float> h(5 * (size_t)nb);
  HCHECK(hipMemcpy(h.data(), B.out, sizeof(float) * 5 * nb, hipMemcpyDeviceToHost));
  results<float>(nb, h, true, loss, grad);
  return 0;
}

}  // namespace

extern "C" size_t amcd_heat_noalloc_bytes(const amcafe_heat_desc* d, int nb) { return noalloc_bytes(d, nb); }
extern "C" size_t amcd_heat_noalloc_max(void) { return kScratchBytes; }
extern "C" int amcd_heat_noalloc(const amcafe_heat_desc* d, int nb, const double* theta, const double* target,
                                 int step0, int nsteps, double* loss, double* grad, char* err, int errlen) {
  return run_noalloc(d, nb, theta, target, step0, nsteps, loss, grad, err, errlen);
}
