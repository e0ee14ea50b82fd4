// heat_adjoint.hip -- the differentiable temperature field (an extension: the reference
// has no adjoint).  The moving Gaussian source of Utilities::Temp_Fun_G (S/Utilities.C:228-249)
// with constant properties (conductivity k, diffusivity alpha) on a straight +x track,
// integrated over the last t_hist seconds with the same 25-point Gauss-Legendre rule as
// Utilities::Integral (S/Utilities.C:105-124):
//
//   T(R, t) = T0 + sum_i W_i f(tau_i),   tau_i = h (1 - xi_i),  W_i = h w_i,  h = t_hist / 2
//   f(tau)  = C tau^-1/2 / (2 alpha tau + sigma^2)
//             * exp(-(dx^2 + dy^2) / (4 alpha tau + 2 sigma^2) - dz^2 / (4 alpha tau))
//   dx = Rx - x_s - v (t - tau),  dy = Ry - y_s,  dz = Rz - z_s,
//   C  = eta P / (pi (k / alpha) sqrt(4 pi alpha)),
//
// parameters theta = (P, v, eta, sigma).  The loss is the squared difference of a
// sigmoid-smoothed melt indicator s(T) = 1 / (1 + exp(-(T - tL) / width)) between theta and a
// target theta*, summed over voxels and steps t_n = n dt, n = step0+1 .. step0+nsteps.
//
// The field at a step is an explicit function of (theta, t): nothing is carried from step
// to step, so the backward pass recomputes the field (and the source-history quadrature)
// step by step instead of storing it, and each launch covers one segment of steps -- the
// segment is the checkpoint interval.  The adjoint of the loss with respect to T is
// 2 (s - s*) s (1 - s) / width per voxel and step; it is contracted with dT/dtheta, whose
// four components come from the same quadrature nodes (dT/dP = (T - T0)/P, dT/deta = (T - T0)/eta,
// df/dv = f 2 dx (t - tau) / D1, df/dsigma = f (-2 sigma / A + 4 sigma (dx^2 + dy^2) / D1^2)).
// FP32 arithmetic with FP32 accumulation per thread and per block; an FP64 instantiation of
// the forward serves finite differences.
#include <hip/hip_runtime.h>

#include <vector>

#include "amcafe_dev.h"
#include "heat_kernels.h"

using namespace amcafe_heat;

namespace {

thread_local unsigned long long g_last_evals = 0;

template <typename F, bool GRAD, bool PRUNE>
int run(const amcafe_heat_desc* d, int nb, const double* theta, const double* target, int step0, int nsteps,
        double* loss, double* grad, float* ms, char* err, int errlen) {
  const size_t nblk = (size_t)n_blocks(d);
  std::vector<Src<F>> hs;
  sources<F>(nb, theta, target, hs);
  Bufs<F> B = {nullptr, nullptr, nullptr, nullptr};
  HCHECK(hipMalloc(&B.src, sizeof(Src<F>) * 2 * nb));
  HCHECK(hipMalloc(&B.part, sizeof(F) * 5 * nblk * nb));
  HCHECK(hipMalloc(&B.out, sizeof(F) * 5 * nb));
  HCHECK(hipMalloc(&B.evals, sizeof(unsigned long long)));
  HCHECK(hipMemcpy(B.src, hs.data(), sizeof(Src<F>) * 2 * nb, hipMemcpyHostToDevice));
  HCHECK(hipMemset(B.evals, 0, sizeof(unsigned long long)));
  hipEvent_t e0, e1;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  HCHECK(hipEventRecord(e0, 0));
  launch_all<F, GRAD, PRUNE>(d, nb, B, step0, nsteps, 0);
  HCHECK(hipGetLastError());
  HCHECK(hipEventRecord(e1, 0));
  std::vector<F> h(5 * (size_t)nb);
  HCHECK(hipMemcpy(h.data(), B.out, sizeof(F) * 5 * nb, hipMemcpyDeviceToHost));
  HCHECK(hipMemcpy(&g_last_evals, B.evals, sizeof(unsigned long long), hipMemcpyDeviceToHost));
  float el = 0;
  HCHECK(hipEventElapsedTime(&el, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HCHECK(hipFree(B.src));
  HCHECK(hipFree(B.part));
  HCHECK(hipFree(B.out));
  HCHECK(hipFree(B.evals));
  results<F>(nb, h, GRAD, loss, grad);
  if (ms) *ms = el;
  return 0;
}

}  // namespace


// mode 0: FP32 loss, 1: FP32 loss and gradient, 2: FP64 loss (finite differences, no pruning)
extern "C" int amcd_heat(const amcafe_heat_desc* d, int nb, const double* theta, const double* target, int step0,
                         int nsteps, int mode, double* loss, double* grad, float* ms, char* err, int errlen) {
  if (mode == 0) return run<float, false, true>(d, nb, theta, target, step0, nsteps, loss, nullptr, ms, err, errlen);
  if (mode == 1) return run<float, true, true>(d, nb, theta, target, step0, nsteps, loss, grad, ms, err, errlen);
  return run<double, false, false>(d, nb, theta, target, step0, nsteps, loss, nullptr, ms, err, errlen);
}
extern "C" unsigned long long amcd_heat_last_evals(void) { return g_last_evals; }
