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

#include <math.h>
#include <stdio.h>

#include <vector>

#include "../../include/amcafe.h"
#include "amcafe_dev.h"
#include "gl25.h"

namespace {

__constant__ double kHx[AMC_GL25_N] = AMC_GL25_X;
__constant__ double kHw[AMC_GL25_N] = AMC_GL25_W;

template <typename F>
struct Src {  // everything a voxel-step evaluation needs, in the arithmetic type
  F P, v, eta, sigma;
};

template <typename F>
struct Geo {
  int nx, ny, nz;
  F dx, xs, ys, zs, dt, T0, tL, width, k, alpha, h;
};

// T at voxel (x, y, z) and time t; with GRAD also dT/dtheta
template <typename F, bool GRAD>
__device__ inline F field(const Geo<F>& g, const Src<F>& s, F x, F y, F z, F t, F* dT) {
  const F pi = (F)M_PI;
  const F C = s.eta * s.P / (pi * (g.k / g.alpha) * sqrt((F)4 * pi * g.alpha));
  const F dy = y - g.ys, dz = z - g.zs;
  const F s2 = s.sigma * s.sigma;
  F sum = 0, sv = 0, ss = 0;
  for (int i = 0; i < AMC_GL25_N; ++i) {
    const F tau = g.h * ((F)1 - (F)kHx[i]);
    const F W = g.h * (F)kHw[i];
    const F dx = x - g.xs - s.v * (t - tau);
    const F A = (F)2 * g.alpha * tau + s2;
    const F D1 = (F)4 * g.alpha * tau + (F)2 * s2;
    const F r2 = dx * dx + dy * dy;
    const F f = C / (sqrt(tau) * A) * exp(-r2 / D1 - dz * dz / ((F)4 * g.alpha * tau));
    sum += W * f;
    if (GRAD) {
      sv += W * f * ((F)2 * dx * (t - tau) / D1);
      ss += W * f * (-(F)2 * s.sigma / A + (F)4 * s.sigma * r2 / (D1 * D1));
    }
  }
  if (GRAD) {
    dT[0] = sum / s.P;
    dT[1] = sv;
    dT[2] = sum / s.eta;
    dT[3] = ss;
  }
  return g.T0 + sum;
}

template <typename F>
__device__ inline F indicator(const Geo<F>& g, F T) {
  return (F)1 / ((F)1 + exp(-(T - g.tL) / g.width));
}

template <typename F>
__device__ inline F block_sum(F v, F* sh) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  F t = 0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sh[q];
  return t;
}

// one thread per voxel and instance (blockIdx.y), steps step0+1 .. step0+nsteps; for
// instance b, out[8b] += loss, out[8b+1..8b+4] += grad
template <typename F, bool GRAD>
__global__ __launch_bounds__(256) void k_heat(Geo<F> g, const Src<F>* S, const Src<F>* Tg, int step0, int nsteps,
                                              F* out) {
  __shared__ F sh[4];
  const Src<F> s = S[blockIdx.y], tgt = Tg[blockIdx.y];
  out += 8 * blockIdx.y;
  const int v = blockIdx.x * 256 + threadIdx.x;
  const int n = g.nx * g.ny * g.nz;
  F loss = 0, gr[4] = {0, 0, 0, 0};
  if (v < n) {
    const int k = v / (g.nx * g.ny), r = v - k * g.nx * g.ny, j = r / g.nx, i = r - j * g.nx;
    const F x = ((F)i + (F)0.5) * g.dx, y = ((F)j + (F)0.5) * g.dx, z = ((F)k + (F)0.5) * g.dx;
    for (int m = step0 + 1; m <= step0 + nsteps; ++m) {
      const F t = (F)m * g.dt;
      F dT[4];
      const F T = field<F, GRAD>(g, s, x, y, z, t, dT);
      const F Ts = field<F, false>(g, tgt, x, y, z, t, nullptr);
      const F a = indicator(g, T), b = indicator(g, Ts);
      loss += (a - b) * (a - b);
      if (GRAD) {
        const F lam = (F)2 * (a - b) * a * ((F)1 - a) / g.width;  // dloss/dT
        for (int q = 0; q < 4; ++q) gr[q] += lam * dT[q];
      }
    }
  }
  F t = block_sum(loss, sh);
  if (threadIdx.x == 0) atomicAdd(&out[0], t);
  if (GRAD) {
    for (int q = 0; q < 4; ++q) {
      t = block_sum(gr[q], sh);
      if (threadIdx.x == 0) atomicAdd(&out[1 + q], t);
    }
  }
}

template <typename F>
Geo<F> geo(const amcafe_heat_desc* d) {
  Geo<F> g;
  g.nx = d->nX[0];
  g.ny = d->nX[1];
  g.nz = d->nX[2];
  g.dx = (F)d->dX;
  g.xs = (F)d->x_start;
  g.ys = (F)d->y_track;
  g.zs = (F)d->z_track;
  g.dt = (F)d->dt;
  g.T0 = (F)d->T0;
  g.tL = (F)d->tL;
  g.width = (F)d->width;
  g.k = (F)d->k;
  g.alpha = (F)d->alpha;
  g.h = (F)(d->t_hist / 2);
  return g;
}
template <typename F>
Src<F> src(const double* th) {
  Src<F> s;
  s.P = (F)th[0];
  s.v = (F)th[1];
  s.eta = (F)th[2];
  s.sigma = (F)th[3];
  return s;
}

#define HCHECK(x)                                                          \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      snprintf(err, errlen, "%s: %s", #x, hipGetErrorString(e_));          \
      return -1;                                                           \
    }                                                                      \
  } while (0)

// segments of at most `seg` steps, one launch each (all instances of the batch in it); sums
// accumulate on the device in F
template <typename F, bool GRAD>
int run(const amcafe_heat_desc* d, int nb, const double* theta, const double* target, int step0, int nsteps,
        double* loss, double* grad, float* ms, char* err, int errlen) {
  const int n = d->nX[0] * d->nX[1] * d->nX[2];
  const int seg = d->checkpoint > 0 ? d->checkpoint : 100;
  F* out = nullptr;
  Src<F>* dsrc = nullptr;
  std::vector<Src<F>> hs(2 * (size_t)nb);
  for (int b = 0; b < nb; ++b) {
    hs[b] = src<F>(theta + 4 * b);
    hs[nb + b] = src<F>(target + 4 * b);
  }
  HCHECK(hipMalloc(&out, 8 * sizeof(F) * nb));
  HCHECK(hipMalloc(&dsrc, sizeof(Src<F>) * 2 * nb));
  HCHECK(hipMemcpy(dsrc, hs.data(), sizeof(Src<F>) * 2 * nb, hipMemcpyHostToDevice));
  HCHECK(hipMemset(out, 0, 8 * sizeof(F) * nb));
  hipEvent_t e0, e1;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  HCHECK(hipEventRecord(e0, 0));
  const Geo<F> g = geo<F>(d);
  for (int m = 0; m < nsteps; m += seg) {
    const int c = nsteps - m < seg ? nsteps - m : seg;
    hipLaunchKernelGGL((k_heat<F, GRAD>), dim3((n + 255) / 256, nb), dim3(256), 0, 0, g, dsrc, dsrc + nb,
                       step0 + m, c, out);
    HCHECK(hipGetLastError());
  }
  HCHECK(hipEventRecord(e1, 0));
  std::vector<F> h(8 * (size_t)nb);
  HCHECK(hipMemcpy(h.data(), out, sizeof(F) * 8 * nb, hipMemcpyDeviceToHost));
  float el = 0;
  HCHECK(hipEventElapsedTime(&el, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HCHECK(hipFree(out));
  HCHECK(hipFree(dsrc));
  for (int b = 0; b < nb; ++b) {
    if (loss) loss[b] = (double)h[8 * b];
    if (GRAD && grad)
      for (int q = 0; q < 4; ++q) grad[4 * b + q] = (double)h[8 * b + 1 + q];
  }
  if (ms) *ms = el;
  return 0;
}

}  // namespace

extern "C" int amcd_heat(const amcafe_heat_desc* d, int nb, const double* theta, const double* target, int step0,
                         int nsteps, int mode, double* loss, double* grad, float* ms, char* err, int errlen) {
  if (mode == 0) return run<float, false>(d, nb, theta, target, step0, nsteps, loss, nullptr, ms, err, errlen);
  if (mode == 1) return run<float, true>(d, nb, theta, target, step0, nsteps, loss, grad, ms, err, errlen);
  return run<double, false>(d, nb, theta, target, step0, nsteps, loss, nullptr, ms, err, errlen);
}
