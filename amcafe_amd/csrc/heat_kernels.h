// heat_kernels.h -- kernels and launch sequence of the differentiable temperature field,
// shared by every entry point of it (each brings its own device buffers).
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>

#include <vector>

#include "../../include/amcafe.h"
#include "gl25.h"

namespace amcafe_heat {


static __constant__ double kHx[AMC_GL25_N] = AMC_GL25_X;
static __constant__ double kHw[AMC_GL25_N] = AMC_GL25_W;

template <typename F>
struct Src {  // everything a voxel-step evaluation needs, in the arithmetic type
  F P, v, eta, sigma;
};

template <typename F>
struct Geo {
  int nx, ny, nz;
  F dx, xs, ys, zs, dt, T0, tL, width, k, alpha, h;
};

// Per-instance constants of the 25 quadrature nodes (the same for every voxel): tau_i,
// v tau_i, 1/D1_i, 1/(4 alpha tau_i), and for the gradient -2 sigma / A_i and 4 sigma / D1_i^2.
template <typename F>
struct NodeK {
  F tau[AMC_GL25_N], vtau[AMC_GL25_N], iD1[AMC_GL25_N], iDz[AMC_GL25_N], k1[AMC_GL25_N], k2[AMC_GL25_N],
      amp[AMC_GL25_N];  // amp = W C / (sqrt(tau) A)
};
template <typename F>
__device__ inline void node_consts(const Geo<F>& g, const Src<F>& s, NodeK<F>* K, int i) {
  const F pi = (F)M_PI;
  const F C = s.eta * s.P / (pi * (g.k / g.alpha) * sqrt((F)4 * pi * g.alpha));
  const F s2 = s.sigma * s.sigma;
  const F tau = g.h * ((F)1 - (F)kHx[i]);
  const F W = g.h * (F)kHw[i];
  const F A = (F)2 * g.alpha * tau + s2;
  const F D1 = (F)4 * g.alpha * tau + (F)2 * s2;
  K->tau[i] = tau;
  K->vtau[i] = s.v * tau;
  K->iD1[i] = (F)1 / D1;
  K->iDz[i] = (F)1 / ((F)4 * g.alpha * tau);
  K->k1[i] = -(F)2 * s.sigma / A;
  K->k2[i] = (F)4 * s.sigma / (D1 * D1);
  K->amp[i] = W * C / (sqrt(tau) * A);
}

template <typename F>
__device__ inline F indicator(const Geo<F>& g, F T) {
  return (F)1 / ((F)1 + exp(-(T - g.tL) / g.width));
}

// Fixed-order block sum (butterfly within each wave, then the waves in order by thread 0).
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

// One thread per voxel and instance (blockIdx.y), steps step0+1 .. step0+nsteps.  The
// Gaussian factors in y and z do not depend on x or t, so each thread keeps
// amp_i exp(-dy^2/D1_i - dz^2/(4 alpha tau_i)) per node (for theta and the target) and
// evaluates only the x factor per step.  Their sum bounds T - T0 from above at every step
// (the x factor is at most 1): with PRUNE, a voxel whose bounds for theta and the target both
// stay below tL - 40 width has indicator values below e^-40 at every step and is skipped
// (each skipped loss term is below 2e-35).  Ahead of the source (x beyond its position at t)
// every node's x offset is at least that distance and D1 is largest at the oldest node, so
// the bound times exp(-(x - x_laser)^2 / D1_max) bounds the step's rise; a step where that
// holds for both parameter sets below the same cut is skipped the same way.  The block's sums go to its own slot of `part`
// (5 values: loss, grad) -- overwritten by the first segment, accumulated by later ones in
// launch order -- so the result does not depend on scheduling.
template <typename F, bool GRAD, bool PRUNE>
__global__ __launch_bounds__(256) void k_heat(Geo<F> g, const Src<F>* S, const Src<F>* Tg, int step0, int nsteps,
                                              F* part, int first, unsigned long long* evals) {
  __shared__ F sh[4];
  __shared__ NodeK<F> Ks, Kt;
  const Src<F> s = S[blockIdx.y], tgt = Tg[blockIdx.y];
  if (threadIdx.x < AMC_GL25_N) {
    node_consts(g, s, &Ks, threadIdx.x);
    node_consts(g, tgt, &Kt, threadIdx.x);
  }
  __syncthreads();
  const int v = blockIdx.x * 256 + threadIdx.x;
  const int n = g.nx * g.ny * g.nz;
  F loss = 0, gr[4] = {0, 0, 0, 0};
  bool live = false;
  unsigned ne = 0;  // voxel-steps evaluated by this thread
  if (v < n) {
    const int k = v / (g.nx * g.ny), r = v - k * g.nx * g.ny, j = r / g.nx, i = r - j * g.nx;
    const F x = ((F)i + (F)0.5) * g.dx, y = ((F)j + (F)0.5) * g.dx, z = ((F)k + (F)0.5) * g.dx;
    const F dy = y - g.ys, dz = z - g.zs;
    const F dy2 = dy * dy, dz2 = dz * dz;
    F as[AMC_GL25_N], at[AMC_GL25_N];
    F bs = 0, bt = 0;
#pragma unroll
    for (int q = 0; q < AMC_GL25_N; ++q) {
      as[q] = Ks.amp[q] * exp(-dy2 * Ks.iD1[q] - dz2 * Ks.iDz[q]);
      at[q] = Kt.amp[q] * exp(-dy2 * Kt.iD1[q] - dz2 * Kt.iDz[q]);
      bs += as[q];
      bt += at[q];
    }
    const F cut = g.tL - (F)40 * g.width - g.T0;
    live = !PRUNE || !(bs < cut && bt < cut);
    if (live) {
      const F x0 = x - g.xs;
      for (int m = step0 + 1; m <= step0 + nsteps; ++m) {
        const F t = (F)m * g.dt;
        const F cs = x0 - s.v * t, ct = x0 - tgt.v * t;  // x - x_laser(t)
        if (PRUNE && cs > (F)0 && ct > (F)0 && bs * exp(-cs * cs * Ks.iD1[0]) < cut &&
            bt * exp(-ct * ct * Kt.iD1[0]) < cut)
          continue;  // node 0 is the oldest (largest tau, largest D1)
        ++ne;
        F sum = 0, sumT = 0, sv = 0, ss = 0;
#pragma unroll
        for (int q = 0; q < AMC_GL25_N; ++q) {
          const F dxs = cs + Ks.vtau[q];  // x - x_s - v (t - tau)
          const F dxt = ct + Kt.vtau[q];
          const F fs = as[q] * exp(-dxs * dxs * Ks.iD1[q]);
          sumT += at[q] * exp(-dxt * dxt * Kt.iD1[q]);
          sum += fs;
          if (GRAD) {
            sv += fs * (dxs * (t - Ks.tau[q]) * ((F)2 * Ks.iD1[q]));
            ss += fs * (Ks.k1[q] + Ks.k2[q] * (dxs * dxs + dy2));
          }
        }
        const F a = indicator(g, g.T0 + sum), b = indicator(g, g.T0 + sumT);
        loss += (a - b) * (a - b);
        if (GRAD) {
          const F lam = (F)2 * (a - b) * a * ((F)1 - a) / g.width;  // dloss/dT
          gr[0] += lam * (sum / s.P);
          gr[1] += lam * sv;
          gr[2] += lam * (sum / s.eta);
          gr[3] += lam * ss;
        }
      }
    }
  }
  if (evals) {
    unsigned c = ne;
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(evals, (unsigned long long)c);
  }
  F* slot = part + 5 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x);
  F t = block_sum(loss, sh);
  if (threadIdx.x == 0) slot[0] = first ? t : slot[0] + t;
  if (!GRAD && threadIdx.x == 0 && first)
    for (int q = 1; q < 5; ++q) slot[q] = 0;
  if (GRAD) {
    for (int q = 0; q < 4; ++q) {
      t = block_sum(gr[q], sh);
      if (threadIdx.x == 0) slot[1 + q] = first ? t : slot[1 + q] + t;
    }
  }
}

// Sum of the block slots of each instance (blockIdx.x) in a fixed order: thread t adds
// slots t, t + 256, ... in order, then the 256 partial sums are added pairwise.
template <typename F>
__global__ __launch_bounds__(256) void k_heat_reduce(const F* part, int nblk, F* out) {
  __shared__ F sh[256];
  const F* p = part + 5 * (size_t)blockIdx.x * nblk;
  for (int q = 0; q < 5; ++q) {
    F acc = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) acc += p[5 * (size_t)b + q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[5 * blockIdx.x + q] = sh[0];
    __syncthreads();
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

// Device buffers of one call: source parameters (theta then target, nb each), the block
// slots (5 values per block and instance), the per-instance sums and the evaluation counter.
template <typename F>
struct Bufs {
  Src<F>* src;
  F* part;
  F* out;
  unsigned long long* evals;
};

static int n_blocks(const amcafe_heat_desc* d) { return (d->nX[0] * d->nX[1] * d->nX[2] + 255) / 256; }

// Segments of at most `checkpoint` steps, one launch each (all instances of the batch in it),
// then the fixed-order reduction; everything on `stream`.
template <typename F, bool GRAD, bool PRUNE>
static void launch_all(const amcafe_heat_desc* d, int nb, const Bufs<F>& B, int step0, int nsteps,
                       hipStream_t stream) {
  const int seg = d->checkpoint > 0 ? d->checkpoint : 100;
  const int nblk = n_blocks(d);
  const Geo<F> g = geo<F>(d);
  if (nsteps <= 0) {  // no steps: every slot holds zero
    hipLaunchKernelGGL((k_heat<F, GRAD, PRUNE>), dim3(nblk, nb), dim3(256), 0, stream, g, B.src, B.src + nb, step0,
                       0, B.part, 1, B.evals);
  }
  for (int m = 0; m < nsteps; m += seg) {
    const int c = nsteps - m < seg ? nsteps - m : seg;
    hipLaunchKernelGGL((k_heat<F, GRAD, PRUNE>), dim3(nblk, nb), dim3(256), 0, stream, g, B.src, B.src + nb,
                       step0 + m, c, B.part, m == 0 ? 1 : 0, B.evals);
  }
  hipLaunchKernelGGL((k_heat_reduce<F>), dim3(nb), dim3(256), 0, stream, B.part, nblk, B.out);
}

template <typename F>
static void sources(int nb, const double* theta, const double* target, std::vector<Src<F>>& hs) {
  hs.resize(2 * (size_t)nb);
  for (int b = 0; b < nb; ++b) {
    hs[b] = src<F>(theta + 4 * b);
    hs[nb + b] = src<F>(target + 4 * b);
  }
}

template <typename F>
static void results(int nb, const std::vector<F>& h, bool grad_on, double* loss, double* grad) {
  for (int b = 0; b < nb; ++b) {
    if (loss) loss[b] = (double)h[5 * b];
    if (grad_on && grad)
      for (int q = 0; q < 4; ++q) grad[4 * b + q] = (double)h[5 * b + 1 + q];
  }
}

}  // namespace amcafe_heat
