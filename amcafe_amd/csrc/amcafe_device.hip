// amcafe_device.hip -- HIP kernels of the CAFE time step for CDNA4 (gfx950) and
// the internal launcher API declared in amcafe_dev.h.
//
// One step of one or many independent domains ("instances") runs as:
//   k_temperature      TempField::AnalyticTempCurr (S/TempField.C:139-217), optionally
//                      fused with SetLiquid3 (S/VoxelsCA.C:1098-1116); k_temp_commit
//                      records the laser plane the field now matches.  (k_easm: the
//                      EASM temperature, S/TempField.C:79-137, when asked for.)
//   k_setliquid        SetLiquid3 alone, when the temperature was a separate call
//   k_solid_to_mushy   ConvertSolid1(1) (S/VoxelsCA.C:780-797)
//   k_active_flags     active-set membership (S/VoxelsCA.C:262-273) + per-block counts
//   k_scan_chunks,     two-level exclusive scan of the block counts
//   k_scan
//   k_compact          ordered active list, voxel->active map, gathers, ruq draws
//                      (S/VoxelsCA.C:274-303, 336-380); k_loop_rng seeds the loop's engine
//   k_active_setup     active-neighbour table (S/VoxelsCA.C:304-335) and growing list;
//   k_grow_lists       the growing cells' sorted candidate lists (the per-pair part of
//                      S/VoxelsCA.C:386-445) with each entry's capture re-centring
//   k_event_loop       the serial capture/nucleation loop (S/VoxelsCA.C:381-610), one
//                      workgroup per instance, FP64 throughout
//   k_scatter          active arrays back to the grid (S/VoxelsCA.C:633-643)
//   k_mushy_to_solid   ConvertSolid1(0) over the active cells (S/VoxelsCA.C:739-758)
// Per-voxel passes take runs of 16 voxels per thread (one 16-byte state load).  Layer
// kernels: k_voronoi, k_site_rank, k_write_layer (base plate and powder layers,
// S/BasePlate.C, S/VoxelsCA.C AddLayer1/CleanLayer); k_grain_presence, k_remap (grain
// renumbering).
// Grid-wide kernels take a flattened block list over all instances, so a batch of
// domains of different sizes is one launch per stage.
//
// Must be compiled with -ffp-contract=off: every floating-point expression that
// restates the reference is evaluated with its own rounding, in its own order.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

#include <atomic>

#include "amcafe_dev.h"
#include "ddmath.h"
#include "gl25.h"

// Phase timers of the capture loop (core-clock cycles summed over events), compiled in
// only with -DAMCD_PHASE_PROFILE; read with amcd_phase_profile().
#ifdef AMCD_PHASE_PROFILE
__device__ unsigned long long g_phase[32];
#define PHASE_INIT(v) unsigned long long v = clock64()
#define PHASE(v, slot, pred)                                \
  do {                                                      \
    unsigned long long t_ = clock64();                      \
    if (pred) atomicAdd(&g_phase[slot], t_ - v);            \
    v = t_;                                                 \
  } while (0)
// barrier arrival of wave w (lane 0) with a flag word of what the wave did this event
#define PHASE_ARRIVE(sh, par, w, flag)         \
  do {                                         \
    int wf_ = 0;                               \
    for (int b_ = 0; b_ < 6; ++b_)             \
      if (__ballot(((flag) >> b_) & 1)) wf_ |= 1 << b_; \
    if ((threadIdx.x & 63) == 0) {             \
      (sh).arr[par][w] = clock64();            \
      (sh).aflag[par][w] = wf_;                \
    }                                          \
  } while (0)
// after the barrier: which wave arrived last, by how much, and what it had done
#define PHASE_LAST(sh, par, NW)                                              \
  do {                                                                       \
    if (threadIdx.x == 0) {                                                  \
      int wl = 0;                                                            \
      unsigned long long mx = 0, mn = ~0ull, m2 = 0;                         \
      for (int q = 0; q < (NW); ++q) {                                       \
        const unsigned long long a = (sh).arr[par][q];                       \
        if (a > mx) { m2 = mx; mx = a; wl = q; } else if (a > m2) m2 = a;    \
        if (a < mn) mn = a;                                                  \
      }                                                                      \
      atomicAdd(&g_phase[16 + wl], 1ull);                                    \
      atomicAdd(&g_phase[24], mx - mn);                                      \
      atomicAdd(&g_phase[25], mx - m2);                                      \
      const int fl = (sh).aflag[par][wl];                                    \
      for (int b = 0; b < 6; ++b)                                            \
        if ((fl >> b) & 1) atomicAdd(&g_phase[26 + b], 1ull);                \
    }                                                                        \
  } while (0)
#else
#define PHASE_INIT(v)
#define PHASE(v, slot, pred)
#define PHASE_ARRIVE(sh, par, w, flag)
#define PHASE_LAST(sh, par, NW)
#endif
// a per-lane word of what a wave did in the event, for the arrival record (profile builds only)
#ifdef AMCD_PHASE_PROFILE
#define PHASE_FLAGS(v) int v = 0
#define PHASE_FLAG(v, bits) (v) |= (bits)
#define PHASE_FLAGS_CLEAR(v) (v) = 0
#else
#define PHASE_FLAGS(v)
#define PHASE_FLAG(v, bits)
#define PHASE_FLAGS_CLEAR(v)
#endif

// A pointer known to address global memory.  The capture loop reads its array pointers
// from a descriptor copied to LDS, so the compiler cannot tell their address space and
// would emit flat accesses, which also count against the LDS wait counter: every LDS
// wait would then wait for the outstanding memory loads, and for the stores to be
// acknowledged, as well.  A flat store may also alias the descriptor in LDS, so each one
// makes the compiler read the next array pointer from LDS again.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}
// 8-byte vector records (int2, float2) read and written through a global pointer as one
// 64-bit access (HIP's vector types have no assignment in that address space)
template <class V>
__device__ __forceinline__ V gload8(const V* p) {
  static_assert(sizeof(V) == 8, "8-byte records only");
  return __builtin_bit_cast(V, *gmem((const unsigned long long*)p));
}
template <class V>
__device__ __forceinline__ void gstore8(V* p, V v) {
  static_assert(sizeof(V) == 8, "8-byte records only");
  *gmem((unsigned long long*)p) = __builtin_bit_cast(unsigned long long, v);
}

// lane l's value of v (l must be the same in every lane)
__device__ static inline double lane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

#define ZB 1024     // voxels per compaction block (one wave, 16 per thread)
#define GB 256      // threads per grid-wide block
#define SB 4096     // voxels per solid->mushy block (GB threads, 16 voxels each)

// -------------------------------------------------------------------------------------
// device structures
struct LoopState {
  int nslots;        // growing-list length
  int vSind2;        // next scan position (S/VoxelsCA.C:346-353, 487-488, 520-521)
  int nGrain;
  int pool_used;
  unsigned int rng;  // minstd_rand0 state
  int nslots_pre;    // slots whose lists the step's set-up built (-1 before the loop runs)
  long long draws;
  double tinc;
  int events, captures, nucs, done;
  int pending;       // 1: a nucleation was decided but the pool was empty
  int pend_tgt;      // active index of the voxel to nucleate
  int Na, nmushy0;
  int zhot0, zhot1;  // z planes the last temperature evaluation could have melted
  // T holds the analytic field of a laser plane at lz_prev (tvalid = 1): every voxel is xi
  // where its hot flag (vs bit 7) is set, else T0 above lz_prev and tS below it
  int tvalid, pad1;
  double lz_prev;
};

struct Inst {
  amcd_params p;
  int N, nxy, nblkZ, nblkG, nblkS, nblkT;
  // grid fields
  uint8_t* vs;
  int* gid;
  double* ext;
  double *cx, *cy, *cz;
  double* T;
  uint8_t* bj;
  uint8_t* build;  // VoxelsCA::Build (changes only where the driver calls SetBuild)
  int* amap;
  unsigned* zliq;  // per z plane: bit 0 = had liquid at the last active-set pass (or unknown)
  int *bcA, *bcM;  // per-ZB-block counts -> exclusive offsets within chunks of 1024 blocks
  int *chA, *chM;  // per chunk: totals -> exclusive offsets
  // active list
  int acap;
  int* aVox;
  uint8_t* aVS;
  int* aG;
  double *aT, *aExt, *aCx, *aCy, *aCz, *ruq;
  int* aNbr;
  double* aRN;     // ruqs .5*(ruq[a] + ruq[aNbr[a][q]]) of every active cell a and neighbour q
  uint8_t* aCold;
  int* aScan;
  // growing list + candidates (26 slots per growing cell)
  int *gAct, *gScan;
  double *gExt, *gVhat;
  double *gR, *gC;  // per slot: the cell's grain rotation (9) and centroid (3)
  int *gV, *gG;     // per slot: the cell's voxel and grain
  float2* gLH;      // per slot: capture-clock bounds (lo, hi) of the memory-resident tier
  int2* gHd;        // per slot: list head (entry, target) of the memory-resident tier
  uint8_t *gN, *gDead, *gGrow;
  unsigned* gdead;  // liveness bitmap of the capture loop in memory (one bit per active cell),
                    // for interfaces larger than the LDS bitmap holds
  double *cD, *cF, *cR;
  double4* cRC;     // per entry of a set-up list: the capture's re-centred centroid and extent
  int *cTgt, *cV;   // per entry: target (active index) and its voxel
  uint8_t* cK;
  // grains
  int gcap;
  double* rot;
  int poolcap;
  double* pool;
  LoopState* ls;
};

// -------------------------------------------------------------------------------------
// minstd_rand0 (x <- 16807 x mod 2^31-1) and libstdc++'s generate_canonical<double,53>
// (/usr/include/c++/11/bits/random.tcc:3345-3381): two engine outputs per double,
//   sum = (x1-1) + (x2-1)*R,  R = 2^31-2,  result = sum / RN(R*R),  clamp below 1.
#define MINSTD_M 2147483647u
__host__ __device__ static inline unsigned int mod_m(unsigned long long x) {
  unsigned long long r = (x & MINSTD_M) + (x >> 31);
  r = (r & MINSTD_M) + (r >> 31);
  return (unsigned int)(r >= MINSTD_M ? r - MINSTD_M : r);
}
__host__ __device__ static inline unsigned int lcg_next(unsigned int x) { return mod_m((unsigned long long)x * 16807u); }
__host__ __device__ static inline unsigned int lcg_pow(unsigned long long e) {  // 16807^e mod m
  unsigned long long b = 16807u, r = 1;
  while (e) {
    if (e & 1) r = mod_m(r * b);
    b = mod_m(b * b);
    e >>= 1;
  }
  return (unsigned int)r;
}
// two steps back: 16807^-2 mod (2^31-1) = 1407677000^2 mod m (16807 * 1407677000 = 1 mod m)
__host__ __device__ static inline unsigned int rng_back(unsigned int x) {
  const unsigned long long inv = 1407677000ull;
  return mod_m(mod_m((unsigned long long)x * inv) * inv);
}
__host__ __device__ static inline unsigned int lcg_seed(unsigned long long s) {
  unsigned int x = (unsigned int)(s % MINSTD_M);
  return x == 0 ? 1u : x;
}
__host__ __device__ static inline double canonical(unsigned int x1, unsigned int x2) {
  double sum = (double)(x1 - 1u);
  sum = sum + (double)(x2 - 1u) * 2147483646.0;
  double ret = sum / 4611686009837453312.0;  // RN((2^31-2)^2)
  if (ret >= 1.0) ret = 0x1.fffffffffffffp-1;
  return ret;
}

// -------------------------------------------------------------------------------------
__device__ static inline int find_inst(const int* start, int n, int b) {
  int lo = 0, hi = n;  // start[lo] <= b < start[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (start[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}
__device__ static inline void decode(const Inst& d, int v, int& i, int& j, int& k) {
  k = v / d.nxy;
  int r = v - k * d.nxy;
  j = r / d.p.nx;
  i = r - j * d.p.nx;
}

// -------------------------------------------------------------------------------------
// TempField::AnalyticTempCurr, S/TempField.C:155-197, with the optional SetLiquid3.
// T takes one of three values, T0 (above the laser plane), xi = tL(1+eps) (inside the melt
// pool) or tS.  Two shortcuts keep the result bit-identical while touching little memory:
//  * a voxel whose |offset| along some semi-axis exceeds that semi-axis by a relative 2^-40
//    is outside the ellipsoid without the divisions (RN is monotone, so the rounded
//    quotient is >= 1 and so is the rounded sum of squares);
//  * while tvalid holds, the previous T of a voxel is known from its hot flag and the
//    previous laser plane, so T, the flag byte and BuildjID are stored only where they change;
//  * a run of voxels whose row segment lies farther from the laser than the melt pool's
//    extent plus two voxels (in the unrotated frame: the rotation about the plate centre
//    preserves distances up to rounding far below that margin) cannot reach tL, so in the
//    planes the pool can reach it changes nothing unless one of its voxels was hot.
#define TV 16  // voxels per thread (their state bytes in 16-byte loads)
__global__ __launch_bounds__(GB) void k_temperature(Inst* I, const int* bstart, int ninst, const amcd_step* S,
                                                    amcd_out* O, int setliquid) {
  int ii = find_inst(bstart, ninst, blockIdx.x);
  // The descriptor and the step record are copied: read through references, every field would
  // be read again from memory after each store of the voxel loop (a store could alias them).
  const Inst d = I[ii];
  const amcd_step st = S[ii];
  if (st.skip) return;
  const int v0 = ((blockIdx.x - bstart[ii]) * GB + threadIdx.x) * TV;
  const amcd_params& p = d.p;
  if (v0 == 0) {  // planes with (k+.5)dz in (lz - max depth, lz] are the only ones T can reach tL in
    const double depth = fmax(p.a1[2], p.a1[5]);
    d.ls->zhot0 = max(0, (int)floor((st.lz - depth) / p.dz - 0.5) - 1);
    d.ls->zhot1 = min(p.nz - 1, (int)floor(st.lz / p.dz - 0.5) + 1);
  }
  bool anyhot = false;
  // the three temperatures as values (selected through their struct fields, the compiler
  // would keep the struct in scratch memory and load through the selected address)
  const double T0 = p.T0, tS = p.tS, tL = p.tL, xi = p.xi;
  if (v0 < d.N) {
    const bool valid = d.ls->tvalid != 0 && tS < tL && T0 < tL;
    const double lzp = d.ls->lz_prev;
    const double e1 = 1 + 0x1p-40;
    const double lim1 = p.a1[1] * e1, lim2 = p.a1[2] * e1, lim3 = p.a1[3] * e1, lim4 = p.a1[4] * e1,
                 lim5 = p.a1[5] * e1;
    uint8_t sv[TV];
    if (v0 + TV <= d.N) {
      unsigned ws[TV / 4];
#pragma unroll
      for (int c = 0; c < TV / 16; ++c) {  // 16-byte loads of the run's state bytes
        const uint4 w = *reinterpret_cast<const uint4*>(d.vs + v0 + 16 * c);
        ws[4 * c] = w.x;
        ws[4 * c + 1] = w.y;
        ws[4 * c + 2] = w.z;
        ws[4 * c + 3] = w.w;
      }
#pragma unroll
      for (int q = 0; q < TV; ++q) sv[q] = (uint8_t)(ws[q >> 2] >> (8 * (q & 3)));
    } else {
#pragma unroll
      for (int q = 0; q < TV; ++q) sv[q] = v0 + q < d.N ? d.vs[v0 + q] : 0;
    }
    // planes outside (lz - max depth, lz] (with a plane of margin) cannot reach tL: the
    // ellipsoid tests need |zp - lz| < depth, so T is T0 or tS there without geometry
    const double depth = fmax(p.a1[2], p.a1[5]);
    const int kh0 = (int)floor((st.lz - depth) / p.dz - 0.5) - 1, kh1 = (int)floor(st.lz / p.dz - 0.5) + 1;
    int i, j, k;
    decode(d, v0, i, j, k);
    // Nothing changes for a run of voxels in one plane outside [kh0, kh1] when none of them
    // is hot (and no state byte carries bits the pass clears) and the plane lies on the
    // same side of the previous and the current laser plane: T stays T0 or tS as before,
    // and no state, grain or melt flag is written (below the laser plane the flag is
    // cleared when T < T0, so tS < T0 rules the shortcut out there).
    bool quiet = false;
    if (valid && v0 + TV <= d.N && (v0 + TV - 1) / d.nxy == k) {
      unsigned busy = 0;
#pragma unroll
      for (int q = 0; q < TV; ++q) busy |= sv[q] & 0xbc;
      const double zp = ((double)k + .5) * p.dz;
      bool cold = k < kh0 || k > kh1 || zp > st.lz;  // no geometry reaches these voxels
      if (!cold && i + TV <= p.nx) {  // a run inside one row: its distance to the laser
        const double Lx = p.hx + st.cg * (st.lx - p.hx) - st.sg * (st.ly - p.hy);
        const double Ly = p.hy + st.sg * (st.lx - p.hx) + st.cg * (st.ly - p.hy);
        const double xa = ((double)i + .5) * p.dx, xb = ((double)(i + TV - 1) + .5) * p.dx;
        const double yc = ((double)j + .5) * p.dy;
        const double ex = fmax(0.0, fmax(xa - Lx, Lx - xb)), ey = yc - Ly;
        const double A = fmax(p.bmDX0, p.a1[3]), B = fmax(p.a1[1], p.a1[4]);
        const double R = sqrt(A * A + B * B) + 2.0 * fmax(p.dx, p.dy);
        cold = ex * ex + ey * ey > R * R;
      }
      quiet = busy == 0 && cold && (zp > st.lz) == (zp > lzp) && ((zp > st.lz) || !(tS < T0));
    }
    if (!quiet)
#pragma unroll
    for (int q = 0; q < TV; ++q) {
      const int v = v0 + q;
      if (v >= d.N) break;
      const double x0 = ((double)i + .5) * p.dx;
      const double y0 = ((double)j + .5) * p.dy;
      const double zp = ((double)k + .5) * p.dz;
      double T = T0;
      if (k < kh0 || k > kh1) {
        if (!(zp > st.lz)) T = tS;
      } else if (!(zp > st.lz)) {
        const double xp = st.cg * (x0 - p.hx) + st.sg * (y0 - p.hy) + p.hx;
        const double yp = (-st.sg) * (x0 - p.hx) + st.cg * (y0 - p.hy) + p.hy;
        const double r0 = xp - st.lx, r1 = yp - st.ly, r2 = zp - st.lz;
        T = tS;
        if (st.dirp * r0 > 0) {
          if (fabs(r0) < p.bmDX0 && !(fabs(r1) > lim1) && !(fabs(r2) > lim2)) {
            const double q1 = r1 / p.a1[1], q2 = r2 / p.a1[2];
            const double dsq = q1 * q1 + q2 * q2;
            if (dsq < 1.0) T = xi;
          }
        } else if (!(fabs(r0) > lim3) && !(fabs(r1) > lim4) && !(fabs(r2) > lim5)) {
          const double q0 = r0 / p.a1[3], q1 = r1 / p.a1[4], q2 = r2 / p.a1[5];
          const double dsq = q0 * q0 + q1 * q1 + q2 * q2;
          if (dsq < 1.0) T = xi;
        }
      }
      const uint8_t s0 = sv[q];
      const bool hot = T >= tL;
      const double Told = (s0 & 0x80) ? xi : (zp > lzp ? T0 : tS);
      if (!valid || T != Told) d.T[v] = T;
      if (!(zp > st.lz)) {
        if (T > tL) {
          if (d.bj[v] != 1) d.bj[v] = 1;
        } else if (T < T0) {
          d.bj[v] = 0;
        }
      }
      uint8_t s = s0 & 3;
      if (setliquid && hot) {
        if (v < d.nxy) {
          s = 2;
        } else if (v < st.ilaserLoc * d.nxy) {
          s = 1;
          d.gid[v] = 0;
        }
        d.ext[v] = 0.0;
      }
      const uint8_t s1 = s | (s0 & 0x40) | (hot ? 0x80 : 0);
      if (s1 != s0) d.vs[v] = s1;
      anyhot = anyhot || hot;
      if (++i == p.nx) {
        i = 0;
        if (++j == p.ny) {
          j = 0;
          ++k;
        }
      }
    }
  }
  if (__any(anyhot) && (threadIdx.x & 63) == 0) atomicOr(&O[ii].any_melt, 1);
}
// After a temperature pass: the field now matches this step's laser plane.
__global__ void k_temp_commit(Inst* I, const amcd_step* S, int valid) {
  const Inst& d = I[blockIdx.x];
  if (threadIdx.x != 0 || S[blockIdx.x].skip) return;
  d.ls->tvalid = valid;
  d.ls->lz_prev = S[blockIdx.x].lz;
}

// -------------------------------------------------------------------------------------
// TempField::EASM (S/TempField.C:79-137): the temperature of a Gaussian moving source with
// 316L properties, Utilities::EASM_Temp_LG (S/Utilities.C:128-140).  Every voxel is an
// independent FP64 fixed-point problem (FP64-VALU bound).
__constant__ double kGLx[AMC_GL25_N] = AMC_GL25_X;
__constant__ double kGLw[AMC_GL25_N] = AMC_GL25_W;

// Utilities::Materials_316L (S/Utilities.C:60-96): cP, k, rho, alpha at temperature T
// (pow(10, -n) written as the literal it folds to)
__device__ static inline void mat316(double T, double& cP, double& k, double& rho, double& alpha) {
  if (T < 1660)
    rho = 8084.2 - 0.42086 * T - 3.8942 * 1e-5 * T * T;
  else if (T >= 1660 && T < 3090)
    rho = 7432.7 + 3.9338 * 1e-2 * T - 1.8007 * 1e-4 * T * T;
  else
    rho = 6124.72;
  cP = T < 1660 ? 458.984 + 0.1328 * T : 769.855;
  if (T < 1660)
    k = 9.248 + 1.571 * 1e-2 * T;
  else if (T >= 1660 && T < 3090)
    k = 12.41 + 3.279 * 1e-3 * T;
  else
    k = 21.6465;
  alpha = k / (rho * cP);
}

// Utilities::Integral (S/Utilities.C:105-124): 300 + the 25-point Gauss-Legendre sum over
// s in [0, t] of Temp_Fun_G (S/Utilities.C:228-249), t = 8e-4 s, sigma 50 um, eta 0.56,
// properties at the iterate T.  Nodes x = t/2 + (t/2) xi, weights (t/2) w, summed in order.
__device__ static inline double easm_integral(const double L[3], double T, const double R[3], double Q, double bmV,
                                              int* evals) {
  const double t = 80.0e-5, sigma = 50.0e-6, eta = 0.56;
  double cP, k, rho, alpha;
  mat316(T, cP, k, rho, alpha);
  const double C = eta * Q / (M_PI * rho * cP * sqrt(4 * M_PI * alpha));
  const double x0 = L[0] - bmV * t;
  const double slp = (t - 0.0) / 2.0, shft = (0.0 + t) / 2.0;
  double sum = 0.0;
  for (int i = 0; i < AMC_GL25_N; ++i) {
    const double x = shft + slp * kGLx[i];
    const double w = kGLw[i] * slp;
    const double Lx = x0 + bmV * x;
    const double dx = R[0] - Lx, dy = R[1] - L[1], dz = fabs(R[2] - L[2]);
    const double f = C * pow(t - x, -0.5) / (2 * alpha * (t - x) + sigma * sigma) *
                     exp(-1 * (dx * dx + dy * dy) / (4 * alpha * (t - x) + 2 * (sigma * sigma)) -
                         dz * dz / (4 * alpha * (t - x)));
    sum += w * f;
  }
  ++*evals;
  return 300 + sum;
}

// Utilities::EASM_Temp_LG (S/Utilities.C:128-140): damped Picard iteration from T
__device__ static inline double easm_temp(const double L[3], const double R[3], double T, double Q, double bmV,
                                          int* evals) {
  double Told = T;
  int n = 0;
  const int nlim = 100;
  const double tol = 10;
  double Tnew = easm_integral(L, T, R, Q, bmV, evals);
  while (fabs(Told - Tnew) > tol && n < nlim) {
    Told = Tnew;
    const double T_temp = easm_integral(L, Told, R, Q, bmV, evals);
    Tnew = 0.5 * (Told + T_temp);
    n++;
  }
  return Tnew;
}

// TempField::EASM's voxel loop (S/TempField.C:90-115): 0 K above the laser plane, 0.99 T0
// ahead of the laser in the top layer, EASM_Temp_LG from 0 K elsewhere; BuildjID as the
// analytic model sets it; the "T >= tL" flag for SetLiquid3; O.any_melt collects whether
// some voxel stayed below tL (the all_of test of S/TempField.C:121-124 is its negation).
__global__ __launch_bounds__(GB) void k_easm(Inst* I, const int* bstart, int ninst, const amcd_step* S, amcd_out* O,
                                             double Q, double layerT, unsigned long long* n_evals) {
  int ii = find_inst(bstart, ninst, blockIdx.x);
  const Inst d = I[ii];  // a copy: stores in the loops cannot alias it
  const amcd_step& st = S[ii];
  if (st.skip) return;
  const int v = (blockIdx.x - bstart[ii]) * GB + threadIdx.x;
  bool cold = false;
  int evals = 0;
  if (v < d.N) {
    int i, j, k;
    decode(d, v, i, j, k);
    const amcd_params& p = d.p;
    const double x0 = ((double)i + .5) * p.dx;
    const double y0 = ((double)j + .5) * p.dy;
    const double zp = ((double)k + .5) * p.dz;
    double T = 0.0;
    if (!(zp > st.lz)) {
      const double xp = st.cg * (x0 - p.hx) + st.sg * (y0 - p.hy) + p.hx;
      const double yp = (-st.sg) * (x0 - p.hx) + st.cg * (y0 - p.hy) + p.hy;
      const double R[3] = {xp, yp, zp}, L[3] = {st.lx, st.ly, st.lz};
      if ((x0 > st.lx) && (zp > (st.lz - layerT)))
        T = .99 * p.T0;
      else
        T = easm_temp(L, R, 0.0, Q, p.bmV, &evals);
      if (T > p.tL) d.bj[v] = 1;
      else if (T < p.T0) d.bj[v] = 0;
    }
    d.T[v] = T;
    const bool hot = T >= p.tL;
    cold = !hot;
    d.vs[v] = (d.vs[v] & 3) | (hot ? 0x80 : 0);
  }
  if (__any(cold) && (threadIdx.x & 63) == 0) atomicOr(&O[ii].any_melt, 1);
  for (int m = 32; m >= 1; m >>= 1) evals += __shfl_xor(evals, m);
  if ((threadIdx.x & 63) == 0 && evals) atomicAdd(n_evals, (unsigned long long)evals);
}

// SetLiquid3 (S/VoxelsCA.C:1098-1116) when the temperature kernel ran without it.
__global__ __launch_bounds__(GB) void k_setliquid(Inst* I, const int* bstart, int ninst, const amcd_step* S) {
  int ii = find_inst(bstart, ninst, blockIdx.x);
  const Inst d = I[ii];  // a copy: stores in the loops cannot alias it
  if (S[ii].skip) return;
  int v = (blockIdx.x - bstart[ii]) * GB + threadIdx.x;
  if (v >= d.N) return;
  uint8_t s = d.vs[v];
  if (!(s & 0x80)) return;
  uint8_t st = s & 3;
  if (v < d.nxy) {
    st = 2;
  } else if (v < S[ii].ilaserLoc * d.nxy) {
    st = 1;
    d.gid[v] = 0;
  }
  d.ext[v] = 0.0;
  d.vs[v] = st | 0x80;
}

// ConvertSolid1(1): solid -> mushy when any Moore neighbour is liquid.  Updating in
// place is equivalent to the reference's collect-then-apply, because the update
// (3 -> 2) never changes the predicate (some neighbour == 1) of any other voxel.
__global__ __launch_bounds__(GB) void k_solid_to_mushy(Inst* I, const int* bstart, int ninst, const amcd_step* S) {
  int ii = find_inst(bstart, ninst, blockIdx.x);
  const Inst d = I[ii];  // a copy: stores in the loops cannot alias it
  if (S[ii].skip) return;
  const int nx = d.p.nx, ny = d.p.ny, nz = d.p.nz;
  const int b0 = (blockIdx.x - bstart[ii]) * SB;
  const int bend = min(d.N, b0 + SB);
  // liquid exists only in planes that held it at the last active-set pass or that the
  // last temperature evaluation could melt: a block with no such plane within one plane
  // of its own leaves at once
  const int h0 = d.ls->zhot0, h1 = d.ls->zhot1;
  const int zb0 = b0 / d.nxy, zb1 = (bend - 1) / d.nxy;
  bool any = false;
  for (int z = max(0, zb0 - 1); z <= min(nz - 1, zb1 + 1); ++z) any = any || (d.zliq[z] & 1u) || (z >= h0 && z <= h1);
  if (!any) return;
  // A solid voxel with a liquid Moore neighbour turns mushy.  Each thread takes 16
  // consecutive voxels; when they share a row it reads the 3x3 rows around them (16 bytes
  // and the two end voxels each): the liquid voxels of each row, widened by one voxel to
  // either side and or-ed over the nine rows, are the solid voxels of the run to mark.
  // Liquid voxels are not changed by the pass, so every test reads settled states, and
  // each thread writes its own voxels only.
  constexpr int PT = SB / GB;
  const int v0 = b0 + (int)threadIdx.x * PT;
  if (v0 >= bend) return;
  int i, j, k;
  decode(d, v0, i, j, k);
  if (nx % PT == 0 && v0 + PT <= d.N) {
    const uint4 c = *reinterpret_cast<const uint4*>(d.vs + v0);
    const unsigned cw[4] = {c.x, c.y, c.z, c.w};
    unsigned sol = 0;
#pragma unroll
    for (int q = 0; q < PT; ++q) sol |= (((cw[q >> 2] >> (8 * (q & 3))) & 3u) == 3u ? 1u : 0u) << q;
    if (!sol) return;
    unsigned near = 0;
    for (int dz = -1; dz <= 1; ++dz) {
      const int z = k + dz;
      if (z < 0 || z >= nz) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int y = j + dy;
        if (y < 0 || y >= ny) continue;
        const uint8_t* row = d.vs + (size_t)z * d.nxy + (size_t)y * nx;
        const uint4 w = *reinterpret_cast<const uint4*>(row + i);
        const unsigned ww[4] = {w.x, w.y, w.z, w.w};
        unsigned liq = 0;  // bit q + 1: voxel i + q of the row is liquid
#pragma unroll
        for (int q = 0; q < PT; ++q) liq |= (((ww[q >> 2] >> (8 * (q & 3))) & 3u) == 1u ? 1u : 0u) << (q + 1);
        if (i > 0 && (row[i - 1] & 3) == 1) liq |= 1u;
        if (i + PT < nx && (row[i + PT] & 3) == 1) liq |= 1u << (PT + 1);
        near |= liq | (liq >> 1) | (liq >> 2);
      }
    }
    const unsigned m = sol & near & ((1u << PT) - 1);
    if (!m) return;
    unsigned nw[4] = {cw[0], cw[1], cw[2], cw[3]};
#pragma unroll
    for (int q = 0; q < PT; ++q)
      if ((m >> q) & 1u) nw[q >> 2] = (nw[q >> 2] & ~(3u << (8 * (q & 3)))) | (2u << (8 * (q & 3)));
    *reinterpret_cast<uint4*>(d.vs + v0) = make_uint4(nw[0], nw[1], nw[2], nw[3]);
    return;
  }
  // rows that are not a multiple of 16 voxels, and the end of the domain: voxel by voxel
  for (int q = 0; q < PT; ++q) {
    const int v = v0 + q;
    if (v >= bend) break;
    const uint8_t sv = d.vs[v];
    if ((sv & 3) == 3) {
      bool hit = false;
      for (int dz = -1; dz <= 1 && !hit; ++dz) {
        const int z = k + dz;
        if (z < 0 || z >= nz) continue;
        for (int dy = -1; dy <= 1 && !hit; ++dy) {
          const int y = j + dy;
          if (y < 0 || y >= ny) continue;
          const uint8_t* row = d.vs + (size_t)z * d.nxy + (size_t)y * nx;
          for (int dx = -1; dx <= 1; ++dx) {
            const int x = i + dx;
            if (x >= 0 && x < nx && (row[x] & 3) == 1) hit = true;
          }
        }
      }
      if (hit) d.vs[v] = (sv & ~3) | 2;
    }
    if (++i == nx) {
      i = 0;
      if (++j == ny) {
        j = 0;
        ++k;
      }
    }
  }
}

// The per-plane liquid flags are rebuilt by every active-set pass.
__global__ void k_zliq_clear(Inst* I, const amcd_step* S) {
  const Inst& d = I[blockIdx.x];
  if (S[blockIdx.x].skip) return;
  for (int z = threadIdx.x; z < d.p.nz; z += blockDim.x) d.zliq[z] = 0u;
}
// Active set: {mushy} U {liquid below tL} U {neighbours at or above tL of those}.
// Writes bit 6 of vs and the per-block (ZB voxels) counts of active and mushy cells.
// A hot voxel that is neither mushy nor cool liquid is active when a Moore neighbour
// is mushy or cool liquid.
__device__ static bool hot_voxel_active(const Inst& d, int v) {
  int i, j, k;
  decode(d, v, i, j, k);
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = k + dz;
    if (z < 0 || z >= d.p.nz) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = j + dy;
      if (y < 0 || y >= d.p.ny) continue;
      const uint8_t* row = d.vs + (size_t)z * d.nxy + (size_t)y * d.p.nx;
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = i + dx;
        if (x < 0 || x >= d.p.nx || (dx == 0 && dy == 0 && dz == 0)) continue;
        const uint8_t sn = row[x];
        const int stn = sn & 3;
        if (stn == 2 || (stn == 1 && !(sn & 0x80))) return true;
      }
    }
  }
  return false;
}
// One wave per block of ZB voxels (four per workgroup), 16 consecutive voxels per thread,
// four state bytes at a time: per byte, bit 0 of lo/hi/hb is state bit 0, state bit 1
// and the hot flag.
__global__ __launch_bounds__(256) void k_active_flags(Inst* I, const int* bstart, int ninst, const amcd_step* S) {
  constexpr int PT = ZB / 64;
  int ii = find_inst(bstart, ninst, blockIdx.x);
  const Inst d = I[ii];  // a copy: stores in the loops cannot alias it
  if (S[ii].skip) return;
  const int blk = (blockIdx.x - bstart[ii]) * 4 + (int)(threadIdx.x >> 6);
  if (blk >= d.nblkZ) return;  // whole waves leave: nothing below synchronises the block
  const int lane = threadIdx.x & 63;
  const int v0 = blk * ZB + lane * PT;
  int na = 0, nm = 0;
  unsigned liq = 0;  // bit q: voxel v0 + q is liquid
  if (v0 < d.N) {
    const int nq = min(PT, d.N - v0);
    unsigned w[4] = {0, 0, 0, 0};
    if (nq == PT) {
      const uint4 x = *reinterpret_cast<const uint4*>(d.vs + v0);
      w[0] = x.x;
      w[1] = x.y;
      w[2] = x.z;
      w[3] = x.w;
    } else {
      for (int q = 0; q < nq; ++q) w[q >> 2] |= (unsigned)d.vs[v0 + q] << (8 * (q & 3));
    }
    const unsigned valid = nq == PT ? 0xffffu : (1u << nq) - 1;
    // hot voxels that are neither mushy nor cool liquid ask their Moore neighbours; when
    // the 16 voxels share a row, the 3x3 rows around them are read whole: their mushy or
    // cool liquid voxels, widened by one voxel either way and or-ed over the nine rows
    unsigned hotpend = 0;  // bit q: voxel v0 + q is hot and not active by itself
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const unsigned x = w[r];
      const unsigned lo = x & 0x01010101u, hi = (x >> 1) & 0x01010101u, hb = (x >> 7) & 0x01010101u;
      const unsigned self = (hi & ~lo) | (lo & ~hi & ~hb);
      const unsigned pend = hb & ~self;
#pragma unroll
      for (int q = 0; q < 4; ++q) hotpend |= ((pend >> (8 * q)) & 1u) << (4 * r + q);
    }
    hotpend &= valid;
    unsigned nearA = 0;  // bit q: a Moore neighbour of voxel v0 + q is mushy or cool liquid
    if (hotpend) {
      if (d.p.nx % PT == 0 && nq == PT) {
        int i, j, k;
        decode(d, v0, i, j, k);
        for (int dz = -1; dz <= 1; ++dz) {
          const int z = k + dz;
          if (z < 0 || z >= d.p.nz) continue;
          for (int dy = -1; dy <= 1; ++dy) {
            const int y = j + dy;
            if (y < 0 || y >= d.p.ny) continue;
            const uint8_t* row = d.vs + (size_t)z * d.nxy + (size_t)y * d.p.nx;
            const uint4 u = *reinterpret_cast<const uint4*>(row + i);
            const unsigned uw[4] = {u.x, u.y, u.z, u.w};
            unsigned am = 0;  // bit q + 1: voxel i + q of the row is mushy or cool liquid
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const unsigned x = uw[r];
              const unsigned lo = x & 0x01010101u, hi = (x >> 1) & 0x01010101u, hb = (x >> 7) & 0x01010101u;
              const unsigned a4 = (hi & ~lo) | (lo & ~hi & ~hb);
#pragma unroll
              for (int q = 0; q < 4; ++q) am |= ((a4 >> (8 * q)) & 1u) << (4 * r + q + 1);
            }
            auto edge = [](uint8_t sn) { const int st = sn & 3; return st == 2 || (st == 1 && !(sn & 0x80)); };
            if (i > 0 && edge(row[i - 1])) am |= 1u;
            if (i + PT < d.p.nx && edge(row[i + PT])) am |= 1u << (PT + 1);
            nearA |= am | (am >> 1) | (am >> 2);
          }
        }
      } else {
        for (unsigned m = hotpend; m; m &= m - 1) {
          const int q = __builtin_ctz(m);
          if (hot_voxel_active(d, v0 + q)) nearA |= 1u << q;
        }
      }
    }
    unsigned nw[4];
    bool changed = false;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const unsigned x = w[r];
      const unsigned lo = x & 0x01010101u, hi = (x >> 1) & 0x01010101u, hb = (x >> 7) & 0x01010101u;
      unsigned vb = 0, nb = 0;  // bytes of this word inside the domain; hot bytes with an active neighbour
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        vb |= ((valid >> (4 * r + q)) & 1u) << (8 * q);
        nb |= (((hotpend & nearA) >> (4 * r + q)) & 1u) << (8 * q);
      }
      const unsigned mu = hi & ~lo & vb, lq = lo & ~hi & vb;
      const unsigned ac = mu | (lq & ~hb) | nb;
      nw[r] = (x & ~0x40404040u) | (ac << 6);
      changed = changed || nw[r] != x;
      na += __popc(ac);
      nm += __popc(mu);
#pragma unroll
      for (int q = 0; q < 4; ++q) liq |= ((lq >> (8 * q)) & 1u) << (4 * r + q);
    }
    if (changed) {
      if (nq == PT) {
        *reinterpret_cast<uint4*>(d.vs + v0) = make_uint4(nw[0], nw[1], nw[2], nw[3]);
      } else {
        for (int q = 0; q < nq; ++q) d.vs[v0 + q] = (uint8_t)(nw[q >> 2] >> (8 * (q & 3)));
      }
    }
  }
  // planes holding liquid now (for the next step's solid->mushy pass): one atomic per
  // wave and plane, and none once the flag is seen set.  A run of 16 spans at most two
  // planes unless a plane holds fewer than 16 voxels.
  int lp0 = -1, lp1 = -1;
  if (liq) {
    lp0 = (v0 + __builtin_ctz(liq)) / d.nxy;
    lp1 = (v0 + 31 - __builtin_clz(liq)) / d.nxy;
    if (d.nxy < PT)
      for (unsigned m = liq; m; m &= m - 1) {
        const int kz = (v0 + __builtin_ctz(m)) / d.nxy;
        if (!(d.zliq[kz] & 1u)) atomicOr(&d.zliq[kz], 1u);
      }
  }
  for (int pass = 0; pass < 2; ++pass) {
    const int pz = pass ? lp1 : lp0;
    unsigned long long pend = __ballot(pz >= 0);
    while (pend) {
      const int leader = __builtin_ctzll(pend);
      const int zl = __builtin_amdgcn_readlane(pz, leader);
      pend &= ~__ballot(pz == zl);
      if (lane == leader && !(d.zliq[zl] & 1u)) atomicOr(&d.zliq[zl], 1u);
    }
  }
  for (int o = 32; o >= 1; o >>= 1) {
    na += __shfl_xor(na, o);
    nm += __shfl_xor(nm, o);
  }
  if (lane == 0) {
    d.bcA[blk] = na;
    d.bcM[blk] = nm;
  }
}

// Exclusive scan of the per-block counts in two levels: chunks of 1024 blocks scanned
// by one workgroup each (offsets within the chunk, chunk totals), then one workgroup
// per instance over the chunk totals.
__global__ __launch_bounds__(1024) void k_scan_chunks(Inst* I, const int* cstart, int ninst) {
  __shared__ int wa[16], wm[16];
  const int ii = find_inst(cstart, ninst, blockIdx.x);
  const Inst& d = I[ii];
  const int c = blockIdx.x - cstart[ii];
  const int b = c * 1024 + (int)threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int a = 0, m = 0;
  if (b < d.nblkZ) {
    a = d.bcA[b];
    m = d.bcM[b];
  }
  int ia = a, im = m;
  for (int o = 1; o < 64; o <<= 1) {
    const int ta = __shfl_up(ia, o), tm = __shfl_up(im, o);
    if (lane >= o) {
      ia += ta;
      im += tm;
    }
  }
  if (lane == 63) {
    wa[w] = ia;
    wm[w] = im;
  }
  __syncthreads();
  int pa = 0, pm = 0;
  for (int q = 0; q < w; ++q) {
    pa += wa[q];
    pm += wm[q];
  }
  if (b < d.nblkZ) {
    d.bcA[b] = pa + ia - a;
    d.bcM[b] = pm + im - m;
  }
  if (threadIdx.x == 1023) {
    d.chA[c] = pa + ia;
    d.chM[c] = pm + im;
  }
}
__global__ __launch_bounds__(1024) void k_scan(Inst* I, const amcd_step* S, amcd_out* O) {
  __shared__ int sA[1024], sM[1024];
  const Inst& d = I[blockIdx.x];
  const int nc = (d.nblkZ + 1023) / 1024, t = threadIdx.x;
  const int per = (nc + 1023) / 1024;
  const int c0 = t * per, c1 = min(nc, c0 + per);
  int a = 0, m = 0;
  for (int c = c0; c < c1; ++c) {
    a += d.chA[c];
    m += d.chM[c];
  }
  sA[t] = a;
  sM[t] = m;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
    int xa = t >= off ? sA[t - off] : 0, xm = t >= off ? sM[t - off] : 0;
    __syncthreads();
    sA[t] += xa;
    sM[t] += xm;
    __syncthreads();
  }
  int ra = sA[t] - a, rm = sM[t] - m;
  for (int c = c0; c < c1; ++c) {
    const int va = d.chA[c], vm = d.chM[c];
    d.chA[c] = ra;
    d.chM[c] = rm;
    ra += va;
    rm += vm;
  }
  if (t == 1023) {
    int Na = sA[1023], Nm = sM[1023];
    if (S[blockIdx.x].skip) Na = Nm = 0;
    LoopState* ls = d.ls;
    ls->Na = Na;
    ls->nmushy0 = Nm;
    ls->nslots = 0;
    ls->nslots_pre = -1;
    ls->vSind2 = Nm;
    ls->tinc = 0.0;
    ls->events = ls->captures = ls->nucs = 0;
    ls->pool_used = 0;
    ls->pending = 0;
    ls->done = 0;
    ls->nGrain = S[blockIdx.x].nGrain;
    amcd_out& o = O[blockIdx.x];
    o.n_active = Na;
    o.n_mushy0 = Nm;
    if (Na > d.acap) {
      o.overflow = 1;
      o.status = -2;
      ls->done = 1;
    }
    if (S[blockIdx.x].skip) ls->done = 1;
  }
}

// Ordered compaction: the active list in ascending voxel id, the voxel -> active map,
// gathers of the per-cell state, the scan position of mushy cells and the uniform
// draws ruq (S/VoxelsCA.C:380): the engine is seeded with 30*tInd+seed1 and active
// cell a takes engine outputs 2a+1 and 2a+2 (jump-ahead 16807^(2a)).
__global__ __launch_bounds__(256) void k_compact(Inst* I, const int* bstart, int ninst, const amcd_step* S) {
  // one wave per block of ZB voxels (four per workgroup), 16 consecutive voxels per thread
  constexpr int PT = ZB / 64;
  int ii = find_inst(bstart, ninst, blockIdx.x);
  const Inst d = I[ii];  // a copy: stores in the loops cannot alias it
  if (d.ls->done) return;
  const int blk = (blockIdx.x - bstart[ii]) * 4 + (int)(threadIdx.x >> 6);
  if (blk >= d.nblkZ) return;
  // a block without active cells (its global offset equals the next block's) has nothing
  // to write; most of the domain is such blocks
  const int o0 = d.bcA[blk] + d.chA[blk >> 10];
  const int o1 = blk + 1 < d.nblkZ ? d.bcA[blk + 1] + d.chA[(blk + 1) >> 10] : d.ls->Na;
  if (o1 == o0) return;
  const int lane = threadIdx.x & 63;
  const int v0 = blk * ZB + lane * PT;
  unsigned am = 0, mm = 0;  // bit q: voxel v0 + q is active / mushy
  if (v0 < d.N) {
    const int nq = min(PT, d.N - v0);
    if (nq == PT) {
      const uint4 x = *reinterpret_cast<const uint4*>(d.vs + v0);
      const unsigned w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int q = 0; q < PT; ++q) {
        const unsigned b = (w[q >> 2] >> (8 * (q & 3))) & 0xffu;
        am |= ((b >> 6) & 1u) << q;
        mm |= ((b & 3u) == 2u ? 1u : 0u) << q;
      }
    } else {
      for (int q = 0; q < nq; ++q) {
        const unsigned b = d.vs[v0 + q];
        am |= ((b >> 6) & 1u) << q;
        mm |= ((b & 3u) == 2u ? 1u : 0u) << q;
      }
    }
  }
  // exclusive prefix over the wave's lanes of the active and mushy counts (voxel order)
  const int ca = __popc(am), cm = __popc(mm);  // (mushy voxels are active)
  int ia = ca, im = cm;
  for (int o = 1; o < 64; o <<= 1) {
    const int ta = __shfl_up(ia, o), tm = __shfl_up(im, o);
    if (lane >= o) {
      ia += ta;
      im += tm;
    }
  }
  const int ea = ia - ca, em = im - cm;  // this lane's exclusive prefixes
  const int total = __shfl(ia, 63);
  const int m0 = d.bcM[blk] + d.chM[blk >> 10];
  const unsigned int x0 = lcg_seed(30ull * (unsigned long long)S[ii].tInd + d.p.seed1);
  // the block's active voxels, 64 at a time: lane L takes the (base + L)-th of them, found
  // by a binary search over the lanes' prefixes and a walk over that lane's mask
  for (int base = 0; base < total; base += 64) {
    const int idx = base + lane;
    int t = 0;  // the lane holding voxel idx: the last lane whose prefix is <= idx
    for (int step = 32; step >= 1; step >>= 1) {
      const int c = t + step;
      const int ec = __shfl(ea, c < 64 ? c : 63);
      if (c < 64 && ec <= idx) t = c;
    }
    const unsigned mt = (unsigned)__shfl((int)am, t), mmt = (unsigned)__shfl((int)mm, t);
    const int et = __shfl(ea, t), emt = __shfl(em, t);
    if (idx >= total) continue;
    unsigned m = mt;
    for (int r = idx - et; r > 0; --r) m &= m - 1;
    const int q = __builtin_ctz(m);
    const int v = blk * ZB + t * PT + q;
    const int a = o0 + idx;
    const bool mushy = (mmt >> q) & 1u;
    d.aVox[a] = v;
    d.amap[v] = a;
    d.aVS[a] = d.vs[v] & 3;
    d.aG[a] = d.gid[v];
    d.aT[a] = d.T[v];
    d.aExt[a] = d.ext[v];
    d.aCx[a] = d.cx[v];
    d.aCy[a] = d.cy[v];
    d.aCz[a] = d.cz[v];
    d.aScan[a] = mushy ? m0 + emt + __popc(mmt & ((1u << q) - 1)) : -1;
    const unsigned int x1 = mod_m((unsigned long long)x0 * lcg_pow(2ull * a + 1));
    const unsigned int x2 = lcg_next(x1);
    d.ruq[a] = (canonical(x1, x2) - .5) * d.p.dx / 100.;
  }
}

// Voxel centre, S/VoxelsCA.C:409-414.
__device__ static inline void centre(const Inst& d, int v, double* c) {
  int i, j, k;
  decode(d, v, i, j, k);
  c[0] = ((double)i + .5) * d.p.dx;
  c[1] = ((double)j + .5) * d.p.dy;
  c[2] = ((double)k + .5) * d.p.dz;
}

// Capture candidate of growing cell a (centroid c, rotation R) towards liquid voxel
// vt: D = pow(cos th sin ph,.5) * |R^T dnx|_1 and the factor f (S/VoxelsCA.C:409-424).
__device__ static inline void candidate_at(const double* ct, const double* c, const double* R, double* D, double* f) {
  double dnx0 = ct[0] - c[0], dnx1 = ct[1] - c[1], dnx2 = ct[2] - c[2];
  double ff = amc_shape_factor(dnx0, dnx1, dnx2);
  double d0 = fabs(R[0] * dnx0 + R[3] * dnx1 + R[6] * dnx2);
  double d1 = fabs(R[1] * dnx0 + R[4] * dnx1 + R[7] * dnx2);
  double d2 = fabs(R[2] * dnx0 + R[5] * dnx1 + R[8] * dnx2);
  *D = ff * (d0 + d1 + d2);
  *f = ff;
}
__device__ static inline void candidate(const Inst& d, const double* c, const double* R, int vt, double* D,
                                        double* f) {
  double ct[3];
  centre(d, vt, ct);
  double dnx0 = ct[0] - c[0], dnx1 = ct[1] - c[1], dnx2 = ct[2] - c[2];
  double ff = amc_shape_factor(dnx0, dnx1, dnx2);
  double d0 = fabs(R[0] * dnx0 + R[3] * dnx1 + R[6] * dnx2);
  double d1 = fabs(R[1] * dnx0 + R[4] * dnx1 + R[7] * dnx2);
  double d2 = fabs(R[2] * dnx0 + R[5] * dnx1 + R[8] * dnx2);
  *D = ff * (d0 + d1 + d2);
  *f = ff;
}

// vhat = Avel * pow(tL - T, nvel) (S/VoxelsCA.C:405), pow rounded to nearest; the two
// temperatures of the analytic model are precomputed, any other (EASM) is evaluated here.
__device__ static inline double growth_velocity(const amcd_params& p, double T) {
  if (T >= p.tL) return 0.0;
  if (T == p.tS) return p.vhat_tS;
  if (T == p.T0) return p.vhat_T0;
  return p.Avel * amc_pow_cr(p.tL - T, p.nvel);
}

// Active-neighbour table and, for every growing cell (mushy, some liquid active
// neighbour, all active neighbours below tL; S/VoxelsCA.C:396-397), its slot in the
// growing list and its capture candidates.  Slot order is irrelevant: the event loop
// compares candidates by (time, ruqs, scan position, neighbour rank).
__device__ static inline bool cand_less(double Da, double Ra, int ka, double Db, double Rb, int kb);

__global__ __launch_bounds__(GB) void k_active_setup(Inst* I, const int* bstart, int ninst, amcd_out* O) {
  int ii = find_inst(bstart, ninst, blockIdx.x);
  const Inst d = I[ii];  // a copy: stores in the loops cannot alias it
  if (d.ls->done) return;
  int a = (blockIdx.x - bstart[ii]) * GB + threadIdx.x;
  int Na = d.ls->Na;
  bool grow = false;
  if (a < Na) {
    int v = d.aVox[a];
    int i, j, k;
    decode(d, v, i, j, k);
    int* nb = d.aNbr + (size_t)a * AMCD_NB;
    int q = 0;
    bool cold = true, liq = false;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0 && dz == 0) continue;
          int x = i + dx, y = j + dy, z = k + dz;
          int n = -1;
          if (x >= 0 && x < d.p.nx && y >= 0 && y < d.p.ny && z >= 0 && z < d.p.nz)
            n = d.amap[v + dz * d.nxy + dy * d.p.nx + dx];
          d.aRN[(size_t)a * AMCD_NB + q] = n >= 0 ? .5 * (d.ruq[a] + d.ruq[n]) : 0.0;
          nb[q++] = n;
          if (n >= 0) {
            if (!(d.aT[n] < d.p.tL)) cold = false;
            if (d.aVS[n] == 1) liq = true;
          }
        }
    d.aCold[a] = cold;
    grow = cold && liq && d.aVS[a] == 2;
  }
  unsigned long long bal = __ballot(grow);
  if (!bal) return;
  int lane = threadIdx.x & 63;
  int leader = __ffsll((long long)bal) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(&d.ls->nslots, __popcll(bal));
  base = __shfl(base, leader);
  if (!grow) return;
  unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int slot = base + __popcll(bal & lt);
  d.gAct[slot] = a;
  d.gScan[slot] = d.aScan[a];
  d.gExt[slot] = d.aExt[a];
  d.gVhat[slot] = growth_velocity(d.p, d.aT[a]);
  d.gDead[slot] = 0;
  double c[3] = {d.aCx[a], d.aCy[a], d.aCz[a]};
  const int G = d.aG[a];
  if (G < 1 || G > d.gcap) {  // a growing cell must carry a valid grain
    atomicOr(&O[ii].err_state, 1);
    d.gN[slot] = 0;
    d.gDead[slot] = 1;
    return;
  }
  const double* R = d.rot + 9 * (size_t)(G - 1);
  double Rl[9];
  for (int q = 0; q < 9; ++q) {
    Rl[q] = R[q];
    d.gR[9 * (size_t)slot + q] = R[q];
  }
  for (int q = 0; q < 3; ++q) d.gC[3 * (size_t)slot + q] = c[q];
  d.gV[slot] = d.aVox[a];
  d.gG[slot] = G;
}

// Re-centring of the decentred octahedron after a capture, S/VoxelsCA.C:528-605 and
// the helpers S/VoxelsCA.h:56-93.  R is row-major (local -> global), f is the
// capturing candidate's shape factor.  Outputs the new centroid and half-diagonal.
__device__ static inline void project_point_line(const double* A, const double* x0, const double* x1, double* xp) {
  double n[3], snorm, dd, t;
  snorm = sqrt((x1[0] - x0[0]) * (x1[0] - x0[0]) + (x1[1] - x0[1]) * (x1[1] - x0[1]) +
               (x1[2] - x0[2]) * (x1[2] - x0[2]));
  n[0] = (x1[0] - x0[0]) / snorm;
  n[1] = (x1[1] - x0[1]) / snorm;
  n[2] = (x1[2] - x0[2]) / snorm;
  dd = -(n[0] * A[0] + n[1] * A[1] + n[2] * A[2]);
  t = (-dd - n[0] * x0[0] - n[1] * x0[1] - n[2] * x0[2]) /
      (n[0] * (x1[0] - x0[0]) + n[1] * (x1[1] - x0[1]) + n[2] * (x1[2] - x0[2]));
  xp[0] = x0[0] + (x1[0] - x0[0]) * t;
  xp[1] = x0[1] + (x1[1] - x0[1]) * t;
  xp[2] = x0[2] + (x1[2] - x0[2]) * t;
}
__device__ static inline double dist3(const double* p, const double* q) {
  return sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]));
}
__device__ static inline void recentre(const Inst& d, int vcap, int vnew, const double* cen, const double* R, double f,
                                double* ncen, double* next) {
  const amcd_params& p = d.p;
  int jx0, jx1, jx2, jy0, jy1, jy2;
  decode(d, vnew, jx0, jx1, jx2);
  decode(d, vcap, jy0, jy1, jy2);
  double l0 = (jx0 - jy0) * p.dx, l1 = (jx1 - jy1) * p.dy, l2 = (jx2 - jy2) * p.dz;
  double l = sqrt(l0 * l0 + l1 * l1 + l2 * l2);
  double dnx[3], locX[3];
  dnx[0] = ((double)jx0 + .5) * p.dx - cen[0];
  dnx[1] = ((double)jx1 + .5) * p.dy - cen[1];
  dnx[2] = ((double)jx2 + .5) * p.dz - cen[2];
  locX[0] = R[0] * dnx[0] + R[3] * dnx[1] + R[6] * dnx[2];
  locX[1] = R[1] * dnx[0] + R[4] * dnx[1] + R[7] * dnx[2];
  locX[2] = R[2] * dnx[0] + R[5] * dnx[1] + R[8] * dnx[2];
  double dr = fabs(locX[0]) + fabs(locX[1]) + fabs(locX[2]);
  // octahedron vertices +-dr e_i (vertex c is +dr e_c, vertex c+3 is -dr e_c); the octant's
  // three vertices in ascending vertex order, the reference's table of S/VoxelsCA.h:61-76
  // ({0,1,2},{1,2,3},{0,2,4},{2,3,4},{0,1,5},{1,3,5},{0,4,5},{3,4,5} by octant), computed
  // instead of looked up so that no memory access sits in the chain
  const int a0 = signbit(locX[0]) ? 3 : 0, a1 = signbit(locX[1]) ? 4 : 1, a2 = signbit(locX[2]) ? 5 : 2;
  const int vlo = min(min(a0, a1), a2), vhi = max(max(a0, a1), a2);
  const int vsel[3] = {vlo, a0 + a1 + a2 - vlo - vhi, vhi};
  double V[3][3], S0[3][3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    int vi = vsel[q];
    for (int c = 0; c < 3; ++c) {
      double sgn = (vi % 3 == c) ? (vi < 3 ? 1.0 : -1.0) : 0.0;
      S0[q][c] = sgn;
      V[q][c] = sgn * dr;
    }
  }
  double dn[3];
  for (int q = 0; q < 3; ++q)
    dn[q] = (locX[0] - V[q][0]) * (locX[0] - V[q][0]) + (locX[1] - V[q][1]) * (locX[1] - V[q][1]) +
            (locX[2] - V[q][2]) * (locX[2] - V[q][2]);
  int k0 = 0, k2 = 0;
  for (int q = 1; q < 3; ++q) {
    if (dn[q] < dn[k0]) k0 = q;   // std::min_element: first minimum
    if (dn[k2] < dn[q]) k2 = q;   // std::max_element: first maximum
  }
  int k1;
  if (k0 == k2) {
    k0 = 0; k1 = 1; k2 = 2;
  } else {
    k1 = 3 - k0 - k2;
  }
  double xI[3], xJ[3];
  project_point_line(locX, V[k0], V[k1], xI);
  project_point_line(locX, V[k0], V[k2], xJ);
  double d1I = dist3(V[k0], xI), dI2 = dist3(V[k1], xI), d1J = dist3(V[k0], xJ), dJ3 = dist3(V[k2], xJ);
  const double sq3 = 0x1.bb67ae8584caap+0;   // pow(3.0,.5), folded by the reference's compiler
  const double s23 = 0x1.a20bd700c2c3ep-1;   // pow(2.0/3.0,.5)
  double m = sq3 * l;
  double L12 = .5 * ((m < d1I ? m : d1I) + (m < dI2 ? m : dI2));
  double L13 = .5 * ((m < d1J ? m : d1J) + (m < dJ3 ? m : dJ3));
  double Lmud = sq3 * (s23 * (L12 < L13 ? L13 : L12));
  *next = f * Lmud * 1.0;
  dnx[0] = V[k0][0] - Lmud * S0[k0][0];
  dnx[1] = V[k0][1] - Lmud * S0[k0][1];
  dnx[2] = V[k0][2] - Lmud * S0[k0][2];
  locX[0] = R[0] * dnx[0] + R[1] * dnx[1] + R[2] * dnx[2];
  locX[1] = R[3] * dnx[0] + R[4] * dnx[1] + R[5] * dnx[2];
  locX[2] = R[6] * dnx[0] + R[7] * dnx[1] + R[8] * dnx[2];
  ncen[0] = cen[0] + locX[0];
  ncen[1] = cen[1] + locX[1];
  ncen[2] = cen[2] + locX[2];
}

// Candidate lists of the growing cells (S/VoxelsCA.C:396-440): one wave per slot, lane q
// evaluates neighbour q when it is liquid, and the entries are placed by rank in
// (D, ruqs, scan key) order -- the order the capture loop reads the list in.
__device__ static inline void grow_list(const Inst& d, int slot, int lane) {
  const int a = d.gAct[slot];
  const int t = lane < AMCD_NB ? d.aNbr[(size_t)a * AMCD_NB + lane] : -1;
  const bool liq = t >= 0 && d.aVS[t] == 1;
  double D = __builtin_inf(), f = 0.0, r = __builtin_inf();
  int key = 0x7fffffff, tv = -1;
  double4 rc = make_double4(0.0, 0.0, 0.0, 0.0);
  if (liq) {
    double c[3], R[9];
    for (int q = 0; q < 3; ++q) c[q] = d.gC[3 * (size_t)slot + q];
    for (int q = 0; q < 9; ++q) R[q] = d.gR[9 * (size_t)slot + q];
    tv = d.aVox[t];
    candidate(d, c, R, tv, &D, &f);
    r = d.aRN[(size_t)a * AMCD_NB + lane];
    key = d.gScan[slot] * 32 + lane;
    // the re-centring a capture through this entry performs depends on the entry and its
    // cell alone (voxels, centroid, rotation, shape factor): evaluated here, in parallel,
    // instead of on the capture loop's serial path
    double nc[3], ne;
    recentre(d, d.gV[slot], tv, c, R, f, nc, &ne);
    rc = make_double4(nc[0], nc[1], nc[2], ne);
  }
  const unsigned long long bal = __ballot(liq);
  int rank = 0;  // (lane j's entry read with readlane: j is a constant of the unrolled loop)
#pragma unroll
  for (int j = 0; j < AMCD_NB; ++j) {
    const double Dj = lane_f64(D, j), rj = lane_f64(r, j);
    const int kj = __builtin_amdgcn_readlane(key, j);
    if (((bal >> j) & 1) && j != lane && cand_less(Dj, rj, kj, D, r, key)) rank++;
  }
  const size_t cb = (size_t)slot * AMCD_NB;
  if (liq) {
    d.cD[cb + rank] = D;
    d.cF[cb + rank] = f;
    d.cR[cb + rank] = r;
    d.cTgt[cb + rank] = t;
    d.cV[cb + rank] = tv;
    d.cK[cb + rank] = (uint8_t)lane;
    d.cRC[cb + rank] = rc;
  }
  if (lane == 0) d.gN[slot] = (uint8_t)__popcll(bal);
}
// every wave of an instance's blocks takes slots wave, wave + (its waves), ...
__global__ __launch_bounds__(GB) void k_grow_lists(Inst* I, const int* bstart, int ninst) {
  int ii = find_inst(bstart, ninst, blockIdx.x);
  const Inst d = I[ii];  // a copy: stores in the loops cannot alias it
  if (d.ls->done) return;
  const int lane = threadIdx.x & 63;
  const int nw = (bstart[ii + 1] - bstart[ii]) * (GB / 64);
  const int ns = d.ls->nslots;
  for (int slot = ((blockIdx.x - bstart[ii]) * GB + (int)threadIdx.x) >> 6; slot < ns; slot += nw)
    if (!d.gDead[slot]) grow_list(d, slot, lane);
}

// The capture / nucleation loop of one instance (S/VoxelsCA.C:381-610): one workgroup
// of NT threads per instance, in two roles.
//   wave 0      the engine: performs each event (nucleation, or capture with octahedron
//               re-centring, S/VoxelsCA.C:482-605) and builds the sorted candidate list
//               of the newly mushy voxel;
//   waves 1..   the owners: thread 64+j owns growing cells j, j+(NT-64), ... (K of them)
//               and keeps each cell's half-diagonal E, velocity V, the head of its
//               sorted candidate list and float bounds on its capture time in registers.
// An event is EVAL | barrier | DECIDE | APPLY:
//   EVAL    each owner wave finds its exact best candidate (t, ruqs, scan order) among
//           the few cells whose time bounds reach the wave's smallest upper bound; the
//           engine has already placed the best candidate of the cell it created last;
//   DECIDE  every wave reads the per-wave bests and the pre-drawn uniform and reaches
//           the same decision independently, so no second barrier is needed;
//   APPLY   the owners grow their cells by V*Dt, take over the cell created in the
//           previous event and re-read lists whose first live target was just captured,
//           while the engine performs the event and evaluates the new cell.
// The exchange areas are double-buffered by event parity: one barrier per event keeps
// every wave within one event of the others.
struct Best {
  double t, r;
  int key;   // scan position * 32 + neighbour rank: the reference's scan order
  int cand;  // candidate entry (slot * AMCD_NB + position)
  int tgt;   // active index of the captured voxel
};
__device__ static inline Best best_none() {
  Best b;
  b.t = __builtin_inf();
  b.r = __builtin_inf();
  b.key = 0x7fffffff;
  b.cand = -1;
  b.tgt = -1;
  return b;
}
// a strictly before b in (t, ruqs, scan order).  The ordering predicates are written
// without short-circuit branches, and a winner is taken field by field with selects on one
// predicate value, so that every field of the result comes from the same operand whatever
// the compiler does with the comparison chain.
__device__ static inline bool better(const Best& a, const Best& b) {
  const bool tl = a.t < b.t, tg = a.t > b.t, rl = a.r < b.r, rg = a.r > b.r;
  return tl | (!tg & (rl | (!rg & (a.key < b.key))));
}
// Wave-wide reductions by DPP row operations: two quad permutes, the half-row and row
// mirrors leave a row's result in all 16 of its lanes, row_bcast:15 and row_bcast:31
// carry rows 0 and 2 into rows 1 and 3 and then row 1 into rows 2 and 3, so that lane
// 63 ends with the result for the whole wave.  Every lane must be active.
template <int CTRL, int ROWS = 0xf>
__device__ static inline int dpp(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS = 0xf>
__device__ static inline float dpp(float v) {
  return __int_as_float(dpp<CTRL, ROWS>(__float_as_int(v)));
}
template <int CTRL, int ROWS = 0xf>
__device__ static inline double dpp(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)dpp<CTRL, ROWS>((int)b), hi = (unsigned)dpp<CTRL, ROWS>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
constexpr int kQuadX1 = 0xb1, kQuadX2 = 0x4e, kHalfMirror = 0x141, kMirror = 0x140, kBcast15 = 0x142,
              kBcast31 = 0x143;
__device__ static inline float wave_min(float v) {
  v = fminf(v, dpp<kQuadX1>(v));
  v = fminf(v, dpp<kQuadX2>(v));
  v = fminf(v, dpp<kHalfMirror>(v));
  v = fminf(v, dpp<kMirror>(v));
  v = fminf(v, dpp<kBcast15, 0xa>(v));
  v = fminf(v, dpp<kBcast31, 0xc>(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ static inline double wave_min_f64(double v) {
  v = fmin(v, dpp<kQuadX1>(v));
  v = fmin(v, dpp<kQuadX2>(v));
  v = fmin(v, dpp<kHalfMirror>(v));
  v = fmin(v, dpp<kMirror>(v));
  v = fmin(v, dpp<kBcast15, 0xa>(v));
  v = fmin(v, dpp<kBcast31, 0xc>(v));
  return lane_f64(v, 63);
}
template <int CTRL, int ROWS = 0xf>
__device__ static inline void best_step(Best& b) {
  Best o;
  o.t = dpp<CTRL, ROWS>(b.t);
  o.r = dpp<CTRL, ROWS>(b.r);
  o.key = dpp<CTRL, ROWS>(b.key);
  o.cand = dpp<CTRL, ROWS>(b.cand);
  o.tgt = dpp<CTRL, ROWS>(b.tgt);
  const bool bt = better(o, b);
  b.t = bt ? o.t : b.t;
  b.r = bt ? o.r : b.r;
  b.key = bt ? o.key : b.key;
  b.cand = bt ? o.cand : b.cand;
  b.tgt = bt ? o.tgt : b.tgt;
}
__device__ static inline Best lane_best(const Best& b, int l) {
  Best o;
  o.t = lane_f64(b.t, l);
  o.r = lane_f64(b.r, l);
  o.key = __builtin_amdgcn_readlane(b.key, l);
  o.cand = __builtin_amdgcn_readlane(b.cand, l);
  o.tgt = __builtin_amdgcn_readlane(b.tgt, l);
  return o;
}
// best over a wave.  (t, ruqs, key) is a strict total order on candidates -- keys are
// unique -- so the order in which lanes are combined does not change the result.  A
// single candidate, or a single one at the smallest time, is the best without the full
// comparison.
__device__ static inline Best wave_best(Best b) {
  const unsigned long long m = __ballot(b.cand >= 0);
  if (m == 0) return best_none();
  if (__popcll(m) == 1) return lane_best(b, __builtin_ctzll(m));
  const double tmin = wave_min_f64(b.cand >= 0 ? b.t : __builtin_inf());
  const unsigned long long at = __ballot((b.cand >= 0) & (b.t == tmin));
  if (__popcll(at) == 1) return lane_best(b, __builtin_ctzll(at));
  best_step<kQuadX1>(b);
  best_step<kQuadX2>(b);
  best_step<kHalfMirror>(b);
  best_step<kMirror>(b);
  best_step<kBcast15, 0xa>(b);
  best_step<kBcast31, 0xc>(b);
  return lane_best(b, 63);
}
// candidate order inside a growing cell's list: (D, ruqs, key), NaN distances last
// (branch-free for the reason given at better())
__device__ static inline bool cand_less(double Da, double Ra, int ka, double Db, double Rb, int kb) {
  const bool na = Da != Da, nb = Db != Db;
  const bool c1 = na != nb;             // exactly one NaN: the other one is first
  const bool c2 = !na & !nb & (Da != Db);  // both numbers, different
  const bool rne = Ra != Rb;
  const bool tail = (rne & (Ra < Rb)) | (!rne & (ka < kb));
  return (c1 & nb) | (!c1 & c2 & (Da < Db)) | (!c1 & !c2 & tail);
}

// Liveness of candidate targets during the loop: an LDS bitmap over active cells with
// bit set = "not liquid" (filled from the state at launch, then set by the engine for
// every voxel it captures or nucleates).  Every list entry was liquid when the lists
// were built, so a clear bit means live.
#define AMCD_DEAD_BITS (1 << 19)
__device__ static inline bool dead_bit(const unsigned* dead, int t) { return (dead[t >> 5] >> (t & 31)) & 1u; }
__device__ static inline bool live_target(const unsigned* dead, int t, int skip) {
  return t >= 0 && t != skip && !dead_bit(dead, t);
}

// First live entry at or after position `from` of slot s's sorted list, read four
// entries at a time (one round of loads in the common case).  Entries before the
// first live one stay dead: a captured voxel never becomes liquid again within the
// step.  `skip` is a voxel that counts as dead although its bit may not be set yet.
struct SlotHead {
  double D1, Dn, r1;  // head distance, the next entry's distance (+inf: none), head ruqs
  int c1, tgt1, k1;   // head entry, its target (-1: no live entry), its neighbour rank
};
// Positions p .. p+3 of a list and the distance at p+4 (p <= 26), read from one base
// address per array with no branches.  Positions past a slot's 26 entries read the next
// slot, or for the last slot the 256 bytes every device array is allocated beyond its
// length; they lie beyond the list length and are never used.
// The four neighbour ranks stay packed in one word until a head is picked: unpacked at
// the load, they would make the loads of a block that is issued early wait there.
struct SlotBlock {
  double D[5], R[4];
  int t[4], n;
  unsigned kp;  // rank of position p + q in bits 8q .. 8q+7
};
typedef unsigned u32_any_align __attribute__((aligned(1)));
__device__ static inline void block_entries(const Inst& d, size_t cb, int p, SlotBlock& b) {
  const size_t e = cb + p;
  b.kp = *gmem((u32_any_align*)(d.cK + e));  // four ranks, little-endian: one load
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    b.D[q] = gmem(d.cD)[e + q];
    if (q < 4) {
      b.t[q] = gmem(d.cTgt)[e + q];
      b.R[q] = gmem(d.cR)[e + q];
    }
  }
}
// The first live entry among positions p .. p+3 of a loaded block into h: true when
// found or when the list ends before p (h unchanged), false to go on at p + 4.  The four
// liveness words are read unconditionally (an entry past the list end reads word 0), so
// that they are requested together: read under each entry's own condition, every read
// waited for the one before.
__device__ static inline bool block_pick(const SlotBlock& b, const unsigned* dead, size_t cb, int p, int skip,
                                         SlotHead& h) {
  if (p >= b.n) return true;
  int tq[4];
  unsigned wq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    tq[q] = (p + q < b.n && b.t[q] >= 0) ? b.t[q] : -1;
    wq[q] = dead[(tq[q] >= 0 ? tq[q] : 0) >> 5];
  }
  int hit = -1;
#pragma unroll
  for (int q = 3; q >= 0; --q) {
    const bool live = (tq[q] >= 0) & (tq[q] != skip) & (((wq[q] >> (tq[q] & 31)) & 1u) == 0);
    hit = live ? q : hit;
  }
  if (hit < 0) return false;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (q == hit) {
      h.c1 = (int)cb + p + q;
      h.tgt1 = b.t[q];
      h.D1 = b.D[q];
      h.Dn = (p + q + 1 < b.n) ? b.D[q + 1] : __builtin_inf();
      h.r1 = b.R[q];
      h.k1 = (int)((b.kp >> (8 * q)) & 0xffu);
    }
  return true;
}
__device__ static inline SlotHead head_none(size_t cb, int n) {
  SlotHead h;
  h.c1 = (int)cb + n;
  h.tgt1 = -1;
  h.k1 = 0;
  h.D1 = h.Dn = h.r1 = __builtin_inf();
  return h;
}
// First live entry at or after position `from` of slot s.  The entries are read four at
// a time together with the list length, so the common case is one round of loads
// (positions past the length hold stale data and are masked afterwards).
__device__ static inline SlotHead slot_head(const Inst& d, const unsigned* dead, int s, int from, int skip) {
  const size_t cb = (size_t)s * AMCD_NB;
  SlotBlock b;
  b.n = gmem(d.gN)[s];
  // the default head is set after the search, so that the length is not waited for
  // before the entries are requested
  SlotHead h = head_none(cb, 0);
  for (int p = from; p < AMCD_NB; p += 4) {
    block_entries(d, cb, p, b);
    if (block_pick(b, dead, cb, p, skip, h)) break;
  }
  if (h.tgt1 < 0) h.c1 = (int)cb + b.n;
  return h;
}
// slot_head's first block (positions p .. p+3, the distance at p+4 and the list length)
// loaded ahead of its use, and the same search started from it.
__device__ static inline SlotBlock slot_block(const Inst& d, int s, int p) {
  SlotBlock b;
  b.n = gmem(d.gN)[s];
  block_entries(d, (size_t)s * AMCD_NB, p, b);
  return b;
}
__device__ static inline SlotBlock slot_block_none() {
  SlotBlock b;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    b.t[q] = -1;
    b.D[q] = b.R[q] = 0.0;
  }
  b.D[4] = 0.0;
  b.n = 0;
  b.kp = 0;
  return b;
}
// slot_head(d, dead, s, from, skip) with its first block b = slot_block(d, s, from)
// already loaded: the same result, the later blocks read only when the first holds no
// live entry
__device__ static inline SlotHead slot_head_pre(const Inst& d, const unsigned* dead, int s, int from, int skip,
                                                const SlotBlock& b) {
  const size_t cb = (size_t)s * AMCD_NB;
  SlotHead h = head_none(cb, b.n);
  if (from >= AMCD_NB || block_pick(b, dead, cb, from, skip, h)) return h;
  return slot_head(d, dead, s, from + 4, skip);
}

// Float bounds [lo, hi] on tau = T + t, the event clock at which the cell's head would
// be captured, t = (D1 - E)/V.  While the head does not change, every event adds Dt to
// T and V*Dt to E, so tau stays fixed up to rounding; the margin, 2^-26 of the
// magnitudes involved, exceeds that drift for more than 2^20 events (the loop
// re-evaluates every cell before that).  lo = hi = +inf: never a candidate (t NaN or
// +inf); lo = -inf, hi = +inf: evaluated at every event (V <= 0 or t = -inf).
__device__ static inline float f_down(double x) {
  float f = (float)x;
  return ((double)f > x) ? nextafterf(f, -__builtin_inff()) : f;
}
__device__ static inline float f_up(double x) {
  float f = (float)x;
  return ((double)f < x) ? nextafterf(f, __builtin_inff()) : f;
}
__device__ static inline void tau_bounds(double D1, double e, double v, double T, float& lo, float& hi) {
  const double t = (D1 - e) / v;
  if (t != t || t == __builtin_inf()) {
    lo = hi = __builtin_inff();
  } else if (!(v > 0.0) || t == -__builtin_inf()) {
    lo = -__builtin_inff();
    hi = __builtin_inff();
  } else {
    const double m = 0x1p-26 * (fabs(t) + fabs(T) + (fabs(D1) + fabs(e)) / v);
    const double tau = T + t;
    lo = f_down(tau - m);
    hi = f_up(tau + m);
  }
}

struct NewCell {  // a growing cell created by the engine, taken over by its owner
  double E, V, D1, Dn1, r1;  // extent, velocity, head distance, next distance, head ruqs
  int c1, tgt1, key1, slot, valid;
};
struct EngineShared {
  Best sb[2][16];   // per-wave bests, by event parity
  double ubuf[2][64];  // uniforms of draws 64b .. 64b+63 of this launch in ubuf[b & 1]
  int susp[2];      // engine suspension request (-1 pool refill, -2 larger tier), by parity
  NewCell mb[2];    // cell created in the event of this parity
  int nslots;
  int r_i1;         // resume: pending nucleation target
  int fin;          // exit code of the loop
#ifdef AMCD_PHASE_PROFILE
  unsigned long long arr[2][16];
  int aflag[2][16];
#endif
};
// The decision every wave takes after the barrier (S/VoxelsCA.C:476-485).
struct Decision {
  int stop, nuc, i1, cand, wave;
  double dt;
};
// Lane q < NW reads wave q's best (one entry per lane, so the eight waves that decide at
// once after the barrier do not queue on the LDS for the whole table each), and the
// tournament runs across lanes with row shifts: at level h lane q compares its entry
// with lane q + h's.  `nd` is the number of draws this launch has consumed, which
// selects the event's uniform.
template <int H>
__device__ static inline int row_from(int v) {  // lane i takes lane i + H of its row of 16
  return __builtin_amdgcn_update_dpp(v, v, 0x100 + H, 0xf, 0xf, false);
}
template <int H>
__device__ static inline double row_from(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)row_from<H>((int)b), hi = (unsigned)row_from<H>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int NW, int H>
__device__ static inline void tournament_level(Best& e, int& wi) {
  if constexpr (H < NW) {
    Best o;
    o.t = row_from<H>(e.t);
    o.r = row_from<H>(e.r);
    o.key = row_from<H>(e.key);
    o.cand = row_from<H>(e.cand);
    o.tgt = row_from<H>(e.tgt);
    const int ow = row_from<H>(wi);
    const bool bt = better(o, e);
    e.t = bt ? o.t : e.t;
    e.r = bt ? o.r : e.r;
    e.key = bt ? o.key : e.key;
    e.cand = bt ? o.cand : e.cand;
    e.tgt = bt ? o.tgt : e.tgt;
    wi = bt ? ow : wi;
    tournament_level<NW, 2 * H>(e, wi);
  }
}
template <int NW>
__device__ static inline Best lane_entry(const EngineShared& sh, int par) {  // wave `lane`'s best
  static_assert(NW <= 16, "one row of lanes");
  const int lane = threadIdx.x & 63;
  Best e = best_none();
  if (lane < NW) e = sh.sb[par][lane];
  return e;
}
template <int NW>
__device__ static inline Decision decide_from(const EngineShared& sh, Best e, int par, double T, double DelT,
                                              double rX, int nd) {
  const int lane = threadIdx.x & 63;
  int wi = lane;
  const double U = sh.ubuf[(nd >> 6) & 1][nd & 63];
  const int susp = sh.susp[par];
  // The smallest time first (three row shifts of one double): when a single entry holds it,
  // that entry is the best whatever its ruqs and key.  Otherwise the pairwise tree, the
  // left entry kept unless the right one is strictly better: the first best entry in wave
  // order, as a sequential scan finds it (the order is a strict weak order: times and ruqs
  // are never NaN here).
  double tm = e.t;
  tm = fmin(tm, row_from<1>(tm));
  tm = fmin(tm, row_from<2>(tm));
  tm = fmin(tm, row_from<4>(tm));
  static_assert(NW <= 8, "three levels cover the entries");
  const double t0 = lane_f64(tm, 0);
  const unsigned long long at_min = __ballot((lane < NW) & (e.cand >= 0) & (e.t == t0));
  int wq;
  Best b;
  if (__popcll(at_min) == 1) {
    wq = __builtin_ctzll(at_min);
    b = lane_best(e, wq);
  } else {
    tournament_level<NW, 1>(e, wi);
    b.t = lane_f64(e.t, 0);
    b.r = lane_f64(e.r, 0);
    b.key = __builtin_amdgcn_readlane(e.key, 0);
    b.cand = __builtin_amdgcn_readlane(e.cand, 0);
    b.tgt = __builtin_amdgcn_readlane(e.tgt, 0);
    wq = __builtin_amdgcn_readlane(wi, 0);
  }
  // branch-free, so that the reads of the uniform and the suspension flag are issued
  // with the read of the bests and not one after another
  Decision x;
  const bool run = T < DelT;                      // loop condition (S/VoxelsCA.C:381)
  const bool none = (b.cand < 0) | (b.t >= 1e6);  // S/VoxelsCA.C:480
  const bool go = run & (susp == 0) & !none;
  const bool nuc = U < rX;
  x.stop = !run ? 1 : (susp != 0 ? susp : (none ? 1 : 0));
  x.nuc = go & nuc;
  x.i1 = b.tgt;
  x.cand = b.cand;
  x.wave = wq;
  x.dt = (go & !nuc) ? b.t : 0.0;
  return x;
}
template <int NW>
__device__ static inline Decision decide(const EngineShared& sh, int par, double T, double DelT, double rX, int nd) {
  return decide_from<NW>(sh, lane_entry<NW>(sh, par), par, T, DelT, rX, nd);
}
// What the engine needs to perform a capture, loaded for every wave's best while the
// decision is being taken: lane q < NW holds the entry of wave q's best (target voxel,
// capturing cell's grain and rotation, the target's cold flag and temperature, the
// set-up's re-centring), and lane l holds positions l, l + 64, l + 128, l + 192 of the
// candidates' neighbour rows laid end to end (26 per candidate).
struct EnginePrefetch {
  double R[9], T1, rpn[4];
  double4 rc;
  int v1, G, cold, nbr[4];
};
template <int NW>
__device__ static inline EnginePrefetch engine_prefetch(const Inst& d, const EngineShared& sh, const Best& e, int par,
                                                        int ns_pre) {
  // branch-free, with every index clamped into its array (a lane without an entry reads
  // entry 0), so that the loads are issued here and not moved past the decision
  const int lane = threadIdx.x & 63;
  EnginePrefetch f;
  const int cand = e.cand >= 0 ? e.cand : 0, tgt = e.tgt >= 0 ? e.tgt : 0;
  const size_t sw = (size_t)(cand / AMCD_NB);
  f.v1 = gmem(d.cV)[cand];
  f.G = gmem(d.gG)[sw];
  f.cold = gmem(d.aCold)[tgt];
  f.T1 = gmem(d.aT)[tgt];
#pragma unroll
  for (int q = 0; q < 9; ++q) f.R[q] = gmem(d.gR)[9 * sw + q];
  {  // meaningful for entries of set-up lists only
    const auto q = gmem((double*)(d.cRC + cand));
    f.rc = make_double4(q[0], q[1], q[2], q[3]);
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int fl = lane + 64 * m, q = min(fl / AMCD_NB, NW - 1), j = fl - (fl / AMCD_NB) * AMCD_NB;
    int tq = sh.sb[par][q].tgt;
    tq = tq >= 0 ? tq : 0;
    f.nbr[m] = gmem(d.aNbr)[(size_t)tq * AMCD_NB + j];
    f.rpn[m] = gmem(d.aRN)[(size_t)tq * AMCD_NB + j];
  }
  return f;
}
// keeps the prefetched registers live at this point: the loads must have been issued
// before it, and their wait is placed here
__device__ static inline void prefetch_landed(const EnginePrefetch& f) {
  asm volatile("" ::"v"(f.v1), "v"(f.G), "v"(f.cold), "v"(f.T1), "v"(f.rc.x), "v"(f.rc.y), "v"(f.rc.z),
               "v"(f.rc.w), "v"(f.nbr[0]), "v"(f.nbr[1]), "v"(f.nbr[2]), "v"(f.nbr[3]), "v"(f.rpn[0]),
               "v"(f.rpn[1]), "v"(f.rpn[2]), "v"(f.rpn[3]));
  asm volatile("" ::"v"(f.R[0]), "v"(f.R[1]), "v"(f.R[2]), "v"(f.R[3]), "v"(f.R[4]), "v"(f.R[5]), "v"(f.R[6]),
               "v"(f.R[7]), "v"(f.R[8]));
}
// lane j < AMCD_NB takes position j of wave wq's neighbour row from the prefetch
__device__ static inline void take_row(const EnginePrefetch& f, int wq, int lane, int& tn, double& rp) {
  const int fl = wq * AMCD_NB + (lane < AMCD_NB ? lane : 0), src = (fl & 63) << 2, m = fl >> 6;
  int t = -1, lo = 0, hi = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int a = __builtin_amdgcn_ds_bpermute(src, f.nbr[k]);
    const long long b = __double_as_longlong(f.rpn[k]);
    const int bl = __builtin_amdgcn_ds_bpermute(src, (int)b), bh = __builtin_amdgcn_ds_bpermute(src, (int)(b >> 32));
    t = m == k ? a : t;
    lo = m == k ? bl : lo;
    hi = m == k ? bh : hi;
  }
  tn = lane < AMCD_NB ? t : -1;
  rp = lane < AMCD_NB ? __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo))
                      : 0.0;
}
// Uniforms 64b .. 64b+63 of a launch whose engine state was x0 at its start: lane j
// jumps to x0 * 16807^(128b + 2j + 1) (aj = 16807^(2j+1), cb = 16807^(128b)).
__device__ static inline double draw_batch(unsigned int x0, unsigned int cb, unsigned int aj) {
  const unsigned int xb = mod_m((unsigned long long)x0 * cb);
  const unsigned int x1 = mod_m((unsigned long long)xb * aj);
  return canonical(x1, lcg_next(x1));
}

// The register tiers keep each cell's head distance and next distance in LDS and its scan
// key in a register, so that EVAL reads no global memory unless a tie has to be walked.  Up
// to eight cells per owner thread the head's ruqs are in LDS too; the sixteen-cell tier
// (7,168 cells) has room for two doubles per cell and reads the ruqs from the list, a load
// issued with the LDS reads and needed only after the divisions.  The memory-resident tier
// reads the head from the lists.
template <int K>
struct HeadCache {
  static constexpr bool on = K > 0 && K <= 16;
  static constexpr bool ruqs = K > 0 && K <= 8;
  static constexpr int words = ruqs ? 3 : 2;  // doubles per cell
};
// Liveness bits in LDS: the 16-cells-per-thread tier gives half of them to its head cache, and
// an interface beyond them runs in the memory-resident tier with the bitmap in memory.
template <int K, bool GD>
struct DeadBits {
  static constexpr int n = GD ? 32 : (K == 16 ? (1 << 18) : AMCD_DEAD_BITS);
};
// Kept out of line: inlined into the kernel beside the engine's loop, the register tiers'
// loops ran about 1% slower per event.
template <int NT, int K>
__device__ __attribute__((noinline)) static void owner_loop(const Inst& d_lds, EngineShared& sh,
                                                            const unsigned* dead, double* hc, int go, double T,
                                                            long long& visits_out, long long& evals_out) {
  // The descriptor's fields in registers: read through the LDS copy, each would be read again
  // after every store to LDS (head cache, bests, liveness), which might alias it.
  const Inst d = d_lds;
  constexpr bool HC = HeadCache<K>::on, HR = HeadCache<K>::ruqs;
  // the refresh's block loaded ahead of its use (its registers spilled in the 16-cell tier)
  constexpr bool PF = HC && K <= 8;
  constexpr int HN = (NT - 64) * K;  // head-cache slots (one per cell of the tier)
  // work counters are kept in registers and handed back once at the end (a reference
  // incremented inside the loop would live in scratch memory)
  long long visits = 0, evals = 0;
  constexpr int NO = NT - 64;  // owner threads
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int me = tid - 64;
  const float fInf = __builtin_inff();
  const amcd_params& p = d.p;
  // loop constants held in registers (a read through the descriptor in LDS would be
  // repeated after every barrier)
  const double DelT = p.DelT, rX = p.rX;
  double E[K], V[K];
  float lo[K], hi[K];
  // c1: the head entry; with the head cache, its position in the slot's list (5 bits)
  // packed with its scan key (scan position * 32 + neighbour rank, < 2^24) above them
  int c1[K], tgt1[K];
  const bool loaded = go != -2;
  const int ns0 = sh.nslots;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = me + k * NO;
    E[k] = V[k] = 0.0;
    lo[k] = hi[k] = fInf;
    c1[k] = 0;
    tgt1[k] = -1;
    if (loaded && s < ns0) {
      E[k] = gmem(d.gExt)[s];
      V[k] = gmem(d.gVhat)[s];
      const SlotHead h = slot_head(d, dead, s, 0, -1);
      c1[k] = h.c1;
      tgt1[k] = h.tgt1;
      if constexpr (HC) {
        c1[k] = (h.c1 - s * AMCD_NB) | ((gmem(d.gScan)[s] * 32 + h.k1) << 5);
        hc[s] = h.D1;
        hc[HN + s] = h.Dn;
        if constexpr (HR) hc[2 * HN + s] = h.r1;
      }
      if (h.tgt1 >= 0) tau_bounds(h.D1, E[k], V[k], T, lo[k], hi[k]);
      // bring the rest of the list (positions 16..25 of the distances and ruqs, the
      // targets' line) into the cache: later head refreshes read it within the event
      const size_t cb = (size_t)s * AMCD_NB;
      const double w0 = gmem(d.cD)[cb + 16], w1 = gmem(d.cR)[cb + 16];
      const int w2 = gmem(d.cTgt)[cb + 25];
      asm volatile("" ::"v"(w0), "v"(w1), "v"(w2));
    }
  }
  int skip = -1;  // voxel the engine may be capturing right now
  int par = 0;
  int nd = 0;     // draws consumed by this launch
  int nev = 0;    // events since the bounds were last recomputed
  Decision x;
  x.stop = 0;
  x.nuc = 1;
  x.i1 = sh.r_i1;
  x.dt = 0.0;
  bool apply_first = go == 2;
  PHASE_FLAGS(aflag);
  int ks = -1;  // the cell whose list block pb is loaded for its head refresh (-1: none)
  SlotBlock pb = slot_block_none();
  PHASE_INIT(ot);
  while (go > 0) {
    // The slot numbers are re-derived from an opaque copy of the thread index in every
    // iteration: otherwise the compiler keeps per-cell addresses live across the loop,
    // which costs more registers than the cell state itself.
    int mo = me;
    asm volatile("" : "+v"(mo));
    if (!apply_first) {
      // ---- EVAL
      float u = fInf;
#pragma unroll
      for (int k = 0; k < K; ++k) u = fminf(u, hi[k]);
      u = wave_min(u);
      Best b = best_none();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (lo[k] != fInf && lo[k] <= u) {
          // exact capture time of the cell's head (S/VoxelsCA.C:420-440)
          const int s = mo + k * NO;
          const int c = HC ? s * AMCD_NB + (c1[k] & 31) : c1[k];
          const int pos = c - s * AMCD_NB;
          const double e = E[k], v = V[k];
          double D1, Dn, rr;
          int kb, ky;
          if constexpr (HC) {
            D1 = hc[s];
            Dn = hc[HN + s];
            rr = HR ? hc[2 * HN + s] : gmem(d.cR)[c];
            ky = c1[k] >> 5;
            kb = ky & ~31;
          } else {
            D1 = gmem(d.cD)[c];
            Dn = (pos + 1 < AMCD_NB) ? gmem(d.cD)[c + 1] : __builtin_inf();
            rr = gmem(d.cR)[c];
            kb = gmem(d.gScan)[s] * 32;
            ky = kb + gmem(d.cK)[c];
          }
          visits++;
          evals++;
          // the head's time and the next entry's, side by side (one round of head reads,
          // two independent divisions)
          const double t1 = (D1 - e) / v;
          const double tn = (Dn - e) / v;
          const bool ok = t1 <= 1e6;
          {
            Best y;
            y.t = t1;
            y.r = rr;
            y.key = ky;
            y.cand = c;
            y.tgt = tgt1[k];
            // times are non-decreasing along the list: a tie needs the next entry's
            // time to equal t1; then walk the tied block
            if (ok & (tn == t1)) {
              PHASE_FLAG(aflag, 2);
              const int n = gmem(d.gN)[s];
              for (int q = pos + 1; q < n; ++q) {
                const int tq = gmem(d.cTgt)[c - pos + q];
                if (!live_target(dead, tq, skip)) continue;
                evals++;
                const double tt = (gmem(d.cD)[c - pos + q] - e) / v;
                if (!(tt == t1)) break;
                Best z;
                z.t = tt;
                z.r = gmem(d.cR)[c - pos + q];
                z.key = kb + gmem(d.cK)[c - pos + q];
                z.cand = c - pos + q;
                z.tgt = tq;
                if (better(z, y)) y = z;
              }
            }
            if (ok && better(y, b)) b = y;
          }
        }
      }
      b = wave_best(b);
      if (lane == 0) sh.sb[par][w] = b;
      PHASE(ot, 7, tid == 64);
      PHASE_ARRIVE(sh, par, w, aflag);
      PHASE_FLAGS_CLEAR(aflag);
      __syncthreads();
      PHASE(ot, 8, tid == 64);
      // ---- DECIDE
      const Best eb = lane_entry<NW>(sh, par);
      x = decide_from<NW>(sh, eb, par, T, DelT, rX, nd);
      if (x.stop) break;
      ++nd;
      PHASE(ot, 5, tid == 64);
    }
    apply_first = false;
    // the cell (first of each thread) whose head is the voxel captured now: the next
    // block of its list is loaded before the growth and the takeover below, and waited
    // for where the refresh of its head uses it
    if constexpr (PF) {
      ks = -1;
      int pk = 0;
#pragma unroll
      for (int k = K - 1; k >= 0; --k) {
        const bool take = (tgt1[k] == x.i1) & (tgt1[k] >= 0);
        ks = take ? k : ks;
        pk = take ? c1[k] : pk;
      }
      if (ks >= 0) pb = slot_block(d, mo + ks * NO, (pk & 31) + 1);
    }
    // ---- APPLY
    const double Dt = x.dt;
    if (!x.nuc) {  // growth of every cell growing at this scan (S/VoxelsCA.C:511-516)
      // straight-line over the K cells, so that their updates interleave
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const bool grows = tgt1[k] >= 0;
        const double e = E[k] + V[k] * Dt;
        const bool clamp = grows & (e < 0.0);
        E[k] = grows ? (clamp ? 0.0 : e) : E[k];
        lo[k] = clamp ? -fInf : lo[k];  // clamped: recompute the bounds below
      }
    }
    T = T + Dt;  // the engine's tinc: same additions in the same order
    const int i1 = x.i1;
    // the cell created by the previous event (its best was evaluated by the engine)
    const NewCell& nc = sh.mb[par];
    if (nc.valid && nc.slot % NO == mo) {
      PHASE_FLAG(aflag, 4);
      const int kk = nc.slot / NO;
      double e = nc.E;
      if (!x.nuc) {
        const double g = e + nc.V * Dt;
        e = (g < 0.0) ? 0.0 : g;
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (k == kk) {
          E[k] = e;
          V[k] = nc.V;
          c1[k] = HC ? nc.key1 << 5 : nc.c1;  // the head is entry 0
          tgt1[k] = nc.tgt1;
          lo[k] = -fInf;  // bounds below (its head may also be the voxel captured now)
        }
      if constexpr (HC) {
        hc[nc.slot] = nc.D1;
        hc[HN + nc.slot] = nc.Dn1;
        if constexpr (HR) hc[2 * HN + nc.slot] = nc.r1;
      }
      if (ks == kk) ks = -1;  // a block loaded for the slot's previous state
    }
    // heads captured now, clamped cells, the new cell, and every cell once per 2^20 events
    ++nev;
    const bool all = nev >= (1 << 20);
    if (all) nev = 0;
    bool any = false;
#pragma unroll
    for (int k = 0; k < K; ++k) any |= (tgt1[k] >= 0) & ((tgt1[k] == i1) | (lo[k] == -fInf) | all);
    if (__ballot(any) != 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const bool hit = tgt1[k] >= 0 && (tgt1[k] == i1 || lo[k] == -fInf || all);
        if (!hit) continue;
        const int s = mo + k * NO;
        if (tgt1[k] == i1) {
          const int pos = HC ? (c1[k] & 31) : c1[k] - s * AMCD_NB;
          PHASE_FLAG(aflag, (PF && k == ks) ? 8 : 1);
          const SlotHead h = (PF && k == ks) ? slot_head_pre(d, dead, s, pos + 1, i1, pb)
                                             : slot_head(d, dead, s, pos + 1, i1);
          tgt1[k] = h.tgt1;
          if constexpr (HC) {
            c1[k] = (h.c1 - s * AMCD_NB) | ((((c1[k] >> 5) & ~31) + h.k1) << 5);
            hc[s] = h.D1;
            hc[HN + s] = h.Dn;
            if constexpr (HR) hc[2 * HN + s] = h.r1;
          } else {
            c1[k] = h.c1;
          }
          lo[k] = hi[k] = fInf;
          if (h.tgt1 >= 0) tau_bounds(h.D1, E[k], V[k], T, lo[k], hi[k]);
        } else {
          PHASE_FLAG(aflag, 16);
          tau_bounds(HC ? hc[s] : gmem(d.cD)[c1[k]], E[k], V[k], T, lo[k], hi[k]);
        }
      }
    }
    skip = i1;
    par ^= 1;
    PHASE(ot, 6, tid == 64);
  }
  // half-diagonals back to the growing list (the cell still in the mailbox is current there)
  if (loaded) {
    const int ns = sh.nslots;
    const NewCell& nc = sh.mb[par];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int s = me + k * NO;
      if (s < ns && !(nc.valid && nc.slot == s)) gmem(d.gExt)[s] = E[k];
    }
  }
  visits_out += visits;
  evals_out += evals;
}

// The owner protocol of owner_loop for growing lists larger than the register tiers hold:
// each cell's state (E in gExt, V in gVhat, bounds in gLH, head in gHd) stays in global
// memory and every event streams the lane's cells (slots me, me + NO, ...).  Same
// decisions, same arithmetic, same order of additions as the register tiers.
template <int NT>
__device__ static __attribute__((always_inline)) void owner_loop_mem(const Inst& d_lds, EngineShared& sh,
                                                                      const unsigned* dead, int go, double T,
                                                                      long long& visits_out, long long& evals_out) {
  const Inst d = d_lds;  // in registers, as in owner_loop
  long long visits = 0, evals = 0;
  constexpr int NO = NT - 64;
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int me = tid - 64;
  const float fInf = __builtin_inff();
  const amcd_params& p = d.p;
  // loop constants held in registers (a read through the descriptor in LDS would be
  // repeated after every barrier)
  const double DelT = p.DelT, rX = p.rX;
  int ns = go != -2 ? sh.nslots : 0;  // slots this lane knows of (new ones arrive by mailbox)
  // The lane's smallest upper bound and its two smallest lower bounds (with their slots)
  // and third smallest lower bound, gathered whenever every cell's bounds pass through the
  // lane (set-up, and the growth pass of each event): the next evaluation needs no pass
  // over the cells unless a third cell reaches the wave's smallest upper bound.
  float hmin = fInf, l0 = fInf, l1 = fInf, l2 = fInf;
  int f0 = -1, f1 = -1;
  auto note = [&](int s, float2 lh) {
    hmin = fminf(hmin, lh.y);
    if (lh.x < l0) {
      l2 = l1;
      l1 = l0;
      f1 = f0;
      l0 = lh.x;
      f0 = s;
    } else if (lh.x < l1) {
      l2 = l1;
      l1 = lh.x;
      f1 = s;
    } else if (lh.x < l2) {
      l2 = lh.x;
    }
  };
  for (int s = me; s < ns; s += NO) {
    const SlotHead h = slot_head(d, dead, s, 0, -1);
    float lo = fInf, hi = fInf;
    if (h.tgt1 >= 0) tau_bounds(h.D1, gmem(d.gExt)[s], gmem(d.gVhat)[s], T, lo, hi);
    gstore8(d.gHd + s, make_int2(h.c1, h.tgt1));
    gstore8(d.gLH + s, make_float2(lo, hi));
    note(s, make_float2(lo, hi));
  }
  int skip = -1;
  int par = 0;
  int nd = 0;
  int nev = 0;
  Decision x;
  x.stop = 0;
  x.nuc = 1;
  x.i1 = sh.r_i1;
  x.dt = 0.0;
  bool apply_first = go == 2;
  // the lane's cells are streamed MB at a time: all loads of a batch are issued before
  // any of them is used
  constexpr int MB = 8;
  while (go > 0) {
    if (!apply_first) {
      // ---- EVAL: the wave's smallest upper bound, then the cells whose lower bound reaches it
      const float u = wave_min(hmin);
      Best b = best_none();
      auto evaluate = [&](int s) {  // exact capture time of slot s's head (S/VoxelsCA.C:420-440)
        const int2 hd = gload8(d.gHd + s);
        const int c = hd.x;
        const int pos = c - s * AMCD_NB;
        const double e = gmem(d.gExt)[s], v = gmem(d.gVhat)[s];
        const double D1 = gmem(d.cD)[c];
        const double Dn = (pos + 1 < AMCD_NB) ? gmem(d.cD)[c + 1] : __builtin_inf();
        const double rr = gmem(d.cR)[c];
        const int kb = gmem(d.gScan)[s] * 32;
        const int kk = gmem(d.cK)[c];
        visits++;
        evals++;
        const double t1 = (D1 - e) / v;
        if (t1 <= 1e6) {
          Best y;
          y.t = t1;
          y.r = rr;
          y.key = kb + kk;
          y.cand = c;
          y.tgt = hd.y;
          if ((Dn - e) / v == t1) {
            const int n = gmem(d.gN)[s];
            for (int q = pos + 1; q < n; ++q) {
              const int tq = gmem(d.cTgt)[c - pos + q];
              if (!live_target(dead, tq, skip)) continue;
              evals++;
              const double tt = (gmem(d.cD)[c - pos + q] - e) / v;
              if (!(tt == t1)) break;
              Best z;
              z.t = tt;
              z.r = gmem(d.cR)[c - pos + q];
              z.key = kb + gmem(d.cK)[c - pos + q];
              z.cand = c - pos + q;
              z.tgt = tq;
              if (better(z, y)) y = z;
            }
          }
          if (better(y, b)) b = y;
        }
      };
      if (l2 <= u) {  // a third cell may reach u: every cell of the lane, MB at a time
        for (int s0 = me; s0 < ns; s0 += MB * NO) {
          float lob[MB];
#pragma unroll
          for (int q = 0; q < MB; ++q) {
            const int s = s0 + q * NO;
            lob[q] = s < ns ? gload8(d.gLH + s).x : fInf;
          }
#pragma unroll
          for (int q = 0; q < MB; ++q)
            if (lob[q] != fInf && lob[q] <= u) evaluate(s0 + q * NO);
        }
      } else {
        if (f0 >= 0 && l0 != fInf && l0 <= u) evaluate(f0);
        if (f1 >= 0 && l1 != fInf && l1 <= u) evaluate(f1);
      }
      b = wave_best(b);
      if (lane == 0) sh.sb[par][w] = b;
      __syncthreads();
      x = decide<NW>(sh, par, T, DelT, rX, nd);
      if (x.stop) break;
      ++nd;
    }
    apply_first = false;
    // ---- APPLY
    const double Dt = x.dt;
    const int i1 = x.i1;
    // the cell created by the previous event joins its owner's slots ungrown; the pass
    // below grows it like every other cell and computes its bounds
    const NewCell& nc = sh.mb[par];
    if (nc.valid && nc.slot % NO == me) {
      gmem(d.gExt)[nc.slot] = nc.E;
      gmem(d.gVhat)[nc.slot] = nc.V;
      gstore8(d.gHd + nc.slot, make_int2(nc.c1, nc.tgt1));
      gstore8(d.gLH + nc.slot, make_float2(-fInf, fInf));
    }
    if (nc.valid) ns = max(ns, nc.slot + 1);
    T = T + Dt;
    ++nev;
    const bool all = nev >= (1 << 20);
    if (all) nev = 0;
    hmin = l0 = l1 = l2 = fInf;
    f0 = f1 = -1;
    for (int s0 = me; s0 < ns; s0 += MB * NO) {
      int2 hdb[MB];
      float2 lhb[MB];
      double eb[MB], vb[MB];
#pragma unroll
      for (int q = 0; q < MB; ++q) {
        const int s = s0 + q * NO;
        const bool in = s < ns;
        hdb[q] = in ? gload8(d.gHd + s) : make_int2(0, -1);
        lhb[q] = in ? gload8(d.gLH + s) : make_float2(fInf, fInf);
        eb[q] = in ? gmem(d.gExt)[s] : 0.0;
        vb[q] = in ? gmem(d.gVhat)[s] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < MB; ++q) {
        const int s = s0 + q * NO;
        int2 hd = hdb[q];
        if (hd.y < 0) continue;
        float2 lh = lhb[q];
        double e = eb[q];
        const double v = vb[q];
        if (!x.nuc) {  // growth (S/VoxelsCA.C:511-516)
          const double g = e + v * Dt;
          if (g < 0.0) {
            e = 0.0;
            lh.x = -fInf;
          } else {
            e = g;
          }
          gmem(d.gExt)[s] = e;
        }
        if (hd.y == i1) {
          const SlotHead h = slot_head(d, dead, s, hd.x - s * AMCD_NB + 1, i1);
          hd = make_int2(h.c1, h.tgt1);
          lh = make_float2(fInf, fInf);
          if (h.tgt1 >= 0) tau_bounds(h.D1, e, v, T, lh.x, lh.y);
          gstore8(d.gHd + s, hd);
          gstore8(d.gLH + s, lh);
        } else if (lh.x == -fInf || all) {
          tau_bounds(gmem(d.cD)[hd.x], e, v, T, lh.x, lh.y);
          gstore8(d.gLH + s, lh);
        }
        note(s, lh);
      }
    }
    skip = i1;
    par ^= 1;
  }
  visits_out += visits;
  evals_out += evals;
}

// GD: the liveness bitmap lives in memory (d.gdead) instead of LDS, for interfaces of more
// than AMCD_DEAD_BITS cells (instantiated with the memory-resident tier only).
template <int NT, int K, bool GD = false>
__global__ __launch_bounds__(NT) void k_event_loop(Inst* I, const amcd_step* S, amcd_out* O) {
  constexpr int NW = NT / 64;
  // slots this tier holds: K per owner thread in registers, or (K = 0) every slot of the
  // domain's arrays in memory
  const int CAP = K > 0 ? (NT - 64) * K : I[blockIdx.x].acap;
  __shared__ EngineShared sh;
  __shared__ int s_go;
  constexpr int DB = DeadBits<K, GD>::n;
  __shared__ unsigned s_dead[DB / 32];
  __shared__ double s_hc[HeadCache<K>::on ? HeadCache<K>::words * (NT - 64) * K : 1];
  // The descriptor and step record are copied to LDS: every event ends in a barrier, after
  // which the compiler must re-read anything in global memory, and each array access would
  // otherwise wait for its base pointer to come back from global memory first.
  __shared__ Inst s_inst;
  __shared__ amcd_step s_step;
  if (I[blockIdx.x].ls->done) return;
  if (threadIdx.x == 0) {
    s_inst = I[blockIdx.x];
    s_step = S[blockIdx.x];
  }
  __syncthreads();
  const Inst& d = s_inst;
  LoopState* ls = d.ls;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const amcd_params& p = d.p;
  // loop constants held in registers (a read through the descriptor in LDS would be
  // repeated after every barrier)
  const double DelT = p.DelT, rX = p.rX;
  const amcd_step& st = s_step;
  const int Na = ls->Na;
  unsigned* const dead = GD ? d.gdead : s_dead;
  if (GD || Na <= DB) {  // liveness bitmap from the current states
    for (int wd = tid; wd < (Na + 31) / 32; wd += NT) {
      unsigned m = 0;
      for (int q = 0; q < 32; ++q) {
        const int a = 32 * wd + q;
        if (a >= Na || d.aVS[a] != 1) m |= 1u << q;
      }
      dead[wd] = m;
    }
  }
  if (tid == 0) {
    if (ls->nslots_pre < 0) ls->nslots_pre = ls->nslots;  // the step's first launch
    sh.nslots = ls->nslots;
    sh.mb[0].valid = sh.mb[1].valid = 0;
    sh.susp[0] = sh.susp[1] = 0;
    for (int q = 0; q < 16; ++q) sh.sb[0][q] = sh.sb[1][q] = best_none();
    sh.fin = 0;
    sh.r_i1 = -1;
    s_go = 1;
    if (!GD && Na > DB) {
      s_go = -3;  // more interface cells than the LDS bitmap holds: the host relaunches with GD
    } else if (ls->nslots > CAP) {
      s_go = -2;  // more growing cells than this tier holds
    } else if (ls->pending) {
      sh.r_i1 = ls->pend_tgt;
      s_go = 2;   // resume with the pending nucleation
    } else if (ls->events == 0 && ls->Na == ls->nmushy0) {
      s_go = 0;   // loop condition vSind2 != countS fails at once (S/VoxelsCA.C:381)
    }
  }
  __syncthreads();
  const int go = s_go;
  long long visits = 0, evals = 0;
  if (w > 0) {
    if constexpr (K > 0)
      owner_loop<NT, K>(d, sh, dead, s_hc, go, ls->tinc, visits, evals);
    else
      owner_loop_mem<NT>(d, sh, dead, go, ls->tinc, visits, evals);
  } else {
    // engine state, identical in every lane of wave 0
    long long e_visits = 0, e_evals = 0, e_units = 0;
    double tinc = ls->tinc;
    int vSind2 = ls->vSind2, nGrain = ls->nGrain, pool_used = ls->pool_used;
    int events = ls->events, captures = ls->captures, nucs = ls->nucs;
    int nslots = ls->nslots;
    const int ns_pre = ls->nslots_pre;
    // the uniforms, 64 per batch, drawn ahead by jump-ahead (S/VoxelsCA.C:483)
    const unsigned int x0 = ls->rng;
    const long long draws0 = ls->draws;
    const unsigned int aj = lcg_pow(2ull * lane + 1), c128 = lcg_pow(128);
    unsigned int cb = 1;  // 16807^(128 b) for the next batch b to write
    int nd = 0;
    if (go > 0) {
      sh.ubuf[0][lane] = draw_batch(x0, cb, aj);
      cb = mod_m((unsigned long long)cb * c128);
    }
    int par = 0, fin = 0;
    const unsigned long long c_start = clock64();  // the loop's own duration (stats)
    Decision x;
    x.stop = 0;
    x.nuc = 1;
    x.i1 = sh.r_i1;
    x.cand = -1;
    x.wave = 0;
    x.dt = 0.0;
    bool apply_first = go == 2;
    EnginePrefetch pf;
    bool have_pf = false;  // the event's inputs were prefetched with its decision
    PHASE_FLAGS(eflag);
    PHASE_INIT(et);
    while (go > 0) {
      if (!apply_first) {
        PHASE(et, 0, tid == 0);
        PHASE_ARRIVE(sh, par, 0, eflag);
        PHASE_FLAGS_CLEAR(eflag);
        __syncthreads();
        PHASE(et, 1, tid == 0);
        PHASE_LAST(sh, par, NW);
        {
          const Best me = lane_entry<NW>(sh, par);
          pf = engine_prefetch<NW>(d, sh, me, par, ns_pre);
          x = decide_from<NW>(sh, me, par, tinc, DelT, rX, nd);
          prefetch_landed(pf);
        }
        have_pf = true;
        if (x.stop) {
          fin = x.stop < 0 ? x.stop : 0;
          break;
        }
        events++;
        e_units += nslots;  // the growing list this event's decision covered
        if ((nd & 63) == 0) {  // first draw of batch nd/64: write batch nd/64 + 1
          sh.ubuf[((nd >> 6) + 1) & 1][lane] = draw_batch(x0, cb, aj);
          cb = mod_m((unsigned long long)cb * c128);
        }
        ++nd;
        PHASE(et, 2, tid == 0);
      }
      apply_first = false;
      // ---- APPLY: the event on voxel i1
      const int i1 = x.i1;
      NewCell& nc = sh.mb[par ^ 1];
      int susp = 0;
      Best nb = best_none();
      if (lane == 0) nc.valid = 0;
      if (x.nuc && (pool_used >= st.n_pool || pool_used >= d.poolcap || nGrain >= d.gcap)) {
        if (lane == 0) {
          *gmem(&ls->pending) = 1;  // the host refills the pool and resumes
          *gmem(&ls->pend_tgt) = i1;
        }
        susp = -1;
      } else {
        // the event's inputs: the captured voxel's neighbour table and pair ruqs, its cold
        // flag and temperature, the winning entry's voxel, and the capturing cell's grain
        // and rotation -- from the lanes that prefetched them with the decision (a resumed
        // nucleation loads them here)
        const int cand = x.cand, sw = cand >= 0 ? cand / AMCD_NB : 0;
        // an entry of a list the set-up built carries its capture's re-centring
        const bool pre = !x.nuc && sw < ns_pre;
        int tn, cold, v1, vjs = 0, G = 1;
        double rpair, T1, fcap = 0.0, R[9], c0[3], cen[3], E0;
        double4 rc = make_double4(0.0, 0.0, 0.0, 0.0);
        if (have_pf) {
          const int wq = x.wave;
          take_row(pf, wq, lane, tn, rpair);
          cold = __builtin_amdgcn_readlane(pf.cold, wq);
          T1 = lane_f64(pf.T1, wq);
          v1 = __builtin_amdgcn_readlane(pf.v1, wq);
          if (!x.nuc) {
#pragma unroll
            for (int q = 0; q < 9; ++q) R[q] = lane_f64(pf.R[q], wq);
            G = __builtin_amdgcn_readlane(pf.G, wq);
            if (pre)
              rc = make_double4(lane_f64(pf.rc.x, wq), lane_f64(pf.rc.y, wq), lane_f64(pf.rc.z, wq),
                                lane_f64(pf.rc.w, wq));
          }
        } else {
          tn = (lane < AMCD_NB) ? gmem(d.aNbr)[(size_t)i1 * AMCD_NB + lane] : -1;
          rpair = (lane < AMCD_NB) ? gmem(d.aRN)[(size_t)i1 * AMCD_NB + lane] : 0.0;
          cold = gmem(d.aCold)[i1];
          T1 = gmem(d.aT)[i1];
          v1 = cand >= 0 ? gmem(d.cV)[cand] : gmem(d.aVox)[i1];  // (no entry: a resumed nucleation)
          if (!x.nuc) {
#pragma unroll
            for (int q = 0; q < 9; ++q) R[q] = gmem(d.gR)[9 * (size_t)sw + q];
            G = gmem(d.gG)[sw];
            if (pre) {
              const auto q = gmem((double*)(d.cRC + cand));
              rc = make_double4(q[0], q[1], q[2], q[3]);
            }
          }
        }
        PHASE(et, 9, tid == 0);
        if (!x.nuc && !pre) {  // a cell made in this launch: re-centred below
#pragma unroll
          for (int q = 0; q < 3; ++q) c0[q] = gmem(d.gC)[3 * (size_t)sw + q];
          fcap = gmem(d.cF)[cand];
          vjs = gmem(d.gV)[sw];
        }
        if (x.nuc) {  // S/VoxelsCA.C:486-506
          PHASE_FLAG(eflag, 8);
          const auto P = gmem(d.pool) + 9 * (size_t)pool_used;
          for (int q = 0; q < 9; ++q) R[q] = P[q];
          G = nGrain + 1;
          centre(d, v1, cen);
          E0 = 0.0;
          if (lane < 9) gmem(d.rot)[9 * (size_t)(G - 1) + lane] = P[lane];
          nGrain = G;
          pool_used++;
          nucs++;
          if (lane == 0) *gmem(&ls->pending) = 0;
        } else {      // S/VoxelsCA.C:507-605
          if (pre) {
            cen[0] = rc.x;
            cen[1] = rc.y;
            cen[2] = rc.z;
            E0 = rc.w;
            PHASE_FLAG(eflag, 1);
          } else {
            PHASE_FLAG(eflag, 2);
            recentre(d, vjs, v1, c0, R, fcap, cen, &E0);
          }
          captures++;
        }
        PHASE(et, 10, tid == 0);
        const int scan = vSind2;
        if (lane == 0) {
          gmem(d.aG)[i1] = G;
          gmem(d.aCx)[i1] = cen[0];
          gmem(d.aCy)[i1] = cen[1];
          gmem(d.aCz)[i1] = cen[2];
          gmem(d.aExt)[i1] = E0;
          gmem(d.aVS)[i1] = 2;
          gmem(d.aScan)[i1] = scan;
          dead[i1 >> 5] |= 1u << (i1 & 31);
        }
        vSind2++;
        tinc += x.dt;
        PHASE(et, 3, tid == 0);
        // the new mushy voxel joins the growing list if it has liquid neighbours
        const bool liq = live_target(dead, tn, i1);
        const unsigned long long bal = __ballot(liq);
        if (cold && bal) {
          PHASE_FLAG(eflag, 4);
          const int slot = nslots;
          const size_t cb = (size_t)slot * AMCD_NB;
          const double V0 = growth_velocity(p, T1);
          double D = __builtin_inf(), f = 0, r = __builtin_inf();
          int key = 0x7fffffff;
          // neighbour lane of v1, in the order of the active-neighbour table
          int vi, vj, vk;
          decode(d, v1, vi, vj, vk);
          const int qq = lane < 13 ? lane : lane + 1;
          const int ddx = qq % 3 - 1, ddy = (qq / 3) % 3 - 1, ddz = qq / 9 - 1;
          const int vt = v1 + ddz * d.nxy + ddy * p.nx + ddx;
          if (liq) {
            const double ct[3] = {((double)(vi + ddx) + .5) * p.dx, ((double)(vj + ddy) + .5) * p.dy,
                                  ((double)(vk + ddz) + .5) * p.dz};
            candidate_at(ct, cen, R, &D, &f);
            r = rpair;
            key = scan * 32 + lane;
          }
          // position in (D, ruqs, key) order among the liquid neighbours (cand_less): a
          // uniform loop over the ballot's set bits that broadcasts only D counts the
          // neighbours strictly before by D (NaN last) and marks those with an equal D;
          // a second loop over the lanes in some tie orders those by (ruqs, key)
          int rank = 0;
          unsigned long long tie = 0;
          const bool nan_me = D != D;
          for (unsigned long long m = bal; m != 0; m &= m - 1) {
            const int j = (int)__builtin_ctzll(m);
            const double Dj = lane_f64(D, j);
            const bool nan_j = Dj != Dj;
            const bool before = (nan_me & !nan_j) | (!nan_me & !nan_j & (Dj < D));
            const bool same = (j != lane) & ((nan_me & nan_j) | (!nan_me & !nan_j & (Dj == D)));
            rank += before ? 1 : 0;
            tie |= (unsigned long long)same << j;
          }
          for (unsigned long long m = __ballot(liq & (tie != 0)); m != 0; m &= m - 1) {
            const int j = (int)__builtin_ctzll(m);
            const double rj = lane_f64(r, j);
            const int kj = __builtin_amdgcn_readlane(key, j);
            const bool rne = rj != r;
            const bool before = (rne & (rj < r)) | (!rne & (kj < key));
            rank += (((tie >> j) & 1ull) != 0 && before) ? 1 : 0;
          }
          if (liq) {
            gmem(d.cD)[cb + rank] = D;
            gmem(d.cF)[cb + rank] = f;
            gmem(d.cR)[cb + rank] = r;
            gmem(d.cTgt)[cb + rank] = tn;
            gmem(d.cV)[cb + rank] = vt;
            gmem(d.cK)[cb + rank] = (uint8_t)lane;
            if (rank == 0) {
              nc.D1 = D;
              nc.r1 = r;
              nc.key1 = key;
              nc.c1 = (int)cb;
              nc.tgt1 = tn;
              if (__popcll(bal) < 2) nc.Dn1 = __builtin_inf();
            }
            if (rank == 1) nc.Dn1 = D;
            // this cell's exact best for the next event (all its targets are liquid)
            const double tq = (D - E0) / V0;
            if (tq <= 1e6) {
              nb.t = tq;
              nb.r = r;
              nb.key = key;
              nb.cand = (int)(cb + rank);
              nb.tgt = tn;
            }
          }
          nb = wave_best(nb);
          e_visits++;
          e_evals += __popcll(bal);
          // the slot's static capture inputs (rotation, centroid, voxel)
          double rv = 0.0;
#pragma unroll
          for (int q = 0; q < 9; ++q)
            if (q == lane) rv = R[q];
#pragma unroll
          for (int q = 0; q < 3; ++q)
            if (q + 9 == lane) rv = cen[q];
          if (lane < 9) gmem(d.gR)[9 * (size_t)slot + lane] = rv;
          if (lane >= 9 && lane < 12) gmem(d.gC)[3 * (size_t)slot + lane - 9] = rv;
          if (lane == 0) {
            gmem(d.gV)[slot] = v1;
            gmem(d.gG)[slot] = G;
            gmem(d.gAct)[slot] = i1;
            gmem(d.gScan)[slot] = scan;
            gmem(d.gExt)[slot] = E0;
            gmem(d.gVhat)[slot] = V0;
            gmem(d.gN)[slot] = (uint8_t)__popcll(bal);
            gmem(d.gDead)[slot] = 0;
            nc.E = E0;
            nc.V = V0;
            nc.slot = slot;
            nc.valid = slot < CAP;  // beyond this tier: resume with a larger one
          }
          nslots = slot + 1;
          if (slot >= CAP && tinc < DelT) susp = -2;
        }
        PHASE(et, 4, tid == 0);
      }
      if (lane == 0) {
        sh.sb[par ^ 1][0] = nb;
        sh.susp[par ^ 1] = susp;
        sh.nslots = nslots;
      }
      par ^= 1;
    }
    if (tid == 0) {
      sh.fin = fin;
      ls->tinc = tinc;
      ls->vSind2 = vSind2;
      ls->nGrain = nGrain;
      ls->pool_used = pool_used;
      ls->events = events;
      ls->captures = captures;
      ls->nucs = nucs;
    }
    if (tid == 1 && go != -2) {  // engine state after the nd draws of this launch
      ls->rng = mod_m((unsigned long long)x0 * lcg_pow(2ull * nd));
      ls->draws = draws0 + 2ll * nd;
      O[blockIdx.x].draws = draws0 + 2ll * nd;
    }
    visits += e_visits;
    evals += e_evals;
    if (tid == 0) {
      atomicAdd((unsigned long long*)&O[blockIdx.x].interface_events, (unsigned long long)e_units);
      atomicAdd((unsigned long long*)&O[blockIdx.x].loop_cycles, clock64() - c_start);
    }
  }
  __syncthreads();
  // ---- loop state and counters
  const int ns = sh.nslots;
  const int fin = sh.fin;  // 0 finished, -1 pool refill, -2 larger tier
  const bool suspended = go == -2 || go == -3 || fin < 0;
  if (!suspended)
    for (int s = tid; s < ns; s += NT) d.aExt[d.gAct[s]] = d.gExt[s];
  atomicAdd((unsigned long long*)&O[blockIdx.x].slot_visits, (unsigned long long)visits);
  atomicAdd((unsigned long long*)&O[blockIdx.x].cand_evals, (unsigned long long)evals);
  if (tid == 0) {
    const int status = go == -3 ? -3 : (go == -2 ? 2 : (fin == -1 ? 1 : (fin == -2 ? 2 : 0)));
    ls->nslots = ns;
    ls->done = status == 0 ? 1 : 0;
    amcd_out& o = O[blockIdx.x];
    o.status = status;
    o.n_events = ls->events;
    o.n_captures = ls->captures;
    o.n_nucleations = ls->nucs;
    o.n_slots = ns;
    o.nGrain = ls->nGrain;
    o.pool_used = ls->pool_used;
    o.tinc = ls->tinc;
  }
}

// Active arrays back to the grid (S/VoxelsCA.C:633-643).
__global__ __launch_bounds__(GB) void k_scatter(Inst* I, const int* astart, int ninst) {
  int ii = find_inst(astart, ninst, blockIdx.x);
  const Inst d = I[ii];  // a copy: stores in the loops cannot alias it
  int a = (blockIdx.x - astart[ii]) * GB + threadIdx.x;
  if (d.ls->pending || a >= d.ls->Na || a >= d.acap) return;
  int v = d.aVox[a];
  d.vs[v] = (d.vs[v] & 0xC0) | d.aVS[a];
  d.gid[v] = d.aG[a];
  d.ext[v] = d.aExt[a];
  d.cx[v] = d.aCx[a];
  d.cy[v] = d.aCy[a];
  d.cz[v] = d.aCz[a];
}

// ConvertSolid1(0) (S/VoxelsCA.C:739-758): mushy -> solid when every Moore neighbour
// is mushy or solid.  Every mushy voxel is in the active list, so only those are
// visited; in-place update is safe because 2 -> 3 keeps "neighbour >= 2" true.
// Also clears the voxel -> active map for the next step.
__global__ __launch_bounds__(GB) void k_mushy_to_solid(Inst* I, const int* astart, int ninst) {
  int ii = find_inst(astart, ninst, blockIdx.x);
  const Inst d = I[ii];  // a copy: stores in the loops cannot alias it
  int a = (blockIdx.x - astart[ii]) * GB + threadIdx.x;
  if (d.ls->pending || a >= d.ls->Na || a >= d.acap) return;
  int v = d.aVox[a];
  d.amap[v] = -1;
  if (d.aVS[a] != 2) return;
  int i, j, k;
  decode(d, v, i, j, k);
  for (int dz = -1; dz <= 1; ++dz) {
    int z = k + dz;
    if (z < 0 || z >= d.p.nz) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      int y = j + dy;
      if (y < 0 || y >= d.p.ny) continue;
      const uint8_t* row = d.vs + (size_t)z * d.nxy + (size_t)y * d.p.nx;
      for (int dx = -1; dx <= 1; ++dx) {
        int x = i + dx;
        if (x < 0 || x >= d.p.nx || (dx == 0 && dy == 0 && dz == 0)) continue;
        if ((row[x] & 3) < 2) return;
      }
    }
  }
  d.vs[v] = (d.vs[v] & 0xFC) | 3;
  d.ext[v] = 0.0;
}

// ---- per-layer helpers ---------------------------------------------------------------
// After a host upload of vState: the "T >= tL" flag (bit 7) follows the device's current
// temperature field, so SetLiquid3 still sees the caller's melt pool.
// Fresh output records inside a captured step (a memset node is not used in graphs).
__global__ void k_clear_outs(unsigned* o, int nwords) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nwords; q += gridDim.x * blockDim.x) o[q] = 0u;
}
__global__ void k_rehot(uint8_t* vs, const double* T, int N, double tL) {
  size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (size_t)N) return;
  vs[v] = (uint8_t)((vs[v] & 0x7f) | (T[v] >= tL ? 0x80 : 0));
}
__global__ void k_grain_presence(const int* gid, int N, int* present, int ngrain) {
  size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (size_t)N) return;
  int g = gid[v];
  if (g > 0 && g <= ngrain && present[g - 1] == 0) present[g - 1] = 1;
}
__global__ void k_remap(int* gid, int N, const int* remap, int ngrain) {
  size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (size_t)N) return;
  int g = gid[v];
  if (g > 0 && g <= ngrain) gid[v] = remap[g - 1];
}
// Brute-force Voronoi labelling of a z-slab: each voxel takes the first site (in draw
// order) at the smallest FP64 squared distance, ((ax*ax + ay*ay) + az*az) with no
// contraction, the order the reference evaluates it in.
//   plate (S/BasePlate.C:74-78, 104-117): voxel centres ((i + .5) dx, ...), planes
//     0..nz-1; the first site is taken unconditionally (start at +inf, strict <).
//   powder layer (S/VoxelsCA.C:920-933): voxel corners (i dx, j dy, (z0 + k) dz), planes
//     z0..z0+nz-1 (planes above the domain included, as the reference labels them);
//     best distance starts at 1e6, strict <.
// Sites are staged through LDS in chunks; every lane of a wave reads the same site
// (an LDS broadcast).  Plate mode writes the base-plate fields of S/VoxelsCA.C:125-161
// directly; layer mode writes the label and marks the site used.
#define VOR_CHUNK 512
template <bool LAYER>
__global__ __launch_bounds__(256) void k_voronoi(Inst* I, int ii, int z0, int nz, const double* __restrict__ xs,
                                                 int ng, int* __restrict__ label, int* __restrict__ used) {
  __shared__ double sx[VOR_CHUNK], sy[VOR_CHUNK], sz[VOR_CHUNK];
  const Inst& d = I[ii];
  const long long n = (long long)d.nxy * nz;
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i = 0, j = 0, k = 0;
  if (v < n) {
    k = (int)(v / d.nxy);
    const int r = (int)(v - (long long)k * d.nxy);
    j = r / d.p.nx;
    i = r - j * d.p.nx;
  }
  double x, y, z;
  if (LAYER) {
    x = i * d.p.dx;
    y = j * d.p.dy;
    z = (z0 + k) * d.p.dz;
  } else {
    x = ((double)i + .5) * d.p.dx;
    y = ((double)j + .5) * d.p.dy;
    z = ((double)(z0 + k) + .5) * d.p.dz;
  }
  double bd = LAYER ? 1e6 : __builtin_huge_val();
  int best = 0;
  for (int c0 = 0; c0 < ng; c0 += VOR_CHUNK) {
    const int m = min(VOR_CHUNK, ng - c0);
    __syncthreads();
    for (int q = threadIdx.x; q < m; q += blockDim.x) {
      sx[q] = xs[3 * (c0 + q)];
      sy[q] = xs[3 * (c0 + q) + 1];
      sz[q] = xs[3 * (c0 + q) + 2];
    }
    __syncthreads();
#pragma unroll 4
    for (int q = 0; q < m; ++q) {
      const double ax = x - sx[q], ay = y - sy[q], az = z - sz[q];
      const double dd = ax * ax + ay * ay + az * az;
      const bool lt = dd < bd;
      bd = lt ? dd : bd;
      best = lt ? c0 + q : best;
    }
  }
  if (v >= n) return;
  if (LAYER) {
    label[v] = best;
    used[best] = 1;
  } else {
    d.gid[v] = best + 1;
    d.vs[v] = (v < (long long)d.nxy * (nz - 1)) ? 3 : 2;
    d.cx[v] = x;
    d.cy[v] = y;
    d.cz[v] = z;
  }
}

// AddLayer1's continuous renumbering (S/VoxelsCA.C:935-947): rank[s] = number of used
// sites before s; the total goes to *n_used.  One workgroup.
__global__ __launch_bounds__(1024) void k_site_rank(const int* __restrict__ used, int ng, int* __restrict__ rank,
                                                    int* __restrict__ n_used) {
  __shared__ int wsum[16];
  __shared__ int carry;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < ng; c0 += 1024) {
    const int s = c0 + threadIdx.x;
    const int u = (s < ng && used[s]) ? 1 : 0;
    int inc = u;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(inc, o);
      inc += lane >= o ? t : 0;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int before = carry;
    for (int q = 0; q < w; ++q) before += wsum[q];
    if (s < ng) rank[s] = before + inc - u;
    __syncthreads();
    if (threadIdx.x == 1023) carry = before + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) *n_used = carry;
}

// AddLayer1's voxel writes (S/VoxelsCA.C:952-981): the new layer takes the renumbered
// Voronoi labels as solid grains gbase + 1.., everything at or above the laser plane is
// emptied.
__global__ void k_write_layer(Inst* I, int ii, int iz1, int ilaserLoc, int gbase, const int* __restrict__ label,
                              const int* __restrict__ rank) {
  const Inst& d = I[ii];
  size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (size_t)d.N) return;
  int i, j, k;
  decode(d, (int)v, i, j, k);
  if (k >= iz1 && k < ilaserLoc) {
    int jind = d.nxy * (k - iz1) + d.p.nx * j + i;
    d.gid[v] = gbase + rank[label[jind]] + 1;
    d.vs[v] = 3;
    d.cx[v] = ((double)i + .5) * d.p.dx;
    d.cy[v] = ((double)j + .5) * d.p.dy;
    d.cz[v] = ((double)k + .5) * d.p.dz;
  }
  if ((long long)v >= (long long)d.nxy * ilaserLoc) {
    d.vs[v] = 0;
    d.gid[v] = 0;
  }
}

// -------------------------------------------------------------------------------------
// host side of the device library
struct amcd_batch {
  int n = 0;
  int dev = 0;
  hipStream_t stream = nullptr;
  std::vector<Inst> h;      // host mirror of the descriptors
  Inst* d_inst = nullptr;
  int *d_bstartG = nullptr, *d_bstartZ = nullptr, *d_astart = nullptr, *d_bstartS = nullptr, *d_bstartT = nullptr,
      *d_lstart = nullptr, *d_cstart = nullptr, *d_bstartF = nullptr;
  std::vector<int> bstartG, bstartZ, astart, bstartS, bstartT, lstart, cstart, bstartF;
  amcd_step* d_steps = nullptr;
  amcd_out* d_outs = nullptr;
  amcd_step* h_steps = nullptr;  // pinned
  amcd_out* h_outs = nullptr;    // pinned
  std::vector<void*> allocs;
  hipEvent_t ev[16];
  double last_ms[16] = {0};
  bool timing = false;
  int tier = 1;  // capture-loop capacity tier (see launch_event_loop)
  int min_tier = 0;  // lowest tier a step starts in (AMCAFE_CAPTURE_TIER_MIN; tests use it to
                     // run small decks in the larger tiers; 5 = memory tier, bitmap in memory)
  // nucleation pools of every instance in one device region (instance i's at pool_off[i]
  // entries of 9 doubles), filled from one pinned staging copy per step
  double* d_pool = nullptr;
  double* h_pool = nullptr;  // pinned
  uint8_t* d_mark = nullptr;  // SetBuild's grain marks (mark_cap bytes)
  size_t mark_cap = 0;
  std::vector<size_t> pool_off;
  hipEvent_t pool_ev = nullptr;  // the last staging copy (the staging area is rewritten after it)
  // The driver's step as HIP graphs (AMCAFE_GRAPH=0 launches the stages one by one): the
  // stages up to the capture loop, per capture tier, and the finishing pair; each is captured
  // again when the block counts it was captured with change (capacity growth)
  bool graphs = true;
  hipGraphExec_t gstep[6] = {};
  long long gstep_key[6] = {};
  hipGraphExec_t gfin = nullptr;
  long long gfin_key = -1;
  std::vector<int> astart_up, lstart_up;  // the block lists the device holds
  // Active-list chunking: blocks per instance of the candidate-list pass, picked by timing
  // the pass (which only rebuilds lists from their inputs) at each candidate, again each time
  // the largest growing list has doubled (AMCAFE_AUTOTUNE=0 keeps the default)
  int list_cap = 2048;
  bool autotune = true, tuned = false;
  bool tune_pending = false;  // the next step tunes (between its list pass and its capture loop)
  int tuned_at = 0;  // largest growing list at the last tuning
  int n_tunings = 0;
  float tune_ms[8] = {0};
};

static thread_local std::string g_dev_err;
#define HIPT(x)                                                                        \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      g_dev_err = std::string(#x) + ": " + hipGetErrorString(e_);                      \
      return -1;                                                                       \
    }                                                                                  \
  } while (0)

// Every device array has 256 bytes beyond its length: the list-block reads of the
// capture loop run up to four entries past the last slot's list.
template <class T>
static int dalloc(amcd_batch* b, T** p, size_t n) {
  void* q = nullptr;
  HIPT(hipMalloc(&q, n * sizeof(T) + 256));
  b->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

static int alloc_active(amcd_batch* b, Inst& d, int acap) {
  d.acap = acap;
  size_t c = (size_t)acap;
  if (dalloc(b, &d.aVox, c) || dalloc(b, &d.aVS, c) || dalloc(b, &d.aG, c) || dalloc(b, &d.aT, c) ||
      dalloc(b, &d.aExt, c) || dalloc(b, &d.aCx, c) || dalloc(b, &d.aCy, c) || dalloc(b, &d.aCz, c) ||
      dalloc(b, &d.ruq, c) || dalloc(b, &d.aNbr, c * AMCD_NB) || dalloc(b, &d.aCold, c) ||
      dalloc(b, &d.aScan, c) || dalloc(b, &d.gAct, c) || dalloc(b, &d.gScan, c) || dalloc(b, &d.gExt, c) ||
      dalloc(b, &d.gVhat, c) || dalloc(b, &d.gN, c) || dalloc(b, &d.gDead, c) || dalloc(b, &d.gGrow, c) ||
      dalloc(b, &d.cD, c * AMCD_NB) || dalloc(b, &d.cF, c * AMCD_NB) || dalloc(b, &d.cR, c * AMCD_NB) ||
      dalloc(b, &d.cTgt, c * AMCD_NB) || dalloc(b, &d.cK, c * AMCD_NB) || dalloc(b, &d.aRN, c * AMCD_NB) ||
      dalloc(b, &d.gR, c * 9) || dalloc(b, &d.gC, c * 3) || dalloc(b, &d.gV, c) || dalloc(b, &d.gG, c) ||
      dalloc(b, &d.cV, c * AMCD_NB) || dalloc(b, &d.gLH, c) || dalloc(b, &d.gHd, c) ||
      dalloc(b, &d.cRC, c * AMCD_NB) || dalloc(b, &d.gdead, (c + 31) / 32))
    return -1;
  return 0;
}

static int push_inst(amcd_batch* b) {
  HIPT(hipMemcpyAsync(b->d_inst, b->h.data(), sizeof(Inst) * b->n, hipMemcpyHostToDevice, b->stream));
  return 0;
}

extern "C" {

const char* amcd_last_error() { return g_dev_err.c_str(); }

int amcd_device_name(char* buf, int len) {
  hipDeviceProp_t pr;
  int dev = 0;
  HIPT(hipGetDevice(&dev));
  HIPT(hipGetDeviceProperties(&pr, dev));
  snprintf(buf, len, "%s (%s, %d CUs)", pr.name, pr.gcnArchName, pr.multiProcessorCount);
  return 0;
}

int amcd_create(const amcd_params* p, int n, amcd_batch** out, char* err, int errlen) {
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    snprintf(err, errlen, "no HIP device");
    return -1;
  }
  amcd_batch* b = new amcd_batch();
  b->n = n;
  b->h.resize(n);
  if (const char* g = getenv("AMCAFE_GRAPH")) b->graphs = atoi(g) != 0;
  if (const char* t = getenv("AMCAFE_AUTOTUNE")) b->autotune = atoi(t) != 0;
  if (const char* mt = getenv("AMCAFE_CAPTURE_TIER_MIN")) {
    b->min_tier = std::max(0, std::min(5, atoi(mt)));  // 5: the memory tier with the bitmap in memory
    b->tier = std::max(b->tier, b->min_tier);
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    snprintf(err, errlen, "hipStreamCreate failed");
    delete b;
    return -1;
  }
  for (int q = 0; q < 16; ++q)
    if (hipEventCreate(&b->ev[q]) != hipSuccess) {
      snprintf(err, errlen, "hipEventCreate failed");
      amcd_destroy(b);
      return -1;
    }
  b->bstartG.assign(n + 1, 0);
  b->bstartS.assign(n + 1, 0);
  b->bstartT.assign(n + 1, 0);
  b->lstart.assign(n + 1, 0);
  b->cstart.assign(n + 1, 0);
  b->bstartZ.assign(n + 1, 0);
  b->bstartF.assign(n + 1, 0);
  b->astart.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    Inst& d = b->h[i];
    memset(&d, 0, sizeof(Inst));
    d.p = p[i];
    long long N = (long long)p[i].nx * p[i].ny * p[i].nz;
    if (N <= 0 || N > 0x7fffffffLL - 4096) {
      snprintf(err, errlen, "domain %d: voxel count %lld out of range", i, N);
      amcd_destroy(b);
      return -1;
    }
    d.N = (int)N;
    d.nxy = p[i].nx * p[i].ny;
    d.nblkZ = (d.N + ZB - 1) / ZB;
    d.nblkG = (d.N + GB - 1) / GB;
    d.nblkS = (d.N + SB - 1) / SB;
    d.nblkT = (d.N + GB * TV - 1) / (GB * TV);
    size_t Nn = (size_t)d.N;
    if (dalloc(b, &d.vs, Nn) || dalloc(b, &d.gid, Nn) || dalloc(b, &d.ext, Nn) || dalloc(b, &d.cx, Nn) ||
        dalloc(b, &d.cy, Nn) || dalloc(b, &d.cz, Nn) || dalloc(b, &d.T, Nn) || dalloc(b, &d.bj, Nn) || dalloc(b, &d.build, Nn) ||
        dalloc(b, &d.amap, Nn) || dalloc(b, &d.bcA, (size_t)d.nblkZ) || dalloc(b, &d.bcM, (size_t)d.nblkZ) ||
        dalloc(b, &d.chA, (size_t)d.nblkZ / 1024 + 1) || dalloc(b, &d.chM, (size_t)d.nblkZ / 1024 + 1) ||
        dalloc(b, &d.zliq, (size_t)p[i].nz) || dalloc(b, &d.ls, 1)) {
      snprintf(err, errlen, "%s", g_dev_err.c_str());
      amcd_destroy(b);
      return -1;
    }
    hipError_t me = hipSuccess;
    auto zero = [&](void* q, int val, size_t bytes) {
      const hipError_t e = hipMemsetAsync(q, val, bytes, b->stream);
      if (me == hipSuccess) me = e;
    };
    zero(d.vs, 0, Nn);
    zero(d.gid, 0, Nn * 4);
    zero(d.ext, 0, Nn * 8);
    zero(d.cx, 0, Nn * 8);
    zero(d.cy, 0, Nn * 8);
    zero(d.cz, 0, Nn * 8);
    zero(d.T, 0, Nn * 8);
    zero(d.bj, 0, Nn);
    zero(d.build, 0, Nn);
    zero(d.zliq, 0xff, (size_t)p[i].nz * 4);  // unknown: every plane may hold liquid
    zero(d.amap, 0xff, Nn * 4);
    zero(d.ls, 0, sizeof(LoopState));
    if (me != hipSuccess) {
      snprintf(err, errlen, "hipMemsetAsync: %s", hipGetErrorString(me));
      amcd_destroy(b);
      return -1;
    }
    int acap = (int)std::min<long long>(N, std::max<long long>(1 << 16, N / 8));
    if (const char* ac = getenv("AMCAFE_ACTIVE_CAP_INIT"))  // tests: start small to exercise the growth path
      acap = (int)std::min<long long>(N, std::max(64, atoi(ac)));
    if (alloc_active(b, d, acap)) {
      snprintf(err, errlen, "%s", g_dev_err.c_str());
      amcd_destroy(b);
      return -1;
    }
    d.gcap = 0;
    d.rot = nullptr;
    d.poolcap = 0;
    d.pool = nullptr;
    b->bstartG[i + 1] = b->bstartG[i] + d.nblkG;
    b->bstartS[i + 1] = b->bstartS[i] + d.nblkS;
    b->bstartT[i + 1] = b->bstartT[i] + d.nblkT;
    b->cstart[i + 1] = b->cstart[i] + (d.nblkZ + 1023) / 1024;
    b->bstartZ[i + 1] = b->bstartZ[i] + d.nblkZ;
    b->bstartF[i + 1] = b->bstartF[i] + (d.nblkZ + 3) / 4;
  }
  if (dalloc(b, &b->d_inst, (size_t)n) || dalloc(b, &b->d_bstartG, (size_t)n + 1) ||
      dalloc(b, &b->d_bstartS, (size_t)n + 1) || dalloc(b, &b->d_bstartT, (size_t)n + 1) ||
      dalloc(b, &b->d_lstart, (size_t)n + 1) || dalloc(b, &b->d_cstart, (size_t)n + 1) ||
      dalloc(b, &b->d_bstartZ, (size_t)n + 1) || dalloc(b, &b->d_astart, (size_t)n + 1) ||
      dalloc(b, &b->d_bstartF, (size_t)n + 1) ||
      dalloc(b, &b->d_steps, (size_t)n) || dalloc(b, &b->d_outs, (size_t)n)) {
    snprintf(err, errlen, "%s", g_dev_err.c_str());
    amcd_destroy(b);
    return -1;
  }
  if (hipHostMalloc((void**)&b->h_steps, sizeof(amcd_step) * n, 0) != hipSuccess ||
      hipHostMalloc((void**)&b->h_outs, sizeof(amcd_out) * n, 0) != hipSuccess) {
    snprintf(err, errlen, "hipHostMalloc failed");
    amcd_destroy(b);
    return -1;
  }
  const std::pair<int*, const std::vector<int>*> starts[] = {{b->d_bstartG, &b->bstartG}, {b->d_bstartS, &b->bstartS},
                                                             {b->d_bstartT, &b->bstartT}, {b->d_cstart, &b->cstart},
                                                             {b->d_bstartZ, &b->bstartZ}, {b->d_bstartF, &b->bstartF}};
  for (const auto& st : starts)
    if (hipMemcpyAsync(st.first, st.second->data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, b->stream) !=
        hipSuccess) {
      snprintf(err, errlen, "hipMemcpyAsync of block starts failed");
      amcd_destroy(b);
      return -1;
    }
  if (push_inst(b) || hipStreamSynchronize(b->stream) != hipSuccess) {
    snprintf(err, errlen, "%s", g_dev_err.c_str());
    amcd_destroy(b);
    return -1;
  }
  *out = b;
  return 0;
}

void amcd_destroy(amcd_batch* b) {  // best effort: errors while tearing down are not reported
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (void* q : b->allocs) (void)hipFree(q);
  for (int q = 0; q < 16; ++q)
    if (b->ev[q]) (void)hipEventDestroy(b->ev[q]);
  if (b->h_steps) (void)hipHostFree(b->h_steps);
  if (b->h_outs) (void)hipHostFree(b->h_outs);
  for (auto& g : b->gstep)
    if (g) (void)hipGraphExecDestroy(g);
  if (b->gfin) (void)hipGraphExecDestroy(b->gfin);
  if (b->h_pool) (void)hipHostFree(b->h_pool);
  if (b->d_pool) (void)hipFree(b->d_pool);
  if (b->d_mark) (void)hipFree(b->d_mark);
  if (b->pool_ev) (void)hipEventDestroy(b->pool_ev);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

void amcd_set_timing(amcd_batch* b, int on) { b->timing = on != 0; }

int amcd_upload(amcd_batch* b, int i, const int* gid, const int* vstate, const double* ext, const double* cen,
                const int* bj) {
  if (i < 0 || i >= b->n) return -1;
  Inst& d = b->h[i];
  HIPT(hipStreamSynchronize(b->stream));
  size_t N = d.N;
  std::vector<uint8_t> tmp(N);
  if (vstate) {
    for (size_t v = 0; v < N; ++v) tmp[v] = (uint8_t)(vstate[v] & 3);
    HIPT(hipMemcpy(d.vs, tmp.data(), N, hipMemcpyHostToDevice));
    HIPT(hipMemset(d.zliq, 0xff, (size_t)d.p.nz * 4));
    HIPT(hipMemset(&d.ls->tvalid, 0, sizeof(int)));
    hipLaunchKernelGGL(k_rehot, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, b->stream, d.vs, d.T, d.N, d.p.tL);
    HIPT(hipGetLastError());
    HIPT(hipStreamSynchronize(b->stream));
  }
  if (bj) {
    for (size_t v = 0; v < N; ++v) tmp[v] = (uint8_t)(bj[v] ? 1 : 0);
    HIPT(hipMemcpy(d.bj, tmp.data(), N, hipMemcpyHostToDevice));
  }
  if (gid) HIPT(hipMemcpy(d.gid, gid, N * 4, hipMemcpyHostToDevice));
  if (ext) HIPT(hipMemcpy(d.ext, ext, N * 8, hipMemcpyHostToDevice));
  if (cen) {
    std::vector<double> c(N);
    double* dst[3] = {d.cx, d.cy, d.cz};
    for (int q = 0; q < 3; ++q) {
      for (size_t v = 0; v < N; ++v) c[v] = cen[3 * v + q];
      HIPT(hipMemcpy(dst[q], c.data(), N * 8, hipMemcpyHostToDevice));
    }
  }
  return 0;
}

int amcd_download(amcd_batch* b, int i, int* gid, int* vstate, double* ext, double* cen, int* bj, double* T) {
  if (i < 0 || i >= b->n) return -1;
  Inst& d = b->h[i];
  HIPT(hipStreamSynchronize(b->stream));
  size_t N = d.N;
  std::vector<uint8_t> tmp(N);
  if (vstate) {
    HIPT(hipMemcpy(tmp.data(), d.vs, N, hipMemcpyDeviceToHost));
    for (size_t v = 0; v < N; ++v) vstate[v] = tmp[v] & 3;
  }
  if (bj) {
    HIPT(hipMemcpy(tmp.data(), d.bj, N, hipMemcpyDeviceToHost));
    for (size_t v = 0; v < N; ++v) bj[v] = tmp[v];
  }
  if (gid) HIPT(hipMemcpy(gid, d.gid, N * 4, hipMemcpyDeviceToHost));
  if (ext) HIPT(hipMemcpy(ext, d.ext, N * 8, hipMemcpyDeviceToHost));
  if (T) HIPT(hipMemcpy(T, d.T, N * 8, hipMemcpyDeviceToHost));
  if (cen) {
    std::vector<double> c(N);
    double* src[3] = {d.cx, d.cy, d.cz};
    for (int q = 0; q < 3; ++q) {
      HIPT(hipMemcpy(c.data(), src[q], N * 8, hipMemcpyDeviceToHost));
      for (size_t v = 0; v < N; ++v) cen[3 * v + q] = c[v];
    }
  }
  return 0;
}

static int ensure_grains(amcd_batch* b, Inst& d, int need) {
  if (need <= d.gcap) return 0;
  int ncap = std::max(need, std::max(1024, d.gcap * 2));
  double* nr = nullptr;
  HIPT(hipMalloc(&nr, sizeof(double) * 9 * (size_t)ncap));
  if (d.rot) {
    HIPT(hipMemcpyAsync(nr, d.rot, sizeof(double) * 9 * (size_t)d.gcap, hipMemcpyDeviceToDevice, b->stream));
    HIPT(hipStreamSynchronize(b->stream));
    for (auto& q : b->allocs)
      if (q == d.rot) q = nr;
    HIPT(hipFree(d.rot));
  } else {
    b->allocs.push_back(nr);
  }
  d.rot = nr;
  d.gcap = ncap;
  return push_inst(b);
}

int amcd_set_rot(amcd_batch* b, int i, const double* rot9, int ngrain) {
  Inst& d = b->h[i];
  if (ensure_grains(b, d, ngrain + 256)) return -1;
  if (ngrain > 0)
    HIPT(hipMemcpyAsync(d.rot, rot9, sizeof(double) * 9 * (size_t)ngrain, hipMemcpyHostToDevice, b->stream));
  HIPT(hipStreamSynchronize(b->stream));
  return 0;
}

int amcd_get_rot(amcd_batch* b, int i, double* rot9, int ngrain) {
  Inst& d = b->h[i];
  if (ngrain > d.gcap) return -1;
  HIPT(hipStreamSynchronize(b->stream));
  HIPT(hipMemcpy(rot9, d.rot, sizeof(double) * 9 * (size_t)ngrain, hipMemcpyDeviceToHost));
  return 0;
}

// Every instance's pool lives in one device region behind one pinned staging area: the
// pools of a step go to the device in one copy instead of one pageable copy per instance
// (each of which the host waits for).  A pool larger than its instance's capacity
// re-lays the region out for every instance (capacities at least double).
int amcd_set_pools(amcd_batch* b, const double* const* rot9, const int* n) {
  bool grow = b->d_pool == nullptr;
  for (int i = 0; i < b->n; ++i) grow = grow || n[i] > b->h[i].poolcap;
  if (grow) {
    HIPT(hipStreamSynchronize(b->stream));  // no launch reads the old pools any more
    std::vector<size_t> off(b->n + 1, 0);
    for (int i = 0; i < b->n; ++i) {
      Inst& d = b->h[i];
      int cap = std::max(64, d.poolcap);
      while (cap < n[i]) cap *= 2;
      d.poolcap = cap;
      off[i + 1] = off[i] + (size_t)cap;
    }
    double *dp = nullptr, *hp = nullptr;
    HIPT(hipMalloc(&dp, sizeof(double) * 9 * off[b->n]));
    if (hipHostMalloc((void**)&hp, sizeof(double) * 9 * off[b->n], 0) != hipSuccess) {
      (void)hipFree(dp);
      g_dev_err = "hipHostMalloc of the pool staging area failed";
      return -1;
    }
    if (b->d_pool) HIPT(hipFree(b->d_pool));
    if (b->h_pool) HIPT(hipHostFree(b->h_pool));
    b->d_pool = dp;
    b->h_pool = hp;
    b->pool_off = off;
    for (int i = 0; i < b->n; ++i) b->h[i].pool = dp + 9 * off[i];
    if (!b->pool_ev) HIPT(hipEventCreateWithFlags(&b->pool_ev, hipEventDisableTiming));
    if (push_inst(b)) return -1;
  } else {
    HIPT(hipEventSynchronize(b->pool_ev));  // the previous copy has read the staging area
  }
  size_t lo = (size_t)-1, hi = 0;
  for (int i = 0; i < b->n; ++i) {
    if (n[i] <= 0 || !rot9[i]) continue;
    const size_t o = b->pool_off[i];
    memcpy(b->h_pool + 9 * o, rot9[i], sizeof(double) * 9 * (size_t)n[i]);
    lo = std::min(lo, o);
    hi = std::max(hi, o + (size_t)n[i]);
  }
  if (hi > lo)
    HIPT(hipMemcpyAsync(b->d_pool + 9 * lo, b->h_pool + 9 * lo, sizeof(double) * 9 * (hi - lo),
                        hipMemcpyHostToDevice, b->stream));
  HIPT(hipEventRecord(b->pool_ev, b->stream));
  return 0;
}

int amcd_set_pool(amcd_batch* b, int i, const double* rot9, int n) {
  if (i < 0 || i >= b->n) return -1;
  std::vector<const double*> r(b->n, nullptr);
  std::vector<int> c(b->n, 0);
  r[i] = rot9;
  c[i] = n;
  return amcd_set_pools(b, r.data(), c.data());
}

// grain-table capacity for the coming step: current grains + pool size
int amcd_reserve_grains(amcd_batch* b, int i, int need) { return ensure_grains(b, b->h[i], need); }

static int put_steps(amcd_batch* b, const amcd_step* steps) {
  memcpy(b->h_steps, steps, sizeof(amcd_step) * b->n);
  HIPT(hipMemcpyAsync(b->d_steps, b->h_steps, sizeof(amcd_step) * b->n, hipMemcpyHostToDevice, b->stream));
  return 0;
}

static void tmark(amcd_batch* b, int q) {
  if (b->timing) (void)hipEventRecord(b->ev[q], b->stream);  // timing only
}

static int run_temperature(amcd_batch* b, int setliquid) {
  int nb = b->bstartT[b->n];
  hipLaunchKernelGGL(k_temperature, dim3(nb), dim3(GB), 0, b->stream, b->d_inst, b->d_bstartT, b->n, b->d_steps,
                     b->d_outs, setliquid);
  hipLaunchKernelGGL(k_temp_commit, dim3(b->n), dim3(64), 0, b->stream, b->d_inst, b->d_steps, 1);
  HIPT(hipGetLastError());
  return 0;
}
// the device T no longer follows the analytic layout (upload, new layer, EASM)
static int invalidate_T(amcd_batch* b, int i) {
  HIPT(hipMemsetAsync(&b->h[i].ls->tvalid, 0, sizeof(int), b->stream));
  return 0;
}

int amcd_temperature(amcd_batch* b, const amcd_step* steps) {
  if (put_steps(b, steps)) return -1;
  HIPT(hipMemsetAsync(b->d_outs, 0, sizeof(amcd_out) * b->n, b->stream));
  return run_temperature(b, 0);
}

static int read_outs(amcd_batch* b, amcd_out* outs);
int amcd_easm(amcd_batch* b, const amcd_step* steps, double Q, double layerT, amcd_out* outs,
              unsigned long long* n_evals) {
  if (put_steps(b, steps)) return -1;
  HIPT(hipMemsetAsync(b->d_outs, 0, sizeof(amcd_out) * b->n, b->stream));
  unsigned long long* dn = nullptr;
  HIPT(hipMalloc(&dn, sizeof(unsigned long long)));
  HIPT(hipMemsetAsync(dn, 0, sizeof(unsigned long long), b->stream));
  tmark(b, 0);
  hipLaunchKernelGGL(k_easm, dim3(b->bstartG[b->n]), dim3(GB), 0, b->stream, b->d_inst, b->d_bstartG, b->n,
                     b->d_steps, b->d_outs, Q, layerT, dn);
  hipLaunchKernelGGL(k_temp_commit, dim3(b->n), dim3(64), 0, b->stream, b->d_inst, b->d_steps, 0);
  HIPT(hipGetLastError());
  tmark(b, 1);
  unsigned long long hn = 0;
  HIPT(hipMemcpyAsync(&hn, dn, sizeof(hn), hipMemcpyDeviceToHost, b->stream));
  if (read_outs(b, outs)) return -1;
  HIPT(hipFree(dn));
  if (b->timing) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
    b->last_ms[0] = ms;
  }
  if (n_evals) *n_evals = hn;
  return 0;
}

__global__ void k_loop_rng(Inst* I, const amcd_step* S) {
  const Inst& d = I[blockIdx.x];
  LoopState* ls = d.ls;
  if (threadIdx.x != 0 || ls->done) return;
  unsigned int x0 = lcg_seed(30ull * (unsigned long long)S[blockIdx.x].tInd + d.p.seed1);
  ls->rng = mod_m((unsigned long long)x0 * lcg_pow(2ull * (unsigned long long)ls->Na));
  ls->draws = 2ll * ls->Na;
}

static int launch_update(amcd_batch* b, bool fused_temperature) {
  const int n = b->n;
  tmark(b, 0);
  if (fused_temperature) {
    if (run_temperature(b, 1)) return -1;
  } else {
    hipLaunchKernelGGL(k_setliquid, dim3(b->bstartG[n]), dim3(GB), 0, b->stream, b->d_inst, b->d_bstartG, n,
                       b->d_steps);
  }
  tmark(b, 1);
  hipLaunchKernelGGL(k_solid_to_mushy, dim3(b->bstartS[n]), dim3(GB), 0, b->stream, b->d_inst, b->d_bstartS, n,
                     b->d_steps);
  tmark(b, 2);
  hipLaunchKernelGGL(k_zliq_clear, dim3(n), dim3(256), 0, b->stream, b->d_inst, b->d_steps);
  hipLaunchKernelGGL(k_active_flags, dim3(b->bstartF[n]), dim3(256), 0, b->stream, b->d_inst, b->d_bstartF, n,
                     b->d_steps);
  hipLaunchKernelGGL(k_scan_chunks, dim3(b->cstart[n]), dim3(1024), 0, b->stream, b->d_inst, b->d_cstart, n);
  hipLaunchKernelGGL(k_scan, dim3(n), dim3(1024), 0, b->stream, b->d_inst, b->d_steps, b->d_outs);
  hipLaunchKernelGGL(k_compact, dim3(b->bstartF[n]), dim3(256), 0, b->stream, b->d_inst, b->d_bstartF, n,
                     b->d_steps);
  hipLaunchKernelGGL(k_loop_rng, dim3(n), dim3(64), 0, b->stream, b->d_inst, b->d_steps);
  tmark(b, 3);
  HIPT(hipGetLastError());
  return 0;
}

// Capacity tiers of the capture loop: threads per workgroup x growing cells per thread.
static int launch_event_loop_wide(amcd_batch* b);
static int launch_event_loop(amcd_batch* b) {
  const int n = b->n;
  if (b->tier >= 5) return launch_event_loop_wide(b);
  switch (b->tier) {
    case 0:
      hipLaunchKernelGGL((k_event_loop<256, 4>), dim3(n), dim3(256), 0, b->stream, b->d_inst, b->d_steps, b->d_outs);
      break;
    case 1:
      hipLaunchKernelGGL((k_event_loop<512, 4>), dim3(n), dim3(512), 0, b->stream, b->d_inst, b->d_steps, b->d_outs);
      break;
    case 2:
      hipLaunchKernelGGL((k_event_loop<512, 8>), dim3(n), dim3(512), 0, b->stream, b->d_inst, b->d_steps,
                         b->d_outs);
      break;
    case 3:
      hipLaunchKernelGGL((k_event_loop<512, 16>), dim3(n), dim3(512), 0, b->stream, b->d_inst, b->d_steps,
                         b->d_outs);
      break;
    default:
      hipLaunchKernelGGL((k_event_loop<512, 0>), dim3(n), dim3(512), 0, b->stream, b->d_inst, b->d_steps,
                         b->d_outs);
      break;
  }
  HIPT(hipGetLastError());
  return 0;
}
// the memory-resident tier with the liveness bitmap in memory (interfaces beyond AMCD_DEAD_BITS)
static int launch_event_loop_wide(amcd_batch* b) {
  hipLaunchKernelGGL((k_event_loop<512, 0, true>), dim3(b->n), dim3(512), 0, b->stream, b->d_inst, b->d_steps,
                     b->d_outs);
  HIPT(hipGetLastError());
  return 0;
}
static const int kNTiers = 5;  // the last one keeps the cells in memory: no limit below the active capacity
static const int kTierCap[kNTiers] = {192 * 4, 448 * 4, 448 * 8, 448 * 16, 0x7fffffff};

static int read_outs(amcd_batch* b, amcd_out* outs) {
  HIPT(hipMemcpyAsync(b->h_outs, b->d_outs, sizeof(amcd_out) * b->n, hipMemcpyDeviceToHost, b->stream));
  HIPT(hipStreamSynchronize(b->stream));
  memcpy(outs, b->h_outs, sizeof(amcd_out) * b->n);
  return 0;
}

// The active-cell block lists, sized by the capacities (idle threads exit, so no host round
// trip is needed mid-step), uploaded when a capacity has changed.
static int prep_active(amcd_batch* b) {
  const int n = b->n;
  b->astart[0] = 0;
  for (int i = 0; i < n; ++i) b->astart[i + 1] = b->astart[i] + (b->h[i].acap + GB - 1) / GB;
  b->lstart[0] = 0;
  for (int i = 0; i < n; ++i) b->lstart[i + 1] = b->lstart[i] + std::min(b->list_cap, (b->h[i].acap + 3) / 4);
  if (b->astart != b->astart_up) {
    HIPT(hipMemcpyAsync(b->d_astart, b->astart.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, b->stream));
    b->astart_up = b->astart;
  }
  if (b->lstart != b->lstart_up) {
    HIPT(hipMemcpyAsync(b->d_lstart, b->lstart.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, b->stream));
    b->lstart_up = b->lstart;
  }
  // the vectors stay unchanged until the next capacity change, so the copies may read them late
  return 0;
}
static int autotune_lists(amcd_batch* b);
static int launch_active_kernels(amcd_batch* b) {
  const int n = b->n;
  hipLaunchKernelGGL(k_active_setup, dim3(b->astart[n]), dim3(GB), 0, b->stream, b->d_inst, b->d_astart, n,
                     b->d_outs);
  hipLaunchKernelGGL(k_grow_lists, dim3(b->lstart[n]), dim3(GB), 0, b->stream, b->d_inst, b->d_lstart, n);
  if (b->tune_pending && !b->timing) {  // stage by stage only: the graphs never tune
    b->tune_pending = false;
    if (autotune_lists(b)) return -1;
  }
  tmark(b, 4);
  if (launch_event_loop(b)) return -1;
  tmark(b, 5);
  HIPT(hipGetLastError());
  return 0;
}
static int launch_active(amcd_batch* b) { return prep_active(b) || launch_active_kernels(b); }

static int launch_finish(amcd_batch* b) {
  const int n = b->n;
  hipLaunchKernelGGL(k_scatter, dim3(b->astart[n]), dim3(GB), 0, b->stream, b->d_inst, b->d_astart, n);
  hipLaunchKernelGGL(k_mushy_to_solid, dim3(b->astart[n]), dim3(GB), 0, b->stream, b->d_inst, b->d_astart, n);
  tmark(b, 6);
  HIPT(hipGetLastError());
  return 0;
}

static void collect_timing(amcd_batch* b) {
  if (!b->timing) return;
  (void)hipEventSynchronize(b->ev[6]);
  for (int q = 0; q < 6; ++q) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, b->ev[q], b->ev[q + 1]);
    b->last_ms[q] = ms;
  }
}

int amcd_read_outs(amcd_batch* b, amcd_out* outs) { return read_outs(b, outs); }

// capture-loop phase timers (zeros unless built with -DAMCD_PHASE_PROFILE); reading resets them
int amcd_phase_profile(unsigned long long* out32) {
#ifdef AMCD_PHASE_PROFILE
  HIPT(hipDeviceSynchronize());
  HIPT(hipMemcpyFromSymbol(out32, HIP_SYMBOL(g_phase), sizeof(unsigned long long) * 32));
  unsigned long long z[32] = {0};
  HIPT(hipMemcpyToSymbol(HIP_SYMBOL(g_phase), z, sizeof(z)));
#else
  for (int q = 0; q < 32; ++q) out32[q] = 0;
#endif
  return 0;
}

int amcd_last_timings(amcd_batch* b, double* ms, int n) {
  for (int q = 0; q < n && q < 6; ++q) ms[q] = b->last_ms[q];
  return 0;
}

// growth of the active capacity after an overflow; the step is then repeated (the
// stages before the compaction are idempotent for the same temperature field)
static int grow_active(amcd_batch* b, const amcd_out* outs) {
  bool any = false;
  for (int i = 0; i < b->n; ++i) {
    if (!outs[i].overflow) continue;
    Inst& d = b->h[i];
    int need = outs[i].n_active;
    int ncap = (int)std::min<long long>(d.N, std::max<long long>((long long)need * 3 / 2, (long long)d.acap * 2));
    // old arrays stay allocated until destroy; they are simply no longer referenced
    if (alloc_active(b, d, ncap)) return -1;
    any = true;
  }
  if (any && push_inst(b)) return -1;
  return 0;
}

// An instance whose interface is larger than the LDS liveness bitmap returns from the
// launch before its loop starts (status -3); the loop then runs for it in the memory-
// resident tier with the bitmap in memory (instances already done return at once).
static int widen(amcd_batch* b, amcd_out* outs) {
  bool need = false;
  for (int i = 0; i < b->n; ++i) need = need || outs[i].status == -3;
  if (!need) return 0;
  if (launch_event_loop_wide(b) || read_outs(b, outs)) return -1;
  return 0;
}

// Relaunch (resume) the capture loop with a larger tier while some instance has more
// growing cells than the current tier holds.
static int regrow_tier(amcd_batch* b, amcd_out* outs) {
  for (;;) {
    bool need = false;
    for (int i = 0; i < b->n; ++i) need = need || outs[i].status == 2;
    if (!need) return 0;
    if (b->tier >= kNTiers - 1) {
      int worst = 0;
      for (int i = 0; i < b->n; ++i)
        if (outs[i].status == 2 && outs[i].n_slots >= outs[worst].n_slots) worst = i;
      g_dev_err = "instance " + std::to_string(worst) + ": " + std::to_string(outs[worst].n_slots) +
                  " growing cells and the memory-resident capture tier still asks for more";
      return -1;
    }
    b->tier++;
    if (launch_event_loop(b) || read_outs(b, outs)) return -1;
  }
}
// smallest tier holding 1/32 more than the largest growing list of the step just done (a
// step that outgrows it resumes in the next tier)
static void pick_tier(amcd_batch* b, const amcd_out* outs) {
  int mx = 0;
  for (int i = 0; i < b->n; ++i) mx = std::max(mx, outs[i].n_slots);
  int t = 0;
  while (t < kNTiers - 1 && kTierCap[t] < mx + mx / 32) ++t;
  b->tier = std::max(t, b->min_tier);
}

// Capture of the step's stages into an executable graph: with `step` the output clear, the
// temperature and state passes, the active-list and candidate-list passes and the capture loop
// of the current tier; otherwise the scatter and mushy->solid pair.  (The step record's upload
// and the outputs' download stay plain stream copies around the graph: as graph nodes they
// made the step slower.)
static int capture_graph(amcd_batch* b, bool step, bool fused, hipGraphExec_t* out) {
  static_assert(sizeof(amcd_out) % 4 == 0, "outputs cleared as words");
  const int n = b->n;
  hipGraph_t g = nullptr;
  HIPT(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  if (step) {
    const int nw = (int)(sizeof(amcd_out) / 4) * n;
    hipLaunchKernelGGL(k_clear_outs, dim3((nw + 255) / 256), dim3(256), 0, b->stream, (unsigned*)b->d_outs, nw);
    rc = launch_update(b, fused) || launch_active_kernels(b);
  } else {
    rc = launch_finish(b);
  }
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (g_dev_err.empty() || e != hipSuccess) g_dev_err = std::string("step graph capture: ") + hipGetErrorString(e);
    return -1;
  }
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    g_dev_err = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei);
    return -1;
  }
  return 0;
}
static std::atomic<long long> g_graph_launches{0};
long long amcd_graph_launches(void) { return g_graph_launches.load(); }
// graph of the step's stages for the current tier, captured again when the block counts change
static int launch_step_graph(amcd_batch* b) {
  const int n = b->n, t = std::min(b->tier, 5);
  const long long key = ((long long)b->astart[n] << 32) | (unsigned)b->lstart[n];
  hipGraphExec_t& g = b->gstep[t];
  if (g && b->gstep_key[t] != key) {
    HIPT(hipGraphExecDestroy(g));
    g = nullptr;
  }
  if (!g) {
    if (capture_graph(b, true, true, &g)) return -1;
    b->gstep_key[t] = key;
  }
  HIPT(hipGraphLaunch(g, b->stream));
  g_graph_launches++;
  return 0;
}
static int launch_finish_graph(amcd_batch* b) {
  const long long key = b->astart[b->n];
  if (b->gfin && b->gfin_key != key) {
    HIPT(hipGraphExecDestroy(b->gfin));
    b->gfin = nullptr;
  }
  if (!b->gfin) {
    if (capture_graph(b, false, false, &b->gfin)) return -1;
    b->gfin_key = key;
  }
  HIPT(hipGraphLaunch(b->gfin, b->stream));
  g_graph_launches++;
  return 0;
}

// Candidate chunkings of the candidate-list pass (blocks per instance).  The pass rebuilds every
// live slot's list from inputs it does not write, so running it again right after itself, before
// the capture loop, changes nothing; each candidate is timed three times after a warm-up on the
// step's interface.  The device's block list is left with the last candidate's and uploaded
// again for the chosen one at the next step.
static const int kListCaps[5] = {256, 512, 1024, 2048, 4096};
static int autotune_lists(amcd_batch* b) {
  const int n = b->n;
  std::vector<int> ls(n + 1);
  float best = 1e30f;
  int pick = b->list_cap;
  for (int c = 0; c < 5; ++c) {
    ls[0] = 0;
    for (int i = 0; i < n; ++i) ls[i + 1] = ls[i] + std::min(kListCaps[c], (b->h[i].acap + 3) / 4);
    HIPT(hipMemcpyAsync(b->d_lstart, ls.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_grow_lists, dim3(ls[n]), dim3(GB), 0, b->stream, b->d_inst, b->d_lstart, n);
    HIPT(hipEventRecord(b->ev[14], b->stream));
    for (int r = 0; r < 3; ++r)
      hipLaunchKernelGGL(k_grow_lists, dim3(ls[n]), dim3(GB), 0, b->stream, b->d_inst, b->d_lstart, n);
    HIPT(hipEventRecord(b->ev[15], b->stream));
    HIPT(hipGetLastError());
    HIPT(hipEventSynchronize(b->ev[15]));
    float ms = 0;
    HIPT(hipEventElapsedTime(&ms, b->ev[14], b->ev[15]));
    b->tune_ms[c] = ms / 3;
    if (ms < best) {
      best = ms;
      pick = kListCaps[c];
    }
  }
  b->list_cap = pick;
  b->tuned = true;
  b->lstart_up.clear();  // the device's list now holds the last candidate's: upload at the next step
  return 0;
}
// after a finished step: the next step tunes once the largest growing list has reached 256
// cells, and again each time it has doubled since
static int maybe_autotune(amcd_batch* b, const amcd_out* outs) {
  if (!b->autotune) return 0;
  int mx = 0;
  for (int i = 0; i < b->n; ++i) mx = std::max(mx, outs[i].n_slots);
  if (mx < 256 || mx < 2 * b->tuned_at) return 0;
  b->tuned_at = mx;
  b->n_tunings++;
  b->tune_pending = true;
  return 0;
}
int amcd_tuning(amcd_batch* b, int* list_cap, int* n_tunings, int* cells, float* ms, int nms) {
  if (list_cap) *list_cap = b->tuned ? b->list_cap : 0;
  if (n_tunings) *n_tunings = b->n_tunings;
  if (cells) *cells = b->tuned_at;
  for (int c = 0; c < nms && c < 5; ++c) ms[c] = b->tune_ms[c];
  return b->tuned ? 1 : 0;
}

// The step with its stages in graphs: one graph launch up to the capture loop's outputs, then,
// when every instance has finished its loop, one for the finishing pair.  Suspensions (pool
// refill, larger tier, wide bitmap) and capacity growth take the same host paths as the
// stage-by-stage form.
static int update_graph(amcd_batch* b, const amcd_step* steps, amcd_out* outs) {
  if (put_steps(b, steps)) return -1;
  for (int attempt = 0; attempt < 8; ++attempt) {
    if (prep_active(b) || launch_step_graph(b) || read_outs(b, outs)) return -1;
    bool ovf = false;
    for (int i = 0; i < b->n; ++i) ovf = ovf || outs[i].overflow;
    if (!ovf) break;
    if (grow_active(b, outs)) return -1;
  }
  if (widen(b, outs) || regrow_tier(b, outs)) return -1;
  bool pend = false;
  for (int i = 0; i < b->n; ++i) pend = pend || outs[i].status == 1;
  if (!pend) {
    if (launch_finish_graph(b)) return -1;
    HIPT(hipStreamSynchronize(b->stream));
    pick_tier(b, outs);
    if (maybe_autotune(b, outs)) return -1;
  }
  return 0;
}

static int update_common(amcd_batch* b, const amcd_step* steps, amcd_out* outs, bool fused) {
  // the driver's fused step runs as graphs; per-stage timing needs events between the stages,
  // and the seam form keeps the output records its temperature call wrote: stage by stage
  if (b->graphs && !b->timing && fused && !b->tune_pending) return update_graph(b, steps, outs);
  if (put_steps(b, steps)) return -1;
  for (int attempt = 0; attempt < 8; ++attempt) {
    if (fused || attempt > 0) {
      // fresh output records; the temperature kernel sets any_melt
      HIPT(hipMemsetAsync(b->d_outs, 0, sizeof(amcd_out) * b->n, b->stream));
    }
    // A retry repeats the step's stages with the same temperature source: the fused form
    // re-evaluates the same laser (idempotent), the seam form keeps the caller's field and
    // re-applies SetLiquid3 to it (idempotent for the same T).
    if (launch_update(b, fused)) return -1;
    if (launch_active(b)) return -1;
    if (read_outs(b, outs)) return -1;
    bool ovf = false;
    for (int i = 0; i < b->n; ++i) ovf = ovf || outs[i].overflow;
    if (!ovf) break;
    if (grow_active(b, outs)) return -1;
  }
  if (widen(b, outs) || regrow_tier(b, outs)) return -1;
  bool pend = false;
  for (int i = 0; i < b->n; ++i) pend = pend || outs[i].status == 1;
  if (!pend) {
    if (launch_finish(b)) return -1;
    HIPT(hipStreamSynchronize(b->stream));
    collect_timing(b);
    pick_tier(b, outs);
    if (maybe_autotune(b, outs)) return -1;
  }
  return 0;
}

int amcd_update_voxels(amcd_batch* b, const amcd_step* steps, amcd_out* outs) {
  return update_common(b, steps, outs, false);
}

int amcd_step_fused(amcd_batch* b, const amcd_step* steps, amcd_out* outs) {
  return update_common(b, steps, outs, true);
}

int amcd_resume(amcd_batch* b, const amcd_step* steps, amcd_out* outs) {
  if (put_steps(b, steps)) return -1;
  if (launch_event_loop(b)) return -1;
  if (read_outs(b, outs)) return -1;
  if (widen(b, outs) || regrow_tier(b, outs)) return -1;
  bool pend = false;
  for (int i = 0; i < b->n; ++i) pend = pend || outs[i].status == 1;
  if (!pend) {
    if (launch_finish(b)) return -1;
    HIPT(hipStreamSynchronize(b->stream));
  }
  return 0;
}

// VoxelsCA::SetBuild (S/VoxelsCA.C:76-115) as two rounds of grain-set membership: mark the
// grains of the flagged voxels (BuildjID == 1, then Build == 1), then flag every voxel of a
// marked grain as built.  Grain ids lie in 1..ngrain.
__global__ void k_build_mark(const int* __restrict__ gid, const uint8_t* __restrict__ flag, int N, int ngrain,
                             uint8_t* __restrict__ mark) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N || flag[v] != 1) return;
  const int g = gid[v];
  if (g > 0 && g <= ngrain) mark[g] = 1;
}
__global__ void k_build_apply(const int* __restrict__ gid, int N, int ngrain, const uint8_t* __restrict__ mark,
                              uint8_t* __restrict__ build) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const int g = gid[v];
  if (g > 0 && g <= ngrain && mark[g]) build[v] = 1;
}
int amcd_set_build(amcd_batch* b, int i, int ngrain) {
  const Inst& d = b->h[i];
  if (b->mark_cap < (size_t)ngrain + 1) {
    HIPT(hipStreamSynchronize(b->stream));
    if (b->d_mark) HIPT(hipFree(b->d_mark));
    b->d_mark = nullptr;
    b->mark_cap = 0;
    const size_t cap = std::max<size_t>(4096, 2 * ((size_t)ngrain + 1));
    HIPT(hipMalloc(&b->d_mark, cap));
    b->mark_cap = cap;
  }
  uint8_t* mk = b->d_mark;
  const dim3 grid((d.N + 255) / 256);
  for (int round = 0; round < 2; ++round) {
    HIPT(hipMemsetAsync(mk, 0, (size_t)ngrain + 1, b->stream));
    hipLaunchKernelGGL(k_build_mark, grid, dim3(256), 0, b->stream, d.gid, round == 0 ? d.bj : d.build, d.N,
                       ngrain, mk);
    hipLaunchKernelGGL(k_build_apply, grid, dim3(256), 0, b->stream, d.gid, d.N, ngrain, mk, d.build);
  }
  HIPT(hipGetLastError());
  return 0;
}
// Build = 1 on the first n_built voxels (the base plate, S/VoxelsCA.C:157-160), 0 elsewhere
int amcd_build_init(amcd_batch* b, int i, long long n_built) {
  const Inst& d = b->h[i];
  HIPT(hipMemsetAsync(d.build, 0, (size_t)d.N, b->stream));
  if (n_built > 0) HIPT(hipMemsetAsync(d.build, 1, (size_t)std::min<long long>(n_built, d.N), b->stream));
  HIPT(hipStreamSynchronize(b->stream));
  return 0;
}
int amcd_get_build(amcd_batch* b, int i, int* out) {
  const Inst& d = b->h[i];
  std::vector<uint8_t> tmp((size_t)d.N);
  HIPT(hipStreamSynchronize(b->stream));
  HIPT(hipMemcpy(tmp.data(), d.build, (size_t)d.N, hipMemcpyDeviceToHost));
  for (size_t v = 0; v < tmp.size(); ++v) out[v] = tmp[v];
  return 0;
}

int amcd_grain_presence(amcd_batch* b, int i, int ngrain, int* present) {
  Inst& d = b->h[i];
  int* dp = nullptr;
  HIPT(hipMalloc(&dp, sizeof(int) * (size_t)(ngrain + 1)));
  HIPT(hipMemsetAsync(dp, 0, sizeof(int) * (size_t)(ngrain + 1), b->stream));
  hipLaunchKernelGGL(k_grain_presence, dim3((d.N + 255) / 256), dim3(256), 0, b->stream, d.gid, d.N, dp, ngrain);
  HIPT(hipMemcpyAsync(present, dp, sizeof(int) * (size_t)ngrain, hipMemcpyDeviceToHost, b->stream));
  HIPT(hipStreamSynchronize(b->stream));
  HIPT(hipFree(dp));
  return 0;
}

int amcd_remap_grains(amcd_batch* b, int i, const int* remap, int ngrain) {
  Inst& d = b->h[i];
  int* dr = nullptr;
  HIPT(hipMalloc(&dr, sizeof(int) * (size_t)(ngrain + 1)));
  HIPT(hipMemcpyAsync(dr, remap, sizeof(int) * (size_t)ngrain, hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(k_remap, dim3((d.N + 255) / 256), dim3(256), 0, b->stream, d.gid, d.N, dr, ngrain);
  HIPT(hipStreamSynchronize(b->stream));
  HIPT(hipFree(dr));
  return 0;
}

int amcd_init_plate(amcd_batch* b, int i, int nzh, const double* xs, int ng) {
  if (i < 0 || i >= b->n || ng <= 0 || nzh <= 0) return -1;
  Inst& d = b->h[i];
  if (nzh > d.p.nz) nzh = d.p.nz;
  double* dx = nullptr;
  HIPT(hipMalloc(&dx, sizeof(double) * 3 * (size_t)ng));
  HIPT(hipMemcpyAsync(dx, xs, sizeof(double) * 3 * (size_t)ng, hipMemcpyHostToDevice, b->stream));
  const long long n = (long long)d.nxy * nzh;
  hipLaunchKernelGGL(k_voronoi<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, b->d_inst, i, 0,
                     nzh, dx, ng, nullptr, nullptr);
  HIPT(hipGetLastError());
  HIPT(hipMemsetAsync(d.zliq, 0xff, (size_t)d.p.nz * 4, b->stream));
  if (invalidate_T(b, i)) return -1;
  HIPT(hipStreamSynchronize(b->stream));
  HIPT(hipFree(dx));
  return 0;
}

int amcd_add_layer(amcd_batch* b, int i, int iz1, int ilaserLoc, int gbase, const double* xs, int ng, int* n_used) {
  if (i < 0 || i >= b->n || ng <= 0) return -1;
  Inst& d = b->h[i];
  const int nl = ilaserLoc - iz1;
  const size_t cnt = (size_t)d.nxy * (size_t)(nl > 0 ? nl : 0);
  double* dx = nullptr;
  int *lab = nullptr, *used = nullptr;
  HIPT(hipMalloc(&dx, sizeof(double) * 3 * (size_t)ng));
  HIPT(hipMalloc(&lab, sizeof(int) * (cnt + 1)));
  HIPT(hipMalloc(&used, sizeof(int) * (2 * (size_t)ng + 1)));  // used[ng], rank[ng], total
  int* rank = used + ng;
  int* tot = used + 2 * ng;
  HIPT(hipMemcpyAsync(dx, xs, sizeof(double) * 3 * (size_t)ng, hipMemcpyHostToDevice, b->stream));
  HIPT(hipMemsetAsync(used, 0, sizeof(int) * (2 * (size_t)ng + 1), b->stream));
  if (cnt)
    hipLaunchKernelGGL(k_voronoi<true>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, b->stream, b->d_inst, i,
                       iz1, nl, dx, ng, lab, used);
  hipLaunchKernelGGL(k_site_rank, dim3(1), dim3(1024), 0, b->stream, used, ng, rank, tot);
  hipLaunchKernelGGL(k_write_layer, dim3((d.N + 255) / 256), dim3(256), 0, b->stream, b->d_inst, i, iz1, ilaserLoc,
                     gbase, lab, rank);
  HIPT(hipGetLastError());
  HIPT(hipMemsetAsync(d.zliq, 0xff, (size_t)d.p.nz * 4, b->stream));
  if (invalidate_T(b, i)) return -1;
  HIPT(hipMemcpyAsync(n_used, tot, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  HIPT(hipStreamSynchronize(b->stream));
  HIPT(hipFree(dx));
  HIPT(hipFree(lab));
  HIPT(hipFree(used));
  return 0;
}

// pow(x, y) rounded to nearest, the host build of the device routine (velocity table)
double amcd_host_pow(double x, double y) { return amc_pow_cr(x, y); }
int amcd_host_pow_n(const double* x, const double* y, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = amc_pow_cr(x[i], y[i]);
  return 0;
}

// the capture shape factor evaluated by the host build of the device math
int amcd_host_shape_factor(const double* dnx, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = amc_shape_factor(dnx[3 * i], dnx[3 * i + 1], dnx[3 * i + 2]);
  return 0;
}

// engine stream of libstdc++'s uniform_real_distribution<double>(0,1) over
// minstd_rand0(seed), as the device computes it (for known-answer tests)
int amcd_uniform_stream(unsigned long long seed, int n, double* out) {
  unsigned int x = lcg_seed(seed);
  for (int q = 0; q < n; ++q) {
    unsigned int x1 = lcg_next(x), x2 = lcg_next(x1);
    x = x2;
    out[q] = canonical(x1, x2);
  }
  return 0;
}
// the same stream from jump-ahead positions (the ruq path of k_compact)
int amcd_uniform_jump(unsigned long long seed, int n, double* out) {
  unsigned int x0 = lcg_seed(seed);
  for (int a = 0; a < n; ++a) {
    unsigned int x1 = mod_m((unsigned long long)x0 * lcg_pow(2ull * a + 1));
    out[a] = canonical(x1, lcg_next(x1));
  }
  return 0;
}

}  // extern "C"
