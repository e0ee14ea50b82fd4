// oracle/cafe_oracle.cpp
//
// TEST INFRASTRUCTURE ONLY -- this is the oracle, the checker, never the product.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// A CPU restatement, operation for operation, of the AMCAFE CAFE time step as the
// reference runs it at one MPI rank (src/src-mpi, cited below as S/):
//   * deck parser and laser state machine        S/Grid.C:15-164, 165-318
//   * analytic melt-pool temperature              S/TempField.C:53-77, 139-217
//   * base plate (Poisson count + Voronoi)        S/BasePlate.C:18-97
//   * orientation sampler (cube -> ball map)      S/SampleOrientation.C:16-161
//   * the CA step UpdateVoxels                    S/VoxelsCA.C:232-655
//       SetLiquid3 S/VoxelsCA.C:1098-1116, ConvertSolid1 :712-799
//   * per-layer routines CleanLayer :800-839, AddLayer1 :887-982
//   * SetBuild :76-115, InitializeVoxels :125-161
//   * the driver loop                             S/main1.C:59-98
// At one rank the partition is the identity (S/Partition.C:269-288), so every
// std::find over icellidLoc / vI in the reference is replaced by a direct index or
// a binary search over the sorted active list; that changes cost, not results.
//
// Compile flags follow the reference build (S/Makefile_gaffney.mk:3-5): -O3, no
// -march, so x86-64 SSE2 doubles with no fused multiply-add.  Random numbers come
// from libstdc++ <random> (minstd_rand0, generate_canonical, poisson_distribution),
// the same third-party library the reference links.
//
// Math backend: by default glibc libm (what the reference calls).  With
// orc_set_crmath(1) the transcendental call sites of the capture loop (atan2, cos, sin,
// pow(x,0.5) of the geometry and pow(tL - T, nvel) of the growth velocity) use correctly
// rounded results from libquadmath
// instead; that mode is a bit-exact model of the GPU kernels' own math and is
// used to measure how much libm rounding matters for the decisions.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>
#include <quadmath.h>

namespace orc {

static int g_crmath = 0;

static double m_atan2(double y, double x) {
  return g_crmath ? (double)atan2q((__float128)y, (__float128)x) : std::atan2(y, x);
}
static double m_cos(double x) { return g_crmath ? (double)cosq((__float128)x) : std::cos(x); }
static double m_sin(double x) { return g_crmath ? (double)sinq((__float128)x) : std::sin(x); }
// pow(x,.5) is a libm call in the reference (GCC does not fold it to sqrt without
// -funsafe-math-optimizations); sqrt is its correctly rounded value.
static double m_powh(double x) { return g_crmath ? std::sqrt(x) : std::pow(x, .5); }
// pow(tL - T, nvel) of the growth velocity (S/VoxelsCA.C:405)
static double m_pow(double x, double y) { return g_crmath ? (double)powq((__float128)x, (__float128)y) : std::pow(x, y); }

// pow(cos th * sin ph, .5) of S/VoxelsCA.C:415-418, 424 (also :548-551, 591), a pure
// function of the centre-to-centroid vector; memoised because the reference's scan
// re-evaluates the same vectors in every iteration of the capture loop.
struct Key3 {
  uint64_t a, b, c;
  bool operator==(const Key3& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct Key3Hash {
  size_t operator()(const Key3& k) const {
    uint64_t h = k.a * 0x9E3779B97F4A7C15ull ^ (k.b + 0x632BE59BD9B4E019ull + (k.a << 6)) ^ (k.c * 0xC2B2AE3D27D4EB4Full);
    return (size_t)(h ^ (h >> 29));
  }
};
static std::unordered_map<Key3, double, Key3Hash> g_shape_memo;
static double shape_factor_raw(const double* dnx) {
  double th = m_atan2(std::fabs(dnx[1]), std::fabs(dnx[0]));
  th > M_PI / 4.0 ? th = M_PI / 2.0 - th : th;
  double ph = m_atan2(m_powh(dnx[0] * dnx[0] + dnx[1] * dnx[1]), std::fabs(dnx[2]));
  ph < M_PI / 4.0 ? ph = M_PI / 2.0 - ph : ph;
  return m_powh(m_cos(th) * m_sin(ph));
}
static double shape_factor(const double* dnx) {
  Key3 k;
  std::memcpy(&k.a, &dnx[0], 8);
  std::memcpy(&k.b, &dnx[1], 8);
  std::memcpy(&k.c, &dnx[2], 8);
  auto it = g_shape_memo.find(k);
  if (it != g_shape_memo.end()) return it->second;
  if (g_shape_memo.size() > (8u << 20)) g_shape_memo.clear();
  double f = shape_factor_raw(dnx);
  g_shape_memo.emplace(k, f);
  return f;
}

typedef std::minstd_rand0 Engine;  // std::default_random_engine in libstdc++ (random.h:1604)

// ---------------------------------------------------------------------------
// Orientation sampler: S/SampleOrientation.C
struct Orient {
  double r, R1, a, aprime, beta;
  int Ngrid;
  double coeffA[7];
  Orient() {  // S/SampleOrientation.C:16-30 (constant pow() calls are folded by GCC)
    r = 1;
    Ngrid = 100;
    R1 = pow(3.0 * M_PI / 4.0, 1.0 / 3.0) * r;
    a = pow(2.0 * M_PI / 3.0, 1.0 / 2.0) * R1;
    aprime = pow(4.0 * M_PI / 3.0, 1.0 / 3.0) * R1;
    const double c[7] = {-0.500009615, -0.024866061, -0.004549382, 0.000511867,
                         -0.001650083, 0.000759335,  -0.000204042};
    for (int i = 0; i < 7; ++i) coeffA[i] = c[i];
    beta = pow(M_PI / 6.0, 1.0 / 2.0);
  }
  // S/SampleOrientation.C:32-58: pyramid permutation, later tests override earlier
  static void perm(const double* p, int* P) {
    double x = p[0], y = p[1], z = p[2];
    if ((fabs(x) <= z) && (fabs(y) <= z)) { P[0] = 1; P[1] = 2; P[2] = 3; }
    if ((z <= -fabs(x)) && (z <= -fabs(y))) { P[0] = -1; P[1] = 2; P[2] = -3; }
    if ((fabs(x) <= y) && (fabs(z) <= y)) { P[0] = 1; P[1] = -3; P[2] = 2; }
    if ((y <= -fabs(x)) && (y <= -fabs(z))) { P[0] = 1; P[1] = 3; P[2] = -2; }
    if ((fabs(z) < x) && (fabs(y) <= x)) { P[0] = -3; P[1] = 2; P[2] = 1; }
    if ((x <= -fabs(z)) && (x <= -fabs(y))) { P[0] = 3; P[1] = 2; P[2] = -1; }
  }
  // S/SampleOrientation.C:68-97 (map 2: square -> disk)
  void map2(const double* p, double* o) const {
    double x = p[0], y = p[1], z = p[2], T1 = 0, T2 = 0;  // T1/T2 set on every path
    if (fabs(y) <= fabs(x)) {
      if (fabs(x) < 1e-10) {
        T1 = 0.0; T2 = 0.0;
      } else {
        T1 = pow(2.0, 1.0 / 2.0) * cos(y * M_PI / 12 / x) - 1;
        T2 = pow(2.0, 1.0 / 2.0) * sin(y * M_PI / 12 / x);
        T1 *= pow(2.0, 1.0 / 4.0) * x / (beta * pow(pow(2.0, 1.0 / 2.0) - cos(y * M_PI / 12 / x), 1.0 / 2.0));
        T2 *= pow(2.0, 1.0 / 4.0) * x / (beta * pow(pow(2.0, 1.0 / 2.0) - cos(y * M_PI / 12 / x), 1.0 / 2.0));
      }
    }
    if (fabs(x) <= fabs(y)) {
      if (fabs(y) < 1e-10) {
        T1 = 0.0; T2 = 0.0;
      } else {
        T1 = pow(2.0, 1.0 / 2.0) * sin(x * M_PI / 12 / y);
        T2 = pow(2.0, 1.0 / 2.0) * cos(x * M_PI / 12 / y) - 1;
        T1 *= pow(2.0, 1.0 / 4.0) * y / (beta * pow(pow(2.0, 1.0 / 2.0) - cos(x * M_PI / 12 / y), 1.0 / 2.0));
        T2 *= pow(2.0, 1.0 / 4.0) * y / (beta * pow(pow(2.0, 1.0 / 2.0) - cos(x * M_PI / 12 / y), 1.0 / 2.0));
      }
    }
    o[0] = T1; o[1] = T2; o[2] = z;
  }
  // S/SampleOrientation.C:99-113 (map 3: disk -> sphere)
  static void map3(const double* p, double* o) {
    double x = p[0], y = p[1], z = p[2];
    if (fabs(z) < 1e-10) {
      o[0] = 0; o[1] = 0; o[2] = 0;
    } else {
      // GCC -O3 vectorises the reference's two identical pow(., 1/2) factors into one
      // packed square root (the pow calls remain only for errno), so its result is
      // the correctly rounded sqrt, which glibc's pow misses by an ulp now and then.
      const double s = sqrt(1 - (pow(x, 2) + pow(y, 2)) * M_PI / (24 * pow(z, 2)));
      o[0] = s * x;
      o[1] = s * y;
      o[2] = pow(6 / M_PI, 1.0 / 2.0) * z - (pow(x, 2) + pow(y, 2)) * pow(M_PI, 1.0 / 2.0) / (pow(24, 1.0 / 2.0) * z);
    }
  }
  // S/SampleOrientation.C:124-161: Ns samples as (omega, axis) from minstd_rand0(seed)
  void generate(int Ns, unsigned int seed, std::vector<double>& out) const {
    std::uniform_real_distribution<double> xrand(0.0, 1.0);
    out.assign(Ns * 4, 0.0);
    Engine g1(seed);
    for (int jn = 0; jn < Ns; ++jn) {
      double xyz0[3];
      xyz0[0] = xrand(g1) - .5;
      xyz0[1] = xrand(g1) - .5;
      xyz0[2] = xrand(g1) - .5;
      double xyz[3] = {aprime * xyz0[0], aprime * xyz0[1], aprime * xyz0[2]};
      int P[3] = {1, 2, 3};
      perm(xyz, P);
      double xyz2[3], m1[3], m2[3], m3[3], xyzb[3];
      for (int j = 0; j < 3; ++j) xyz2[j] = P[j] / abs(P[j]) * xyz[abs(P[j]) - 1];
      for (int j = 0; j < 3; ++j) m1[j] = a / aprime * xyz2[j];
      map2(m1, m2);
      map3(m2, m3);
      for (int j = 0; j < 3; ++j) xyzb[abs(P[j]) - 1] = P[j] / abs(P[j]) * m3[j];
      double pmax = std::max(std::max(std::abs(xyz0[0]), std::abs(xyz0[1])), std::abs(xyz0[2]));
      double rhoP = R1 * pmax / .5;
      double t = 1;
      for (int j = 1; j < 8; ++j) t += coeffA[j - 1] * pow(rhoP, 2.0 * j);
      double omega = 2.0 * acos(t);
      double tnorm = pow(pow(xyzb[0], 2.0) + pow(xyzb[1], 2.0) + pow(xyzb[2], 2.0), .5);
      out[4 * jn] = omega;
      out[4 * jn + 1] = xyzb[0] / tnorm;
      out[4 * jn + 2] = xyzb[1] / tnorm;
      out[4 * jn + 3] = xyzb[2] / tnorm;
    }
  }
};

// ---------------------------------------------------------------------------
// Grid: S/Grid.C
struct Grid {
  std::vector<int> nX{128, 128, 64};
  std::vector<double> dX{1e-6, 1e-6, 1e-6};
  int outint = 1, patternID = 1, outNL = 0, nTsd = 0, ictrl = 3;
  double tL = 1620, tS = 1531.5, T0 = 300, layerT = 25e-6, bpH = 0.0, mu = .01;
  std::vector<double> meltparam{75e-6, 162.75e-6, 75e-6, 75e-6}, offset{0.0, 0.0, 0.0}, gsize{0, 0};
  double bhatch = 0, bmDelT = 0, bmV = 500e-3, bmP = 0, rNmax = .002, lrate = 0, gth0 = 0., gth = 0.;
  double Avel = 0., nvel = 0.;
  bool set_bmdelt = false, set_lrate = false;
  // derived (S/Grid.C:69-108)
  std::vector<double> LX, lcoor, lcoor2;
  double gbox[4];
  int nZlayer = 0, isp = 0, inewscanflg = 1, inewlayerflg = 1, Nzhg = 0, indlayer = 1, ilaserLoc = 0;
  int Ntd = 0, Nsd = 0, NpT = 0, tInd = 0;
  double time = 0.0, dt = 0.0;

  // S/Grid.C:170-190: '#' comments dropped up to the newline (the newline is
  // kept); the failed get() at end of file leaves the last character in place,
  // so the last character is stored twice.
  static std::string strip(const std::string& raw) {
    std::string out;
    size_t i = 0, n = raw.size();
    char c = 0;
    while (i < n) {
      c = raw[i++];
      if (c == '#') {
        while (c != '\n' && i < n) c = raw[i++];
      }
      out.push_back(c);
    }
    if (n > 0) out.push_back(c);
    return out;
  }

  int parse(const std::string& raw, std::string& err) {  // S/Grid.C:195-317
    std::istringstream in(strip(raw));
    std::string kw;
    in >> kw;
    while (in) {
      std::transform(kw.begin(), kw.end(), kw.begin(), [](unsigned char s) { return std::tolower(s); });
      if (kw == "dx") { in >> kw; dX.assign(3, std::stod(kw)); }
      if (kw == "bmv") { in >> kw; bmV = std::stod(kw); }
      if (kw == "bmp") { in >> kw; bmP = std::stod(kw); }
      if (kw == "bmdelt") { in >> kw; bmDelT = std::stod(kw); set_bmdelt = true; }
      if (kw == "nx") {
        nX.resize(3);
        in >> kw; nX[0] = std::stoi(kw);
        in >> kw; nX[1] = std::stoi(kw);
        in >> kw; nX[2] = std::stoi(kw);
      }
      if (kw == "r") { in >> kw; rNmax = std::stod(kw); }
      if (kw == "lt") { in >> kw; layerT = std::stod(kw); }
      if (kw == "tl") { in >> kw; tL = std::stod(kw); }
      if (kw == "ts") { in >> kw; tS = std::stod(kw); }
      if (kw == "bhatch") { in >> kw; bhatch = std::stod(kw); }
      if (kw == "mu") { in >> kw; mu = std::stod(kw); }
      if (kw == "bph") { in >> kw; bpH = std::stod(kw); }
      if (kw == "lrate") { in >> kw; lrate = std::stod(kw); set_lrate = true; }
      if (kw == "outnl") { in >> kw; outNL = std::stoi(kw); }
      if (kw == "ntsd") { in >> kw; nTsd = std::stoi(kw); }
      if (kw == "offset") {
        in >> kw; offset[0] = std::stod(kw);
        in >> kw; offset[1] = std::stod(kw);
        in >> kw; offset[2] = std::stod(kw);
      }
      if (kw == "meltparam") {
        meltparam.resize(4);
        for (int q = 0; q < 4; ++q) { in >> kw; meltparam[q] = std::stod(kw); }
      }
      if (kw == "patternid") { in >> kw; patternID = std::stoi(kw); }
      if (kw == "outint") { in >> kw; outint = std::stoi(kw); }
      if (kw == "gridsize") { in >> kw; gsize[0] = std::stod(kw); in >> kw; gsize[1] = std::stod(kw); }
      if (kw == "gridtheta") { in >> kw; gth0 = std::stod(kw); }
      if (kw == "dendritegrowth2") { in >> kw; Avel = std::stod(kw); in >> kw; nvel = std::stod(kw); }
      in >> kw;
    }
    if (!set_bmdelt) { err = "deck must set bmDelT (the reference default reads bmV before it is set)"; return -1; }
    if (!set_lrate) { err = "deck must set lrate (the reference leaves it uninitialised)"; return -1; }
    return 0;
  }

  void init() {  // S/Grid.C:68-108 (bhatch default 1.53*meltparam[2] only if deck lacks it)
    LX = {nX[0] * dX[0], nX[1] * dX[1], nX[2] * dX[2]};
    bpH < std::numeric_limits<double>::epsilon() ? bpH = layerT : bpH = bpH;
    nZlayer = round(layerT / dX[2]);
    isp = 0;
    inewscanflg = 1;
    inewlayerflg = 1;
    Nzhg = std::min(nX[2], int(ceil(bpH * .9999 / dX[2])));
    indlayer = 1;
    ilaserLoc = Nzhg + indlayer * nZlayer;
    Ntd = int(ceil(gsize[1] / bhatch)) + 1;
    Nsd = int(ceil(gsize[0] / (bmV * bmDelT))) + 1;
    NpT = Nsd * Ntd;
    lcoor.assign(2 * NpT, 0.);
    lcoor2.assign(2 * NpT, 0.);
    double xlmin = (LX[0] - gsize[0]) / 2., ylmin = (LX[1] - gsize[1]) / 2.;
    double gmid[2] = {LX[0] / 2., LX[1] / 2.};
    gth += gth0;
    for (int j2 = 0; j2 < Ntd; ++j2)
      for (int j1 = 0; j1 < Nsd; ++j1) {
        int k = Nsd * j2 + j1;
        lcoor[2 * k] = j1 * bmDelT * bmV + xlmin;
        lcoor[2 * k + 1] = j2 * bhatch + ylmin;
      }
    rotate(gmid);
    gbox[0] = -bhatch / 2.;
    gbox[1] = LX[0] + bhatch / 2.;
    gbox[2] = -bhatch / 2.;
    gbox[3] = LX[1] + bhatch / 2.;
  }
  void rotate(const double* gmid) {
    for (int j2 = 0; j2 < Ntd; ++j2)
      for (int j1 = 0; j1 < Nsd; ++j1) {
        int k = Nsd * j2 + j1;
        lcoor2[2 * k] = cos(gth) * (lcoor[2 * k] - gmid[0]) - sin(gth) * (lcoor[2 * k + 1] - gmid[1]) + gmid[0];
        lcoor2[2 * k + 1] = sin(gth) * (lcoor[2 * k] - gmid[0]) + cos(gth) * (lcoor[2 * k + 1] - gmid[1]) + gmid[1];
      }
  }
  void update_laser() {  // S/Grid.C:110-164
    int itmp, iflg = 0, irep = 0;
    double x, y;
    if (fmod(isp + 1, Nsd) == 0) inewscanflg = 1;
    if (isp == (NpT - 1)) {
      inewscanflg = 1;
      inewlayerflg = 1;
      isp = 0;
      indlayer += 1;
    }
    while (irep == 0 || isp == 0) {
      irep += 1;
      itmp = isp;
      if (inewscanflg == 0 && itmp < (NpT - 1)) {
        isp += 1;
      } else {
        while (itmp < (NpT - 1) && iflg == 0) {
          itmp += 1;
          x = lcoor2[2 * itmp];
          y = lcoor2[2 * itmp + 1];
          if (x > gbox[0] && x < gbox[1] && y > gbox[2] && y < gbox[3]) iflg = 1;
        }
        if (itmp < (NpT - 1)) {
          isp = itmp;
          inewscanflg = 0;
        } else {
          inewlayerflg = 1;
          inewscanflg = 1;
          indlayer += 1;
          ilaserLoc = Nzhg + indlayer * nZlayer;
          isp = 0;
          double gmid[2] = {LX[0] / 2., LX[1] / 2.};
          gth += gth0;
          rotate(gmid);
        }
      }
    }
    if (irep == 1) inewlayerflg = 0;
  }
  void update_time2(double dtIn) { dt = dtIn; time += dt; tInd += 1; }  // S/Grid.C:334-339
};

// ---------------------------------------------------------------------------
// Per-run state (TempField + BasePlate + VoxelsCA + driver)
struct StepStats {
  long Na = 0, n_mushy0 = 0, iters = 0, captures = 0, nucleations = 0, draws = 0;
  double tinc = 0;
  // event log of the step: captured voxel, capturing voxel (-1: nucleation), Dt
  std::vector<int> ev_vox, ev_src;
  std::vector<double> ev_dt;
};

struct Sim {
  Grid g;
  Orient sa;
  int N = 0, nXY = 0;
  // TempField (S/TempField.C:16-51)
  std::vector<int> ispvec;
  double a1[6], DelT = 0, bmDX[3];
  int TtInd = 0;
  std::vector<double> TempCurr;
  std::vector<int> BuildjID;
  // BasePlate (S/BasePlate.C:18-30): the member engine stays default-seeded
  Engine bp_generator;
  int Nzh = 0, bpNgrain = 0;
  double bp_height = 0;
  // VoxelsCA (S/VoxelsCA.C:34-74)
  std::vector<int> gID, vState, Build;
  std::vector<double> extents, centroidOct, cTheta;
  int nGrain = 0, seed0 = 2132512, seed1 = 2912351;
  Engine genlayer;
  // driver
  int nlayerTot = 0;
  bool done = false;
  std::vector<int> out_steps;
  // what WriteToHDF1 would write at each output (S/main1.C:83-84, 95-97), when recording
  struct Snapshot {
    int tInd;
    std::vector<int> gID, vState, Build;
    std::vector<double> T, cTheta;
  };
  bool record_outputs = false;
  std::vector<Snapshot> snaps;
  void snapshot() {
    if (!record_outputs) return;
    snaps.push_back(Snapshot{TtInd, gID, vState, Build, TempCurr, cTheta});
  }
  StepStats st;
  // sites of the last Voronoi tessellation (base plate or powder layer) and, for the last
  // AddLayer1: iz1, ilaserLoc, sites drawn, labels used (i2), orientation seed, nGrain before
  std::vector<double> last_sites;
  long last_layer[6] = {0, 0, 0, 0, 0, 0};
  // Moore offsets in ascending linear-index order (S/Grid.C:399-414 + sort at :375)
  int noff = 0;

  void nbrs(int j, int* out, int& cnt) const {
    const int nx = g.nX[0], ny = g.nX[1], nz = g.nX[2];
    int k = j / nXY, r = j - k * nXY, jy = r / nx, ix = r - jy * nx;
    cnt = 0;
    for (int dz = -1; dz <= 1; ++dz) {
      int z = k + dz;
      if (z < 0 || z >= nz) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        int y = jy + dy;
        if (y < 0 || y >= ny) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          int x = ix + dx;
          if (x < 0 || x >= nx) continue;
          if (dz == 0 && dy == 0 && dx == 0) continue;
          out[cnt++] = z * nXY + y * nx + x;
        }
      }
    }
  }

  int setup(const std::string& deck, std::string& err, bool bare = false) {
    if (g.parse(deck, err)) return -1;
    g.init();
    N = g.nX[0] * g.nX[1] * g.nX[2];
    nXY = g.nX[0] * g.nX[1];
    // TempField ctor + InitializeAnalytic (S/TempField.C:16-77)
    DelT = g.bmDelT;
    bmDX[0] = DelT * g.bmV; bmDX[1] = g.bhatch; bmDX[2] = g.layerT;
    ispvec.assign(g.NpT, 0);
    if (g.patternID == 1 || g.patternID == 3) std::iota(ispvec.begin(), ispvec.end(), 0);
    if (g.patternID == 2 || g.patternID == 4) {
      for (int j = 0; j < g.Ntd; ++j) {
        int k = g.Nsd * j;
        if (fmod(j, 2) == 0) std::iota(ispvec.begin() + k, ispvec.begin() + k + g.Nsd, k);
        else for (int j1 = 0; j1 < g.Nsd; ++j1) ispvec[k + j1] = k + g.Nsd - 1 - j1;
      }
    }
    a1[0] = g.meltparam[0]; a1[1] = g.meltparam[2]; a1[2] = g.meltparam[3];
    a1[3] = g.meltparam[1]; a1[4] = a1[1]; a1[5] = a1[2];
    TempCurr.assign(N, 0.0);
    BuildjID.assign(N, 0);
    // BasePlate ctor
    bp_height = ceil(g.bpH / g.dX[2]) * g.dX[2];
    Nzh = std::min(g.nX[2], int(ceil(g.bpH * .9999 / g.dX[2])));
    bp_height = Nzh * (g.dX[2]);
    // VoxelsCA ctor
    gID.assign(N, 0); vState.assign(N, 0); Build.assign(N, 0);
    extents.assign(N, 0.0); centroidOct.assign(3 * (size_t)N, 0.0);
    genlayer.seed(seed1);
    // driver (S/main1.C:59-61)
    nlayerTot = int(ceil((double)(g.nX[2] - Nzh) / (double)g.nZlayer));
    // bare: state, grains and laser are set by the caller (the powder-layer engine stays
    // at its seed, so a bare instance is exact only until its next AddLayer1)
    if (bare) return 0;
    initialize_voxels();
    add_layer1();
    g.update_laser();
    return 0;
  }

  // the base plate's Poisson grain count and Voronoi sites (S/BasePlate.C:31-58), drawn from
  // the member engine
  int draw_plate_sites(std::vector<double>& Xs) {
    double rate = g.mu * (g.dX[0] * 1e6) * (g.dX[1] * 1e6) * g.nX[0] * g.nX[1] * (bp_height * 1e6);
    std::poisson_distribution<int> ng(rate);
    int Ngrain = 0;
    while (Ngrain == 0) Ngrain = ng(bp_generator);
    std::uniform_real_distribution<double> xrand(0.0, 1.0);
    double Lx = g.dX[0] * g.nX[0], Ly = g.dX[1] * g.nX[1];
    Xs.assign(3 * (size_t)Ngrain, 0.0);
    for (int j = 0; j < Ngrain; ++j) {  // braced list: evaluated left to right
      Xs[3 * j] = xrand(bp_generator) * Lx;
      Xs[3 * j + 1] = xrand(bp_generator) * Ly;
      Xs[3 * j + 2] = xrand(bp_generator) * bp_height;
    }
    return Ngrain;
  }
  // the plate's orientation seed (S/BasePlate.C:83-86)
  static unsigned int plate_sdloc() {
    Engine g1(531231);
    return unsigned(double(g1()) / double(g1.max()) * pow(2.0, 32.0));
  }

  // S/BasePlate.C:31-97 + S/VoxelsCA.C:125-161
  void initialize_voxels() {
    std::vector<double> Xs;
    const int Ngrain = draw_plate_sites(Xs);
    last_sites = Xs;
    int nplate = Nzh * g.nX[0] * g.nX[1];
    std::vector<int> X(nplate, 0);
    for (int j = 0; j < nplate; ++j) {
      int j3 = j / nXY, j2 = (j - nXY * j3) / g.nX[0], j1 = j - nXY * j3 - g.nX[0] * j2;
      double zloc = (j3 + .5) * g.dX[2], yloc = (j2 + .5) * g.dX[1], xloc = (j1 + .5) * g.dX[0];
      int best = 0;
      double bd = 0;
      for (int q = 0; q < Ngrain; ++q) {
        double ax = xloc - Xs[3 * q], ay = yloc - Xs[3 * q + 1], az = zloc - Xs[3 * q + 2];
        double d = ax * ax + ay * ay + az * az;
        if (q == 0 || d < bd) { bd = d; best = q; }  // std::min_element: first minimum
      }
      X[j] = best + 1;
    }
    unsigned int sdloc = plate_sdloc();
    std::vector<double> aa;
    sa.generate(Ngrain, sdloc, aa);
    bpNgrain = Ngrain;
    nGrain = Ngrain;
    for (int j = 0; j < nplate; ++j) {
      gID[j] = X[j];
      vState[j] = (j < (Nzh - 1) * nXY) ? 3 : 2;
    }
    cTheta = aa;
    for (int j = 0; j < N; ++j) {
      if (gID[j] == 0) continue;
      int k = j / nXY, r = j - k * nXY, jy = r / g.nX[0], ix = r - jy * g.nX[0];
      centroidOct[3 * (size_t)j] = (double(ix) + .5) * g.dX[0];
      centroidOct[3 * (size_t)j + 1] = (double(jy) + .5) * g.dX[1];
      centroidOct[3 * (size_t)j + 2] = (double(k) + .5) * g.dX[2];
    }
    for (int k = 0; k < N; ++k)
      if (gID[k] != 0) Build[k] = 1;
  }

  // S/TempField.C:139-217
  void analytic_temp() {
    TtInd = int(round(g.time / DelT));  // S/main1.C:64
    const double T0 = g.T0, xi = g.tL * (1 + std::numeric_limits<double>::epsilon());
    double x = g.lcoor[2 * ispvec[g.isp]] - g.offset[0];
    double y = g.lcoor[2 * ispvec[g.isp] + 1] - g.offset[1];
    double z = g.ilaserLoc * g.dX[2] - g.offset[2];
    double cg = cos(g.gth), sg = sin(g.gth);
    double dirp;
    if (fmod(g.isp, g.Nsd) == 0) dirp = (g.lcoor[2 * ispvec[g.isp + 1]] - g.lcoor[2 * ispvec[g.isp]]);
    else dirp = (g.lcoor[2 * ispvec[g.isp]] - g.lcoor[2 * ispvec[g.isp - 1]]);
    for (int j = 0; j < N; ++j) {
      TempCurr[j] = T0;
      int j3 = j / nXY, j2 = (j - nXY * j3) / g.nX[0], j1 = j - nXY * j3 - g.nX[0] * j2;
      double x0 = (double(j1) + .5) * (g.dX[0]);
      double y0 = (double(j2) + .5) * (g.dX[1]);
      double zp = (double(j3) + .5) * (g.dX[2]);
      if (zp > z) continue;
      double xp = cg * (x0 - g.LX[0] / 2.) + sg * (y0 - g.LX[1] / 2.) + g.LX[0] / 2.;
      double yp = -sg * (x0 - g.LX[0] / 2.) + cg * (y0 - g.LX[1] / 2.) + g.LX[1] / 2.;
      double r0 = xp - x, r1 = yp - y, r2 = zp - z, dsq;
      if (dirp * r0 > 0) {
        double q1 = r1 / a1[1], q2 = r2 / a1[2];
        dsq = q1 * q1 + q2 * q2;
        TempCurr[j] = (dsq < 1.0 && (fabs(r0) < bmDX[0])) ? xi : g.tS;
      } else {
        double q0 = r0 / a1[3], q1 = r1 / a1[4], q2 = r2 / a1[5];
        dsq = q0 * q0 + q1 * q1 + q2 * q2;
        TempCurr[j] = (dsq < 1.0) ? xi : g.tS;
      }
      if (TempCurr[j] > g.tL) BuildjID[j] = 1;
      else if (TempCurr[j] < T0) BuildjID[j] = 0;
    }
    // end-of-scan check (S/TempField.C:199-216)
    x = g.lcoor2[2 * ispvec[g.isp]] - g.offset[0];
    y = g.lcoor2[2 * ispvec[g.isp] + 1] - g.offset[1];
    if (x < g.gbox[0] || x > g.gbox[1] || y < g.gbox[2] || y > g.gbox[3]) {
      int icheck = 0;
      for (int j = 0; j < N; ++j)
        if (TempCurr[j] >= g.tL) { icheck = 1; break; }
      if (icheck == 0 || fmod(g.isp + 1, g.Nsd) == 0) g.inewscanflg = 1;
      if (g.isp == (g.NpT - 1)) {
        g.inewscanflg = 1;
        g.inewlayerflg = 1;
        g.isp = 0;
        g.indlayer += 1;
      }
    }
  }

  // ---- TempField::EASM (S/TempField.C:79-137) with Utilities::EASM_Temp_LG, Integral,
  // Temp_Fun_G and Materials_316L (S/Utilities.C:60-140, 228-249).  The reference's rule is
  // GSL's fixed 25-point Gauss-Legendre (not in this image); gl_rule() computes the same
  // rule independently (Newton on P25 in quad precision, rounded to double).
  static void mat316(double T, double* M) {  // M = {cP, k, rho, alpha}
    double rho, cP, k;
    if (T < 1660) rho = 8084.2 - 0.42086 * T - 3.8942 * pow(10, -5) * T * T;
    else if (T >= 1660 && T < 3090) rho = 7432.7 + 3.9338 * pow(10, -2) * T - 1.8007 * pow(10, -4) * T * T;
    else rho = 6124.72;
    if (T < 1660) cP = 458.984 + 0.1328 * T;
    else cP = 769.855;
    if (T < 1660) k = 9.248 + 1.571 * pow(10, -2) * T;
    else if (T >= 1660 && T < 3090) k = 12.41 + 3.279 * pow(10, -3) * T;
    else k = 21.6465;
    M[0] = cP; M[1] = k; M[2] = rho; M[3] = k / (rho * cP);
  }
  double easm_integral(const double* L, double T, const double* R, long* nint) const {
    const double t = 80.0e-5, to = 0.0, sigma = 50.0e-6, eta = 0.56, Q = g.bmP;
    double M[4];
    mat316(T, M);
    const double C = eta * Q / (M_PI * M[2] * M[0] * sqrt(4 * M_PI * M[3]));
    const double x0 = L[0] - g.bmV * t;
    const double shft = (to + t) / 2.0, slp = (t - to) / 2.0;
    double sum = 0.0;
    for (int i = 0; i < 25; ++i) {
      const double x = shft + slp * gl_x[i], w = gl_w[i] * slp;
      const double Lx = x0 + g.bmV * x, Ly = L[1], Lz = L[2];
      const double f = C * pow(t - x, -0.5) / (2 * M[3] * (t - x) + pow(sigma, 2)) *
                       exp(-1 * (pow(R[0] - Lx, 2) + pow(R[1] - Ly, 2)) / (4 * M[3] * (t - x) + 2 * pow(sigma, 2)) -
                           pow(std::fabs(R[2] - Lz), 2) / (4 * M[3] * (t - x)));
      sum += w * f;
    }
    ++*nint;
    return 300 + sum;
  }
  double easm_temp_lg(const double* L, const double* R, double T, long* nint) const {
    double Told = T;
    int n = 0, nlim = 100;
    double tol = 10;
    double Tnew = easm_integral(L, T, R, nint);
    while (std::fabs(Told - Tnew) > tol && n < nlim) {
      Told = Tnew;
      double T_temp = easm_integral(L, Told, R, nint);
      Tnew = 0.5 * (Told + T_temp);
      n++;
    }
    return Tnew;
  }
  // the voxel loop with an explicit laser (x, y, z, cos gth, sin gth); returns all_of(T >= tL)
  int easm(double x, double y, double z, double cg, double sg, long* nint) {
    for (int j = 0; j < N; ++j) {
      TempCurr[j] = 0;
      int j3 = j / nXY, j2 = (j - nXY * j3) / g.nX[0], j1 = j - nXY * j3 - g.nX[0] * j2;
      double x0 = (double(j1) + .5) * (g.dX[0]);
      double y0 = (double(j2) + .5) * (g.dX[1]);
      double zp = (double(j3) + .5) * (g.dX[2]);
      if (zp > z) continue;
      double xp = cg * (x0 - g.LX[0] / 2.) + sg * (y0 - g.LX[1] / 2.) + g.LX[0] / 2.;
      double yp = -sg * (x0 - g.LX[0] / 2.) + cg * (y0 - g.LX[1] / 2.) + g.LX[1] / 2.;
      const double R[3] = {xp, yp, zp}, L[3] = {x, y, z};
      if ((x0 > x) && (zp > (z - g.layerT))) TempCurr[j] = .99 * g.T0;
      else TempCurr[j] = easm_temp_lg(L, R, TempCurr[j], nint);
      if (TempCurr[j] > g.tL) BuildjID[j] = 1;
      else if (TempCurr[j] < g.T0) BuildjID[j] = 0;
    }
    for (int j = 0; j < N; ++j)
      if (!(TempCurr[j] >= g.tL)) return 0;
    return 1;
  }
  double gl_x[25], gl_w[25];
  void gl_rule() {
    const int n = 25;
    const __float128 pi = 4 * atanq((__float128)1), tiny = (__float128)1e-33;
    for (int i = 0; i < n; ++i) {
      __float128 x = cosq(pi * (i + (__float128)0.75) / (n + (__float128)0.5)), p0, p1, dp;
      for (int it = 0; it < 100; ++it) {
        p0 = 1; p1 = x;
        for (int k = 2; k <= n; ++k) { __float128 p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
        dp = n * (x * p1 - p0) / (x * x - 1);
        __float128 dx = p1 / dp;
        x -= dx;
        if (fabsq(dx) < tiny) break;
      }
      p0 = 1; p1 = x;
      for (int k = 2; k <= n; ++k) { __float128 p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
      dp = n * (x * p1 - p0) / (x * x - 1);
      if (fabsq(x) < tiny) x = 0;  // the middle root of the odd P25 is exactly 0
      gl_x[n - 1 - i] = (double)x;  // roots come out descending
      gl_w[n - 1 - i] = (double)(2 / ((1 - x * x) * dp * dp));
    }
  }

  void set_liquid3() {  // S/VoxelsCA.C:1098-1116
    int n1 = nXY, iz1 = g.ilaserLoc * nXY;
    for (int j = 0; j < N; ++j) {
      if (TempCurr[j] >= g.tL) {
        if (j < n1) {
          vState[j] = 2;
        } else if (j < iz1) {
          vState[j] = 1;
          gID[j] = 0;
        }
        extents[j] = 0.0;
      }
    }
  }

  void convert_solid1(int iswitch) {  // S/VoxelsCA.C:712-799 (collect, then apply)
    std::vector<int> hit;
    int nb[26], cnt;
    for (int j = 0; j < N; ++j) {
      if (iswitch == 0 ? vState[j] != 2 : vState[j] != 3) continue;
      nbrs(j, nb, cnt);
      bool ok;
      if (iswitch == 0) {
        ok = true;
        for (int q = 0; q < cnt; ++q) if (vState[nb[q]] < 2) { ok = false; break; }
      } else {
        ok = false;
        for (int q = 0; q < cnt; ++q) if (vState[nb[q]] == 1) { ok = true; break; }
      }
      if (ok) hit.push_back(j);
    }
    for (int j : hit) {
      if (iswitch == 0) { vState[j] = 3; extents[j] = 0.0; }
      else vState[j] = 2;
    }
  }

  // S/VoxelsCA.h:94-107 (matrix is local -> global)
  static void rot_mat(double omega, const double* ax, double R[3][3]) {
    R[0][0] = cos(omega) + pow(ax[0], 2.0) * (1 - cos(omega));
    R[0][1] = ax[0] * ax[1] * (1 - cos(omega)) - ax[2] * sin(omega);
    R[0][2] = ax[0] * ax[2] * (1 - cos(omega)) + ax[1] * sin(omega);
    R[1][0] = ax[0] * ax[1] * (1 - cos(omega)) + ax[2] * sin(omega);
    R[1][1] = cos(omega) + pow(ax[1], 2.0) * (1 - cos(omega));
    R[1][2] = ax[1] * ax[2] * (1 - cos(omega)) - ax[0] * sin(omega);
    R[2][0] = ax[2] * ax[0] * (1 - cos(omega)) - ax[1] * sin(omega);
    R[2][1] = ax[2] * ax[1] * (1 - cos(omega)) + ax[0] * sin(omega);
    R[2][2] = cos(omega) + pow(ax[2], 2.0) * (1 - cos(omega));
  }
  void grain_rot(int G, double R[3][3]) const {
    double ax[3] = {cTheta[4 * (G - 1) + 1], cTheta[4 * (G - 1) + 2], cTheta[4 * (G - 1) + 3]};
    rot_mat(cTheta[4 * (G - 1)], ax, R);
  }

  // S/VoxelsCA.h:78-93
  static void project_point_line(const double* A, const double* x0, const double* x1, double* xp) {
    double n[3], snorm, d, t;
    snorm = m_powh((x1[0] - x0[0]) * (x1[0] - x0[0]) + (x1[1] - x0[1]) * (x1[1] - x0[1]) +
                   (x1[2] - x0[2]) * (x1[2] - x0[2]));
    n[0] = (x1[0] - x0[0]) / snorm;
    n[1] = (x1[1] - x0[1]) / snorm;
    n[2] = (x1[2] - x0[2]) / snorm;
    d = -(n[0] * A[0] + n[1] * A[1] + n[2] * A[2]);
    t = (-d - n[0] * x0[0] - n[1] * x0[1] - n[2] * x0[2]) /
        (n[0] * (x1[0] - x0[0]) + n[1] * (x1[1] - x0[1]) + n[2] * (x1[2] - x0[2]));
    xp[0] = x0[0] + (x1[0] - x0[0]) * t;
    xp[1] = x0[1] + (x1[1] - x0[1]) * t;
    xp[2] = x0[2] + (x1[2] - x0[2]) * t;
  }

  // Active set (S/VoxelsCA.C:262-275): sorted unique {mushy, or liquid below tL} plus their
  // neighbours at or above tL; active-space CSR (S/VoxelsCA.C:304-335): each active cell's
  // active neighbours as active indices, neighbours in ascending id order.
  void active_set(std::vector<int>& indA, std::vector<int>& nptr, std::vector<int>& nidx) const {
    const double tL = g.tL;
    int nb[26], cnt;
    indA.clear();
    for (int j = 0; j < N; ++j) {
      if (vState[j] == 2 || (vState[j] == 1 && TempCurr[j] < tL)) {
        indA.push_back(j);
        nbrs(j, nb, cnt);
        for (int q = 0; q < cnt; ++q)
          if (TempCurr[nb[q]] >= tL && nb[q] < N) indA.push_back(nb[q]);
      }
    }
    std::sort(indA.begin(), indA.end());
    indA.erase(std::unique(indA.begin(), indA.end()), indA.end());
    const int Na = (int)indA.size();
    nptr.assign(Na + 1, 0);
    nidx.clear();
    nidx.reserve(8 * (size_t)Na);
    for (int j = 0; j < Na; ++j) {
      nptr[j] = (int)nidx.size();
      nbrs(indA[j], nb, cnt);
      for (int q = 0; q < cnt; ++q) {
        auto it = std::lower_bound(indA.begin(), indA.end(), nb[q]);
        if (it != indA.end() && *it == nb[q]) nidx.push_back((int)(it - indA.begin()));
      }
    }
    nptr[Na] = (int)nidx.size();
  }

  // S/VoxelsCA.C:232-655 at one rank
  void update_voxels() {
    StepStats s;
    set_liquid3();
    convert_solid1(1);
    const double tL = g.tL;
    std::vector<int> indA, nptr, nidx;
    active_set(indA, nptr, nidx);
    const int Na = (int)indA.size();
    s.Na = Na;
    std::vector<double> T(Na), ExtA(Na), Cen(3 * (size_t)Na);
    std::vector<int> vS(Na), G(Na), vI(Na);
    for (int j = 0; j < Na; ++j) {
      int i1 = indA[j];
      vS[j] = vState[i1]; vI[j] = i1; G[j] = gID[i1]; T[j] = TempCurr[i1]; ExtA[j] = extents[i1];
      for (int q = 0; q < 3; ++q) Cen[3 * (size_t)j + q] = centroidOct[3 * (size_t)i1 + q];
    }
    // random numbers (S/VoxelsCA.C:336-380)
    unsigned int sdloc = seed0 + 32 * g.tInd + 64 * nGrain;
    Engine g1(30 * g.tInd + seed1);
    std::uniform_real_distribution<double> xrand1(0.0, 1.0);
    double rX = g.rNmax * pow(g.dX[0] * 1e6, 3.);
    std::vector<int> vs2ind(Na, 0);
    int vSind2 = 0;
    for (int j = 0; j < Na; ++j)
      if (vS[j] == 2) vs2ind[vSind2++] = j;
    s.n_mushy0 = vSind2;
    int ivlc = vSind2;  // one rank: ivl is the identity (S/VoxelsCA.C:354-364)
    std::vector<double> ruq(Na, 0.0);
    for (int j = 0; j < Na; ++j) ruq[j] = (xrand1(g1) - .5) * g.dX[0] / 100.;
    s.draws = Na;
    const double sq3 = pow(3.0, .5), s23 = pow(2.0 / 3.0, .5);  // folded to CR constants by GCC
    static const double S0[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    static const int sInd[8][3] = {{0, 1, 2}, {1, 2, 3}, {0, 2, 4}, {2, 3, 4},
                                   {0, 1, 5}, {1, 3, 5}, {0, 4, 5}, {3, 4, 5}};
    const int i2 = g.nX[0], i3 = g.nX[0] * g.nX[1];
    int js = 0, j1s = 0, countS = Na;
    double tinc = 0.0;
    double R[3][3];
    while (vSind2 != countS && tinc < DelT) {
      double ruqflag = 1000, DtMin = 1e6;
      countS = vSind2;
      std::vector<double> vhatvec(ivlc, 0.0);
      for (int j = 0; j < ivlc; ++j) {
        int i1 = vs2ind[j];
        if (vS[i1] != 2) continue;
        int i5 = nptr[i1], i6 = nptr[i1 + 1] - i5;
        bool anyliq = false, allcold = true;
        for (int j1 = 0; j1 < i6; ++j1) {
          int i4 = nidx[i5 + j1];
          if (vS[i4] == 1) anyliq = true;
          if (!(T[i4] < tL)) allcold = false;
        }
        if (!(anyliq && allcold)) continue;
        grain_rot(G[i1], R);
        T[i1] >= tL ? vhatvec[j] = 0.0 : vhatvec[j] = g.Avel * m_pow(tL - T[i1], g.nvel);
        for (int j1 = 0; j1 < i6; ++j1) {
          int i4 = nidx[i5 + j1];
          if (vS[i4] != 1) continue;
          int jx2 = vI[i4] / i3, jx1 = (vI[i4] - i3 * jx2) / i2, jx0 = vI[i4] - i3 * jx2 - i2 * jx1;
          double dnx[3];
          dnx[0] = (double(jx0) + .5) * g.dX[0] - Cen[3 * (size_t)i1];
          dnx[1] = (double(jx1) + .5) * g.dX[1] - Cen[3 * (size_t)i1 + 1];
          dnx[2] = (double(jx2) + .5) * g.dX[2] - Cen[3 * (size_t)i1 + 2];
          double d0 = std::fabs(R[0][0] * dnx[0] + R[1][0] * dnx[1] + R[2][0] * dnx[2]);
          double d1 = std::fabs(R[0][1] * dnx[0] + R[1][1] * dnx[1] + R[2][1] * dnx[2]);
          double d2 = std::fabs(R[0][2] * dnx[0] + R[1][2] * dnx[1] + R[2][2] * dnx[2]);
          double dr = shape_factor(dnx) * (d0 + d1 + d2) - ExtA[i1];
          double timeUntil = dr / vhatvec[j];
          double ruqs = .5 * (ruq[i1] + ruq[i4]);
          if (timeUntil <= DtMin) {
            if (timeUntil < DtMin) {
              DtMin = timeUntil; js = i1; j1s = j1; ruqflag = ruqs;
            } else if (ruqs < ruqflag) {
              DtMin = timeUntil; js = i1; j1s = j1; ruqflag = ruqs;
            }
          }
        }
      }
      // one-rank pick (S/VoxelsCA.C:451-474); no candidate leaves DtMin at 1e6
      double Dt = 1e6;
      if (DtMin < 1e6) Dt = DtMin;
      if (Dt >= 1e6) break;
      s.iters++;
      s.draws++;
      if (xrand1(g1) < rX) {  // nucleation (S/VoxelsCA.C:483-505)
        int i1 = nidx[nptr[js] + j1s];
        std::vector<double> q;
        sa.generate(1, sdloc, q);
        vS[i1] = 2;
        vs2ind[vSind2] = i1;
        vSind2 += 1;
        nGrain += 1;
        ivlc += 1;
        G[i1] = nGrain;
        cTheta.insert(cTheta.end(), q.begin(), q.end());
        int jx2 = vI[i1] / i3, jx1 = (vI[i1] - i3 * jx2) / i2, jx0 = vI[i1] - i3 * jx2 - i2 * jx1;
        Cen[3 * (size_t)i1] = (double(jx0) + .5) * g.dX[0];
        Cen[3 * (size_t)i1 + 1] = (double(jx1) + .5) * g.dX[1];
        Cen[3 * (size_t)i1 + 2] = (double(jx2) + .5) * g.dX[2];
        sdloc += 10;
        Dt = 0.0;
        ExtA[i1] = 0.0;
        s.nucleations++;
        s.ev_vox.push_back(indA[i1]);
        s.ev_src.push_back(-1);
        s.ev_dt.push_back(Dt);
      } else {  // growth + capture + re-centring (S/VoxelsCA.C:506-607)
        for (int j = 0; j < (int)vhatvec.size(); ++j) {
          int i1 = vs2ind[j];
          if (vS[i1] != 2) continue;
          ExtA[i1] += vhatvec[j] * Dt;
          ExtA[i1] = std::max(ExtA[i1], 0.0);
        }
        int i1 = nidx[nptr[js] + j1s];
        vS[i1] = 2;
        vs2ind[vSind2] = i1;
        vSind2 += 1;
        ivlc += 1;
        G[i1] = G[js];
        int jx[3], jy[3];
        jx[2] = vI[i1] / i3; jx[1] = (vI[i1] - i3 * jx[2]) / i2; jx[0] = vI[i1] - i3 * jx[2] - i2 * jx[1];
        jy[2] = vI[js] / i3; jy[1] = (vI[js] - i3 * jy[2]) / i2; jy[0] = vI[js] - i3 * jy[2] - i2 * jy[1];
        double l0 = (jx[0] - jy[0]) * g.dX[0], l1 = (jx[1] - jy[1]) * g.dX[1], l2 = (jx[2] - jy[2]) * g.dX[2];
        double l = m_powh(l0 * l0 + l1 * l1 + l2 * l2);
        double dnx[3], locX[3];
        dnx[0] = (double(jx[0]) + .5) * g.dX[0] - Cen[3 * (size_t)js];
        dnx[1] = (double(jx[1]) + .5) * g.dX[1] - Cen[3 * (size_t)js + 1];
        dnx[2] = (double(jx[2]) + .5) * g.dX[2] - Cen[3 * (size_t)js + 2];
        grain_rot(G[js], R);
        locX[0] = R[0][0] * dnx[0] + R[1][0] * dnx[1] + R[2][0] * dnx[2];
        locX[1] = R[0][1] * dnx[0] + R[1][1] * dnx[1] + R[2][1] * dnx[2];
        locX[2] = R[0][2] * dnx[0] + R[1][2] * dnx[1] + R[2][2] * dnx[2];
        const double fshape = shape_factor(dnx);
        double dr = std::fabs(locX[0]) + std::fabs(locX[1]) + std::fabs(locX[2]);
        double sd[6][3];
        for (int q = 0; q < 6; ++q)
          for (int c = 0; c < 3; ++c) sd[q][c] = S0[q][c] * dr;
        int jInd = 4 * std::signbit(locX[2]) + 2 * std::signbit(locX[1]) + std::signbit(locX[0]);
        double dnorm[3];
        for (int q = 0; q < 3; ++q) {
          const double* v = sd[sInd[jInd][q]];
          dnorm[q] = (locX[0] - v[0]) * (locX[0] - v[0]) + (locX[1] - v[1]) * (locX[1] - v[1]) +
                     (locX[2] - v[2]) * (locX[2] - v[2]);
        }
        int k0 = (int)(std::min_element(dnorm, dnorm + 3) - dnorm);
        int k2 = (int)(std::max_element(dnorm, dnorm + 3) - dnorm);
        int k1;
        if (k0 == k2) { k0 = 0; k1 = 1; k2 = 2; }
        else k1 = 3 - k0 - k2;
        const double* v0 = sd[sInd[jInd][k0]];
        const double* v1 = sd[sInd[jInd][k1]];
        const double* v2 = sd[sInd[jInd][k2]];
        double xI[3], xJ[3];
        project_point_line(locX, v0, v1, xI);
        project_point_line(locX, v0, v2, xJ);
        auto dist = [](const double* p, const double* q) {
          return m_powh((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]));
        };
        double d1I = dist(v0, xI), dI2 = dist(v1, xI), d1J = dist(v0, xJ), dJ3 = dist(v2, xJ);
        double L12 = .5 * (std::min(d1I, sq3 * l) + std::min(dI2, sq3 * l));
        double L13 = .5 * (std::min(d1J, sq3 * l) + std::min(dJ3, sq3 * l));
        double Lmud = sq3 * (s23 * std::max(L12, L13));
        double xiL = 1.0;
        ExtA[i1] = fshape * Lmud * xiL;
        const double* s0 = S0[sInd[jInd][k0]];
        dnx[0] = v0[0] - Lmud * s0[0];
        dnx[1] = v0[1] - Lmud * s0[1];
        dnx[2] = v0[2] - Lmud * s0[2];
        locX[0] = R[0][0] * dnx[0] + R[0][1] * dnx[1] + R[0][2] * dnx[2];
        locX[1] = R[1][0] * dnx[0] + R[1][1] * dnx[1] + R[1][2] * dnx[2];
        locX[2] = R[2][0] * dnx[0] + R[2][1] * dnx[1] + R[2][2] * dnx[2];
        Cen[3 * (size_t)i1] = Cen[3 * (size_t)js] + locX[0];
        Cen[3 * (size_t)i1 + 1] = Cen[3 * (size_t)js + 1] + locX[1];
        Cen[3 * (size_t)i1 + 2] = Cen[3 * (size_t)js + 2] + locX[2];
        s.captures++;
        s.ev_vox.push_back(indA[i1]);
        s.ev_src.push_back(indA[js]);
        s.ev_dt.push_back(Dt);
      }
      tinc += Dt;
    }
    s.tinc = tinc;
    // scatter back (S/VoxelsCA.C:633-643)
    for (int j = 0; j < Na; ++j) {
      int i1 = indA[j];
      vState[i1] = vS[j]; gID[i1] = G[j]; extents[i1] = ExtA[j];
      for (int q = 0; q < 3; ++q) centroidOct[3 * (size_t)i1 + q] = Cen[3 * (size_t)j + q];
    }
    convert_solid1(0);
    st = s;
  }

  void clean_layer() {  // S/VoxelsCA.C:800-839 (volume > 0 <=> at least one voxel)
    std::vector<int> cnt(nGrain, 0), remap(nGrain, 0);
    for (int j = 0; j < N; ++j)
      if (gID[j] > 0) cnt[gID[j] - 1] += 1;
    int nG = 0;
    std::vector<double> ct;
    for (int j = 0; j < nGrain; ++j)
      if (cnt[j] > 0) {
        for (int q = 0; q < 4; ++q) ct.push_back(cTheta[4 * j + q]);
        remap[j] = nG + 1;
        nG += 1;
      }
    cTheta = ct;
    nGrain = nG;
    for (int j = 0; j < N; ++j)
      if (gID[j] != 0) gID[j] = remap[gID[j] - 1];
  }

  // AddLayer1's orientation seed, Poisson site count and sites (S/VoxelsCA.C:895-919), drawn
  // from the layer engine; returns the site count
  int draw_layer_sites(std::vector<double>& Xs, unsigned int& sdloc) {
    const int nx = g.nX[0], ny = g.nX[1];
    sdloc = unsigned(double(genlayer()) / double(genlayer.max()) * pow(2.0, 32.0));
    double rate = g.lrate * (g.dX[0] * 1e6) * (g.dX[1] * 1e6) * nx * ny * (g.layerT * 1e6);
    std::poisson_distribution<int> ngd(rate);
    int ng2 = 0;
    while (ng2 == 0) ng2 = ngd(genlayer);
    int iz1 = g.ilaserLoc - g.nZlayer;
    double zloc0 = iz1 * g.dX[2];
    Xs.assign(ng2 * 3, 0.0);
    std::uniform_real_distribution<double> xrand(0.0, 1.0);
    double Lx = nx * g.dX[0], Ly = ny * g.dX[1];
    for (int j = 0; j < ng2; ++j) {
      Xs[3 * j] = xrand(genlayer) * Lx;
      Xs[3 * j + 1] = xrand(genlayer) * Ly;
      Xs[3 * j + 2] = xrand(genlayer) * g.layerT + zloc0;
    }
    return ng2;
  }

  void add_layer1() {  // S/VoxelsCA.C:887-982
    const int nx = g.nX[0], ny = g.nX[1];
    unsigned int sdloc = 0;
    std::vector<double> Xs;
    const int ng2 = draw_layer_sites(Xs, sdloc);
    int iz1 = g.ilaserLoc - g.nZlayer;
    last_sites = Xs;
    const int nGrain_before = nGrain;
    int i3 = nx * ny;
    std::vector<int> itmp(ng2, 0), gtmp((size_t)i3 * g.nZlayer, 0);
    int i2 = 0;
    for (int j3a = iz1; j3a < g.ilaserLoc; ++j3a)
      for (int j2a = 0; j2a < ny; ++j2a)
        for (int j1a = 0; j1a < nx; ++j1a) {
          double dsqc = 1e6;
          for (int jz = 0; jz < ng2; ++jz) {
            double ax = j1a * g.dX[0] - Xs[3 * jz], ay = j2a * g.dX[1] - Xs[3 * jz + 1],
                   az = j3a * g.dX[2] - Xs[3 * jz + 2];
            double dsq = ax * ax + ay * ay + az * az;
            if (dsq < dsqc) { i2 = jz; dsqc = dsq; }
          }
          int jind = i3 * (j3a - iz1) + nx * j2a + j1a;
          if (jind >= 0 && jind < (int)gtmp.size()) gtmp[jind] = i2 + 1;
          itmp[i2] = 1;
        }
    i2 = 0;
    for (int j = 0; j < ng2; ++j)
      if (itmp[j] == 1) { itmp[j] = i2; i2 += 1; }
    if (i2 != ng2)
      for (size_t j = 0; j < gtmp.size(); ++j) gtmp[j] = itmp[gtmp[j] - 1] + 1;
    for (int j = 0; j < N; ++j) {
      int j3 = j / i3;
      if (j3 >= iz1 && j3 < g.ilaserLoc) {
        int j2 = (j - i3 * j3) / nx, j1 = j - i3 * j3 - nx * j2;
        int jind = i3 * (j3 - iz1) + nx * j2 + j1;
        gID[j] = nGrain + gtmp[jind];
        vState[j] = 3;
        centroidOct[3 * (size_t)j] = (double(j1) + .5) * g.dX[0];
        centroidOct[3 * (size_t)j + 1] = (double(j2) + .5) * g.dX[1];
        centroidOct[3 * (size_t)j + 2] = (double(j3) + .5) * g.dX[2];
      }
    }
    nGrain += i2 + 1;
    last_layer[0] = iz1; last_layer[1] = g.ilaserLoc; last_layer[2] = ng2; last_layer[3] = i2;
    last_layer[4] = sdloc; last_layer[5] = nGrain_before;
    std::vector<double> aa;
    sa.generate(i2 + 1, sdloc, aa);
    cTheta.insert(cTheta.end(), aa.begin(), aa.end());
    int izz = nXY * g.ilaserLoc;
    for (int j = 0; j < N; ++j)
      if (j >= izz) { vState[j] = 0; gID[j] = 0; }
  }

  void set_build() {  // S/VoxelsCA.C:76-115 as set membership over grain ids
    int maxg = 0;
    for (int j = 0; j < N; ++j) maxg = std::max(maxg, gID[j]);
    std::vector<char> in1(maxg + 1, 0), in2(maxg + 1, 0);
    for (int j = 0; j < N; ++j)
      if (BuildjID[j] == 1 && gID[j] > 0) in1[gID[j]] = 1;
    for (int j = 0; j < N; ++j)
      if (gID[j] > 0 && in1[gID[j]]) Build[j] = 1;
    for (int j = 0; j < N; ++j)
      if (Build[j] == 1 && gID[j] > 0) in2[gID[j]] = 1;
    for (int j = 0; j < N; ++j)
      if (gID[j] > 0 && in2[gID[j]]) Build[j] = 1;
  }

  // one iteration of S/main1.C:62-93 after the temperature and UpdateVoxels calls
  int finish_step() {
    g.update_laser();
    done = g.indlayer > nlayerTot;
    if (g.inewlayerflg == 1) clean_layer();
    int indOut = TtInd % g.outint;
    if (indOut == 0 || done || (g.inewlayerflg == 1 && g.outNL == 0)) {
      out_steps.push_back(TtInd);
      set_build();
      snapshot();
    }
    g.update_time2(DelT);
    if (g.inewlayerflg == 1) add_layer1();
    if (done) {
      clean_layer();  // S/main1.C:94 after the loop
      snapshot();     // S/main1.C:95-97
    }
    return done ? 1 : 0;
  }
};

}  // namespace orc

// ---------------------------------------------------------------------------
// C interface for the tests (ctypes)
extern "C" {

static std::string g_err;
const char* orc_last_error() { return g_err.c_str(); }
void orc_set_crmath(int on) {
  orc::g_crmath = on;
  orc::g_shape_memo.clear();
}

void* orc_create(const char* deck_text) {
  orc::Sim* s = new orc::Sim();
  if (s->setup(deck_text, g_err)) { delete s; return nullptr; }
  return s;
}
void* orc_create_bare(const char* deck_text) {
  orc::Sim* s = new orc::Sim();
  if (s->setup(deck_text, g_err, true)) { delete s; return nullptr; }
  return s;
}
void orc_destroy(void* h) { delete (orc::Sim*)h; }
void orc_temperature(void* h) { ((orc::Sim*)h)->analytic_temp(); }
void orc_update_voxels(void* h) { ((orc::Sim*)h)->update_voxels(); }
void orc_state_in(void* h) {  // SetLiquid3 + ConvertSolid1(1) only, for kernel tests
  orc::Sim* s = (orc::Sim*)h;
  s->set_liquid3();
  s->convert_solid1(1);
}
void orc_state_out(void* h) { ((orc::Sim*)h)->convert_solid1(0); }
int orc_finish_step(void* h) { return ((orc::Sim*)h)->finish_step(); }
int orc_step(void* h) {
  orc::Sim* s = (orc::Sim*)h;
  s->analytic_temp();
  s->update_voxels();
  return s->finish_step();
}
int orc_done(void* h) { return ((orc::Sim*)h)->done ? 1 : 0; }
// TempField::EASM on the oracle's grid with an explicit laser; returns all_of(T >= tL)
int orc_easm(void* h, double x, double y, double z, double cg, double sg, long* n_integrals) {
  orc::Sim* s = (orc::Sim*)h;
  s->gl_rule();
  long n = 0;
  int r = s->easm(x, y, z, cg, sg, &n);
  if (n_integrals) *n_integrals = n;
  return r;
}
// the oracle's 25-point Gauss-Legendre rule on [-1, 1]
void orc_gl25(double* x, double* w) {
  orc::Sim s;
  s.gl_rule();
  for (int i = 0; i < 25; ++i) { x[i] = s.gl_x[i]; w[i] = s.gl_w[i]; }
}
// output recording: snapshots of gID, vState, Build, TempCurr and cTheta at every output
void orc_record_outputs(void* h, int on) { ((orc::Sim*)h)->record_outputs = on != 0; }
int orc_n_outputs(void* h) { return (int)((orc::Sim*)h)->snaps.size(); }
// k-th snapshot: returns its tInd and sets *ngrain; buffers may be null
int orc_output(void* h, int k, int* gid, int* vs, int* build, double* T, double* ct, int* ngrain) {
  const orc::Sim::Snapshot& o = ((orc::Sim*)h)->snaps[k];
  const size_t n = o.gID.size();
  if (gid) std::memcpy(gid, o.gID.data(), n * 4);
  if (vs) std::memcpy(vs, o.vState.data(), n * 4);
  if (build) std::memcpy(build, o.Build.data(), n * 4);
  if (T) std::memcpy(T, o.T.data(), n * 8);
  if (ct) std::memcpy(ct, o.cTheta.data(), o.cTheta.size() * 8);
  if (ngrain) *ngrain = (int)(o.cTheta.size() / 4);
  return o.tInd;
}
void orc_dims(void* h, int* d) {
  orc::Sim* s = (orc::Sim*)h;
  d[0] = s->g.nX[0]; d[1] = s->g.nX[1]; d[2] = s->g.nX[2];
}
int orc_ngrain(void* h) { return ((orc::Sim*)h)->nGrain; }
void orc_get_ctheta(void* h, double* out) {
  orc::Sim* s = (orc::Sim*)h;
  std::memcpy(out, s->cTheta.data(), sizeof(double) * s->cTheta.size());
}
void orc_set_ctheta(void* h, const double* in, int ngrain) {
  orc::Sim* s = (orc::Sim*)h;
  s->cTheta.assign(in, in + 4 * (size_t)ngrain);
  s->nGrain = ngrain;
}
// any pointer may be null
void orc_get_fields(void* h, int* gid, int* vs, double* T, double* ext, double* cen, int* bj, int* build) {
  orc::Sim* s = (orc::Sim*)h;
  size_t n = s->N;
  if (gid) std::memcpy(gid, s->gID.data(), n * 4);
  if (vs) std::memcpy(vs, s->vState.data(), n * 4);
  if (T) std::memcpy(T, s->TempCurr.data(), n * 8);
  if (ext) std::memcpy(ext, s->extents.data(), n * 8);
  if (cen) std::memcpy(cen, s->centroidOct.data(), 3 * n * 8);
  if (bj) std::memcpy(bj, s->BuildjID.data(), n * 4);
  if (build) std::memcpy(build, s->Build.data(), n * 4);
}
void orc_set_fields(void* h, const int* gid, const int* vs, const double* T, const double* ext, const double* cen,
                    const int* bj) {
  orc::Sim* s = (orc::Sim*)h;
  size_t n = s->N;
  if (gid) std::memcpy(s->gID.data(), gid, n * 4);
  if (vs) std::memcpy(s->vState.data(), vs, n * 4);
  if (T) std::memcpy(s->TempCurr.data(), T, n * 8);
  if (ext) std::memcpy(s->extents.data(), ext, n * 8);
  if (cen) std::memcpy(s->centroidOct.data(), cen, 3 * n * 8);
  if (bj) std::memcpy(s->BuildjID.data(), bj, n * 4);
}
// scalars: time, tInd, isp, ilaserLoc, indlayer, inewscanflg, inewlayerflg, gth, nlayerTot, Nzh, NpT, Nsd, Ntd, TtInd
void orc_get_scalars(void* h, double* o) {
  orc::Sim* s = (orc::Sim*)h;
  o[0] = s->g.time; o[1] = s->g.tInd; o[2] = s->g.isp; o[3] = s->g.ilaserLoc; o[4] = s->g.indlayer;
  o[5] = s->g.inewscanflg; o[6] = s->g.inewlayerflg; o[7] = s->g.gth; o[8] = s->nlayerTot; o[9] = s->Nzh;
  o[10] = s->g.NpT; o[11] = s->g.Nsd; o[12] = s->g.Ntd; o[13] = s->TtInd;
}
void orc_set_laser(void* h, int isp, int ilaserLoc, int tInd) {
  orc::Sim* s = (orc::Sim*)h;
  s->g.isp = isp; s->g.ilaserLoc = ilaserLoc; s->g.tInd = tInd;
  s->g.time = tInd * s->DelT;
}
void orc_set_grid(void* h, double time, double gth, int indlayer, int inewscanflg, int inewlayerflg) {
  orc::Sim* s = (orc::Sim*)h;
  s->g.time = time; s->g.gth = gth; s->g.indlayer = indlayer;
  s->g.inewscanflg = inewscanflg; s->g.inewlayerflg = inewlayerflg;
}
// stats of the last UpdateVoxels: Na, mushy at loop start, iterations, captures, nucleations, draws
void orc_stats(void* h, long* o, double* tinc) {
  orc::Sim* s = (orc::Sim*)h;
  o[0] = s->st.Na; o[1] = s->st.n_mushy0; o[2] = s->st.iters; o[3] = s->st.captures;
  o[4] = s->st.nucleations; o[5] = s->st.draws;
  if (tinc) *tinc = s->st.tinc;
}
// laser inputs of the next temperature evaluation, as TempField.C:146-171 forms them:
// x, y, z (grid frame), dirp, gth, ilaserLoc
void orc_laser_inputs(void* h, double* o) {
  orc::Sim* s = (orc::Sim*)h;
  const orc::Grid& g = s->g;
  const std::vector<int>& isv = s->ispvec;
  o[0] = g.lcoor[2 * isv[g.isp]] - g.offset[0];
  o[1] = g.lcoor[2 * isv[g.isp] + 1] - g.offset[1];
  o[2] = g.ilaserLoc * g.dX[2] - g.offset[2];
  if (fmod(g.isp, g.Nsd) == 0) o[3] = (g.lcoor[2 * isv[g.isp + 1]] - g.lcoor[2 * isv[g.isp]]);
  else o[3] = (g.lcoor[2 * isv[g.isp]] - g.lcoor[2 * isv[g.isp - 1]]);
  o[4] = g.gth;
  o[5] = g.ilaserLoc;
}
// event log of the last UpdateVoxels (grid indices); returns the number of events
long orc_events(void* h, int* vox, int* src, double* dt, long cap) {
  orc::Sim* s = (orc::Sim*)h;
  long n = (long)s->st.ev_vox.size();
  for (long q = 0; q < n && q < cap; ++q) {
    vox[q] = s->st.ev_vox[q];
    src[q] = s->st.ev_src[q];
    dt[q] = s->st.ev_dt[q];
  }
  return n;
}
// standalone sampler and engine draws, for known-answer tests
// the active set and its CSR for the current state (call orc_state_in first to apply
// SetLiquid3 and ConvertSolid1(1) as UpdateVoxels does); returns Na, sets *nnz.
// Arrays may be null to query the sizes.
long orc_active_set(void* h, int* indA, int* nptr, int* nidx, long* nnz) {
  std::vector<int> a, p, x;
  ((orc::Sim*)h)->active_set(a, p, x);
  if (indA) std::copy(a.begin(), a.end(), indA);
  if (nptr) std::copy(p.begin(), p.end(), nptr);
  if (nidx) std::copy(x.begin(), x.end(), nidx);
  if (nnz) *nnz = (long)x.size();
  return (long)a.size();
}
// sites (x, y, z per site) of the last Voronoi tessellation; returns the site count
long orc_last_sites(void* h, double* xs) {
  const orc::Sim* s = (const orc::Sim*)h;
  if (xs) std::copy(s->last_sites.begin(), s->last_sites.end(), xs);
  return (long)(s->last_sites.size() / 3);
}
// The initial state's random draws of a deck, without the Voronoi labelling (for checks at
// sizes where the oracle's brute-force labelling is too slow): the base plate's sites and
// orientation seed, then the first powder layer's sites and orientation seed.  counts gets
// {plate sites, layer sites, Nzh, first layer's iz1, first layer's ilaserLoc}, seeds gets
// {plate seed, layer seed}.  Arrays may be null to query the counts.
void orc_initial_sites(const char* deck, double* plate, double* layer, long* counts, unsigned int* seeds) {
  orc::Sim s;
  std::string err;
  if (s.setup(deck, err, true)) {
    counts[0] = -1;
    return;
  }
  std::vector<double> xp, xl;
  const int ng = s.draw_plate_sites(xp);
  unsigned int sdl = 0;
  const int ng2 = s.draw_layer_sites(xl, sdl);
  counts[0] = ng;
  counts[1] = ng2;
  counts[2] = s.Nzh;
  counts[3] = s.g.ilaserLoc - s.g.nZlayer;
  counts[4] = s.g.ilaserLoc;
  if (seeds) {
    seeds[0] = orc::Sim::plate_sdloc();
    seeds[1] = sdl;
  }
  if (plate) std::copy(xp.begin(), xp.end(), plate);
  if (layer) std::copy(xl.begin(), xl.end(), layer);
}

// the last AddLayer1: iz1, ilaserLoc, sites drawn, labels used, orientation seed, nGrain before
void orc_last_layer(void* h, long* out6) {
  const orc::Sim* s = (const orc::Sim*)h;
  for (int q = 0; q < 6; ++q) out6[q] = s->last_layer[q];
}
void orc_set_build(void* h) { ((orc::Sim*)h)->set_build(); }
void orc_clean_layer(void* h) { ((orc::Sim*)h)->clean_layer(); }
void orc_convert_solid(void* h, int iswitch) { ((orc::Sim*)h)->convert_solid1(iswitch); }
void orc_set_liquid(void* h) { ((orc::Sim*)h)->set_liquid3(); }
void orc_set_build_flags(void* h, const int* build) {
  orc::Sim* s = (orc::Sim*)h;
  if (build) std::copy(build, build + s->N, s->Build.begin());
}

void orc_generate_samples(int n, unsigned int seed, double* out) {
  orc::Orient o;
  std::vector<double> v;
  o.generate(n, seed, v);
  std::memcpy(out, v.data(), v.size() * 8);
}
void orc_uniform_stream(unsigned long seed, int n, double* out) {
  orc::Engine e(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (int i = 0; i < n; ++i) out[i] = u(e);
}
// capture-geometry kernels in isolation (rotation, D factor) for unit tests
void orc_rot_mat(double omega, const double* ax, double* R9) {
  double R[3][3];
  orc::Sim::rot_mat(omega, ax, R);
  for (int i = 0; i < 9; ++i) R9[i] = R[i / 3][i % 3];
}
void orc_shape_factor_n(const double* dnx, int n, double* out);
// the growth velocity's pow in the current math mode
void orc_pow_n(const double* x, const double* y, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = orc::m_pow(x[i], y[i]);
}
double orc_shape_factor(const double* dnx) { return orc::shape_factor_raw(dnx); }
void orc_shape_factor_n(const double* dnx, int n, double* out) {
  for (int i = 0; i < n; ++i) out[i] = orc_shape_factor(dnx + 3 * i);
}
}
