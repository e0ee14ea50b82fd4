// h5out.cpp -- VoxelsCA::WriteToHDF1 (S/VoxelsCA.C:1592-1856): the reference's output file.
//
// The reference writes through ADIOS2's HDF5 engine: one group per step ("Step0") holding
// dims int32[3], VoxelDX float32[3], gID / vState / Build int32[N], Temperature float32[N],
// IPFz / IPFx / IPFy float32[3N] and angleAxis float32[4 nGrain], all 1-D, x fastest.
// This writer produces the same group, names, types and shapes with the HDF5 C library,
// which is loaded at run time (dlopen) so the library itself has no link dependency on it.
//
// IPF colours depend on a voxel's grain only, so they are evaluated once per grain, in
// double with glibc exactly as the reference's per-voxel loop writes them (including the
// y-direction hue that reuses the x-direction H, S/VoxelsCA.C:1763), and gathered per voxel.
#include <dlfcn.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "h5out.h"

namespace amcafe {

// ---------------------------------------------------------------------------------
// IPF colouring (S/VoxelsCA.C:1597-1817)
namespace {
struct IpfConst {
  double y0, x0, sMax, hoff;
};
IpfConst ipf_const() {
  double tri[2][3];
  tri[0][0] = 0.0;
  tri[0][1] = 2. / pow(2, .5) / (1. + 1. / pow(2, .5));
  tri[0][2] = 2. / pow(3, .5) / (1. + 1. / pow(3, .5));
  tri[1][0] = 0.0;
  tri[1][1] = 0.0;
  tri[1][2] = 2. / pow(3, .5) / (1. + 1. / pow(3, .5));
  const double m = tan(1. / 2. * atan2(tri[1][2], tri[0][2]));
  const double a = pow(pow(tri[1][2] - tri[1][1], 2.) + pow(tri[0][2] - tri[0][1], 2.), .5);
  const double b = pow(pow(tri[0][1], 2.) + pow(tri[1][1], 2.), 0.5);
  const double c = pow(pow(tri[0][2], 2.) + pow(tri[1][2], 2.), 0.5);
  IpfConst k;
  k.y0 = 1. / 2. * pow((b + c - a) * (c + a - b) * (a + b - c) / (a + b + c), .5);
  k.x0 = k.y0 / m;
  k.sMax = pow(pow(k.x0, 2.) + pow(k.y0, 2.), .5);
  k.hoff = atan((tri[1][2] - k.y0) / (tri[0][2] - k.x0)) * 180 / M_PI;
  return k;
}

// |R^T v| sorted, first two swapped, last clipped at 1, projected (S/VoxelsCA.C:1646-1653)
void project(const double R[3][3], const double v[3], double& xp, double& yp) {
  double vCD[3];
  vCD[0] = std::fabs(R[0][0] * v[0] + R[1][0] * v[1] + R[2][0] * v[2]);
  vCD[1] = std::fabs(R[0][1] * v[0] + R[1][1] * v[1] + R[2][1] * v[2]);
  vCD[2] = std::fabs(R[0][2] * v[0] + R[1][2] * v[1] + R[2][2] * v[2]);
  std::sort(vCD, vCD + 3);
  std::swap(vCD[0], vCD[1]);
  vCD[2] = std::min(vCD[2], 1.0);
  xp = 2. * vCD[0] / (1. + vCD[2]);
  yp = 2. * vCD[1] / (1. + vCD[2]);
}

// HSV -> RGB with the reference's six cases; H is the hue already reduced to [0, 360)
void hsv(const IpfConst& k, double H, double xp, double yp, float* out) {
  const double V = 1.;
  double S = pow(pow(xp - k.x0, 2.) + pow(yp - k.y0, 2.), .5);
  S = S / k.sMax * 0.8 + 0.2;
  const int icase = (int)floor(H / 60.0);
  const double ff = H / 6.0 - icase;
  const double p = V * (1. - S), q = V * (1. - S * ff), t = V * (1. - (S * (1. - ff)));
  double r = 0, g = 0, b = 0;
  if (S <= 0.0) {
    r = g = b = 0.0;
  } else {
    switch (icase) {
      case 0: r = V; g = t; b = p; break;
      case 1: r = q; g = V; b = p; break;
      case 2: r = p; g = V; b = t; break;
      case 3: r = p; g = q; b = V; break;
      case 4: r = t; g = p; b = V; break;
      case 5: r = V; g = p; b = q; break;
      default: break;  // the reference leaves the zero initial value
    }
  }
  out[0] = (float)r;
  out[1] = (float)g;
  out[2] = (float)b;
}

double reduce_hue(double H) { return H >= 360.0 ? 0.0 : H; }
}  // namespace

void ipf_colours(const double* cTheta, int nGrain, float* rgb9) {
  const IpfConst k = ipf_const();
  const double vBD[3] = {0.0, 0.0, 1.0}, vX[3] = {1.0, 0.0, 0.0}, vY[3] = {0.0, 1.0, 0.0};
  for (int g = 0; g < nGrain; ++g) {
    const double omega = cTheta[4 * (size_t)g];
    const double ax[3] = {cTheta[4 * (size_t)g + 1], cTheta[4 * (size_t)g + 2], cTheta[4 * (size_t)g + 3]};
    double R[3][3];  // S/VoxelsCA.C:1634-1642 (local -> global)
    R[0][0] = cos(omega) + pow(ax[0], 2.0) * (1 - cos(omega));
    R[0][1] = ax[0] * ax[1] * (1 - cos(omega)) - ax[2] * sin(omega);
    R[0][2] = ax[0] * ax[2] * (1 - cos(omega)) + ax[1] * sin(omega);
    R[1][0] = ax[0] * ax[1] * (1 - cos(omega)) + ax[2] * sin(omega);
    R[1][1] = cos(omega) + pow(ax[1], 2.0) * (1 - cos(omega));
    R[1][2] = ax[1] * ax[2] * (1 - cos(omega)) - ax[0] * sin(omega);
    R[2][0] = ax[2] * ax[0] * (1 - cos(omega)) - ax[1] * sin(omega);
    R[2][1] = ax[2] * ax[1] * (1 - cos(omega)) + ax[0] * sin(omega);
    R[2][2] = cos(omega) + pow(ax[2], 2.0) * (1 - cos(omega));
    float* o = rgb9 + 9 * (size_t)g;
    double xp, yp, H;
    // build direction (IPFz)
    project(R, vBD, xp, yp);
    H = atan((yp - k.y0) / (xp - k.x0)) * 180. / M_PI;
    if (xp < k.x0) H += 180;
    H = reduce_hue(H + 240 - k.hoff);
    hsv(k, H, xp, yp, o);
    // x direction (IPFx)
    project(R, vX, xp, yp);
    H = atan((yp - k.y0) / (xp - k.x0)) * 180. / M_PI;
    if (xp < k.x0) H += 180;
    H = reduce_hue(H + 240 - k.hoff);
    hsv(k, H, xp, yp, o + 3);
    // y direction (IPFy): the hue continues from the x direction's (S/VoxelsCA.C:1763)
    project(R, vY, xp, yp);
    H = reduce_hue(H + 240 - k.hoff);
    hsv(k, H, xp, yp, o + 6);
  }
}

// ---------------------------------------------------------------------------------
// HDF5 through dlopen.  Types as in HDF5 >= 1.10 (hid_t is 64-bit there).
namespace {
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
const hid_t kDefault = 0;          // H5P_DEFAULT
const unsigned kAccTrunc = 0x0002u;  // H5F_ACC_TRUNC
const hid_t kAll = 0;              // H5S_ALL

struct H5 {
  void* so = nullptr;
  herr_t (*open)();
  herr_t (*libversion)(unsigned*, unsigned*, unsigned*);
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*gcreate)(hid_t, const char*, hid_t, hid_t, hid_t);
  herr_t (*gclose)(hid_t);
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
  herr_t (*sclose)(hid_t);
  hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  herr_t (*dclose)(hid_t);
  hid_t (*acreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t);
  herr_t (*awrite)(hid_t, hid_t, const void*);
  herr_t (*aclose)(hid_t);
  hid_t *t_int, *t_float, *t_i32le, *t_f32le;
};

template <class F>
bool sym(void* so, const char* name, F& f) {
  f = reinterpret_cast<F>(dlsym(so, name));
  return f != nullptr;
}

bool load(H5& h, std::string& err) {
  static H5 cache;
  static bool ok = false;
  if (ok) {
    h = cache;
    return true;
  }
  const char* env = getenv("AMCAFE_HDF5_LIB");
  const char* names[] = {env, "libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
                         "/opt/conda/lib/libhdf5.so"};
  for (const char* n : names) {
    if (!n || !*n) continue;
    cache.so = dlopen(n, RTLD_NOW | RTLD_LOCAL);
    if (cache.so) break;
  }
  if (!cache.so) {
    err = "the HDF5 shared library (1.10 or later) was not found; set AMCAFE_HDF5_LIB";
    return false;
  }
  void* so = cache.so;
  bool all = sym(so, "H5open", cache.open) && sym(so, "H5get_libversion", cache.libversion) &&
             sym(so, "H5Fcreate", cache.fcreate) && sym(so, "H5Fclose", cache.fclose) &&
             sym(so, "H5Gcreate2", cache.gcreate) && sym(so, "H5Gclose", cache.gclose) &&
             sym(so, "H5Screate_simple", cache.screate_simple) && sym(so, "H5Sclose", cache.sclose) &&
             sym(so, "H5Dcreate2", cache.dcreate) && sym(so, "H5Dwrite", cache.dwrite) &&
             sym(so, "H5Dclose", cache.dclose) && sym(so, "H5Acreate2", cache.acreate) &&
             sym(so, "H5Awrite", cache.awrite) && sym(so, "H5Aclose", cache.aclose) &&
             sym(so, "H5T_NATIVE_INT_g", cache.t_int) && sym(so, "H5T_NATIVE_FLOAT_g", cache.t_float) &&
             sym(so, "H5T_STD_I32LE_g", cache.t_i32le) && sym(so, "H5T_IEEE_F32LE_g", cache.t_f32le);
  if (!all) {
    err = "the HDF5 shared library lacks an expected symbol";
    return false;
  }
  unsigned maj = 0, mi = 0, rel = 0;
  if (cache.open() < 0 || cache.libversion(&maj, &mi, &rel) < 0 || maj < 1 || (maj == 1 && mi < 10)) {
    err = "HDF5 1.10 or later is needed (64-bit identifiers)";
    return false;
  }
  ok = true;
  h = cache;
  return true;
}

struct Writer {
  Writer(H5& h_, hid_t g_) : h(h_), grp(g_), ok(g_ >= 0) {}
  H5& h;
  hid_t grp;
  bool ok;
  template <class T>
  void put(const char* name, const T* data, size_t n, bool is_float) {
    if (!ok) return;
    const hsize_t dim = (hsize_t)n;
    hid_t sp = h.screate_simple(1, &dim, nullptr);
    hid_t ds = sp < 0 ? -1 : h.dcreate(grp, name, is_float ? *h.t_f32le : *h.t_i32le, sp, kDefault, kDefault, kDefault);
    if (ds < 0 || (n && h.dwrite(ds, is_float ? *h.t_float : *h.t_int, kAll, kAll, kDefault, data) < 0)) ok = false;
    if (ds >= 0) h.dclose(ds);
    if (sp >= 0) h.sclose(sp);
  }
};
}  // namespace

int write_hdf5(const char* path, const int nX[3], const double dX[3], const int* gID, const int* vState,
               const int* Build, const double* TempCurr, const double* cTheta, int nGrain, std::string& err) {
  H5 h;
  if (!load(h, err)) return -1;
  const size_t N = (size_t)nX[0] * nX[1] * nX[2];
  std::vector<float> rgb((size_t)9 * std::max(nGrain, 0));
  ipf_colours(cTheta, nGrain, rgb.data());
  std::vector<float> T(N), z(3 * N, 0.f), x(3 * N, 0.f), y(3 * N, 0.f), cth(4 * (size_t)std::max(nGrain, 0));
  for (size_t j = 0; j < N; ++j) {
    T[j] = (float)TempCurr[j];
    const int g = gID[j];
    if (g < 1 || g > nGrain) continue;  // gID < 1: black (S/VoxelsCA.C:1619-1628)
    const float* c = &rgb[9 * (size_t)(g - 1)];
    for (int q = 0; q < 3; ++q) {
      z[3 * j + q] = c[q];
      x[3 * j + q] = c[3 + q];
      y[3 * j + q] = c[6 + q];
    }
  }
  for (size_t j = 0; j < cth.size(); ++j) cth[j] = (float)cTheta[j];
  const float dxf[3] = {(float)dX[0], (float)dX[1], (float)dX[2]};
  hid_t f = h.fcreate(path, kAccTrunc, kDefault, kDefault);
  if (f < 0) {
    err = std::string("cannot create ") + path;
    return -1;
  }
  hid_t g = h.gcreate(f, "Step0", kDefault, kDefault, kDefault);
  Writer w(h, g);
  w.put("dims", nX, 3, false);
  w.put("VoxelDX", dxf, 3, true);
  w.put("gID", gID, N, false);
  w.put("vState", vState, N, false);
  w.put("Build", Build, N, false);
  w.put("Temperature", T.data(), N, true);
  w.put("IPFz", z.data(), 3 * N, true);
  w.put("IPFx", x.data(), 3 * N, true);
  w.put("IPFy", y.data(), 3 * N, true);
  w.put("angleAxis", cth.data(), cth.size(), true);
  if (w.ok) {  // the step count ADIOS2's HDF5 engine leaves on the root group
    const int one = 1;
    hid_t sp = h.screate_simple(0, nullptr, nullptr);
    hid_t a = sp < 0 ? -1 : h.acreate(f, "NumSteps", *h.t_i32le, sp, kDefault, kDefault);
    if (a < 0 || h.awrite(a, *h.t_int, &one) < 0) w.ok = false;
    if (a >= 0) h.aclose(a);
    if (sp >= 0) h.sclose(sp);
  }
  if (g >= 0) h.gclose(g);
  if (h.fclose(f) < 0) w.ok = false;
  if (!w.ok) {
    err = std::string("HDF5 write failed for ") + path;
    return -1;
  }
  return 0;
}

}  // namespace amcafe
