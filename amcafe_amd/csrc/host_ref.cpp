// host_ref.cpp -- see host_ref.h.  Every routine cites the reference lines it
// restates; expressions keep the reference's operation order because the compiler
// (g++, -O3, no -march, -ffp-contract=off) must round them exactly as it does there.
#include "host_ref.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace amcafe {

// ---------------------------------------------------------------------------------
// Deck grammar (S/Grid.C:165-318): "#" starts a comment that runs to the end of the
// line (the newline is kept), keywords are matched case-insensitively, values are
// read with stod/stoi, unknown tokens are skipped.  The reference's reader stores
// the last character of the file twice (the failed get() at end of file leaves it
// in place, S/Grid.C:172-188), so a deck must end in whitespace, as its own decks
// note; that behaviour is kept.
int Deck::parse(const std::string& raw, std::string& err) {
  std::string text;
  text.reserve(raw.size() + 1);
  char c = 0;
  for (size_t i = 0; i < raw.size();) {
    c = raw[i++];
    if (c == '#')
      while (c != '\n' && i < raw.size()) c = raw[i++];
    text.push_back(c);
  }
  if (!raw.empty()) text.push_back(c);
  std::istringstream in(text);
  std::string tok;
  auto val = [&](const char* key) -> std::string {
    std::string v;
    if (!(in >> v)) throw std::runtime_error(std::string("deck: missing value after ") + key);
    return v;
  };
  try {
    while (in >> tok) {
      std::string k = tok;
      std::transform(k.begin(), k.end(), k.begin(), [](unsigned char s) { return (char)std::tolower(s); });
      if (k == "dx") {
        dX.assign(3, std::stod(val("dx")));
      } else if (k == "bmv") {
        bmV = std::stod(val("bmV"));
      } else if (k == "bmp") {
        bmP = std::stod(val("bmP"));
      } else if (k == "bmdelt") {
        bmDelT = std::stod(val("bmDelT"));
        has_bmdelt = true;
      } else if (k == "nx") {
        nX.resize(3);
        for (int q = 0; q < 3; ++q) nX[q] = std::stoi(val("nX"));
      } else if (k == "r") {
        rNmax = std::stod(val("r"));
      } else if (k == "lt") {
        layerT = std::stod(val("lT"));
      } else if (k == "tl") {
        tL = std::stod(val("tL"));
      } else if (k == "ts") {
        tS = std::stod(val("tS"));
      } else if (k == "bhatch") {
        bhatch = std::stod(val("bhatch"));
        has_bhatch = true;
      } else if (k == "mu") {
        mu = std::stod(val("mu"));
      } else if (k == "bph") {
        bpH = std::stod(val("bpH"));
      } else if (k == "lrate") {
        lrate = std::stod(val("lrate"));
        has_lrate = true;
      } else if (k == "outnl") {
        outNL = std::stoi(val("outNL"));
      } else if (k == "ntsd") {
        nTsd = std::stoi(val("nTsd"));
      } else if (k == "offset") {
        for (int q = 0; q < 3; ++q) offset[q] = std::stod(val("offset"));
      } else if (k == "meltparam") {
        meltparam.resize(4);
        for (int q = 0; q < 4; ++q) meltparam[q] = std::stod(val("meltparam"));
      } else if (k == "patternid") {
        patternID = std::stoi(val("patternID"));
      } else if (k == "outint") {
        outint = std::stoi(val("outint"));
      } else if (k == "gridsize") {
        gsize[0] = std::stod(val("gridsize"));
        gsize[1] = std::stod(val("gridsize"));
      } else if (k == "gridtheta") {
        gth0 = std::stod(val("gridtheta"));
      } else if (k == "dendritegrowth2") {
        Avel = std::stod(val("dendritegrowth2"));
        nvel = std::stod(val("dendritegrowth2"));
      }
      // "easm" is compared against the upper-case literal in the reference
      // (S/Grid.C:276) after lower-casing, so it never matches: ignored here too.
    }
  } catch (const std::exception& e) {
    err = std::string("deck: ") + e.what();
    return -1;
  }
  if (!has_bmdelt) {
    err = "deck must set bmDelT (the reference's default reads bmV before it is assigned, S/Grid.C:57)";
    return -1;
  }
  if (!has_lrate) {
    err = "deck must set lrate (the reference leaves it uninitialised)";
    return -1;
  }
  if (!has_bhatch) bhatch = 1.53 * meltparam[2];  // S/Grid.C:56
  if (nX[0] <= 0 || nX[1] <= 0 || nX[2] <= 0 || dX[0] <= 0) {
    err = "deck: nX and dx must be positive";
    return -1;
  }
  return 0;
}

void Deck::rotate_grid() {  // S/Grid.C:93-103, 150-159
  double gmid[2] = {LX[0] / 2., LX[1] / 2.};
  for (int j2 = 0; j2 < Ntd; ++j2)
    for (int j1 = 0; j1 < Nsd; ++j1) {
      int k = Nsd * j2 + j1;
      lcoor2[2 * k] = cos(gth) * (lcoor[2 * k] - gmid[0]) - sin(gth) * (lcoor[2 * k + 1] - gmid[1]) + gmid[0];
      lcoor2[2 * k + 1] = sin(gth) * (lcoor[2 * k] - gmid[0]) + cos(gth) * (lcoor[2 * k + 1] - gmid[1]) + gmid[1];
    }
}

void Deck::init() {
  // S/Grid.C:69-108
  LX = {nX[0] * dX[0], nX[1] * dX[1], nX[2] * dX[2]};
  if (bpH < std::numeric_limits<double>::epsilon()) bpH = layerT;
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
  gth += gth0;
  for (int j2 = 0; j2 < Ntd; ++j2)
    for (int j1 = 0; j1 < Nsd; ++j1) {
      int k = Nsd * j2 + j1;
      lcoor[2 * k] = j1 * bmDelT * bmV + xlmin;
      lcoor[2 * k + 1] = j2 * bhatch + ylmin;
    }
  rotate_grid();
  gbox[0] = -bhatch / 2.;
  gbox[1] = LX[0] + bhatch / 2.;
  gbox[2] = -bhatch / 2.;
  gbox[3] = LX[1] + bhatch / 2.;
  // scan order (S/TempField.C:33-50): patterns 1/3 in order, 2/4 serpentine
  ispvec.assign(NpT, 0);
  if (patternID == 1 || patternID == 3) std::iota(ispvec.begin(), ispvec.end(), 0);
  if (patternID == 2 || patternID == 4)
    for (int j = 0; j < Ntd; ++j) {
      int k = Nsd * j;
      if (fmod(j, 2) == 0) std::iota(ispvec.begin() + k, ispvec.begin() + k + Nsd, k);
      else
        for (int j1 = 0; j1 < Nsd; ++j1) ispvec[k + j1] = k + Nsd - 1 - j1;
    }
  // melt-pool semi-axes (S/TempField.C:71-76)
  a1[0] = meltparam[0];
  a1[1] = meltparam[2];
  a1[2] = meltparam[3];
  a1[3] = meltparam[1];
  a1[4] = a1[1];
  a1[5] = a1[2];
}

void Deck::update_laser() {  // S/Grid.C:110-164
  int itmp, iflg = 0, irep = 0;
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
        double x = lcoor2[2 * itmp], y = lcoor2[2 * itmp + 1];
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
        gth += gth0;
        rotate_grid();
      }
    }
  }
  if (irep == 1) inewlayerflg = 0;
}

void Deck::laser_inputs(double* x, double* y, double* z, double* dirp, double* cg, double* sg) const {
  *x = lcoor[2 * ispvec[isp]] - offset[0];  // S/TempField.C:152-154
  *y = lcoor[2 * ispvec[isp] + 1] - offset[1];
  *z = ilaserLoc * dX[2] - offset[2];
  if (fmod(isp, Nsd) == 0) *dirp = (lcoor[2 * ispvec[isp + 1]] - lcoor[2 * ispvec[isp]]);  // :167-171
  else *dirp = (lcoor[2 * ispvec[isp]] - lcoor[2 * ispvec[isp - 1]]);
  *cg = cos(gth);
  *sg = sin(gth);
}

bool Deck::laser_outside_box() const {
  double x = lcoor2[2 * ispvec[isp]] - offset[0], y = lcoor2[2 * ispvec[isp] + 1] - offset[1];
  return x < gbox[0] || x > gbox[1] || y < gbox[2] || y > gbox[3];
}

void Deck::end_of_scan(int any_melt) {
  if (!laser_outside_box()) return;
  if (any_melt == 0 || fmod(isp + 1, Nsd) == 0) inewscanflg = 1;
  if (isp == (NpT - 1)) {
    inewscanflg = 1;
    inewlayerflg = 1;
    isp = 0;
    indlayer += 1;
  }
}

// ---------------------------------------------------------------------------------
// SampleOrientation: uniform orientations on SO(3) by the cube -> ball area-
// preserving map (S/SampleOrientation.C:16-161).  Constant pow() expressions are
// written as in the reference so the compiler folds them to the same doubles.
Orientation::Orientation() {
  r = 1;
  R1 = pow(3.0 * M_PI / 4.0, 1.0 / 3.0) * r;
  a = pow(2.0 * M_PI / 3.0, 1.0 / 2.0) * R1;
  aprime = pow(4.0 * M_PI / 3.0, 1.0 / 3.0) * R1;
  const double c[7] = {-0.500009615, -0.024866061, -0.004549382, 0.000511867, -0.001650083, 0.000759335,
                       -0.000204042};
  for (int i = 0; i < 7; ++i) coeffA[i] = c[i];
  beta = pow(M_PI / 6.0, 1.0 / 2.0);
}

void Orientation::map2(const double* p, double* o) const {  // S/SampleOrientation.C:68-97
  double x = p[0], y = p[1], z = p[2], T1 = 0, T2 = 0;
  if (fabs(y) <= fabs(x)) {
    if (fabs(x) < 1e-10) {
      T1 = 0.0;
      T2 = 0.0;
    } else {
      T1 = pow(2.0, 1.0 / 2.0) * cos(y * M_PI / 12 / x) - 1;
      T2 = pow(2.0, 1.0 / 2.0) * sin(y * M_PI / 12 / x);
      T1 *= pow(2.0, 1.0 / 4.0) * x / (beta * pow(pow(2.0, 1.0 / 2.0) - cos(y * M_PI / 12 / x), 1.0 / 2.0));
      T2 *= pow(2.0, 1.0 / 4.0) * x / (beta * pow(pow(2.0, 1.0 / 2.0) - cos(y * M_PI / 12 / x), 1.0 / 2.0));
    }
  }
  if (fabs(x) <= fabs(y)) {
    if (fabs(y) < 1e-10) {
      T1 = 0.0;
      T2 = 0.0;
    } else {
      T1 = pow(2.0, 1.0 / 2.0) * sin(x * M_PI / 12 / y);
      T2 = pow(2.0, 1.0 / 2.0) * cos(x * M_PI / 12 / y) - 1;
      T1 *= pow(2.0, 1.0 / 4.0) * y / (beta * pow(pow(2.0, 1.0 / 2.0) - cos(x * M_PI / 12 / y), 1.0 / 2.0));
      T2 *= pow(2.0, 1.0 / 4.0) * y / (beta * pow(pow(2.0, 1.0 / 2.0) - cos(x * M_PI / 12 / y), 1.0 / 2.0));
    }
  }
  o[0] = T1;
  o[1] = T2;
  o[2] = z;
}

void Orientation::map3(const double* p, double* o) {  // S/SampleOrientation.C:99-113
  double x = p[0], y = p[1], z = p[2];
  if (fabs(z) < 1e-10) {
    o[0] = o[1] = o[2] = 0;
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

void Orientation::generate(int n, unsigned int seed, std::vector<double>& out) const {  // :124-161
  std::uniform_real_distribution<double> xrand(0.0, 1.0);
  out.assign((size_t)n * 4, 0.0);
  Engine g1(seed);
  for (int jn = 0; jn < n; ++jn) {
    double xyz0[3];
    for (int q = 0; q < 3; ++q) xyz0[q] = xrand(g1) - .5;
    double xyz[3] = {aprime * xyz0[0], aprime * xyz0[1], aprime * xyz0[2]};
    // pyramid permutation (S/SampleOrientation.C:32-58): later tests override earlier
    int P[3] = {1, 2, 3};
    {
      double x = xyz[0], y = xyz[1], z = xyz[2];
      if ((fabs(x) <= z) && (fabs(y) <= z)) { P[0] = 1; P[1] = 2; P[2] = 3; }
      if ((z <= -fabs(x)) && (z <= -fabs(y))) { P[0] = -1; P[1] = 2; P[2] = -3; }
      if ((fabs(x) <= y) && (fabs(z) <= y)) { P[0] = 1; P[1] = -3; P[2] = 2; }
      if ((y <= -fabs(x)) && (y <= -fabs(z))) { P[0] = 1; P[1] = 3; P[2] = -2; }
      if ((fabs(z) < x) && (fabs(y) <= x)) { P[0] = -3; P[1] = 2; P[2] = 1; }
      if ((x <= -fabs(z)) && (x <= -fabs(y))) { P[0] = 3; P[1] = 2; P[2] = -1; }
    }
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

void rot_from_axis_angle(double omega, const double* ax, double* R) {  // S/VoxelsCA.h:94-107
  R[0] = cos(omega) + pow(ax[0], 2.0) * (1 - cos(omega));
  R[1] = ax[0] * ax[1] * (1 - cos(omega)) - ax[2] * sin(omega);
  R[2] = ax[0] * ax[2] * (1 - cos(omega)) + ax[1] * sin(omega);
  R[3] = ax[0] * ax[1] * (1 - cos(omega)) + ax[2] * sin(omega);
  R[4] = cos(omega) + pow(ax[1], 2.0) * (1 - cos(omega));
  R[5] = ax[1] * ax[2] * (1 - cos(omega)) - ax[0] * sin(omega);
  R[6] = ax[2] * ax[0] * (1 - cos(omega)) - ax[1] * sin(omega);
  R[7] = ax[2] * ax[1] * (1 - cos(omega)) + ax[0] * sin(omega);
  R[8] = cos(omega) + pow(ax[2], 2.0) * (1 - cos(omega));
}

void rot_table(const std::vector<double>& cTheta, int first, int count, std::vector<double>& rot9) {
  rot9.resize((size_t)9 * count);
  for (int g = 0; g < count; ++g) {
    const double* c = &cTheta[4 * (size_t)(first + g)];
    double ax[3] = {c[1], c[2], c[3]};
    rot_from_axis_angle(c[0], ax, &rot9[9 * (size_t)g]);
  }
}

}  // namespace amcafe
