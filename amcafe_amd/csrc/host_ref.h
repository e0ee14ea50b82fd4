// host_ref.h -- host-side parts of the CAFE path that stay on the CPU: the deck,
// the laser scan state machine, the orientation sampler, grain rotations and the
// per-layer Voronoi labelling.  Compiled with g++ and the reference's flags
// (-O3, no -march) so constant folding and libm calls round as the reference does.
#pragma once
#include <random>
#include <string>
#include <vector>

namespace amcafe {

typedef std::minstd_rand0 Engine;  // std::default_random_engine of libstdc++

// Grid (S/Grid.h:42-70, S/Grid.C:15-164) plus the scan bookkeeping of TempField
// (S/TempField.C:16-51) needed to place the laser.
struct Deck {
  // keywords and defaults (S/Grid.C:18-65)
  std::vector<int> nX{128, 128, 64};
  std::vector<double> dX{1e-6, 1e-6, 1e-6};
  int outint = 1, patternID = 1, outNL = 0, nTsd = 0;
  double tL = 1620, tS = 1531.5, T0 = 300, layerT = 25e-6, bpH = 0.0, mu = .01;
  std::vector<double> meltparam{75e-6, 162.75e-6, 75e-6, 75e-6}, offset{0.0, 0.0, 0.0}, gsize{0, 0};
  double bhatch = 0, bmDelT = 0, bmV = 500e-3, bmP = 0, rNmax = .002, lrate = 0, gth0 = 0., gth = 0.;
  double Avel = 0., nvel = 0.;
  bool has_bmdelt = false, has_lrate = false, has_bhatch = false;
  // derived
  std::vector<double> LX, lcoor, lcoor2;
  double gbox[4] = {0, 0, 0, 0};
  int nZlayer = 0, isp = 0, inewscanflg = 1, inewlayerflg = 1, Nzhg = 0, indlayer = 1, ilaserLoc = 0;
  int Ntd = 0, Nsd = 0, NpT = 0, tInd = 0;
  double time = 0.0;
  std::vector<int> ispvec;
  double a1[6] = {0, 0, 0, 0, 0, 0};

  // parse deck text (S/Grid.C:165-318); returns 0 or sets err
  int parse(const std::string& text, std::string& err);
  void init();          // S/Grid.C:68-108 + TempField ctor/InitializeAnalytic
  void update_laser();  // S/Grid.C:110-164
  void update_time2(double dt) {  // S/Grid.C:334-339
    time += dt;
    tInd += 1;
  }
  // laser position and direction for the temperature (S/TempField.C:146-171)
  void laser_inputs(double* x, double* y, double* z, double* dirp, double* cg, double* sg) const;
  bool laser_outside_box() const;       // S/TempField.C:199-201
  void end_of_scan(int any_melt);       // S/TempField.C:202-215
  long long nvox() const { return (long long)nX[0] * nX[1] * nX[2]; }

 private:
  void rotate_grid();
};

// SampleOrientation (S/SampleOrientation.C)
struct Orientation {
  double r, R1, a, aprime, beta;
  double coeffA[7];
  Orientation();
  void generate(int n, unsigned int seed, std::vector<double>& out) const;

 private:
  void map2(const double* p, double* o) const;
  static void map3(const double* p, double* o);
};

// loadRotMat (S/VoxelsCA.h:94-107): row-major 3x3, local -> global
void rot_from_axis_angle(double omega, const double* ax, double* R9);
void rot_table(const std::vector<double>& cTheta, int first, int count, std::vector<double>& rot9);

}  // namespace amcafe
