// oracle/ref_harness.cpp
//
// TEST INFRASTRUCTURE ONLY.  Driver for the reference's own sources that compile
// here unchanged: src/src-mpi/Grid.C (deck parser + laser state machine) and
// src/src-mpi/SampleOrientation.C (orientation sampler).  The Makefile in this
// directory compiles them from /root/reference into oracle/_ref/; nothing from the
// reference is copied into the repository.
//
// The rest of the reference's path (TempField.C, VoxelsCA.C, Utilities.C,
// Partition.C, main1.C) includes GSL, ADIOS2 or METIS headers that this image
// does not have, so it is not built.
//
// Also built: src/src-mpi/BasePlate.C, which compiles unchanged.  Its constructor takes a
// Partition and reads only the two members icellidLoc and ncellLoc; Partition's own
// constructor lives in Partition.C, which needs METIS and is not built.  The harness
// therefore sets up those two members of a Partition-typed object itself (the one-rank
// identity partition of S/Partition.C:281-288) and calls no Partition member function.
//
// Usage:
//   ref_harness samples <n> <seed>            -> n lines "omega nx ny nz" (%a)
//   ref_harness nucs <n> <seed> <stride>      -> n lines: GenerateSamples(1, seed + stride*k)
//   ref_harness plate <deck>                  -> BasePlate::GenerateVoxel at one rank: Ngrain,
//                                                then the label of every plate voxel, then
//                                                the orientations (omega nx ny nz, %a)
//   ref_harness grid <deck> <flags-file>      -> parsed deck, then one line per step
// The flags file holds, per step, the reference's end-of-scan input "is any voxel
// at or above the liquidus" (0/1); the harness applies the end-of-scan rule of
// S/TempField.C:199-216 before each Grid::UpdateLaser call, as main1.C does.

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>
#include <cmath>
#include "mpi.h"
#include <new>
#include <numeric>
#include "Grid.h"
#include "BasePlate.h"
#include "Partition.h"
#include "SampleOrientation.h"

static void print_grid(const Grid& g) {
  printf("nX %d %d %d\n", g.nX[0], g.nX[1], g.nX[2]);
  printf("dX %a %a %a\n", g.dX[0], g.dX[1], g.dX[2]);
  printf("LX %a %a %a\n", g.LX[0], g.LX[1], g.LX[2]);
  printf("bmV %a bmDelT %a bhatch %a layerT %a bpH %a\n", g.bmV, g.bmDelT, g.bhatch, g.layerT, g.bpH);
  printf("tL %a tS %a T0 %a mu %a lrate %a rNmax %a\n", g.tL, g.tS, g.T0, g.mu, g.lrate, g.rNmax);
  printf("meltparam %a %a %a %a\n", g.meltparam[0], g.meltparam[1], g.meltparam[2], g.meltparam[3]);
  printf("offset %a %a %a\n", g.offset[0], g.offset[1], g.offset[2]);
  printf("gsize %a %a gth0 %a gth %a Avel %a nvel %a\n", g.gsize[0], g.gsize[1], g.gth0, g.gth, g.Avel, g.nvel);
  printf("patternID %d outint %d outNL %d nZlayer %d Nzhg %d Ntd %d Nsd %d NpT %d ilaserLoc %d\n", g.patternID,
         g.outint, g.outNL, g.nZlayer, g.Nzhg, g.Ntd, g.Nsd, g.NpT, g.ilaserLoc);
  printf("gbox %a %a %a %a\n", g.gbox[0], g.gbox[1], g.gbox[2], g.gbox[3]);
  double s = 0;
  for (int k = 0; k < 2 * g.NpT; ++k) s += g.lcoor[k] * (k + 1) + g.lcoor2[k] * (k + 7);
  printf("lcoor_sum %a\n", s);
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::string mode = argv[1];
  if (mode == "samples" && argc == 4) {
    int n = atoi(argv[2]);
    unsigned int seed = (unsigned int)strtoul(argv[3], nullptr, 10);
    SampleOrientation sa;
    std::vector<double> out;
    sa.GenerateSamples(n, seed, out);
    for (int i = 0; i < n; ++i) printf("%a %a %a %a\n", out[4 * i], out[4 * i + 1], out[4 * i + 2], out[4 * i + 3]);
    return 0;
  }
  if (mode == "nucs" && argc == 5) {
    int n = atoi(argv[2]);
    unsigned int seed = (unsigned int)strtoul(argv[3], nullptr, 10);
    unsigned int stride = (unsigned int)strtoul(argv[4], nullptr, 10);
    SampleOrientation sa;
    std::vector<double> out;
    for (int k = 0; k < n; ++k) {
      sa.GenerateSamples(1, seed + stride * (unsigned int)k, out);
      printf("%a %a %a %a\n", out[0], out[1], out[2], out[3]);
    }
    return 0;
  }
  if (mode == "plate" && argc == 3) {
    MPI_Init(NULL, NULL);
    int myid = 0, np = 1;
    std::string deck = argv[2];
    Grid g(deck, myid, np);
    const int N = g.nX[0] * g.nX[1] * g.nX[2];
    // zeroed storage: every std::vector member of libstdc++ starts as an empty vector
    alignas(Partition) static unsigned char pbuf[sizeof(Partition)];
    Partition* part = reinterpret_cast<Partition*>(pbuf);
    new (&part->icellidLoc) std::vector<int>(N);
    std::iota(part->icellidLoc.begin(), part->icellidLoc.end(), 0);
    part->ncellLoc = N;
    part->nGhost = 0;
    BasePlate bp(g, *part);
    bp.GenerateVoxel();
    printf("Ngrain %d Nzh %d nlabels %zu\n", bp.Ngrain, bp.Nzh, bp.X.size());
    for (size_t j = 0; j < bp.X.size(); ++j) printf("%d\n", bp.X[j]);
    for (int j = 0; j < bp.Ngrain; ++j)
      printf("%a %a %a %a\n", bp.cTheta[4 * j], bp.cTheta[4 * j + 1], bp.cTheta[4 * j + 2], bp.cTheta[4 * j + 3]);
    MPI_Finalize();
    return 0;
  }
  if (mode == "grid" && argc == 4) {
    MPI_Init(NULL, NULL);
    int myid = 0, np = 1;
    std::string deck = argv[2];
    Grid g(deck, myid, np);
    print_grid(g);
    // the TempField scan pattern (S/TempField.C:33-50) is needed to map isp -> point
    std::vector<int> ispvec(g.NpT, 0);
    for (int i = 0; i < g.NpT; ++i) ispvec[i] = i;
    if (g.patternID == 2 || g.patternID == 4)
      for (int j = 0; j < g.Ntd; ++j)
        if (j % 2 == 1)
          for (int j1 = 0; j1 < g.Nsd; ++j1) ispvec[g.Nsd * j + j1] = g.Nsd * j + g.Nsd - 1 - j1;
    std::ifstream ff(argv[3]);
    std::vector<int> flags;
    int f;
    while (ff >> f) flags.push_back(f);
    g.UpdateLaser();  // S/main1.C:61
    for (size_t st = 0; st < flags.size(); ++st) {
      printf("step %zu isp %d ilaserLoc %d indlayer %d scan %d layer %d gth %a\n", st, g.isp, g.ilaserLoc,
             g.indlayer, g.inewscanflg, g.inewlayerflg, g.gth);
      double x = g.lcoor2[2 * ispvec[g.isp]] - g.offset[0];
      double y = g.lcoor2[2 * ispvec[g.isp] + 1] - g.offset[1];
      if (x < g.gbox[0] || x > g.gbox[1] || y < g.gbox[2] || y > g.gbox[3]) {
        if (flags[st] == 0 || fmod(g.isp + 1, g.Nsd) == 0) g.inewscanflg = 1;
        if (g.isp == (g.NpT - 1)) {
          g.inewscanflg = 1;
          g.inewlayerflg = 1;
          g.isp = 0;
          g.indlayer += 1;
        }
      }
      g.UpdateLaser();
      g.UpdateTime2(g.bmDelT);
    }
    MPI_Finalize();
    return 0;
  }
  return 2;
}
