// capi.cpp -- the public C ABI (include/amcafe.h) over the device library
// (amcafe_dev.h) and the host routines (host_ref.h).  Compiled with g++.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "../../include/amcafe.h"
#include "amcafe_dev.h"
#include "h5out.h"
#include "host_ref.h"

using namespace amcafe;

static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
static int dev_fail() { return fail(AMCAFE_ERR_DEVICE, std::string("device: ") + amcd_last_error()); }

extern "C" const char* amcafe_last_error(void) { return g_err.c_str(); }
extern "C" int amcafe_device_name(char* buf, int len) {
  if (!buf || len <= 0) return fail(AMCAFE_ERR_ARG, "null buffer");
  return amcd_device_name(buf, len) ? dev_fail() : AMCAFE_OK;
}

static const int kSeed0 = 2132512;          // VoxelsCA::seed0 (S/VoxelsCA.C:47)
static const unsigned int kSeed1 = 2912351; // VoxelsCA::seed1 (S/VoxelsCA.C:48)
static const int kPoolChunk = 8;            // nucleation orientations prepared per step

// per-domain parameters the kernels read
static amcd_params make_params(const int* nX, double dx, double tL, double tS, double T0, const double* a1,
                               double bmV, double DelT, double rNmax, double Avel, double nvel) {
  amcd_params p;
  memset(&p, 0, sizeof(p));
  p.nx = nX[0];
  p.ny = nX[1];
  p.nz = nX[2];
  p.dx = p.dy = p.dz = dx;
  p.hx = nX[0] * dx / 2.;
  p.hy = nX[1] * dx / 2.;
  p.tL = tL;
  p.tS = tS;
  p.T0 = T0;
  p.xi = tL * (1 + std::numeric_limits<double>::epsilon());  // S/TempField.C:146
  for (int q = 0; q < 6; ++q) p.a1[q] = a1[q];
  p.bmDX0 = DelT * bmV;                     // S/TempField.C:31
  p.DelT = DelT;
  p.rX = rNmax * pow(dx * 1e6, 3.);         // S/VoxelsCA.C:344
  // S/VoxelsCA.C:405 at T = tS and T = T0, pow rounded to nearest (the device evaluates
  // the same routine at any other temperature)
  p.vhat_tS = Avel * amcd_host_pow(tL - tS, nvel);
  p.vhat_T0 = Avel * amcd_host_pow(tL - T0, nvel);
  p.Avel = Avel;
  p.nvel = nvel;
  p.seed1 = kSeed1;
  p.bmV = bmV;
  return p;
}

// Nucleation orientations of one step: the m-th nucleation samples one orientation
// from seed sdloc0 + 10 m, sdloc0 = seed0 + 32 tInd + 64 nGrain (S/VoxelsCA.C:338, 485, 503).
struct Pool {
  std::vector<double> ct, rot;  // 4 and 9 doubles per entry
  void build(const Orientation& sa, int tInd, int nGrain, int n) {
    unsigned int sd0 = kSeed0 + 32 * tInd + 64 * nGrain;
    int have = (int)(ct.size() / 4);
    ct.resize((size_t)4 * n);
    rot.resize((size_t)9 * n);
    std::vector<double> one;
    for (int m = have; m < n; ++m) {
      sa.generate(1, sd0 + 10u * (unsigned int)m, one);
      for (int q = 0; q < 4; ++q) ct[4 * (size_t)m + q] = one[q];
      double ax[3] = {one[1], one[2], one[3]};
      rot_from_axis_angle(one[0], ax, &rot[9 * (size_t)m]);
    }
  }
  void clear() {
    ct.clear();
    rot.clear();
  }
};

// Run UpdateVoxels (optionally with the fused temperature) for a batch, refilling
// the nucleation pools of instances that used them up.
// Pool size for an instance whose previous step nucleated `last` grains: twice that plus a
// margin, so that a refill (another launch of the capture loop) is rare.
// AMCAFE_POOL_CHUNK overrides the margin (tests set 1 so that steps refill their pools)
static int pool_size(int last) {
  const char* e = getenv("AMCAFE_POOL_CHUNK");
  const int chunk = e && atoi(e) > 0 ? atoi(e) : kPoolChunk;
  return chunk + 2 * (last > 0 ? last : 0);
}

static int run_update(amcd_batch* dev, std::vector<amcd_step>& st, std::vector<amcd_out>& out,
                      std::vector<Pool>& pools, const Orientation& sa, bool fused, const std::vector<int>& last_nucs) {
  const int n = (int)st.size();
  std::vector<const double*> prot(n, nullptr);
  std::vector<int> pn(n, 0);
  for (int i = 0; i < n; ++i) {
    if (st[i].skip) continue;
    const int np = pool_size(last_nucs[i]);
    pools[i].clear();
    pools[i].build(sa, st[i].tInd, st[i].nGrain, np);
    st[i].n_pool = np;
    if (amcd_reserve_grains(dev, i, st[i].nGrain + 4 * np + 64)) return dev_fail();
    prot[i] = pools[i].rot.data();
    pn[i] = np;
  }
  if (amcd_set_pools(dev, prot.data(), pn.data())) return dev_fail();
  int rc = fused ? amcd_step_fused(dev, st.data(), out.data()) : amcd_update_voxels(dev, st.data(), out.data());
  if (rc) return dev_fail();
  for (int round = 0; round < 1 << 20; ++round) {
    bool pend = false;
    std::fill(prot.begin(), prot.end(), nullptr);
    std::fill(pn.begin(), pn.end(), 0);
    for (int i = 0; i < n; ++i) {
      if (out[i].status < 0 && !st[i].skip) {
        if (out[i].overflow) return fail(AMCAFE_ERR_DEVICE, "active list capacity could not be grown");
        if (out[i].status == -3)
          return fail(AMCAFE_ERR_STATE, "the interface has more than 2^19 active cells (capture-loop limit)");
        return fail(AMCAFE_ERR_STATE, "device step failed");
      }
      if (out[i].status != 1) continue;
      pend = true;
      int nn = 2 * st[i].n_pool;
      pools[i].build(sa, st[i].tInd, st[i].nGrain, nn);
      st[i].n_pool = nn;
      if (amcd_reserve_grains(dev, i, st[i].nGrain + nn + 64)) return dev_fail();
      prot[i] = pools[i].rot.data();
      pn[i] = nn;
    }
    if (!pend) break;
    if (amcd_set_pools(dev, prot.data(), pn.data())) return dev_fail();
    if (amcd_resume(dev, st.data(), out.data())) return dev_fail();
  }
  for (int i = 0; i < n; ++i)
    if (!st[i].skip && out[i].err_state)
      return fail(AMCAFE_ERR_STATE, "a mushy cell with liquid neighbours carries no valid grain id; the "
                                    "domain state is corrupt and must be uploaded again");
  return AMCAFE_OK;
}

static void to_stats(const amcd_out& o, amcafe_stats* s) {
  s->n_active = o.n_active;
  s->n_mushy0 = o.n_mushy0;
  s->n_events = o.n_events;
  s->n_captures = o.n_captures;
  s->n_nucleations = o.n_nucleations;
  s->n_growing = o.n_slots;
  s->nGrain = o.nGrain;
  s->draws = o.draws;
  s->tinc = o.tinc;
  s->slot_visits = o.slot_visits;
  s->cand_evals = o.cand_evals;
  s->interface_events = o.interface_events;
  s->loop_cycles = o.loop_cycles;
}

// ====================================================================================
// 1. Seam level
struct amcafe_domain {
  amcafe_domain_desc desc;
  amcd_batch* dev = nullptr;
  std::vector<double> cTheta;
  int nGrain = 0;
  Orientation sa;
  bool have_T = false;
  int last_nucs = 0;  // nucleations of the previous step (sizes the orientation pool)
};

extern "C" int amcafe_domain_create(const amcafe_domain_desc* d, amcafe_domain** out) {
  if (!d || !out) return fail(AMCAFE_ERR_ARG, "null argument");
  *out = nullptr;
  if (d->nX[0] <= 0 || d->nX[1] <= 0 || d->nX[2] <= 0 || !(d->dX > 0))
    return fail(AMCAFE_ERR_ARG, "nX and dX must be positive");
  amcafe_domain* h = new (std::nothrow) amcafe_domain();
  if (!h) return fail(AMCAFE_ERR_ARG, "out of host memory");
  h->desc = *d;
  double a1[6] = {d->meltparam[0], d->meltparam[2], d->meltparam[3], d->meltparam[1], d->meltparam[2],
                  d->meltparam[3]};
  amcd_params p = make_params(d->nX, d->dX, d->tL, d->tS, d->T0, a1, d->bmV, d->bmDelT, d->rNmax, d->Avel, d->nvel);
  char err[256];
  if (amcd_create(&p, 1, &h->dev, err, sizeof err)) {
    delete h;
    return fail(AMCAFE_ERR_DEVICE, err);
  }
  *out = h;
  return AMCAFE_OK;
}

extern "C" void amcafe_domain_destroy(amcafe_domain* d) {
  if (!d) return;
  amcd_destroy(d->dev);
  delete d;
}

extern "C" int amcafe_domain_upload(amcafe_domain* d, const int* gID, const int* vState, const double* extents,
                                    const double* centroidOct, const int* BuildjID) {
  if (!d) return fail(AMCAFE_ERR_ARG, "null domain");
  return amcd_upload(d->dev, 0, gID, vState, extents, centroidOct, BuildjID) ? dev_fail() : AMCAFE_OK;
}

extern "C" int amcafe_domain_download(amcafe_domain* d, int* gID, int* vState, double* extents, double* centroidOct,
                                      int* BuildjID, double* TempCurr) {
  if (!d) return fail(AMCAFE_ERR_ARG, "null domain");
  return amcd_download(d->dev, 0, gID, vState, extents, centroidOct, BuildjID, TempCurr) ? dev_fail() : AMCAFE_OK;
}

extern "C" int amcafe_domain_set_grains(amcafe_domain* d, const double* cTheta, int nGrain) {
  if (!d || (nGrain > 0 && !cTheta) || nGrain < 0) return fail(AMCAFE_ERR_ARG, "bad grain table");
  d->cTheta.assign(cTheta, cTheta + 4 * (size_t)nGrain);
  d->nGrain = nGrain;
  std::vector<double> rot;
  rot_table(d->cTheta, 0, nGrain, rot);
  return amcd_set_rot(d->dev, 0, rot.data(), nGrain) ? dev_fail() : AMCAFE_OK;
}

extern "C" int amcafe_domain_get_grains(amcafe_domain* d, double* cTheta, int* nGrain) {
  if (!d) return fail(AMCAFE_ERR_ARG, "null domain");
  if (nGrain) *nGrain = d->nGrain;
  if (cTheta) memcpy(cTheta, d->cTheta.data(), sizeof(double) * d->cTheta.size());
  return AMCAFE_OK;
}

static amcd_step laser_step(const amcafe_laser* l) {
  amcd_step s;
  memset(&s, 0, sizeof(s));
  s.lx = l->x;
  s.ly = l->y;
  s.lz = l->z;
  s.dirp = l->dirp;
  s.cg = cos(l->gth);
  s.sg = sin(l->gth);
  s.ilaserLoc = l->ilaserLoc;
  return s;
}

extern "C" int amcafe_analytic_temp_curr(amcafe_domain* d, const amcafe_laser* laser, int* any_melt) {
  if (!d || !laser) return fail(AMCAFE_ERR_ARG, "null argument");
  amcd_step s = laser_step(laser);
  if (amcd_temperature(d->dev, &s)) return dev_fail();
  d->have_T = true;
  if (any_melt) {
    // the any-melt flag is read back with the state (synchronises the stream)
    std::vector<amcd_out> o(1);
    if (amcd_read_outs(d->dev, o.data())) return dev_fail();
    *any_melt = o[0].any_melt;
  }
  return AMCAFE_OK;
}

extern "C" int amcafe_easm_temp_curr(amcafe_domain* d, const amcafe_laser* laser, double bmP, double layerT,
                                     int* all_melt, long long* n_integrals) {
  if (!d || !laser) return fail(AMCAFE_ERR_ARG, "null argument");
  amcd_step s = laser_step(laser);
  std::vector<amcd_out> o(1);
  unsigned long long ne = 0;
  if (amcd_easm(d->dev, &s, bmP, layerT, o.data(), &ne)) return dev_fail();
  d->have_T = true;
  if (all_melt) *all_melt = o[0].any_melt ? 0 : 1;
  if (n_integrals) *n_integrals = (long long)ne;
  return AMCAFE_OK;
}

extern "C" int amcafe_update_voxels(amcafe_domain* d, int tInd, int ilaserLoc, amcafe_stats* stats) {
  if (!d) return fail(AMCAFE_ERR_ARG, "null domain");
  if (!d->have_T) return fail(AMCAFE_ERR_ARG, "amcafe_analytic_temp_curr must run before amcafe_update_voxels");
  std::vector<amcd_step> st(1);
  memset(st.data(), 0, sizeof(amcd_step));
  st[0].ilaserLoc = ilaserLoc;
  st[0].tInd = tInd;
  st[0].nGrain = d->nGrain;
  std::vector<amcd_out> out(1);
  std::vector<Pool> pools(1);
  int rc = run_update(d->dev, st, out, pools, d->sa, false, std::vector<int>(1, d->last_nucs));
  if (rc) return rc;
  d->have_T = false;
  int used = out[0].pool_used;
  d->cTheta.insert(d->cTheta.end(), pools[0].ct.begin(), pools[0].ct.begin() + 4 * (size_t)used);
  d->nGrain = out[0].nGrain;
  d->last_nucs = out[0].n_nucleations;
  if (stats) to_stats(out[0], stats);
  return AMCAFE_OK;
}

// ====================================================================================
// 2. Driver level: S/main1.C:22-98 for a batch of decks
struct Run {
  Deck g;
  std::vector<double> cTheta;
  int nGrain = 0;
  Engine genlayer, bp_generator;  // VoxelsCA::genlayer; BasePlate's member engine (default seed)
  int Nzh = 0, nlayerTot = 0, TtInd = 0;
  double bp_height = 0;
  bool done = false;
  std::vector<int> out_steps;
  std::string out_prefix;  // output files <prefix><tInd>.h5 when set (S/main1.C:53)
  amcd_out last = {};
};

struct amcafe_sim {
  std::vector<Run> runs;
  amcd_batch* dev = nullptr;
  Orientation sa;
};

static int upload_rot(amcafe_sim* s, int i) {
  Run& r = s->runs[i];
  std::vector<double> rot;
  rot_table(r.cTheta, 0, r.nGrain, rot);
  return amcd_set_rot(s->dev, i, rot.data(), r.nGrain) ? dev_fail() : AMCAFE_OK;
}

// BasePlate::GenerateVoxel + VoxelsCA::InitializeVoxels (S/BasePlate.C:18-97,
// S/VoxelsCA.C:125-161)
static int init_plate(amcafe_sim* s, int i) {
  Run& r = s->runs[i];
  const Deck& g = r.g;
  r.bp_height = ceil(g.bpH / g.dX[2]) * g.dX[2];
  r.Nzh = std::min(g.nX[2], int(ceil(g.bpH * .9999 / g.dX[2])));
  r.bp_height = r.Nzh * (g.dX[2]);
  double rate = g.mu * (g.dX[0] * 1e6) * (g.dX[1] * 1e6) * g.nX[0] * g.nX[1] * (r.bp_height * 1e6);
  std::poisson_distribution<int> ng(rate);
  int Ngrain = 0;
  while (Ngrain == 0) Ngrain = ng(r.bp_generator);
  std::uniform_real_distribution<double> xrand(0.0, 1.0);
  double Lx = g.dX[0] * g.nX[0], Ly = g.dX[1] * g.nX[1];
  std::vector<double> Xs(3 * (size_t)Ngrain);
  for (int j = 0; j < Ngrain; ++j) {
    Xs[3 * j] = xrand(r.bp_generator) * Lx;
    Xs[3 * j + 1] = xrand(r.bp_generator) * Ly;
    Xs[3 * j + 2] = xrand(r.bp_generator) * r.bp_height;
  }
  if (amcd_init_plate(s->dev, i, r.Nzh, Xs.data(), Ngrain)) return dev_fail();
  Engine g1(531231);
  unsigned int sdloc = unsigned(double(g1()) / double(g1.max()) * pow(2.0, 32.0));
  s->sa.generate(Ngrain, sdloc, r.cTheta);
  r.nGrain = Ngrain;
  // InitializeVoxels marks every voxel with a grain as built: the whole plate (S/VoxelsCA.C:157-160)
  if (amcd_build_init(s->dev, i, (long long)r.Nzh * g.nX[0] * g.nX[1])) return dev_fail();
  return upload_rot(s, i);
}

// VoxelsCA::AddLayer1 (S/VoxelsCA.C:887-982)
static int add_layer(amcafe_sim* s, int i) {
  Run& r = s->runs[i];
  const Deck& g = r.g;
  unsigned int sdloc = unsigned(double(r.genlayer()) / double(r.genlayer.max()) * pow(2.0, 32.0));
  double rate = g.lrate * (g.dX[0] * 1e6) * (g.dX[1] * 1e6) * g.nX[0] * g.nX[1] * (g.layerT * 1e6);
  std::poisson_distribution<int> ngd(rate);
  int ng2 = 0;
  while (ng2 == 0) ng2 = ngd(r.genlayer);
  int iz1 = g.ilaserLoc - g.nZlayer;
  double zloc0 = iz1 * g.dX[2];
  std::vector<double> Xs(3 * (size_t)ng2);
  std::uniform_real_distribution<double> xrand(0.0, 1.0);
  double Lx = g.nX[0] * g.dX[0], Ly = g.nX[1] * g.dX[1];
  for (int j = 0; j < ng2; ++j) {
    Xs[3 * j] = xrand(r.genlayer) * Lx;
    Xs[3 * j + 1] = xrand(r.genlayer) * Ly;
    Xs[3 * j + 2] = xrand(r.genlayer) * g.layerT + zloc0;
  }
  int i2 = 0;  // sites that label a voxel of the layer
  if (amcd_add_layer(s->dev, i, iz1, g.ilaserLoc, r.nGrain, Xs.data(), ng2, &i2)) return dev_fail();
  r.nGrain += i2 + 1;  // one more than the labels used (S/VoxelsCA.C:967-972)
  std::vector<double> aa;
  s->sa.generate(i2 + 1, sdloc, aa);
  r.cTheta.insert(r.cTheta.end(), aa.begin(), aa.end());
  return upload_rot(s, i);
}

// VoxelsCA::CleanLayer (S/VoxelsCA.C:800-839): drop grains with no voxel, renumber
static int clean_layer(amcafe_sim* s, int i) {
  Run& r = s->runs[i];
  std::vector<int> present(r.nGrain + 1, 0), remap(r.nGrain + 1, 0);
  if (r.nGrain > 0 && amcd_grain_presence(s->dev, i, r.nGrain, present.data())) return dev_fail();
  std::vector<double> ct;
  int nG = 0;
  for (int j = 0; j < r.nGrain; ++j)
    if (present[j]) {
      for (int q = 0; q < 4; ++q) ct.push_back(r.cTheta[4 * (size_t)j + q]);
      remap[j] = ++nG;
    }
  if (r.nGrain > 0 && amcd_remap_grains(s->dev, i, remap.data(), r.nGrain)) return dev_fail();
  r.cTheta.swap(ct);
  r.nGrain = nG;
  return upload_rot(s, i);
}

// VoxelsCA::WriteToHDF1 of instance i's current state
static int write_run(amcafe_sim* s, int i, const char* path) {
  Run& r = s->runs[i];
  const size_t N = (size_t)r.g.nvox();
  std::vector<int> gid(N), vs(N), build(N);
  std::vector<double> T(N);
  if (amcd_download(s->dev, i, gid.data(), vs.data(), nullptr, nullptr, nullptr, T.data()) ||
      amcd_get_build(s->dev, i, build.data()))
    return dev_fail();
  std::string err;
  const int nX[3] = {r.g.nX[0], r.g.nX[1], r.g.nX[2]};
  const double dX[3] = {r.g.dX[0], r.g.dX[1], r.g.dX[2]};
  if (write_hdf5(path, nX, dX, gid.data(), vs.data(), build.data(), T.data(), r.cTheta.data(), r.nGrain, err))
    return fail(AMCAFE_ERR_IO, err);
  return AMCAFE_OK;
}
static int write_scheduled(amcafe_sim* s, int i) {
  const Run& r = s->runs[i];
  if (r.out_prefix.empty()) return AMCAFE_OK;
  return write_run(s, i, (r.out_prefix + std::to_string(r.TtInd) + ".h5").c_str());
}

// VoxelsCA::SetBuild (S/VoxelsCA.C:76-115) as set membership over grain ids, on the device:
// voxels of every grain that has a melted voxel, then of every grain with a Build voxel.
static int set_build(amcafe_sim* s, int i) {
  return amcd_set_build(s->dev, i, s->runs[i].nGrain) ? dev_fail() : AMCAFE_OK;
}

extern "C" int amcafe_sim_create(const char* const* decks, int n, amcafe_sim** out) {
  if (!decks || n <= 0 || !out) return fail(AMCAFE_ERR_ARG, "need at least one deck");
  *out = nullptr;
  amcafe_sim* s = new (std::nothrow) amcafe_sim();
  if (!s) return fail(AMCAFE_ERR_ARG, "out of host memory");
  s->runs.resize(n);
  std::vector<amcd_params> ps(n);
  for (int i = 0; i < n; ++i) {
    Run& r = s->runs[i];
    std::string err;
    if (!decks[i] || r.g.parse(decks[i], err)) {
      delete s;
      return fail(AMCAFE_ERR_DECK, "deck " + std::to_string(i) + ": " + err);
    }
    r.g.init();
    if (r.g.NpT < 2 || r.g.nZlayer <= 0) {
      delete s;
      return fail(AMCAFE_ERR_DECK, "deck " + std::to_string(i) + ": need at least two scan points and lT >= dx");
    }
    const Deck& g = r.g;
    ps[i] = make_params(g.nX.data(), g.dX[0], g.tL, g.tS, g.T0, g.a1, g.bmV, g.bmDelT, g.rNmax, g.Avel, g.nvel);
    r.genlayer.seed(kSeed1);
  }
  char err[256];
  if (amcd_create(ps.data(), n, &s->dev, err, sizeof err)) {
    delete s;
    return fail(AMCAFE_ERR_DEVICE, err);
  }
  for (int i = 0; i < n; ++i) {
    Run& r = s->runs[i];
    int rc = init_plate(s, i);
    if (!rc) {
      r.nlayerTot = int(ceil((double)(r.g.nX[2] - r.Nzh) / (double)r.g.nZlayer));  // S/main1.C:59
      rc = add_layer(s, i);                                                        // S/main1.C:60
    }
    if (rc) {
      amcafe_sim_destroy(s);
      return rc;
    }
    r.g.update_laser();  // S/main1.C:61
  }
  *out = s;
  return AMCAFE_OK;
}

extern "C" void amcafe_sim_destroy(amcafe_sim* s) {
  if (!s) return;
  amcd_destroy(s->dev);
  delete s;
}

extern "C" int amcafe_sim_step(amcafe_sim* s, int* n_running) {
  if (!s) return fail(AMCAFE_ERR_ARG, "null sim");
  const int n = (int)s->runs.size();
  std::vector<amcd_step> st(n);
  std::vector<amcd_out> out(n);
  std::vector<Pool> pools(n);
  int running = 0;
  for (int i = 0; i < n; ++i) {
    Run& r = s->runs[i];
    amcd_step& p = st[i];
    memset(&p, 0, sizeof(p));
    if (r.done) {
      p.skip = 1;
      continue;
    }
    running++;
    r.TtInd = int(round(r.g.time / r.g.bmDelT));  // S/main1.C:64
    r.g.laser_inputs(&p.lx, &p.ly, &p.lz, &p.dirp, &p.cg, &p.sg);
    p.ilaserLoc = r.g.ilaserLoc;
    p.tInd = r.g.tInd;
    p.nGrain = r.nGrain;
  }
  if (running == 0) {
    if (n_running) *n_running = 0;
    return AMCAFE_OK;
  }
  std::vector<int> last(n, 0);
  for (int i = 0; i < n; ++i) last[i] = s->runs[i].last.n_nucleations;
  int rc = run_update(s->dev, st, out, pools, s->sa, true, last);
  if (rc) return rc;
  running = 0;
  for (int i = 0; i < n; ++i) {
    Run& r = s->runs[i];
    if (r.done) continue;
    r.last = out[i];
    int used = out[i].pool_used;
    r.cTheta.insert(r.cTheta.end(), pools[i].ct.begin(), pools[i].ct.begin() + 4 * (size_t)used);
    r.nGrain = out[i].nGrain;
    Deck& g = r.g;
    g.end_of_scan(out[i].any_melt);  // S/TempField.C:199-216
    // S/main1.C:74-88
    g.update_laser();
    r.done = g.indlayer > r.nlayerTot;
    if (g.inewlayerflg == 1 && (rc = clean_layer(s, i))) return rc;
    int indOut = r.TtInd % g.outint;
    if (indOut == 0 || r.done || (g.inewlayerflg == 1 && g.outNL == 0)) {
      r.out_steps.push_back(r.TtInd);
      if ((rc = set_build(s, i))) return rc;
      if ((rc = write_scheduled(s, i))) return rc;  // S/main1.C:83-84
    }
    g.update_time2(g.bmDelT);
    if (g.inewlayerflg == 1 && (rc = add_layer(s, i))) return rc;
    if (r.done && (rc = clean_layer(s, i))) return rc;  // S/main1.C:94
    if (r.done && (rc = write_scheduled(s, i))) return rc;  // S/main1.C:95-97
    if (!r.done) running++;
  }
  if (n_running) *n_running = running;
  return AMCAFE_OK;
}

extern "C" int amcafe_sim_info(amcafe_sim* s, int i, amcafe_info* info) {
  if (!s || i < 0 || i >= (int)s->runs.size() || !info) return fail(AMCAFE_ERR_ARG, "bad instance");
  const Run& r = s->runs[i];
  for (int q = 0; q < 3; ++q) info->nX[q] = r.g.nX[q];
  info->nGrain = r.nGrain;
  info->tInd = r.g.tInd;
  info->isp = r.g.isp;
  info->indlayer = r.g.indlayer;
  info->ilaserLoc = r.g.ilaserLoc;
  info->nlayerTot = r.nlayerTot;
  info->done = r.done ? 1 : 0;
  info->inewscanflg = r.g.inewscanflg;
  info->inewlayerflg = r.g.inewlayerflg;
  info->time = r.g.time;
  info->gth = r.g.gth;
  return AMCAFE_OK;
}

extern "C" int amcafe_sim_tuning(amcafe_sim* s, amcafe_tuning* t) {
  if (!s || !t) return fail(AMCAFE_ERR_ARG, "null sim or output");
  static const int caps[5] = {256, 512, 1024, 2048, 4096};
  for (int c = 0; c < 5; ++c) t->candidates[c] = caps[c];
  amcd_tuning(s->dev, &t->list_blocks, &t->tunings, &t->growing_cells, t->ms, 5);
  return AMCAFE_OK;
}

extern "C" int amcafe_sim_stats(amcafe_sim* s, int i, amcafe_stats* st) {
  if (!s || i < 0 || i >= (int)s->runs.size() || !st) return fail(AMCAFE_ERR_ARG, "bad instance");
  to_stats(s->runs[i].last, st);
  return AMCAFE_OK;
}

extern "C" int amcafe_sim_get_fields(amcafe_sim* s, int i, int* gID, int* vState, double* T, double* ext,
                                     double* cen, int* bj, int* Build) {
  if (!s || i < 0 || i >= (int)s->runs.size()) return fail(AMCAFE_ERR_ARG, "bad instance");
  if (amcd_download(s->dev, i, gID, vState, ext, cen, bj, T)) return dev_fail();
  if (Build && amcd_get_build(s->dev, i, Build)) return dev_fail();
  return AMCAFE_OK;
}

extern "C" int amcafe_sim_get_grains(amcafe_sim* s, int i, double* cTheta, int* nGrain) {
  if (!s || i < 0 || i >= (int)s->runs.size()) return fail(AMCAFE_ERR_ARG, "bad instance");
  const Run& r = s->runs[i];
  if (nGrain) *nGrain = r.nGrain;
  if (cTheta) memcpy(cTheta, r.cTheta.data(), sizeof(double) * r.cTheta.size());
  return AMCAFE_OK;
}

extern "C" int amcafe_sim_set_timing(amcafe_sim* s, int on) {
  if (!s) return fail(AMCAFE_ERR_ARG, "null sim");
  amcd_set_timing(s->dev, on);
  return AMCAFE_OK;
}

extern "C" int amcafe_sim_timings(amcafe_sim* s, double* ms, int n) {
  if (!s || !ms) return fail(AMCAFE_ERR_ARG, "null argument");
  return amcd_last_timings(s->dev, ms, n) ? dev_fail() : AMCAFE_OK;
}

extern "C" int amcafe_sim_write_hdf5(amcafe_sim* s, int i, const char* path) {
  if (!s || i < 0 || i >= (int)s->runs.size() || !path) return fail(AMCAFE_ERR_ARG, "bad instance or path");
  return write_run(s, i, path);
}

extern "C" int amcafe_sim_set_output(amcafe_sim* s, int i, const char* prefix) {
  if (!s || i < 0 || i >= (int)s->runs.size()) return fail(AMCAFE_ERR_ARG, "bad instance");
  s->runs[i].out_prefix = prefix ? prefix : "";
  return AMCAFE_OK;
}

extern "C" int amcafe_write_hdf5_fields(const char* path, const int* nX, const double* dX, const int* gID,
                                        const int* vState, const int* Build, const double* TempCurr,
                                        const double* cTheta, int nGrain) {
  if (!path || !nX || !dX || !gID || !vState || !Build || !TempCurr || (nGrain > 0 && !cTheta) || nGrain < 0 ||
      nX[0] <= 0 || nX[1] <= 0 || nX[2] <= 0)
    return fail(AMCAFE_ERR_ARG, "bad output arguments");
  std::string err;
  if (amcafe::write_hdf5(path, nX, dX, gID, vState, Build, TempCurr, cTheta, nGrain, err))
    return fail(AMCAFE_ERR_IO, err);
  return AMCAFE_OK;
}

extern "C" int amcafe_ipf_colours(const double* cTheta, int nGrain, float* rgb9) {
  if (nGrain < 0 || (nGrain > 0 && (!cTheta || !rgb9))) return fail(AMCAFE_ERR_ARG, "bad grain table");
  amcafe::ipf_colours(cTheta, nGrain, rgb9);
  return AMCAFE_OK;
}

// ====================================================================================
// 3. Differentiable temperature field
static bool heat_args_ok(const amcafe_heat_desc* d, int nb, const double* theta, const double* target, int step0,
                         int nsteps, const double* loss) {
  return d && nb > 0 && nb <= 65535 && theta && target && loss && nsteps >= 0 && step0 >= 0 && d->nX[0] > 0 &&
         d->nX[1] > 0 && d->nX[2] > 0 && (long long)d->nX[0] * d->nX[1] * d->nX[2] <= 0x7fffffffLL && d->dX > 0 &&
         d->dt > 0 && d->t_hist > 0 && d->width > 0 && d->k > 0 && d->alpha > 0;
}
static int heat(const amcafe_heat_desc* d, int nb, const double* theta, const double* target, int step0, int nsteps,
                int mode, double* loss, double* grad, float* ms) {
  if (!heat_args_ok(d, nb, theta, target, step0, nsteps, loss) || (mode == 1 && !grad))
    return fail(AMCAFE_ERR_ARG, "bad heat-field arguments");
  char err[256] = {0};
  if (amcd_heat(d, nb, theta, target, step0, nsteps, mode, loss, grad, ms, err, sizeof err))
    return fail(AMCAFE_ERR_DEVICE, err);
  return AMCAFE_OK;
}
extern "C" int amcafe_heat_loss_f32(const amcafe_heat_desc* d, const double* theta, const double* target, int step0,
                                    int nsteps, double* loss, float* ms) {
  return heat(d, 1, theta, target, step0, nsteps, 0, loss, nullptr, ms);
}
extern "C" int amcafe_heat_loss_grad_f32(const amcafe_heat_desc* d, const double* theta, const double* target,
                                         int step0, int nsteps, double* loss, double* grad, float* ms) {
  return heat(d, 1, theta, target, step0, nsteps, 1, loss, grad, ms);
}
extern "C" int amcafe_heat_loss_f64(const amcafe_heat_desc* d, const double* theta, const double* target, int step0,
                                    int nsteps, double* loss, float* ms) {
  return heat(d, 1, theta, target, step0, nsteps, 2, loss, nullptr, ms);
}
extern "C" int amcafe_heat_loss_grad_batch_f32_noalloc(const amcafe_heat_desc* d, int n, const double* theta,
                                                       const double* target, int step0, int nsteps, double* loss,
                                                       double* grad) {
  if (!heat_args_ok(d, n, theta, target, step0, nsteps, loss) || !grad)
    return fail(AMCAFE_ERR_ARG, "bad heat-field arguments");
  char err[256] = {0};
  const int rc = amcd_heat_noalloc(d, n, theta, target, step0, nsteps, loss, grad, err, sizeof err);
  if (rc == -2) return fail(AMCAFE_ERR_ARG, err);
  if (rc) return fail(AMCAFE_ERR_DEVICE, err);
  return AMCAFE_OK;
}
extern "C" long long amcafe_heat_noalloc_bytes(const amcafe_heat_desc* d, int n) {
  if (!d || n <= 0 || d->nX[0] <= 0 || d->nX[1] <= 0 || d->nX[2] <= 0) return -1;
  return (long long)amcd_heat_noalloc_bytes(d, n);
}
extern "C" long long amcafe_heat_noalloc_max_bytes(void) { return (long long)amcd_heat_noalloc_max(); }
extern "C" long long amcafe_heat_last_evaluations(void) { return (long long)amcd_heat_last_evals(); }
extern "C" int amcafe_heat_loss_grad_batch_f32(const amcafe_heat_desc* d, int n, const double* theta,
                                               const double* target, int step0, int nsteps, double* loss, double* grad,
                                               float* ms) {
  return heat(d, n, theta, target, step0, nsteps, 1, loss, grad, ms);
}
