// capi.cpp -- the extern "C" boundary declared in include/kofm.h.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/kofm.h"
#include "engine.hpp"
#include "geometry.hpp"
#include "init.hpp"
#include "io.hpp"

struct kofm_ctx {
  int nX = 0, nY = 0, storage = 0;
  bool komega = false;
  std::unique_ptr<kofm::Engine> engine;
  kofm::Geometry geo;
  kofm::Params params;
  double pipeRadius = 0.;
  bool haveGeo = false;
  std::string err;
};

namespace {
thread_local std::string g_createError;

#define NEED_ENGINE(ctx)                                                              \
  do {                                                                                \
    if (!(ctx)->engine) {                                                             \
      (ctx)->err = "host-only context (kofm_create_host) has no device state";        \
      return KOFM_ERR_STATE;                                                          \
    }                                                                                 \
  } while (0)

int setErr(kofm_ctx* c, int rc, const std::string& msg) {
  if (c) c->err = msg;
  return rc;
}
}  // namespace

namespace kofm {
void setGlobalError(const std::string& msg) { g_createError = msg; }
}  // namespace kofm

extern "C" {

int kofm_create(const kofm_grid* grid, kofm_ctx** out) {
  if (!grid || !out) {
    g_createError = "null argument";
    return KOFM_ERR_ARG;
  }
  *out = nullptr;
  if (grid->nX < 4 || grid->nY < 4) {
    g_createError = "lattice must be at least 4 x 4";
    return KOFM_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_createError = "no HIP device available";
    return KOFM_ERR_DEVICE;
  }
  auto* c = new kofm_ctx;
  c->nX = grid->nX;
  c->nY = grid->nY;
  c->storage = grid->storage;
  c->komega = grid->use_komega != 0;
  std::string err;
  int rc = 0;
  c->engine = kofm::makeEngine(c->nX, c->nY, c->storage, c->komega, rc, err);
  if (!c->engine) {
    g_createError = err.empty() ? "engine creation failed" : err;
    delete c;
    return rc < 0 ? rc : KOFM_ERR_DEVICE;
  }
  c->engine->setParams(c->params);
  *out = c;
  return KOFM_OK;
}

int kofm_create_host(const kofm_grid* grid, kofm_ctx** out) {
  if (!grid || !out) {
    g_createError = "null argument";
    return KOFM_ERR_ARG;
  }
  *out = nullptr;
  if (grid->nX < 4 || grid->nY < 4) {
    g_createError = "lattice must be at least 4 x 4";
    return KOFM_ERR_ARG;
  }
  auto* c = new kofm_ctx;
  c->nX = grid->nX;
  c->nY = grid->nY;
  c->storage = grid->storage;
  c->komega = grid->use_komega != 0;
  *out = c;
  return KOFM_OK;
}

void kofm_destroy(kofm_ctx* ctx) { delete ctx; }

const char* kofm_last_error(const kofm_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_createError.c_str();
}

int kofm_set_params(kofm_ctx* ctx, const kofm_params* p) {
  if (!ctx || !p) return KOFM_ERR_ARG;
  if (!(p->tau0 > 0. && p->tau0 < 2.))
    return setErr(ctx, KOFM_ERR_ARG, "tau0 (relaxation frequency) must lie in (0, 2)");
  kofm::Params& q = ctx->params;
  q.tau0 = p->tau0;
  q.Fval = p->Fval;
  q.nu = p->nu;
  q.rho0 = p->rho0;
  q.maxNut = p->max_turb_visc;
  q.koDt = p->ko_dt;
  q.relTol[0] = p->kappa_rel_tol;
  q.absTol[0] = p->kappa_abs_tol;
  q.maxIters[0] = p->kappa_max_iters;
  q.relTol[1] = p->omega_rel_tol;
  q.absTol[1] = p->omega_abs_tol;
  q.maxIters[1] = p->omega_max_iters;
  if (q.maxIters[0] < 1 || q.maxIters[1] < 1)
    return setErr(ctx, KOFM_ERR_ARG, "solver iteration limits must be >= 1");
  if (ctx->engine) ctx->engine->setParams(q);
  return KOFM_OK;
}

int kofm_build_geometry(kofm_ctx* ctx, const double* C0, int32_t nN, const int32_t* BL,
                        int32_t nBL, const kofm_geom_params* gp) {
  if (!ctx || !C0 || !BL || !gp || nN < 2 || nBL < 1) return KOFM_ERR_ARG;
  kofm::GeometryInput in;
  in.nX = ctx->nX;
  in.nY = ctx->nY;
  in.C0.assign(C0, C0 + 2 * size_t(nN));
  in.BL.assign(BL, BL + 2 * size_t(nBL));
  in.pipeRadius = gp->pipe_radius;
  in.eps = gp->eps;
  in.minDx = gp->min_dx;
  in.wallDSearchRad = gp->wall_d_search_rad;
  in.xRelease = gp->x_release_pos;
  in.xMin = gp->x_min_val;
  in.xStop = gp->x_stop_pos;
  in.cutoff = gp->cutoff_radius;
  in.indexRadiusSearch = gp->index_radius_search;
  std::string err;
  int rc = kofm::buildGeometry(in, ctx->geo, err);
  if (rc) return setErr(ctx, rc, err);
  if (ctx->engine) {
    rc = ctx->engine->uploadGeometry(ctx->geo, err);
    if (rc) return setErr(ctx, rc, err);
  }
  ctx->pipeRadius = gp->pipe_radius;
  ctx->haveGeo = true;
  return KOFM_OK;
}

int kofm_update_walls(kofm_ctx* ctx, const double* C, int32_t nN) {
  if (!ctx || !C || nN < 2) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  if (!ctx->haveGeo) return setErr(ctx, KOFM_ERR_STATE, "build the geometry first");
  if (size_t(2 * int64_t(nN)) != ctx->geo.C0.size())
    return setErr(ctx, KOFM_ERR_ARG, "the wall update needs the built geometry's number of wall points");
  std::string err;
  int rc = ctx->engine->updateWalls(std::vector<double>(C, C + 2 * size_t(nN)), err);
  if (rc) return setErr(ctx, rc, err);
  ctx->geo = ctx->engine->geometry();
  return KOFM_OK;
}

int kofm_geometry_counts(const kofm_ctx* ctx, kofm_geom_counts* out) {
  if (!ctx || !out) return KOFM_ERR_ARG;
  const kofm::Geometry& g = ctx->geo;
  out->nX = ctx->nX;
  out->nY = ctx->nY;
  out->XsizeFull = ((ctx->nX + 127) / 128) * 128;
  out->D = int64_t(out->XsizeFull) * ctx->nY;
  out->nIbb = int32_t(g.ibb.size());
  out->nSS = int32_t(g.ss.size());
  out->nBLshear = g.nBLshear();
  int nodes = 0;
  {
    std::vector<char> seen(size_t(out->D), 0);
    for (int32_t l : g.ibb) {
      size_t node = size_t(l % out->D);
      if (!seen[node]) {
        seen[node] = 1;
        nodes++;
      }
    }
  }
  out->nIbbNodes = nodes;
  return KOFM_OK;
}

int kofm_geometry_array(const kofm_ctx* ctx, int32_t which, void* out) {
  if (!ctx || !out) return KOFM_ERR_ARG;
  if (!ctx->haveGeo) return KOFM_ERR_STATE;
  const kofm::Geometry& g = ctx->geo;
  auto put = [&](const auto& v) {
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
    return KOFM_OK;
  };
  switch (which) {
    case KOFM_GEOM_NTYPE: return put(g.nType);
    case KOFM_GEOM_DX: return put(g.dX);
    case KOFM_GEOM_IBB: return put(g.ibb);
    case KOFM_GEOM_IBBQ: return put(g.ibbQ);
    case KOFM_GEOM_SS: return put(g.ss);
    case KOFM_GEOM_SSMAP: return put(g.ssMap);
    case KOFM_GEOM_SSINDS: return put(g.ssInds);
    case KOFM_GEOM_LSMAP: return put(g.lsMap);
    case KOFM_GEOM_WALLD: return put(g.wallD);
    case KOFM_GEOM_CFSS: return put(g.cfss);
    case KOFM_GEOM_SIND: return put(g.sInd);
    case KOFM_GEOM_WEIGHTS: return put(g.weights);
    case KOFM_GEOM_BOUNDS: std::memcpy(out, g.bounds, sizeof(g.bounds)); return KOFM_OK;
    case KOFM_GEOM_MASSES: std::memcpy(out, g.masses, sizeof(g.masses)); return KOFM_OK;
    case KOFM_GEOM_IBB_SEG: return put(g.ibbSeg);
    case KOFM_GEOM_IBB_DQDY: return put(g.ibbDqDy);
    default: return KOFM_ERR_ARG;
  }
}

int kofm_ini_poiseuille(const kofm_ctx* ctx, const kofm_ini_flow* p, double* rho, double* ux,
                        double* uy) {
  if (!ctx || !p || !rho || !ux || !uy) return KOFM_ERR_ARG;
  if (!ctx->haveGeo) return KOFM_ERR_STATE;
  if (!(p->pipe_radius > 0.) || !(p->nu > 0.) || !(p->Re > 0.)) return KOFM_ERR_ARG;
  kofm::iniPoiseuille(ctx->geo, *p, rho, ux, uy);
  return KOFM_OK;
}

int kofm_init_state(kofm_ctx* ctx, const double* rho, const double* ux, const double* uy) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  if (!ctx->haveGeo) return setErr(ctx, KOFM_ERR_STATE, "build the geometry first");
  std::string err;
  int rc = ctx->engine->initState(rho, ux, uy, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_init_komega(kofm_ctx* ctx, double k_ini, double omega_ini, double Re,
                     double geom_penalty_factor) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->initKOmega(k_ini, omega_ini, Re, geom_penalty_factor, ctx->pipeRadius, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_upload_state(kofm_ctx* ctx, const double* f, const double* k, const double* omega,
                      const double* nut, const double* S) {
  if (!ctx || !f) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->uploadState(f, k, omega, nut, S, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_step(kofm_ctx* ctx, int32_t nsteps, uint32_t options) {
  if (!ctx || nsteps < 0) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->step(nsteps, options, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_time_steps(kofm_ctx* ctx, int32_t nsteps, uint32_t options, double* ms_per_step,
                    double* ms_per_kernel) {
  if (!ctx || !ms_per_step) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->timeSteps(nsteps, options, *ms_per_step, ms_per_kernel, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_wall_shear(kofm_ctx* ctx) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->wallShear(err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_komega_time_step(kofm_ctx* ctx, int32_t action, int32_t* steps_per_lb,
                          double* time_step) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int spl = 0;
  double dt = 0.;
  int rc = ctx->engine->komegaTimeStep(action, spl, dt, err);
  if (rc) return setErr(ctx, rc, err);
  if (steps_per_lb) *steps_per_lb = spl;
  if (time_step) *time_step = dt;
  return KOFM_OK;
}

int kofm_thermal_init(kofm_ctx* ctx, const kofm_thermal_params* q) {
  if (!ctx || !q) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  kofm::ThermalParams t;
  t.alphaFluid = q->alpha_fluid;
  t.alphaFoul = q->alpha_foul;
  t.prTurb = q->pr_turb;
  t.tIn = q->t_in;
  t.tInitial = q->t_initial;
  t.dt = q->dt;
  t.relTol = q->rel_tol;
  t.absTol = q->abs_tol;
  t.maxIters = q->max_iters;
  std::string err;
  int rc = ctx->engine->thermalInit(t, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_thermal_step(kofm_ctx* ctx, int32_t n) {
  if (!ctx || n < 0) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->thermalStep(n, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_thermal_time_step(kofm_ctx* ctx, int32_t action, int32_t* steps_per_lb, double* time_step) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int spl = 0;
  double dt = 0.;
  int rc = ctx->engine->thermalTimeStep(action, spl, dt, err);
  if (rc) return setErr(ctx, rc, err);
  if (steps_per_lb) *steps_per_lb = spl;
  if (time_step) *time_step = dt;
  return KOFM_OK;
}

int kofm_set_temperature(kofm_ctx* ctx, const double* T) {
  if (!ctx || !T) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->setTemperature(T, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_nusselt_init(kofm_ctx* ctx, const kofm_nusselt_params* q) {
  if (!ctx || !q) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  kofm::NusseltParams p;
  p.height = q->height;
  p.wavyStart = q->wavy_start;
  p.wavyPeriod = q->wavy_period;
  p.wavyPeriods = q->wavy_periods;
  p.kSoot = q->k_soot;
  p.kAir = q->k_air;
  std::string err;
  int rc = ctx->engine->nusseltInit(p, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_nusselt_sizes(kofm_ctx* ctx, int32_t* columns, int32_t* locations) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int c = 0, l = 0;
  int rc = ctx->engine->nusseltSizes(c, l, err);
  if (rc) return setErr(ctx, rc, err);
  if (columns) *columns = c;
  if (locations) *locations = l;
  return KOFM_OK;
}

int kofm_nusselt(kofm_ctx* ctx, double* nu, double* mean, int32_t* nodes, double* vec) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->nusselt(nu, mean, nodes, vec, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_shear_removal(kofm_ctx* ctx, const kofm_removal_params* q, const kofm_particles* par,
                       int64_t offset, int64_t count, uint32_t* bl_dep) {
  if (!ctx || !q || !par) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  kofm::RemovalParams p;
  p.nSizes = q->n_sizes;
  p.tauCrit = q->tau_crit;
  p.dcoeff = q->drag_coeff;
  p.lcoeff = q->lift_coeff;
  p.mp = q->mass;
  p.dtWall = q->dt_wall;
  p.foulSwitch = q->foul_size_switch;
  p.locStride = q->loc_stride;
  p.timerStart = q->timer_start;
  kofm::ParticleArrays P;
  P.pos = par->pos;
  P.numRep = par->num_rep;
  P.type = par->type;
  P.depFlag = par->dep_flag;
  P.depTimer = par->dep_timer;
  P.loc = par->loc;
  P.timer = par->timer;
  std::string err;
  int rc = ctx->engine->shearRemoval(p, P, offset, count, bl_dep, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_removal_links(kofm_ctx* ctx, double foul_size_switch, double* p0, double* vN, double* len,
                       int16_t* surf) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->removalLinks(foul_size_switch, p0, vN, len, surf, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_get_solver_stats(kofm_ctx* ctx, kofm_solver_stats* out) {
  if (!ctx || !out) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  kofm::SolverStats s;
  std::string err;
  int rc = ctx->engine->solverStats(s, err);
  if (rc) return setErr(ctx, rc, err);
  out->kappa_iters = s.iters[0];
  out->omega_iters = s.iters[1];
  out->kappa_res0 = s.res0[0];
  out->kappa_res = s.res[0];
  out->omega_res0 = s.res0[1];
  out->omega_res = s.res[1];
  out->nan_flag = s.nan;
  out->steps_per_lb = s.stepsPerLB;
  out->time_step = s.timeStep;
  out->batches = s.batches;
  out->batch_rollbacks = s.rollbacks;
  out->temp_iters = s.tIters;
  out->temp_res0 = s.tRes0;
  out->temp_res = s.tRes;
  out->temp_steps_per_lb = s.tStepsPerLB;
  out->temp_time_step = s.tTimeStep;
  return KOFM_OK;
}

int kofm_reduce(kofm_ctx* ctx, int32_t field, int32_t op, double* out) {
  if (!ctx || !out) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->reduce(field, op, *out, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_download(kofm_ctx* ctx, int32_t field, double* host) {
  if (!ctx || !host) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->download(field, host, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_save2file(kofm_ctx* ctx, const char* dir) {
  if (!ctx || !dir) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = kofm::saveFields(*ctx->engine, ctx->nX, ctx->nY, ctx->komega, dir, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_save_restart(kofm_ctx* ctx, const char* dir) {
  if (!ctx || !dir) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  if (!ctx->haveGeo) return setErr(ctx, KOFM_ERR_STATE, "build the geometry first");
  std::string err;
  int rc = kofm::saveRestart(*ctx->engine, ctx->nX, ctx->nY, ctx->komega, ctx->geo.C0, dir, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_save_debug(kofm_ctx* ctx, const char* dir) {
  if (!ctx || !dir) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  if (!ctx->haveGeo) return setErr(ctx, KOFM_ERR_STATE, "build the geometry first");
  std::string err;
  int rc = kofm::saveDebug(*ctx->engine, ctx->geo, ctx->komega, dir, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_set_launch(kofm_ctx* ctx, int32_t block_x, int32_t nodes_per_lane,
                    int32_t steps_per_pass, int64_t plane_pad) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->setLaunch(block_x, nodes_per_lane, steps_per_pass, plane_pad, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_autotune(kofm_ctx* ctx, int32_t steps, int32_t* block_x, int32_t* nodes_per_lane,
                  int32_t* steps_per_pass, int64_t* plane_pad, double* ms_per_step) {
  if (!ctx || !block_x || !nodes_per_lane || !steps_per_pass || !plane_pad || !ms_per_step)
    return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int b = 0, npt = 1, spp = 1;
  int64_t pad = 0;
  double ms = 0.;
  int rc = ctx->engine->autotune(steps, b, npt, spp, pad, ms, err);
  if (rc) return setErr(ctx, rc, err);
  *block_x = b;
  *nodes_per_lane = npt;
  *steps_per_pass = spp;
  *plane_pad = pad;
  *ms_per_step = ms;
  return KOFM_OK;
}

int kofm_objective_run(kofm_ctx* ctx, int32_t nsteps, int32_t checkpoint_interval,
                       const kofm_objective* obj, double* J) {
  if (!ctx || !obj || !J) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  kofm::Objective o;
  o.window = obj->window;
  o.xIn = obj->x_in;
  o.xOut = obj->x_out;
  o.shearWeight = obj->shear_weight;
  o.flags = obj->flags;
  if (o.flags & ~KOFM_ADJOINT_FROZEN_TURBULENCE)
    return setErr(ctx, KOFM_ERR_ARG, "unknown kofm_objective flags");
  std::string err;
  int rc = ctx->engine->objectiveRun(nsteps, checkpoint_interval, o, *J, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_adjoint(kofm_ctx* ctx, double* dJ_dy, double* dJ_dq) {
  if (!ctx || !dJ_dy) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->adjoint(dJ_dy, dJ_dq, err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

int kofm_geometry_field_gradient(const kofm_ctx* ctx, const double* wall_d_bar,
                                 const double* dx_bar, double* dJ_dy) {
  if (!ctx || !dJ_dy) return KOFM_ERR_ARG;
  if (!ctx->haveGeo) return KOFM_ERR_STATE;
  std::vector<double> dy;
  kofm::wallFieldGradient(ctx->geo, wall_d_bar, dx_bar, dy);
  for (size_t i = 0; i < dy.size(); i++) dJ_dy[i] += dy[i];
  return KOFM_OK;
}

int kofm_set_graphs(kofm_ctx* ctx, int32_t on) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  ctx->engine->setGraphs(on != 0);
  return KOFM_OK;
}

int kofm_synchronize(kofm_ctx* ctx) {
  if (!ctx) return KOFM_ERR_ARG;
  NEED_ENGINE(ctx);
  std::string err;
  int rc = ctx->engine->synchronize(err);
  return rc ? setErr(ctx, rc, err) : KOFM_OK;
}

struct kofm_ensemble {
  std::unique_ptr<kofm::Ensemble> impl;
  std::string err;
};

int kofm_ensemble_create(kofm_ctx* const* members, int32_t n, kofm_ensemble** out) {
  if (!members || n < 1 || !out) return KOFM_ERR_ARG;
  *out = nullptr;
  std::vector<kofm::Engine*> es;
  for (int i = 0; i < n; i++) {
    if (!members[i] || !members[i]->engine) {
      g_createError = "ensemble members must be device contexts";
      return KOFM_ERR_ARG;
    }
    es.push_back(members[i]->engine.get());
  }
  std::string err;
  auto impl = kofm::makeEnsemble(es, err);
  if (!impl) {
    g_createError = err;
    return KOFM_ERR_STATE;
  }
  auto* e = new kofm_ensemble;
  e->impl = std::move(impl);
  *out = e;
  return KOFM_OK;
}

int kofm_ensemble_step(kofm_ensemble* e, int32_t nsteps, uint32_t options) {
  if (!e || nsteps < 0) return KOFM_ERR_ARG;
  int rc = e->impl->step(nsteps, options, e->err);
  return rc ? rc : KOFM_OK;
}

int kofm_ensemble_time_steps(kofm_ensemble* e, int32_t nsteps, uint32_t options,
                             double* ms_per_step, int64_t* fluid_nodes) {
  if (!e || nsteps < 1 || !ms_per_step) return KOFM_ERR_ARG;
  int rc = e->impl->timeSteps(nsteps, options, *ms_per_step, e->err);
  if (fluid_nodes) *fluid_nodes = e->impl->fluidNodes();
  return rc ? rc : KOFM_OK;
}

int kofm_ensemble_set_block(kofm_ensemble* e, int32_t block_x) {
  if (!e) return KOFM_ERR_ARG;
  int rc = e->impl->setBlock(block_x, e->err);
  return rc ? rc : KOFM_OK;
}

int kofm_ensemble_set_launch(kofm_ensemble* e, int32_t block_x, int32_t steps_per_launch) {
  if (!e) return KOFM_ERR_ARG;
  int rc = e->impl->setLaunch(block_x, steps_per_launch, e->err);
  return rc ? rc : KOFM_OK;
}

int kofm_ensemble_set_cache_budget(kofm_ensemble* e, int64_t bytes, int32_t* groups) {
  if (!e) return KOFM_ERR_ARG;
  int rc = e->impl->setCacheBudget(bytes, e->err);
  if (!rc && groups) *groups = e->impl->groups();
  return rc ? rc : KOFM_OK;
}

int kofm_ensemble_autotune(kofm_ensemble* e, int32_t steps, int32_t* block_x,
                           int32_t* steps_per_launch, int64_t* cache_budget, double* ms_per_step) {
  if (!e || steps < 1) return KOFM_ERR_ARG;
  int block = 0, spp = 0;
  int64_t budget = 0;
  double ms = 0.;
  int rc = e->impl->autotune(steps, block, spp, budget, ms, e->err);
  if (rc) return rc;
  if (block_x) *block_x = block;
  if (steps_per_launch) *steps_per_launch = spp;
  if (cache_budget) *cache_budget = budget;
  if (ms_per_step) *ms_per_step = ms;
  return KOFM_OK;
}

const char* kofm_ensemble_last_error(const kofm_ensemble* e) {
  return e ? e->err.c_str() : g_createError.c_str();
}

void kofm_ensemble_destroy(kofm_ensemble* e) { delete e; }

}  // extern "C"
