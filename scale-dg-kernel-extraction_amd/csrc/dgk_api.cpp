// C-ABI entry points of libdgk: solver lifetime, field transfer, the
// piecewise reference stage loop (update_halo / cal_elembnd_flux / cal_dqdt /
// RK update) and the fused device time loop.
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "solver.hpp"

namespace dgk {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
void clear_error() { g_last_error.clear(); }


}  // namespace dgk

using namespace dgk;


namespace {

hipEvent_t take_event(dgk_solver *s) {
  if (!s->free_events.empty()) {
    hipEvent_t e = s->free_events.back();
    s->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// One event pair on the solver's stream around a sequence of `count`
// launches, when timing is on.  A sequence left before done() (an error
// return) gives its events back to the pool.
struct Timed {
  dgk_solver *s;
  int slot;
  int64_t count;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(dgk_solver *s_, int slot_, bool on, int64_t count_ = 1)
      : s(s_), slot(slot_), count(count_) {
    if (!on) return;
    a = take_event(s);
    b = take_event(s);
    if (a && b) (void)hipEventRecord(a, s->stream);
  }
  Timed(const Timed &) = delete;
  Timed &operator=(const Timed &) = delete;
  ~Timed() {
    if (a) s->free_events.push_back(a);
    if (b) s->free_events.push_back(b);
  }
  void done() {
    if (a && b) {
      (void)hipEventRecord(b, s->stream);
      s->pending.push_back({slot, a, b, count});
      a = b = nullptr;
    }
  }
};

// Fold the event pairs that have completed, oldest first, without waiting;
// keeps the pending list short for callers that step for a long time
// without synchronising.
int poll_timers(dgk_solver *s) {
  size_t done = 0;
  for (; done < s->pending.size(); ++done) {
    const auto &p = s->pending[done];
    const hipError_t q = hipEventQuery(p.b);
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) return set_error(DGK_ERR_DEVICE, std::string("hipEventQuery: ") +
                                                              hipGetErrorString(q));
    float ms = 0.f;
    DGK_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    s->sec[p.slot] += 1e-3 * (double)ms;
    s->cnt[p.slot] += p.count;
    s->free_events.push_back(p.a);
    s->free_events.push_back(p.b);
  }
  s->pending.erase(s->pending.begin(), s->pending.begin() + (std::ptrdiff_t)done);
  return DGK_OK;
}

int resolve_timers(dgk_solver *s) {
  if (s->pending.empty()) return DGK_OK;
  DGK_HIP(hipStreamSynchronize(s->stream));
  for (auto &p : s->pending) {
    float ms = 0.f;
    DGK_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    s->sec[p.slot] += 1e-3 * (double)ms;
    s->cnt[p.slot] += p.count;
    s->free_events.push_back(p.a);
    s->free_events.push_back(p.b);
  }
  s->pending.clear();
  return DGK_OK;
}

int check_solver(const dgk_solver *s) {
  if (!s) return set_error(DGK_ERR_ARG, "solver handle is NULL");
  return DGK_OK;
}

int enter(const dgk_solver *s) {
  clear_error();
  DGK_TRY(check_solver(s));
  DGK_HIP(hipSetDevice(s->device));
  return DGK_OK;
}

int require_fields(const dgk_solver *s) {
  for (int f = 0; f < 4; ++f)
    if (!s->set[f])
      return set_error(DGK_ERR_STATE, "fields q, u, v, w must be set before computing");
  return DGK_OK;
}

}  // namespace

namespace dgk {

double *field_ptr(dgk_solver *s, int which) {
  switch (which) {
  case DGK_FIELD_Q: return s->q[s->cur];
  case DGK_FIELD_U: return s->u;
  case DGK_FIELD_V: return s->v;
  case DGK_FIELD_W: return s->w;
  case DGK_FIELD_DQDT: return s->dqdt;
  default: return nullptr;
  }
}

int refresh_f32(dgk_solver *s, int which) {
  if (s->evaltype != DGK_EVAL_HIP_MP) return DGK_OK;
  float *dst = which == DGK_FIELD_U ? s->u32 : which == DGK_FIELD_V ? s->v32 : s->w32;
  return launch_to_f32(field_ptr(s, which), dst, (int64_t)s->Np * s->Ne, s->stream);
}

// Backwards launches walk each XCD's own slot range backwards where the
// order is XCD-interleaved: an XCD then starts a launch on the elements it
// ended the previous one with (per launch at about 1.7e7 DOF, against
// walking the whole order backwards: p = 7 141.1 -> 139.4 us FP64,
// 104.8 -> 103.2 us mixed; p = 1..4, 15 0-1% faster).  DGK_ALT=global walks
// the whole order backwards, DGK_ALT=0 always forwards.
int alternate_launches() {
  const char *alt = std::getenv("DGK_ALT");
  if (alt && alt[0] == '0') return 0;
  if (alt && std::string(alt) == "global") return 1;
  return 2;
}

StageArgs stage_args(dgk_solver *s, const double *qin, const double *q0,
                     double *out, int mode, double ca, double cb, double dt) {
  StageArgs a;
  a.q = qin; a.q0 = q0; a.u = s->u; a.v = s->v; a.w = s->w; a.out = out;
  a.flux_in = nullptr; a.nbr = s->nbr; a.vmapP = s->vmapP; a.D = s->D; a.Lift = s->Lift;
  a.ell = s->ell; a.geo = s->geo; a.geoid = s->geoid; a.order = s->order; a.reverse = 0;
  a.zperiod = s->zperiod; a.Ne = s->Ne;
  a.ca = ca; a.cb = cb; a.dt = dt; a.mode = mode; a.axis_aligned = s->axis_aligned;
  a.dt_geo = 0;
  a.tile = s->tile;
  a.geo_per_run = s->geo_per_run;
  a.cvel = s->cvel;
  return a;
}

// One stage launch on `stream` (the solver's own stream unless a batch
// supplies one); dt_geo takes the time step from each element's geometry
// record instead of `dt`.
StageArgsMP stage_args_mp(dgk_solver *s, const double *qin, const double *q0,
                          double *out, int mode, double ca, double cb, double dt) {
  StageArgsMP a;
  a.q = qin; a.q0 = q0; a.u = s->u32; a.v = s->v32; a.w = s->w32; a.out = out;
  a.nbr = s->nbr; a.vmapP = s->vmapP; a.D = s->D32; a.ell = s->ell32;
  a.D64 = s->D; a.ell64 = s->ell;
  a.geo = s->geo; a.geoid = s->geoid; a.order = s->order; a.reverse = 0;
  a.zperiod = s->zperiod; a.Ne = s->Ne;
  a.ca = ca; a.cb = cb; a.dt = dt; a.mode = mode; a.axis_aligned = s->axis_aligned;
  a.dt_geo = 0;
  a.tile = s->tile;
  a.geo_per_run = s->geo_per_run;
  a.cvel = s->cvel;
  return a;
}

int launch_stage(dgk_solver *s, const double *qin, const double *q0, double *out,
                 int mode, double ca, double cb, double dt, hipStream_t stream, int dt_geo) {
  if (!stream) stream = s->stream;
  const int reverse = (s->launches++ & 1) ? s->alternate : 0;
  if (s->evaltype == DGK_EVAL_HIP_MP) {
    StageArgsMP a = stage_args_mp(s, qin, q0, out, mode, ca, cb, dt);
    a.reverse = reverse;
    a.dt_geo = dt_geo;
    return launch_stage_mp(s->Nq, a, stream);
  }
  StageArgs a = stage_args(s, qin, q0, out, mode, ca, cb, dt);
  a.reverse = reverse;
  a.dt_geo = dt_geo;
  return launch_stage_f64(s->Nq, a, stream);
}

}  // namespace dgk

namespace dgk {

// Separable 1D form of the stored Lift: Lift(:,:,:,f) = I (x) l-/+ (x) I
// along the face normal (faces -y,+x,+y,-x,-z,+z use l-(j), l+(i), l+(j),
// l-(i), l-(k), l+(k)).  l-/+ are read from the -x and +x blocks at
// tangential node (0,0); returns the largest deviation of any stored entry
// from that form relative to max|Lift|.
double separable_lift(int Nq, const std::vector<double> &Lift, std::vector<double> &ell) {
  const int Np = Nq * Nq * Nq;
  ell.assign(2 * Nq, 0.0);
  for (int a = 0; a < Nq; ++a) {
    ell[a] = Lift[a + 3 * Np];       // face -x: Lift(a,0,0,4) = l-(a)
    ell[Nq + a] = Lift[a + 1 * Np];  // face +x: Lift(a,0,0,2) = l+(a)
  }
  double scale = 0.0;
  for (double x : Lift) scale = std::fmax(scale, std::fabs(x));
  double dev = 0.0;
  for (int f = 0; f < 6; ++f)
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int axis_idx = (f == 1 || f == 3) ? i : (f == 0 || f == 2) ? j : k;
          const bool plus = (f == 1 || f == 2 || f == 5);
          const double ref = ell[(plus ? Nq : 0) + axis_idx];
          dev = std::fmax(dev, std::fabs(Lift[i + j * Nq + k * Nq * Nq + f * Np] - ref));
        }
  return scale > 0.0 ? dev / scale : dev;
}

}  // namespace dgk

namespace dgk {

int create_from_tables(int evaltype, MeshTables &t, dgk_solver **out) {
  if (!out) return set_error(DGK_ERR_ARG, "output handle pointer is NULL");
  *out = nullptr;
  if (evaltype == DGK_EVAL_GENERAL || evaltype == DGK_EVAL_OPT1)
    return set_error(DGK_ERR_UNSUPPORTED,
                     "GENERAL and OPT1 name the reference CPU kernels; use HIP or HIP_MP");
  if (evaltype != DGK_EVAL_HIP && evaltype != DGK_EVAL_HIP_MP)
    return set_error(DGK_ERR_ARG, "unknown evaluation type id " + std::to_string(evaltype));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return set_error(DGK_ERR_DEVICE, "no HIP device available (the path has no CPU fallback)");
  }
  std::vector<double> ell;
  const double sep_dev = separable_lift(t.Nq, t.Lift, ell);
  if (evaltype == DGK_EVAL_HIP_MP && !(sep_dev <= kSeparableMP))
    return set_error(DGK_ERR_UNSUPPORTED,
                     "Lift_mat is not of tensor-product form (relative deviation " +
                         std::to_string(sep_dev) + "); HIP_MP needs the separable Lift");
  const bool use_ell = sep_dev <= kSeparableF64;

  auto *s = new dgk_solver();
  auto fail = [&](int rc) {
    dgk_destroy(s);
    return rc;
  };
  int rc;
  s->evaltype = evaltype;
  if (hipGetDevice(&s->device) != hipSuccess) return fail(set_error(DGK_ERR_DEVICE, "hipGetDevice failed"));
  s->Nq = t.Nq; s->Np = t.Np; s->Nfp = t.Nfp; s->NfpTot = t.NfpTot;
  s->Ne = t.Ne; s->NeA = t.NeA; s->NhaloNode = t.NhaloNode;
  s->nfield = (int64_t)t.Np * t.Ne + t.NhaloNode;
  s->axis_aligned = t.axis_aligned ? 1 : 0;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_error(DGK_ERR_DEVICE, "hipStreamCreate failed"));
  for (int b = 0; b < 3; ++b)
    if ((rc = dalloc(s, &s->q[b], s->nfield + kStagePad))) return fail(rc);
  if ((rc = dalloc(s, &s->u, s->nfield + kStagePad)) ||
      (rc = dalloc(s, &s->v, s->nfield + kStagePad)) ||
      (rc = dalloc(s, &s->w, s->nfield + kStagePad)))
    return fail(rc);
  if ((rc = dalloc(s, &s->nbr, (int64_t)t.nbr.size())) ||
      (rc = dalloc(s, &s->vmapP, (int64_t)t.vmapP.size())) ||
      (rc = dalloc(s, &s->halo, (int64_t)t.halo_src.size())) ||
      (rc = dalloc(s, &s->geo, (int64_t)t.geo.size())) ||
      (rc = dalloc(s, &s->geoid, (int64_t)t.geoid.size())) ||
      (rc = dalloc(s, &s->D, (int64_t)t.D.size())) ||
      (rc = dalloc(s, &s->Lift, (int64_t)t.Lift.size())) ||
      (rc = dalloc(s, &s->mm, 2 * kMinmaxBlocks + 2)) ||
      (rc = dalloc(s, &s->order, (int64_t)t.order.size())))
    return fail(rc);
  if (evaltype == DGK_EVAL_HIP_MP) {
    const int64_t n = (int64_t)t.Np * t.Ne;
    if ((rc = dalloc(s, &s->u32, n)) || (rc = dalloc(s, &s->v32, n)) ||
        (rc = dalloc(s, &s->w32, n)) || (rc = dalloc(s, &s->D32, (int64_t)t.D.size())) ||
        (rc = dalloc(s, &s->ell32, (int64_t)ell.size())))
      return fail(rc);
    std::vector<float> D32(t.D.begin(), t.D.end()), ell32(ell.begin(), ell.end());
    if ((rc = upload(s, s->D32, D32)) || (rc = upload(s, s->ell32, ell32))) return fail(rc);
  }
  if (use_ell) {
    if ((rc = dalloc(s, &s->ell, (int64_t)ell.size())) || (rc = upload(s, s->ell, ell)))
      return fail(rc);
  }
  if ((rc = upload(s, s->nbr, t.nbr)) || (rc = upload(s, s->vmapP, t.vmapP)) ||
      (rc = upload(s, s->halo, t.halo_src)) || (rc = upload(s, s->geo, t.geo)) ||
      (rc = upload(s, s->geoid, t.geoid)) || (rc = upload(s, s->D, t.D)) ||
      (rc = upload(s, s->Lift, t.Lift)) || (rc = upload(s, s->order, t.order)))
    return fail(rc);
  // Consecutive stage launches walk the elements in opposite directions: a
  // launch starts where the previous one ended, on the fields it touched
  // last, which are still in the memory-side cache (DGK_ALT=0: always
  // forwards).  The arithmetic of an element does not depend on when it runs.
  s->alternate = alternate_launches();
  s->zperiod = t.zperiod;
  if (!t.geoid.empty()) {
    const int T = run_elements(t.Nq);  // the mixed run kernel's runs are the same
    for (int64_t e = 0; T > 0 && e < t.Ne; ++e)
      if (t.geoid[e] != t.geoid[e - e % T]) {
        s->geo_per_run = 0;
        break;
      }
  }
  for (int b = 0; b < 3; ++b)
    if (hipMemsetAsync(s->q[b], 0, sizeof(double) * s->nfield, s->stream) != hipSuccess)
      return fail(set_error(DGK_ERR_DEVICE, "hipMemsetAsync failed"));
  if (hipStreamSynchronize(s->stream) != hipSuccess)
    return fail(set_error(DGK_ERR_DEVICE, "stream synchronize failed after upload"));
  if ((rc = evaltype == DGK_EVAL_HIP_MP ? prepare_stage_mp(s->Nq) : prepare_stage_f64(s->Nq)))
    return fail(rc);
  *out = s;
  return DGK_OK;
}

}  // namespace dgk

extern "C" {

const char *dgk_last_error(void) { return g_last_error.c_str(); }

int dgk_version(void) { return 100; }

int dgk_device_count(int *count) {
  clear_error();
  if (!count) return set_error(DGK_ERR_ARG, "count is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *count = n;
  return DGK_OK;
}

int dgk_evaltype_id(const char *name, int *id) {
  clear_error();
  if (!name || !id) return set_error(DGK_ERR_ARG, "null argument");
  // trim trailing blanks, as Fortran's trim() does (mod_dg_optr_kernel.f90:50)
  std::string s(name);
  while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
  if (s == "GENERAL") *id = DGK_EVAL_GENERAL;
  else if (s == "OPT1") *id = DGK_EVAL_OPT1;
  else if (s == "HIP") *id = DGK_EVAL_HIP;
  else if (s == "HIP_MP") *id = DGK_EVAL_HIP_MP;
  else return set_error(DGK_ERR_ARG, "Unsupported dgoptr_evaltype: " + s);
  return DGK_OK;
}

int dgk_create(int evaltype, int Nq, int64_t Ne, int64_t NeA, int64_t NhaloNode,
               const double *D1D, const double *Lift_mat, const int *VMapM,
               const int *VMapP, const int *halo_src_map, const double *normal_fn,
               const double *Escale, const double *Fscale, dgk_solver **out) {
  clear_error();
  if (!out) return set_error(DGK_ERR_ARG, "output handle pointer is NULL");
  *out = nullptr;
  MeshTables t;
  DGK_TRY(build_tables_from_reference(Nq, Ne, NeA, NhaloNode, D1D, Lift_mat, VMapM,
                                      VMapP, halo_src_map, normal_fn, Escale, Fscale, t));
  return create_from_tables(evaltype, t, out);
}

int dgk_create_structured(int evaltype, int NeX, int NeY, int NeZ, int p,
                          const char *operator_dir, double dom_x, double dom_y,
                          double dom_z, dgk_solver **out) {
  clear_error();
  if (!out) return set_error(DGK_ERR_ARG, "output handle pointer is NULL");
  *out = nullptr;
  std::vector<double> x, D, L;
  DGK_TRY(load_operators(operator_dir, p, x, D, L));
  MeshTables t;
  DGK_TRY(build_tables_structured(NeX, NeY, NeZ, p, D.data(), L.data(), dom_x, dom_y,
                                  dom_z, t));
  DGK_TRY(create_from_tables(evaltype, t, out));
  dgk_solver *s = *out;
  s->box.NeX = NeX; s->box.NeY = NeY; s->box.NeZ = NeZ;
  s->box.dom[0] = dom_x; s->box.dom[1] = dom_y; s->box.dom[2] = dom_z;
  int rc;
  if ((rc = dalloc(s, &s->x1D, (int64_t)x.size())) || (rc = upload(s, s->x1D, x)) ||
      (rc = (hipStreamSynchronize(s->stream) == hipSuccess ? DGK_OK
                                                            : set_error(DGK_ERR_DEVICE, "sync")))) {
    dgk_destroy(s);
    *out = nullptr;
    return rc;
  }
  return DGK_OK;
}

int dgk_init_gaussian(dgk_solver *s, double cx, double cy, double cz, double sigma,
                      double ux, double uy, double uz) {
  DGK_TRY(enter(s));
  if (!s->x1D || s->box.NeX < 1)
    return set_error(DGK_ERR_UNSUPPORTED, "device initial conditions need a structured solver");
  if (!(sigma > 0)) return set_error(DGK_ERR_ARG, "sigma must be positive");
  const double h[3] = {s->box.dom[0] / s->box.NeX, s->box.dom[1] / s->box.NeY,
                       s->box.dom[2] / s->box.NeZ};
  const double c[3] = {cx, cy, cz}, vel[3] = {ux, uy, uz};
  DGK_TRY(launch_init_gaussian(s->q[s->cur], s->u, s->v, s->w, s->x1D, s->Nq, s->box.NeX,
                               s->box.NeY, s->box.NeZ, h, c, sigma, vel, s->stream));
  const int64_t owned = (int64_t)s->Np * s->Ne;
  for (int which = DGK_FIELD_U; which <= DGK_FIELD_W; ++which) {
    DGK_TRY(launch_halo_gather(field_ptr(s, which), s->halo, owned, s->NhaloNode, s->stream));
    DGK_TRY(refresh_f32(s, which));
  }
  for (int f = 0; f < 4; ++f) s->set[f] = true;
  s->q0_valid = s->flux_valid = s->dqdt_valid = false;
  DGK_HIP(hipStreamSynchronize(s->stream));
  return DGK_OK;
}

int dgk_destroy(dgk_solver *s) {
  clear_error();
  if (!s) return DGK_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto &p : s->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  for (void *p : s->allocs) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return DGK_OK;
}

int dgk_solver_info(const dgk_solver *s, int *Nq, int64_t *Ne, int64_t *NeA,
                    int64_t *NhaloNode) {
  clear_error();
  DGK_TRY(check_solver(s));
  if (Nq) *Nq = s->Nq;
  if (Ne) *Ne = s->Ne;
  if (NeA) *NeA = s->NeA;
  if (NhaloNode) *NhaloNode = s->NhaloNode;
  return DGK_OK;
}

int dgk_device_bytes(const dgk_solver *s, int64_t *bytes) {
  clear_error();
  DGK_TRY(check_solver(s));
  if (!bytes) return set_error(DGK_ERR_ARG, "bytes is NULL");
  *bytes = s->bytes;
  return DGK_OK;
}

int dgk_set_field(dgk_solver *s, int which, const double *f, int64_t n) {
  DGK_TRY(enter(s));
  if (!f) return set_error(DGK_ERR_ARG, "field array is NULL");
  if (which < DGK_FIELD_Q || which > DGK_FIELD_W)
    return set_error(DGK_ERR_ARG, "field id must be Q, U, V or W");
  const int64_t owned = (int64_t)s->Np * s->Ne;
  if (n != owned && n != (int64_t)s->Np * s->NeA)
    return set_error(DGK_ERR_ARG, "field length must be Np*Ne or Np*NeA");
  double *dst = field_ptr(s, which);
  DGK_HIP(hipMemcpyAsync(dst, f, sizeof(double) * owned, hipMemcpyHostToDevice, s->stream));
  if (which != DGK_FIELD_Q) {
    DGK_TRY(launch_halo_gather(dst, s->halo, owned, s->NhaloNode, s->stream));
    DGK_TRY(refresh_f32(s, which));
  } else {
    s->q0_valid = false;
  }
  s->flux_valid = s->dqdt_valid = false;
  s->set[which] = true;
  DGK_HIP(hipStreamSynchronize(s->stream));
  return DGK_OK;
}

int dgk_get_field(dgk_solver *s, int which, double *f, int64_t n) {
  DGK_TRY(enter(s));
  if (!f) return set_error(DGK_ERR_ARG, "field array is NULL");
  if (which < DGK_FIELD_Q || which > DGK_FIELD_DQDT)
    return set_error(DGK_ERR_ARG, "unknown field id");
  const int64_t owned = (int64_t)s->Np * s->Ne;
  const bool with_halo = n == (int64_t)s->Np * s->NeA;
  if (n != owned && !with_halo) return set_error(DGK_ERR_ARG, "field length must be Np*Ne or Np*NeA");
  if (which == DGK_FIELD_DQDT && !s->dqdt) return set_error(DGK_ERR_STATE, "dqdt has not been computed");
  const double *src = field_ptr(s, which);
  DGK_HIP(hipMemcpyAsync(f, src, sizeof(double) * owned, hipMemcpyDeviceToHost, s->stream));
  if (with_halo) {
    const int64_t tail = (int64_t)s->Np * s->NeA - owned;
    if (which == DGK_FIELD_DQDT) {
      std::memset(f + owned, 0, sizeof(double) * tail);
    } else {
      DGK_HIP(hipMemcpyAsync(f + owned, src + owned, sizeof(double) * s->NhaloNode,
                             hipMemcpyDeviceToHost, s->stream));
      if (tail > s->NhaloNode)
        std::memset(f + owned + s->NhaloNode, 0, sizeof(double) * (tail - s->NhaloNode));
    }
  }
  DGK_HIP(hipStreamSynchronize(s->stream));
  return DGK_OK;
}

int dgk_update_halo(dgk_solver *s, int which) {
  DGK_TRY(enter(s));
  if (which < DGK_FIELD_Q || which > DGK_FIELD_W)
    return set_error(DGK_ERR_ARG, "field id must be Q, U, V or W");
  return launch_halo_gather(field_ptr(s, which), s->halo, (int64_t)s->Np * s->Ne,
                            s->NhaloNode, s->stream);
}

int dgk_copy_q0(dgk_solver *s) {
  DGK_TRY(enter(s));
  DGK_TRY(require_fields(s));
  const int64_t owned = (int64_t)s->Np * s->Ne;
  if (!s->q0) DGK_TRY(dalloc(s, &s->q0, owned));
  DGK_HIP(hipMemcpyAsync(s->q0, s->q[s->cur], sizeof(double) * owned,
                         hipMemcpyDeviceToDevice, s->stream));
  s->q0_valid = true;
  return DGK_OK;
}

int dgk_cal_elembnd_flux(dgk_solver *s, double *flux_host) {
  DGK_TRY(enter(s));
  DGK_TRY(require_fields(s));
  const int64_t n = (int64_t)s->NfpTot * s->Ne;
  if (!s->flux) DGK_TRY(dalloc(s, &s->flux, n));
  Timed tm(s, T_FLUX, s->kernel_timing);
  DGK_TRY(launch_elembnd_flux_f64(s->Nq, s->q[s->cur], s->u, s->v, s->w, s->nbr, s->vmapP,
                                  s->geo, s->geoid, s->Ne, s->flux, s->stream));
  tm.done();
  s->flux_valid = true;
  if (flux_host) {
    DGK_HIP(hipMemcpyAsync(flux_host, s->flux, sizeof(double) * n, hipMemcpyDeviceToHost,
                           s->stream));
    DGK_HIP(hipStreamSynchronize(s->stream));
  }
  return DGK_OK;
}

int dgk_cal_dqdt(dgk_solver *s, const double *flux_host) {
  DGK_TRY(enter(s));
  DGK_TRY(require_fields(s));
  if (s->evaltype != DGK_EVAL_HIP)
    return set_error(DGK_ERR_UNSUPPORTED, "the unfused cal_dqdt is provided by the FP64 path only");
  const int64_t n = (int64_t)s->NfpTot * s->Ne;
  if (!s->flux) DGK_TRY(dalloc(s, &s->flux, n));
  if (flux_host) {
    DGK_HIP(hipMemcpyAsync(s->flux, flux_host, sizeof(double) * n, hipMemcpyHostToDevice,
                           s->stream));
    s->flux_valid = true;
  }
  if (!s->flux_valid) return set_error(DGK_ERR_STATE, "no element boundary flux computed");
  if (!s->dqdt) DGK_TRY(dalloc(s, &s->dqdt, (int64_t)s->Np * s->Ne));
  StageArgs a = stage_args(s, s->q[s->cur], nullptr, s->dqdt, 0, 0.0, 1.0, 0.0);
  a.flux_in = s->flux;
  Timed tm(s, T_DQDT, s->kernel_timing);
  DGK_TRY(launch_stage_f64(s->Nq, a, s->stream));
  tm.done();
  s->dqdt_valid = true;
  if (flux_host) DGK_HIP(hipStreamSynchronize(s->stream));
  return DGK_OK;
}

int dgk_cal_tend(dgk_solver *s, double *dqdt_host) {
  DGK_TRY(enter(s));
  DGK_TRY(require_fields(s));
  const int64_t n = (int64_t)s->Np * s->Ne;
  if (!s->dqdt) DGK_TRY(dalloc(s, &s->dqdt, n));
  Timed tm(s, T_CALTEND, s->kernel_timing);
  DGK_TRY(launch_stage(s, s->q[s->cur], nullptr, s->dqdt, 0, 0.0, 1.0, 0.0));
  tm.done();
  s->dqdt_valid = true;
  if (dqdt_host) {
    DGK_HIP(hipMemcpyAsync(dqdt_host, s->dqdt, sizeof(double) * n, hipMemcpyDeviceToHost,
                           s->stream));
    DGK_HIP(hipStreamSynchronize(s->stream));
  }
  return DGK_OK;
}

int dgk_rk_update(dgk_solver *s, int stage, double dt) {
  DGK_TRY(enter(s));
  if (stage < 1 || stage > 3) return set_error(DGK_ERR_ARG, "stage must be 1, 2 or 3");
  if (!s->q0_valid) return set_error(DGK_ERR_STATE, "q0 not set (dgk_copy_q0)");
  if (!s->dqdt_valid) return set_error(DGK_ERR_STATE, "dqdt not computed");
  DGK_TRY(launch_rk_update(s->q[s->cur], s->q0, s->dqdt, kRkA[stage - 1], kRkB[stage - 1], dt,
                           (int64_t)s->Np * s->Ne, s->stream));
  s->flux_valid = s->dqdt_valid = false;
  return DGK_OK;
}

int dgk_step(dgk_solver *s, double dt, int nstep) {
  DGK_TRY(enter(s));
  DGK_TRY(require_fields(s));
  if (nstep < 0) return set_error(DGK_ERR_ARG, "nstep must be >= 0");
  if (nstep == 0) return DGK_OK;
  // Buffer rotation: A holds q at the step start and serves as q0; stage 1
  // A->B, stage 2 (B, q0=A)->C, stage 3 (C, q0=A)->A.  Stage 3 overwrites
  // A node-for-node after reading q0 there, and no element reads another
  // element's q0, so no copy of q0 is needed.
  DGK_TRY(poll_timers(s));
  double *A = s->q[s->cur], *B = s->q[(s->cur + 1) % 3], *C = s->q[(s->cur + 2) % 3];
  Timed main_tm(s, T_MAIN, true);
  // with kernel timing on, one event pair spans the call's 3 nstep stage
  // launches, so timing adds nothing between them
  Timed tm(s, T_STAGE, s->kernel_timing, 3 * (int64_t)nstep);
  for (int it = 0; it < nstep; ++it) {
    const double *in[3] = {A, B, C};
    double *out[3] = {B, C, A};
    for (int st = 0; st < 3; ++st)
      DGK_TRY(launch_stage(s, in[st], A, out[st], 1, kRkA[st], kRkB[st], dt));
  }
  tm.done();
  main_tm.done();
  s->q0_valid = s->flux_valid = s->dqdt_valid = false;
  return DGK_OK;
}

int dgk_minmax(dgk_solver *s, double *qmin, double *qmax) {
  DGK_TRY(enter(s));
  if (!s->set[DGK_FIELD_Q]) return set_error(DGK_ERR_STATE, "q not set");
  const int64_t n = (int64_t)s->Np * s->Ne;
  DGK_TRY(launch_minmax(s->q[s->cur], n, s->mm, kMinmaxBlocks, s->mm + 2 * kMinmaxBlocks,
                        s->stream));
  double r[2];
  DGK_HIP(hipMemcpyAsync(r, s->mm + 2 * kMinmaxBlocks, sizeof(r), hipMemcpyDeviceToHost,
                         s->stream));
  DGK_HIP(hipStreamSynchronize(s->stream));
  if (qmin) *qmin = r[0];
  if (qmax) *qmax = r[1];
  return resolve_timers(s);  // the stream is idle: folding costs no wait
}

// Bounds of a device allocation holding [p, p + bytes), or an error: a host
// pointer or a range past the allocation's end would fault the GPU.
static int check_device_range(const void *p, int64_t bytes, const char *what) {
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(DGK_ERR_ARG, std::string(what) + " is not device memory");
  }
  if (at.type != hipMemoryTypeDevice)
    return set_error(DGK_ERR_ARG, std::string(what) + " is not device memory");
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(DGK_ERR_ARG, std::string(what) + ": no device allocation holds it");
  }
  const int64_t off = (const char *)p - (const char *)base;
  if (off < 0 || off + bytes > (int64_t)size)
    return set_error(DGK_ERR_ARG, std::string(what) + ": " + std::to_string(bytes) +
                                      " bytes run past its allocation");
  return DGK_OK;
}

int dgk_minmax_noalloc(const double *q, int64_t n, double *result, void *stream) {
  clear_error();
  if (!q || !result) return set_error(DGK_ERR_ARG, "null argument");
  if (n < 1 || n > kNoallocMaxN)
    return set_error(DGK_ERR_ARG, "n = " + std::to_string(n) + " outside 1.." +
                                      std::to_string(kNoallocMaxN));
  DGK_TRY(check_device_range(q, n * (int64_t)sizeof(double), "q"));
  DGK_TRY(check_device_range(result, 2 * (int64_t)sizeof(double), "result"));
  return launch_minmax_noalloc(q, n, result, (hipStream_t)stream);
}

int dgk_synchronize(dgk_solver *s) {
  DGK_TRY(enter(s));
  DGK_HIP(hipStreamSynchronize(s->stream));
  return resolve_timers(s);
}

int dgk_set_kernel_timing(dgk_solver *s, int on) {
  DGK_TRY(enter(s));
  s->kernel_timing = on != 0;
  return DGK_OK;
}

int dgk_timer_report(dgk_solver *s, double *seconds, int64_t *counts) {
  DGK_TRY(enter(s));
  DGK_TRY(resolve_timers(s));
  for (int i = 0; i < T_N; ++i) {
    if (seconds) seconds[i] = s->sec[i];
    if (counts) counts[i] = s->cnt[i];
  }
  return DGK_OK;
}

int dgk_timer_reset(dgk_solver *s) {
  DGK_TRY(enter(s));
  DGK_TRY(resolve_timers(s));
  for (int i = 0; i < T_N; ++i) {
    s->sec[i] = 0.0;
    s->cnt[i] = 0;
  }
  return DGK_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched ensembles: instances of one polynomial order are concatenated into
// one element space (build_tables_batch) and advanced by the same stage
// kernels in one launch per stage; the per-instance time step comes from the
// geometry record.  One SSP-RK3 step of every group is captured once into a
// hipGraph (one parallel branch per order group) and replayed.
// ---------------------------------------------------------------------------
struct dgk_batch {
  int evaltype = 0, device = 0;
  hipStream_t stream = nullptr;  // capture origin
  std::vector<dgk_solver *> groups;
  std::vector<int> group_p;
  // rec: the instance's geometry record in its group; cv[d]: velocity
  // component d is one constant, held at word kGeoVel + d of that record
  struct Loc {
    int group; int64_t elem_off; int64_t Ne; int p; BoxShape box;
    int rec = 0; bool cv[3] = {false, false, false};
  };
  std::vector<Loc> loc;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  // Two captured steps, replayed in turn: stage launches alternate their
  // direction, so step 2k runs forwards-backwards-forwards and step 2k+1
  // backwards-forwards-backwards, as an uncaptured sequence of steps would.
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  int next = 0;  // the captured step to replay next
  bool use_graph = true;
  double step_seconds = 0.0;
  int64_t step_calls = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

namespace {

int batch_enter(const dgk_batch *b) {
  clear_error();
  if (!b) return set_error(DGK_ERR_ARG, "batch handle is NULL");
  DGK_HIP(hipSetDevice(b->device));
  return DGK_OK;
}

// The stages of one RK step of group g on `stream` (buffer rotation as in
// dgk_step: A -> B -> C -> A).
int batch_group_step(dgk_solver *s, hipStream_t stream) {
  double *A = s->q[s->cur], *B = s->q[(s->cur + 1) % 3], *C = s->q[(s->cur + 2) % 3];
  const double *in[3] = {A, B, C};
  double *out[3] = {B, C, A};
  for (int st = 0; st < 3; ++st)
    DGK_TRY(launch_stage(s, in[st], A, out[st], 1, kRkA[st], kRkB[st], 0.0, stream, 1));
  return DGK_OK;
}

int batch_enqueue_step(dgk_batch *b) {
  DGK_HIP(hipEventRecord(b->fork, b->stream));
  for (size_t g = 0; g < b->groups.size(); ++g) {
    hipStream_t gs = b->groups[g]->stream;
    DGK_HIP(hipStreamWaitEvent(gs, b->fork, 0));
    DGK_TRY(batch_group_step(b->groups[g], gs));
    DGK_HIP(hipEventRecord(b->join[g], gs));
    DGK_HIP(hipStreamWaitEvent(b->stream, b->join[g], 0));
  }
  return DGK_OK;
}

void batch_release_graphs(dgk_batch *b) {
  for (int k = 0; k < 2; ++k) {
    if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    b->exec[k] = nullptr;
    b->graph[k] = nullptr;
  }
  b->next = 0;
}

int batch_capture(dgk_batch *b) {
  for (int k = 0; k < 2; ++k) {
    DGK_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_enqueue_step(b);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc != DGK_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      batch_release_graphs(b);
      return rc;
    }
    if (e != hipSuccess) {
      batch_release_graphs(b);
      return set_error(DGK_ERR_DEVICE, std::string("stream capture: ") + hipGetErrorString(e));
    }
    b->graph[k] = graph;
    if (hipGraphInstantiate(&b->exec[k], b->graph[k], nullptr, nullptr, 0) != hipSuccess) {
      batch_release_graphs(b);
      return set_error(DGK_ERR_DEVICE, "hipGraphInstantiate failed");
    }
  }
  b->next = 0;
  return DGK_OK;
}

// Fold the batch's completed timing pairs, oldest first; with wait, all of
// them after the batch stream has drained.
int batch_fold_timers(dgk_batch *b, bool wait) {
  if (wait) DGK_HIP(hipStreamSynchronize(b->stream));
  size_t done = 0;
  for (; done < b->pending.size(); ++done) {
    auto &p = b->pending[done];
    if (!wait) {
      const hipError_t q = hipEventQuery(p.second);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) return set_error(DGK_ERR_DEVICE, std::string("hipEventQuery: ") +
                                                                hipGetErrorString(q));
    }
    float ms = 0.f;
    DGK_HIP(hipEventElapsedTime(&ms, p.first, p.second));
    b->step_seconds += 1e-3 * ms;
    b->step_calls += 1;
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  b->pending.erase(b->pending.begin(), b->pending.begin() + (std::ptrdiff_t)done);
  return DGK_OK;
}

// Record instance inst's constant velocity components (vel[d], for the d with
// set[d]) in its geometry record, or mark them non-constant (set[d] and not
// constant[d]); a group launches its constant-velocity kernels when every one
// of its boxes has all three components constant.  A change of the group's
// form releases the captured steps.
int batch_set_velocity(dgk_batch *b, int inst, const bool set[3], const bool constant[3],
                       const double vel[3]) {
  dgk_batch::Loc &l = b->loc[inst];
  dgk_solver *s = b->groups[l.group];
  for (int d = 0; d < 3; ++d) {
    if (!set[d]) continue;
    l.cv[d] = constant[d];
    if (constant[d])
      DGK_HIP(hipMemcpy(s->geo + (int64_t)l.rec * kGeoStride + kGeoVel + d, &vel[d],
                        sizeof(double), hipMemcpyHostToDevice));
  }
  // DGK_CVEL=0 keeps the per-node velocity reads (for A/B timings)
  const char *env = std::getenv("DGK_CVEL");
  int all = (s->axis_aligned && !(env && env[0] == '0')) ? 1 : 0;
  for (const auto &m : b->loc)
    if (m.group == l.group && !(m.cv[0] && m.cv[1] && m.cv[2])) all = 0;
  if (all != s->cvel) {
    s->cvel = all;
    batch_release_graphs(b);
  }
  return DGK_OK;
}

int batch_loc(const dgk_batch *b, int inst, const dgk_batch::Loc **l) {
  if (inst < 0 || inst >= (int)b->loc.size())
    return set_error(DGK_ERR_ARG, "instance index " + std::to_string(inst) + " out of range");
  *l = &b->loc[inst];
  return DGK_OK;
}

}  // namespace

extern "C" {

int dgk_batch_destroy(dgk_batch *b) {
  clear_error();
  if (!b) return DGK_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto &p : b->pending) {
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  batch_release_graphs(b);
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (dgk_solver *s : b->groups) dgk_destroy(s);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return DGK_OK;
}

int dgk_batch_create(int evaltype, int ninst, const int *ne_xyz, const int *p,
                     const double *dom_xyz, const double *dt, const char *operator_dir,
                     dgk_batch **out) {
  clear_error();
  if (!out) return set_error(DGK_ERR_ARG, "output handle pointer is NULL");
  *out = nullptr;
  if (ninst < 1 || !ne_xyz || !p || !dt) return set_error(DGK_ERR_ARG, "empty batch or NULL array");
  if (evaltype != DGK_EVAL_HIP && evaltype != DGK_EVAL_HIP_MP)
    return set_error(DGK_ERR_UNSUPPORTED, "batches run on HIP or HIP_MP");
  auto *b = new dgk_batch();
  auto fail = [&](int rc) {
    std::string msg = dgk_last_error();
    dgk_batch_destroy(b);
    set_error(rc, msg);
    return rc;
  };
  b->evaltype = evaltype;
  if (hipGetDevice(&b->device) != hipSuccess) return fail(set_error(DGK_ERR_DEVICE, "hipGetDevice failed"));
  b->loc.resize(ninst);
  // group instances by order, in order of first appearance
  std::vector<std::vector<int>> members;
  for (int k = 0; k < ninst; ++k) {
    int g = -1;
    for (size_t j = 0; j < b->group_p.size(); ++j)
      if (b->group_p[j] == p[k]) g = (int)j;
    if (g < 0) {
      g = (int)b->group_p.size();
      b->group_p.push_back(p[k]);
      members.emplace_back();
    }
    members[g].push_back(k);
  }
  for (size_t g = 0; g < b->group_p.size(); ++g) {
    std::vector<double> x, D, L;
    int rc = load_operators(operator_dir, b->group_p[g], x, D, L);
    if (rc) return fail(rc);
    std::vector<BatchInstance> inst;
    for (int k : members[g]) {
      BatchInstance bi;
      bi.NeX = ne_xyz[3 * k]; bi.NeY = ne_xyz[3 * k + 1]; bi.NeZ = ne_xyz[3 * k + 2];
      for (int d = 0; d < 3; ++d) bi.dom[d] = dom_xyz ? dom_xyz[3 * k + d] : 1.0;
      bi.dt = dt[k];
      inst.push_back(bi);
    }
    MeshTables t;
    std::vector<int64_t> off;
    if ((rc = build_tables_batch(b->group_p[g], D.data(), L.data(), inst, t, off))) return fail(rc);
    dgk_solver *s = nullptr;
    if ((rc = create_from_tables(evaltype, t, &s))) return fail(rc);
    b->groups.push_back(s);
    if ((rc = dalloc(s, &s->x1D, (int64_t)x.size())) || (rc = upload(s, s->x1D, x))) return fail(rc);
    for (size_t j = 0; j < members[g].size(); ++j) {
      const int k = members[g][j];
      BoxShape bs;
      bs.NeX = inst[j].NeX; bs.NeY = inst[j].NeY; bs.NeZ = inst[j].NeZ;
      for (int d = 0; d < 3; ++d) bs.dom[d] = inst[j].dom[d];
      b->loc[k].group = (int)g;
      b->loc[k].elem_off = off[j];
      b->loc[k].Ne = off[j + 1] - off[j];
      b->loc[k].p = b->group_p[g];
      b->loc[k].box = bs;
      b->loc[k].rec = (int)j;
    }
    s->set[DGK_FIELD_Q] = s->set[DGK_FIELD_U] = s->set[DGK_FIELD_V] = s->set[DGK_FIELD_W] = true;
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    return fail(set_error(DGK_ERR_DEVICE, "stream/event creation failed"));
  b->join.resize(b->groups.size(), nullptr);
  for (auto &e : b->join)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return fail(set_error(DGK_ERR_DEVICE, "event creation failed"));
  *out = b;
  return DGK_OK;
}

int dgk_batch_info(const dgk_batch *b, int inst, int *p, int64_t *Ne, int64_t *dof) {
  DGK_TRY(batch_enter(b));
  const dgk_batch::Loc *l;
  DGK_TRY(batch_loc(b, inst, &l));
  const int64_t np = (int64_t)(l->p + 1) * (l->p + 1) * (l->p + 1);
  if (p) *p = l->p;
  if (Ne) *Ne = l->Ne;
  if (dof) *dof = l->Ne * np;
  return DGK_OK;
}

int dgk_batch_size(const dgk_batch *b, int *ninst, int *ngroups, int64_t *bytes) {
  DGK_TRY(batch_enter(b));
  if (ninst) *ninst = (int)b->loc.size();
  if (ngroups) *ngroups = (int)b->groups.size();
  if (bytes) {
    int64_t t = 0;
    for (const dgk_solver *s : b->groups) t += s->bytes;
    *bytes = t;
  }
  return DGK_OK;
}

int dgk_batch_set_field(dgk_batch *b, int inst, int which, const double *f, int64_t n) {
  DGK_TRY(batch_enter(b));
  const dgk_batch::Loc *l;
  DGK_TRY(batch_loc(b, inst, &l));
  if (!f) return set_error(DGK_ERR_ARG, "field array is NULL");
  if (which < DGK_FIELD_Q || which > DGK_FIELD_W) return set_error(DGK_ERR_ARG, "field id must be Q, U, V or W");
  dgk_solver *s = b->groups[l->group];
  const int64_t off = l->elem_off * s->Np, cnt = l->Ne * s->Np;
  if (n != cnt) return set_error(DGK_ERR_ARG, "field length must be Np*Ne of the instance");
  // a replayed step runs on the batch stream: let it finish before overwriting
  DGK_HIP(hipStreamSynchronize(b->stream));
  double *dst = field_ptr(s, which) + off;
  DGK_HIP(hipMemcpyAsync(dst, f, sizeof(double) * cnt, hipMemcpyHostToDevice, s->stream));
  if (which != DGK_FIELD_Q && s->evaltype == DGK_EVAL_HIP_MP) {
    float *d32 = (which == DGK_FIELD_U ? s->u32 : which == DGK_FIELD_V ? s->v32 : s->w32) + off;
    DGK_TRY(launch_to_f32(dst, d32, cnt, s->stream));
  }
  DGK_HIP(hipStreamSynchronize(s->stream));
  if (which != DGK_FIELD_Q) {
    // a velocity component that is one constant over the box is read from
    // the instance's geometry record instead of the field
    const int d = which - DGK_FIELD_U;
    bool set[3] = {false, false, false}, constant[3] = {false, false, false};
    double vel[3] = {0.0, 0.0, 0.0};
    set[d] = true;
    constant[d] = std::all_of(f, f + cnt, [&](double x) {
      return std::memcmp(&x, f, sizeof(double)) == 0;
    });
    vel[d] = f[0];
    DGK_TRY(batch_set_velocity(b, inst, set, constant, vel));
  }
  return DGK_OK;
}

int dgk_batch_get_field(dgk_batch *b, int inst, int which, double *f, int64_t n) {
  DGK_TRY(batch_enter(b));
  const dgk_batch::Loc *l;
  DGK_TRY(batch_loc(b, inst, &l));
  if (!f) return set_error(DGK_ERR_ARG, "field array is NULL");
  if (which < DGK_FIELD_Q || which > DGK_FIELD_W) return set_error(DGK_ERR_ARG, "field id must be Q, U, V or W");
  dgk_solver *s = b->groups[l->group];
  const int64_t off = l->elem_off * s->Np, cnt = l->Ne * s->Np;
  if (n != cnt) return set_error(DGK_ERR_ARG, "field length must be Np*Ne of the instance");
  DGK_HIP(hipStreamSynchronize(b->stream));
  DGK_HIP(hipMemcpyAsync(f, field_ptr(s, which) + off, sizeof(double) * cnt,
                         hipMemcpyDeviceToHost, s->stream));
  DGK_HIP(hipStreamSynchronize(s->stream));
  return DGK_OK;
}

int dgk_batch_init_gaussian(dgk_batch *b, int inst, double cx, double cy, double cz,
                            double sigma, double ux, double uy, double uz) {
  DGK_TRY(batch_enter(b));
  const dgk_batch::Loc *l;
  DGK_TRY(batch_loc(b, inst, &l));
  if (!(sigma > 0)) return set_error(DGK_ERR_ARG, "sigma must be positive");
  dgk_solver *s = b->groups[l->group];
  const int64_t off = l->elem_off * s->Np, cnt = l->Ne * s->Np;
  const double h[3] = {l->box.dom[0] / l->box.NeX, l->box.dom[1] / l->box.NeY,
                       l->box.dom[2] / l->box.NeZ};
  const double c[3] = {cx, cy, cz}, vel[3] = {ux, uy, uz};
  DGK_HIP(hipStreamSynchronize(b->stream));
  DGK_TRY(launch_init_gaussian(s->q[s->cur] + off, s->u + off, s->v + off, s->w + off, s->x1D,
                               s->Nq, l->box.NeX, l->box.NeY, l->box.NeZ, h, c, sigma, vel,
                               s->stream));
  if (s->evaltype == DGK_EVAL_HIP_MP) {
    DGK_TRY(launch_to_f32(s->u + off, s->u32 + off, cnt, s->stream));
    DGK_TRY(launch_to_f32(s->v + off, s->v32 + off, cnt, s->stream));
    DGK_TRY(launch_to_f32(s->w + off, s->w32 + off, cnt, s->stream));
  }
  DGK_HIP(hipStreamSynchronize(s->stream));
  const bool all[3] = {true, true, true};
  return batch_set_velocity(b, inst, all, all, vel);
}

int dgk_batch_set_graph(dgk_batch *b, int on) {
  DGK_TRY(batch_enter(b));
  b->use_graph = on != 0;
  return DGK_OK;
}

int dgk_batch_step(dgk_batch *b, int nstep) {
  DGK_TRY(batch_enter(b));
  if (nstep < 0) return set_error(DGK_ERR_ARG, "nstep must be >= 0");
  if (nstep == 0) return DGK_OK;
  // fields uploaded on the group streams are ordered before the step
  for (dgk_solver *s : b->groups) DGK_HIP(hipStreamSynchronize(s->stream));
  if (b->use_graph && !b->exec[0]) DGK_TRY(batch_capture(b));
  DGK_TRY(batch_fold_timers(b, false));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  DGK_HIP(hipEventCreate(&t0));
  DGK_HIP(hipEventCreate(&t1));
  DGK_HIP(hipEventRecord(t0, b->stream));
  for (int it = 0; it < nstep; ++it) {
    if (b->use_graph) {
      DGK_HIP(hipGraphLaunch(b->exec[b->next], b->stream));
      b->next ^= 1;
    } else {
      DGK_TRY(batch_enqueue_step(b));
    }
  }
  DGK_HIP(hipEventRecord(t1, b->stream));
  b->pending.push_back({t0, t1});
  return DGK_OK;
}

int dgk_batch_synchronize(dgk_batch *b) {
  DGK_TRY(batch_enter(b));
  return batch_fold_timers(b, true);
}

int dgk_batch_timer(dgk_batch *b, double *seconds, int64_t *calls) {
  DGK_TRY(dgk_batch_synchronize(b));
  if (seconds) *seconds = b->step_seconds;
  if (calls) *calls = b->step_calls;
  b->step_seconds = 0.0;
  b->step_calls = 0;
  return DGK_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Tile shapes and the autotuner.  A stage kernel has one or more tile shapes
// per order (elements per workgroup); the selection is per process, per form
// and order.  The autotuner times every candidate on a solver's own mesh with
// HIP events around `reps` RK-stage launches (stage-2 configuration: RK base
// read, output to the solver's scratch rotation buffer) and keeps the
// fastest.  Per-element arithmetic does not depend on the shape.
// ---------------------------------------------------------------------------
namespace {

int tile_shapes(int evaltype, int Nq, int *epb) {
  if (evaltype == DGK_EVAL_HIP_MP) return stage_mp_shapes(Nq, epb, kMaxTileShapes);
  if (evaltype == DGK_EVAL_HIP) return stage_f64_shapes(Nq, epb, kMaxTileShapes);
  return 0;
}

int check_form_order(int evaltype, int p) {
  if (evaltype != DGK_EVAL_HIP && evaltype != DGK_EVAL_HIP_MP)
    return set_error(DGK_ERR_UNSUPPORTED, "tile shapes exist for HIP and HIP_MP only");
  if (p < kMinNq - 1 || p > kMaxNq - 1)
    return set_error(DGK_ERR_UNSUPPORTED, "polynomial order outside 1.." + std::to_string(kMaxNq - 1));
  return DGK_OK;
}

int tile_select(int evaltype, int Nq, int cand) {
  if (evaltype == DGK_EVAL_HIP_MP) return stage_mp_select(Nq, cand);
  return stage_f64_select(Nq, cand);
}

int tile_selected(int evaltype, int Nq) {
  return evaltype == DGK_EVAL_HIP_MP ? stage_mp_selected(Nq) : stage_f64_selected(Nq);
}

// The process-wide selection that launches of s take: constant-velocity
// launches (batches, served requests) have defaults of their own.
int solver_selected(const dgk_solver *s) {
  if (!s->cvel) return tile_selected(s->evaltype, s->Nq);
  return s->evaltype == DGK_EVAL_HIP_MP ? stage_mp_selected_cvel(s->Nq)
                                        : stage_f64_selected_cvel(s->Nq);
}

// Average device time of one stage launch of s, in microseconds.
int time_stage(dgk_solver *s, int reps, double *us) {
  double *A = s->q[s->cur], *C = s->q[(s->cur + 2) % 3];
  hipEvent_t a = nullptr, b = nullptr;
  DGK_HIP(hipEventCreate(&a));
  DGK_HIP(hipEventCreate(&b));
  int rc = launch_stage(s, A, A, C, 1, kRkA[1], kRkB[1], 0.0);
  if (rc == DGK_OK && hipEventRecord(a, s->stream) != hipSuccess) rc = set_error(DGK_ERR_DEVICE, "event");
  for (int r = 0; r < reps && rc == DGK_OK; ++r) rc = launch_stage(s, A, A, C, 1, kRkA[1], kRkB[1], 0.0);
  if (rc == DGK_OK && hipEventRecord(b, s->stream) != hipSuccess) rc = set_error(DGK_ERR_DEVICE, "event");
  if (rc == DGK_OK && hipEventSynchronize(b) != hipSuccess) rc = set_error(DGK_ERR_DEVICE, "event sync");
  float ms = 0.f;
  if (rc == DGK_OK && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = set_error(DGK_ERR_DEVICE, "elapsed");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc == DGK_OK) *us = 1e3 * ms / reps;
  return rc;
}

constexpr double kTuneMargin = 0.02;

// Whether candidate c launches its own kernel for s's stage launches.
bool candidate_runs(dgk_solver *s, int c) {
  double *A = s->q[s->cur], *C = s->q[(s->cur + 2) % 3];
  if (s->evaltype == DGK_EVAL_HIP_MP) {
    StageArgsMP a = stage_args_mp(s, A, A, C, 1, kRkA[1], kRkB[1], 0.0);
    return stage_mp_candidate_runs(s->Nq, c, a);
  }
  return stage_f64_candidate_runs(s->Nq, c, stage_args(s, A, A, C, 1, kRkA[1], kRkB[1], 0.0));
}

int autotune(dgk_solver *s, int reps, int *best, double *us, int max, int *n) {
  int epb[kMaxTileShapes];
  const int nc = tile_shapes(s->evaltype, s->Nq, epb);
  if (nc < 1) return set_error(DGK_ERR_UNSUPPORTED, "no tile shapes for this form and order");
  double t[kMaxTileShapes] = {};
  int arg = -1;
  for (int c = 0; c < nc; ++c) {
    // a candidate that would launch another shape here is not timed (0)
    if (!candidate_runs(s, c)) {
      if (us && c < max) us[c] = 0.0;
      continue;
    }
    s->tile = c;
    const int rc = time_stage(s, reps, &t[c]);
    if (rc != DGK_OK) {
      s->tile = -1;
      return rc;
    }
    if (us && c < max) us[c] = t[c];
    if (arg < 0 || t[c] < t[arg]) arg = c;
  }
  if (arg < 0) return set_error(DGK_ERR_UNSUPPORTED, "no tile shape applies");
  // keep the default unless another shape is at least kTuneMargin faster
  const int def = solver_selected(s);
  if (def >= 0 && def < nc && t[def] > 0.0 && t[arg] > (1.0 - kTuneMargin) * t[def]) arg = def;
  s->tile = arg;
  if (best) *best = arg;
  if (n) *n = nc;
  return DGK_OK;
}

}  // namespace

extern "C" {

int dgk_tile_shapes(int evaltype, int p, int *epb, int max, int *n) {
  clear_error();
  DGK_TRY(check_form_order(evaltype, p));
  if (!n || (max > 0 && !epb)) return set_error(DGK_ERR_ARG, "NULL argument");
  int e[kMaxTileShapes];
  const int nc = tile_shapes(evaltype, p + 1, e);
  for (int c = 0; c < nc && c < max; ++c) epb[c] = e[c];
  *n = nc;
  return DGK_OK;
}

int dgk_tile_select(int evaltype, int p, int cand) {
  clear_error();
  DGK_TRY(check_form_order(evaltype, p));
  return tile_select(evaltype, p + 1, cand);
}

int dgk_tile_selected(int evaltype, int p, int *cand) {
  clear_error();
  DGK_TRY(check_form_order(evaltype, p));
  if (!cand) return set_error(DGK_ERR_ARG, "NULL argument");
  *cand = tile_selected(evaltype, p + 1);
  return DGK_OK;
}

int dgk_solver_tile(const dgk_solver *s, int *cand) {
  clear_error();
  DGK_TRY(check_solver(s));
  if (!cand) return set_error(DGK_ERR_ARG, "NULL argument");
  *cand = s->tile >= 0 ? s->tile : solver_selected(s);
  return DGK_OK;
}

int dgk_autotune(dgk_solver *s, int reps, int *best, double *us_per_launch, int max, int *n) {
  DGK_TRY(enter(s));
  DGK_TRY(require_fields(s));
  if (reps < 1) return set_error(DGK_ERR_ARG, "reps must be >= 1");
  return autotune(s, reps, best, us_per_launch, max, n);
}

int dgk_batch_autotune(dgk_batch *b, int reps, int *p, int *best, double *us_best,
                       double *us_default, int max, int *n) {
  DGK_TRY(batch_enter(b));
  if (reps < 1) return set_error(DGK_ERR_ARG, "reps must be >= 1");
  // the timing launches write a rotation buffer that a replayed step on the
  // batch stream may still be using
  DGK_HIP(hipStreamSynchronize(b->stream));
  for (dgk_solver *s : b->groups) DGK_HIP(hipStreamSynchronize(s->stream));
  for (size_t g = 0; g < b->groups.size(); ++g) {
    dgk_solver *s = b->groups[g];
    double t[kMaxTileShapes] = {};
    int arg = 0, nc = 0;
    DGK_TRY(autotune(s, reps, &arg, t, kMaxTileShapes, &nc));
    if ((int)g < max) {
      if (p) p[g] = b->group_p[g];
      if (best) best[g] = arg;
      if (us_best) us_best[g] = t[arg];
      const int def = solver_selected(s);
      if (us_default) us_default[g] = t[def >= 0 && def < nc ? def : 0];  // 0: not applicable
    }
  }
  if (n) *n = (int)b->groups.size();
  // the captured steps used the previous shapes
  batch_release_graphs(b);
  return DGK_OK;
}

}  // extern "C"
