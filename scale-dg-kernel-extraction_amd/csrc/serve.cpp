// Serving loop of libdgk: continuous batching of independent advection solves
// (dgk_serve_*).  The reference driver runs one mesh to completion
// (main.f90:51-79); here many periodic boxes of mixed size and order share one
// device:
//
//  * memory   one arena of the configured HBM budget, allocated once and cut
//             first-fit into per-request blocks (fields, connectivity,
//             geometry record, summary partials); operators are shared per
//             order.  Admission never calls hipMalloc, so it neither blocks
//             nor synchronises the device.
//  * admission arrival order, a request enters when its block fits and one of
//             max_active slots is free; later requests that fit may pass a
//             waiting one at most kMaxBypass times.
//  * ticks    every tick enqueues SSP-RK3 steps (three fused stage launches
//             each, the kernels of dgk_step) for every admitted request on
//             that request's stream, streams chosen by least resident DOF:
//             one step each (step-fair), or, by default, about the same
//             device time each (time-fair: a request gets the largest step
//             cost of the tick over its own step cost, in steps, so small
//             solves are not paced by large ones).  At most two ticks are in
//             flight; the worker retires a tick's finished requests when its
//             events complete and admits the arrivals that came meanwhile, so
//             new work joins a running mix.
//  * results  a device summary (min, max, quadrature mass and norm) per
//             request, optionally the whole final field, posted with submit,
//             admission and completion times.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "solver.hpp"

using namespace dgk;

namespace {

constexpr int kMaxBypass = 8;
constexpr int kMaxInflight = 2;
constexpr int kMaxStreams = 8;
constexpr int64_t kGranule = 1 << 16;  // arena block granularity
constexpr int64_t kAlign = 256;        // sub-array alignment inside a block
constexpr int kSlotDoubles = 4 + kGeoStride;  // pinned: summary + geometry record
// Time-fair ticks: a step's cost is its DOF, at least kCostFloor (the launch
// latency of three stage kernels is worth about that many DOF); the tick's
// quantum is the largest cost among the admitted requests, at least
// kMinQuantum.
constexpr int64_t kCostFloor = 1 << 18;
constexpr int64_t kMinQuantum = 1 << 22;

int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// First-fit allocator over one device allocation; free blocks coalesce.
class Arena {
 public:
  char *base = nullptr;
  int64_t size = 0, used = 0, peak = 0;

  void reset(char *b, int64_t n) {
    base = b;
    size = n;
    used = peak = 0;
    free_.clear();
    if (n > 0) free_[0] = n;
  }
  bool alloc(int64_t n, int64_t *off) {
    n = round_up(n, kGranule);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const int64_t rest = it->second - n, at = it->first + n;
      free_.erase(it);
      if (rest > 0) free_[at] = rest;
      used += n;
      peak = std::max(peak, used);
      return true;
    }
    return false;
  }
  void release(int64_t off, int64_t n) {
    n = round_up(n, kGranule);
    used -= n;
    auto next = free_.lower_bound(off);
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        off = prev->first;
        n += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && off + n == next->first) {
      n += next->second;
      free_.erase(next);
    }
    free_[off] = n;
  }

 private:
  std::map<int64_t, int64_t> free_;  // offset -> size
};

// Byte offsets of one request's arrays inside its arena block.
struct Layout {
  int64_t q[3], u, v, w, u32 = -1, v32 = -1, w32 = -1, nbr, geo, sum, total;
};

Layout layout(int64_t Ne, int Np, bool mp) {
  Layout L;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t o = off;
    off = round_up(off + bytes, kAlign);
    return o;
  };
  const int64_t dof = Ne * Np;
  // q in three rotation buffers; the velocity is one constant per request,
  // read by the stage kernels from the geometry record, so no velocity fields
  (void)mp;
  // (each with the tail pad the staged reads of the FP64 kernel may touch)
  for (int b = 0; b < 3; ++b) L.q[b] = take((dof + kStagePad) * 8);
  L.u = L.v = L.w = -1;
  L.nbr = take(Ne * 6 * 4);
  L.geo = take(kGeoStride * 8);
  L.sum = take((4 * kSummaryBlocks + 4) * 8);
  L.total = round_up(off, kGranule);
  return L;
}

// Operators of one order, shared by every request of that order.
struct OrderOps {
  bool ready = false;
  bool separable = false;  // Lift within kSeparableMP of the 1D form
  bool use_ell = false;    // within kSeparableF64: the FP64 kernel applies it
  double gap = 0.0;        // x1D[1] - x1D[0]
  double *D = nullptr, *Lift = nullptr, *ell = nullptr, *x1D = nullptr, *w3 = nullptr;
  float *D32 = nullptr, *ell32 = nullptr;
};

// LGL quadrature weights on the nodes x1D (degree p = size - 1):
// w_i = 2 / (p (p+1) P_p(x_i)^2).
std::vector<double> lgl_weights(const std::vector<double> &x) {
  const int p = (int)x.size() - 1;
  std::vector<double> w(x.size());
  for (size_t i = 0; i < x.size(); ++i) {
    double pm = 1.0, pc = x[i];
    for (int n = 1; n < p; ++n) {
      const double pn = ((2.0 * n + 1.0) * x[i] * pc - n * pm) / (n + 1.0);
      pm = pc;
      pc = pn;
    }
    w[i] = 2.0 / (p * (p + 1.0) * pc * pc);
  }
  return w;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct dgk_serve_req {
  int64_t id = 0;
  dgk_request r{};
  int evaltype = 0, Nq = 0, Np = 0;
  int64_t Ne = 0, dof = 0, bytes = 0;
  double dt = 0.0;
  double t_submit = 0.0, t_admit = 0.0;
  int bypassed = 0;
  int64_t off = -1;
  int slot = -1, stream = 0;
  int steps_left = 0;
  double *sum_dev = nullptr;  // summary partials, result at [4*kSummaryBlocks]
  dgk_solver s;               // views into the arena block and shared operators
};

struct dgk_server {
  int device = 0;
  std::string opdir;
  std::chrono::steady_clock::time_point t0;
  Arena arena;
  int max_active = 0;
  std::atomic<int> policy{DGK_SERVE_TIME_FAIR};
  std::atomic<int> max_steps{64};
  std::vector<hipStream_t> streams;
  hipStream_t copy_stream = nullptr;
  double *pinned = nullptr;  // max_active slots of kSlotDoubles
  OrderOps ops[kMaxNq];
  std::mutex ops_mu;

  std::mutex mu;  // guards everything below up to `worker`
  std::condition_variable cv_in, cv_out;
  std::deque<dgk_serve_req *> incoming;
  std::deque<dgk_result> results;
  bool stop = false, failed = false;
  std::string fail_msg;
  int64_t next_id = 1;
  dgk_serve_stats st{};

  std::thread worker;
  // worker-only state
  std::deque<dgk_serve_req *> waiting;
  std::vector<dgk_serve_req *> active;
  std::vector<int> free_slots;
  std::vector<int64_t> stream_dof;
  struct Tick {
    std::vector<hipEvent_t> ev;
    std::vector<dgk_serve_req *> finishing;
  };
  std::deque<Tick> inflight;
  std::vector<hipEvent_t> event_pool;
};

namespace {

using Req = dgk_serve_req;

int load_order(dgk_server *sv, int p) {
  std::lock_guard<std::mutex> lk(sv->ops_mu);
  OrderOps &op = sv->ops[p];
  if (op.ready) return DGK_OK;
  std::vector<double> x, D, L, ell;
  DGK_TRY(load_operators(sv->opdir.c_str(), p, x, D, L));
  const int Nq = p + 1;
  const double dev = separable_lift(Nq, L, ell);
  const std::vector<double> w = lgl_weights(x);
  std::vector<double> w3((size_t)Nq * Nq * Nq);
  for (int c = 0; c < Nq; ++c)
    for (int b = 0; b < Nq; ++b)
      for (int a = 0; a < Nq; ++a) w3[a + b * Nq + c * Nq * Nq] = w[a] * w[b] * w[c];
  std::vector<float> D32(D.begin(), D.end()), ell32(ell.begin(), ell.end());
  auto put = [](auto **dst, const auto &src) -> int {
    using T = typename std::remove_reference<decltype(src)>::type::value_type;
    DGK_HIP(hipMalloc((void **)dst, sizeof(T) * src.size()));
    DGK_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return DGK_OK;
  };
  DGK_TRY(put(&op.D, D));
  DGK_TRY(put(&op.Lift, L));
  DGK_TRY(put(&op.ell, ell));
  DGK_TRY(put(&op.x1D, x));
  DGK_TRY(put(&op.w3, w3));
  DGK_TRY(put(&op.D32, D32));
  DGK_TRY(put(&op.ell32, ell32));
  DGK_TRY(prepare_stage_f64(Nq));
  DGK_TRY(prepare_stage_mp(Nq));
  op.separable = dev <= kSeparableMP;
  op.use_ell = dev <= kSeparableF64;
  op.gap = x[1] - x[0];
  op.ready = true;
  return DGK_OK;
}

hipEvent_t take_event(dgk_server *sv) {
  if (!sv->event_pool.empty()) {
    hipEvent_t e = sv->event_pool.back();
    sv->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

// Bind request r to arena block `off` and enqueue its set-up on its stream:
// geometry record, device-built neighbour table, initial condition.
int start(dgk_server *sv, Req *r, int64_t off) {
  r->off = off;
  r->slot = sv->free_slots.back();
  sv->free_slots.pop_back();
  int k = 0;
  for (int i = 1; i < (int)sv->streams.size(); ++i)
    if (sv->stream_dof[i] < sv->stream_dof[k]) k = i;
  r->stream = k;
  sv->stream_dof[k] += r->dof;
  hipStream_t stream = sv->streams[k];

  const bool mp = r->evaltype == DGK_EVAL_HIP_MP;
  const Layout L = layout(r->Ne, r->Np, mp);
  char *blk = sv->arena.base + off;
  const OrderOps &op = sv->ops[r->r.p];
  dgk_solver &s = r->s;
  s.evaltype = r->evaltype;
  s.device = sv->device;
  s.Nq = r->Nq; s.Np = r->Np; s.Nfp = r->Nq * r->Nq; s.NfpTot = 6 * s.Nfp;
  s.Ne = s.NeA = r->Ne;
  s.NhaloNode = 0;
  s.nfield = r->dof;
  for (int b = 0; b < 3; ++b) s.q[b] = (double *)(blk + L.q[b]);
  s.u = s.v = s.w = nullptr;
  s.u32 = s.v32 = s.w32 = nullptr;
  s.cvel = 1;  // the stage kernels read the velocity from the geometry record
  s.nbr = (int32_t *)(blk + L.nbr);
  s.geo = (double *)(blk + L.geo);
  r->sum_dev = (double *)(blk + L.sum);
  s.D = op.D; s.Lift = op.Lift; s.ell = op.use_ell ? op.ell : nullptr;
  s.D32 = op.D32; s.ell32 = op.ell32; s.x1D = op.x1D;
  s.axis_aligned = 1;
  // the tile shape is fixed at admission, so a later dgk_tile_select from
  // another thread never changes a request in flight
  s.tile = mp ? stage_mp_selected_cvel(r->Nq) : stage_f64_selected_cvel(r->Nq);
  s.alternate = alternate_launches();
  s.launches = 0;
  // the XCD-interleaved order of plane-numbered meshes (box_nbr numbers
  // elements plane by plane, x fastest)
  {
    const char *env = std::getenv("DGK_ORDER");
    const bool keep = env && std::string(env) == "identity";
    s.zperiod = keep ? 0 : xcd_period((int64_t)r->r.ne[0] * r->r.ne[1]);
  }
  s.cur = 0;

  const int *ne = r->r.ne;
  const double h[3] = {r->r.dom[0] / ne[0], r->r.dom[1] / ne[1], r->r.dom[2] / ne[2]};
  double *slot = sv->pinned + (int64_t)r->slot * kSlotDoubles;
  fill_box_geo(h[0], h[1], h[2], slot + 4);
  for (int d = 0; d < 3; ++d) slot[4 + kGeoVel + d] = r->r.vel[d];
  DGK_HIP(hipMemcpyAsync(s.geo, slot + 4, sizeof(double) * kGeoStride, hipMemcpyHostToDevice,
                         stream));
  DGK_TRY(launch_box_nbr(s.nbr, ne[0], ne[1], ne[2], 0, stream));
  DGK_TRY(launch_init_gaussian(s.q[0], nullptr, nullptr, nullptr, op.x1D, r->Nq, ne[0], ne[1],
                               ne[2], h, r->r.centre, r->r.sigma, r->r.vel, stream));
  r->steps_left = r->r.nstep;
  r->t_admit = seconds_since(sv->t0);
  sv->active.push_back(r);
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->st.admitted++;
  sv->st.queued--;
  sv->st.active++;
  sv->st.active_peak = std::max(sv->st.active_peak, sv->st.active);
  sv->st.bytes_in_use = sv->arena.used;
  sv->st.bytes_peak = sv->arena.peak;
  return DGK_OK;
}

// Admission in arrival order; a request that does not fit may be passed by
// later ones that do, at most kMaxBypass times.
int admit(dgk_server *sv) {
  std::vector<Req *> blocked;
  for (auto it = sv->waiting.begin(); it != sv->waiting.end();) {
    if (sv->free_slots.empty()) break;
    Req *r = *it;
    int64_t off = 0;
    if (!sv->arena.alloc(r->bytes, &off)) {
      if (r->bypassed >= kMaxBypass) break;
      blocked.push_back(r);
      ++it;
      continue;
    }
    for (Req *b : blocked) b->bypassed++;
    if (!blocked.empty()) {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->st.bypasses++;
    }
    it = sv->waiting.erase(it);
    DGK_TRY(start(sv, r, off));
    if (std::any_of(blocked.begin(), blocked.end(),
                    [](const Req *b) { return b->bypassed >= kMaxBypass; }))
      break;
  }
  return DGK_OK;
}

// Summary of the final field, copied to the request's pinned slot.
int finish(dgk_server *sv, Req *r, hipStream_t stream) {
  const OrderOps &op = sv->ops[r->r.p];
  const int *ne = r->r.ne;
  const double jac = (r->r.dom[0] / ne[0]) * (r->r.dom[1] / ne[1]) * (r->r.dom[2] / ne[2]) / 8.0;
  double *res = r->sum_dev + 4 * kSummaryBlocks;
  DGK_TRY(launch_field_summary(r->s.q[0], op.w3, r->Np, r->dof, jac, r->sum_dev, res, stream));
  double *slot = sv->pinned + (int64_t)r->slot * kSlotDoubles;
  DGK_HIP(hipMemcpyAsync(slot, res, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
  return DGK_OK;
}

// SSP-RK3 steps of every admitted request (buffer rotation of dgk_step:
// stages read A, B, C and write B, C, A, with A the RK base), as many per
// request as the tick policy gives it.
int enqueue_tick(dgk_server *sv) {
  if (sv->active.empty()) return DGK_OK;
  const bool time_fair = sv->policy.load() == DGK_SERVE_TIME_FAIR;
  const int max_steps = std::max(1, sv->max_steps.load());
  int64_t quantum = kMinQuantum;
  for (const Req *r : sv->active) quantum = std::max(quantum, std::max(r->dof, kCostFloor));
  dgk_server::Tick t;
  t.ev.assign(sv->streams.size(), nullptr);
  std::vector<char> used(sv->streams.size(), 0);
  std::vector<Req *> still;
  still.reserve(sv->active.size());
  for (Req *r : sv->active) {
    hipStream_t stream = sv->streams[r->stream];
    used[r->stream] = 1;
    int n = 1;
    if (time_fair)
      n = (int)std::min<int64_t>(max_steps, quantum / std::max(r->dof, kCostFloor));
    n = std::max(1, std::min(n, r->steps_left));
    dgk_solver &s = r->s;
    double *A = s.q[0], *B = s.q[1], *C = s.q[2];
    const double *in[3] = {A, B, C};
    double *out[3] = {B, C, A};
    for (int k = 0; k < n && r->steps_left > 0; ++k) {
      for (int st = 0; st < 3; ++st)
        DGK_TRY(launch_stage(&s, in[st], A, out[st], 1, kRkA[st], kRkB[st], r->dt, stream, 0));
      r->steps_left--;
    }
    if (r->steps_left == 0) {
      DGK_TRY(finish(sv, r, stream));
      t.finishing.push_back(r);
    } else {
      still.push_back(r);
    }
  }
  sv->active.swap(still);
  for (size_t k = 0; k < sv->streams.size(); ++k) {
    if (!used[k]) continue;
    t.ev[k] = take_event(sv);
    if (!t.ev[k]) return set_error(DGK_ERR_DEVICE, "hipEventCreate failed");
    DGK_HIP(hipEventRecord(t.ev[k], sv->streams[k]));
  }
  sv->inflight.push_back(std::move(t));
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->st.ticks++;
  return DGK_OK;
}

int retire(dgk_server *sv, Req *r) {
  dgk_result res{};
  res.id = r->id;
  res.tag = r->r.tag;
  res.status = DGK_OK;
  res.evaltype = r->evaltype;
  res.p = r->r.p;
  res.nstep = r->r.nstep;
  res.dof = r->dof;
  res.bytes = r->bytes;
  res.dt = r->dt;
  res.t_submit = r->t_submit;
  res.t_admit = r->t_admit;
  const double *slot = sv->pinned + (int64_t)r->slot * kSlotDoubles;
  res.qmin = slot[0];
  res.qmax = slot[1];
  res.mass = slot[2];
  res.l2sq = slot[3];
  if (r->r.q_out) {
    DGK_HIP(hipMemcpyAsync(r->r.q_out, r->s.q[0], sizeof(double) * r->dof,
                           hipMemcpyDeviceToHost, sv->copy_stream));
    DGK_HIP(hipStreamSynchronize(sv->copy_stream));
  }
  res.t_done = seconds_since(sv->t0);
  sv->arena.release(r->off, r->bytes);
  sv->free_slots.push_back(r->slot);
  sv->stream_dof[r->stream] -= r->dof;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->results.push_back(res);
    sv->st.completed++;
    sv->st.active--;
    sv->st.dof_updates += (double)r->dof * r->r.nstep;
    sv->st.bytes_in_use = sv->arena.used;
  }
  sv->cv_out.notify_all();
  delete r;
  return DGK_OK;
}

// Wait for the oldest tick in flight when two are queued (or when nothing is
// left to enqueue) and retire its finished requests.
int throttle(dgk_server *sv) {
  if (sv->inflight.empty()) return DGK_OK;
  if ((int)sv->inflight.size() < kMaxInflight && !sv->active.empty()) return DGK_OK;
  dgk_server::Tick t = std::move(sv->inflight.front());
  sv->inflight.pop_front();
  for (hipEvent_t e : t.ev) {
    if (!e) continue;
    DGK_HIP(hipEventSynchronize(e));
    sv->event_pool.push_back(e);
  }
  for (size_t k = 0; k < t.finishing.size(); ++k) {
    const int rc = retire(sv, t.finishing[k]);
    if (rc != DGK_OK) {
      // the failed request and those after it have no result yet: leave them
      // where fail_all posts one for each
      sv->active.insert(sv->active.end(), t.finishing.begin() + (std::ptrdiff_t)k,
                        t.finishing.end());
      return rc;
    }
  }
  return DGK_OK;
}

void post_failure(dgk_server *sv, Req *r, int rc, const std::string &msg) {
  dgk_result res{};
  res.id = r->id;
  res.tag = r->r.tag;
  res.status = rc;
  res.evaltype = r->evaltype;
  res.p = r->r.p;
  res.nstep = r->r.nstep;
  res.dof = r->dof;
  res.bytes = r->bytes;
  res.t_submit = r->t_submit;
  res.t_admit = r->t_admit;
  res.t_done = seconds_since(sv->t0);
  std::snprintf(res.message, sizeof(res.message), "%s", msg.c_str());
  sv->results.push_back(res);
  sv->st.failed++;
}

// A device error ends the worker: every request not yet posted gets a failed
// result and later submissions are refused.
void fail_all(dgk_server *sv, int rc) {
  const std::string msg = std::string("serving worker stopped: ") + dgk_last_error();
  std::vector<Req *> all(sv->waiting.begin(), sv->waiting.end());
  all.insert(all.end(), sv->active.begin(), sv->active.end());
  for (auto &t : sv->inflight) all.insert(all.end(), t.finishing.begin(), t.finishing.end());
  sv->waiting.clear();
  sv->active.clear();
  sv->inflight.clear();
  std::lock_guard<std::mutex> lk(sv->mu);
  all.insert(all.end(), sv->incoming.begin(), sv->incoming.end());
  sv->incoming.clear();
  sv->failed = true;
  sv->fail_msg = msg;
  for (Req *r : all) {
    post_failure(sv, r, rc, msg);
    delete r;
  }
  sv->cv_out.notify_all();
}

void worker_main(dgk_server *sv) {
  if (hipSetDevice(sv->device) != hipSuccess) {
    set_error(DGK_ERR_DEVICE, "hipSetDevice failed in the serving worker");
    fail_all(sv, DGK_ERR_DEVICE);
    return;
  }
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      for (;;) {
        while (!sv->incoming.empty()) {
          sv->waiting.push_back(sv->incoming.front());
          sv->incoming.pop_front();
        }
        if (!sv->waiting.empty() || !sv->active.empty() || !sv->inflight.empty()) break;
        if (sv->stop) return;
        sv->cv_in.wait(lk);
      }
    }
    int rc = admit(sv);
    if (rc == DGK_OK && sv->active.empty() && sv->inflight.empty() && !sv->waiting.empty())
      rc = set_error(DGK_ERR_STATE, "admission stalled with an empty device");
    if (rc == DGK_OK) rc = enqueue_tick(sv);
    if (rc == DGK_OK) rc = throttle(sv);
    if (rc != DGK_OK) {
      fail_all(sv, rc);
      return;
    }
  }
}

int check_request(const dgk_request *r) {
  if (!r) return set_error(DGK_ERR_ARG, "request is NULL");
  for (int d = 0; d < 3; ++d) {
    if (r->ne[d] < 1) return set_error(DGK_ERR_ARG, "elements per axis must be >= 1");
    if (!(r->dom[d] > 0) || !std::isfinite(r->dom[d]))
      return set_error(DGK_ERR_ARG, "box sizes must be positive");
    if (!std::isfinite(r->vel[d]) || !std::isfinite(r->centre[d]))
      return set_error(DGK_ERR_ARG, "velocity and centre must be finite");
  }
  if (r->p < kMinNq - 1 || r->p > kMaxNq - 1)
    return set_error(DGK_ERR_UNSUPPORTED, "polynomial order " + std::to_string(r->p) +
                                              " outside 1.." + std::to_string(kMaxNq - 1));
  if (r->nstep < 0) return set_error(DGK_ERR_ARG, "nstep must be >= 0");
  if (r->tol_class != DGK_TOL_FP64 && r->tol_class != DGK_TOL_1E5)
    return set_error(DGK_ERR_ARG, "unknown tolerance class " + std::to_string(r->tol_class));
  if (!(r->sigma > 0) || !std::isfinite(r->sigma))
    return set_error(DGK_ERR_ARG, "sigma must be positive");
  if (!std::isfinite(r->dt)) return set_error(DGK_ERR_ARG, "dt must be finite");
  const int64_t Ne = (int64_t)r->ne[0] * r->ne[1] * r->ne[2];
  const int64_t Np = (int64_t)(r->p + 1) * (r->p + 1) * (r->p + 1);
  if (Ne * Np >= (int64_t)INT32_MAX)
    return set_error(DGK_ERR_UNSUPPORTED, "request exceeds 2^31 nodes");
  return DGK_OK;
}

}  // namespace

extern "C" {

int dgk_serve_footprint(const int32_t ne[3], int p, int tol_class, int64_t *bytes) {
  clear_error();
  if (!ne || !bytes) return set_error(DGK_ERR_ARG, "NULL argument");
  if (ne[0] < 1 || ne[1] < 1 || ne[2] < 1) return set_error(DGK_ERR_ARG, "elements per axis must be >= 1");
  if (p < kMinNq - 1 || p > kMaxNq - 1) return set_error(DGK_ERR_UNSUPPORTED, "order outside 1.." + std::to_string(kMaxNq - 1));
  const int Np = (p + 1) * (p + 1) * (p + 1);
  *bytes = layout((int64_t)ne[0] * ne[1] * ne[2], Np, tol_class == DGK_TOL_1E5).total;
  return DGK_OK;
}

int dgk_serve_create(const char *operator_dir, int64_t hbm_budget, int max_active,
                     int nstreams, dgk_server **out) {
  clear_error();
  if (!out) return set_error(DGK_ERR_ARG, "output handle pointer is NULL");
  *out = nullptr;
  if (!operator_dir) return set_error(DGK_ERR_ARG, "operator directory is NULL");
  if (hbm_budget < kGranule) return set_error(DGK_ERR_ARG, "HBM budget below 64 KiB");
  if (max_active < 1) return set_error(DGK_ERR_ARG, "max_active must be >= 1");
  if (nstreams < 1 || nstreams > kMaxStreams) return set_error(DGK_ERR_ARG, "nstreams must be 1..8");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return set_error(DGK_ERR_DEVICE, "no HIP device available (the path has no CPU fallback)");
  }
  auto *sv = new dgk_server();
  auto fail = [&](int rc) {
    for (hipStream_t s : sv->streams) (void)hipStreamDestroy(s);
    if (sv->copy_stream) (void)hipStreamDestroy(sv->copy_stream);
    if (sv->arena.base) (void)hipFree(sv->arena.base);
    if (sv->pinned) (void)hipHostFree(sv->pinned);
    delete sv;
    return rc;
  };
  if (hipGetDevice(&sv->device) != hipSuccess) return fail(set_error(DGK_ERR_DEVICE, "hipGetDevice failed"));
  sv->opdir = operator_dir;
  sv->max_active = max_active;
  const int64_t budget = hbm_budget / kGranule * kGranule;
  void *mem = nullptr;
  if (hipMalloc(&mem, (size_t)budget) != hipSuccess) {
    (void)hipGetLastError();
    return fail(set_error(DGK_ERR_NOMEM, "arena of " + std::to_string(budget) +
                                             " bytes does not fit in device memory"));
  }
  sv->arena.reset((char *)mem, budget);
  void *pin = nullptr;
  if (hipHostMalloc(&pin, sizeof(double) * kSlotDoubles * (size_t)max_active, 0) != hipSuccess) {
    (void)hipGetLastError();
    return fail(set_error(DGK_ERR_NOMEM, "pinned host slots"));
  }
  sv->pinned = (double *)pin;
  for (int i = max_active - 1; i >= 0; --i) sv->free_slots.push_back(i);
  for (int k = 0; k < nstreams; ++k) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
      return fail(set_error(DGK_ERR_DEVICE, "hipStreamCreate failed"));
    sv->streams.push_back(s);
  }
  sv->stream_dof.assign(nstreams, 0);
  if (hipStreamCreateWithFlags(&sv->copy_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_error(DGK_ERR_DEVICE, "hipStreamCreate failed"));
  sv->st.budget = budget;
  sv->t0 = std::chrono::steady_clock::now();
  sv->worker = std::thread(worker_main, sv);
  *out = sv;
  return DGK_OK;
}

int dgk_serve_submit(dgk_server *sv, const dgk_request *r, int64_t *id) {
  clear_error();
  if (!sv) return set_error(DGK_ERR_ARG, "server handle is NULL");
  DGK_TRY(check_request(r));
  DGK_HIP(hipSetDevice(sv->device));
  DGK_TRY(load_order(sv, r->p));
  const OrderOps &op = sv->ops[r->p];
  auto *q = new Req();
  q->r = *r;
  q->evaltype = r->tol_class == DGK_TOL_1E5 ? DGK_EVAL_HIP_MP : DGK_EVAL_HIP;
  if ((q->evaltype == DGK_EVAL_HIP_MP && !op.separable) ||
      (q->evaltype == DGK_EVAL_HIP && !op.use_ell)) {
    delete q;
    return set_error(DGK_ERR_UNSUPPORTED, "the Lift of this order is not of tensor-product "
                                          "form; served solves need it");
  }
  q->Nq = r->p + 1;
  q->Np = q->Nq * q->Nq * q->Nq;
  q->Ne = (int64_t)r->ne[0] * r->ne[1] * r->ne[2];
  q->dof = q->Ne * q->Np;
  q->bytes = layout(q->Ne, q->Np, q->evaltype == DGK_EVAL_HIP_MP).total;
  if (r->dt > 0) {
    q->dt = r->dt;
  } else {
    const double speed = std::fabs(r->vel[0]) + std::fabs(r->vel[1]) + std::fabs(r->vel[2]);
    if (!(speed > 0)) {
      delete q;
      return set_error(DGK_ERR_ARG, "dt <= 0 asks for the CFL step, which needs a nonzero velocity");
    }
    double hmin = INFINITY;
    for (int d = 0; d < 3; ++d) hmin = std::fmin(hmin, op.gap * (r->dom[d] / r->ne[d]) / 2.0);
    q->dt = 0.1 * hmin / speed;
  }
  if (q->bytes > sv->arena.size) {
    const int64_t need = q->bytes;
    delete q;
    return set_error(DGK_ERR_NOMEM, "request needs " + std::to_string(need) +
                                        " device bytes; the server's arena holds " +
                                        std::to_string(sv->arena.size));
  }
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->failed || sv->stop) {
      delete q;
      return set_error(DGK_ERR_STATE, sv->failed ? sv->fail_msg : "server is shutting down");
    }
    q->id = sv->next_id++;
    q->t_submit = seconds_since(sv->t0);
    sv->incoming.push_back(q);
    sv->st.submitted++;
    sv->st.queued++;
    if (id) *id = q->id;
  }
  sv->cv_in.notify_one();
  return DGK_OK;
}

int dgk_serve_poll(dgk_server *sv, dgk_result *out, int max, int *n, double timeout_s) {
  clear_error();
  if (!sv || !n || (max > 0 && !out)) return set_error(DGK_ERR_ARG, "NULL argument");
  *n = 0;
  std::unique_lock<std::mutex> lk(sv->mu);
  if (sv->results.empty() && timeout_s > 0)
    sv->cv_out.wait_for(lk, std::chrono::duration<double>(timeout_s),
                        [&] { return !sv->results.empty(); });
  while (*n < max && !sv->results.empty()) {
    out[(*n)++] = sv->results.front();
    sv->results.pop_front();
  }
  return DGK_OK;
}

int dgk_serve_set_policy(dgk_server *sv, int policy, int max_steps_per_tick) {
  clear_error();
  if (!sv) return set_error(DGK_ERR_ARG, "server handle is NULL");
  if (policy != DGK_SERVE_STEP_FAIR && policy != DGK_SERVE_TIME_FAIR)
    return set_error(DGK_ERR_ARG, "unknown tick policy " + std::to_string(policy));
  if (max_steps_per_tick < 1 || max_steps_per_tick > 1024)
    return set_error(DGK_ERR_ARG, "max_steps_per_tick must be 1..1024");
  sv->max_steps.store(max_steps_per_tick);
  sv->policy.store(policy);
  return DGK_OK;
}

int dgk_serve_get_stats(dgk_server *sv, dgk_serve_stats *st) {
  clear_error();
  if (!sv || !st) return set_error(DGK_ERR_ARG, "NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  *st = sv->st;
  return DGK_OK;
}

int dgk_serve_destroy(dgk_server *sv) {
  clear_error();
  if (!sv) return DGK_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_in.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  (void)hipSetDevice(sv->device);
  for (hipStream_t s : sv->streams) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  if (sv->copy_stream) (void)hipStreamDestroy(sv->copy_stream);
  for (hipEvent_t e : sv->event_pool) (void)hipEventDestroy(e);
  for (auto &t : sv->inflight)
    for (hipEvent_t e : t.ev)
      if (e) (void)hipEventDestroy(e);
  for (OrderOps &op : sv->ops) {
    for (double *p : {op.D, op.Lift, op.ell, op.x1D, op.w3})
      if (p) (void)hipFree(p);
    for (float *p : {op.D32, op.ell32})
      if (p) (void)hipFree(p);
  }
  if (sv->arena.base) (void)hipFree(sv->arena.base);
  if (sv->pinned) (void)hipHostFree(sv->pinned);
  delete sv;
  return DGK_OK;
}

}  // extern "C"
