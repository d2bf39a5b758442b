// The device solver object shared by the C entry points, the batch layer and
// the serving loop of libdgk.
#pragma once

#include <vector>

#include "dgk_internal.hpp"

namespace dgk {

// SSP-RK3 Shu-Osher coefficients (mod_common.f90:28-30).
constexpr double kRkA[3] = {0.0, 0.75, 1.0 / 3.0};
constexpr double kRkB[3] = {1.0, 0.25, 2.0 / 3.0};

// Timer slots (names of main.f90:175-177 and mod_advect3d_eq.f90:34-35).
enum { T_MAIN = 0, T_CALTEND = 1, T_FLUX = 2, T_DQDT = 3, T_STAGE = 4, T_N = 5 };

}  // namespace dgk

using dgk::BoxShape;

struct dgk_solver {
  // (all device pointers are owned: freed by dgk_destroy)
  int evaltype = 0;
  int device = 0;
  int Nq = 0, Np = 0, Nfp = 0, NfpTot = 0;
  int64_t Ne = 0, NeA = 0, NhaloNode = 0;
  int64_t nfield = 0;  // Np*Ne + NhaloNode: owned part + halo tail
  hipStream_t stream = nullptr;
  double *q[3] = {nullptr, nullptr, nullptr};
  int cur = 0;
  double *u = nullptr, *v = nullptr, *w = nullptr;
  double *q0 = nullptr, *dqdt = nullptr, *flux = nullptr;
  bool flux_valid = false, dqdt_valid = false, q0_valid = false;
  float *u32 = nullptr, *v32 = nullptr, *w32 = nullptr;
  float *D32 = nullptr, *ell32 = nullptr;
  double *ell = nullptr;     // separable Lift (nullptr: stored-Lift kernel)
  int32_t *order = nullptr;  // element processing order
  int alternate = 0;         // reverse mode of every other stage launch (slot_element)
  int64_t zperiod = 0;       // period of the XCD-interleaved order (slot_element)
  uint64_t launches = 0;     // stage launches so far (picks the direction)
  int axis_aligned = 0;
  int tile = -1;             // tile-shape candidate chosen for this mesh (-1: process default)
  int geo_per_run = 1;       // every run of run_elements(Nq) elements shares a geometry record
  int cvel = 0;              // stage launches read each element's constant velocity from its
                             // geometry record (batch groups whose boxes all have one)
  double *x1D = nullptr;     // LGL nodes (device), for device-side initial conditions
  BoxShape box;              // structured box (NeX == 0 when built from arrays)
  int32_t *nbr = nullptr, *vmapP = nullptr, *halo = nullptr;
  double *geo = nullptr, *D = nullptr, *Lift = nullptr;
  uint32_t *geoid = nullptr;
  double *mm = nullptr;  // min/max partials + result
  bool set[4] = {false, false, false, false};
  int64_t bytes = 0;
  std::vector<void *> allocs;
  // timers
  bool kernel_timing = false;
  double sec[dgk::T_N] = {0, 0, 0, 0, 0};
  int64_t cnt[dgk::T_N] = {0, 0, 0, 0, 0};
  struct Pending { int slot; hipEvent_t a, b; int64_t count; };  // count: launches spanned
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
};

namespace dgk {

template <typename T>
int dalloc(dgk_solver *s, T **p, int64_t n) {
  *p = nullptr;
  if (n <= 0) return DGK_OK;
  void *ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, sizeof(T) * (size_t)n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(DGK_ERR_NOMEM, "hipMalloc of " + std::to_string(sizeof(T) * n) +
                                        " bytes failed: " + hipGetErrorString(e));
  }
  s->allocs.push_back(ptr);
  s->bytes += (int64_t)sizeof(T) * n;
  *p = (T *)ptr;
  return DGK_OK;
}

template <typename T>
int upload(dgk_solver *s, T *dst, const std::vector<T> &src) {
  if (src.empty()) return DGK_OK;
  DGK_HIP(hipMemcpyAsync(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice,
                         s->stream));
  return DGK_OK;
}

double *field_ptr(dgk_solver *s, int which);
int refresh_f32(dgk_solver *s, int which);
// One stage launch on `stream` (the solver's own stream when nullptr);
// dt_geo takes the time step from each element's geometry record.
// Consecutive stage launches alternate direction unless DGK_ALT=0: the
// reverse mode of every other launch (slot_element).
int alternate_launches();
int launch_stage(dgk_solver *s, const double *qin, const double *q0, double *out,
                 int mode, double ca, double cb, double dt, hipStream_t stream = nullptr,
                 int dt_geo = 0);
int create_from_tables(int evaltype, MeshTables &t, dgk_solver **out);

// Separable 1D form of a stored Lift (Nq^3*6): ell = l-(0..Nq-1), l+(0..Nq-1);
// returns the largest deviation of a stored entry from that form relative to
// max|Lift|.
double separable_lift(int Nq, const std::vector<double> &Lift, std::vector<double> &ell);
// Lift deviations from the tensor-product form below this (relative) are
// rounding in the stored data (2.8e-14 for p = 1..15); the FP64 path then
// applies the 1D form, otherwise it reads Lift as stored.
constexpr double kSeparableF64 = 1e-12;
constexpr double kSeparableMP = 1e-10;

}  // namespace dgk
