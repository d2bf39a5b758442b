// Internal declarations shared by the host side and the HIP kernels of libdgk.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "dgk.h"

namespace dgk {

// ---------------------------------------------------------------------------
// Error reporting: every C entry point returns a status and leaves a message
// in a thread-local buffer (dgk_last_error).
// ---------------------------------------------------------------------------
int set_error(int code, const std::string &msg);
void clear_error();

#define DGK_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::dgk::set_error(DGK_ERR_DEVICE, std::string(#call) + ": " +      \
                                                  hipGetErrorString(e_));      \
  } while (0)

#define DGK_TRY(call)                                                          \
  do {                                                                         \
    int rc_ = (call);                                                          \
    if (rc_ != DGK_OK) return rc_;                                             \
  } while (0)

constexpr int kMinNq = 2;
constexpr int kMaxNq = 20;      // p <= 19: operators above p = 15 are generated
constexpr int kMaxNqFile = 16;  // operator_data/p{p}.dat exist for p <= 15
constexpr int kMaxNqValu = 16;  // the VALU FP64 kernels; above, MFMA kernels only
// Doubles of readable memory after the last node of every field the FP64 VALU
// stage kernel stages to LDS (q, u, v, w): with an odd node count per element
// its 16-byte copy granules end 8 bytes past an even-numbered element.
constexpr int kStagePad = 2;

// Per-element geometry record: Escale[3] (2/dx_d, constant per element,
// mod_mesh.f90:147-149), Fscale[6] and the outward normal of each face
// (set_face_geometry, mod_mesh.f90:356-397).  Identical records are shared.
constexpr int kGeoStride = 32;  // doubles per record (31 used)
constexpr int kGeoEscale = 0;
constexpr int kGeoFscale = 3;
constexpr int kGeoNormal = 9;  // + 3*f + d
constexpr int kGeoDt = 27;     // time step of the element's instance (batched solves)
constexpr int kGeoVel = 28;    // constant velocity u, v, w of the element's box (28..30),
                               // read instead of the fields by launches with cvel set

// Node index (0-based) inside an element of face node (a,b) of face f on the
// element's own side, in the reference's Fmask order (mod_mesh.f90:278-324):
// faces 0..5 = -y, +x, +y, -x, -z, +z.
__host__ __device__ inline int own_face_node(int f, int a, int b, int Nq) {
  const int Nq2 = Nq * Nq;
  switch (f) {
  case 0: return a + b * Nq2;                       // (a, 0, b)
  case 1: return (Nq - 1) + a * Nq + b * Nq2;       // (Nq-1, a, b)
  case 2: return a + (Nq - 1) * Nq + b * Nq2;       // (a, Nq-1, b)
  case 3: return a * Nq + b * Nq2;                  // (0, a, b)
  case 4: return a + b * Nq;                        // (a, b, 0)
  default: return a + b * Nq + (Nq - 1) * Nq2;      // (a, b, Nq-1)
  }
}
// Opposite face (mod_mesh.f90:420-480): 0<->2, 1<->3, 4<->5.
__host__ __device__ inline int opp_face(int f) {
  return f < 4 ? (f + 2) % 4 : (f == 4 ? 5 : 4);
}
// Node of the neighbour that face node (a,b) of face f meets: the same (a,b)
// on the neighbour's opposite face (interior VMapP, mod_mesh.f90:163).
__host__ __device__ inline int nbr_face_node(int f, int a, int b, int Nq) {
  return own_face_node(opp_face(f), a, b, Nq);
}

#if defined(__HIPCC__)
// Neighbour lookups of face node (i,j) of element e on its six faces: raw
// entries of the resolved per-face-node table when there is one, else of the
// element neighbour table.  All six entries are loaded before any is used, so
// their latencies overlap; nbr_node turns an entry into a flat node index.
template <int NQ>
__device__ __forceinline__ void nbr_entries(const int32_t *__restrict__ vmapP,
                                            const int32_t *__restrict__ nbr, int64_t e, int ij,
                                            int32_t (&t)[6]) {
  constexpr int NQ2 = NQ * NQ, NFT = 6 * NQ2;
  if (vmapP) {
#pragma unroll
    for (int f = 0; f < 6; ++f) t[f] = vmapP[e * NFT + f * NQ2 + ij];
  } else {
#pragma unroll
    for (int f = 0; f < 6; ++f) t[f] = nbr[e * 6 + f];
  }
}
// Element of launch slot `slot` (< Ne): the processing order, walked
// forwards (reverse = 0) or backwards (1: the whole order; 2: with a period,
// each XCD's own slot range, so that an XCD walks the same elements in both
// directions).  Without an order array, zperiod = Q > 0 (8 <= Q, Ne a
// multiple of Q) deals every Q consecutive elements of the numbering to the
// 8 XCDs in eighths: with n = Ne / Q periods and o_x = floor(x Q / 8), the
// slots [n o_x, n o_{x+1}), which the XCD x runs, are the elements
// [o_x, o_{x+1}) of every period, period after period.  Q = P (planes of P
// elements) gives XCD x the eighth x of every plane; Q = 8 G deals groups of
// G elements round-robin.
__device__ __forceinline__ int64_t slot_element(int64_t slot, int64_t Ne,
                                                const int32_t *__restrict__ order, int reverse,
                                                int64_t zperiod) {
  const bool per_xcd = reverse == 2 && !order && zperiod > 0;
  if (reverse && !per_xcd) slot = Ne - 1 - slot;
  if (order) return (int64_t)order[slot];
  if (zperiod > 0) {
    const uint32_t Q = (uint32_t)zperiod, n = (uint32_t)Ne / Q, s = (uint32_t)slot;
    uint32_t x = 0;
#pragma unroll
    for (uint32_t j = 1; j < 8; ++j) x += s >= n * (uint32_t)(((uint64_t)j * Q) >> 3) ? 1u : 0u;
    const uint32_t o = (uint32_t)(((uint64_t)x * Q) >> 3);
    const uint32_t g = (uint32_t)(((uint64_t)(x + 1) * Q) >> 3) - o;
    uint32_t t = s - n * o;
    if (per_xcd) t = n * g - 1 - t;
    const uint32_t k = t / g;
    return (int64_t)k * Q + o + (t - k * g);
  }
  return slot;
}
template <int NQ>
__device__ __forceinline__ int64_t nbr_node(bool resolved, int32_t t, int f, int i, int j) {
  return resolved ? (int64_t)t : (int64_t)t * (NQ * NQ * NQ) + nbr_face_node(f, i, j, NQ);
}
#endif

// Compact device description of the mesh, built from the reference arrays.
struct MeshTables {
  int Nq = 0, Np = 0, Nfp = 0, NfpTot = 0;
  int64_t Ne = 0, NeA = 0, NhaloNode = 0;
  // Connectivity: either a neighbour table nbr[e*6+f] (0-based element) when
  // every face meets the opposite face of one element node-for-node, or the
  // resolved 0-based P index of every face node (halo folded in).
  std::vector<int32_t> nbr;
  std::vector<int32_t> vmapP;
  std::vector<int32_t> halo_src;  // 0-based, NhaloNode
  std::vector<double> geo;        // ngeo * kGeoStride
  std::vector<uint32_t> geoid;    // Ne, empty when ngeo == 1
  std::vector<double> D;          // Nq*Nq column-major
  std::vector<double> Lift;       // Nq^3*6 column-major
  std::vector<int32_t> order;     // element processing order (empty: identity)
  int64_t zperiod = 0;            // > 0: period of the XCD-interleaved order (slot_element)
  bool axis_aligned = false;      // every face normal lies along its face axis
};

// Normal axis of faces -y, +x, +y, -x, -z, +z (set_face_geometry,
// mod_mesh.f90:364-394).
constexpr int kFaceAxis[6] = {1, 0, 1, 0, 2, 2};
// True when every geometry record's face normals have zero components off
// the face's axis.
bool geo_axis_aligned(const std::vector<double> &geo);
// Geometry record of an axis-aligned hexahedron of size dx x dy x dz.
void fill_box_geo(double dx, double dy, double dz, double *rec);

// Processing order of a NeX x NeY x NeZ box in bricks of bx x by x bz
// elements (bricks x-fastest, elements x-fastest inside a brick).  Only the
// assignment of elements to workgroups changes, never the arithmetic.
std::vector<int32_t> brick_order(int NeX, int NeY, int NeZ, int bx, int by, int bz);
// Element processing order.  The stage kernels deal consecutive slots of a
// launch to one XCD (an eighth of the slots each).  For a mesh numbered
// plane by plane along z with periodic +z neighbours e -> (e + P) mod Ne
// (the reference's numbering, mod_mesh.f90:491) and P >= 8 the default order
// is slot_element's with period P (xcd_period): XCD x walks the eighth x of
// every plane in turn, so its z-neighbours come round again on the same XCD
// and the 8 XCDs stream neighbouring addresses at any time.  Only the
// assignment of elements to workgroups changes, never the arithmetic.
// DGK_ORDER=identity keeps the element numbering; DGK_ORDER=zperiod:Q takes
// the period Q instead (8 <= Q, Ne a multiple of Q; else the numbering).
int64_t z_plane_size(const MeshTables &t);  // P, or 0 when not of that form
int64_t xcd_period(int64_t P);               // the default period for planes of P, or 0
void default_order(MeshTables &t);           // sets zperiod (the kernels compute the order)

int build_tables_from_reference(int Nq, int64_t Ne, int64_t NeA,
                                int64_t NhaloNode, const double *D1D,
                                const double *Lift_mat, const int *VMapM,
                                const int *VMapP, const int *halo_src_map,
                                const double *normal_fn, const double *Escale,
                                const double *Fscale, MeshTables &t);
int build_tables_structured(int NeX, int NeY, int NeZ, int p,
                            const double *D1D, const double *Lift_mat,
                            double dom_x, double dom_y, double dom_z,
                            MeshTables &t);
int load_operators(const char *dir, int p, std::vector<double> &x1D,
                   std::vector<double> &D, std::vector<double> &Lift);
// Closed-form operators of order p (operators.cpp) and the p{p}.dat writer.
int make_operators(int p, std::vector<double> &x1D, std::vector<double> &D,
                   std::vector<double> &Lift);
int write_operators(const char *dir, int p);

// One instance of a batch: a periodic box of NeX x NeY x NeZ elements.
struct BatchInstance {
  int NeX, NeY, NeZ;
  double dom[3];
  double dt;
};

// Shape of a structured periodic box (kept by solvers built from one, for
// device-side initial conditions).
struct BoxShape {
  int NeX = 0, NeY = 0, NeZ = 0;
  double dom[3] = {0, 0, 0};
};
// Tables of several same-order instances concatenated into one element
// space: element indices (and neighbour entries) of instance k are offset by
// the elements of instances 0..k-1; each instance gets its own geometry
// record, carrying its time step at kGeoDt.  No halo (the neighbour table
// addresses the periodic partner directly).
int build_tables_batch(int p, const double *D1D, const double *Lift_mat,
                       const std::vector<BatchInstance> &inst, MeshTables &t,
                       std::vector<int64_t> &elem_offset);

// Function attributes (dynamic LDS limits) of the stage kernels of order Nq,
// set outside any stream capture.
int prepare_stage_f64(int Nq);
int prepare_stage_mp(int Nq);

// ---------------------------------------------------------------------------
// Kernel parameter blocks and launchers (kernels_f64.hip, kernels_mfma.hip).
// ---------------------------------------------------------------------------
struct StageArgs {
  const double *q;        // stage input, owned part (Np*Ne)
  const double *q0;       // RK base (may alias out); unused when ca == 0
  const double *u, *v, *w;
  double *out;            // dqdt (mode 0) or RK result (mode 1)
  const double *flux_in;  // face flux (NfpTot,Ne) to use instead of computing it
  const int32_t *nbr;     // compact connectivity or nullptr
  const int32_t *vmapP;   // node connectivity or nullptr
  const double *D;        // Nq*Nq
  const double *Lift;     // Nq^3*6 (stored-Lift kernel)
  const double *ell;      // 2*Nq separable Lift l-, l+ (nullptr: stored Lift)
  const double *geo;
  const uint32_t *geoid;  // nullptr when one record
  const int32_t *order;   // processing order of elements (nullptr: identity)
  int reverse;            // walk the order backwards
  int64_t zperiod;        // order == nullptr and > 0: period of the XCD-interleaved order
  int64_t Ne;
  double ca, cb, dt;      // RK stage: out = ca*q0 + cb*(q + dt*tend)
  int mode;               // 0 = tendency, 1 = RK stage
  int axis_aligned;       // face normals along the face axes (see MeshTables)
  int dt_geo;             // RK time step from the geometry record (per instance)
  int tile;               // tile-shape candidate, -1: the process-wide selection
  int geo_per_run = 0;    // with geoid: every run of run_elements(Nq) elements from a
                          // multiple of it shares one geometry record
  int cvel = 0;           // velocity constant per element, from its geometry record
};

// Mixed-precision operands (FP32 copies of the constant inputs).
struct StageArgsMP {
  const double *q;
  const double *q0;
  const float *u, *v, *w;
  double *out;
  const int32_t *nbr;
  const int32_t *vmapP;
  const float *D;         // Nq*Nq, FP32
  const float *ell;       // 2*Nq: l-(x) then l+(x), the separable 1D Lift
  const double *D64 = nullptr;    // the same in FP64, for the VALU kernel
  const double *ell64 = nullptr;
  const double *geo;
  const uint32_t *geoid;
  const int32_t *order;
  int reverse;
  int64_t zperiod;
  int64_t Ne;
  double ca, cb, dt;
  int mode;
  int axis_aligned;
  int dt_geo;
  int tile;               // tile-shape candidate, -1: the process-wide selection
  int geo_per_run = 0;    // as StageArgs::geo_per_run, for the mixed run kernel
  int cvel = 0;           // as StageArgs::cvel
};

int launch_stage_f64(int Nq, const StageArgs &a, hipStream_t s);
int launch_stage_mp(int Nq, const StageArgsMP &a, hipStream_t s);
// Whether tile candidate `cand` launches its own kernel for these arguments
// (false: the run-kernel candidate on launches it cannot take, which run
// another shape instead).
bool stage_f64_candidate_runs(int Nq, int cand, const StageArgs &a);
bool stage_mp_candidate_runs(int Nq, int cand, const StageArgsMP &a);
// The mixed form's VALU kernel (FP32 velocity fields, FP64 arithmetic) at
// p <= 7: elements per workgroup (0 where there is none), launch, set-up.
int valu_mp_elements(int Nq);
int launch_stage_valu_mp(int Nq, const StageArgsMP &a, hipStream_t s);
int prepare_stage_valu_mp(int Nq);
// Tile shapes (elements per block) of the mixed-precision stage kernel of
// order Nq (the MFMA shapes, then at p <= 4 the run kernel): fills epb and
// returns their number (n >= 5).  stage_mp_select picks the
// candidate that later launches of that Nq use.
int stage_mp_shapes(int Nq, int *epb, int n);
int stage_mp_select(int Nq, int cand);
int stage_mp_selected(int Nq);
int stage_mp_selected_cvel(int Nq);   // the selection for constant-velocity launches
// FP64: candidates 0, 1 are the VALU kernels, 2.. the FP64 MFMA kernel's shapes.
constexpr int kMaxTileShapes = 8;
int stage_f64_shapes(int Nq, int *epb, int n);
int stage_f64_select(int Nq, int cand);
int stage_f64_selected(int Nq);
int stage_f64_selected_cvel(int Nq);  // the selection for constant-velocity launches
// FP64 MFMA stage kernel (kernels_mfma.hip): shapes, attributes, launch of
// shape candidate `cand`.
int tile_f64_shapes(int Nq, int *epb, int n);
int prepare_tile_f64(int Nq);
int launch_tile_f64(int Nq, int cand, const StageArgs &a, hipStream_t s);
// FP64 stage kernel over runs of consecutive elements (kernels_run.hip):
// the last FP64 candidate at p <= 7.  run_elements gives the elements per
// workgroup (0: no run kernel for this Nq); launches it cannot take
// (run_applicable false) run the VALU kernel instead, with the same bits.
int run_elements(int Nq);
bool run_applicable(int Nq, const StageArgs &a);
int launch_run_f64(int Nq, const StageArgs &a, hipStream_t s);
int launch_elembnd_flux_f64(int Nq, const double *q, const double *u,
                            const double *v, const double *w,
                            const int32_t *nbr, const int32_t *vmapP,
                            const double *geo, const uint32_t *geoid,
                            int64_t Ne, double *flux, hipStream_t s);
int launch_halo_gather(double *f, const int32_t *src, int64_t base,
                       int64_t n, hipStream_t s);
int launch_rk_update(double *q, const double *q0, const double *dqdt,
                     double a, double b, double dt, int64_t n, hipStream_t s);
int launch_minmax(const double *q, int64_t n, double *partial, int nblk,
                  double *result, hipStream_t s);
int launch_to_f32(const double *src, float *dst, int64_t n, hipStream_t s);
int launch_init_gaussian(double *q, double *u, double *v, double *w, const double *x1D,
                         int Nq, int NeX, int NeY, int NeZ, const double h[3],
                         const double c[3], double sigma, const double vel[3],
                         hipStream_t s);
constexpr int kMinmaxBlocks = 1024;
// min/max of n doubles into result[0..1] with the partials in the library's
// static device scratch (kNoallocBlocks pairs); allocates nothing.
constexpr int kNoallocBlocks = 1024;
constexpr int64_t kNoallocMaxN = (int64_t)1 << 40;  // 8 TiB of doubles
int launch_minmax_noalloc(const double *q, int64_t n, double *result, hipStream_t s);

// Periodic neighbour table of a NeX x NeY x NeZ box built on the device, with
// element indices offset by elem_off.
int launch_box_nbr(int32_t *nbr, int NeX, int NeY, int NeZ, int64_t elem_off,
                   hipStream_t s);
// result[0..3] = min q, max q, jac * sum w3[n % Np] q[n], jac * sum w3[n % Np] q[n]^2
// (partial: 4 * kSummaryBlocks doubles).
constexpr int kSummaryBlocks = 256;
int launch_field_summary(const double *q, const double *w3, int Np, int64_t n, double jac,
                         double *partial, double *result, hipStream_t s);

}  // namespace dgk
