// MFMA stage kernels (gfx950), one template for two precisions:
//  * mixed (HIP_MP): the 1D-operator contractions and the Lift run in FP32 on
//    v_mfma_f32_16x16x4f32 with FP32 accumulation; the state q, the face-flux
//    jumps, the volume fluxes q*u_d and the RK stage combination stay FP64;
//    velocities are read from FP32 copies;
//  * FP64 (HIP, tile candidates): the same data flow in FP64 throughout, on
//    v_mfma_f64_16x16x4f64.
//
// For each direction d the sum-factorised derivative and that direction's two
// Lift faces form one GEMM over the block's "pencils" r (node lines along d):
//   out_d(r, c) = sum_l E_d F_d(r, l) D(c, l)
//               + l-(c) flux(face d-, r) + l+(c) flux(face d+, r)
// i.e. A_d = [E_d F_d | flux_d- | flux_d+ | 0]  (pencils x KPAD),
//      B   = [D^T ; l- ; l+ ; 0]               (KPAD x NQ), the same for
// every direction and element.  The reference's face-node order (Fmask,
// mod_mesh.f90:278-324) is the pencil order of the face's normal direction,
// so the Lift (tensorprod_Lift_hexahedral, mod_dg_optr_kernel.f90:262-288, in
// its tensor-product form) is two extra K columns of the contraction
// (tensorprod_divlike_dirXYZ, mod_dg_optr_kernel.f90:197-258).
//
// Data flow of one block (EPB elements, one lane per node column (i,j)):
//   1. the lane loads its q column (FP64), its velocity columns, the own
//      traces of the x/y faces and the neighbour traces straight into
//      registers (coalesced along (i,j));
//   2. FP64 upwind fluxes and FP64 volume fluxes E_d q u_d, each rounded once
//      to the operand precision, are written into augmented A arrays in LDS,
//      one pencil per row of KPAD values;
//   3. MFMA over 16-pencil tiles; K is permuted so that lane (col, kq) keeps
//      rows l = kq*KSTEPS + ks of B in registers and reads A at fixed offsets
//      from its own base; the result overwrites the tile's rows in place;
//   4. the three direction results are summed in FP64, dqdt = -((X + Y) + Z),
//      and RK-combined in FP64 (main.f90:69-73).
// The direction-sequential variant (SEQ) runs steps 2-3 one direction at a
// time through one LDS area, loading each velocity column only when its
// direction comes up: a third of the LDS and fewer live registers, for large
// orders; it adds the directions in the same order, so its results are the
// same bits.
#include <algorithm>
#include <atomic>

#include "dgk_internal.hpp"

namespace dgk {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Operand precision of the contraction GEMMs.
// Both MFMAs take A(row = lane%16, k = lane/16) and B(k = lane/16,
// col = lane%16); result value t of a lane lies in column lane%16 and row
// RSTRIDE_Q*(lane/16) + RSTRIDE_T*t (checked on the device for both).
struct PrecMixed {
  using vel_t = float;   // velocity copies read by the kernel
  using op_t = float;    // A, B operands in LDS / registers
  using acc_t = f32x4;   // 16x16 result, 4 values per lane
  static constexpr int RSTRIDE_Q = 4, RSTRIDE_T = 1;
  static __device__ __forceinline__ acc_t mfma(op_t a, op_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
struct PrecF64 {
  using vel_t = double;
  using op_t = double;
  using acc_t = f64x4;
  static constexpr int RSTRIDE_Q = 1, RSTRIDE_T = 4;
  static __device__ __forceinline__ acc_t mfma(op_t a, op_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

__device__ __forceinline__ int xcd_block_mp(int b, int n) {
  const int per = n >> 3, rem = n & 7;
  const int x = b & 7, local = b >> 3;
  return x * per + min(x, rem) + local;
}

__device__ __forceinline__ int64_t uniform_mp(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Tile shape of the MFMA stage kernel: EPB elements per block, all three
// directions at once or one after the other (SEQ).
template <class P, int NQ, int EPB_, bool SEQ_>
struct MP {
  static constexpr int NQ2 = NQ * NQ;
  static constexpr int NP = NQ2 * NQ;
  static constexpr int EPB = EPB_;
  static constexpr bool SEQ = SEQ_;
  static constexpr int LANES = EPB * NQ2;                   // node columns per block
  // Above p = 15 (289..400 columns) the block is eight wavefronts, two per
  // SIMD, the idle lanes taking GEMM tiles, and held to 4 waves per SIMD, so
  // two blocks share a CU (per launch at about 1.7e7 DOF, p = 16..19, against
  // 5..7 wavefronts and no cap: FP64 314/276/308/349 -> 278/270/263/320 us,
  // mixed 171/153/180/209 -> 151/166/176/193 us)
  static constexpr bool HI = SEQ && NQ > 16;
  static constexpr int THREADS = HI ? 512 : ((LANES + 63) / 64) * 64;
  static constexpr int WAVES = THREADS / 64;
  static constexpr int ROWS = ((LANES + 15) / 16) * 16;     // pencils per direction
  static constexpr int MTILES = ROWS / 16;
  static constexpr int NT = (NQ + 15) / 16;                // 16-wide column tiles of B
  static constexpr int KAUG = NQ + 2;
  static constexpr int KSTEPS = (KAUG + 3) / 4;
  static constexpr int KPAD = 4 * KSTEPS;
  static constexpr int AREA = ROWS * KPAD;                  // operands per direction
  static constexpr int NAREA = SEQ ? 1 : 3;
  static constexpr size_t A_BYTES = (size_t)NAREA * AREA * sizeof(typename P::op_t);
  // One-wavefront all-at-once shapes first pass the q, u, v columns through
  // LDS, where each lane picks up its own x/y face traces (same area as A,
  // used before it); the others load those traces from global memory (the
  // exchange would cost them registers and a wave per SIMD)
  static constexpr bool XCH = !SEQ && WAVES == 1;
  static constexpr size_t XCH_BYTES =
      XCH ? (size_t)LANES * NQ * (sizeof(double) + 2 * sizeof(typename P::vel_t)) : 0;
  // Multi-wavefront all-at-once shapes up to p = 9: the face lanes put the
  // x/y face traces into a buffer after the A area, and the x/y fluxes are
  // formed after the barrier that already precedes the GEMMs (one barrier
  // more; above p = 9 it cost more than the global trace loads it saves)
  static constexpr bool XD = !SEQ && WAVES > 1 && NQ <= 10;
  static constexpr size_t XD_BYTES =
      XD ? (size_t)EPB * 4 * NQ2 * (sizeof(double) + sizeof(typename P::vel_t)) : 0;
  static constexpr size_t LDS_BYTES = (A_BYTES > XCH_BYTES ? A_BYTES : XCH_BYTES) + XD_BYTES;
  // Register budget: the mixed exchange shapes keep 4 waves per SIMD (at most
  // 128 registers; left free, p = 7 lands two above and drops to 3), and so
  // do the shapes above p = 15 (FP64 spills 12..60 bytes per lane)
  static constexpr int MIN_WAVES_EU = ((sizeof(typename P::op_t) == 4 && XCH) || HI) ? 4 : 1;
  // Shapes whose face-normal form is tested at run time (see stage_mfma).
  static constexpr bool AX_RUNTIME = SEQ || MIN_WAVES_EU > 1;
  static constexpr int JOBS = NAREA * MTILES;               // (direction, tile) GEMMs per pass
  static constexpr int JPW = (JOBS + WAVES - 1) / WAVES;
};

// Candidate kinds per order: for Nq^2 <= 64 one, two or four wavefronts of
// whole elements, then one wavefront direction by direction; above, one
// element, two elements, and one element direction by direction; above
// p = 15 one element, all at once or direction by direction.  Kinds whose
// LDS exceeds a CU's 160 KiB are dropped (all at once above p = 15: FP64 up
// to p = 17, mixed up to p = 19); the remaining ones are the candidates 0,
// 1, ... (per-order defaults in kMpDefault below).
constexpr int kKinds = 4;
template <int NQ>
constexpr int kind_epb(int k) {
  constexpr int NQ2 = NQ * NQ;
  if (NQ2 <= 64) {
    const int waves = k == 1 ? 2 : k == 2 ? 4 : 1;
    const int e = 64 * waves / NQ2;
    return e < 1 ? 1 : e;
  }
  return k == 1 ? 2 : 1;
}
template <int NQ>
constexpr bool kind_seq(int k) {
  return NQ * NQ > 64 ? k == 2 : k == 3;
}
template <class P, int NQ>
constexpr bool kind_ok(int k) {
  if (NQ * NQ > 64 && k == 3) return false;
  if (NQ > kMaxNqValu && k == 1) return false;  // above p = 15: one element per workgroup
  return k == 0   ? MP<P, NQ, kind_epb<NQ>(0), kind_seq<NQ>(0)>::LDS_BYTES <= 160 * 1024
         : k == 1 ? MP<P, NQ, kind_epb<NQ>(1), kind_seq<NQ>(1)>::LDS_BYTES <= 160 * 1024
         : k == 2 ? MP<P, NQ, kind_epb<NQ>(2), kind_seq<NQ>(2)>::LDS_BYTES <= 160 * 1024
                  : MP<P, NQ, kind_epb<NQ>(3), kind_seq<NQ>(3)>::LDS_BYTES <= 160 * 1024;
}
template <class P, int NQ>
constexpr int ncand() {
  return (int)kind_ok<P, NQ>(0) + (int)kind_ok<P, NQ>(1) + (int)kind_ok<P, NQ>(2) +
         (int)kind_ok<P, NQ>(3);
}
// kind of candidate c (-1 if none)
template <class P, int NQ>
constexpr int kind_of(int c) {
  int n = 0;
  for (int k = 0; k < kKinds; ++k) {
    if (!kind_ok<P, NQ>(k)) continue;
    if (n == c) return k;
    ++n;
  }
  return -1;
}

// One 16-pencil tile: A rows from `tile`, B (NT column tiles) in registers;
// all A values are read before any result overwrites the tile's first NQ
// columns.
template <class P, int KSTEPS, int KPAD, int NQ, int NT>
__device__ __forceinline__ void mfma_tile(typename P::op_t *tile,
                                          const typename P::op_t (&bop)[NT][KSTEPS], int rd,
                                          int wr, int col) {
  typename P::op_t a[KSTEPS];
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ++ks) a[ks] = tile[rd + ks];
  typename P::acc_t acc[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    acc[nt] = typename P::acc_t{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks) acc[nt] = P::mfma(a[ks], bop[nt][ks], acc[nt]);
  }
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    if (16 * nt + col < NQ) {
#pragma unroll
      for (int t4 = 0; t4 < 4; ++t4) tile[wr + 16 * nt + t4 * P::RSTRIDE_T * KPAD] = acc[nt][t4];
    }
  }
}

// AXAL: every face normal lies along its face's axis (checked on the host); a
// compile-time form, so the axis-aligned kernel holds no registers for the
// general normal velocity (one more wave per SIMD for several multi-wavefront
// shapes, e.g. the mixed p = 10, 11, 15 defaults).  Two kinds of shape keep
// the test at run time (AX_RUNTIME), because decided at compile time their
// allocations grow: the direction-by-direction (SEQ) shapes (p = 5..9:
// 5 -> 4 waves per SIMD) and the mixed one-wavefront shapes held to 4 waves
// (p = 2..4 then spill 2 to 24 VGPRs).
// CVEL: constant velocity per element, (u, v, w) = words kGeoVel..+2 of its
// geometry record (batched and served boxes); no velocity field is read, and
// the operand-precision value of a constant is the value its per-node copy
// holds, so the results are the same bits.
// GU: several elements per workgroup and one geometry record for the whole
// launch (no per-element record ids): the record is read through the kernel
// argument itself, i.e. through the scalar cache (as in stage_f64).
template <class P, int NQ, int EPB, bool SEQ, bool AXAL, bool CVEL = false, bool GU = false>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, MP<P, NQ, EPB, SEQ>::THREADS),
                               amdgpu_waves_per_eu(MP<P, NQ, EPB, SEQ>::MIN_WAVES_EU)))
stage_mfma(const double *__restrict__ q_, const double *q0_,
           const typename P::vel_t *__restrict__ u_, const typename P::vel_t *__restrict__ v_,
           const typename P::vel_t *__restrict__ w_, double *out_,
           const int32_t *__restrict__ nbr_, const int32_t *__restrict__ vmapP_,
           const typename P::op_t *__restrict__ D_, const typename P::op_t *__restrict__ ell_,
           const double *__restrict__ geo_, const uint32_t *__restrict__ geoid_,
           const int32_t *__restrict__ order_, const int64_t Ne_, const double ca_,
           const double cb_, const double dt_, const int mode_, const int axis_aligned_,
           const int dt_geo_, const int rev_, const int64_t zp_) {
  using C = MP<P, NQ, EPB, SEQ>;
  using vel_t = typename P::vel_t;
  using op_t = typename P::op_t;
  constexpr int NQ2 = C::NQ2, NP = C::NP, NFT = 6 * NQ2, KPAD = C::KPAD, KSTEPS = C::KSTEPS;
  constexpr int AREA = C::AREA;
  extern __shared__ double smem_mp[];
  op_t *const As = (op_t *)smem_mp;

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int64_t slot0 = (int64_t)xcd_block_mp(blockIdx.x, gridDim.x) * EPB;
  const bool lane_ok = tid < C::LANES;
  const int el = lane_ok ? tid / NQ2 : 0;
  const int ij = lane_ok ? tid - el * NQ2 : 0;
  const int i = ij % NQ, j = ij / NQ;
  const int64_t slot = slot0 + el;
  const bool active = lane_ok && slot < Ne_;
  int64_t e;
  const double *g_lane;
  if (EPB == 1) {
    e = slot < Ne_ ? slot_element(slot, Ne_, order_, rev_, zp_) : 0;
    e = uniform_mp(e);
    g_lane = geo_ + (geoid_ ? (int64_t)geoid_[e] * kGeoStride : 0);
  } else {
    e = active ? slot_element(slot, Ne_, order_, rev_, zp_) : 0;
    g_lane = geo_ + (active && geoid_ ? (int64_t)geoid_[e] * kGeoStride : 0);
  }
  static_assert(!GU || EPB > 1, "one element per workgroup reads its record uniformly anyway");
  auto body = [&](const double *__restrict__ g) {
    const int64_t ebase = e * NP;
    const int64_t base = ebase + ij;
    const bool rk = mode_ == 1;
    const bool use_q0 = rk && ca_ != 0.0;
    const bool axal = C::AX_RUNTIME ? axis_aligned_ != 0 : AXAL;
    const int R = el * NQ2;
    const vel_t *const velp[3] = {u_, v_, w_};
    vel_t cv[3] = {(vel_t)0, (vel_t)0, (vel_t)0};
    if constexpr (CVEL) {
      cv[0] = (vel_t)g[kGeoVel];
      cv[1] = (vel_t)g[kGeoVel + 1];
      cv[2] = (vel_t)g[kGeoVel + 2];
    }

    // ---- 1. loads straight into registers ------------------------------------
    // Neighbour-table entries: with several elements per workgroup they are
    // per-lane loads, issued first so that waiting for them before the trace
    // gathers does not also wait for the column loads behind them; with one
    // they are scalar loads read where used (fewer live registers at high
    // order).
    int32_t nb6[6] = {};
    const bool resolved = vmapP_ != nullptr;
    if constexpr (EPB > 1) {
      if (active) nbr_entries<NQ>(vmapP_, nbr_, e, ij, nb6);
    }
    auto nbr_at = [&](int f) -> int64_t {
      if constexpr (EPB > 1) return nbr_node<NQ>(resolved, nb6[f], f, i, j);
      return resolved ? (int64_t)vmapP_[e * NFT + f * NQ2 + ij]
                      : (int64_t)nbr_[e * 6 + f] * NP + nbr_face_node(f, i, j, NQ);
    };
    double qc[NQ];
    vel_t uc[SEQ ? 1 : NQ], vc[SEQ ? 1 : NQ], wc[NQ];  // SEQ: w only at k = 0, NQ-1
    double qM[4], qP[6];
    vel_t vM[4], vP[6];  // velocity component along the face axis
    if (active) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        qc[k] = q_[base + k * NQ2];
        if constexpr (!SEQ) {
          uc[k] = CVEL ? cv[0] : u_[base + k * NQ2];
          vc[k] = CVEL ? cv[1] : v_[base + k * NQ2];
          wc[k] = CVEL ? cv[2] : w_[base + k * NQ2];
        }
      }
      if constexpr (SEQ) {
        wc[0] = CVEL ? cv[2] : w_[base];
        wc[NQ - 1] = CVEL ? cv[2] : w_[base + (NQ - 1) * NQ2];
      }
      if constexpr (!C::XCH && !C::XD) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {  // own traces of the x and y faces
          const int64_t m = ebase + own_face_node(f, i, j, NQ);
          qM[f] = q_[m];
          vM[f] = CVEL ? cv[kFaceAxis[f]] : (kFaceAxis[f] == 0 ? u_ : v_)[m];
        }
      }
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const int64_t pidx = nbr_at(f);
        qP[f] = q_[pidx];
        vP[f] = CVEL ? cv[kFaceAxis[f]]
                     : (kFaceAxis[f] == 0 ? u_ : kFaceAxis[f] == 1 ? v_ : w_)[pidx];
      }
    }

    // ---- 1b. own x/y face traces from the lanes that hold them (XCH) --------
    // Face node (a,b) of face f lies in the column of lane (a,0), (NQ-1,a),
    // (a,NQ-1) or (0,a) at k = b; every lane stores its q, u, v columns in LDS
    // and reads its four traces back, with no global loads (they would be
    // gathers of one cache line per two lanes on the x faces).
    if constexpr (C::XCH) {
      double *xq = (double *)smem_mp;
      vel_t *xu = (vel_t *)(xq + C::LANES * NQ), *xv = xu + C::LANES * NQ;
      if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
          xq[tid * NQ + k] = qc[k];
          if constexpr (!CVEL) {
            xu[tid * NQ + k] = uc[k];
            xv[tid * NQ + k] = vc[k];
          }
        }
      }
      __syncthreads();
      if (active) {
        const int src[4] = {R + i, R + (NQ - 1) + i * NQ, R + i + (NQ - 1) * NQ, R + i * NQ};
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          qM[f] = xq[src[f] * NQ + j];
          vM[f] = CVEL ? cv[kFaceAxis[f]] : (kFaceAxis[f] == 0 ? xu : xv)[src[f] * NQ + j];
        }
      }
      __syncthreads();  // the A operands below reuse the area
    }

    double *xd_q = (double *)((char *)smem_mp + C::A_BYTES) + el * 4 * NQ2;
    vel_t *xd_v =
        (vel_t *)((double *)((char *)smem_mp + C::A_BYTES) + EPB * 4 * NQ2) + el * 4 * NQ2;
    if constexpr (C::XD) {
      if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
          if (j == 0) {
            xd_q[0 * NQ2 + i + k * NQ] = qc[k];
            xd_v[0 * NQ2 + i + k * NQ] = vc[k];
          }
          if (i == NQ - 1) {
            xd_q[1 * NQ2 + j + k * NQ] = qc[k];
            xd_v[1 * NQ2 + j + k * NQ] = uc[k];
          }
          if (j == NQ - 1) {
            xd_q[2 * NQ2 + i + k * NQ] = qc[k];
            xd_v[2 * NQ2 + i + k * NQ] = vc[k];
          }
          if (i == 0) {
            xd_q[3 * NQ2 + j + k * NQ] = qc[k];
            xd_v[3 * NQ2 + j + k * NQ] = uc[k];
          }
        }
      }
    }
    // ---- 2a. FP64 upwind fluxes -------------------------------------------------
    op_t fl[6];
    auto face_flux = [&](int f) {
      {
        const double qm = f == 4 ? qc[0] : f == 5 ? qc[NQ - 1] : qM[f];
        double VelM, VelP;
        if (axal) {
          const int ax = kFaceAxis[f];
          const double na = g[kGeoNormal + 3 * f + ax];
          const vel_t vm = f == 4 ? wc[0] : f == 5 ? wc[NQ - 1] : vM[f];
          VelM = (double)vm * na;
          VelP = (double)vP[f] * na;
        } else {
          // general normals: all three components on both sides
          const int64_t m = ebase + own_face_node(f, i, j, NQ);
          const int64_t pidx = nbr_at(f);
          const double n1 = g[kGeoNormal + 3 * f], n2 = g[kGeoNormal + 3 * f + 1],
                       n3 = g[kGeoNormal + 3 * f + 2];
          if constexpr (CVEL) {
            VelM = (double)cv[0] * n1 + (double)cv[1] * n2 + (double)cv[2] * n3;
            VelP = VelM;
            (void)m;
            (void)pidx;
          } else {
            VelM = (double)u_[m] * n1 + (double)v_[m] * n2 + (double)w_[m] * n3;
            VelP = (double)u_[pidx] * n1 + (double)v_[pidx] * n2 + (double)w_[pidx] * n3;
          }
        }
        const double alpha = 0.5 * fabs(VelP + VelM);
        fl[f] = (op_t)(0.5 * g[kGeoFscale + f] * (qP[f] * VelP - qm * VelM - alpha * (qP[f] - qm)));
      }
    };
    if (active) {
#pragma unroll
      for (int f = C::XD ? 4 : 0; f < 6; ++f) face_flux(f);
    }

    // B = [D^T ; l- ; l+ ; 0] rows of this lane (k permuted, see above)
    constexpr int NT = C::NT;
    const int col = lane & 15, kq = lane >> 4;
    op_t bop[NT][KSTEPS];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int c = 16 * nt + col;
#pragma unroll
      for (int ks = 0; ks < KSTEPS; ++ks) {
        const int l = kq * KSTEPS + ks;
        bop[nt][ks] = c >= NQ ? (op_t)0 : l < NQ ? D_[c + l * NQ] : l == NQ ? ell_[c]
                      : l == NQ + 1 ? ell_[NQ + c] : (op_t)0;
      }
    }
    const int rd = col * KPAD + kq * KSTEPS;        // A read offset of this lane in a tile
    const int wr = P::RSTRIDE_Q * kq * KPAD + col;  // result offset of value 0
    // Node (i,j,k) sits in the x pencil (j,k) at position i, in the y pencil
    // (i,k) at position j and in the z pencil (i,j) at position k; face node
    // (i,j) of the faces normal to d is pencil ij of direction d.
    const int pos_off[3] = {(R + j) * KPAD + i, (R + i) * KPAD + j, (R + ij) * KPAD};
    const int pos_step[3] = {NQ * KPAD, NQ * KPAD, 1};
    // non-temporal RK base and output where k-slabs are whole 128-byte lines;
    // cached where one instruction covers pieces of lines (p <= 6), so the L2
    // merges them (as stage_f64; mixed p = 2, 4: 6% faster per launch)
    constexpr bool NT_SLAB = (NQ2 % 16 == 0) || NQ > 7;
    const int face_m[3] = {3, 0, 4}, face_p[3] = {1, 2, 5};  // -d and +d faces

    double q0v[NQ];
    double tend[NQ];
    if constexpr (!SEQ) {
      // ---- 2b. volume fluxes of all three directions -> A arrays ------------
      if (active) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double Ed = g[kGeoEscale + d];
          op_t *a = As + d * AREA + pos_off[d];
#pragma unroll
          for (int k = 0; k < NQ; ++k) {
            const vel_t vel = d == 0 ? uc[k] : d == 1 ? vc[k] : wc[k];
            a[k * pos_step[d]] = (op_t)(Ed * (qc[k] * (double)vel));
          }
          op_t *r = As + d * AREA + (R + ij) * KPAD;
          if (!C::XD || d == 2) {
            r[NQ] = fl[face_m[d]];
            r[NQ + 1] = fl[face_p[d]];
          }
#pragma unroll
          for (int l = NQ + 2; l < KPAD; ++l) r[l] = (op_t)0;
        }
      }
      // RK base column, in flight during the GEMMs (the RK base and the output
      // are touched once per launch: non-temporal, so L2 keeps the neighbour
      // traces)
      if (active && use_q0) {
#pragma unroll
        for (int k = 0; k < NQ; ++k)
          q0v[k] = NT_SLAB ? __builtin_nontemporal_load(q0_ + base + k * NQ2) : q0_[base + k * NQ2];
      }
      __syncthreads();
      if constexpr (C::XD) {
        if (active) {
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            qM[f] = xd_q[f * NQ2 + ij];
            vM[f] = CVEL ? cv[kFaceAxis[f]] : xd_v[f * NQ2 + ij];
            face_flux(f);
          }
#pragma unroll
          for (int d = 0; d < 2; ++d) {
            op_t *r = As + d * AREA + (R + ij) * KPAD;
            r[NQ] = fl[face_m[d]];
            r[NQ + 1] = fl[face_p[d]];
          }
        }
        __syncthreads();
      }

      // ---- 3. MFMA GEMMs over (direction, 16-pencil tile) jobs ---------------
      // Rows of A beyond the block's active pencils hold stale LDS; each output
      // row depends only on its own A row, and those rows are never read back.
#pragma unroll
      for (int jj = 0; jj < C::JPW; ++jj) {
        const int job = wave + jj * C::WAVES;
        if (job >= C::JOBS) break;  // wave-uniform
        const int d = job / C::MTILES, t = job - d * C::MTILES;
        mfma_tile<P, KSTEPS, KPAD, NQ, NT>(As + d * AREA + t * 16 * KPAD, bop, rd, wr, col);
      }
      __syncthreads();

      // ---- 4a. FP64 sum of the three directions ------------------------------
      if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k)
          tend[k] = ((double)As[pos_off[0] + k * pos_step[0]] +
                     (double)As[AREA + pos_off[1] + k * pos_step[1]]) +
                    (double)As[2 * AREA + pos_off[2] + k * pos_step[2]];
      }
    } else {
      // ---- 2b-4a, one direction at a time through one LDS area ---------------
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (active) {
          const double Ed = g[kGeoEscale + d];
          const vel_t *vel = velp[d];
          op_t *a = As + pos_off[d];
#pragma unroll
          for (int k = 0; k < NQ; ++k)
            a[k * pos_step[d]] =
                (op_t)(Ed * (qc[k] * (double)(CVEL ? cv[d] : vel[base + k * NQ2])));
          op_t *r = As + (R + ij) * KPAD;
          r[NQ] = fl[face_m[d]];
          r[NQ + 1] = fl[face_p[d]];
#pragma unroll
          for (int l = NQ + 2; l < KPAD; ++l) r[l] = (op_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < C::JPW; ++jj) {
          const int t = wave + jj * C::WAVES;
          if (t >= C::MTILES) break;  // wave-uniform
          mfma_tile<P, KSTEPS, KPAD, NQ, NT>(As + t * 16 * KPAD, bop, rd, wr, col);
        }
        __syncthreads();
        if (active) {
#pragma unroll
          for (int k = 0; k < NQ; ++k) {
            const double x = (double)As[pos_off[d] + k * pos_step[d]];
            tend[k] = d == 0 ? x : tend[k] + x;
          }
        }
        if (d < 2) __syncthreads();  // the next direction overwrites the area
      }
    }

    // ---- 4b. RK epilogue (FP64) -------------------------------------------------
    // SEQ reads the RK base only here: held through the last direction's GEMM it
    // cost p = 16, 17 a wave per SIMD (FP64 Nq = 17: 154 -> 124 VGPRs; p = 16
    // 399 -> 313 us per launch, p = 17 344 -> 276, mixed p = 19 271 -> 208)
    if (SEQ && active && use_q0) {
#pragma unroll
      for (int k = 0; k < NQ; ++k)
        q0v[k] = NT_SLAB ? __builtin_nontemporal_load(q0_ + base + k * NQ2) : q0_[base + k * NQ2];
    }
    if (active) {
      const double dtv = dt_geo_ ? g[kGeoDt] : dt_;
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        const double td = -tend[k];
        const int64_t idx = base + (int64_t)k * NQ2;
        double res = td;
        if (rk) {
          const double upd = cb_ * (qc[k] + dtv * td);
          res = use_q0 ? ca_ * q0v[k] + upd : upd;
        }
        if constexpr (NT_SLAB) __builtin_nontemporal_store(res, out_ + idx);
        else out_[idx] = res;
      }
    }
  };
  if constexpr (GU) body(geo_);
  else body(g_lane);
}

// ---- mixed precision over runs of consecutive elements (p <= 4) -----------
// stage_run_mp<NQ, AXAL>: the data flow of the FP64 run kernel
// (kernels_run.hip) with the mixed form's arithmetic.  A workgroup advances
// T consecutive elements: their q (FP64) and FP32 u, v, w blocks are one
// contiguous range per field, copied to LDS 16 bytes per thread; faces whose
// neighbour is one of the tile's elements take its traces from LDS, the
// others gather them.  The A operands (FP64 volume fluxes E_d q u_d and the
// FP64 face fluxes, each rounded once to FP32), the B operand, the MFMA tile
// GEMMs and the FP64 combination and RK update are those of stage_mfma, so
// every shape of the mixed form gives the same bits.
typedef double dbl2_mp __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int face_axis_mp(int f) { return f >= 4 ? 2 : ((f & 1) ? 0 : 1); }

template <int NQ>
struct RunMP {
  static constexpr int NQ2 = NQ * NQ, NP = NQ2 * NQ, NFT = 6 * NQ2;
  static constexpr int T = NQ == 2 ? 32 : NQ == 3 ? 16 : NQ == 4 ? 8 : 4;
  static constexpr int TNP = T * NP;
  static_assert(TNP % 4 == 0, "16-byte aligned FP32 tiles");
  static constexpr int PAIRS = TNP / 2;
  static constexpr int THREADS = ((PAIRS + 63) / 64) * 64;
  static constexpr int WAVES = THREADS / 64;
  static constexpr int QF = 2 * THREADS;                    // LDS doubles of q
  static constexpr int VCH = (TNP / 4 + 63) / 64;           // 1 KiB chunks of an FP32 field
  static constexpr int VF = 256 * VCH;                      // LDS floats of u, v or w
  static constexpr int ITEMS = T * NFT;
  static constexpr int R = (ITEMS + THREADS - 1) / THREADS;
  static constexpr int SLAB = R * THREADS;
  static constexpr int NBR_REGS = (6 * T + 63) / 64;
  static constexpr int ROWS = ((T * NQ2 + 15) / 16) * 16;  // pencils per direction
  static constexpr int MTILES = ROWS / 16;
  static constexpr int KSTEPS = (NQ + 2 + 3) / 4;
  static constexpr int KPAD = 4 * KSTEPS;
  static constexpr int AREA = ROWS * KPAD;
  static constexpr int JOBS = 3 * MTILES;
  static constexpr int JPW = (JOBS + WAVES - 1) / WAVES;
};

template <int NQ, bool AXAL>
__global__ void __launch_bounds__(RunMP<NQ>::THREADS)
stage_run_mp(const double *__restrict__ q_, const double *q0_, const float *__restrict__ u_,
             const float *__restrict__ v_, const float *__restrict__ w_, double *out_,
             const int32_t *__restrict__ nbr_, const float *__restrict__ D_,
             const float *__restrict__ ell_, const double *__restrict__ geo_,
             const uint32_t *__restrict__ geoid_, const int64_t Ne_, const double ca_,
             const double cb_, const double dt_, const int mode_, const int dt_geo_,
             const int rev_, const int64_t zpt_) {
  using C = RunMP<NQ>;
  using P = PrecMixed;
  constexpr int NQ2 = C::NQ2, NP = C::NP, T = C::T, NFT = C::NFT, R = C::R;
  constexpr int KPAD = C::KPAD, KSTEPS = C::KSTEPS, AREA = C::AREA;
  constexpr int NTV = AXAL ? 1 : 3;  // traced FP32 velocity fields
  __shared__ double sdbl[C::QF + C::SLAB + 24];
  __shared__ float sflt[3 * C::VF + NTV * C::SLAB + 3 * AREA];
  double *sq = sdbl, *tq = sq + C::QF, *sgeo = tq + C::SLAB;
  float *su = sflt, *sv = su + C::VF, *sw = sv + C::VF;
  float *tv = sw + C::VF;              // traced velocities, NTV slabs
  float *As = tv + NTV * C::SLAB;      // A operands of the three directions
  const float *svel[3] = {su, sv, sw};

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t ntile = (Ne_ + T - 1) / T;
  const int64_t tile =
      slot_element(xcd_block_mp(blockIdx.x, gridDim.x), ntile, nullptr, rev_, zpt_);
  const int64_t e0 = tile * T;
  const int nel = (int)min((int64_t)T, Ne_ - e0);
  const int rem = nel * NP;
  const double *g = geo_ + (geoid_ ? (int64_t)geoid_[e0] * kGeoStride : 0);

  int32_t nbv[C::NBR_REGS];
#pragma unroll
  for (int c = 0; c < C::NBR_REGS; ++c) {
    const int x = c * 64 + lane;
    nbv[c] = x < 6 * nel ? nbr_[e0 * 6 + x] : -1;
  }

  // ---- own fields: q 16 bytes (two nodes) and u, v, w 16 bytes (four nodes)
  // per thread, async into LDS ------------------------------------------------
  const bool qfull = 2 * tid + 1 < rem, qhalf = 2 * tid + 1 == rem;
  const bool vfull = 4 * tid + 3 < rem;
  const int vpart = vfull ? 0 : max(0, rem - 4 * tid);  // 1..3 nodes of a last granule
  const int64_t own = e0 * NP + 2 * tid, own4 = e0 * NP + 4 * tid;
#if defined(__HIP_DEVICE_COMPILE__)
  if (qfull)
    __builtin_amdgcn_global_load_lds((const void *)(q_ + own),
                                     (__attribute__((address_space(3))) void *)(sq + wave * 128),
                                     16, 0, 0);
  if (vfull) {
    const float *src[3] = {u_, v_, w_};
#pragma unroll
    for (int d = 0; d < 3; ++d)
      __builtin_amdgcn_global_load_lds(
          (const void *)(src[d] + own4),
          (__attribute__((address_space(3))) void *)((float *)svel[d] + wave * 256), 16, 0, 0);
  }
#endif
  double qtail = 0.0;
  if (qhalf) qtail = q_[own];
  float vtail[3][3] = {};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c < vpart) {
      vtail[0][c] = u_[own4 + c];
      vtail[1][c] = v_[own4 + c];
      vtail[2][c] = w_[own4 + c];
    }
  }

  // ---- neighbour traces of faces that leave the tile -----------------------
  int32_t nb_r[R];
  double tq_r[R];
  float tv_r[R][NTV];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    tq_r[r] = 0.0;
#pragma unroll
    for (int c = 0; c < NTV; ++c) tv_r[r][c] = 0.f;
    const int it = tid + r * C::THREADS;
    const int el = it / NFT, f = (it - el * NFT) / NQ2, ab = it - el * NFT - f * NQ2;
    const int x = el * 6 + f;
    int32_t nb = -1;
#pragma unroll
    for (int c = 0; c < C::NBR_REGS; ++c) {
      const int32_t sv_ = __shfl(nbv[c], x & 63);
      if ((x >> 6) == c) nb = sv_;
    }
    nb_r[r] = el < nel ? nb : -1;
    if (el < nel && (nb < e0 || nb >= e0 + nel)) {
      const int64_t pidx = (int64_t)nb * NP + nbr_face_node(f, ab % NQ, ab / NQ, NQ);
      tq_r[r] = q_[pidx];
      if constexpr (AXAL) {
        const int ax = face_axis_mp(f);
        tv_r[r][0] = (ax == 0 ? u_ : ax == 1 ? v_ : w_)[pidx];
      } else {
        tv_r[r][0] = u_[pidx];
        tv_r[r][1] = v_[pidx];
        tv_r[r][2] = w_[pidx];
      }
    }
  }

  // ---- RK base, B operand, geometry ------------------------------------------
  const bool rk = mode_ == 1;
  const bool use_q0 = rk && ca_ != 0.0;
  dbl2_mp q0v = {0.0, 0.0};
  if (use_q0) {
    if (qfull) q0v = __builtin_nontemporal_load((const dbl2_mp *)(q0_ + own));
    else if (qhalf) q0v.x = __builtin_nontemporal_load(q0_ + own);
  }
  const int col = lane & 15, kq = lane >> 4;
  float bop[1][KSTEPS];
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ++ks) {
    const int l = kq * KSTEPS + ks;
    bop[0][ks] = col >= NQ ? 0.f : l < NQ ? D_[col + l * NQ] : l == NQ ? ell_[col]
                 : l == NQ + 1 ? ell_[NQ + col] : 0.f;
  }
  static_assert(kGeoNormal == kGeoFscale + 6, "Fscale and normals are adjacent");
  if (tid < 24) sgeo[tid] = g[kGeoFscale + tid];
  __builtin_amdgcn_s_waitcnt(0);  // the async copies have landed in LDS

  if (qhalf) sq[2 * tid] = qtail;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c < vpart) {
      su[4 * tid + c] = vtail[0][c];
      sv[4 * tid + c] = vtail[1][c];
      sw[4 * tid + c] = vtail[2][c];
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int it = tid + r * C::THREADS;
    tq[it] = tq_r[r];
#pragma unroll
    for (int c = 0; c < NTV; ++c) tv[c * C::SLAB + it] = tv_r[r][c];
  }
  __syncthreads();

  // ---- FP64 upwind fluxes, rounded into the Lift columns of A ---------------
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int it = tid + r * C::THREADS;
    const int el = it / NFT, f = (it - el * NFT) / NQ2, ab = it - el * NFT - f * NQ2;
    const int32_t nb = nb_r[r];
    if (nb >= 0) {
      const int a = ab % NQ, b = ab / NQ;
      const bool inside = nb >= e0 && nb < e0 + nel;
      const int m = el * NP + own_face_node(f, a, b, NQ);
      const int pm = inside ? (int)(nb - e0) * NP + nbr_face_node(f, a, b, NQ) : 0;
      const double qM = sq[m];
      const double qP = inside ? sq[pm] : tq[it];
      const int ax = face_axis_mp(f);
      const double n1 = sgeo[6 + 3 * f], n2 = sgeo[7 + 3 * f], n3 = sgeo[8 + 3 * f];
      double VelM, VelP;
      if constexpr (AXAL) {
        const float *sf = ax == 0 ? su : ax == 1 ? sv : sw;
        const double na = ax == 0 ? n1 : ax == 1 ? n2 : n3;
        VelM = (double)sf[m] * na;
        VelP = (double)(inside ? sf[pm] : tv[it]) * na;
      } else {
        VelM = (double)su[m] * n1 + (double)sv[m] * n2 + (double)sw[m] * n3;
        const float up = inside ? su[pm] : tv[it];
        const float vp = inside ? sv[pm] : tv[C::SLAB + it];
        const float wp = inside ? sw[pm] : tv[2 * C::SLAB + it];
        VelP = (double)up * n1 + (double)vp * n2 + (double)wp * n3;
      }
      const double alpha = 0.5 * fabs(VelP + VelM);
      const float fl = (float)(0.5 * sgeo[f] * (qP * VelP - qM * VelM - alpha * (qP - qM)));
      // face f is the minus (faces 3, 0, 4) or plus (1, 2, 5) face of its axis
      const bool plus = f == 1 || f == 2 || f == 5;
      float *row = As + ax * AREA + (el * NQ2 + ab) * KPAD;
      row[NQ + (plus ? 1 : 0)] = fl;
      if (!plus) {
#pragma unroll
        for (int l = NQ + 2; l < KPAD; ++l) row[l] = 0.f;
      }
    }
  }

  // ---- FP64 volume fluxes E_d q u_d, rounded into A ---------------------------
  const double E3[3] = {g[kGeoEscale], g[kGeoEscale + 1], g[kGeoEscale + 2]};
  double qc[2] = {0.0, 0.0};
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int n = 2 * tid + r;
    if (n < rem) {
      const int el = n / NP, nn = n - el * NP;
      const int i = nn % NQ, j = (nn / NQ) % NQ, k = nn / NQ2, R0 = el * NQ2;
      const double qq = sq[n];
      qc[r] = qq;
      As[0 * AREA + (R0 + j + k * NQ) * KPAD + i] = (float)(E3[0] * (qq * (double)su[n]));
      As[1 * AREA + (R0 + i + k * NQ) * KPAD + j] = (float)(E3[1] * (qq * (double)sv[n]));
      As[2 * AREA + (R0 + i + j * NQ) * KPAD + k] = (float)(E3[2] * (qq * (double)sw[n]));
    }
  }
  __syncthreads();

  // ---- MFMA GEMMs over (direction, 16-pencil tile) jobs ---------------------
  const int rd = col * KPAD + kq * KSTEPS;
  const int wr = P::RSTRIDE_Q * kq * KPAD + col;
#pragma unroll
  for (int jj = 0; jj < C::JPW; ++jj) {
    const int job = wave + jj * C::WAVES;
    if (job >= C::JOBS) break;  // wave-uniform
    const int d = job / C::MTILES, t = job - d * C::MTILES;
    mfma_tile<P, KSTEPS, KPAD, NQ, 1>(As + d * AREA + t * 16 * KPAD, bop, rd, wr, col);
  }
  __syncthreads();

  // ---- FP64 sum of the directions, RK epilogue ---------------------------------
  const double dtv = dt_geo_ ? g[kGeoDt] : dt_;
  double res[2] = {0.0, 0.0};
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int n = 2 * tid + r;
    if (n < rem) {
      const int el = n / NP, nn = n - el * NP;
      const int i = nn % NQ, j = (nn / NQ) % NQ, k = nn / NQ2, R0 = el * NQ2;
      const double tend = ((double)As[0 * AREA + (R0 + j + k * NQ) * KPAD + i] +
                           (double)As[1 * AREA + (R0 + i + k * NQ) * KPAD + j]) +
                          (double)As[2 * AREA + (R0 + i + j * NQ) * KPAD + k];
      const double td = -tend;
      if (!rk) {
        res[r] = td;
      } else {
        const double upd = cb_ * (qc[r] + dtv * td);
        res[r] = use_q0 ? ca_ * (r == 0 ? q0v.x : q0v.y) + upd : upd;
      }
    }
  }
  if (qfull) {
    const dbl2_mp v = {res[0], res[1]};
    __builtin_nontemporal_store(v, (dbl2_mp *)(out_ + own));
  } else if (qhalf) {
    __builtin_nontemporal_store(res[0], out_ + own);
  }
}

template <int NQ, bool AXAL>
int launch_run_mp_nq(const StageArgsMP &a, hipStream_t s) {
  using C = RunMP<NQ>;
  const int64_t ntile = (a.Ne + C::T - 1) / C::T;
  if (ntile > INT32_MAX) return set_error(DGK_ERR_UNSUPPORTED, "too many elements");
  int64_t zpt = 0;  // the XCD-interleaved order at tile granularity, as stage_run
  if (a.zperiod > 0 && a.zperiod % C::T == 0) {
    const int64_t Qt = a.zperiod / C::T;
    if (Qt >= 8 && ntile % Qt == 0) zpt = Qt;
  }
  hipLaunchKernelGGL((stage_run_mp<NQ, AXAL>), dim3((unsigned)ntile), dim3(C::THREADS), 0, s,
                     a.q, a.q0, a.u, a.v, a.w, a.out, a.nbr, a.D, a.ell, a.geo, a.geoid, a.Ne,
                     a.ca, a.cb, a.dt, a.mode, a.dt_geo, a.reverse, zpt);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

// the FP64 run kernel's elements per workgroup at the same orders (the
// solver's per-run geometry check covers both)
constexpr int run_mp_elements(int NQ) {
  return NQ == 2 ? 32 : NQ == 3 ? 16 : NQ == 4 ? 8 : NQ == 5 ? 4 : 0;
}

bool aligned16_mp(const void *p) { return ((uintptr_t)p & 15) == 0; }

// The mixed run kernel takes the launch (true), or the launch needs another
// kernel (false): node connectivity, an order array, per-element geometry
// within a tile, unaligned fields, or constant velocities (no fields).
bool run_mp_applicable(int Nq, const StageArgsMP &a) {
  const int T = run_mp_elements(Nq);
  if (T < 1 || !a.nbr || a.vmapP || a.order || a.cvel) return false;
  if (a.geoid && !a.geo_per_run) return false;
  const bool q0_ok = a.mode != 1 || a.ca == 0.0 || aligned16_mp(a.q0);
  return aligned16_mp(a.q) && aligned16_mp(a.u) && aligned16_mp(a.v) && aligned16_mp(a.w) &&
         aligned16_mp(a.out) && q0_ok;
}

int launch_run_mp(int Nq, const StageArgsMP &a, hipStream_t s) {
  const bool ax = a.axis_aligned != 0;
  switch (Nq) {
  case 2: return ax ? launch_run_mp_nq<2, true>(a, s) : launch_run_mp_nq<2, false>(a, s);
  case 3: return ax ? launch_run_mp_nq<3, true>(a, s) : launch_run_mp_nq<3, false>(a, s);
  case 4: return ax ? launch_run_mp_nq<4, true>(a, s) : launch_run_mp_nq<4, false>(a, s);
  case 5: return ax ? launch_run_mp_nq<5, true>(a, s) : launch_run_mp_nq<5, false>(a, s);
  default: return set_error(DGK_ERR_UNSUPPORTED, "mixed run kernel: Nq outside 2..5");
  }
}

template <class P, int NQ, int EPB, bool SEQ, bool AXAL, bool CVEL, bool GU>
int prepare_kernel() {
  // set once per device; concurrent callers may both set it (idempotent)
  static std::atomic<int> prepared_dev{-1};
  int dev = 0;
  DGK_HIP(hipGetDevice(&dev));
  if (prepared_dev.load(std::memory_order_acquire) != dev) {
    DGK_HIP(hipFuncSetAttribute((const void *)stage_mfma<P, NQ, EPB, SEQ, AXAL, CVEL, GU>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)MP<P, NQ, EPB, SEQ>::LDS_BYTES));
    prepared_dev.store(dev, std::memory_order_release);
  }
  return DGK_OK;
}

// a form with its uniform-geometry variant where a workgroup holds several
// elements
template <class P, int NQ, int EPB, bool SEQ, bool AXAL, bool CVEL = false>
int prepare_form() {
  DGK_TRY((prepare_kernel<P, NQ, EPB, SEQ, AXAL, CVEL, false>()));
  if constexpr (EPB > 1) DGK_TRY((prepare_kernel<P, NQ, EPB, SEQ, AXAL, CVEL, true>()));
  return DGK_OK;
}

// Both face-normal forms of a shape (one, testing at run time, for the
// AX_RUNTIME shapes).
template <class P, int NQ, int EPB, bool SEQ>
int prepare_shape() {
  constexpr bool AXR = MP<P, NQ, EPB, SEQ>::AX_RUNTIME;
  DGK_TRY((prepare_form<P, NQ, EPB, SEQ, false>()));
  if constexpr (!AXR) DGK_TRY((prepare_form<P, NQ, EPB, SEQ, true>()));
  DGK_TRY((prepare_form<P, NQ, EPB, SEQ, !AXR, true>()));
  return DGK_OK;
}

template <class P, int NQ, int EPB, bool SEQ, bool AXAL, bool CVEL, bool GU, class Args>
int launch_kernel(const Args &a, hipStream_t s) {
  using T = MP<P, NQ, EPB, SEQ>;
  const int64_t nblk = (a.Ne + T::EPB - 1) / T::EPB;
  if (nblk > INT32_MAX) return set_error(DGK_ERR_UNSUPPORTED, "too many elements");
  DGK_TRY((prepare_kernel<P, NQ, EPB, SEQ, AXAL, CVEL, GU>()));
  hipLaunchKernelGGL((stage_mfma<P, NQ, EPB, SEQ, AXAL, CVEL, GU>), dim3((unsigned)nblk),
                     dim3(T::THREADS), T::LDS_BYTES, s, a.q, a.q0, a.u, a.v, a.w, a.out, a.nbr,
                     a.vmapP, a.D, a.ell, a.geo, a.geoid, a.order, a.Ne, a.ca, a.cb, a.dt,
                     a.mode, a.axis_aligned, a.dt_geo, a.reverse, a.zperiod);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

template <class P, int NQ, int EPB, bool SEQ, bool AXAL, class Args, bool CVEL = false>
int launch_form(const Args &a, hipStream_t s) {
  if constexpr (EPB > 1)
    if (!a.geoid) return launch_kernel<P, NQ, EPB, SEQ, AXAL, CVEL, true>(a, s);
  return launch_kernel<P, NQ, EPB, SEQ, AXAL, CVEL, false>(a, s);
}

template <class P, int NQ, int EPB, bool SEQ, class Args>
int launch_shape(const Args &a, hipStream_t s) {
  constexpr bool AXR = MP<P, NQ, EPB, SEQ>::AX_RUNTIME;
  // constant velocity: batched and served boxes, whose faces are axis-aligned
  if (a.cvel && a.axis_aligned && !a.vmapP)
    return launch_form<P, NQ, EPB, SEQ, !AXR, Args, true>(a, s);
  if constexpr (!AXR)
    if (a.axis_aligned) return launch_form<P, NQ, EPB, SEQ, true>(a, s);
  return launch_form<P, NQ, EPB, SEQ, false>(a, s);
}

// Candidate c of order NQ (compile-time), as a kernel instantiation.
template <class P, int NQ, int CAND>
struct Cand {
  static constexpr int K = kind_of<P, NQ>(CAND);
  static constexpr int EPB = kind_epb<NQ>(K < 0 ? 0 : K);
  static constexpr bool SEQ = kind_seq<NQ>(K < 0 ? 0 : K);
};

// Tile-shape candidate selected per Nq for the mixed form.  Defaults from
// per-order timings on MI355X at about 1.7e7 DOF (whole-step runs, every
// candidate, with a launch-wide geometry record read through the scalar
// cache): the VALU kernel at p = 1 (154 -> 149 us per launch against one
// wavefront of MFMA tiles; level at p = 2, slower above), the two-wavefront
// shape at p = 4, the four-wavefront shape at
// p = 5 (136 -> 126 us per launch), the direction-by-direction shape at
// p = 14, 16, 17 (p = 16: 203 -> 157 us, p = 17: 163 -> 138 us against one
// element all at once), the one-wavefront shape else (p = 2: 178 -> 174 us,
// p = 3: 133 -> 121 us against two wavefronts; p = 18: 165 -> 160 us,
// p = 19: 193 -> 177 us against direction by direction).
constexpr int kMpDefault[kMaxNq + 1] = {0, 0, 5, 0, 0, 1, 2, 0, 0, 0, 0,
                                        0, 0, 0, 0, 2, 0, 1, 1, 0, 0};
std::atomic<int> g_mp_shape[kMaxNq + 1] = {0, 0, 5, 0, 0, 1, 2, 0, 0, 0, 0,
                                           0, 0, 0, 0, 2, 0, 1, 1, 0, 0};
// Defaults of constant-velocity launches (batches, served requests), from the
// same timings with the velocity read from the geometry record, among the
// shapes that give the bits of the default above (the MFMA shapes and the
// run kernel), so a served request equals its solve on a plain solver: two
// wavefronts at p = 3 (161 -> 152 us), one wavefront at p = 5 (119 -> 105 us),
// the direction-by-direction shape at p = 6, 9, 12 (109 -> 102, 105 -> 98,
// 112 -> 106 us).
constexpr int kMpDefaultCvel[kMaxNq + 1] = {0, 0, 5, 0, 1, 1, 0, 3, 0, 0, 2,
                                            0, 0, 2, 0, 2, 0, 1, 1, 0, 0};
std::atomic<int> g_mp_shape_cvel[kMaxNq + 1] = {0, 0, 5, 0, 1, 1, 0, 3, 0, 0, 2,
                                                0, 0, 2, 0, 2, 0, 1, 1, 0, 0};

template <class P, int NQ>
int prepare_nq_all() {
  DGK_TRY((prepare_shape<P, NQ, Cand<P, NQ, 0>::EPB, Cand<P, NQ, 0>::SEQ>()));
  if constexpr (ncand<P, NQ>() > 1)
    DGK_TRY((prepare_shape<P, NQ, Cand<P, NQ, 1>::EPB, Cand<P, NQ, 1>::SEQ>()));
  if constexpr (ncand<P, NQ>() > 2)
    DGK_TRY((prepare_shape<P, NQ, Cand<P, NQ, 2>::EPB, Cand<P, NQ, 2>::SEQ>()));
  if constexpr (ncand<P, NQ>() > 3)
    DGK_TRY((prepare_shape<P, NQ, Cand<P, NQ, 3>::EPB, Cand<P, NQ, 3>::SEQ>()));
  return DGK_OK;
}

template <class P, int NQ, class Args>
int launch_nq_cand(int cand, const Args &a, hipStream_t s) {
  if constexpr (ncand<P, NQ>() > 1)
    if (cand == 1) return launch_shape<P, NQ, Cand<P, NQ, 1>::EPB, Cand<P, NQ, 1>::SEQ>(a, s);
  if constexpr (ncand<P, NQ>() > 2)
    if (cand == 2) return launch_shape<P, NQ, Cand<P, NQ, 2>::EPB, Cand<P, NQ, 2>::SEQ>(a, s);
  if constexpr (ncand<P, NQ>() > 3)
    if (cand == 3) return launch_shape<P, NQ, Cand<P, NQ, 3>::EPB, Cand<P, NQ, 3>::SEQ>(a, s);
  return launch_shape<P, NQ, Cand<P, NQ, 0>::EPB, Cand<P, NQ, 0>::SEQ>(a, s);
}

template <class P, int NQ>
int shapes_of(int *epb) {
  constexpr int n = ncand<P, NQ>();
  if constexpr (n > 0) epb[0] = Cand<P, NQ, 0>::EPB;
  if constexpr (n > 1) epb[1] = Cand<P, NQ, 1>::EPB;
  if constexpr (n > 2) epb[2] = Cand<P, NQ, 2>::EPB;
  if constexpr (n > 3) epb[3] = Cand<P, NQ, 3>::EPB;
  return n;
}

}  // namespace

#define DGK_NQ_SWITCH(NQV, CALL)                                                         \
  switch (NQV) {                                                                         \
  case 2: return CALL(2); case 3: return CALL(3); case 4: return CALL(4);                \
  case 5: return CALL(5); case 6: return CALL(6); case 7: return CALL(7);                \
  case 8: return CALL(8); case 9: return CALL(9); case 10: return CALL(10);              \
  case 11: return CALL(11); case 12: return CALL(12); case 13: return CALL(13);          \
  case 14: return CALL(14); case 15: return CALL(15); case 16: return CALL(16);          \
  case 17: return CALL(17); case 18: return CALL(18); case 19: return CALL(19);          \
  case 20: return CALL(20);                                                              \
  default: return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2.." + std::to_string(kMaxNq));                   \
  }

static int mp_mfma_shapes(int Nq, int *epb);

// Candidates of the mixed form: the MFMA shapes, then the run kernel at
// p <= 4, then the VALU kernel at p <= 7 (-1: none at this order).
static int mp_run_index(int Nq) {
  int e[kMaxTileShapes];
  return run_mp_elements(Nq) > 0 ? mp_mfma_shapes(Nq, e) : -1;
}
static int mp_valu_index(int Nq) {
  int e[kMaxTileShapes];
  if (Nq < kMinNq || Nq > kMaxNq || valu_mp_elements(Nq) <= 0) return -1;
  return mp_mfma_shapes(Nq, e) + (run_mp_elements(Nq) > 0 ? 1 : 0);
}

int prepare_stage_mp(int Nq) {
  DGK_TRY(prepare_stage_valu_mp(Nq));
#define DGK_CALL(N) prepare_nq_all<PrecMixed, N>()
  DGK_NQ_SWITCH(Nq, DGK_CALL)
#undef DGK_CALL
}

bool stage_mp_candidate_runs(int Nq, int cand, const StageArgsMP &a) {
  if (cand == mp_run_index(Nq)) return run_mp_applicable(Nq, a);
  if (cand == mp_valu_index(Nq)) return a.D64 && a.ell64;
  return true;
}

int launch_stage_mp(int Nq, const StageArgsMP &a, hipStream_t s) {
  if (a.Ne < 1) return DGK_OK;
  int cand = a.tile;
  if (cand < 0 && Nq >= kMinNq && Nq <= kMaxNq)
    cand = (a.cvel ? g_mp_shape_cvel : g_mp_shape)[Nq].load(std::memory_order_relaxed);
  // the VALU kernel (without the FP64 operators, the first shape runs) and the
  // run kernel (launches it cannot take run the first shape, with the same bits)
  if (cand >= 0 && cand == mp_valu_index(Nq)) {
    if (a.D64 && a.ell64) return launch_stage_valu_mp(Nq, a, s);
    cand = 0;
  }
  if (cand >= 0 && cand == mp_run_index(Nq)) {
    if (run_mp_applicable(Nq, a)) return launch_run_mp(Nq, a, s);
    cand = 0;
  }
#define DGK_CALL(N) launch_nq_cand<PrecMixed, N>(cand, a, s)
  DGK_NQ_SWITCH(Nq, DGK_CALL)
#undef DGK_CALL
}

int prepare_tile_f64(int Nq) {
#define DGK_CALL(N) prepare_nq_all<PrecF64, N>()
  DGK_NQ_SWITCH(Nq, DGK_CALL)
#undef DGK_CALL
}

int launch_tile_f64(int Nq, int cand, const StageArgs &a, hipStream_t s) {
  if (a.Ne < 1) return DGK_OK;
#define DGK_CALL(N) launch_nq_cand<PrecF64, N>(cand, a, s)
  DGK_NQ_SWITCH(Nq, DGK_CALL)
#undef DGK_CALL
}

static int mp_mfma_shapes(int Nq, int *epb) {
#define DGK_CALL(N) shapes_of<PrecMixed, N>(epb)
  DGK_NQ_SWITCH(Nq, DGK_CALL)
#undef DGK_CALL
}

int stage_mp_shapes(int Nq, int *epb, int n) {
  if (Nq < kMinNq || Nq > kMaxNq || !epb || n < kKinds + 2) return 0;
  int nc = mp_mfma_shapes(Nq, epb);
  if (nc > 0 && run_mp_elements(Nq) > 0) epb[nc++] = run_mp_elements(Nq);
  if (nc > 0 && valu_mp_elements(Nq) > 0) epb[nc++] = valu_mp_elements(Nq);
  return nc;
}

int tile_f64_shapes(int Nq, int *epb, int n) {
  if (Nq < kMinNq || Nq > kMaxNq || !epb || n < kKinds) return 0;
#define DGK_CALL(N) shapes_of<PrecF64, N>(epb)
  DGK_NQ_SWITCH(Nq, DGK_CALL)
#undef DGK_CALL
}

int stage_mp_select(int Nq, int cand) {
  int e[kMaxTileShapes];
  const int nc = stage_mp_shapes(Nq, e, kMaxTileShapes);
  if (nc < 1) return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2.." + std::to_string(kMaxNq));
  if (cand < -1 || cand >= nc)
    return set_error(DGK_ERR_ARG, "tile-shape candidate outside 0.." + std::to_string(nc - 1));
  g_mp_shape[Nq].store(cand == -1 ? kMpDefault[Nq] : cand, std::memory_order_relaxed);
  g_mp_shape_cvel[Nq].store(cand == -1 ? kMpDefaultCvel[Nq] : cand, std::memory_order_relaxed);
  return DGK_OK;
}

int stage_mp_selected(int Nq) {
  return (Nq < kMinNq || Nq > kMaxNq) ? 0 : g_mp_shape[Nq].load(std::memory_order_relaxed);
}

int stage_mp_selected_cvel(int Nq) {
  return (Nq < kMinNq || Nq > kMaxNq) ? 0 : g_mp_shape_cvel[Nq].load(std::memory_order_relaxed);
}

#undef DGK_NQ_SWITCH

}  // namespace dgk
