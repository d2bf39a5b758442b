// FP64 device kernels of the DG advection path (gfx950).
//
// One RK stage of the whole mesh is one launch.  The elements of a launch are
// independent: an element reads its own nodes, its neighbours' face traces of
// the stage input, and writes only its own nodes.  Both stage kernels fuse
// cal_elembnd_flux (mod_advect3d_eq.f90:89-142), cal_dqdt
// (mod_advect3d_eq.f90:146-201) with tensorprod_divlike_dirXYZ and
// tensorprod_Lift_hexahedral (mod_dg_optr_kernel.f90:197-288), and the RK
// update (main.f90:69-73).  Sums over l run in the reference's order
// (l ascending); products are fused into FMAs.
//
// stage_f64<NQ> (separable Lift): NQ*NQ lanes per element, lane (i,j) owns
// the k-column (i,j,:).  At p <= 7 a workgroup is one wavefront holding
// 64/(NQ*NQ) elements; above, an element spans ceil(NQ*NQ/64) wavefronts.
//   1. the workgroup's q,u,v,w blocks are copied HBM -> LDS with
//      global_load_lds_dwordx4 (no VGPRs held, 1 KiB per wave instruction);
//      meanwhile each lane gathers the neighbour traces of its face node (i,j)
//      on all six faces into registers;
//   2. upwind flux of face node (i,j) of every face from the LDS copy;
//   3. F = q*(u,v,w) in place in LDS;
//   With VREG only q and u are staged: the lane loads its v and w columns
//   (and the v traces of its y faces) into registers, and Fy takes the place
//   of q.  That halves the LDS per element (more workgroups per CU) for a
//   few more registers; the arithmetic is the same.
//   4. per node: the three sum-factorised contractions from LDS (D rows in
//      registers), the Lift in tensor-product form -- face f contributes
//      l-/+(x_n) * flux_f(tangential node) along its normal n, with l-/+ taken
//      from the stored Lift_mat (checked separable at solver creation) -- and
//      dqdt = -(((Ex Dx + Ey Dy) + Ez Dz) + Lift), stored or RK-combined.
//
// stage_f64_lift<NQ> (stored Lift): the same arithmetic with Lift_mat read as
// stored; used when Lift_mat is not of tensor-product form.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>

#include "dgk_internal.hpp"

namespace dgk {

namespace {

// Workgroups are dealt round-robin to the 8 XCDs; give each XCD a
// contiguous range of element blocks so neighbours share its L2.
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int per = n >> 3, rem = n & 7;
  const int x = b & 7, local = b >> 3;
  return x * per + min(x, rem) + local;
}

// Broadcast lane 0's value: marks a value the compiler cannot prove uniform
// as wave-uniform, so loads addressed by it become scalar loads.
__device__ __forceinline__ int64_t uniform(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

template <int NQ, bool VREG_>
struct Staged {
  static constexpr bool VREG = VREG_;
  static constexpr int NQ2 = NQ * NQ;
  static constexpr int NP = NQ2 * NQ;
  static constexpr int EPB = NQ2 <= 64 ? 64 / NQ2 : 1;  // elements per workgroup
  static constexpr int THREADS = NQ2 <= 64 ? 64 : ((NQ2 + 63) / 64) * 64;
  static constexpr int WAVES = THREADS / 64;
  // Async copy granule: 16 B per lane.  With NP odd (p = 2, 4, 6) an element
  // block starts 16-byte aligned or 8 bytes past; each element takes a slot of
  // NP + 1 doubles in LDS, filled from the aligned 16-byte granules that cover
  // it, so its nodes start at slot offset e & 1.  The granules reach 8 bytes
  // before an odd element or after an even one: into the neighbouring element
  // or the field allocation's tail pad (kStagePad).  Against 4-byte granules:
  // p = 2 225 -> 209 us per launch, p = 4 and 6 2-3% faster.
  static constexpr bool ODD = (NP % 2) == 1;
  static constexpr int NPP = ODD ? NP + 1 : NP;   // LDS doubles per element slot
  static constexpr int CHUNK = 128;               // doubles per wave instruction
  static constexpr int NCHUNK = (EPB * NPP + CHUNK - 1) / CHUNK;
  static constexpr int FIELD = NCHUNK * CHUNK;    // padded staged field (doubles)
  static constexpr int NSTAGED = VREG ? 2 : 4;     // staged fields: q, u (, v, w)
  static constexpr int OFF_G = NSTAGED * FIELD;   // x/y-face fluxes [EPB][4][NQ2]
  static constexpr int OFF_D = OFF_G + EPB * 4 * NQ2;
  static constexpr int OFF_L = OFF_D + NQ2;
  static constexpr int LDS_DOUBLES = OFF_L + 2 * NQ;
  static constexpr size_t LDS_BYTES = sizeof(double) * LDS_DOUBLES;
  // In a one-wavefront workgroup, Fx can be written slab-transposed in place
  // (lane (i,j) writes where lane (j,i) read in the same instruction slot);
  // the x-contraction then reads consecutive addresses.
  static constexpr bool FX_T = WAVES == 1;
  // Waves per SIMD the register allocation must leave room for: 3 for the
  // register form at p = 11, which otherwise lands one VGPR above the
  // 168 that 3 waves allow (and drops to 2).
  static constexpr int MIN_WAVES_EU = (VREG && NQ == 12) ? 3 : 1;
  // The axis-aligned register form also at p = 3, 4 (5 waves: 100 and 118
  // VGPRs -> 83-92, no spills; launches 6% faster) and p = 13 (3 waves: 183
  // VGPRs -> 168 with 28 bytes per lane spilled; 4% faster).  The staged
  // form ran slower held the same way.
  static constexpr int MIN_WAVES_EU_AX = !VREG ? 1 : NQ == 12 ? 3 : (NQ == 4 || NQ == 5) ? 5
                                         : NQ == 14 ? 3 : 1;
};

// AXAL: every face normal lies along its face's axis (checked on the host), so
// only q and the normal velocity component of the neighbour are gathered (the
// other components enter the normal velocity as exact zeros).  The general
// form holds three more traces per face in registers; a compile-time choice
// lets the axis-aligned kernel drop them (p = 1: 85 -> 62 VGPRs, 5 -> 8 waves
// per SIMD).  The register form at p = 11 keeps the test at run time: with
// it decided at compile time the allocation lands at 168 VGPRs with 2 spilled.
// CVEL: every element's velocity is constant, (u, v, w) = words kGeoVel..+2
// of its geometry record (batched and served boxes); no velocity field is
// read.  The constants enter the same expressions as the per-node values
// they equal, so the results are the same bits.
// GU: several elements per workgroup (p <= 4) and one geometry record for the
// whole launch (no per-element record ids).
// VT: the velocity fields' type.  double for the FP64 form (stage_f64);
// float for the mixed form's VALU kernel (stage_f64_mp): the velocities are
// stored in FP32 (4 of the 100 bytes per DOF-update of that form) and every
// operation is FP64, so its results are those of the FP64 form with the
// velocities rounded to FP32.  With float velocities only q is copied to LDS
// asynchronously; each lane stages its own u column as doubles.
template <int NQ, bool VREG, bool AXAL, bool CVEL, bool GU, typename VT>
__device__ __forceinline__ void
stage_valu(const double *__restrict__ q_, const double *q0_, const VT *__restrict__ u_,
           const VT *__restrict__ v_, const VT *__restrict__ w_, double *out_,
           const double *__restrict__ flux_in_, const int32_t *__restrict__ nbr_,
           const int32_t *__restrict__ vmapP_, const double *__restrict__ D_,
           const double *__restrict__ ell_, const double *__restrict__ geo_,
           const uint32_t *__restrict__ geoid_, const int32_t *__restrict__ order_,
           const int64_t Ne_, const double ca_, const double cb_, const double dt_,
           const int mode_, const int axis_aligned_, const int dt_geo_, const int rev_,
           const int64_t zp_) {
  // Read-only operands are __restrict__ kernel parameters, so wave-uniform
  // reads of them (operators, geometry, neighbour table) are scalar loads.
  static_assert(!CVEL || AXAL, "constant velocity on axis-aligned faces");
  constexpr bool F64_VEL = std::is_same<VT, double>::value;
  static_assert(F64_VEL || VREG, "FP32 velocities: the register form only");
  using C = Staged<NQ, VREG>;
  constexpr int NQ2 = C::NQ2, NP = C::NP, EPB = C::EPB, NFT = 6 * NQ2;
  extern __shared__ double smem[];
  // VREG: sv, sw are not staged (the pointers are unused)
  double *sq = smem, *su = smem + C::FIELD, *sv = smem + 2 * C::FIELD,
         *sw = smem + 3 * C::FIELD;
  double *sG = smem + C::OFF_G, *sD = smem + C::OFF_D, *sL = smem + C::OFF_L;

  const int tid = threadIdx.x;
  const int64_t slot0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * EPB;

  const bool lane_ok = tid < EPB * NQ2;
  const int el = lane_ok ? tid / NQ2 : 0;
  const int ij = lane_ok ? tid - el * NQ2 : 0;
  const int i = ij % NQ, j = ij / NQ;
  const int64_t slot = slot0 + el;
  const bool active = lane_ok && slot < Ne_;
  int64_t e;
  const double *g_lane;
  if (EPB == 1) {
    // one element per workgroup: its index, neighbour table entries and
    // geometry record are wave-uniform and read through the scalar cache
    e = slot < Ne_ ? slot_element(slot, Ne_, order_, rev_, zp_) : 0;
    e = uniform(e);
    g_lane = geo_ + (geoid_ ? (int64_t)geoid_[e] * kGeoStride : 0);
  } else {
    e = active ? slot_element(slot, Ne_, order_, rev_, zp_) : 0;
    g_lane = geo_ + (active && geoid_ ? (int64_t)geoid_[e] * kGeoStride : 0);
  }
  // The rest runs with g = the element's geometry record.  With several
  // elements per workgroup and one record for the whole launch (GU) it is
  // entered with the kernel argument itself, so the record's reads are
  // scalar-cache loads: as per-lane loads they were a memory round trip of
  // their own after the field loads had landed (3.7-4.3 us of a 12-16 us
  // workgroup life at p = 1..4, from in-kernel timestamps; launches 7-9%
  // faster without it).  A compile-time choice, so each form holds only the
  // registers of its own path.
  auto body = [&](const double *__restrict__ g) {
    const int64_t base = e * NP + ij;
    // element block in the staged fields
    const int lb = el * C::NPP + (C::ODD ? (int)(e & 1) : 0);
    const bool rk = mode_ == 1;
    const bool use_q0 = rk && ca_ != 0.0;
    double cvel[3] = {0.0, 0.0, 0.0};
    if constexpr (CVEL) {
      cvel[0] = g[kGeoVel];
      cvel[1] = g[kGeoVel + 1];
      cvel[2] = g[kGeoVel + 2];
    }

    // With several elements per workgroup the neighbour-table entries are
    // per-lane loads: issued first, so waiting for them before the trace
    // gathers does not also wait for the async copies behind them.  With one
    // element they are scalar loads, read where used (loading them ahead costs
    // a register, which at p = 11 drops the kernel from 3 to 2 waves per SIMD).
    const bool resolved = vmapP_ != nullptr;
    int32_t nb6[6];
    if constexpr (EPB > 1) {
      if (active && !flux_in_) nbr_entries<NQ>(vmapP_, nbr_, e, ij, nb6);
    }

    // ---- 1. async HBM -> LDS copy of q, u, v, w -------------------------------
    {
      const int wave = tid >> 6, lane = tid & 63;
      const double *src[4] = {q_, nullptr, nullptr, nullptr};
      if constexpr (F64_VEL) {
        src[1] = u_;
        src[2] = v_;
        src[3] = w_;
      }
      constexpr int NASYNC = (CVEL || !F64_VEL) ? 1 : C::NSTAGED;
#pragma unroll
      for (int c0 = 0; c0 < C::NCHUNK; c0 += C::WAVES) {
        const int c = c0 + wave;
        if (c < C::NCHUNK) {
          const int d = c * C::CHUNK + 2 * lane;  // first slot double of the granule
          const int del = d / C::NPP, dn = d - del * C::NPP;
          const int64_t sl = slot0 + del;
          int64_t off = 0;
          if (del < EPB && sl < Ne_) {
            const int64_t ge = slot_element(sl, Ne_, order_, rev_, zp_);
            off = ge * NP + dn - (C::ODD ? (ge & 1) : 0);
          }
#pragma unroll
          for (int fld = 0; fld < NASYNC; ++fld) {
            const double *gp = src[fld] + off;
            __attribute__((address_space(3))) void *lp =
                (__attribute__((address_space(3))) void *)(smem + fld * C::FIELD + c * C::CHUNK);
#if defined(__HIP_DEVICE_COMPILE__)
            __builtin_amdgcn_global_load_lds((const void *)gp, lp, 16, 0, 0);
#else
            (void)gp;
            (void)lp;
#endif
          }
        }
      }
    }

    // FP32 velocities: this lane's u column, as doubles, into the staged u
    // field at the slots the async copy would fill
    if constexpr (!F64_VEL && !CVEL) {
      if (active) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) su[lb + ij + k * NQ2] = (double)u_[base + (int64_t)k * NQ2];
      }
    }

    // neighbour traces of face node (i,j) on all six faces, and the RK base
    // column, in flight together with the async copy (axis-aligned: q and the
    // normal velocity component, fixed per face at compile time)
    constexpr bool AX_RUNTIME = VREG && NQ == 12;
    const bool axal = AX_RUNTIME ? axis_aligned_ != 0 : AXAL;
    double qP[6], uP[6], vP[6], wP[6], q0v[NQ];
    if (active && !flux_in_) {
      auto nbr_at = [&](int f) -> int64_t {
        if constexpr (EPB > 1)
          return nbr_node<NQ>(resolved, nb6[f], f, i, j);
        else
          return resolved ? (int64_t)vmapP_[e * NFT + f * NQ2 + ij]
                          : (int64_t)nbr_[e * 6 + f] * NP + nbr_face_node(f, i, j, NQ);
      };
      if (axal) {
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          const int64_t pidx = nbr_at(f);
          const VT *vel = kFaceAxis[f] == 0 ? u_ : kFaceAxis[f] == 1 ? v_ : w_;
          qP[f] = q_[pidx];
          uP[f] = CVEL ? cvel[kFaceAxis[f]] : vel[pidx];
        }
      } else {
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          const int64_t pidx = nbr_at(f);
          qP[f] = q_[pidx];
          uP[f] = u_[pidx];
          vP[f] = v_[pidx];
          wP[f] = w_[pidx];
        }
      }
    }
    // VREG: the v and w columns and the v traces of the y faces
    double vc[VREG ? NQ : 1], wc[VREG ? NQ : 1], vM[2];
    if (VREG && active) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        vc[k] = CVEL ? cvel[1] : v_[base + (int64_t)k * NQ2];
        wc[k] = CVEL ? cvel[2] : w_[base + (int64_t)k * NQ2];
      }
      vM[0] = CVEL ? cvel[1] : v_[e * NP + own_face_node(0, i, j, NQ)];
      vM[1] = CVEL ? cvel[1] : v_[e * NP + own_face_node(2, i, j, NQ)];
    }
    // the RK base and the output are touched once per launch: non-temporal,
    // so L2 keeps the neighbour traces -- where a k-slab of an element is a
    // run of whole 128-byte lines.  At p <= 6 with other slab sizes one
    // instruction covers pieces of lines, which non-temporal stores send to
    // memory as partial writes (twice the output's write requests at p = 1, 2)
    // and non-temporal loads fetch again per piece; cached accesses let the
    // L2 merge them (FP64 p = 1, 2, 4: 12%, 16%, 8% faster per launch).
    constexpr bool NT_SLAB = (NQ2 % 16 == 0) || NQ > 7;
    if (active && use_q0) {
#pragma unroll
      for (int k = 0; k < NQ; ++k)
        q0v[k] = NT_SLAB ? __builtin_nontemporal_load(q0_ + base + (int64_t)k * NQ2)
                    : q0_[base + (int64_t)k * NQ2];
    }
    // D and the 1D Lift go to LDS last: the wait for them comes after every
    // other load is in flight
    for (int t = tid; t < NQ2; t += C::THREADS) sD[t] = D_[t];
    for (int t = tid; t < 2 * NQ; t += C::THREADS) sL[t] = ell_[t];
    __builtin_amdgcn_s_waitcnt(0);  // the async copies have landed in LDS
    __syncthreads();

    // ---- 2. upwind flux of face node (i,j) on each face ----------------------
    double gf[6];
    if (active) {
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        if (flux_in_) {
          gf[f] = flux_in_[e * NFT + f * NQ2 + ij];
        } else {
          const double n1 = g[kGeoNormal + 3 * f], n2 = g[kGeoNormal + 3 * f + 1],
                       n3 = g[kGeoNormal + 3 * f + 2];
          const int m = lb + own_face_node(f, i, j, NQ);
          const double qM = sq[m];
          double VelM, VelP;
          if (axal) {
            double vm;
            if constexpr (CVEL)
              vm = cvel[kFaceAxis[f]];
            else if constexpr (VREG)
              vm = kFaceAxis[f] == 0 ? su[m] : f == 0 ? vM[0] : f == 2 ? vM[1]
                   : f == 4 ? wc[0] : wc[NQ - 1];
            else
              vm = (kFaceAxis[f] == 0 ? su : kFaceAxis[f] == 1 ? sv : sw)[m];
            const double na = g[kGeoNormal + 3 * f + kFaceAxis[f]];
            VelM = vm * na;
            VelP = uP[f] * na;
          } else {
            if constexpr (VREG) {
              const int64_t gm = e * NP + own_face_node(f, i, j, NQ);
              VelM = su[m] * n1 + v_[gm] * n2 + w_[gm] * n3;
            } else {
              VelM = su[m] * n1 + sv[m] * n2 + sw[m] * n3;
            }
            VelP = uP[f] * n1 + vP[f] * n2 + wP[f] * n3;
          }
          const double alpha = 0.5 * fabs(VelP + VelM);
          gf[f] = 0.5 * g[kGeoFscale + f] * (qP[f] * VelP - qM * VelM - alpha * (qP[f] - qM));
        }
      }
#pragma unroll
      for (int f = 0; f < 4; ++f) sG[(el * 4 + f) * NQ2 + ij] = gf[f];
    }
    __syncthreads();

    // ---- 3. volume fluxes: Fx, Fy in LDS, Fz and q in registers ---------------
    double qc[NQ], fz[NQ];
    if (active) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        const int n = lb + ij + k * NQ2;
        const double qq = sq[n], uu = CVEL ? cvel[0] : su[n];
        const double vv = CVEL ? cvel[1] : VREG ? vc[k] : sv[n];
        const double ww = CVEL ? cvel[2] : VREG ? wc[k] : sw[n];
        qc[k] = qq;
        fz[k] = qq * ww;
        if (C::FX_T) su[lb + j + i * NQ + k * NQ2] = qq * uu;
        else su[n] = qq * uu;
        (VREG ? sq : sv)[n] = qq * vv;  // Fy (in place of q with VREG)
      }
    }
    __syncthreads();

    // ---- 4. contractions, Lift, combine, epilogue -----------------------------
    if (active) {
      double Di[NQ], Dj[NQ];
#pragma unroll
      for (int l = 0; l < NQ; ++l) {
        Di[l] = sD[i + l * NQ];
        Dj[l] = sD[j + l * NQ];
      }
      const double Ex = g[kGeoEscale], Ey = g[kGeoEscale + 1], Ez = g[kGeoEscale + 2];
      const double lmi = sL[i], lpi = sL[NQ + i], lmj = sL[j], lpj = sL[NQ + j];
      const double *G = sG + el * 4 * NQ2;
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        double dx = 0.0, dy = 0.0, dz = 0.0;
#pragma unroll
        for (int l = 0; l < NQ; ++l) {
          dx += Di[l] * (C::FX_T ? su[lb + j + l * NQ + k * NQ2] : su[lb + l + j * NQ + k * NQ2]);
          dy += Dj[l] * (VREG ? sq : sv)[lb + i + l * NQ + k * NQ2];
          dz += sD[k + l * NQ] * fz[l];  // wave-uniform LDS broadcast
        }
        const double lift = lmj * G[0 * NQ2 + i + k * NQ] + lpi * G[1 * NQ2 + j + k * NQ] +
                            lpj * G[2 * NQ2 + i + k * NQ] + lmi * G[3 * NQ2 + j + k * NQ] +
                            sL[k] * gf[4] + sL[NQ + k] * gf[5];
        const double tend = -(((Ex * dx + Ey * dy) + Ez * dz) + lift);
        const int64_t idx = base + (int64_t)k * NQ2;
        double res = tend;
        if (rk) {
          const double upd = cb_ * (qc[k] + (dt_geo_ ? g[kGeoDt] : dt_) * tend);
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

template <int NQ, bool VREG, bool AXAL, bool CVEL, bool GU>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, Staged<NQ, VREG>::THREADS),
                               amdgpu_waves_per_eu(AXAL ? Staged<NQ, VREG>::MIN_WAVES_EU_AX
                                                        : Staged<NQ, VREG>::MIN_WAVES_EU)))
stage_f64(const double *__restrict__ q_, const double *q0_, const double *__restrict__ u_,
          const double *__restrict__ v_, const double *__restrict__ w_, double *out_,
          const double *__restrict__ flux_in_, const int32_t *__restrict__ nbr_,
          const int32_t *__restrict__ vmapP_, const double *__restrict__ D_,
          const double *__restrict__ ell_, const double *__restrict__ geo_,
          const uint32_t *__restrict__ geoid_, const int32_t *__restrict__ order_,
          const int64_t Ne_, const double ca_, const double cb_, const double dt_,
          const int mode_, const int axis_aligned_, const int dt_geo_, const int rev_,
          const int64_t zp_) {
  stage_valu<NQ, VREG, AXAL, CVEL, GU, double>(q_, q0_, u_, v_, w_, out_, flux_in_, nbr_,
                                               vmapP_, D_, ell_, geo_, geoid_, order_, Ne_,
                                               ca_, cb_, dt_, mode_, axis_aligned_, dt_geo_,
                                               rev_, zp_);
}

// The mixed form's VALU kernel (register form, FP32 velocity fields).
template <int NQ, bool AXAL, bool CVEL, bool GU>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, Staged<NQ, true>::THREADS),
                               amdgpu_waves_per_eu(AXAL ? Staged<NQ, true>::MIN_WAVES_EU_AX
                                                        : Staged<NQ, true>::MIN_WAVES_EU)))
stage_f64_mp(const double *__restrict__ q_, const double *q0_, const float *__restrict__ u_,
             const float *__restrict__ v_, const float *__restrict__ w_, double *out_,
             const int32_t *__restrict__ nbr_, const int32_t *__restrict__ vmapP_,
             const double *__restrict__ D_, const double *__restrict__ ell_,
             const double *__restrict__ geo_, const uint32_t *__restrict__ geoid_,
             const int32_t *__restrict__ order_, const int64_t Ne_, const double ca_,
             const double cb_, const double dt_, const int mode_, const int axis_aligned_,
             const int dt_geo_, const int rev_, const int64_t zp_) {
  stage_valu<NQ, true, AXAL, CVEL, GU, float>(q_, q0_, u_, v_, w_, out_, nullptr, nbr_, vmapP_,
                                              D_, ell_, geo_, geoid_, order_, Ne_, ca_, cb_,
                                              dt_, mode_, axis_aligned_, dt_geo_, rev_, zp_);
}

// Stored-Lift kernel: Lift_mat read as stored (any Lift_mat).
template <int NQ>
struct Tile {
  static constexpr int NQ2 = NQ * NQ;
  static constexpr int NP = NQ2 * NQ;
  // Elements per workgroup: fill ~256 lanes.
  static constexpr int EPB = (256 / NQ2) > 0 ? (256 / NQ2) : 1;
  static constexpr int THREADS = EPB * NQ2;
};

template <int NQ>
__global__ void __launch_bounds__(Tile<NQ>::THREADS)
stage_f64_lift(const StageArgs P) {
  constexpr int NQ2 = Tile<NQ>::NQ2, NP = Tile<NQ>::NP, EPB = Tile<NQ>::EPB;
  constexpr int NFT = 6 * NQ2;
  __shared__ double sF[EPB][NP];        // Fx (slab-transposed), then Fy
  __shared__ double sQ[EPB][6][NQ2];    // own face traces of q, then face flux
  __shared__ double sV[EPB][6][NQ2];    // own normal velocity on each face

  const int tid = threadIdx.x;
  const int el = tid / NQ2, ij = tid - el * NQ2;
  const int i = ij % NQ, j = ij / NQ;
  const int64_t slot = (int64_t)xcd_block(blockIdx.x, gridDim.x) * EPB + el;
  const bool active = slot < P.Ne;
  const int64_t e = active ? slot_element(slot, P.Ne, P.order, P.reverse, P.zperiod) : 0;

  const double *g = P.geo + (active && P.geoid ? (int64_t)P.geoid[e] * kGeoStride : 0);
  const int64_t base = e * NP + ij;

  double q[NQ], fy[NQ], fz[NQ], t[NQ];
  double Di[NQ], Dj[NQ];
#pragma unroll
  for (int l = 0; l < NQ; ++l) {
    Di[l] = P.D[i + l * NQ];
    Dj[l] = P.D[j + l * NQ];
  }

  // ---- A: columns, volume fluxes, own face traces -----------------------
  if (active) {
    double nrm[6][3];
#pragma unroll
    for (int f = 0; f < 6; ++f)
#pragma unroll
      for (int d = 0; d < 3; ++d) nrm[f][d] = g[kGeoNormal + 3 * f + d];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const int64_t idx = base + (int64_t)k * NQ2;
      const double qq = P.q[idx], uu = P.u[idx], vv = P.v[idx], ww = P.w[idx];
      q[k] = qq;
      sF[el][j + i * NQ + k * NQ2] = qq * uu;
      fy[k] = qq * vv;
      fz[k] = qq * ww;
      auto trace = [&](int f, int fp) {
        sQ[el][f][fp] = qq;
        sV[el][f][fp] = uu * nrm[f][0] + vv * nrm[f][1] + ww * nrm[f][2];
      };
      if (j == 0) trace(0, i + k * NQ);
      if (i == NQ - 1) trace(1, j + k * NQ);
      if (j == NQ - 1) trace(2, i + k * NQ);
      if (i == 0) trace(3, j + k * NQ);
      if (k == 0) trace(4, ij);
      if (k == NQ - 1) trace(5, ij);
    }
  }
  __syncthreads();

  // ---- B: face flux of node (i,j) on every face; x-derivative ------------
  if (active) {
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      double flux;
      if (P.flux_in) {
        flux = P.flux_in[e * NFT + f * NQ2 + ij];
      } else {
        const int64_t pidx = P.vmapP ? (int64_t)P.vmapP[e * NFT + f * NQ2 + ij]
                                     : (int64_t)P.nbr[e * 6 + f] * NP +
                                           nbr_face_node(f, i, j, NQ);
        const double qP = P.q[pidx], uP = P.u[pidx], vP = P.v[pidx], wP = P.w[pidx];
        const double n1 = g[kGeoNormal + 3 * f], n2 = g[kGeoNormal + 3 * f + 1],
                     n3 = g[kGeoNormal + 3 * f + 2];
        const double qM = sQ[el][f][ij], VelM = sV[el][f][ij];
        const double VelP = uP * n1 + vP * n2 + wP * n3;
        const double alpha = 0.5 * fabs(VelP + VelM);
        flux = 0.5 * g[kGeoFscale + f] * (qP * VelP - qM * VelM - alpha * (qP - qM));
      }
      sQ[el][f][ij] = flux;
    }
    const double Ex = g[kGeoEscale];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < NQ; ++l) s += Di[l] * sF[el][j + l * NQ + k * NQ2];
      t[k] = Ex * s;
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) sF[el][i + j * NQ + k * NQ2] = fy[k];
  }
  __syncthreads();

  // ---- C: y, z derivatives, Lift, combine, epilogue ----------------------
  if (active) {
    const double Ey = g[kGeoEscale + 1], Ez = g[kGeoEscale + 2];
    const bool rk = P.mode == 1;
    const bool use_q0 = rk && P.ca != 0.0;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      double sy = 0.0, sz = 0.0;
#pragma unroll
      for (int l = 0; l < NQ; ++l) {
        sy += Dj[l] * sF[el][i + l * NQ + k * NQ2];
        sz += P.D[k + l * NQ] * fz[l];
      }
      const int n = ij + k * NQ2;
      const double *L = P.Lift + n;
      const double lift = L[0 * NP] * sQ[el][0][i + k * NQ] +
                          L[1 * NP] * sQ[el][1][j + k * NQ] +
                          L[2 * NP] * sQ[el][2][i + k * NQ] +
                          L[3 * NP] * sQ[el][3][j + k * NQ] +
                          L[4 * NP] * sQ[el][4][ij] + L[5 * NP] * sQ[el][5][ij];
      const double tend = -(((t[k] + Ey * sy) + Ez * sz) + lift);
      const int64_t idx = base + (int64_t)k * NQ2;
      if (!rk) {
        P.out[idx] = tend;
      } else {
        const double upd = P.cb * (q[k] + (P.dt_geo ? g[kGeoDt] : P.dt) * tend);
        P.out[idx] = use_q0 ? P.ca * P.q0[idx] + upd : upd;
      }
    }
  }
}

template <int NQ, bool VREG, bool AXAL, bool CVEL, bool GU>
int prepare_form_valu() {
  // set once per device; concurrent callers may both set it (idempotent)
  static std::atomic<int> prepared_dev{-1};
  int dev = 0;
  DGK_HIP(hipGetDevice(&dev));
  if (prepared_dev.load(std::memory_order_acquire) != dev) {
    DGK_HIP(hipFuncSetAttribute((const void *)stage_f64<NQ, VREG, AXAL, CVEL, GU>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)Staged<NQ, VREG>::LDS_BYTES));
    prepared_dev.store(dev, std::memory_order_release);
  }
  return DGK_OK;
}

// the uniform-geometry form exists where a workgroup holds several elements
template <int NQ, bool VREG>
constexpr bool has_gu() {
  return Staged<NQ, VREG>::EPB > 1;
}

template <int NQ, bool VREG, bool AXAL, bool CVEL = false>
int prepare_shape_valu() {
  DGK_TRY((prepare_form_valu<NQ, VREG, AXAL, CVEL, false>()));
  if constexpr (has_gu<NQ, VREG>()) DGK_TRY((prepare_form_valu<NQ, VREG, AXAL, CVEL, true>()));
  return DGK_OK;
}

template <int NQ>
int prepare_nq() {
  DGK_TRY((prepare_shape_valu<NQ, false, false>()));
  DGK_TRY((prepare_shape_valu<NQ, false, true>()));
  DGK_TRY((prepare_shape_valu<NQ, true, false>()));
  DGK_TRY((prepare_shape_valu<NQ, false, true, true>()));
  DGK_TRY((prepare_shape_valu<NQ, true, true, true>()));
  return prepare_shape_valu<NQ, true, true>();
}

// The VALU kernels keep the element numbering where an XCD's eighth of a
// period is under 96 elements (per launch at about 1.7e7 DOF: p = 9..13 with
// 84..40 elements per eighth ran 1-4% slower in the XCD-interleaved order;
// p = 2..8 with 903..98 elements 2-11% faster).  The MFMA kernels take the
// interleaved order at every size.
constexpr int64_t kValuMinEighth = 96;
int64_t valu_period(int64_t q) { return q / 8 >= kValuMinEighth ? q : 0; }

template <int NQ, bool VREG, bool AXAL, bool CVEL, bool GU>
int launch_form_valu(const StageArgs &a, hipStream_t s) {
  using T = Staged<NQ, VREG>;
  const int64_t nblk = (a.Ne + T::EPB - 1) / T::EPB;
  if (nblk > INT32_MAX) return set_error(DGK_ERR_UNSUPPORTED, "too many elements");
  DGK_TRY((prepare_form_valu<NQ, VREG, AXAL, CVEL, GU>()));
  hipLaunchKernelGGL((stage_f64<NQ, VREG, AXAL, CVEL, GU>), dim3((unsigned)nblk),
                     dim3(T::THREADS), T::LDS_BYTES, s, a.q, a.q0, a.u, a.v, a.w, a.out,
                     a.flux_in, a.nbr, a.vmapP, a.D, a.ell, a.geo, a.geoid, a.order, a.Ne, a.ca,
                     a.cb, a.dt, a.mode, a.axis_aligned, a.dt_geo, a.reverse,
                     valu_period(a.zperiod));
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

template <int NQ, bool VREG, bool AXAL, bool CVEL = false>
int launch_valu(const StageArgs &a, hipStream_t s) {
  if constexpr (has_gu<NQ, VREG>())
    if (!a.geoid) return launch_form_valu<NQ, VREG, AXAL, CVEL, true>(a, s);
  return launch_form_valu<NQ, VREG, AXAL, CVEL, false>(a, s);
}

// VALU stage kernel (VREG: v, w in registers; the axis-aligned or general
// normal form), or the stored-Lift kernel.
template <int NQ>
int launch_nq(const StageArgs &a, hipStream_t s, bool vreg) {
  if (a.ell) {
    const bool ax = a.axis_aligned != 0;
    if (a.cvel && ax && !a.flux_in && !a.vmapP)
      return vreg ? launch_valu<NQ, true, true, true>(a, s) : launch_valu<NQ, false, true, true>(a, s);
    if (vreg) return ax ? launch_valu<NQ, true, true>(a, s) : launch_valu<NQ, true, false>(a, s);
    return ax ? launch_valu<NQ, false, true>(a, s) : launch_valu<NQ, false, false>(a, s);
  }
  using T = Tile<NQ>;
  const int64_t nblk = (a.Ne + T::EPB - 1) / T::EPB;
  if (nblk > INT32_MAX) return set_error(DGK_ERR_UNSUPPORTED, "too many elements");
  hipLaunchKernelGGL(stage_f64_lift<NQ>, dim3((unsigned)nblk), dim3(T::THREADS), 0, s, a);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

template <int NQ, bool AXAL, bool CVEL, bool GU>
int prepare_form_valu_mp() {
  // set once per device; concurrent callers may both set it (idempotent)
  static std::atomic<int> prepared_dev{-1};
  int dev = 0;
  DGK_HIP(hipGetDevice(&dev));
  if (prepared_dev.load(std::memory_order_acquire) != dev) {
    DGK_HIP(hipFuncSetAttribute((const void *)stage_f64_mp<NQ, AXAL, CVEL, GU>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)Staged<NQ, true>::LDS_BYTES));
    prepared_dev.store(dev, std::memory_order_release);
  }
  return DGK_OK;
}

template <int NQ, bool AXAL, bool CVEL = false>
int prepare_shape_valu_mp() {
  DGK_TRY((prepare_form_valu_mp<NQ, AXAL, CVEL, false>()));
  if constexpr (has_gu<NQ, true>()) DGK_TRY((prepare_form_valu_mp<NQ, AXAL, CVEL, true>()));
  return DGK_OK;
}

template <int NQ, bool AXAL, bool CVEL, bool GU>
int launch_form_valu_mp(const StageArgsMP &a, hipStream_t s) {
  using T = Staged<NQ, true>;
  const int64_t nblk = (a.Ne + T::EPB - 1) / T::EPB;
  if (nblk > INT32_MAX) return set_error(DGK_ERR_UNSUPPORTED, "too many elements");
  DGK_TRY((prepare_form_valu_mp<NQ, AXAL, CVEL, GU>()));
  hipLaunchKernelGGL((stage_f64_mp<NQ, AXAL, CVEL, GU>), dim3((unsigned)nblk),
                     dim3(T::THREADS), T::LDS_BYTES, s, a.q, a.q0, a.u, a.v, a.w, a.out,
                     a.nbr, a.vmapP, a.D64, a.ell64, a.geo, a.geoid, a.order, a.Ne, a.ca,
                     a.cb, a.dt, a.mode, a.axis_aligned, a.dt_geo, a.reverse,
                     valu_period(a.zperiod));
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

template <int NQ, bool AXAL, bool CVEL = false>
int launch_valu_mp(const StageArgsMP &a, hipStream_t s) {
  if constexpr (has_gu<NQ, true>())
    if (!a.geoid) return launch_form_valu_mp<NQ, AXAL, CVEL, true>(a, s);
  return launch_form_valu_mp<NQ, AXAL, CVEL, false>(a, s);
}

// cal_elembnd_flux: one thread per face node (mod_advect3d_eq.f90:116-140).
__global__ void elembnd_flux_f64(const double *__restrict__ q,
                                 const double *__restrict__ u,
                                 const double *__restrict__ v,
                                 const double *__restrict__ w,
                                 const int32_t *__restrict__ nbr,
                                 const int32_t *__restrict__ vmapP,
                                 const double *__restrict__ geo,
                                 const uint32_t *__restrict__ geoid, int Nq,
                                 int64_t Ne, double *__restrict__ flux) {
  const int Nfp = Nq * Nq, NfpTot = 6 * Nfp, Np = Nfp * Nq;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (int64_t)NfpTot * Ne) return;
  const int64_t e = r / NfpTot;
  const int rr = (int)(r - e * NfpTot), f = rr / Nfp, fp = rr - f * Nfp;
  const int a = fp % Nq, b = fp / Nq;
  const double *g = geo + (geoid ? (int64_t)geoid[e] * kGeoStride : 0);
  const int64_t iM = e * Np + own_face_node(f, a, b, Nq);
  const int64_t iP = vmapP ? (int64_t)vmapP[r] : (int64_t)nbr[e * 6 + f] * Np + nbr_face_node(f, a, b, Nq);
  const double n1 = g[kGeoNormal + 3 * f], n2 = g[kGeoNormal + 3 * f + 1],
               n3 = g[kGeoNormal + 3 * f + 2];
  const double qM = q[iM], qP = q[iP];
  const double VelM = u[iM] * n1 + v[iM] * n2 + w[iM] * n3;
  const double VelP = u[iP] * n1 + v[iP] * n2 + w[iP] * n3;
  const double alpha = 0.5 * fabs(VelP + VelM);
  flux[r] = 0.5 * g[kGeoFscale + f] * (qP * VelP - qM * VelM - alpha * (qP - qM));
}

// update_halo (mod_mesh.f90:216-218): f(base + i) = f(src(i)).
__global__ void halo_gather(double *f, const int32_t *__restrict__ src,
                            int64_t base, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) f[base + t] = f[src[t]];
}

// RK stage update (main.f90:69-73), in place on q.
__global__ void rk_update(double *q, const double *__restrict__ q0,
                          const double *__restrict__ dqdt, double a, double b,
                          double dt, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) q[t] = a * q0[t] + b * (q[t] + dt * dqdt[t]);
}

// min/max of q: per-block partials, then one block folds them.
__global__ void minmax_partial(const double *__restrict__ q, int64_t n,
                               double *partial) {
  __shared__ double smin[256], smax[256];
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double x = q[t];
    mn = fmin(mn, x);
    mx = fmax(mx, x);
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = smin[0];
    partial[2 * blockIdx.x + 1] = smax[0];
  }
}

__global__ void minmax_final(const double *partial, int nblk, double *result) {
  __shared__ double smin[256], smax[256];
  double mn = INFINITY, mx = -INFINITY;
  for (int t = threadIdx.x; t < nblk; t += blockDim.x) {
    mn = fmin(mn, partial[2 * t]);
    mx = fmax(mx, partial[2 * t + 1]);
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    result[0] = smin[0];
    result[1] = smax[0];
  }
}

// Periodic Gaussian blob exp(-r^2/(2 sigma^2)) (minimum-image distance) and
// constant velocity on a NeX x NeY x NeZ box, node coordinates as in
// set_element_position (mod_mesh.f90:332-353).
__global__ void init_gaussian(double *q, double *u, double *v, double *w,
                              const double *__restrict__ x1D, int Nq, int NeX, int NeY,
                              double hx, double hy, double hz, double cx, double cy,
                              double cz, double sigma, double ux, double uy, double uz,
                              int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int Np = Nq * Nq * Nq;
  const int64_t e = t / Np;
  const int nd = (int)(t - e * Np);
  const int64_t ie = e % NeX, je = (e / NeX) % NeY, ke = e / ((int64_t)NeX * NeY);
  const int a = nd % Nq, b = (nd / Nq) % Nq, c = nd / (Nq * Nq);
  const double x = (((double)(ie + 1)) - 0.5) * hx + (0.5 * hx) * x1D[a];
  const double y = (((double)(je + 1)) - 0.5) * hy + (0.5 * hy) * x1D[b];
  const double z = (((double)(ke + 1)) - 0.5) * hz + (0.5 * hz) * x1D[c];
  double r2 = 0.0, d;
  d = x - cx; d = d - rint(d); r2 += d * d;
  d = y - cy; d = d - rint(d); r2 += d * d;
  d = z - cz; d = d - rint(d); r2 += d * d;
  q[t] = exp(-r2 / (2.0 * sigma * sigma));
  if (u) {  // no velocity fields when the launches read the constant itself
    u[t] = ux;
    v[t] = uy;
    w[t] = uz;
  }
}

__global__ void to_f32(const double *__restrict__ src, float *__restrict__ dst,
                       int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (float)src[t];
}

unsigned grid1d(int64_t n, int threads) {
  return (unsigned)((n + threads - 1) / threads);
}

}  // namespace

// FP64 tile shapes: candidate 0 is the VALU kernel with q, u, v, w staged in
// LDS, 1 the VALU kernel with v, w in registers (VREG), 2.. the FP64 MFMA
// kernel's shapes (kernels_mfma.hip); above p = 15 only the MFMA shapes.  The
// MFMA kernel computes the flux itself and applies the separable Lift, so
// launches that pass a flux (cal_dqdt) or need the stored Lift take the VALU
// kernels (stage_f64 / stage_f64_lift).
constexpr int kValuShapes = 2;
// Default per Nq, from whole-step timings of every candidate on MI355X at
// about 1.7e7 DOF: the register VALU kernel at p = 4 and 6..13, the
// direction-by-direction MFMA kernel (the last candidate) at p = 14..19, the
// staged VALU kernel else (p = 3: 158 -> 153 us per launch once both read a
// launch-wide geometry record through the scalar cache).  At p = 16, 17 the
// one-element all-at-once MFMA kernel is candidate 0 (313 and 260 us per
// launch against 276 and 249 direction by direction: at Nq = 17 its GEMMs
// take two 16-wide column tiles).
constexpr int kF64Default[kMaxNq + 1] = {0, 0, 0, 0, 0, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3,
                                         1, 1, 0, 0};
// Selected FP64 tile shape per Nq (see stage_f64_shapes); atomic because a
// serving worker may launch while another thread selects.
std::atomic<int> g_f64_shape[kMaxNq + 1] = {0, 0, 0, 0, 0, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1,
                                            3, 3, 1, 1, 0, 0};
// Defaults of constant-velocity launches (batches, served requests), from the
// same timings with the velocity read from the geometry record: the register
// VALU kernel also at p = 2, 3 (185 -> 165 and 155 -> 140 us per launch).
// Only shapes that give the bits of the default above are taken, so a served
// request equals its solve on a plain solver.
constexpr int kF64DefaultCvel[kMaxNq + 1] = {0, 0, 0, 1, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3,
                                             1, 1, 0, 0};
std::atomic<int> g_f64_shape_cvel[kMaxNq + 1] = {0, 0, 0, 1, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1,
                                                 3, 3, 1, 1, 0, 0};

bool stage_f64_candidate_runs(int Nq, int cand, const StageArgs &a) {
  if (Nq < kMinNq || Nq > kMaxNqValu) return true;
  int mfma[kMaxTileShapes];
  const int nm = tile_f64_shapes(Nq, mfma, kMaxTileShapes);
  return !(cand == kValuShapes + nm && run_elements(Nq) > 0 && !run_applicable(Nq, a));
}

int launch_stage_f64(int Nq, const StageArgs &a, hipStream_t s) {
  if (a.Ne < 1) return DGK_OK;
  int cand = (Nq < kMinNq || Nq > kMaxNq) ? 0
             : a.tile >= 0                ? a.tile
                                          : (a.cvel ? g_f64_shape_cvel : g_f64_shape)[Nq].load(
                                                std::memory_order_relaxed);
  if (Nq > kMaxNqValu && Nq <= kMaxNq) {
    if (a.flux_in || !a.ell)
      return set_error(DGK_ERR_UNSUPPORTED, "above p = 15 the stage runs on the MFMA kernel "
                                            "only: no given flux, separable Lift only");
    return launch_tile_f64(Nq, cand, a, s);
  }
  int mfma[kMaxTileShapes];
  const int nm = tile_f64_shapes(Nq, mfma, kMaxTileShapes);
  if (cand == kValuShapes + nm && run_elements(Nq) > 0) {
    if (run_applicable(Nq, a)) return launch_run_f64(Nq, a, s);
    cand = 1;  // the register VALU kernel: the same arithmetic
  }
  if (cand >= kValuShapes && !a.flux_in && a.ell)
    return launch_tile_f64(Nq, cand - kValuShapes, a, s);
  const bool vreg = cand == 1;
  switch (Nq) {
#define DGK_VALU(N) \
  case N: return launch_nq<N>(a, s, vreg);
  DGK_VALU(2) DGK_VALU(3) DGK_VALU(4) DGK_VALU(5) DGK_VALU(6) DGK_VALU(7) DGK_VALU(8)
  DGK_VALU(9) DGK_VALU(10) DGK_VALU(11) DGK_VALU(12) DGK_VALU(13) DGK_VALU(14) DGK_VALU(15)
  DGK_VALU(16)
#undef DGK_VALU
  default: return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2.." + std::to_string(kMaxNq));
  }
}

int prepare_stage_f64(int Nq) {
  DGK_TRY(prepare_tile_f64(Nq));
  if (Nq > kMaxNqValu) return DGK_OK;
  switch (Nq) {
#define DGK_VALU(N) \
  case N: return prepare_nq<N>();
  DGK_VALU(2) DGK_VALU(3) DGK_VALU(4) DGK_VALU(5) DGK_VALU(6) DGK_VALU(7) DGK_VALU(8)
  DGK_VALU(9) DGK_VALU(10) DGK_VALU(11) DGK_VALU(12) DGK_VALU(13) DGK_VALU(14) DGK_VALU(15)
  DGK_VALU(16)
#undef DGK_VALU
  default: return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2.." + std::to_string(kMaxNq));
  }
}

int valu_mp_elements(int Nq) {
  switch (Nq) {
#define DGK_VALU_MP(N) \
  case N: return Staged<N, true>::EPB;
  DGK_VALU_MP(2) DGK_VALU_MP(3) DGK_VALU_MP(4) DGK_VALU_MP(5) DGK_VALU_MP(6) DGK_VALU_MP(7)
  DGK_VALU_MP(8)
#undef DGK_VALU_MP
  default: return 0;
  }
}

int launch_stage_valu_mp(int Nq, const StageArgsMP &a, hipStream_t s) {
  if (a.Ne < 1) return DGK_OK;
  if (!a.D64 || !a.ell64)
    return set_error(DGK_ERR_UNSUPPORTED, "the mixed VALU kernel needs the FP64 operators");
  const bool ax = a.axis_aligned != 0;
  switch (Nq) {
#define DGK_VALU_MP(N)                                                     \
  case N:                                                                  \
    if (a.cvel && ax && !a.vmapP) return launch_valu_mp<N, true, true>(a, s); \
    return ax ? launch_valu_mp<N, true>(a, s) : launch_valu_mp<N, false>(a, s);
  DGK_VALU_MP(2) DGK_VALU_MP(3) DGK_VALU_MP(4) DGK_VALU_MP(5) DGK_VALU_MP(6) DGK_VALU_MP(7)
  DGK_VALU_MP(8)
#undef DGK_VALU_MP
  default: return set_error(DGK_ERR_UNSUPPORTED, "no mixed VALU kernel at this order");
  }
}

int prepare_stage_valu_mp(int Nq) {
  switch (Nq) {
#define DGK_VALU_MP(N)                                          \
  case N:                                                       \
    DGK_TRY((prepare_shape_valu_mp<N, false>()));               \
    DGK_TRY((prepare_shape_valu_mp<N, true, true>()));          \
    return prepare_shape_valu_mp<N, true>();
  DGK_VALU_MP(2) DGK_VALU_MP(3) DGK_VALU_MP(4) DGK_VALU_MP(5) DGK_VALU_MP(6) DGK_VALU_MP(7)
  DGK_VALU_MP(8)
#undef DGK_VALU_MP
  default: return DGK_OK;
  }
}

int stage_f64_shapes(int Nq, int *epb, int n) {
  if (!epb || n < kMaxTileShapes) return 0;
  if (Nq > kMaxNqValu) return tile_f64_shapes(Nq, epb, n - kValuShapes);
  switch (Nq) {
#define DGK_F64_SHAPE(N) \
  case N: epb[0] = Staged<N, false>::EPB; epb[1] = Staged<N, true>::EPB; break;
  DGK_F64_SHAPE(2) DGK_F64_SHAPE(3) DGK_F64_SHAPE(4) DGK_F64_SHAPE(5) DGK_F64_SHAPE(6)
  DGK_F64_SHAPE(7) DGK_F64_SHAPE(8) DGK_F64_SHAPE(9) DGK_F64_SHAPE(10) DGK_F64_SHAPE(11)
  DGK_F64_SHAPE(12) DGK_F64_SHAPE(13) DGK_F64_SHAPE(14) DGK_F64_SHAPE(15) DGK_F64_SHAPE(16)
#undef DGK_F64_SHAPE
  default: return 0;
  }
  const int nc = kValuShapes + tile_f64_shapes(Nq, epb + kValuShapes, n - kValuShapes);
  if (run_elements(Nq) > 0 && nc < n) {
    epb[nc] = run_elements(Nq);
    return nc + 1;
  }
  return nc;
}

int stage_f64_select(int Nq, int cand) {
  int e[kMaxTileShapes];
  const int nc = stage_f64_shapes(Nq, e, kMaxTileShapes);
  if (nc < 1) return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2.." + std::to_string(kMaxNq));
  if (cand < -1 || cand >= nc)
    return set_error(DGK_ERR_ARG, "tile-shape candidate outside 0.." + std::to_string(nc - 1));
  g_f64_shape[Nq].store(cand == -1 ? kF64Default[Nq] : cand, std::memory_order_relaxed);
  g_f64_shape_cvel[Nq].store(cand == -1 ? kF64DefaultCvel[Nq] : cand, std::memory_order_relaxed);
  return DGK_OK;
}

int stage_f64_selected(int Nq) {
  return (Nq < kMinNq || Nq > kMaxNq) ? 0 : g_f64_shape[Nq].load(std::memory_order_relaxed);
}

int stage_f64_selected_cvel(int Nq) {
  return (Nq < kMinNq || Nq > kMaxNq) ? 0 : g_f64_shape_cvel[Nq].load(std::memory_order_relaxed);
}

int launch_elembnd_flux_f64(int Nq, const double *q, const double *u,
                            const double *v, const double *w,
                            const int32_t *nbr, const int32_t *vmapP,
                            const double *geo, const uint32_t *geoid,
                            int64_t Ne, double *flux, hipStream_t s) {
  const int64_t n = (int64_t)6 * Nq * Nq * Ne;
  if (n == 0) return DGK_OK;
  hipLaunchKernelGGL(elembnd_flux_f64, dim3(grid1d(n, 256)), dim3(256), 0, s, q, u,
                     v, w, nbr, vmapP, geo, geoid, Nq, Ne, flux);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

int launch_halo_gather(double *f, const int32_t *src, int64_t base, int64_t n,
                       hipStream_t s) {
  if (n == 0) return DGK_OK;
  hipLaunchKernelGGL(halo_gather, dim3(grid1d(n, 256)), dim3(256), 0, s, f, src, base, n);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

int launch_rk_update(double *q, const double *q0, const double *dqdt, double a,
                     double b, double dt, int64_t n, hipStream_t s) {
  if (n == 0) return DGK_OK;
  hipLaunchKernelGGL(rk_update, dim3(grid1d(n, 256)), dim3(256), 0, s, q, q0, dqdt,
                     a, b, dt, n);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

int launch_minmax(const double *q, int64_t n, double *partial, int nblk,
                  double *result, hipStream_t s) {
  hipLaunchKernelGGL(minmax_partial, dim3(nblk), dim3(256), 0, s, q, n, partial);
  DGK_HIP(hipGetLastError());
  hipLaunchKernelGGL(minmax_final, dim3(1), dim3(256), 0, s, partial, nblk, result);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

// ---- min/max without allocation (dgk_minmax_noalloc) ----------------------
// The per-block partials live in `scratch`, a file-scope device array of
// kNoallocBlocks (min, max) pairs = 16 KiB.  Every GPU the library runs on
// holds its own copy until the process ends.  In host code the array's name
// is not a device address: the address of each device's copy is looked up
// once per device (hipGetSymbolAddress, with that device current) and only
// that address is handed to kernels.  A call writes the pairs of the blocks
// it launches before its final fold reads exactly those pairs, so nothing has
// to start at zero; calls that overlap on one GPU share the array and race.
static __device__ __align__(256) double scratch[2 * kNoallocBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device
}  // namespace

static int scratch_address(double **out) {
  int dev = 0;
  DGK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(DGK_ERR_DEVICE, "device index " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    double *p = nullptr;
    DGK_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(DGK_ERR_DEVICE, "hipGetSymbolAddress(scratch) gave NULL");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return DGK_OK;
}

int launch_minmax_noalloc(const double *q, int64_t n, double *result, hipStream_t s) {
  double *partial = nullptr;
  DGK_TRY(scratch_address(&partial));
  // one block per 2048 values, at most kNoallocBlocks; blocks stride over the rest
  const int64_t want = (n + 2047) / 2048;
  const int nblk = (int)std::min<int64_t>(kNoallocBlocks, std::max<int64_t>(1, want));
  return launch_minmax(q, n, partial, nblk, result, s);
}

int launch_init_gaussian(double *q, double *u, double *v, double *w, const double *x1D,
                         int Nq, int NeX, int NeY, int NeZ, const double h[3],
                         const double c[3], double sigma, const double vel[3],
                         hipStream_t s) {
  const int64_t n = (int64_t)NeX * NeY * NeZ * Nq * Nq * Nq;
  if (n == 0) return DGK_OK;
  hipLaunchKernelGGL(init_gaussian, dim3(grid1d(n, 256)), dim3(256), 0, s, q, u, v, w, x1D, Nq,
                     NeX, NeY, h[0], h[1], h[2], c[0], c[1], c[2], sigma, vel[0], vel[1],
                     vel[2], n);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

int launch_to_f32(const double *src, float *dst, int64_t n, hipStream_t s) {
  if (n == 0) return DGK_OK;
  hipLaunchKernelGGL(to_f32, dim3(grid1d(n, 256)), dim3(256), 0, s, src, dst, n);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

}  // namespace dgk
