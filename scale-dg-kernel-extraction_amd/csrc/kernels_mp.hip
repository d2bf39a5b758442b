// Mixed-precision stage kernel (gfx950): the 1D-operator contractions and the
// Lift run in FP32 on MFMA with FP32 accumulation; the state q, the face-flux
// jumps and the RK stage combination stay FP64.  Velocities are read from
// FP32 copies.
//
// For each direction d the sum-factorised derivative and that direction's two
// Lift faces form one GEMM over the element's NQ*NQ "pencils" r (the two
// indices other than d):
//   out_d(r, c) = sum_l F_d(r, l) * E_d * D(c, l)
//               + l-(c) * flux(face d-, r) + l+(c) * flux(face d+, r)
// i.e. A_d = [F_d | flux_d- | flux_d+]  (pencils x (NQ+2)),
//      B_d = [E_d D^T ; l- ; l+]        ((NQ+2) x NQ).
// The reference's face-node order (Fmask, mod_mesh.f90:278-324) is exactly
// the pencil order of the face's normal direction, so the Lift
// (tensorprod_Lift_hexahedral, mod_dg_optr_kernel.f90:262-288, in its
// tensor-product form) is two extra K columns of the contraction
// (tensorprod_divlike_dirXYZ, mod_dg_optr_kernel.f90:197-258).  The GEMMs use
// v_mfma_f32_16x16x4f32: M = pencils (16 per tile), K = NQ+2 (4 per step),
// N = NQ (one 16-column tile).  The three direction results are summed in
// FP64: dqdt = -((X + Y) + Z), then RK-combined in FP64 (main.f90:69-73).
#include <algorithm>

#include "dgk_internal.hpp"

namespace dgk {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

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

template <int NQ>
struct MP {
  static constexpr int NQ2 = NQ * NQ;
  static constexpr int NP = NQ2 * NQ;
  static constexpr int EPB = NQ2 <= 64 ? 64 / NQ2 : 1;
  static constexpr int THREADS = NQ2 <= 64 ? 64 : ((NQ2 + 63) / 64) * 64;
  static constexpr int WAVES = THREADS / 64;
  static constexpr int PPAD = ((NQ2 + 15) / 16) * 16;  // pencils per element, whole M tiles
  static constexpr int MTILES = EPB * PPAD / 16;        // M tiles per direction
  static constexpr int KAUG = NQ + 2;
  static constexpr int KSTEPS = (KAUG + 3) / 4;
  // async copy granules: q (FP64) 16 B when NP is even, else 4 B;
  // u, v, w (FP32) 16 B when NP % 4 == 0, else 4 B
  static constexpr bool QV16 = (NP % 2) == 0;
  static constexpr int QCHUNK = QV16 ? 128 : 32;  // doubles per wave instruction
  static constexpr int QNCH = (EPB * NP + QCHUNK - 1) / QCHUNK;
  static constexpr int QFIELD = QNCH * QCHUNK;     // doubles
  static constexpr bool FV16 = (NP % 4) == 0;
  static constexpr int FCHUNK = FV16 ? 256 : 64;   // floats per wave instruction
  static constexpr int FNCH = (EPB * NP + FCHUNK - 1) / FCHUNK;
  static constexpr int FFIELD = FNCH * FCHUNK;     // floats
  // LDS layout in bytes
  static constexpr int OFF_Q = 0;
  static constexpr int OFF_U = OFF_Q + 8 * QFIELD;
  static constexpr int OFF_V = OFF_U + 4 * FFIELD;
  static constexpr int OFF_W = OFF_V + 4 * FFIELD;
  static constexpr int OFF_G = OFF_W + 4 * FFIELD;          // [EPB][6][PPAD] floats
  static constexpr int OFF_D = OFF_G + 4 * EPB * 6 * PPAD;  // [NQ2] floats
  static constexpr int OFF_L = OFF_D + 4 * NQ2;             // [2*NQ] floats
  static constexpr int OFF_Z = OFF_L + 4 * 2 * NQ;          // one 0.0f (padding operand)
  static constexpr size_t LDS_BYTES = OFF_Z + 16;
  static constexpr int JOBS = 3 * MTILES;                   // (direction, M tile) GEMM tiles
  static constexpr int JPW = (JOBS + WAVES - 1) / WAVES;    // tiles per wavefront
};

// node of pencil r, position l along direction d (0 = x, 1 = y, 2 = z)
template <int NQ>
__device__ __forceinline__ int pencil_node(int d, int r, int l) {
  constexpr int NQ2 = NQ * NQ;
  return d == 0 ? l + r * NQ : d == 1 ? (r % NQ) + l * NQ + (r / NQ) * NQ2 : r + l * NQ2;
}

template <int NQ>
__global__ void __launch_bounds__(MP<NQ>::THREADS)
stage_mp(const double *__restrict__ q_, const double *q0_, const float *__restrict__ u_,
         const float *__restrict__ v_, const float *__restrict__ w_, double *out_,
         const int32_t *__restrict__ nbr_, const int32_t *__restrict__ vmapP_,
         const float *__restrict__ D_, const float *__restrict__ ell_,
         const double *__restrict__ geo_, const uint32_t *__restrict__ geoid_,
         const int32_t *__restrict__ order_, const int64_t Ne_, const double ca_,
         const double cb_, const double dt_, const int mode_, const int axis_aligned_, const int dt_geo_) {
  using C = MP<NQ>;
  constexpr int NQ2 = C::NQ2, NP = C::NP, EPB = C::EPB, NFT = 6 * NQ2, PPAD = C::PPAD;
  extern __shared__ double smem_mp[];
  char *sm = (char *)smem_mp;
  double *sq = (double *)(sm + C::OFF_Q);
  float *sf[3] = {(float *)(sm + C::OFF_U), (float *)(sm + C::OFF_V), (float *)(sm + C::OFF_W)};
  float *sG = (float *)(sm + C::OFF_G), *sD = (float *)(sm + C::OFF_D), *sL = (float *)(sm + C::OFF_L);

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int64_t slot0 = (int64_t)xcd_block_mp(blockIdx.x, gridDim.x) * EPB;

  auto elem_at = [&](int del) -> int64_t {  // element of staged block `del` (0 if none)
    const int64_t sl = slot0 + del;
    return (del < EPB && sl < Ne_) ? (order_ ? (int64_t)order_[sl] : sl) : -1;
  };

  // ---- 1. async HBM -> LDS copies: q (FP64), u, v, w (FP32) ----------------
#pragma unroll
  for (int c0 = 0; c0 < C::QNCH; c0 += C::WAVES) {
    const int c = c0 + wave;
    if (c < C::QNCH) {
      const int d = c * C::QCHUNK + (C::QV16 ? 2 * lane : lane / 2);
      const int del = d / NP;
      const int64_t ee = elem_at(del);
      const int64_t off = ee >= 0 ? ee * NP + (d - del * NP) : 0;
      const char *gp = (const char *)(q_ + off) + (C::QV16 ? 0 : 4 * (lane & 1));
      __attribute__((address_space(3))) void *lp =
          (__attribute__((address_space(3))) void *)(sq + c * C::QCHUNK);
#if defined(__HIP_DEVICE_COMPILE__)
      __builtin_amdgcn_global_load_lds((const void *)gp, lp, C::QV16 ? 16 : 4, 0, 0);
#else
      (void)gp;
      (void)lp;
#endif
    }
  }
  {
    const float *src[3] = {u_, v_, w_};
#pragma unroll
    for (int c0 = 0; c0 < C::FNCH; c0 += C::WAVES) {
      const int c = c0 + wave;
      if (c < C::FNCH) {
        const int d = c * C::FCHUNK + (C::FV16 ? 4 * lane : lane);
        const int del = d / NP;
        const int64_t ee = elem_at(del);
        const int64_t off = ee >= 0 ? ee * NP + (d - del * NP) : 0;
#pragma unroll
        for (int fld = 0; fld < 3; ++fld) {
          const float *gp = src[fld] + off;
          __attribute__((address_space(3))) void *lp =
              (__attribute__((address_space(3))) void *)(sf[fld] + c * C::FCHUNK);
#if defined(__HIP_DEVICE_COMPILE__)
          __builtin_amdgcn_global_load_lds((const void *)gp, lp, C::FV16 ? 16 : 4, 0, 0);
#else
          (void)gp;
          (void)lp;
#endif
        }
      }
    }
  }
  for (int t = tid; t < NQ2; t += C::THREADS) sD[t] = D_[t];
  for (int t = tid; t < 2 * NQ; t += C::THREADS) sL[t] = ell_[t];
  if (tid == 0) *(float *)(sm + C::OFF_Z) = 0.0f;

  const bool lane_ok = tid < EPB * NQ2;
  const int el = lane_ok ? tid / NQ2 : 0;
  const int ij = lane_ok ? tid - el * NQ2 : 0;
  const int i = ij % NQ, j = ij / NQ;
  const int64_t slot = slot0 + el;
  const bool active = lane_ok && slot < Ne_;
  int64_t e;
  const double *g;
  if (EPB == 1) {
    e = slot < Ne_ ? (order_ ? (int64_t)order_[slot] : slot) : 0;
    e = uniform_mp(e);
    g = geo_ + (geoid_ ? (int64_t)geoid_[e] * kGeoStride : 0);
  } else {
    e = active ? (order_ ? (int64_t)order_[slot] : slot) : 0;
    g = geo_ + (active && geoid_ ? (int64_t)geoid_[e] * kGeoStride : 0);
  }
  const int64_t base = e * NP + ij;
  const int lb = el * NP;
  const bool rk = mode_ == 1;
  const bool use_q0 = rk && ca_ != 0.0;
  const bool axal = axis_aligned_ != 0;

  // neighbour traces (FP64 q, FP32 velocity) and the RK base column
  double qP[6];
  float uP[6], vP[6], wP[6];
  double q0v[NQ];
  if (active) {
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int64_t pidx = vmapP_ ? (int64_t)vmapP_[e * NFT + f * NQ2 + ij]
                                  : (int64_t)nbr_[e * 6 + f] * NP + nbr_face_node(f, i, j, NQ);
      qP[f] = q_[pidx];
      if (axal) {
        const float *vel = kFaceAxis[f] == 0 ? u_ : kFaceAxis[f] == 1 ? v_ : w_;
        uP[f] = vel[pidx];
      } else {
        uP[f] = u_[pidx];
        vP[f] = v_[pidx];
        wP[f] = w_[pidx];
      }
    }
    if (use_q0) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) q0v[k] = q0_[base + (int64_t)k * NQ2];
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  // ---- 2. FP64 upwind flux of face node (i,j) on each face -> FP32 ---------
  if (active) {
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int m = lb + own_face_node(f, i, j, NQ);
      const double qM = sq[m];
      double VelM, VelP;
      if (axal) {
        const int ax = kFaceAxis[f];
        const double na = g[kGeoNormal + 3 * f + ax];
        VelM = (double)sf[ax][m] * na;
        VelP = (double)uP[f] * na;
      } else {
        const double n1 = g[kGeoNormal + 3 * f], n2 = g[kGeoNormal + 3 * f + 1],
                     n3 = g[kGeoNormal + 3 * f + 2];
        VelM = (double)sf[0][m] * n1 + (double)sf[1][m] * n2 + (double)sf[2][m] * n3;
        VelP = (double)uP[f] * n1 + (double)vP[f] * n2 + (double)wP[f] * n3;
      }
      const double alpha = 0.5 * fabs(VelP + VelM);
      const double flux =
          0.5 * g[kGeoFscale + f] * (qP[f] * VelP - qM * VelM - alpha * (qP[f] - qM));
      sG[(el * 6 + f) * PPAD + ij] = (float)flux;
    }
  }
  __syncthreads();

  // ---- 3. volume fluxes F_d = q u_d in place (FP32), q column kept --------
  double qc[NQ];
  if (active) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const int n = lb + ij + k * NQ2;
      const double qq = sq[n];
      qc[k] = qq;
#pragma unroll
      for (int d = 0; d < 3; ++d) sf[d][n] = (float)(qq * (double)sf[d][n]);
    }
  }
  __syncthreads();

  // ---- 4. MFMA GEMMs over (direction, 16-pencil tile) jobs ---------------
  // B = [E_d D^T ; l- ; l+] is the same for every direction up to E_d on its
  // first NQ rows: lane (col = lane%16, k = lane/16) keeps its raw entries.
  {
    const int row = lane & 15, kq = lane >> 4;
    float braw[C::KSTEPS];
    bool bscaled[C::KSTEPS];
#pragma unroll
    for (int ks = 0; ks < C::KSTEPS; ++ks) {
      const int l = 4 * ks + kq;
      braw[ks] = row >= NQ ? 0.f : l < NQ ? sD[row + l * NQ] : l == NQ ? sL[row]
                                 : l == NQ + 1 ? sL[NQ + row] : 0.f;
      bscaled[ks] = l < NQ;
    }
    const float *lds = (const float *)sm;
    constexpr int ZERO = C::OFF_Z / 4;
    f32x4 acc[C::JPW];
#pragma unroll
    for (int jj = 0; jj < C::JPW; ++jj) {
      acc[jj] = f32x4{0.f, 0.f, 0.f, 0.f};
      const int job = wave + jj * C::WAVES;
      const int d = job / C::MTILES, t = job - d * C::MTILES;
      const int tel = (t * 16) / PPAD, rb = (t * 16) - tel * PPAD;
      if (job >= C::JOBS || slot0 + tel >= Ne_) continue;  // wave-uniform
      const double *tg = geo_;
      if (geoid_) {
        const int64_t ts = slot0 + tel;
        tg = geo_ + (int64_t)geoid_[order_ ? (int64_t)order_[ts] : ts] * kGeoStride;
      }
      const float Ed = (float)tg[kGeoEscale + d];
      const int fm = d == 0 ? 3 : d == 1 ? 0 : 4;  // -x, -y, -z faces
      const int fp = d == 0 ? 1 : d == 1 ? 2 : 5;  // +x, +y, +z faces
      const int fbase = (C::OFF_U / 4) + d * C::FFIELD + tel * NP;
      const int gm = C::OFF_G / 4 + (tel * 6 + fm) * PPAD, gp = C::OFF_G / 4 + (tel * 6 + fp) * PPAD;
      const int r = rb + row;
#pragma unroll
      for (int ks = 0; ks < C::KSTEPS; ++ks) {
        const int l = 4 * ks + kq;
        int addr = ZERO;
        if (r < NQ2) addr = l < NQ ? fbase + pencil_node<NQ>(d, r, l) : l == NQ ? gm + r
                          : l == NQ + 1 ? gp + r : ZERO;
        const float a = lds[addr];
        const float b = bscaled[ks] ? Ed * braw[ks] : braw[ks];
        acc[jj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[jj], 0, 0, 0);
      }
    }
    // results (rows rb + 4*kq + t4, column row) go back in place: a tile
    // alone read those nodes of F_d
#pragma unroll
    for (int jj = 0; jj < C::JPW; ++jj) {
      const int job = wave + jj * C::WAVES;
      const int d = job / C::MTILES, t = job - d * C::MTILES;
      const int tel = (t * 16) / PPAD, rb = (t * 16) - tel * PPAD;
      if (job >= C::JOBS || slot0 + tel >= Ne_) continue;
#pragma unroll
      for (int t4 = 0; t4 < 4; ++t4) {
        const int ro = rb + 4 * kq + t4;
        if (row < NQ && ro < NQ2) sf[d][tel * NP + pencil_node<NQ>(d, ro, row)] = acc[jj][t4];
      }
    }
  }
  __syncthreads();

  // ---- 5. FP64 combine and RK epilogue -------------------------------------
  if (active) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const int n = lb + ij + k * NQ2;
      const double tend = -(((double)sf[0][n] + (double)sf[1][n]) + (double)sf[2][n]);
      const int64_t idx = base + (int64_t)k * NQ2;
      if (!rk) {
        out_[idx] = tend;
      } else {
        const double upd = cb_ * (qc[k] + (dt_geo_ ? g[kGeoDt] : dt_) * tend);
        out_[idx] = use_q0 ? ca_ * q0v[k] + upd : upd;
      }
    }
  }
}

template <int NQ>
int prepare_mp_nq() {
  static int prepared_dev = -1;
  int dev = 0;
  DGK_HIP(hipGetDevice(&dev));
  if (prepared_dev != dev) {
    DGK_HIP(hipFuncSetAttribute((const void *)stage_mp<NQ>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)MP<NQ>::LDS_BYTES));
    prepared_dev = dev;
  }
  return DGK_OK;
}

template <int NQ>
int launch_mp_nq(const StageArgsMP &a, hipStream_t s) {
  using T = MP<NQ>;
  const int64_t nblk = (a.Ne + T::EPB - 1) / T::EPB;
  if (nblk > INT32_MAX) return set_error(DGK_ERR_UNSUPPORTED, "too many elements");
  DGK_TRY(prepare_mp_nq<NQ>());
  hipLaunchKernelGGL(stage_mp<NQ>, dim3((unsigned)nblk), dim3(T::THREADS), T::LDS_BYTES, s,
                     a.q, a.q0, a.u, a.v, a.w, a.out, a.nbr, a.vmapP, a.D, a.ell, a.geo,
                     a.geoid, a.order, a.Ne, a.ca, a.cb, a.dt, a.mode, a.axis_aligned, a.dt_geo);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

}  // namespace

int prepare_stage_mp(int Nq) {
  switch (Nq) {
  case 2: return prepare_mp_nq<2>();
  case 3: return prepare_mp_nq<3>();
  case 4: return prepare_mp_nq<4>();
  case 5: return prepare_mp_nq<5>();
  case 6: return prepare_mp_nq<6>();
  case 7: return prepare_mp_nq<7>();
  case 8: return prepare_mp_nq<8>();
  case 9: return prepare_mp_nq<9>();
  case 10: return prepare_mp_nq<10>();
  case 11: return prepare_mp_nq<11>();
  case 12: return prepare_mp_nq<12>();
  case 13: return prepare_mp_nq<13>();
  case 14: return prepare_mp_nq<14>();
  case 15: return prepare_mp_nq<15>();
  case 16: return prepare_mp_nq<16>();
  default: return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2..16");
  }
}

int launch_stage_mp(int Nq, const StageArgsMP &a, hipStream_t s) {
  if (a.Ne < 1) return DGK_OK;
  switch (Nq) {
  case 2: return launch_mp_nq<2>(a, s);
  case 3: return launch_mp_nq<3>(a, s);
  case 4: return launch_mp_nq<4>(a, s);
  case 5: return launch_mp_nq<5>(a, s);
  case 6: return launch_mp_nq<6>(a, s);
  case 7: return launch_mp_nq<7>(a, s);
  case 8: return launch_mp_nq<8>(a, s);
  case 9: return launch_mp_nq<9>(a, s);
  case 10: return launch_mp_nq<10>(a, s);
  case 11: return launch_mp_nq<11>(a, s);
  case 12: return launch_mp_nq<12>(a, s);
  case 13: return launch_mp_nq<13>(a, s);
  case 14: return launch_mp_nq<14>(a, s);
  case 15: return launch_mp_nq<15>(a, s);
  case 16: return launch_mp_nq<16>(a, s);
  default: return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2..16");
  }
}

}  // namespace dgk
