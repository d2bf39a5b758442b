// Mixed-precision stage kernel (gfx950): the 1D-operator contractions and the
// Lift run in FP32 on MFMA with FP32 accumulation; the state q, the face-flux
// jumps, the volume fluxes q*u_d and the RK stage combination stay FP64.
// Velocities are read from FP32 copies.
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
//   1. the lane loads its q column (FP64), its u, v, w columns (FP32), the
//      own traces of the x/y faces and the neighbour traces straight into
//      registers (coalesced along (i,j));
//   2. FP64 upwind fluxes and FP64 volume fluxes E_d q u_d, each rounded once
//      to FP32, are written into three augmented A arrays in LDS, one pencil
//      per row of KPAD floats;
//   3. v_mfma_f32_16x16x4f32 over 16-pencil tiles; K is permuted so that lane
//      (col, kq) keeps rows l = kq*KSTEPS + ks of B in registers and reads A
//      at fixed offsets from its own base; the result overwrites the tile's
//      rows in place;
//   4. the three direction results are summed in FP64, dqdt = -((X + Y) + Z),
//      and RK-combined in FP64 (main.f90:69-73).
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

// Tile shape of the mixed-precision stage kernel: EPB elements per block.
template <int NQ, int EPB_>
struct MP {
  static constexpr int NQ2 = NQ * NQ;
  static constexpr int NP = NQ2 * NQ;
  static constexpr int EPB = EPB_;
  static constexpr int LANES = EPB * NQ2;                   // node columns per block
  static constexpr int THREADS = ((LANES + 63) / 64) * 64;
  static constexpr int WAVES = THREADS / 64;
  static constexpr int ROWS = ((LANES + 15) / 16) * 16;     // pencils per direction
  static constexpr int MTILES = ROWS / 16;
  static constexpr int KAUG = NQ + 2;
  static constexpr int KSTEPS = (KAUG + 3) / 4;
  static constexpr int KPAD = 4 * KSTEPS;
  static constexpr int AREA = ROWS * KPAD;                  // floats per direction
  static constexpr size_t LDS_BYTES = (size_t)3 * AREA * 4;
  static constexpr int JOBS = 3 * MTILES;                   // (direction, tile) GEMMs
  static constexpr int JPW = (JOBS + WAVES - 1) / WAVES;
};

// Candidate tile shapes per order (index 0 is the default): one, two or four
// wavefronts of whole elements for Nq^2 <= 64, one or two elements above.
template <int NQ>
constexpr int mp_epb(int cand) {
  constexpr int NQ2 = NQ * NQ;
  if (NQ2 <= 64) {
    const int waves = cand == 0 ? 1 : cand == 1 ? 2 : 4;
    const int e = 64 * waves / NQ2;
    return e < 1 ? 1 : e;
  }
  return cand == 0 ? 1 : 2;
}

template <int NQ, int EPB>
__global__ void __launch_bounds__((MP<NQ, EPB>::THREADS))
stage_mp(const double *__restrict__ q_, const double *q0_, const float *__restrict__ u_,
         const float *__restrict__ v_, const float *__restrict__ w_, double *out_,
         const int32_t *__restrict__ nbr_, const int32_t *__restrict__ vmapP_,
         const float *__restrict__ D_, const float *__restrict__ ell_,
         const double *__restrict__ geo_, const uint32_t *__restrict__ geoid_,
         const int32_t *__restrict__ order_, const int64_t Ne_, const double ca_,
         const double cb_, const double dt_, const int mode_, const int axis_aligned_,
         const int dt_geo_) {
  using C = MP<NQ, EPB>;
  constexpr int NQ2 = C::NQ2, NP = C::NP, NFT = 6 * NQ2, KPAD = C::KPAD, KSTEPS = C::KSTEPS;
  constexpr int AREA = C::AREA;
  extern __shared__ float smem_mp[];
  float *const Ax = smem_mp, *const Ay = smem_mp + AREA, *const Az = smem_mp + 2 * AREA;

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
  const double *g;
  if (EPB == 1) {
    e = slot < Ne_ ? (order_ ? (int64_t)order_[slot] : slot) : 0;
    e = uniform_mp(e);
    g = geo_ + (geoid_ ? (int64_t)geoid_[e] * kGeoStride : 0);
  } else {
    e = active ? (order_ ? (int64_t)order_[slot] : slot) : 0;
    g = geo_ + (active && geoid_ ? (int64_t)geoid_[e] * kGeoStride : 0);
  }
  const int64_t ebase = e * NP;
  const int64_t base = ebase + ij;
  const bool rk = mode_ == 1;
  const bool use_q0 = rk && ca_ != 0.0;
  const bool axal = axis_aligned_ != 0;

  // ---- 1. loads straight into registers ------------------------------------
  double qc[NQ];
  float uc[NQ], vc[NQ], wc[NQ];
  double qM[4], qP[6];
  float vM[4], vP[6];  // velocity component along the face axis
  if (active) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      qc[k] = q_[base + k * NQ2];
      uc[k] = u_[base + k * NQ2];
      vc[k] = v_[base + k * NQ2];
      wc[k] = w_[base + k * NQ2];
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {  // own traces of the x and y faces
      const int64_t m = ebase + own_face_node(f, i, j, NQ);
      qM[f] = q_[m];
      vM[f] = (kFaceAxis[f] == 0 ? u_ : v_)[m];
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int64_t pidx = vmapP_ ? (int64_t)vmapP_[e * NFT + f * NQ2 + ij]
                                  : (int64_t)nbr_[e * 6 + f] * NP + nbr_face_node(f, i, j, NQ);
      qP[f] = q_[pidx];
      vP[f] = (kFaceAxis[f] == 0 ? u_ : kFaceAxis[f] == 1 ? v_ : w_)[pidx];
    }
  }

  // ---- 2. FP64 fluxes and volume fluxes -> augmented A rows (FP32) ---------
  if (active) {
    float fl[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const double qm = f == 4 ? qc[0] : f == 5 ? qc[NQ - 1] : qM[f];
      double VelM, VelP;
      if (axal) {
        const int ax = kFaceAxis[f];
        const double na = g[kGeoNormal + 3 * f + ax];
        const float vm = f == 4 ? wc[0] : f == 5 ? wc[NQ - 1] : vM[f];
        VelM = (double)vm * na;
        VelP = (double)vP[f] * na;
      } else {
        // general normals: all three components on both sides
        const int64_t m = ebase + own_face_node(f, i, j, NQ);
        const int64_t pidx = vmapP_ ? (int64_t)vmapP_[e * NFT + f * NQ2 + ij]
                                    : (int64_t)nbr_[e * 6 + f] * NP + nbr_face_node(f, i, j, NQ);
        const double n1 = g[kGeoNormal + 3 * f], n2 = g[kGeoNormal + 3 * f + 1],
                     n3 = g[kGeoNormal + 3 * f + 2];
        VelM = (double)u_[m] * n1 + (double)v_[m] * n2 + (double)w_[m] * n3;
        VelP = (double)u_[pidx] * n1 + (double)v_[pidx] * n2 + (double)w_[pidx] * n3;
      }
      const double alpha = 0.5 * fabs(VelP + VelM);
      fl[f] = (float)(0.5 * g[kGeoFscale + f] * (qP[f] * VelP - qm * VelM - alpha * (qP[f] - qm)));
    }
    const double Ex = g[kGeoEscale], Ey = g[kGeoEscale + 1], Ez = g[kGeoEscale + 2];
    const int R = el * NQ2;
    float *ax = Ax + (R + j) * KPAD + i;  // node (i,j,k): x pencil (j,k), position i
    float *ay = Ay + (R + i) * KPAD + j;  //               y pencil (i,k), position j
    float *az = Az + (R + ij) * KPAD;     //               z pencil (i,j), position k
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      ax[k * NQ * KPAD] = (float)(Ex * (qc[k] * (double)uc[k]));
      ay[k * NQ * KPAD] = (float)(Ey * (qc[k] * (double)vc[k]));
      az[k] = (float)(Ez * (qc[k] * (double)wc[k]));
    }
    // face node (i,j) of the faces normal to d is pencil ij of direction d
    float *rx = Ax + (R + ij) * KPAD, *ry = Ay + (R + ij) * KPAD;
    rx[NQ] = fl[3];      // -x
    rx[NQ + 1] = fl[1];  // +x
    ry[NQ] = fl[0];      // -y
    ry[NQ + 1] = fl[2];  // +y
    az[NQ] = fl[4];      // -z
    az[NQ + 1] = fl[5];  // +z
#pragma unroll
    for (int l = NQ + 2; l < KPAD; ++l) {
      rx[l] = 0.f;
      ry[l] = 0.f;
      az[l] = 0.f;
    }
  }
  // RK base column, in flight during the GEMMs
  double q0v[NQ];
  if (active && use_q0) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) q0v[k] = q0_[base + k * NQ2];
  }
  __syncthreads();

  // ---- 3. MFMA GEMMs over (direction, 16-pencil tile) jobs -----------------
  // Rows of A beyond the block's active pencils hold stale LDS; each output
  // row depends only on its own A row, and those rows are never read back.
  {
    const int col = lane & 15, kq = lane >> 4;
    float bop[KSTEPS];
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks) {
      const int l = kq * KSTEPS + ks;
      bop[ks] = col >= NQ ? 0.f : l < NQ ? D_[col + l * NQ] : l == NQ ? ell_[col]
                : l == NQ + 1 ? ell_[NQ + col] : 0.f;
    }
    const int rd = col * KPAD + kq * KSTEPS;  // A read offset of this lane in a tile
    const int wr = 4 * kq * KPAD + col;       // result offset (tile rows 4kq + t4)
#pragma unroll
    for (int jj = 0; jj < C::JPW; ++jj) {
      const int job = wave + jj * C::WAVES;
      if (job >= C::JOBS) break;  // wave-uniform
      const int d = job / C::MTILES, t = job - d * C::MTILES;
      float *tile = smem_mp + d * AREA + t * 16 * KPAD;
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KSTEPS; ++ks)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(tile[rd + ks], bop[ks], acc, 0, 0, 0);
      if (col < NQ) {
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) tile[wr + t4 * KPAD] = acc[t4];
      }
    }
  }
  __syncthreads();

  // ---- 4. FP64 combine and RK epilogue -------------------------------------
  if (active) {
    const int R = el * NQ2;
    const float *ax = Ax + (R + j) * KPAD + i;
    const float *ay = Ay + (R + i) * KPAD + j;
    const float *az = Az + (R + ij) * KPAD;
    const double dtv = dt_geo_ ? g[kGeoDt] : dt_;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const double tend =
          -(((double)ax[k * NQ * KPAD] + (double)ay[k * NQ * KPAD]) + (double)az[k]);
      const int64_t idx = base + (int64_t)k * NQ2;
      if (!rk) {
        out_[idx] = tend;
      } else {
        const double upd = cb_ * (qc[k] + dtv * tend);
        out_[idx] = use_q0 ? ca_ * q0v[k] + upd : upd;
      }
    }
  }
}

template <int NQ, int EPB>
int prepare_mp_shape() {
  static int prepared_dev = -1;
  int dev = 0;
  DGK_HIP(hipGetDevice(&dev));
  if (prepared_dev != dev) {
    DGK_HIP(hipFuncSetAttribute((const void *)stage_mp<NQ, EPB>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)MP<NQ, EPB>::LDS_BYTES));
    prepared_dev = dev;
  }
  return DGK_OK;
}

template <int NQ, int EPB>
int launch_mp_shape(const StageArgsMP &a, hipStream_t s) {
  using T = MP<NQ, EPB>;
  const int64_t nblk = (a.Ne + T::EPB - 1) / T::EPB;
  if (nblk > INT32_MAX) return set_error(DGK_ERR_UNSUPPORTED, "too many elements");
  DGK_TRY((prepare_mp_shape<NQ, EPB>()));
  hipLaunchKernelGGL((stage_mp<NQ, EPB>), dim3((unsigned)nblk), dim3(T::THREADS),
                     T::LDS_BYTES, s, a.q, a.q0, a.u, a.v, a.w, a.out, a.nbr, a.vmapP, a.D,
                     a.ell, a.geo, a.geoid, a.order, a.Ne, a.ca, a.cb, a.dt, a.mode,
                     a.axis_aligned, a.dt_geo);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

// Tile-shape candidate selected per Nq (index into mp_epb).
int g_mp_shape[kMaxNq + 1] = {0};

template <int NQ>
int prepare_mp_nq() {
  DGK_TRY((prepare_mp_shape<NQ, mp_epb<NQ>(0)>()));
  DGK_TRY((prepare_mp_shape<NQ, mp_epb<NQ>(1)>()));
  DGK_TRY((prepare_mp_shape<NQ, mp_epb<NQ>(2)>()));
  return DGK_OK;
}

template <int NQ>
int launch_mp_nq(const StageArgsMP &a, hipStream_t s) {
  switch (g_mp_shape[NQ]) {
  case 1: return launch_mp_shape<NQ, mp_epb<NQ>(1)>(a, s);
  case 2: return launch_mp_shape<NQ, mp_epb<NQ>(2)>(a, s);
  default: return launch_mp_shape<NQ, mp_epb<NQ>(0)>(a, s);
  }
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

int stage_mp_shapes(int Nq, int *epb, int n) {
  if (Nq < kMinNq || Nq > kMaxNq || !epb || n < 3) return 0;
  switch (Nq) {
#define DGK_MP_SHAPES(N) \
  case N: epb[0] = mp_epb<N>(0); epb[1] = mp_epb<N>(1); epb[2] = mp_epb<N>(2); break;
  DGK_MP_SHAPES(2) DGK_MP_SHAPES(3) DGK_MP_SHAPES(4) DGK_MP_SHAPES(5) DGK_MP_SHAPES(6)
  DGK_MP_SHAPES(7) DGK_MP_SHAPES(8) DGK_MP_SHAPES(9) DGK_MP_SHAPES(10) DGK_MP_SHAPES(11)
  DGK_MP_SHAPES(12) DGK_MP_SHAPES(13) DGK_MP_SHAPES(14) DGK_MP_SHAPES(15) DGK_MP_SHAPES(16)
#undef DGK_MP_SHAPES
  }
  return Nq * Nq <= 64 ? 3 : 2;
}

int stage_mp_select(int Nq, int cand) {
  if (Nq < kMinNq || Nq > kMaxNq) return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2..16");
  const int nc = Nq * Nq <= 64 ? 3 : 2;
  if (cand < 0 || cand >= nc)
    return set_error(DGK_ERR_ARG, "tile-shape candidate outside 0.." + std::to_string(nc - 1));
  g_mp_shape[Nq] = cand;
  return DGK_OK;
}

int stage_mp_selected(int Nq) {
  return (Nq < kMinNq || Nq > kMaxNq) ? 0 : g_mp_shape[Nq];
}

}  // namespace dgk
