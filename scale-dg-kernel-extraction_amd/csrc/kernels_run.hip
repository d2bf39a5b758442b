// FP64 stage kernel over runs of consecutive elements (p <= 7): stage_run.
//
// The VALU and MFMA stage kernels give each lane a node column and gather
// the six neighbour traces of its face node with one 8-byte load per face
// and field: at low order a wavefront spends most of its memory
// instructions on those gathers, each touching several cache lines for
// 512 useful bytes.  This kernel cuts the memory instructions per node:
//
//   * a workgroup advances T consecutive elements [e0, e0 + T) (about 512
//     nodes); their q, u, v, w blocks are one contiguous range per field,
//     copied HBM -> LDS with global_load_lds_dwordx4, thread t taking nodes
//     2t and 2t+1 (T is even or Np is, so the range is 16-byte aligned);
//   * a face whose neighbour is one of the tile's elements (in the
//     reference's numbering, element e's +x neighbour is e + 1) takes the
//     neighbour's traces from LDS; only the other faces' traces are gathered,
//     one face node per thread and field;
//   * the RK base is read and the result written 16 bytes per thread.
//
// Per node the arithmetic is that of stage_f64 (flux of mod_advect3d_eq.f90
// :121-139, contractions l ascending as tensorprod_divlike_dirXYZ_general,
// mod_dg_optr_kernel.f90:197-258, the separable Lift in the order of
// mod_dg_optr_kernel.f90:277-283, dqdt = -(((Ex Dx + Ey Dy) + Ez Dz) + Lift),
// mod_advect3d_eq.f90:194-198, and the RK update of main.f90:69-73), written
// as the same expressions, so the two kernels give the same bits.
//
// Requirements (checked by run_applicable on the host): the compact
// neighbour table, the separable Lift, no given flux, no order array, field
// pointers 16-byte aligned, one geometry record per tile, and velocity
// fields (launches with constant velocities run the VALU kernel).
#include <algorithm>

#include "dgk_internal.hpp"

namespace dgk {
namespace {

__device__ __forceinline__ int xcd_run(int b, int n) {
  // workgroups are dealt round-robin to the 8 XCDs: give each XCD a
  // contiguous range of tiles so neighbouring tiles share its L2
  const int per = n >> 3, rem = n & 7;
  const int x = b & 7, local = b >> 3;
  return x * per + min(x, rem) + local;
}

// normal axis of face f (-y, +x, +y, -x, -z, +z -> 1, 0, 1, 0, 2, 2)
__device__ __forceinline__ int face_axis(int f) { return f >= 4 ? 2 : ((f & 1) ? 0 : 1); }

typedef double dbl2 __attribute__((ext_vector_type(2)));

template <int NQ>
struct Run {
  static constexpr int NQ2 = NQ * NQ, NP = NQ2 * NQ, NFT = 6 * NQ2;
  // elements per workgroup: about 512 nodes, an even node count
  static constexpr int T = NQ == 2 ? 32 : NQ == 3 ? 16 : NQ == 4 ? 8 : NQ == 5 ? 4
                         : NQ <= 7 ? 2 : 1;
  static constexpr int TNP = T * NP;
  static_assert(TNP % 2 == 0 && (T % 2 == 0 || NP % 2 == 0), "16-byte aligned tiles");
  static constexpr int PAIRS = TNP / 2;                     // two nodes per thread
  static constexpr int THREADS = ((PAIRS + 63) / 64) * 64;  // one 16-byte chunk per thread
  static constexpr int FIELD = 2 * THREADS;                 // LDS doubles per staged field
  static constexpr int ITEMS = T * NFT;                     // face nodes of the tile
  static constexpr int R = (ITEMS + THREADS - 1) / THREADS; // face nodes per thread
  static constexpr int SLAB = R * THREADS;                  // LDS doubles per traced field
  static constexpr int NBR_REGS = (6 * T + 63) / 64;        // neighbour entries per lane
};

template <int NQ, bool AXAL>
__global__ void __launch_bounds__(Run<NQ>::THREADS)
stage_run(const double *__restrict__ q_, const double *q0_, const double *__restrict__ u_,
          const double *__restrict__ v_, const double *__restrict__ w_, double *out_,
          const int32_t *__restrict__ nbr_, const double *__restrict__ D_,
          const double *__restrict__ ell_, const double *__restrict__ geo_,
          const uint32_t *__restrict__ geoid_, const int64_t Ne_, const double ca_,
          const double cb_, const double dt_, const int mode_, const int dt_geo_,
          const int rev_, const int64_t zpt_) {
  using C = Run<NQ>;
  constexpr int NQ2 = C::NQ2, NP = C::NP, T = C::T, NFT = C::NFT, R = C::R;
  constexpr int NTR = AXAL ? 2 : 4;  // traced fields: q and the normal velocity, or q, u, v, w
  __shared__ double smem[4 * C::FIELD + NTR * C::SLAB + NQ2 + 2 * NQ + 24];
  double *sq = smem, *su = sq + C::FIELD, *sv = su + C::FIELD, *sw = sv + C::FIELD;
  // neighbour traces of faces whose neighbour lies outside the tile, per face
  // node item it = el*NFT + f*NQ2 + (a + b*NQ); the q slab then holds the flux
  double *tq = sw + C::FIELD;
  double *t1 = tq + C::SLAB;
  double *sD = tq + NTR * C::SLAB, *sL = sD + NQ2;
  // the tile's Fscale[6] and face normals [6][3] (geometry record words 3..26)
  double *sgeo = sL + 2 * NQ;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t ntile = (Ne_ + T - 1) / T;
  const int64_t tile = slot_element(xcd_run(blockIdx.x, gridDim.x), ntile, nullptr, rev_, zpt_);
  const int64_t e0 = tile * T;
  const int nel = (int)min((int64_t)T, Ne_ - e0);
  const int rem = nel * NP;  // nodes of this tile
  // one geometry record per tile (checked on the host): scalar loads
  const double *g = geo_ + (geoid_ ? (int64_t)geoid_[e0] * kGeoStride : 0);

  // ---- neighbour entries of the tile: entry c*64 + lane of every wave -----
  int32_t nbv[C::NBR_REGS];
#pragma unroll
  for (int c = 0; c < C::NBR_REGS; ++c) {
    const int x = c * 64 + lane;
    nbv[c] = x < 6 * nel ? nbr_[e0 * 6 + x] : -1;
  }

  // ---- own fields: 16-byte async copies, thread t -> nodes 2t, 2t+1 ---------
  const bool full = 2 * tid + 1 < rem;  // both nodes in the tile
  const bool half = 2 * tid + 1 == rem; // node 2t only (odd node count)
  const int64_t own = e0 * NP + 2 * tid;
  {
    const double *src[4] = {q_, u_, v_, w_};
#pragma unroll
    for (int fld = 0; fld < 4; ++fld) {
      if (full) {
        __attribute__((address_space(3))) void *lp =
            (__attribute__((address_space(3))) void *)(smem + fld * C::FIELD + wave * 128);
#if defined(__HIP_DEVICE_COMPILE__)
        __builtin_amdgcn_global_load_lds((const void *)(src[fld] + own), lp, 16, 0, 0);
#else
        (void)lp;
#endif
      }
    }
  }
  double tail[4];
  if (half) {
    tail[0] = q_[own];
    tail[1] = u_[own];
    tail[2] = v_[own];
    tail[3] = w_[own];
  }

  // ---- neighbour traces of faces that leave the tile -----------------------
  int32_t nb_r[R];
  double tq_r[R], t1_r[R], t2_r[R], t3_r[R];
#pragma unroll
  for (int r = 0; r < R; ++r) tq_r[r] = t1_r[r] = t2_r[r] = t3_r[r] = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int it = tid + r * C::THREADS;
    const int el = it / NFT, f = (it - el * NFT) / NQ2, ab = it - el * NFT - f * NQ2;
    const int x = el * 6 + f;  // entry of (el, f)
    int32_t nb = -1;
#pragma unroll
    for (int c = 0; c < C::NBR_REGS; ++c) {
      const int32_t s = __shfl(nbv[c], x & 63);
      if ((x >> 6) == c) nb = s;
    }
    nb_r[r] = el < nel ? nb : -1;
    const bool outside = el < nel && (nb < e0 || nb >= e0 + nel);
    if (outside) {
      const int64_t pidx = (int64_t)nb * NP + nbr_face_node(f, ab % NQ, ab / NQ, NQ);
      tq_r[r] = q_[pidx];
      if constexpr (AXAL) {
        const int ax = face_axis(f);
        const double *vel = ax == 0 ? u_ : ax == 1 ? v_ : w_;
        t1_r[r] = vel[pidx];
      } else {
        t1_r[r] = u_[pidx];
        t2_r[r] = v_[pidx];
        t3_r[r] = w_[pidx];
      }
    }
  }

  // ---- RK base, operators, geometry ------------------------------------------
  const bool rk = mode_ == 1;
  const bool use_q0 = rk && ca_ != 0.0;
  dbl2 q0v = {0.0, 0.0};
  if (use_q0) {
    if (full) q0v = __builtin_nontemporal_load((const dbl2 *)(q0_ + own));
    else if (half) q0v.x = __builtin_nontemporal_load(q0_ + own);
  }
  if (tid < NQ2) sD[tid] = D_[tid];
  if (tid < 2 * NQ) sL[tid] = ell_[tid];
  static_assert(kGeoNormal == kGeoFscale + 6, "Fscale and normals are adjacent");
  if (tid < 24) sgeo[tid] = g[kGeoFscale + tid];
  __builtin_amdgcn_s_waitcnt(0);  // the async copies have landed in LDS

  if (half) {
    sq[2 * tid] = tail[0];
    su[2 * tid] = tail[1];
    sv[2 * tid] = tail[2];
    sw[2 * tid] = tail[3];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int it = tid + r * C::THREADS;
    tq[it] = tq_r[r];
    t1[it] = t1_r[r];
    if constexpr (!AXAL) {
      t1[C::SLAB + it] = t2_r[r];
      t1[2 * C::SLAB + it] = t3_r[r];
    }
  }
  __syncthreads();

  // ---- upwind flux of each face node of the tile -----------------------------
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
      double VelM, VelP;
      const double n1 = sgeo[6 + 3 * f], n2 = sgeo[7 + 3 * f], n3 = sgeo[8 + 3 * f];
      if constexpr (AXAL) {
        const int ax = face_axis(f);
        const double *sf = ax == 0 ? su : ax == 1 ? sv : sw;
        const double nf = ax == 0 ? n1 : ax == 1 ? n2 : n3;
        VelM = sf[m] * nf;
        VelP = (inside ? sf[pm] : t1[it]) * nf;
      } else {
        VelM = su[m] * n1 + sv[m] * n2 + sw[m] * n3;
        const double up = inside ? su[pm] : t1[it];
        const double vp = inside ? sv[pm] : t1[C::SLAB + it];
        const double wp = inside ? sw[pm] : t1[2 * C::SLAB + it];
        VelP = up * n1 + vp * n2 + wp * n3;
      }
      const double alpha = 0.5 * fabs(VelP + VelM);
      tq[it] = 0.5 * sgeo[f] * (qP * VelP - qM * VelM - alpha * (qP - qM));
    }
  }
  __syncthreads();

  // ---- volume fluxes in place: Fx over u, Fy over v, Fz over w -------------
  double qc[2] = {0.0, 0.0};
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int n = 2 * tid + r;
    if (n < rem) {
      const double qq = sq[n];
      qc[r] = qq;
      su[n] = qq * su[n];
      sv[n] = qq * sv[n];
      sw[n] = qq * sw[n];
    }
  }
  __syncthreads();

  // ---- contractions, Lift, combine, epilogue --------------------------------
  const double Ex = g[kGeoEscale], Ey = g[kGeoEscale + 1], Ez = g[kGeoEscale + 2];
  const double dtv = dt_geo_ ? g[kGeoDt] : dt_;
  double res[2] = {0.0, 0.0};
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int n = 2 * tid + r;
    if (n < rem) {
      const int el = n / NP, nn = n - el * NP;
      const int i = nn % NQ, j = (nn / NQ) % NQ, k = nn / NQ2, lb = el * NP;
      double dx = 0.0, dy = 0.0, dz = 0.0;
#pragma unroll
      for (int l = 0; l < NQ; ++l) {
        dx += sD[i + l * NQ] * su[lb + l + j * NQ + k * NQ2];
        dy += sD[j + l * NQ] * sv[lb + i + l * NQ + k * NQ2];
        dz += sD[k + l * NQ] * sw[lb + i + j * NQ + l * NQ2];
      }
      const double *G = tq + el * NFT;
      const double lift = sL[j] * G[0 * NQ2 + i + k * NQ] + sL[NQ + i] * G[1 * NQ2 + j + k * NQ] +
                          sL[NQ + j] * G[2 * NQ2 + i + k * NQ] + sL[i] * G[3 * NQ2 + j + k * NQ] +
                          sL[k] * G[4 * NQ2 + i + j * NQ] + sL[NQ + k] * G[5 * NQ2 + i + j * NQ];
      const double tend = -(((Ex * dx + Ey * dy) + Ez * dz) + lift);
      if (!rk) {
        res[r] = tend;
      } else {
        const double upd = cb_ * (qc[r] + dtv * tend);
        res[r] = use_q0 ? ca_ * (r == 0 ? q0v.x : q0v.y) + upd : upd;
      }
    }
  }
  if (full) {
    const dbl2 v = {res[0], res[1]};
    __builtin_nontemporal_store(v, (dbl2 *)(out_ + own));
  } else if (half) {
    __builtin_nontemporal_store(res[0], out_ + own);
  }
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <int NQ, bool AXAL>
int launch_run(const StageArgs &a, hipStream_t s) {
  using C = Run<NQ>;
  const int64_t ntile = (a.Ne + C::T - 1) / C::T;
  if (ntile > INT32_MAX) return set_error(DGK_ERR_UNSUPPORTED, "too many elements");
  // the XCD-interleaved order at tile granularity when periods hold whole tiles
  int64_t zpt = 0;
  if (a.zperiod > 0 && a.zperiod % C::T == 0) {
    const int64_t Qt = a.zperiod / C::T;
    if (Qt >= 8 && ntile % Qt == 0) zpt = Qt;
  }
  hipLaunchKernelGGL((stage_run<NQ, AXAL>), dim3((unsigned)ntile), dim3(C::THREADS), 0, s, a.q,
                     a.q0, a.u, a.v, a.w, a.out, a.nbr, a.D, a.ell, a.geo, a.geoid, a.Ne, a.ca,
                     a.cb, a.dt, a.mode, a.dt_geo, a.reverse, zpt);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

}  // namespace

int run_elements(int Nq) {
  switch (Nq) {
#define DGK_RUN_T(N) \
  case N: return Run<N>::T;
    DGK_RUN_T(2) DGK_RUN_T(3) DGK_RUN_T(4) DGK_RUN_T(5) DGK_RUN_T(6) DGK_RUN_T(7) DGK_RUN_T(8)
#undef DGK_RUN_T
  default: return 0;
  }
}

bool run_applicable(int Nq, const StageArgs &a) {
  const int T = run_elements(Nq);
  if (T < 1 || !a.ell || a.flux_in || !a.nbr || a.vmapP || a.order || a.cvel) return false;
  if (a.geoid && !a.geo_per_run) return false;
  const bool q0_ok = a.mode != 1 || a.ca == 0.0 || aligned16(a.q0);
  return aligned16(a.q) && aligned16(a.u) && aligned16(a.v) && aligned16(a.w) &&
         aligned16(a.out) && q0_ok;
}

int launch_run_f64(int Nq, const StageArgs &a, hipStream_t s) {
  if (a.Ne < 1) return DGK_OK;
  const bool ax = a.axis_aligned != 0;
  switch (Nq) {
#define DGK_RUN(N) \
  case N: return ax ? launch_run<N, true>(a, s) : launch_run<N, false>(a, s);
    DGK_RUN(2) DGK_RUN(3) DGK_RUN(4) DGK_RUN(5) DGK_RUN(6) DGK_RUN(7) DGK_RUN(8)
#undef DGK_RUN
  default: return set_error(DGK_ERR_UNSUPPORTED, "run kernel: Nq outside 2..8");
  }
}

}  // namespace dgk
