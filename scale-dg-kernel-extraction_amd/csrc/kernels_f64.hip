// FP64 device kernels of the DG advection path (gfx950).
//
// stage_f64<NQ,EPB>: one RK stage of the whole mesh in one launch.  Each
// element is owned by NQ*NQ threads; thread (i,j) holds the k-column of its
// element in registers (coalesced: a k-slab of one element is NQ*NQ
// consecutive doubles).  Per element:
//   A. load q,u,v,w columns; F = q*(u,v,w); Fx slab-transposed into LDS;
//      the element's own face traces (q, normal velocity) into LDS;
//   B. upwind flux of face node (i,j) of all six faces (neighbour traces are
//      gathered from HBM/L2 through the folded neighbour map), and the
//      x-derivative from LDS;
//   C. Fy into LDS, y-derivative, z-derivative in registers, Lift from the
//      stored Lift_mat, dqdt = -(((Ex Dx + Ey Dy) + Ez Dz) + Lift), then either
//      store dqdt or the RK stage value a*q0 + b*(q + dt*dqdt).
// This fuses cal_elembnd_flux (mod_advect3d_eq.f90:89-142), cal_dqdt
// (mod_advect3d_eq.f90:146-201), tensorprod_divlike_dirXYZ_general and
// tensorprod_Lift_hexahedral_general (mod_dg_optr_kernel.f90:197-288) and the
// RK update (main.f90:69-73).  The contraction order is the reference's
// (l ascending); products are fused into FMAs.
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
stage_f64(const StageArgs P) {
  constexpr int NQ2 = Tile<NQ>::NQ2, NP = Tile<NQ>::NP, EPB = Tile<NQ>::EPB;
  constexpr int NFT = 6 * NQ2;
  __shared__ double sF[EPB][NP];        // Fx (slab-transposed), then Fy
  __shared__ double sQ[EPB][6][NQ2];    // own face traces of q, then face flux
  __shared__ double sV[EPB][6][NQ2];    // own normal velocity on each face

  const int tid = threadIdx.x;
  const int el = tid / NQ2, ij = tid - el * NQ2;
  const int i = ij % NQ, j = ij / NQ;
  const int64_t e = (int64_t)xcd_block(blockIdx.x, gridDim.x) * EPB + el;
  const bool active = e < P.Ne;

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
        const double upd = P.cb * (q[k] + P.dt * tend);
        P.out[idx] = use_q0 ? P.ca * P.q0[idx] + upd : upd;
      }
    }
  }
}

template <int NQ>
int launch_nq(const StageArgs &a, hipStream_t s) {
  using T = Tile<NQ>;
  const int64_t nblk = (a.Ne + T::EPB - 1) / T::EPB;
  if (nblk > INT32_MAX) return set_error(DGK_ERR_UNSUPPORTED, "too many elements");
  hipLaunchKernelGGL(stage_f64<NQ>, dim3((unsigned)nblk), dim3(T::THREADS), 0, s, a);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
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

__global__ void to_f32(const double *__restrict__ src, float *__restrict__ dst,
                       int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (float)src[t];
}

unsigned grid1d(int64_t n, int threads) {
  return (unsigned)((n + threads - 1) / threads);
}

}  // namespace

int launch_stage_f64(int Nq, const StageArgs &a, hipStream_t s) {
  if (a.Ne < 1) return DGK_OK;
  switch (Nq) {
  case 2: return launch_nq<2>(a, s);
  case 3: return launch_nq<3>(a, s);
  case 4: return launch_nq<4>(a, s);
  case 5: return launch_nq<5>(a, s);
  case 6: return launch_nq<6>(a, s);
  case 7: return launch_nq<7>(a, s);
  case 8: return launch_nq<8>(a, s);
  case 9: return launch_nq<9>(a, s);
  case 10: return launch_nq<10>(a, s);
  case 11: return launch_nq<11>(a, s);
  case 12: return launch_nq<12>(a, s);
  case 13: return launch_nq<13>(a, s);
  case 14: return launch_nq<14>(a, s);
  case 15: return launch_nq<15>(a, s);
  case 16: return launch_nq<16>(a, s);
  default: return set_error(DGK_ERR_UNSUPPORTED, "Nq outside 2..16");
  }
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

int launch_to_f32(const double *src, float *dst, int64_t n, hipStream_t s) {
  if (n == 0) return DGK_OK;
  hipLaunchKernelGGL(to_f32, dim3(grid1d(n, 256)), dim3(256), 0, s, src, dst, n);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

}  // namespace dgk
