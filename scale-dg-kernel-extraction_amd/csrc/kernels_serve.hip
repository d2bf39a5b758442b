// Device-side set-up and read-out for requests admitted by the serving loop:
//
//  * box_nbr       : the periodic neighbour table of a NeX x NeY x NeZ box,
//                    built in HBM (eval_neighbor_element, mod_mesh.f90:407-483,
//                    face order -y,+x,+y,-x,-z,+z), so that admission uploads
//                    no connectivity;
//  * field_summary : min, max, the quadrature integral sum_e J sum_n w_n q and
//                    the quadrature norm sum_e J sum_n w_n q^2 of a finished
//                    field, by a fixed two-level reduction (the partition and
//                    the summation order depend only on Ne and Nq, so a
//                    request's summary does not depend on what ran beside it).
#include <hip/hip_runtime.h>

#include "dgk_internal.hpp"

namespace dgk {

namespace {

__global__ void box_nbr(int32_t *__restrict__ nbr, int NeX, int NeY, int NeZ,
                        int64_t off) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ne = (int64_t)NeX * NeY * NeZ;
  if (e >= ne) return;
  const int i = (int)(e % NeX), j = (int)((e / NeX) % NeY), k = (int)(e / ((int64_t)NeX * NeY));
  const int jm = j > 0 ? j - 1 : NeY - 1, jp = j < NeY - 1 ? j + 1 : 0;
  const int im = i > 0 ? i - 1 : NeX - 1, ip = i < NeX - 1 ? i + 1 : 0;
  const int km = k > 0 ? k - 1 : NeZ - 1, kp = k < NeZ - 1 ? k + 1 : 0;
  const int64_t sy = NeX, sz = (int64_t)NeX * NeY;
  int32_t *r = nbr + e * 6;
  r[0] = (int32_t)(off + i + jm * sy + k * sz);
  r[1] = (int32_t)(off + ip + j * sy + k * sz);
  r[2] = (int32_t)(off + i + jp * sy + k * sz);
  r[3] = (int32_t)(off + im + j * sy + k * sz);
  r[4] = (int32_t)(off + i + j * sy + km * sz);
  r[5] = (int32_t)(off + i + j * sy + kp * sz);
}

constexpr int kSumThreads = 256;

// Block b folds nodes t = b*256 + tid + m*gridDim*256 (m = 0, 1, ...) into
// per-thread partials, then a fixed tree over the block.
__global__ void summary_partial(const double *__restrict__ q, const double *__restrict__ w3,
                                int Np, int64_t n, double *partial) {
  __shared__ double s[4][kSumThreads];
  double mn = INFINITY, mx = -INFINITY, m1 = 0.0, m2 = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * kSumThreads + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * kSumThreads) {
    const double x = q[t];
    const double wt = w3[t % Np];
    mn = fmin(mn, x);
    mx = fmax(mx, x);
    m1 += wt * x;
    m2 += wt * (x * x);
  }
  s[0][threadIdx.x] = mn;
  s[1][threadIdx.x] = mx;
  s[2][threadIdx.x] = m1;
  s[3][threadIdx.x] = m2;
  __syncthreads();
  for (int h = kSumThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + h]);
      s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + h]);
      s[2][threadIdx.x] += s[2][threadIdx.x + h];
      s[3][threadIdx.x] += s[3][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[4 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void summary_final(const double *__restrict__ partial, int nblk, double jac,
                              double *result) {
  __shared__ double s[4][kSumThreads];
  double mn = INFINITY, mx = -INFINITY, m1 = 0.0, m2 = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kSumThreads) {
    mn = fmin(mn, partial[4 * b]);
    mx = fmax(mx, partial[4 * b + 1]);
    m1 += partial[4 * b + 2];
    m2 += partial[4 * b + 3];
  }
  s[0][threadIdx.x] = mn;
  s[1][threadIdx.x] = mx;
  s[2][threadIdx.x] = m1;
  s[3][threadIdx.x] = m2;
  __syncthreads();
  for (int h = kSumThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + h]);
      s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + h]);
      s[2][threadIdx.x] += s[2][threadIdx.x + h];
      s[3][threadIdx.x] += s[3][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    result[0] = s[0][0];
    result[1] = s[1][0];
    result[2] = jac * s[2][0];
    result[3] = jac * s[3][0];
  }
}

}  // namespace

int launch_box_nbr(int32_t *nbr, int NeX, int NeY, int NeZ, int64_t elem_off,
                   hipStream_t s) {
  const int64_t ne = (int64_t)NeX * NeY * NeZ;
  if (ne == 0) return DGK_OK;
  if (elem_off + ne > (int64_t)INT32_MAX)
    return set_error(DGK_ERR_UNSUPPORTED, "element index exceeds int32");
  hipLaunchKernelGGL(box_nbr, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, nbr, NeX,
                     NeY, NeZ, elem_off);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

int launch_field_summary(const double *q, const double *w3, int Np, int64_t n, double jac,
                         double *partial, double *result, hipStream_t s) {
  if (n < 1 || Np < 1) return set_error(DGK_ERR_ARG, "empty field");
  hipLaunchKernelGGL(summary_partial, dim3(kSummaryBlocks), dim3(kSumThreads), 0, s, q, w3,
                     Np, n, partial);
  DGK_HIP(hipGetLastError());
  hipLaunchKernelGGL(summary_final, dim3(1), dim3(kSumThreads), 0, s, partial,
                     kSummaryBlocks, jac, result);
  DGK_HIP(hipGetLastError());
  return DGK_OK;
}

}  // namespace dgk
