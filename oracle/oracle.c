/*
 * CPU oracle for the 3D DG advection tendency and its SSP-RK3 time loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C with OpenMP, the
 * reference Fortran GENERAL path (and the OPT1 summation order for p > 7) so
 * that the HIP kernels can be checked against it.  Only tests/, the smoke()
 * check and bench.py's cpu_baseline leg may load it.  The product library
 * (libdgk.so) never links or calls it.
 *
 * Conventions follow the reference exactly:
 *   * fields are f(Np, NeA), column-major, node n = i + j*Nq + k*Nq^2
 *     (0-based here), element ke = i + j*NeX + k*NeX*NeY;
 *   * VMapM / VMapP / halo_src_map hold 1-based flat indices into f(Np*NeA)
 *     (mod_mesh.f90:159-184), stored as int32;
 *   * every floating-point expression is evaluated in the reference's order,
 *     compiled with -ffp-contract=off so that no FMA is formed (the reference
 *     build, flang -O3 for x86-64, forms none either).
 *
 * Parity pins (tests/test_oracle_*.py): operators vs LGL closed forms and SBP
 * identities; maps vs geometric invariants; tendency vs a dense Kronecker
 * brute force; the whole loop vs golden fixtures written by the reference
 * Fortran (oracle/ref_harness.f90 + oracle/Makefile), conservation, free
 * stream, energy decay and the SSP-RK3 amplification factor.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define IDX3(i, j, k, Nq) ((i) + (j) * (Nq) + (k) * (Nq) * (Nq))

/* ------------------------------------------------------------------------ */
/* Operator file: element_init (mod_mesh.f90:253-263).                       */
/* Three list-directed records: x1D(Nq), D1D(Nq,Nq), Lift_mat(Nq,Nq,Nq,6).   */
/* Each read starts on a fresh line; values beyond the count are skipped.    */
/* ------------------------------------------------------------------------ */
static int read_record(FILE *fp, double *dst, long n) {
  long got = 0;
  char *line = NULL;
  size_t cap = 0;
  while (got < n) {
    ssize_t len = getline(&line, &cap, fp);
    if (len < 0) { free(line); return -1; }
    char *s = line, *end;
    while (got < n) {
      double v = strtod(s, &end);
      if (end == s) break;
      dst[got++] = v;
      s = end;
    }
  }
  free(line);
  return 0;
}

int orc_load_operators(const char *path, int Nq, double *x1D, double *D1D,
                       double *Lift) {
  FILE *fp = fopen(path, "r");
  if (!fp) return 1;
  int rc = 0;
  if (read_record(fp, x1D, Nq)) rc = 2;
  if (!rc && read_record(fp, D1D, (long)Nq * Nq)) rc = 3;
  if (!rc && read_record(fp, Lift, (long)Nq * Nq * Nq * 6)) rc = 4;
  fclose(fp);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Face node indices Fmask(Nfp,6), 1-based: element_init                      */
/* (mod_mesh.f90:278-324).  Faces 1..6 = -y, +x, +y, -x, -z, +z.              */
/* ------------------------------------------------------------------------ */
void orc_fmask(int Nq, int *Fmask) {
  int Nfp = Nq * Nq, fp;
  fp = 0;
  for (int k = 1; k <= Nq; ++k)
    for (int i = 1; i <= Nq; ++i) Fmask[fp++ + 0 * Nfp] = i + (k - 1) * Nq * Nq;
  fp = 0;
  for (int k = 1; k <= Nq; ++k)
    for (int j = 1; j <= Nq; ++j)
      Fmask[fp++ + 1 * Nfp] = Nq + (j - 1) * Nq + (k - 1) * Nq * Nq;
  fp = 0;
  for (int k = 1; k <= Nq; ++k)
    for (int i = 1; i <= Nq; ++i)
      Fmask[fp++ + 2 * Nfp] = i + (Nq - 1) * Nq + (k - 1) * Nq * Nq;
  fp = 0;
  for (int k = 1; k <= Nq; ++k)
    for (int j = 1; j <= Nq; ++j)
      Fmask[fp++ + 3 * Nfp] = 1 + (j - 1) * Nq + (k - 1) * Nq * Nq;
  fp = 0;
  for (int j = 1; j <= Nq; ++j)
    for (int i = 1; i <= Nq; ++i) Fmask[fp++ + 4 * Nfp] = i + (j - 1) * Nq;
  fp = 0;
  for (int j = 1; j <= Nq; ++j)
    for (int i = 1; i <= Nq; ++i)
      Fmask[fp++ + 5 * Nfp] = i + (j - 1) * Nq + (Nq - 1) * Nq * Nq;
}

/* elem_id (mod_mesh.f90:486-493), 1-based. */
static int elem_id(int i, int j, int k, int NeX, int NeY) {
  return i + (j - 1) * NeX + (k - 1) * NeX * NeY;
}

/* eval_neighbor_element (mod_mesh.f90:407-483). */
static void neighbor(int i, int j, int k, int f, int NeX, int NeY, int NeZ,
                     int *keP, int *opp, int *bnd) {
  *bnd = 0;
  switch (f) {
  case 1:
    *opp = 3;
    if (j > 1) *keP = elem_id(i, j - 1, k, NeX, NeY);
    else { *keP = elem_id(i, NeY, k, NeX, NeY); *bnd = 1; }
    break;
  case 2:
    *opp = 4;
    if (i < NeX) *keP = elem_id(i + 1, j, k, NeX, NeY);
    else { *keP = elem_id(1, j, k, NeX, NeY); *bnd = 1; }
    break;
  case 3:
    *opp = 1;
    if (j < NeY) *keP = elem_id(i, j + 1, k, NeX, NeY);
    else { *keP = elem_id(i, 1, k, NeX, NeY); *bnd = 1; }
    break;
  case 4:
    *opp = 2;
    if (i > 1) *keP = elem_id(i - 1, j, k, NeX, NeY);
    else { *keP = elem_id(NeX, j, k, NeX, NeY); *bnd = 1; }
    break;
  case 5:
    *opp = 6;
    if (k > 1) *keP = elem_id(i, j, k - 1, NeX, NeY);
    else { *keP = elem_id(i, j, NeZ, NeX, NeY); *bnd = 1; }
    break;
  default:
    *opp = 5;
    if (k < NeZ) *keP = elem_id(i, j, k + 1, NeX, NeY);
    else { *keP = elem_id(i, j, 1, NeX, NeY); *bnd = 1; }
    break;
  }
}

/* ------------------------------------------------------------------------ */
/* mesh_setup (mod_mesh.f90:74-193) with set_element_position (:332-353) and */
/* set_face_geometry (:356-397).  Array shapes (column-major):                */
/*   VMapM, VMapP (NfpTot,Ne); halo_src_map (NhaloNode);                      */
/*   normal_fn (NfpTot,Ne,3); Escale (Np,Ne,3); Fscale (NfpTot,Ne);           */
/*   pos_en (Np,Ne,3) -- pos_en may be NULL.                                  */
/* ------------------------------------------------------------------------ */
void orc_mesh_build(int NeX, int NeY, int NeZ, int Nq, const double *x1D,
                    double domx, double domy, double domz, int *VMapM,
                    int *VMapP, int *halo_src_map, double *normal_fn,
                    double *Escale, double *Fscale, double *pos_en) {
  const long Np = (long)Nq * Nq * Nq, Nfp = (long)Nq * Nq, NfpTot = 6 * Nfp;
  const long Ne = (long)NeX * NeY * NeZ;
  int *Fmask = (int *)malloc(sizeof(int) * Nfp * 6);
  orc_fmask(Nq, Fmask);
  const double dx = domx / (double)NeX, dy = domy / (double)NeY,
               dz = domz / (double)NeZ;

#pragma omp parallel for collapse(3) schedule(static)
  for (int k = 1; k <= NeZ; ++k)
    for (int j = 1; j <= NeY; ++j)
      for (int i = 1; i <= NeX; ++i) {
        const long ke = elem_id(i, j, k, NeX, NeY);
        const long e0 = ke - 1;
        const double xc = ((double)i - 0.5) * dx;
        const double yc = ((double)j - 0.5) * dy;
        const double zc = ((double)k - 0.5) * dz;
        if (pos_en) {
          for (int c = 0; c < Nq; ++c)
            for (int b = 0; b < Nq; ++b)
              for (int a = 0; a < Nq; ++a) {
                long n = IDX3(a, b, c, Nq);
                pos_en[n + e0 * Np + 0 * Np * Ne] = xc + 0.5 * dx * x1D[a];
                pos_en[n + e0 * Np + 1 * Np * Ne] = yc + 0.5 * dy * x1D[b];
                pos_en[n + e0 * Np + 2 * Np * Ne] = zc + 0.5 * dz * x1D[c];
              }
        }
        for (long n = 0; n < Np; ++n) {
          Escale[n + e0 * Np + 0 * Np * Ne] = 2.0 / dx;
          Escale[n + e0 * Np + 1 * Np * Ne] = 2.0 / dy;
          Escale[n + e0 * Np + 2 * Np * Ne] = 2.0 / dz;
        }
        /* set_face_geometry */
        for (long fp = 0; fp < NfpTot; ++fp)
          for (int d = 0; d < 3; ++d) normal_fn[fp + e0 * NfpTot + d * NfpTot * Ne] = 0.0;
        static const int ndir[6] = {1, 0, 1, 0, 2, 2};
        static const double nsign[6] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0};
        const double fsc[6] = {2.0 / dy, 2.0 / dx, 2.0 / dy, 2.0 / dx, 2.0 / dz, 2.0 / dz};
        for (int f = 0; f < 6; ++f)
          for (long fp = 0; fp < Nfp; ++fp) {
            long r = f * Nfp + fp;
            normal_fn[r + e0 * NfpTot + ndir[f] * NfpTot * Ne] = nsign[f];
            Fscale[r + e0 * NfpTot] = fsc[f];
          }
        for (int f = 1; f <= 6; ++f) {
          int keP, opp, bnd;
          neighbor(i, j, k, f, NeX, NeY, NeZ, &keP, &opp, &bnd);
          for (long fp = 0; fp < Nfp; ++fp)
            VMapM[(f - 1) * Nfp + fp + e0 * NfpTot] =
                Fmask[fp + (f - 1) * Nfp] + (int)(e0 * Np);
          if (!bnd)
            for (long fp = 0; fp < Nfp; ++fp)
              VMapP[(f - 1) * Nfp + fp + e0 * NfpTot] =
                  Fmask[fp + (opp - 1) * Nfp] + (int)((keP - 1) * Np);
        }
      }

  /* serialized halo map (mod_mesh.f90:171-190) */
  long ibuf = 0;
  for (int k = 1; k <= NeZ; ++k)
    for (int j = 1; j <= NeY; ++j)
      for (int i = 1; i <= NeX; ++i) {
        const long e0 = elem_id(i, j, k, NeX, NeY) - 1;
        for (int f = 1; f <= 6; ++f) {
          int keP, opp, bnd;
          neighbor(i, j, k, f, NeX, NeY, NeZ, &keP, &opp, &bnd);
          for (long fp = 0; fp < Nfp; ++fp)
            if (bnd) {
              ibuf += 1;
              VMapP[(f - 1) * Nfp + fp + e0 * NfpTot] = (int)(Np * Ne + ibuf);
              halo_src_map[ibuf - 1] = Fmask[fp + (opp - 1) * Nfp] + (int)((keP - 1) * Np);
            }
        }
      }
  free(Fmask);
}

/* update_halo (mod_mesh.f90:209-221). */
void orc_update_halo(double *f, long Np, long Ne, long NhaloNode,
                     const int *halo_src_map) {
  for (long ib = 0; ib < NhaloNode; ++ib)
    f[Np * Ne + ib] = f[halo_src_map[ib] - 1];
}

/* ------------------------------------------------------------------------ */
/* cal_elembnd_flux (mod_advect3d_eq.f90:89-142).  flux(NfpTot,Ne).          */
/* ------------------------------------------------------------------------ */
void orc_elembnd_flux(double *flux, const double *q, const double *u,
                      const double *v, const double *w, const int *VMapM,
                      const int *VMapP, const double *normal_fn,
                      const double *Fscale, long NfpTot, long Ne) {
#pragma omp parallel for schedule(static)
  for (long ke = 0; ke < Ne; ++ke) {
    for (long fp = 0; fp < NfpTot; ++fp) {
      const long r = fp + ke * NfpTot;
      const long iM = VMapM[r] - 1, iP = VMapP[r] - 1;
      const double n1 = normal_fn[r], n2 = normal_fn[r + NfpTot * Ne],
                   n3 = normal_fn[r + 2 * NfpTot * Ne];
      const double qM = q[iM], qP = q[iP];
      const double VelM = u[iM] * n1 + v[iM] * n2 + w[iM] * n3;
      const double VelP = u[iP] * n1 + v[iP] * n2 + w[iP] * n3;
      const double alpha = 0.5 * fabs(VelP + VelM);
      flux[r] = 0.5 * Fscale[r] * (qP * VelP - qM * VelM - alpha * (qP - qM));
    }
  }
}

/* ------------------------------------------------------------------------ */
/* tensorprod_divlike_dirXYZ_general (mod_dg_optr_kernel.f90:197-258) and   */
/* the OPT1 generated form (mod_dg_optr_kernel_opt1.F90.erb:228-344):        */
/* variant 1 = GENERAL (0 + sum over l, in order);                           */
/* variant 2 = OPT1 (same order for p <= 7; for p > 7 three partial sums over */
/* l-blocks of round(Nq/3), round(Nq/3), rest, added as (t1 + t2) + t3).     */
/* Mat(i,l) = D[i + l*Nq]; Mat_tr(l,j) = Mat(j,l).                           */
/* ------------------------------------------------------------------------ */
static inline double dot_general(const double *a, long sa, const double *b,
                                 long sb, int Nq) {
  double s = 0.0;
  for (int l = 0; l < Nq; ++l) s = s + a[l * sa] * b[l * sb];
  return s;
}

static inline double dot_opt1(const double *a, long sa, const double *b,
                              long sb, int Nq) {
  if (Nq <= 8) {
    double s = a[0] * b[0];
    for (int l = 1; l < Nq; ++l) s = s + a[l * sa] * b[l * sb];
    return s;
  }
  const int nd3 = (int)floor((double)Nq / 3.0 + 0.5);
  double t1 = a[0] * b[0];
  for (int l = 1; l < nd3; ++l) t1 = t1 + a[l * sa] * b[l * sb];
  double t2 = a[nd3 * sa] * b[nd3 * sb];
  for (int l = nd3 + 1; l < 2 * nd3; ++l) t2 = t2 + a[l * sa] * b[l * sb];
  double t3 = a[2 * nd3 * sa] * b[2 * nd3 * sb];
  for (int l = 2 * nd3 + 1; l < Nq; ++l) t3 = t3 + a[l * sa] * b[l * sb];
  return t1 + t2 + t3;
}

void orc_divlike(double *ox, double *oy, double *oz, const double *D,
                 const double *Fx, const double *Fy, const double *Fz, int Nq,
                 int variant) {
  const long Nq2 = (long)Nq * Nq;
  for (int k = 0; k < Nq; ++k)
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        const long n = IDX3(i, j, k, Nq);
        /* x: sum_l Mat(i,l) * Fx(l,jk) */
        const double *ax = D + i, *bx = Fx + j * Nq + k * Nq2;
        /* y: sum_l Fy(i,l,k) * Mat_tr(l,j),  Mat_tr(l,j) = D[j + l*Nq] */
        const double *ay = Fy + i + k * Nq2, *by = D + j;
        /* z: sum_l Fz(i,j,l) * Mat_tr(l,k) */
        const double *az = Fz + i + j * Nq, *bz = D + k;
        if (variant == 2) {
          ox[n] = dot_opt1(ax, Nq, bx, 1, Nq);
          oy[n] = dot_opt1(ay, Nq, by, Nq, Nq);
          oz[n] = dot_opt1(az, Nq2, bz, Nq, Nq);
        } else {
          ox[n] = dot_general(ax, Nq, bx, 1, Nq);
          oy[n] = dot_general(ay, Nq, by, Nq, Nq);
          oz[n] = dot_general(az, Nq2, bz, Nq, Nq);
        }
      }
}

/* tensorprod_Lift_hexahedral_general (mod_dg_optr_kernel.f90:262-288);     */
/* the OPT1 form (mod_dg_optr_kernel_opt1.F90.erb:134-157) is identical.     */
void orc_lift(double *out, const double *Lift, const double *g, int Nq) {
  const long Nq2 = (long)Nq * Nq, Np = Nq2 * Nq;
  for (int k = 0; k < Nq; ++k)
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        const long n = IDX3(i, j, k, Nq);
        out[n] = Lift[n + 0 * Np] * g[i + k * Nq + 0 * Nq2] +
                 Lift[n + 1 * Np] * g[j + k * Nq + 1 * Nq2] +
                 Lift[n + 2 * Np] * g[i + k * Nq + 2 * Nq2] +
                 Lift[n + 3 * Np] * g[j + k * Nq + 3 * Nq2] +
                 Lift[n + 4 * Np] * g[i + j * Nq + 4 * Nq2] +
                 Lift[n + 5 * Np] * g[i + j * Nq + 5 * Nq2];
      }
}

/* cal_dqdt (mod_advect3d_eq.f90:146-201). */
void orc_cal_dqdt(double *dqdt, const double *q, const double *u,
                  const double *v, const double *w, const double *flux_bnd,
                  const double *D, const double *Lift, const double *Escale,
                  int Nq, long Ne, int variant) {
  const long Np = (long)Nq * Nq * Nq, NfpTot = 6L * Nq * Nq;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * Np * 7);
    double *Fx = buf, *Fy = buf + Np, *Fz = buf + 2 * Np, *Dx = buf + 3 * Np,
           *Dy = buf + 4 * Np, *Dz = buf + 5 * Np, *L = buf + 6 * Np;
#pragma omp for schedule(static)
    for (long ke = 0; ke < Ne; ++ke) {
      const long o = ke * Np;
      for (long n = 0; n < Np; ++n) {
        Fx[n] = q[o + n] * u[o + n];
        Fy[n] = q[o + n] * v[o + n];
        Fz[n] = q[o + n] * w[o + n];
      }
      orc_divlike(Dx, Dy, Dz, D, Fx, Fy, Fz, Nq, variant);
      orc_lift(L, Lift, flux_bnd + ke * NfpTot, Nq);
      for (long n = 0; n < Np; ++n)
        dqdt[o + n] = -(Escale[o + n] * Dx[n] + Escale[o + n + Np * Ne] * Dy[n] +
                        Escale[o + n + 2 * Np * Ne] * Dz[n] + L[n]);
    }
    free(buf);
  }
}

/* advect3d_eq_cal_tend (mod_advect3d_eq.f90:41-85).  ebnd_flux is a          */
/* caller-visible scratch of NfpTot*Ne doubles (may be NULL: allocated here).  */
void orc_cal_tend(double *dqdt, const double *q, const double *u,
                  const double *v, const double *w, const double *D,
                  const double *Lift, const int *VMapM, const int *VMapP,
                  const double *normal_fn, const double *Escale,
                  const double *Fscale, int Nq, long Ne, double *ebnd_flux,
                  int variant) {
  const long NfpTot = 6L * Nq * Nq;
  double *fl = ebnd_flux ? ebnd_flux : (double *)malloc(sizeof(double) * NfpTot * Ne);
  orc_elembnd_flux(fl, q, u, v, w, VMapM, VMapP, normal_fn, Fscale, NfpTot, Ne);
  orc_cal_dqdt(dqdt, q, u, v, w, fl, D, Lift, Escale, Nq, Ne, variant);
  if (!ebnd_flux) free(fl);
}

/* SSP-RK3 coefficients (mod_common.f90:28-30). */
static const double RK_A[3] = {0.0, 0.75, 1.0 / 3.0};
static const double RK_B[3] = {1.0, 0.25, 2.0 / 3.0};

/* One RK stage update (main.f90:69-73). */
void orc_rk_update(double *q, const double *q0, const double *dqdt, double dt,
                   int stage, long Np, long Ne) {
  const double a = RK_A[stage], b = RK_B[stage];
#pragma omp parallel for schedule(static)
  for (long n = 0; n < Np * Ne; ++n) q[n] = a * q0[n] + b * (q[n] + dt * dqdt[n]);
}

/* ------------------------------------------------------------------------ */
/* Time loop (main.f90:51-80).  q,u,v,w are (Np,NeA) with u,v,w halos       */
/* already filled (main.f90:159-162).  minmax (optional, 2*nstep doubles)    */
/* receives min/max of the owned q after each step (main.f90:76-79 prints    */
/* them every output_interval).  Returns the wall time of the loop (the      */
/* reference "Main" timer span, main.f90:164->174).                          */
/* ------------------------------------------------------------------------ */
double orc_run(double *q, const double *u, const double *v, const double *w,
               const double *D, const double *Lift, const int *VMapM,
               const int *VMapP, const int *halo_src_map,
               const double *normal_fn, const double *Escale,
               const double *Fscale, int Nq, long Ne, long NeA, long NhaloNode,
               double dt, int nstep, int variant, double *minmax) {
  const long Np = (long)Nq * Nq * Nq, NfpTot = 6L * Nq * Nq;
  double *q0 = (double *)malloc(sizeof(double) * Np * NeA);
  double *dqdt = (double *)malloc(sizeof(double) * Np * NeA);
  double *fl = (double *)malloc(sizeof(double) * NfpTot * Ne);
  double t0 = omp_get_wtime();
  for (int istep = 0; istep < nstep; ++istep) {
#pragma omp parallel for schedule(static)
    for (long n = 0; n < Np * Ne; ++n) q0[n] = q[n];
    for (int stage = 0; stage < 3; ++stage) {
      orc_update_halo(q, Np, Ne, NhaloNode, halo_src_map);
      orc_cal_tend(dqdt, q, u, v, w, D, Lift, VMapM, VMapP, normal_fn, Escale,
                   Fscale, Nq, Ne, fl, variant);
      orc_rk_update(q, q0, dqdt, dt, stage, Np, Ne);
    }
    if (minmax) {
      double mn = q[0], mx = q[0];
      for (long n = 1; n < Np * Ne; ++n) {
        if (q[n] < mn) mn = q[n];
        if (q[n] > mx) mx = q[n];
      }
      minmax[2 * istep] = mn;
      minmax[2 * istep + 1] = mx;
    }
  }
  double t1 = omp_get_wtime();
  free(q0);
  free(dqdt);
  free(fl);
  return t1 - t0;
}

/* ------------------------------------------------------------------------ */
/* Deterministic synthetic inputs: splitmix64 (Vigna, public domain) over a  */
/* counter, u01 = (x >> 11) * 2^-53.  Shared with the product's generator so  */
/* that every component sees bit-identical inputs.                           */
/* ------------------------------------------------------------------------ */
static inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double orc_u01(uint64_t seed, uint64_t counter) {
  return (double)(splitmix64((seed << 40) ^ counter) >> 11) * 0x1.0p-53;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
