/*
 * dgk.h -- C ABI of the MI355X DG advection library (libdgk.so).
 *
 * The library is a drop-in for the reference's operator selection
 * (DGOptrKernel_OptType = GENERAL | OPT1, main.f90:101-107 ->
 * dg_optr_kernel_setup, mod_dg_optr_kernel.f90:45-58) one level up: it takes
 * over the whole-mesh tendency advect3d_eq_cal_tend
 * (mod_advect3d_eq.f90:41-85) together with the stage loop's update_halo and
 * RK update (main.f90:51-75), with the state resident in device memory.
 *
 * ABI rules
 *   * Plain C: pointers and sizes only, no ownership transfer.  Arrays are the
 *     reference's Fortran arrays passed by reference (column-major, real64,
 *     default integer = int32, 1-based flat indices in VMapM/VMapP/
 *     halo_src_map).  The library copies what it needs; the caller keeps its
 *     host memory.
 *   * Every entry point returns an int status (DGK_OK = 0).  On failure the
 *     message of the calling thread's last error is in dgk_last_error().
 *   * A solver handle is single-thread-affine and owns one HIP stream; its
 *     calls are ordered on that stream.  Calls that return host data
 *     synchronize the stream.
 *   * The device path has no CPU fallback: when no gfx950 device is present,
 *     dgk_create* fail with DGK_ERR_DEVICE.
 */
#ifndef DGK_H
#define DGK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  DGK_OK = 0,
  DGK_ERR_ARG = 1,         /* null pointer, size or range error              */
  DGK_ERR_UNSUPPORTED = 2, /* valid for the reference, outside the GPU path  */
  DGK_ERR_DEVICE = 3,      /* HIP runtime error or no usable device          */
  DGK_ERR_NOMEM = 4,       /* device memory exhausted / over budget          */
  DGK_ERR_IO = 5,          /* operator file missing or malformed             */
  DGK_ERR_STATE = 6        /* call out of order                              */
};

/* ---- operator evaluation types -----------------------------------------
 * Replaces the string -> eval_typeid switch of dg_optr_kernel_setup
 * (mod_dg_optr_kernel.f90:45-58).  "GENERAL" and "OPT1" are recognised (they
 * name the reference CPU kernels and are rejected by dgk_create with
 * DGK_ERR_UNSUPPORTED); "HIP" selects the FP64 device path and "HIP_MP" the
 * mixed-precision device path.  Unknown names are an error (DGK_ERR_ARG),
 * where the reference prints and continues (mod_dg_optr_kernel.f90:54-55). */
enum { DGK_EVAL_GENERAL = 1, DGK_EVAL_OPT1 = 2, DGK_EVAL_HIP = 3, DGK_EVAL_HIP_MP = 4 };

/* ---- field ids ---------------------------------------------------------- */
enum { DGK_FIELD_Q = 0, DGK_FIELD_U = 1, DGK_FIELD_V = 2, DGK_FIELD_W = 3,
       DGK_FIELD_DQDT = 4 };

typedef struct dgk_solver dgk_solver;

/* Message of the calling thread's last failed call ("" if none). */
const char *dgk_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int dgk_version(void);
/* Number of usable HIP devices (0 on a host without one). */
int dgk_device_count(int *count);

/* Name -> evaluation type id (mod_dg_optr_kernel.f90:45-58). */
int dgk_evaltype_id(const char *name, int *id);

/* ---- element and operators ------------------------------------------- */
/* Nq = p+1, Np = Nq^3, Nfp = Nq^2, NfpTot = 6*Nfp (mod_mesh.f90:236-241). */
int dgk_element_sizes(int p, int *Nq, int *Np, int *Nfp, int *NfpTot);
/* Read operator_data/p{p}.dat from dir: x1D(Nq), D1D(Nq,Nq),
 * Lift_mat(Nq,Nq,Nq,6), column-major (element_init, mod_mesh.f90:253-261).
 * Orders 1..19 are accepted; above p = 15, for which the reference ships no
 * file, a missing file is replaced by dgk_make_operators. */
int dgk_load_operators(const char *dir, int p, double *x1D, double *D1D,
                       double *Lift_mat);
/* The same three arrays from their closed forms, for any p >= 1: x1D the
 * Legendre-Gauss-Lobatto nodes; D1D(i,j) = l_j'(x_i); Lift_mat(:,:,:,f) the
 * tensor product of the identity with l-/+(x) = sum_n (2n+1)/2 P_n(x)
 * P_n(-/+1) along face f's normal.  For p <= 15 they equal the reference's
 * files to the files' rounding.  No counterpart in the reference, which
 * stops at p = 15 (mod_dg_optr_kernel_opt1.F90.erb:1). */
int dgk_make_operators(int p, double *x1D, double *D1D, double *Lift_mat);
/* Write dir/p{p}.dat in the reference's format (three records of
 * ES25.16E3 values) from dgk_make_operators. */
int dgk_write_operators(const char *dir, int p);

/* ---- structured periodic mesh (mesh_setup, mod_mesh.f90:74-193) ------- */
/* Ne = NeX*NeY*NeZ, Nhalo = 2(NeY*NeZ + NeX*NeZ + NeX*NeY), NeA = Ne + Nhalo,
 * NhaloNode = Nfp*Nhalo (mod_mesh.f90:96-100). */
int dgk_mesh_sizes(int NeX, int NeY, int NeZ, int p, int64_t *Ne,
                   int64_t *NeA, int64_t *NhaloNode);
/* Fill the reference's mesh arrays for the periodic box [0,dom_x]x[0,dom_y]x
 * [0,dom_z] (mesh_setup, mod_mesh.f90:74-193; set_element_position :332-353;
 * set_face_geometry :356-397; eval_neighbor_element :407-483).  Shapes:
 * VMapM, VMapP (NfpTot,Ne); halo_src_map (NhaloNode); normal_fn (NfpTot,Ne,3);
 * Escale (Np,Ne,3); Fscale (NfpTot,Ne); pos_en (Np,Ne,3) (may be NULL). */
int dgk_mesh_build(int NeX, int NeY, int NeZ, int p, const double *x1D,
                   double dom_x, double dom_y, double dom_z, int *VMapM,
                   int *VMapP, int *halo_src_map, double *normal_fn,
                   double *Escale, double *Fscale, double *pos_en);

/* ---- solver ------------------------------------------------------------ */
/* Create a device solver from the reference's arrays: the inputs of
 * advect3d_eq_cal_tend (mod_advect3d_eq.f90:41-66) plus the halo map of
 * update_halo (mod_mesh.f90:209-221).  The connectivity and geometry are
 * compressed at creation (per-element neighbour table, deduplicated per-face
 * geometry); arrays the GPU path cannot represent (non-affine geometry, or
 * VMapM not addressing the element's own Fmask nodes) give
 * DGK_ERR_UNSUPPORTED.  "HIP" applies Lift_mat in its separable 1D form
 * (l-/l+ along each face normal) when every stored entry lies within 1e-12
 * of max|Lift_mat| of that form, and reads Lift_mat as stored otherwise;
 * "HIP_MP" always uses the 1D form and rejects a Lift_mat that deviates from
 * it by more than 1e-10 relative.
 * Processing order (never the arithmetic, so never the results): a mesh
 * numbered plane by plane along z with periodic +z neighbours and planes of
 * at least 8 elements is dealt to the GPU's 8 XCDs in eighths of every
 * plane (XCD x walks the eighth x of every plane in turn; the FP64 VALU
 * kernels keep the element numbering where an eighth holds fewer than 96
 * elements), and consecutive stage launches walk the elements in opposite
 * directions (each XCD its own elements, in that order).  Environment:
 * DGK_ORDER=identity keeps the element numbering, DGK_ALT=0 keeps one
 * direction, DGK_ALT=global reverses the whole order instead
 * (DGK_ORDER=zperiod:Q deals eighths of every Q elements instead, 8 <= Q, Q
 * dividing Ne; DGK_ORDER=brick:bx,by,bz, structured solvers only, walks
 * bricks of elements instead). */
int dgk_create(int evaltype, int Nq, int64_t Ne, int64_t NeA,
               int64_t NhaloNode, const double *D1D, const double *Lift_mat,
               const int *VMapM, const int *VMapP, const int *halo_src_map,
               const double *normal_fn, const double *Escale,
               const double *Fscale, dgk_solver **out);
/* Same solver for the structured periodic box, built without materialising
 * the per-node arrays (operators read from operator_dir/p{p}.dat). */
int dgk_create_structured(int evaltype, int NeX, int NeY, int NeZ, int p,
                          const char *operator_dir, double dom_x,
                          double dom_y, double dom_z, dgk_solver **out);
int dgk_destroy(dgk_solver *s);

/* Sizes of a solver: Nq, Ne, NeA, NhaloNode. */
int dgk_solver_info(const dgk_solver *s, int *Nq, int64_t *Ne, int64_t *NeA,
                    int64_t *NhaloNode);
/* Device bytes held by the solver. */
int dgk_device_bytes(const dgk_solver *s, int64_t *bytes);

/* Device-side initial condition for a solver made by dgk_create_structured:
 * q = exp(-r^2 / (2 sigma^2)) with r the periodic minimum-image distance of
 * the node (set_element_position, mod_mesh.f90:332-353) to (cx,cy,cz), and
 * constant velocity (ux,uy,uz); halos of u, v, w refreshed. */
int dgk_init_gaussian(dgk_solver *s, double cx, double cy, double cz, double sigma,
                      double ux, double uy, double uz);

/* Upload a field, n = Np*Ne (owned part) or Np*NeA (the halo tail is
 * ignored: it is derived from the owned part).  Setting U, V or W refreshes
 * their halos once, as main.f90:160-162 does. */
int dgk_set_field(dgk_solver *s, int which, const double *f, int64_t n);
/* Download a field, n = Np*Ne or Np*NeA.  With n = Np*NeA the halo tail holds
 * what the last dgk_update_halo wrote (zeros before any). */
int dgk_get_field(dgk_solver *s, int which, double *f, int64_t n);

/* update_halo (mod_mesh.f90:209-221) on the device copy of a field. */
int dgk_update_halo(dgk_solver *s, int which);
/* q0 <- q over the owned elements (main.f90:53-56). */
int dgk_copy_q0(dgk_solver *s);
/* cal_elembnd_flux (mod_advect3d_eq.f90:89-142) of the current q into the
 * solver's flux buffer (NfpTot,Ne); copied to flux_host when not NULL. */
int dgk_cal_elembnd_flux(dgk_solver *s, double *flux_host);
/* cal_dqdt (mod_advect3d_eq.f90:146-201) from the flux buffer filled by
 * dgk_cal_elembnd_flux (or uploaded from flux_host when not NULL) into the
 * device dqdt. */
int dgk_cal_dqdt(dgk_solver *s, const double *flux_host);
/* advect3d_eq_cal_tend (mod_advect3d_eq.f90:41-85), fused flux + volume +
 * Lift in one kernel, into the device dqdt; copied to dqdt_host (Np*Ne) when
 * not NULL.  Reads the halo through the folded neighbour map, so it needs no
 * prior dgk_update_halo. */
int dgk_cal_tend(dgk_solver *s, double *dqdt_host);
/* RK stage update q = a_s q0 + b_s (q + dt dqdt), stage = 1..3
 * (main.f90:69-73, coefficients mod_common.f90:28-30). */
int dgk_rk_update(dgk_solver *s, int stage, double dt);
/* nstep full SSP-RK3 steps on the device (main.f90:51-75), each stage one
 * fused kernel (flux + volume + Lift + RK update, halo folded). */
int dgk_step(dgk_solver *s, double dt, int nstep);
/* min/max of the owned q (main.f90:76-79). */
int dgk_minmax(dgk_solver *s, double *qmin, double *qmax);
/* Wait for the solver's stream. */
int dgk_synchronize(dgk_solver *s);
/* min/max of n doubles at the device address q into result[0] (min) and
 * result[1] (max), also a device address, enqueued on stream (a hipStream_t,
 * NULL for the null stream) without waiting, allocating nothing: the partials
 * of the reduction live in a 16 KiB static device array that each GPU holds
 * until the process ends.  No counterpart in the reference; it is for callers
 * that cannot allocate (code under HIP graph capture -- make the first call
 * on each device outside a capture, it looks up the array's address -- and
 * tight serving loops).  Any n from 1 to 2^40 fits the fixed partials;
 * larger or smaller n, NULL pointers, host pointers and ranges that run past
 * their device allocation are refused with DGK_ERR_ARG.  Calls that overlap
 * on one GPU share the array and race; calls ordered on one stream do not.
 * Same values as dgk_minmax of the same field. */
int dgk_minmax_noalloc(const double *q, int64_t n, double *result, void *stream);

/* ---- timers (mod_common.f90:33-71, names of main.f90:175-177 and
 * mod_advect3d_eq.f90:34-35) ---------------------------------------------
 * seconds[0] Main (host wall time inside dgk_step),
 * seconds[1] Cal_tend (device time of tendency kernels),
 * seconds[2] Element boundary flux (device, dgk_cal_elembnd_flux),
 * seconds[3] Volume derivate + surface lift (device, dgk_cal_dqdt),
 * seconds[4] fused stage kernels (device, dgk_step),
 * counts[0..3] number of timed intervals, counts[4] number of stage launches.
 * Device times are collected only while dgk_set_kernel_timing(s, 1) is on:
 * HIP events around each tendency launch, and one event pair around all
 * 3 nstep stage launches of a dgk_step call (nothing is recorded between
 * them, so seconds[4] / counts[4] is the mean launch duration of a
 * back-to-back sequence, the gaps between launches included). */
int dgk_set_kernel_timing(dgk_solver *s, int on);
int dgk_timer_report(dgk_solver *s, double *seconds, int64_t *counts);
int dgk_timer_reset(dgk_solver *s);

/* ---- batched ensembles ----------------------------------------------------
 * No reference counterpart: the reference runs one mesh per process.  A
 * batch holds ninst independent periodic boxes (instance k: ne_xyz[3k..3k+2]
 * elements, order p[k], domain dom_xyz[3k..3k+2] or the unit cube when NULL,
 * time step dt[k]) resident in device memory.  Instances of one order are
 * advanced together by the stage kernels (one launch per stage and order);
 * dgk_batch_step captures one SSP-RK3 step of every order group in a
 * hipGraph on first use and replays it nstep times.  Per-instance arithmetic
 * is identical to a dgk_solver of that instance alone.  A velocity component
 * that is one constant over an instance (dgk_batch_init_gaussian, or a
 * constant field given to dgk_batch_set_field) is kept in the instance's
 * geometry record; an order group whose instances all have constant
 * velocities runs stage kernels that read it there instead of the fields
 * (the same bits, 24 bytes less per DOF and stage). */
typedef struct dgk_batch dgk_batch;
int dgk_batch_create(int evaltype, int ninst, const int *ne_xyz, const int *p,
                     const double *dom_xyz, const double *dt,
                     const char *operator_dir, dgk_batch **out);
int dgk_batch_destroy(dgk_batch *b);
/* order, elements and DOF of instance inst */
int dgk_batch_info(const dgk_batch *b, int inst, int *p, int64_t *Ne, int64_t *dof);
/* number of instances, order groups and device bytes */
int dgk_batch_size(const dgk_batch *b, int *ninst, int *ngroups, int64_t *bytes);
/* fields of one instance, n = Np*Ne of that instance (Q, U, V or W) */
int dgk_batch_set_field(dgk_batch *b, int inst, int which, const double *f, int64_t n);
int dgk_batch_get_field(dgk_batch *b, int inst, int which, double *f, int64_t n);
/* device-side Gaussian-blob initial condition of one instance (as
 * dgk_init_gaussian) */
int dgk_batch_init_gaussian(dgk_batch *b, int inst, double cx, double cy, double cz,
                            double sigma, double ux, double uy, double uz);
/* graph replay (default) or eager launches */
int dgk_batch_set_graph(dgk_batch *b, int on);
int dgk_batch_step(dgk_batch *b, int nstep);
int dgk_batch_synchronize(dgk_batch *b);
/* device seconds and number of dgk_batch_step calls since the last read */
int dgk_batch_timer(dgk_batch *b, double *seconds, int64_t *calls);

/* ---- tile shapes and the autotuner --------------------------------------
 * The stage kernels come in several tile shapes per order, given as the
 * number of elements one workgroup advances.
 *   HIP_MP: the MFMA kernel with one, two or four wavefronts of whole
 *           elements, then one wavefront processed direction by direction
 *           (a third of the LDS), for p <= 7; for p >= 8 one element, two
 *           elements, and one element direction by direction; above p = 15
 *           one element, all at once and direction by direction.  Shapes
 *           are limited to 160 KiB of LDS.  At p <= 4 the next candidate
 *           advances runs of consecutive elements, and at p <= 7 the last
 *           is the FP64 register VALU kernel reading the FP32 velocity
 *           fields (FP64 arithmetic: its bits differ from the MFMA shapes',
 *           which agree with each other).  Defaults: the VALU kernel at
 *           p = 1, two wavefronts (candidate 1) at p = 4, four wavefronts
 *           (candidate 2) at p = 5, the direction-by-direction shape at
 *           p = 14, 16, 17, candidate 0 otherwise.
 *   HIP:    candidate 0 is the VALU kernel with q, u, v, w staged in LDS,
 *           1 the VALU kernel with v, w in registers (half the LDS);
 *           candidates 2.. are the FP64 MFMA kernel's shapes as for HIP_MP,
 *           limited to 160 KiB of LDS; above p = 15 only the MFMA shapes
 *           (one element all at once up to p = 17, direction by direction).
 *           Defaults: the register VALU kernel at p = 4 and 6..13, the
 *           direction-by-direction MFMA shape (the last candidate) at
 *           p = 14..19, the staged VALU kernel otherwise.
 * Constant-velocity launches (batch groups of constant-velocity instances,
 * served requests) start from defaults of their own, taken only among shapes
 * that give the bits of the defaults above: FP64 the register VALU kernel
 * also at p = 2, 3; mixed two wavefronts at p = 3, one at p = 5, the
 * direction-by-direction shape at p = 6, 9, 12.
 * dgk_tile_select sets the process-wide choice for that form and order, for
 * both kinds of launch (-1 restores the defaults); it holds for every solver
 * that has not been autotuned; dgk_tile_selected reports the choice for
 * launches with velocity fields.  Launches with a given flux (dgk_cal_dqdt)
 * or a stored, non-separable Lift always use the VALU kernels.  dgk_autotune
 * times each candidate with HIP events over reps RK-stage launches on the
 * solver's own mesh (writing only the solver's scratch rotation buffer) and
 * keeps the fastest for that solver alone, since the best shape depends on
 * the mesh (the solver's default stays unless another shape is at least 2%
 * faster);
 * us_per_launch[c] receives candidate c's time (0 for a candidate that would
 * launch another shape on this solver: the kernel over runs of elements, on
 * launches it cannot take -- constant velocities, a resolved face-node table,
 * an explicit order).  dgk_batch_autotune does this
 * for every order group of a batch and reports per group: the order, the
 * chosen candidate, its time and the default's time (the batch's captured
 * step is re-captured at the next dgk_batch_step).  No counterpart
 * in the reference, whose OPT1 kernels are generated per order
 * (mod_dg_optr_kernel_opt1.F90.erb).  The two FP64 VALU kernels give the
 * same bits, and so do the MFMA shapes of one form among themselves; the
 * FP64 VALU and MFMA kernels differ by rounding only, both within the FP64
 * parity bar. */
int dgk_tile_shapes(int evaltype, int p, int *epb, int max, int *n);
int dgk_tile_select(int evaltype, int p, int cand);
int dgk_tile_selected(int evaltype, int p, int *cand);
int dgk_autotune(dgk_solver *s, int reps, int *best, double *us_per_launch, int max,
                 int *n);
/* Candidate the solver's launches use (its own, else the process-wide one). */
int dgk_solver_tile(const dgk_solver *s, int *cand);
int dgk_batch_autotune(dgk_batch *b, int reps, int *p, int *best, double *us_best,
                       double *us_default, int max, int *n);

/* ---- serving: continuous batching of advection solves -------------------
 * No counterpart in the reference, whose driver runs one mesh to completion
 * (main.f90:51-79); these entry points are for callers that serve many
 * solves at once.  A server owns one device arena of hbm_budget bytes, a
 * worker thread and nstreams HIP streams (default priority).  A request is a
 * periodic box with a Gaussian-blob initial condition and constant velocity
 * (the inputs of dgk_init_gaussian).  Requests wait in arrival order and are
 * admitted when their device footprint fits in the arena and one of
 * max_active slots is free; a request may be passed by later ones that fit,
 * at most 8 times, after which admission waits for it.  Every tick advances
 * each admitted request by SSP-RK3 steps on its stream (see
 * dgk_serve_set_policy), so a request that arrives while others run joins at
 * the next tick and its kernels overlap theirs.  A request runs the same
 * stage arithmetic on the same inputs as a dgk_solver of its box, so its
 * result does not depend on what else is admitted.  Its arena block holds
 * q only: the velocity is one constant per request, which the stage kernels
 * read from the request's geometry record.  All dgk_serve_* calls may come
 * from any thread. */
enum { DGK_TOL_FP64 = 0, /* plain FP64 path (HIP)                         */
       DGK_TOL_1E5 = 1   /* mixed precision (HIP_MP), 1e-5 relative L2     */ };

typedef struct dgk_request {
  int32_t ne[3];     /* elements per axis (>= 1)                           */
  int32_t p;         /* polynomial order 1..15                             */
  int32_t nstep;     /* SSP-RK3 steps (>= 0)                               */
  int32_t tol_class; /* DGK_TOL_FP64 or DGK_TOL_1E5                        */
  double dom[3];     /* box size per axis (> 0)                            */
  double dt;         /* time step; <= 0: CFL 0.1 on the smallest node gap,
                        dt = 0.1 * min_d (x1(1)-x1(0)) L_d/(2 Ne_d) / sum|vel| */
  double vel[3];     /* constant velocity                                  */
  double centre[3];  /* blob centre                                        */
  double sigma;      /* blob width (> 0)                                   */
  double *q_out;     /* optional host buffer of Np*Ne doubles that receives
                        the final q before the result is posted (NULL: the
                        summary only)                                      */
  int64_t tag;       /* caller's label, returned in the result             */
} dgk_request;

typedef struct dgk_result {
  int64_t id;        /* from dgk_serve_submit                              */
  int64_t tag;
  int32_t status;    /* DGK_OK or the failing call's code                  */
  int32_t evaltype;  /* DGK_EVAL_HIP or DGK_EVAL_HIP_MP                    */
  int32_t p, nstep;
  int64_t dof;       /* Np*Ne                                              */
  int64_t bytes;     /* device footprint                                   */
  double dt;
  double t_submit, t_admit, t_done; /* seconds since dgk_serve_create    */
  double qmin, qmax; /* of the final q                                     */
  double mass;       /* quadrature integral of the final q over the box    */
  double l2sq;       /* quadrature integral of q^2                         */
  char message[160]; /* error text when status != DGK_OK                   */
} dgk_result;

typedef struct dgk_serve_stats {
  int64_t submitted, admitted, completed, failed;
  int64_t queued, active, active_peak, ticks;
  int64_t bytes_in_use, bytes_peak, budget;
  int64_t bypasses;      /* admissions that passed a waiting request       */
  double dof_updates;    /* sum of dof * nstep of completed requests       */
} dgk_serve_stats;

typedef struct dgk_server dgk_server;
/* hbm_budget: arena bytes (allocated once); max_active >= 1; nstreams 1..8 */
int dgk_serve_create(const char *operator_dir, int64_t hbm_budget, int max_active,
                     int nstreams, dgk_server **out);
/* Device bytes a request of this shape occupies while admitted. */
int dgk_serve_footprint(const int32_t ne[3], int p, int tol_class, int64_t *bytes);
/* Queue a request (copied).  DGK_ERR_NOMEM when its footprint exceeds the
 * whole arena, DGK_ERR_ARG / DGK_ERR_IO for a malformed request or a missing
 * operator file, DGK_ERR_STATE after a device failure of the worker. */
int dgk_serve_submit(dgk_server *sv, const dgk_request *r, int64_t *id);
/* Move up to max finished results into out (completion order); waits up to
 * timeout_s seconds for the first one (0: do not wait). */
int dgk_serve_poll(dgk_server *sv, dgk_result *out, int max, int *n, double timeout_s);
/* Tick policy.  STEP_FAIR: one step per admitted request per tick.
 * TIME_FAIR (the default, max_steps_per_tick 64): about the same device time
 * per request per tick; a request of D DOF gets Q / max(D, 2^18) steps,
 * 1..max_steps_per_tick, with Q the largest max(D, 2^18) among the admitted
 * requests and at least 2^22.  Results do not depend on the policy. */
enum { DGK_SERVE_STEP_FAIR = 0, DGK_SERVE_TIME_FAIR = 1 };
int dgk_serve_set_policy(dgk_server *sv, int policy, int max_steps_per_tick);
int dgk_serve_get_stats(dgk_server *sv, dgk_serve_stats *st);
/* Finish every submitted request, stop the worker, free the arena.
 * Results not yet polled are discarded. */
int dgk_serve_destroy(dgk_server *sv);

/* ---- streaming calibration -----------------------------------------------
 * No counterpart in the reference.  The HBM rate, in GB/s of bytes read plus
 * written, of a kernel that reads five arrays of n doubles and writes one --
 * the stage kernels' mix of about five bytes read per byte written -- with
 * 16-byte loads into LDS (gbs_async, one per lane and array, as the stage
 * kernels stage their fields) and 8-byte loads into registers (gbs_regs, one
 * per lane and array, the fastest streaming form measured), non-temporal
 * stores, averaged over reps launches (after one warm-up) timed with HIP
 * events.  n is a multiple of 2048; the call
 * allocates 48 n bytes of device memory and frees them before returning. */
int dgk_calibrate_stream(int64_t n, int reps, double *gbs_async, double *gbs_regs);

#ifdef __cplusplus
}
#endif

#endif /* DGK_H */
