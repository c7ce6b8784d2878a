/*
 * fvm_gpu.h -- C ABI of the MI355X Voronoi-face pipeline of fvmhd3d.
 *
 * The accelerated path is the reference's per-step fluid pipeline, five
 * non-virtual members of the Charm++ chare System (fvmhd3d.h:514-559) plus
 * System::computeFlux.  Each has one entry point here with the same meaning;
 * the System member arrays they read and write (fvmhd3d.h:343-406: the ptcl_act
 * pointer tables, cell_list, face_list, face_active_list) are uploaded as
 * plain arrays in the reference's index order:
 *
 *   [0, nactive)                          active cells (cell_list exists for these)
 *   [nactive, nactive + nimport_loc)      local inactive neighbours
 *   [nactive + nimport_loc, np)           remote ghosts (periodic images included)
 *
 * (localMesh.cpp:216-322).  Host arrays are arrays of the reference's own
 * records: Fluid = double[10], vec3 = double[3], Face = fvm_face (56 B).
 * Float-stored fields of the shipped build (Fluid_flt, fvec3) are passed as
 * doubles; the context's storage mode decides whether the device rounds them to
 * float at the reference's rounding points.
 *
 * Rules: every call returns an int status (FVM_OK or a negative code, detail in
 * fvm_last_error); no exceptions cross the ABI; host buffers are caller-owned
 * and device buffers library-owned; calls on one context are stream-ordered and
 * a context is used by one thread at a time (one context per chare/thread).
 */
#ifndef FVM_GPU_H
#define FVM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fvm_ctx fvm_ctx;

/* status codes */
#define FVM_OK 0
#define FVM_EARG (-1)    /* bad argument / size / missing upload */
#define FVM_EHIP (-2)    /* HIP runtime error */
#define FVM_EPOS (-3)    /* non-positive density or thermal energy at a face (the reference asserts, computeFluidUpdate.cpp:179-189) */
#define FVM_ESTATE (-4)  /* call out of order (e.g. no mesh uploaded) */

/* limiter modes: Phi of computeReconstruction.cpp:5-17 / slopeLimiter.cpp:5-18 */
#define FVM_LIMITER_SHIPPED 0 /* Phi = 1, the shipped code */
#define FVM_LIMITER_MINMOD 1  /* Phi = min(1, x), the dead branch of the same function */
#define FVM_LIMITER_SMOOTH 2  /* Phi of MPI/compute_reconstruction.cpp:5-15 */

/* storage of the reference's float fields (Fluid_flt, fvec3, FluidD; Fluid.h:140-206) */
#define FVM_STORAGE_F64 0     /* all-FP64: every field double */
#define FVM_STORAGE_SHIPPED 1 /* float, as the reference ships */

/* boundary kinds, MeshPoint.h:10 */
#define FVM_NO_BOUNDARY 0
#define FVM_REFLECTING 1
#define FVM_OUTFLOW 2
#define FVM_DIOD 3
#define FVM_INFLOW 4

typedef struct {
  double gamma_gas;      /* System::gamma_gas */
  double courant_no;     /* System::courant_no */
  double domain_size[3]; /* global_domain_size (periodic minimum image in the reconstruction) */
  int limiter_mode;      /* FVM_LIMITER_* */
  int storage;           /* FVM_STORAGE_* */
} fvm_params;

/* fvmhd3d::Face (fvmhd3d.h:262-284): n is the area vector oriented s1 -> s2 */
typedef struct {
  int s1, s2;
  double centroid[3];
  double n[3];
} fvm_face;

/* per-cell fields, host layout [np][width] */
enum fvm_field {
  FVM_FIELD_W = 0,          /* double[10] Fluid_rec::w */
  FVM_FIELD_GRAD_X = 1,     /* double[10] Fluid_rec::x */
  FVM_FIELD_GRAD_Y = 2,     /* double[10] Fluid_rec::y */
  FVM_FIELD_GRAD_Z = 3,     /* double[10] Fluid_rec::z */
  FVM_FIELD_DWDT = 4,       /* double[10] Fluid_rec::t */
  FVM_FIELD_POS = 5,        /* double[3]  Fluid_rec::pos */
  FVM_FIELD_VEL = 6,        /* double[3]  Fluid_rec::vel (face-velocity input) */
  FVM_FIELD_J = 7,          /* double[3]  Fluid_rec::J */
  FVM_FIELD_TEND = 8,       /* double     Fluid_rec::tend */
  FVM_FIELD_BND = 9,        /* int        Fluid_rec::bnd == MeshPoint::boundary */
  FVM_FIELD_IS_NGB = 10,    /* int        Particle::is_ngb() */
  FVM_FIELD_IS_ACTIVE = 11, /* int        Particle::is_active() */
  FVM_FIELD_MINMAX_LO = 12, /* double[10] Wrec_minmax.first */
  FVM_FIELD_MINMAX_HI = 13, /* double[10] Wrec_minmax.second */
  FVM_FIELD_LIMITER = 14,   /* double[10] slopeLimiter_act */
  FVM_FIELD_DU = 15,        /* double[10] FluidD::U */
  FVM_FIELD_DGRADPSI = 16,  /* double[3]  FluidD::gradPsi */
  FVM_FIELD_DDIVB = 17,     /* double     FluidD::divB */
  FVM_FIELD_U = 18,         /* double[10] U_act (conserved) */
  FVM_FIELD_MESH_VOLUME = 19, /* double   MeshPoint::Volume */
  FVM_FIELD_TBEG = 20,      /* double     MeshPoint::tbeg */
  FVM_FIELD_MESH_VEL = 21,  /* double[3]  MeshPoint::vel */
  FVM_FIELD_ACC0 = 22,      /* double[3]  MeshPoint::acc0 */
  FVM_FIELD_ACC1 = 23,      /* double[3]  MeshPoint::acc1 */
  FVM_FIELD_DT_NEW = 24,    /* double     MeshPoint::dt_new */
  FVM_FIELD_DIVB = 25,      /* double     divBi (output of computePrimitives) */
  FVM_NFIELDS = 26
};

/* diagnostics: how often a reference assert would have fired */
enum fvm_diag {
  FVM_DIAG_POSITIVITY = 0, /* computeFluidUpdate.cpp:179-189 */
  FVM_DIAG_CLOSURE = 1,    /* slopeLimiter.cpp:218-237 */
  FVM_DIAG_ETHM_FAILED = 2, /* entropy fallback, computeFluidUpdate.cpp:462-466 */
  FVM_DIAG_MASS_FAILED = 3, /* failed-mass fallback, computeFluidUpdate.cpp:377-386 */
  FVM_DIAG_DSL = 4,        /* dsl > 0, slopeLimiter.cpp:81, computeFluidUpdate.cpp:143-144 */
  FVM_DIAG_SG_FALLBACK = 5, /* surrogate flux step: faces the gate sent to the HLLD (no reference counterpart) */
  FVM_NDIAG = 8
};

/* ---- context (replaces the System chare's member storage) ---- */
int fvm_create(fvm_ctx **ctx, int device, const fvm_params *params);
int fvm_destroy(fvm_ctx *ctx);
const char *fvm_last_error(const fvm_ctx *ctx);
int fvm_set_params(fvm_ctx *ctx, const fvm_params *params);
/* System::t_global */
int fvm_set_time(fvm_ctx *ctx, double t_global);

/* ---- mesh of this step (System::localMesh_build output, localMesh.cpp:333-614) ----
 * faces: face_list; cell_face_off/cell_face: cell_list[i].ngb for i < nactive
 * (CSR, nactive+1 offsets); cell_volume: cell_list[i].Volume. */
int fvm_upload_mesh(fvm_ctx *ctx, int np, int nactive, int nimport_loc, int nface,
                    const fvm_face *faces, const int *cell_face_off, const int *cell_face,
                    const double *cell_volume);

/* ---- per-cell state (the pointer tables Wrec_act, mesh_act, U_act, dU_act, ...) ---- */
/* fvm_upload_mesh without the cell lists: the face records (converted to SoA
 * on the device) and each active cell's face list built on the device the way
 * localMesh.cpp:586-589 builds cell_list[i].ngb: every face is listed by its
 * s1 and, when s2 is active, by its s2, in ascending face index.  Every face's
 * s1 must be an active cell (the cell that created it). */
int fvm_upload_faces(fvm_ctx *ctx, int np, int nactive, int nimport_loc, int nface, const fvm_face *faces,
                     const double *cell_volume);
/* the uploaded mesh's cell lists: nactive+1 offsets, then offsets[nactive]
 * face indices (either pointer may be NULL) */
int fvm_download_cell_faces(fvm_ctx *ctx, int *cell_face_off, int *cell_face);
int fvm_upload_field(fvm_ctx *ctx, int field, const void *host, int np);
int fvm_download_field(fvm_ctx *ctx, int field, void *host, int np);

/* ---- the path (System members, fvmhd3d.h:514-559) ---- */
/* System::get_active_faces (Iterate.cpp:298-324); returns the count in *nface_active */
int fvm_get_active_faces(fvm_ctx *ctx, int *nface_active);
/* face_active_list given by the caller (face indices), replacing fvm_get_active_faces */
int fvm_set_active_faces(fvm_ctx *ctx, int nface_active, const int *face_active);
int fvm_download_active_faces(fvm_ctx *ctx, int *face_active, int nface_active);
/* System::computeReconstruction (computeReconstruction.cpp:42-264) */
int fvm_compute_reconstruction(fvm_ctx *ctx);
/* System::slopeLimiter compute part (slopeLimiter.cpp:27-237) followed by
 * System::computeFluidUpdateI (computeFluidUpdate.cpp:25-75) */
int fvm_slope_limiter(fvm_ctx *ctx);
/* System::computeFluidUpdate (computeFluidUpdate.cpp:88-247); returns FVM_EPOS
 * when a face state is not positive (the results are still written) */
int fvm_compute_fluid_update(fvm_ctx *ctx);
/* System::computePrimitives (computeFluidUpdate.cpp:347-542) */
int fvm_compute_primitives(fvm_ctx *ctx, int predictor_flag);
/* System::computeTimestep (computeTimestep.cpp:6-62); *dt_min = min over active cells */
int fvm_compute_timestep(fvm_ctx *ctx, double *dt_min);
/* System::computeTimestepLimiter (computeTimestep.cpp:64-110) */
int fvm_compute_timestep_limiter(fvm_ctx *ctx);
/* System::computePredictor (computePredictor.cpp:11-391); dtmax/min_rung are the
 * Scheduler's (Scheduler.h:15-25, 94-111) for the dt/128 probe */
int fvm_compute_predictor(fvm_ctx *ctx, double dtmax, int min_rung);
/* System::computeFlux for n independent faces (computeFlux.cpp:6-148), host
 * arrays: wij[n][3], Wi/Wj[n][10], jbnd[n], normal[n][3] -> psi[n], Bn[n], flux[n][10] */
int fvm_compute_flux(fvm_ctx *ctx, int n, const double *wij, const double *Wi, const double *Wj,
                     const int *jbnd, const double *normal, double *psi, double *Bn, double *flux);
/* fvm_compute_flux without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as fvm_compute_flux; the device staging is a
 * fixed array of FVM_FLUX_NOALLOC_BYTES in the library, one per GPU, held
 * until the process ends.  n > FVM_FLUX_NOALLOC_MAX is refused with FVM_EARG.
 * Calls that overlap on one GPU (from any context on that device) share that
 * array and race: serialise them. */
#define FVM_FLUX_BYTES_PER_FACE 308 /* wij 3, Wi 10, Wj 10, normal 3, psi, Bn, flux 10 doubles + jbnd int */
#define FVM_FLUX_NOALLOC_MAX 65536
#define FVM_FLUX_NOALLOC_BYTES (FVM_FLUX_BYTES_PER_FACE * FVM_FLUX_NOALLOC_MAX)
int fvm_compute_flux_noalloc(fvm_ctx *ctx, int n, const double *wij, const double *Wi, const double *Wj,
                             const int *jbnd, const double *normal, double *psi, double *Bn, double *flux);
/* The arguments computeFluidUpdate passes to computeFlux (computeFluidUpdate.cpp:99-211:
 * boundary swap, space-time extrapolated face states, face velocity) for active
 * faces [first, first + count) of the active list, in fvm_compute_flux's layout.
 * Needs the reconstruction (and limiter) of the current step. */
int fvm_get_face_states(fvm_ctx *ctx, int first, int count, double *wij, double *Wi, double *Wj, int *jbnd,
                        double *normal);

/* ---- diagnostics and snapshots ------------------------------------------- */
/* Energy totals over the active cells (Energy.h:27-79), from the U, MESH_VOLUME
 * and POS fields, in Energy's order: MASS ETHM EPOT ETOT EKINX EKINY EKINZ EKIN
 * EMAGX EMAGY EMAGZ EMAG MOMX MOMY MOMZ MOM ANMX ANMY ANMZ ANM VOLUME.  EPOT is 0
 * as in the reference.  Deterministic (fixed summation order). */
#define FVM_ENERGY_NVAR 21
int fvm_energy(fvm_ctx *ctx, double *out);
/* System::dump_binary (IO.cpp:7-99): one per-chare float32 snapshot file, a
 * 20-word header and one 26-word record per active cell, each wrapped in
 * 4-byte length markers.  idx (global cell ids) and dt_cell (each cell's rung
 * dt) may be NULL: ids 0..n-1 and dt_global. */
typedef struct fvm_snapshot_info {
  int chare_index;   /* thisIndex */
  int num_elements;  /* numElements */
  int global_n;
  int iteration;
  unsigned long long tsysU;
  double dt_global;
  double rmin[3], rmax[3];
} fvm_snapshot_info;
int fvm_dump_binary(fvm_ctx *ctx, const char *path, const fvm_snapshot_info *info, const int *idx,
                    const double *dt_cell);

/* ---- device-resident stepping: static mesh, every cell active, one global
 * timestep (all cells on the smallest rung of the Scheduler, dt = rung of the
 * minimum CFL step).  No counterpart entry point in the reference: it chains
 * computeTimestep -> computePredictor -> computeReconstruction ->
 * slopeLimiter -> computeFluidUpdate -> computePrimitives (Iterate.cpp:92-202)
 * on the device with the state resident in HBM.  Uses the mesh and the fields
 * W, U, POS, MESH_VEL uploaded before fvm_fused_setup; W and U are updated in
 * place. */
#define FVM_FUSED_FP64 0 /* FP64 face arithmetic: the plain FP64 path */
#define FVM_FUSED_FP32 1 /* FP32 face reconstruction + HLLD; FP64 flux sums, pressure recovery, CFL */
int fvm_fused_setup(fvm_ctx *ctx, int precision, int moving_mesh);
int fvm_fused_run(fvm_ctx *ctx, int nsteps, double dtmax, double *t_out);
/* dt of each step of the last fvm_fused_run */
int fvm_fused_dt_history(fvm_ctx *ctx, double *dts, int n);
/* copy the last step's gradients, Wt and dt_new into the per-cell fields */
int fvm_fused_sync(fvm_ctx *ctx);
/* per-kernel HIP-event timing of fvm_fused_run on the context's stream:
 * ms[0] CFL+predictor+reconstruction (FP32 path: with the opt-in limiters),
 * ms[1] FP32 path: the second launch that checks the limiters' dt and redoes
 *       the step when it changed; FP64 path: the limiter launch,
 * ms[2] flux+apply, ms[3] advance */
int fvm_fused_timing(fvm_ctx *ctx, int enable);
int fvm_fused_kernel_ms(fvm_ctx *ctx, double *ms, long long *launches, int n);
/* per-kernel HIP-event timing of the reference-order entry points (and so of
 * fvm_substep_run): with timing enabled, each launch of the kernels below is
 * bracketed by events on the context's stream and waited for, so enable it only
 * to measure.  fvm_kernel_ms returns the total milliseconds and launch counts
 * since timing was (re-)enabled, indexed by FVM_KT_*.  No counterpart in the
 * reference. */
#define FVM_KT_RECONSTRUCTION 0  /* k_reconstruction (computeReconstruction) */
#define FVM_KT_SLOPE_LIMITER 1   /* k_slope_limiter (slopeLimiter + computeFluidUpdateI) */
#define FVM_KT_FACE_FLUX 2       /* k_face_flux (computeFluidUpdate, face part) */
#define FVM_KT_SCATTER 3         /* k_scatter (computeFluidUpdate, per-cell sums) */
#define FVM_KT_PRIMITIVES 4      /* k_primitives (computePrimitives) */
#define FVM_KT_TIMESTEP 5        /* k_timestep (computeTimestep) */
#define FVM_KT_PREDICTOR_FACE 6  /* k_predictor_face (computePredictor, face part) */
#define FVM_KT_PREDICTOR_CELL 7  /* k_predictor_cell (computePredictor, cell part) */
#define FVM_NKT 8
/* Face arithmetic of fvm_compute_fluid_update and fvm_compute_predictor (and so
 * of the substeps, fvm_substep_run): bits = 64 (default) is the reference's
 * FP64; bits = 32 forms the face states (W + Wt dti + grad . dr, with dt, dti
 * and the geometry in FP64) and runs computeFlux's HLLD in FP32, while the
 * records times dt and area, their per-cell sums, pressure recovery and the
 * CFL step stay FP64.  No counterpart in the reference: it selects the FP32
 * face arithmetic of the mixed-precision plan for the individual-timestep
 * calls.  FVM_EARG for any other bits. */
int fvm_set_face_precision(fvm_ctx *ctx, int bits);

int fvm_kernel_timing(fvm_ctx *ctx, int enable);
int fvm_kernel_ms(fvm_ctx *ctx, double *ms, long long *launches, int n);

/* Individual timesteps with the per-cell state resident on the device, in
 * mesh order.  One fvm_substep_run is one substep of the reference's driver
 * loop (Iterate.cpp:92-202): the local layout of active cells and their
 * inactive neighbours with faces turned so that s1 is active
 * (localMesh.cpp:216-322, :497-591), the prediction of the neighbours to the
 * current time (Iterate.cpp:99-106), the path (computePrimitives,
 * computeTimestep, computeTimestepLimiter, computePredictor,
 * computeReconstruction, slopeLimiter, computeFluidUpdate), the new rungs of
 * the active cells (Scheduler.h get_rung, Iterate.cpp:146-168) and the return
 * of the neighbours' dU to their owners (computeFluidUpdate.cpp:320-345).  The
 * rung clock (Scheduler.h pull) stays with the caller, which passes the
 * bitmask of the rungs that are active (bit r = rung r), the time, dtmax and
 * min_rung.  The mesh is the whole periodic mesh; it and the context's mesh
 * and fields are overwritten.  Needs FVM_STORAGE_F64.
 *   counts[5]: active cells, active faces, inactive neighbours, local faces,
 *              1 if the fluid update reported a non-positive state
 *   rung_hist[32]: cells per rung after the substep */
#define FVM_SUBSTEP_RUNG FVM_NFIELDS /* int rung per cell, for upload/download */
int fvm_substep_setup(fvm_ctx *ctx, int ncell, int nface, const fvm_face *faces, const int *cell_face_off,
                      const int *cell_face, const double *cell_volume);
/* per-cell fields kept between substeps (W, gradients, Wt, POS, J, TEND, BND,
 * MINMAX_LO/HI, DU, DGRADPSI, DDIVB, U, MESH_VOLUME, TBEG, DT_NEW) and the
 * rungs, host layout [ncell][width] in mesh order */
int fvm_substep_upload(fvm_ctx *ctx, int field, const void *host, int ncell);
int fvm_substep_download(fvm_ctx *ctx, int field, void *host, int ncell);
/* first rungs from the CFL timestep of every cell (System.cpp:255-300, per cell) */
int fvm_substep_init_rungs(fvm_ctx *ctx, double dtmax, int nrung, int *rung_hist);
int fvm_substep_run(fvm_ctx *ctx, unsigned rung_mask, double t_global, double dtmax, int min_rung, int nrung,
                    int *counts, int *rung_hist);

/* The sequential timestep-limiter sweep of computeTimestep.cpp:64-110 on host
 * arrays of a whole mesh (no context, no GPU): active cells in ascending index,
 * each cell's neighbours nbr[cell_off[i] .. cell_off[i+1]) in list order,
 * dti[j] = min(dti[j], 2 dti[i]) then dti[i] = min(dti[i], 2 dti[j]).  Visits
 * only the cells a change can reach (the same routine fvm_substep_run uses);
 * the result equals the full sweep's.  dti is updated in place. */
int fvm_timestep_limiter_sweep(int ncell, const int *cell_off, const int *nbr, const unsigned char *active,
                               double *dti);

/* read+write bandwidth of a plain 16-byte-vector copy of `bytes` bytes, GB/s */
int fvm_measure_copy_bandwidth(fvm_ctx *ctx, long long bytes, int iters, double *gbps);
/* the same measurement with the HIP runtime's own device-to-device copy
 * (hipMemcpyAsync) in place of the library's copy kernel: a second opinion on
 * what a copy reaches on the device.  No counterpart in the reference. */
int fvm_measure_runtime_copy_bandwidth(fvm_ctx *ctx, long long bytes, int iters, double *gbps);

/* diagnostics counters accumulated since the last reset */
int fvm_get_diagnostics(fvm_ctx *ctx, long long *out, int n);
int fvm_reset_diagnostics(fvm_ctx *ctx);
int fvm_synchronize(fvm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* FVM_GPU_H */
