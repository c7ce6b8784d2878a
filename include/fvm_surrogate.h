/* Learned HLLD surrogate: C ABI.
 *
 * No counterpart in the reference.  It stands in for riemannSolver_HLLD
 * (computeFlux.cpp:178-425) inside computeFlux (computeFlux.cpp:6-148), with
 * the same batch interface as fvm_compute_flux (fvm_gpu.h) and a gated
 * fallback to the FP64 HLLD for every face the gate rejects.
 *
 * The network is 16 -> 128 -> 128 -> 128 -> 8 (ReLU), bfloat16 operands with
 * FP32 accumulation on the matrix cores.  Inputs are the HLLD inputs of a face
 * in the face frame after a Galilean boost (face speed along the normal, mean
 * tangential velocity) and the MHD scaling (density, fast-speed scale) are
 * divided out: the mean state of the two sides, the unit direction and the
 * size J of the normalised jump between them, the log density and pressure
 * ratios and the larger fast speed.  Outputs 0..6 are (G' - Fb)/J for the
 * normalised moving-frame fluxes G' (mass, three momenta, energy, two
 * tangential induction components), Fb the HLL flux of the two sides with the
 * HLLD's own outer wave speeds: the flux error scales with the jump.  Output 7
 * predicts log10 of the weighted squared error of outputs 0..6 (the
 * confidence gate).  The normal induction flux -w_n B_n, the Dedner B_n/psi,
 * and the psi and entropy fluxes are formed exactly as computeFlux forms them.
 *
 * Parameters are one FP32 vector, layer by layer: W0[16][128] b0[128]
 * W1[128][128] b1[128] W2[128][128] b2[128] W3[128][8] b3[8] (W[in][out]).
 *
 * All functions return FVM_OK (0) or a negative FVM_E* code (fvm_gpu.h);
 * fvm_sg_last_error gives the detail.  Host arrays are caller-owned; calls are
 * synchronous on the surrogate's own stream.
 */
#ifndef FVM_SURROGATE_H
#define FVM_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fvm_sg fvm_sg;

enum { FVM_SG_NIN = 16, FVM_SG_NH = 128, FVM_SG_NOUT = 8, FVM_SG_NPARAM = 36232 };

/* Faces the network may answer; the rest go to the FP64 HLLD.
 *   max_log_rho_jump, max_log_p_jump  bounds on |ln(rho_R/rho_L)|, |ln(p_R/p_L)|
 *   max_vn                            bound on |v_n - w_n| / c_s of both sides
 *   max_log10_err                     bound on the confidence output
 * Boundary faces (jbnd != FVM_NO_BOUNDARY) and non-positive states always fall back. */
typedef struct fvm_sg_gate {
  double max_log_rho_jump;
  double max_log_p_jump;
  double max_vn;
  double max_log10_err;
} fvm_sg_gate;

int fvm_sg_create(fvm_sg **s, int device, double gamma_gas);
void fvm_sg_destroy(fvm_sg *s);
const char *fvm_sg_last_error(const fvm_sg *s);

int fvm_sg_set_params(fvm_sg *s, const float *params, int n);
int fvm_sg_get_params(fvm_sg *s, float *params, int n);
int fvm_sg_set_gate(fvm_sg *s, const fvm_sg_gate *gate);

/* Training set: faces in the computeFlux batch layout (wij[n][3], Wi[n][10],
 * Wj[n][10], jbnd[n], normal[n][3]; host arrays).  Each face that passes the
 * domain part of the gate becomes one sample (16 features, 7 targets from the
 * FP64 HLLD); *n_kept receives how many did. */
int fvm_sg_add_faces(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                     const double *normal, int *n_kept);
int fvm_sg_clear_samples(fvm_sg *s);
int fvm_sg_num_samples(fvm_sg *s, long long *n);
int fvm_sg_get_samples(fvm_sg *s, long long first, int n, float *x, float *t);

/* Adam training on the device: `steps` minibatches of `batch` samples drawn by
 * a counter hash of (seed, step, slot); with e2 = sum_k w_k (y_k - t_k)^2 over
 * outputs 0..6, loss = mean e2 + lam * (y[7] - log10(e2 + 1e-10))^2.
 * loss_hist[steps] (may be NULL) receives each step's loss.  The weights w
 * (default 1) are set by fvm_sg_set_loss_weights. */
int fvm_sg_set_loss_weights(fvm_sg *s, const double *w, int n);
int fvm_sg_train(fvm_sg *s, int steps, int batch, double lr, double lam, unsigned long long seed, float *loss_hist);

/* One forward and backward pass over the given samples, no update: network
 * outputs y[n][8], gradients grads[FVM_SG_NPARAM] and the loss. */
int fvm_sg_gradients(fvm_sg *s, int n, const float *x, const float *t, double lam, float *y, float *grads,
                     double *loss);

/* The fused inference kernel on raw features x[n][16] -> y[n][8]. */
int fvm_sg_forward(fvm_sg *s, int n, const float *x, float *y);

/* computeFlux on a batch, surrogate + FP64 fallback.  Host arrays; the device
 * variant takes device pointers and runs on the surrogate's stream.
 * *n_fallback (may be NULL) receives the number of faces the FP64 HLLD
 * answered. */
int fvm_sg_compute_flux(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                        const double *normal, double *psi, double *Bn, double *flux, int *n_fallback);
int fvm_sg_compute_flux_device(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj,
                               const int *jbnd, const double *normal, double *psi, double *Bn, double *flux,
                               int *n_fallback);

/* The surrogate in the device-resident FP32 face path (fvm_fused_*, fvm_gpu.h):
 * after fvm_fused_setup(ctx, FVM_FUSED_FP32, 0), every later fvm_fused_run
 * evaluates each face's flux once, by this surrogate where its gate accepts the
 * face and by the FP32 HLLD (compute_flux) where it does not, in place of the
 * FP32 HLLD evaluated from both sides.  Everything else of the step is
 * unchanged (FP64 sums in each cell's face order, FP64 computePrimitives and
 * CFL).  The surrogate must live on the context's device and stay alive while
 * set; sg = NULL switches back.  A later fvm_fused_setup or mesh upload clears
 * it.  Static meshes only.  Diagnostics count the faces the gate rejected
 * (FVM_DIAG_SG_FALLBACK). */
typedef struct fvm_ctx fvm_ctx;
int fvm_fused_set_surrogate(fvm_ctx *ctx, fvm_sg *sg);

/* Kernel timing on device-resident copies of one batch: average milliseconds
 * per launch of the surrogate flux (featurise + network + gate, then the
 * fallback launch), of the plain FP64 HLLD (computeFlux in double) and of the
 * FP32 HLLD over the same faces.  ms[3] = {surrogate, hlld_f64, hlld_f32}. */
int fvm_sg_time_flux(fvm_sg *s, int n, const double *wij, const double *Wi, const double *Wj, const int *jbnd,
                     const double *normal, int iters, double *ms, int *n_fallback);

#ifdef __cplusplus
}
#endif
#endif
