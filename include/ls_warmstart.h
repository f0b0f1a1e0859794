/* C ABI of the learned warm start (part of liblippmannschwinger.so): the
 * kernels that evaluate and train a periodic 3-D convolutional
 * encoder-decoder in BF16 on MFMA with FP32 accumulation.  None of these has
 * a counterpart in the reference: its solver always starts from the uniform
 * strain (src/jaxmaterials/solver/lippmann_schwinger.py:181-183,
 * cuda/src/lippmann_schwinger.cu:8-16, 181).
 *
 * All pointers are device pointers; `stream` is a hipStream_t (NULL = the
 * default stream); every call is asynchronous on that stream and returns 0,
 * or -1 after printing the reason for invalid arguments or a launch error.
 *
 * Tensors: bf16 (uint16_t) activations and gradients are channels last,
 * [instance][x][y][z][channel] with channel stride ld (a multiple of 8, base
 * pointers 16-byte aligned).  dims = {instances, nx, ny, nz}.  Master
 * weights are fp32 [cout][27][cin], tap = 9 (dx+1) + 3 (dy+1) + (dz+1), all
 * convolutions 3x3x3 and periodic.
 */
#ifndef LS_WARMSTART_H
#define LS_WARMSTART_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flags of ls_ws_conv3d. */
#define LS_WS_ACT 1    /* leaky ReLU (slope 0.1) on the output */
#define LS_WS_DACT 2   /* multiply the output by the leaky ReLU derivative of tsrc */
#define LS_WS_PLANAR 4 /* fp32 output [instance][cvalid][voxel] plus addc[instance][c] */

/* y = act(conv(x, W) + bias).  wt = packed bf16 weights, 16*ceil(cout/16)
 * rows of kpad = roundup(27*cin, 32) entries (ls_ws_pack).  cin a multiple of
 * 8; cout <= 192 and a multiple of 4.  With LS_WS_DACT the call computes a
 * data gradient: x = output gradient, wt = the data-gradient packing, tsrc =
 * the forward input (channel stride ldt). */
int ls_ws_conv3d(const int *dims, const uint16_t *x, int ldx, int cin,
                 const uint16_t *wt, int kpad, const float *bias, int cout,
                 void *out, int ldo, int flags, const uint16_t *tsrc, int ldt,
                 const float *addc, int cvalid, void *stream);

/* Weight and bias gradients of one convolution: dw[cout][27][cin] =
 * sum_v dz[v][co] x[v + tap][ci], db[cout] = sum_v dz[v][co] (db may be NULL).
 * cout, cin multiples of 8, the voxel count a multiple of 8.  planar =
 * bf16 scratch of
 * planar_cap >= (cout + cin) * voxels elements (planar copies of dz and x);
 * part = fp32 scratch of part_cap floats (deterministic two-stage sums). */
int ls_ws_conv3d_wgrad(const int *dims, const uint16_t *dz, int ldz, int cout,
                       const uint16_t *x, int ldx, int cin, uint16_t *planar,
                       long long planar_cap, float *part, long long part_cap,
                       float *dw, float *db, void *stream);

/* 2x average pooling; dims = the fine grid (even sizes). */
int ls_ws_pool(const int *dims, const uint16_t *in, int ldi, int c,
               uint16_t *out, int ldo, void *stream);

/* Gradient of a skip tensor s used by a pooling and a concatenation:
 * dz = act'(s) * upsample(dcoarse) / 8 + dcat (dcat may be NULL). */
int ls_ws_pool_backward(const int *dims, const uint16_t *dcoarse, int ldc,
                        int c, const uint16_t *s, int lds,
                        const uint16_t *dcat, int ldcat, uint16_t *dz, int ldz,
                        void *stream);

/* Nearest 2x upsampling into a fine tensor (dims = the fine grid). */
int ls_ws_upsample(const int *dims, const uint16_t *in, int ldi, int c,
                   uint16_t *out, int ldo, void *stream);

/* Gradient of the upsampling: sum of the 8 children (dims = the fine grid). */
int ls_ws_upsample_backward(const int *dims, const uint16_t *dfine, int ldf,
                            int c, uint16_t *dcoarse, int ldc, void *stream);

/* Network input [B][voxel][8]: (log lambda - norm[0]) * norm[1],
 * (log mu - norm[2]) * norm[3], eps_bar[b][0..5]; instance b reads the Lame
 * fields of material b / per_mat ([nmat][nper] fp32). */
int ls_ws_input(const float *lam, const float *mu, int per_mat,
                const float *eps_bar, int B, long long nper, const float *norm,
                uint16_t *out, void *stream);

/* Training target: bf16 [B][6][nper] = eps - eps_bar (eps fp32 [B][6][nper]). */
int ls_ws_target(const float *eps, const float *eps_bar, int B, long long nper,
                 uint16_t *out, void *stream);

/* Squared error of pred (fp32 [B][6][nper]) against target (bf16): loss[0]
 * (FP64, device) = sum of squares; grad [B*nper][8] bf16 = gscale * (pred -
 * target), channels 6 and 7 zero.  part = FP64 scratch of part_cap. */
int ls_ws_mse(const float *pred, const uint16_t *target, int B, long long nper,
              float gscale, uint16_t *grad, double *part, long long part_cap,
              double *loss, void *stream);

/* Compatible strains from predicted displacements: per load case k of S
 * samples, eps_k = sym(D^+ u_k), D^+ = -(D^-)^T the forward derivative of the
 * rotated staggered grid (D^+_d = i xi~_d times a phase common to the axes
 * in Fourier, so eps lies in the range of the solver's Gamma0 and has zero
 * mean).  dims = {S, nx, ny, nz}; u fp32 [S][18][nper] (3 k + c); eps fp32
 * [S][36][nper] (6 k + c, Voigt with tensor shear), plus e_k if add_unit. */
int ls_ws_symgrad(const int *dims, const float *extents, const float *u,
                  int add_unit, float *eps, void *stream);

/* Adjoint of ls_ws_symgrad: grad_u bf16 [S][nper][ldu] (channel 3 k + c,
 * channels >= 18 zero) from grad_eps bf16 [S][nper][ldg] (channel 6 k + c). */
int ls_ws_symgrad_backward(const int *dims, const float *extents,
                           const uint16_t *grad_eps, int ldg,
                           uint16_t *grad_u, int ldu, void *stream);

/* Training loss on the solver's own terms for S samples of the six unit load
 * cases: dims = {S, nx, ny, nz}, extents = {Lx, Ly, Lz}; pred fp32 and target
 * bf16 [S][36][nper] (strain fluctuation, channel 6 k + c = component c of
 * load case k); lam, mu fp32 [S][nper].  With e = pred - target,
 *   L = sum_{s,k} [ wm <e, C e>_w / <t, C t>_w + wr |D^- . C e|^2 / |D^- . C e_k|^2 ] / (6 S):
 * the relative strain energy of the error (the norm conjugate gradients
 * reduce) and the squared equilibrium residual of the predicted strain
 * relative to the uniform start's (the solver's stop measure; the target is
 * in equilibrium), D^- the backward divergence of the rotated staggered grid
 * (src/jaxmaterials/solver/derivatives.py:8-96).  grad bf16 [S][nper][ldg] =
 * dL/dpred (ldg >= 36, a multiple of 8; channels >= 36 zero).  scratch fp32
 * of >= 54 S nper floats, part FP64 of >= 24 S ceil(nper / 256), scale fp32
 * [S][6][2] (the two gradient factors), loss FP64 [1] (device). */
int ls_ws_residual_loss(const int *dims, const float *extents,
                        const float *pred, const uint16_t *target,
                        const float *lam, const float *mu, float wm,
                        float wr, uint16_t *grad, int ldg, float *scratch,
                        long long scratch_cap, double *part,
                        long long part_cap, float *scale, double *loss,
                        void *stream);

/* Adam update of n fp32 parameters (step >= 1, bias-corrected). */
int ls_ws_adam(float *w, const float *g, float *m, float *v, long long n,
               float lr, float b1, float b2, float eps, int step, void *stream);

/* Pack fp32 W[cout][27][cin] into the bf16 forward layout wt (rows_t rows
 * of kpad = roundup(27*cin, 32), K index tap*cin + ci) and, if wd != NULL,
 * the data-gradient layout wd (rows_d rows of kpad_d = roundup(27*cs, 32),
 * K index tap'*cs + co with the tap mirrored, cs = channel stride of the
 * output gradient). */
int ls_ws_pack(const float *w, int cout, int cin, int cs, uint16_t *wt,
               int kpad, int rows_t, uint16_t *wd, int kpad_d, int rows_d,
               void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LS_WARMSTART_H */
