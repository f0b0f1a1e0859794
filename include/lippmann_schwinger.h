/* C ABI of liblippmannschwinger.so, the MI355X (gfx950) Lippmann-Schwinger
 * solver.  All functions are extern "C", take plain pointers and sizes, and
 * run synchronously on the calling thread's current HIP device.
 *
 * Conventions (same as the reference solver):
 *   - scalar fields are float32 [nx][ny][nz], C order (z fastest);
 *   - strain / stress fields are float32 [6][nx][ny][nz], Voigt order
 *     (xx, yy, zz, xy, xz, yz) with tensor (not engineering) shear components;
 *   - voxels = {nx, ny, nz} (int32), extents = {Lx, Ly, Lz} (float32).
 */
#ifndef LIPPMANN_SCHWINGER_H
#define LIPPMANN_SCHWINGER_H

#ifdef __cplusplus
extern "C" {
#endif

/* Drop-in replacement of the reference entry point
 * (cuda/src/lippmann_schwinger.hh:176-184, implemented at
 * cuda/src/lippmann_schwinger.cu:236-258; bound by ctypes in
 * src/jaxmaterials/solver/lippmann_schwinger.py:268-329).
 *
 * Host pointers.  Solves for the strain with mean epsilon_bar[6] in the medium
 * with Lame fields lambda[N], mu[N]; writes epsilon and sigma ([6][N] each).
 * Stops when the relative residual R < max(rtol * R0, atol), R0 = residual of
 * the uniform start.  Returns the number of strain updates; maxiter means not
 * converged (sigma is then the stress before the last update, as in the
 * reference).  Differences: a HIP error or invalid argument prints a message
 * to stderr and returns -1 (the reference prints and carries on); a
 * non-finite residual stops the loop and returns maxiter. */
int lippmann_schwinger_solve(float *lambda, float *mu, float *epsilon_bar,
                             float *epsilon, float *sigma, int *voxels,
                             float *extents, float rtol, float atol,
                             int maxiter, int verbose);

/* ---- additive entry points (no counterpart in the reference) ---------- */

/* Effective 6x6 tensor of one microstructure (host pointers).  Row k is the
 * volume average of sigma under epsilon_bar = e_k (the C_eff loop of
 * src/jaxmaterials/data/data.py:178-186), all six load cases batched in one
 * solve; averages in FP64.  depth = Anderson depth (0 = plain iteration).
 * iters[6] receives the iterations per load case.  Returns 0 on success,
 * 1 if a load case did not converge, -1 on error. */
int lippmann_schwinger_effective_tensor(const float *lambda, const float *mu,
                                        const int *voxels, const float *extents,
                                        float rtol, float atol, int maxiter,
                                        int depth, int verbose, double *ceff,
                                        int *iters);

/* Reusable solver on device memory (used by the Python/PyTorch layer).
 * nrhs load cases share nmat Lame-field pairs; depth = Anderson depth (<= 8).
 * stream = hipStream_t to run on (NULL: the solver creates its own). */
typedef struct ls_solver ls_solver;

ls_solver *ls_solver_create(const int *voxels, const float *extents, int nrhs,
                            int nmat, int depth, void *stream);

/* Solve all load cases.  Device pointers: d_lambda/d_mu [nmat][N];
 * d_tau [nrhs][6][N] or NULL (polarisation added to the stress);
 * d_eps [nrhs][6][N] in/out (input = warm start when warm_start != 0, it is
 * projected onto compatible fields with mean eps_bar and replaced by the
 * uniform start if its equilibrium residual norm ||xi . sigma_hat|| exceeds
 * the uniform start's); d_sigma
 * [nrhs][6][N] or NULL.  Host arrays: rhs_mat[nrhs] (material of each load
 * case), eps_bar[nrhs][6]; outputs iters[nrhs], residual[nrhs] (final
 * relative residual), sigma_mean[nrhs][6] (FP64 volume averages), each may
 * be NULL.  The relative tolerance is measured against the residual of the
 * uniform start R(eps_bar) in every case; rel_ref > 0 overrides it.
 * flags: LS_ABSOLUTE_RESIDUAL = residual ||xi . sigma_hat|| without the
 * division by ||sigma_hat(0)|| (for polarisation-driven solves such as the
 * adjoint, whose stress may have zero mean); LS_GRAPH = run the iterations
 * between two host synchronisations as one replay of a captured HIP graph
 * (captured on first use, again when the fields or tolerances change; the
 * results are identical to the launch-by-launch path); LS_CG = conjugate
 * gradients instead of the Lippmann-Schwinger / Anderson update (the same
 * equation, residual, stop test and iteration count semantics: Gamma0 C is
 * self-adjoint and positive definite in the reference medium's inner
 * product on compatible fields; needs Hooke's law, any depth);
 * LS_COMPATIBLE_START = the warm start in d_eps is compatible with mean
 * eps_bar already (e.g. eps_bar + sym(D^+ u) on the rotated grid, as the
 * learned warm start writes it), so its projection is skipped.  Returns 0 if
 * all converged, 1 if some did not, -1 on error. */
#define LS_ABSOLUTE_RESIDUAL 1
#define LS_GRAPH 2
#define LS_CG 4
#define LS_COMPATIBLE_START 8
int ls_solver_solve(ls_solver *s, const float *d_lambda, const float *d_mu,
                    const int *rhs_mat, const float *eps_bar,
                    const float *d_tau, float *d_eps, float *d_sigma,
                    int warm_start, float rtol, float atol, int maxiter,
                    double rel_ref, int flags, int verbose, int *iters,
                    double *residual, double *sigma_mean);

/* Autotuner: time every candidate tile of the z-line, y and x passes (lines
 * per block, register elements and threads of the LDS FFT engine, and the
 * register-FFT kernels compiled for the axis length) on the solver's own
 * buffers and keep the fastest for later solves.  d_lambda, d_mu [nmat][N] and d_eps
 * [nrhs][6][N] are device buffers (d_eps is overwritten); reps launches per
 * candidate.  report (may be NULL) receives the chosen tiles and times.
 * Returns 0, or -1 on error. */
int ls_solver_autotune(ls_solver *s, const float *d_lambda, const float *d_mu,
                       float *d_eps, int reps, char *report, int report_cap);

/* Kernel time of the last solve (recorded with HIP events on the solver's
 * stream when the environment variable LS_TIMING is set), ms and launch
 * counts per pass kind: [0] z-line passes (plain), [1] y passes, [2] x passes,
 * [3] everything else, [4] z-line Anderson passes (update, mixing, Hooke,
 * forward transform), [5] Anderson Gram-row reductions, [6] z-line
 * conjugate-gradient passes.  Fills the first
 * nkinds kinds; returns the number of base re-evaluations of the last solve. */
int ls_solver_last_timing(ls_solver *s, double *ms, int *count, int nkinds);

void ls_solver_destroy(ls_solver *s);

/* Real-space relative divergence of a stress field (device pointers):
 * ||D^- . sigma|| / (sqrt(N) ||<sigma>||_V) with the backward derivative of
 * the rotated staggered grid, the reference's `relative_divergence`
 * (src/jaxmaterials/solver/lippmann_schwinger.py:103-115) over the stencil of
 * src/jaxmaterials/solver/derivatives.py:8-96 and
 * cuda/src/derivatives.cu:9-124.  d_sigma [6][N]; d_div [3][N] receives
 * D^- . sigma (may be NULL); *rel the relative divergence (sums in FP64).
 * stream = hipStream_t (NULL: the default stream); synchronous.  Returns 0,
 * or -1 on error.  An independent check of the solver's Fourier residual,
 * equal to it by Parseval. */
int ls_relative_divergence(const int *voxels, const float *extents,
                           const float *d_sigma, float *d_div, double *rel,
                           void *stream);

/* Seeded cross-ply fibre microstructure (random sequential adsorption) and
 * its rasterisation into Lame fields on the GPU.
 *   nlayers layers along z; layer k has fibres along x (k odd) or y (k even);
 *   fibre radius, target volume fraction vf per layer; seed of the
 *   splitmix64 stream.  Writes d_lambda, d_mu [N] (device pointers) and
 *   returns the number of fibres placed (or -1 on error); achieved_vf
 *   receives the rasterised fibre volume fraction. */
int ls_crossply_lame(const int *voxels, const float *extents, int nlayers,
                     double radius, double vf, unsigned long long seed,
                     double lam_matrix, double mu_matrix, double lam_fibre,
                     double mu_fibre, float *d_lambda, float *d_mu,
                     double *achieved_vf);

/* Host-only RSA placement (same stream and rules as ls_crossply_lame).
 * Writes up to cap fibres as (layer, in-plane coordinate, z) triples into
 * out[cap][3]; returns the count. */
int ls_crossply_place(const int *voxels, const float *extents, int nlayers,
                      double radius, double vf, unsigned long long seed,
                      double *out, int cap);

/* Lame fields of the reference's layered-fibre generator, rasterised on the
 * GPU (LayeredFibresDatasetGenerator.material_properties,
 * src/jaxmaterials/data/data.py:267-314; the reference evaluates it with
 * NumPy on the host).  voxels = {nx, ny, nz}; extents = {Lx, Ly, Lz} in
 * float64 (the reference's grid spec holds Python floats).  fibres[n][4] =
 * (in-plane coordinate p0, z coordinate p1, radius r, orientation): every
 * fibre of every layer, boundary copies included, as generate_fibre_positions
 * returns them (data.py:188-265); orientation 0 tests p0 against x, 1
 * against y.  mu3 / lambda3 = (void, fibre, matrix) values.  d_out = device
 * [2][nx][ny][nz] of float32 (dtype_bytes 4) or float64 (8): channel 0 mu,
 * channel 1 lambda; bit-identical to the reference in both dtypes.
 * stream = hipStream_t (NULL: default stream); returns after the fields are
 * written.  0 on success, -1 on error. */
int ls_rasterise_layered_fibres(const int *voxels, const double *extents,
                                double d_void, int nfibres,
                                const double *fibres, const double *mu3,
                                const double *lambda3, int dtype_bytes,
                                void *d_out, void *stream);

/* ---- phase-field fracture (src/jaxmaterials/solver/lippmann_schwinger.py:36-100, 334-547) */

/* Material law of the handle's next solves: the damaged law of
 * compute_sigma_damaged (lippmann_schwinger.py:36-74),
 * sigma = g sigma+ + sigma-, g = (1 - d)^2 + k, with the tension part from
 * the eigen-split of the strain matrix as the reference forms it; d_damage =
 * device [nmat][N] damage fields (float32), or NULL for Hooke's law.  With
 * it, ls_solver_solve performs elasticity_solve (lippmann_schwinger.py:
 * 334-462): the same iteration, stop test and Anderson mixing, the reference
 * medium still from lambda and mu.  The field must stay valid while set.
 * 0 on success. */
int ls_solver_set_damage(ls_solver *h, const float *d_damage, float k);

/* phase_field_solve (lippmann_schwinger.py:465-547) on the handle's grid, in
 * FP64: the fixed point d = IFFT[FFT(chi) / (A0 + |xi~|^2)],
 * chi = B - (A - A0) d, A = 1/lc^2 + 2H/(gc lc), B = 2H/(gc lc),
 * A0 = (min A + max A) / 2, from chi = B - (A - A0) d_old until
 * ||chi_new - chi|| / ||chi_new|| <= tol or maxiter iterations.  Device
 * FP64 fields d_H (history), d_dold, d_out [N]; the iteration count and the
 * last residual go to *iters and *residual.  extents: host double[3], the
 * domain size in FP64 as the reference builds xi~ for this solve (NULL: the
 * handle's float extents).  Every axis <= 3840.  0 on success. */
int ls_solver_phase_field(ls_solver *h, const double *d_H, const double *d_dold,
                          double gc, double lc, double tol, int maxiter,
                          double *d_out, int *iters, double *residual,
                          const double *extents);

/* compute_sigma_damaged (lippmann_schwinger.py:36-74) pointwise on device
 * arrays: d_eps, d_sigma [6][N] float32, d_lambda, d_mu [N], d_damage [N]
 * or NULL (undamaged: g = 1).  Synchronous on `stream`. */
int ls_damaged_stress(const int *voxels, const float *d_lambda,
                      const float *d_mu, const float *d_eps,
                      const float *d_damage, float k, float *d_sigma,
                      void *stream);

/* compute_strain_energy (lippmann_schwinger.py:77-100): psi+ per voxel into
 * d_psi (FP64 [N], may be NULL) and, if d_history is not NULL, the history
 * update H = max(H, psi+) of the staggered scheme (tests/phasefield.py:
 * 221-223).  Synchronous on `stream`. */
int ls_strain_energy(const int *voxels, const float *d_lambda, const float *d_mu,
                     const float *d_eps, double *d_psi, double *d_history,
                     void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same solve as ls_solver_create(voxels, extents,
 * nrhs, 1, depth, stream) + ls_solver_solve(..., rhs_mat NULL, d_tau NULL,
 * warm_start 0, rel_ref 0, flags 0, verbose 0, ...) + ls_solver_destroy, with
 * every internal buffer (spectrum, strain base, base divergence, Anderson
 * rings, reduction partials, FFT tables) carved from one file-scope device
 * array of LS_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from its
 * load until the process ends.  It calls no hipMalloc, hipFree, hipHostMalloc
 * or other allocating HIP routine and creates no stream, event or graph; the
 * host-side plan is rebuilt (ordinary heap memory) only when the grid,
 * extents, nrhs or depth differ from the previous call on the device.
 * Device pointers d_lambda, d_mu [N], d_eps [nrhs][6][N] (out), d_sigma
 * [nrhs][6][N] or NULL; host eps_bar[nrhs][6]; iters, residual, sigma_mean as
 * in ls_solver_solve (may be NULL).  stream = hipStream_t (NULL: the null
 * stream); synchronous.  Returns 0 / 1 (not all converged) / -1 on error,
 * including a problem whose buffers exceed the array (the message states the
 * bytes needed).  Every call on a GPU uses that GPU's one array, which keeps
 * the last call's contents (the solve initialises everything it reads).  Two
 * calls that overlapped on one GPU would race on the array; the library
 * serialises the calls of all threads of a process with a lock and returns
 * only after the solve's work on `stream` has finished, so they cannot
 * overlap within a process (each process has its own copy of the array). */
#define LS_NOALLOC_SCRATCH_BYTES (128ull << 20)
int ls_solve_noalloc(const int *voxels, const float *extents, int nrhs,
                     int depth, const float *d_lambda, const float *d_mu,
                     const float *eps_bar, float *d_eps, float *d_sigma,
                     float rtol, float atol, int maxiter, int *iters,
                     double *residual, double *sigma_mean, void *stream);

/* Measured HBM streaming ceilings of the current device (a measurement aid,
 * no counterpart in the reference): grid-stride float4 kernels with
 * non-temporal loads and stores over `bytes`-sized arrays (three are
 * allocated), best of `reps` runs each;
 * gbps[4] = GB/s of copy (read + write counted), read only, write only, and
 * two reads + one write.  0 on success, -1 on error. */
int ls_hbm_stream_probe(long long bytes, int reps, double *gbps, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LIPPMANN_SCHWINGER_H */
