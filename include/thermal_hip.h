/* thermal_hip.h - C ABI of the MI355X (gfx950) MG-PCG thermal solver.
 *
 * Drop-in boundary for the GPU path of cuda-thermal: the model builder of
 * src/stack3d.h and src/case3d.h, and the steady / backward-Euler solver
 * concept shared by CpuPcg and GpuPcg (src/pcg3d.h:60-163), which
 * run_transient (src/transient3d.h:48-90) consumes. Plain pointers and sizes
 * only. Every function returns TH_OK (0) or an error code; th_last_error()
 * returns the message of the last failure on the calling thread. No function
 * exits the process: the CLI turns a failure into message + exit(1), as the
 * reference's CUDA_CHECK does (src/cuda_check.h:8-16).
 *
 * Arrays crossing this boundary are in the reference's natural layout:
 * cell (i, j, l) at ((l * ny + j) * nx + i); gB has nx * ny entries.
 * Temperatures are rises above ambient (K), as in the reference.
 */
#ifndef THERMAL_HIP_H
#define THERMAL_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define TH_API __attribute__((visibility("default")))

enum {
    TH_OK = 0,
    TH_ERR_ARG = 1,           /* invalid argument */
    TH_ERR_HIP = 2,           /* a HIP runtime call failed */
    TH_ERR_NOT_FINALIZED = 3, /* model used before th_model_finalize (pcg3d_gpu.cu:347-350) */
    TH_ERR_CASE = 4,          /* malformed case folder (case3d.h:86-155) */
    TH_ERR_ZLIMIT = 5,        /* more than TH_MAXZ z-planes (stack3d.h:95-98) */
    TH_ERR_NO_CELLS = 6       /* hotspot covers no cells (stack3d.h:242) */
};

#define TH_MAXZ 64 /* stack3d.h:34 */

TH_API const char* th_last_error(void);
TH_API const char* th_version(void);

/* ---- model: stack3d.h:36-298, case3d.h:24-190 -------------------------- */
typedef struct th_model th_model;

/* build_model (stack3d.h:81-102): layers bottom to top. names may be NULL. */
TH_API int th_model_create(int nlayers, const char* const* names, const double* thickness_m,
                           const double* k, const double* rho_c, const int* planes,
                           const int* active, int nx, int ny, double Lx_m, double Ly_m,
                           double h_top, double h_bot, th_model** out);
/* default_model (stack3d.h:153-168): the two-tier stack, 10 mm x 10 mm. */
TH_API int th_model_create_default(int nx, int ny, th_model** out);
/* load_case + build_case_model (case3d.h:86-166). On TH_ERR_CASE the
 * message names the file and the problem, like the reference's `err`. */
TH_API int th_model_load_case(const char* dir, int nx, int ny, int with_tsv, th_model** out);
/* example_case (case3d.h:170-190) built on nx x ny and finalized. */
TH_API int th_model_example_case(int nx, int ny, int with_tsv, th_model** out);
TH_API void th_model_destroy(th_model* m);

TH_API int th_model_add_power_block(th_model* m, int layer, double x0, double y0, double w,
                                    double h, double watts);          /* stack3d.h:196 */
TH_API int th_model_add_uniform_power(th_model* m, int layer, double watts); /* :223 */
TH_API int th_model_add_hotspot(th_model* m, int layer, double xc, double yc, double size,
                                double watts);                         /* stack3d.h:230 */
TH_API int th_model_add_tsv_region(th_model* m, int layer, double x0, double y0, double w,
                                   double h, double fraction);         /* stack3d.h:210 */
TH_API int th_model_finalize(th_model* m);                             /* stack3d.h:112 */
TH_API int th_model_set_power(th_model* m, const double* P);           /* replace P (W/cell) */

TH_API int th_model_dims(const th_model* m, int* nx, int* ny, int* nz, double* dx, double* dy);
TH_API int th_model_layer_count(const th_model* m);
TH_API int th_model_plane_layer(const th_model* m, int* plane_layer /* nz */);
TH_API int th_model_layer_info(const th_model* m, int layer, double* thickness_m, double* k,
                               double* rho_c, int* planes, int* active, const char** name);
TH_API int th_model_settings(const th_model* m, double* Lx, double* Ly, double* h_top,
                             double* h_bot, double* T_amb);
/* which: "gE" "gN" "gU" "cap" (float, n), "gB" (float, nx*ny). */
TH_API int th_model_get_f32(const th_model* m, const char* which, float* out);
/* which: "P" "tsv_frac" (double, n), "dz" (double, nz). */
TH_API int th_model_get_f64(const th_model* m, const char* which, double* out);
TH_API double th_model_total_power(const th_model* m);                /* stack3d.h:246 */
TH_API int th_model_analytic_1d(const th_model* m, double* T /* nz */); /* stack3d.h:279 */
TH_API double th_model_heat_out(const th_model* m, const double* T);  /* stack3d.h:254 */

/* Power blocks and TSV regions of the case a model was built from
 * (th_model_load_case / th_model_example_case; CaseSpec, case3d.h:24-33);
 * 0 for other models. */
TH_API int th_model_block_count(const th_model* m);
TH_API int th_model_tsv_count(const th_model* m);
TH_API int th_model_block_info(const th_model* m, int k, const char** name, int* layer, double* x0,
                               double* y0, double* w, double* h, double* watts);
/* Max and area-weighted mean of T (natural layout) over block k's rectangle
 * on its layer's planes (block_temps, main3d.cu:286-301). */
TH_API int th_model_block_temps(const th_model* m, int k, const double* T, double* tmax, double* tavg);

/* ---- 3D-ICE interchange: case3d.h:192-400, main3d.cu:380-416 ------------
 * export_case_3dice / export_3dice: the case folder `dir` described for
 * 3D-ICE on square cells of cell_um micrometres, written into the existing
 * folder outdir: case.stk, L<k>.flp per active layer (die) and LT<k>_<p>.lay
 * per plane with TSVs; the same text, byte for byte, as the reference's
 * exporter. Refused (TH_ERR_CASE, with the reference's message) when the
 * cell size does not divide the die, an active layer has TSVs or more than
 * one plane, there is a bottom heat sink, or two power blocks of one layer
 * overlap. nx, ny (may be NULL): the lateral grid used. */
TH_API int th_case_export_3dice(const char* dir, double cell_um, const char* outdir, int* nx, int* ny);
/* read_3dice_tmap (case3d.h:387-400): a Tmap result file, rows of values
 * (K) with '%' comment lines; *rows x *cols values, row-major, into values
 * (capacity doubles; values may be NULL to ask for the size). Rows of
 * different lengths are an error (TH_ERR_CASE). */
TH_API int th_3dice_read_tmap(const char* path, int* rows, int* cols, double* values, long capacity);
/* read_3dice_tflp (main3d.cu:380-403): a Tflp result file, the block names
 * of its "% Time(s) name(K) ..." header and one value (K) per block. Names
 * go to names + k * name_len (NUL-terminated), values to values[k], for up
 * to capacity blocks; *count receives the number of blocks (names and
 * values may both be NULL to ask for it). */
TH_API int th_3dice_read_tflp(const char* path, char* names, int name_len, double* values, int capacity,
                              int* count);

/* What the solver reads from a finalized Model3D (pcg3d_gpu.cu:372-391).
 * All arrays are copied by th_solver_create. */
typedef struct {
    int nx, ny, nz;
    double dx, dy;           /* lateral cell size (m), for the coarse operators */
    const float* gE;         /* n: conductance to +x (0 on the last column) */
    const float* gN;         /* n: to +y (0 on the last row) */
    const float* gU;         /* n: to +z, or to ambient on the top plane */
    const float* gB;         /* nx * ny: bottom plane to ambient */
    const float* cap;        /* n: heat capacity (J/K) */
    const double* P;         /* n: power (W) */
} th_model_desc;

/* Descriptor pointing into a finalized model (valid while the model lives). */
TH_API int th_model_describe(const th_model* m, th_model_desc* out);

/* ---- solver: pcg3d.h:60-163 ---------------------------------------------- */
typedef struct th_solver th_solver;

enum { TH_PRECOND_JACOBI = 0, TH_PRECOND_MG = 1 };           /* pcg3d.h:54 */
enum {
    TH_PREC_MIXED = 0, /* FP32 conductances, factors and V-cycle; FP64 PCG */
    TH_PREC_FP64 = 1   /* FP64 everywhere (conductances widened from float) */
};
enum { TH_BETA_FLETCHER_REEVES = 0, TH_BETA_FLEXIBLE = 1 };

typedef struct {
    int precond;    /* TH_PRECOND_* */
    int precision;  /* TH_PREC_* */
    int beta;       /* TH_BETA_*; the reference uses Fletcher-Reeves (pcg3d_cpu.cpp:179) */
    int use_graphs; /* replay iterations from hipGraphs (gpu_use_graphs, pcg3d.h:167) */
    int device;     /* HIP device ordinal */
    int true_residual; /* 1: on convergence recompute b - A x in FP64 and restart
                          from x while it is above tol (up to 4 restarts); with the
                          multigrid preconditioner and tol < 1e-8 the recurrence
                          residual is also replaced once by b - A x on the way (when
                          it falls below 1e4 tol), which keeps the two close enough
                          that the check rarely restarts; 0: stop on the recurrence
                          residual, as the reference does */
    int restriction;   /* TH_RESTRICT_*: how the V-cycle restricts its residual */
} th_solver_opts;

/* TH_RESTRICT_LATERAL: coarse b = sum over the red children of their lateral
 * terms sum G x, the residual left after the red (x = 0) and black sweeps,
 * equal to the reference's sum of (b - A x) over the children in exact
 * arithmetic and reading half the bytes (default). TH_RESTRICT_RESIDUAL: the
 * reference's form (pcg3d_cpu.cpp:87-100), operation for operation. */
enum { TH_RESTRICT_LATERAL = 0, TH_RESTRICT_RESIDUAL = 1 };

TH_API void th_solver_default_opts(th_solver_opts* o);

typedef struct {
    int iters;
    double relres;  /* ||r|| / ||b|| of the PCG recurrence */
    int converged;
} th_stats;          /* PcgStats, pcg3d.h:45-49 */

TH_API int th_solver_create(const th_model_desc* m, const th_solver_opts* o, th_solver** out);
/* Where th_solver_create's time went, in ms (GpuSetupTimes and
 * gpu_last_setup, pcg3d.h:101-109; here per solver, not process-global):
 * device allocations, the fine operator and power map to the device, the
 * multigrid levels (coarse operators built on the host in parallel, then
 * re-laid out on the device), the column types (found on the host, their
 * table to the device), the factors for dt = 0, and the recorded graphs. */
typedef struct {
    double alloc_ms, upload_ms, levels_ms, types_ms, factor_ms, graphs_ms, total_ms;
} th_setup_times;
TH_API int th_solver_setup_times(th_solver* s, th_setup_times* out);
TH_API void th_solver_destroy(th_solver* s); /* never fails loudly (pcg3d_gpu.cu:495-509) */
TH_API int th_solver_set_dt(th_solver* s, double dt);       /* 0 = steady */
TH_API int th_solver_set_x_zero(th_solver* s);
TH_API int th_solver_set_x(th_solver* s, const double* x);  /* warm start / initial guess */
TH_API int th_solver_set_rhs_steady(th_solver* s);          /* b = P */
TH_API int th_solver_set_rhs_transient(th_solver* s);       /* b = P + C/dt x */
TH_API int th_solver_solve(th_solver* s, double tol, int maxit, th_stats* st);
TH_API int th_solver_get_x(th_solver* s, double* x);
TH_API int th_solver_set_power(th_solver* s, const double* P);
/* Device pointer to x in the natural layout (refreshed by this call; valid
 * until the next solver call). */
TH_API int th_solver_device_x(th_solver* s, const double** d_x);
/* ||P - A T|| / ||P|| of the steady operator at the current x, in FP64 on
 * the device (pcg3d_cpu.cpp:207-220). */
TH_API int th_solver_true_relres(th_solver* s, double* out);
/* heat_out (stack3d.h:254) at the current x, on the device. */
TH_API int th_solver_heat_out(th_solver* s, double* out);
TH_API int th_solver_nlevels(th_solver* s);
/* Number of distinct fine-level column types the solver found (no reference
 * counterpart; introspection): > 0 when the fine level's kernels read a
 * 2-byte type per column and a table of the types' faces instead of the
 * per-cell arrays, 0 when they read the arrays, -1 for a null solver. */
TH_API int th_solver_column_types(th_solver* s);
/* Multigrid level k (0 = fine, th_solver_nlevels - 1 = the single column;
 * no reference counterpart; introspection): its grid, its number of column
 * types (0: per-cell arrays) and which kernels run it (*kind: 0 the column
 * walk, 1 the plane-parallel sweeps, 2 the single-workgroup coarse tail).
 * Any output may be NULL. */
TH_API int th_solver_level_info(th_solver* s, int k, int* nx, int* ny, int* nz, int* ntypes, int* kind);
/* Die map of the current x (the viewer's k_colorize, viewer.cu:68-87): for
 * each column (i, j) the hottest temperature rise v over planes l0..l1 (at
 * least 0), coloured on [0, rise_hi] with the viewer's five-stop palette
 * (colormap, viewer.cu:53-66) into rgba[4 * (j * nx + i) + 0..3] (R, G, B,
 * 255; host memory, nx * ny * 4 bytes, or NULL for no image); and the hottest
 * cell of those planes, *hot_rise = its v and *hot_cell = j * nx + i (the
 * largest index among equal values). rise_hi must be > 0. */
TH_API int th_solver_colorize(th_solver* s, int l0, int l1, float rise_hi, unsigned char* rgba, float* hot_rise,
                              int* hot_cell);
/* The same with the image written to device memory (d_rgba, nx * ny * 4
 * bytes, e.g. a buffer a display maps), as the viewer writes its PBO. */
TH_API int th_solver_colorize_device(th_solver* s, int l0, int l1, float rise_hi, unsigned char* d_rgba,
                                     float* hot_rise, int* hot_cell);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under graph capture, tight display or serving loops,
 * harnesses that forbid hipMalloc): th_solver_colorize_device without any
 * allocation. Its per-block partials and hottest-cell key live in a static
 * device array of the library (136 KB per GPU, held until the process ends)
 * instead of memory allocated on the first call. Limit: grids of at most
 * 16384 blocks of 64 x 4 columns (2048 x 2048 columns); larger grids fail
 * with TH_ERR_ARG. d_rgba may be NULL (hottest cell only). Calls on one GPU
 * share that array: calls that overlap in time on one GPU (other threads,
 * other solvers' streams) race and must be serialised by the caller. */
TH_API int th_solver_colorize_device_noalloc(th_solver* s, int l0, int l1, float rise_hi, unsigned char* d_rgba,
                                             float* hot_rise, int* hot_cell);
/* z = M^-1 r: one preconditioner application (the multigrid V-cycle in the
 * solver's precision, or Jacobi) for the current dt, host arrays in the
 * natural layout (CpuPcg::precondition, pcg3d_cpu.cpp:115-124). */
TH_API int th_solver_precondition(th_solver* s, const double* r, double* z);
/* y = (A + C/dt) x in FP64 for the current dt (operator3d.h:109-112). */
TH_API int th_solver_apply_operator(th_solver* s, const double* x, double* y);

/* Measurement (no reference counterpart; the reference times kernels with
 * Nsight only): average duration of `reps` back-to-back launches of one
 * hot-path kernel on the solver's stream, timed with HIP events, and the
 * kernel's algorithmic bytes per launch. Call after a solve (the kernels run
 * on the solver's current vectors; x and b of the PCG are not modified).
 *   TH_K_SWEEP    fine-level colour sweep (red columns, black neighbours)
 *   TH_K_APPLY    fine-level FP64 operator apply Ap = (A + C/dt) p
 *   TH_K_RESTRICT fine-to-first-coarse restriction
 *   TH_K_COPY     streaming copy of one FP64 vector (bandwidth reference)
 *   TH_K_VCYCLE   one whole V-cycle (bytes reported as 0)
 *   TH_K_UPDATE   the iteration's residual update: r -= alpha Ap, the V-cycle's
 *                 right-hand side, ||r||^2 and r.z_old (multigrid solvers) */
enum { TH_K_SWEEP = 0, TH_K_APPLY = 1, TH_K_RESTRICT = 2, TH_K_COPY = 3, TH_K_VCYCLE = 4, TH_K_UPDATE = 5 };
TH_API int th_solver_time_kernel(th_solver* s, int which, int reps, double* avg_ms, double* bytes);

/* Steady solve from T = 0 with timing (gpu_pcg, pcg3d_gpu.cu:686-719). */
typedef struct {
    int iters;
    double relres;
    int converged;
    double setup_ms, solve_ms, total_ms;
} th_solve_result;
TH_API int th_steady_solve(const th_model_desc* m, const th_solver_opts* o, double tol, int maxit,
                           double* T, th_solve_result* res);

/* ---- transient: transient3d.h:19-90 -------------------------------------- */
typedef struct {
    double dt0, growth, dt_max, t_end; /* TransientOptions (transient3d.h:19-27) */
    double tol;
    int maxit;
    int max_steps;                     /* > 0: also stop after this many steps */
    int predictor;                     /* start of each step: 0 = T_old (the reference's warm
                                          start); 1 = linear extrapolation of the last two steps;
                                          2 = learned: T_old + w (G - T_old) towards the guide field
                                          G (th_solver_set_guide*, e.g. the surrogate's predicted
                                          field), w >= 0 minimising the step's starting residual
                                          ||b - (C/dt + A) x||, w = 0 (T_old, the plain warm start)
                                          when no w > 0 lowers it; 3 = 1, then 2 from the
                                          extrapolated start */
} th_transient_opts;
TH_API void th_transient_default_opts(th_transient_opts* o);

typedef struct {
    double t, dt;
    int iters;
    double T_max, q_out, q_stored;
} th_step_record; /* StepRecord, transient3d.h:29-35 */

/* run_transient: starts at ambient; one backward-Euler step per record.
 * Energy terms are computed on the device in FP64. T_final (optional, n)
 * receives the last field. */
TH_API int th_run_transient(th_solver* s, const th_transient_opts* o, th_step_record* recs,
                            int rec_cap, int* nsteps, double* T_final, int* all_converged,
                            double* solve_ms);
/* The guide field G of the learned step start (no counterpart in the
 * reference): from a host array (natural layout, K above ambient), or from a
 * coarse device field lifted as th_solver_set_x_lifted lifts its guess (for
 * the surrogate's output; the same arguments, without the judging). */
TH_API int th_solver_set_guide(th_solver* s, const double* G);
TH_API int th_solver_set_guide_lifted(th_solver* s, const float* field, int Hc, int Wc, int C, int nzc,
                                      const double* zc_coarse, const double* zc_fine, double scale, int passes);
/* Over the last th_run_transient: the steps whose start took the guide
 * direction (w > 0) and their mean w. */
TH_API int th_solver_guide_stats(th_solver* s, int* steps, double* w_mean);

/* ---- batched steady solves (no reference counterpart) ---------------------
 * Many independent stacks (grid sizes may differ; plane counts <= 16 in one
 * bucket of 4/8/12/16) solved from T = 0 together: every stage of a PCG
 * iteration is one launch over all instances, each instance masked by its
 * own convergence state. Multigrid, mixed precision, flexible beta and the
 * lateral restriction only. The descriptors are copied at creation. */
typedef struct th_batch th_batch;
TH_API int th_batch_create(int count, const th_model_desc* const* models, const th_solver_opts* o, th_batch** out);
TH_API void th_batch_destroy(th_batch* b);
TH_API int th_batch_count(th_batch* b);
TH_API int th_batch_set_power(th_batch* b, int i, const double* P);
/* Per-instance iterations, recurrence rel-res and convergence (arrays of
 * count); ms: wall time of the whole batch. */
TH_API int th_batch_solve(th_batch* b, double tol, int maxit, int* iters, double* relres, int* converged,
                          double* ms);
TH_API int th_batch_get_x(th_batch* b, int i, double* x);
/* Autotuner: times the V-cycle with blocks of 1, 2, 4 or 8 wavefronts
 * (*rows) whose wavefronts each cover 64 columns of one row, 32 columns of
 * two rows or 16 of four (*cols: columns per wavefront) and the vector stages
 * with 256- or 1024-thread blocks on this batch, keeps the fastest; reports
 * the choice and the V-cycle time. */
TH_API int th_batch_autotune(th_batch* b, int reps, int* rows, int* cols, int* vec_threads, double* vcycle_ms);
TH_API int th_batch_shape(th_batch* b, int* rows, int* cols, int* vec_threads);

/* ---- learned warm start (no reference counterpart) ------------------------
 * A U-Net surrogate (BF16 MFMA kernels, FP32 accumulation) mapping, on an
 * H x W grid over the die, 3 input features per cell (power density of die
 * 0 and die 1 in MW/m^2, TSV copper fraction) to 10 outputs per cell (the
 * temperature rise of the 10 planes of the two-tier stack, in units of
 * 50 K). Feature and output arrays are host float arrays [B][H][W][C].
 * Parameters are one flat float vector (per layer W[cout][tap][cin], then
 * bias[cout]). One handle owns one HIP stream and is not thread-safe. */
typedef struct th_surrogate th_surrogate;
TH_API long th_surrogate_param_count(void);
TH_API int th_surrogate_create(int H, int W, unsigned long long seed, th_surrogate** out);
TH_API void th_surrogate_destroy(th_surrogate* s);
TH_API int th_surrogate_get_params(th_surrogate* s, float* out);
TH_API int th_surrogate_set_params(th_surrogate* s, const float* in);
TH_API int th_surrogate_predict(th_surrogate* s, const float* in, int batch, float* out);
/* One image; the output stays on the device: float [H][W][16], the first 10
 * channels used, valid until the next call on s. */
TH_API int th_surrogate_predict_device(th_surrogate* s, const float* in, const float** d_out);
/* The same with the input already in device memory (d_in: float [H][W][in
 * channels], the layout of `in`; writes to it must be complete). No
 * counterpart in the reference: for callers that keep their power maps on
 * the device between solves, so the upload is not paid per prediction. */
TH_API int th_surrogate_predict_device_input(th_surrogate* s, const float* d_in, const float** d_out);
/* Loss (mean squared error over the 10 outputs) and its gradient with
 * respect to the flat parameters (grad may be NULL). */
TH_API int th_surrogate_grad(th_surrogate* s, const float* in, const float* tgt, int batch, double* loss,
                             float* grad);
/* Training set held in device memory: features in BF16, targets in FP32. */
TH_API int th_surrogate_dataset_reserve(th_surrogate* s, int capacity);
TH_API int th_surrogate_dataset_add(th_surrogate* s, const float* in, const float* tgt, int count);
TH_API int th_surrogate_dataset_size(th_surrogate* s);
/* Adam on random batches of samples [0, n_train), learning rate decaying
 * from lr to lr_final (cosine); per-step losses (steps floats) to host. */
TH_API int th_surrogate_train(th_surrogate* s, int steps, int batch, float lr, float lr_final,
                              unsigned long long seed, int n_train, float* losses, double* ms);
TH_API int th_surrogate_eval(th_surrogate* s, int first, int count, double* mse);
/* Average forward time of a batch and its MFMA work (flop per forward). */
TH_API int th_surrogate_time_forward(th_surrogate* s, int batch, int reps, double* avg_ms, double* flops);

/* Initial guess from a coarse field (for instance the surrogate's output):
 * field is a device float array [Hc][Wc][C] whose channels 0..nzc-1 are
 * planes with centres zc_coarse (m); each of the solver's planes (centres
 * zc_fine, nz entries) is interpolated linearly in z and bilinearly over
 * the die, times `scale`, then `passes` red-black column line relaxations
 * of the steady operator (each a tridiagonal solve per column) follow.
 * Fallback: if ||P - A x|| of the result is not below that of the current
 * x (the plain warm start), x is left as it was. *used = 1 if the lifted
 * guess was kept; relres_plain / relres_lift receive both starting
 * residuals ||P - A x|| / ||P||. */
TH_API int th_solver_set_x_lifted(th_solver* s, const float* field, int Hc, int Wc, int C, int nzc,
                                  const double* zc_coarse, const double* zc_fine, double scale, int passes,
                                  int* used, double* relres_plain, double* relres_lift);
/* The learned steady start in one asynchronous call (no counterpart in the
 * reference): the surrogate's forward pass on d_in (its input in device
 * memory, as th_surrogate_predict_device_input takes it), the lift of its
 * output and th_solver_set_x_lifted's fallback, all queued on the solver's
 * stream and judged on the device, so the host waits for nothing before the
 * next th_solver_solve. The surrogate must have been created on the
 * solver's device; its output channels 0..nzc-1 (nzc <= 16) are the planes
 * with centres zc_coarse. th_solver_lifted_stats then reads the judgement
 * (waiting for the device): *used = 1 if the lifted guess was kept, and
 * both starting relative residuals. */
TH_API int th_solver_set_x_learned(th_solver* s, th_surrogate* sur, const float* d_in, int nzc,
                                   const double* zc_coarse, const double* zc_fine, double scale, int passes);
TH_API int th_solver_lifted_stats(th_solver* s, int* used, double* relres_plain, double* relres_lift);

#ifdef __cplusplus
}
#endif
#endif /* THERMAL_HIP_H */
