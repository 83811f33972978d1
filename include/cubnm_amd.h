/*
 * cubnm_amd — C ABI of the MI355X simulation core.
 *
 * The reference exposes its core as the CPython extension `cubnm.core`
 * (src/ext/core.cpp:495-626) with three functions: run_simulations, init and
 * set_const.  This header is the plain-C form of that boundary: pointers,
 * sizes and strings only.  `cubnm_amd.core` (a CPython module built on top of
 * this library) re-exports the exact Python signatures, and ctypes / cgo /
 * any FFI can bind these symbols directly.
 *
 * Array layouts are the reference's (core.cpp:194-493, sim/base.py:705-750):
 *   SC               [N_SCs][nodes*nodes], rows = sources, cols = targets
 *   SC_indices       [N_SIMS] int32
 *   SC_dist          [nodes*nodes]
 *   global_params    [n_global_params][N_SIMS]
 *   regional_params  [n_regional_params][N_SIMS*nodes]
 *   v_list           [N_SIMS]
 *   sim_bold         [N_SIMS][bold_len*nodes]
 *   sim_fc_trils     [N_SIMS][n_pairs]          lower triangle, row-major
 *   sim_fcd_trils    [N_SIMS][n_window_pairs]
 *   sim_states       [n_state_vars][N_SIMS][nodes (or nodes*states_len)]
 *   global_bools     [n_global_out_bool][N_SIMS] (uint8)
 *   global_ints      [n_global_out_int][N_SIMS]
 *   noise            [noise_size]; shuffled_nodes [noise_repeats][nodes];
 *   shuffled_ts      [noise_repeats][noise_bw_it]
 *
 * Every function returns 0 on success and a negative value on error; the
 * message is then available from cubnm_last_error().  Not re-entrant: the
 * session is process-global, as in the reference.
 */
#ifndef CUBNM_AMD_H
#define CUBNM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cubnm_dims {
    int bold_len, n_pairs, n_windows, n_window_pairs, states_len, n_state_vars;
    int n_global_params, n_regional_params, n_global_out_bool, n_global_out_int;
    int noise_bw_it, noise_repeats, bw_it, inner_it, n_vols_remove, corr_len;
    int do_fc, do_fcd, modifies_params, N_SIMS, nodes;
    long long noise_size;
    double init_time; /* seconds spent initialising the session in cubnm_prepare */
} cubnm_dims;

typedef struct cubnm_outputs {
    double* sim_bold;             /* required */
    double* sim_fc_trils;         /* may be NULL when FC is off or not wanted */
    double* sim_fcd_trils;        /* may be NULL */
    double* sim_states;           /* may be NULL (ext_out off) */
    unsigned char* global_bools;  /* may be NULL */
    int* global_ints;             /* may be NULL */
    double* noise;                /* may be NULL (noise_out off) */
    int* shuffled_nodes;          /* may be NULL */
    int* shuffled_ts;             /* may be NULL */
    double run_time;              /* out: seconds of the run phase */
} cubnm_outputs;

/* init(): initialise the Balloon-Windkessel constants (core.cpp:147-155). */
int cubnm_init(void);

/* set_const(key, value): set BW k1, k2, k3 or V_0 and refresh the derived
 * constants (core.cpp:157-192).  Unknown keys are ignored, as upstream. */
int cubnm_set_const(const char* key, double value);

/* First half of run_simulations (core.cpp:194-386): select or rebuild the
 * model session, apply the string config (base.hpp:271-297, rww.cpp:54-71),
 * (re)generate the noise segment when seed/duration/nodes/segment changed or
 * force_reinit, and report the output dimensions the caller must allocate.
 * use_cpu != 0 is rejected: this library is the GPU core. */
int cubnm_prepare(const char* model_name, int n_config, const char* const* config_keys,
                  const char* const* config_values, int ext_out, int states_ts, int noise_out,
                  int do_delay, int force_reinit, int use_cpu, int N_SIMS, int nodes, int N_SCs,
                  int time_steps, int BOLD_TR, int states_sampling, int sim_seed, double dt,
                  double bw_dt, cubnm_dims* dims);

/* Second half of run_simulations (core.cpp:388-450) on host buffers: model
 * parameter preparation (rWW analytical FIC writes wIE into regional_params
 * in place, rww.cpp:73-146), the simulations, FC/FCD, and write-back of
 * parameters modified during the run (numerical FIC). */
int cubnm_run(const double* SC, const int* SC_indices, const double* SC_dist,
              double* global_params, double* regional_params, const double* v_list,
              cubnm_outputs* out);

/* Same as cubnm_run, but every array pointer (inputs and outputs) is device
 * memory already resident in HBM (hipMalloc'd or a torch CUDA tensor).  For
 * callers that keep data on the GPU; no reference counterpart. The noise
 * outputs are ignored (host-generated, see cubnm_run). */
int cubnm_run_device(const double* d_SC, const int* d_SC_indices, const double* d_SC_dist,
                     double* d_global_params, double* d_regional_params, const double* d_v_list,
                     cubnm_outputs* d_out);

/* Heterogeneous batches (no counterpart in the reference, whose
 * run_simulations takes one model, node count and duration per call): many
 * groups, each a complete run_simulations request (its own model, node
 * count, duration, config, SCs and parameters), run in one call.  Each group
 * gets its own session; all groups are staged before any kernel runs.  Groups
 * on the batched (shared-SC) integrator launch on their own streams; all
 * groups on the per-simulation integrator, any model, run as ONE launch with
 * their simulations ordered longest first.  With "verbose" in a group's
 * config, progress (BOLD volumes written, simulations completed) is printed
 * to stderr while the launch runs.  Host buffers, with
 * the same layouts and in-place parameter write-back as cubnm_run.  Sessions
 * persist between calls (group i reuses session i, and its noise when the
 * seed, node count and segment agree).  Separate from the single-request
 * session behind cubnm_prepare / cubnm_run. */
typedef struct cubnm_group {
    /* request, as cubnm_prepare */
    const char* model_name;
    int n_config;
    const char* const* config_keys;
    const char* const* config_values;
    int ext_out, states_ts, noise_out, do_delay, force_reinit;
    int N_SIMS, nodes, N_SCs, time_steps, BOLD_TR, states_sampling, sim_seed;
    double dt, bw_dt;
    int batched_tpw;              /* launch shape for this group (0 = auto) */
    /* inputs, as cubnm_run */
    const double* SC;
    const int* SC_indices;
    const double* SC_dist;
    double* global_params;
    double* regional_params;
    const double* v_list;
    /* output sizes, filled by cubnm_prepare_groups */
    cubnm_dims dims;
    /* outputs, allocated by the caller from dims */
    cubnm_outputs out;
} cubnm_group;

int cubnm_prepare_groups(int n_groups, cubnm_group* groups);
int cubnm_run_groups(int n_groups, cubnm_group* groups);
/* Timings of group `group` of the last cubnm_run_groups call, as
 * cubnm_last_timings; [2] is the integration kernel time of the launch that
 * ran the group (the one merged launch for per-simulation groups). */
int cubnm_group_timings(int group, double* out6);
/* How the last cubnm_run_groups call ran its merged per-simulation launch
 * (no counterpart in the reference): out6 = [graph: 0 not captured, 1
 * captured and launched, 2 replayed; k partitions of the coupling sum; pair
 * mode (1 used by some request); launch shape from: 0 the rule, 1 autotuner
 * trials, 2 the autotuner's cached choice; best trial ms; shapes timed]. */
int cubnm_run_groups_info(double* out6);

/* Timings of the last run (milliseconds): [0] parameter prep on host,
 * [1] host->device copies, [2] integration kernel, [3] FC/FCD kernel,
 * [4] device->host copies; [5] integration kernel launches. */
int cubnm_last_timings(double* out6);

/* Which integrator the next/last run uses: 0 = general (workgroup per
 * simulation), 1 = batched MFMA.  Set force to -1 for automatic choice. */
int cubnm_force_integrator(int force);
int cubnm_last_integrator(void);

/* Launch shape of the last run (no counterpart in the reference): out5[0]
 * integrator (0 per-simulation, 1 batched, 2 many-node), [1] 16-node tiles
 * per wave of the batched integrator (up to 256 nodes), [2] its SC fragments
 * in LDS (1) or streamed from L2 (0), [3] k partitions of the per-simulation
 * integrator, [4] chosen by the rule (0), autotuner trials (1) or the
 * autotuner's cache (2).  The autotuner (knob "autotune") times the
 * batched shapes and the per-simulation integrator's k partitions on the
 * request's own launch cut to 50 BW steps. */
int cubnm_last_launch_shape(int* out5);

/* Launch-shape and arithmetic knobs (no counterpart in the reference):
 *   "batched_tpw" = 16-node tiles per wave of the batched integrator
 *                   (0 = auto, 1, 2);
 *   "general_ksplit" = partitions of the per-simulation integrator's
 *                   coupling sum over source nodes (0 = auto, 1..8; capped
 *                   at 1024 lanes per workgroup); the partial sums are added
 *                   in partition order;
 *   "general_pair" = per-simulation integrator sums two adjacent SC columns
 *                   per lane with 16-byte loads: 0 = auto (launches bound by
 *                   latency: SCs shared or few simulations), 1 = whenever
 *                   eligible (even N >= 96, SC in global memory, no delay),
 *                   -1 = never;
 *   "autotune"    = time short trials of the launch shapes and keep the
 *                   fastest, cached per launch kind in memory and in
 *                   $CUBNM_AMD_TUNE_CACHE (else ~/.cache/cubnm_amd/tune.txt):
 *                   for a single request the batched integrator can run,
 *                   simulations per workgroup (16 batched or 1), tiles per
 *                   wave and SC-fragment layout (batched), k partitions
 *                   (per-simulation); for the merged per-simulation launch
 *                   of cubnm_run_groups, its k partitions and pair mode;
 *                   -1 never, 0 = launches of >= 2e10 node-steps, 1 always;
 *   "group_graph" = capture that launch and its table uploads into a
 *                   hipGraph and replay it while the launch is unchanged
 *                   (1, default) or launch directly (0);
 *   "fc_mode"     = FC/FCD path: 0 = auto (per-simulation LDS kernel when a
 *                   simulation's BOLD fits in LDS, else FP64 Gram products),
 *                   1 = FP64 Gram products on MFMA, 2 = split-BF16 Gram
 *                   products (3 BF16 MFMA terms, FP32 accumulation flushed to
 *                   FP64; normalisations in FP64). */
int cubnm_set_tuning(const char* key, int value);
/* FC/FCD path of the last run: 0 LDS kernel, 1 FP64 Gram, 2 split-BF16 Gram. */
int cubnm_last_fc_path(void);

const char* cubnm_last_error(void);

/* Host copies of the noise segment and shuffles of the last cubnm_prepare
 * (the _noise, _shuffled_nodes and _shuffled_ts outputs of run_simulations
 * without running it).  Sizes: dims.noise_size, noise_repeats*nodes,
 * noise_repeats*noise_bw_it. */
int cubnm_copy_noise(double* noise, int* shuffled_nodes, int* shuffled_ts);

/* Gradients through the rWW simulation (no counterpart in the reference):
 * a discrete adjoint of the Euler-Maruyama rWW steps, coupling,
 * Balloon-Windkessel steps and BOLD read-out with the noise held fixed
 * (noise and shuffles as from cubnm_copy_noise, noise width 2).  Device
 * pointers; one SC [nodes][nodes] shared by all simulations; G [n_sims];
 * w_p, J_N, wIE, sigma [n_sims][nodes]; BOLD and dL/dBOLD
 * [n_sims][bold_len][nodes].  Timing arguments as cubnm_prepare.  The
 * forward pass stores a checkpoint every `ck` BW steps; the backward pass
 * recomputes one segment at a time.  Buffer sizes (doubles) come from
 * cubnm_rww_grad_sizes.  No FIC, no delay (wIE is an input). */
int cubnm_rww_grad_sizes(int n_sims, int nodes, int time_steps, int BOLD_TR, double dt, double bw_dt, int ck,
                         long long* bold_doubles, long long* ckpt_doubles, long long* scratch_doubles);
int cubnm_rww_grad_forward(int n_sims, int nodes, int time_steps, int BOLD_TR, double dt, double bw_dt, int ck,
                           int noise_time_steps, const double* d_SC, const double* d_G, const double* d_wp,
                           const double* d_JN, const double* d_wIE, const double* d_sigma, const double* d_noise,
                           const int* d_sh_nodes, const int* d_sh_ts, double* d_bold, double* d_ckpt);
int cubnm_rww_grad_backward(int n_sims, int nodes, int time_steps, int BOLD_TR, double dt, double bw_dt, int ck,
                            int noise_time_steps, const double* d_SC, const double* d_G, const double* d_wp,
                            const double* d_JN, const double* d_wIE, const double* d_sigma, const double* d_noise,
                            const int* d_sh_nodes, const int* d_sh_ts, const double* d_ckpt, const double* d_dbold,
                            double* d_scratch, double* d_gG, double* d_gwp, double* d_gJN);

/* The regional-fit loss L = (1 - r(FC_sim, FC_emp)) + mean_q |Q_q(FCD_sim) -
 * E_q| of each simulation's BOLD and, when d_dbold is not NULL, its gradient
 * dL/dBOLD by a hand-written adjoint (FC over the volumes after
 * n_vols_remove; FCD between the FCs of windows [d_win_start[k],
 * d_win_end[k]] (inclusive, within [0, bold_len)); Q_q the linear-interpolation
 * quantile at d_qs[q] (n - 1) of the sorted FCD values).  No counterpart in
 * the reference (the differentiable objective of gradient fitting, checked
 * against torch autograd of the same loss).  Device pointers; d_scratch holds
 * cubnm_fc_fcd_loss_scratch(...) doubles per simulation; queued on `stream`
 * (a hipStream_t, NULL = null stream) without waiting.  A NaN FC or FCD gives
 * a NaN loss and a zero gradient. */
long long cubnm_fc_fcd_loss_scratch(int bold_len, int nodes, int n_vols_remove, int n_windows, int max_window_len);
int cubnm_fc_fcd_loss(const double* d_bold, int n_sims, int bold_len, int nodes, int n_vols_remove,
                      const int* d_win_start, const int* d_win_end, int n_windows, int max_window_len,
                      const double* d_emp_fc, const double* d_emp_q, const double* d_qs, int n_quantiles,
                      double* d_loss, double* d_dbold, double* d_scratch, void* stream);

/* Goodness of fit on the GPU (utils.py:306-507 fc_corr_device and
 * fcd_ks_device, sim/base.py:1084-1207): Pearson r between each simulated FC
 * tril and the empirical one, and the two-sample Kolmogorov-Smirnov
 * statistic between each simulated FCD tril and the empirical FCD tril.
 * Host buffers; NaN inputs yield NaN (callers map NaN to the worst score). */
int cubnm_gof_fc_corr(const double* sim_fc, int n_sims, int n_pairs, const double* emp_fc, double* out);
int cubnm_gof_fcd_ks(const double* sim_fcd, int n_sims, int n_sim_values, const double* emp_fcd,
                     int n_emp_values, double* out);

/* Device-pointer forms of the two GOF reductions (inputs and outputs resident
 * in HBM); the empirical FCD values must already be sorted ascending.  Both
 * return after the kernel has completed. */
int cubnm_gof_fc_corr_device(const double* d_sim_fc, int n_sims, int n_pairs, const double* d_emp_fc,
                             double* d_out);
int cubnm_gof_fcd_ks_device(const double* d_sim_fcd, int n_sims, int n_sim_values,
                            const double* d_emp_fcd_sorted, int n_emp_values, double* d_out);

/* Allocation-free variant of cubnm_gof_fcd_ks_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its sort and reduction workspace is a fixed 16 MiB array in the library's
 * device image (one per GPU, held until the process ends), so it calls no
 * hipMalloc/hipFree; simulations are processed in batches that fit it, and a
 * simulation with more than 2,097,150 FCD values is rejected with an error.
 * Work is queued on `stream` (a hipStream_t; NULL = the null stream) and the
 * call returns without waiting.  Every call on one GPU shares that GPU's
 * workspace: calls that overlap on one GPU (other streams, other threads)
 * race and must be ordered by the caller. */
int cubnm_gof_fcd_ks_device_static(const double* d_sim_fcd, int n_sims, int n_sim_values,
                                   const double* d_emp_fcd_sorted, int n_emp_values, double* d_out,
                                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CUBNM_AMD_H */
