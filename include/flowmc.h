/*
 * flowmc.h — C ABI of the MI355X (gfx950) flowMC sampling hot path.
 *
 * Every entry point:
 *   - takes device pointers owned by the caller (no allocation inside hot calls),
 *   - is stream-ordered on the `stream` argument (a hipStream_t passed as void*),
 *   - returns FMC_OK (0) or a non-zero fmc_status; the message of the last
 *     failure on the calling thread is available from fmc_last_error().
 * Nothing throws across this boundary.
 *
 * Reference citations are relative to the flowMC v0.4.5 source tree.
 */
#ifndef FLOWMC_H_
#define FLOWMC_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMC_ABI_VERSION 3

typedef enum fmc_status {
  FMC_OK = 0,
  FMC_ERR_INVALID_ARG = 1,  /* null pointer, bad size, bad enum            */
  FMC_ERR_UNSUPPORTED = 2,  /* valid request the library cannot serve      */
  FMC_ERR_LAUNCH = 3,       /* kernel launch failed                        */
  FMC_ERR_HIP = 4           /* other HIP runtime error                     */
} fmc_status;

/* ------------------------------------------------------------------------ */
/* Lifecycle / introspection                                                */
/* ------------------------------------------------------------------------ */
int fmc_abi_version(void);
/* Thread-local message of the last failing call on this thread ("" if none). */
const char* fmc_last_error(void);
/* Device properties used by the host-side tile chooser. */
int fmc_device_query(int device, int* n_cu, int* lds_bytes_per_block,
                     int* warp_size, int* arch_major, int* arch_minor);

/* ------------------------------------------------------------------------ */
/* Built-in log-densities (value + hand-written gradient on device).        */
/* Replaces the user `logpdf(x, data)` + jax.value_and_grad pair that       */
/* MALA.kernel (src/flowMC/resource/kernel/MALA.py:59) and HMC              */
/* (HMC.py:71-80) differentiate with autodiff.                              */
/* ------------------------------------------------------------------------ */
typedef enum fmc_target_kind {
  FMC_TARGET_GAUSS_DIAG = 0,  /* params: mean[d], prec[d]:  -1/2 sum prec (x-mean)^2          */
  FMC_TARGET_GAUSS_DENSE = 1, /* params: mean[d], P[d*d]:  -1/2 (x-mean)^T P (x-mean)         */
  FMC_TARGET_ROSENBROCK = 2,  /* params: scale[1]: -sum_i [100(x_{i+1}-x_i^2)^2+(1-x_i)^2]*scale */
  FMC_TARGET_FUNNEL = 3,      /* params: none; x_0 = v ~ N(0,9), x_i|v ~ N(0,e^v)              */
  FMC_TARGET_GMM_DIAG = 4,    /* params: mean[K*d], inv_sigma[K*d], c[K] (c = log w - sum log s) */
  FMC_TARGET_DUAL_MOON = 5    /* params: center[d]; docs/tutorials/dualmoon.ipynb target        */
} fmc_target_kind;

typedef struct fmc_target {
  int32_t kind;        /* fmc_target_kind                                   */
  int32_t n_dim;       /* d                                                 */
  int32_t n_modes;     /* K for FMC_TARGET_GMM_DIAG, else 0                 */
  int32_t reserved;
  const float* params; /* device pointer to the parameter block (may be 0 for FUNNEL) */
  int64_t n_params;    /* number of floats in params                        */
} fmc_target;

/* Lane layout of the chain kernels: each chain is held by `lanes_per_chain`
 * lanes (power of two, 1..64) of one wavefront; each lane holds
 * `vec_per_lane` float4 chunks (1,2,4,8). d <= 4*vec_per_lane*lanes_per_chain.
 * Zero fields mean "choose by heuristic". */
typedef struct fmc_launch_cfg {
  int32_t lanes_per_chain;
  int32_t vec_per_lane;
  int32_t block_size; /* threads per workgroup (64..1024, multiple of 64;  */
                      /* fused HMC launches: at most 512)                */
  int32_t reserved;
} fmc_launch_cfg;

/* Host-side validation of a target descriptor (kind, n_dim, n_modes and the
 * parameter count the kind needs), the check every chain entry point makes
 * first; no device work. For callers that place descriptors in device
 * memory (fmc_hmc_run_batched's slots), where the library cannot read them. */
int fmc_target_check(const fmc_target* target);

/* Value (and optionally gradient) of a built-in target at C points.
 * x [C,d] -> logp [C], grad [C,d] (grad may be NULL). */
int fmc_target_eval(void* stream, const fmc_target* target, const float* x,
                    int64_t n_points, float* logp, float* grad,
                    const fmc_launch_cfg* cfg);

/* ------------------------------------------------------------------------ */
/* Fused K-step MALA.                                                       */
/* Replaces TakeSerialSteps.sample (src/flowMC/strategy/take_steps.py:164-  */
/* 180) scanning MALA.kernel (src/flowMC/resource/kernel/MALA.py:26-89)     */
/* vmapped over chains (take_steps.py:127-132), plus the thinned buffer     */
/* writes of TakeSteps.__call__ (take_steps.py:134-143).                    */
/*                                                                          */
/* Steps step0 .. step0+n_steps-1 of one strategy call are run. The chain   */
/* state x[C,d] / logp[C] is read at entry and written back at exit; the    */
/* kernel evaluates (logp, grad) of the entry state itself, as the          */
/* reference does at take_steps.py:177 and MALA.py:59.                      */
/* Noise: if noise != NULL the Langevin noise of step s, chain c is         */
/* noise[(s*C + c)*d + j] and the MH log-uniform is log_u[s*C + c]          */
/* (injected-noise mode); otherwise both come from Philox4x32-10 keyed by   */
/* (key0, key1) with counter (dim_chunk, chain, step0+s, stream_id).        */
/* Outputs (any may be NULL): rows of the [C, buf_len, d] / [C, buf_len]    */
/* buffers: global step t (= step0 + s) with t % thinning == 0 is written    */
/* to row cursor + t / thinning; acceptance is stored as 1.0f / 0.0f.       */
/* ------------------------------------------------------------------------ */
int fmc_mala_run(void* stream, const fmc_target* target, float* x, float* logp,
                 int64_t n_chains, int32_t n_steps, int32_t step0, float step_size,
                 uint32_t key0, uint32_t key1, const float* noise, const float* log_u,
                 float* out_pos, float* out_logp, float* out_acc, int64_t buf_len,
                 int64_t cursor, int32_t thinning, const fmc_launch_cfg* cfg);

/* ------------------------------------------------------------------------ */
/* Fused K-step Gaussian random walk.                                       */
/* Replaces TakeSerialSteps.sample (take_steps.py:164-180) scanning          */
/* GaussianRandomWalk.kernel (src/flowMC/resource/kernel/                   */
/* Gaussian_random_walk.py:25-61): x' = x + step_size * xi, accept iff      */
/* log u < logp(x') - logp(x). Same arguments, state handling, noise layout */
/* and buffer rows as fmc_mala_run; the log-density of the entry state is   */
/* evaluated once (take_steps.py:177) and carried. Philox streams 7 (xi)    */
/* and 8 (u).                                                               */
/* ------------------------------------------------------------------------ */
int fmc_grw_run(void* stream, const fmc_target* target, float* x, float* logp,
                int64_t n_chains, int32_t n_steps, int32_t step0, float step_size,
                uint32_t key0, uint32_t key1, const float* noise, const float* log_u,
                float* out_pos, float* out_logp, float* out_acc, int64_t buf_len,
                int64_t cursor, int32_t thinning, const fmc_launch_cfg* cfg);

/* ------------------------------------------------------------------------ */
/* Fused K-step HMC.                                                        */
/* Replaces HMC.kernel (src/flowMC/resource/kernel/HMC.py:98-151) with its  */
/* leapfrog scan (HMC.py:71-96) and coefficient table (HMC.py:47-50:        */
/* n_leapfrog+1 drifts, n_leapfrog+2 kicks), scanned by TakeSerialSteps.    */
/* metric: NULL = identity, else diag(metric[d]) (the condition matrix M);  */
/* momentum p = xi / sqrt(metric), kinetic = 1/2 sum metric p^2.            */
/* Dense M: momentum_chol = chol(inv(M)) [d, d] row-major lower triangle and */
/* metric = colsum(M): p = xi @ momentum_chol^T (HMC.py:133-136) and the    */
/* reference's broadcast kinetic 1/2 sum(p^2 * M) = 1/2 sum_j colsum_j p_j^2 */
/* (HMC.py:118-121), whose p-gradient colsum * p is the drift. NULL when M  */
/* is diagonal.                                                             */
/* The Hamiltonian difference H0 - H1 is formed in FP64.                    */
/* Noise: noise[(s*C + c)*d + j] is the momentum draw, log_u[s*C + c].      */
/* ------------------------------------------------------------------------ */
int fmc_hmc_run(void* stream, const fmc_target* target, float* x, float* logp,
                int64_t n_chains, int32_t n_steps, int32_t step0, float step_size,
                int32_t n_leapfrog, const float* metric, const float* momentum_chol,
                uint32_t key0, uint32_t key1,
                const float* noise, const float* log_u, float* out_pos, float* out_logp,
                float* out_acc, int64_t buf_len, int64_t cursor, int32_t thinning,
                const fmc_launch_cfg* cfg);

/* ------------------------------------------------------------------------ */
/* Finite-row mask of a row-major [n_rows, n_dim] buffer: out[r] = 1 iff    */
/* every entry of row r is finite, else 0. Replaces                         */
/* jnp.isfinite(data).all(axis=-1) in TrainModel's training-data selection  */
/* (src/flowMC/strategy/train_model.py:66-80).                              */
/* ------------------------------------------------------------------------ */
int fmc_finite_rows(void* stream, const float* data, int64_t n_rows, int32_t n_dim, int32_t* out);

/* ------------------------------------------------------------------------ */
/* Parallel tempering.                                                      */
/* fmc_tempered_run replaces ParallelTempering._ensemble_step               */
/* (src/flowMC/strategy/parallel_tempering.py:135-290): n_steps local steps */
/* of `kernel` (fmc_kernel_kind) for every (chain, temperature) row; x is   */
/* [n_chains, n_temps, d] (row c*n_temps + t), and row r targets            */
/* TemperedPDF.tempered_log_pdf (src/flowMC/resource/logPDF.py:104-106)     */
/* with a zero log prior: log p(x) / temperatures[r % n_temps]. logp [rows] */
/* receives the tempered log-density of the final state (may be NULL);      */
/* out_acc [rows, acc_len] (may be NULL) the decision of step step0 + s at  */
/* column step0 + s. Noise layout, Philox streams, metric, momentum_chol   */
/* and step_size are those of fmc_mala_run / fmc_hmc_run / fmc_grw_run over */
/* the rows (momentum_chol: HMC only).                                      */
/* ------------------------------------------------------------------------ */
typedef enum fmc_kernel_kind {
  FMC_KERNEL_MALA = 0,
  FMC_KERNEL_HMC = 1,
  FMC_KERNEL_GRW = 2
} fmc_kernel_kind;

int fmc_tempered_run(void* stream, int32_t kernel, const fmc_target* target, float* x, float* logp,
                     int64_t n_chains, int32_t n_temps, const float* temperatures, int32_t n_steps,
                     int32_t step0, float step_size, int32_t n_leapfrog, const float* metric,
                     const float* momentum_chol, uint32_t key0, uint32_t key1, const float* noise,
                     const float* log_u, float* out_acc, int64_t acc_len, const fmc_launch_cfg* cfg);

/* Replaces ParallelTempering._exchange (parallel_tempering.py:291-372) for */
/* all chains: per chain, adjacent pairs i = 0 .. n_temps-2 in order, swap  */
/* iff log u < (1/T[i+1] - 1/T[i]) (L_i - L_{i+1}) on the current slots.    */
/* x_in / x_out [n_chains, n_temps, d] (no aliasing); loglik [n_chains,     */
/* n_temps] holds the untempered log-likelihoods of x_in on entry and of    */
/* x_out on exit; accepts [n_chains, n_temps-1] receives 1.0f / 0.0f.       */
/* log_u [n_chains, n_temps-1] or NULL (Philox stream 9, counter (chain,    */
/* pair)). n_temps <= 64.                                                   */
int fmc_pt_exchange(void* stream, const float* x_in, float* x_out, float* loglik,
                    const float* temperatures, int64_t n_chains, int32_t n_temps, int32_t n_dim,
                    uint32_t key0, uint32_t key1, const float* log_u, float* accepts);

/* Replaces ParallelTempering._adapt_temperature (parallel_tempering.py:     */
/* 400-436), in place on temperatures[n_temps] from the exchange decisions  */
/* accepts [n_chains, n_temps-1]. n_temps <= 64.                            */
int fmc_pt_adapt(void* stream, float* temperatures, const float* accepts, int64_t n_chains,
                 int32_t n_temps);

/* Many independent HMC problems of uneven dimension in ONE launch.        */
/* slots[n_slots] (device memory) describe each problem; items[n_items]    */
/* (device memory) are the (slot, first chain, n chains) tiles, one per    */
/* workgroup. Each slot carries its own lanes_per_chain.                   */
typedef struct fmc_hmc_slot {
  fmc_target target;
  float* x;        /* [C_t, d_t] */
  float* logp;     /* [C_t]      */
  const float* metric; /* [d_t] or NULL (kinetic weights colsum(M) with momentum_chol) */
  const float* momentum_chol; /* [d_t, d_t] chol(inv(M)) of a dense M, or NULL */
  const float* noise;  /* [n_steps, C_t, d_t] or NULL */
  const float* log_u;  /* [n_steps, C_t] or NULL */
  float* out_pos;  /* [C_t, buf_len, d_t] or NULL */
  float* out_logp; /* [C_t, buf_len] or NULL */
  float* out_acc;  /* [C_t, buf_len] or NULL */
  int64_t n_chains;
  int64_t buf_len;
  int64_t cursor;
  uint32_t key0, key1;
  int32_t lanes_per_chain;
  int32_t reserved;
} fmc_hmc_slot;

typedef struct fmc_work_item {
  int32_t slot;
  int32_t chain0;
  int32_t n_chains;
  int32_t reserved;
} fmc_work_item;

/* block_size: a multiple of 64 in [64, 512]; every slot's lanes_per_chain   */
/* divides it.                                                               */
/* uniform_kind: -1 for any mix of targets and condition matrices, or the    */
/* FMC_TARGET_* kind that EVERY slot holds when no slot has a metric or      */
/* momentum factor (HMC's default unit condition matrix). With it, and       */
/* vec_per_lane 1 or 2, a kernel specialised on that kind and on each        */
/* slot's lanes_per_chain runs the steps (same results, no run-time target   */
/* dispatch); a slot that breaks the promise is not advanced and its logp    */
/* state is set to NaN.                                                      */
int fmc_hmc_run_batched(void* stream, const fmc_hmc_slot* slots, int32_t n_slots,
                        const fmc_work_item* items, int32_t n_items, int32_t max_dim,
                        int32_t n_steps, int32_t step0, float step_size, int32_t n_leapfrog,
                        int32_t thinning, int32_t vec_per_lane, int32_t block_size,
                        int32_t uniform_kind);

/* ------------------------------------------------------------------------ */
/* Per-step building blocks for user-supplied (non built-in) targets whose  */
/* gradient is computed outside this library.                              */
/* ------------------------------------------------------------------------ */
/* MALA proposal: prop = x + dt^2/2 grad + dt * xi   (MALA.py:60-63).     */
int fmc_mala_propose(void* stream, const float* x, const float* grad, float* prop,
                     int64_t n_chains, int32_t n_dim, float step_size, uint32_t key0,
                     uint32_t key1, int32_t step, const float* noise);
/* MALA MH accept (MALA.py:75-87) and thinned buffer write (take_steps.py:134-143).
 * x/logp/grad are updated in place to the accepted state. */
int fmc_mala_accept(void* stream, float* x, float* logp, float* grad, const float* prop,
                    const float* prop_logp, const float* prop_grad, int64_t n_chains,
                    int32_t n_dim, float step_size, uint32_t key0, uint32_t key1,
                    int32_t step, const float* log_u, float* out_pos, float* out_logp,
                    float* out_acc, int64_t buf_len, int64_t cursor, int32_t thinning);
/* Gaussian random walk proposal prop = x + step_size * xi
 * (Gaussian_random_walk.py:48-53) and its MH accept with thinned buffer
 * write (:55-60); x/logp are updated in place. */
int fmc_grw_propose(void* stream, const float* x, float* prop, int64_t n_chains, int32_t n_dim,
                    float step_size, uint32_t key0, uint32_t key1, int32_t step, const float* noise);
int fmc_grw_accept(void* stream, float* x, float* logp, const float* prop, const float* prop_logp,
                   int64_t n_chains, int32_t n_dim, uint32_t key0, uint32_t key1, int32_t step,
                   const float* log_u, float* out_pos, float* out_logp, float* out_acc,
                   int64_t buf_len, int64_t cursor, int32_t thinning);
/* HMC pieces (HMC.py:128-149): momentum draw + kinetic energy (FP64),
 * drift x += a * metric * p, kick p -= b * grad, and the FP64 accept.
 * momentum_chol / metric as for fmc_hmc_run (dense M: n_dim <= 4096). */
int fmc_hmc_momentum(void* stream, float* p, double* kinetic, int64_t n_chains,
                     int32_t n_dim, const float* metric, const float* momentum_chol,
                     uint32_t key0, uint32_t key1, int32_t step, const float* noise);
int fmc_hmc_drift(void* stream, float* x, const float* p, const float* metric,
                  int64_t n_elem_chains, int32_t n_dim, float coef);
int fmc_hmc_kick(void* stream, float* p, const float* grad_logp, int64_t n_chains,
                 int32_t n_dim, float coef);
int fmc_hmc_accept(void* stream, float* x, float* logp, const float* x_new,
                   const float* logp_new, const float* p_new, const double* kinetic0,
                   const float* metric, int64_t n_chains, int32_t n_dim, uint32_t key0,
                   uint32_t key1, int32_t step, const float* log_u, float* out_pos,
                   float* out_logp, float* out_acc, int64_t buf_len, int64_t cursor,
                   int32_t thinning);

/* ------------------------------------------------------------------------ */
/* Normalizing-flow independence Metropolis–Hastings                        */
/* (NFProposal.kernel's scan, src/flowMC/resource/kernel/NF_proposal.py:    */
/* 91-126). Per chain c, proposals s = 0..S-1:                             */
/*   r = (lp_prop - lp_cur) - (q_prop - q_cur); accept iff log_u < r.       */
/* Writes, for each step, the index of the proposal the chain sits on       */
/* (-1 = still at the entry state) plus logp and acceptance, then a gather  */
/* kernel fills the [C, buf_len, d] position buffer rows.                  */
/* Proposals are laid out [C, S, d] / [C, S].                              */
/* ------------------------------------------------------------------------ */
int fmc_nf_mh(void* stream, const float* x0, const float* lp0, const float* q0,
              const float* prop, const float* lp_prop, const float* q_prop,
              int64_t n_chains, int32_t n_steps, int32_t n_dim, uint32_t key0,
              uint32_t key1, const float* log_u, int32_t* out_idx, float* x_last,
              float* lp_last, float* out_pos, float* out_logp, float* out_acc,
              int64_t buf_len, int64_t cursor, int32_t thinning);

/* fmc_nf_mh without the out_idx argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid allocation). No counterpart in the reference. The [C, S]
 * proposal indices live in a fixed array of FMC_NF_MH_NOALLOC_MAX int32
 * (16 MiB) that the library holds on each device it runs on until the
 * process ends; n_chains * n_steps above that is refused with
 * FMC_ERR_INVALID_ARG. The call allocates nothing. Calls that overlap on one
 * device share the array and race: order them on one stream. All other
 * arguments and outputs are those of fmc_nf_mh; the stream must belong to
 * the current device. */
#define FMC_NF_MH_NOALLOC_MAX (1 << 22)
int fmc_nf_mh_noalloc(void* stream, const float* x0, const float* lp0, const float* q0,
                      const float* prop, const float* lp_prop, const float* q_prop,
                      int64_t n_chains, int32_t n_steps, int32_t n_dim, uint32_t key0,
                      uint32_t key1, const float* log_u, float* x_last, float* lp_last,
                      float* out_pos, float* out_logp, float* out_acc, int64_t buf_len,
                      int64_t cursor, int32_t thinning);

/* ------------------------------------------------------------------------ */
/* Random numbers (the exact streams the fused kernels use).                */
/* ------------------------------------------------------------------------ */
/* Raw Philox4x32-10 blocks: out[4*i .. 4*i+3] = philox(ctr_i, key) with    */
/* ctr_i = (ctr0 + i, ctr1, ctr2, ctr3). */
int fmc_philox_blocks(void* stream, uint32_t key0, uint32_t key1, uint32_t ctr0,
                      uint32_t ctr1, uint32_t ctr2, uint32_t ctr3, int64_t n_blocks,
                      uint32_t* out);
/* Standard normals the chain kernels draw for (chain c, step t, dim j):    */
/* out[(s*C + c)*d + j] for t = step0 + s, using stream id `stream_id`.    */
int fmc_rng_normal(void* stream, uint32_t key0, uint32_t key1, int64_t n_chains,
                   int32_t n_dim, int32_t n_steps, int32_t step0, int32_t stream_id,
                   float* out);
/* log-uniforms the chain kernels draw: out[s*C + c]. */
int fmc_rng_log_uniform(void* stream, uint32_t key0, uint32_t key1, int64_t n_chains,
                        int32_t n_steps, int32_t step0, int32_t stream_id, float* out);


/* ------------------------------------------------------------------------ */
/* RealNVP (src/flowMC/resource/model/nf_model/realNVP.py:102-228).        */
/* Per layer a MaskedCouplingLayer around an MLPAffine (common.py:127-208):  */
/* scale and shift MLPs [d, hidden, d] with ReLU read the conditioning dims; */
/* s = tanh(MLP_s) dt, t = MLP_t dt, y = (x + t) e^s on the transformed     */
/* dims, log-det = sum s. FP32 throughout.                                  */
/* theta (FP32): per layer, for the scale then the shift MLP: W1 [h][d],    */
/* b1 [h], W2 [d][h], b2 [d]; then data_mean [d], data_cov [d][d],          */
/* base_mean [d], base_cov [d][d] (only the diagonals of the covariances    */
/* are used). layout (device int32): per layer [n_cond, n_trans,            */
/* cond[n_dim], trans[n_dim]] (realNVP.py:155-158 masks).                   */
/* ------------------------------------------------------------------------ */
typedef struct fmc_rnvp {
  int32_t n_dim, n_layers, n_hidden;
  float dt;               /* MLPAffine dt (1 in RealNVP)                    */
  const float* theta;
  float* packed;          /* packed_size floats, written by fmc_rnvp_pack   */
  const int32_t* layout;
} fmc_rnvp;

/* Sizes: theta and packed (floats), weight-gradient partials per sample     */
/* tile (floats) and the tile (samples).                                    */
int fmc_rnvp_sizes(const fmc_rnvp* flow, int64_t* theta_size, int64_t* packed_size,
                   int64_t* partial_per_tile, int32_t* tile);
/* theta -> packed first-layer weights and biases (after every update).     */
int fmc_rnvp_pack(void* stream, const fmc_rnvp* flow);
/* One pass over n samples: dir 0 = RealNVP.forward (+ log N(y; 0, I) with  */
/* FMC_FLOW_BASE_LOGP: RealNVP.log_prob, realNVP.py:207-214), dir 1 =       */
/* RealNVP.inverse; flags as fmc_flow_pass (STANDARDIZE: data statistics;   */
/* SAMPLE_BASE with x_in NULL: base_dist draws from Philox stream 5 at      */
/* sample index sample0 + row, RealNVP.sample realNVP.py:200-205). logdet   */
/* [n] receives the summed log-dets (+ base log-density); saved [L, n, d]   */
/* the input of every layer (forward only); gy_out [n, d] with g_logp [n]   */
/* the gradient of g_logp * log N(y) at the output.                         */
int fmc_rnvp_pass(void* stream, const fmc_rnvp* flow, int32_t dir, int32_t flags,
                  const float* x_in, int64_t n, uint32_t key0, uint32_t key1, int64_t sample0,
                  float* x_out, float* logdet, float* saved, const float* g_logp, float* gy_out);
/* Adjoint of one layer (forward direction) for the loss sum g_logp *       */
/* log_prob: replaces eqx.filter_value_and_grad in NFModel.train_step       */
/* (nf_model/base.py:98-125) for RealNVP. gy [n, d] = dL/d(layer output)    */
/* -> gx [n, d] (no aliasing); partials [rows, partial_per_tile] receive    */
/* the per-tile weight gradients (rows 0 .. tiles-1), summed in a fixed     */
/* order by fmc_rnvp_weight_grad, which uses up to 16 more rows             */
/* (partial_rows >= tiles + min(16, ceil(tiles / 16))).                     */
int fmc_rnvp_backward_layer(void* stream, const fmc_rnvp* flow, int32_t layer, const float* x_in,
                            int64_t n, const float* g_logp, const float* gy, float* gx,
                            float* partials, int64_t n_partial_tiles);
int fmc_rnvp_weight_grad(void* stream, const fmc_rnvp* flow, int32_t layer, float* partials,
                         int64_t n_tiles, int64_t partial_rows, float* grad);

/* ------------------------------------------------------------------------ */
/* MaskedCouplingRQSpline flow                                              */
/* (src/flowMC/resource/model/nf_model/rqSpline.py:364-504 with the spline  */
/* of rqSpline.py:20-361 and the blocks of common.py:68-293).               */
/* The master parameters `theta` are FP32 in the leaf order documented in   */
/* flowmc_amd/csrc/fmc_flow_layout.h; `packed` / `packed_bias` hold the     */
/* MFMA operand images rebuilt from theta by fmc_flow_pack.                 */
/* ------------------------------------------------------------------------ */
typedef enum fmc_precision { FMC_PREC_F32 = 0, FMC_PREC_BF16 = 1 } fmc_precision;

/* fmc_flow.flags. With FMC_FLOW_DENSE_BASE the base Gaussian uses its full
 * covariance (Gaussian.log_prob / sample, common.py:285-293:
 * multivariate_normal.logpdf and mean + chol(cov) xi); fmc_flow_pack then also
 * writes chol(base_cov) and its log-determinant after the layers' packed
 * biases (packed_bias holds d*d + 1 more floats). Without it only the
 * diagonal of base_cov is read (exact for a diagonal covariance). */
enum { FMC_FLOW_DENSE_BASE = 1 };

#define FMC_FLOW_MAX_HIDDEN 4

typedef struct fmc_flow {
  int32_t n_dim, n_layers, n_bins;
  int32_t n_hidden;           /* hidden layers of every conditioner MLP, 1..4   */
                              /* (hidden_size of rqSpline.py:395, common.py:86-107) */
  int32_t hidden[FMC_FLOW_MAX_HIDDEN]; /* their widths; entries >= n_hidden unused */
  int32_t precision;          /* fmc_precision of the conditioner GEMMs */
  int32_t flags;              /* FMC_FLOW_DENSE_BASE: base_cov has off-diagonal entries */
  float range_min, range_max; /* spline range (rqSpline.py:399)        */
  float min_bin_size, min_knot_slope; /* rqSpline.py:299-300           */
  const float* theta;
  void* packed;
  float* packed_bias;
} fmc_flow;

/* Sizes (elements) of theta, packed (of the precision's type), packed_bias,
 * the rows of one backward layer's activation store, the sample tile of the
 * pass kernel and the samples per ScalarAffine partial of the backward (32). */
int fmc_flow_sizes(const fmc_flow* flow, int64_t* theta_size, int64_t* packed_size,
                   int64_t* packed_bias_size, int64_t* store_rows, int32_t* tile_fwd,
                   int32_t* tile_bwd);
/* theta -> packed MFMA operands + biases, and chol(base_cov) with
 * FMC_FLOW_DENSE_BASE (after every parameter update). */
int fmc_flow_pack(void* stream, const fmc_flow* flow);

enum { FMC_FLOW_FORWARD = 0, FMC_FLOW_INVERSE = 1, FMC_FLOW_INVERSE_EXACT = 2 };
enum { FMC_FLOW_STANDARDIZE = 1, FMC_FLOW_BASE_LOGP = 2, FMC_FLOW_SAMPLE_BASE = 4 };

/* One pass over n samples [n, d].
 * FORWARD  (MaskedCouplingRQSpline.forward :450-468): per layer ScalarAffine
 *   then RQ-spline coupling; with STANDARDIZE the input is first mapped by
 *   (x - data_mean) / sqrt(diag data_cov); with BASE_LOGP `logdet` receives
 *   log_prob = logdet + log N(y; base_mean, diag base_cov) (log_prob :498-504).
 * INVERSE  (inverse :470-488): layers in reverse, within a layer
 *   ScalarAffine^-1 then spline^-1 (the reference's order);
 *   INVERSE_EXACT applies spline^-1 then ScalarAffine^-1.
 *   With SAMPLE_BASE the input is standard-normal noise (x_in, or Philox
 *   stream 5 keyed by (key0,key1) with counter (dim/4, sample0+i, 0, 5) when
 *   x_in is NULL) mapped to N(base_mean, diag base_cov); with STANDARDIZE
 *   the output is y * sqrt(diag data_cov) + data_mean (sample :490-496).
 *   INVERSE_EXACT with BASE_LOGP: `logdet` receives the flow log-density of
 *   the output (log_prob's value, without the -sum log sigma the reference
 *   omits too): log N(z; base) - sum of the inverse log-dets, where z is the
 *   pass input after SAMPLE_BASE. Refused for the reference-order INVERSE,
 *   whose output density is not the forward map's.
 * saved (FORWARD only, may be NULL): [n_layers, n, d] input of every layer.
 * gy_out (FORWARD + BASE_LOGP): g_logp[i] * d log N / dy. */
int fmc_flow_pass(void* stream, const fmc_flow* flow, int32_t direction, int32_t flags,
                  const float* x_in, int64_t n, uint32_t key0, uint32_t key1, int64_t sample0,
                  float* x_out, float* logdet, float* saved, const float* g_logp, float* gy_out);

/* Adjoint of forward layer `layer` (autodiff of NFModel.loss_fn,
 * nf_model/base.py:98-100, done by hand): given the layer input x_layer,
 * dL/dlog_prob per sample (g_logp) and dL/d(layer output) (gy), writes
 * dL/d(layer input) (gx), the activation / gradient store (fragment-blocked,
 * layout in flowmc_amd/csrc/fmc_flow_layout.h)
 * [store_rows, store_ld] (store_ld = n rounded up to 64) and ScalarAffine
 * partial gradients per 32 samples (n_partials >= store_ld / 32 slots of 2). */
int fmc_flow_backward_layer(void* stream, const fmc_flow* flow, int32_t layer,
                            const float* x_layer, int64_t n, const float* g_logp, const float* gy,
                            float* gx, void* store, int64_t store_ld, float* affine_partials,
                            int32_t n_partials);
/* The backward sweep of a training step: for layer = n_layers-1 .. 0 the
 * adjoint (fmc_flow_backward_layer, reading saved[layer] of the forward's
 * saved layer inputs [n_layers, n, d], gy -> gx, the two swapping between
 * layers: the input gradient ends in gy for an even layer count, else gx),
 * then that layer's weight gradient into grad (fmc_flow_weight_grad_split
 * with wg_workspace when split != 0, else fmc_flow_weight_grad). grad
 * (theta_size floats) is cleared by the call first, in stream order ahead of
 * the first weight gradient, so entries no layer writes (the masked input
 * columns, the data and base statistics) read zero. With
 * side_stream != NULL the weight gradient of each layer runs there beside
 * the next adjoint, on store0 / aff0 and store1 / aff1 in turn, and
 * `stream` joins side_stream before the call returns; NULL runs everything
 * on `stream` with store0 / aff0. The split weight gradient of layer 0,
 * which nothing runs beside, spreads its batch ranges over the whole device
 * when wg_workspace_bytes holds that (fmc_flow_weight_grad_ws_size reports
 * the larger size). Replaces the per-layer launches of the gradient of
 * NFModel.loss_fn (nf_model/base.py:98-125) with one call. */
int fmc_flow_backward(void* stream, void* side_stream, const fmc_flow* flow, const float* saved, int64_t n,
                      const float* g_logp, float* gy, float* gx, void* store0, void* store1,
                      int64_t store_ld, float* aff0, float* aff1, int32_t n_partials, float* grad,
                      int32_t split, float* wg_workspace, int64_t wg_workspace_bytes);
/* Weight / bias / ScalarAffine gradients of `layer` into grad (theta layout),
 * from the store and the store_ld / 32 affine partials of the backward. */
int fmc_flow_weight_grad(void* stream, const fmc_flow* flow, int32_t layer, const void* store,
                         int64_t store_ld, const float* affine_partials, int32_t n_partials,
                         float* grad);
/* The same gradients as a split-K product over the batch (the same
 * eqx.filter_value_and_grad in NFModel.train_step, nf_model/base.py:98-125):
 * one workgroup per (128 x 256 block of a weight matrix, batch range), FP32
 * partial blocks in workspace, then a fixed-order sum of the partials
 * (deterministic; equal to fmc_flow_weight_grad up to FP32 rounding).
 * workspace: at least fmc_flow_weight_grad_ws_size(flow, store_ld) bytes. */
int fmc_flow_weight_grad_ws_size(const fmc_flow* flow, int64_t store_ld, int64_t* bytes);
int fmc_flow_weight_grad_split(void* stream, const fmc_flow* flow, int32_t layer, const void* store,
                               int64_t store_ld, const float* affine_partials, int32_t n_partials,
                               float* grad, float* workspace, int64_t workspace_bytes);

/* optax.chain(clip_by_global_norm(max_norm), adamw(lr, b1, b2, eps, weight_decay))
 * applied to the flat vector theta (src/flowMC/resource/optimizer.py:19-22);
 * step counts from 1. workspace: workspace_len >= 2 doubles (up to 1024 partial
 * sums and the norm are kept there); theta, grad, m, v 16-byte aligned.
 * grad_norm (may be NULL) receives the global gradient norm. */
int fmc_adamw_clip(void* stream, float* theta, const float* grad, float* m, float* v, int64_t n,
                   float lr, float b1, float b2, float eps, float weight_decay, float max_norm,
                   int32_t step, double* workspace, int32_t workspace_len, float* grad_norm);
/* The same update with the step count read from device memory (*step_dev,
 * counting from 1) when the kernel runs: a launch recorded once in a HIP
 * graph and replayed takes each replay's count (the bias corrections
 * 1 - b^t are the same FP64 products as fmc_adamw_clip's). No reference
 * counterpart beyond fmc_adamw_clip's: for callers that replay a captured
 * training step. */
int fmc_adamw_clip_dev(void* stream, float* theta, const float* grad, float* m, float* v, int64_t n,
                       float lr, float b1, float b2, float eps, float weight_decay, float max_norm,
                       const int32_t* step_dev, double* workspace, int32_t workspace_len,
                       float* grad_norm);

#ifdef __cplusplus
}
#endif
#endif /* FLOWMC_H_ */
