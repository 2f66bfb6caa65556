/* szsnn_hip.h -- C ABI of the MI355X (gfx950) SZSNN simulation core.
 *
 * This library replaces the reference's OpenCL device runtime and kernels:
 *   szsnn1.1/src/Context.hpp:77-303       (setup, createKernel, createBuffer,
 *                                          writeBuffer/readBuffer, setKernelArg,
 *                                          enqueueKernel, finish)
 *   szsnn1.1/src/Simulation.ipp:42-138    (kernel construction, device data upload)
 *   szsnn1.1/src/Simulation.ipp:359-530   (per-step gather -> update -> scatter)
 *   szsnn1.1/src/cl/ kernels             (gather, scatter, X_update)
 *
 * Conventions
 *   - Plain C types only.  Every function returns 0 (SZSNN_OK) or a negative
 *     SZSNN_ERR_* code; szsnn_last_error() gives the message.  The C++ host
 *     layer turns a negative code into szsnn::SzsnnException, the reference's
 *     error contract (szsnn1.1/src/utils.h:144-155).
 *   - The library owns device memory; callers own host buffers.
 *   - One HIP stream per context.  Calls on one context are not reentrant;
 *     each call returns after its work is complete unless documented otherwise.
 *   - Ids: neurons 0..N-1, input sources N..N+M-1 (the reference's input
 *     partition follows the neuron partitions, NetworkGraph.hpp:153-188).
 *   - Delays: an API delay k (integer steps, k < max_delay <= 64) delivers a
 *     spike emitted in step c into the update of step c+k+1
 *     (NetworkGraph.ipp:416 stores k+1; scatter_kernel.cl:154-169,204-213;
 *     gather_kernel.cl:55-58,104).
 *   - Model parameters are in per-step units (times in steps, capacitance
 *     divided by dt): the reference kernels carry no dt.
 */
#ifndef SZSNN_HIP_H_
#define SZSNN_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SZSNN_OK 0
#define SZSNN_ERR_INVALID (-1)  /* bad argument or description */
#define SZSNN_ERR_HIP (-2)      /* HIP runtime failure */
#define SZSNN_ERR_OVERFLOW (-3) /* a device queue overflowed (spikes would be lost) */
#define SZSNN_ERR_NOMEM (-4)    /* allocation failed */
#define SZSNN_ERR_STATE (-5)    /* call not valid in the current state (also: after a persistent launch's
                                   grid barrier timed out, the network's state is lost and stepping calls
                                   fail until it is uploaded again) */

#define SZSNN_MODEL_IF_CURR_EXP 0 /* params: v_rest cm tau_m tau_refrac tau_syn_E tau_syn_I i_offset v_reset v_thresh; vars: v ie ii fired */
#define SZSNN_MODEL_IZHIKEVICH 1  /* params: a b c d; vars: v u */
#define SZSNN_MODEL_IF_COND_EXP 2 /* params: the IF_curr_exp nine, e_rev_E e_rev_I; vars: v ge gi fired */

#define SZSNN_SOURCE_NONE 0
#define SZSNN_SOURCE_CONST_RATE 1 /* szsnn1.1/src/models/Const_rate_InputModel.hpp:11-32 */
#define SZSNN_SOURCE_POISSON 2    /* spike iff draw(key, t, 3) < src_thresh */

typedef struct szsnn_ctx szsnn_ctx;
typedef struct szsnn_net szsnn_net;

/* Network description handed to szsnn_net_upload.  Replaces the reference's
 * NeuronData / SynapseData / SynapseIndex / GroupData initialisation
 * (szsnn1.1/src/Simulation.ipp:108-118): the library builds a delay-bucketed
 * CSR (rows keyed by (source id, delay)) from the synapse list. */
typedef struct szsnn_net_desc {
  int32_t model;              /* SZSNN_MODEL_* */
  uint32_t n_neurons;         /* N > 0 */
  uint32_t n_sources;         /* M >= 0 */
  int32_t source_model;       /* SZSNN_SOURCE_* */
  const float* params;        /* [n_params][N], SoA */
  const float* vars;          /* [n_vars][N], SoA, initial state */
  const float* src_rate;      /* [M] Const_rate rate (steps) or NULL */
  const float* src_last;      /* [M] Const_rate initial `last` or NULL (0) */
  const uint64_t* src_thresh; /* [M] Poisson thresholds or NULL */
  uint64_t n_syn;
  const uint32_t* syn_src;    /* [n_syn] source id (neuron or source) */
  const uint32_t* syn_tgt;    /* [n_syn] target neuron id */
  const float* syn_w;         /* [n_syn] weight; sign picks the E/I channel, 0 is inert */
  const uint16_t* syn_delay;  /* [n_syn] API delay k < max_delay */
  uint32_t max_delay;         /* D, 1..64 */
  int32_t frac_bits;          /* F: weights are used as multiples of 2^-F (rounded to nearest) */
  uint64_t seed;              /* random drive seed; key(id) = splitmix64 output #id of seed */
  const float* noise_amp;     /* [N] Gaussian current amplitude or NULL */
  uint32_t drive_k;           /* binomial drive: number of CDF thresholds (0 = none, <= 32) */
  const uint64_t* drive_cdf;  /* [drive_k] events = #{j : draw(key, t, 2) >= cdf[j]} */
  float drive_w;              /* weight of one drive event, E channel */
  uint32_t queue_capacity;    /* ignored: delivery goes through a ring of per-step accumulator
                                 slots and has no queue that could overflow (reference D4) */
  int32_t pull_layout;        /* 0: also build the target-blocked layout for dense gathers when the
                                 network allows it (ids < 2^23, bounded source spans), with 2-byte
                                 entries when it has at most 3 distinct weights, 4-byte entries up to
                                 255; 1: the same with 8-byte entries always; -1: never */
} szsnn_net_desc;

typedef struct szsnn_net_info {
  uint32_t n_neurons, n_sources, max_delay, n_params, n_vars;
  int32_t frac_bits;
  uint64_t n_syn_stored;      /* nonzero-weight synapses in the CSR */
  uint64_t n_rows;            /* (source, delay) rows = (N+M) * D */
  uint64_t pull_groups;       /* delay groups of the pull gather (0: no pull layout) */
  uint64_t device_bytes;      /* device memory held by the network */
  uint64_t raster_words;      /* u64 words per raster row = ceil((N+M)/64) */
  uint64_t pull_entry_bytes;  /* bytes per synapse entry of the pull layout (2, 4, 8; 0: none) */
  uint64_t pull_slices;       /* slices of 64 targets per pull block (4..8, chosen at upload so the
                                 gather's work items spread evenly over the CUs; 0: no pull layout) */
} szsnn_net_info;

/* Counters since upload (or the last szsnn_reset_counters). */
typedef struct szsnn_counters {
  uint64_t cycle;         /* completed steps (szsnn1.1/src/SzsnnTimer.cpp:66-71) */
  uint64_t events;        /* synaptic events delivered (counted at the consuming step) */
  uint64_t spikes;        /* spikes of neurons and sources */
  uint64_t drive_events;  /* binomial drive events (not synaptic) */
  uint64_t overflow;      /* >= 2^32 after a persistent-loop barrier timed out; else 0 */
  uint64_t rows;          /* (source, delay) rows delivered */
  uint64_t deferred;      /* steps deferred to the pull gather (predicted dense, or late) */
  uint64_t late_deferred; /* of them, deferred late: posted as heavy pushes, then pulled */
} szsnn_counters;

/* ---- context / device runtime (replaces Context.hpp:77-303) ---- */
int szsnn_create(int device, szsnn_ctx** out);          /* Context::setup, Context.cpp:136-224 */
int szsnn_destroy(szsnn_ctx* ctx);
const char* szsnn_last_error(const szsnn_ctx* ctx);     /* ctx may be NULL: last global error */
int szsnn_device_name(szsnn_ctx* ctx, char* buf, size_t len);
int szsnn_alloc(szsnn_ctx* ctx, size_t bytes, void** dptr);   /* Context::createBuffer */
int szsnn_free(szsnn_ctx* ctx, void* dptr);
int szsnn_memcpy_h2d(szsnn_ctx* ctx, void* dst, const void* src, size_t bytes); /* writeBuffer(CL_TRUE) */
int szsnn_memcpy_d2h(szsnn_ctx* ctx, void* dst, const void* src, size_t bytes); /* readBuffer(CL_TRUE) */
int szsnn_synchronize(szsnn_ctx* ctx);                  /* getCommandQueue().finish() */
void* szsnn_stream(szsnn_ctx* ctx);                     /* the context's hipStream_t */

/* ---- network (replaces Simulation ctor + prepareGPUMemory, Simulation.ipp:42-138) ---- */
int szsnn_net_upload(szsnn_ctx* ctx, const szsnn_net_desc* desc, szsnn_net** out);
int szsnn_net_destroy(szsnn_net* net);
int szsnn_net_get_info(szsnn_net* net, szsnn_net_info* out);

/* Batched mode: n_instances independent networks of one model (uneven sizes,
 * own seeds, own parameters) become one device network that every step
 * kernel covers in a single launch.  Neurons of all instances come first, in
 * instance order, then their input sources.  Each instance keeps the random
 * streams of its own seed and local ids, so it evolves exactly as its plain
 * upload would.  Instances must share model, frac_bits, drive table and
 * source model.  A batch starts with heavy pushes off and a gather policy
 * that never defers (every step pushed by the update blocks).  No reference
 * counterpart: the reference runs one network per Simulation
 * (Simulation.hpp:109-112). */
int szsnn_batch_upload(szsnn_ctx* ctx, const szsnn_net_desc* instances, uint32_t n_instances,
                       szsnn_net** out);
int szsnn_batch_instance(szsnn_net* net, uint32_t instance, uint32_t* neuron_offset, uint32_t* n_neurons,
                         uint32_t* source_offset, uint32_t* n_sources);

/* One step: gather -> update(+input model) -> scatter (Simulation.ipp:471-530).
 * forced_ids: ids to fire this step regardless of state (SpikeBuffer::setInput,
 * SpikeBuffer.cpp:97-129; OR-ed into the spike condition, also during the
 * refractory period).  out_bitmap: NULL or raster_words u64 receiving this
 * step's dense spike bitmap (bit i = id i fired; SpikeBuffer::getOutput,
 * SpikeBuffer.cpp:136-165). */
int szsnn_step(szsnn_net* net, const uint32_t* forced_ids, uint32_t n_forced, uint64_t* out_bitmap);

/* Run n steps with no host input: HIP-graph-captured step loop.  When a raster
 * buffer is set (szsnn_set_recording), steps with cycle in
 * [rec_first, rec_first + rec_steps) write their bitmap to it.  Returns after
 * completion; the error flag of the persistent loop is checked at the end. */
int szsnn_run(szsnn_net* net, uint64_t n_steps);
/* Same, but launched asynchronously: returns once queued on the stream. */
int szsnn_run_async(szsnn_net* net, uint64_t n_steps);

/* Volley input: every step of the following szsnn_run / szsnn_step fires, in
 * addition to the dynamics, the ids in [0, n_forced_prefix).  0 disables. */
int szsnn_set_volley(szsnn_net* net, uint32_t n_forced_prefix);

int szsnn_set_recording(szsnn_net* net, uint64_t rec_first, uint64_t rec_steps);
int szsnn_read_raster(szsnn_net* net, uint64_t* host_out); /* rec_steps * raster_words u64 */

int szsnn_get_vars(szsnn_net* net, float* host_out);       /* [n_vars][N] */
int szsnn_set_vars(szsnn_net* net, const float* host_in);  /* [n_vars][N] */
int szsnn_get_counters(szsnn_net* net, szsnn_counters* out);
int szsnn_reset_counters(szsnn_net* net);

/* Per-kernel timing: runs n steps launched one kernel at a time (no graph)
 * on the context stream and returns the summed durations of the gather and
 * update kernels in milliseconds, from HIP events that each kernel's dispatch
 * records at its start and end (hipExtLaunchKernelGGL). */
int szsnn_time_kernels(szsnn_net* net, uint64_t n_steps, double* gather_ms, double* update_ms);

/* One kernel of the current step, launched alone and waited for: the
 * reference's single-kernel harnesses Simulation::testUpdate / testScatter /
 * testGather (szsnn1.1/src/Simulation.ipp:142-210).  SZSNN_KERNEL_UPDATE is
 * the neuron (or source) update of the step and the scatter of its spikes,
 * which this library runs as one kernel; SZSNN_KERNEL_GATHER is the pull of
 * the synapses due at the step (it does nothing unless a deferred step has
 * synapses due).  The step counter does not advance, as in the reference's
 * harnesses: a following szsnn_step runs the same step again, so these are
 * for inspecting one kernel's effect (szsnn_get_vars) and time (ms), not
 * for stepping. */
#define SZSNN_KERNEL_GATHER 0
#define SZSNN_KERNEL_UPDATE 1
int szsnn_test_kernel(szsnn_net* net, int32_t kernel, double* ms);

/* Launch configuration of the graph mode (update block, gather block and
 * grid); 0 keeps the current value.  szsnn_autotune times candidates of both
 * modes on the live network for `steps` steps each, keeps the fastest, and
 * restores the state it found. */
int szsnn_set_tiling(szsnn_net* net, uint32_t update_block, uint32_t gather_block, uint32_t gather_grid);
int szsnn_get_tiling(szsnn_net* net, uint32_t* update_block, uint32_t* gather_block, uint32_t* gather_grid);

/* The whole launch configuration at once.  The reference's only tiling knobs
 * are compile-time constants (work-items per group, NEURONS_PER_WI neurons
 * per work-item, szsnn1.1/src/locales.h:17-46); here they are per network
 * and chosen by szsnn_autotune:
 *   update_block, update_ids_per_lane: the update kernel's block and ids per
 *     lane (1, 2 or 4: a block serves that many consecutive blocks of ids);
 *   gather_block, gather_grid: the pull gather's block and grid (grid 0: a
 *     default from the work items and CUs);
 *   gather_stage: delays whose spike bitmaps one pull pass stages in LDS at
 *     once (1 .. gather_stage_max; 0 on set: the maximum);
 *   gather_batch: 16-byte entry vectors per lane in each batch of the pull's
 *     two-batch load pipeline (2 or 4: deeper prefetch, fewer waves);
 *   mode, persist_block, persist_grid: as szsnn_set_mode. */
typedef struct szsnn_tiling {
  uint32_t update_block;
  uint32_t update_ids_per_lane;
  uint32_t gather_block;
  uint32_t gather_grid;
  uint32_t gather_stage;
  uint32_t gather_stage_max;  /* get only: the most delays one pass can stage */
  uint32_t gather_batch;      /* 16-byte entry vectors per lane loaded ahead together: 2 or 4 */
  int32_t mode;
  uint32_t persist_block;
  uint32_t persist_grid;
} szsnn_tiling;
int szsnn_set_tiling_ex(szsnn_net* net, const szsnn_tiling* tiling);
int szsnn_get_tiling_ex(szsnn_net* net, szsnn_tiling* tiling);

/* Gather policy: the spikes of a step are either pushed by the update that
 * produced them, each synapse into the accumulator slot of its delivery step
 * (cost ~ events, global atomics), or, with the pull layout, deferred: each
 * later step then pulls, per target block, the synapses of the deferred steps
 * due at it against their staged spike bitmaps (cost ~ synapses, no global
 * atomics).  A step is deferred when its predicted spikes (the previous
 * step's less the ids forced in it, plus the ids forced in this one) reach
 * dense_fraction x the ids with synapses;
 * dense_fraction <= 0 always defers, > 1 never does.  Both give identical
 * sums.  The default follows the pull entry width (where pushing a step
 * costs as much as pulling): 0.012 for 2-byte entries, 0.02 for 4-byte,
 * 0.04 for 8-byte (a batch: never).  With no heavy pushes and a policy
 * that never defers, steps from R = max_delay + 1 steps after the change
 * on launch no gather kernel.  has_pull reports whether the layout was
 * built. */
int szsnn_set_gather_policy(szsnn_net* net, float dense_fraction);
int szsnn_get_gather_policy(szsnn_net* net, float* dense_fraction, int32_t* has_pull);

/* Heavy pushes (networks with a pull layout; no counterpart in the
 * reference, whose push is one work-item per spiking id): on a pushed step,
 * an update block whose spiking ids carry at least min_entries synapse
 * entries posts their synapse ranges instead of adding them itself, and the
 * next step's gather launch adds them with its whole grid, so a burst
 * concentrated in a few blocks is not limited by one CU's atomic rate.  The
 * sums are identical either way.  Default 1024 (a batch: 0); 0 = every
 * block pushes its own ids.  Setting it on a network without a pull layout is an error
 * unless min_entries is 0. */
int szsnn_set_heavy_push(szsnn_net* net, uint32_t min_entries);
int szsnn_get_heavy_push(szsnn_net* net, uint32_t* min_entries);

/* Execution mode of szsnn_run.  GRAPH: each step is a gather launch (a no-op
 * unless something is due for the pull) and an update launch, replayed from a
 * captured HIP graph.  PERSISTENT: one cooperative launch runs all steps, a
 * grid barrier after each update and one after each pull; block 256 or 512
 * (the step loop keeps an id's state in registers: 220-241 VGPRs, 2 waves
 * per SIMD; a 1024-thread block would have to spill), grid 0 = the
 * co-resident maximum (clamped to it). */
#define SZSNN_MODE_GRAPH 0
#define SZSNN_MODE_PERSISTENT 1
int szsnn_set_mode(szsnn_net* net, int32_t mode, uint32_t block, uint32_t grid);
int szsnn_get_mode(szsnn_net* net, int32_t* mode, uint32_t* block, uint32_t* grid);
int szsnn_autotune(szsnn_net* net, uint32_t steps);

/* ---- discrete adjoint of the IF_curr_exp step (spike pattern held fixed) ----
 * No reference counterpart (the reference has no gradients); these are the
 * kernels behind the PyTorch autograd Function szsnn_amd.adjoint.  All
 * pointers are device pointers, laid out [.][B*H] (trial-major, neuron
 * fastest); `stream` is a hipStream_t (NULL: default stream); calls are
 * asynchronous on it.  K <= 64 steps per segment.
 *   forward_segment : K steps from `state` ([4][B*H]: v, se, si, fired; updated
 *                     in place) given synaptic inputs AE/AI [K][B*H]; writes
 *                     h = v - v_rest [K][B*H] and spike bits [K][ceil(B*H/64)].
 *   backward_segment: adjoint over the segment; ckpt = state at its start;
 *                     g = dL/dh [K][B*H]; carry [3][B*H] in/out; writes
 *                     LE/LI = dL/dse, dL/dsi [K][B*H].
 *   input_currents  : AE[r][j] = sum_i X[r][i] max(W1[i][j], 0) and AI[r][j] the
 *                     same with min(W1[i][j], 0), for rows r < rows of X
 *                     [rows][n_in] and W1 [n_in][H], over the nonzero X in
 *                     ascending i (spiking inputs are sparse).
 *   input_grad      : dW1[i][j] = sum_r X[r][i] * (W1[i][j] >= 0 ? LE[r][j] : LI[r][j])
 *                     over the rows r with X[r][i] != 0, ascending (the weight
 *                     gradient of input_currents' sign-split currents); needs a
 *                     device workspace of szsnn_adj_input_grad_workspace(rows,
 *                     n_in) bytes (the input spikes as a bit matrix, and a
 *                     word per tile of 64 inputs x 64 rows saying whether
 *                     the tile holds nonzero X other than 1); rows at most
 *                     65535 * 64.
 *   readout_forward : y_t = alpha y_{t-1} + drive_t over [T][BR]; y0 may be NULL.
 *   readout_backward: lam_t = scale (y_t - target_t) + alpha lam_{t+1}. */
int szsnn_adj_forward_segment(void* stream, uint32_t B, uint32_t H, uint32_t K, const float* params, float* state,
                              const float* AE, const float* AI, float* h, uint64_t* spike_bits);
int szsnn_adj_backward_segment(void* stream, uint32_t B, uint32_t H, uint32_t K, const float* params,
                               const float* ckpt, const uint64_t* spike_bits, const float* g, float* carry,
                               float* LE, float* LI);
int szsnn_adj_input_currents(void* stream, uint32_t rows, uint32_t n_in, uint32_t H, const float* X, const float* W1,
                             float* AE, float* AI);
size_t szsnn_adj_input_grad_workspace(uint32_t rows, uint32_t n_in);
int szsnn_adj_input_grad(void* stream, uint32_t rows, uint32_t n_in, uint32_t H, const float* X, const float* W1,
                         const float* LE, const float* LI, float* dW1, void* workspace, size_t workspace_bytes);
/* szsnn_adj_input_grad without a workspace argument and without allocating:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * it).  The input bit matrix and the tile words live in a 16 MiB device
 * array of the library (one per GPU, held until the process ends):
 * (n_in + ceil(n_in / 64)) * ceil(rows / 64) must not exceed 2^21, else
 * SZSNN_ERR_INVALID with a message in szsnn_last_error(NULL).
 * Every word of the scratch a call uses is written before it is read, so
 * nothing is cleared.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (different streams or threads) race.  Results equal szsnn_adj_input_grad's
 * bit for bit. */
int szsnn_adj_input_grad_noalloc(void* stream, uint32_t rows, uint32_t n_in, uint32_t H, const float* X,
                                 const float* W1, const float* LE, const float* LI, float* dW1);
int szsnn_readout_forward(void* stream, uint32_t BR, uint32_t T, float alpha, const float* y0, const float* drive,
                          float* y);
int szsnn_readout_backward(void* stream, uint32_t BR, uint32_t T, float alpha, float scale, const float* y,
                           const float* target, float* lam);

#ifdef __cplusplus
}
#endif

#endif /* SZSNN_HIP_H_ */
