// szsnn_runtime.hip -- host side of the C ABI (include/szsnn_hip.h).
//
// Replaces the reference's OpenCL runtime (szsnn1.1/src/Context.cpp,
// DataBuffer.hpp, KernelWrapper.cpp) and the device-format builders
// (NeuronData.ipp, SynapseData.ipp, SynapseIndex.cpp, GroupData.cpp,
// GlobalQueue.cpp, LocalQueue.cpp): one HIP stream per context, device
// memory owned by the library, a delay-bucketed CSR in place of the bundle
// matrix, and HIP-graph-captured step loops in place of blocking per-kernel
// enqueue/wait/finish (Simulation.ipp:402-460).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/szsnn_hip.h"
#include "szsnn_internal.h"
#include "szsnn_layout.h"

using namespace szsnn;

namespace {
std::mutex g_err_mu;
std::string g_err;  // last error of any context

void set_global_error(const std::string& m) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = m;
}
}  // namespace

namespace szsnn {
void set_global_error(const char* msg) { ::set_global_error(std::string(msg)); }
}  // namespace szsnn

struct szsnn_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  std::string err;
  std::string name;
};

struct Tiling {
  uint32_t update_block = 256;
  uint32_t update_npl = 1;       // ids per lane of the update kernel
  uint32_t gather_block = 256;
  uint32_t gather_grid = 0;
  uint32_t gather_stage = 1;     // delays whose bitmaps a pull pass stages at once
  uint32_t gather_batch = 2;     // 16-byte vectors per lane and batch of the pull
  int32_t mode = 0;              // SZSNN_MODE_GRAPH or SZSNN_MODE_PERSISTENT
  uint32_t persist_block = 256;
  uint32_t persist_grid = 0;     // 0: the co-resident maximum
};

struct szsnn_net {
  szsnn_ctx* ctx = nullptr;
  DevNet dn{};
  szsnn_net_info info{};
  uint64_t t_host = 0;
  Tiling tiling;
  std::vector<void*> allocs;
  uint64_t* forced_bits = nullptr;
  uint64_t* step_bitmap = nullptr;
  uint64_t* t_base = nullptr;
  // graph of kGraphSteps steps, valid for `graph_dn`/`graph_tiling`
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  bool graph_dirty = true;
  bool graph_gathers = true;      // the graph's steps launch gather_kernel
  // With no heavy pushes and a policy that never defers, gather_kernel has
  // nothing to do from this step on (no deferred step is still due), and
  // graph steps and single steps leave it out.
  uint64_t gather_quiet_from = 0;
  // autotune snapshot buffers
  std::vector<std::pair<void*, size_t>> snap;
  uint64_t active_ids = 0;        // ids with at least one synapse
  uint32_t last_forced = 0;       // ids forced by the last szsnn_step, at step last_forced_t
  uint64_t last_forced_t = 0;
  float dense_fraction = 0.04f;   // set per pull entry format at upload
  uint32_t stage_max = 1;         // the most delays a pull pass can stage (LDS)
  // set when a persistent launch's grid barrier timed out: its blocks left
  // the step loop at different steps, so the state is no longer one step's
  bool lost = false;
  // batched networks: per-instance id ranges (empty for a plain network)
  std::vector<uint32_t> inst_noff, inst_soff, inst_n, inst_m;
};

namespace {

constexpr uint32_t kGraphSteps = 50;

int fail(szsnn_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  set_global_error(msg);
  return code;
}

#define HIP_TRY(ctx, expr)                                                               \
  do {                                                                                   \
    hipError_t e__ = (expr);                                                             \
    if (e__ != hipSuccess) {                                                             \
      std::ostringstream os__;                                                           \
      os__ << #expr << " failed: " << hipGetErrorString(e__) << " (" << (int)e__ << ")"; \
      return fail((ctx), SZSNN_ERR_HIP, os__.str());                                     \
    }                                                                                    \
  } while (0)

int n_params_of(int model) {
  return model == kIfCurrExp ? 9 : model == kIzhikevich ? 4 : model == kIfCondExp ? 11 : -1;
}
int n_vars_of(int model) {
  return model == kIfCurrExp ? 4 : model == kIzhikevich ? 2 : model == kIfCondExp ? 4 : -1;
}

template <typename T>
int dev_alloc(szsnn_net* net, T** out, size_t count) {
  void* p = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    std::ostringstream os;
    os << "hipMalloc(" << bytes << ") failed: " << hipGetErrorString(e);
    return fail(net->ctx, SZSNN_ERR_NOMEM, os.str());
  }
  net->allocs.push_back(p);
  net->info.device_bytes += bytes;
  *out = static_cast<T*>(p);
  return SZSNN_OK;
}

template <typename T>
int upload(szsnn_net* net, T* dst, const T* src, size_t count) {
  if (count == 0) return SZSNN_OK;
  HIP_TRY(net->ctx, hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, net->ctx->stream));
  return SZSNN_OK;
}

void free_graph(szsnn_net* net) {
  if (net->graph_exec) (void)hipGraphExecDestroy(net->graph_exec);
  if (net->graph) (void)hipGraphDestroy(net->graph);
  net->graph_exec = nullptr;
  net->graph = nullptr;
  net->graph_dirty = true;
}

// The launch configuration takes effect: the staging depth travels in the
// device description, and the captured step graph is rebuilt.
void apply_tiling(szsnn_net* net) {
  net->dn.pull_stage = net->tiling.gather_stage;
  net->dn.pull_batch = net->tiling.gather_batch;
  net->graph_dirty = true;
}

// The most delays one pull pass can stage: their bitmaps with the sums and
// weight table in at most 64 KiB of LDS, no more than a delay group holds,
// and a size the gather kernel can be resident with.
uint32_t stage_limit(const DevNet& n) {
  if (!n.pull_off) return 1;
  const uint32_t per_group = (n.D + n.pull_groups - 1) / n.pull_groups;
  uint32_t s = 1;
  while (s < per_group && pull_dyn_bytes_for(n.pull_words, n.pull_fmt, n.pull_wn, s + 1) <= 65536) ++s;
  while (s > 1 && gather_max_blocks_per_cu(pull_dyn_bytes_for(n.pull_words, n.pull_fmt, n.pull_wn, s), 256, s) <= 0) --s;
  return s;
}

uint32_t default_gather_grid(const szsnn_net* net, uint32_t block) {
  (void)block;
  const uint64_t items = (uint64_t)net->dn.pull_nb * net->dn.pull_groups;
  const uint64_t cap = (uint64_t)std::max(net->ctx->cu_count, 1) * 8;
  return (uint32_t)std::max<uint64_t>(1, std::min(items, cap));
}

// Default gather policy per pull entry format: the fraction of ids spiking
// at which pushing a step's synapses (atomics, about 2.3e10 per second with
// heavy pushes) costs as much as pulling all synapses (measured on 1e8
// synapses: 2-byte entries 1.6e12 per second, 4-byte 1.1e12, 8-byte 6.1e11).
float default_dense_fraction(uint32_t fmt) {
  return fmt == kPullCount ? 0.012f : fmt == kPullTable ? 0.02f : 0.04f;
}

// Deferral threshold of the gather policy: a step is deferred to the pull
// when its predicted spikes reach dense_fraction x the ids with synapses.
uint32_t defer_threshold(const szsnn_net* net, float dense_fraction) {
  if (!net->dn.pull_off || dense_fraction > 1.0f) return 0xffffffffu;
  if (dense_fraction <= 0.0f) return 0;
  return (uint32_t)std::max<double>(1.0, std::ceil((double)dense_fraction * (double)net->active_ids));
}

// Whether the policy can leave work for gather_kernel: heavy pushes, or
// steps deferred to the pull.
bool gather_policy_active(const szsnn_net* net) {
  return net->dn.heavy_min != 0 || net->dn.defer_spikes != 0xffffffffu;
}

// Whether step t needs its gather_kernel launch: the policy is active, or
// was until fewer than R steps ago (a step deferred then is pulled up to D
// steps later).
bool gather_needed(const szsnn_net* net, uint64_t t) {
  return gather_policy_active(net) || t < net->gather_quiet_from;
}

// Before a policy change: if the old policy was active, steps before now
// may still be due.
void policy_changing(szsnn_net* net) {
  if (gather_policy_active(net)) net->gather_quiet_from = net->t_host + net->dn.R;
  net->graph_dirty = true;
}

int launch_step(szsnn_net* net, const DevNet& dn, uint32_t j, bool gather) {
  szsnn_ctx* ctx = net->ctx;
  if (gather) HIP_TRY(ctx, launch_gather(dn, j, net->tiling.gather_block, net->tiling.gather_grid, ctx->stream));
  HIP_TRY(ctx, launch_update(dn, j, net->tiling.update_block, net->tiling.update_npl, ctx->stream));
  return SZSNN_OK;
}

int build_graph(szsnn_net* net, bool gather) {
  free_graph(net);
  szsnn_ctx* ctx = net->ctx;
  HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = SZSNN_OK;
  for (uint32_t j = 0; j < kGraphSteps && rc == SZSNN_OK; ++j) rc = launch_step(net, net->dn, j, gather);
  if (rc == SZSNN_OK) {
    hipError_t e = launch_advance(net->t_base, kGraphSteps, ctx->stream);
    if (e != hipSuccess) rc = fail(ctx, SZSNN_ERR_HIP, std::string("advance launch: ") + hipGetErrorString(e));
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc != SZSNN_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(ctx, SZSNN_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  net->graph = g;
  HIP_TRY(ctx, hipGraphInstantiate(&net->graph_exec, g, nullptr, nullptr, 0));
  net->graph_dirty = false;
  net->graph_gathers = gather;
  return SZSNN_OK;
}

uint32_t persistent_grid(szsnn_net* net, uint32_t block, uint32_t want) {
  const int per_cu = persistent_max_blocks_per_cu(net->dn, block);
  const uint32_t maxg = (uint32_t)std::max(per_cu, 0) * (uint32_t)std::max(net->ctx->cu_count, 1);
  const uint32_t needed = (uint32_t)std::max<uint64_t>(1, ((uint64_t)net->dn.T + block - 1) / block);
  uint32_t g = want ? want : maxg;
  g = std::min(g, maxg);
  g = std::min(g, std::max(needed, (uint32_t)std::max(net->ctx->cu_count, 1)));
  return g;
}

int run_steps(szsnn_net* net, uint64_t n) {
  szsnn_ctx* ctx = net->ctx;
  if (net->tiling.mode == SZSNN_MODE_PERSISTENT) {
    if (n == 0) return SZSNN_OK;
    const uint32_t g = persistent_grid(net, net->tiling.persist_block, net->tiling.persist_grid);
    if (g == 0) return fail(ctx, SZSNN_ERR_INVALID, "persistent mode: block size has no co-resident occupancy");
    HIP_TRY(ctx, launch_heavy(net->dn, net->tiling.gather_grid, ctx->stream));
    HIP_TRY(ctx, launch_persistent(net->dn, n, net->t_base, net->tiling.persist_block, g, ctx->stream));
    net->t_host += n;
    return SZSNN_OK;
  }
  uint64_t done = 0;
  while (n - done >= kGraphSteps) {
    const bool gather = gather_needed(net, net->t_host + done);
    if (net->graph_dirty || gather != net->graph_gathers) {
      int rc = build_graph(net, gather);
      if (rc) return rc;
    }
    HIP_TRY(ctx, hipGraphLaunch(net->graph_exec, ctx->stream));
    done += kGraphSteps;
  }
  const uint64_t rem = n - done;
  for (uint32_t j = 0; j < rem; ++j) {
    int rc = launch_step(net, net->dn, j, gather_needed(net, net->t_host + done + j));
    if (rc) return rc;
  }
  if (rem) HIP_TRY(ctx, launch_advance(net->t_base, rem, ctx->stream));
  net->t_host += n;
  return SZSNN_OK;
}

// Sum of the stripes of every counter slot.
int read_counters(szsnn_net* net, unsigned long long* c) {
  std::vector<unsigned long long> raw(kCntWords);
  HIP_TRY(net->ctx, hipMemcpyAsync(raw.data(), net->dn.counters, kCntWords * 8, hipMemcpyDeviceToHost,
                                   net->ctx->stream));
  HIP_TRY(net->ctx, hipStreamSynchronize(net->ctx->stream));
  for (int k = 0; k < kCntSlots; ++k) {
    c[k] = 0;
    for (uint32_t st = 0; st < kCntStripes; ++st) c[k] += *counter(raw.data(), k, st);
  }
  return SZSNN_OK;
}

int check_overflow(szsnn_net* net) {
  unsigned long long c[kCntSlots];
  int rc = read_counters(net, c);
  if (rc) return rc;
  if (c[kCntOverflow] >= kBarTimeout) {
    // the blocks left the step loop where their waits ended: the step
    // counter is the device's (the launch did not advance it), and the
    // state is not one step's, so the network must be uploaded again
    (void)hipMemset(net->dn.barrier, 0, kBarWords * 4);
    uint64_t tb = net->t_host;
    if (hipMemcpy(&tb, net->t_base, 8, hipMemcpyDeviceToHost) == hipSuccess) net->t_host = tb;
    net->lost = true;
    return fail(net->ctx, SZSNN_ERR_STATE,
                "persistent loop: a grid barrier wait timed out (grid not co-resident); the network state is lost, "
                "upload it again");
  }
  return SZSNN_OK;
}

int check_usable(szsnn_net* net, const char* what) {
  if (net->lost)
    return fail(net->ctx, SZSNN_ERR_STATE,
                std::string(what) + ": the network state was lost when a grid barrier timed out; upload it again");
  return SZSNN_OK;
}

}  // namespace

extern "C" {

const char* szsnn_last_error(const szsnn_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  std::lock_guard<std::mutex> lk(g_err_mu);
  return g_err.c_str();
}

int szsnn_create(int device, szsnn_ctx** out) {
  if (!out) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_create: out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(nullptr, SZSNN_ERR_HIP, std::string("no HIP device: ") + hipGetErrorString(e));
  if (device < 0 || device >= count) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_create: device index out of range");
  auto* ctx = new (std::nothrow) szsnn_ctx();
  if (!ctx) return fail(nullptr, SZSNN_ERR_NOMEM, "szsnn_create: out of host memory");
  ctx->device = device;
  e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  hipDeviceProp_t prop;
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete ctx;
    return fail(nullptr, SZSNN_ERR_HIP, std::string("szsnn_create: ") + hipGetErrorString(e));
  }
  ctx->cu_count = prop.multiProcessorCount;
  // the marketing name can be empty (no libdrm name table on the host)
  const std::string pname = prop.name[0] ? std::string(prop.name) : "HIP device " + std::to_string(device);
  ctx->name = pname + " (" + prop.gcnArchName + ")";
  *out = ctx;
  return SZSNN_OK;
}

int szsnn_destroy(szsnn_ctx* ctx) {
  if (!ctx) return SZSNN_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SZSNN_OK;
}

int szsnn_device_name(szsnn_ctx* ctx, char* buf, size_t len) {
  if (!ctx || !buf || len == 0) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_device_name: bad argument");
  std::snprintf(buf, len, "%s cus=%d", ctx->name.c_str(), ctx->cu_count);
  return SZSNN_OK;
}

void* szsnn_stream(szsnn_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int szsnn_alloc(szsnn_ctx* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_alloc: bad argument");
  hipError_t e = hipMalloc(dptr, std::max<size_t>(bytes, 1));
  if (e != hipSuccess) return fail(ctx, SZSNN_ERR_NOMEM, std::string("szsnn_alloc: ") + hipGetErrorString(e));
  return SZSNN_OK;
}
int szsnn_free(szsnn_ctx* ctx, void* dptr) {
  if (!ctx) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_free: ctx is NULL");
  HIP_TRY(ctx, hipFree(dptr));
  return SZSNN_OK;
}
int szsnn_memcpy_h2d(szsnn_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_memcpy_h2d: ctx is NULL");
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SZSNN_OK;
}
int szsnn_memcpy_d2h(szsnn_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_memcpy_d2h: ctx is NULL");
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SZSNN_OK;
}
int szsnn_synchronize(szsnn_ctx* ctx) {
  if (!ctx) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_synchronize: ctx is NULL");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SZSNN_OK;
}

int szsnn_net_destroy(szsnn_net* net) {
  if (!net) return SZSNN_OK;
  (void)hipSetDevice(net->ctx->device);
  (void)hipStreamSynchronize(net->ctx->stream);
  free_graph(net);
  for (void* p : net->allocs) (void)hipFree(p);
  for (auto& s : net->snap) (void)hipFree(s.first);
  delete net;
  return SZSNN_OK;
}

}  // extern "C"

namespace {

// Shared body of szsnn_net_upload and szsnn_batch_upload.  keys: per-id
// random-stream keys, or null to derive them from d->seed.
int upload_impl(szsnn_ctx* ctx, const szsnn_net_desc* d, const uint64_t* id_keys, szsnn_net** out) {
  if (!ctx || !d || !out) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: bad argument");
  *out = nullptr;
  // Simulation.ipp:97-101
  if (d->n_neurons == 0) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: network has 0 neurons");
  const int np = n_params_of(d->model), nv = n_vars_of(d->model);
  if (np < 0) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: unknown model");
  if (d->max_delay == 0 || d->max_delay > 64)
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: max_delay must be in 1..64");
  if (d->frac_bits < 0 || d->frac_bits > 62)
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: frac_bits must be in 0..62");
  if (!d->params || !d->vars) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: params/vars are NULL");
  if (d->n_sources) {
    if (d->source_model == kSrcConstRate && !d->src_rate)
      return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: Const_rate sources need src_rate");
    if (d->source_model == kSrcPoisson && !d->src_thresh)
      return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: Poisson sources need src_thresh");
    if (d->source_model < 0 || d->source_model > 2)
      return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: unknown source model");
  }
  if (d->drive_k > 32 || (d->drive_k && !d->drive_cdf))
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: drive_k must be <= 32 with drive_cdf");
  if (d->n_syn && (!d->syn_src || !d->syn_tgt || !d->syn_w || !d->syn_delay))
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: synapse arrays are NULL");
  const uint64_t N = d->n_neurons, M = d->n_sources, T = N + M, D = d->max_delay;
  if (T * D >= 0xffffffffull) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: (N+M)*max_delay too large");
  const double scale = std::ldexp(1.0, d->frac_bits);
  const uint64_t nrows = T * D;

  // --- device formats: the push entries with per-id ranges and the pull layout (szsnn_layout.h)
  // gather slots: resident blocks per CU at the default block size times CUs
  // (an occupancy query on the context's device)
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, SZSNN_ERR_HIP, "szsnn_net_upload: hipSetDevice failed");
  const uint64_t cus = (uint64_t)std::max(ctx->cu_count, 1);
  auto gather_slots = [&](uint32_t words, uint32_t fmt, uint32_t wn) {
    const int per_cu = gather_max_blocks_per_cu(pull_dyn_bytes_for(words, fmt, wn, 1), 256, 1);
    return (uint64_t)std::max(per_cu, 1) * cus;
  };
  HostLayout L;
  const SliceChooser choose = [&](uint32_t fmt, uint32_t wn, uint32_t words) {
    return choose_pull_slices(N, D, gather_slots(words, fmt, wn), cus);
  };
  if (!build_layout(*d, L, choose)) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: " + L.error);
  const uint64_t S = L.row_off[nrows];
  const std::vector<uint32_t>& syn_off = L.syn_off;
  const uint64_t active = L.active;
  const std::vector<uint64_t>& delay_bits = L.delay_bits;
  const std::vector<uint32_t>& pull_off = L.pull_off;
  const std::vector<uint32_t>& pull_lo = L.pull_lo;
  const uint32_t pull_words = L.pull_words;
  const uint64_t NB = L.NB;
  static_assert(sizeof(LayoutPair) == sizeof(uint2), "LayoutPair must match uint2");
  const uint2* syn_data = reinterpret_cast<const uint2*>(L.syn.data());
  const uint2* pull_syn_data = reinterpret_cast<const uint2*>(L.pull_syn.data());
  const size_t pull_syn_size = L.pull_syn.size();

  const double drive_q = std::nearbyint((double)d->drive_w * scale);
  if (!(std::fabs(drive_q) < 9.2e18)) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_net_upload: drive_w too large");

  auto* net = new (std::nothrow) szsnn_net();
  if (!net) return fail(ctx, SZSNN_ERR_NOMEM, "szsnn_net_upload: out of host memory");
  net->ctx = ctx;
  hipError_t he = hipSetDevice(ctx->device);
  if (he != hipSuccess) {
    delete net;
    return fail(ctx, SZSNN_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(he));
  }
  DevNet& n = net->dn;
  n.model = d->model;
  n.source_model = M ? d->source_model : kSrcNone;
  n.N = (uint32_t)N;
  n.M = (uint32_t)M;
  n.T = (uint32_t)T;
  n.D = (uint32_t)D;
  n.R = (uint32_t)D + 1;
  n.W = (uint32_t)((T + 63) / 64);
  n.npad = (uint32_t)((N + 63) / 64 * 64);
  n.drive_k = d->drive_k;
  n.frac_bits = d->frac_bits;
  n.acc_scale = (float)std::ldexp(1.0, -d->frac_bits);
  n.drive_w = (long long)drive_q;

  int rc = SZSNN_OK;
  auto guard = [&](int r) {
    if (r != SZSNN_OK && rc == SZSNN_OK) rc = r;
    return rc == SZSNN_OK;
  };
  float *params = nullptr, *vars = nullptr, *noise = nullptr, *src_rate = nullptr, *src_last = nullptr;
  long long *acc = nullptr, *acc_pull = nullptr;
  uint32_t* barrier = nullptr;
  // delay groups of the pull gather, for the staged words the layout needs
  const uint32_t groups =
      pull_off.empty() ? 1u : pull_groups_for(NB, D, gather_slots(L.pull_words, L.pull_fmt, (uint32_t)L.pull_wtab.size()));
  uint64_t *keys = nullptr, *src_thresh = nullptr, *drive_cdf = nullptr, *dbits = nullptr, *dstamp = nullptr;
  uint32_t* soff = nullptr;
  uint2* dsyn = nullptr;
  unsigned long long *counters = nullptr, *ring_cnt = nullptr, *spk_ring = nullptr;
  guard(dev_alloc(net, &params, (size_t)np * n.npad)) && guard(dev_alloc(net, &vars, (size_t)nv * n.npad)) &&
      guard(dev_alloc(net, &acc, (size_t)n.R * 2 * n.npad)) && guard(dev_alloc(net, &barrier, kBarWords)) &&
      guard(dev_alloc(net, &keys, T)) && guard(dev_alloc(net, &soff, T + 1)) && guard(dev_alloc(net, &dsyn, S)) &&
      guard(dev_alloc(net, &dbits, T)) && guard(dev_alloc(net, &dstamp, n.R)) &&
      guard(dev_alloc(net, &ring_cnt, kRingCntWords(n.R))) && guard(dev_alloc(net, &spk_ring, kSpkRingWords)) &&
      guard(dev_alloc(net, &counters, kCntWords)) && guard(dev_alloc(net, &net->t_base, 1)) &&
      guard(dev_alloc(net, &net->forced_bits, n.W)) && guard(dev_alloc(net, &net->step_bitmap, n.W));
  if (rc == SZSNN_OK && d->noise_amp) guard(dev_alloc(net, &noise, N));
  if (rc == SZSNN_OK && M && n.source_model == kSrcConstRate)
    guard(dev_alloc(net, &src_rate, M)) && guard(dev_alloc(net, &src_last, M));
  if (rc == SZSNN_OK && M && n.source_model == kSrcPoisson) guard(dev_alloc(net, &src_thresh, M));
  if (rc == SZSNN_OK && d->drive_k) guard(dev_alloc(net, &drive_cdf, d->drive_k));
  uint64_t* hist = nullptr;
  uint32_t *d_pull_off = nullptr, *d_pull_lo = nullptr;
  uint2* d_pull_syn = nullptr;
  uint32_t* d_pull_syn32 = nullptr;
  uint16_t* d_pull_syn16 = nullptr;
  int32_t* d_pull_wtab = nullptr;
  uint16_t* d_pull_lane = nullptr;
  uint32_t* heavy_cnt = nullptr;
  uint2* heavy_list = nullptr;
  uint64_t* hvy = nullptr;
  if (rc == SZSNN_OK) guard(dev_alloc(net, &hist, (size_t)n.R * n.W));
  // heavy pushes are drained by the gather launch of the next step, so they
  // exist with a pull layout only
  if (rc == SZSNN_OK && !pull_off.empty())
    guard(dev_alloc(net, &heavy_cnt, 2)) && guard(dev_alloc(net, &heavy_list, (size_t)2 * T)) &&
        guard(dev_alloc(net, &hvy, (size_t)n.R * n.W));
  if (rc == SZSNN_OK && !pull_off.empty())
    guard(dev_alloc(net, &d_pull_off, pull_off.size())) &&
        (L.pull_fmt == kPullWide    ? guard(dev_alloc(net, &d_pull_syn, pull_syn_size))
         : L.pull_fmt == kPullTable ? guard(dev_alloc(net, &d_pull_syn32, L.pull_syn32.size()))
                                    : guard(dev_alloc(net, &d_pull_syn16, L.pull_syn16.size()))) &&
        (L.pull_fmt == kPullWide || guard(dev_alloc(net, &d_pull_wtab, L.pull_wtab.size()))) &&
        guard(dev_alloc(net, &d_pull_lo, pull_lo.size())) && guard(dev_alloc(net, &d_pull_lane, L.pull_lane.size())) &&
        guard(dev_alloc(net, &acc_pull, (size_t)groups * 2 * n.npad));
  if (rc != SZSNN_OK) {
    szsnn_net_destroy(net);
    return rc;
  }
  // host staging with the padded SoA stride
  std::vector<float> hp((size_t)np * n.npad, 0.0f), hv((size_t)nv * n.npad, 0.0f);
  for (int k = 0; k < np; ++k) std::memcpy(&hp[(size_t)k * n.npad], d->params + (size_t)k * N, N * 4);
  if (d->model == kIfCurrExp || d->model == kIfCondExp) {
    // rows 1, 2, 4, 5 -> 1/cm, 1/tau_m, 1 - 1/tau_syn_E, 1 - 1/tau_syn_I (FP32, as the step evaluates them)
    for (uint64_t i = 0; i < N; ++i) {
      float* c = &hp[1 * (size_t)n.npad + i];
      float* m = &hp[2 * (size_t)n.npad + i];
      float* e = &hp[4 * (size_t)n.npad + i];
      float* q = &hp[5 * (size_t)n.npad + i];
      *c = 1.0f / *c;
      *m = 1.0f / *m;
      *e = 1.0f - 1.0f / *e;
      *q = 1.0f - 1.0f / *q;
    }
  }
  for (int k = 0; k < nv; ++k) std::memcpy(&hv[(size_t)k * n.npad], d->vars + (size_t)k * N, N * 4);
  std::vector<uint64_t> hk(T);
  for (uint64_t i = 0; i < T; ++i) hk[i] = id_keys ? id_keys[i] : splitmix_at(d->seed, i);
  std::vector<float> hsl;
  if (src_last) {
    hsl.assign(M, 0.0f);
    if (d->src_last) std::memcpy(hsl.data(), d->src_last, M * 4);
  }
  guard(upload(net, params, hp.data(), hp.size())) && guard(upload(net, vars, hv.data(), hv.size())) &&
      guard(upload(net, keys, hk.data(), T)) && guard(upload(net, soff, syn_off.data(), T + 1)) &&
      guard(upload(net, dsyn, syn_data, S)) && guard(upload(net, dbits, delay_bits.data(), T));
  if (rc == SZSNN_OK && noise) guard(upload(net, noise, d->noise_amp, N));
  if (rc == SZSNN_OK && src_rate) guard(upload(net, src_rate, d->src_rate, M)) && guard(upload(net, src_last, hsl.data(), M));
  if (rc == SZSNN_OK && src_thresh) guard(upload(net, src_thresh, d->src_thresh, M));
  if (rc == SZSNN_OK && drive_cdf) guard(upload(net, drive_cdf, d->drive_cdf, d->drive_k));
  if (rc == SZSNN_OK && d_pull_off)
    guard(upload(net, d_pull_off, pull_off.data(), pull_off.size())) &&
        (L.pull_fmt == kPullWide    ? guard(upload(net, d_pull_syn, pull_syn_data, pull_syn_size))
         : L.pull_fmt == kPullTable ? guard(upload(net, d_pull_syn32, L.pull_syn32.data(), L.pull_syn32.size()))
                                    : guard(upload(net, d_pull_syn16, L.pull_syn16.data(), L.pull_syn16.size()))) &&
        (L.pull_fmt == kPullWide || guard(upload(net, d_pull_wtab, L.pull_wtab.data(), L.pull_wtab.size()))) &&
        guard(upload(net, d_pull_lo, pull_lo.data(), pull_lo.size())) &&
        guard(upload(net, d_pull_lane, L.pull_lane.data(), L.pull_lane.size()));
  if (rc == SZSNN_OK) {
    hipStream_t s = ctx->stream;
    he = hipMemsetAsync(acc, 0, (size_t)n.R * 2 * n.npad * 8, s);
    if (he == hipSuccess && acc_pull) he = hipMemsetAsync(acc_pull, 0, (size_t)groups * 2 * n.npad * 8, s);
    if (he == hipSuccess) he = hipMemsetAsync(barrier, 0, kBarWords * 4, s);
    if (he == hipSuccess) he = hipMemsetAsync(dstamp, 0, (size_t)n.R * 8, s);
    if (he == hipSuccess) he = hipMemsetAsync(ring_cnt, 0, kRingCntWords(n.R) * 8, s);
    if (he == hipSuccess) he = hipMemsetAsync(spk_ring, 0, kSpkRingWords * 8, s);
    if (he == hipSuccess) he = hipMemsetAsync(counters, 0, kCntWords * 8, s);
    if (he == hipSuccess) he = hipMemsetAsync(net->t_base, 0, 8, s);
    if (he == hipSuccess) he = hipMemsetAsync(net->forced_bits, 0, (size_t)n.W * 8, s);
    if (he == hipSuccess) he = hipMemsetAsync(hist, 0, (size_t)n.R * n.W * 8, s);
    if (he == hipSuccess && heavy_cnt) he = hipMemsetAsync(heavy_cnt, 0, 2 * 4, s);
    if (he == hipSuccess && hvy) he = hipMemsetAsync(hvy, 0, (size_t)n.R * n.W * 8, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) rc = fail(ctx, SZSNN_ERR_HIP, std::string("szsnn_net_upload: ") + hipGetErrorString(he));
  }
  if (rc != SZSNN_OK) {
    szsnn_net_destroy(net);
    return rc;
  }
  n.params = params;
  n.vars = vars;
  n.acc = acc;
  n.acc_pull = acc_pull;
  n.pull_groups = groups;
  n.barrier = barrier;
  n.keys = keys;
  n.noise_amp = noise;
  n.src_rate = src_rate;
  n.src_last = src_last;
  n.src_thresh = src_thresh;
  n.drive_cdf = drive_cdf;
  n.syn_off = soff;
  n.syn = dsyn;
  n.delay_bits = dbits;
  n.dstamp = dstamp;
  n.ring_cnt = ring_cnt;
  n.spk_ring = spk_ring;
  n.forced_count = 0;
  n.forced_prev = 0;
  n.forced_prev_t = 0;
  n.volley_t0 = 0;
  n.raster = nullptr;
  n.rec_first = 0;
  n.rec_steps = 0;
  n.counters = counters;
  n.t_base = net->t_base;
  n.volley = 0;
  n.forced_bits = nullptr;
  n.hist = hist;
  n.pull_off = d_pull_off;
  n.pull_lane = d_pull_lane;
  n.pull_syn = d_pull_syn;
  n.pull_syn32 = d_pull_syn32;
  n.pull_syn16 = d_pull_syn16;
  n.pull_wtab = d_pull_wtab;
  n.pull_fmt = L.pull_fmt;
  n.pull_wn = (uint32_t)L.pull_wtab.size();
  n.pull_lo = d_pull_lo;
  n.pull_nb = d_pull_off ? (uint32_t)NB : 0;
  n.pull_slices = L.pull_slices;
  n.pull_tb = 64 * L.pull_slices;
  // the null vector follows the last slice (its entries per vector by format)
  n.pull_null = d_pull_off ? pull_off.back() / (L.pull_fmt == kPullWide ? 2u : L.pull_fmt == kPullTable ? 4u : 8u) : 0;
  n.pull_words = pull_words;
  n.heavy_min = heavy_cnt ? kHeavyPushMin : 0;
  n.heavy_cnt = heavy_cnt;
  n.heavy_list = heavy_list;
  n.hvy = hvy;
  net->active_ids = active;
  net->dense_fraction = default_dense_fraction(d_pull_off ? L.pull_fmt : kPullWide);
  n.defer_spikes = defer_threshold(net, net->dense_fraction);

  net->info.n_neurons = (uint32_t)N;
  net->info.n_sources = (uint32_t)M;
  net->info.max_delay = (uint32_t)D;
  net->info.n_params = (uint32_t)np;
  net->info.n_vars = (uint32_t)nv;
  net->info.frac_bits = d->frac_bits;
  net->info.n_syn_stored = S;
  net->info.n_rows = nrows;
  net->info.pull_groups = d_pull_off ? groups : 0;
  net->info.pull_entry_bytes = !d_pull_off ? 0 : L.pull_fmt == kPullWide ? 8 : L.pull_fmt == kPullTable ? 4 : 2;
  net->info.raster_words = n.W;
  net->info.pull_slices = d_pull_off ? L.pull_slices : 0;
  net->tiling.gather_grid = default_gather_grid(net, net->tiling.gather_block);
  net->stage_max = stage_limit(n);
  n.pull_stage = 1;
  n.pull_batch = 2;
  *out = net;
  return SZSNN_OK;
}

}  // namespace

extern "C" {

int szsnn_net_upload(szsnn_ctx* ctx, const szsnn_net_desc* d, szsnn_net** out) {
  return upload_impl(ctx, d, nullptr, out);
}

int szsnn_batch_upload(szsnn_ctx* ctx, const szsnn_net_desc* inst, uint32_t n_inst, szsnn_net** out) {
  if (!ctx || !inst || !out || n_inst == 0) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_batch_upload: bad argument");
  *out = nullptr;
  const szsnn_net_desc& d0 = inst[0];
  const int np = n_params_of(d0.model), nv = n_vars_of(d0.model);
  if (np < 0) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_batch_upload: unknown model");
  uint64_t N = 0, M = 0, S = 0;
  uint32_t D = 0;
  bool noise = false;
  for (uint32_t b = 0; b < n_inst; ++b) {
    const szsnn_net_desc& d = inst[b];
    if (d.model != d0.model || d.frac_bits != d0.frac_bits || d.drive_k != d0.drive_k ||
        d.drive_w != d0.drive_w || (d.n_sources && d0.n_sources && d.source_model != d0.source_model))
      return fail(ctx, SZSNN_ERR_INVALID,
                  "szsnn_batch_upload: instances must share model, frac_bits, drive and source model");
    if (d.n_neurons == 0) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_batch_upload: instance with 0 neurons");
    if (d.drive_k && std::memcmp(d.drive_cdf, d0.drive_cdf, d.drive_k * 8) != 0)
      return fail(ctx, SZSNN_ERR_INVALID, "szsnn_batch_upload: instances must share the drive table");
    N += d.n_neurons;
    M += d.n_sources;
    S += d.n_syn;
    D = std::max(D, d.max_delay);
    noise = noise || d.noise_amp;
  }
  if (N + M >= 0xffffffffull) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_batch_upload: too many ids");
  int src_model = kSrcNone;
  for (uint32_t b = 0; b < n_inst; ++b)
    if (inst[b].n_sources) src_model = inst[b].source_model;
  // concatenation: all instances' neurons, then all instances' sources
  std::vector<float> params((size_t)np * N), vars((size_t)nv * N), amp(noise ? N : 0, 0.0f);
  std::vector<float> rate(M, 0.0f), last(M, 0.0f);
  std::vector<uint64_t> thresh(M, 0), keys(N + M);
  std::vector<uint32_t> ssrc(S), stgt(S);
  std::vector<float> sw(S);
  std::vector<uint16_t> sd(S);
  std::vector<uint32_t> noff(n_inst), soff(n_inst);
  uint64_t no = 0, so = 0, sy = 0;
  for (uint32_t b = 0; b < n_inst; ++b) {
    const szsnn_net_desc& d = inst[b];
    noff[b] = (uint32_t)no;
    soff[b] = (uint32_t)(N + so);
    const uint64_t n = d.n_neurons, m = d.n_sources;
    for (int k = 0; k < np; ++k) std::memcpy(&params[(size_t)k * N + no], d.params + (size_t)k * n, n * 4);
    for (int k = 0; k < nv; ++k) std::memcpy(&vars[(size_t)k * N + no], d.vars + (size_t)k * n, n * 4);
    if (d.noise_amp) std::memcpy(&amp[no], d.noise_amp, n * 4);
    if (m && d.src_rate) std::memcpy(&rate[so], d.src_rate, m * 4);
    if (m && d.src_last) std::memcpy(&last[so], d.src_last, m * 4);
    if (m && d.src_thresh) std::memcpy(&thresh[so], d.src_thresh, m * 8);
    // keys follow each instance's own seed and local ids, as in its plain run
    for (uint64_t i = 0; i < n; ++i) keys[no + i] = splitmix_at(d.seed, i);
    for (uint64_t j = 0; j < m; ++j) keys[N + so + j] = splitmix_at(d.seed, n + j);
    for (uint64_t q = 0; q < d.n_syn; ++q) {
      const uint32_t s = d.syn_src[q], t = d.syn_tgt[q];
      if (s >= n + m || t >= n) {
        std::ostringstream os;
        os << "szsnn_batch_upload: instance " << b << " synapse " << q << " connects nonexisting ids";
        return fail(ctx, SZSNN_ERR_INVALID, os.str());
      }
      ssrc[sy] = s < n ? (uint32_t)(no + s) : (uint32_t)(N + so + (s - n));
      stgt[sy] = (uint32_t)(no + t);
      sw[sy] = d.syn_w[q];
      sd[sy] = d.syn_delay[q];
      ++sy;
    }
    no += n;
    so += m;
  }
  szsnn_net_desc c = d0;
  c.n_neurons = (uint32_t)N;
  c.n_sources = (uint32_t)M;
  c.source_model = src_model;
  c.params = params.data();
  c.vars = vars.data();
  c.src_rate = src_model == kSrcConstRate ? rate.data() : nullptr;
  c.src_last = src_model == kSrcConstRate ? last.data() : nullptr;
  c.src_thresh = src_model == kSrcPoisson ? thresh.data() : nullptr;
  c.n_syn = S;
  c.syn_src = ssrc.data();
  c.syn_tgt = stgt.data();
  c.syn_w = sw.data();
  c.syn_delay = sd.data();
  c.max_delay = D;
  c.noise_amp = noise ? amp.data() : nullptr;
  c.queue_capacity = 0;
  int rc = upload_impl(ctx, &c, keys.data(), out);
  if (rc) return rc;
  (*out)->inst_noff = std::move(noff);
  (*out)->inst_soff = std::move(soff);
  (*out)->inst_n.resize(n_inst);
  (*out)->inst_m.resize(n_inst);
  for (uint32_t b = 0; b < n_inst; ++b) {
    (*out)->inst_n[b] = inst[b].n_neurons;
    (*out)->inst_m[b] = inst[b].n_sources;
  }
  // Default policy of a batch: push every step from the update blocks.  The
  // instances are independent networks, so a step's spikes are spread over
  // all of them and pulling every synapse of the batch does not pay; and
  // with rows of ~100 synapses the update grid pushes a burst's spikes
  // faster than the gather grid would as heavy pushes (256 Izhikevich
  // instances: 84.3 us per step against 85.7 with heavy pushes at 1024
  // entries).  With neither, graph steps leave gather_kernel out.
  (*out)->dn.heavy_min = 0;
  (*out)->dense_fraction = 2.0f;
  (*out)->dn.defer_spikes = defer_threshold(*out, 2.0f);
  return SZSNN_OK;
}

int szsnn_batch_instance(szsnn_net* net, uint32_t b, uint32_t* neuron_offset, uint32_t* n_neurons,
                         uint32_t* source_offset, uint32_t* n_sources) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_batch_instance: net is NULL");
  if (b >= net->inst_n.size())
    return fail(net->ctx, SZSNN_ERR_INVALID, "szsnn_batch_instance: instance index out of range");
  if (neuron_offset) *neuron_offset = net->inst_noff[b];
  if (n_neurons) *n_neurons = net->inst_n[b];
  if (source_offset) *source_offset = net->inst_soff[b];
  if (n_sources) *n_sources = net->inst_m[b];
  return SZSNN_OK;
}

int szsnn_net_get_info(szsnn_net* net, szsnn_net_info* out) {
  if (!net || !out) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_net_get_info: bad argument");
  *out = net->info;
  return SZSNN_OK;
}

int szsnn_step(szsnn_net* net, const uint32_t* forced_ids, uint32_t n_forced, uint64_t* out_bitmap) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_step: net is NULL");
  if (int rc = check_usable(net, "szsnn_step")) return rc;
  szsnn_ctx* ctx = net->ctx;
  DevNet dn = net->dn;
  if (n_forced) {
    if (!forced_ids) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_step: forced_ids is NULL");
    std::vector<uint64_t> bits(dn.W, 0);
    for (uint32_t f = 0; f < n_forced; ++f) {
      const uint32_t id = forced_ids[f];
      if (id >= dn.T) {  // SpikeBuffer.cpp:110-113
        std::ostringstream os;
        os << "szsnn_step: neuron with idx " << id << " not in the network";
        return fail(ctx, SZSNN_ERR_INVALID, os.str());
      }
      bits[id >> 6] |= 1ull << (id & 63);
    }
    HIP_TRY(ctx, hipMemcpyAsync(net->forced_bits, bits.data(), dn.W * 8, hipMemcpyHostToDevice, ctx->stream));
    dn.forced_bits = net->forced_bits;
    dn.forced_count = n_forced;
  }
  const uint64_t t = net->t_host;
  // the previous host step's forced ids, taken out of its spike count in
  // this step's deferral prediction
  dn.forced_prev = net->last_forced;
  dn.forced_prev_t = net->last_forced_t;
  uint64_t* row = nullptr;
  if (dn.raster && t >= dn.rec_first && t - dn.rec_first < dn.rec_steps) {
    row = dn.raster + (t - dn.rec_first) * dn.W;
  } else if (out_bitmap) {
    dn.raster = net->step_bitmap;
    dn.rec_first = t;
    dn.rec_steps = 1;
    row = net->step_bitmap;
  }
  int rc = launch_step(net, dn, 0, gather_needed(net, t));
  if (rc) return rc;
  HIP_TRY(ctx, launch_advance(net->t_base, 1, ctx->stream));
  net->last_forced = n_forced;
  net->last_forced_t = t;
  net->t_host += 1;
  if (out_bitmap)
    HIP_TRY(ctx, hipMemcpyAsync(out_bitmap, row, dn.W * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return check_overflow(net);
}

int szsnn_run_async(szsnn_net* net, uint64_t n_steps) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_run_async: net is NULL");
  if (int rc = check_usable(net, "szsnn_run")) return rc;
  return run_steps(net, n_steps);
}

int szsnn_run(szsnn_net* net, uint64_t n_steps) {
  int rc = szsnn_run_async(net, n_steps);
  if (rc) return rc;
  return check_overflow(net);
}

int szsnn_set_volley(szsnn_net* net, uint32_t n_forced_prefix) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_set_volley: net is NULL");
  if (n_forced_prefix > net->dn.T) return fail(net->ctx, SZSNN_ERR_INVALID, "szsnn_set_volley: prefix beyond the network");
  if (n_forced_prefix != net->dn.volley) net->dn.volley_t0 = net->t_host;
  net->dn.volley = n_forced_prefix;
  net->graph_dirty = true;
  return SZSNN_OK;
}

int szsnn_set_recording(szsnn_net* net, uint64_t rec_first, uint64_t rec_steps) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_set_recording: net is NULL");
  szsnn_ctx* ctx = net->ctx;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (net->dn.raster) {
    auto it = std::find(net->allocs.begin(), net->allocs.end(), (void*)net->dn.raster);
    if (it != net->allocs.end()) net->allocs.erase(it);
    (void)hipFree(net->dn.raster);
    net->info.device_bytes -= std::max<uint64_t>(net->dn.rec_steps * net->dn.W, 1) * 8;
    net->dn.raster = nullptr;
  }
  net->dn.rec_first = rec_first;
  net->dn.rec_steps = rec_steps;
  if (rec_steps) {
    uint64_t* r = nullptr;
    int rc = dev_alloc(net, &r, rec_steps * net->dn.W);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemsetAsync(r, 0, rec_steps * net->dn.W * 8, ctx->stream));
    net->dn.raster = r;
  }
  net->graph_dirty = true;
  return SZSNN_OK;
}

int szsnn_read_raster(szsnn_net* net, uint64_t* host_out) {
  if (!net || !host_out) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_read_raster: bad argument");
  if (!net->dn.raster) return fail(net->ctx, SZSNN_ERR_STATE, "szsnn_read_raster: recording is off");
  szsnn_ctx* ctx = net->ctx;
  HIP_TRY(ctx, hipMemcpyAsync(host_out, net->dn.raster, net->dn.rec_steps * net->dn.W * 8, hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SZSNN_OK;
}

int szsnn_get_vars(szsnn_net* net, float* host_out) {
  if (!net || !host_out) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_get_vars: bad argument");
  szsnn_ctx* ctx = net->ctx;
  const size_t N = net->dn.N;
  for (uint32_t k = 0; k < net->info.n_vars; ++k)
    HIP_TRY(ctx, hipMemcpyAsync(host_out + k * N, net->dn.vars + (size_t)k * net->dn.npad, N * 4,
                                hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SZSNN_OK;
}

int szsnn_set_vars(szsnn_net* net, const float* host_in) {
  if (!net || !host_in) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_set_vars: bad argument");
  szsnn_ctx* ctx = net->ctx;
  const size_t N = net->dn.N;
  for (uint32_t k = 0; k < net->info.n_vars; ++k)
    HIP_TRY(ctx, hipMemcpyAsync(net->dn.vars + (size_t)k * net->dn.npad, host_in + k * N, N * 4,
                                hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SZSNN_OK;
}

int szsnn_get_counters(szsnn_net* net, szsnn_counters* out) {
  if (!net || !out) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_get_counters: bad argument");
  unsigned long long c[kCntSlots];
  int rc = read_counters(net, c);
  if (rc) return rc;
  out->cycle = net->t_host;
  out->events = c[kCntEvents];
  out->spikes = c[kCntSpikes];
  out->drive_events = c[kCntDrive];
  out->overflow = c[kCntOverflow];
  out->rows = c[kCntRows];
  out->deferred = c[kCntDeferred];
  out->late_deferred = c[kCntLate];
  return SZSNN_OK;
}

int szsnn_reset_counters(szsnn_net* net) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_reset_counters: net is NULL");
  HIP_TRY(net->ctx, hipMemsetAsync(net->dn.counters, 0, kCntWords * 8, net->ctx->stream));
  HIP_TRY(net->ctx, hipStreamSynchronize(net->ctx->stream));
  return SZSNN_OK;
}

int szsnn_time_kernels(szsnn_net* net, uint64_t n_steps, double* gather_ms, double* update_ms) {
  if (!net || !gather_ms || !update_ms) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_time_kernels: bad argument");
  if (int rc = check_usable(net, "szsnn_time_kernels")) return rc;
  szsnn_ctx* ctx = net->ctx;
  // four events per step, recorded by the kernels' dispatches: gather start
  // and end, update start and end
  std::vector<hipEvent_t> ev(4 * n_steps);
  for (auto& e : ev) HIP_TRY(ctx, hipEventCreate(&e));
  int rc = SZSNN_OK;
  for (uint64_t s = 0; s < n_steps && rc == SZSNN_OK; ++s) {
    hipEvent_t* e4 = &ev[4 * s];
    hipError_t e = launch_gather(net->dn, 0, net->tiling.gather_block, net->tiling.gather_grid, ctx->stream, e4[0], e4[1]);
    if (e == hipSuccess)
      e = launch_update(net->dn, 0, net->tiling.update_block, net->tiling.update_npl, ctx->stream, e4[2], e4[3]);
    if (e == hipSuccess) e = launch_advance(net->t_base, 1, ctx->stream);
    if (e != hipSuccess) rc = fail(ctx, SZSNN_ERR_HIP, std::string("szsnn_time_kernels: ") + hipGetErrorString(e));
    net->t_host += 1;
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  double g = 0, u = 0;
  if (rc == SZSNN_OK && e == hipSuccess) {
    for (uint64_t s = 0; s < n_steps && e == hipSuccess; ++s) {
      float a = 0, b = 0;
      e = hipEventElapsedTime(&a, ev[4 * s], ev[4 * s + 1]);
      if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[4 * s + 2], ev[4 * s + 3]);
      g += a;
      u += b;
    }
  }
  for (auto& x : ev) (void)hipEventDestroy(x);
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, SZSNN_ERR_HIP, std::string("szsnn_time_kernels: ") + hipGetErrorString(e));
  *gather_ms = g;
  *update_ms = u;
  return check_overflow(net);
}

int szsnn_test_kernel(szsnn_net* net, int32_t kernel, double* ms) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_test_kernel: net is NULL");
  if (int rc = check_usable(net, "szsnn_test_kernel")) return rc;
  szsnn_ctx* ctx = net->ctx;
  if (kernel != SZSNN_KERNEL_GATHER && kernel != SZSNN_KERNEL_UPDATE)
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_test_kernel: kernel must be SZSNN_KERNEL_GATHER or SZSNN_KERNEL_UPDATE");
  hipEvent_t e0, e1;
  HIP_TRY(ctx, hipEventCreate(&e0));
  hipError_t e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(ctx, SZSNN_ERR_HIP, std::string("szsnn_test_kernel: ") + hipGetErrorString(e));
  }
  // the events are recorded by the kernel's dispatch: its own duration
  e = kernel == SZSNN_KERNEL_GATHER
          ? launch_gather(net->dn, 0, net->tiling.gather_block, net->tiling.gather_grid, ctx->stream, e0, e1)
          : launch_update(net->dn, 0, net->tiling.update_block, net->tiling.update_npl, ctx->stream, e0, e1);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(ctx, SZSNN_ERR_HIP, std::string("szsnn_test_kernel: ") + hipGetErrorString(e));
  if (ms) *ms = t;
  return check_overflow(net);
}

int szsnn_set_tiling(szsnn_net* net, uint32_t update_block, uint32_t gather_block, uint32_t gather_grid) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_set_tiling: net is NULL");
  if (update_block && update_block != 64 && update_block != 128 && update_block != 256 && update_block != 512 &&
      update_block != 1024)
    return fail(net->ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling: update_block must be 64, 128, 256, 512 or 1024");
  if (gather_block && gather_block != 128 && gather_block != 256 && gather_block != 512 && gather_block != 1024)
    return fail(net->ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling: gather_block must be 128, 256, 512 or 1024");
  if (update_block) net->tiling.update_block = update_block;
  if (gather_block) net->tiling.gather_block = gather_block;
  net->tiling.gather_grid = gather_grid ? gather_grid : default_gather_grid(net, net->tiling.gather_block);
  apply_tiling(net);
  return SZSNN_OK;
}

int szsnn_set_tiling_ex(szsnn_net* net, const szsnn_tiling* t) {
  if (!net || !t) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_set_tiling_ex: bad argument");
  szsnn_ctx* ctx = net->ctx;
  auto pow2_in = [](uint32_t v, uint32_t lo, uint32_t hi) { return v >= lo && v <= hi && (v & (v - 1)) == 0; };
  if (!pow2_in(t->update_block, 64, 1024))
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling_ex: update_block must be 64, 128, 256, 512 or 1024");
  if (t->update_ids_per_lane != 1 && t->update_ids_per_lane != 2 && t->update_ids_per_lane != 4)
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling_ex: update_ids_per_lane must be 1, 2 or 4");
  if (!pow2_in(t->gather_block, 128, 1024))
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling_ex: gather_block must be 128, 256, 512 or 1024");
  const uint32_t stage = t->gather_stage ? t->gather_stage : net->stage_max;
  if (stage > net->stage_max) {
    std::ostringstream os;
    os << "szsnn_set_tiling_ex: gather_stage " << stage << " exceeds " << net->stage_max
       << " (the delays whose bitmaps fit the LDS of one gather block)";
    return fail(ctx, SZSNN_ERR_INVALID, os.str());
  }
  if (t->gather_batch != 2 && t->gather_batch != 4)
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling_ex: gather_batch must be 2 or 4");
  if (net->dn.pull_off &&
      gather_max_blocks_per_cu(pull_dyn_bytes(net->dn, stage), t->gather_block, stage, t->gather_batch) <= 0)
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling_ex: that gather block and stage cannot be resident");
  if (t->mode != SZSNN_MODE_GRAPH && t->mode != SZSNN_MODE_PERSISTENT)
    return fail(ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling_ex: unknown mode");
  if (t->mode == SZSNN_MODE_PERSISTENT) {
    if (t->persist_block != 256 && t->persist_block != 512)
      return fail(ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling_ex: persist_block must be 256 or 512");
    if (persistent_grid(net, t->persist_block, t->persist_grid) == 0)
      return fail(ctx, SZSNN_ERR_INVALID, "szsnn_set_tiling_ex: no co-resident occupancy for that persistent block");
  }
  Tiling& tl = net->tiling;
  tl.update_block = t->update_block;
  tl.update_npl = t->update_ids_per_lane;
  tl.gather_block = t->gather_block;
  tl.gather_grid = t->gather_grid ? t->gather_grid : default_gather_grid(net, t->gather_block);
  tl.gather_stage = stage;
  tl.gather_batch = t->gather_batch;
  tl.mode = t->mode;
  if (t->mode == SZSNN_MODE_PERSISTENT) {
    tl.persist_block = t->persist_block;
    tl.persist_grid = t->persist_grid;
  }
  apply_tiling(net);
  return SZSNN_OK;
}

int szsnn_get_tiling_ex(szsnn_net* net, szsnn_tiling* t) {
  if (!net || !t) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_get_tiling_ex: bad argument");
  const Tiling& tl = net->tiling;
  t->update_block = tl.update_block;
  t->update_ids_per_lane = tl.update_npl;
  t->gather_block = tl.gather_block;
  t->gather_grid = tl.gather_grid;
  t->gather_stage = tl.gather_stage;
  t->gather_stage_max = net->stage_max;
  t->gather_batch = tl.gather_batch;
  t->mode = tl.mode;
  t->persist_block = tl.persist_block;
  t->persist_grid = persistent_grid(net, tl.persist_block, tl.persist_grid);
  return SZSNN_OK;
}

int szsnn_set_gather_policy(szsnn_net* net, float dense_fraction) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_set_gather_policy: net is NULL");
  policy_changing(net);
  net->dense_fraction = dense_fraction;
  net->dn.defer_spikes = defer_threshold(net, dense_fraction);
  return SZSNN_OK;
}

int szsnn_set_heavy_push(szsnn_net* net, uint32_t min_entries) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_set_heavy_push: net is NULL");
  szsnn_ctx* ctx = net->ctx;
  if (!net->dn.heavy_cnt) {
    if (min_entries) return fail(ctx, SZSNN_ERR_INVALID, "szsnn_set_heavy_push: the network has no pull layout");
    return SZSNN_OK;
  }
  // the ranges the last step posted are pushed under the old threshold,
  // then both lists start empty
  HIP_TRY(ctx, launch_heavy(net->dn, net->tiling.gather_grid, ctx->stream));
  policy_changing(net);
  net->dn.heavy_min = min_entries;
  return SZSNN_OK;
}

int szsnn_get_heavy_push(szsnn_net* net, uint32_t* min_entries) {
  if (!net || !min_entries) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_get_heavy_push: bad argument");
  *min_entries = net->dn.heavy_min;
  return SZSNN_OK;
}

int szsnn_get_gather_policy(szsnn_net* net, float* dense_fraction, int32_t* has_pull) {
  if (!net || !dense_fraction || !has_pull)
    return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_get_gather_policy: bad argument");
  *dense_fraction = net->dense_fraction;
  *has_pull = net->dn.pull_off != nullptr;
  return SZSNN_OK;
}

int szsnn_set_mode(szsnn_net* net, int32_t mode, uint32_t block, uint32_t grid) {
  if (!net) return fail(nullptr, SZSNN_ERR_INVALID, "szsnn_set_mode: net is NULL");
  if (mode != SZSNN_MODE_GRAPH && mode != SZSNN_MODE_PERSISTENT)
    return fail(net->ctx, SZSNN_ERR_INVALID, "szsnn_set_mode: unknown mode");
  if (mode == SZSNN_MODE_PERSISTENT) {
    if (block == 0) block = net->tiling.persist_block;
    if (block != 256 && block != 512)
      return fail(net->ctx, SZSNN_ERR_INVALID, "szsnn_set_mode: persistent block must be 256 or 512");
    if (persistent_grid(net, block, grid) == 0)
      return fail(net->ctx, SZSNN_ERR_INVALID, "szsnn_set_mode: no co-resident occupancy for that block");
    net->tiling.persist_block = block;
    net->tiling.persist_grid = grid;
  }
  net->tiling.mode = mode;
  return SZSNN_OK;
}

int szsnn_get_mode(szsnn_net* net, int32_t* mode, uint32_t* block, uint32_t* grid) {
  if (!net || !mode || !block || !grid) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_get_mode: bad argument");
  *mode = net->tiling.mode;
  *block = net->tiling.persist_block;
  *grid = persistent_grid(net, net->tiling.persist_block, net->tiling.persist_grid);
  return SZSNN_OK;
}

int szsnn_get_tiling(szsnn_net* net, uint32_t* update_block, uint32_t* gather_block, uint32_t* gather_grid) {
  if (!net || !update_block || !gather_block || !gather_grid)
    return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_get_tiling: bad argument");
  *update_block = net->tiling.update_block;
  *gather_block = net->tiling.gather_block;
  *gather_grid = net->tiling.gather_grid;
  return SZSNN_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- autotuner
namespace {

struct Region {
  void* p;
  size_t bytes;
};

std::vector<Region> state_regions(szsnn_net* net) {
  const DevNet& n = net->dn;
  std::vector<Region> r = {
      {n.vars, (size_t)net->info.n_vars * n.npad * 4},
      {n.acc, (size_t)n.R * 2 * n.npad * 8},
      {n.dstamp, (size_t)n.R * 8},
      {n.ring_cnt, kRingCntWords(n.R) * 8},
      {n.spk_ring, kSpkRingWords * 8},
      {n.counters, kCntWords * 8},
      {net->t_base, 8},
      {n.hist, (size_t)n.R * n.W * 8},
  };
  if (n.src_last) r.push_back({n.src_last, (size_t)n.M * 4});
  if (n.acc_pull) r.push_back({n.acc_pull, (size_t)n.pull_groups * 2 * n.npad * 8});
  if (n.heavy_cnt) {
    r.push_back({n.heavy_cnt, 2 * 4});
    r.push_back({n.heavy_list, (size_t)2 * n.T * 8});
    r.push_back({n.hvy, (size_t)n.R * n.W * 8});
  }
  return r;
}

int snapshot(szsnn_net* net, bool restore) {
  szsnn_ctx* ctx = net->ctx;
  auto regs = state_regions(net);
  if (net->snap.empty()) {
    for (auto& r : regs) {
      void* p = nullptr;
      hipError_t e = hipMalloc(&p, r.bytes);
      if (e != hipSuccess) return fail(ctx, SZSNN_ERR_NOMEM, "autotune: snapshot allocation failed");
      net->snap.push_back({p, r.bytes});
    }
  }
  for (size_t i = 0; i < regs.size(); ++i) {
    if (restore)
      HIP_TRY(ctx, hipMemcpyAsync(regs[i].p, net->snap[i].first, regs[i].bytes, hipMemcpyDeviceToDevice, ctx->stream));
    else
      HIP_TRY(ctx, hipMemcpyAsync(net->snap[i].first, regs[i].p, regs[i].bytes, hipMemcpyDeviceToDevice, ctx->stream));
  }
  // a restored state starts with clean grid-barrier counters (a trial whose
  // barrier timed out may have left leaf counts behind)
  if (restore) HIP_TRY(ctx, hipMemsetAsync(net->dn.barrier, 0, kBarWords * 4, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return SZSNN_OK;
}

}  // namespace

extern "C" int szsnn_autotune(szsnn_net* net, uint32_t steps) {
  if (!net || steps == 0) return fail(net ? net->ctx : nullptr, SZSNN_ERR_INVALID, "szsnn_autotune: bad argument");
  if (int rc = check_usable(net, "szsnn_autotune")) return rc;
  szsnn_ctx* ctx = net->ctx;
  const uint64_t t0 = net->t_host;
  int rc = snapshot(net, false);
  if (rc) return rc;
  const int cus = std::max(ctx->cu_count, 1);
  const uint32_t ublocks[] = {64, 128, 256, 512, 1024};
  const uint32_t npls[] = {1, 2, 4};
  const uint32_t gblocks[] = {128, 256, 512, 1024};
  const uint32_t gmult[] = {1, 2, 4, 8, 16};
  Tiling best = net->tiling;
  float best_ms = 1e30f;
  hipEvent_t e0, e1;
  HIP_TRY(ctx, hipEventCreate(&e0));
  HIP_TRY(ctx, hipEventCreate(&e1));
  auto trial = [&](const Tiling& tl, float* ms) -> int {
    net->tiling = tl;
    apply_tiling(net);
    free_graph(net);
    int r = snapshot(net, true);
    net->t_host = t0;
    if (r) return r;
    r = run_steps(net, kGraphSteps);  // warm-up + graph build
    if (r) return r;
    HIP_TRY(ctx, hipEventRecord(e0, ctx->stream));
    r = run_steps(net, steps);
    if (r) return r;
    HIP_TRY(ctx, hipEventRecord(e1, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(e1));
    HIP_TRY(ctx, hipEventElapsedTime(ms, e0, e1));
    // a persistent trial whose grid barrier timed out is not a candidate
    unsigned long long c[kCntSlots];
    r = read_counters(net, c);
    if (r) return r;
    if (c[kCntOverflow] >= kBarTimeout) {
      HIP_TRY(ctx, hipMemsetAsync(net->dn.barrier, 0, kBarWords * 4, ctx->stream));
      *ms = 1e30f;
    }
    return SZSNN_OK;
  };
  auto consider = [&](const Tiling& tl) -> int {
    float ms = 0;
    int r = trial(tl, &ms);
    if (r == SZSNN_OK && ms < best_ms) {
      best_ms = ms;
      best = tl;
    }
    return r;
  };
  // stage 1: graph mode, update block x ids per lane at the current gather tiling
  for (uint32_t ub : ublocks) {
    for (uint32_t np : npls) {
      if (rc) break;
      Tiling tl = best;
      tl.mode = SZSNN_MODE_GRAPH;
      tl.update_block = ub;
      tl.update_npl = np;
      rc = consider(tl);
    }
  }
  // stage 2: graph mode, gather block x grid (grid in CU multiples) x
  // staging depth (one delay, or as many as fit)
  const uint32_t stages[] = {1, net->stage_max};
  const uint32_t batches[] = {2, 4};
  for (uint32_t st : stages) {
    for (uint32_t ba : batches) {
      for (uint32_t gb : gblocks) {
        if (net->dn.pull_off && gather_max_blocks_per_cu(pull_dyn_bytes(net->dn, st), gb, st, ba) <= 0) continue;
        for (uint32_t gm : gmult) {
          if (rc) break;
          Tiling tl = best;
          tl.mode = SZSNN_MODE_GRAPH;
          tl.gather_block = gb;
          tl.gather_grid = (uint32_t)cus * gm;
          tl.gather_stage = st;
          tl.gather_batch = ba;
          rc = consider(tl);
        }
      }
      if (!net->dn.pull_off) break;
    }
    if (net->stage_max == 1) break;
  }
  // stage 3: persistent mode, block x grid (CUs/4, CUs/2, CUs, co-resident max)
  const uint32_t pblocks[] = {256, 512};
  const uint32_t pdiv[] = {4, 2, 1};
  for (uint32_t pb : pblocks) {
    if (persistent_max_blocks_per_cu(net->dn, pb) <= 0) continue;
    for (int gi = 0; gi < 4; ++gi) {
      if (rc) break;
      Tiling tl = best;
      tl.mode = SZSNN_MODE_PERSISTENT;
      tl.persist_block = pb;
      tl.persist_grid = gi < 3 ? std::max<uint32_t>(1, (uint32_t)cus / pdiv[gi]) : 0;
      rc = consider(tl);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  net->tiling = best;
  apply_tiling(net);
  free_graph(net);
  int r2 = snapshot(net, true);
  net->t_host = t0;
  if (rc) return rc;
  return r2;
}
