// szsnn_adjoint.hip -- discrete adjoint of the IF_curr_exp step with the spike
// pattern held fixed (no reference counterpart: the reference has no
// gradients).  Used by the PyTorch autograd Function in szsnn_amd/adjoint.py.
//
// Hidden layer, one thread per (trial b, neuron j), state SoA [var][b][j]:
//   se_t = (1 - 1/tau_e) se_{t-1} + AE_t        (AE_t = W1+^T x_{t-1}, delay 1)
//   si_t = (1 - 1/tau_i) si_{t-1} + AI_t        (AI_t = W1-^T x_{t-1})
//   refr_t = fired_{t-1} <= tau_refrac; fired += 1
//   v'_t = refr_t ? v_{t-1} : v_{t-1} + I*(1/cm) + (v_rest - v_{t-1})*(1/tau_m)
//   spike_t: v_t = v_reset, fired = 1; else v_t = v'_t
//   h_t = v_t - v_rest drives the readout
// (the IF_curr_exp update of IF_curr_exp_update_kernel.cl:230-297 in this
// library's FP32 evaluation order).
//
// Adjoint, backwards over t with masks s_t (spike) and r_t (refractory):
//   lv_t   = g_t + Cv            (g_t = dL/dh_t from the readout)
//   lv'_t  = s_t ? 0 : lv_t
//   lse_t  = lv'_t * (r_t ? 0 : 1/cm) + Ce,  lsi_t likewise with Ci
//   Cv = lv'_t * (r_t ? 1 : 1 - 1/tau_m);  Ce = (1-1/tau_e) lse_t;  Ci = (1-1/tau_i) lsi_t
// The carries (Cv, Ce, Ci) cross segment boundaries through `carry`.
// dL/dW1[i][j] = sum_t x_{t-1}[i] (lse_t[j] or lsi_t[j] by the sign of
// W1[i][j]) comes from input_grad_kernel, which walks only the steps where
// input i spiked; the input currents AE, AI of all steps come from
// input_currents_kernel, which adds only the nonzero inputs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "../../include/szsnn_hip.h"

namespace szsnn {
void set_global_error(const char* msg);  // szsnn_runtime.hip: the message szsnn_last_error(NULL) returns
}

namespace {

constexpr int kMaxSeg = 64;
constexpr uint32_t kStepBatch = 8;      // recurrence steps whose loads are issued together
constexpr uint32_t kReadoutBatch = 16;  // the same for the readouts (fewer threads: more steps in
                                        // flight) and for the input bit matrix's rows

// params: [9][H] IF_curr_exp parameters per hidden neuron (per-step units)
struct HidParams {
  float v_rest, rc, rt, tau_refrac, ae, ai, i_offset, v_reset, v_thresh;
};

__device__ inline HidParams load_params(const float* p, uint32_t H, uint32_t j) {
  HidParams q;
  q.v_rest = p[0 * H + j];
  q.rc = 1.0f / p[1 * H + j];
  q.rt = 1.0f / p[2 * H + j];
  q.tau_refrac = p[3 * H + j];
  q.ae = 1.0f - 1.0f / p[4 * H + j];
  q.ai = 1.0f - 1.0f / p[5 * H + j];
  q.i_offset = p[6 * H + j];
  q.v_reset = p[7 * H + j];
  q.v_thresh = p[8 * H + j];
  return q;
}

// Forward over K steps from `state` ([4][B*H]: v, se, si, fired), writing
// h [K][B*H] and spike bits [K][ceil(B*H/64)] (u64 words).
__global__ void adj_forward_kernel(uint32_t BH, uint32_t H, uint32_t K, const float* __restrict__ params,
                                   float* __restrict__ state, const float* __restrict__ AE,
                                   const float* __restrict__ AI, float* __restrict__ h, uint64_t* __restrict__ bits) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t words = (BH + 63) / 64;
  const bool live = idx < BH;
  const uint32_t j = live ? idx % H : 0;
  HidParams q = load_params(params, H, j);
  float v = 0, se = 0, si = 0, fired = 0;
  if (live) {
    v = state[idx];
    se = state[BH + idx];
    si = state[2 * BH + idx];
    fired = state[3 * BH + idx];
  }
  // kStepBatch steps at a time: their inputs' loads in flight together (the
  // recurrence itself does not wait on memory)
  for (uint32_t t0 = 0; t0 < K; t0 += kStepBatch) {
    float ae[kStepBatch], ai[kStepBatch];
#pragma unroll
    for (uint32_t u = 0; u < kStepBatch; ++u) {
      const bool ok = live && t0 + u < K;
      ae[u] = ok ? AE[(size_t)(t0 + u) * BH + idx] : 0.0f;
      ai[u] = ok ? AI[(size_t)(t0 + u) * BH + idx] : 0.0f;
    }
#pragma unroll
    for (uint32_t u = 0; u < kStepBatch; ++u) {
      const uint32_t t = t0 + u;
      if (t >= K) break;  // uniform
      bool spike = false;
      if (live) {
        se = (q.ae * se) + ae[u];
        si = (q.ai * si) + ai[u];
        const float I = (se + si) + q.i_offset;
        const bool refr = fired <= q.tau_refrac;
        fired = fired + 1.0f;
        if (!refr) v = v + ((I * q.rc) + ((q.v_rest - v) * q.rt));
        spike = !refr && v >= q.v_thresh;
        if (spike) {
          v = q.v_reset;
          fired = 1.0f;
        }
        h[(size_t)t * BH + idx] = v - q.v_rest;
      }
      const uint64_t m = __ballot(spike);
      if (__lane_id() == 0 && (idx >> 6) < words) bits[(size_t)t * words + (idx >> 6)] = m;
    }
  }
  if (live) {
    state[idx] = v;
    state[BH + idx] = se;
    state[2 * BH + idx] = si;
    state[3 * BH + idx] = fired;
  }
}

// Backward over one segment of K steps.  ckpt: state at the segment start
// ([4][B*H], only `fired` is needed: the masks follow from it and the fixed
// spike bits).  g: dL/dh [K][B*H].  carry [3][B*H] in/out (Cv, Ce, Ci).
// LE/LI out: dL/dse_t, dL/dsi_t [K][B*H].
__global__ void adj_backward_kernel(uint32_t BH, uint32_t H, uint32_t K, const float* __restrict__ params,
                                    const float* __restrict__ ckpt, const uint64_t* __restrict__ bits,
                                    const float* __restrict__ g, float* __restrict__ carry,
                                    float* __restrict__ LE, float* __restrict__ LI) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= BH) return;
  const uint32_t words = (BH + 63) / 64;
  const uint32_t j = idx % H;
  HidParams q = load_params(params, H, j);
  // recompute the refractory mask of each step from the checkpoint (the
  // spike words of kStepBatch steps loaded together)
  uint64_t rmask = 0, smask = 0;
  float fired = ckpt[3 * BH + idx];
  for (uint32_t t0 = 0; t0 < K; t0 += kStepBatch) {
    uint64_t w[kStepBatch];
#pragma unroll
    for (uint32_t u = 0; u < kStepBatch; ++u)
      w[u] = t0 + u < K ? bits[(size_t)(t0 + u) * words + (idx >> 6)] : 0ull;
#pragma unroll
    for (uint32_t u = 0; u < kStepBatch; ++u) {
      const uint32_t t = t0 + u;
      if (t >= K) break;  // uniform
      const bool refr = fired <= q.tau_refrac;
      fired = fired + 1.0f;
      const bool spike = (w[u] >> (idx & 63)) & 1ull;
      if (spike) fired = 1.0f;
      rmask |= (uint64_t)refr << t;
      smask |= (uint64_t)spike << t;
    }
  }
  float cv = carry[idx], ce = carry[BH + idx], ci = carry[2 * BH + idx];
  // backwards, kStepBatch steps at a time: their g loads in flight together
  for (int t1 = (int)K - 1; t1 >= 0; t1 -= (int)kStepBatch) {
    float gv[kStepBatch];
#pragma unroll
    for (int u = 0; u < (int)kStepBatch; ++u) gv[u] = t1 - u >= 0 ? g[(size_t)(t1 - u) * BH + idx] : 0.0f;
#pragma unroll
    for (int u = 0; u < (int)kStepBatch; ++u) {
      const int t = t1 - u;
      if (t < 0) break;  // uniform
      const bool refr = (rmask >> t) & 1ull;
      const bool spike = (smask >> t) & 1ull;
      const float lv = gv[u] + cv;
      const float lvp = spike ? 0.0f : lv;
      const float dse = refr ? 0.0f : q.rc;
      const float lse = lvp * dse + ce;
      const float lsi = lvp * dse + ci;
      LE[(size_t)t * BH + idx] = lse;
      LI[(size_t)t * BH + idx] = lsi;
      cv = lvp * (refr ? 1.0f : 1.0f - q.rt);
      ce = q.ae * lse;
      ci = q.ai * lsi;
    }
  }
  carry[idx] = cv;
  carry[BH + idx] = ce;
  carry[2 * BH + idx] = ci;
}

// Leaky readout y_t = alpha y_{t-1} + drive_t, [T][B*R]; y0 the state before step 0.
__global__ void readout_forward_kernel(uint32_t BR, uint32_t T, float alpha, const float* __restrict__ y0,
                                       const float* __restrict__ drive, float* __restrict__ y) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= BR) return;
  float acc = y0 ? y0[idx] : 0.0f;
  for (uint32_t t0 = 0; t0 < T; t0 += kReadoutBatch) {  // the batch's loads in flight together
    float d[kReadoutBatch];
#pragma unroll
    for (uint32_t u = 0; u < kReadoutBatch; ++u) d[u] = t0 + u < T ? drive[(size_t)(t0 + u) * BR + idx] : 0.0f;
#pragma unroll
    for (uint32_t u = 0; u < kReadoutBatch; ++u) {
      if (t0 + u >= T) break;
      acc = (alpha * acc) + d[u];
      y[(size_t)(t0 + u) * BR + idx] = acc;
    }
  }
}

// lam_t = scale (y_t - target_t) + alpha lam_{t+1}; scale = 2 / (T*B*R) for the mean
// squared error times the incoming gradient of the loss.
__global__ void readout_backward_kernel(uint32_t BR, uint32_t T, float alpha, float scale,
                                        const float* __restrict__ y, const float* __restrict__ target,
                                        float* __restrict__ lam) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= BR) return;
  float acc = 0.0f;
  for (int t1 = (int)T - 1; t1 >= 0; t1 -= (int)kReadoutBatch) {  // the batch's loads in flight together
    float yv[kReadoutBatch], tv[kReadoutBatch];
#pragma unroll
    for (int u = 0; u < (int)kReadoutBatch; ++u) {
      const bool ok = t1 - u >= 0;
      yv[u] = ok ? y[(size_t)(t1 - u) * BR + idx] : 0.0f;
      tv[u] = ok ? target[(size_t)(t1 - u) * BR + idx] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < (int)kReadoutBatch; ++u) {
      if (t1 - u < 0) break;
      acc = (scale * (yv[u] - tv[u])) + (alpha * acc);
      lam[(size_t)(t1 - u) * BR + idx] = acc;
    }
  }
}

// Input currents of one input row r (a (step, trial) pair): AE[r][j] =
// sum_i X[r][i] * max(W1[i][j], 0), AI[r][j] likewise with min(W1[i][j], 0),
// over the nonzero X[r][i] only (input spikes: a few percent of the inputs),
// in ascending i.  One block per row; the inputs are scanned in chunks of
// the block size, each chunk's nonzeros listed in LDS in order, then every
// thread adds them into its kCols columns (strided by the block size, so
// each W1 row is read coalesced).
constexpr int kInBlock = 256;
constexpr int kCols = 8;

__global__ void __launch_bounds__(kInBlock) input_currents_kernel(uint32_t n_in, uint32_t H, uint32_t col0,
                                                                  const float* __restrict__ X,
                                                                  const float* __restrict__ W1, float* __restrict__ AE,
                                                                  float* __restrict__ AI) {
  __shared__ uint32_t s_idx[kInBlock];
  __shared__ float s_val[kInBlock];
  __shared__ uint32_t s_wave[kInBlock / 64];
  const uint32_t r = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const float* x = X + (size_t)r * n_in;
  float e[kCols], in[kCols];
#pragma unroll
  for (int c = 0; c < kCols; ++c) e[c] = in[c] = 0.0f;
  for (uint32_t base = 0; base < n_in; base += kInBlock) {
    const uint32_t i = base + tid;
    const float xv = i < n_in ? x[i] : 0.0f;
    const uint64_t m = __ballot(xv != 0.0f);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t pos = 0, n = 0;
    for (uint32_t w = 0; w < kInBlock / 64; ++w) {
      pos += w < wave ? s_wave[w] : 0u;
      n += s_wave[w];
    }
    if (xv != 0.0f) {
      pos += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      s_idx[pos] = i;
      s_val[pos] = xv;
    }
    __syncthreads();
    for (uint32_t q = 0; q < n; ++q) {
      const float* wr = W1 + (size_t)s_idx[q] * H;
      const float xq = s_val[q];
#pragma unroll
      for (int c = 0; c < kCols; ++c) {
        const uint32_t j = col0 + c * kInBlock + tid;
        const float w = j < H ? wr[j] : 0.0f;
        e[c] += xq * (w > 0.0f ? w : 0.0f);
        in[c] += xq * (w < 0.0f ? w : 0.0f);
      }
    }
    __syncthreads();  // the list is rewritten by the next chunk
  }
#pragma unroll
  for (int c = 0; c < kCols; ++c) {
    const uint32_t j = col0 + c * kInBlock + tid;
    if (j < H) {
      AE[(size_t)r * H + j] = e[c];
      AI[(size_t)r * H + j] = in[c];
    }
  }
}

// Input spikes as a bit matrix: bit k of XT[i][w] is X[64 w + k][i] != 0
// (rows x n_in row-major X; XT [n_in][ceil(rows / 64)]), and one word per
// tile of 64 inputs x 64 rows, XF[i / 64][w]: nonzero when a nonzero X of the
// tile is not exactly 1 (so the gradient reads X there, and takes 1
// elsewhere).  Thread (i, w) reads its 64 rows of column i: each of its loads
// is one contiguous 256-byte piece of a row across the wave (the wave's 64
// inputs).  Every word of XT and XF is written.
__global__ void __launch_bounds__(64) input_bits_kernel(uint32_t rows, uint32_t n_in, uint32_t words,
                                                        const float* __restrict__ X, uint64_t* __restrict__ XT,
                                                        uint64_t* __restrict__ XF) {
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  const uint32_t w = blockIdx.y;
  const uint32_t r0 = w * 64;
  const uint32_t nr = i < n_in ? min(64u, rows - r0) : 0u;
  uint64_t m = 0;
  bool other = false;
  for (uint32_t k0 = 0; k0 < nr; k0 += kReadoutBatch) {  // the batch's loads in flight together
    float x[kReadoutBatch];
#pragma unroll
    for (uint32_t u = 0; u < kReadoutBatch; ++u) x[u] = k0 + u < nr ? X[(size_t)(r0 + k0 + u) * n_in + i] : 0.0f;
#pragma unroll
    for (uint32_t u = 0; u < kReadoutBatch; ++u) {
      m |= (uint64_t)(x[u] != 0.0f) << (k0 + u);
      other = other || (x[u] != 0.0f && x[u] != 1.0f);
    }
  }
  if (i < n_in) XT[(size_t)i * words + w] = m;
  const uint64_t tile = __ballot(other);
  if (threadIdx.x == 0) XF[(size_t)blockIdx.x * words + w] = tile;
}

// Lane j's value of a 64-bit register (j uniform).
__device__ inline uint64_t readlane64(uint64_t x, uint32_t j) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, (int)j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), (int)j);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

// dW1[i][j] = sum over rows r with X[r][i] != 0, ascending, of X[r][i] *
// (W1[i][j] >= 0 ? LE[r][j] : LI[r][j]): the gradient of the sign-split
// input currents (the E current takes the nonnegative weights).  The work is
// split into chunks of 64 columns, one wave per (input, chunk), lane = column,
// walking the input's rows (from XT) kGradRows at a time so that their loads
// are in flight together; both adjoints of a column are loaded and one is
// selected in registers.  Each input's rows are about 2 % of the rows, and an
// LE or LI row is read by every input that spiked in it (about 20 for
// B = 64, T = 200, 1000 inputs), so where the repeats are served from
// decides the cost.  The grid is chunk-major per XCD (block b runs on XCD
// b % 8, which takes chunks x, x + 8, ... in turn, each for all inputs): the
// blocks of an XCD walk one chunk of LE and LI together and its L2 serves
// part of the repeats (for the sizes above 2.9 GB per call come from beyond
// the L2, against 4.4 GB with one block per input over all columns).  The
// per-column sums add the same terms in the same order as one thread per
// (input, column) walking all its rows would.
constexpr int kGradBlock = 256;
constexpr int kGradWaves = kGradBlock / 64;  // inputs per block
constexpr int kGradRows = 8;
constexpr uint32_t kXcds = 8;

__global__ void __launch_bounds__(kGradBlock) input_grad_kernel(uint32_t n_in, uint32_t H, uint32_t words,
                                                                uint32_t chunks, const float* __restrict__ X,
                                                                const float* __restrict__ W1,
                                                                const float* __restrict__ LE,
                                                                const float* __restrict__ LI,
                                                                const uint64_t* __restrict__ XT,
                                                                const uint64_t* __restrict__ XF,
                                                                float* __restrict__ dW1) {
  const uint32_t groups = (n_in + kGradWaves - 1) / kGradWaves;
  const uint32_t xcd = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
  const uint32_t chunk = (k / groups) * kXcds + xcd;
  if (chunk >= chunks) return;  // no barrier in this kernel
  const uint32_t i = (k % groups) * kGradWaves + (threadIdx.x >> 6);
  if (i >= n_in) return;
  const uint32_t j0 = chunk * 64 + __lane_id();
  const uint32_t j = min(j0, H - 1);  // past H: a valid column, never stored
  const bool pos = W1[(size_t)i * H + j] >= 0.0f;
  float acc = 0.0f;
  const uint64_t* xt = XT + (size_t)i * words;
  const uint64_t* xf = XF + (size_t)(i / 64) * words;
  // The input's bit words are read 64 at a time, one per lane, and walked
  // from registers (the wave serves one input, so the walk is uniform):
  // lw/lf the lane's word and tile flag of the round starting at word w0,
  // nz the round's words not yet taken that hold rows.
  const uint32_t lane = __lane_id();
  uint32_t w0 = 0;
  uint64_t lw = 0, lf = 0, nz = 0;
  auto load_round = [&]() {
    const uint32_t wl = w0 + lane;
    lw = wl < words ? xt[wl] : 0ull;
    lf = wl < words ? xf[wl] : 0ull;
    nz = __ballot(lw != 0ull);
  };
  load_round();
  uint64_t m = 0;    // rows left in the current word (bit k: row 64 w + k)
  uint32_t wc = 0;   // the current word
  bool fc = false;   // its tile holds values other than 1
  for (;;) {
    // the next kGradRows rows of input i (r = 0xffffffff: none), and whether
    // each one's X has to be read (else it is 1)
    uint32_t r[kGradRows];
    bool rd[kGradRows];
#pragma unroll
    for (int q = 0; q < kGradRows; ++q) {
      while (!m) {
        if (!nz) {
          if (w0 + 64 >= words) break;
          w0 += 64;
          load_round();
          continue;
        }
        const uint32_t l = (uint32_t)__ffsll((unsigned long long)nz) - 1;
        nz &= nz - 1;
        m = readlane64(lw, l);
        fc = readlane64(lf, l) != 0ull;
        wc = w0 + l;
      }
      rd[q] = fc;
      if (m) {
        r[q] = wc * 64 + (uint32_t)__ffsll((unsigned long long)m) - 1;
        m &= m - 1;
      } else {
        r[q] = 0xffffffffu;
      }
    }
    if (r[0] == 0xffffffffu) break;
    // their loads, all in flight together, then the adds in row order
    float x[kGradRows], v[kGradRows];
#pragma unroll
    for (int q = 0; q < kGradRows; ++q) {
      const bool ok = r[q] != 0xffffffffu;
      const uint32_t rr = ok ? r[q] : r[0];
      x[q] = !ok ? 0.0f : rd[q] ? X[(size_t)rr * n_in + i] : 1.0f;
      const float e = LE[(size_t)rr * H + j], in = LI[(size_t)rr * H + j];
      v[q] = pos ? e : in;
    }
#pragma unroll
    for (int q = 0; q < kGradRows; ++q) acc += x[q] * v[q];
  }
  if (j0 < H) dW1[(size_t)i * H + j0] = acc;
}

// Bytes of the input gradient's scratch for `rows` rows of n_in inputs: the
// input bit matrix and the tiles' flags.
size_t input_grad_bytes(uint32_t rows, uint32_t n_in) {
  return ((size_t)n_in + (n_in + 63) / 64) * ((rows + 63) / 64) * 8;
}

int launch_input_grad(hipStream_t st, uint32_t rows, uint32_t n_in, uint32_t H, const float* X, const float* W1,
                      const float* LE, const float* LI, float* dW1, uint64_t* XT) {
  const uint32_t words = (rows + 63) / 64;
  uint64_t* XF = XT + (size_t)n_in * words;
  input_bits_kernel<<<dim3((n_in + 63) / 64, words), 64, 0, st>>>(rows, n_in, words, X, XT, XF);
  hipError_t e = hipGetLastError();
  const uint32_t chunks = (H + 63) / 64;
  const uint32_t groups = (n_in + kGradWaves - 1) / kGradWaves;
  const uint32_t blocks = kXcds * ((chunks + kXcds - 1) / kXcds) * groups;
  if (e == hipSuccess)
    input_grad_kernel<<<blocks, kGradBlock, 0, st>>>(n_in, H, words, chunks, X, W1, LE, LI, XT, XF, dW1);
  if (e == hipSuccess) e = hipGetLastError();
  return e == hipSuccess ? SZSNN_OK : SZSNN_ERR_HIP;
}

// Scratch of szsnn_adj_input_grad_noalloc: the input bit matrix and the tile
// words, in a
// file-scope device array (each device holds its own copy at its own
// address, looked up once per device with hipGetSymbolAddress).
constexpr size_t kScratchWords = (size_t)2 << 20;  // 16 MiB
static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
uint64_t* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

int launch_check(hipError_t e) { return e == hipSuccess ? SZSNN_OK : SZSNN_ERR_HIP; }

// The input bit matrix's kernel takes one grid row per 64 rows: at most
// 65535 of them.
constexpr uint32_t kMaxGradRows = 65535u * 64u;
bool input_grad_rows_ok(uint32_t rows, const char* who) {
  if (rows <= kMaxGradRows) return true;
  szsnn::set_global_error((std::string(who) + ": rows exceeds 65535 * 64").c_str());
  return false;
}

}  // namespace

extern "C" {

int szsnn_adj_forward_segment(void* stream, uint32_t B, uint32_t H, uint32_t K, const float* params, float* state,
                              const float* AE, const float* AI, float* h, uint64_t* spike_bits) {
  if (!B || !H || !K || !params || !state || !AE || !AI || !h || !spike_bits) return SZSNN_ERR_INVALID;
  const uint32_t BH = B * H;
  const uint32_t blocks = (BH + 255) / 256;
  adj_forward_kernel<<<blocks, 256, 0, (hipStream_t)stream>>>(BH, H, K, params, state, AE, AI, h, spike_bits);
  return launch_check(hipGetLastError());
}

int szsnn_adj_backward_segment(void* stream, uint32_t B, uint32_t H, uint32_t K, const float* params,
                               const float* ckpt, const uint64_t* spike_bits, const float* g, float* carry,
                               float* LE, float* LI) {
  if (!B || !H || !K || K > kMaxSeg || !params || !ckpt || !spike_bits || !g || !carry || !LE || !LI)
    return SZSNN_ERR_INVALID;
  const uint32_t BH = B * H;
  const uint32_t blocks = (BH + 255) / 256;
  adj_backward_kernel<<<blocks, 256, 0, (hipStream_t)stream>>>(BH, H, K, params, ckpt, spike_bits, g, carry, LE, LI);
  return launch_check(hipGetLastError());
}

int szsnn_adj_input_currents(void* stream, uint32_t rows, uint32_t n_in, uint32_t H, const float* X, const float* W1,
                             float* AE, float* AI) {
  if (!rows || !n_in || !H || !X || !W1 || !AE || !AI) return SZSNN_ERR_INVALID;
  for (uint32_t col0 = 0; col0 < H; col0 += kCols * kInBlock) {
    input_currents_kernel<<<rows, kInBlock, 0, (hipStream_t)stream>>>(n_in, H, col0, X, W1, AE, AI);
    const int rc = launch_check(hipGetLastError());
    if (rc != SZSNN_OK) return rc;
  }
  return SZSNN_OK;
}

size_t szsnn_adj_input_grad_workspace(uint32_t rows, uint32_t n_in) { return input_grad_bytes(rows, n_in); }

int szsnn_adj_input_grad(void* stream, uint32_t rows, uint32_t n_in, uint32_t H, const float* X, const float* W1,
                         const float* LE, const float* LI, float* dW1, void* workspace, size_t workspace_bytes) {
  if (!rows || !n_in || !H || !X || !W1 || !LE || !LI || !dW1 || !workspace) return SZSNN_ERR_INVALID;
  if (!input_grad_rows_ok(rows, "szsnn_adj_input_grad")) return SZSNN_ERR_INVALID;
  if (workspace_bytes < input_grad_bytes(rows, n_in)) return SZSNN_ERR_INVALID;
  return launch_input_grad((hipStream_t)stream, rows, n_in, H, X, W1, LE, LI, dW1, (uint64_t*)workspace);
}

int szsnn_adj_input_grad_noalloc(void* stream, uint32_t rows, uint32_t n_in, uint32_t H, const float* X,
                                 const float* W1, const float* LE, const float* LI, float* dW1) {
  if (!rows || !n_in || !H || !X || !W1 || !LE || !LI || !dW1) return SZSNN_ERR_INVALID;
  if (!input_grad_rows_ok(rows, "szsnn_adj_input_grad_noalloc")) return SZSNN_ERR_INVALID;
  if (input_grad_bytes(rows, n_in) > kScratchWords * 8) {
    szsnn::set_global_error(
        "szsnn_adj_input_grad_noalloc: (n_in + ceil(n_in / 64)) * ceil(rows / 64) exceeds the 2^21 words (16 MiB) "
        "of its scratch; "
        "use szsnn_adj_input_grad with a workspace");
    return SZSNN_ERR_INVALID;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return SZSNN_ERR_HIP;
  // the device's copy of the scratch array, looked up once with the device current
  std::call_once(g_scratch_once[dev], [dev] {
    g_scratch_err[dev] = hipGetSymbolAddress(reinterpret_cast<void**>(&g_scratch_ptr[dev]), HIP_SYMBOL(scratch));
  });
  if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev]) return SZSNN_ERR_HIP;
  return launch_input_grad((hipStream_t)stream, rows, n_in, H, X, W1, LE, LI, dW1, g_scratch_ptr[dev]);  // the device copy
}

int szsnn_readout_forward(void* stream, uint32_t BR, uint32_t T, float alpha, const float* y0, const float* drive,
                          float* y) {
  if (!BR || !T || !drive || !y) return SZSNN_ERR_INVALID;
  readout_forward_kernel<<<(BR + 255) / 256, 256, 0, (hipStream_t)stream>>>(BR, T, alpha, y0, drive, y);
  return launch_check(hipGetLastError());
}

int szsnn_readout_backward(void* stream, uint32_t BR, uint32_t T, float alpha, float scale, const float* y,
                           const float* target, float* lam) {
  if (!BR || !T || !y || !target || !lam) return SZSNN_ERR_INVALID;
  readout_backward_kernel<<<(BR + 255) / 256, 256, 0, (hipStream_t)stream>>>(BR, T, alpha, scale, y, target, lam);
  return launch_check(hipGetLastError());
}

}  // extern "C"
