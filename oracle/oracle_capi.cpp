// oracle_capi.cpp -- C entry points over szsnn_oracle.hpp for the Python tests.
//
// TEST INFRASTRUCTURE ONLY (see szsnn_oracle.hpp).  Built by oracle/Makefile
// into oracle/libszsnn_oracle.so with -ffp-contract=off.
#include <chrono>
#include <cstdio>
#include <exception>
#include <memory>

#include "szsnn_oracle.hpp"

using namespace szsnn_oracle;

namespace {
thread_local std::string g_err;

struct Handle {
  int precision;
  std::unique_ptr<Engine<float>> f;
  std::unique_ptr<Engine<double>> d;
};

template <typename E>
int run_engine(E& e, uint64_t steps, const uint64_t* forced_t, const uint32_t* forced_id,
               uint64_t n_forced, uint64_t* raster, double* v_trace, uint32_t n_neurons, int threads = 1) {
  uint64_t fpos = 0;
  // skip schedule entries before the current cycle
  while (fpos < n_forced && forced_t[fpos] < e.cycle()) ++fpos;
  for (uint64_t s = 0; s < steps; ++s) {
    const uint64_t t = e.cycle();
    const uint64_t f0 = fpos;
    while (fpos < n_forced && forced_t[fpos] == t) ++fpos;
    uint64_t* row = raster ? raster + s * e.words() : nullptr;
    double* vrow = v_trace ? v_trace + s * (uint64_t)n_neurons : nullptr;
    if (threads > 1)
      e.step_parallel(threads, forced_id + f0, fpos - f0, row, vrow);
    else
      e.step(forced_id + f0, fpos - f0, row, vrow);
  }
  return 0;
}
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

void* oracle_create(const NetDesc* d, int precision) {
  try {
    auto h = new Handle();
    h->precision = precision;
    if (precision == 32)
      h->f.reset(new Engine<float>(*d));
    else
      h->d.reset(new Engine<double>(*d));
    return h;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return nullptr;
  }
}

void oracle_destroy(void* p) { delete static_cast<Handle*>(p); }

// forced schedule sorted by step: (forced_t[i], forced_id[i]).
// raster: steps * words u64 (or null); v_trace: steps * n_neurons doubles (or null).
int oracle_run(void* p, uint64_t steps, const uint64_t* forced_t, const uint32_t* forced_id,
               uint64_t n_forced, uint64_t* raster, double* v_trace, uint32_t n_neurons) {
  auto h = static_cast<Handle*>(p);
  try {
    if (h->f) return run_engine(*h->f, steps, forced_t, forced_id, n_forced, raster, v_trace, n_neurons);
    return run_engine(*h->d, steps, forced_t, forced_id, n_forced, raster, v_trace, n_neurons);
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return -1;
  }
}

// oracle_run with the FP32 engine's parallel step on `threads` OpenMP
// threads (the same results bit for bit); -1 for the FP64 engine.
int oracle_run_parallel(void* p, uint64_t steps, const uint64_t* forced_t, const uint32_t* forced_id,
                        uint64_t n_forced, uint64_t* raster, double* v_trace, uint32_t n_neurons, int threads) {
  auto h = static_cast<Handle*>(p);
  if (!h->f || threads < 1) {
    g_err = "oracle_run_parallel: FP32 engine only";
    return -1;
  }
  try {
    return run_engine(*h->f, steps, forced_t, forced_id, n_forced, raster, v_trace, n_neurons, threads);
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return -1;
  }
}

// Wall time (seconds, steady_clock) of `steps` steps with no forced input.
double oracle_time_steps(void* p, uint64_t steps) {
  auto h = static_cast<Handle*>(p);
  const auto t0 = std::chrono::steady_clock::now();
  if (h->f)
    run_engine(*h->f, steps, nullptr, nullptr, 0, nullptr, nullptr, 0);
  else
    run_engine(*h->d, steps, nullptr, nullptr, 0, nullptr, nullptr, 0);
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// Wall time (seconds) of `steps` steps of the FP32 engine on `threads`
// OpenMP threads (step_parallel); -1 for the FP64 engine.
double oracle_time_steps_parallel(void* p, uint64_t steps, int threads) {
  auto h = static_cast<Handle*>(p);
  if (!h->f || threads < 1) return -1.0;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint64_t s = 0; s < steps; ++s) h->f->step_parallel(threads);
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

uint64_t oracle_words(void* p) {
  auto h = static_cast<Handle*>(p);
  return h->f ? h->f->words() : h->d->words();
}
uint64_t oracle_events(void* p) {
  auto h = static_cast<Handle*>(p);
  return h->f ? h->f->events() : h->d->events();
}
uint64_t oracle_drive_events(void* p) {
  auto h = static_cast<Handle*>(p);
  return h->f ? h->f->drive_events() : h->d->drive_events();
}

// vars: n_vars * n_neurons doubles (exact images of the engine's scalars).
void oracle_get_vars(void* p, double* out) {
  auto h = static_cast<Handle*>(p);
  if (h->f) {
    const auto& v = h->f->vars();
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
  } else {
    const auto& v = h->d->vars();
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
  }
}

// Single-neuron update driven by given per-step inputs (known-answer tests).
// vars in/out (n_vars doubles, rounded to the precision first); trace:
// steps * n_vars; spikes: steps bytes.
int oracle_neuron_trace(int model, int precision, const double* params, double* vars,
                        const double* accE, const double* accI, const double* iinj,
                        const uint8_t* forced, uint32_t steps, double* trace, uint8_t* spikes) {
  const int np = n_params(model), nv = n_vars(model);
  if (np < 0) return -1;
  if (precision == 32) {
    float pf[16], xf[8];
    for (int k = 0; k < np; ++k) pf[k] = (float)params[k];
    for (int k = 0; k < nv; ++k) xf[k] = (float)vars[k];
    for (uint32_t s = 0; s < steps; ++s) {
      spikes[s] = update_neuron<float>(model, pf, xf, (float)accE[s], (float)accI[s],
                                       iinj ? (float)iinj[s] : 0.0f, forced && forced[s]);
      for (int k = 0; k < nv; ++k) trace[(size_t)s * nv + k] = xf[k];
    }
    for (int k = 0; k < nv; ++k) vars[k] = xf[k];
  } else {
    double pd[16], xd[8];
    for (int k = 0; k < np; ++k) pd[k] = params[k];
    for (int k = 0; k < nv; ++k) xd[k] = vars[k];
    for (uint32_t s = 0; s < steps; ++s) {
      spikes[s] = update_neuron<double>(model, pd, xd, accE[s], accI[s], iinj ? iinj[s] : 0.0,
                                        forced && forced[s]);
      for (int k = 0; k < nv; ++k) trace[(size_t)s * nv + k] = xd[k];
    }
    for (int k = 0; k < nv; ++k) vars[k] = xd[k];
  }
  return 0;
}

// Random-drive primitives, exposed for their own known-answer tests.
uint64_t oracle_splitmix_at(uint64_t seed, uint64_t idx) { return splitmix_at(seed, idx); }
uint64_t oracle_draw(uint64_t key, uint64_t t, unsigned s) { return draw(key, t, s); }
int32_t oracle_gauss_int(uint64_t key, uint64_t t) { return gauss_int(key, t); }
float oracle_gauss_f32(uint64_t key, uint64_t t) { return gauss<float>(key, t); }
double oracle_gauss_f64(uint64_t key, uint64_t t) { return gauss<double>(key, t); }

}  // extern "C"
