// server.hip — serving form: continuous batching of reconstruction jobs.
//
// A worker thread owns the device. Jobs are queued by fpmhip_server_submit,
// admitted while tile slots and the HBM budget allow, initialised on the
// device, and then advanced together with every other resident tile by one
// launch of the fused sweep kernel per tick. Within a tick every tile runs a
// quota of sweeps inversely proportional to its per-sweep work (LEDs x the
// measured time of one LED update at its size), so that a small tile is not
// held to the pace of the largest one resident. A job that reaches its iteration count is finished (result
// transform, download) and leaves. Each tick is planned before it is
// launched; the jobs it completes are known then, so queued jobs are
// selected into the room they leave and admitted (uploaded, initialised)
// on streams of their own while the tick runs, together with the downloads
// of the jobs the previous tick finished, and join the next tick. Job
// buffers come from a device memory pool of the server, recycled between
// ticks, so that the frees of finished jobs and the allocations of new ones
// do not call hipFree / hipMalloc between sweep launches.
// Tiles are independent workgroups and sweeps are sequential within a tile,
// so a job's result does not depend on which other jobs share its ticks.
#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "fused.h"

using namespace fpmhip;

namespace {

template <class F>
int guarded(F&& body) {
  clear_error();
  try {
    body();
    return FPMHIP_OK;
  } catch (const Failure& f) {
    return f.code;
  } catch (const std::bad_alloc&) {
    return capture_bad_alloc();
  }
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum JobState { kQueued = 0, kResident = 1, kDone = 2, kFailed = 3 };

struct Job {
  uint64_t id = 0;
  // host copies of the inputs
  std::vector<uint64_t> crop, row_ptr, msrc;
  std::vector<double> offset, pupil, gain, bg, kv, mw;
  std::vector<uint8_t> support;
  std::vector<uint16_t> counts;
  std::vector<double> frames, weights;
  fpmhip_model_desc desc{};
  fpmhip_algorithm alg{};
  uint64_t iterations = 0, done_iterations = 0;
  uint64_t bytes = 0;
  uint32_t wgs = 1;        // workgroups of its tile in a tick (a cooperative group for large tiles)
  bool finishing = false;  // iterations done, results not yet downloaded
  // device objects while resident
  fpmhip_model* model = nullptr;
  fpmhip_measurements* meas = nullptr;
  fpmhip_session* sess = nullptr;
  DeviceBuffer scratch;  // [16 B per-tile state][trace][frame losses]
  // results
  int state = kQueued;
  int error_code = 0;
  std::string error;
  double t_submit = 0, t_admit = 0, t_done = 0;
  std::vector<double> object, spectrum, trace;

  void release_device() {
    if (sess) fpmhip_session_destroy(sess);
    if (meas) fpmhip_measurements_destroy(meas);
    if (model) fpmhip_model_destroy(model);
    sess = nullptr;
    meas = nullptr;
    model = nullptr;
    scratch.release();
  }
};

}  // namespace

struct fpmhip_server {
  static constexpr int kAdmitThreads = 4;
  // Device blocks of the jobs' buffers (declared first: it outlives every
  // job). Blocks given back become reusable at recycle_pool().
  std::unique_ptr<DevicePool> pool;
  fpmhip_context own;                   // private stream of the ticks: never the caller's stream
  fpmhip_context stage[kAdmitThreads];  // streams of the admission threads, while a tick runs
  fpmhip_context drain;                 // stream of the result downloads, while a tick runs
  Context* ctx = nullptr;
  fpmhip_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;
  std::map<uint64_t, std::unique_ptr<Job>> jobs;
  std::vector<Job*> resident;
  uint64_t resident_bytes = 0;
  uint64_t resident_wgs = 0;  // workgroups of the resident tiles
  uint32_t wg_limit = 0;      // workgroups of k_sweep resident at once on the device
  uint64_t next_id = 1;
  bool stop = false;
  fpmhip_server_statistics stats{};
  std::thread worker;
  DeviceBuffer tw, tiles;
  uint32_t prepared_nmax = 0;

  void run();
  bool admit(Job& j, fpmhip_context& c);
  // Admits one job on `c` (admit) and marks it resident or failed.
  void admit_one(Job& j, fpmhip_context& c);
  // Admits jobs on up to kAdmitThreads threads and empties the list.
  void admit_all(std::vector<Job*>& to_admit);
  void finish(Job& j);
  // Finishes jobs (finish), releases their device objects and marks them
  // done or failed; empties the list.
  void finish_all(std::vector<Job*>& done);
  // Moves queued jobs to `to_admit` while the next tick has tile slots,
  // co-resident workgroups and HBM budget for them. The tiles of finishing
  // jobs and of `leaving` (the jobs the tick being planned completes) take
  // no part in the next tick, so their slots and workgroups count as free;
  // their bytes still count until they are released.
  void select_admissions(std::vector<Job*>& to_admit, const std::vector<Job*>& leaving);
  // Makes the blocks given back to the pool reusable. Called between ticks,
  // after every stream that job buffers are used on has been synchronised.
  void recycle_pool();
};

namespace {

uint64_t job_bytes(const fpmhip_model_desc& d, uint64_t iterations) {
  const uint64_t N = d.image_h * d.image_w, M = d.recon_h * d.recon_w;
  const uint64_t F = d.mux_row_ptr ? d.frame_count : d.source_count;
  // spectrum and scratch (FP64 complex), velocity and anchor (FP32 complex)
  // + per-pixel session buffers + counts + model arrays + per-tile scratch
  return 2 * M * 16 + 2 * M * 8 + 12 * N * 16 + F * N * 2 + N * 25 + F * 64 + 16 + (iterations + F) * 8;
}

// A tick lasts about this many sweeps of its heaviest tile: every other tile
// runs floor(kTickSweeps x heaviest work / its work) sweeps, so the rounding
// leaves a tile idle for less than 1 / kTickSweeps of the tick.
constexpr double kTickSweeps = 2.0;

// Time of one LED update of the fused sweep by tile size, in microseconds:
// batches of one tile size at a time, 13 x 13 LEDs, on an MI355X (256^2
// tiles as cooperative groups of kGroupMembers workgroups, or on one
// workgroup). It is not proportional to the pixels: a small tile's LED is
// bound by the latency of its dependent phases. Used only to balance the
// per-tick sweep quotas.
double led_cost_us(uint32_t n, bool grouped) {
  switch (n) {
    case 32: return 9.1;
    case 64: return 12.2;
    case 128: return 26.2;
    default: return grouped ? 35.7 : 199.0;  // 256^2
  }
}

// Work of one sweep of a job's tile: LED updates x the cost of one.
double sweep_work(const Job& j, bool grouped) {
  const double F = (double)(j.desc.mux_row_ptr ? j.desc.frame_count : j.desc.source_count);
  return F * led_cost_us((uint32_t)j.desc.image_h, grouped);
}

}  // namespace

bool fpmhip_server::admit(Job& j, fpmhip_context& c) {
  int rc = fpmhip_model_create(&c, &j.desc, &j.model);
  if (!rc)
    rc = fpmhip_measurements_create(&c, j.model, j.counts.empty() ? nullptr : j.counts.data(),
                                    j.frames.empty() ? nullptr : j.frames.data(),
                                    j.weights.empty() ? nullptr : j.weights.data(), nullptr, &j.meas);
  if (!rc) rc = fpmhip_session_create(j.model, j.meas, FPMHIP_MIXED_F32, &j.alg, &j.sess);
  const char* why = "";
  if (!rc && j.alg.kind == FPMHIP_EPRY) {  // the tick's sweep has no per-tile pupil recovery or gauge
    rc = FPMHIP_UNSUPPORTED;
    j.error = "served jobs do not support EPRY";
  }
  if (!rc && !sweep_supports(*j.sess, &why)) {
    rc = FPMHIP_UNSUPPORTED;
    j.error = why;
  }
  if (!rc) rc = fpmhip_session_initialize(j.sess);
  if (rc) {
    j.error_code = rc;
    if (j.error.empty()) j.error = fpmhip_last_error();
    j.release_device();
    return false;
  }
  j.scratch.allocate(16 + (j.iterations + j.sess->model->F) * sizeof(double));
  HIP_CHECK(hipMemsetAsync(j.scratch.ptr, 0, j.scratch.bytes, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  // the counter starts from the session's (zero after initialisation)
  return true;
}

void fpmhip_server::finish(Job& j) {
  PoolScope scope(pool.get());
  fpmhip_session& s = *j.sess;
  s.ctx = &drain;  // the result transform and downloads run on the drain stream
  const uint64_t M = s.model->M();
  j.object.resize(2 * M);
  j.spectrum.resize(2 * M);
  j.trace.resize(j.iterations);
  int rc = fpmhip_session_get_spectrum(&s, j.spectrum.data());
  if (!rc) rc = fpmhip_session_get_object(&s, j.object.data());
  if (!rc) {
    HIP_CHECK(hipMemcpyAsync(j.trace.data(), j.scratch.as<char>() + 16, j.iterations * sizeof(double),
                             hipMemcpyDeviceToHost, drain.stream));
    int32_t status = 0;
    HIP_CHECK(hipMemcpyAsync(&status, j.scratch.as<char>() + 12, sizeof(status), hipMemcpyDeviceToHost, drain.stream));
    HIP_CHECK(hipStreamSynchronize(drain.stream));
    if (status == 2) {
      rc = FPMHIP_DEVICE_ERROR;
      j.error = "a cooperative tile group stalled at a group barrier";
    } else if (status) {
      rc = FPMHIP_NUMERICAL;
      j.error = "mPIE momentum produced a non-finite object or velocity";
    }
  } else {
    j.error = fpmhip_last_error();
  }
  j.error_code = rc;
}

void fpmhip_server::finish_all(std::vector<Job*>& done) {
  for (Job* j : done) {
    try {
      finish(*j);
    } catch (const Failure& f) {
      j->error_code = f.code;
      j->error = fpmhip_last_error();
    } catch (...) {
      j->error_code = FPMHIP_DEVICE_ERROR;
      j->error = "result download failed";
    }
    j->release_device();  // into the pool: no hipFree while a tick runs
    std::lock_guard<std::mutex> lock(mu);
    j->t_done = now_s();
    j->state = j->error_code ? kFailed : kDone;
    if (j->error_code)
      stats.failed++;
    else
      stats.completed++;
    resident.erase(std::find(resident.begin(), resident.end(), j));
    resident_bytes -= j->bytes;
    resident_wgs -= j->wgs;
    cv_done.notify_all();
  }
  done.clear();
}

void fpmhip_server::select_admissions(std::vector<Job*>& to_admit, const std::vector<Job*>& leaving) {
  std::lock_guard<std::mutex> lock(mu);
  uint64_t bytes = resident_bytes, wgs = resident_wgs;
  size_t slots = resident.size();
  auto free_up = [&](const Job* j) {
    wgs -= j->wgs;
    --slots;
  };
  for (const Job* j : resident)
    if (j->finishing) free_up(j);
  for (const Job* j : leaving) free_up(j);
  // the resident tiles of the next tick, large ones as cooperative groups,
  // fit on the device at once, so that no large tile falls back to a single
  // workgroup (a job over the HBM or workgroup budget is still admitted when
  // nothing else would run)
  while (!queue.empty() && slots < cfg.max_tiles) {
    Job* j = jobs[queue.front()].get();
    if (slots > 0 && (bytes + j->bytes > cfg.hbm_budget_bytes || wgs + j->wgs > wg_limit)) break;
    queue.pop_front();
    to_admit.push_back(j);
    bytes += j->bytes;
    wgs += j->wgs;
    ++slots;
  }
}

void fpmhip_server::admit_one(Job& j, fpmhip_context& c) {
  bool ok = false;
  try {
    ok = admit(j, c);
  } catch (const Failure& f) {
    j.error_code = f.code;
    j.error = fpmhip_last_error();
    j.release_device();
  } catch (...) {
    j.error_code = FPMHIP_DEVICE_ERROR;
    j.error = "admission failed";
    j.release_device();
  }
  std::lock_guard<std::mutex> lock(mu);
  if (ok) {
    j.state = kResident;
    j.t_admit = now_s();
    resident.push_back(&j);
    resident_bytes += j.bytes;
    resident_wgs += j.wgs;
    stats.admitted++;
  } else {
    j.state = kFailed;
    j.t_done = now_s();
    stats.failed++;
    cv_done.notify_all();
  }
}

void fpmhip_server::admit_all(std::vector<Job*>& to_admit) {
  // jobs are independent: thread k admits jobs k, k + t, ... on stage[k]
  const size_t n = to_admit.size();
  const size_t t = std::min<size_t>(n, kAdmitThreads);
  auto work = [&](size_t k) {
    (void)hipSetDevice(ctx->device);
    PoolScope scope(pool.get());
    for (size_t i = k; i < n; i += t) admit_one(*to_admit[i], stage[k]);
  };
  std::vector<std::thread> helpers;
  for (size_t k = 1; k < t; ++k) helpers.emplace_back(work, k);
  if (n) work(0);
  for (std::thread& th : helpers) th.join();
  to_admit.clear();
}

void fpmhip_server::recycle_pool() {
  // an admission that failed part-way may have left work on its stream
  (void)hipStreamSynchronize(own.stream);
  (void)hipStreamSynchronize(drain.stream);
  for (fpmhip_context& c : stage) (void)hipStreamSynchronize(c.stream);
  pool->recycle();
}

void fpmhip_server::run() {
  (void)hipSetDevice(ctx->device);
  double busy_since = -1.0;
  std::vector<Job*> finished;  // jobs whose last tick has run
  for (;;) {
    std::vector<Job*> active;
    {
      std::unique_lock<std::mutex> lock(mu);
      cv_work.wait(lock, [&] { return stop || !queue.empty() || !resident.empty(); });
      if (stop) break;
      for (Job* j : resident)
        if (!j->finishing) active.push_back(j);
    }
    std::vector<Job*> to_admit;
    if (active.empty()) {
      select_admissions(to_admit, {});
      finish_all(finished);
      admit_all(to_admit);
      recycle_pool();
      if (busy_since >= 0.0) {
        std::lock_guard<std::mutex> lock(mu);
        stats.busy_seconds += now_s() - busy_since;
        busy_since = -1.0;
      }
      continue;
    }
    if (busy_since < 0.0) busy_since = now_s();
    // one tick: every resident tile advances by its quota of sweeps
    std::vector<uint32_t> quota(active.size());
    uint64_t grid = 0;
    try {
      std::vector<TileDesc> desc(active.size());
      uint32_t nmax = 0;
      std::map<const double2*, size_t> index;  // tile (its spectrum) -> resident job
      for (size_t i = 0; i < active.size(); ++i) {
        Job& j = *active[i];
        char* base = j.scratch.as<char>();
        fill_tile_desc(desc[i], *j.sess, reinterpret_cast<double*>(base + 16),
                       reinterpret_cast<double*>(base + 16 + j.iterations * sizeof(double)), base,
                       (uint32_t)j.iterations);
        nmax = std::max(nmax, desc[i].n);
        index[desc[i].spec] = i;
      }
      if (nmax != prepared_nmax) {
        prepare_sweep(nmax);
        prepared_nmax = nmax;
      }
      std::vector<TileDesc> launch = order_tiles(desc, kGroupMembers, nullptr, true, sweep_coresident_limit(nmax));
      // quotas from each tile's per-sweep work (as a cooperative group when
      // the launch made it one)
      std::vector<char> grouped(active.size(), 0);
      for (const TileDesc& t : launch)
        if (t.members > 1) grouped[index[t.spec]] = 1;
      std::vector<double> work(active.size());
      for (size_t i = 0; i < active.size(); ++i) work[i] = std::max(sweep_work(*active[i], grouped[i] != 0), 1e-3);
      const double wmax = *std::max_element(work.begin(), work.end());
      std::vector<Job*> leaving;  // jobs this tick completes
      for (size_t i = 0; i < active.size(); ++i) {
        const double share = std::max(std::floor(kTickSweeps * wmax / work[i]), 1.0);
        const uint64_t left = active[i]->iterations - active[i]->done_iterations;
        quota[i] = (uint32_t)std::min<double>(share, (double)left);
        if (quota[i] >= left) leaving.push_back(active[i]);
      }
      for (TileDesc& t : launch) t.sweeps = quota[index[t.spec]];
      bool cooperative = false;
      for (const TileDesc& t : launch) cooperative |= t.members > 1;
      // the jobs admitted while this tick runs join the next one, in the
      // room the finished and the leaving jobs make
      select_admissions(to_admit, leaving);
      if (tiles.bytes < launch.size() * sizeof(TileDesc))
        tiles.allocate(std::max<size_t>(launch.size(), cfg.max_tiles * kGroupMembers) * sizeof(TileDesc));
      HIP_CHECK(hipMemcpyAsync(tiles.ptr, launch.data(), launch.size() * sizeof(TileDesc), hipMemcpyHostToDevice,
                               ctx->stream));
      launch_sweep(tiles.as<TileDesc>(), (uint32_t)launch.size(), tw.as<float2>(), nmax, cooperative, ctx->stream);
      grid = launch.size();
      // while the tick runs: the results of the jobs finished by the last
      // tick are downloaded, and the selected jobs are admitted (uploads and
      // initialisation), on streams of their own
      std::thread drainer;
      if (!finished.empty())
        drainer = std::thread([&] {
          (void)hipSetDevice(ctx->device);
          finish_all(finished);
        });
      try {
        admit_all(to_admit);
      } catch (...) {
        if (drainer.joinable()) drainer.join();
        throw;
      }
      if (drainer.joinable()) drainer.join();
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
      recycle_pool();
    } catch (const Failure& f) {
      // no-ops for the lists handled before the failure
      finish_all(finished);
      admit_all(to_admit);
      {
        std::lock_guard<std::mutex> lock(mu);
        for (Job* j : active) {
          j->error_code = f.code;
          j->error = "device failure during a sweep";
          j->release_device();
          j->state = kFailed;
          j->t_done = now_s();
          stats.failed++;
          resident.erase(std::find(resident.begin(), resident.end(), j));
          resident_bytes -= j->bytes;
          resident_wgs -= j->wgs;
        }
        cv_done.notify_all();
      }
      recycle_pool();
      continue;
    }
    {
      std::lock_guard<std::mutex> lock(mu);
      stats.ticks++;
      stats.peak_tick_workgroups = std::max(stats.peak_tick_workgroups, grid);
      for (size_t i = 0; i < active.size(); ++i) {
        Job* j = active[i];
        j->done_iterations += quota[i];
        if (j->done_iterations >= j->iterations) {
          j->finishing = true;
          finished.push_back(j);
        }
      }
    }
  }
  // drop resident device state on shutdown
  std::lock_guard<std::mutex> lock(mu);
  for (Job* j : resident) j->release_device();
  resident.clear();
}

extern "C" {

int fpmhip_server_create(fpmhip_context* c, const fpmhip_server_config* config, fpmhip_server** out) {
  return guarded([&] {
    if (!c || !config || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "server");
    if (config->max_tiles == 0 || config->max_tiles > 65535)
      fail(FPMHIP_INVALID_PARAMETER, "max_tiles must be in 1..65535", "max_tiles");
    HIP_CHECK(hipSetDevice(c->device));
    auto s = std::make_unique<fpmhip_server>();
    s->own.device = c->device;
    s->own.props = c->props;
    HIP_CHECK(hipStreamCreateWithFlags(&s->own.stream, hipStreamNonBlocking));
    std::vector<fpmhip_context*> aux = {&s->drain};
    for (fpmhip_context& x : s->stage) aux.push_back(&x);
    for (fpmhip_context* x : aux) {
      x->device = c->device;
      x->props = c->props;
      HIP_CHECK(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));
    }
    s->ctx = &s->own;
    s->cfg = *config;
    // idle pooled blocks kept: a quarter of the admission budget, at most 4 GiB
    s->pool = std::make_unique<DevicePool>(std::min<uint64_t>(config->hbm_budget_bytes / 4, 4ull << 30));
    s->wg_limit = std::max<uint32_t>(sweep_coresident_limit(kFusedMaxN), kGroupMembers);  // tiles up to 256^2
    make_sweep_twiddles(s->tw, s->own.stream);
    s->tiles.allocate(config->max_tiles * sizeof(TileDesc));
    fpmhip_server* raw = s.get();
    s->worker = std::thread([raw] { raw->run(); });
    *out = s.release();
  });
}

int fpmhip_server_destroy(fpmhip_server* s) {
  return guarded([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto& kv : s->jobs) kv.second->release_device();
    std::vector<hipStream_t> streams = {s->own.stream, s->drain.stream};
    for (const fpmhip_context& x : s->stage) streams.push_back(x.stream);
    delete s;
    for (hipStream_t st : streams)
      if (st) (void)hipStreamDestroy(st);
  });
}

int fpmhip_server_submit(fpmhip_server* s, const fpmhip_model_desc* d, const uint16_t* counts, const double* frames,
                         const double* weights, const fpmhip_algorithm* alg, uint64_t* job_id) {
  return guarded([&] {
    if (!s || !d || !alg || !job_id) fail(FPMHIP_INVALID_PARAMETER, "null argument", "submit");
    if ((counts == nullptr) == (frames == nullptr))
      fail(FPMHIP_INVALID_MEASUREMENTS, "exactly one of uint16 counts and f64 frames is required");
    if (int rc = fpmhip_algorithm_validate(alg)) throw Failure{rc};
    if (alg->iterations == 0) fail(FPMHIP_INVALID_PARAMETER, "a served job needs at least one iteration", "iterations");
    // the tick balance is calibrated for groups of up to 256^2 tiles; larger
    // tiles run in a Batch
    if (d->image_h > 256 || d->image_w > 256)
      fail(FPMHIP_UNSUPPORTED, "served jobs take images of at most 256 x 256 pixels");
    auto j = std::make_unique<Job>();
    const uint64_t N = d->image_h * d->image_w, S = d->source_count;
    const uint64_t F = d->mux_row_ptr ? d->frame_count : S;
    j->crop.assign(d->crop_start, d->crop_start + 2 * S);
    if (d->offset) j->offset.assign(d->offset, d->offset + 2 * S);
    j->pupil.assign(d->pupil, d->pupil + 2 * N);
    j->support.assign(d->support, d->support + N);
    if (d->frame_gain) j->gain.assign(d->frame_gain, d->frame_gain + F);
    if (d->background && d->background_len) j->bg.assign(d->background, d->background + d->background_len);
    if (d->k_vectors) j->kv.assign(d->k_vectors, d->k_vectors + 2 * S);
    if (d->mux_row_ptr) {
      j->row_ptr.assign(d->mux_row_ptr, d->mux_row_ptr + F + 1);
      j->msrc.assign(d->mux_source, d->mux_source + j->row_ptr.back());
      j->mw.assign(d->mux_weight, d->mux_weight + j->row_ptr.back());
    }
    if (counts) j->counts.assign(counts, counts + F * N);
    if (frames) j->frames.assign(frames, frames + F * N);
    if (weights) j->weights.assign(weights, weights + F);
    j->desc = *d;
    j->desc.crop_start = j->crop.data();
    j->desc.offset = j->offset.empty() ? nullptr : j->offset.data();
    j->desc.pupil = j->pupil.data();
    j->desc.support = j->support.data();
    j->desc.frame_gain = j->gain.empty() ? nullptr : j->gain.data();
    j->desc.background = j->bg.empty() ? nullptr : j->bg.data();
    j->desc.background_len = j->bg.size();
    j->desc.k_vectors = j->kv.empty() ? nullptr : j->kv.data();
    j->desc.mux_row_ptr = j->row_ptr.empty() ? nullptr : j->row_ptr.data();
    j->desc.mux_source = j->msrc.empty() ? nullptr : j->msrc.data();
    j->desc.mux_weight = j->mw.empty() ? nullptr : j->mw.data();
    j->alg = *alg;
    j->iterations = alg->iterations;
    j->bytes = job_bytes(*d, alg->iterations);
    j->wgs = d->image_h > kFusedLdsMaxN ? kGroupMembers : 1;
    j->t_submit = now_s();
    {
      std::lock_guard<std::mutex> lock(s->mu);
      j->id = s->next_id++;
      *job_id = j->id;
      s->queue.push_back(j->id);
      s->jobs[j->id] = std::move(j);
      s->stats.submitted++;
    }
    s->cv_work.notify_all();
  });
}

int fpmhip_server_poll(fpmhip_server* s, uint64_t id, int32_t* state) {
  return guarded([&] {
    if (!s || !state) fail(FPMHIP_INVALID_PARAMETER, "null argument", "poll");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) fail(FPMHIP_INVALID_PARAMETER, "unknown job id", "job_id");
    *state = it->second->state;
  });
}

int fpmhip_server_wait(fpmhip_server* s, uint64_t id, double timeout_s) {
  return guarded([&] {
    if (!s) fail(FPMHIP_INVALID_PARAMETER, "null argument", "wait");
    std::unique_lock<std::mutex> lock(s->mu);
    auto done = [&] {
      auto it = s->jobs.find(id);
      return it == s->jobs.end() || it->second->state >= kDone;
    };
    if (timeout_s > 0) {
      if (!s->cv_done.wait_for(lock, std::chrono::duration<double>(timeout_s), done))
        fail(FPMHIP_UNSUPPORTED, "timed out waiting for the job");
    } else {
      s->cv_done.wait(lock, done);
    }
    if (s->jobs.find(id) == s->jobs.end()) fail(FPMHIP_INVALID_PARAMETER, "unknown job id", "job_id");
  });
}

int fpmhip_server_result(fpmhip_server* s, uint64_t id, double* object, double* spectrum, double* trace,
                         double* latency_s, double* queue_s) {
  return guarded([&] {
    if (!s) fail(FPMHIP_INVALID_PARAMETER, "null argument", "result");
    std::unique_ptr<Job> j;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      auto it = s->jobs.find(id);
      if (it == s->jobs.end()) fail(FPMHIP_INVALID_PARAMETER, "unknown job id", "job_id");
      if (it->second->state < kDone) fail(FPMHIP_UNSUPPORTED, "job is not finished");
      j = std::move(it->second);
      s->jobs.erase(it);
    }
    if (j->state == kFailed) fail(j->error_code ? j->error_code : FPMHIP_DEVICE_ERROR, j->error);
    if (object) std::memcpy(object, j->object.data(), j->object.size() * sizeof(double));
    if (spectrum) std::memcpy(spectrum, j->spectrum.data(), j->spectrum.size() * sizeof(double));
    if (trace) std::memcpy(trace, j->trace.data(), j->trace.size() * sizeof(double));
    if (latency_s) *latency_s = j->t_done - j->t_submit;
    if (queue_s) *queue_s = j->t_admit - j->t_submit;
  });
}

int fpmhip_server_stats(fpmhip_server* s, fpmhip_server_statistics* out) {
  return guarded([&] {
    if (!s || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "stats");
    std::lock_guard<std::mutex> lock(s->mu);
    *out = s->stats;
    out->resident_tiles = s->resident.size();
    out->resident_bytes = s->resident_bytes;
    out->pool_reuses = s->pool->reuses();
    out->pool_allocations = s->pool->allocations();
  });
}

}  // extern "C"
