// batch.hip — batched form: many independent tiles (own model, pupil and
// measurements; uneven sizes) advanced by one launch of the fused sweep
// kernel per iteration, captured once in a HIP graph and replayed.
#include <algorithm>
#include <cmath>
#include <limits>
#include <utility>

#include "fused.h"
#include "gauge.h"

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

__global__ void k_set_sweeps(TileDesc* tiles, uint32_t count, uint32_t sweeps) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) tiles[i].sweeps = sweeps;
}

__global__ void k_reset_tiles(TileDesc* tiles, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count && tiles[i].member == 0) {
    *tiles[i].iteration = 0;
    tiles[i].status[0] = 0;
    if (tiles[i].group) {
      tiles[i].group[0] = 0;
      tiles[i].group[1] = 0;
      tiles[i].group[2] = 0;
    }
  }
}

}  // namespace

namespace fpmhip {

void fill_tile_desc(TileDesc& t, fpmhip_session& s, double* trace, double* frame_loss, char* per16,
                    uint32_t trace_len) {
  const Model& m = *s.model;
  const Measurements& me = *s.meas;
  t = TileDesc{};
  t.spec = s.b.spec.as<double2>();
  t.vel = s.b.vel.as<float2>();
  t.anc = s.b.anc.as<float2>();
  t.pupil = s.b.pupil.as<float2>();
  t.extent = s.b.extent.as<uint32_t>();
  t.dest = s.b.dest.as<uint32_t>();
  t.field = m.nh > kFusedLdsMaxN ? s.b.field.as<float2>() : nullptr;  // the session's N-pixel field buffer
  t.group = s.b.group.as<uint64_t>();
  t.partial = s.b.group.ptr ? reinterpret_cast<double2*>(s.b.group.as<char>() + 64) : nullptr;
  t.member = 0;
  t.members = 1;
  t.frames = me.d_frames.ptr;
  t.masks = me.d_masks.as<uint8_t>();
  t.weights = me.d_weights.as<double>();
  t.gains = s.b.gains.as<double>();
  t.bg = s.b.bg.as<double>();
  t.bg_len = s.bg_len;
  t.src = m.d_src.as<SrcParam>();
  t.order = nullptr;
  t.mspan = s.b.mspan.as<uint32_t>();
  t.trace = trace;
  t.frame_loss = frame_loss;
  t.counter = reinterpret_cast<uint64_t*>(per16);
  t.iteration = reinterpret_cast<uint32_t*>(per16 + 8);
  t.status = reinterpret_cast<int32_t*>(per16 + 12);
  t.trace_len = trace_len;
  t.sweeps = 1;
  t.n = m.nh;
  t.rh = m.rh;
  t.rw = m.rw;
  t.F = (uint32_t)m.F;
  t.interval = (uint32_t)s.cfg.interval;
  t.is_u16 = me.is_u16 ? 1 : 0;
  t.denom = s.cfg.denom;
  t.momentum = s.cfg.momentum ? 1 : 0;
  t.loss = s.cfg.loss;
  t.object_step = s.cfg.object_step;
  t.alpha = s.cfg.alpha;
  t.eps = s.cfg.epsilon;
  double maxp = 0.0;
  for (uint64_t p = 0; p < m.N(); ++p)
    maxp = std::max(maxp, m.pupil[2 * p] * m.pupil[2 * p] + m.pupil[2 * p + 1] * m.pupil[2 * p + 1]);
  t.max_p = maxp;
  t.friction = s.cfg.friction;
  t.feedback = s.cfg.feedback;
  t.recover_pupil = s.cfg.has_pupil_step ? 1 : 0;
  t.patch = t.recover_pupil ? s.b.patch.as<float2>() : nullptr;
  t.pupil64 = s.b.pupil64.as<double2>();
  t.bounds = t.recover_pupil ? s.b.bounds.as<uint64_t>() : nullptr;
  t.bcols = (m.rw + kBoundBlock - 1) / kBoundBlock;
  t.pupil_step = s.cfg.pupil_step;
}

std::vector<TileDesc> order_tiles(const std::vector<TileDesc>& tiles, uint32_t members, const uint32_t* radix,
                                  bool largest_first, uint32_t grid_limit) {
  std::vector<TileDesc> sorted(tiles);
  if (largest_first)
    std::stable_sort(sorted.begin(), sorted.end(), [](const TileDesc& a, const TileDesc& b) { return a.n > b.n; });
  // large tiles that become cooperative groups: as many as the resident-grid
  // limit leaves room for, those with the most LEDs first
  std::vector<uint32_t> big;
  for (uint32_t i = 0; i < sorted.size(); ++i)
    if (sorted[i].n > kFusedLdsMaxN && sorted[i].group) big.push_back(i);
  // workgroups of each tile: 1, or its cooperative group's size
  std::vector<uint32_t> gsize(sorted.size(), 1);
  if (members != 1 && !big.empty()) {
    // each grouped tile adds its group size - 1 workgroups to the grid
    uint64_t room = grid_limit > sorted.size() ? grid_limit - sorted.size() : 0;
    std::stable_sort(big.begin(), big.end(), [&](uint32_t x, uint32_t y) { return sorted[x].F > sorted[y].F; });
    auto wanted = [&](uint32_t n) { return n == 256 && members > 1 ? members : group_members(n); };
    uint32_t n256 = 0;
    for (uint32_t k : big) {
      const uint64_t extra = wanted(sorted[k].n) - 1;
      if (extra > room) continue;
      gsize[k] = wanted(sorted[k].n);
      room -= extra;
      n256 += sorted[k].n == 256;
    }
    if (members == 0 && n256 > 0) {
      // automatic: the co-resident room left over goes to the 256^2 groups,
      // the same number of members each, up to kMaxGroupMembers
      const uint32_t add = (uint32_t)std::min<uint64_t>(room / n256, kMaxGroupMembers - group_members(256));
      for (uint32_t k : big)
        if (sorted[k].n == 256 && gsize[k] > 1) gsize[k] += add;
    }
  }
  // Grid positions. Workgroups are dispatched to the XCDs round-robin
  // (workgroup i to XCD i mod xccs), so the members of a cooperative group
  // take positions xccs apart, all on one XCD, on the XCD with the most free
  // positions; single-workgroup tiles fill the other positions in order.
  uint32_t total = 0;
  for (uint32_t i = 0; i < sorted.size(); ++i) total += gsize[i];
  const uint32_t xccs = std::max<uint32_t>(device_xccs(), 1);
  std::vector<std::vector<uint32_t>> free_at(xccs);  // free positions per XCD, ascending
  for (uint32_t p = total; p-- > 0;) free_at[p % xccs].push_back(p);  // back() = lowest
  std::vector<TileDesc> out(total);
  std::vector<char> used(total, 0);
  std::vector<uint32_t> singles;
  for (uint32_t i = 0; i < sorted.size(); ++i) {
    TileDesc t = sorted[i];
    if (radix && t.n <= 256) t.plan = radix[t.n == 32 ? 0 : t.n == 64 ? 1 : t.n == 128 ? 2 : 3];
    const uint32_t g = gsize[i];
    if (g == 1) {
      sorted[i] = t;
      singles.push_back(i);
      continue;
    }
    uint32_t x = 0;
    for (uint32_t k = 1; k < xccs; ++k)
      if (free_at[k].size() > free_at[x].size()) x = k;
    for (uint32_t m = 0; m < g; ++m) {
      uint32_t p;
      if (!free_at[x].empty()) {
        p = free_at[x].back();
        free_at[x].pop_back();
      } else {  // no XCD has room for the whole group: any free position
        uint32_t y = 0;
        for (uint32_t k = 1; k < xccs; ++k)
          if (free_at[k].size() > free_at[y].size()) y = k;
        p = free_at[y].back();
        free_at[y].pop_back();
      }
      used[p] = 1;
      out[p] = t;
      out[p].member = m;
      out[p].members = g;
    }
  }
  uint32_t p = 0;
  for (uint32_t i : singles) {
    while (used[p]) ++p;
    used[p] = 1;
    out[p] = sorted[i];
    out[p].member = 0;
    out[p].members = 1;
  }
  return out;
}

uint32_t device_xccs() {
  int dev = 0, n = 1;
  HIP_CHECK(hipGetDevice(&dev));
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess) n = 1;
  return (uint32_t)std::max(n, 1);
}

void make_sweep_twiddles(DeviceBuffer& out, hipStream_t stream) {
  // packed tables W_n^j (forward, inverse) for n = 32, 64, 128, 256, 512
  std::vector<float> tw(2 * kTwTotal);
  for (uint32_t n : {32u, 64u, 128u, 256u, 512u}) {
    const uint32_t off = tw_offset(n);
    for (uint32_t j = 0; j < n; ++j)
      for (int d = 0; d < 2; ++d) {
        const int sign = d == 0 ? -1 : 1;
        const long double th = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
        double re, im;
        unit_root(j, n, sign, &re, &im);
        const bool exact = (4ull * j) % n == 0;
        tw[2 * (off + d * n + j)] = exact ? (float)re : (float)cosl(th);
        tw[2 * (off + d * n + j) + 1] = exact ? (float)im : (float)(sign * sinl(th));
      }
  }
  out.allocate(tw.size() * sizeof(float));
  HIP_CHECK(hipMemcpyAsync(out.ptr, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
}

bool sweep_supports(const fpmhip_session& s, const char** why) {
  const Model& m = *s.model;
  if (s.precision != FPMHIP_MIXED_F32) {
    *why = "batched tiles must use MIXED_F32 sessions";
    return false;
  }
  if (m.nh != m.nw || (m.nh != 32 && m.nh != 64 && m.nh != 128 && m.nh != 256 && m.nh != 512)) {
    *why = "batched tiles must be square 32, 64, 128, 256 or 512 pixel images";
    return false;
  }
  if (m.mux) {
    *why = "batched tiles do not support multiplexed frames";
    return false;
  }
  if (s.alg.kind == FPMHIP_EPRY) {
    if (m.nh > kFusedLdsMaxN) {
      *why = "batched EPRY tiles must be at most 128 x 128 pixels";
      return false;
    }
    if (s.cfg.gain_update || s.cfg.bg_update) {
      *why = "batched EPRY tiles do not recover frame gains or background";
      return false;
    }
    if (s.cfg.has_pupil_step) {
      if (!s.cfg.constrain) {
        *why = "batched EPRY pupil recovery needs the pupil support constraint";
        return false;
      }
      // the fused sweep updates the pupil on its support strips: the model
      // pupil (where every run starts) must be non-zero exactly on the support
      for (uint64_t p = 0; p < m.N(); ++p)
        if (((float)m.pupil[2 * p] != 0.0f || (float)m.pupil[2 * p + 1] != 0.0f) != (m.support[p] != 0)) {
          *why = "batched EPRY pupil recovery needs a model pupil that is non-zero exactly on its support";
          return false;
        }
    }
  }
  if (s.cfg.momentum && !s.b.mspan.ptr) {  // chunk starts are 25-bit pixel indices
    *why = "batched mPIE tiles need an object grid below 2^25 pixels";
    return false;
  }
  return true;
}

}  // namespace fpmhip

struct fpmhip_batch {
  Context* ctx = nullptr;
  std::vector<fpmhip_session*> sessions;
  std::vector<TileDesc> host_tiles;    // in session order
  std::vector<TileDesc> device_order;  // launch order (largest tiles first, groups expanded)
  fpmhip_batch_schedule schedule{0, 1, {0, 0, 0, 0}, 0};
  uint32_t launched = 0;               // workgroups per sweep launch
  bool cooperative = false;
  DeviceBuffer tiles;      // TileDesc[launched]
  DeviceBuffer tw;         // float2[kTwTotal]
  DeviceBuffer per_tile;   // counters, iterations, status
  DeviceBuffer traces;     // double[count * trace_cap]
  DeviceBuffer losses;     // double[count * Fmax]
  uint32_t nmax = 0;
  uint32_t trace_cap = 0;
  uint64_t fmax = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DeviceBuffer gauges;    // GaugeTile[gauge_count] of the tiles that canonicalize (EPRY), in session order
  uint32_t gauge_count = 0;
  double* pinned = nullptr;  // page-locked result objects (fpmhip_batch_objects_pinned)
  size_t pinned_bytes = 0;
  // batched initialisation and results: one InitTile per session, and one
  // R-array slot per tile in `slots`, tiles of equal R consecutive so that
  // each group is transformed by one batched 2-D FFT
  struct SlotGroup {
    const Fft2* plan;
    double2* base;
    uint32_t count;
  };
  DeviceBuffer init_tiles;  // InitTile[count], session order
  DeviceBuffer slots;       // double2[sum of R_h * R_w]
  DeviceBuffer init_flags;  // int[2 * count]
  std::vector<SlotGroup> slot_groups;
  std::vector<double2*> slot_of;  // per session
  uint32_t init_nmax = 0, init_mmax = 0;
  uint64_t init_cmax = 0;
  double ms_mean = 0.0, ms_max = 0.0;
  uint64_t sweeps = 0;
  ~fpmhip_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (pinned) (void)hipHostFree(pinned);
  }
};

namespace {

// InitTile table and result slots of the batch (the per-tile counters live
// in per_tile, which build_tiles reallocates: rebuilt with it).
void build_init(fpmhip_batch& b) {
  const uint32_t count = (uint32_t)b.sessions.size();
  std::vector<std::pair<uint32_t, uint32_t>> shapes;  // distinct R, in order of first appearance
  for (fpmhip_session* s : b.sessions) {
    const std::pair<uint32_t, uint32_t> r{s->model->rh, s->model->rw};
    if (std::find(shapes.begin(), shapes.end(), r) == shapes.end()) shapes.push_back(r);
  }
  uint64_t total = 0;
  for (fpmhip_session* s : b.sessions) total += s->model->M();
  if (b.slots.bytes < total * sizeof(double2)) b.slots.allocate(total * sizeof(double2));
  b.init_flags.allocate(std::max<size_t>(2 * (size_t)count, 1) * sizeof(int));
  b.slot_of.assign(count, nullptr);
  b.slot_groups.clear();
  uint64_t off = 0;
  for (const auto& r : shapes) {
    fpmhip_batch::SlotGroup g{nullptr, b.slots.as<double2>() + off, 0};
    for (uint32_t i = 0; i < count; ++i) {
      const Model& m = *b.sessions[i]->model;
      if (m.rh != r.first || m.rw != r.second) continue;
      if (!g.plan) g.plan = &*m.fft_high;
      b.slot_of[i] = b.slots.as<double2>() + off;
      off += m.M();
      ++g.count;
    }
    b.slot_groups.push_back(g);
  }
  std::vector<InitTile> tiles(count);
  b.init_nmax = b.init_mmax = 0;
  b.init_cmax = 0;
  for (uint32_t i = 0; i < count; ++i) {
    fpmhip_session& s = *b.sessions[i];
    const Model& m = *s.model;
    const Measurements& me = *s.meas;
    InitTile& t = tiles[i];
    t = InitTile{};
    t.frames = me.d_frames.ptr;
    t.weights = me.d_weights.as<double>();
    t.masks = me.d_masks.as<uint8_t>();
    t.model_gains = m.d_gains.as<double>();
    t.model_bg = m.d_background.as<double>();
    t.model_pupil = m.d_pupil64.as<double2>();
    t.part = s.b.work_m.as<double2>();
    t.avg = s.b.avg.as<double>();
    t.aw = s.b.aw.as<double>();
    t.scal = s.b.scal.as<double>();
    t.flag = s.b.flag.as<int>();
    t.counter = b.host_tiles[i].counter;
    t.slot = b.slot_of[i];
    t.spec = s.b.spec.as<double2>();
    t.pupil = s.b.pupil.as<float2>();
    t.pupil64 = s.b.pupil64.as<double2>();
    t.vel = s.cfg.momentum ? s.b.vel.as<float2>() : nullptr;
    t.anc = s.cfg.momentum ? s.b.anc.as<float2>() : nullptr;
    t.gains = s.b.gains.as<double>();
    t.bg = s.b.bg.as<double>();
    t.F = m.F;
    t.chunks = init_chunks(s);
    t.chunk = (m.F + t.chunks - 1) / t.chunks;
    t.bg_len = m.background.size();
    t.N = (uint32_t)m.N();
    t.nh = m.nh;
    t.nw = m.nw;
    t.rh = m.rh;
    t.rw = m.rw;
    t.is_u16 = me.is_u16 ? 1 : 0;
    b.init_nmax = std::max<uint32_t>(b.init_nmax, t.N);
    b.init_mmax = std::max<uint32_t>(b.init_mmax, (uint32_t)m.M());
    b.init_cmax = std::max<uint64_t>({b.init_cmax, (uint64_t)t.N, t.F, t.bg_len});
  }
  b.init_tiles.allocate(std::max<size_t>(count, 1) * sizeof(InitTile));
  HIP_CHECK(hipMemcpyAsync(b.init_tiles.ptr, tiles.data(), count * sizeof(InitTile), hipMemcpyHostToDevice,
                           b.ctx->stream));
  HIP_CHECK(hipStreamSynchronize(b.ctx->stream));  // `tiles` is pageable host memory
}

void build_tiles(fpmhip_batch& b, uint32_t trace_cap) {
  const uint32_t count = (uint32_t)b.sessions.size();
  b.trace_cap = trace_cap;
  b.traces.allocate((size_t)count * trace_cap * sizeof(double));
  b.losses.allocate((size_t)count * b.fmax * sizeof(double));
  // per tile: counter (u64), iteration (u32), status (i32)
  b.per_tile.allocate((size_t)count * 16);
  b.host_tiles.resize(count);
  for (uint32_t i = 0; i < count; ++i)
    fill_tile_desc(b.host_tiles[i], *b.sessions[i], b.traces.as<double>() + (size_t)i * trace_cap,
                   b.losses.as<double>() + (size_t)i * b.fmax, b.per_tile.as<char>() + (size_t)i * 16, trace_cap);
  prepare_sweep(b.nmax);  // the LDS attribute first: the residency query depends on it
  b.device_order = order_tiles(b.host_tiles, b.schedule.group_members, b.schedule.radix,
                               b.schedule.largest_first != 0, sweep_coresident_limit(b.nmax));
  b.launched = (uint32_t)b.device_order.size();
  b.cooperative = false;
  for (const TileDesc& t : b.device_order) b.cooperative |= t.members > 1;
  b.tiles.allocate(b.launched * sizeof(TileDesc));
  HIP_CHECK(hipMemcpyAsync(b.tiles.ptr, b.device_order.data(), b.launched * sizeof(TileDesc), hipMemcpyHostToDevice,
                           b.ctx->stream));
  // capture the sweep launch once
  if (b.exec) {
    (void)hipGraphExecDestroy(b.exec);
    b.exec = nullptr;
  }
  if (b.graph) {
    (void)hipGraphDestroy(b.graph);
    b.graph = nullptr;
  }
  HIP_CHECK(hipStreamSynchronize(b.ctx->stream));
  // the EPRY gauge of every tile that canonicalizes, after each of its sweeps
  std::vector<GaugeTile> gauges;
  for (uint32_t i = 0; i < count; ++i) {
    fpmhip_session& s = *b.sessions[i];
    if (!s.cfg.gauge) continue;
    GaugeTile g = gauge_tile(s);
    g.status = b.host_tiles[i].status;
    g.maxp_out = nullptr;  // the fused sweep takes max |P|^2 at the start of each sweep
    gauges.push_back(g);
  }
  b.gauge_count = (uint32_t)gauges.size();
  if (b.gauge_count) {
    b.gauges.allocate(gauges.size() * sizeof(GaugeTile));
    HIP_CHECK(hipMemcpyAsync(b.gauges.ptr, gauges.data(), gauges.size() * sizeof(GaugeTile), hipMemcpyHostToDevice,
                             b.ctx->stream));
    HIP_CHECK(hipStreamSynchronize(b.ctx->stream));  // `gauges` is pageable host memory
  }
  // capture the launch (the LDS attribute was set above, outside capture);
  // with a gauge, a launch runs one sweep and the gauge follows it
  HIP_CHECK(hipStreamBeginCapture(b.ctx->stream, hipStreamCaptureModeThreadLocal));
  // EPRY tiles: exact block bounds of |O|^2 before the sweep (their inserts
  // raise them during it)
  if (b.host_tiles[0].recover_pupil) launch_patch_bounds(b.tiles.as<TileDesc>(), b.launched, b.ctx->stream);
  launch_sweep(b.tiles.as<TileDesc>(), b.launched, b.tw.as<float2>(), b.nmax, b.cooperative, b.ctx->stream,
               b.host_tiles[0].recover_pupil != 0);
  if (b.gauge_count) launch_gauge(b.gauges.as<GaugeTile>(), b.gauge_count, b.ctx->stream);
  HIP_CHECK(hipStreamEndCapture(b.ctx->stream, &b.graph));
  HIP_CHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
  build_init(b);
}

// ReconstructionState::initialize (+ mPIE auxiliary state) of every tile and
// a reset of the per-tile sweep state: each pass one launch for all tiles,
// one batched FFT per distinct R, one wait for the checks.
void initialize_tiles(fpmhip_batch& b) {
  hipStream_t st = b.ctx->stream;
  const uint32_t count = (uint32_t)b.sessions.size();
  for (fpmhip_session* s : b.sessions) {
    const Model& m = *s->model;
    for (uint64_t f = 0; f < m.F; ++f) {
      const double g = m.has_gain ? m.gain[f] : 1.0;
      if (!std::isfinite(g) || g <= 0.0) fail(FPMHIP_INVALID_MODEL, "frame has invalid gain");
    }
    s->initialized = false;
  }
  const InitTile* tiles = b.init_tiles.as<InitTile>();
  launch_init_batch_pre(tiles, count, b.init_nmax, b.init_mmax, b.init_cmax, st);
  for (fpmhip_session* s : b.sessions)
    if (!s->extent_from_model) {  // after the pupil copy: extents of the model's pupil
      launch_pupil_extent(*s);
      upload_momentum_spans(*s, s->model->pupil.data());
      s->extent_from_model = true;
    }
  for (const fpmhip_batch::SlotGroup& g : b.slot_groups) fft2_f64(*g.plan, g.base, g.count, false, st);
  launch_init_batch_post(tiles, count, b.init_mmax, st);
  launch_init_batch_flags(tiles, count, b.init_flags.as<int>(), st);
  std::vector<int> flags(2 * (size_t)count);
  HIP_CHECK(hipMemcpyAsync(flags.data(), b.init_flags.ptr, flags.size() * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  for (uint32_t i = 0; i < count; ++i) {
    if (flags[2 * i]) fail(FPMHIP_NUMERICAL, "mPIE momentum produced a non-finite object or velocity");
    if (flags[2 * i + 1]) fail(FPMHIP_INVALID_MEASUREMENTS, "no positive-weight, unmasked measurements are available");
  }
  for (fpmhip_session* s : b.sessions) {  // host side of load_calibration and init_aux
    const Model& m = *s->model;
    s->has_gain = m.has_gain;
    s->gain = m.gain;
    s->bg_len = m.background.size();
    s->counter = 0;
    s->initialized = true;
  }
  hipLaunchKernelGGL(k_reset_tiles, dim3((b.launched + 255) / 256), dim3(256), 0, st, b.tiles.as<TileDesc>(),
                     b.launched);
  HIP_CHECK(hipGetLastError());
  // canonicalize_state after initialisation (runner.rs:122)
  if (b.gauge_count) launch_gauge(b.gauges.as<GaugeTile>(), b.gauge_count, st);
}

// 50: against 10, one dispatch's tail (tiles that finish before the slowest
// idle their CUs) is paid a fifth as often: 256 tiles of 225 LEDs x 128^2
// ran 636 -> 641 tiles/s.
constexpr uint32_t kDefaultSweepsPerLaunch = 50;

// Sweeps one graph replay runs: the schedule's, or one when the tiles'
// gauge must follow each sweep.
uint32_t sweeps_per_launch(const fpmhip_batch& b) {
  if (b.gauge_count) return 1;
  return b.schedule.sweeps_per_launch ? b.schedule.sweeps_per_launch : kDefaultSweepsPerLaunch;
}

// `sweeps` sweeps as graph replays of up to `per_launch` sweeps each (the
// sweep count of a launch is set in the device descriptors before it), each
// replay bracketed by HIP events on the batch stream. mean / mx: per-sweep
// time (launch time / its sweeps), averaged / the largest.
void replay(fpmhip_batch& b, uint64_t sweeps, double* mean, double* mx, uint32_t per_launch = 1) {
  hipStream_t st = b.ctx->stream;
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  double total = 0.0, worst = 0.0;
  uint32_t current = 0;
  for (uint64_t done = 0; done < sweeps;) {
    const uint32_t k = (uint32_t)std::min<uint64_t>(std::max<uint32_t>(per_launch, 1), sweeps - done);
    if (k != current) {
      hipLaunchKernelGGL(k_set_sweeps, dim3((b.launched + 255) / 256), dim3(256), 0, st, b.tiles.as<TileDesc>(),
                         b.launched, k);
      HIP_CHECK(hipGetLastError());
      current = k;
    }
    HIP_CHECK(hipEventRecord(e0, st));
    HIP_CHECK(hipGraphLaunch(b.exec, st));
    HIP_CHECK(hipEventRecord(e1, st));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    total += ms;
    worst = std::max(worst, (double)ms / k);
    done += k;
  }
  if (current != 1) {  // leave single-sweep launches in place
    hipLaunchKernelGGL(k_set_sweeps, dim3((b.launched + 255) / 256), dim3(256), 0, st, b.tiles.as<TileDesc>(),
                       b.launched, 1u);
    HIP_CHECK(hipGetLastError());
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *mean = sweeps ? total / (double)sweeps : 0.0;
  *mx = worst;
}

void check_tile_status(int32_t status) {
  if (status == 2) fail(FPMHIP_DEVICE_ERROR, "a cooperative tile group stalled at a group barrier");
  int code = FPMHIP_NUMERICAL;
  if (const char* msg = gauge_failure(status, &code)) fail(code, msg);
  if (status) fail(FPMHIP_NUMERICAL, "mPIE momentum produced a non-finite object or velocity");
}

void validate_schedule(const fpmhip_batch_schedule& s) {
  if (s.group_members > kMaxGroupMembers || (s.group_members > 1 && s.group_members < kGroupMembers))
    fail(FPMHIP_INVALID_PARAMETER, "group_members must be 0, 1 or 8..32", "group_members");
  if (s.largest_first > 1) fail(FPMHIP_INVALID_PARAMETER, "largest_first must be 0 or 1", "largest_first");
  if (s.sweeps_per_launch > 1000)
    fail(FPMHIP_INVALID_PARAMETER, "sweeps_per_launch must be in 0..1000", "sweeps_per_launch");
  const uint32_t sizes[4] = {32, 64, 128, 256};
  for (int k = 0; k < 4; ++k) {
    if (s.radix[k] == 0) continue;
    uint32_t choices[3];
    const uint32_t nc = sweep_radix_choices(sizes[k], choices);
    bool ok = false;
    for (uint32_t i = 0; i < nc; ++i) ok |= choices[i] == s.radix[k];
    if (!ok) fail(FPMHIP_INVALID_PARAMETER, "radix is not a plan of the fused sweep for that tile size", "radix");
  }
}

}  // namespace

extern "C" {

int fpmhip_batch_create(fpmhip_context* c, fpmhip_session* const* sessions, uint64_t count, fpmhip_batch** out) {
  return guarded([&] {
    if (!c || !sessions || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "batch");
    if (count == 0 || count > 65535) fail(FPMHIP_INVALID_PARAMETER, "tile count must be in 1..65535", "count");
    HIP_CHECK(hipSetDevice(c->device));
    auto b = std::make_unique<fpmhip_batch>();
    b->ctx = c;
    int kind = -1;
    for (uint64_t i = 0; i < count; ++i) {
      fpmhip_session* s = sessions[i];
      if (!s) fail(FPMHIP_INVALID_PARAMETER, "null session", "sessions");
      if (s->ctx != c) fail(FPMHIP_INVALID_PARAMETER, "sessions must share the batch context", "sessions");
      const char* why = "";
      if (!sweep_supports(*s, &why)) fail(FPMHIP_UNSUPPORTED, why);
      const Model& m = *s->model;
      if (kind >= 0 && kind != s->alg.kind) fail(FPMHIP_INVALID_PARAMETER, "tiles must share one algorithm", "sessions");
      if (!b->sessions.empty() && b->sessions[0]->cfg.has_pupil_step != s->cfg.has_pupil_step)
        fail(FPMHIP_INVALID_PARAMETER, "tiles must all recover their pupil or none", "sessions");
      kind = s->alg.kind;
      b->nmax = std::max(b->nmax, m.nh);
      b->fmax = std::max<uint64_t>(b->fmax, m.F);
      b->sessions.push_back(s);
    }
    if (sweep_lds_bytes(b->nmax) > (size_t)c->props.sharedMemPerBlockOptin && c->props.sharedMemPerBlockOptin > 0)
      fail(FPMHIP_UNSUPPORTED, "tile does not fit the workgroup's LDS");
    make_sweep_twiddles(b->tw, c->stream);
    build_tiles(*b, 64);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    *out = b.release();
  });
}

int fpmhip_batch_destroy(fpmhip_batch* b) {
  return guarded([&] {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    delete b;
  });
}

int fpmhip_batch_run(fpmhip_batch* b, uint64_t iterations, double* trace) {
  return guarded([&] {
    if (!b) fail(FPMHIP_INVALID_PARAMETER, "null batch", "batch");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const uint32_t count = (uint32_t)b->sessions.size();
    if (iterations > b->trace_cap) build_tiles(*b, (uint32_t)iterations);
    initialize_tiles(*b);
    replay(*b, iterations, &b->ms_mean, &b->ms_max, sweeps_per_launch(*b));
    b->sweeps = iterations;
    std::vector<char> per(count * 16);
    HIP_CHECK(hipMemcpyAsync(per.data(), b->per_tile.ptr, per.size(), hipMemcpyDeviceToHost, st));
    std::vector<double> tr((size_t)count * b->trace_cap);
    HIP_CHECK(hipMemcpyAsync(tr.data(), b->traces.ptr, tr.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < count; ++i) {
      const char* base = per.data() + (size_t)i * 16;
      check_tile_status(*reinterpret_cast<const int32_t*>(base + 12));
      b->sessions[i]->counter = *reinterpret_cast<const uint64_t*>(base);
      if (trace)
        for (uint64_t it = 0; it < iterations; ++it) trace[i * iterations + it] = tr[(size_t)i * b->trace_cap + it];
    }
  });
}

int fpmhip_batch_get_schedule(const fpmhip_batch* b, fpmhip_batch_schedule* out) {
  return guarded([&] {
    if (!b || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "batch");
    *out = b->schedule;
  });
}

int fpmhip_batch_set_schedule(fpmhip_batch* b, const fpmhip_batch_schedule* schedule) {
  return guarded([&] {
    if (!b || !schedule) fail(FPMHIP_INVALID_PARAMETER, "null argument", "batch");
    validate_schedule(*schedule);
    HIP_CHECK(hipSetDevice(b->ctx->device));
    b->schedule = *schedule;
    build_tiles(*b, b->trace_cap);
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  });
}

int fpmhip_batch_tune(fpmhip_batch* b, uint32_t trial_sweeps, fpmhip_batch_tuning* report) {
  return guarded([&] {
    if (!b) fail(FPMHIP_INVALID_PARAMETER, "null batch", "batch");
    if (trial_sweeps == 0 || trial_sweeps > 1000)
      fail(FPMHIP_INVALID_PARAMETER, "trial_sweeps must be in 1..1000", "trial_sweeps");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    uint32_t timed = 0;
    // one candidate: rebuild the launch, start every tile from its initial
    // state, one untimed sweep, then the mean of `trial_sweeps` sweeps
    auto measure = [&](const fpmhip_batch_schedule& cand) {
      b->schedule = cand;
      build_tiles(*b, b->trace_cap);
      initialize_tiles(*b);
      double mean = 0.0, mx = 0.0;
      replay(*b, 1, &mean, &mx);
      replay(*b, trial_sweeps, &mean, &mx);
      ++timed;
      return mean;
    };
    fpmhip_batch_schedule best = b->schedule;
    const double first = measure(best);
    double best_ms = first;
    auto consider = [&](fpmhip_batch_schedule cand) {
      const double ms = measure(cand);
      if (ms < 0.98 * best_ms) {  // keep a change only for a clear gain
        best = cand;
        best_ms = ms;
      }
    };
    bool big = false, present[4] = {false, false, false, false};
    for (const TileDesc& t : b->host_tiles) {
      big |= t.n > kFusedLdsMaxN;
      present[t.n == 32 ? 0 : t.n == 64 ? 1 : t.n == 128 ? 2 : 3] = true;
    }
    // coordinate descent: group size, radix per tile size, launch order
    if (big)
      for (uint32_t g : {0u, 1u, kGroupMembers})
        if (g != best.group_members) {
          fpmhip_batch_schedule c = best;
          c.group_members = g;
          consider(c);
        }
    const uint32_t sizes[4] = {32, 64, 128, 256};
    for (int k = 0; k < 4; ++k) {
      if (!present[k]) continue;
      uint32_t choices[3];
      const uint32_t nc = sweep_radix_choices(sizes[k], choices);
      const uint32_t cur = best.radix[k] ? best.radix[k] : choices[0];
      for (uint32_t i = 0; i < nc; ++i)
        if (choices[i] != cur) {
          fpmhip_batch_schedule c = best;
          c.radix[k] = choices[i];
          consider(c);
        }
    }
    {
      fpmhip_batch_schedule c = best;
      c.largest_first = best.largest_first ? 0 : 1;
      consider(c);
    }
    b->schedule = best;
    build_tiles(*b, b->trace_cap);
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    if (report) {
      report->candidates = timed;
      report->first_ms = first;
      report->best_ms = best_ms;
    }
  });
}

namespace {

// state_object of every tile (result.rs:397-413) into its slot: one shift
// launch for all tiles and one inverse FFT per distinct R.
void batch_objects_compute(fpmhip_batch& b) {
  HIP_CHECK(hipSetDevice(b.ctx->device));
  hipStream_t st = b.ctx->stream;
  launch_object_batch_shift(b.init_tiles.as<InitTile>(), (uint32_t)b.sessions.size(), b.init_mmax, st);
  for (const fpmhip_batch::SlotGroup& g : b.slot_groups) fft2_f64(*g.plan, g.base, g.count, true, st);
}

void batch_objects_into(fpmhip_batch& b, double* out) {
  batch_objects_compute(b);
  uint64_t off = 0;
  for (size_t i = 0; i < b.sessions.size(); ++i) {
    const uint64_t M = b.sessions[i]->model->M();
    HIP_CHECK(hipMemcpyAsync(out + 2 * off, b.slot_of[i], M * sizeof(double2), hipMemcpyDeviceToHost, b.ctx->stream));
    off += M;
  }
  HIP_CHECK(hipStreamSynchronize(b.ctx->stream));
}

}  // namespace

int fpmhip_batch_objects(fpmhip_batch* b, double* out) {
  return guarded([&] {
    if (!b || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "batch");
    batch_objects_into(*b, out);
  });
}

int fpmhip_batch_objects_device(fpmhip_batch* b, const double** objects) {
  return guarded([&] {
    if (!b || !objects) fail(FPMHIP_INVALID_PARAMETER, "null argument", "batch");
    batch_objects_compute(*b);
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    for (size_t i = 0; i < b->sessions.size(); ++i) objects[i] = reinterpret_cast<const double*>(b->slot_of[i]);
  });
}

int fpmhip_batch_objects_pinned(fpmhip_batch* b, const double** out) {
  return guarded([&] {
    if (!b || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "batch");
    size_t bytes = 0;
    for (fpmhip_session* s : b->sessions) bytes += s->model->M() * sizeof(double2);
    if (b->pinned_bytes < bytes) {
      HIP_CHECK(hipSetDevice(b->ctx->device));
      if (b->pinned) HIP_CHECK(hipHostFree(b->pinned));
      b->pinned = nullptr;
      b->pinned_bytes = 0;
      void* p = nullptr;
      if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p)
        fail(FPMHIP_OUT_OF_MEMORY, "page-locked result buffer could not be allocated");
      b->pinned = static_cast<double*>(p);
      b->pinned_bytes = bytes;
    }
    batch_objects_into(*b, b->pinned);
    *out = b->pinned;
  });
}

int fpmhip_batch_timing(const fpmhip_batch* b, double* mean, double* mx, uint64_t* sweeps) {
  return guarded([&] {
    if (!b) fail(FPMHIP_INVALID_PARAMETER, "null batch", "batch");
    if (mean) *mean = b->ms_mean;
    if (mx) *mx = b->ms_max;
    if (sweeps) *sweeps = b->sweeps;
  });
}

}  // extern "C"
