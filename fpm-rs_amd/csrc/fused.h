// fused.h — descriptor and launcher of the fused per-tile sweep kernel.
#pragma once

#include <vector>

#include "session.h"

namespace fpmhip {

// Everything one workgroup needs to run one sweep of one tile.
struct TileDesc {
  double2* spec;          // [R] FP64 centred object spectrum
  float2* vel;            // [R] mPIE velocity (or null), complex64 storage
  float2* anc;            // [R] mPIE anchor (or null), complex64 storage
  float2* pupil;          // [N] complex64 pupil (updated in place by EPRY pupil recovery)
  const uint32_t* extent; // [n] per-row support of the pupil: columns [lo, hi) as lo | hi << 16
  const uint32_t* dest;   // live-pixel lists, tiles <= 128^2: [0] insert destinations, [1] support
                          // strips, then (n+1)^2 slots of destinations a | b << 16 and the strips
                          // r | c << 10 | mask << 20 (SessionBuffers::dest)
  float2* field;          // [N] global field scratch (tiles above kFusedLdsMaxN)
  float2* patch;          // [N] the patch of the support pixels (EPRY pupil recovery)
  double2* pupil64;       // [N] FP64 master of the recovered pupil (EPRY pupil recovery)
  uint64_t* bounds;       // [ceil(rh/8)][bcols] |O|^2 upper bound per 8 x 8 spectrum block, or null
                          // (EPRY pupil recovery: exact before each sweep, raised by the inserts)
  const void* frames;     // [F*N] uint16 counts or f64 frames
  const uint8_t* masks;   // [F*N] or null
  const double* weights;  // [F]
  const double* gains;    // [F]
  const double* bg;       // [N] or [F*N] or null
  uint64_t bg_len;
  const SrcParam* src;    // [F] per-source crop parameters
  const uint32_t* order;  // [F] frame order, or null for sequential
  const uint32_t* mspan;  // momentum chunks (null: the event visits every object pixel)
  double* trace;          // [trace_len] iteration objectives
  double* frame_loss;     // [F] objective of each order position, last sweep
  uint64_t* counter;      // mPIE effective-frame counter (persists across sweeps)
  uint32_t* iteration;    // sweeps completed
  int32_t* status;        // [0] 1: non-finite momentum, 2: cooperative group stalled
  // cooperative tile group (tiles above kFusedLdsMaxN): `members` workgroups
  // share one tile; this descriptor is workgroup `member` of the group
  uint64_t* group;        // [0] barrier arrivals, [1] barriers completed before this launch, [2] stall flag
  double2* partial;       // [members] per-member objective sums (loss, count)
  uint32_t member, members;
  uint32_t plan;          // first-stage radix of the LDS line transforms (0: default)
  uint32_t trace_len;
  uint32_t sweeps;        // sweeps this launch runs (>= 1)
  uint32_t n, rh, rw, F;
  uint32_t interval;
  int32_t is_u16, denom, momentum, loss;
  int32_t recover_pupil;  // EPRY pupil recovery (common.rs:271-300; tiles up to kFusedLdsMaxN)
  uint32_t bcols;         // blocks per spectrum row of `bounds`
  double object_step, alpha, eps, max_p, friction, feedback;
  double pupil_step;
};

// Offset (in float2 units) of the forward/inverse twiddle tables of length n
// inside the packed table: [32 fwd, 32 inv, 64 fwd, 64 inv, ..., 512 inv].
__host__ __device__ inline uint32_t tw_offset(uint32_t n) {
  return n == 32 ? 0u : n == 64 ? 64u : n == 128 ? 192u : n == 256 ? 448u : 960u;
}
constexpr uint32_t kTwTotal = 2 * (32 + 64 + 128 + 256 + 512);
// Largest tile whose whole field is LDS-resident; larger tiles (up to
// kFusedMaxN) keep it in TileDesc::field.
constexpr uint32_t kFusedLdsMaxN = 128;
constexpr uint32_t kFusedMaxN = 512;

size_t sweep_lds_bytes(uint32_t nmax);
// Descriptor of one MIXED_F32 session's tile; per16 points at 16 device bytes
// (counter u64, iteration u32, status i32).
void fill_tile_desc(TileDesc& t, fpmhip_session& s, double* trace, double* frame_loss, char* per16,
                    uint32_t trace_len);
// Launch order of a sweep's tiles: largest first (stable), so the longest
// workgroups are dispatched first and spread over the XCDs, or session
// order. Unless `members` is 1, tiles above kFusedLdsMaxN become cooperative
// groups as far as the whole grid stays within `grid_limit` resident
// workgroups, tiles with more LEDs first; the others run on one workgroup.
// Group sizes: group_members(n) (512^2: 32); at 256^2 `members` (8..32), or
// with `members` 0 group_members(256) plus an equal share of the resident
// room left over, up to kMaxGroupMembers. A group's members are placed on
// one XCD's dispatch slots (positions device_xccs() apart). radix[4]:
// first-stage radix per tile size
// 32, 64, 128, 256 (0: default; null: keep the descriptors' plans; 512^2
// tiles have one plan).
std::vector<TileDesc> order_tiles(const std::vector<TileDesc>& tiles, uint32_t members, const uint32_t* radix,
                                  bool largest_first, uint32_t grid_limit);
// XCDs (accelerator complex dies) of the current device; workgroups are
// dispatched to them round-robin.
uint32_t device_xccs();
// First-stage radix choices of the fused sweep for an n x n tile; the first
// entry is the default. Returns the number of choices (0: size unsupported).
uint32_t sweep_radix_choices(uint32_t n, uint32_t out[3]);
// Lines per chunk of a large tile's field (256^2: 32, 512^2: 16, so that
// a chunk fits LDS), and the workgroups of its cooperative group (one chunk
// each). kGroupMembers is the group size of a 256^2 tile, the value that
// turns cooperative groups on in a batch schedule.
__host__ __device__ constexpr uint32_t big_chunk(uint32_t n) { return n >= 512 ? 16u : 32u; }
__host__ __device__ constexpr uint32_t group_members(uint32_t n) { return n / big_chunk(n); }
constexpr uint32_t kGroupMembers = group_members(256);
constexpr uint32_t kMaxGroupMembers = group_members(kFusedMaxN);
// Bytes of a session's cooperative-group state (counters and per-member sums).
size_t group_state_bytes();
// Largest grid whose workgroups are all resident at once (one per CU at the
// sweep kernel's LDS size).
uint32_t sweep_coresident_limit(uint32_t nmax);
// Packed FP32 twiddle tables of the sweep kernel (kTwTotal float2).
void make_sweep_twiddles(DeviceBuffer& out, hipStream_t stream);
// Whether the fused sweep can run this session (else *why says why not).
bool sweep_supports(const fpmhip_session& s, const char** why);
// Raises the kernel's dynamic LDS limit (outside any graph capture).
void prepare_sweep(uint32_t nmax);
// Launches the sweep over `count` descriptors; a grid holding cooperative
// groups is launched cooperatively (all workgroups co-resident).
// `epry`: every tile recovers its pupil (EPRY), through the kernel variant
// that does (tiles up to kFusedLdsMaxN).
void launch_sweep(const TileDesc* tiles_dev, uint32_t count, const float2* tw_all, uint32_t nmax, bool cooperative,
                  hipStream_t stream, bool epry = false);
// Side of the spectrum blocks of TileDesc::bounds.
constexpr uint32_t kBoundBlock = 8;
// Recomputes TileDesc::bounds exactly (max |O|^2 per block) for every
// descriptor that has them; one workgroup per descriptor.
void launch_patch_bounds(const TileDesc* tiles_dev, uint32_t count, hipStream_t stream);

}  // namespace fpmhip
