// Work-tile descriptors shared by the temporal-blocked tile kernels and their
// host-side planner.
#pragma once

#include <cstdint>

namespace npbx {

enum GridKind : int32_t { kJacobi2D = 0, kHeat3D = 1 };

// One grid in one round: `depth` sweeps from level L0 (interior in X) to
// level L0+depth (interior written to Y).  Global-boundary values of level L
// are A's when L is even and B's when L is odd (the reference's ping-pong
// leaves each array's boundary untouched).
struct GridRound {
  void* A;
  void* B;
  const void* X;
  void* Y;
  int32_t n;        // side length
  int32_t kind;     // GridKind
  int32_t parity0;  // L0 & 1
  int32_t depth;    // 1 .. T
  // tiling (fixed per grid for a given T)
  int32_t tiles_k;  // tiles along the contiguous axis
  int32_t tiles_j;  // tiles along the row axis
  int32_t seg;      // heat: output planes per tile along i
  int32_t rows;     // heat: rows per wave (a tile is 4*rows rows)
  // FP32 heat: packed boundary columns of A and B (npbx_pack_cols layout
  // [array][plane][row][2]), or null to read them from A and B directly
  const void* cols;
};

// Up to kMaxBatch grid rounds passed by value as one launch's kernel
// argument (< 4 KB); cum[g] is the number of tiles of rounds 0..g-1, so tile
// t of the launch belongs to the round g with cum[g] <= t < cum[g+1].
constexpr int kMaxBatch = 48;
struct RoundBatch {
  int32_t count;
  int32_t cum[kMaxBatch + 1];
  GridRound round[kMaxBatch];
};

// Per-launch queue state, self-resetting: `next` hands out tile indices and
// `done` counts workgroups that found the queue empty; the last one resets
// both to zero for the next launch.
struct QueueState {
  int32_t next;
  int32_t done;
  int32_t pad[62];  // one 256-byte line per queue
};

}  // namespace npbx
