// sdtfe_internal.h -- device data layout and kernels shared by the
// translation units of libsdtfe_amd.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

// Pure arithmetic helpers are host-callable too, so the same walk code can
// be exercised on the CPU.
#define SDTFE_HD __host__ __device__

namespace sdtfe {

// ---------------------------------------------------------------- layout

// Walk record of the FP64 walk, one per tet, 144 B.  Everything one step of
// the walk reads sits in one record, so a step costs a single dependent load
// (the record named by the previous record's neighbour slot) issued as nine
// 16-byte loads.
//   nb[f]  tet across the face opposite vertex f (-1 on the hull)
//   v      A, B, C, D coordinates (x, y, z each)
//   rho_a  density at vertex A; g the tet's density gradient
struct alignas(16) Rec64 {
  int nb[4];
  double v[12];
  double rho_a;
  double g[3];
};
static_assert(sizeof(Rec64) == 144, "Rec64 layout");

// FP64 walk record of a tet seen from one of its faces, one per (tet t,
// face f), 144 B read as nine 16-byte loads: what a ray entering t through
// face f (the face opposite vertex f) needs to cross t and leave it.
//   n     vertex f (N): x, y, z
//   p     face f's three vertices P0, P1, P2 (x, y, z each) in the order the
//         reference forms that face's barycentric height (src/dtfe.c:67-213):
//         f0 BDC: D, C, B; f1 ACD: C, D, A; f2 ADB: B, A, D; f3 ABC: A, B, C
//   rho_a, g  the density at vertex A and the tet's gradient
//   next  for k = 0..2, the face opposite P_k (one of the faces meeting at
//         N): 4 x the tet across it + that face's index there, -1 on the
//         hull, -2 when the mesh lacks the back link
//   code  bits 0-1: which of face f's sides its denominator adds last;
//         bits 2+4k (2 bits): rotation of the face opposite P_k (its order
//         in the reference is (P_{k+2}, P_{k+1}, N) rotated left by this);
//         bits 4+4k (2 bits): which of that face's sides is added last;
//         bits 14+2i (2 bits): position of P_i in the reference order of
//         the same face in the tet across f; bit 30: no such position (the
//         ray's sides are then formed afresh); bit 31: a ray entering here
//         is handed to the walk over Rec64 (missing or ambiguous links, or
//         the two tets do not see the face with opposite orientations)
struct alignas(16) FRec64 {
  double n[3];
  double p[9];
  double rho_a, g[3];
  int next[3];
  unsigned code;
};
static_assert(sizeof(FRec64) == 144, "FRec64 layout");
constexpr unsigned kFRecFresh = 1u << 30;
constexpr unsigned kFRecHand = 1u << 31;

// Mixed-precision walk record, 96 B: neighbours, vertex A in FP64 and the
// edges B - A, C - A, D - A plus the four vertex densities in FP32.  Offsets
// from a ray are built from A in FP64, so FP32 rounding stays a fraction of
// the local tet size.
struct alignas(16) Rec32 {
  int nb[4];
  double a[3];
  float e[9];
  float rho[4];
  float pad;
};
static_assert(sizeof(Rec32) == 96, "Rec32 layout");

// Face record of the mixed walk, one per (tet, face), 32 B read as two
// 16-byte loads: where a ray that leaves tet t through face f (the face
// opposite vertex f) goes.  The walk carries the face's three vertices, so
// the next tet only adds its fourth vertex, stored here.
//   q     the new vertex: x and y each as an FP32 pair (hi = x rounded,
//         lo = x - hi rounded), z minus the mean z of face f's three vertices
//         ((z1 + z2) + z3) / 3 and the density, in FP32
//   nb    tet across the face (-1 on the hull)
//   code  one byte per label of t, each value at bits 3-4 of its byte:
//         byte m (m != f) holds 8 x the label in nb of t's vertex m, byte f
//         holds 8 x the face's index in nb (nb's vertex opposite it, i.e.
//         the new vertex), so that the walk, which keeps labels as
//         8 x label, relabels a vertex, and finds the new vertex's label,
//         with one bit-field extract at offset 8 x label; bit 31
//         (kFaceStop): the ray stops here (nb < 0, or the mesh lacks the
//         back link)
struct alignas(16) FaceRec {
  float qxh, qxl, qyh, qyl;
  int nb;
  unsigned code;
  float dqz;
  float rho;
};
static_assert(sizeof(FaceRec) == 32, "FaceRec layout");
constexpr unsigned kFaceStop = 1u << 31;
// Face records of up to this many tets are addressed with 32-bit offsets.
constexpr int64_t kFaceRecNarrowTets = (int64_t)1 << 25;

constexpr int kDeferWindow = 256;  // handed-over samples one entry can list

// A pixel whose mixed walk handed sample-rays to the FP64 walk (64 B); the
// second pass of a mixed map finishes it.
struct alignas(16) Deferred {
  double sum;         // over the finished samples below `tail`, in sample order
  uint64_t pix;       // pixel index in its map
  uint64_t mask[4];   // bit k: sample base + k is walked in FP64
  int base;           // first handed-over sample
  int tail;           // samples tail .. samples-1 are all walked in FP64
  int n_ok;           // finished samples that crossed a tet
  int item;           // batch item (0 for a single map)
};
static_assert(sizeof(Deferred) == 64, "Deferred layout");

// Queue of a launch: count may exceed cap, which the host reads as overflow
// (the launch is then repeated with a larger queue).
struct DeferQueue {
  Deferred *q;
  int *count;
  int cap;
};

// Entry lookup: forward-hull triangles binned on a uniform 2D grid.
struct EntryGrid {
  const double *tri_xy;   // F x 6: a.x a.y b.x b.y c.x c.y
  const int *tri_tet;     // F
  const int *tri_rid;     // F: 4 x tet + face (the face record a ray enters by)
  const int *cell_off;    // nx*ny + 1
  const int *cell_tri;    // triangle ids per cell
  double x0, y0, x1, y1;  // grid extent (bounding box of the triangles)
  double sx, sy;          // cells per unit length
  int nx, ny, n_tri;
};

// Everything the walk needs to place samples (src/dtfe.c:579-589, 613-622).
struct MapGeom {
  double dxy, off_x, off_y, off_z, lz, scale, soff;
  double org_x, org_y, org_z;  // origin of the FP32 coordinates
  double ray_span;             // max |ray x or y - origin| over the launch
  uint64_t seed;
  int grid, samples, max_steps;
};

// A volumetric grid (src/dtfe.c:667-735): grid^3 cells, ss x ss rays per
// column and ss fine z points per cell.
struct VolGeom {
  double dxy, ds;              // cell width, ray spacing (dxy / ss)
  double off_x, off_y, off_z;  // grid corner
  double lz;                   // ray length (the cube's side)
  double dzs;                  // fine z spacing (lz / grid / ss)
  double ss_weight;            // 1 / ss^3
  double pert_step;            // (dxy / 2) / PERTURB_MAX
  uint64_t seed;
  int grid, ss, nzs, max_steps;
};

enum StatIdx {
  ST_SAMPLES = 0,
  ST_MISSED,
  ST_FAILED,
  ST_NO_ISECT,
  ST_PERT_EXHAUSTED,
  ST_NO_ISECT_PERT,
  ST_PERT_STEPS,
  ST_TET_STEPS,
  ST_STEP_CAP,
  ST_FP64_REDO,
  ST_COUNT
};

// ---------------------------------------------------------------- stream

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  // SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014).
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

// Key of one (seed, pixel, sample); draws are draw31(key, d).
__host__ __device__ inline uint64_t sample_key(uint64_t seed, uint64_t pixel, uint32_t sample) {
  uint64_t k = mix64(seed + kGamma);
  k = mix64(k ^ pixel);
  return mix64(k ^ (uint64_t)sample);
}

__host__ __device__ inline uint32_t draw31(uint64_t key, uint32_t draw) {
  return (uint32_t)(mix64(key + (uint64_t)(draw + 1u) * kGamma) >> 33);
}

// --------------------------------------------------------- wave helpers

__device__ inline void add_stat(unsigned long long *stats, int idx, unsigned long long v) {
  // one atomic per wave: sum over the wave's lanes first
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(stats + idx, v);
}

// --------------------------------------------------------- launch helpers

// Blocks are dealt to the 8 XCDs round-robin; remap so that each XCD takes
// whole rows of `row` consecutive tiles in turn (row r on XCD r mod 8):
// neighbouring tiles of a row, which walk the same tets, share an L2, and
// each XCD's rows spread over the whole map, so that its share of the work
// follows the map's mean density (contiguous eighths of the map left some
// XCDs with more of the dense halos than others).
__device__ inline int xcd_tile(int block, int n_blocks, int row) {
  const int n_xcd = 8, group = n_xcd * row;
  const int full = row > 0 ? (n_blocks / group) * group : 0;
  if (block >= full) return block;
  const int x = block % n_xcd, k = block / n_xcd;
  return (k / row) * group + x * row + k % row;
}

}  // namespace sdtfe
