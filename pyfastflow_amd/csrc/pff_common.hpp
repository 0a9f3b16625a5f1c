// pff_common.hpp -- device-side grid topology and packed keys shared by every
// kernel of libpff_hip.so.  gfx950 (CDNA4) only.
//
// Layout: 1-D row-major, i = row*nx + col, row 0 = north/top, the layout of
// pyfastflow/flow/neighbourer_flat.py:22-50.  Directions k: 0 top (i-nx),
// 1 left (i-1), 2 right (i+1), 3 bottom (i+nx) (neighbourer_flat.py:123-141).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pff {

struct GridD {
  int32_t nx, ny, N;
  float dx;
  int32_t mode;         // 0 normal, 1 periodic_EW, 2 periodic_NS, 3 custom, 4 tile canvas
  const uint8_t* B;     // per-cell codes (device): custom codes in mode 3, canvas codes in mode 4
  // mode 4 only: f42[b] != 0 marks the basin b that holds cell 42 of its own
  // tile (the index the "no outlet" sentinel decodes to); nullptr elsewhere
  const uint8_t* F42 = nullptr;
};

// Edge class of which_edge (neighbourer_flat.py:75-110) from (row, col); valid
// for nx >= 2 and ny >= 2, which the host checks.
__device__ __forceinline__ int edge_class(const GridD& g, int row, int col) {
  if (row == 0) return col == 0 ? 1 : (col < g.nx - 1 ? 2 : 3);
  if (row < g.ny - 1) return col == 0 ? 4 : (col == g.nx - 1 ? 5 : 0);
  return col == 0 ? 6 : (col == g.nx - 1 ? 8 : 7);
}

// validate_link_n / _pew / _pns (neighbourer_flat.py:146-159, 228-237, 305-314)
__device__ __forceinline__ bool valid_n(int e, int k) {
  if (e == 0) return true;
  if (k == 0) return e > 3;
  if (k == 1) return !(e == 1 || e == 4 || e == 6);
  if (k == 2) return !(e == 3 || e == 5 || e == 8);
  return e < 6;
}
__device__ __forceinline__ bool valid_pew(int e, int k) {
  if (e == 0) return true;
  if (k == 0) return e > 3;
  if (k == 3) return e < 6;
  return true;
}
__device__ __forceinline__ bool valid_pns(int e, int k) {
  if (e == 0) return true;
  if (k == 1) return !(e == 1 || e == 4 || e == 6);
  if (k == 2) return !(e == 3 || e == 5 || e == 8);
  return true;
}

// neighbour(i, k) for boundary mode MODE (neighbourer_flat.py:542-569).
// Returns -1 for an invalid link.
template <int MODE>
__device__ __forceinline__ int neighbour(const GridD& g, int i, int row, int col, int k) {
  const int nx = g.nx, ny = g.ny;
  if constexpr (MODE == 0) {  // normal: links leaving the grid are invalid (:146-180)
    switch (k) {
      case 0: return row > 0 ? i - nx : -1;
      case 1: return col > 0 ? i - 1 : -1;
      case 2: return col < nx - 1 ? i + 1 : -1;
      default: return row < ny - 1 ? i + nx : -1;
    }
  } else if constexpr (MODE == 1) {  // periodic east-west (:193-256)
    switch (k) {
      case 0: return row > 0 ? i - nx : -1;
      case 1: return col > 0 ? i - 1 : i + nx - 1;
      case 2: return col < nx - 1 ? i + 1 : i - nx + 1;
      default: return row < ny - 1 ? i + nx : -1;
    }
  } else if constexpr (MODE == 2) {  // periodic north-south (:270-334)
    switch (k) {
      case 0: return row > 0 ? i - nx : i + nx * (ny - 1);
      case 1: return col > 0 ? i - 1 : -1;
      case 2: return col < nx - 1 ? i + 1 : -1;
      default: return row < ny - 1 ? i + nx : i - nx * (ny - 1);
    }
  } else if constexpr (MODE == 4) {
    // tile canvas: several 'normal' grids packed into one array, separated by
    // padding.  Codes: 0 padding, 1 tile interior, 3 tile edge.  A link is
    // valid when both ends are tile cells, which inside a tile is exactly the
    // normal-mode rule since every tile is framed by padding or the canvas
    // border.
    if (g.B[i] == 0) return -1;
    int node;
    switch (k) {
      case 0: node = row > 0 ? i - nx : -1; break;
      case 1: node = col > 0 ? i - 1 : -1; break;
      case 2: node = col < nx - 1 ? i + 1 : -1; break;
      default: node = row < ny - 1 ? i + nx : -1; break;
    }
    return (node < 0 || g.B[node] == 0) ? -1 : node;
  } else {  // custom per-node codes 0/1/3/7/9 (:360-462)
    const uint8_t tb = g.B[i];
    if (tb == 0) return -1;
    const int e = edge_class(g, row, col);
    bool ok;
    if (tb != 9 || e == 0) ok = valid_n(e, k);
    else if (e <= 3 || e >= 6) ok = valid_pns(e, k);
    else ok = valid_pew(e, k);
    if (!ok) return -1;
    int node;
    if (tb == 9) {
      switch (k) {
        case 0: node = row > 0 ? i - nx : i + nx * (ny - 1); break;
        case 1: node = col > 0 ? i - 1 : i + nx - 1; break;
        case 2: node = col < nx - 1 ? i + 1 : i - nx + 1; break;
        default: node = row < ny - 1 ? i + nx : i - nx * (ny - 1); break;
      }
    } else {
      switch (k) {
        case 0: node = i - nx; break;
        case 1: node = i - 1; break;
        case 2: node = i + 1; break;
        default: node = i + nx; break;
      }
    }
    if (node < 0 || node >= g.N) return -1;
    return g.B[node] == 0 ? -1 : node;
  }
}

// can_leave_domain (neighbourer_flat.py:183-185, 259-262, 337-340, 465-467)
template <int MODE>
__device__ __forceinline__ bool can_leave(const GridD& g, int i, int row, int col) {
  if constexpr (MODE == 0) return row == 0 || row == g.ny - 1 || col == 0 || col == g.nx - 1;
  else if constexpr (MODE == 1) return row == 0 || row == g.ny - 1;
  else if constexpr (MODE == 2) return col == 0 || col == g.nx - 1;
  else if constexpr (MODE == 4) return g.B[i] != 1;  // tile edges, and padding (never a pit)
  else return g.B[i] == 3;
}

// Packed (f32, i32) key, f32_i32_struct.py:14-103.  As a signed 64-bit
// integer it orders by the float first (with -0 < +0), then the index.
__device__ __forceinline__ int64_t pack_key(float f, int32_t idx) {
  const uint32_t u = __float_as_uint(f);
  const uint32_t hi = (u & 0x80000000u) ? (0x80000000u | (~u & 0x7FFFFFFFu)) : u;
  return (int64_t)(((uint64_t)hi << 32) | (uint64_t)(uint32_t)idx);
}
__device__ __forceinline__ float key_float(int64_t key) {
  const uint32_t hi = (uint32_t)((uint64_t)key >> 32);
  const uint32_t u = (hi & 0x80000000u) ? (0x80000000u | (~hi & 0x7FFFFFFFu)) : hi;
  return __uint_as_float(u);
}
__device__ __forceinline__ int32_t key_index(int64_t key) { return (int32_t)(uint32_t)((uint64_t)key & 0xFFFFFFFFull); }

// pack_float_index(1e8, 42): the reference's "no saddle / no outlet" sentinel
// (lakeflow.py:127)
constexpr int64_t INVALID_KEY = (int64_t)0x4CBEBC200000002Aull;

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs (each with its own L2) in linear block order; this remaps the linear
// id so that XCD k processes the k-th contiguous eighth of the 2-D grid in
// row-major order.  A stencil block then finds the rows of its vertical
// neighbours in its own L2.
constexpr int kXCD = 8;
__device__ __forceinline__ void xcd_block(int& bx, int& by) {
  const int gx = gridDim.x;
  const int total = gx * gridDim.y;
  const int p = blockIdx.y * gx + blockIdx.x;
  const int per = total / kXCD;
  const int l = (p < per * kXCD) ? (p % kXCD) * per + p / kXCD : p;
  bx = l % gx;
  by = l / gx;
}

__device__ __forceinline__ void rc_of(const GridD& g, int i, int& row, int& col) {
  row = i / g.nx;
  col = i - row * g.nx;
}

}  // namespace pff

// Dispatch a templated kernel launch on the runtime boundary mode.
#define PFF_MODE_DISPATCH(mode, ...)            \
  do {                                          \
    switch (mode) {                             \
      case 0: { constexpr int M = 0; __VA_ARGS__; } break; \
      case 1: { constexpr int M = 1; __VA_ARGS__; } break; \
      case 2: { constexpr int M = 2; __VA_ARGS__; } break; \
      case 4: { constexpr int M = 4; __VA_ARGS__; } break; \
      default: { constexpr int M = 3; __VA_ARGS__; } break; \
    }                                           \
  } while (0)
