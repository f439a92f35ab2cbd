// fmr_stencil.hpp -- hand-written CDNA4 (gfx950) kernel templates for the
// explicit step of a Formura program, and the host-side step library they run
// under.  A generated `<name>.hip` supplies a program struct `P`:
//
//   struct P {
//     using real = double;                 // device element type
//     static constexpr int DIM, NF;        // dimension, number of state fields
//     static constexpr int RX, RY, RZ;     // stencil radius per axis
//     static constexpr bool STAR;          // every load offset is axis-aligned
//     template <class A> __device__ static void point(A& a);
//   };
//
// `point` reads state fields through `a.template ld<F, D0, D1, D2>()`
// (field F at offset (D0, D1, D2) from the site; for DIM == 2 the offsets are
// (dx, dy, 0); axis DIM-1 is the contiguous one), scalar parameters through
// `a.param(i)`, global coordinates through `a.index(axis)`, and writes the
// new value of field F with `a.template st<F>(v)`.  The arithmetic inside
// `point` is the per-site statement list of the reference's generated C
// (src/Formura/MPICxx/Translate.hs:481-629), so every kernel below computes
// the same bits as that C for the same inputs; kernels only differ in where the
// neighbour values come from (global memory, registers, cross-lane moves).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#define FMR_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t fmr_e_ = (expr);                                                      \
    if (fmr_e_ != hipSuccess) {                                                      \
      fprintf(stderr, "formura_amd: HIP error '%s' at %s:%d (%s)\n",                 \
              hipGetErrorString(fmr_e_), __FILE__, __LINE__, #expr);                 \
      abort();                                                                       \
    }                                                                                \
  } while (0)

// Extern functions of a program are defined in its `<stem>.extern.h` with
// FMR_HD in front, so that one definition serves the GPU kernels and the
// host (and, with FMR_HD empty, the CPU oracle).
#ifndef FMR_HD
#define FMR_HD __host__ __device__
#endif

namespace fmr {

constexpr int WAVE = 64;
constexpr int MAX_FIELDS = 16;
constexpr int MAX_PARAMS = 16;
constexpr int N_XCD = 8;   // MI355X: 8 accelerator complex dies, each with its own L2

// Workgroup b of G is dispatched to XCD b % N_XCD.  xcd_order numbers the
// workgroups so that each XCD holds one contiguous range of [0, G) (ranges of
// G / N_XCD or one more, the longer ones first): consecutive work items share
// an XCD and its L2.  A bijection of [0, G) for every G >= 1.
__device__ __forceinline__ int xcd_order(int b, int G) {
  const int x = b % N_XCD, q = G / N_XCD, rem = G % N_XCD;
  return x * q + (x < rem ? x : rem) + b / N_XCD;
}

// Work items of a march launch: (chunk, tile, plane), listed chunk-major, x
// cut into chunks of L planes and a remainder chunk.  march_item_range gives
// workgroup g of G its items [a, e) and returns L positive.  L > 0: an even
// share, [g * total / G, (g + 1) * total / G).  L < 0 (whole runs with spare
// workgroups): chunks of -L planes; the first R = ntile * (nx / -L)
// workgroups take one whole run of a full chunk each, the other G - R (at
// least one whenever g >= R) share the remainder chunk's items evenly.
__host__ __device__ __forceinline__ void march_item_range(int g, int G, long long ntile, int nx, int& L,
                                                          long long& a, long long& e) {
  const long long total = ntile * nx;
  if (L > 0) {
    a = (long long)g * total / G;
    e = (long long)(g + 1) * total / G;
    return;
  }
  L = -L;
  const long long R = ntile * (nx / L);
  if (g < R) {
    a = (long long)g * L;
    e = a + L;
  } else {
    const long long full = R * L, rem = total - full, S = G - R;
    a = full + (g - R) * rem / S;
    e = full + (g - R + 1) * rem / S;
  }
}

// The runs of workgroup g of G: fn(tile, x0, xend) for each stretch of
// consecutive planes [x0, xend) of one tile within one chunk, in item order.
// Everything here is uniform over the workgroup, so every wave runs the same
// runs and reaches the same barriers.
template <class Fn>
__host__ __device__ __forceinline__ void march_runs(int g, int G, long long ntile, int nx, int L, Fn&& fn) {
  long long a, e;
  march_item_range(g, G, ntile, nx, L, a, e);
  const int nfull = nx / L;                              // chunks of exactly L planes
  const long long full_items = ntile * nfull * L;
  while (a < e) {
    int c0, len;         // first plane and length of the item's chunk
    long long r;         // item index within the chunk
    if (a < full_items) {
      const long long per_chunk = ntile * L;
      c0 = (int)(a / per_chunk) * L;
      len = L;
      r = a % per_chunk;
    } else {
      c0 = nfull * L;
      len = nx - c0;
      r = a - full_items;
    }
    const int tile = (int)(r / len);
    const int x0 = c0 + (int)(r % len);
    const long long stop = x0 + (e - a);
    const int xend = stop < c0 + len ? (int)stop : c0 + len;
    fn(tile, x0, xend);
    a += xend - x0;
  }
}

// Work split of a march launch over ntile tiles of nx planes with W-wave
// workgroups at fusion depth T on a device of `simds` SIMDs: the workgroup
// count and the chunk argument of the kernels (march_item_range).  lx > 0:
// x cut into chunks of lx planes, one workgroup per (chunk, tile).
// -200 > lx >= -300: x cut into the most chunks k whose k * ntile workgroups
// are resident at once with (-lx - 200) waves per SIMD (at least one), one
// workgroup per (chunk, tile): every workgroup one run, all of them in step.
// -300 > lx: the same with (-lx - 300) waves per SIMD, and when it leaves
// resident workgroups idle (k * ntile short of the count), those take the
// planes at the end of every tile: the k runs per tile shrink to L planes
// and the spare workgroups share the ntile * (nx - k L) planes left over
// (chunk -L), L chosen so that the longer of a run (L + 2T iterations) and a
// spare workgroup's share (its planes plus 2T per run it starts) is
// shortest.  -100 > lx >= -200: (-lx - 100) waves per SIMD of the device, x
// cut into the chunks that give each workgroup one run of its own (ntile *
// nx / workgroups planes).  Other lx < 0: x whole, -lx waves per SIMD,
// balanced to a plane.  lx = 0 is not a split.
inline void march_split_of(long long ntile, int nx, int W, int T, int lx, int simds, int& groups, int& chunk) {
  long long g;
  if (lx < -200) {
    const long long res = (long long)(lx < -300 ? -lx - 300 : -lx - 200) * simds / W;
    long long k = res / ntile;
    if (k < 1) k = 1;
    if (k > nx) k = nx;
    chunk = (int)((nx + k - 1) / k);
    g = ntile * ((nx + chunk - 1) / chunk);
    const long long spare = res - ntile * k;
    if (lx < -300 && spare > 0) {
      long long best = chunk + 2 * T;
      int best_len = 0;
      for (int len = (int)(nx / k); len > 0 && (k + 1) * len > nx; --len) {
        const long long left = nx - k * len;
        const long long share = (ntile * left + spare - 1) / spare;
        const long long runs = left > 0 ? (share + left - 1) / left + 1 : 0;
        const long long cost = std::max<long long>(len + 2 * T, share + 2 * T * runs);
        if (cost < best) {
          best = cost;
          best_len = len;
        }
      }
      if (best_len > 0) {
        chunk = -best_len;
        g = res;
      }
    }
  } else if (lx > 0) {
    chunk = lx < nx ? lx : nx;
    g = ntile * ((nx + chunk - 1) / chunk);
  } else if (lx < -100) {
    g = (long long)(-lx - 100) * simds / W;
    if (g > ntile * nx) g = ntile * nx;
    const long long c = (ntile * nx + g - 1) / g;
    chunk = (int)(c < 1 ? 1 : (c > nx ? nx : c));
  } else {
    chunk = nx;
    g = (long long)(-lx) * simds / W;
    if (g > ntile * nx) g = ntile * nx;
  }
  groups = (int)(g < 1 ? 1 : g);
}

template <class real>
struct Fields {
  real* f[MAX_FIELDS];
};

struct Params {
  double v[MAX_PARAMS];
};

__device__ __forceinline__ int wrap(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

// ---- cross-lane moves over the whole 64-lane wavefront (DPP wave_shr / wave_shl)
// lane i receives lane i-1 (shr) or lane i+1 (shl); the lane with no source reads 0
// (bound_ctrl set, so no zero-initialised destination register is needed).
__device__ __forceinline__ int dpp_shr1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_shl1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, true); }

__device__ __forceinline__ float lane_prev(float v) { return __int_as_float(dpp_shr1(__float_as_int(v))); }
__device__ __forceinline__ float lane_next(float v) { return __int_as_float(dpp_shl1(__float_as_int(v))); }
__device__ __forceinline__ double lane_prev(double v) {
  const long long q = __double_as_longlong(v);
  const int lo = dpp_shr1((int)(q & 0xffffffffLL));
  const int hi = dpp_shr1((int)(q >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_next(double v) {
  const long long q = __double_as_longlong(v);
  const int lo = dpp_shl1((int)(q & 0xffffffffLL));
  const int hi = dpp_shl1((int)(q >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---- the same moves as rotations (DPP wave_ror / wave_rol): lane 0 receives
// lane 63 (ror) and lane 63 lane 0 (rol).  When a wavefront holds a whole
// periodic z-line these are the periodic z-neighbours; otherwise the wrapped
// lane is a ghost lane like the zero of the shift.  Every lane has a source,
// so bound_ctrl changes no value; set, it spares the compiler a zeroed
// destination register per move.
__device__ __forceinline__ int dpp_ror1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x13C, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_rol1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x134, 0xF, 0xF, true); }

__device__ __forceinline__ float lane_rot_prev(float v) { return __int_as_float(dpp_ror1(__float_as_int(v))); }
__device__ __forceinline__ float lane_rot_next(float v) { return __int_as_float(dpp_rol1(__float_as_int(v))); }
__device__ __forceinline__ double lane_rot_prev(double v) {
  const long long q = __double_as_longlong(v);
  const int lo = dpp_ror1((int)(q & 0xffffffffLL));
  const int hi = dpp_ror1((int)(q >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_rot_next(double v) {
  const long long q = __double_as_longlong(v);
  const int lo = dpp_rol1((int)(q & 0xffffffffLL));
  const int hi = dpp_rol1((int)(q >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// the value of lane l (uniform l) in every lane
__device__ __forceinline__ float read_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
  const long long q = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(q & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(q >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---- buffer loads and stores (raw buffer: uniform base in a resource, 32-bit
// per-lane byte offset).  A resource spans one plane of a field, so the base
// moves in scalar registers and every row keeps one 32-bit offset.
constexpr int BUF_DWORD3 = 0x00020000;   // gfx9 raw buffer: 32-bit data format, no swizzle
constexpr int BUF_NT = 2;                // cache policy bit: non-temporal
// A vector offset >= the resource's record count makes a raw buffer access
// out of range: a load returns 0, a store is dropped.  Resources here span at
// most 2 GiB, so this offset masks a lane without a branch.
constexpr unsigned BUF_OOB = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, BUF_DWORD3);
}
// off: per-lane byte offset (VGPR); soff: uniform byte offset (SGPR)
__device__ __forceinline__ double buf_load(double*, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff = 0) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
}
__device__ __forceinline__ float buf_load(float*, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff = 0) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
}
__device__ __forceinline__ void buf_store_nt(double v, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff = 0) {
  using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, off, soff, BUF_NT);
}
__device__ __forceinline__ void buf_store_nt(float v, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff = 0) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, soff, BUF_NT);
}

// ===========================================================================
// Kernel 1: one site per thread, neighbours straight from global memory with
// periodic index wrap.  Any stencil shape and radius; the reference kernel the
// tiled kernels are checked against on the GPU.
// ===========================================================================
// Shift reuse hook called by `point` (the compiler's shift analysis finds op
// pairs with B(s) = A(s - e) along the lane axis): a kernel that holds the
// previous site of that axis in the neighbouring lane returns lane_prev(A);
// the default evaluates B.
struct EvalReuse {
  template <class V, class Fn>
  __device__ __forceinline__ V lane_prev_of(V, Fn f) const { return f(); }
};

template <class P>
struct DirectAcc : EvalReuse {
  using real = typename P::real;
  const Fields<real>& in;
  const Params& prm;
  int c[3];
  int n[3];
  real out[P::NF];

  template <int F, int D0, int D1, int D2>
  __device__ __forceinline__ real ld() const {
    const int x = (D0 == 0) ? c[0] : wrap(c[0] + D0, n[0]);
    const int y = (D1 == 0) ? c[1] : wrap(c[1] + D1, n[1]);
    const int z = (D2 == 0) ? c[2] : wrap(c[2] + D2, n[2]);
    return in.f[F][((size_t)x * n[1] + y) * n[2] + z];
  }
  __device__ __forceinline__ double param(int i) const { return prm.v[i]; }
  __device__ __forceinline__ int index(int axis) const { return c[axis]; }
  template <int F>
  __device__ __forceinline__ void st(real v) { out[F] = v; }
};

// Grid of (n0, n1, n2): for DIM == 2 the grid is (NX, NY, 1) and offsets
// (dx, dy, 0); for DIM == 3 it is (NX, NY, NZ).
template <class P>
__global__ void __launch_bounds__(256) direct_step_kernel(Fields<typename P::real> in,
                                                          Fields<typename P::real> out,
                                                          int n0, int n1, int n2, Params prm) {
  const size_t total = (size_t)n0 * n1 * n2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    DirectAcc<P> a{{}, in, prm, {(int)(i / ((size_t)n1 * n2)), (int)((i / n2) % n1), (int)(i % n2)}, {n0, n1, n2}, {}};
    P::point(a);
#pragma unroll
    for (int f = 0; f < P::NF; ++f) out.f[f][i] = a.out[f];
  }
}

// ===========================================================================
// Kernel 1b: manifest nodes of a multi-stage step (a program with
// `manifest ::` declarations, e.g. Runge-Kutta stages).  One kernel evaluates
// a group of nodes that do not read each other's arrays; `S::point` reads the
// state fields and the arrays of earlier nodes through
// `a.template ld<ARRAY, D0, D1, D2>()` (ARRAY indexes in.a: state fields
// first, then the stage arrays) and writes member k's value with
// `a.template st<k>(v)`.  Each node is evaluated once per site and stored, as
// the reference's loop nest per manifest node does, instead of being
// recomputed inside its users.
//
// Thread mapping: the contiguous axis LANE (z in 3D, y in 2D) across the 64
// lanes of a wavefront, ROWS wavefronts on consecutive positions of the next
// axis, the remaining axis over grid.z -- no index division, coalesced rows.
// ===========================================================================
constexpr int MAX_ARRAYS = 128;   // state fields + stage arrays of one program
constexpr int MAX_OUTS = 16;      // manifest nodes per stage kernel

template <class real>
struct Arrays {
  real* a[MAX_ARRAYS];
};

template <class real>
struct Outs {
  real* o[MAX_OUTS];
};

// neighbour coordinate i + d with |d| <= n on a periodic axis of extent n
__device__ __forceinline__ int wrap_near(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Addressing: the site's element offset `base`, plus one signed delta per
// axis and neighbour distance (the same delta for every array, computed once
// per distance: the compiler merges the identical expressions).  WIDE = false
// (arrays below 4 GiB): 32-bit byte offsets added to the array's base in
// scalar registers; WIDE = true: 64-bit element offsets.
template <class S, bool WIDE>
struct StageAcc : EvalReuse {
  using real = typename S::real;
  using off_t = typename std::conditional<WIDE, long long, unsigned>::type;
  const Arrays<real>& in;
  const Params& prm;
  int c[3];
  int n[3];
  off_t base;      // bytes (WIDE: elements) of the site from an array's start
  off_t stride[3];
  real out[S::NOUT];

  template <int A, int D>
  __device__ __forceinline__ off_t delta() const {
    if constexpr (D == 0) {
      return 0;
    } else {
      return (off_t)(wrap_near(c[A] + D, n[A]) - c[A]) * stride[A];
    }
  }
  template <int F, int D0, int D1, int D2>
  __device__ __forceinline__ real ld() const {
    const off_t o = base + delta<0, D0>() + delta<1, D1>() + delta<2, D2>();
    if constexpr (WIDE) {
      return in.a[F][o];
    } else {
      return *(const real*)((const char*)in.a[F] + o);
    }
  }
  __device__ __forceinline__ double param(int i) const { return prm.v[i]; }
  __device__ __forceinline__ int index(int axis) const { return c[axis]; }
  template <int F>
  __device__ __forceinline__ void st(real v) { out[F] = v; }
};

// ROWS wavefronts per workgroup, one row each (4 or 16)
template <class S, int LANE, bool WIDE, int ROWS>
__global__ void __launch_bounds__(64 * ROWS) stage_kernel(Arrays<typename S::real> in, Outs<typename S::real> out,
                                                         int n0, int n1, int n2, Params prm) {
  using real = typename S::real;
  using A = StageAcc<S, WIDE>;
  using off_t = typename A::off_t;
  constexpr int ROW = LANE == 2 ? 1 : 0;    // axis of the ROWS rows
  constexpr int SLAB = LANE == 2 ? 0 : 2;   // axis over grid.z (extent 1 in 2D)
  const int n[3] = {n0, n1, n2};
  int c[3];
  c[LANE] = blockIdx.x * 64 + threadIdx.x;
  c[ROW] = blockIdx.y * ROWS + threadIdx.y;
  c[SLAB] = blockIdx.z;
  if (c[LANE] >= n[LANE] || c[ROW] >= n[ROW]) return;
  const off_t unit = WIDE ? 1 : (off_t)sizeof(real);
  const off_t st2 = unit, st1 = (off_t)n2 * st2, st0 = (off_t)n1 * st1;
  const off_t base = (off_t)c[0] * st0 + (off_t)c[1] * st1 + (off_t)c[2] * st2;
  A a{{}, in, prm, {c[0], c[1], c[2]}, {n0, n1, n2}, base, {st0, st1, st2}, {}};
  S::point(a);
  const size_t i = ((size_t)c[0] * n1 + c[1]) * n2 + c[2];
#pragma unroll
  for (int k = 0; k < S::NOUT; ++k) out.o[k][i] = a.out[k];
}

// ===========================================================================
// Kernel 2: 3D temporally blocked march (star stencils of radius 1).
//
// A workgroup of W wavefronts owns a tile of 64 z-columns (one per lane; z is
// the contiguous axis) by W*BY y-rows (wave w holds rows [w*BY, (w+1)*BY) in
// its registers) and marches along x over a chunk of planes.  T time levels
// are pipelined along the march: level l computes plane (xin - l) as soon as
// level l-1 has produced plane (xin - l + 1).  Per level a lane keeps a
// 3-plane window (x-1, x, x+1) of its BY rows in registers.  Neighbours:
//   x +-1 : the window (registers),
//   y +-1 : the neighbouring row of the same lane (registers); across a wave
//           boundary, the neighbouring wave's edge row, exchanged through LDS,
//   z +-1 : the neighbouring lane (DPP wave_shr:1 / wave_shl:1).
// The edge rows every level needs in one iteration are all known when the
// iteration starts (they are the window centres, produced one iteration
// earlier), so one LDS exchange and one barrier per plane serve all T levels;
// the LDS buffer alternates between two halves so the next plane's writes
// cannot overtake this plane's reads.  With W == 1 the kernel uses no LDS
// and no barrier.
// Ghost zones: the tile carries T ghost rows / lanes on each side, recomputed
// redundantly; after level l the valid part is rows [l, W*BY-l) and lanes
// [l, 64-l).  Only level T's valid part is written.  Reads wrap periodically,
// so the ghost zones also implement the periodic boundary.  With T = 1 there
// are no ghost lanes: a wave's 64 lanes are 64 consecutive z sites from a
// multiple of 64, so every row it loads or stores is whole, aligned cache
// lines, and the two sites just outside its z range come from one extra
// load per plane (the z neighbours of lanes 0 and 63).  The window slot
// of the newest plane rotates with the iteration phase (0,1,2), so the march
// loop is unrolled by 3 and no register moves are needed.
//
// The input planes are prefetched PF (1 or 2) iterations ahead, so that a
// wavefront alone on its SIMD has one or two planes of compute to hide each
// load; results are written with non-temporal stores so they do not evict
// the ghost rows the neighbouring tiles are about to read from L2.
//
// Addressing: loads and stores are raw buffer operations on a resource that
// spans one plane of a field.  The plane base lives in scalar registers and
// moves with the march; the row offsets (uniform) are scalar offsets and the
// lane's z offset is one vector register, so a row costs no address
// arithmetic and no vector register.  A plane must be smaller than 2 GiB.
//
// OCC asks the compiler for at least OCC waves per SIMD (smaller register
// budget per wave, more waves to hide latency).
//
// Workgroup -> work mapping is XCD-aware: consecutive workgroups go to
// different XCDs, so run g is given to workgroup (g % per)*8 + g / per, which
// keeps neighbouring runs -- same tile or adjacent tiles, whose ghost rows
// overlap -- on one XCD and its L2.
// ===========================================================================
template <class P, int T, int BY, int W, int PF>
struct March3D {
  using real = typename P::real;
  static constexpr int NF = P::NF;
  static constexpr int H = T;
  static constexpr int ROWS = BY * W;
  static_assert(ROWS > 2 * T, "tile must be taller than its ghost rows");
  static_assert(PF == 1 || PF == 2, "prefetch depth 1 or 2");
  using Edge = real[2][T][W][2][NF][WAVE];   // [half][level][wave][top,bottom][field][lane]

  // T = 1: a wave's 64 lanes are 64 consecutive z sites from a multiple of 64
  // (whole, aligned cache lines; no ghost lanes).  The z neighbours of lanes
  // 0 and 63 come from a halo load per plane: lane l of halo register j holds
  // row 32j + l/2's site just before (l even) or just after (l odd) the
  // wave's z range.  T > 1: 2T ghost lanes, whose sites other tiles store.
  static constexpr bool ZFULL = (T == 1);
  static constexpr int NH = ZFULL ? (BY + 31) / 32 : 1;

  real w[T][3][BY][NF];
  real nxt[PF][BY][NF];   // planes prefetched PF iterations ahead
  real hw[3][NH][NF];     // halo registers of the window planes (ZFULL)
  real hnxt[PF][NH][NF];  // and of the prefetched planes
  Edge* edge;   // LDS (W > 1 only)
  int wave, lane, half;
  int nx, ny;          // grid extents (index values wrap into them)
  unsigned zoff;       // byte offset of this lane's z-column within a row
  unsigned zst;        // store offset: zoff, or BUF_OOB for a ghost lane (the store is dropped)
  unsigned roff[BY];   // byte offset of row r within a plane (uniform: scalar registers)
  unsigned hoff[NH];   // byte offset of this lane's halo site within a plane (ZFULL)

  // z neighbours of a row of plane phase CEN at level 0 (ZFULL): the lane
  // shift, with the halo site in lane 0 (before) or lane 63 (after)
  template <int CEN, int ROW, int F>
  __device__ __forceinline__ real z_before(real shifted) const {
    const real h = read_lane(hw[CEN][ROW / 32][F], 2 * (ROW % 32));
    return lane == 0 ? h : shifted;
  }
  template <int CEN, int ROW, int F>
  __device__ __forceinline__ real z_after(real shifted) const {
    const real h = read_lane(hw[CEN][ROW / 32][F], 2 * (ROW % 32) + 1);
    return lane == WAVE - 1 ? h : shifted;
  }

  template <int L, int CEN, int ROW, int F>
  __device__ __forceinline__ real row_above() const {
    if constexpr (ROW > 0) return w[L][CEN][ROW - 1][F];
    else if constexpr (W == 1) return w[L][CEN][ROW][F];   // a ghost row: any finite value
    else return wave > 0 ? (*edge)[half][L][wave - 1][1][F][lane] : w[L][CEN][ROW][F];
  }
  template <int L, int CEN, int ROW, int F>
  __device__ __forceinline__ real row_below() const {
    if constexpr (ROW < BY - 1) return w[L][CEN][ROW + 1][F];
    else if constexpr (W == 1) return w[L][CEN][ROW][F];
    else return wave < W - 1 ? (*edge)[half][L][wave + 1][0][F][lane] : w[L][CEN][ROW][F];
  }

  template <int L, int PH, int ROW>
  struct Acc {
    const March3D& m;
    const Params& prm;
    int cx, cy, cz;
    real out[NF];
    template <int F, int D0, int D1, int D2>
    __device__ __forceinline__ real ld() const {
      static_assert(((D0 != 0) + (D1 != 0) + (D2 != 0)) <= 1, "march kernel needs a star stencil");
      static_assert(D0 >= -1 && D0 <= 1 && D1 >= -1 && D1 <= 1 && D2 >= -1 && D2 <= 1, "radius 1");
      constexpr int NEW = PH, CEN = (PH + 2) % 3, OLD = (PH + 1) % 3;
      if constexpr (D0 == -1) return m.w[L][OLD][ROW][F];
      else if constexpr (D0 == 1) return m.w[L][NEW][ROW][F];
      else if constexpr (D1 == -1) return m.template row_above<L, CEN, ROW, F>();
      else if constexpr (D1 == 1) return m.template row_below<L, CEN, ROW, F>();
      else if constexpr (D2 == -1) {
        if constexpr (ZFULL) return m.template z_before<CEN, ROW, F>(lane_prev(m.w[L][CEN][ROW][F]));
        else return lane_prev(m.w[L][CEN][ROW][F]);
      } else if constexpr (D2 == 1) {
        if constexpr (ZFULL) return m.template z_after<CEN, ROW, F>(lane_next(m.w[L][CEN][ROW][F]));
        else return lane_next(m.w[L][CEN][ROW][F]);
      } else {
        return m.w[L][CEN][ROW][F];
      }
    }
    __device__ __forceinline__ double param(int i) const { return prm.v[i]; }
    // global coordinates of the site, wrapped into [0, n): ghost planes, rows
    // and lanes and the rows past the grid end of the last tile are periodic
    // images of interior sites and take those sites' index values
    __device__ __forceinline__ int index(int axis) const {
      return axis == 0 ? wrap(cx, m.nx) : (axis == 1 ? wrap(cy, m.ny) : cz);
    }
    template <int F>
    __device__ __forceinline__ void st(real v) { out[F] = v; }
    // the lane axis is z: the previous site's value is in the previous lane;
    // lane 0 of a ZFULL wave has none there and evaluates it (the same
    // operations on the same values: the same bits)
    template <class V, class Fn>
    __device__ __forceinline__ V lane_prev_of(V v, Fn fallback) const {
      if constexpr (ZFULL) {
        const V shifted = lane_prev(v);
        const V own = fallback();
        return m.lane == 0 ? own : shifted;
      } else {
        return lane_prev(v);
      }
    }
  };

  // load plane x (already wrapped) into prefetch slot K
  template <int K>
  __device__ __forceinline__ void load_plane(const Fields<real>& in, int x, unsigned plane_bytes) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const auto rs = plane_rsrc(reinterpret_cast<const char*>(in.f[f]) + (size_t)x * plane_bytes, plane_bytes);
#pragma unroll
      for (int r = 0; r < BY; ++r) nxt[K][r][f] = buf_load((real*)nullptr, rs, zoff, roff[r]);
      if constexpr (ZFULL) {
#pragma unroll
        for (int j = 0; j < NH; ++j) hnxt[K][j][f] = buf_load((real*)nullptr, rs, hoff[j]);
      }
    }
  }

  // rows of level L+1 computed from window L: the tile's rows [L+1,
  // ROWS-L-1).  With one wave the range is known at compile time; with
  // several, the first and last wave skip their share by a uniform test on
  // the rows near their edge (the middle rows need none).
  static constexpr int row_lo(int L) { return W == 1 ? L + 1 : 0; }
  static constexpr int row_hi(int L) { return W == 1 ? BY - L - 1 : BY; }

  template <int L, int ROW>
  __device__ __forceinline__ bool row_needed() const {
    if constexpr (W == 1 || (ROW >= L + 1 && ROW < BY - L - 1)) return true;
    else return (wave > 0 || ROW >= L + 1) && (wave < W - 1 || ROW < BY - L - 1);
  }

  // Stores of level T: every valid row (rows [T, ROWS-T)) is written; ghost
  // lanes and lanes past the grid end are masked by an out-of-range offset.
  // Rows past the grid end (last y-tile) are periodic duplicates of rows
  // another tile owns -- the same operations on the same values, so the same
  // bits -- and are written too (identical values, in any order).
  template <int L, int PH, int ROW>
  __device__ __forceinline__ void level_row(const Params& prm, int q, int y0, int zg, const __amdgpu_buffer_rsrc_t (&orsrc)[NF],
                                            bool st_ok) {
    if constexpr (ROW < row_hi(L)) {
      if (row_needed<L, ROW>()) {
        const int trow = wave * BY + ROW;   // row within the tile
        Acc<L, PH, ROW> a{*this, prm, q, y0 + trow, zg, {}};
        P::point(a);
        if constexpr (L + 1 < T) {
#pragma unroll
          for (int f = 0; f < NF; ++f) w[L + 1][PH][ROW][f] = a.out[f];
        } else if (st_ok && (W == 1 || (trow >= T && trow < ROWS - T))) {
#pragma unroll
          for (int f = 0; f < NF; ++f) buf_store_nt(a.out[f], orsrc[f], zst, roff[ROW]);
        }
      }
      level_row<L, PH, ROW + 1>(prm, q, y0, zg, orsrc, st_ok);
    }
  }

  // levels L+1 .. T for one iteration.  Level T's plane q is stored when it
  // lies in [x0, xend) and not computed otherwise (nothing else reads it).
  template <int L, int PH>
  __device__ __forceinline__ void levels(const Params& prm, int xin, int x0, int xend, int y0, int zg,
                                         const Fields<real>& out, unsigned plane_bytes) {
    if constexpr (L < T) {
      const int q = xin - (L + 1);   // plane computed by level L+1
      if constexpr (L + 1 == T) {
        if (q >= x0 && q < xend) {
          __amdgpu_buffer_rsrc_t orsrc[NF];
#pragma unroll
          for (int f = 0; f < NF; ++f)
            orsrc[f] = plane_rsrc(reinterpret_cast<char*>(out.f[f]) + (size_t)q * plane_bytes, plane_bytes);
          level_row<L, PH, row_lo(L)>(prm, q, y0, zg, orsrc, true);
        }
      } else {
        __amdgpu_buffer_rsrc_t none[NF];
        level_row<L, PH, row_lo(L)>(prm, q, y0, zg, none, false);
        levels<L + 1, PH>(prm, xin, x0, xend, y0, zg, out, plane_bytes);
      }
    }
  }

  // one step of the march: plane xin enters level 0, plane xl (= xin + PF,
  // wrapped) is requested
  template <int PH>
  __device__ __forceinline__ void iter(const Fields<real>& in, const Fields<real>& out, const Params& prm, int xin,
                                       int xl, int x0, int xend, int y0, int zg, unsigned plane_bytes) {
    constexpr int CEN = (PH + 2) % 3;
#pragma unroll
    for (int r = 0; r < BY; ++r)
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        w[0][PH][r][f] = nxt[0][r][f];
#pragma unroll
        for (int k = 0; k + 1 < PF; ++k) nxt[k][r][f] = nxt[k + 1][r][f];
      }
    if constexpr (ZFULL) {
#pragma unroll
      for (int j = 0; j < NH; ++j)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          hw[PH][j][f] = hnxt[0][j][f];
#pragma unroll
          for (int k = 0; k + 1 < PF; ++k) hnxt[k][j][f] = hnxt[k + 1][j][f];
        }
    }
    if constexpr (W > 1) {
      // publish this wave's edge rows of every window centre, then one barrier
#pragma unroll
      for (int l = 0; l < T; ++l)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          (*edge)[half][l][wave][0][f][lane] = w[l][CEN][0][f];
          (*edge)[half][l][wave][1][f][lane] = w[l][CEN][BY - 1][f];
        }
      __syncthreads();
    }
    load_plane<PF - 1>(in, xl, plane_bytes);
    levels<0, PH>(prm, xin, x0, xend, y0, zg, out, plane_bytes);
    half ^= 1;
  }
};

__device__ __forceinline__ int next_plane(int x, int nx) { return x + 1 == nx ? 0 : x + 1; }

// Work split: the x axis is cut into chunks of L planes (the last one may be
// shorter) and the (chunk, tile, plane) items are listed chunk-major, so that
// the tiles of one chunk -- whose ghost rows overlap -- are marched at the same
// time and share their reads in L2.  Workgroup g of the G launched (G a
// multiple of 8) marches the items [g * total / G, (g + 1) * total / G): a run
// of planes of one (chunk, tile), or the end of one and the start of the next
// (the pipeline restarts there).  With G = chunks x tiles every workgroup
// marches exactly one chunk of one tile; with fewer, larger workgroup counts
// are balanced to one plane.  A plane (ny * nz sites) must be < 2 GiB.
template <class P, int T, int BY, int W, int PF>
__device__ __forceinline__ void march_run(March3D<P, T, BY, W, PF>& m, const Fields<typename P::real>& in,
                                          const Fields<typename P::real>& out, int nx, int ny, int nz, int ntz,
                                          int tile, int x0, int xend, const Params& prm) {
  using M = March3D<P, T, BY, W, PF>;
  using real = typename P::real;
  constexpr int H = T;
  const int tz = tile % ntz;
  const int ty = tile / ntz;
  const int y0 = ty * (M::ROWS - 2 * H) - H;          // tile row 0
  const int z0 = M::ZFULL ? tz * WAVE : tz * (WAVE - 2 * H) - H;   // unwrapped z of lane 0
  const int zu = z0 + m.lane;                           // unwrapped z of this lane
  const int zg = wrap(zu, nz);

  const unsigned plane_bytes = (unsigned)((size_t)ny * nz * sizeof(real));
#pragma unroll
  for (int r = 0; r < BY; ++r)
    m.roff[r] = (unsigned)__builtin_amdgcn_readfirstlane(wrap(y0 + m.wave * BY + r, ny) * nz * (int)sizeof(real));
  m.zoff = (unsigned)(zg * sizeof(real));
  // lanes past the grid end hold periodic duplicates and are not written
  if constexpr (M::ZFULL) {
    m.zst = zu < nz ? m.zoff : BUF_OOB;
    const int zh = wrap((m.lane & 1) ? z0 + WAVE : z0 - 1, nz);
#pragma unroll
    for (int j = 0; j < M::NH; ++j) {
      const int r = min(32 * j + m.lane / 2, BY - 1);
      m.hoff[j] = (unsigned)((wrap(y0 + m.wave * BY + r, ny) * nz + zh) * (int)sizeof(real));
    }
  } else {
    m.zst = (m.lane >= H && m.lane < WAVE - H && zu < nz) ? m.zoff : BUF_OOB;
  }
  int xl = wrap(x0 - H, nx);
  m.template load_plane<0>(in, xl, plane_bytes);
  if constexpr (PF > 1) {
    xl = next_plane(xl, nx);
    m.template load_plane<PF - 1>(in, xl, plane_bytes);
  }
  xl = next_plane(xl, nx);   // plane requested by the first iteration: x0 - H + PF
  const int iters = xend - x0 + 2 * T;
  for (int s = 0; s < iters; s += 3) {
    const int xin = x0 - H + s;
    m.template iter<0>(in, out, prm, xin, xl, x0, xend, y0, zg, plane_bytes);
    xl = next_plane(xl, nx);
    m.template iter<1>(in, out, prm, xin + 1, xl, x0, xend, y0, zg, plane_bytes);
    xl = next_plane(xl, nx);
    m.template iter<2>(in, out, prm, xin + 2, xl, x0, xend, y0, zg, plane_bytes);
    xl = next_plane(xl, nx);
  }
}

template <class P, int T, int BY, int W, int PF, int OCC>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(OCC)))
march3d_kernel(Fields<typename P::real> in, Fields<typename P::real> out, int nx, int ny, int nz, int nty, int ntz,
               int L, Params prm) {
  using M = March3D<P, T, BY, W, PF>;
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int g = xcd_order(b, G);   // neighbouring runs on one XCD (shared L2)
  const long long ntile = (long long)nty * ntz;
  M m;
  m.wave = threadIdx.x / WAVE;
  m.lane = threadIdx.x % WAVE;
  m.half = 0;
  m.nx = nx;
  m.ny = ny;
  if constexpr (W > 1) {
    __shared__ typename M::Edge edge_buf;
    m.edge = &edge_buf;
  } else {
    m.edge = nullptr;
  }
  march_runs(g, G, ntile, nx, L, [&](int tile, int x0, int xend) {
    march_run<P, T, BY, W, PF>(m, in, out, nx, ny, nz, ntz, tile, x0, xend, prm);
  });
}

}  // namespace fmr
