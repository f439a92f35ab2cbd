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

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define FMR_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t fmr_e_ = (expr);                                                      \
    if (fmr_e_ != hipSuccess) {                                                      \
      fprintf(stderr, "formura_amd: HIP error '%s' at %s:%d (%s)\n",                 \
              hipGetErrorString(fmr_e_), __FILE__, __LINE__, #expr);                 \
      abort();                                                                       \
    }                                                                                \
  } while (0)

namespace fmr {

constexpr int WAVE = 64;
constexpr int MAX_FIELDS = 8;
constexpr int MAX_PARAMS = 16;
constexpr int N_XCD = 8;   // MI355X: 8 accelerator complex dies, each with its own L2

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
// lane i receives lane i-1 (shr) or lane i+1 (shl); the lane with no source keeps 0.
__device__ __forceinline__ int dpp_shr1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_shl1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, false); }

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

// ===========================================================================
// Kernel 1: one site per thread, neighbours straight from global memory with
// periodic index wrap.  Any stencil shape and radius; the reference kernel the
// tiled kernels are checked against on the GPU.
// ===========================================================================
template <class P>
struct DirectAcc {
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
    DirectAcc<P> a{in, prm, {(int)(i / ((size_t)n1 * n2)), (int)((i / n2) % n1), (int)(i % n2)}, {n0, n1, n2}, {}};
    P::point(a);
#pragma unroll
    for (int f = 0; f < P::NF; ++f) out.f[f][i] = a.out[f];
  }
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
// so the ghost zones also implement the periodic boundary.  The window slot
// of the newest plane rotates with the iteration phase (0,1,2), so the march
// loop is unrolled by 3 and no register moves are needed.
//
// The input planes are prefetched PF (1 or 2) iterations ahead, so that a
// wavefront alone on its SIMD has one or two planes of compute to hide each
// load; results are written with non-temporal stores so they do not evict
// the ghost rows the neighbouring tiles are about to read from L2.
//
// OCC asks the compiler for at least OCC waves per SIMD (smaller register
// budget per wave, more waves to hide latency).
//
// Workgroup -> tile mapping is XCD-aware: consecutive workgroups go to
// different XCDs, so tile t is given to workgroup (t % per)*8 + t / per, which
// keeps neighbouring tiles -- whose ghost rows overlap -- on one XCD and its L2.
// ===========================================================================
template <class P, int T, int BY, int W, int PF>
struct March3D {
  using real = typename P::real;
  static constexpr int NF = P::NF;
  static constexpr int H = T;
  static constexpr int ROWS = BY * W;
  static_assert(ROWS > 2 * T, "tile must be taller than its ghost rows");
  using Edge = real[2][T][W][2][NF][WAVE];   // [half][level][wave][top,bottom][field][lane]

  real w[T][3][BY][NF];
  real nxt[PF][BY][NF];   // planes prefetched PF iterations ahead
  Edge* edge;   // LDS (W > 1 only)
  int wave, lane, half;
  int ylim;     // tile rows >= ylim lie past the grid end (periodic duplicates): never stored

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
      else if constexpr (D2 == -1) return lane_prev(m.w[L][CEN][ROW][F]);
      else if constexpr (D2 == 1) return lane_next(m.w[L][CEN][ROW][F]);
      else return m.w[L][CEN][ROW][F];
    }
    __device__ __forceinline__ double param(int i) const { return prm.v[i]; }
    __device__ __forceinline__ int index(int axis) const { return axis == 0 ? cx : (axis == 1 ? cy : cz); }
    template <int F>
    __device__ __forceinline__ void st(real v) { out[F] = v; }
  };

  template <int K>
  __device__ __forceinline__ void load_plane(const Fields<real>& in, size_t xoff, const int (&yoff)[BY]) {
#pragma unroll
    for (int r = 0; r < BY; ++r)
#pragma unroll
      for (int f = 0; f < NF; ++f) nxt[K][r][f] = in.f[f][xoff + yoff[r]];
  }

  // rows of level L+1 computed from window L: with one wave the valid rows
  // [L+1, BY-L-1) are known at compile time; with several, every row is
  // computed and the stores filter.
  static constexpr int row_lo(int L) { return W == 1 ? L + 1 : 0; }
  static constexpr int row_hi(int L) { return W == 1 ? BY - L - 1 : BY; }

  template <int L, int PH, int ROW>
  __device__ __forceinline__ void level_row(const Params& prm, int q, int y0, int zg, Fields<real>& out,
                                            size_t qoff, const int (&yoff)[BY], bool st_ok) {
    if constexpr (ROW < row_hi(L)) {
      const int trow = wave * BY + ROW;   // row within the tile
      Acc<L, PH, ROW> a{*this, prm, q, y0 + trow, zg, {}};
      P::point(a);
      if constexpr (L + 1 < T) {
#pragma unroll
        for (int f = 0; f < NF; ++f) w[L + 1][PH][ROW][f] = a.out[f];
      } else {
        if (st_ok && trow >= T && trow < ROWS - T && trow < ylim) {
#pragma unroll
          for (int f = 0; f < NF; ++f) __builtin_nontemporal_store(a.out[f], &out.f[f][qoff + yoff[ROW]]);
        }
      }
      level_row<L, PH, ROW + 1>(prm, q, y0, zg, out, qoff, yoff, st_ok);
    }
  }

  template <int L, int PH>
  __device__ __forceinline__ void levels(const Params& prm, int xin, int x0, int xend, int y0, int zg,
                                         bool lane_ok, Fields<real>& out, size_t plane, const int (&yoff)[BY]) {
    if constexpr (L < T) {
      const int q = xin - (L + 1);   // plane computed by level L+1
      bool st_ok = false;
      size_t qoff = 0;
      if constexpr (L + 1 == T) {
        st_ok = lane_ok && q >= x0 && q < xend;
        qoff = (size_t)(st_ok ? q : 0) * plane;
      }
      level_row<L, PH, row_lo(L)>(prm, q, y0, zg, out, qoff, yoff, st_ok);
      levels<L + 1, PH>(prm, xin, x0, xend, y0, zg, lane_ok, out, plane, yoff);
    }
  }

  template <int PH>
  __device__ __forceinline__ void iter(const Fields<real>& in, Fields<real>& out, const Params& prm, int xin,
                                       int x0, int xend, int nx, int y0, int zg, bool lane_ok, size_t plane,
                                       const int (&yoff)[BY]) {
    constexpr int CEN = (PH + 2) % 3;
#pragma unroll
    for (int r = 0; r < BY; ++r)
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        w[0][PH][r][f] = nxt[0][r][f];
#pragma unroll
        for (int k = 0; k + 1 < PF; ++k) nxt[k][r][f] = nxt[k + 1][r][f];
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
    load_plane<PF - 1>(in, (size_t)wrap(xin + PF, nx) * plane, yoff);
    levels<0, PH>(prm, xin, x0, xend, y0, zg, lane_ok, out, plane, yoff);
    half ^= 1;
  }
};

template <class P, int T, int BY, int W, int PF, int OCC>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(OCC))) march3d_kernel(Fields<typename P::real> in, Fields<typename P::real> out,
                                                         int nx, int ny, int nz, int lx, int nty, int ntz, int ntx,
                                                         Params prm) {
  using M = March3D<P, T, BY, W, PF>;
  constexpr int H = T;
  const int ntot = nty * ntz * ntx;
  const int per = (ntot + N_XCD - 1) / N_XCD;
  const int b = blockIdx.x;
  const int tile = (b % N_XCD) * per + b / N_XCD;
  if (tile >= ntot) return;   // uniform over the workgroup: no barrier is skipped by part of it
  const int tz = tile % ntz;
  const int ty = (tile / ntz) % nty;
  const int tx = tile / (ntz * nty);
  M m;
  m.wave = threadIdx.x / WAVE;
  m.lane = threadIdx.x % WAVE;
  m.half = 0;
  if constexpr (W > 1) {
    __shared__ typename M::Edge edge_buf;
    m.edge = &edge_buf;
  } else {
    m.edge = nullptr;
  }
  const int y0 = ty * (M::ROWS - 2 * H) - H;          // tile row 0
  const int zu = tz * (WAVE - 2 * H) - H + m.lane;     // unwrapped z of this lane
  const int zg = wrap(zu, nz);
  const int x0 = tx * lx;
  const int xend = min(x0 + lx, nx);
  const bool lane_ok = m.lane >= H && m.lane < WAVE - H && zu < nz;
  const size_t plane = (size_t)ny * nz;
  int yoff[BY];   // offset of row r within a plane (a plane holds < 2^31 sites)
#pragma unroll
  for (int r = 0; r < BY; ++r) yoff[r] = wrap(y0 + m.wave * BY + r, ny) * nz + zg;
  m.ylim = ny - y0;
  m.template load_plane<0>(in, (size_t)wrap(x0 - H, nx) * plane, yoff);
  if constexpr (PF > 1) m.template load_plane<PF - 1>(in, (size_t)wrap(x0 - H + 1, nx) * plane, yoff);
  static_assert(PF == 1 || PF == 2, "prefetch depth 1 or 2");
  const int iters = xend - x0 + 2 * T;
  for (int s = 0; s < iters; s += 3) {
    const int xin = x0 - H + s;
    m.template iter<0>(in, out, prm, xin, x0, xend, nx, y0, zg, lane_ok, plane, yoff);
    m.template iter<1>(in, out, prm, xin + 1, x0, xend, nx, y0, zg, lane_ok, plane, yoff);
    m.template iter<2>(in, out, prm, xin + 2, x0, xend, nx, y0, zg, lane_ok, plane, yoff);
  }
}

}  // namespace fmr
