// fmr_march.hpp -- the temporally blocked 3D march with VZ z-sites per lane
// and the x-difference carried from plane to plane.
//
// Same scheme as march3d_kernel (fmr_stencil.hpp): a workgroup of W
// wavefronts owns a tile of z-sites by W*BY y-rows and marches along x over a
// run of planes; T time levels are pipelined along the march, each in a
// register window; y-neighbours across a wave boundary come from an LDS edge
// exchange (one barrier per plane); ghost zones of T rows and T (rounded up
// to whole lanes) z-sites on each side are recomputed, and loads wrap
// periodically.  What differs:
//
//  * z-lines held whole.  The lane moves are rotations, so with one site per
//    lane and NZ = 64 a wavefront row is a whole periodic z-line: the tiling
//    has one z-tile with no ghost lanes (the host passes gz = 0) and every
//    lane stores.  Otherwise the wrapped lane is a ghost lane, as the zero
//    of a shift was.
//
//  * VZ consecutive z-sites per lane (1 or 2).  A lane loads and stores its
//    VZ sites with one 8 * VZ-byte (FP64) access, and of the 2 VZ z-neighbours
//    of its sites only two come from other lanes (DPP wave rotations): with
//    VZ = 2 a site costs half the cross-lane moves and half the memory
//    instructions of VZ = 1.  The point function of the program is emitted
//    for VZ sites at once (`P::template mpoint<VZ>`), so that a value one
//    site computes and the next site needs along z (the lane-axis shift
//    reuse) is taken from the register of the site before, or across the
//    lane boundary from the previous lane.
//
//  * x-carry (XC).  When every use of the plane behind (load offset x-1) is
//    an operation that equals an operation of the plane before, shifted by
//    one along x (for the Laplacian: c - c[x-1] at x is c[x+1] - c at x-1),
//    that value is kept from the previous plane in a carry register instead
//    of being recomputed: one FP64 operation less per field and site, and
//    the window holds two planes (centre, ahead) plus the carry per level
//    instead of three planes.  Same IEEE operation on the same operands:
//    the same bits.
//
// Every kernel computes the per-site statement list of the reference's
// generated C (src/Formura/MPICxx/Translate.hs:481-629) on the same operands
// as the direct kernel, so every variant gives the same bits.
#pragma once

#include "fmr_stencil.hpp"

namespace fmr {

// cache policy of the march's state stores: non-temporal (streamed past the
// caches; plain stores measured the same)
constexpr int MV_STORE = BUF_NT;

template <class real, int VZ>
struct VecIO;

template <>
struct VecIO<double, 1> {
  __device__ __forceinline__ static void load(double (&v)[1], __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    v[0] = buf_load((double*)nullptr, r, off, soff);
  }
  __device__ __forceinline__ static void store(const double (&v)[1], __amdgpu_buffer_rsrc_t r, unsigned off,
                                               unsigned soff) {
    typedef int V2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v[0]), r, off, soff, MV_STORE);
  }
};

template <>
struct VecIO<double, 2> {
  typedef int V4 __attribute__((ext_vector_type(4)));
  typedef double D2 __attribute__((ext_vector_type(2)));
  __device__ __forceinline__ static void load(double (&v)[2], __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    const D2 d = __builtin_bit_cast(D2, __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0));
    v[0] = d.x;
    v[1] = d.y;
  }
  __device__ __forceinline__ static void store(const double (&v)[2], __amdgpu_buffer_rsrc_t r, unsigned off,
                                               unsigned soff) {
    D2 d;
    d.x = v[0];
    d.y = v[1];
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, d), r, off, soff, MV_STORE);
  }
};

template <>
struct VecIO<float, 1> {
  __device__ __forceinline__ static void load(float (&v)[1], __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    v[0] = buf_load((float*)nullptr, r, off, soff);
  }
  __device__ __forceinline__ static void store(const float (&v)[1], __amdgpu_buffer_rsrc_t r, unsigned off,
                                               unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v[0]), r, off, soff, MV_STORE);
  }
};

template <>
struct VecIO<float, 2> {
  typedef int V2 __attribute__((ext_vector_type(2)));
  typedef float F2 __attribute__((ext_vector_type(2)));
  __device__ __forceinline__ static void load(float (&v)[2], __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    const F2 d = __builtin_bit_cast(F2, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
    v[0] = d.x;
    v[1] = d.y;
  }
  __device__ __forceinline__ static void store(const float (&v)[2], __amdgpu_buffer_rsrc_t r, unsigned off,
                                               unsigned soff) {
    F2 d;
    d.x = v[0];
    d.y = v[1];
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, d), r, off, soff, MV_STORE);
  }
};

// Wave roles of a multi-wave tile with edge rows (EY > 0): the first wave
// (ROLE_TOP) and the last (ROLE_BOTTOM) hold BY + EY rows, the others
// (ROLE_INNER) BY.  Level l (1..T) is needed only on tile rows [l, ROWS - l),
// so the first wave computes level l on its rows [l, BY + EY) and the last
// on [0, BY + EY - l): with EY = 2 at T = 4 they evaluate 26 row-levels
// against the inner waves' 28, and the tile stores 4 rows more for about the
// same work.  ROLE_ALL (EY = 0): every wave BY rows, all rows at every level
// (the waves meet at a barrier each plane, so rows the first and last wave
// skipped would save no time).
enum MarchRole { ROLE_ALL = 0, ROLE_TOP = 1, ROLE_INNER = 2, ROLE_BOTTOM = 3 };

template <class P, int T, int BY, int W, int VZ, bool XC, int PF, int EY = 0, int ROLE = ROLE_ALL>
struct MarchV {
  using real = typename P::real;
  static constexpr int NF = P::NF;
  static constexpr int NC = XC ? P::NC : 0;     // carried values per site
  static constexpr int NP = XC ? 2 : 3;         // window planes per level
  static constexpr int RB = (ROLE == ROLE_TOP || ROLE == ROLE_BOTTOM) ? BY + EY : BY;   // rows of this wave
  static constexpr int ROWS = BY * W + 2 * EY;
  static constexpr int G = (T + VZ - 1) / VZ * VZ;   // fewest ghost z-sites per side (whole lanes)
  static constexpr int ZW = WAVE * VZ;               // z-sites per row of a tile
  static_assert(T >= 1 && ROWS > 2 * T, "tile must be taller than its ghost rows");
  static_assert(ZW > 2 * G, "tile must be wider than its ghost lanes");
  static_assert(VZ == 1 || VZ == 2, "one or two z-sites per lane");
  static_assert(PF == 1 || PF == 2, "prefetch depth 1 or 2");
  static_assert(!XC || P::NC > 0, "x-carry needs carried operations");
  static_assert(EY == 0 ? ROLE == ROLE_ALL : (W > 1 && ROLE != ROLE_ALL && BY + EY > T), "edge rows need roles");
  using Edge = real[2][T][W][2][NF][VZ][WAVE];   // [half][level][wave][top,bottom][field][site][lane]

  real w[T][NP][RB][VZ][NF];
  real c[T][RB][VZ][NC > 0 ? NC : 1];
  real nxt[PF][RB][VZ][NF];
  Edge* edge;
  int wave, lane, half;
  int wave_above, wave_below;   // (wave - 1) and (wave + 1) modulo W
  int nx, ny, nz;
  int gz;              // ghost z-sites per side: G, or more (a multiple of VZ) so that the stored
                       // range of every tile starts on a 128-byte line, or 0 for whole lines (VZ = 1, NZ = 64)
  int zg;              // wrapped z of this lane's first site
  unsigned zoff;       // byte offset of this lane's first site within a row
  unsigned zst;        // store offset: zoff, or BUF_OOB for a ghost lane or a lane past the grid end
  unsigned roff[RB];   // byte offset of row r within a plane (uniform)

  // tile row of this wave's row 0
  __device__ __forceinline__ int row0() const {
    if constexpr (ROLE == ROLE_ALL) return wave * BY;
    else if constexpr (ROLE == ROLE_TOP) return 0;
    else if constexpr (ROLE == ROLE_INNER) return BY + EY + (wave - 1) * BY;
    else return BY + EY + (W - 2) * BY;
  }

  // y-neighbours across a wave boundary: the edge row the neighbouring wave
  // published.  The tile's first and last rows are ghost rows whose values
  // are never stored, so wave 0 reads wave W-1's row there (and vice versa):
  // any finite value will do, and no select is needed.
  template <int L, int CEN, int ROW, int V, int F>
  __device__ __forceinline__ real row_above() const {
    if constexpr (ROW > 0) return w[L][CEN][ROW - 1][V][F];
    else if constexpr (W == 1) return w[L][CEN][ROW][V][F];   // a ghost row: any finite value
    else return (*edge)[half][L][wave_above][1][F][V][lane];
  }
  template <int L, int CEN, int ROW, int V, int F>
  __device__ __forceinline__ real row_below() const {
    if constexpr (ROW < RB - 1) return w[L][CEN][ROW + 1][V][F];
    else if constexpr (W == 1) return w[L][CEN][ROW][V][F];
    else return (*edge)[half][L][wave_below][0][F][V][lane];
  }

  template <int L, int PH, int ROW>
  struct Acc {
    MarchV& m;
    const Params& prm;
    int cx, cy;
    real out[VZ][NF];
    static constexpr int NEW = PH, CEN = (PH + NP - 1) % NP, OLD = (PH + NP - 2) % NP;

    template <int F, int D0, int D1, int D2, int V>
    __device__ __forceinline__ real ld() const {
      static_assert(((D0 != 0) + (D1 != 0) + (D2 != 0)) <= 1, "march kernel needs a star stencil");
      static_assert(D0 >= -1 && D0 <= 1 && D1 >= -1 && D1 <= 1 && D2 >= -1 && D2 <= 1, "radius 1");
      static_assert(!(XC && D0 == -1), "with the x-carry the plane behind is not kept");
      if constexpr (D0 == -1) return m.w[L][OLD][ROW][V][F];
      else if constexpr (D0 == 1) return m.w[L][NEW][ROW][V][F];
      else if constexpr (D1 == -1) return m.template row_above<L, CEN, ROW, V, F>();
      else if constexpr (D1 == 1) return m.template row_below<L, CEN, ROW, V, F>();
      else if constexpr (D2 != 0) {
        constexpr int S = V + D2;
        if constexpr (S >= 0 && S < VZ) return m.w[L][CEN][ROW][S][F];
        else if constexpr (S < 0) return lane_prev(m.w[L][CEN][ROW][VZ - 1][F]);
        else return fmr::lane_rot_next(m.w[L][CEN][ROW][0][F]);
      } else {
        return m.w[L][CEN][ROW][V][F];
      }
    }
    // the value of the previous lane (the z-site before this lane's first;
    // lane 0 takes lane 63's)
    __device__ __forceinline__ real lane_prev(real v) const { return fmr::lane_rot_prev(v); }
    // carried operation K of site V: its value at the previous plane
    template <int K, int V>
    __device__ __forceinline__ real xprev() const {
      return m.c[L][ROW][V][K];
    }
    template <int K, int V>
    __device__ __forceinline__ void xkeep(real v) {
      m.c[L][ROW][V][K] = v;
    }
    __device__ __forceinline__ double param(int i) const { return prm.v[i]; }
    // global coordinates wrapped into [0, n): ghost planes, rows and lanes
    // and rows past the grid end are periodic images of interior sites and
    // take those sites' index values
    template <int AXIS, int V>
    __device__ __forceinline__ int index() const {
      return AXIS == 0 ? wrap(cx, m.nx) : (AXIS == 1 ? wrap(cy, m.ny) : m.zg + V);
    }
    template <int F, int V>
    __device__ __forceinline__ void st(real v) {
      out[V][F] = v;
    }
  };

  template <int K>
  __device__ __forceinline__ void load_plane(const Fields<real>& in, int x, unsigned plane_bytes) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const auto rs = plane_rsrc(reinterpret_cast<const char*>(in.f[f]) + (size_t)x * plane_bytes, plane_bytes);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        real v[VZ];
        VecIO<real, VZ>::load(v, rs, zoff, roff[r]);
#pragma unroll
        for (int s = 0; s < VZ; ++s) nxt[K][r][s][f] = v[s];
      }
    }
  }

  // rows of level L+1 computed from window L: with one wave, the tile's rows
  // [L+1, ROWS-L-1) (known at compile time); with several and no edge rows,
  // every row of every wave -- the waves meet at a barrier each plane, so
  // rows the first and last wave could skip would save no time, only add
  // branches; with edge rows, the first and last wave skip them (MarchRole)
  static constexpr int row_lo(int L) { return (W == 1 || ROLE == ROLE_TOP) ? L + 1 : 0; }
  static constexpr int row_hi(int L) { return (W == 1 || ROLE == ROLE_BOTTOM) ? RB - L - 1 : RB; }

  template <int L, int PH, int ROW>
  __device__ __forceinline__ void level_row(const Params& prm, int q, int y0, const __amdgpu_buffer_rsrc_t (&orsrc)[NF],
                                            bool st_ok) {
    if constexpr (ROW < row_hi(L)) {
      const int trow = row0() + ROW;
      Acc<L, PH, ROW> a{*this, prm, q, y0 + trow, {}};
      P::template mpoint<VZ>(a);
      if constexpr (L + 1 < T) {
#pragma unroll
        for (int s = 0; s < VZ; ++s)
#pragma unroll
          for (int f = 0; f < NF; ++f) w[L + 1][PH][ROW][s][f] = a.out[s][f];
      } else if (st_ok && (W == 1 || ROLE != ROLE_ALL || (trow >= T && trow < ROWS - T))) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          real v[VZ];
#pragma unroll
          for (int s = 0; s < VZ; ++s) v[s] = a.out[s][f];
          VecIO<real, VZ>::store(v, orsrc[f], zst, roff[ROW]);
        }
      }
      level_row<L, PH, ROW + 1>(prm, q, y0, orsrc, st_ok);
    }
  }

  // levels L+1 .. T for one iteration.  Level T computes plane q when it
  // lies in [x0, xend) -- from x0 - 1 with the x-carry, whose value at x0 is
  // the one computed at x0 - 1 -- and stores it when it lies in [x0, xend).
  template <int L, int PH>
  __device__ __forceinline__ void levels(const Params& prm, int xin, int x0, int xend, int y0, const Fields<real>& out,
                                         unsigned plane_bytes) {
    if constexpr (L < T) {
      const int q = xin - (L + 1);
      if constexpr (L + 1 == T) {
        if (q >= x0 - (XC ? 1 : 0) && q < xend) {
          const bool st_ok = q >= x0;
          const int qs = st_ok ? q : x0;
          __amdgpu_buffer_rsrc_t orsrc[NF];
#pragma unroll
          for (int f = 0; f < NF; ++f)
            orsrc[f] = plane_rsrc(reinterpret_cast<char*>(out.f[f]) + (size_t)wrap(qs, nx) * plane_bytes, plane_bytes);
          level_row<L, PH, row_lo(L)>(prm, q, y0, orsrc, st_ok);
        }
      } else {
        __amdgpu_buffer_rsrc_t none[NF];
        level_row<L, PH, row_lo(L)>(prm, q, y0, none, false);
        levels<L + 1, PH>(prm, xin, x0, xend, y0, out, plane_bytes);
      }
    }
  }

  template <int PH>
  __device__ __forceinline__ void iter(const Fields<real>& in, const Fields<real>& out, const Params& prm, int xin,
                                       int xl, int x0, int xend, int y0, unsigned plane_bytes) {
    constexpr int CEN = (PH + NP - 1) % NP;
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int s = 0; s < VZ; ++s)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          w[0][PH][r][s][f] = nxt[0][r][s][f];
#pragma unroll
          for (int k = 0; k + 1 < PF; ++k) nxt[k][r][s][f] = nxt[k + 1][r][s][f];
        }
    if constexpr (W > 1) {
#pragma unroll
      for (int l = 0; l < T; ++l)
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int s = 0; s < VZ; ++s) {
            // the first wave's first row and the last wave's last row are
            // read by no wave when they are edge rows
            if constexpr (ROLE != ROLE_TOP) (*edge)[half][l][wave][0][f][s][lane] = w[l][CEN][0][s][f];
            if constexpr (ROLE != ROLE_BOTTOM) (*edge)[half][l][wave][1][f][s][lane] = w[l][CEN][RB - 1][s][f];
          }
      __syncthreads();
    }
    load_plane<PF - 1>(in, xl, plane_bytes);
    levels<0, PH>(prm, xin, x0, xend, y0, out, plane_bytes);
    half ^= 1;
  }

  template <int PH>
  __device__ __forceinline__ void iter_group(const Fields<real>& in, const Fields<real>& out, const Params& prm,
                                             int xin, int& xl, int x0, int xend, int y0, unsigned plane_bytes) {
    if constexpr (PH < NP) {
      iter<PH>(in, out, prm, xin + PH, xl, x0, xend, y0, plane_bytes);
      xl = xl + 1 == nx ? 0 : xl + 1;
      iter_group<PH + 1>(in, out, prm, xin, xl, x0, xend, y0, plane_bytes);
    }
  }

  // one run: planes [x0, xend) of tile `tile`
  __device__ __forceinline__ void run(const Fields<real>& in, const Fields<real>& out, int ntz, int tile, int x0,
                                      int xend, const Params& prm) {
    const int tz = tile % ntz;
    const int ty = tile / ntz;
    const int y0 = ty * (ROWS - 2 * T) - T;                // unwrapped y of tile row 0
    const int z0 = tz * (ZW - 2 * gz) - gz;                // unwrapped z of lane 0's first site
    const int zu = z0 + lane * VZ;                         // unwrapped z of this lane's first site
    zg = wrap(zu, nz);
    const unsigned plane_bytes = (unsigned)((size_t)ny * nz * sizeof(real));
#pragma unroll
    for (int r = 0; r < RB; ++r)
      roff[r] = (unsigned)__builtin_amdgcn_readfirstlane(wrap(y0 + row0() + r, ny) * nz * (int)sizeof(real));
    zoff = (unsigned)(zg * sizeof(real));
    zst = (lane * VZ >= gz && lane * VZ < ZW - gz && zu < nz) ? zoff : BUF_OOB;
    int xl = wrap(x0 - T, nx);
    load_plane<0>(in, xl, plane_bytes);
    if constexpr (PF > 1) {
      xl = xl + 1 == nx ? 0 : xl + 1;
      load_plane<PF - 1>(in, xl, plane_bytes);
    }
    xl = xl + 1 == nx ? 0 : xl + 1;
    const int iters = xend - x0 + 2 * T;
    for (int s = 0; s < iters; s += NP) iter_group<0>(in, out, prm, x0 - T + s, xl, x0, xend, y0, plane_bytes);
  }
};

// The (chunk, tile, plane) items of workgroup g of G, marched by m.
template <class M>
__device__ __forceinline__ void marchv_items(M& m, const Fields<typename M::real>& in,
                                             const Fields<typename M::real>& out, int nx, int nty, int ntz, int L,
                                             const Params& prm) {
  const int G = gridDim.x;
  const int g = xcd_order(blockIdx.x, G);
  march_runs(g, G, (long long)nty * ntz, nx, L,
             [&](int tile, int x0, int xend) { m.run(in, out, ntz, tile, x0, xend, prm); });
}

template <class M, int W>
__device__ __forceinline__ void marchv_setup(M& m, int wave, int nx, int ny, int nz, int gz, typename M::Edge* edge) {
  m.wave = wave;
  m.lane = threadIdx.x % WAVE;
  m.wave_above = wave == 0 ? W - 1 : wave - 1;
  m.wave_below = wave == W - 1 ? 0 : wave + 1;
  m.half = 0;
  m.nx = nx;
  m.ny = ny;
  m.nz = nz;
  m.gz = gz;
  m.edge = edge;
}

// Work split and workgroup order as march3d_kernel: x cut into chunks of L
// planes, (chunk, tile, plane) items listed chunk-major, workgroup g of G
// marches items [g * total / G, (g + 1) * total / G); consecutive workgroups
// on one XCD.  Requires nz to be a multiple of VZ (a lane's sites never wrap).
// With edge rows (EY > 0) the first and last wave run their own role of the
// march (MarchRole); every role meets the same barriers in the same order.
template <class P, int T, int BY, int W, int VZ, bool XC, int PF, int OCC, int EY>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(OCC)))
marchv_kernel(Fields<typename P::real> in, Fields<typename P::real> out, int nx, int ny, int nz, int nty, int ntz,
              int L, int gz, Params prm) {
  using M0 = MarchV<P, T, BY, W, VZ, XC, PF>;
  using Edge = typename M0::Edge;
  Edge* edge = nullptr;
  if constexpr (W > 1) {
    __shared__ Edge edge_buf;
    edge = &edge_buf;
  }
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  if constexpr (EY == 0) {
    M0 m;
    marchv_setup<M0, W>(m, wave, nx, ny, nz, gz, edge);
    marchv_items(m, in, out, nx, nty, ntz, L, prm);
  } else {
    using MT = MarchV<P, T, BY, W, VZ, XC, PF, EY, ROLE_TOP>;
    using MI = MarchV<P, T, BY, W, VZ, XC, PF, EY, ROLE_INNER>;
    using MB = MarchV<P, T, BY, W, VZ, XC, PF, EY, ROLE_BOTTOM>;
    if (wave == 0) {
      MT m;
      marchv_setup<MT, W>(m, wave, nx, ny, nz, gz, edge);
      marchv_items(m, in, out, nx, nty, ntz, L, prm);
    } else if (wave == W - 1) {
      MB m;
      marchv_setup<MB, W>(m, wave, nx, ny, nz, gz, edge);
      marchv_items(m, in, out, nx, nty, ntz, L, prm);
    } else {
      MI m;
      marchv_setup<MI, W>(m, wave, nx, ny, nz, gz, edge);
      marchv_items(m, in, out, nx, nty, ntz, L, prm);
    }
  }
}

}  // namespace fmr
