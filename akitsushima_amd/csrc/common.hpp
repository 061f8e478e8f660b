// Shared device-side definitions of the akitsushima_amd force engine (gfx950).
//
// Layout conventions (all FP64, row-major, C-contiguous):
//   lattice rows are lattice vectors, cart = frac @ lattice; rij = r_j - r_i;
//   positions (N, 3); forces (N, 3); stress (3, 3) = -(1/V) sum_bonds G (x) r.
// A "set" is one or more periodic structures concatenated atom-wise; atom a
// belongs to structure sid[a], whose geometry is geoms[sid[a]].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ak {

constexpr int kWave = 64;

struct Mat3 {
  double m[9];
};

struct Geom {
  double lat[9];
  double inv[9];
  double volume;
  int nc[3];      // cell-grid dimensions of the neighbour list
  int cell_off;   // first global cell id of this structure
  int atom_off;   // first atom of this structure
  int natoms;
  int pad;
};

// Minimum image with contraction allowed (force evaluation).
__device__ __forceinline__ void mic_fast(const Geom& g, const double* __restrict__ pos, int i,
                                         int j, double& dx, double& dy, double& dz) {
  double rx = pos[3 * j] - pos[3 * i], ry = pos[3 * j + 1] - pos[3 * i + 1],
         rz = pos[3 * j + 2] - pos[3 * i + 2];
  double sx = rx * g.inv[0] + ry * g.inv[3] + rz * g.inv[6];
  double sy = rx * g.inv[1] + ry * g.inv[4] + rz * g.inv[7];
  double sz = rx * g.inv[2] + ry * g.inv[5] + rz * g.inv[8];
  sx -= rint(sx);
  sy -= rint(sy);
  sz -= rint(sz);
  dx = sx * g.lat[0] + sy * g.lat[3] + sz * g.lat[6];
  dy = sx * g.lat[1] + sy * g.lat[4] + sz * g.lat[7];
  dz = sx * g.lat[2] + sy * g.lat[5] + sz * g.lat[8];
}

// A cell whose lattice and inverse have exactly zero off-diagonal entries
// (uniform over a block: evaluate once).
__device__ __forceinline__ bool geom_diagonal(const Geom& g) {
  return g.lat[1] == 0.0 && g.lat[2] == 0.0 && g.lat[3] == 0.0 && g.lat[5] == 0.0 &&
         g.lat[6] == 0.0 && g.lat[7] == 0.0 && g.inv[1] == 0.0 && g.inv[2] == 0.0 &&
         g.inv[3] == 0.0 && g.inv[5] == 0.0 && g.inv[6] == 0.0 && g.inv[7] == 0.0;
}

// mic_fast, with the products by the zero off-diagonal entries of a
// diagonal cell left out (the same values up to the sign of zero).
__device__ __forceinline__ void mic_cell(const Geom& g, bool diagonal,
                                         const double* __restrict__ pos, int i, int j,
                                         double& dx, double& dy, double& dz) {
  if (!diagonal) {
    mic_fast(g, pos, i, j, dx, dy, dz);
    return;
  }
  double sx = (pos[3 * j] - pos[3 * i]) * g.inv[0];
  double sy = (pos[3 * j + 1] - pos[3 * i + 1]) * g.inv[4];
  double sz = (pos[3 * j + 2] - pos[3 * i + 2]) * g.inv[8];
  dx = (sx - rint(sx)) * g.lat[0];
  dy = (sy - rint(sy)) * g.lat[4];
  dz = (sz - rint(sz)) * g.lat[8];
}

// Minimum-image displacement pj - pi in the structure's cell, with the
// reference's operation order (engine/neighborlist.py:350-361): fractional
// components rounded half-to-even (rint), no fused multiply-add, so the
// neighbour-list predicate r2 < rl2 matches the reference bit for bit.
__device__ __forceinline__ void mic_exact(const double* __restrict__ lat,
                                          const double* __restrict__ inv, double pix,
                                          double piy, double piz, double pjx, double pjy,
                                          double pjz, double& dx, double& dy, double& dz) {
#pragma clang fp contract(off)
  double rx = pjx - pix, ry = pjy - piy, rz = pjz - piz;
  double sx = rx * inv[0] + ry * inv[3] + rz * inv[6];
  double sy = rx * inv[1] + ry * inv[4] + rz * inv[7];
  double sz = rx * inv[2] + ry * inv[5] + rz * inv[8];
  sx -= rint(sx);
  sy -= rint(sy);
  sz -= rint(sz);
  dx = sx * lat[0] + sy * lat[3] + sz * lat[6];
  dy = sx * lat[1] + sy * lat[4] + sz * lat[7];
  dz = sx * lat[2] + sy * lat[5] + sz * lat[8];
}

__device__ __forceinline__ double r2_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy + dz * dz;
}

// Clamped exp of the reference (engine/_numba_common.py:12-32).
__device__ __forceinline__ double nb_exp(double x) {
  if (x > 500.0) return exp(500.0);
  if (x < -500.0) return 0.0;
  return exp(x);
}

// Reciprocal and reciprocal square root from the hardware estimates with two
// Newton steps each (a few ulp; the correctly rounded division and square
// root cost about three times the instructions).  x > 0 and finite.
__device__ __forceinline__ double fast_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ double fast_rsq(double x) {
  const double hx = 0.5 * x;
  double y = __builtin_amdgcn_rsq(x);
  y *= fma(-hx, y * y, 1.5);
  y *= fma(-hx, y * y, 1.5);
  return y;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (uint64_t(1) << __lane_id()) - 1;
}

// DPP quad permutation of a double: lane q of every quad receives lane
// ((CTRL >> 2q) & 3) of its quad.  A VALU move, no LDS round trip.
template <int CTRL>
__device__ __forceinline__ double dpp_quad(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ int dpp_quad(int x) {
  return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, false);
}

// Value of lane `seg + s` of the calling lane's segment of L lanes.  For
// L = 4 the segment is a DPP quad and s must be a compile-time constant after
// unrolling (broadcast of lane s: quad_perm [s, s, s, s]).
template <int L, class T>
__device__ __forceinline__ T seg_get(T x, int seg, int s) {
  if constexpr (L == 4) {
    switch (s) {
      case 0: return dpp_quad<0x00>(x);
      case 1: return dpp_quad<0x55>(x);
      case 2: return dpp_quad<0xAA>(x);
      default: return dpp_quad<0xFF>(x);
    }
  } else {
    return __shfl(x, seg + s, kWave);
  }
}

// Sum over a segment of L consecutive lanes (L a power of two <= 64); every
// lane of the segment receives the total.  The last two butterfly steps are
// DPP quad permutations (xor 1: [1,0,3,2], xor 2: [2,3,0,1]).
template <int L>
__device__ __forceinline__ double seg_sum(double v) {
#pragma unroll
  for (int o = L / 2; o >= 4; o >>= 1) v += __shfl_xor(v, o, kWave);
  if constexpr (L >= 4) v += dpp_quad<0x4E>(v);
  if constexpr (L >= 2) v += dpp_quad<0xB1>(v);
  return v;
}

// DPP row rotation of a double within rows of 16 lanes (row_ror:N is
// 0x120 + N).
template <int CTRL>
__device__ __forceinline__ double dpp_row(double x) {
  int lo = __double2loint(x), hi = __double2hiint(x);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Position of atom a in row j of a neighbour list (rows ascending, stride a
// power of two of at least four, cnt[j] clamped to the stride), or -1 when
// a is not in the row: the number of entries below a, counted over 16-byte
// loads, the first 16 entries read with cnt[j] rather than after it.
__device__ __forceinline__ int row_rank(int a, int j, const int* __restrict__ cnt,
                                        const int* __restrict__ nbr, int stride) {
  const int4* row = reinterpret_cast<const int4*>(nbr + (size_t)j * stride);
  constexpr int kHead = 4;
  const int head = min(kHead, stride >> 2);
  int4 h[kHead];
#pragma unroll
  for (int u = 0; u < kHead; ++u) h[u] = u < head ? row[u] : make_int4(0, 0, 0, 0);
  const int cj = min(cnt[j], stride);
  int below = 0;
  bool found = false;
  auto scan4 = [&](const int4& q, int m) {
    const int v[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool in = m + e < cj;
      below += (in && v[e] < a) ? 1 : 0;
      found |= in && v[e] == a;
    }
  };
#pragma unroll
  for (int u = 0; u < kHead; ++u) scan4(h[u], 4 * u);
#pragma unroll 4
  for (int m = 4 * kHead; m < cj; m += 4) scan4(row[m >> 2], m);
  return found ? below : -1;
}

// Site gradients in HBM: slot s = a * stride + k of a neighbour list holds
// G[a, k] = dU_a / dv_ak as the record (x, y, z, 0) at G + kGRec * s.  A row
// holds every candidate within r_cut + skin and often only a few bonds
// inside r_cut, so the bond kernels write and the gathers read scattered
// slots: one 32-byte record costs one memory sector per slot where three
// component planes cost three.
constexpr int kGRec = 4;
__device__ __forceinline__ void g_put(double* __restrict__ G, size_t s, double x, double y,
                                      double z) {
  double2* p = reinterpret_cast<double2*>(G + kGRec * s);
  p[0] = make_double2(x, y);
  p[1] = make_double2(z, 0.0);
}
__device__ __forceinline__ void g_get(const double* __restrict__ G, size_t s, double& x,
                                      double& y, double& z) {
  const double2 u = reinterpret_cast<const double2*>(G + kGRec * s)[0];
  x = u.x;
  y = u.y;
  z = G[kGRec * s + 2];
}

// Lane l's value of x (l wave-uniform).
__device__ __forceinline__ double readlane_d(double x, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}

// Maximum over the 64 lanes of a wavefront, every lane receives it.
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

// Exchange of one recursive-halving step: the value of the partner lane
// (row_half_mirror: lane i <-> 7 - i within 8; quad xor 1, xor 2; row_ror 8:
// xor 8), all within a row of 16 lanes.
template <int STEP>
__device__ __forceinline__ double halving_partner(double x) {
  if constexpr (STEP == 0) return dpp_row<0x141>(x);
  else if constexpr (STEP == 1) return dpp_quad<0xB1>(x);
  else if constexpr (STEP == 2) return dpp_quad<0x4E>(x);
  else return dpp_row<0x128>(x);
}

// Keep half of the H + H values w (the upper half where `upper`) and add
// the partner lane's copy of that half.
template <int STEP, int H>
__device__ __forceinline__ void halving_step(double (&w)[16], bool upper) {
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double lo = halving_partner<STEP>(w[i]), hi = halving_partner<STEP>(w[H + i]);
    w[i] = upper ? w[H + i] + hi : w[i] + lo;
  }
}

// Totals over the 64 lanes of a wavefront of NV <= 16 values per lane, by
// recursive halving: four exchange steps inside each row of 16 lanes (lane
// bit 2, 0, 1, 3 selects the half a lane keeps) leave each lane one value,
// two cross-row shuffles sum it over the rows.  Lane l then holds the total
// of value idx = 8 b2 + 4 b0 + 2 b1 + b3 (b the bits of l); lanes 0-15 hold
// the 16 indices once each.  Fixed order, so bitwise reproducible.
template <int NV>
__device__ __forceinline__ double wave_sums(const double (&v)[NV], int& idx) {
  static_assert(NV >= 1 && NV <= 16, "at most 16 values");
  const int lane = __lane_id();
  double w[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) w[i] = i < NV ? v[i] : 0.0;
  halving_step<0, 8>(w, (lane & 4) != 0);
  halving_step<1, 4>(w, (lane & 1) != 0);
  halving_step<2, 2>(w, (lane & 2) != 0);
  halving_step<3, 1>(w, (lane & 8) != 0);
  double t = w[0];
  t += __shfl_xor(t, 16, kWave);
  t += __shfl_xor(t, 32, kWave);
  idx = ((lane & 4) << 1) | ((lane & 1) << 2) | (lane & 2) | ((lane >> 3) & 1);
  return t;
}

// The dispatcher hands consecutive workgroups to the 8 XCDs in turn, and each
// XCD has its own L2.  Chunks (and atom blocks) are numbered so that XCD x
// gets the contiguous x-th eighth of them: neighbouring atoms, whose
// positions and bond gradients a chunk reads, then stay in one L2.  A
// bijection on [0, nb): the last nb mod 8 blocks keep their numbers.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int q = nb / kXcds;
  return b < q * kXcds ? (b % kXcds) * q + b / kXcds : b;
}

// Deterministic block reduction of `n` doubles per thread (blockDim.x a
// multiple of 64, <= 1024); thread 0 receives the sums in out[].
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* smem /* NV*16 */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    v[k] = x;
  }
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) smem[k * 16 + wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += smem[k * 16 + w];
      v[k] = s;
    }
  }
}

}  // namespace ak
