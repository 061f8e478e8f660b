// Cell-list neighbour list on the device (after the reference's
// NeighbourList._build_cell_linked, engine/neighborlist.py:935-1024, and its
// pair search _cell_linked_pairs_numpy :312-373).
//
// 1. cell_of (in the first histogram pass): per atom, its cell in the
//    structure's grid (nc_i = floor(d_i / r_list), fractional coordinates
//    wrapped into [0, 1)).
// 2. Stable LSD radix sort of (cell, atom) by cell with 8-bit digits, free of
//    atomics: digit counts per wavefront come from ballot match masks, block
//    histograms are scanned globally, and each item's rank is its position
//    among equal digits earlier in its block.  Atoms stay in ascending order
//    inside each cell.
// 3. k_nl_search: one wavefront per cell.  The wave collects the candidates
//    of the canonical stencil (per-axis offsets {0} / {0, 1} / {-1, 0, 1} for
//    nc = 1 / 2 / >= 3, engine/neighborlist.py:270-309), sorts them by atom
//    index (bitonic, in LDS), then tests every atom of the cell against them
//    with the reference's exact predicate r2 < rl2 and r2 > 1e-20, compacting
//    hits by ballot.  Rows therefore come out sorted by neighbour index, the
//    canonical order of engine/neighborlist.py:1166-1181, with no per-row
//    sort.
// 4. k_nl_rev: reverse slot of every entry (binary search in the sorted row).
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ak {

// ---------------------------------------------------------------------------
// Geometry (host)
// ---------------------------------------------------------------------------

void invert3(const double* m, double* inv) {
  double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7],
         i = m[8];
  double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  double det = a * A + b * B + c * C;
  AK_REQUIRE(det != 0.0 && std::isfinite(det), AK_ERR_ARG, "singular lattice");
  double id = 1.0 / det;
  inv[0] = A * id;
  inv[1] = -(b * i - c * h) * id;
  inv[2] = (b * f - c * e) * id;
  inv[3] = B * id;
  inv[4] = (a * i - c * g) * id;
  inv[5] = -(a * f - c * d) * id;
  inv[6] = C * id;
  inv[7] = -(a * h - b * g) * id;
  inv[8] = (a * e - b * d) * id;
}

CellInfo cell_info(const double* L) {
  CellInfo ci;
  double det = L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) +
               L[2] * (L[3] * L[7] - L[4] * L[6]);
  ci.volume = std::fabs(det);
  double cn[3];
  for (int ax = 0; ax < 3; ++ax) {
    const double* u = L + 3 * ((ax + 1) % 3);
    const double* v = L + 3 * ((ax + 2) % 3);
    double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2],
           z = u[0] * v[1] - u[1] * v[0];
    cn[ax] = std::sqrt(x * x + y * y + z * z);
  }
  AK_REQUIRE(ci.volume > 0.0 && cn[0] > 0.0 && cn[1] > 0.0 && cn[2] > 0.0, AK_ERR_ARG,
             "degenerate (singular) lattice: the three lattice vectors must be linearly "
             "independent");
  for (int ax = 0; ax < 3; ++ax) ci.perp[ax] = ci.volume / cn[ax];
  ci.min_width = std::min(ci.perp[0], std::min(ci.perp[1], ci.perp[2]));
  return ci;
}

void SetDev::setup(int S, const int64_t* offsets, const double* lattices, const double* invs,
                   double r_list, hipStream_t st, bool check_min_image, double r_grid) {
  AK_REQUIRE(S >= 1, AK_ERR_ARG, "need at least one structure");
  // Repeated calls on an unchanged set (the common MD / benchmark case)
  // skip the host-side geometry work entirely.
  if (S == n_struct && r_list == setup_r_list && r_grid == setup_r_grid &&
      check_min_image == setup_checked &&
      (invs != nullptr) == setup_had_inv && (int)geoms.size() == S && d_geoms.p && d_sid.p) {
    bool same = true;
    for (int s = 0; s < S && same; ++s) {
      same = std::memcmp(geoms[s].lat, lattices + 9 * s, 9 * sizeof(double)) == 0 &&
             (!invs || std::memcmp(geoms[s].inv, invs + 9 * s, 9 * sizeof(double)) == 0) &&
             offsets[s] == atom_off[s] && offsets[s + 1] == atom_off[s + 1];
    }
    if (same) return;
  }
  setup_r_list = -1.0;  // set again only once this setup has succeeded
  n_struct = S;
  geoms.assign(S, Geom{});
  atom_off.assign(S + 1, 0);
  int cells = 0;
  for (int s = 0; s < S; ++s) {
    Geom& g = geoms[s];
    std::memcpy(g.lat, lattices + 9 * s, 9 * sizeof(double));
    if (invs)
      std::memcpy(g.inv, invs + 9 * s, 9 * sizeof(double));
    else
      invert3(g.lat, g.inv);
    CellInfo ci = cell_info(g.lat);
    g.volume = ci.volume;
    if (check_min_image && r_list >= ci.min_width / 2.0) {
      char buf[512];
      snprintf(buf, sizeof buf,
               "structure %d: r_cut + skin_width (%.4f A) must be < half the smallest "
               "perpendicular cell width (%.4f A) for the minimum-image convention to be "
               "valid. Perpendicular widths are [%.4f, %.4f, %.4f] A",
               s, r_list, ci.min_width / 2.0, ci.perp[0], ci.perp[1], ci.perp[2]);
      throw Error(AK_ERR_MIN_IMAGE, buf);
    }
    for (int ax = 0; ax < 3; ++ax) {
      int nc = (int)std::floor(ci.perp[ax] / std::max(r_list, r_grid));
      g.nc[ax] = nc < 1 ? 1 : nc;
    }
    g.cell_off = cells;
    long long nc_tot = (long long)g.nc[0] * g.nc[1] * g.nc[2];
    AK_REQUIRE(cells + nc_tot < (1LL << 30), AK_ERR_CAPACITY, "too many cells");
    cells += (int)nc_tot;
    AK_REQUIRE(offsets[s + 1] >= offsets[s], AK_ERR_ARG, "atom offsets must not decrease");
    g.atom_off = (int)offsets[s];
    g.natoms = (int)(offsets[s + 1] - offsets[s]);
    atom_off[s] = (int)offsets[s];
  }
  atom_off[S] = (int)offsets[S];
  AK_REQUIRE(offsets[0] == 0, AK_ERR_ARG, "atom offsets must start at 0");
  n_atoms = (int)offsets[S];
  n_cells = cells;
  upload_geoms(st);
  if (uploaded_off != atom_off || d_sid.p == nullptr) {
    uploaded_off = atom_off;
    std::vector<int> sid(n_atoms > 0 ? n_atoms : 1, 0);
    for (int s = 0; s < S; ++s)
      for (int a = atom_off[s]; a < atom_off[s + 1]; ++a) sid[a] = s;
    d_sid.ensure(sizeof(int) * sid.size());
    AK_HIP(hipMemcpyAsync(d_sid.p, sid.data(), sizeof(int) * sid.size(),
                          hipMemcpyHostToDevice, st));
    AK_HIP(hipStreamSynchronize(st));
    chunk_tables.clear();
  }
  setup_r_list = r_list;
  setup_r_grid = r_grid;
  setup_checked = check_min_image;
  setup_had_inv = invs != nullptr;
}

SetDev::Chunks& SetDev::chunks(int size, hipStream_t st) {
  for (auto& c : chunk_tables)
    if (c->size == size) return *c;
  auto c = std::make_unique<Chunks>();
  c->size = size;
  std::vector<int> rng, first(n_struct + 1, 0);
  for (int s = 0; s < n_struct; ++s) {
    first[s] = (int)rng.size() / 2;
    for (int a = atom_off[s]; a < atom_off[s + 1]; a += size) {
      rng.push_back(a);
      rng.push_back(std::min(a + size, atom_off[s + 1]));
    }
  }
  first[n_struct] = (int)rng.size() / 2;
  c->n = first[n_struct];
  c->rng.ensure(sizeof(int) * std::max<size_t>(2, rng.size()));
  c->first.ensure(sizeof(int) * first.size());
  c->partial.ensure(sizeof(double) * 10 * std::max(1, c->n));
  if (!rng.empty())
    AK_HIP(hipMemcpyAsync(c->rng.p, rng.data(), sizeof(int) * rng.size(), hipMemcpyHostToDevice,
                          st));
  AK_HIP(hipMemcpyAsync(c->first.p, first.data(), sizeof(int) * first.size(),
                        hipMemcpyHostToDevice, st));
  AK_HIP(hipStreamSynchronize(st));
  chunk_tables.push_back(std::move(c));
  return *chunk_tables.back();
}

void SetDev::upload_geoms(hipStream_t st) {
  if (uploaded_geoms.size() == geoms.size() && d_geoms.p &&
      std::memcmp(uploaded_geoms.data(), geoms.data(), sizeof(Geom) * geoms.size()) == 0)
    return;
  uploaded_geoms = geoms;
  d_geoms.ensure(sizeof(Geom) * geoms.size());
  AK_HIP(hipMemcpyAsync(d_geoms.p, geoms.data(), sizeof(Geom) * geoms.size(),
                        hipMemcpyHostToDevice, st));
  std::vector<int> co(n_struct + 1);
  for (int s = 0; s < n_struct; ++s) co[s] = geoms[s].cell_off;
  co[n_struct] = n_cells;
  d_cell_offs.ensure(sizeof(int) * co.size());
  AK_HIP(hipMemcpyAsync(d_cell_offs.p, co.data(), sizeof(int) * co.size(),
                        hipMemcpyHostToDevice, st));
  // the host vectors are copied by the time the stream reaches them only if
  // they outlive the call; synchronise to keep it simple
  AK_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool gated_off(const int* gate) { return gate && *gate == 0; }

// Cell of atom a in its structure's grid.
__device__ __forceinline__ unsigned cell_of(int a, const double* __restrict__ pos,
                                            const int* __restrict__ sid,
                                            const Geom* __restrict__ geoms) {
  const Geom& g = geoms[sid[a]];
  double p0 = pos[3 * a], p1 = pos[3 * a + 1], p2 = pos[3 * a + 2];
  int c[3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    double f = p0 * g.inv[b] + p1 * g.inv[3 + b] + p2 * g.inv[6 + b];
    f = f - floor(f);
    double t = floor(f * g.nc[b]);
    int k = isfinite(t) ? (int)t : 0;
    k = k < 0 ? 0 : (k > g.nc[b] - 1 ? g.nc[b] - 1 : k);
    c[b] = k;
  }
  return (unsigned)(g.cell_off + (c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2]);
}

// Blocks of the gated (captured MD) cell search and reverse-slot kernels,
// which loop over their work: a step without a rebuild dispatches this many
// blocks that return at once instead of one per cell or per 256 entries.
constexpr int kGatedBlocks = 4096;
// The cell search's cap (one wavefront per block): a rebuild step gains more
// from the wider grid than the other steps lose.
constexpr int kGatedSearchBlocks = 8192;
// The reverse-slot pass's cap (256 threads per block).
constexpr int kGatedRevBlocks = 2048;

// Counting sort of the atoms by cell: k_cell_count (cell and arrival rank of
// every atom, counts by atomic add), then k_cell_scan_place when the counts
// fit in one block's LDS, else k_cell_scan (exclusive prefix of the counts
// inside blocks of CS_SCAN cells) and k_cell_place (block offsets,
// cell_start and the atom lists).  Atoms inside
// a cell end in arrival order, which varies from run to run; the search
// sorts each stencil's candidates by atom index and writes every atom's row
// on its own, so the list does not depend on that order.
constexpr int CS_T = 256;      // threads per block of the atom kernels
constexpr int CS_SCAN = 1024;  // cells per scan block
// block sums that k_cell_place scans in LDS itself; more take a k_scan_excl
constexpr int CS_PLACE_LDS = 4096;

// key[a] = cell of atom a, rank[a] = its slot among the atoms of that cell;
// count[] is zero on entry (k_cell_scan clears it again).  With sg.flag
// (a captured MD step), every block evaluates the skin test itself and
// block 0 writes the flag the later kernels read; else *gate decides.
__global__ __launch_bounds__(CS_T) void k_cell_count(int n, const double* __restrict__ pos,
                                                     const int* __restrict__ sid,
                                                     const Geom* __restrict__ geoms,
                                                     unsigned* __restrict__ key,
                                                     unsigned* __restrict__ rank,
                                                     unsigned* __restrict__ count,
                                                     const int* gate, SkinGate sg) {
  if (sg.flag) {
    if (!skin_gate_eval(sg, blockIdx.x == 0)) return;
  } else if (gated_off(gate)) {
    return;
  }
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
    const unsigned c = cell_of(a, pos, sid, geoms);
    key[a] = c;
    rank[a] = atomicAdd(&count[c], 1u);
  }
}

// pre[c] = sum of count[c'] for c' in [CS_SCAN * b, c), b = c / CS_SCAN, for
// c in [0, C] (count[C] taken as 0); blk[b] = the block's total.  Clears
// count[] for the next sort.
__global__ __launch_bounds__(CS_SCAN) void k_cell_scan(int C, unsigned* __restrict__ count,
                                                       unsigned* __restrict__ pre,
                                                       unsigned* __restrict__ blk,
                                                       const int* gate) {
  if (gated_off(gate)) return;
  __shared__ unsigned wsum[CS_SCAN / 64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int c = blockIdx.x * CS_SCAN + tid;
  unsigned v = 0u;
  if (c < C) {
    v = count[c];
    count[c] = 0u;
  }
  unsigned inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  unsigned before = 0u;
  for (int w = 0; w < wid; ++w) before += wsum[w];
  if (c <= C) pre[c] = before + inc - v;
  if (tid == CS_SCAN - 1) blk[blockIdx.x] = before + inc;
}

// cell_start[c] = pre[c] + (offset of block c / CS_SCAN) for c in [0, C],
// and cell_atoms[cell_start[key[a]] + rank[a]] = a.  The nb block sums are
// scanned here in LDS (dynamic, nb entries), or, with scanned != 0, blk
// already holds their exclusive prefix.
__global__ __launch_bounds__(CS_T) void k_cell_place(int n, int C, const unsigned* __restrict__ key,
                                                     const unsigned* __restrict__ rank,
                                                     const unsigned* __restrict__ pre,
                                                     const unsigned* __restrict__ blk, int nb,
                                                     int scanned, int* __restrict__ cell_start,
                                                     unsigned* __restrict__ cell_atoms,
                                                     const int* gate) {
  if (gated_off(gate)) return;
  extern __shared__ unsigned boff[];
  __shared__ unsigned wsum[CS_T / 64];
  __shared__ unsigned carry;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (!scanned) {
    if (tid == 0) carry = 0u;
    __syncthreads();
    for (int b0 = 0; b0 < nb; b0 += CS_T) {
      const int b = b0 + tid;
      const unsigned v = b < nb ? blk[b] : 0u;
      unsigned inc = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
      }
      if (lane == 63) wsum[wid] = inc;
      __syncthreads();
      unsigned before = carry;
      for (int w = 0; w < wid; ++w) before += wsum[w];
      if (b < nb) boff[b] = before + inc - v;
      __syncthreads();
      if (tid == CS_T - 1) carry = before + inc;
      __syncthreads();
    }
  }
  const unsigned* off = scanned ? blk : boff;
  const int64_t end = max((int64_t)n, (int64_t)C + 1);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + tid; t < end;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t <= C) cell_start[t] = (int)(pre[t] + off[t / CS_SCAN]);
    if (t < n) {
      const unsigned c = key[t];
      cell_atoms[pre[c] + off[c / CS_SCAN] + rank[t]] = (unsigned)t;
    }
  }
}

// k_cell_scan and k_cell_place in one pass, for cell counts that fit in LDS
// (C + 1 entries, dynamic): every block copies the counts into LDS, scans
// them, writes its share of cell_start (grid-stride) and places its atoms.
// The counts stay as they are; k_nl_rev, which follows every sort, clears
// them.
__global__ __launch_bounds__(CS_T) void k_cell_scan_place(int n, int C,
                                                          const unsigned* __restrict__ key,
                                                          const unsigned* __restrict__ rank,
                                                          const unsigned* __restrict__ count,
                                                          int* __restrict__ cell_start,
                                                          unsigned* __restrict__ cell_atoms,
                                                          const int* gate) {
  if (gated_off(gate)) return;
  extern __shared__ unsigned pre[];
  __shared__ unsigned wsum[CS_T / 64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int m = C + 1;
  // counts -> LDS, 16 bytes per load and eight loads in flight per thread
  // (count[] holds a multiple of four entries; those from C on read as 0)
  const int m4 = (m + 3) >> 2;
  const uint4* c4 = reinterpret_cast<const uint4*>(count);
  for (int i0 = tid; i0 < m4; i0 += 8 * CS_T) {
    uint4 q[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * CS_T;
      q[u] = i < m4 ? c4[i] : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = 4 * (i0 + u * CS_T);
      if (i < m) pre[i] = i < C ? q[u].x : 0u;
      if (i + 1 < m) pre[i + 1] = i + 1 < C ? q[u].y : 0u;
      if (i + 2 < m) pre[i + 2] = i + 2 < C ? q[u].z : 0u;
      if (i + 3 < m) pre[i + 3] = i + 3 < C ? q[u].w : 0u;
    }
  }
  __syncthreads();
  // thread tid owns the run [lo, hi); exclusive scan of the run sums
  const int per = (m + CS_T - 1) / CS_T;
  const int lo = min(m, tid * per), hi = min(m, lo + per);
  unsigned s = 0u;
  for (int i = lo; i < hi; ++i) s += pre[i];
  unsigned inc = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  unsigned run = inc - s;
  for (int w = 0; w < wid; ++w) run += wsum[w];
  for (int i = lo; i < hi; ++i) {
    const unsigned v = pre[i];
    pre[i] = run;
    run += v;
  }
  __syncthreads();
  const int64_t end = max((int64_t)n, (int64_t)m);
  for (int64_t t = (int64_t)blockIdx.x * CS_T + tid; t < end; t += (int64_t)gridDim.x * CS_T) {
    if (t < m) cell_start[t] = (int)pre[t];
    if (t < n) cell_atoms[pre[key[t]] + rank[t]] = (unsigned)t;
  }
}

// In-place exclusive scan of m unsigned values by one block of 1024 threads;
// *total (if given) receives the sum.  Thread t owns the contiguous run
// [t * per, (t + 1) * per) and reads it eight independent loads at a time;
// the per-thread sums are scanned with wavefront shuffles and one LDS pass.
__global__ __launch_bounds__(1024) void k_scan_excl(unsigned* __restrict__ data, int m,
                                                    unsigned* total, const int* gate) {
  if (gated_off(gate)) return;
  __shared__ unsigned wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int per = (m + 1023) / 1024;
  const int lo = min(m, tid * per), hi = min(m, lo + per);
  unsigned s = 0;
  for (int i = lo; i < hi; i += 8) {
    unsigned v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = i + q < hi ? data[i + q] : 0u;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += v[q];
  }
  // inclusive scan of s over the block
  unsigned inc = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    unsigned y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  unsigned before = 0;
  for (int w = 0; w < wid; ++w) before += wsum[w];
  unsigned run = before + inc - s;  // exclusive prefix of this thread's run
  for (int i = lo; i < hi; i += 8) {
    unsigned v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = i + q < hi ? data[i + q] : 0u;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (i + q < hi) data[i + q] = run;
      run += v[q];
    }
  }
  if (total && tid == 1023) *total = before + inc;
}

__device__ __forceinline__ int find_struct(const int* __restrict__ offs, int S, int c) {
  int lo = 0, hi = S;  // offs[lo] <= c < offs[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (offs[mid] <= c)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// One wavefront (one 64-thread block) per cell.
__device__ __forceinline__ void search_cell(
    int c, int S, const int* __restrict__ cell_offs, const Geom* __restrict__ geoms,
    const int* __restrict__ cell_start, const unsigned* __restrict__ cell_atoms,
    const double* __restrict__ pos, double rl2, int cap, int cmax, int* __restrict__ cnt,
    int* __restrict__ nbr, int* __restrict__ stats, int* __restrict__ cand_n,
    unsigned* __restrict__ cand, int* __restrict__ spre, int* __restrict__ sstart) {
  const int lane = threadIdx.x;
  const int s = find_struct(cell_offs, S, c);
  const Geom& g = geoms[s];
  const int nc0 = g.nc[0], nc1 = g.nc[1], nc2 = g.nc[2];
  const int l = c - g.cell_off;
  const int c0 = l / (nc1 * nc2), c1 = (l / nc2) % nc1, c2 = l % nc2;
  const int n0 = nc0 == 1 ? 1 : (nc0 == 2 ? 2 : 3);
  const int n1 = nc1 == 1 ? 1 : (nc1 == 2 ? 2 : 3);
  const int n2 = nc2 == 1 ? 1 : (nc2 == 2 ? 2 : 3);
  const int nst = n0 * n1 * n2;
  int size = 0, cid = 0, cst = 0;
  if (lane < nst) {
    int a0 = lane / (n1 * n2), a1 = (lane / n2) % n1, a2 = lane % n2;
    int o0 = n0 == 3 ? a0 - 1 : a0, o1 = n1 == 3 ? a1 - 1 : a1, o2 = n2 == 3 ? a2 - 1 : a2;
    int j0 = (c0 + o0 + nc0) % nc0, j1 = (c1 + o1 + nc1) % nc1, j2 = (c2 + o2 + nc2) % nc2;
    cid = g.cell_off + (j0 * nc1 + j1) * nc2 + j2;
    cst = cell_start[cid];
    size = cell_start[cid + 1] - cst;
  }
  // inclusive wave prefix of the stencil cell sizes
  int inc = size;
#pragma unroll
  for (int o = 1; o < 32; o <<= 1) {
    int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  const int ncand = __shfl(inc, 31, 64);
  if (lane < nst) {
    spre[lane + 1] = inc;
    sstart[lane] = cst;  // first sorted position of the stencil cell
  }
  if (lane == 0) spre[0] = 0;
  const int first = cell_start[c], last = cell_start[c + 1];
  if (cand_n && lane == 0) cand_n[c] = ncand;
  // the cell's own atoms, one per lane (the first 64), read alongside the
  // candidates rather than one after another in the row loop
  int h_i = 0;
  double h_x = 0.0, h_y = 0.0, h_z = 0.0;
  if (lane < last - first) {
    h_i = (int)cell_atoms[first + lane];
    h_x = pos[3 * h_i];
    h_y = pos[3 * h_i + 1];
    h_z = pos[3 * h_i + 2];
  }
  if (ncand > cmax) {
    if (lane == 0) stats[1] = max(stats[1], ncand);  // candidate overflow
    for (int t = first + lane; t < last; t += 64) cnt[cell_atoms[t]] = 0;
    return;
  }
  __syncthreads();
  int npow = 1;
  while (npow < ncand) npow <<= 1;
  for (int q = lane; q < npow; q += 64) {
    unsigned v = 0xFFFFFFFFu;
    if (q < ncand) {
      int lo = 0, hi = nst;  // largest st with spre[st] <= q
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (spre[mid] <= q)
          lo = mid;
        else
          hi = mid;
      }
      v = cell_atoms[sstart[lo] + (q - spre[lo])];
    }
    cand[q] = v;
  }
  __syncthreads();
  // bitonic sort of cand[0, npow) ascending
  for (int k = 2; k <= npow; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int e = lane; e < npow; e += 64) {
        int p = e ^ j;
        if (p > e) {
          unsigned x = cand[e], y = cand[p];
          bool up = (e & k) == 0;
          if ((x > y) == up) {
            cand[e] = y;
            cand[p] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int t = first; t < last; ++t) {
    const int h = t - first;  // wave-uniform
    int i;
    double pix, piy, piz;
    if (h < 64) {
      i = __builtin_amdgcn_readlane(h_i, h);
      pix = readlane_d(h_x, h);
      piy = readlane_d(h_y, h);
      piz = readlane_d(h_z, h);
    } else {
      i = (int)cell_atoms[t];
      pix = pos[3 * i];
      piy = pos[3 * i + 1];
      piz = pos[3 * i + 2];
    }
    int count = 0;
    for (int q0 = 0; q0 < ncand; q0 += 64) {
      int q = q0 + lane;
      bool hit = false;
      int j = -1;
      if (q < ncand) {
        j = (int)cand[q];
        double dx, dy, dz;
        mic_exact(g.lat, g.inv, pix, piy, piz, pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], dx,
                  dy, dz);
        double r2 = r2_exact(dx, dy, dz);
        hit = r2 < rl2 && r2 > 1e-20;
      }
      uint64_t m = __ballot(hit);
      if (hit) {
        int slot = count + __popcll(m & lanemask_lt());
        if (slot < cap) nbr[(size_t)i * cap + slot] = j;
      }
      count += __popcll(m);
    }
    if (lane == 0) {
      cnt[i] = count;
      if (count > cap) stats[2] = max(stats[2], count);  // row overflow
    }
  }
}

// One wavefront per cell; a grid smaller than the cell count loops over the
// cells (the captured MD step launches few blocks, so that a step without a
// rebuild, whose blocks return at once, dispatches few of them).
__global__ __launch_bounds__(64) void k_nl_search(
    int C, int S, const int* __restrict__ cell_offs, const Geom* __restrict__ geoms,
    const int* __restrict__ cell_start, const unsigned* __restrict__ cell_atoms,
    const double* __restrict__ pos, double rl2, int cap, int cmax, int* __restrict__ cnt,
    int* __restrict__ nbr, int* __restrict__ stats, int* __restrict__ cand_n, const int* gate) {
  if (gated_off(gate)) return;
  extern __shared__ unsigned cand[];  // cmax entries (power of two)
  __shared__ int spre[28];
  __shared__ int sstart[27];
  for (int cc = blockIdx.x; cc < C; cc += gridDim.x) {
    search_cell(xcd_remap(cc, C), S, cell_offs, geoms, cell_start, cell_atoms, pos, rl2, cap,
                cmax, cnt, nbr, stats, cand_n, cand, spre, sstart);
    __syncthreads();  // the shared arrays are filled again for the next cell
  }
}

// Register-resident variant of k_nl_search for stencils of at most 64 * EPL
// candidates: lane l holds candidates l, l + 64, ...; the bitonic network
// runs on registers (shuffles across lanes, swaps within a lane), and each
// candidate's position is loaded once per cell.  Same pairs, same order.
template <int EPL>
__device__ __forceinline__ void search_cell_reg(
    int c, int S, const int* __restrict__ cell_offs, const Geom* __restrict__ geoms,
    const int* __restrict__ cell_start, const unsigned* __restrict__ cell_atoms,
    const double* __restrict__ pos, double rl2, int cap, int* __restrict__ cnt,
    int* __restrict__ nbr, int* __restrict__ stats, int* __restrict__ cand_n,
    int* __restrict__ spre, int* __restrict__ sstart) {
  constexpr int NP = 64 * EPL;
  const int lane = threadIdx.x;
  const int s = find_struct(cell_offs, S, c);
  const Geom& g = geoms[s];
  const int nc0 = g.nc[0], nc1 = g.nc[1], nc2 = g.nc[2];
  const int l = c - g.cell_off;
  const int c0 = l / (nc1 * nc2), c1 = (l / nc2) % nc1, c2 = l % nc2;
  const int n0 = nc0 == 1 ? 1 : (nc0 == 2 ? 2 : 3);
  const int n1 = nc1 == 1 ? 1 : (nc1 == 2 ? 2 : 3);
  const int n2 = nc2 == 1 ? 1 : (nc2 == 2 ? 2 : 3);
  const int nst = n0 * n1 * n2;
  int size = 0, cid = 0, cst = 0;
  if (lane < nst) {
    int a0 = lane / (n1 * n2), a1 = (lane / n2) % n1, a2 = lane % n2;
    int o0 = n0 == 3 ? a0 - 1 : a0, o1 = n1 == 3 ? a1 - 1 : a1, o2 = n2 == 3 ? a2 - 1 : a2;
    int j0 = (c0 + o0 + nc0) % nc0, j1 = (c1 + o1 + nc1) % nc1, j2 = (c2 + o2 + nc2) % nc2;
    cid = g.cell_off + (j0 * nc1 + j1) * nc2 + j2;
    cst = cell_start[cid];
    size = cell_start[cid + 1] - cst;
  }
  int inc = size;
#pragma unroll
  for (int o = 1; o < 32; o <<= 1) {
    int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  const int ncand = __shfl(inc, 31, 64);
  if (lane < nst) {
    spre[lane + 1] = inc;
    sstart[lane] = cst;  // first sorted position of the stencil cell
  }
  if (lane == 0) spre[0] = 0;
  const int first = cell_start[c], last = cell_start[c + 1];
  if (cand_n && lane == 0) cand_n[c] = ncand;
  // the cell's own atoms, one per lane (the first 64), read alongside the
  // candidates rather than one after another in the row loop
  int h_i = 0;
  double h_x = 0.0, h_y = 0.0, h_z = 0.0;
  if (lane < last - first) {
    h_i = (int)cell_atoms[first + lane];
    h_x = pos[3 * h_i];
    h_y = pos[3 * h_i + 1];
    h_z = pos[3 * h_i + 2];
  }
  if (ncand > NP) {
    if (lane == 0) stats[1] = max(stats[1], ncand);  // candidate overflow
    for (int t = first + lane; t < last; t += 64) cnt[cell_atoms[t]] = 0;
    return;
  }
  __syncthreads();
  unsigned v[EPL];
#pragma unroll
  for (int r = 0; r < EPL; ++r) {
    const int q = lane + 64 * r;
    v[r] = 0xFFFFFFFFu;
    if (q < ncand) {
      // stencil cell holding candidate q: largest st with spre[st] <= q
      int lo = 0, hi = nst;
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (spre[mid] <= q)
          lo = mid;
        else
          hi = mid;
      }
      v[r] = cell_atoms[sstart[lo] + (q - spre[lo])];
    }
  }
  // bitonic sort, ascending, of the elements e = lane + 64 r below P, the
  // power of two covering the candidates (the rest hold the sentinel and
  // are never paired with an element below P)
  int P = 1;
  while (P < ncand) P <<= 1;
#pragma unroll
  for (int k = 2; k <= NP; k <<= 1) {
    if (k <= P) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        if (j >= 64) {
#pragma unroll
          for (int r = 0; r < EPL; ++r) {
            const int pr = r ^ (j >> 6);
            if (pr > r && 64 * pr < P) {
              const bool up = ((lane + 64 * r) & k) == 0;
              const unsigned x = v[r], y = v[pr];
              if ((x > y) == up) {
                v[r] = y;
                v[pr] = x;
              }
            }
          }
        } else {
          const bool lower = (lane & j) == 0;
#pragma unroll
          for (int r = 0; r < EPL; ++r) {
            if (64 * r < P) {
              const bool up = ((lane + 64 * r) & k) == 0;
              const unsigned x = v[r];
              const unsigned y = (unsigned)__shfl_xor((int)x, j, 64);
              v[r] = (lower == up) ? min(x, y) : max(x, y);
            }
          }
        }
      }
    }
  }
  double px[EPL], py[EPL], pz[EPL];
#pragma unroll
  for (int r = 0; r < EPL; ++r) {
    const bool ok = lane + 64 * r < ncand;
    const int j = ok ? (int)v[r] : 0;
    px[r] = ok ? pos[3 * j] : 0.0;
    py[r] = ok ? pos[3 * j + 1] : 0.0;
    pz[r] = ok ? pos[3 * j + 2] : 0.0;
  }
  for (int t = first; t < last; ++t) {
    const int h = t - first;  // wave-uniform
    int i;
    double pix, piy, piz;
    if (h < 64) {
      i = __builtin_amdgcn_readlane(h_i, h);
      pix = readlane_d(h_x, h);
      piy = readlane_d(h_y, h);
      piz = readlane_d(h_z, h);
    } else {
      i = (int)cell_atoms[t];
      pix = pos[3 * i];
      piy = pos[3 * i + 1];
      piz = pos[3 * i + 2];
    }
    int count = 0;
#pragma unroll
    for (int r = 0; r < EPL; ++r) {
      if (64 * r >= ncand) continue;
      bool hit = false;
      if (lane + 64 * r < ncand) {
        double dx, dy, dz;
        mic_exact(g.lat, g.inv, pix, piy, piz, px[r], py[r], pz[r], dx, dy, dz);
        double r2 = r2_exact(dx, dy, dz);
        hit = r2 < rl2 && r2 > 1e-20;
      }
      uint64_t m = __ballot(hit);
      if (hit) {
        int slot = count + __popcll(m & lanemask_lt());
        if (slot < cap) nbr[(size_t)i * cap + slot] = (int)v[r];
      }
      count += __popcll(m);
    }
    if (lane == 0) {
      cnt[i] = count;
      if (count > cap) stats[2] = max(stats[2], count);  // row overflow
    }
  }
}

template <int EPL>
__global__ __launch_bounds__(64) void k_nl_search_reg(
    int C, int S, const int* __restrict__ cell_offs, const Geom* __restrict__ geoms,
    const int* __restrict__ cell_start, const unsigned* __restrict__ cell_atoms,
    const double* __restrict__ pos, double rl2, int cap, int* __restrict__ cnt,
    int* __restrict__ nbr, int* __restrict__ stats, int* __restrict__ cand_n, const int* gate) {
  if (gated_off(gate)) return;
  __shared__ int spre[28];
  __shared__ int sstart[27];
  // neighbouring cells on one XCD; a smaller grid loops over the cells
  for (int cc = blockIdx.x; cc < C; cc += gridDim.x) {
    search_cell_reg<EPL>(xcd_remap(cc, C), S, cell_offs, geoms, cell_start, cell_atoms, pos, rl2,
                         cap, cnt, nbr, stats, cand_n, spre, sstart);
    __syncthreads();  // the shared arrays are filled again for the next cell
  }
}

// stats[0] = max row length over the n atoms and, with cand_n, stats[4] =
// the most candidates around one of the C cells: block maxima, one atomic
// per block.
__global__ __launch_bounds__(1024) void k_nl_maxcnt(int n, const int* __restrict__ cnt, int C,
                                                    const int* __restrict__ cand_n,
                                                    int* __restrict__ stats, const int* gate) {
  if (gated_off(gate)) return;
  __shared__ int m[2][16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int v = 0, w = 0;
  for (int a = blockIdx.x * 1024 + tid; a < n; a += gridDim.x * 1024) v = max(v, cnt[a]);
  if (cand_n)
    for (int c = blockIdx.x * 1024 + tid; c < C; c += gridDim.x * 1024) w = max(w, cand_n[c]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v = max(v, __shfl_xor(v, o, 64));
    w = max(w, __shfl_xor(w, o, 64));
  }
  if (lane == 0) {
    m[0][wid] = v;
    m[1][wid] = w;
  }
  __syncthreads();
  if (tid == 0) {
    for (int k = 1; k < 16; ++k) {
      v = max(v, m[0][k]);
      w = max(w, m[1][k]);
    }
    atomicMax(&stats[0], v);
    if (cand_n) atomicMax(&stats[4], w);
  }
}

__global__ void k_nl_repack(int n, const int* __restrict__ cnt, const int* __restrict__ src,
                            int cap, int* __restrict__ dst, int stride) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * stride) return;
  const int sh = 31 - __clz(stride);  // the stride is a power of two
  const int a = (int)(t >> sh), k = (int)(t & (stride - 1));
  dst[t] = k < cnt[a] ? src[(size_t)a * cap + k] : -1;
}

// Reverse slot of entry t = (a, k): the position of a in the row of its
// neighbour j.  Rows are sorted ascending, so that position is the number of
// entries of row j below a; the row is read 16 bytes at a time (the stride
// is a power of two of at least four).  A row whose search overflowed its
// capacity is read up to the stride only.
__device__ __forceinline__ void rev_entry(int64_t t, const int* __restrict__ cnt,
                                          const int* __restrict__ nbr, int stride,
                                          int* __restrict__ rev, int* stats) {
  const int sh = 31 - __clz(stride);  // the stride is a power of two
  const int a = (int)(t >> sh), k = (int)(t & (stride - 1));
  const int ca = cnt[a];
  const int j = nbr[t];  // read with cnt[a] (in the row's storage either way)
  if (k >= ca) {
    rev[t] = -1;
    return;
  }
  const int below = row_rank(a, j, cnt, nbr, stride);
  if (below >= 0) {
    rev[t] = j * stride + below;
  } else {
    rev[t] = -1;
    stats[3] = 1;  // asymmetric list: should never happen
  }
}

// Entries t of the grid-stride loop (a grid smaller than n * stride / 256
// loops); with copy_src, also copy_dst[t] = copy_src[t] for t < 3 n; and
// clear[t] = 0 for t < n_clear (the cell counts of the sort before).
__global__ void k_nl_rev(int n, const int* __restrict__ cnt, const int* __restrict__ nbr,
                         int stride, int* __restrict__ rev, int* stats, const int* gate,
                         const double* __restrict__ copy_src, double* __restrict__ copy_dst,
                         unsigned* __restrict__ clear, int n_clear) {
  if (gated_off(gate)) return;
  const int64_t total = (int64_t)n * stride;
  int64_t end = copy_src ? max(total, 3 * (int64_t)n) : total;
  end = max(end, (int64_t)n_clear);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < end;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (copy_src && t < 3 * (int64_t)n) copy_dst[t] = copy_src[t];
    if (t < total) rev_entry(t, cnt, nbr, stride, rev, stats);
    if (t < n_clear) clear[t] = 0u;
  }
}

// Per-block max of |MIC(pos - ref @ M)|^2: the non-affine displacement since
// the last build, M mapping the build cell onto the current one
// (engine/neighborlist.py:662-669).
__global__ __launch_bounds__(256) void k_disp_max(int n, const double* __restrict__ pos,
                                                  const double* __restrict__ ref,
                                                  const int* __restrict__ sid,
                                                  const Geom* __restrict__ geoms, Mat3 M,
                                                  double* __restrict__ part) {
  __shared__ double m[256];
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (a < n) {
    const Geom& g = geoms[sid[a]];
    double r0 = ref[3 * a], r1 = ref[3 * a + 1], r2 = ref[3 * a + 2];
    double q0 = r0 * M.m[0] + r1 * M.m[3] + r2 * M.m[6];
    double q1 = r0 * M.m[1] + r1 * M.m[4] + r2 * M.m[7];
    double q2 = r0 * M.m[2] + r1 * M.m[5] + r2 * M.m[8];
    double dx, dy, dz;
    mic_exact(g.lat, g.inv, q0, q1, q2, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], dx, dy,
              dz);
    v = dx * dx + dy * dy + dz * dz;
  }
  m[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = m[0];
}

__global__ __launch_bounds__(256) void k_disp_flag(int nb, const double* __restrict__ part,
                                                   double sigma_min, double r_list,
                                                   double r_cut, int* __restrict__ flag,
                                                   double* __restrict__ dmax_out) {
  __shared__ double m[256];
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) v = fmax(v, part[b]);
  m[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double dmax = sqrt(m[0]);
    if (flag) *flag = (sigma_min * r_list - 2.0 * dmax <= r_cut) ? 1 : 0;
    if (dmax_out) *dmax_out = dmax;
  }
}

__global__ void k_half_count(int n, const int* __restrict__ cnt, const int* __restrict__ nbr,
                             int stride, unsigned* __restrict__ hc) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  int c = 0;
  for (int k = 0; k < cnt[a]; ++k) c += nbr[(size_t)a * stride + k] > a;
  hc[a] = (unsigned)c;
}

__global__ void k_half_write(int n, const int* __restrict__ cnt, const int* __restrict__ nbr,
                             int stride, const unsigned* __restrict__ off,
                             int32_t* __restrict__ oi, int32_t* __restrict__ oj) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  unsigned p = off[a];
  for (int k = 0; k < cnt[a]; ++k) {
    int j = nbr[(size_t)a * stride + k];
    if (j > a) {
      oi[p] = a;
      oj[p] = j;
      ++p;
    }
  }
}

// ---------------------------------------------------------------------------
// Host drivers
// ---------------------------------------------------------------------------

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// LDS a block may allocate on the current device (cached per device).
static size_t lds_per_block() {
  static int cached[64] = {0};
  int dev = 0;
  AK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return 64 * 1024;
  if (cached[dev] == 0) {
    int v = 0;
    AK_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    cached[dev] = v > 0 ? v : 64 * 1024;
  }
  return (size_t)cached[dev];
}

static void sort_cells(NList& nl, SetDev& set, const double* d_pos, hipStream_t st,
                       const int* gate, const SkinGate* sg = nullptr) {
  const int n = set.n_atoms;
  const int C = set.n_cells;
  const int nb = (int)nblk((int64_t)C + 1, CS_SCAN);
  nl.cell_key.ensure(sizeof(unsigned) * (n + 1));
  nl.cell_rank.ensure(sizeof(unsigned) * (n + 1));
  nl.cell_atoms.ensure(sizeof(unsigned) * (n + 1));
  nl.cell_pre.ensure(sizeof(unsigned) * (C + 2));
  nl.cell_blk.ensure(sizeof(unsigned) * (nb + 1));
  nl.cell_start.ensure(sizeof(int) * (C + 2));
  // the counts are zero between sorts (k_cell_scan clears them, or k_nl_rev
  // after k_cell_scan_place); a sort outside a captured step clears them here
  // as well
  // (C + 1 rounded up to whole 16-byte groups, read as such by
  // k_cell_scan_place)
  const size_t count_bytes = sizeof(unsigned) * (((size_t)C + 4) & ~(size_t)3);
  const bool fresh = nl.cell_count.bytes < count_bytes;
  nl.cell_count.ensure(count_bytes);
  if (!gate || fresh) {
    AK_REQUIRE(!gate || !fresh, AK_ERR_STATE, "captured cell sort without prepared counts");
    AK_HIP(hipMemsetAsync(nl.cell_count.p, 0, count_bytes, st));
  }
  unsigned* key = nl.cell_key.as<unsigned>();
  unsigned* rank = nl.cell_rank.as<unsigned>();
  unsigned* blk = nl.cell_blk.as<unsigned>();
  const unsigned cap = gate ? (unsigned)kGatedBlocks : 65535u;
  if (n > 0)
    hipLaunchKernelGGL(k_cell_count, dim3(std::min(nblk(n, CS_T), cap)), dim3(CS_T), 0, st, n,
                       d_pos, (const int*)set.d_sid.as<int>(), (const Geom*)set.d_geoms.as<Geom>(),
                       key, rank, nl.cell_count.as<unsigned>(), gate, sg ? *sg : SkinGate{});
  const int64_t items = std::max((int64_t)n, (int64_t)C + 1);
  const size_t lds_all = sizeof(unsigned) * ((size_t)C + 1);
  if (lds_all + 1024 <= lds_per_block()) {
    // one pass: every block scans the counts in LDS; k_nl_rev clears them
    hipLaunchKernelGGL(k_cell_scan_place, dim3(std::min(nblk(items, CS_T), cap)), dim3(CS_T),
                       lds_all, st, n, C, (const unsigned*)key, (const unsigned*)rank,
                       (const unsigned*)nl.cell_count.as<unsigned>(), nl.cell_start.as<int>(),
                       nl.cell_atoms.as<unsigned>(), gate);
    return;
  }
  hipLaunchKernelGGL(k_cell_scan, dim3(nb), dim3(CS_SCAN), 0, st, C,
                     nl.cell_count.as<unsigned>(), nl.cell_pre.as<unsigned>(), blk, gate);
  const int scanned = nb > CS_PLACE_LDS ? 1 : 0;
  if (scanned)
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, blk, nb, (unsigned*)nullptr,
                       gate);
  const size_t lds = scanned ? 0 : sizeof(unsigned) * (size_t)nb;
  hipLaunchKernelGGL(k_cell_place, dim3(std::min(nblk(items, CS_T), cap)), dim3(CS_T), lds, st,
                     n, C, (const unsigned*)key, (const unsigned*)rank,
                     (const unsigned*)nl.cell_pre.as<unsigned>(), (const unsigned*)blk, nb,
                     scanned, nl.cell_start.as<int>(), nl.cell_atoms.as<unsigned>(), gate);
}

static void launch_search(NList& nl, SetDev& set, const double* d_pos, int* d_cnt, int* d_nbr,
                          int cap, hipStream_t st, const int* gate) {
  const int C = set.n_cells;
  const double rl2 = nl.r_list * nl.r_list;
  const int* co = set.d_cell_offs.as<int>();
  const Geom* ge = set.d_geoms.as<Geom>();
  const int* cs = nl.cell_start.as<int>();
  const unsigned* ca = nl.cell_atoms.as<unsigned>();
  int* stt = nl.stats.as<int>();
  // candidates per cell, kept by full builds (a captured step reads none)
  int* cn = nullptr;
  if (!gate) {
    nl.cell_cand.ensure(sizeof(int) * (size_t)std::max(1, C));
    cn = nl.cell_cand.as<int>();
  }
  const int S = set.n_struct;
  // the captured (gated) search runs on a capped grid that loops over cells
  const unsigned G = gate ? (unsigned)std::min(C, kGatedSearchBlocks) : (unsigned)C;
  switch (nl.cmax) {
    case 64:
      hipLaunchKernelGGL(k_nl_search_reg<1>, dim3(G), dim3(64), 0, st, C, S, co, ge, cs, ca,
                         d_pos, rl2, cap, d_cnt, d_nbr, stt, cn, gate);
      return;
    case 128:
      hipLaunchKernelGGL(k_nl_search_reg<2>, dim3(G), dim3(64), 0, st, C, S, co, ge, cs, ca,
                         d_pos, rl2, cap, d_cnt, d_nbr, stt, cn, gate);
      return;
    case 256:
      hipLaunchKernelGGL(k_nl_search_reg<4>, dim3(G), dim3(64), 0, st, C, S, co, ge, cs, ca,
                         d_pos, rl2, cap, d_cnt, d_nbr, stt, cn, gate);
      return;
    case 512:
      hipLaunchKernelGGL(k_nl_search_reg<8>, dim3(G), dim3(64), 0, st, C, S, co, ge, cs, ca,
                         d_pos, rl2, cap, d_cnt, d_nbr, stt, cn, gate);
      return;
    default: {
      size_t lds = sizeof(unsigned) * (size_t)nl.cmax;
      hipLaunchKernelGGL(k_nl_search, dim3(G), dim3(64), lds, st, C, S, co, ge, cs, ca, d_pos,
                         rl2, cap, nl.cmax, d_cnt, d_nbr, stt, cn, gate);
    }
  }
}

// Candidate capacity per cell: twice the expected stencil population
// (27 cells, or fewer along axes with fewer than three cells), a power of two
// of at least 64.
static int est_cmax(SetDev& set) {
  double most = 0.0;
  for (auto& g : set.geoms) {
    const long long cells = (long long)g.nc[0] * g.nc[1] * g.nc[2];
    if (cells <= 0) continue;
    const int n0 = std::min(3, g.nc[0]), n1 = std::min(3, g.nc[1]), n2 = std::min(3, g.nc[2]);
    most = std::max(most, (double)g.natoms / (double)cells * n0 * n1 * n2);
  }
  return pow2ceil(std::max(64, (int)(2.0 * most) + 8));
}

static int est_cap(SetDev& set, double r_list) {
  double dens = 0.0;
  for (auto& g : set.geoms) dens = std::max(dens, g.natoms / g.volume);
  double est = dens * 4.18879020478639 * r_list * r_list * r_list;
  return pow2ceil(std::max(16, (int)(2.0 * est) + 16));
}

void nl_build(NList& nl, SetDev& set, const double* d_pos, hipStream_t st) {
  const int n = set.n_atoms;
  nl.n_atoms = n;
  nl.stats.ensure(sizeof(int) * 8);
  sort_cells(nl, set, d_pos, st, nullptr);
  nl.cnt.ensure(sizeof(int) * (n + 1));
  if (nl.cap == 0) nl.cap = est_cap(set, nl.r_list);
  if (nl.cmax == 0) nl.cmax = est_cmax(set);
  int h_stats[8];
  for (int attempt = 0; attempt < 8; ++attempt) {
    AK_HIP(hipMemsetAsync(nl.stats.p, 0, sizeof(int) * 8, st));
    nl.tmp_nbr.ensure(sizeof(int) * (size_t)std::max(1, n) * nl.cap);
    if (n > 0) {
      launch_search(nl, set, d_pos, nl.cnt.as<int>(), nl.tmp_nbr.as<int>(), nl.cap, st,
                    nullptr);
      const int C = set.n_cells;
      const unsigned g = std::min(64u, nblk(std::max(n, C), 1024));
      hipLaunchKernelGGL(k_nl_maxcnt, dim3(g), dim3(1024), 0, st, n, nl.cnt.as<int>(), C,
                         (const int*)nl.cell_cand.as<int>(), nl.stats.as<int>(),
                         (const int*)nullptr);
    }
    AK_HIP(hipMemcpyAsync(h_stats, nl.stats.p, sizeof(int) * 8, hipMemcpyDeviceToHost, st));
    AK_HIP(hipStreamSynchronize(st));
    bool again = false;
    if (h_stats[1] > 0) {  // some cell's stencil held more candidates than cmax
      nl.cmax = std::max(2 * nl.cmax, pow2ceil(h_stats[1]));
      AK_REQUIRE(nl.cmax <= 8192, AK_ERR_CAPACITY,
                 "more than 8192 candidate atoms around one cell: density too high for "
                 "the cell list");
      again = true;
    }
    if (h_stats[2] > 0) {  // rows longer than cap; counts are exact, use their max
      nl.cap = pow2ceil(std::max(h_stats[0], h_stats[2]));
      again = true;
    }
    if (!again) break;
    AK_REQUIRE(attempt < 7, AK_ERR_CAPACITY, "neighbour list capacity search did not converge");
  }
  nl.maxcnt = h_stats[0];
  nl.max_cand = h_stats[4];
  int stride = std::max(4, pow2ceil(std::max(1, nl.maxcnt)));
  if (nl.fixed_stride && nl.stride > 0) {
    AK_REQUIRE(stride <= nl.stride, AK_ERR_CAPACITY,
               "neighbour row exceeds the fixed stride of a captured list");
    stride = nl.stride;
  }
  AK_REQUIRE(stride >= 4 && (stride & (stride - 1)) == 0, AK_ERR_STATE,
             "neighbour rows need a power-of-two stride");
  nl.stride = stride;
  nl.nbr.ensure(sizeof(int) * (size_t)std::max(1, n) * stride);
  nl.rev.ensure(sizeof(int) * (size_t)std::max(1, n) * stride);
  if (n > 0) {
    int64_t tot = (int64_t)n * stride;
    hipLaunchKernelGGL(k_nl_repack, dim3(nblk(tot, 256)), dim3(256), 0, st, n, nl.cnt.as<int>(),
                       nl.tmp_nbr.as<int>(), nl.cap, nl.nbr.as<int>(), stride);
    hipLaunchKernelGGL(k_nl_rev, dim3(nblk(tot, 256)), dim3(256), 0, st, n, nl.cnt.as<int>(),
                       nl.nbr.as<int>(), stride, nl.rev.as<int>(), nl.stats.as<int>(),
                       (const int*)nullptr, (const double*)nullptr, (double*)nullptr,
                       nl.cell_count.as<unsigned>(), set.n_cells + 1);
  }
  nl.ref_pos.ensure(sizeof(double) * 3 * (size_t)std::max(1, n));
  if (n > 0)
    AK_HIP(hipMemcpyAsync(nl.ref_pos.p, d_pos, sizeof(double) * 3 * n,
                          hipMemcpyDeviceToDevice, st));
  nl.built_geoms = set.geoms;
  nl.valid = true;
  nl.n_builds++;
}

void nl_build_gated(NList& nl, SetDev& set, const double* d_pos, const int* d_flag,
                    hipStream_t st, const SkinGate* sg) {
  const int n = set.n_atoms;
  AK_REQUIRE(nl.valid && nl.fixed_stride && nl.stride > 0, AK_ERR_STATE,
             "gated rebuild needs a list built with a fixed stride");
  AK_REQUIRE(!sg || (n > 0 && sg->flag == d_flag), AK_ERR_STATE,
             "a skin gate needs atoms and writes the list's own flag");
  sort_cells(nl, set, d_pos, st, d_flag, sg);
  // search straight into the final layout (cap = stride)
  launch_search(nl, set, d_pos, nl.cnt.as<int>(), nl.nbr.as<int>(), nl.stride, st, d_flag);
  const int n_clear = set.n_cells + 1;
  int64_t tot = std::max(std::max((int64_t)n * nl.stride, 3 * (int64_t)n), (int64_t)n_clear);
  const unsigned grid = (unsigned)std::min<int64_t>(nblk(tot, 256), kGatedRevBlocks);
  hipLaunchKernelGGL(k_nl_rev, dim3(grid), dim3(256), 0, st, n, nl.cnt.as<int>(),
                     nl.nbr.as<int>(), nl.stride, nl.rev.as<int>(), nl.stats.as<int>(), d_flag,
                     d_pos, nl.ref_pos.as<double>(), nl.cell_count.as<unsigned>(), n_clear);
}

void nl_check_displacement(NList& nl, SetDev& set, const double* d_pos, const Mat3& M,
                           double sigma_min, int* d_flag, double* d_dmax, hipStream_t st) {
  const int n = set.n_atoms;
  int nb = (int)nblk(std::max(1, n), 256);
  nl.disp_part.ensure(sizeof(double) * nb);
  hipLaunchKernelGGL(k_disp_max, dim3(nb), dim3(256), 0, st, n, d_pos, nl.ref_pos.as<double>(),
                     set.d_sid.as<int>(), set.d_geoms.as<Geom>(), M, nl.disp_part.as<double>());
  hipLaunchKernelGGL(k_disp_flag, dim3(1), dim3(256), 0, st, nb, nl.disp_part.as<double>(),
                     sigma_min, nl.r_list, nl.r_cut, d_flag, d_dmax);
}

int nl_stats_overflow(NList& nl, hipStream_t st) {
  int h[8];
  AK_HIP(hipMemcpyAsync(h, nl.stats.p, sizeof(int) * 8, hipMemcpyDeviceToHost, st));
  AK_HIP(hipStreamSynchronize(st));
  return (h[1] > 0 || h[2] > 0 || h[3] > 0) ? 1 : 0;
}

int64_t nl_half_pairs(NList& nl, SetDev& set, hipStream_t st, int32_t* h_i, int32_t* h_j) {
  const int n = set.n_atoms;
  AK_REQUIRE(nl.valid, AK_ERR_STATE, "neighbour list has not been built");
  if (n == 0) return 0;
  nl.half_off.ensure(sizeof(unsigned) * (n + 1));
  DevBuf total;
  total.ensure(sizeof(unsigned));
  hipLaunchKernelGGL(k_half_count, dim3(nblk(n, 256)), dim3(256), 0, st, n, nl.cnt.as<int>(),
                     nl.nbr.as<int>(), nl.stride, nl.half_off.as<unsigned>());
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, nl.half_off.as<unsigned>(), n,
                     total.as<unsigned>(), (const int*)nullptr);
  unsigned np = 0;
  AK_HIP(hipMemcpyAsync(&np, total.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  AK_HIP(hipStreamSynchronize(st));
  if (h_i && np > 0) {
    DevBuf di, dj;
    di.ensure(sizeof(int32_t) * np);
    dj.ensure(sizeof(int32_t) * np);
    hipLaunchKernelGGL(k_half_write, dim3(nblk(n, 256)), dim3(256), 0, st, n, nl.cnt.as<int>(),
                       nl.nbr.as<int>(), nl.stride, nl.half_off.as<unsigned>(),
                       di.as<int32_t>(), dj.as<int32_t>());
    AK_HIP(hipMemcpyAsync(h_i, di.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, st));
    AK_HIP(hipMemcpyAsync(h_j, dj.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, st));
    AK_HIP(hipStreamSynchronize(st));
  }
  return (int64_t)np;
}

}  // namespace ak
