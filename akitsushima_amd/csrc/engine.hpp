// Host-side state of the akitsushima_amd engine: device buffers, structure
// sets, neighbour lists, potentials, and the launchers each .hip file exports.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/akitsushima_amd.h"
#include "common.hpp"

namespace ak {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define AK_HIP(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      throw ::ak::Error(AK_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define AK_REQUIRE(cond, code, msg)              \
  do {                                           \
    if (!(cond)) throw ::ak::Error((code), (msg)); \
  } while (0)

// Grow-only device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void ensure(size_t b) {
    if (b <= bytes && p) return;
    if (p) AK_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
    size_t nb = b < 256 ? 256 : b;
    AK_HIP(hipMalloc(&p, nb));
    bytes = nb;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Host geometry of one periodic cell (C++ restatement of the reference's
// _cell_metrics, engine/neighborlist.py:170-267).
struct CellInfo {
  double perp[3];
  double volume;
  double min_width;
};
CellInfo cell_info(const double* lat);
void invert3(const double* m, double* inv);

// A set of structures laid out atom-wise on the device.
struct SetDev {
  int n_atoms = 0;
  int n_struct = 0;
  int n_cells = 0;
  double setup_r_list = -1.0, setup_r_grid = 0.0;
  bool setup_checked = true;
  bool setup_had_inv = false;
  std::vector<Geom> geoms;
  std::vector<int> atom_off;  // n_struct + 1
  std::vector<int> uploaded_off;  // atom_off of the uploaded sid array
  std::vector<Geom> uploaded_geoms;
  DevBuf d_sid, d_geoms, d_cell_offs;
  // Launch chunks: contiguous atom ranges of at most `size` atoms inside one
  // structure, one thread block each, so that block partial sums never mix
  // structures.  rng (n, 2) = [first atom, end); first (S + 1) = first chunk
  // of each structure; partial (n, 10) = block sums of energy and virial.
  struct Chunks {
    int size = 0;
    int n = 0;
    DevBuf rng, first, partial;
  };
  std::vector<std::unique_ptr<Chunks>> chunk_tables;
  Chunks& chunks(int size, hipStream_t st);
  // (re)builds geometry for r_list; lattices (S, 9), invs may be null;
  // r_grid > r_list sizes the cell grid for cells that may shrink later.
  void setup(int n_struct, const int64_t* offsets, const double* lattices, const double* invs,
             double r_list, hipStream_t st, bool check_min_image = true, double r_grid = 0.0);
  void upload_geoms(hipStream_t st);
};

// Full (two-direction) neighbour list in ELL layout: row a holds cnt[a]
// neighbour indices sorted ascending, padded to `stride`; rev[a*stride+k]
// is the slot of a in the row of its k-th neighbour.
struct NList {
  double r_cut = 0, skin = 0, r_list = 0;
  int stride = 0;   // ELL row stride
  int maxcnt = 0;   // largest row
  int cap = 0;      // search capacity per atom
  int cmax = 0;     // candidate capacity per cell
  int max_cand = 0; // most candidates around one cell at the last full build
  int n_atoms = 0;
  int n_builds = 0, n_checks = 0;
  bool valid = false;
  bool fixed_stride = false;  // graph mode: stride never changes
  std::vector<Geom> built_geoms;
  DevBuf cnt, nbr, rev, ref_pos;
  // cell sort: cell and rank per atom, counts / block prefixes / block sums
  // per cell, the atoms grouped by cell and each cell's first position
  DevBuf cell_key, cell_rank, cell_count, cell_pre, cell_blk, cell_atoms, cell_start;
  DevBuf cell_cand;  // candidates around each cell (full builds)
  DevBuf tmp_nbr, stats, half_off;
  DevBuf disp_part, flag;
};

void nl_build(NList& nl, SetDev& set, const double* d_pos, hipStream_t st);
// The Verlet-skin test and the record of a captured MD step, for one block
// of 256 threads (all of them call it).  With disp_part, the test
// (engine/neighborlist.py:653-673, fixed cell): rebuild when
// r_list - 2 sqrt(max disp_part) <= r_cut.  With `owner`, *flag receives the
// result and counters[1] counts it, and with e_pot the state the previous
// kick left is recorded at index counters[0], which advances: E_pot =
// es[0], E_kin = the sum of the nb kinetic partials in fixed order, their
// sum, the pressure from E_kin and the virial diagonal (es[1], es[5],
// es[9]) and the rebuild count.  Returns the test's result (0 without
// disp_part).
struct SkinGate {
  int nb = 0;
  const double* ke_part = nullptr;
  const double* es = nullptr;
  const double* disp_part = nullptr;
  double r_list = 0.0, r_cut = 0.0;
  int* flag = nullptr;
  int* counters = nullptr;
  double* e_pot = nullptr;
  double* e_kin = nullptr;
  double* e_cons = nullptr;
  double* h_press = nullptr;
  int* h_nb = nullptr;
  double V = 0.0;
};

__device__ __forceinline__ int skin_gate_eval(const SkinGate& g, bool owner) {
  __shared__ double smem[16];
  __shared__ double wm[16];
  __shared__ int s_fl;
  double v[1] = {0.0};
  double d = 0.0;
  for (int b = threadIdx.x; b < g.nb; b += blockDim.x) {
    if (owner) v[0] += g.ke_part[b];
    if (g.disp_part) d = fmax(d, g.disp_part[b]);
  }
  if (owner) block_sum<1>(v, smem);  // block-uniform branch
  d = wave_max(d);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) d = fmax(d, wm[w]);
    const int fl = (g.disp_part && g.r_list - 2.0 * sqrt(d) <= g.r_cut) ? 1 : 0;
    s_fl = fl;
    if (owner) {
      if (g.e_pot) {
        const int step = g.counters[0];
        g.e_pot[step] = g.es[0];
        g.e_kin[step] = v[0];
        g.e_cons[step] = g.es[0] + v[0];
        g.h_press[step] = 2.0 * v[0] / (3.0 * g.V) + ((g.es[1] + g.es[5]) + g.es[9]) / 3.0;
        g.h_nb[step] = g.counters[1];
        g.counters[0] = step + 1;
      }
      if (g.disp_part) {
        *g.flag = fl;
        g.counters[1] += fl;
      }
    }
  }
  __syncthreads();
  return s_fl;
}

// Device-gated rebuild for captured MD steps: runs the cell sort and search
// only when *d_flag is nonzero; rows keep nl.stride.  With sg (and n > 0),
// the skin test is evaluated by the sort's first kernel in every block
// (block 0 owns the flag and the record) instead of read from *d_flag,
// which the later kernels read.
void nl_build_gated(NList& nl, SetDev& set, const double* d_pos, const int* d_flag,
                    hipStream_t st, const SkinGate* sg = nullptr);
// Largest minimum-image displacement since the last build -> *d_flag = 1 when
// sigma_min * r_list - 2 d_max <= r_cut (fixed cell: sigma_min = 1, M = I);
// d_flag and d_dmax may be null.
void nl_check_displacement(NList& nl, SetDev& set, const double* d_pos, const Mat3& M,
                           double sigma_min, int* d_flag, double* d_dmax, hipStream_t st);
int64_t nl_half_pairs(NList& nl, SetDev& set, hipStream_t st, int32_t* h_i, int32_t* h_j);
int nl_stats_overflow(NList& nl, hipStream_t st);

// Potentials ------------------------------------------------------------
enum Kind { KIND_TERSOFF = 1, KIND_SW = 2, KIND_EAM = 3 };

struct Potential {
  int kind = 0;
  int ns = 0;
  double r_cut = 0;
  std::vector<double> h_params;
  DevBuf d_params;
  // EAM function-table indices; eam_fast: the table is the Zhou alloy form
  // (per-species density and pair rows, cross rows of two species), so the
  // force pass may evaluate each pair's functions with shared exponentials
  std::vector<int> h_map;
  DevBuf d_map, d_eam_fast;
  bool eam_fast = false;
  int eam_nf = 0;  // function rows of an EAM table
  // SW with multi-species tables (pair, leg and triple records) in d_params
  bool sw_multi = false;
  // per-call scratch; amask (n) = bonds inside r_c per atom (rows >= 8)
  DevBuf G, aux, out, amask;
  // EAM: per launch chunk, 1 when any of its bonds lies inside r_c
  DevBuf pair_flag;
  // launch shape chosen by the autotuner (0 = default)
  int eam_lanes = 0;
  // Tersoff site kernel form: 2 = one lane per (bond, partner) pair (list
  // strides up to 8), otherwise one lane per bond
  int bond_tile = 0;
  // threads per block of the site / force kernels (64, 128, 256), and the
  // largest grid (0 = one block per chunk; else blocks stride over chunks)
  int block = 256;
  int grid_cap = 0;
  // fitting session (fit.hip)
  std::shared_ptr<void> fit;
  // list of the pairs inside r_cut alone, for force calls whose list rows
  // are longer than a wavefront (compute_forces)
  std::unique_ptr<NList> tight;
  // optional HIP-event timing of the dominant kernel (site kernel for
  // Tersoff/SW, force pass for EAM), recorded on the launch stream
  struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
  };
  bool prof_on = false;
  std::vector<EventPair> prof_pool;
  size_t prof_used = 0;
  EventPair* prof_next() {
    if (!prof_on) return nullptr;
    if (prof_used == prof_pool.size()) {
      EventPair p;
      AK_HIP(hipEventCreate(&p.a));
      AK_HIP(hipEventCreate(&p.b));
      prof_pool.push_back(p);
    }
    return &prof_pool[prof_used++];
  }
  ~Potential() {
    for (auto& p : prof_pool) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
  }
};

// Per-species / per-pair table of the shared-exponential EAM path, or false
// when the function table is not of the Zhou alloy form (forces.hip).
bool eam_fast_table(int ns, const std::vector<double>& rows, const std::vector<int>& map,
                    std::vector<double>& out);

// Function `f` of an EAM table at m device arguments (forces.hip).
void eam_feval(const Potential& pot, int f, int64_t m, const double* d_x, double* d_v,
               double* d_dv, hipStream_t st);

// Forces on all atoms of `set`, per-structure energy and stress.
// d_out: (n_struct, 10) = [E, S00..S22]; d_F (n_atoms, 3).
// Boltzmann constant in eV/K, K_B_SI * EV_PER_JOULE (constants.py:4-13)
constexpr double kBoltzmannEv = 1.380649e-23 * 6.241509074e18;

// eV/A per amu -> A/fs^2: (1.60217663e-19 / 1e-10) / 1.66053906660e-27 * 1e-20
// (the reference's EV_ANG_TO_AMU_ANG_FS2, constants.py:20)
constexpr double kConvAmu = (1.60217663e-19 / 1.0e-10) / (1.66053906660e-27) * 1.0e-20;

// d_esite (n, nullable): per-atom site energies U_a.
void compute_forces(Potential& pot, NList& nl, SetDev& set, const double* d_pos,
                    const int* d_spc, double* d_F, double* d_out, hipStream_t st,
                    double* d_esite = nullptr);

// Launch shape of a force evaluation: lanes per atom of the site (or EAM)
// kernels, whether the Tersoff lane-per-(bond, partner) form runs, threads
// per block and atoms per launch chunk.
struct ForcePlan {
  int lanes = 0, blk = 256, chunk = 0;
  bool trip = false;
};
ForcePlan force_plan(const Potential& pot, const NList& nl, bool esite);

// Device memory of one force evaluation besides its inputs and outputs:
// structure geometry and atom -> structure map, the launch chunks (ranges
// (n_chunks, 2), first chunk of each structure (n_struct + 1), block
// partials (10, n_chunks)), and the kernels' scratch: bond gradients G
// (3 n stride) and active-bond masks (n) for Tersoff / SW, the embedding
// derivatives (2 n) and per-chunk pair flags for EAM.
struct ForceWork {
  int n = 0, n_struct = 1, n_chunks = 0;
  const int* sid = nullptr;
  const Geom* geoms = nullptr;
  const int* rng = nullptr;
  const int* first = nullptr;
  double* partial = nullptr;
  double* G = nullptr;
  uint64_t* amask = nullptr;
  double* aux = nullptr;
  int* pair_flag = nullptr;
};
void compute_forces_work(Potential& pot, NList& nl, const ForcePlan& plan, const ForceWork& w,
                         const double* d_pos, const int* d_spc, double* d_F, double* d_out,
                         hipStream_t st, double* d_esite = nullptr);

// Forces, per-atom site energies and the site gradients G[a, k] = dU_a/dv_ak
// in pot.G for every kind (forces.hip); site_masked: whether only the slots
// whose amask bit is set hold G (else every slot below cnt does).
void site_grads(Potential& pot, NList& nl, SetDev& set, const double* d_pos, const int* d_spc,
                double* d_F, double* d_out, double* d_esite, hipStream_t st);
bool site_masked(const Potential& pot, const NList& nl);

// Device MD options: thermostat 0 = none (NVE), 1 = Nose-Hoover chain of
// `chain` links at `temperature` K with time constant `tau` fs, 2 = Langevin
// and 3 = CSVR with damping `tau` fs, 4 = NPT (isotropic barostat at
// pressure_gpa with time constant tau_baro, and the chain of 1); heat flux
// of the state after every step s with s % flux_every == 0 (0 = none).
struct MDOptions {
  // 0 NVE, 1 Nose-Hoover chain, 2 Langevin (BAOAB), 3 CSVR, 4 NPT, 5 NEMD
  // (velocity Verlet with Ornstein-Uhlenbeck heat baths)
  int thermostat = 0;
  double temperature = 0.0, tau = 0.0;  // tau: chain time constant or damping, fs
  int chain = 3;
  uint64_t seed = 0;  // random stream of the stochastic thermostats
  double pressure_gpa = 0.0, tau_baro = 0.0;  // NPT target pressure and barostat time
  int cell_dof = 0;         // NPT cell degrees of freedom: 0 iso
  int baro_thermostat = 0;  // NPT barostat thermostat: 0 Nose-Hoover chain
  int flux_every = 0;
  int rescale_steps = 0;            // velocity_rescale_steps (chain and NPT)
  const uint8_t* fixed = nullptr;   // fix_atoms: 3N flags, 1 = frozen component
  // trajectory: positions (n_steps / frame_every, N, 3) after every
  // frame_every steps, and under NPT the cell of each frame (host outputs)
  int frame_every = 0;
  double* frames = nullptr;
  double* frame_lattices = nullptr;
  // list rebuilds of the run up to each recorded state (n_steps + 1)
  int32_t* rebuild_counts = nullptr;
  // the cell (n_steps + 1, 9) and, for flexible NPT cells, the pressure
  // tensor (n_steps + 1, 6: xx, yy, zz, yz, xz, xy; NaN otherwise) of each record
  double* cell_history = nullptr;
  double* ptensor_history = nullptr;
  // heat flux per group of atoms (with flux_every): static labels (N, values
  // below n_labels) and/or regions along region_axis between region_bounds
  // (n_regions + 1); out (frames, n_labels * n_regions, 3)
  const int32_t* flux_labels = nullptr;
  int n_labels = 1;
  int region_axis = -1;
  const double* region_bounds = nullptr;
  int n_regions = 1;
  double* flux_groups = nullptr;
  // per-atom heat currents J_a / V of every recorded frame, float32
  // (frames, N, 3); may be given without flux_groups
  float* atom_flux = nullptr;
  // record the flux on the steps s with (s + flux_phase) % flux_every == 0;
  // regions as region_bounds == nullptr equal bins; group sums of five
  // components (J, the kinetic energy, the atom count)
  int flux_phase = 0;
  bool uniform_bins = false;
  bool group_kinetic = false;
  // NEMD (thermostat 5): bath per atom (0 none, 1 hot at `temperature`, 2
  // cold at T_cold), damping `tau`; per record (dQ_hot, dQ_cold, K_hot,
  // K_cold) into bath_history (n_steps + 1, 4)
  const uint8_t* bath = nullptr;
  double T_cold = 0.0;
  double* bath_history = nullptr;
};

// Thermostat / barostat state carried between runs (in and out of md_run):
// the particle chain xi[16], vxi[16]; the isotropic barostat momentum p_eps,
// its chain xi_b[2], vxi_b[2]; six strain momenta (reserved); the heat the
// stochastic thermostat put in; the count of random draws already used; the
// steps already run (for velocity_rescale_steps).
constexpr int kMdStateLen = 48;
constexpr int kStXi = 0, kStVxi = 16, kStPeps = 32, kStXiB = 33, kStVxiB = 35, kStPg = 37,
              kStHeat = 43, kStRng = 44, kStSteps = 45;

// eV/A^3 -> GPa (constants.py EV_ANG3_TO_GPA)
constexpr double kEvAng3ToGpa = 160.2176621;

// md.hip: n_steps of velocity Verlet on one cell, positions and velocities
// updated in place; e_pot, e_kin, e_cons (n_steps + 1) and flux
// ((n_steps + flux_every - 1) / flux_every, 3) as recorded on the device.
// state (kMdStateLen, may be null = start from rest) is read at the start
// and holds the final thermostat state on return.
void md_run(Potential& pot, int n, const double* lattice, const double* inv, double* pos,
            double* vel, const double* mass, const int32_t* spc, double dt, int n_steps,
            double skin, bool use_graph, const MDOptions& opt, double* state, double* e_pot,
            double* e_kin, double* e_cons, double* flux, double* lattice_out, double* volume,
            double* pressure, int* n_rebuilds, double* seconds, hipStream_t st);

// Per-group heat fluxes (device pointers): group of atom a = labels[a] (or
// 0) * n_reg + its region along `axis` (-1: none, n_reg = 1) between
// bounds[0..n_reg]; ng groups; part (grid * ng * 3) scratch; out
// (frames, ng, nc); atoms (frames, n, 3, float32, may be null) the per-atom
// currents / V.  bounds == nullptr with axis >= 0: n_reg equal bins.  nc = 3
// (the group fluxes) or 5 (also the group's kinetic energy and atom count).
struct FluxGroups {
  const int* labels = nullptr;
  int axis = -1;
  const double* bounds = nullptr;
  int n_reg = 1;
  int ng = 0;
  double* part = nullptr;
  double* out = nullptr;
  float* atoms = nullptr;
  int nc = 3;
};

// Heat flux J = [sum_a E_a v_a - sum_(a,k) v_ak (G[a,k] . v_nbr)] / V of one
// cell after site_grads (d_part: one triple per 32 atoms).  With d_gate, only
// on the steps s = counters[2] with (s + phase) % every == 0, into
// d_out[3 * ((s + phase) / every - (phase > 0))], and counters[2] advances;
// with groups, also the per-group sums of the per-atom heat currents.
void heat_flux(Potential& pot, NList& nl, SetDev& set, const double* d_pos, const double* d_vel,
               const double* d_mass, const double* d_esite, const double* d_vscale,
               double conv, int* d_gate, int every, double* d_part, double* d_out,
               hipStream_t st, const FluxGroups* groups = nullptr, int phase = 0);

inline int pow2ceil(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

}  // namespace ak
