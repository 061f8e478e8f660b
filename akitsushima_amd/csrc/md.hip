// Device-resident NVE molecular dynamics (velocity Verlet), after
// NVESimulation.run (solvers/molecular_dynamics.py:1969-1996):
//     v += dt/2 a;  x += dt v;  wrap into the cell;  list check/rebuild;
//     F = F(x);  a = F conv / m;  v += dt/2 a
// with conv = EV_ANG_TO_AMU_ANG_FS2 (constants.py:20).  The whole step runs on
// the device: the Verlet-skin test (engine/neighborlist.py:653-673, fixed
// cell) sets a device flag and the list kernels return at once when it is
// clear, so a step can be captured once in a hipGraph and replayed.
#include <chrono>
#include <cstring>
#include <vector>

#include "engine.hpp"

namespace ak {

// (1.60217663e-19 / 1e-10) / 1.66053906660e-27 * 1e-20
constexpr double kConv = (1.60217663e-19 / 1.0e-10) / (1.66053906660e-27) * 1.0e-20;

// Half kick, drift and wrap into the cell, and this block's largest squared
// displacement since the list was built (minimum image of x - x_build):
//   v += dt/2 a; x += dt v; frac = x @ inv; frac -= floor(frac);
//   frac[frac >= 1] = 0; x = frac @ lat.
__global__ __launch_bounds__(256) void k_md_drift(int n, double* __restrict__ pos,
                                                  double* __restrict__ vel,
                                                  const double* __restrict__ acc, double dt,
                                                  Geom g, const double* __restrict__ ref,
                                                  double* __restrict__ part) {
  __shared__ double m[256];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (a < n) {
    double x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double v = vel[3 * a + d] + 0.5 * dt * acc[3 * a + d];
      vel[3 * a + d] = v;
      x[d] = pos[3 * a + d] + dt * v;
    }
    double f[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double s = x[0] * g.inv[b] + x[1] * g.inv[3 + b] + x[2] * g.inv[6 + b];
      s -= floor(s);
      f[b] = s >= 1.0 ? 0.0 : s;
    }
    double y[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      y[b] = f[0] * g.lat[b] + f[1] * g.lat[3 + b] + f[2] * g.lat[6 + b];
      pos[3 * a + b] = y[b];
    }
    double dx, dy, dz;
    mic_exact(g.lat, g.inv, ref[3 * a], ref[3 * a + 1], ref[3 * a + 2], y[0], y[1], y[2], dx,
              dy, dz);
    d2 = dx * dx + dy * dy + dz * dz;
  }
  m[threadIdx.x] = d2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = m[0];
}

// a = F conv / m; v += dt/2 a (kick != 0); per-block kinetic-energy sums
// 0.5 m v^2 / conv.
__global__ __launch_bounds__(256) void k_md_kick(int n, double* __restrict__ vel,
                                                 double* __restrict__ acc,
                                                 const double* __restrict__ F,
                                                 const double* __restrict__ mass, double dt,
                                                 int kick, double* __restrict__ ke_part) {
  __shared__ double smem[16];
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  double ke[1] = {0.0};
  if (a < n) {
    double m = mass[a];
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double ad = F[3 * a + d] * kConv / m;
      acc[3 * a + d] = ad;
      double v = vel[3 * a + d];
      if (kick) {
        v += 0.5 * dt * ad;
        vel[3 * a + d] = v;
      }
      s += v * v;
    }
    ke[0] = 0.5 * m * s / kConv;
  }
  block_sum<1>(ke, smem);
  if (threadIdx.x == 0) ke_part[blockIdx.x] = ke[0];
}

// One block.  Records the state the previous kick left -- E_pot = es[0] and
// E_kin = sum of the kick's block sums in fixed order -- at index
// counters[0] and advances it; with disp_part != nullptr, also the Verlet-
// skin test (engine/neighborlist.py:653-673, fixed cell: rebuild when
// r_list - 2 max|MIC(x - x_build)| <= r_cut) -> *flag, counted in
// counters[1].
__global__ __launch_bounds__(256) void k_md_flag(int nb, const double* __restrict__ ke_part,
                                                 const double* __restrict__ es,
                                                 const double* __restrict__ disp_part,
                                                 double r_list, double r_cut,
                                                 int* __restrict__ flag,
                                                 int* __restrict__ counters,
                                                 double* __restrict__ e_pot,
                                                 double* __restrict__ e_kin) {
  __shared__ double smem[16];
  __shared__ double m[256];
  double v[1] = {0.0};
  double d = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    v[0] += ke_part[b];
    if (disp_part) d = fmax(d, disp_part[b]);
  }
  block_sum<1>(v, smem);
  m[threadIdx.x] = d;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int step = counters[0];
    e_pot[step] = es[0];
    e_kin[step] = v[0];
    counters[0] = step + 1;
    if (disp_part) {
      const int fl = (r_list - 2.0 * sqrt(m[0]) <= r_cut) ? 1 : 0;
      *flag = fl;
      counters[1] += fl;
    }
  }
}

struct MDState {
  DevBuf pos, vel, acc, F, mass, spc, es, ke, hist_ep, hist_ek, counters, flag;
};

// One step: drift (+ block displacement maxima); record of the previous state
// and the skin test; the gated list rebuild; forces; kick.  The state after
// the last step is recorded by one more k_md_flag after the loop.
static void md_step(Potential& pot, NList& nl, SetDev& set, MDState& S, int n, double dt,
                    hipStream_t st) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  nl.disp_part.ensure(sizeof(double) * nb);
  hipLaunchKernelGGL(k_md_drift, dim3(nb), dim3(256), 0, st, n, S.pos.as<double>(),
                     S.vel.as<double>(), (const double*)S.acc.as<double>(), dt, set.geoms[0],
                     (const double*)nl.ref_pos.as<double>(), nl.disp_part.as<double>());
  hipLaunchKernelGGL(k_md_flag, dim3(1), dim3(256), 0, st, (int)nb,
                     (const double*)S.ke.as<double>(), (const double*)S.es.as<double>(),
                     (const double*)nl.disp_part.as<double>(), nl.r_list, nl.r_cut,
                     S.flag.as<int>(), S.counters.as<int>(), S.hist_ep.as<double>(),
                     S.hist_ek.as<double>());
  nl_build_gated(nl, set, S.pos.as<double>(), S.flag.as<int>(), st);
  compute_forces(pot, nl, set, S.pos.as<double>(), S.spc.as<int>(), S.F.as<double>(),
                 S.es.as<double>(), st);
  hipLaunchKernelGGL(k_md_kick, dim3(nb), dim3(256), 0, st, n, S.vel.as<double>(),
                     S.acc.as<double>(), (const double*)S.F.as<double>(),
                     (const double*)S.mass.as<double>(), dt, 1, S.ke.as<double>());
}

static void md_record(MDState& S, int n, hipStream_t st) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_md_flag, dim3(1), dim3(256), 0, st, (int)nb,
                     (const double*)S.ke.as<double>(), (const double*)S.es.as<double>(),
                     (const double*)nullptr, 0.0, 0.0, (int*)nullptr, S.counters.as<int>(),
                     S.hist_ep.as<double>(), S.hist_ek.as<double>());
}

void md_nve_run(Potential& pot, int n, const double* lattice, const double* inv, double* pos,
                double* vel, const double* mass, const int32_t* spc, double dt, int n_steps,
                double skin, bool use_graph, double* e_pot, double* e_kin, int* n_rebuilds,
                double* seconds, hipStream_t st) {
  AK_REQUIRE(n > 0 && n_steps >= 0 && dt > 0.0, AK_ERR_ARG, "bad MD arguments");
  SetDev set;
  NList nl;
  nl.r_cut = pot.r_cut;
  nl.skin = skin;
  nl.r_list = pot.r_cut + skin;
  set.setup(1, std::vector<int64_t>{0, (int64_t)n}.data(), lattice, inv, nl.r_list, st);
  MDState S;
  S.pos.ensure(sizeof(double) * 3 * n);
  S.vel.ensure(sizeof(double) * 3 * n);
  S.acc.ensure(sizeof(double) * 3 * n);
  S.F.ensure(sizeof(double) * 3 * n);
  S.mass.ensure(sizeof(double) * n);
  S.spc.ensure(sizeof(int) * n);
  S.es.ensure(sizeof(double) * 10);
  const unsigned nb = (unsigned)((n + 255) / 256);
  S.ke.ensure(sizeof(double) * nb);
  S.hist_ep.ensure(sizeof(double) * (n_steps + 1));
  S.hist_ek.ensure(sizeof(double) * (n_steps + 1));
  S.counters.ensure(sizeof(int) * 4);
  S.flag.ensure(sizeof(int));
  AK_HIP(hipMemcpyAsync(S.pos.p, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  AK_HIP(hipMemcpyAsync(S.vel.p, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  AK_HIP(hipMemcpyAsync(S.mass.p, mass, sizeof(double) * n, hipMemcpyHostToDevice, st));
  std::vector<int32_t> sp(n, 0);
  if (spc) sp.assign(spc, spc + n);
  AK_HIP(hipMemcpyAsync(S.spc.p, sp.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  AK_HIP(hipMemsetAsync(S.counters.p, 0, sizeof(int) * 4, st));
  AK_HIP(hipMemsetAsync(S.flag.p, 0, sizeof(int), st));

  // initial list with head room for the captured rebuilds: rows up to twice
  // the current longest, candidate buffers twice the current need
  nl_build(nl, set, S.pos.as<double>(), st);
  nl.cmax = std::max(nl.cmax, 2 * nl.cmax);
  int stride = pow2ceil(std::max(4, nl.maxcnt + std::max(4, nl.maxcnt / 2)));
  AK_REQUIRE(stride <= 64 || pot.kind == KIND_EAM, AK_ERR_CAPACITY,
             "too many neighbours within r_cut + skin for a captured MD step");
  nl.fixed_stride = true;
  nl.stride = stride;
  nl.cap = stride;
  nl.valid = false;
  nl_build(nl, set, S.pos.as<double>(), st);

  // initial forces: acceleration and the kinetic-energy sums of state 0
  compute_forces(pot, nl, set, S.pos.as<double>(), S.spc.as<int>(), S.F.as<double>(),
                 S.es.as<double>(), st);
  hipLaunchKernelGGL(k_md_kick, dim3(nb), dim3(256), 0, st, n, S.vel.as<double>(),
                     S.acc.as<double>(), (const double*)S.F.as<double>(),
                     (const double*)S.mass.as<double>(), dt, 0, S.ke.as<double>());
  AK_HIP(hipMemsetAsync(nl.stats.p, 0, sizeof(int) * 8, st));
  // make sure every scratch buffer a step touches exists before capture: a
  // step with dt = 0 leaves positions, velocities and the state-0 sums as
  // they were
  md_step(pot, nl, set, S, n, 0.0, st);
  AK_HIP(hipStreamSynchronize(st));
  // restart the record and the rebuild count: step k records state k - 1
  int reset[4] = {0, 0, 0, 0};
  AK_HIP(hipMemcpy(S.counters.as<int>(), reset, sizeof(reset), hipMemcpyHostToDevice));

  auto t0 = std::chrono::steady_clock::now();
  if (use_graph && n_steps > 0) {
    hipGraph_t graph;
    hipGraphExec_t exec;
    AK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    md_step(pot, nl, set, S, n, dt, st);
    AK_HIP(hipStreamEndCapture(st, &graph));
    AK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < n_steps; ++s) AK_HIP(hipGraphLaunch(exec, st));
    md_record(S, n, st);
    AK_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
  } else {
    for (int s = 0; s < n_steps; ++s) md_step(pot, nl, set, S, n, dt, st);
    md_record(S, n, st);
    AK_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  }
  AK_REQUIRE(nl_stats_overflow(nl, st) == 0, AK_ERR_CAPACITY,
             "neighbour capacity of the captured MD step was exceeded");
  int counters[4];
  AK_HIP(hipMemcpy(counters, S.counters.p, sizeof(counters), hipMemcpyDeviceToHost));
  AK_REQUIRE(counters[0] == n_steps + 1, AK_ERR_STATE, "MD step counter mismatch");
  AK_HIP(hipMemcpy(pos, S.pos.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  AK_HIP(hipMemcpy(vel, S.vel.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  if (e_pot)
    AK_HIP(hipMemcpy(e_pot, S.hist_ep.p, sizeof(double) * (n_steps + 1),
                     hipMemcpyDeviceToHost));
  if (e_kin)
    AK_HIP(hipMemcpy(e_kin, S.hist_ek.p, sizeof(double) * (n_steps + 1),
                     hipMemcpyDeviceToHost));
  if (n_rebuilds) *n_rebuilds = counters[1];
}

}  // namespace ak
