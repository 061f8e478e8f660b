// Device-resident molecular dynamics (velocity Verlet), after
// NVESimulation.run (solvers/molecular_dynamics.py:1969-1996):
//     v += dt/2 a;  x += dt v;  wrap into the cell;  list check/rebuild;
//     F = F(x);  a = F conv / m;  v += dt/2 a
// with conv = EV_ANG_TO_AMU_ANG_FS2 (constants.py:20), and NVTSimulation.run
// (:1113-1160): the same step between two half-steps of an M-link
// Nose-Hoover chain (_nhc_step, :635-727).  The whole step runs on the
// device: the Verlet-skin test (engine/neighborlist.py:653-673, fixed cell)
// sets a device flag and the list kernels return at once when it is clear,
// so a step can be captured once in a hipGraph and replayed.  Optionally the
// heat flux of the state after each step is recorded on the device every
// `every` steps (ThermalConductivityCalculator.run_md,
// solvers/thermal_conductivity.py:575-610).
#include <chrono>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "sym3.hpp"

namespace ak {

constexpr double kConv = kConvAmu;

// x wrapped into the cell (frac = x @ inv; frac -= floor(frac);
// frac[frac >= 1] = 0; x = frac @ lat) and stored as atom a's position;
// returns the squared minimum-image displacement from ref.
__device__ __forceinline__ double wrap_store(int a, const double (&x)[3], const Geom& g,
                                             const double* __restrict__ ref,
                                             double* __restrict__ pos) {
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
  mic_exact(g.lat, g.inv, ref[3 * a], ref[3 * a + 1], ref[3 * a + 2], y[0], y[1], y[2], dx, dy,
            dz);
  return dx * dx + dy * dy + dz * dz;
}

// part[blockIdx.x] = the block's largest d2 (wavefront maxima, then one
// LDS exchange).
__device__ __forceinline__ void block_max_store(double d2, double* __restrict__ part) {
  __shared__ double wm[16];
  d2 = wave_max(d2);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = wm[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) x = fmax(x, wm[w]);
    part[blockIdx.x] = x;
  }
}

// Half kick, drift and wrap into the cell, and this block's largest squared
// displacement since the list was built (minimum image of x - x_build):
//   v += dt/2 a; x += dt v; wrap (wrap_store).
// vscale (nullable): the thermostat's velocity scale, applied first.
// keep (nullable, 3N of 0 or 1): frozen components (fix_atoms) are zeroed
// after the kick, as `velocities *= ~fix_mask` does.
__global__ __launch_bounds__(256) void k_md_drift(int n, double* __restrict__ pos,
                                                  double* __restrict__ vel,
                                                  const double* __restrict__ acc, double dt,
                                                  Geom g, const double* __restrict__ ref,
                                                  double* __restrict__ part,
                                                  const double* __restrict__ vscale,
                                                  const double* __restrict__ keep) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (a < n) {
    const double sc = vscale ? vscale[0] : 1.0;
    double x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double v0 = vscale ? vel[3 * a + d] * sc : vel[3 * a + d];
      double v = v0 + 0.5 * dt * acc[3 * a + d];
      if (keep) v *= keep[3 * a + d];
      vel[3 * a + d] = v;
      x[d] = pos[3 * a + d] + dt * v;
    }
    d2 = wrap_store(a, x, g, ref, pos);
  }
  block_max_store(d2, part);
}

// Plain NVE with the closing half kick of a step deferred to the next
// step's first pass: with counters[0] > 0 (a step has run since the records
// restarted) v += dt/2 a first, a = F conv / m; the kinetic-energy sums of
// that state go to ke_part (one per block, as k_md_kick<1> writes them);
// then the half kick and the drift of k_md_drift.  The same operations in
// the same order as k_md_kick<1> followed by k_md_drift, so the trajectory
// is bitwise the same; the run's last closing kick is a k_md_kick.
__global__ __launch_bounds__(256) void k_md_kick_drift(
    int n, double* __restrict__ pos, double* __restrict__ vel, const double* __restrict__ F,
    const double* __restrict__ mass, double dt, Geom g, const double* __restrict__ ref,
    double* __restrict__ part, double* __restrict__ ke_part, const int* __restrict__ counters,
    const double* __restrict__ keep) {
  __shared__ double smem[16];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const bool pending = counters[0] > 0;
  double d2 = 0.0;
  double ke[1] = {0.0};
  if (a < n) {
    const double ms = mass[a];
    double ad[3], v[3], s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      ad[d] = F[3 * a + d] * kConv / ms;
      v[d] = vel[3 * a + d];
      if (pending) {
        v[d] += 0.5 * dt * ad[d];
        if (keep) v[d] *= keep[3 * a + d];
      }
      s += v[d] * v[d];
    }
    ke[0] = 0.5 * ms * s / kConv;
    double x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double w = v[d] + 0.5 * dt * ad[d];
      if (keep) w *= keep[3 * a + d];
      vel[3 * a + d] = w;
      x[d] = pos[3 * a + d] + dt * w;
    }
    d2 = wrap_store(a, x, g, ref, pos);
  }
  block_sum<1>(ke, smem);
  if (threadIdx.x == 0) ke_part[blockIdx.x] = ke[0];
  block_max_store(d2, part);
}

// a = F conv / m; v += dt/2 a (kick != 0); per-block kinetic-energy sums
// 0.5 m v^2 / conv (NT = 1), and with NT = 7 also the kinetic tensor
// sum m v_a v_b / conv as xx, yy, zz, yz, xz, xy (ke_part[NT b + c]).
template <int NT>
__global__ __launch_bounds__(256) void k_md_kick(int n, double* __restrict__ vel,
                                                 double* __restrict__ acc,
                                                 const double* __restrict__ F,
                                                 const double* __restrict__ mass, double dt,
                                                 int kick, double* __restrict__ ke_part,
                                                 const double* __restrict__ keep) {
  __shared__ double smem[16 * NT];
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  double ke[NT];
#pragma unroll
  for (int c = 0; c < NT; ++c) ke[c] = 0.0;
  if (a < n) {
    double m = mass[a];
    double s = 0.0;
    double v3[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double ad = F[3 * a + d] * kConv / m;
      acc[3 * a + d] = ad;
      double v = vel[3 * a + d];
      if (kick) {
        v += 0.5 * dt * ad;
        if (keep) v *= keep[3 * a + d];
        vel[3 * a + d] = v;
      }
      s += v * v;
      v3[d] = v;
    }
    ke[0] = 0.5 * m * s / kConv;
    if constexpr (NT == 7) {
      const double mc = m / kConv;
      ke[1] = mc * v3[0] * v3[0];
      ke[2] = mc * v3[1] * v3[1];
      ke[3] = mc * v3[2] * v3[2];
      ke[4] = mc * v3[1] * v3[2];
      ke[5] = mc * v3[0] * v3[2];
      ke[6] = mc * v3[0] * v3[1];
    }
  }
  block_sum<NT>(ke, smem);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NT; ++c) ke_part[NT * blockIdx.x + c] = ke[c];
  }
}

// One block of 256 threads: the skin test and the record (skin_gate_eval).
__global__ __launch_bounds__(256) void k_md_flag(SkinGate g) { (void)skin_gate_eval(g, true); }

// Nose-Hoover chain state on the device: nhc[0] = velocity scale of the last
// end-of-step half-step not yet applied to the velocities, nhc[1] = the
// scale the next drift applies, then Q[M], xi[M], vxi[M].
constexpr int kNhcMax = 16;

// One M-link chain half-step of dt/2 (_nhc_step, solvers/molecular_dynamics.py:
// 635-727, the same operations in the same order) for kinetic energy ek;
// returns the particle velocity scale exp(-vxi_1 dt/2).  E_kin after the
// scaling (the reference recomputes it from the velocities) is ek scale^2.
__device__ double nhc_half(double* __restrict__ nhc, int M, double dt, double kT, double ndof,
                           double ek) {
  double* Q = nhc + 2;
  double* xi = Q + M;
  double* v = xi + M;
  auto force = [&](int j, double e) {
    return j == 0 ? (2.0 * e - ndof * kT) / Q[0] : (Q[j - 1] * v[j - 1] * v[j - 1] - kT) / Q[j];
  };
  v[M - 1] += force(M - 1, ek) * (dt * 0.25);
  double s[kNhcMax];
  for (int j = M - 2; j >= 0; --j) {
    const double arg = fmin(fmax(-v[j + 1] * dt * 0.125, -350.0), 350.0);
    s[j] = exp(arg);
    v[j] = s[j] * (s[j] * v[j] + force(j, ek) * dt * 0.25);
  }
  const double scale = exp(fmin(fmax(-v[0] * dt * 0.5, -350.0), 350.0));
  for (int j = 0; j < M; ++j) xi[j] += v[j] * dt * 0.5;
  ek *= scale * scale;
  for (int j = 0; j < M - 1; ++j) v[j] = s[j] * (s[j] * v[j] + force(j, ek) * dt * 0.25);
  v[M - 1] += force(M - 1, ek) * (dt * 0.25);
  return scale;
}

// velocity_rescale_steps (NVTSimulation.run :1154-1167, NPTSimulation.run
// :2907-2920): for the steps s < vr_until of this call, after the end-of-step
// chain half-step, v *= sqrt(T / T_now), T_now = 2 E_kin / (n_dof k_B), when
// T_now > 1e-12 -- folded into the pending velocity scale nhc[0].  The step
// is counters[0] - 1 (state 0 is record 0); the set-up step before the
// record restarts runs at counters[0] = 0 and is never rescaled.
__device__ __forceinline__ void vel_rescale(double* __restrict__ nhc, double& ek, double ndof,
                                            int vr_until, double T,
                                            const int* __restrict__ counters) {
  const int s = counters[0] - 1;
  if (s < 0 || s >= vr_until) return;
  const double T_now = ek > 0.0 ? 2.0 * ek / (ndof * kBoltzmannEv) : 0.0;
  if (T_now > 1e-12) {
    const double f = sqrt(T / T_now);
    nhc[0] *= f;
    ek *= f * f;
  }
}

// One block.  E_kin of the velocities the last kick left = the kick's block
// sums in fixed order.  mode 1 (start of a step): half-step with E_kin of the
// velocities after the pending scale; the drift applies pending x new scale.
// mode 2 (end of a step): half-step; its scale becomes pending; records E_pot,
// E_kin and H_NHC = E_kin + E_pot + sum Q vxi^2 / 2 + n_dof kT xi_1
// + sum_{j>1} kT xi_j (_nhc_conserved, :947-969) at counters[0].  mode 0:
// records only (the initial state).
__global__ __launch_bounds__(256) void k_nhc(int nb, const double* __restrict__ ke_part,
                                             double* __restrict__ nhc, int M, double dt,
                                             double kT, double ndof, int vr_until, double T,
                                             int mode,
                                             const double* __restrict__ es,
                                             int* __restrict__ counters,
                                             double* __restrict__ e_pot,
                                             double* __restrict__ e_kin,
                                             double* __restrict__ e_cons, double V,
    double* __restrict__ h_press, int* __restrict__ h_nb) {
  __shared__ double smem[16];
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += ke_part[b];
  block_sum<1>(v, smem);
  if (threadIdx.x != 0) return;
  double ek = v[0];
  if (mode == 1) {
    const double p = nhc[0];
    ek *= p * p;
    const double s = nhc_half(nhc, M, dt, kT, ndof, ek);
    nhc[1] = p * s;
    nhc[0] = 1.0;
    return;
  }
  if (mode == 2) {
    const double s = nhc_half(nhc, M, dt, kT, ndof, ek);
    nhc[0] = s;
    ek *= s * s;
    vel_rescale(nhc, ek, ndof, vr_until, T, counters);
  }
  const double* Q = nhc + 2;
  const double* xi = Q + M;
  const double* w = xi + M;
  double t = 0.5 * Q[0] * w[0] * w[0];
  for (int j = 1; j < M; ++j) t = t + 0.5 * Q[j] * w[j] * w[j];
  double u = ndof * kT * xi[0];
  for (int j = 1; j < M; ++j) u = u + kT * xi[j];
  const int step = counters[0];
  e_pot[step] = es[0];
  e_kin[step] = ek;
  e_cons[step] = ek + es[0] + t + u;
  h_press[step] = 2.0 * ek / (3.0 * V) + ((es[1] + es[5]) + es[9]) / 3.0;
  h_nb[step] = counters[1];
  counters[0] = step + 1;
}

// ---------------------------------------------------------------------------
// Stochastic thermostats.  Random numbers come from Philox4x32-10 (Salmon,
// Moraes, Dror & Shaw, SC'11), a counter-based generator: the draw of atom a
// at step s is a pure function of (seed, a, s, stream), so the device and a
// host restatement produce the same numbers in any order.  Normals by the
// Box-Muller transform of two 53-bit uniforms.
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = U4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
           (uint32_t)p0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 53-bit uniform in [0, 1) from two words.
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Two standard normals from one Philox block: r = sqrt(-2 ln(1 - u1)),
// (r cos 2 pi u2, r sin 2 pi u2).
__device__ __forceinline__ void normal2(U4 q, double& z0, double& z1) {
  const double u1 = u53(q.x, q.y), u2 = u53(q.z, q.w);
  const double r = sqrt(-2.0 * log(1.0 - u1));
  const double t = 6.283185307179586 * u2;
  z0 = r * cos(t);
  z1 = r * sin(t);
}

// Streams of the counter's third word.
constexpr uint32_t kStreamLangevin = 1, kStreamCsvr = 3;

// Gamma(a) deviate, a >= 1 (Marsaglia & Tsang, ACM TOMS 26, 363 (2000)),
// drawing block k of the counter (k, h, stream) for its k-th normal and
// (k + 1) for its uniform; at most 64 rounds.
__device__ double gamma_mt(double a, uint32_t h, uint32_t k0, uint32_t k1) {
  const double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  for (uint32_t it = 0; it < 64; ++it) {
    double x, unused;
    normal2(philox4x32(U4{2 * it + 1, h, kStreamCsvr, 0}, k0, k1), x, unused);
    double v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    const U4 q = philox4x32(U4{2 * it + 2, h, kStreamCsvr, 0}, k0, k1);
    const double u = u53(q.x, q.y);
    if (u < 1.0 - 0.0331 * (x * x) * (x * x)) return d * v;
    if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v;
  }
  return d;
}

// CSVR half-step (CSVRSimulation._rescale, solvers/molecular_dynamics.py:
// 1648-1667; Bussi, Donadio & Parrinello, JCP 126, 014101 (2007)): the
// kinetic energy K propagated over dt/2 with c = exp(-dt/2 / damping);
// returns alpha, with *dq = alpha^2 K - K.  Draws: R_1 = normal of block
// (0, h), S = 2 Gamma((n_f - 1)/2).
__device__ double csvr_alpha(double K, double c, double nf, double kT, uint32_t h, uint32_t k0,
                             uint32_t k1, double* dq) {
  *dq = 0.0;
  if (!(K > 0.0) || nf < 1.0) return 1.0;
  const double Kt = 0.5 * nf * kT;
  double r1, unused;
  normal2(philox4x32(U4{0, h, kStreamCsvr, 0}, k0, k1), r1, unused);
  const double sq = nf > 1.0 ? 2.0 * gamma_mt(0.5 * (nf - 1.0), h, k0, k1) : 0.0;
  const double Kn = c * K + (1.0 - c) * Kt * (r1 * r1 + sq) / nf +
                    2.0 * r1 * sqrt(c * (1.0 - c) * K * Kt / nf);
  double alpha = sqrt(fmax(Kn, 0.0) / K);
  if (r1 + sqrt(c * nf * K / ((1.0 - c) * Kt)) < 0.0) alpha = -alpha;
  *dq = alpha * alpha * K - K;
  return alpha;
}

// One block, CSVR (R B A F B R, CSVRSimulation._step :1669-1687) with the
// pending-scale bookkeeping of k_nhc; st[0] pending, st[1] drift scale, st[2]
// heat Q added by the rescalings.  counters[3] counts the half-steps (the
// random stream).  mode 2 records E_pot, E_kin and E_tot - Q.
__global__ __launch_bounds__(256) void k_csvr(int nb, const double* __restrict__ ke_part,
                                              double* __restrict__ th, double c, double nf,
                                              double kT, uint32_t k0, uint32_t k1, int mode,
                                              const double* __restrict__ es,
                                              int* __restrict__ counters,
                                              double* __restrict__ e_pot,
                                              double* __restrict__ e_kin,
                                              double* __restrict__ e_cons, double V,
    double* __restrict__ h_press, int* __restrict__ h_nb) {
  __shared__ double smem[16];
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += ke_part[b];
  block_sum<1>(v, smem);
  if (threadIdx.x != 0) return;
  double ek = v[0];
  if (mode == 1 || mode == 2) {
    const double p = mode == 1 ? th[0] : 1.0;
    ek *= p * p;
    double dq;
    const double a = csvr_alpha(ek, c, nf, kT, (uint32_t)counters[3], k0, k1, &dq);
    counters[3] += 1;
    th[2] += dq;
    if (mode == 1) {
      th[1] = p * a;
      th[0] = 1.0;
      return;
    }
    th[0] = a;
    ek *= a * a;
  }
  const int step = counters[0];
  e_pot[step] = es[0];
  e_kin[step] = ek;
  e_cons[step] = ek + es[0] - th[2];
  h_press[step] = 2.0 * ek / (3.0 * V) + ((es[1] + es[5]) + es[9]) / 3.0;
  h_nb[step] = counters[1];
  counters[0] = step + 1;
}

// Langevin BAOAB (LangevinSimulation._step, :1570-1587): B, A and O of one
// step: v += dt/2 a; x += dt/2 v; (kinetic-energy block sums here);
// v = c v + sqrt((1 - c^2) kT / m) xi, xi of Philox blocks (a, s, 1, 0/1);
// block sums of m v for the centre-of-mass removal.
__global__ __launch_bounds__(256) void k_lang_bao(int n, double* __restrict__ pos,
                                                  double* __restrict__ vel,
                                                  const double* __restrict__ acc,
                                                  const double* __restrict__ mass, double dt,
                                                  double c, double kTc, uint32_t k0, uint32_t k1,
                                                  const int* __restrict__ counters,
                                                  double* __restrict__ ke_part,
                                                  double* __restrict__ p_part,
                                                  const double* __restrict__ keep) {
  __shared__ double smem[16 * 4];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  if (a < n) {
    const double m = mass[a];
    double v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[d] = vel[3 * a + d] + 0.5 * dt * acc[3 * a + d];
      if (keep) v[d] *= keep[3 * a + d];
      pos[3 * a + d] += 0.5 * dt * v[d];
    }
    r[0] = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / kConv;
    const uint32_t s = (uint32_t)counters[3];
    double z[4];
    normal2(philox4x32(U4{(uint32_t)a, s, kStreamLangevin, 0}, k0, k1), z[0], z[1]);
    normal2(philox4x32(U4{(uint32_t)a, s, kStreamLangevin, 1}, k0, k1), z[2], z[3]);
    const double sig = sqrt((1.0 - c * c) * kTc / m);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[d] = c * v[d] + sig * z[d];
      if (keep) v[d] *= keep[3 * a + d];
      vel[3 * a + d] = v[d];
      r[1 + d] = m * v[d];
    }
  }
  block_sum<4>(r, smem);
  if (threadIdx.x == 0) {
    ke_part[blockIdx.x] = r[0];
    p_part[3 * blockIdx.x] = r[1];
    p_part[3 * blockIdx.x + 1] = r[2];
    p_part[3 * blockIdx.x + 2] = r[3];
  }
}

// One block: centre-of-mass velocity P / M into th[4..6] (zero when remove
// is 0, the setup step).
__global__ __launch_bounds__(256) void k_lang_com(int nb, const double* __restrict__ p_part,
                                                  double inv_mass, int remove,
                                                  double* __restrict__ th) {
  __shared__ double smem[16 * 3];
  double p[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    p[0] += p_part[3 * b];
    p[1] += p_part[3 * b + 1];
    p[2] += p_part[3 * b + 2];
  }
  block_sum<3>(p, smem);
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; ++d) th[4 + d] = remove ? p[d] * inv_mass : 0.0;
}

// v -= v_com (the end of the O step); block sums of the kinetic energy after
// it; x += dt/2 v; wrap; the block's largest squared displacement (as
// k_md_drift).
__global__ __launch_bounds__(256) void k_lang_a2(int n, double* __restrict__ pos,
                                                 double* __restrict__ vel,
                                                 const double* __restrict__ mass, double dt,
                                                 Geom g, const double* __restrict__ ref,
                                                 const double* __restrict__ th,
                                                 double* __restrict__ part,
                                                 double* __restrict__ ke_part) {
  __shared__ double m[256];
  __shared__ double smem[16];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  double ke[1] = {0.0};
  if (a < n) {
    double x[3], v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[d] = vel[3 * a + d] - th[4 + d];
      vel[3 * a + d] = v[d];
      x[d] = pos[3 * a + d] + 0.5 * dt * v[d];
    }
    ke[0] = 0.5 * mass[a] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / kConv;
    double f[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double sb = x[0] * g.inv[b] + x[1] * g.inv[3 + b] + x[2] * g.inv[6 + b];
      sb -= floor(sb);
      f[b] = sb >= 1.0 ? 0.0 : sb;
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
  block_sum<1>(ke, smem);
  if (threadIdx.x == 0) ke_part[blockIdx.x] = ke[0];
  m[threadIdx.x] = d2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = m[0];
}

// One block: Q += K(after O) - K(before O); records E_pot, E_kin (the kick's
// sums) and E_tot - Q (the E-Q column); advances the random step.  mode 0:
// records only.
__global__ __launch_bounds__(256) void k_lang_rec(int nb, const double* __restrict__ ke_kick,
                                                  const double* __restrict__ ke_b,
                                                  const double* __restrict__ ke_a, int mode,
                                                  double* __restrict__ th,
                                                  const double* __restrict__ es,
                                                  int* __restrict__ counters,
                                                  double* __restrict__ e_pot,
                                                  double* __restrict__ e_kin,
                                                  double* __restrict__ e_cons, double V,
    double* __restrict__ h_press, int* __restrict__ h_nb) {
  __shared__ double smem[16 * 3];
  double v[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    v[0] += ke_kick[b];
    if (mode) {
      v[1] += ke_b[b];
      v[2] += ke_a[b];
    }
  }
  block_sum<3>(v, smem);
  if (threadIdx.x != 0) return;
  if (mode) {
    th[2] += v[2] - v[1];
    counters[3] += 1;
  }
  const int step = counters[0];
  e_pot[step] = es[0];
  e_kin[step] = v[0];
  e_cons[step] = v[0] + es[0] - th[2];
  h_press[step] = 2.0 * v[0] / (3.0 * V) + ((es[1] + es[5]) + es[9]) / 3.0;
  h_nb[step] = counters[1];
  counters[0] = step + 1;
}

// ---------------------------------------------------------------------------
// NEMD with two Langevin heat baths (NEMDSimulation.run, solvers/nemd.py:
// 150-232): after each velocity Verlet step the bath atoms take an exact
// Ornstein-Uhlenbeck update v <- c v + sqrt((1 - c^2) kT_bath / m) xi,
// c = exp(-dt / damping) (_o_step :113-117), xi of Philox blocks
// (a, draw, 7, 0/1), draw = counters[3]; the kinetic energy each bath gains
// is its heat.
// ---------------------------------------------------------------------------
constexpr uint32_t kStreamNemd = 7;

// Per block: the kinetic energy of all atoms after the update and of each
// bath before and after it, part[5 * block + (0: all, 1, 2: hot before,
// after, 3, 4: cold before, after)].  apply = 0 measures only.
__global__ __launch_bounds__(256) void k_nemd_o(int n, double* __restrict__ vel,
                                                const double* __restrict__ mass,
                                                const uint8_t* __restrict__ bath, double c,
                                                double kTc_hot, double kTc_cold, uint32_t k0,
                                                uint32_t k1, const int* __restrict__ counters,
                                                int apply, double* __restrict__ part) {
  __shared__ double smem[16 * 5];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (a < n) {
    const double m = mass[a];
    double v[3] = {vel[3 * a], vel[3 * a + 1], vel[3 * a + 2]};
    const int b = bath[a];
    if (b == 1 || b == 2) {
      const double before = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / kConv;
      if (apply) {
        const uint32_t s = (uint32_t)counters[3];
        double z[4];
        normal2(philox4x32(U4{(uint32_t)a, s, kStreamNemd, 0}, k0, k1), z[0], z[1]);
        normal2(philox4x32(U4{(uint32_t)a, s, kStreamNemd, 1}, k0, k1), z[2], z[3]);
        const double sig = sqrt((1.0 - c * c) * (b == 1 ? kTc_hot : kTc_cold) / m);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          v[d] = c * v[d] + sig * z[d];
          vel[3 * a + d] = v[d];
        }
      }
      const double after = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / kConv;
      r[b == 1 ? 1 : 3] = before;
      r[b == 1 ? 2 : 4] = after;
    }
    r[0] = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / kConv;
  }
  block_sum<5>(r, smem);
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; ++k) part[5 * blockIdx.x + k] = r[k];
}

// One block: the bath heats of the step (mode 1: Q += dQ_hot + dQ_cold in
// th[2], the random draw advances) and the record of E_pot, E_kin, E_tot - Q
// (the E_minus_Q drift), P_inst and (dQ_hot, dQ_cold, K_hot, K_cold).  mode
// 0: record only.
__global__ __launch_bounds__(256) void k_nemd_rec(int nb, const double* __restrict__ part,
                                                  int mode, double* __restrict__ th,
                                                  const double* __restrict__ es,
                                                  int* __restrict__ counters,
                                                  double* __restrict__ e_pot,
                                                  double* __restrict__ e_kin,
                                                  double* __restrict__ e_cons, double V,
                                                  double* __restrict__ h_press,
                                                  int* __restrict__ h_nb,
                                                  double* __restrict__ h_bath) {
  __shared__ double smem[16 * 5];
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] += part[5 * b + k];
  block_sum<5>(v, smem);
  if (threadIdx.x != 0) return;
  const double dh = mode ? v[2] - v[1] : 0.0, dc = mode ? v[4] - v[3] : 0.0;
  if (mode) {
    th[2] += dh + dc;
    counters[3] += 1;
  }
  const int step = counters[0];
  e_pot[step] = es[0];
  e_kin[step] = v[0];
  e_cons[step] = v[0] + es[0] - th[2];
  h_press[step] = 2.0 * v[0] / (3.0 * V) + ((es[1] + es[5]) + es[9]) / 3.0;
  h_nb[step] = counters[1];
  h_bath[4 * (size_t)step] = dh;
  h_bath[4 * (size_t)step + 1] = dc;
  h_bath[4 * (size_t)step + 2] = v[2];
  h_bath[4 * (size_t)step + 3] = v[4];
  counters[0] = step + 1;
}

// ---------------------------------------------------------------------------
// NPT: isotropic MTTK barostat with a 2-link chain on its momentum and the
// particle Nose-Hoover chain (NPTSimulation, solvers/molecular_dynamics.py:
// 2141-2555, cell_dof "iso", baro_thermostat "nhc"; step :2871-2895).  The
// cell is s L_0; the device Geom of the set is scaled in place each step, so
// the force and list kernels see the current cell.  State np[]: 0 s, 1 s at
// the last list build, 2 p_eps, 3 xi1_b, 4 xi2_b, 5 vxi1_b, 6 vxi2_b,
// 7 this step's cell scale exp(eps_dot dt), 8 its velocity factor
// exp(-(1 + 3/n_dof) eps_dot dt), 9 the count of steps whose cell left the
// range the cell grid was sized for.
// ---------------------------------------------------------------------------
struct NptParams {
  double kT, ndof, W, Qb, P_ext, s_min;
  // flexible cell (cell_dof other than iso): nb strain generators E_k as
  // symmetric xx, yy, zz, yz, xz, xy, their traces, the barostat chain's
  // first mass nb kT tau_b^2 (Qb is the second)
  int nb;
  double gen[6][6], gtr[6], Qb1;
  // Langevin barostat (baro_thermostat "langevin") with friction time tau_b
  int lv;
  double tau_b;
  uint32_t k0, k1;
  // velocity_rescale_steps of this call and the target temperature
  int vr_until;
  double T;
};

// np[] slots beyond the isotropic ones (0-9): 10 the heat the Langevin
// barostat added; flexible cell: 16 p_k (6), 24 the step's cell map
// expm(G dt) (9), 33 the velocity map expm(-(G + Tr G / n_dof) dt) (9), 48
// the new cell (9), 57 its inverse (9), 66 its volume, 72 the cell at the
// last list build (9), 81 its inverse (9), 96 the build-to-new cell map
// inv(L_b) L (9).
constexpr int kNpHeat = 10, kNpPg = 16, kNpMx = 24, kNpMv = 33, kNpLat = 48, kNpInv = 57,
              kNpVol = 66, kNpLb = 72, kNpLbInv = 81, kNpMap = 96, kNpLen = 112;
constexpr uint32_t kStreamBaro = 5;

// _baro_langevin (:2596-2617): p_k <- c p_k + sqrt((1 - c^2) W kT) xi_k,
// c = exp(-dt / (2 tau_b)), xi_k the first normal of Philox block
// (k, draw, 5, 0), draw = counters[3]; the kinetic energy the noise adds,
// (p'.p' - p.p) / 2W, goes to np[10].
__device__ void npt_baro_lv(double* __restrict__ pb, int nb, double* __restrict__ np,
                            const NptParams& p, double dt, int* __restrict__ counters) {
  const uint32_t draw = (uint32_t)counters[3];
  const double c = exp(-0.5 * dt / p.tau_b);
  const double sig = sqrt((1.0 - c * c) * p.W * p.kT);
  double old2 = 0.0, new2 = 0.0;
  for (int k = 0; k < nb; ++k) old2 += pb[k] * pb[k];
  for (int k = 0; k < nb; ++k) {
    double z, unused;
    normal2(philox4x32(U4{(uint32_t)k, draw, kStreamBaro, 0}, p.k0, p.k1), z, unused);
    pb[k] = c * pb[k] + sig * z;
    new2 += pb[k] * pb[k];
  }
  np[kNpHeat] += 0.5 * (new2 - old2) / p.W;
  counters[3] = (int)draw + 1;
}

// _nhc_baro_step (:2456-2494): two links on the barostat momentum.
__device__ void npt_baro_nhc(double* __restrict__ np, const NptParams& p, double dt) {
  const double e1 = 0.5 * np[2] * np[2] / p.W;
  double G1 = (2.0 * e1 - p.kT) / p.Qb;
  double G2 = (p.Qb * np[5] * np[5] - p.kT) / p.Qb;
  np[6] += G2 * (dt * 0.25);
  const double sc = exp(fmin(fmax(-np[6] * dt * 0.125, -350.0), 350.0));
  np[5] = sc * (sc * np[5] + G1 * dt * 0.25);
  np[2] *= exp(fmin(fmax(-np[5] * dt * 0.5, -350.0), 350.0));
  np[3] += np[5] * dt * 0.5;
  np[4] += np[6] * dt * 0.5;
  const double e2 = 0.5 * np[2] * np[2] / p.W;
  G1 = (2.0 * e2 - p.kT) / p.Qb;
  np[5] = sc * (sc * np[5] + G1 * dt * 0.25);
  G2 = (p.Qb * np[5] * np[5] - p.kT) / p.Qb;
  np[6] += G2 * (dt * 0.25);
}

// _barostat_half_step (:2496-2515): p_eps += dt/2 [3 V (P - P_ext) + (3/n_f) 2 E_kin].
__device__ __forceinline__ void npt_baro_half(double* __restrict__ np, const NptParams& p,
                                              double dt, double V, double P, double ek) {
  np[2] += 0.5 * dt * (3.0 * V * (P - p.P_ext) + (3.0 / p.ndof) * 2.0 * ek);
}

// One block.  mode 1 (start of a step): P_inst of the current state, the
// particle chain, the barostat chain and half-step, and this step's cell and
// velocity factors.  mode 2 (end): P_inst of the new state, barostat
// half-step and chain, particle chain (its scale pending), record of E_pot,
// E_kin, H_NPT (_npt_conserved :2425-2454), V and P_inst.  mode 0: record.
__global__ __launch_bounds__(256) void k_npt(int nb, const double* __restrict__ ke_part,
                                             double* __restrict__ nhc, int M,
                                             double* __restrict__ np, NptParams p, double dt,
                                             int mode, const double* __restrict__ es,
                                             const Geom* __restrict__ geom,
                                             int* __restrict__ counters,
                                             double* __restrict__ e_pot,
                                             double* __restrict__ e_kin,
                                             double* __restrict__ e_cons,
                                             double* __restrict__ h_vol,
                                             double* __restrict__ h_press,
    int* __restrict__ h_nb, double* __restrict__ h_lat) {
  __shared__ double smem[16];
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += ke_part[b];
  block_sum<1>(v, smem);
  if (threadIdx.x != 0) return;
  const double V = geom->volume;
  const double trs = (es[1] + es[5]) + es[9];
  double ek = v[0];
  if (mode == 1) {
    const double pend = nhc[0];
    ek *= pend * pend;
    const double P = 2.0 * ek / (3.0 * V) + trs / 3.0;
    const double s1 = nhc_half(nhc, M, dt, p.kT, p.ndof, ek);
    ek *= s1 * s1;
    if (p.lv)
      npt_baro_lv(np + 2, 1, np, p, dt, counters);
    else
      npt_baro_nhc(np, p, dt);
    npt_baro_half(np, p, dt, V, P, ek);
    const double ed = np[2] / p.W;
    np[7] = exp(ed * dt);
    np[8] = exp(-(1.0 + 3.0 / p.ndof) * ed * dt);
    nhc[1] = pend * s1;
    nhc[0] = 1.0;
    return;
  }
  const double P = 2.0 * ek / (3.0 * V) + trs / 3.0;
  if (mode == 2) {
    npt_baro_half(np, p, dt, V, P, ek);
    if (p.lv)
      npt_baro_lv(np + 2, 1, np, p, dt, counters);
    else
      npt_baro_nhc(np, p, dt);
    const double s2 = nhc_half(nhc, M, dt, p.kT, p.ndof, ek);
    nhc[0] = s2;
    ek *= s2 * s2;
    vel_rescale(nhc, ek, p.ndof, p.vr_until, p.T, counters);
  }
  const double* Q = nhc + 2;
  const double* xi = Q + M;
  const double* w = xi + M;
  double t = 0.5 * Q[0] * w[0] * w[0];
  for (int j = 1; j < M; ++j) t = t + 0.5 * Q[j] * w[j] * w[j];
  double u = p.ndof * p.kT * xi[0];
  for (int j = 1; j < M; ++j) u = u + p.kT * xi[j];
  const double ed = np[2] / p.W;
  const double hb = 0.5 * p.Qb * np[5] * np[5] + 0.5 * p.Qb * np[6] * np[6] + p.kT * np[3] +
                    p.kT * np[4];
  const int step = counters[0];
  e_pot[step] = es[0];
  e_kin[step] = ek;
  e_cons[step] = (ek + es[0] + p.P_ext * V + 0.5 * p.W * ed * ed + t + u + hb) - np[kNpHeat];
  h_vol[step] = V;
  h_press[step] = P;
  if (h_lat)
    for (int c = 0; c < 9; ++c) h_lat[9 * (size_t)step + c] = geom->lat[c];
  h_nb[step] = counters[1];
  counters[0] = step + 1;
}

// Drift of the NPT step (_scale_step :2517-2555): v = v d + dt/2 a;
// x += dt v; wrapped in the current cell; frac @ (cell x scale); v *= vfac;
// and the block's largest squared non-affine displacement from the build
// positions mapped to the new cell (engine/neighborlist.py:646-668).
__global__ __launch_bounds__(256) void k_npt_drift(int n, double* __restrict__ pos,
                                                   double* __restrict__ vel,
                                                   const double* __restrict__ acc, double dt,
                                                   const Geom* __restrict__ geom,
                                                   const double* __restrict__ ref,
                                                   const double* __restrict__ nhc,
                                                   const double* __restrict__ np,
                                                   double* __restrict__ part,
                                                   const double* __restrict__ keep) {
  __shared__ double m[256];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (a < n) {
    const double d = nhc[1], sc = np[7], vf = np[8];
    const double aff = np[0] * sc / np[1];
    double lat[9], inv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      lat[k] = geom->lat[k] * sc;
      inv[k] = geom->inv[k] / sc;
    }
    double x[3], v[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      v[q] = vel[3 * a + q] * d + 0.5 * dt * acc[3 * a + q];
      if (keep) v[q] *= keep[3 * a + q];
      x[q] = pos[3 * a + q] + dt * v[q];
    }
    double f[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double sb = x[0] * geom->inv[b] + x[1] * geom->inv[3 + b] + x[2] * geom->inv[6 + b];
      sb -= floor(sb);
      f[b] = sb >= 1.0 ? 0.0 : sb;
    }
    double y[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      y[b] = f[0] * lat[b] + f[1] * lat[3 + b] + f[2] * lat[6 + b];
      pos[3 * a + b] = y[b];
      vel[3 * a + b] = v[b] * vf;
    }
    double dx, dy, dz;
    mic_exact(lat, inv, ref[3 * a] * aff, ref[3 * a + 1] * aff, ref[3 * a + 2] * aff, y[0],
              y[1], y[2], dx, dy, dz);
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

// One block: the cell of the set scaled in place (lattice x sc, inverse /
// sc, volume x sc^3), s advanced, and the list test with the affine bound
// sigma_min = s / s_build (engine/neighborlist.py:613-673) -> *flag; a
// rebuild records s_build = s.  Cells below the grid's range are counted.
__global__ __launch_bounds__(256) void k_npt_flag(int nb, const double* __restrict__ disp_part,
                                                  double r_list, double r_cut, double s_min,
                                                  Geom* __restrict__ geom,
                                                  double* __restrict__ np,
                                                  int* __restrict__ flag,
                                                  int* __restrict__ counters) {
  __shared__ double m[256];
  double d = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) d = fmax(d, disp_part[b]);
  m[threadIdx.x] = d;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double sc = np[7];
  for (int k = 0; k < 9; ++k) {
    geom->lat[k] *= sc;
    geom->inv[k] /= sc;
  }
  geom->volume *= sc * sc * sc;
  const double s_new = np[0] * sc;
  np[0] = s_new;
  const int fl = ((s_new / np[1]) * r_list - 2.0 * sqrt(m[0]) <= r_cut) ? 1 : 0;
  if (fl) np[1] = s_new;
  if (s_new < s_min) np[9] += 1.0;
  *flag = fl;
  counters[1] += fl;
}

// ---------------------------------------------------------------------------
// NPT with a flexible cell (ADR-0036: cell_dof aniso, tri, xy, xy_iso, z;
// solvers/molecular_dynamics.py:2113-2138, 2566-2692): one momentum p_k per
// strain generator E_k, the cell moving as L -> L expm(G dt) with
// G = sum_k (p_k / W) E_k.  The kick kernels add the kinetic tensor to their
// block sums, so the pressure tensor sum m v v^T / V + sym(sigma) of the
// current velocities is formed without another pass over the atoms.
// ---------------------------------------------------------------------------

// _nhc_baro_step_g (:2619-2646): the 2-link chain over all p_k, target
// n_k kT, every p_k scaled by the same factor.
__device__ void nptg_baro_nhc(double* __restrict__ np, const NptParams& p, double dt) {
  double* pg = np + kNpPg;
  double e1 = 0.0;
  for (int k = 0; k < p.nb; ++k) e1 += pg[k] * pg[k];
  e1 = 0.5 * e1 / p.W;
  const double nkT = p.nb * p.kT;
  double G1 = (2.0 * e1 - nkT) / p.Qb1;
  double G2 = (p.Qb1 * np[5] * np[5] - p.kT) / p.Qb;
  np[6] += G2 * (dt * 0.25);
  const double sc = exp(fmin(fmax(-np[6] * dt * 0.125, -350.0), 350.0));
  np[5] = sc * (sc * np[5] + G1 * dt * 0.25);
  const double f = exp(fmin(fmax(-np[5] * dt * 0.5, -350.0), 350.0));
  for (int k = 0; k < p.nb; ++k) pg[k] *= f;
  np[3] += np[5] * dt * 0.5;
  np[4] += np[6] * dt * 0.5;
  double e2 = 0.0;
  for (int k = 0; k < p.nb; ++k) e2 += pg[k] * pg[k];
  e2 = 0.5 * e2 / p.W;
  G1 = (2.0 * e2 - nkT) / p.Qb1;
  np[5] = sc * (sc * np[5] + G1 * dt * 0.25);
  G2 = (p.Qb1 * np[5] * np[5] - p.kT) / p.Qb;
  np[6] += G2 * (dt * 0.25);
}

// _barostat_half_step_g (:2648-2656): p_k += dt/2 [V Tr(E_k (P_int - P I))
// + Tr(E_k) 2 E_kin / n_dof], P_int = K / V + sym(sigma), K the kinetic
// tensor (xx, yy, zz, yz, xz, xy) and sigma = es[1..9] row-major.
__device__ void nptg_baro_half(double* __restrict__ np, const NptParams& p, double dt, double V,
                               const double* K, const double* __restrict__ es, double ek) {
  double D[6];
  D[0] = K[0] / V + es[1] - p.P_ext;
  D[1] = K[1] / V + es[5] - p.P_ext;
  D[2] = K[2] / V + es[9] - p.P_ext;
  D[3] = K[3] / V + 0.5 * (es[6] + es[8]);
  D[4] = K[4] / V + 0.5 * (es[3] + es[7]);
  D[5] = K[5] / V + 0.5 * (es[2] + es[4]);
  double* pg = np + kNpPg;
  for (int k = 0; k < p.nb; ++k) {
    const double* E = p.gen[k];
    const double tr = E[0] * D[0] + E[1] * D[1] + E[2] * D[2] +
                      2.0 * (E[3] * D[3] + E[4] * D[4] + E[5] * D[5]);
    pg[k] = pg[k] + 0.5 * dt * (V * tr + p.gtr[k] * (2.0 * ek) / p.ndof);
  }
}

// One block, as k_npt for the flexible cell.  mode 1: the particle chain,
// the barostat thermostat and half-step, then G, the maps expm(G dt) and
// expm(-(G + Tr G / n_dof) dt) (elementwise for a diagonal G, by Jacobi
// eigenvectors otherwise; _scale_step_g :2658-2692), the new cell, its
// inverse and volume, and the map inv(L_b) L_new.  mode 2: barostat
// half-step and thermostat, particle chain, record.  mode 0: record.
__global__ __launch_bounds__(256) void k_nptg(int nb, const double* __restrict__ part,
                                              double* __restrict__ nhc, int M,
                                              double* __restrict__ np, NptParams p, double dt,
                                              int mode, const double* __restrict__ es,
                                              const Geom* __restrict__ geom,
                                              int* __restrict__ counters,
                                              double* __restrict__ e_pot,
                                              double* __restrict__ e_kin,
                                              double* __restrict__ e_cons,
                                              double* __restrict__ h_vol,
                                              double* __restrict__ h_press,
    int* __restrict__ h_nb, double* __restrict__ h_lat, double* __restrict__ h_pt) {
  __shared__ double smem[16 * 7];
  double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int c = 0; c < 7; ++c) v[c] += part[7 * b + c];
  block_sum<7>(v, smem);
  if (threadIdx.x != 0) return;
  const double V = geom->volume;
  const double trs = (es[1] + es[5]) + es[9];
  double ek = v[0];
  double K[6] = {v[1], v[2], v[3], v[4], v[5], v[6]};
  double* pg = np + kNpPg;
  if (mode == 1) {
    const double pend = nhc[0];
    ek *= pend * pend;
    for (int c = 0; c < 6; ++c) K[c] *= pend * pend;
    const double s1 = nhc_half(nhc, M, dt, p.kT, p.ndof, ek);
    ek *= s1 * s1;
    for (int c = 0; c < 6; ++c) K[c] *= s1 * s1;
    if (p.lv)
      npt_baro_lv(pg, p.nb, np, p, dt, counters);
    else
      nptg_baro_nhc(np, p, dt);
    nptg_baro_half(np, p, dt, V, K, es, ek);
    // G (symmetric) from the momenta
    double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < p.nb; ++k)
      for (int c = 0; c < 6; ++c) g[c] += (pg[k] / p.W) * p.gen[k][c];
    const double a = ((g[0] + g[1]) + g[2]) / p.ndof;
    double* Mx = np + kNpMx;
    double* Mv = np + kNpMv;
    if (dt == 0.0) {
      // a step of zero length leaves the cell exactly as it is
      for (int c = 0; c < 9; ++c) Mx[c] = Mv[c] = (c % 4 == 0) ? 1.0 : 0.0;
    } else if (g[3] == 0.0 && g[4] == 0.0 && g[5] == 0.0) {
      for (int c = 0; c < 9; ++c) Mx[c] = Mv[c] = 0.0;
      for (int d = 0; d < 3; ++d) {
        Mx[4 * d] = exp(g[d] * dt);
        Mv[4 * d] = exp(-(g[d] + a) * dt);
      }
    } else {
      const double G[9] = {g[0], g[5], g[4], g[5], g[1], g[3], g[4], g[3], g[2]};
      double w[3], U[9], fx[3], fv[3];
      sym3_eig(G, w, U);
      for (int d = 0; d < 3; ++d) {
        fx[d] = exp(w[d] * dt);
        fv[d] = exp(-(w[d] + a) * dt);
      }
      sym3_compose(U, fx, Mx);
      sym3_compose(U, fv, Mv);
    }
    double* L = np + kNpLat;
    mat3_mul(geom->lat, Mx, L);
    if (dt == 0.0) {
      for (int c = 0; c < 9; ++c) np[kNpInv + c] = geom->inv[c];
      np[kNpVol] = geom->volume;
    } else {
      np[kNpVol] = fabs(mat3_inv(L, np + kNpInv));
    }
    mat3_mul(np + kNpLbInv, L, np + kNpMap);
    nhc[1] = pend * s1;
    nhc[0] = 1.0;
    return;
  }
  const double P = 2.0 * ek / (3.0 * V) + trs / 3.0;
  if (mode == 2) {
    nptg_baro_half(np, p, dt, V, K, es, ek);
    if (p.lv)
      npt_baro_lv(pg, p.nb, np, p, dt, counters);
    else
      nptg_baro_nhc(np, p, dt);
    const double s2 = nhc_half(nhc, M, dt, p.kT, p.ndof, ek);
    nhc[0] = s2;
    ek *= s2 * s2;
    vel_rescale(nhc, ek, p.ndof, p.vr_until, p.T, counters);
  }
  const double* Q = nhc + 2;
  const double* xi = Q + M;
  const double* w = xi + M;
  double t = 0.5 * Q[0] * w[0] * w[0];
  for (int j = 1; j < M; ++j) t = t + 0.5 * Q[j] * w[j] * w[j];
  double u = p.ndof * p.kT * xi[0];
  for (int j = 1; j < M; ++j) u = u + p.kT * xi[j];
  double pp = 0.0;
  for (int k = 0; k < p.nb; ++k) pp += pg[k] * pg[k];
  const double hb = 0.5 * p.Qb1 * np[5] * np[5] + 0.5 * p.Qb * np[6] * np[6] +
                    p.nb * p.kT * np[3] + p.kT * np[4];
  const int step = counters[0];
  e_pot[step] = es[0];
  e_kin[step] = ek;
  e_cons[step] = (ek + es[0] + p.P_ext * V + 0.5 * pp / p.W + (t + u) + hb) - np[kNpHeat];
  h_vol[step] = V;
  h_press[step] = P;
  if (h_lat)
    for (int c = 0; c < 9; ++c) h_lat[9 * (size_t)step + c] = geom->lat[c];
  if (h_pt) {
    // pressure tensor of the recorded velocities (the pending scale applied):
    // sum m v v^T / V + sym(sigma) as xx, yy, zz, yz, xz, xy (_pressure_tensor)
    const double f2 = nhc[0] * nhc[0];
    double* q = h_pt + 6 * (size_t)step;
    q[0] = K[0] * f2 / V + es[1];
    q[1] = K[1] * f2 / V + es[5];
    q[2] = K[2] * f2 / V + es[9];
    q[3] = K[3] * f2 / V + 0.5 * (es[6] + es[8]);
    q[4] = K[4] * f2 / V + 0.5 * (es[3] + es[7]);
    q[5] = K[5] * f2 / V + 0.5 * (es[2] + es[4]);
  }
  h_nb[step] = counters[1];
  counters[0] = step + 1;
}

// Drift of the flexible-cell step (_scale_step_g :2658-2692): v = v d +
// dt/2 a; x += dt v, wrapped into the current cell; x @ expm(G dt);
// v @ expm(-(G + Tr G / n_dof) dt); and the block's largest squared
// non-affine displacement, x - x_build inv(L_b) L_new minimum-imaged in the
// new cell (engine/neighborlist.py:646-668).
__global__ __launch_bounds__(256) void k_nptg_drift(int n, double* __restrict__ pos,
                                                    double* __restrict__ vel,
                                                    const double* __restrict__ acc, double dt,
                                                    const Geom* __restrict__ geom,
                                                    const double* __restrict__ ref,
                                                    const double* __restrict__ nhc,
                                                    const double* __restrict__ np,
                                                    double* __restrict__ part,
                                                    const double* __restrict__ keep) {
  __shared__ double m[256];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (a < n) {
    const double d = nhc[1];
    const double* Mx = np + kNpMx;
    const double* Mv = np + kNpMv;
    const double* Mp = np + kNpMap;
    double x[3], v[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      v[q] = vel[3 * a + q] * d + 0.5 * dt * acc[3 * a + q];
      if (keep) v[q] *= keep[3 * a + q];
      x[q] = pos[3 * a + q] + dt * v[q];
    }
    double f[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double sb = x[0] * geom->inv[b] + x[1] * geom->inv[3 + b] + x[2] * geom->inv[6 + b];
      sb -= floor(sb);
      f[b] = sb >= 1.0 ? 0.0 : sb;
    }
    double xw[3], y[3], r[3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
      xw[b] = f[0] * geom->lat[b] + f[1] * geom->lat[3 + b] + f[2] * geom->lat[6 + b];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      y[b] = xw[0] * Mx[b] + xw[1] * Mx[3 + b] + xw[2] * Mx[6 + b];
      pos[3 * a + b] = y[b];
      vel[3 * a + b] = v[0] * Mv[b] + v[1] * Mv[3 + b] + v[2] * Mv[6 + b];
      r[b] = ref[3 * a] * Mp[b] + ref[3 * a + 1] * Mp[3 + b] + ref[3 * a + 2] * Mp[6 + b];
    }
    double dx, dy, dz;
    mic_exact(np + kNpLat, np + kNpInv, r[0], r[1], r[2], y[0], y[1], y[2], dx, dy, dz);
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

// One block: the new cell becomes the set's cell; the list test of
// check_and_rebuild with a cell change (engine/neighborlist.py:613-673):
// sigma_min(inv(L_b) L) (1 when L is allclose to L_b: |L - L_b| <= 1e-12 +
// 1e-5 |L_b| elementwise), rebuild when sigma_min r_list - 2 d_max <= r_cut,
// a rebuild recording L_b = L.  A cell whose perpendicular widths leave the
// cell grid or the minimum image is counted in np[9].
__global__ __launch_bounds__(256) void k_nptg_flag(int nb, const double* __restrict__ disp_part,
                                                   double r_list, double r_cut,
                                                   Geom* __restrict__ geom,
                                                   double* __restrict__ np,
                                                   int* __restrict__ flag,
                                                   int* __restrict__ counters) {
  __shared__ double m[256];
  double d = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) d = fmax(d, disp_part[b]);
  m[threadIdx.x] = d;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double* L = np + kNpLat;
  double* Lb = np + kNpLb;
  for (int k = 0; k < 9; ++k) {
    geom->lat[k] = L[k];
    geom->inv[k] = np[kNpInv + k];
  }
  geom->volume = np[kNpVol];
  bool close = true;
  for (int k = 0; k < 9; ++k) close = close && fabs(L[k] - Lb[k]) <= 1e-12 + 1e-5 * fabs(Lb[k]);
  double sig = 1.0;
  if (!close) {
    const double* Mp = np + kNpMap;
    double Mt[9], B[9], w[3], U[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Mt[3 * i + j] = Mp[3 * j + i];
    mat3_mul(Mt, Mp, B);
    sym3_eig(B, w, U);
    sig = sqrt(fmax(fmin(fmin(w[0], w[1]), w[2]), 0.0));
  }
  const int fl = (sig * r_list - 2.0 * sqrt(m[0]) <= r_cut) ? 1 : 0;
  if (fl) {
    for (int k = 0; k < 9; ++k) {
      Lb[k] = L[k];
      np[kNpLbInv + k] = np[kNpInv + k];
    }
  }
  // perpendicular widths V / |a_j x a_k| against the grid and the minimum image
  double wmin = 1e300;
  for (int ax = 0; ax < 3; ++ax) {
    const double* u = L + 3 * ((ax + 1) % 3);
    const double* q = L + 3 * ((ax + 2) % 3);
    const double cx = u[1] * q[2] - u[2] * q[1], cy = u[2] * q[0] - u[0] * q[2],
                 cz = u[0] * q[1] - u[1] * q[0];
    const double wd = np[kNpVol] / sqrt(cx * cx + cy * cy + cz * cz);
    wmin = fmin(wmin, wd);
    if (wd / geom->nc[ax] < r_list) np[9] += 1.0;
  }
  if (r_list >= 0.5 * wmin) np[9] += 1.0;
  *flag = fl;
  counters[1] += fl;
}

// Trajectory frame of the state after every `every` steps: positions (and
// the cell) copied into frame f = completed / every - 1, completed = the
// record index counters[0] less rec_off (1 when state 0 is recorded before
// the loop).  Steps outside [1, n_frames * every] (the set-up step) write
// nothing.
__global__ __launch_bounds__(256) void k_frame(int n, const double* __restrict__ pos,
                                               const Geom* __restrict__ geom,
                                               const int* __restrict__ counters, int rec_off,
                                               int every, int n_frames,
                                               double* __restrict__ frames,
                                               double* __restrict__ frame_lat) {
  const int done = counters[0] - rec_off;
  if (done < 1 || done % every != 0) return;
  const int f = done / every - 1;
  if (f >= n_frames) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * (int64_t)n) frames[(int64_t)f * 3 * n + t] = pos[t];
  if (frame_lat && t < 9) frame_lat[(int64_t)f * 9 + t] = geom->lat[t];
}

// v *= scale[0] (the pending thermostat scale at the end of a run).
__global__ __launch_bounds__(256) void k_scale_vel(int n, double* __restrict__ vel,
                                                   const double* __restrict__ scale) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * n) vel[t] *= scale[0];
}

struct MDState {
  DevBuf pos, vel, acc, F, mass, spc, es, ke, hist_ep, hist_ek, hist_ec, counters, flag;
  DevBuf nhc, esite, flux_part, hist_J, ke_b, ke_a, p_part, np, hist_V, hist_P, keep, hist_nb;
  // thermostat: 0 none (NVE), 1 Nose-Hoover chain, 2 Langevin, 3 CSVR,
  // 4 NPT (barostat + Nose-Hoover chain), 5 NEMD (two heat baths); cell: 0
  // iso, else a flexible cell
  // whose kicks also sum the kinetic tensor
  int thermo = 0, M = 0, every = 0, cell = 0, n = 0;
  double kT = 0.0, ndof = 0.0, damping = 0.0, inv_mass = 0.0;
  uint32_t k0 = 0, k1 = 0;
  // fix_atoms: keep = 1 - frozen per component (null when nothing is frozen)
  bool fixed = false;
  // velocity_rescale_steps left for this call, target temperature
  int vr_until = 0;
  double T = 0.0;
  double V = 0.0;  // volume of a fixed cell (the pressure record)
  // trajectory frames every frame_every steps (0 = none)
  int frame_every = 0, n_frames = 0;
  DevBuf frames, frame_lat;
  // flexible-cell NPT: the cell and the pressure tensor of every record
  bool cell_hist = false;
  DevBuf hist_L, hist_Pt;
  // heat flux per group of atoms
  FluxGroups fg;
  DevBuf fg_labels, fg_bounds, fg_part, hist_Jg, hist_Ja;
  int flux_phase = 0;
  // NEMD: bath per atom, the cold bath's kT, the per-record bath heats
  DevBuf bath, hist_bath;
  double kT_cold = 0.0;
  const double* keep_ptr() const { return fixed ? keep.as<double>() : nullptr; }
  NptParams npt{};
  Geom* geom = nullptr;  // the set's device cell
  // velocity-scaling thermostats keep a pending scale in nhc[0]
  bool scaled() const { return thermo == 1 || thermo == 3 || thermo == 4; }
  // plain NVE without flux or frame records: each step's closing half kick
  // runs in the next step's drift (k_md_kick_drift)
  bool lazy = false;
  bool flex() const { return thermo == 4 && cell != 0; }
};

static void kick(MDState& S, int n, double dt, int do_kick, hipStream_t st) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (S.flex())
    hipLaunchKernelGGL(k_md_kick<7>, dim3(nb), dim3(256), 0, st, n, S.vel.as<double>(),
                       S.acc.as<double>(), (const double*)S.F.as<double>(),
                       (const double*)S.mass.as<double>(), dt, do_kick, S.ke.as<double>(),
                       S.keep_ptr());
  else
    hipLaunchKernelGGL(k_md_kick<1>, dim3(nb), dim3(256), 0, st, n, S.vel.as<double>(),
                       S.acc.as<double>(), (const double*)S.F.as<double>(),
                       (const double*)S.mass.as<double>(), dt, do_kick, S.ke.as<double>(),
                       S.keep_ptr());
}

static void thermo_kernel(MDState& S, int mode, double dt, hipStream_t st, unsigned nb) {
  if (S.thermo == 1)
    hipLaunchKernelGGL(k_nhc, dim3(1), dim3(256), 0, st, (int)nb,
                       (const double*)S.ke.as<double>(), S.nhc.as<double>(), S.M, dt, S.kT,
                       S.ndof, S.vr_until, S.T, mode, (const double*)S.es.as<double>(),
                       S.counters.as<int>(), S.hist_ep.as<double>(), S.hist_ek.as<double>(),
                       S.hist_ec.as<double>(), S.V, S.hist_P.as<double>(), S.hist_nb.as<int>());
  else if (S.thermo == 3)
    hipLaunchKernelGGL(k_csvr, dim3(1), dim3(256), 0, st, (int)nb,
                       (const double*)S.ke.as<double>(), S.nhc.as<double>(),
                       exp(-0.5 * dt / S.damping), S.ndof, S.kT, S.k0, S.k1, mode,
                       (const double*)S.es.as<double>(), S.counters.as<int>(),
                       S.hist_ep.as<double>(), S.hist_ek.as<double>(), S.hist_ec.as<double>(),
                       S.V, S.hist_P.as<double>(), S.hist_nb.as<int>());
  else if (S.flex())
    hipLaunchKernelGGL(k_nptg, dim3(1), dim3(256), 0, st, (int)nb,
                       (const double*)S.ke.as<double>(), S.nhc.as<double>(), S.M,
                       S.np.as<double>(), S.npt, dt, mode, (const double*)S.es.as<double>(),
                       (const Geom*)S.geom, S.counters.as<int>(), S.hist_ep.as<double>(),
                       S.hist_ek.as<double>(), S.hist_ec.as<double>(), S.hist_V.as<double>(),
                       S.hist_P.as<double>(), S.hist_nb.as<int>(),
                       S.cell_hist ? S.hist_L.as<double>() : (double*)nullptr,
                       S.cell_hist ? S.hist_Pt.as<double>() : (double*)nullptr);
  else if (S.thermo == 4)
    hipLaunchKernelGGL(k_npt, dim3(1), dim3(256), 0, st, (int)nb,
                       (const double*)S.ke.as<double>(), S.nhc.as<double>(), S.M,
                       S.np.as<double>(), S.npt, dt, mode, (const double*)S.es.as<double>(),
                       (const Geom*)S.geom, S.counters.as<int>(), S.hist_ep.as<double>(),
                       S.hist_ek.as<double>(), S.hist_ec.as<double>(), S.hist_V.as<double>(),
                       S.hist_P.as<double>(), S.hist_nb.as<int>(),
                       S.cell_hist ? S.hist_L.as<double>() : (double*)nullptr);
  else if (S.thermo == 5) {
    hipLaunchKernelGGL(k_nemd_o, dim3(nb), dim3(256), 0, st, S.n, S.vel.as<double>(),
                       (const double*)S.mass.as<double>(), (const uint8_t*)S.bath.as<uint8_t>(),
                       exp(-dt / S.damping), S.kT * kConv, S.kT_cold * kConv, S.k0, S.k1,
                       (const int*)S.counters.as<int>(), mode != 0 ? 1 : 0, S.ke_b.as<double>());
    hipLaunchKernelGGL(k_nemd_rec, dim3(1), dim3(256), 0, st, (int)nb,
                       (const double*)S.ke_b.as<double>(), mode != 0 ? 1 : 0, S.nhc.as<double>(),
                       (const double*)S.es.as<double>(), S.counters.as<int>(),
                       S.hist_ep.as<double>(), S.hist_ek.as<double>(), S.hist_ec.as<double>(),
                       S.V, S.hist_P.as<double>(), S.hist_nb.as<int>(),
                       S.hist_bath.as<double>());
  } else if (S.thermo == 2)
    hipLaunchKernelGGL(k_lang_rec, dim3(1), dim3(256), 0, st, (int)nb,
                       (const double*)S.ke.as<double>(), (const double*)S.ke_b.as<double>(),
                       (const double*)S.ke_a.as<double>(), mode, S.nhc.as<double>(),
                       (const double*)S.es.as<double>(), S.counters.as<int>(),
                       S.hist_ep.as<double>(), S.hist_ek.as<double>(), S.hist_ec.as<double>(),
                       S.V, S.hist_P.as<double>(), S.hist_nb.as<int>());
}

// The skin test of list nl (nullable: none) and, for NVE, the record of a
// step, over the nb kinetic and displacement partials of the step's passes.
static SkinGate md_gate(MDState& S, const NList* nl, int nb) {
  SkinGate g;
  g.nb = nb;
  g.ke_part = S.ke.as<double>();
  g.es = S.es.as<double>();
  g.counters = S.counters.as<int>();
  if (nl) {
    g.disp_part = nl->disp_part.as<double>();
    g.r_list = nl->r_list;
    g.r_cut = nl->r_cut;
    g.flag = S.flag.as<int>();
  }
  if (S.thermo == 0) {
    g.e_pot = S.hist_ep.as<double>();
    g.e_kin = S.hist_ek.as<double>();
    g.e_cons = S.hist_ec.as<double>();
    g.h_press = S.hist_P.as<double>();
    g.h_nb = S.hist_nb.as<int>();
  }
  g.V = S.V;
  return g;
}

// One step: [thermostat half-step;] drift (+ block displacement maxima) --
// for Langevin the B A O A sequence; the skin test (NVE: also the record of
// the previous state); the gated list rebuild; forces; kick; [thermostat
// half-step and record;] [heat flux].  For NVE the state after the last step
// is recorded by md_record after the loop.
static void md_step(Potential& pot, NList& nl, SetDev& set, MDState& S, int n, double dt,
                    hipStream_t st, bool flux_now = true) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  nl.disp_part.ensure(sizeof(double) * nb);
  if (S.scaled()) thermo_kernel(S, 1, dt, st, nb);
  if (S.thermo == 2) {
    hipLaunchKernelGGL(k_lang_bao, dim3(nb), dim3(256), 0, st, n, S.pos.as<double>(),
                       S.vel.as<double>(), (const double*)S.acc.as<double>(),
                       (const double*)S.mass.as<double>(), dt, exp(-dt / S.damping),
                       S.kT * kConv, S.k0, S.k1, (const int*)S.counters.as<int>(),
                       S.ke_b.as<double>(), S.p_part.as<double>(), S.keep_ptr());
    // with frozen components the centre-of-mass velocity is not removed
    // (LangevinSimulation._o_step :1552-1565)
    hipLaunchKernelGGL(k_lang_com, dim3(1), dim3(256), 0, st, (int)nb,
                       (const double*)S.p_part.as<double>(), S.inv_mass,
                       (dt > 0.0 && !S.fixed) ? 1 : 0, S.nhc.as<double>());
    hipLaunchKernelGGL(k_lang_a2, dim3(nb), dim3(256), 0, st, n, S.pos.as<double>(),
                       S.vel.as<double>(), (const double*)S.mass.as<double>(), dt, set.geoms[0],
                       (const double*)nl.ref_pos.as<double>(), (const double*)S.nhc.as<double>(),
                       nl.disp_part.as<double>(), S.ke_a.as<double>());
  } else if (S.flex()) {
    hipLaunchKernelGGL(k_nptg_drift, dim3(nb), dim3(256), 0, st, n, S.pos.as<double>(),
                       S.vel.as<double>(), (const double*)S.acc.as<double>(), dt,
                       (const Geom*)S.geom, (const double*)nl.ref_pos.as<double>(),
                       (const double*)S.nhc.as<double>(), (const double*)S.np.as<double>(),
                       nl.disp_part.as<double>(), S.keep_ptr());
  } else if (S.thermo == 4) {
    hipLaunchKernelGGL(k_npt_drift, dim3(nb), dim3(256), 0, st, n, S.pos.as<double>(),
                       S.vel.as<double>(), (const double*)S.acc.as<double>(), dt,
                       (const Geom*)S.geom, (const double*)nl.ref_pos.as<double>(),
                       (const double*)S.nhc.as<double>(), (const double*)S.np.as<double>(),
                       nl.disp_part.as<double>(), S.keep_ptr());
  } else if (S.lazy) {
    hipLaunchKernelGGL(k_md_kick_drift, dim3(nb), dim3(256), 0, st, n, S.pos.as<double>(),
                       S.vel.as<double>(), (const double*)S.F.as<double>(),
                       (const double*)S.mass.as<double>(), dt, set.geoms[0],
                       (const double*)nl.ref_pos.as<double>(), nl.disp_part.as<double>(),
                       S.ke.as<double>(), (const int*)S.counters.as<int>(), S.keep_ptr());
  } else {
    hipLaunchKernelGGL(k_md_drift, dim3(nb), dim3(256), 0, st, n, S.pos.as<double>(),
                       S.vel.as<double>(), (const double*)S.acc.as<double>(), dt, set.geoms[0],
                       (const double*)nl.ref_pos.as<double>(), nl.disp_part.as<double>(),
                       (const double*)(S.scaled() ? S.nhc.as<double>() + 1 : nullptr),
                       S.keep_ptr());
  }
  if (S.flex())
    hipLaunchKernelGGL(k_nptg_flag, dim3(1), dim3(256), 0, st, (int)nb,
                       (const double*)nl.disp_part.as<double>(), nl.r_list, nl.r_cut, S.geom,
                       S.np.as<double>(), S.flag.as<int>(), S.counters.as<int>());
  else if (S.thermo == 4)
    hipLaunchKernelGGL(k_npt_flag, dim3(1), dim3(256), 0, st, (int)nb,
                       (const double*)nl.disp_part.as<double>(), nl.r_list, nl.r_cut,
                       S.npt.s_min, S.geom, S.np.as<double>(), S.flag.as<int>(),
                       S.counters.as<int>());
  // fixed cell: the skin test and the record run in the list's first kernel
  const bool fixed_cell = !S.flex() && S.thermo != 4;
  const SkinGate sg = md_gate(S, &nl, (int)nb);
  nl_build_gated(nl, set, S.pos.as<double>(), S.flag.as<int>(), st,
                 fixed_cell ? &sg : nullptr);
  // site energies and gradients only on the steps that record the flux
  if (S.every > 0 && flux_now)
    site_grads(pot, nl, set, S.pos.as<double>(), S.spc.as<int>(), S.F.as<double>(),
               S.es.as<double>(), S.esite.as<double>(), st);
  else
    compute_forces(pot, nl, set, S.pos.as<double>(), S.spc.as<int>(), S.F.as<double>(),
                   S.es.as<double>(), st);
  if (!S.lazy) kick(S, n, dt, 1, st);
  if (S.thermo) thermo_kernel(S, S.thermo == 2 ? 1 : 2, dt, st, nb);
  // a recording step: the gate counts the records, so frame = counters[2]
  if (S.every > 0 && flux_now)
    heat_flux(pot, nl, set, S.pos.as<double>(), S.vel.as<double>(), S.mass.as<double>(),
              S.esite.as<double>(), S.scaled() ? S.nhc.as<double>() : nullptr, kConv,
              S.counters.as<int>(), 1, S.flux_part.as<double>(), S.hist_J.as<double>(), st,
              S.fg.ng > 0 ? &S.fg : nullptr, 0);
  if (S.frame_every > 0)
    hipLaunchKernelGGL(k_frame, dim3((unsigned)((3 * (int64_t)n + 255) / 256)), dim3(256), 0,
                       st, n, (const double*)S.pos.as<double>(), (const Geom*)S.geom,
                       (const int*)S.counters.as<int>(), S.thermo ? 1 : 0, S.frame_every,
                       S.n_frames, S.frames.as<double>(),
                       S.thermo == 4 ? S.frame_lat.as<double>() : (double*)nullptr);
}

static void md_record(MDState& S, int n, hipStream_t st) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (S.thermo) {
    thermo_kernel(S, 0, 0.0, st, nb);
    return;
  }
  hipLaunchKernelGGL(k_md_flag, dim3(1), dim3(256), 0, st, md_gate(S, nullptr, (int)nb));
}

// NVE: the closing half kick still pending after the last step (deferred
// kicks), then the record of the final state.
static void md_finish(MDState& S, int n, double dt, int n_steps, hipStream_t st) {
  if (S.lazy && n_steps > 0) kick(S, n, dt, 1, st);
  md_record(S, n, st);
}

void md_run(Potential& pot, int n, const double* lattice, const double* inv, double* pos,
            double* vel, const double* mass, const int32_t* spc, double dt, int n_steps,
            double skin, bool use_graph, const MDOptions& opt, double* state, double* e_pot,
            double* e_kin, double* e_cons, double* flux, double* lattice_out, double* volume,
            double* pressure, int* n_rebuilds, double* seconds, hipStream_t st) {
  AK_REQUIRE(n > 0 && n_steps >= 0 && dt > 0.0, AK_ERR_ARG, "bad MD arguments");
  AK_REQUIRE(opt.thermostat >= 0 && opt.thermostat <= 5, AK_ERR_ARG,
             "thermostat must be 0 (NVE), 1 (Nose-Hoover chain), 2 (Langevin), 3 (CSVR), "
             "4 (NPT) or 5 (NEMD)");
  AK_REQUIRE(opt.cell_dof >= 0 && opt.cell_dof <= 5, AK_ERR_ARG,
             "NPT cell_dof must be 0 (iso), 1 (aniso), 2 (tri), 3 (xy), 4 (xy_iso) or 5 (z)");
  AK_REQUIRE(opt.baro_thermostat == 0 || opt.baro_thermostat == 1, AK_ERR_ARG,
             "NPT baro_thermostat must be 0 (Nose-Hoover chain) or 1 (Langevin)");
  if (state)
    for (int k = 0; k < kMdStateLen; ++k)
      AK_REQUIRE(std::isfinite(state[k]), AK_ERR_ARG, "thermostat state must be finite");
  const auto st_at = [&](int k) { return state ? state[k] : 0.0; };
  AK_REQUIRE(opt.flux_every >= 0, AK_ERR_ARG, "flux_every must be >= 0");
  AK_REQUIRE(opt.thermostat != 4 || opt.flux_every == 0, AK_ERR_ARG,
             "the heat flux is recorded at fixed volume (not under NPT)");
  SetDev set;
  NList nl;
  nl.r_cut = pot.r_cut;
  nl.skin = skin;
  nl.r_list = pot.r_cut + skin;
  // NPT: a cell grid that stays valid while the cell shrinks by up to 5 %
  const double kNptShrink = 0.95;
  set.setup(1, std::vector<int64_t>{0, (int64_t)n}.data(), lattice, inv, nl.r_list, st, true,
            opt.thermostat == 4 ? nl.r_list / kNptShrink : 0.0);
  MDState S;
  S.n = n;
  S.geom = set.d_geoms.as<Geom>();
  S.thermo = opt.thermostat;
  S.cell = opt.thermostat == 4 ? opt.cell_dof : 0;
  // fix_atoms (_build_fix_mask :248-267): frozen components start at zero
  // and stay there; _compute_n_dof (:341-361): 3N - 3 with nothing frozen,
  // else 3N minus the frozen components
  int n_fixed = 0;
  if (opt.fixed)
    for (int64_t k = 0; k < 3 * (int64_t)n; ++k) n_fixed += opt.fixed[k] ? 1 : 0;
  S.fixed = n_fixed > 0;
  const double ndof = (double)std::max(n_fixed ? 3 * n - n_fixed : 3 * n - 3, 1);
  AK_REQUIRE(opt.rescale_steps >= 0, AK_ERR_ARG, "velocity_rescale_steps must be >= 0");
  AK_REQUIRE(opt.rescale_steps == 0 || opt.thermostat == 1 || opt.thermostat == 4, AK_ERR_ARG,
             "velocity_rescale_steps applies to the Nose-Hoover chain and NPT only");
  const double steps_done = st_at(kStSteps);
  AK_REQUIRE(steps_done >= 0.0 && steps_done == std::floor(steps_done), AK_ERR_ARG,
             "thermostat state: bad step count");
  S.vr_until = (int)std::max(0.0, std::min((double)opt.rescale_steps - steps_done,
                                           (double)n_steps + 1.0));
  S.T = opt.temperature;
  S.every = opt.flux_every;
  AK_REQUIRE(opt.flux_phase >= 0 && (S.every == 0 ? opt.flux_phase == 0 : opt.flux_phase < S.every),
             AK_ERR_ARG, "flux_phase must lie in [0, flux_every)");
  S.flux_phase = opt.flux_phase;
  S.pos.ensure(sizeof(double) * 3 * n);
  S.vel.ensure(sizeof(double) * 3 * n);
  S.acc.ensure(sizeof(double) * 3 * n);
  S.F.ensure(sizeof(double) * 3 * n);
  S.mass.ensure(sizeof(double) * n);
  S.spc.ensure(sizeof(int) * n);
  S.es.ensure(sizeof(double) * 10);
  const unsigned nb = (unsigned)((n + 255) / 256);
  S.ke.ensure(sizeof(double) * (S.flex() ? 7 : 1) * nb);
  // one record slot per step, plus the two set-up steps (dt = 0) before
  // the counters are reset, which record at slots 0 and 1 even when
  // n_steps is 0
  S.hist_ep.ensure(sizeof(double) * (n_steps + 2));
  S.hist_ek.ensure(sizeof(double) * (n_steps + 2));
  S.hist_ec.ensure(sizeof(double) * (n_steps + 2));
  S.hist_P.ensure(sizeof(double) * (n_steps + 2));
  S.hist_nb.ensure(sizeof(int) * (n_steps + 2));
  S.V = set.geoms[0].volume;
  if (S.thermo == 4 && (opt.cell_history || opt.ptensor_history)) {
    S.cell_hist = true;
    S.hist_L.ensure(sizeof(double) * 9 * (size_t)(n_steps + 2));
    S.hist_Pt.ensure(sizeof(double) * 6 * (size_t)(n_steps + 2));
  }
  AK_REQUIRE(opt.frame_every >= 0, AK_ERR_ARG, "frame_every must be >= 0");
  if (opt.frame_every > 0 && opt.frames) {
    S.frame_every = opt.frame_every;
    S.n_frames = n_steps / opt.frame_every;
    S.frames.ensure(sizeof(double) * 3 * (size_t)n * std::max(S.n_frames, 1));
    S.frame_lat.ensure(sizeof(double) * 9 * std::max(S.n_frames, 1));
  }
  S.lazy = S.thermo == 0 && S.every == 0 && S.frame_every == 0;
  S.counters.ensure(sizeof(int) * 4);
  S.flag.ensure(sizeof(int));
  // records at the steps s < n_steps with (s + phase) % every == 0
  const int frames = (S.every > 0 && n_steps > 0)
                         ? (n_steps - 1 + S.flux_phase) / S.every - (S.flux_phase > 0 ? 1 : 0) + 1
                         : 0;
  if (S.every > 0 && (opt.flux_groups || opt.atom_flux)) {
    // groups of the per-atom heat currents (ThermalConductivityCalculator.run_md
    // decompose, solvers/thermal_conductivity.py:339-363)
    AK_REQUIRE(opt.n_labels >= 1 && opt.n_regions >= 1 &&
                   (long long)opt.n_labels * opt.n_regions <= 1024,
               AK_ERR_ARG, "heat-flux groups: 1 to 1024 groups (labels x regions)");
    AK_REQUIRE(opt.region_axis >= -1 && opt.region_axis <= 2, AK_ERR_ARG,
               "heat-flux regions: axis must be 0, 1, 2 (or -1 for none)");
    AK_REQUIRE(opt.region_axis >= 0 || opt.n_regions == 1, AK_ERR_ARG,
               "heat-flux regions need an axis");
    S.fg.nc = opt.group_kinetic ? 5 : 3;
    S.fg.ng = opt.n_labels * opt.n_regions;
    S.fg.n_reg = opt.n_regions;
    S.fg.axis = opt.region_axis;
    if (opt.flux_labels) {
      for (int a = 0; a < n; ++a)
        AK_REQUIRE(opt.flux_labels[a] >= 0 && opt.flux_labels[a] < opt.n_labels, AK_ERR_ARG,
                   "heat-flux group labels must lie in [0, n_labels)");
      S.fg_labels.ensure(sizeof(int) * n);
      AK_HIP(hipMemcpy(S.fg_labels.p, opt.flux_labels, sizeof(int) * n, hipMemcpyHostToDevice));
      S.fg.labels = S.fg_labels.as<int>();
    } else {
      AK_REQUIRE(opt.n_labels == 1, AK_ERR_ARG, "heat-flux groups: n_labels > 1 needs labels");
    }
    S.fg.bounds = nullptr;
    if (opt.region_axis >= 0 && !opt.uniform_bins) {
      AK_REQUIRE(opt.region_bounds != nullptr, AK_ERR_ARG, "heat-flux regions need bounds");
      for (int k = 0; k <= opt.n_regions; ++k)
        AK_REQUIRE(std::isfinite(opt.region_bounds[k]) &&
                       (k == 0 || opt.region_bounds[k] >= opt.region_bounds[k - 1]),
                   AK_ERR_ARG, "heat-flux region bounds must be finite and ascending");
      S.fg_bounds.ensure(sizeof(double) * (opt.n_regions + 1));
      AK_HIP(hipMemcpy(S.fg_bounds.p, opt.region_bounds, sizeof(double) * (opt.n_regions + 1),
                       hipMemcpyHostToDevice));
      S.fg.bounds = S.fg_bounds.as<double>();
    }
    const size_t grid = ((size_t)n * 8 + 255) / 256;
    S.fg_part.ensure(sizeof(double) * S.fg.nc * grid * S.fg.ng);
    S.fg.part = S.fg_part.as<double>();
    const int fr = frames;
    S.hist_Jg.ensure(sizeof(double) * S.fg.nc * (size_t)std::max(fr, 1) * S.fg.ng);
    S.fg.out = S.hist_Jg.as<double>();
    S.fg.atoms = nullptr;
    if (opt.atom_flux) {
      S.hist_Ja.ensure(sizeof(float) * 3 * (size_t)n * std::max(fr, 1));
      S.fg.atoms = S.hist_Ja.as<float>();
    }
  }
  if (S.every > 0) {
    S.esite.ensure(sizeof(double) * n);
    S.flux_part.ensure(sizeof(double) * 3 * ((size_t)n / 32 + 1));
    S.hist_J.ensure(sizeof(double) * 3 * (size_t)std::max(frames, 1));
  }
  if (S.thermo == 1 || S.thermo == 4) {
    // Martyna, Tuckerman & Klein masses: Q_1 = n_dof kT tau^2, Q_j = kT tau^2
    // (_init_nhc_chain, solvers/molecular_dynamics.py:575-588), n_dof = 3N - 3
    AK_REQUIRE(opt.chain >= 2 && opt.chain <= kNhcMax, AK_ERR_ARG,
               "the Nose-Hoover chain needs 2 to 16 links");
    AK_REQUIRE(opt.temperature > 0.0 && opt.tau > 0.0, AK_ERR_ARG,
               "thermostat temperature and tau must be positive");
    S.M = opt.chain;
    S.kT = kBoltzmannEv * opt.temperature;
    S.ndof = ndof;
    std::vector<double> h(2 + 3 * S.M, 0.0);
    h[0] = h[1] = 1.0;
    h[2] = S.ndof * kBoltzmannEv * opt.temperature * opt.tau * opt.tau;
    for (int k = 1; k < S.M; ++k) h[2 + k] = kBoltzmannEv * opt.temperature * opt.tau * opt.tau;
    for (int k = 0; k < S.M; ++k) {
      h[2 + S.M + k] = st_at(kStXi + k);
      h[2 + 2 * S.M + k] = st_at(kStVxi + k);
    }
    S.nhc.ensure(sizeof(double) * h.size());
    AK_HIP(hipMemcpyAsync(S.nhc.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                          st));
    if (S.thermo == 4) {
      // barostat mass W = (n_dof + 3) kT tau_b^2 and chain masses kT tau_b^2
      // (NPTSimulation.__init__, :2370-2395); P_ext GPa -> eV/A^3
      AK_REQUIRE(opt.tau_baro > 0.0, AK_ERR_ARG, "tau_baro must be positive");
      S.npt.kT = S.kT;
      S.npt.ndof = S.ndof;
      S.npt.W = (S.ndof + 3.0) * kBoltzmannEv * opt.temperature * opt.tau_baro * opt.tau_baro;
      S.npt.Qb = kBoltzmannEv * opt.temperature * opt.tau_baro * opt.tau_baro;
      S.npt.P_ext = opt.pressure_gpa / kEvAng3ToGpa;
      // the cell may shrink while the grid and the minimum image hold
      const Geom& g = set.geoms[0];
      CellInfo ci = cell_info(g.lat);
      double s_min = 2.0 * nl.r_list / ci.min_width;
      for (int ax = 0; ax < 3; ++ax)
        s_min = std::max(s_min, nl.r_list * g.nc[ax] / ci.perp[ax]);
      S.npt.s_min = s_min;
      // flexible cell: the strain generators of _cell_generators (:2113-2138)
      // as xx, yy, zz, yz, xz, xy; iso keeps one momentum p_eps
      static const double kGen[6][6] = {{1, 0, 0, 0, 0, 0}, {0, 1, 0, 0, 0, 0},
                                        {0, 0, 1, 0, 0, 0}, {0, 0, 0, 1, 0, 0},
                                        {0, 0, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 1}};
      static const double kXyIso[6] = {1, 1, 0, 0, 0, 0};
      int ng = 1;
      const double* gsrc[6] = {kGen[0], kGen[1], kGen[2], kGen[3], kGen[4], kGen[5]};
      switch (opt.cell_dof) {
        case 1: ng = 3; break;                                   // aniso: xx, yy, zz
        case 2: ng = 6; break;                                   // tri: + yz, xz, xy
        case 3: ng = 2; break;                                   // xy: xx, yy
        case 4: ng = 1; gsrc[0] = kXyIso; break;                 // xy_iso: xx + yy
        case 5: ng = 1; gsrc[0] = kGen[2]; break;                // z: zz
        default: ng = 1; break;                                  // iso
      }
      S.npt.nb = ng;
      for (int k = 0; k < 6; ++k)
        for (int c = 0; c < 6; ++c) S.npt.gen[k][c] = k < ng ? gsrc[k][c] : 0.0;
      for (int k = 0; k < 6; ++k)
        S.npt.gtr[k] = k < ng ? S.npt.gen[k][0] + S.npt.gen[k][1] + S.npt.gen[k][2] : 0.0;
      S.npt.Qb1 = (opt.cell_dof == 0 ? 1.0 : (double)ng) * S.npt.Qb;
      S.npt.lv = opt.baro_thermostat == 1 ? 1 : 0;
      S.npt.tau_b = opt.tau_baro;
      S.npt.k0 = (uint32_t)(opt.seed & 0xffffffffu);
      S.npt.k1 = (uint32_t)(opt.seed >> 32);
      S.npt.vr_until = S.vr_until;
      S.npt.T = S.T;
      std::vector<double> q(kNpLen, 0.0);
      q[0] = q[1] = q[7] = q[8] = 1.0;
      q[2] = st_at(kStPeps);
      q[3] = st_at(kStXiB);
      q[4] = st_at(kStXiB + 1);
      q[5] = st_at(kStVxiB);
      q[6] = st_at(kStVxiB + 1);
      q[kNpHeat] = st_at(kStHeat);
      for (int k = 0; k < 6; ++k) q[kNpPg + k] = k < ng && opt.cell_dof != 0 ? st_at(kStPg + k)
                                                                             : 0.0;
      for (int k = 0; k < 9; ++k) {
        q[kNpLb + k] = lattice[k];
        q[kNpLbInv + k] = inv[k];
      }
      S.np.ensure(sizeof(double) * q.size());
      AK_HIP(hipMemcpy(S.np.p, q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice));
      S.hist_V.ensure(sizeof(double) * (n_steps + 2));
    }
    AK_HIP(hipStreamSynchronize(st));
  } else if (S.thermo == 2 || S.thermo == 3) {
    // Langevin (damping = 1/gamma) and CSVR (damping = tau of K): state
    // [pending scale, drift scale, heat Q, -, v_com (3)]; n_dof = 3N - 3
    AK_REQUIRE(opt.temperature > 0.0 && opt.tau > 0.0, AK_ERR_ARG,
               "thermostat temperature and damping must be positive");
    AK_REQUIRE(n >= 2, AK_ERR_ARG, "a stochastic thermostat needs at least two atoms");
    S.kT = kBoltzmannEv * opt.temperature;
    S.ndof = ndof;
    S.damping = opt.tau;
    S.k0 = (uint32_t)(opt.seed & 0xffffffffu);
    S.k1 = (uint32_t)(opt.seed >> 32);
    double M = 0.0;
    for (int a = 0; a < n; ++a) M += mass[a];
    S.inv_mass = 1.0 / M;
    std::vector<double> h(8, 0.0);
    h[0] = h[1] = 1.0;
    h[2] = st_at(kStHeat);
    S.nhc.ensure(sizeof(double) * h.size());
    AK_HIP(hipMemcpyAsync(S.nhc.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                          st));
    S.ke_b.ensure(sizeof(double) * nb);
    S.ke_a.ensure(sizeof(double) * nb);
    S.p_part.ensure(sizeof(double) * 3 * nb);
    AK_HIP(hipStreamSynchronize(st));
  } else if (S.thermo == 5) {
    // NEMD: hot bath at `temperature`, cold at T_cold, damping tau; state
    // [-, -, heat Q = Q_hot + Q_cold]
    AK_REQUIRE(opt.bath != nullptr, AK_ERR_ARG, "NEMD needs the bath of every atom");
    AK_REQUIRE(opt.temperature >= 0.0 && opt.T_cold >= 0.0 && opt.tau > 0.0, AK_ERR_ARG,
               "NEMD bath temperatures must be >= 0 and the damping positive");
    AK_REQUIRE(!S.fixed, AK_ERR_ARG, "NEMD does not take frozen atoms");
    int n_hot = 0, n_cold = 0;
    for (int a = 0; a < n; ++a) {
      AK_REQUIRE(opt.bath[a] <= 2, AK_ERR_ARG, "NEMD bath codes are 0 (none), 1 hot, 2 cold");
      n_hot += opt.bath[a] == 1;
      n_cold += opt.bath[a] == 2;
    }
    AK_REQUIRE(n_hot > 0 && n_cold > 0, AK_ERR_ARG, "NEMD: a heat bath has no atoms");
    S.kT = kBoltzmannEv * opt.temperature;
    S.kT_cold = kBoltzmannEv * opt.T_cold;
    S.ndof = ndof;
    S.damping = opt.tau;
    S.k0 = (uint32_t)(opt.seed & 0xffffffffu);
    S.k1 = (uint32_t)(opt.seed >> 32);
    std::vector<double> h(8, 0.0);
    h[0] = h[1] = 1.0;
    h[2] = st_at(kStHeat);
    S.nhc.ensure(sizeof(double) * h.size());
    AK_HIP(hipMemcpy(S.nhc.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    S.bath.ensure((size_t)n);
    AK_HIP(hipMemcpy(S.bath.p, opt.bath, (size_t)n, hipMemcpyHostToDevice));
    S.ke_b.ensure(sizeof(double) * 5 * nb);
    S.hist_bath.ensure(sizeof(double) * 4 * (size_t)(n_steps + 2));
  }
  AK_HIP(hipMemcpyAsync(S.pos.p, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  if (S.fixed) {
    std::vector<double> kp(3 * (size_t)n);
    for (size_t k = 0; k < kp.size(); ++k) {
      kp[k] = opt.fixed[k] ? 0.0 : 1.0;
      if (opt.fixed[k]) vel[k] = 0.0;
    }
    S.keep.ensure(sizeof(double) * kp.size());
    AK_HIP(hipMemcpy(S.keep.p, kp.data(), sizeof(double) * kp.size(), hipMemcpyHostToDevice));
  }
  AK_HIP(hipMemcpyAsync(S.vel.p, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  AK_HIP(hipMemcpyAsync(S.mass.p, mass, sizeof(double) * n, hipMemcpyHostToDevice, st));
  std::vector<int32_t> sp(n, 0);
  if (spc) sp.assign(spc, spc + n);
  AK_HIP(hipMemcpyAsync(S.spc.p, sp.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  AK_HIP(hipMemsetAsync(S.counters.p, 0, sizeof(int) * 4, st));
  AK_HIP(hipMemsetAsync(S.flag.p, 0, sizeof(int), st));

  // initial list with head room for the captured rebuilds: rows of one and a
  // half times the current longest (at least four more), candidate buffers of
  // at least twice the densest stencil
  nl_build(nl, set, S.pos.as<double>(), st);
  nl.cmax = std::max(nl.cmax, pow2ceil(std::max(64, 2 * nl.max_cand)));
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
  kick(S, n, dt, 0, st);
  AK_HIP(hipMemsetAsync(nl.stats.p, 0, sizeof(int) * 8, st));
  // make sure every scratch buffer a step touches exists before capture: a
  // step with dt = 0 leaves positions, velocities, the chain and the state-0
  // sums as they were
  md_step(pot, nl, set, S, n, 0.0, st);
  // with a sampled flux, the steps between records take the plain force path
  // (its launch shape and chunk table may differ): set that up too
  if (S.every > 0) md_step(pot, nl, set, S, n, 0.0, st, false);
  AK_HIP(hipStreamSynchronize(st));
  // restart the record, the rebuild count and the flux gate: step k records
  // state k - 1 (NVE) or k (chain; state 0 recorded here); the random
  // stream continues after the draws a previous run used
  const double rng0 = st_at(kStRng);
  AK_REQUIRE(rng0 >= 0.0 && rng0 == std::floor(rng0) && rng0 + 2.0 * n_steps + 4.0 < 2147483647.0,
             AK_ERR_ARG, "thermostat state: bad random-draw count");
  int reset[4] = {0, 0, 0, (int)rng0};
  AK_HIP(hipMemcpy(S.counters.as<int>(), reset, sizeof(reset), hipMemcpyHostToDevice));
  if (S.thermo) md_record(S, n, st);

  // the steps s with (s + phase) % every == 0 record the heat flux; the
  // others run the plain force path and launch no flux kernel
  const auto records = [&](int s) {
    return S.every > 0 && (s + S.flux_phase) % S.every == 0;
  };
  auto t0 = std::chrono::steady_clock::now();
  if (use_graph && n_steps > 0) {
    // one captured step of each kind that occurs; without flux records, a
    // graph of kBurst steps carries the run and single steps the remainder
    constexpr int kBurst = 8;
    const bool burst = S.every == 0 && n_steps >= kBurst;
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    bool need[3] = {false, false, burst};
    const int singles_from = burst ? n_steps - n_steps % kBurst : 0;
    for (int s = singles_from; s < n_steps && !(need[0] && need[1]); ++s)
      need[records(s) ? 1 : 0] = true;
    for (int k = 0; k < 3; ++k) {
      if (!need[k]) continue;
      AK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      for (int r = 0; r < (k == 2 ? kBurst : 1); ++r) md_step(pot, nl, set, S, n, dt, st, k == 1);
      AK_HIP(hipStreamEndCapture(st, &graph[k]));
      AK_HIP(hipGraphInstantiate(&exec[k], graph[k], nullptr, nullptr, 0));
    }
    AK_HIP(hipStreamSynchronize(st));
    t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < singles_from; s += kBurst) AK_HIP(hipGraphLaunch(exec[2], st));
    for (int s = singles_from; s < n_steps; ++s)
      AK_HIP(hipGraphLaunch(exec[records(s) ? 1 : 0], st));
    if (S.thermo == 0) md_finish(S, n, dt, n_steps, st);
    AK_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    for (int k = 0; k < 3; ++k) {
      if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
      if (graph[k]) (void)hipGraphDestroy(graph[k]);
    }
  } else {
    for (int s = 0; s < n_steps; ++s) md_step(pot, nl, set, S, n, dt, st, records(s));
    if (S.thermo == 0) md_finish(S, n, dt, n_steps, st);
    AK_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  }
  if (S.scaled())
    hipLaunchKernelGGL(k_scale_vel, dim3((unsigned)((3 * n + 255) / 256)), dim3(256), 0, st, n,
                       S.vel.as<double>(), (const double*)S.nhc.as<double>());
  AK_HIP(hipStreamSynchronize(st));
  AK_REQUIRE(nl_stats_overflow(nl, st) == 0, AK_ERR_CAPACITY,
             "neighbour capacity of the captured MD step was exceeded");
  int counters[4];
  AK_HIP(hipMemcpy(counters, S.counters.p, sizeof(counters), hipMemcpyDeviceToHost));
  AK_REQUIRE(counters[0] == n_steps + 1, AK_ERR_STATE, "MD step counter mismatch");
  AK_REQUIRE(S.every == 0 || counters[2] == frames, AK_ERR_STATE, "heat-flux gate mismatch");
  if (state) state[kStSteps] = steps_done + n_steps;
  if (state && (S.thermo == 1 || S.thermo == 4)) {
    std::vector<double> h(2 + 3 * S.M);
    AK_HIP(hipMemcpy(h.data(), S.nhc.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < S.M; ++k) {
      state[kStXi + k] = h[2 + S.M + k];
      state[kStVxi + k] = h[2 + 2 * S.M + k];
    }
  } else if (state && (S.thermo == 2 || S.thermo == 3 || S.thermo == 5)) {
    double h[3];
    AK_HIP(hipMemcpy(h, S.nhc.p, sizeof(h), hipMemcpyDeviceToHost));
    state[kStHeat] = h[2];
    state[kStRng] = (double)counters[3];
  }
  if (S.thermo == 4) {
    double q[kNpLen];
    AK_HIP(hipMemcpy(q, S.np.p, sizeof(q), hipMemcpyDeviceToHost));
    if (state) {
      state[kStPeps] = q[2];
      state[kStXiB] = q[3];
      state[kStXiB + 1] = q[4];
      state[kStVxiB] = q[5];
      state[kStVxiB + 1] = q[6];
      for (int k = 0; k < 6; ++k) state[kStPg + k] = q[kNpPg + k];
      state[kStHeat] = q[kNpHeat];
      state[kStRng] = (double)counters[3];
    }
    AK_REQUIRE(q[9] == 0.0, AK_ERR_CAPACITY,
               S.flex() ? "the NPT cell left the range its neighbour grid and minimum image were "
                          "set up for (perpendicular widths 5 % below the start); restart from "
                          "the current cell"
                        : "the NPT cell shrank beyond the range its neighbour grid and minimum "
                          "image were set up for (5 % linear); restart from the current cell");
    Geom g;
    AK_HIP(hipMemcpy(&g, S.geom, sizeof(Geom), hipMemcpyDeviceToHost));
    if (lattice_out) std::memcpy(lattice_out, g.lat, sizeof(g.lat));
    if (volume)
      AK_HIP(hipMemcpy(volume, S.hist_V.p, sizeof(double) * (n_steps + 1),
                       hipMemcpyDeviceToHost));
    if (pressure)
      AK_HIP(hipMemcpy(pressure, S.hist_P.p, sizeof(double) * (n_steps + 1),
                       hipMemcpyDeviceToHost));
  } else {
    if (lattice_out) std::memcpy(lattice_out, lattice, 9 * sizeof(double));
    if (volume)
      for (int k = 0; k <= n_steps; ++k) volume[k] = S.V;
    if (pressure)
      AK_HIP(hipMemcpy(pressure, S.hist_P.p, sizeof(double) * (n_steps + 1),
                       hipMemcpyDeviceToHost));
  }
  AK_HIP(hipMemcpy(pos, S.pos.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  AK_HIP(hipMemcpy(vel, S.vel.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  if (e_pot)
    AK_HIP(hipMemcpy(e_pot, S.hist_ep.p, sizeof(double) * (n_steps + 1),
                     hipMemcpyDeviceToHost));
  if (e_kin)
    AK_HIP(hipMemcpy(e_kin, S.hist_ek.p, sizeof(double) * (n_steps + 1),
                     hipMemcpyDeviceToHost));
  if (e_cons)
    AK_HIP(hipMemcpy(e_cons, S.hist_ec.p, sizeof(double) * (n_steps + 1),
                     hipMemcpyDeviceToHost));
  if (flux && frames > 0)
    AK_HIP(hipMemcpy(flux, S.hist_J.p, sizeof(double) * 3 * frames, hipMemcpyDeviceToHost));
  if (opt.flux_groups && S.fg.ng > 0 && frames > 0)
    AK_HIP(hipMemcpy(opt.flux_groups, S.hist_Jg.p,
                     sizeof(double) * S.fg.nc * (size_t)frames * S.fg.ng, hipMemcpyDeviceToHost));
  if (S.thermo == 5 && opt.bath_history)
    AK_HIP(hipMemcpy(opt.bath_history, S.hist_bath.p, sizeof(double) * 4 * (n_steps + 1),
                     hipMemcpyDeviceToHost));
  if (opt.atom_flux && S.fg.atoms && frames > 0)
    AK_HIP(hipMemcpy(opt.atom_flux, S.hist_Ja.p, sizeof(float) * 3 * (size_t)n * frames,
                     hipMemcpyDeviceToHost));
  if (S.frame_every > 0 && S.n_frames > 0) {
    AK_HIP(hipMemcpy(opt.frames, S.frames.p, sizeof(double) * 3 * (size_t)n * S.n_frames,
                     hipMemcpyDeviceToHost));
    if (opt.frame_lattices) {
      if (S.thermo == 4)
        AK_HIP(hipMemcpy(opt.frame_lattices, S.frame_lat.p, sizeof(double) * 9 * S.n_frames,
                         hipMemcpyDeviceToHost));
      else
        for (int f = 0; f < S.n_frames; ++f)
          std::memcpy(opt.frame_lattices + 9 * (size_t)f, lattice, 9 * sizeof(double));
    }
  }
  if (n_rebuilds) *n_rebuilds = counters[1];
  if (opt.rebuild_counts)
    AK_HIP(hipMemcpy(opt.rebuild_counts, S.hist_nb.p, sizeof(int) * (n_steps + 1),
                     hipMemcpyDeviceToHost));
  for (int k = 0; k <= n_steps; ++k) {
    // a fixed cell on the host; the pressure tensor only for flexible cells
    if (opt.cell_history && !S.cell_hist)
      std::memcpy(opt.cell_history + 9 * (size_t)k, lattice, 9 * sizeof(double));
    if (opt.ptensor_history && !S.flex())
      for (int c = 0; c < 6; ++c) opt.ptensor_history[6 * (size_t)k + c] = NAN;
  }
  if (S.cell_hist && opt.cell_history)
    AK_HIP(hipMemcpy(opt.cell_history, S.hist_L.p, sizeof(double) * 9 * (n_steps + 1),
                     hipMemcpyDeviceToHost));
  if (S.flex() && S.cell_hist && opt.ptensor_history)
    AK_HIP(hipMemcpy(opt.ptensor_history, S.hist_Pt.p, sizeof(double) * 6 * (n_steps + 1),
                     hipMemcpyDeviceToHost));
}

}  // namespace ak
