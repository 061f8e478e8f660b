// Tersoff parameter gradient of the fitting loss (hand-written backward).
//
// Loss over structures s (fitting/loss.py:212-242, squared):
//   L = sum_s w_E (E_s - N_s e_s)^2 + w_F r0^2/(3 N_s) |F_s - F_s^ref|^2
//           + w_S (V_s eps0)^2 / 9 |sigma_s - sigma_s^ref|^2
// With kappa_s = dL/dE_s, lambda_a = dL/dF_a, mu_s = dL/dsigma_s and the site
// form of the forces (F_a = sum_j G_aj - G_ja, sigma = -(1/V) sum G (x) r,
// G_ij = dE_i/dr_ij):
//   sum_a lambda_a . F_a + mu : sigma = sum_{bonds i->j} G_ij . w_ij,
//   w_ij = lambda_i - lambda_j - mu r_ij / V,
// which is the derivative of the site energies when every bond vector moves
// along w_ij.  Hence
//   dL/dtheta_p = sum_s kappa_s dE_s/dtheta_p + sum_i d^2 E_i / (d eps d theta_p),
// and both terms come from one evaluation of each site energy in hyper-dual
// numbers (value, d/d eps, d/d theta_p, d^2/d eps d theta_p), one parameter
// per grid row: the mixed second derivative needs no stored bond-order
// intermediates, the bond orders are recomputed per structure and parameter.
//
// theta = A, B, lambda1, lambda2, lambda3, beta, n, c, d, h, R, D, m, gamma
// (the single-species record of engine/tersoff.py:528-855 with explicit m and
// gamma).  (r_ij - r_ik)^m is taken as sign(x)|x|^m, which equals the
// reference's pow for the integer m it admits (m = 1 or 3).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "engine.hpp"

namespace ak {

// ---------------------------------------------------------------------------
// Hyper-dual numbers: x = v + e eps + t tau + et eps tau
// ---------------------------------------------------------------------------
struct HD {
  double v, e, t, et;
};

__device__ __forceinline__ HD hd(double v) { return {v, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ HD operator+(HD a, HD b) {
  return {a.v + b.v, a.e + b.e, a.t + b.t, a.et + b.et};
}
__device__ __forceinline__ HD operator-(HD a, HD b) {
  return {a.v - b.v, a.e - b.e, a.t - b.t, a.et - b.et};
}
__device__ __forceinline__ HD operator-(HD a) { return {-a.v, -a.e, -a.t, -a.et}; }
__device__ __forceinline__ HD operator+(HD a, double b) { return {a.v + b, a.e, a.t, a.et}; }
__device__ __forceinline__ HD operator+(double b, HD a) { return a + b; }
__device__ __forceinline__ HD operator-(double b, HD a) { return {b - a.v, -a.e, -a.t, -a.et}; }
__device__ __forceinline__ HD operator*(HD a, HD b) {
  return {a.v * b.v, a.e * b.v + a.v * b.e, a.t * b.v + a.v * b.t,
          a.et * b.v + a.e * b.t + a.t * b.e + a.v * b.et};
}
__device__ __forceinline__ HD operator*(double s, HD a) {
  return {s * a.v, s * a.e, s * a.t, s * a.et};
}
__device__ __forceinline__ HD operator*(HD a, double s) { return s * a; }
// f(a) from f, f', f''
__device__ __forceinline__ HD chain(HD a, double f0, double f1, double f2) {
  return {f0, f1 * a.e, f1 * a.t, f1 * a.et + f2 * a.e * a.t};
}
__device__ __forceinline__ HD hinv(HD a) {
  double i = 1.0 / a.v;
  return chain(a, i, -i * i, 2.0 * i * i * i);
}
__device__ __forceinline__ HD operator/(HD a, HD b) { return a * hinv(b); }
__device__ __forceinline__ HD hexp(HD a) {
  double x = exp(a.v);
  return chain(a, x, x, x);
}
__device__ __forceinline__ HD hlog(HD a) {
  double i = 1.0 / a.v;
  return chain(a, log(a.v), i, -i * i);
}
__device__ __forceinline__ HD hsqrt(HD a) {
  double s = sqrt(a.v);
  return chain(a, s, 0.5 / s, -0.25 / (s * a.v));
}
__device__ __forceinline__ HD hsin(HD a) {
  double s, c;
  sincos(a.v, &s, &c);
  return chain(a, s, c, -s);
}
// sign(x) |x|^m for m > 0
__device__ __forceinline__ HD oddpow(HD x, HD m) {
  if (x.v == 0.0) return m.v == 1.0 ? x : hd(0.0);
  double sg = x.v > 0.0 ? 1.0 : -1.0;
  HD ax = sg * x;
  return sg * hexp(m * hlog(ax));
}

__device__ __forceinline__ HD fc_hd(HD r, HD R, HD D) {
  if (r.v < R.v - D.v) return hd(1.0);
  if (r.v > R.v + D.v) return hd(0.0);
  return 0.5 - 0.5 * hsin((0.5 * M_PI) * (r - R) / D);
}

struct Theta {
  HD A, B, l1, l2, l3, beta, n, c, d, h, R, D, m, gam;
};

__device__ __forceinline__ Theta load_theta(const double* __restrict__ th, int p) {
  HD x[14];
#pragma unroll
  for (int q = 0; q < 14; ++q) x[q] = {th[q], 0.0, q == p ? 1.0 : 0.0, 0.0};
  return {x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8], x[9], x[10], x[11], x[12],
          x[13]};
}

// One parameter per grid row (blockIdx.y); L lanes per atom as the forward
// site kernel.  Writes dE[p*n + a] = dE_a/dtheta_p and dD[p*n + a] =
// d^2 E_a / (d eps d theta_p).
template <int L>
__global__ __launch_bounds__(256) void k_tersoff_hd(
    int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, const double* __restrict__ theta,
    const double* __restrict__ lam, const double* __restrict__ muv,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    double* __restrict__ dE, double* __restrict__ dD) {
  const int p = blockIdx.y;
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = gt / L;
  const int k = gt & (L - 1);
  const int seg = (threadIdx.x & 63) & ~(L - 1);
  const bool live = a < n;
  const int ac = live ? a : 0;
  const int nk = live ? cnt[ac] : 0;
  const bool has = live && k < nk;
  const int j = has ? nbr[(size_t)ac * stride + k] : ac;
  const Theta P = load_theta(theta, p);
  const double rc = theta[10] + theta[11];

  double ux = 0, uy = 0, uz = 0, wx = 0, wy = 0, wz = 0;
  bool act = false;
  if (has) {
    const int s = sid[ac];
    const Geom& g = geoms[s];
    double rx = pos[3 * j] - pos[3 * ac], ry = pos[3 * j + 1] - pos[3 * ac + 1],
           rz = pos[3 * j + 2] - pos[3 * ac + 2];
    double sx = rx * g.inv[0] + ry * g.inv[3] + rz * g.inv[6];
    double sy = rx * g.inv[1] + ry * g.inv[4] + rz * g.inv[7];
    double sz = rx * g.inv[2] + ry * g.inv[5] + rz * g.inv[8];
    sx -= rint(sx);
    sy -= rint(sy);
    sz -= rint(sz);
    ux = sx * g.lat[0] + sy * g.lat[3] + sz * g.lat[6];
    uy = sx * g.lat[1] + sy * g.lat[4] + sz * g.lat[7];
    uz = sx * g.lat[2] + sy * g.lat[5] + sz * g.lat[8];
    double r = sqrt(ux * ux + uy * uy + uz * uz);
    act = r < rc && r > 1e-12;
    if (act) {
      const double* M = muv + 9 * (size_t)s;  // mu_s / V_s
      wx = lam[3 * ac] - lam[3 * j] - (M[0] * ux + M[1] * uy + M[2] * uz);
      wy = lam[3 * ac + 1] - lam[3 * j + 1] - (M[3] * ux + M[4] * uy + M[5] * uz);
      wz = lam[3 * ac + 2] - lam[3 * j + 2] - (M[6] * ux + M[7] * uy + M[8] * uz);
    } else {
      ux = uy = uz = 0.0;
    }
  }
  // |u + eps w|
  const double uu = ux * ux + uy * uy + uz * uz;
  // eps^2 terms drop out: only eps, tau and eps*tau orders are carried
  HD r = act ? hsqrt(HD{uu, 2.0 * (ux * wx + uy * wy + uz * wz), 0.0, 0.0}) : hd(0.0);
  HD fcl = act ? fc_hd(r, P.R, P.D) : hd(0.0);

  HD zeta = hd(0.0);
  for (int s = 0; s < L; ++s) {
    const int src = seg + s;
    double vx = __shfl(ux, src, 64), vy = __shfl(uy, src, 64), vz = __shfl(uz, src, 64);
    double qx = __shfl(wx, src, 64), qy = __shfl(wy, src, 64), qz = __shfl(wz, src, 64);
    int pact = __shfl((int)act, src, 64);
    if (s == k || !act || !pact) continue;
    double vv = vx * vx + vy * vy + vz * vz;
    HD rs = hsqrt(HD{vv, 2.0 * (vx * qx + vy * qy + vz * qz), 0.0, 0.0});
    HD fcs = fc_hd(rs, P.R, P.D);
    if (fcs.v == 0.0) continue;
    HD dot = {ux * vx + uy * vy + uz * vz, wx * vx + wy * vy + wz * vz + ux * qx + uy * qy +
                                               uz * qz,
              0.0, 0.0};
    HD cs = dot / (r * rs);
    HD hc = P.h - cs;
    HD c2 = P.c * P.c, d2 = P.d * P.d;
    HD g = 1.0 + c2 / d2 - c2 / (d2 + hc * hc);
    HD ex = hexp(oddpow(P.l3 * (r - rs), P.m));
    zeta = zeta + P.gam * fcs * g * ex;
  }
  HD Ek = hd(0.0);
  if (act) {
    HD b = hd(1.0);
    if (zeta.v > 0.0 && P.beta.v != 0.0) {
      HD bz = hexp(P.n * hlog(P.beta * zeta));
      b = hexp(-(hlog(1.0 + bz) / (2.0 * P.n)));
    }
    HD fr = P.A * hexp(-(P.l1 * r));
    HD fa = -(P.B * hexp(-(P.l2 * r)));
    Ek = 0.5 * fcl * (fr + b * fa);
  }
  double de = seg_sum<L>(Ek.t), dd = seg_sum<L>(Ek.et);
  if (live && k == 0) {
    dE[(size_t)p * n + ac] = de;
    dD[(size_t)p * n + ac] = dd;
  }
}

// Per atom: lambda_a = 2 w_F r0^2 / (3 N_s) (F_a - F_a^ref) and the atom's
// share of the force term of the loss.
__global__ void k_fit_atoms(int n, const int* __restrict__ sid, const Geom* __restrict__ geoms,
                            const double* __restrict__ F, const double* __restrict__ Fref,
                            double wF, double* __restrict__ lam, double* __restrict__ lf) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const Geom& g = geoms[sid[a]];
  double scale = wF / (3.0 * g.natoms);
  double s2 = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double r = F[3 * a + d] - Fref[3 * a + d];
    lam[3 * a + d] = 2.0 * scale * r;
    s2 += r * r;
  }
  lf[a] = scale * s2;
}

// Per structure (one block): loss, kappa_s, mu_s / V_s.
//   out_es (S, 10) from the forward pass; ref (S, 10) = [e per atom, sigma];
//   loss_s (S), kappa (S), muv (S, 9).
__global__ __launch_bounds__(256) void k_fit_structs(
    const Geom* __restrict__ geoms, const double* __restrict__ es,
    const double* __restrict__ ref, const double* __restrict__ lf, double wE, double wS,
    double eps0, double* __restrict__ loss_s, double* __restrict__ kappa,
    double* __restrict__ muv) {
  __shared__ double smem[16];
  const int s = blockIdx.x;
  const Geom& g = geoms[s];
  double v[1] = {0.0};
  for (int a = g.atom_off + threadIdx.x; a < g.atom_off + g.natoms; a += blockDim.x)
    v[0] += lf[a];
  block_sum<1>(v, smem);
  if (threadIdx.x == 0) {
    const double* E = es + 10 * (size_t)s;
    const double* R = ref + 10 * (size_t)s;
    double N = g.natoms;
    double dEs = E[0] - N * R[0];
    double sc = wS * (g.volume * eps0) * (g.volume * eps0) / 9.0;
    double ls = 0.0;
    for (int q = 0; q < 9; ++q) {
      double d = E[1 + q] - R[1 + q];
      ls += d * d;
      muv[9 * (size_t)s + q] = 2.0 * sc * d / g.volume;
    }
    loss_s[s] = wE * dEs * dEs + v[0] + sc * ls;
    kappa[s] = 2.0 * wE * dEs;
  }
}

// grad_s[p] = kappa_s sum_a dE[p, a] + sum_a dD[p, a] (one block per (s, p)).
__global__ __launch_bounds__(256) void k_fit_reduce(const Geom* __restrict__ geoms,
                                                    const double* __restrict__ dE,
                                                    const double* __restrict__ dD,
                                                    const double* __restrict__ kappa, int n,
                                                    double* __restrict__ grad_s) {
  __shared__ double smem[32];
  const int s = blockIdx.x, p = blockIdx.y;
  const Geom& g = geoms[s];
  double v[2] = {0.0, 0.0};
  for (int a = g.atom_off + threadIdx.x; a < g.atom_off + g.natoms; a += blockDim.x) {
    v[0] += dE[(size_t)p * n + a];
    v[1] += dD[(size_t)p * n + a];
  }
  block_sum<2>(v, smem);
  if (threadIdx.x == 0) grad_s[14 * (size_t)s + p] = kappa[s] * v[0] + v[1];
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------

struct FitState {
  int S = 0, n = 0;
  std::vector<int64_t> off;
  std::vector<double> h_lat, h_pos, h_ref, h_fref;
  SetDev set;
  NList nl;
  DevBuf pos, spc, fref, ref, F, es, lam, lf, loss_s, kappa, muv, dE, dD, grad_s, theta;
};

static FitState& fit_state(Potential& pot) {
  if (!pot.fit) pot.fit = std::make_shared<FitState>();
  return *static_cast<FitState*>(pot.fit.get());
}

template <int L>
static void launch_hd(FitState& F, int n, hipStream_t st) {
  const int64_t threads = (int64_t)n * L;
  dim3 grid((unsigned)((threads + 255) / 256), 14);
  hipLaunchKernelGGL(k_tersoff_hd<L>, grid, dim3(256), 0, st, n, F.pos.as<double>(),
                     F.set.d_sid.as<int>(), F.set.d_geoms.as<Geom>(), F.theta.as<double>(),
                     F.lam.as<double>(), F.muv.as<double>(), F.nl.cnt.as<int>(),
                     F.nl.nbr.as<int>(), F.nl.stride, F.dE.as<double>(), F.dD.as<double>());
}

void tersoff_loss_grad_run(Potential& pot, int S, const int64_t* off, const double* lattices,
                           const double* positions, const double* theta,
                           const double* ref_energy, const double* ref_forces,
                           const double* ref_stress, const double* weights, double* loss,
                           double* grad, double* energies, double* forces, double* stresses,
                           hipStream_t st) {
  for (int q = 0; q < 14; ++q)
    AK_REQUIRE(std::isfinite(theta[q]), AK_ERR_ARG, "non-finite Tersoff parameter");
  const double A = theta[0], B = theta[1], beta = theta[5], nn = theta[6], c = theta[7],
               d = theta[8], R = theta[10], D = theta[11], m = theta[12];
  AK_REQUIRE(A > 0 && B > 0 && beta > 0 && nn > 0 && c > 0 && d > 0 && D > 0 && R > D &&
                 m > 0 && theta[4] >= 0,
             AK_ERR_ARG, "Tersoff parameters out of their domain");
  const int n = (int)off[S];
  AK_REQUIRE(n > 0, AK_ERR_ARG, "no atoms");
  FitState& F = fit_state(pot);
  const double rc = R + D;

  // parameter tables of the forward kernels (one species)
  double tb[9] = {1.0, A, B, theta[2], theta[3], R, D, beta, nn};
  double tri[7] = {1.0, c, d, theta[9], theta[4], m, theta[13]};
  std::memcpy(pot.h_params.data(), tb, sizeof tb);
  std::memcpy(pot.h_params.data() + 9, tri, sizeof tri);
  AK_HIP(hipMemcpyAsync(pot.d_params.p, pot.h_params.data(), sizeof(double) * 16,
                        hipMemcpyHostToDevice, st));
  pot.r_cut = rc;
  F.theta.ensure(sizeof(double) * 14);
  AK_HIP(hipMemcpyAsync(F.theta.p, theta, sizeof(double) * 14, hipMemcpyHostToDevice, st));

  // structures: upload once, reuse while the inputs are unchanged
  std::vector<double> ref(10 * (size_t)S);
  for (int s = 0; s < S; ++s) {
    ref[10 * s] = ref_energy[s];
    std::memcpy(ref.data() + 10 * s + 1, ref_stress + 9 * s, 9 * sizeof(double));
  }
  bool same = F.S == S && F.n == n && std::equal(F.off.begin(), F.off.end(), off) &&
              F.off.size() == (size_t)S + 1 &&
              std::memcmp(F.h_lat.data(), lattices, sizeof(double) * 9 * S) == 0 &&
              std::memcmp(F.h_pos.data(), positions, sizeof(double) * 3 * n) == 0 &&
              std::memcmp(F.h_fref.data(), ref_forces, sizeof(double) * 3 * n) == 0 &&
              F.h_ref == ref;
  if (!same) {
    F.S = S;
    F.n = n;
    F.off.assign(off, off + S + 1);
    F.h_lat.assign(lattices, lattices + 9 * (size_t)S);
    F.h_pos.assign(positions, positions + 3 * (size_t)n);
    F.h_fref.assign(ref_forces, ref_forces + 3 * (size_t)n);
    F.h_ref = ref;
    F.pos.ensure(sizeof(double) * 3 * n);
    F.fref.ensure(sizeof(double) * 3 * n);
    F.ref.ensure(sizeof(double) * 10 * S);
    F.spc.ensure(sizeof(int) * n);
    AK_HIP(hipMemcpyAsync(F.pos.p, positions, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                          st));
    AK_HIP(hipMemcpyAsync(F.fref.p, ref_forces, sizeof(double) * 3 * n,
                          hipMemcpyHostToDevice, st));
    AK_HIP(hipMemcpyAsync(F.ref.p, ref.data(), sizeof(double) * 10 * S, hipMemcpyHostToDevice,
                          st));
    AK_HIP(hipMemsetAsync(F.spc.p, 0, sizeof(int) * n, st));
    F.nl.valid = false;
  }
  if (!F.nl.valid || F.nl.r_cut < rc) {
    F.nl.r_cut = rc * 1.05;  // head room for fitted R and D
    F.nl.skin = 0.0;
    F.nl.r_list = F.nl.r_cut;
    F.set.setup(S, off, lattices, nullptr, F.nl.r_list, st);
    nl_build(F.nl, F.set, F.pos.as<double>(), st);
  }
  AK_REQUIRE(F.nl.stride <= 64, AK_ERR_CAPACITY, "more than 64 neighbours per atom");

  // forward
  F.F.ensure(sizeof(double) * 3 * n);
  F.es.ensure(sizeof(double) * 10 * S);
  compute_forces(pot, F.nl, F.set, F.pos.as<double>(), F.spc.as<int>(), F.F.as<double>(),
                 F.es.as<double>(), st);
  // loss and adjoints
  const double wE = weights[0], wF = weights[1] * weights[3] * weights[3], wS = weights[2],
               eps0 = weights[4];
  F.lam.ensure(sizeof(double) * 3 * n);
  F.lf.ensure(sizeof(double) * n);
  F.loss_s.ensure(sizeof(double) * S);
  F.kappa.ensure(sizeof(double) * S);
  F.muv.ensure(sizeof(double) * 9 * S);
  hipLaunchKernelGGL(k_fit_atoms, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     F.set.d_sid.as<int>(), F.set.d_geoms.as<Geom>(),
                     (const double*)F.F.as<double>(), (const double*)F.fref.as<double>(), wF,
                     F.lam.as<double>(), F.lf.as<double>());
  hipLaunchKernelGGL(k_fit_structs, dim3(S), dim3(256), 0, st, F.set.d_geoms.as<Geom>(),
                     (const double*)F.es.as<double>(), (const double*)F.ref.as<double>(),
                     (const double*)F.lf.as<double>(), wE, wS, eps0, F.loss_s.as<double>(),
                     F.kappa.as<double>(), F.muv.as<double>());
  std::vector<double> ls(S), gs;
  if (grad) {
    F.dE.ensure(sizeof(double) * 14 * (size_t)n);
    F.dD.ensure(sizeof(double) * 14 * (size_t)n);
    F.grad_s.ensure(sizeof(double) * 14 * (size_t)S);
    switch (F.nl.stride) {
      case 4: launch_hd<4>(F, n, st); break;
      case 8: launch_hd<8>(F, n, st); break;
      case 16: launch_hd<16>(F, n, st); break;
      case 32: launch_hd<32>(F, n, st); break;
      case 64: launch_hd<64>(F, n, st); break;
      default: throw Error(AK_ERR_STATE, "unsupported neighbour stride");
    }
    hipLaunchKernelGGL(k_fit_reduce, dim3(S, 14), dim3(256), 0, st, F.set.d_geoms.as<Geom>(),
                       (const double*)F.dE.as<double>(), (const double*)F.dD.as<double>(),
                       (const double*)F.kappa.as<double>(), n, F.grad_s.as<double>());
    gs.resize(14 * (size_t)S);
    AK_HIP(hipMemcpyAsync(gs.data(), F.grad_s.p, sizeof(double) * gs.size(),
                          hipMemcpyDeviceToHost, st));
  }
  AK_HIP(hipMemcpyAsync(ls.data(), F.loss_s.p, sizeof(double) * S, hipMemcpyDeviceToHost, st));
  std::vector<double> es;
  if (energies || stresses) {
    es.resize(10 * (size_t)S);
    AK_HIP(hipMemcpyAsync(es.data(), F.es.p, sizeof(double) * es.size(), hipMemcpyDeviceToHost,
                          st));
  }
  if (forces)
    AK_HIP(hipMemcpyAsync(forces, F.F.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
  AK_HIP(hipStreamSynchronize(st));
  double L = 0.0;
  for (int s = 0; s < S; ++s) L += ls[s];
  *loss = L;
  if (grad) {
    for (int q = 0; q < 14; ++q) grad[q] = 0.0;
    for (int s = 0; s < S; ++s)
      for (int q = 0; q < 14; ++q) grad[q] += gs[14 * (size_t)s + q];
  }
  for (int s = 0; s < S; ++s) {
    if (energies) energies[s] = es[10 * s];
    if (stresses) std::memcpy(stresses + 9 * s, es.data() + 10 * s + 1, 9 * sizeof(double));
  }
}

}  // namespace ak
