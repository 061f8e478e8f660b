// Many-body force kernels: Tersoff, Stillinger-Weber, Zhou EAM.
//
// Tersoff and SW use a site formulation.  The energy is split into site
// energies E_a = E_a({r_aj}) that depend only on the bond vectors of atom a to
// its neighbours (for Tersoff E_a = 1/2 sum_j fc [fR + b_aj fA]; the pair sum
// of the reference, engine/tersoff.py:1399-1414, with b_avg = (b_ij + b_ji)/2,
// is the same total).  One segment of L lanes handles one atom, one lane per
// neighbour slot; the lanes exchange bond vectors by wavefront shuffles.  Each
// lane writes G_aj = dE_a/dr_aj for its bond, and a second pass gathers
//     F_a = sum_j (G_aj - G_ja),      sigma = -(1/V) sum_bonds G_aj (x) r_aj,
// reading G_ja through the reverse-slot index of the neighbour list.  No
// atomics, so results are bitwise reproducible run to run.
//
// EAM runs two per-atom passes over the full neighbour rows: density and
// embedding derivative, then forces (after the reference's numba kernel,
// engine/_eam_kernels.py:231-307).
#include <hip/hip_ext.h>

#include "engine.hpp"

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>

namespace ak {

// ---------------------------------------------------------------------------
// Scalar helpers
// ---------------------------------------------------------------------------

// Tersoff cutoff and its derivative (engine/tersoff.py:48-58, 127-135).
__device__ __forceinline__ void te_fc(double r, double R, double D, double& fc, double& dfc) {
  if (r < R - D) {
    fc = 1.0;
    dfc = 0.0;
  } else if (r > R + D) {
    fc = 0.0;
    dfc = 0.0;
  } else {
    double s, c;
    sincos(0.5 * M_PI * (r - R) / D, &s, &c);
    fc = 0.5 - 0.5 * s;
    dfc = -0.25 * M_PI / D * c;
  }
}

// exp(l3^m du^m) and its derivative in du (the zeta-term factor of
// engine/tersoff.py:161-202); m = 3 and m = 1 are multiplied out, other m go
// through pow out of line (rare, and kept out of the unrolled partner loops).
__device__ __attribute__((noinline)) void te_exp_general(double l3, double m, double du,
                                                         double& e, double& de) {
  double lm = pow(l3, m);
  e = nb_exp(lm * pow(du, m));
  de = e * m * lm * pow(du, m - 1.0);
}

__device__ __forceinline__ void te_exp(double l3, double m, double du, double& e, double& de) {
  if (m == 3.0) {
    if (l3 == 0.0) {
      e = 1.0;
      de = 0.0;
      return;
    }
    double lm = l3 * l3 * l3, u2 = du * du;
    e = nb_exp(lm * u2 * du);
    de = e * 3.0 * lm * u2;
  } else if (m == 1.0) {
    e = nb_exp(l3 * du);
    de = e * l3;
  } else {
    te_exp_general(l3, m, du, e, de);
  }
}

// pow for a base that may be non-positive (reference semantics), out of line
__device__ __attribute__((noinline)) double pow_any(double x, double y) { return pow(x, y); }

// x^e with small integer exponents multiplied out.
__device__ __forceinline__ double xpow(double x, double e) {
  if (e == 0.0) return 1.0;
  if (e == 4.0) {
    double x2 = x * x;
    return x2 * x2;
  }
  if (e == -4.0) {
    double x2 = x * x;
    return 1.0 / (x2 * x2);
  }
  if (e == -5.0) {
    double x2 = x * x;
    return 1.0 / (x2 * x2 * x);
  }
  if (e == -1.0) return 1.0 / x;
  if (e == 1.0) return x;
  return pow(x, e);
}

// x^e for the exponents of the Stillinger-Weber pair term, given xi = 1 / x:
// the negative integer powers are products of xi, other exponents pow().
__device__ __forceinline__ double xpow_r(double x, double xi, double e) {
  if (e == 0.0) return 1.0;
  if (e == -4.0) {
    const double q = xi * xi;
    return q * q;
  }
  if (e == -5.0) {
    const double q = xi * xi;
    return q * q * xi;
  }
  if (e == -1.0) return xi;
  if (e == 4.0) {
    const double q = x * x;
    return q * q;
  }
  if (e == 1.0) return x;
  return pow(x, e);
}

// Stillinger-Weber two-body term of one leg (x = r / sigma < a), with
// rinv = 1 / r: e2 = V2 / 2, su = (dV2/dr) / (2 r), dlnE = d ln E / dr for
// E = exp(gamma / (x - a)), where
//   V2 = eps A (B x^-p - x^-q) exp(1 / (x - a)).
// Two divisions (x and 1 / (x - a)); the powers of x use 1 / x = sigma / r.
__device__ __forceinline__ void sw_leg(double r, double rinv, double eps, double sig,
                                       double inv_sig, double Asw, double Bsw, double pp,
                                       double qq, double asw, double gam, double& e2, double& su,
                                       double& dlnE) {
  const double x = r * inv_sig;
  const double xi = sig * rinv;
  const double t = fast_rcp(x - asw);
  dlnE = -gam * inv_sig * (t * t);
  const double ex = nb_exp(t);
  const double bx = Bsw * xpow_r(x, xi, -pp) - xpow_r(x, xi, -qq);
  const double dV2 = eps * Asw * ex *
                     ((-pp * Bsw * xpow_r(x, xi, -pp - 1.0) + qq * xpow_r(x, xi, -qq - 1.0)) +
                      bx * (-t * t));
  e2 = 0.5 * (eps * Asw * bx * ex);
  su = 0.5 * (dV2 * inv_sig) * rinv;
}

// Block sums of 10 values (energy, virial) of launch chunk c: wavefront
// totals by recursive halving (wave_sums), one LDS exchange of them,
// combined in fixed order;
// thread q < 10 writes value q to partial[q * n_chunks + c].  A chunk is the
// atoms of one block (of one structure).
__device__ __forceinline__ void wave_partial(double (&v)[10], double* __restrict__ partial, int c,
                                             int n_chunks) {
  __shared__ double sm[10][4];
  int idx;
  const double t = wave_sums<10>(v, idx);
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) < 16 && idx < 10) sm[idx][wid] = t;
  __syncthreads();
  if (threadIdx.x < 10) {
    const int q = threadIdx.x;
    double t = sm[q][0];
    for (int w = 1; w < nw; ++w) t += sm[q][w];
    partial[(size_t)q * n_chunks + c] = t;
  }
  __syncthreads();  // sm is written again by the block's next chunk
}

// Bond energies and virials (e, -G (x) u) of the lanes.
__device__ __forceinline__ void block_partial(double e, double Gx, double Gy, double Gz,
                                              double ux, double uy, double uz,
                                              double* __restrict__ partial, int c,
                                              int n_chunks) {
  double v[10] = {e,        -Gx * ux, -Gx * uy, -Gx * uz, -Gy * ux,
                  -Gy * uy, -Gy * uz, -Gz * ux, -Gz * uy, -Gz * uz};
  wave_partial(v, partial, c, n_chunks);
}

// Active legs of an atom's segment of L lanes (L >= 8): the segment's
// ballot bits, this lane's rank among the active ones (slot order), their
// count, and the largest count over the wavefront's segments (the bound of
// a wave-uniform partner loop).
template <int L>
struct ActiveLegs {
  uint64_t mask;
  int rank, count, wave_max;
  __device__ __forceinline__ ActiveLegs(bool active, int k, int seg) {
    const uint64_t bal = __ballot(active);
    mask = L == 64 ? bal : (bal >> seg) & ((uint64_t(1) << (L & 63)) - 1);
    rank = __popcll(mask & ((uint64_t(1) << k) - 1));
    count = __popcll(mask);
    int m = count;
#pragma unroll
    for (int o = L; o < 64; o <<= 1) m = max(m, __shfl_xor(m, o, kWave));
    wave_max = m;
  }
};

// ---------------------------------------------------------------------------
// Tersoff site kernel
//   tb : (ns, ns, 9) [valid, A, B, lambda1, lambda2, R, D, beta, n]
//   tri: (ns, ns, ns, 7) [valid, c, d, h, lambda3, m, gamma]
// Pair terms use the record of (species of the lower-indexed atom, other),
// bond orders the record of (centre, neighbour), angular terms the triple
// (centre, bond neighbour, third atom) with the leg cutoff (centre, third):
// the table semantics of engine/tersoff.py:528-855 (the flat single-species
// form is the ns = 1 case with m = 3, gamma = 1).
// ---------------------------------------------------------------------------
template <int L, bool SP1>
__device__ __forceinline__ void k_tersoff_site_chunk(int c, int n_chunks, 
    int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns_arg,
    const double* __restrict__ tb, const double* __restrict__ tri, double rc,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const int* __restrict__ chunk_rng, double* __restrict__ G, double* __restrict__ partial,
    uint64_t* __restrict__ amask, double* __restrict__ esite) {
  // one species: every table row is the first, read through uniform (scalar)
  // loads, and the leg cutoff equals the pair cutoff
  const int ns = SP1 ? 1 : ns_arg;
  const int a = chunk_rng[2 * c] + (int)threadIdx.x / L;
  const int k = threadIdx.x & (L - 1);
  const int seg = (threadIdx.x & 63) & ~(L - 1);
  const bool live = a < chunk_rng[2 * c + 1];
  const int ac = live ? a : 0;
  const int nk_row = cnt[ac];  // loaded with the neighbour index
  const int nk = live ? nk_row : 0;
  const bool has = live && k < nk;
  const size_t slot = (size_t)ac * stride + k;
  const int jr = nbr[slot];  // used when has
  const int j = has ? jr : ac;
  const int sa = ns > 1 ? spc[ac] : 0;
  const int sj = ns > 1 ? spc[j] : 0;

  double ux = 0.0, uy = 0.0, uz = 0.0, r = 0.0, rinv = 0.0;
  bool act = false;
  if (has) {
    const Geom& g = geoms[single ? 0 : sid[chunk_rng[2 * c]]];
    mic_fast(g, pos, ac, j, ux, uy, uz);
    r = sqrt(ux * ux + uy * uy + uz * uz);
    act = r < rc && r > 1e-12;
    rinv = act ? 1.0 / r : 0.0;
  }
  const int lo = ac < j ? sa : sj, hi = ac < j ? sj : sa;
  const double* P2 = tb + ((size_t)lo * ns + hi) * 9;
  const double* PL = tb + ((size_t)sa * ns + sj) * 9;
  const bool pair_ok = act && P2[0] != 0.0;
  // cutoff of this bond as the third leg of an angle centred on a
  double fcl = 0.0, dfcl = 0.0;
  if (act && PL[0] != 0.0) te_fc(r, PL[5], PL[6], fcl, dfcl);

  // Per partner leg s: cos of the angle, and g, g', the exponential factor
  // of the triple (a; this bond, s).  Kept in registers for the gradient
  // pass when the segment is short; recomputed otherwise.
  constexpr bool kCache = L == 4;
  constexpr int NC = kCache ? L : 1;
  double c_cs[NC], c_g[NC], c_dg[NC], c_e[NC], c_de[NC];
  double c_wx[NC], c_wy[NC], c_wz[NC], c_wr[NC], c_wri[NC], c_wfc[NC];
  int c_wsp[NC];

  // Rows of 8 or more slots (a skin, or liquid-like cells): the active legs
  // are staged in LDS in slot order and both partner loops run over them
  // only -- the same terms in the same order as a loop over every slot that
  // skips the inactive ones.
  const ActiveLegs<L> al(act, k, seg);
  __shared__ double t_x[256], t_y[256], t_z[256], t_r[256], t_ri[256], t_fc[256], t_cw[256];
  __shared__ int t_sp[256];
  const int lbase = (int)threadIdx.x & ~(L - 1);

  // zeta of this bond: sum over the other legs of the segment
  double zeta = 0.0;
  if constexpr (L == 4) {
  #pragma unroll
    for (int s = 0; s < L; ++s) {
      double wx = seg_get<L>(ux, seg, s), wy = seg_get<L>(uy, seg, s), wz = seg_get<L>(uz, seg, s);
      double wr = seg_get<L>(r, seg, s), wri = seg_get<L>(rinv, seg, s);
      double wfc = seg_get<L>(fcl, seg, s);
      int wsp = ns > 1 ? seg_get<L>(sj, seg, s) : 0;
      double cs = 0.0, g = 0.0, dg = 0.0, ex = 0.0, dex = 0.0;
      const bool geo = s != k && act && wri != 0.0;
      if (geo) {
        cs = (ux * wx + uy * wy + uz * wz) * rinv * wri;
        const double* T = tri + (((size_t)sa * ns + sj) * ns + wsp) * 7;
        if (T[0] != 0.0) {
          double c = T[1], d = T[2], hc = T[3] - cs;
          double c2 = c * c, d2 = d * d;
          double q = d2 + hc * hc;
          double iq = 1.0 / q;
          g = 1.0 + c2 / d2 - c2 * iq;
          dg = -2.0 * c2 * hc * iq * iq;
          te_exp(T[4], T[5], r - wr, ex, dex);
          if (pair_ok && wfc != 0.0) zeta += T[6] * wfc * g * ex;
        }
      }
      if constexpr (kCache) {
        c_cs[s] = cs;
        c_g[s] = g;
        c_dg[s] = dg;
        c_e[s] = ex;
        c_de[s] = dex;
        c_wx[s] = wx;
        c_wy[s] = wy;
        c_wz[s] = wz;
        c_wr[s] = wr;
        c_wri[s] = wri;
        c_wfc[s] = wfc;
        c_wsp[s] = wsp;
      }
    }
  } else {
    if (act) {
      const int q = lbase + al.rank;
      t_x[q] = ux;
      t_y[q] = uy;
      t_z[q] = uz;
      t_r[q] = r;
      t_ri[q] = rinv;
      t_fc[q] = fcl;
      t_sp[q] = sj;
    }
    __syncthreads();
    for (int t = 0; t < al.wave_max; ++t) {
      if (!act || t >= al.count || t == al.rank) continue;
      const int q = lbase + t;
      const double wx = t_x[q], wy = t_y[q], wz = t_z[q], wr = t_r[q], wri = t_ri[q];
      const double wfc = t_fc[q];
      const int wsp = ns > 1 ? t_sp[q] : 0;
      const double cs = (ux * wx + uy * wy + uz * wz) * rinv * wri;
      const double* T = tri + (((size_t)sa * ns + sj) * ns + wsp) * 7;
      if (T[0] != 0.0) {
        double c = T[1], d = T[2], hc = T[3] - cs;
        double c2 = c * c, d2 = d * d;
        double qq = d2 + hc * hc;
        double iq = 1.0 / qq;
        double g = 1.0 + c2 / d2 - c2 * iq;
        double ex, dex;
        te_exp(T[4], T[5], r - wr, ex, dex);
        if (pair_ok && wfc != 0.0) zeta += T[6] * wfc * g * ex;
      }
    }
  }

  // bond order (beta, n of the record (a, j)) and the two-body term;
  // (beta zeta)^n and (1 + bz)^(-1/2n) through exp/log of positive bases
  double b = 1.0, db = 0.0;
  if (pair_ok && zeta > 0.0 && PL[7] != 0.0) {
    const double beta = PL[7], nn = PL[8];
    double x = beta * zeta;
    double bz = x > 0.0 ? exp(nn * log(x)) : pow_any(x, nn);
    b = exp((-0.5 / nn) * log1p(bz));
    db = -0.5 * b * bz / (zeta * (1.0 + bz));
  }
  double e = 0.0, dVdr = 0.0, cown = 0.0;
  if (pair_ok) {
    double fc2 = fcl, dfc2 = dfcl;
    if (!SP1) te_fc(r, P2[5], P2[6], fc2, dfc2);
    if (fc2 != 0.0 || dfc2 != 0.0) {
      double fr = P2[1] * nb_exp(-P2[3] * r);
      double fa = -P2[2] * nb_exp(-P2[4] * r);
      e = 0.5 * fc2 * (fr + b * fa);
      dVdr = 0.5 * (dfc2 * (fr + b * fa) + fc2 * (-P2[3] * fr + b * (-P2[4] * fa)));
      if (fabs(db) > 1e-30) cown = 0.5 * fc2 * fa * db;
    }
  }

  // gradient of E_a with respect to this bond vector u:
  //   own zeta (this bond's b depends on the angle to every other leg w) and
  //   every other bond's zeta in which u is the third leg.
  double su = dVdr * rinv, swx = 0.0, swy = 0.0, swz = 0.0;
  if constexpr (L == 4) {
  #pragma unroll
    for (int s = 0; s < L; ++s) {
      const double wc = seg_get<L>(cown, seg, s);
      double wx, wy, wz, wr, wri, wfc, cs, gA, dgA, eA, deA;
      int wsp;
      if constexpr (kCache) {
        wx = c_wx[s];
        wy = c_wy[s];
        wz = c_wz[s];
        wr = c_wr[s];
        wri = c_wri[s];
        wfc = c_wfc[s];
        wsp = c_wsp[s];
        cs = c_cs[s];
        gA = c_g[s];
        dgA = c_dg[s];
        eA = c_e[s];
        deA = c_de[s];
      } else {
        wx = seg_get<L>(ux, seg, s);
        wy = seg_get<L>(uy, seg, s);
        wz = seg_get<L>(uz, seg, s);
        wr = seg_get<L>(r, seg, s);
        wri = seg_get<L>(rinv, seg, s);
        wfc = seg_get<L>(fcl, seg, s);
        wsp = ns > 1 ? seg_get<L>(sj, seg, s) : 0;
        cs = gA = dgA = eA = deA = 0.0;
        if (s != k && act && wri != 0.0) {
          cs = (ux * wx + uy * wy + uz * wz) * rinv * wri;
          const double* T = tri + (((size_t)sa * ns + sj) * ns + wsp) * 7;
          if (T[0] != 0.0) {
            double c = T[1], d = T[2], hc = T[3] - cs;
            double c2 = c * c, q = d * d + hc * hc, iq = 1.0 / q;
            gA = 1.0 + c2 / (d * d) - c2 * iq;
            dgA = -2.0 * c2 * hc * iq * iq;
            te_exp(T[4], T[5], r - wr, eA, deA);
          }
        }
      }
      if (s == k || !act || wri == 0.0) continue;
      const double* TA = tri + (((size_t)sa * ns + sj) * ns + wsp) * 7;
      const double* TB = tri + (((size_t)sa * ns + wsp) * ns + sj) * 7;
      double alpha = 0.0, beta = 0.0;
      if (cown != 0.0 && wfc != 0.0 && TA[0] != 0.0) {
        alpha += cown * TA[6] * wfc * dgA * eA;
        beta += cown * TA[6] * wfc * gA * deA;
      }
      if (wc != 0.0 && fcl != 0.0 && TB[0] != 0.0) {
        double gB, dgB;
        if (TB == TA) {
          gB = gA;
          dgB = dgA;
        } else {
          double c = TB[1], d = TB[2], hc = TB[3] - cs;
          double c2 = c * c, q = d * d + hc * hc, iq = 1.0 / q;
          gB = 1.0 + c2 / (d * d) - c2 * iq;
          dgB = -2.0 * c2 * hc * iq * iq;
        }
        double eB, deB;
        te_exp(TB[4], TB[5], wr - r, eB, deB);
        alpha += wc * TB[6] * fcl * dgB * eB;
        beta += wc * TB[6] * (dfcl * gB * eB - fcl * gB * deB);
      }
      // d cos / du = w/(|u||w|) - cos u/|u|^2
      su += beta * rinv - alpha * cs * rinv * rinv;
      double aw = alpha * rinv * wri;
      swx += aw * wx;
      swy += aw * wy;
      swz += aw * wz;
    }
  } else {
    if (act) t_cw[lbase + al.rank] = cown;
    __syncthreads();
    for (int t = 0; t < al.wave_max; ++t) {
      if (!act || t >= al.count || t == al.rank) continue;
      const int q = lbase + t;
      const double wx = t_x[q], wy = t_y[q], wz = t_z[q], wr = t_r[q], wri = t_ri[q];
      const double wfc = t_fc[q], wc = t_cw[q];
      const int wsp = ns > 1 ? t_sp[q] : 0;
      const double cs = (ux * wx + uy * wy + uz * wz) * rinv * wri;
      const double* TA = tri + (((size_t)sa * ns + sj) * ns + wsp) * 7;
      const double* TB = tri + (((size_t)sa * ns + wsp) * ns + sj) * 7;
      double gA = 0.0, dgA = 0.0, eA = 0.0, deA = 0.0;
      if (TA[0] != 0.0) {
        double c = TA[1], d = TA[2], hc = TA[3] - cs;
        double c2 = c * c, qq = d * d + hc * hc, iq = 1.0 / qq;
        gA = 1.0 + c2 / (d * d) - c2 * iq;
        dgA = -2.0 * c2 * hc * iq * iq;
        te_exp(TA[4], TA[5], r - wr, eA, deA);
      }
      double alpha = 0.0, beta = 0.0;
      if (cown != 0.0 && wfc != 0.0 && TA[0] != 0.0) {
        alpha += cown * TA[6] * wfc * dgA * eA;
        beta += cown * TA[6] * wfc * gA * deA;
      }
      if (wc != 0.0 && fcl != 0.0 && TB[0] != 0.0) {
        double gB, dgB;
        if (TB == TA) {
          gB = gA;
          dgB = dgA;
        } else {
          double c = TB[1], d = TB[2], hc = TB[3] - cs;
          double c2 = c * c, qq = d * d + hc * hc, iq = 1.0 / qq;
          gB = 1.0 + c2 / (d * d) - c2 * iq;
          dgB = -2.0 * c2 * hc * iq * iq;
        }
        double eB, deB;
        te_exp(TB[4], TB[5], wr - r, eB, deB);
        alpha += wc * TB[6] * fcl * dgB * eB;
        beta += wc * TB[6] * (dfcl * gB * eB - fcl * gB * deB);
      }
      su += beta * rinv - alpha * cs * rinv * rinv;
      double aw = alpha * rinv * wri;
      swx += aw * wx;
      swy += aw * wy;
      swz += aw * wz;
    }
  }
  const double Gx = su * ux + swx, Gy = su * uy + swy, Gz = su * uz + swz;
  // bonds outside r_c have G = 0 exactly; with rows of 8 or more slots only
  // the active ones are written and the gather skips the rest
  if constexpr (L >= 8) {
    if (live && k == 0) amask[ac] = al.mask;
  }
  if (has && (L == 4 || act)) {
    g_put(G, slot, Gx, Gy, Gz);
  }
  if (esite) {  // per-atom site energy U_a (uniform branch: every lane joins the sum)
    const double ea = seg_sum<L>(e);
    if (live && k == 0) esite[a] = ea;
  }
  block_partial(e, Gx, Gy, Gz, ux, uy, uz, partial, c, n_chunks);
}

template <int L, bool SP1>
__global__ __launch_bounds__(256) void k_tersoff_site(int n_chunks, int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns_arg,
    const double* __restrict__ tb, const double* __restrict__ tri, double rc,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const int* __restrict__ chunk_rng, double* __restrict__ G, double* __restrict__ partial,
    uint64_t* __restrict__ amask, double* __restrict__ esite) {
  k_tersoff_site_chunk<L, SP1>(xcd_remap((int)blockIdx.x, n_chunks), n_chunks, n, pos, sid, geoms, single, spc, ns_arg, tb, tri, rc, cnt,
                                nbr, stride, chunk_rng, G, partial, amask, esite);
}

// ---------------------------------------------------------------------------
// Wavefront totals of 8 values per lane by recursive halving: three exchange
// steps inside each group of 8 lanes (row half-mirror, quad xor 1, quad
// xor 2), each keeping half of a lane's values and adding the partner
// lane's copy of that half, then the groups summed (row rotation by 8, two
// cross-row shuffles).  Lane l < 8 holds the total of value
// idx = 4 b2 + 2 b0 + b1 (b the bits of l).  Fixed order.
__device__ __forceinline__ double wave_sums8(const double (&v)[8], int& idx) {
  const int lane = __lane_id();
  double w[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = v[i];
  auto step = [&](auto partner, int H, bool upper) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i < H) {
        const double keep = upper ? w[H + i] : w[i];
        const double send = upper ? w[i] : w[H + i];
        w[i] = keep + partner(send);
      }
    }
  };
  step([](double x) { return dpp_row<0x141>(x); }, 4, (lane & 4) != 0);
  step([](double x) { return dpp_quad<0xB1>(x); }, 2, (lane & 1) != 0);
  step([](double x) { return dpp_quad<0x4E>(x); }, 1, (lane & 2) != 0);
  double t = w[0];
  t += dpp_row<0x128>(t);
  t += __shfl_xor(t, 16, kWave);
  t += __shfl_xor(t, 32, kWave);
  idx = ((lane & 4)) | ((lane & 1) << 1) | ((lane >> 1) & 1);
  return t;
}

// Block sums of the energy and a symmetric virial (e, xx, xy, xz, yy, yz, zz;
// v[7] unused) of launch chunk c, written as the ten values (e, xx, xy, xz,
// yx, yy, yz, zx, zy, zz) the structure reduction reads: wavefront totals
// by wave_sums8, one LDS exchange, combined in fixed order.  Every thread of
// the block calls it once (one chunk per block).
__device__ __forceinline__ void block_partial7(const double (&v)[8], double* __restrict__ partial,
                                               int c, int n_chunks) {
  __shared__ double sm[8][8];  // up to 8 wavefronts
  int idx;
  const double t = wave_sums8(v, idx);
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) < 8) sm[idx][wid] = t;
  __syncthreads();
  if (threadIdx.x < 10) {
    // value q of the ten from the seven sums: 4-bit fields of 0x6535423210,
    // field q = source of value q
    const int q = (int)((0x6535423210ull >> (4 * threadIdx.x)) & 0xF);
    double s = sm[q][0];
    for (int w = 1; w < nw; ++w) s += sm[q][w];
    partial[(size_t)threadIdx.x * n_chunks + c] = s;
  }
}

// Tersoff site kernel for the flat one-species form with rows of four slots
// (diamond-like solids, list skin below the second shell): m = 3, gamma = 1
// as the reference's flat path hard-codes them (engine/tersoff.py:86, 245,
// 494, 1431).  One DPP quad per atom, lane k = bond k, one chunk of
// blockDim / 4 atoms per block.  The partner legs arrive by quad rotations
// (lane k reads lane k+P mod 4, P = 1, 2, 3), so every lane evaluates the
// same three partner terms with no self slot and no out-of-line calls.
// g(theta), g'(theta) and the two exponential factors of each (bond,
// partner) pair are formed once: the factor of the partner's zeta in which
// this bond is the third leg is the partner lane's own factor for this bond,
// brought over by the inverse rotation.  Terms and gating are those of
// k_tersoff_site; the summation order of the zeta sums and the bond
// gradients differs, and 1 / r, |r| and the three 1 / (d^2 + (h - cos)^2)
// come from the hardware estimates refined by Newton steps.
//
// Chunk c is atoms [chunk_rng[2c], chunk_rng[2c + 1]) of one structure, so
// the cell geometry is uniform over the block (scalar loads); with a single
// structure the chunk is c * blockDim / 4 onwards and chunk_rng is not read.
// Row length and neighbour index are loaded together.  Block partials: the
// energy and the six components of the symmetrised virial (the site energy
// of each atom is rotation invariant, so sum_k u_k (x) G_k is symmetric per
// atom), written as the ten values the structure reduction reads.
// ---------------------------------------------------------------------------

// Lane k of every quad receives lane (k + P) mod 4 of its quad.
template <int P>
__device__ __forceinline__ double qrot(double x) {
  static_assert(P >= 1 && P <= 3, "rotation of a quad");
  constexpr int ctrl = P == 1 ? 0x39 : (P == 2 ? 0x4E : 0x93);
  return dpp_quad<ctrl>(x);
}

template <bool L3Z>
__device__ __forceinline__ void q4_chunk(
    int n_chunks, int apb, int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const double* __restrict__ tb,
    const double* __restrict__ tri, double rc, const int* __restrict__ cnt,
    const int* __restrict__ nbr, const int* __restrict__ chunk_rng, double* __restrict__ G,
    double* __restrict__ partial, double* __restrict__ esite) {
  const int c = xcd_remap((int)blockIdx.x, n_chunks);
  const int a0 = single ? c * apb : chunk_rng[2 * c];
  const int a1 = single ? min(a0 + apb, n) : chunk_rng[2 * c + 1];
  const Geom& g = geoms[single ? 0 : sid[a0]];
  const int a = a0 + (int)threadIdx.x / 4;
  const int k = threadIdx.x & 3;
  const bool live = a < a1;
  const int ac = live ? a : a0;
  const size_t slot = (size_t)ac * 4 + k;
  const int nk = cnt[ac];
  const int jr = nbr[slot];
  const bool has = live && k < nk;
  const int j = has ? jr : ac;
  // one record: tb = [valid, A, B, lambda1, lambda2, R, D, beta, n],
  // tri = [valid, c, d, h, lambda3, m, gamma]
  const double A = tb[1], B = tb[2], l1 = tb[3], l2 = tb[4], R = tb[5], D = tb[6];
  const double beta = tb[7], nn = tb[8];
  const double cc = tri[1], dd = tri[2], hh = tri[3], l3 = tri[4];
  const double c2 = cc * cc, d2 = dd * dd, g0 = 1.0 + c2 * fast_rcp(d2), l3c = l3 * l3 * l3;

  double ux, uy, uz;
  mic_fast(g, pos, ac, j, ux, uy, uz);
  const double r2 = ux * ux + uy * uy + uz * uz;
  const bool act = has && r2 > 1e-24 && r2 < rc * rc;
  const double rinv = act ? fast_rsq(r2) : 0.0;
  const double r = r2 * rinv;
  // unit vector (zero for slots outside r_c or past the row)
  const double ex = ux * rinv, ey = uy * rinv, ez = uz * rinv;
  double fc = 0.0, dfc = 0.0;
  if (act) te_fc(r, R, D, fc, dfc);

  // zeta of this bond over the three partner legs; per partner P: cos,
  // g, g' and exp(l3^3 (r - r_P)^3) with its derivative
  double cs[4], gv[4], dg[4], xe[4], xd[4];
  double zeta = 0.0;
#define AK_Q4_PARTNER(P)                                                  \
  {                                                                       \
    const double wx = qrot<P>(ex), wy = qrot<P>(ey), wz = qrot<P>(ez);    \
    const double wfc = qrot<P>(fc);                                       \
    cs[P] = ex * wx + ey * wy + ez * wz;                                  \
    const double hc = hh - cs[P];                                         \
    const double iq = fast_rcp(d2 + hc * hc);                             \
    gv[P] = g0 - c2 * iq;                                                 \
    dg[P] = -2.0 * c2 * hc * iq * iq;                                     \
    if constexpr (L3Z) {                                                  \
      xe[P] = 1.0;                                                        \
      xd[P] = 0.0;                                                        \
      zeta += wfc * gv[P];                                                \
    } else {                                                              \
      const double du = r - qrot<P>(r), u2 = du * du;                     \
      xe[P] = nb_exp(l3c * u2 * du);                                      \
      xd[P] = xe[P] * 3.0 * l3c * u2;                                     \
      zeta += wfc * gv[P] * xe[P];                                        \
    }                                                                     \
  }
  AK_Q4_PARTNER(1)
  AK_Q4_PARTNER(2)
  AK_Q4_PARTNER(3)
#undef AK_Q4_PARTNER

  // bond order and the two-body term (engine/tersoff.py:1380-1414); with
  // bz = (beta zeta)^n, b = (1 + bz)^(-1 / 2n): log(1 + bz) in place of
  // log1p(bz) changes b by less than an ulp of 1
  double b = 1.0, db = 0.0;
  if (act && zeta > 0.0) {
    const double bz = exp(nn * log(beta * zeta));
    b = exp(-0.5 * fast_rcp(nn) * log(1.0 + bz));
    db = -0.5 * b * bz * fast_rcp(zeta * (1.0 + bz));
  }
  double e = 0.0, dVdr = 0.0, cown = 0.0;
  if (act) {
    const double fr = A * nb_exp(-l1 * r);
    const double fa = -B * nb_exp(-l2 * r);
    e = 0.5 * fc * (fr + b * fa);
    dVdr = 0.5 * (dfc * (fr + b * fa) + fc * (-l1 * fr + b * (-l2 * fa)));
    if (fabs(db) > 1e-30) cown = 0.5 * fc * fa * db;
  }

  // gradient of E_a in this bond vector: the own bond order (angles to the
  // partners) and each partner's bond order, in which this bond is the
  // third leg (its cutoff, angle and exponential factor)
  double sr = dVdr, swx = 0.0, swy = 0.0, swz = 0.0;
#define AK_Q4_GRAD(P)                                                                  \
  {                                                                                    \
    const double wx = qrot<P>(ex), wy = qrot<P>(ey), wz = qrot<P>(ez);                 \
    const double wfc = qrot<P>(fc), wc = qrot<P>(cown);                                \
    double alpha, bet;                                                                 \
    if constexpr (L3Z) {                                                               \
      alpha = (cown * wfc + wc * fc) * dg[P];                                          \
      bet = wc * gv[P] * dfc;                                                          \
    } else {                                                                           \
      const double xr = qrot<P>(xe[4 - P]), xrd = qrot<P>(xd[4 - P]);                  \
      alpha = (cown * wfc * xe[P] + wc * fc * xr) * dg[P];                             \
      bet = cown * wfc * gv[P] * xd[P] + wc * gv[P] * (dfc * xr - fc * xrd);           \
    }                                                                                  \
    sr += bet - alpha * cs[P] * rinv;                                                  \
    const double ar = alpha * rinv;                                                    \
    swx += ar * wx;                                                                    \
    swy += ar * wy;                                                                    \
    swz += ar * wz;                                                                    \
  }
  AK_Q4_GRAD(1)
  AK_Q4_GRAD(2)
  AK_Q4_GRAD(3)
#undef AK_Q4_GRAD
  const double Gx = sr * ex + swx, Gy = sr * ey + swy, Gz = sr * ez + swz;
  if (has) {
    g_put(G, slot, Gx, Gy, Gz);
  }
  if (esite) {  // per-atom site energy U_a (uniform branch)
    const double ea = seg_sum<4>(e);
    if (live && k == 0) esite[a] = ea;
  }
  // energy and -sum G (x) u, symmetrised: xy = -(Gx uy + Gy ux) / 2
  double v[8] = {e,        -Gx * ux,           -0.5 * (Gx * uy + Gy * ux),
                 -0.5 * (Gx * uz + Gz * ux), -Gy * uy, -0.5 * (Gy * uz + Gz * uy),
                 -Gz * uz, 0.0};
  block_partial7(v, partial, c, n_chunks);
}

// Batched sets: chunks from the chunk table, geometry by structure.  The
// arguments the first loads need (chunk table, structure ids, geometry,
// positions, rows) come first, inside the preloaded 56 bytes.
template <bool L3Z>
__global__ __launch_bounds__(512) void k_tersoff_q4(
    int n_chunks, int n, const int* __restrict__ chunk_rng, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, const double* __restrict__ pos,
    const int* __restrict__ cnt, const int* __restrict__ nbr, const double* __restrict__ tb,
    double rc, double* __restrict__ G, double* __restrict__ partial,
    double* __restrict__ esite) {
  q4_chunk<L3Z>(n_chunks, 0, n, pos, sid, geoms, 0, tb, tb + 9, rc, cnt, nbr, chunk_rng, G,
                partial, esite);
}

// One structure: chunk c is atoms c * BLK / 4 onwards, one chunk per block.
// Fewer kernel arguments than the batched form: the host's launch cost
// grows with the bytes of arguments, and a single force call on a prebuilt
// list is paced by it.  The block size is a template parameter and the
// chunk count an argument, so that nothing is read from the dispatch's
// hidden arguments, and the arguments the first loads need come first
// (with kernel-argument preloading they arrive in SGPRs).
template <bool L3Z, int BLK>
__global__ __launch_bounds__(BLK) void k_tersoff_q4_one(
    int n, int n_chunks, const double* __restrict__ pos, const Geom* __restrict__ geoms,
    const double* __restrict__ tb, double rc, const int* __restrict__ cnt,
    const int* __restrict__ nbr, double* __restrict__ G, double* __restrict__ partial,
    double* __restrict__ esite) {
  q4_chunk<L3Z>(n_chunks, BLK / 4, n, pos, nullptr, geoms, 1, tb, tb + 9, rc, cnt, nbr, nullptr,
                G, partial, esite);
}


// ---------------------------------------------------------------------------
// Tersoff site kernel, one lane per (bond, partner) pair.
// L*L lanes serve one atom (L = list stride, L*L <= 64): lane (k, p) of the
// atom handles bond k and its partner leg s = (k + p) mod L.  Lane p = 0 of
// each bond group has no partner and evaluates the bond order and the
// two-body term of bond k instead.  zeta_k and the gradient of bond k are
// sums over the group's L lanes (DPP butterflies); the coefficient of the
// partner's bond order comes from lane (s, 0) by one shuffle.  Same
// arithmetic and gating as k_tersoff_site; only the summation order of
// zeta and of the gradient differs.  Four times more wavefronts than the
// lane-per-bond form for the same atoms, so the latency of the partner
// loop is hidden by occupancy instead of being serialised in one lane.
// ---------------------------------------------------------------------------
struct Bond {
  double x = 0.0, y = 0.0, z = 0.0, r = 0.0, rinv = 0.0, fcl = 0.0, dfcl = 0.0;
  int j = 0, sj = 0;
  bool act = false;
};

__device__ __forceinline__ Bond te_bond(const Geom& g, const double* __restrict__ pos,
                                        const int* __restrict__ spc, int ns,
                                        const double* __restrict__ tb, double rc, int a, int sa,
                                        int jr, bool has) {
  Bond b;
  b.j = has ? jr : a;
  b.sj = ns > 1 ? spc[b.j] : 0;
  if (has) {
    mic_fast(g, pos, a, b.j, b.x, b.y, b.z);
    b.r = sqrt(b.x * b.x + b.y * b.y + b.z * b.z);
    b.act = b.r < rc && b.r > 1e-12;
    b.rinv = b.act ? 1.0 / b.r : 0.0;
    const double* PL = tb + ((size_t)sa * ns + b.sj) * 9;
    if (b.act && PL[0] != 0.0) te_fc(b.r, PL[5], PL[6], b.fcl, b.dfcl);
  }
  return b;
}

template <int L, bool SP1>
__device__ __forceinline__ void k_tersoff_trip_chunk(int c, int n_chunks, 
    int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns_arg,
    const double* __restrict__ tb, const double* __restrict__ tri, double rc,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const int* __restrict__ chunk_rng, double* __restrict__ G, double* __restrict__ partial,
    uint64_t* __restrict__ amask) {
  static_assert(L * L <= 64, "one atom must fit a wavefront");
  constexpr int LL = L * L;
  const int ns = SP1 ? 1 : ns_arg;
  const int a = chunk_rng[2 * c] + (int)threadIdx.x / LL;
  const int k = ((int)threadIdx.x / L) & (L - 1);
  const int p = threadIdx.x & (L - 1);
  const int s = (k + p) & (L - 1);
  const int wl = threadIdx.x & 63;
  const int gseg = wl & ~(L - 1);    // first lane of this bond group
  const int aseg = wl & ~(LL - 1);   // first lane of this atom
  const bool live = a < chunk_rng[2 * c + 1];
  const int ac = live ? a : 0;
  const int nk_row = cnt[ac];  // loaded with the neighbour index
  const int nk = live ? nk_row : 0;
  const bool has = live && k < nk;
  const bool has_s = live && p != 0 && s < nk;
  const size_t slot = (size_t)ac * stride + k;
  const int jr = nbr[slot];
  const int jsr = nbr[(size_t)ac * stride + s];
  const int sa = ns > 1 ? spc[ac] : 0;
  const Geom& g = geoms[single ? 0 : sid[chunk_rng[2 * c]]];
  const Bond u = te_bond(g, pos, spc, ns, tb, rc, ac, sa, jr, has);
  const Bond w = te_bond(g, pos, spc, ns, tb, rc, ac, sa, jsr, has_s);

  const int lo = ac < u.j ? sa : u.sj, hi = ac < u.j ? u.sj : sa;
  const double* P2 = tb + ((size_t)lo * ns + hi) * 9;
  const double* PL = tb + ((size_t)sa * ns + u.sj) * 9;
  const bool pair_ok = u.act && P2[0] != 0.0;

  // angle (a; k, s): cos, g, g', exponential factor and its derivative
  const bool tri_ok = p != 0 && u.act && w.act;
  const double* TA = tri + (((size_t)sa * ns + u.sj) * ns + w.sj) * 7;
  double cs = 0.0, gA = 0.0, dgA = 0.0, eA = 0.0, deA = 0.0, zt = 0.0;
  if (tri_ok) {
    cs = (u.x * w.x + u.y * w.y + u.z * w.z) * u.rinv * w.rinv;
    if (TA[0] != 0.0) {
      double c = TA[1], d = TA[2], hc = TA[3] - cs;
      double c2 = c * c, d2 = d * d;
      double iq = 1.0 / (d2 + hc * hc);
      gA = 1.0 + c2 / d2 - c2 * iq;
      dgA = -2.0 * c2 * hc * iq * iq;
      te_exp(TA[4], TA[5], u.r - w.r, eA, deA);
      if (pair_ok && w.fcl != 0.0) zt = TA[6] * w.fcl * gA * eA;
    }
  }
  const double zeta = seg_sum<L>(zt);

  // bond order and two-body term of bond k, in lane p = 0
  double e = 0.0, dVdr = 0.0, cown = 0.0;
  if (p == 0 && pair_ok) {
    double b = 1.0, db = 0.0;
    if (zeta > 0.0 && PL[7] != 0.0) {
      const double beta = PL[7], nn = PL[8];
      double x = beta * zeta;
      double bz = x > 0.0 ? exp(nn * log(x)) : pow_any(x, nn);
      b = exp((-0.5 / nn) * log1p(bz));
      db = -0.5 * b * bz / (zeta * (1.0 + bz));
    }
    double fc2 = u.fcl, dfc2 = u.dfcl;
    if (!SP1) te_fc(u.r, P2[5], P2[6], fc2, dfc2);
    if (fc2 != 0.0 || dfc2 != 0.0) {
      double fr = P2[1] * nb_exp(-P2[3] * u.r);
      double fa = -P2[2] * nb_exp(-P2[4] * u.r);
      e = 0.5 * fc2 * (fr + b * fa);
      dVdr = 0.5 * (dfc2 * (fr + b * fa) + fc2 * (-P2[3] * fr + b * (-P2[4] * fa)));
      if (fabs(db) > 1e-30) cown = 0.5 * fc2 * fa * db;
    }
  }
  const double ck = seg_get<L>(cown, gseg, 0);              // bond k's coefficient
  const double cw = __shfl(cown, aseg + s * L, kWave);       // bond s's coefficient

  // gradient share of this lane: own zeta through the angle to leg s (A),
  // and the zeta of bond s in which bond k is the third leg (B)
  double su = p == 0 ? dVdr * u.rinv : 0.0, aw = 0.0;
  if (tri_ok) {
    const double* TB = tri + (((size_t)sa * ns + w.sj) * ns + u.sj) * 7;
    double alpha = 0.0, beta = 0.0;
    if (ck != 0.0 && w.fcl != 0.0 && TA[0] != 0.0) {
      alpha += ck * TA[6] * w.fcl * dgA * eA;
      beta += ck * TA[6] * w.fcl * gA * deA;
    }
    if (cw != 0.0 && u.fcl != 0.0 && TB[0] != 0.0) {
      double gB, dgB;
      if (TB == TA) {
        gB = gA;
        dgB = dgA;
      } else {
        double c = TB[1], d = TB[2], hc = TB[3] - cs;
        double c2 = c * c, iq = 1.0 / (d * d + hc * hc);
        gB = 1.0 + c2 / (d * d) - c2 * iq;
        dgB = -2.0 * c2 * hc * iq * iq;
      }
      double eB, deB;
      te_exp(TB[4], TB[5], w.r - u.r, eB, deB);
      alpha += cw * TB[6] * u.fcl * dgB * eB;
      beta += cw * TB[6] * (u.dfcl * gB * eB - u.fcl * gB * deB);
    }
    // d cos / du = w/(|u||w|) - cos u/|u|^2
    su = beta * u.rinv - alpha * cs * u.rinv * u.rinv;
    aw = alpha * u.rinv * w.rinv;
  }
  su = seg_sum<L>(su);
  const double swx = seg_sum<L>(aw * w.x), swy = seg_sum<L>(aw * w.y),
               swz = seg_sum<L>(aw * w.z);
  double Gx = 0.0, Gy = 0.0, Gz = 0.0;
  if constexpr (L >= 8) {
    // bit k of the atom's mask: bond k is inside r_c (lane (k, 0) holds it)
    const uint64_t bal = __ballot(p == 0 && u.act);
    uint64_t m = 0;
#pragma unroll
    for (int kk = 0; kk < L; ++kk) m |= ((bal >> (aseg + kk * L)) & 1) << kk;
    if (live && wl == aseg) amask[ac] = m;
  }
  if (p == 0) {
    Gx = su * u.x + swx;
    Gy = su * u.y + swy;
    Gz = su * u.z + swz;
    if (has && (L == 4 || u.act)) {
      g_put(G, slot, Gx, Gy, Gz);
    }
  }
  block_partial(e, Gx, Gy, Gz, u.x, u.y, u.z, partial, c, n_chunks);
}

template <int L, bool SP1>
__global__ __launch_bounds__(256) void k_tersoff_trip(int n_chunks, int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns_arg,
    const double* __restrict__ tb, const double* __restrict__ tri, double rc,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const int* __restrict__ chunk_rng, double* __restrict__ G, double* __restrict__ partial,
    uint64_t* __restrict__ amask) {
  k_tersoff_trip_chunk<L, SP1>(xcd_remap((int)blockIdx.x, n_chunks), n_chunks, n, pos, sid, geoms, single, spc, ns_arg, tb, tri, rc, cnt,
                                nbr, stride, chunk_rng, G, partial, amask);
}

// ---------------------------------------------------------------------------
// Stillinger-Weber site kernel (flat single-species form,
// engine/stillinger_weber.py:23-255):
//   E_a = 1/2 sum_j V2(r_aj) + 1/2 sum_{j != k} h(r_aj, r_ak)
// prm = epsilon, sigma, A, B, p, q, a, lambda, gamma, costheta0.
// ---------------------------------------------------------------------------
template <int L>
__device__ __forceinline__ void k_sw_site_chunk(int c, int n_chunks, 
    int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const double* __restrict__ prm,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const int* __restrict__ chunk_rng, double* __restrict__ G, double* __restrict__ partial,
    uint64_t* __restrict__ amask, double* __restrict__ esite) {
  const int a = chunk_rng[2 * c] + (int)threadIdx.x / L;
  const int k = threadIdx.x & (L - 1);
  const int seg = (threadIdx.x & 63) & ~(L - 1);
  const bool live = a < chunk_rng[2 * c + 1];
  const int ac = live ? a : 0;
  const int nk_row = cnt[ac];  // loaded with the neighbour index
  const int nk = live ? nk_row : 0;
  const bool has = live && k < nk;
  const size_t slot = (size_t)ac * stride + k;
  const int jr = nbr[slot];  // used when has
  const int j = has ? jr : ac;
  const double eps = prm[0], sig = prm[1], Asw = prm[2], Bsw = prm[3], pp = prm[4],
               qq = prm[5], asw = prm[6], lam = prm[7], gam = prm[8], cth0 = prm[9];
  const double rc = asw * sig;

  double ux = 0.0, uy = 0.0, uz = 0.0, r = 0.0, rinv = 0.0;
  bool act = false;
  if (has) {
    const Geom& g = geoms[single ? 0 : sid[chunk_rng[2 * c]]];
    mic_fast(g, pos, ac, j, ux, uy, uz);
    r = sqrt(ux * ux + uy * uy + uz * uz);
    act = r < rc && r > 1e-12;
  }
  double x = act ? r / sig : 0.0;
  bool leg = act && x < asw;  // leg of a three-body term
  rinv = leg ? 1.0 / r : 0.0;
  double E = 0.0, dlnE = 0.0;  // exp(gamma/(x - a)) and d ln / dr
  double e = 0.0, su = 0.0;
  if (leg) {
    E = nb_exp(gam / (x - asw));
    sw_leg(r, rinv, eps, sig, 1.0 / sig, Asw, Bsw, pp, qq, asw, gam, e, su, dlnE);
  }
  const double c3 = lam * eps;
  double swx = 0.0, swy = 0.0, swz = 0.0;
  // one three-body term h(a; this leg, leg w)
  auto add_leg = [&](double wx, double wy, double wz, double wri, double wE) {
    double cs = (ux * wx + uy * wy + uz * wz) * rinv * wri;
    double dc = cs - cth0;
    double pre = c3 * E * wE * dc;
    e += 0.5 * pre * dc;
    // dh/du = pre [dc dlnE u_hat + 2 (w_hat - cos u_hat) / r]
    double rad = pre * dc * dlnE, ang = 2.0 * pre * rinv;
    su += (rad - ang * cs) * rinv;
    double aw = ang * wri;
    swx += aw * wx;
    swy += aw * wy;
    swz += aw * wz;
  };
  if constexpr (L == 4) {
#pragma unroll
    for (int s = 0; s < L; ++s) {
      double wx = seg_get<L>(ux, seg, s), wy = seg_get<L>(uy, seg, s), wz = seg_get<L>(uz, seg, s);
      double wri = seg_get<L>(rinv, seg, s), wE = seg_get<L>(E, seg, s);
      if (s == k || !leg || wri == 0.0) continue;
      add_leg(wx, wy, wz, wri, wE);
    }
  } else {
    // Long rows (a Verlet skin, or second shells beyond r_c): the legs
    // inside r_c are compacted into LDS in slot order, and the partner loop
    // runs over them only -- the same terms in the same order as a loop over
    // every slot that skips the inactive ones.
    const ActiveLegs<L> al(leg, k, seg);
    __shared__ double s_x[256], s_y[256], s_z[256], s_ri[256], s_E[256];
    const int base = (int)threadIdx.x & ~(L - 1);
    if (leg) {
      s_x[base + al.rank] = ux;
      s_y[base + al.rank] = uy;
      s_z[base + al.rank] = uz;
      s_ri[base + al.rank] = rinv;
      s_E[base + al.rank] = E;
    }
    __syncthreads();
    for (int t = 0; t < al.wave_max; ++t) {
      if (!leg || t >= al.count || t == al.rank) continue;
      add_leg(s_x[base + t], s_y[base + t], s_z[base + t], s_ri[base + t], s_E[base + t]);
    }
    if (live && k == 0) amask[ac] = al.mask;
  }
  const double Gx = su * ux + swx, Gy = su * uy + swy, Gz = su * uz + swz;
  // inactive legs have G = 0 exactly; with long rows the gather skips them
  if (has && (L == 4 || leg)) {
    g_put(G, slot, Gx, Gy, Gz);
  }
  if (esite) {  // per-atom site energy U_a (uniform branch: every lane joins the sum)
    const double ea = seg_sum<L>(e);
    if (live && k == 0) esite[a] = ea;
  }
  block_partial(e, Gx, Gy, Gz, ux, uy, uz, partial, c, n_chunks);
}

template <int L>
__global__ __launch_bounds__(256) void k_sw_site(
    int n_chunks, int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const double* __restrict__ prm,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const int* __restrict__ chunk_rng, double* __restrict__ G, double* __restrict__ partial,
    uint64_t* __restrict__ amask, double* __restrict__ esite) {
  k_sw_site_chunk<L>(xcd_remap((int)blockIdx.x, n_chunks), n_chunks, n, pos, sid, geoms, single, prm, cnt, nbr, stride,
                       chunk_rng, G, partial, amask, esite);
}

// Per-structure energy and stress of structure s by one block of 256
// threads: the chunk partials of s in fixed order; out[s] = [E, W/V (9)].
// Per-structure energy and stress from the block partials of chunks
// [first[s], first[s + 1]): each thread sums its chunks in chunk order, U
// chunks' loads in flight at a time (a structure of many chunks is reduced
// by one block, and one load round trip per chunk was the cost), then the
// wavefront sums and the waves in order.  Any multiple of 64 threads up to
// 1024.
template <int U>
__device__ __forceinline__ void struct_reduce_block(int s, const Geom* __restrict__ geoms,
                                                    const int* __restrict__ first,
                                                    int n_chunks,
                                                    const double* __restrict__ partial,
                                                    double* __restrict__ out,
                                                    const int* __restrict__ pair_flag = nullptr,
                                                    int threads = 0) {
  __shared__ double sm[10][16];
  const int c0 = first[s], c1 = first[s + 1];
  // threads per block: blockDim, or the caller's compile-time block size
  const int nt = threads > 0 ? threads : (int)blockDim.x, nw = nt >> 6;
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int c = c0 + (int)threadIdx.x; c < c1; c += U * nt) {
    double x[U][10];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int cu = c + u * nt;
#pragma unroll
      for (int q = 0; q < 10; ++q) x[u][q] = cu < c1 ? partial[(size_t)q * n_chunks + cu] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < 10; ++q) v[q] += x[u][q];
  }
  int idx;
  const double t = wave_sums<10>(v, idx);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) < 16 && idx < 10) sm[idx][wid] = t;
  __syncthreads();
  // EAM: a structure without any bond inside r_c gives zeros, its isolated
  // atoms' embedding energy F(0) included (EAMPotential.compute,
  // engine/eam.py:236-237); pair_flag holds one flag per chunk
  int any = 0;
  if (pair_flag)
    for (int c = c0 + (int)threadIdx.x; c < c1; c += nt) any |= pair_flag[c];
  const bool none = pair_flag && !__syncthreads_or(any);
  if (threadIdx.x < 10) {
    const int q = threadIdx.x;
    double t = sm[q][0];
    for (int w = 1; w < nw; ++w) t += sm[q][w];
    out[10 * (size_t)s + q] = none ? 0.0 : (q == 0 ? t : t / geoms[s].volume);
  }
}

__global__ __launch_bounds__(1024) void k_struct_final(const Geom* __restrict__ geoms,
                                                      const int* __restrict__ first,
                                                      int n_chunks,
                                                      const double* __restrict__ partial,
                                                      double* __restrict__ out,
                                                      const int* __restrict__ pair_flag) {
  struct_reduce_block<4>(blockIdx.x, geoms, first, n_chunks, partial, out, pair_flag);
}

// F_a = sum_k G[a, k] - G[rev(a, k)].  The first gridDim - atom_blocks
// blocks (one per structure) reduce the site kernel's partials to
// per-structure energy and stress (the site kernel has completed, so any
// block may read them).
template <int L>
__device__ __forceinline__ void gather_blocks(int n, const int* __restrict__ cnt,
                                              const int* __restrict__ rev, int stride,
                                              const double* __restrict__ G,
                                              double* __restrict__ F, int atom_blocks,
                                              int n_struct, const Geom* __restrict__ geoms,
                                              const int* __restrict__ first, int n_chunks,
                                              const double* __restrict__ partial,
                                              double* __restrict__ out,
                                              const uint64_t* __restrict__ amask, int sym,
                                              int lead, int threads) {
  // the structure blocks come first, so that they are dispatched before
  // the atom blocks instead of forming the launch's tail; gather_lead pads
  // them to a whole number of XCD rounds, so that atom tile b runs on the
  // XCD whose L2 holds the site gradients of site-kernel chunk b
  if ((int)blockIdx.x < lead) {
    if ((int)blockIdx.x < n_struct)
      struct_reduce_block<2>((int)blockIdx.x, geoms, first, n_chunks, partial, out, nullptr,
                             threads);
    return;
  }
  const int gt = xcd_remap((int)blockIdx.x - lead, atom_blocks) * threads + threadIdx.x;
  const int a = gt / L;
  const int k = gt & (L - 1);
  const bool live = a < n;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (live) {
    const size_t s = (size_t)a * stride + k;
    const int c = cnt[a];
    const int r = rev[s];  // issued alongside cnt[]; used when k < c
    if (k < c && r >= 0) {
      const size_t rs = (size_t)r;
      if constexpr (L == 4) {
        double ax, ay, az, bx, by, bz;
        g_get(G, s, ax, ay, az);
        g_get(G, rs, bx, by, bz);
        fx = ax - bx;
        fy = ay - by;
        fz = az - bz;
      } else {
        // rows of stride L: slot r is entry r % L of atom r / L; only the
        // bonds inside r_c were written (the others are zero)
        const bool own = (amask[a] >> k) & 1;
        if (own) g_get(G, s, fx, fy, fz);
        // sym: the cutoff test is symmetric in the bond, so the reverse bond
        // is active exactly when this one is and its mask need not be read
        if (sym ? own : ((amask[r / L] >> (r & (L - 1))) & 1)) {
          double bx, by, bz;
          g_get(G, rs, bx, by, bz);
          fx -= bx;
          fy -= by;
          fz -= bz;
        }
      }
    }
  }
  fx = seg_sum<L>(fx);
  fy = seg_sum<L>(fy);
  fz = seg_sum<L>(fz);
  if (live && k == 0) {
    F[3 * (size_t)a] = fx;
    F[3 * (size_t)a + 1] = fy;
    F[3 * (size_t)a + 2] = fz;
  }
}

template <int L>
__global__ __launch_bounds__(256) void k_gather(int n, const int* __restrict__ cnt,
                                                const int* __restrict__ rev, int stride,
                                                const double* __restrict__ G,
                                                double* __restrict__ F, int atom_blocks,
                                                int n_struct, const Geom* __restrict__ geoms,
                                                const int* __restrict__ first, int n_chunks,
                                                const double* __restrict__ partial,
                                                double* __restrict__ out,
                                                const uint64_t* __restrict__ amask, int sym) {
  gather_blocks<L>(n, cnt, rev, stride, G, F, atom_blocks, n_struct, geoms, first, n_chunks,
                   partial, out, amask, sym, (int)gridDim.x - atom_blocks, (int)blockDim.x);
}

// One structure on rows of four slots, with the arguments of that case only
// (the launch cost grows with the bytes of arguments), blocks of 256
// threads.  The lead-block count is an argument, so that nothing is read
// from the dispatch's hidden arguments, and the atom blocks' arguments come
// first (with kernel-argument preloading they arrive in SGPRs).
__global__ __launch_bounds__(256) void k_gather_one4(int n, int atom_blocks, int lead,
                                                     int n_chunks, const int* __restrict__ cnt,
                                                     const int* __restrict__ rev,
                                                     const double* __restrict__ G,
                                                     double* __restrict__ F,
                                                     const Geom* __restrict__ geoms,
                                                     const int* __restrict__ first,
                                                     const double* __restrict__ partial,
                                                     double* __restrict__ out) {
  gather_blocks<4>(n, cnt, rev, 4, G, F, atom_blocks, 1, geoms, first, n_chunks, partial, out,
                   nullptr, 1, lead, 256);
}

// ---------------------------------------------------------------------------
// Stillinger-Weber with long rows (stride L >= 16, e.g. a Verlet skin that
// reaches the second shell): LA lanes per atom instead of L.  Phase 1: lane q
// tests slots q, q + LA, ... and the legs inside r_c are appended, in slot
// order, to the atom's region of LDS (bond vector, length, exp(gamma/(x-a)),
// slot).  Phase 2: lane q evaluates legs q, q + LA, ...: two-body term and
// the three-body terms against every other active leg of the atom, in slot
// order -- the same terms in the same order as k_sw_site -- and writes G to
// the leg's slot.  Dynamic LDS: (block / LA) * L legs of 5 doubles + 1 int.
// ---------------------------------------------------------------------------
template <int L, int LA>
__global__ __launch_bounds__(256) void k_sw_site_c(
    int n_chunks, int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const double* __restrict__ prm,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const int* __restrict__ chunk_rng, double* __restrict__ G, double* __restrict__ partial,
    uint64_t* __restrict__ amask, double* __restrict__ esite) {
  static_assert(L >= LA && L % LA == 0 && LA <= 64, "LA lanes cover the row in L / LA rounds");
  extern __shared__ double lds[];
  const int c = xcd_remap((int)blockIdx.x, n_chunks);
  const int apb = blockDim.x / LA;          // atoms per block
  const int al = (int)threadIdx.x / LA;     // atom of this lane within the block
  const int q = threadIdx.x & (LA - 1);
  const int seg = (threadIdx.x & 63) & ~(LA - 1);
  double* lx = lds;
  double* ly = lx + apb * L;
  double* lz = ly + apb * L;
  double* lri = lz + apb * L;  // 1 / r
  double* lE = lri + apb * L;
  int* lslot = (int*)(lE + apb * L);
  const int reg = al * L;
  const int a0 = single ? c * apb : chunk_rng[2 * c];
  const int a1 = single ? min(a0 + apb, n) : chunk_rng[2 * c + 1];
  const Geom& g = geoms[single ? 0 : sid[a0]];
  const int a = a0 + al;
  const bool live = a < a1;
  const int ac = live ? a : a0;
  const int nk = live ? cnt[ac] : 0;
  const double eps = prm[0], sig = prm[1], Asw = prm[2], Bsw = prm[3], pp = prm[4],
               qq = prm[5], asw = prm[6], lam = prm[7], gam = prm[8], cth0 = prm[9];
  const double rc = asw * sig;
  const double inv_sig = fast_rcp(sig);

  // phase 1: active legs, compacted in slot order (the row's indices are
  // loaded for every round at once)
  int jr[L / LA];
#pragma unroll
  for (int i = 0; i < L / LA; ++i) jr[i] = nbr[(size_t)ac * stride + q + LA * i];
  int count = 0;
  uint64_t mask = 0;
#pragma unroll
  for (int i = 0; i < L / LA; ++i) {
    const int k = q + LA * i;
    const bool has = live && k < nk;
    double ux, uy, uz;
    mic_fast(g, pos, ac, has ? jr[i] : ac, ux, uy, uz);
    const double r2 = ux * ux + uy * uy + uz * uz;
    const double rinv = r2 > 1e-24 ? fast_rsq(r2) : 0.0;
    const double r = r2 * rinv, x = r * inv_sig;
    const bool leg = has && r < rc && r > 1e-12 && x < asw;
    const uint64_t bal = __ballot(leg);
    const uint64_t sm = LA == 64 ? bal : (bal >> seg) & ((uint64_t(1) << (LA & 63)) - 1);
    if (leg) {
      const int t = reg + count + __popcll(sm & ((uint64_t(1) << q) - 1));
      lx[t] = ux;
      ly[t] = uy;
      lz[t] = uz;
      lri[t] = rinv;
      lE[t] = nb_exp(gam * fast_rcp(x - asw));
      lslot[t] = k;
    }
    mask |= sm << (LA * i);
    count += __popcll(sm);
  }
  if (live && q == 0) amask[ac] = mask;
  __syncthreads();

  // phase 2: legs q, q + LA, ... of this atom
  const double c3 = lam * eps;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // e, -sum G (x) u symmetrised
  for (int t = q; t < count; t += LA) {
    const double ux = lx[reg + t], uy = ly[reg + t], uz = lz[reg + t], rinv = lri[reg + t];
    const double E = lE[reg + t];
    // the distance of phase 1, from the same operations
    const double r = (ux * ux + uy * uy + uz * uz) * rinv;
    double e, su, dlnE;
    sw_leg(r, rinv, eps, sig, inv_sig, Asw, Bsw, pp, qq, asw, gam, e, su, dlnE);
    double swx = 0.0, swy = 0.0, swz = 0.0;
    for (int s2 = 0; s2 < count; ++s2) {
      if (s2 == t) continue;
      const double wx = lx[reg + s2], wy = ly[reg + s2], wz = lz[reg + s2];
      const double wri = lri[reg + s2], wE = lE[reg + s2];
      double cs = (ux * wx + uy * wy + uz * wz) * rinv * wri;
      double dc = cs - cth0;
      double pre = c3 * E * wE * dc;
      e += 0.5 * pre * dc;
      double rad = pre * dc * dlnE, ang = 2.0 * pre * rinv;
      su += (rad - ang * cs) * rinv;
      double aw = ang * wri;
      swx += aw * wx;
      swy += aw * wy;
      swz += aw * wz;
    }
    const double Gx = su * ux + swx, Gy = su * uy + swy, Gz = su * uz + swz;
    const size_t slot = (size_t)ac * stride + lslot[reg + t];
    g_put(G, slot, Gx, Gy, Gz);
    v[0] += e;
    v[1] -= Gx * ux;
    v[2] -= 0.5 * (Gx * uy + Gy * ux);
    v[3] -= 0.5 * (Gx * uz + Gz * ux);
    v[4] -= Gy * uy;
    v[5] -= 0.5 * (Gy * uz + Gz * uy);
    v[6] -= Gz * uz;
  }
  if (esite) {  // per-atom site energy U_a
    const double ea = seg_sum<LA>(v[0]);
    if (live && q == 0) esite[a] = ea;
  }
  block_partial7(v, partial, c, n_chunks);
}

// Gather for long rows (stride L >= 16): LA lanes per atom, lane q summing
// slots q, q + LA, ... (only bonds inside r_c hold G; see the masks).
template <int L, int LA>
__global__ __launch_bounds__(256) void k_gather_c(int n, const int* __restrict__ cnt,
                                                  const int* __restrict__ rev, int stride,
                                                  const double* __restrict__ G,
                                                  double* __restrict__ F, int atom_blocks,
                                                  int n_struct, const Geom* __restrict__ geoms,
                                                  const int* __restrict__ first, int n_chunks,
                                                  const double* __restrict__ partial,
                                                  double* __restrict__ out,
                                                  const uint64_t* __restrict__ amask, int sym) {
  // the structure blocks come first, so that they are dispatched before
  // the atom blocks instead of forming the launch's tail; gather_lead pads
  // them to a whole number of XCD rounds, so that atom tile b runs on the
  // XCD whose L2 holds the site gradients of site-kernel chunk b
  const int lead = (int)gridDim.x - atom_blocks;
  if ((int)blockIdx.x < lead) {
    if ((int)blockIdx.x < n_struct)
      struct_reduce_block<2>((int)blockIdx.x, geoms, first, n_chunks, partial, out);
    return;
  }
  const int gt = xcd_remap((int)blockIdx.x - lead, atom_blocks) * blockDim.x + threadIdx.x;
  const int a = gt / LA;
  const int q = gt & (LA - 1);
  const bool live = a < n;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (live) {
    const int c = cnt[a];
    const uint64_t ma = amask[a];
    // all rounds unrolled so that the reverse-slot, mask and G loads of the
    // rounds are in flight together
    constexpr int R = L / LA;
    int rv[R];
    bool own[R], oth[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int k = q + LA * i;
      own[i] = k < c && ((ma >> k) & 1);
      // sym (cutoff symmetric in the bond): only active bonds have an active
      // reverse, so only their reverse slots are read
      rv[i] = (sym ? own[i] : k < c) ? rev[(size_t)a * stride + k] : -1;
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
      oth[i] = rv[i] >= 0 &&
               (sym ? own[i] : (bool)((amask[rv[i] / L] >> (rv[i] & (L - 1))) & 1));
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if (rv[i] < 0) continue;
      const size_t s = (size_t)a * stride + q + LA * i;
      if (own[i]) {
        double gx, gy, gz;
        g_get(G, s, gx, gy, gz);
        fx += gx;
        fy += gy;
        fz += gz;
      }
      if (oth[i]) {
        double gx, gy, gz;
        g_get(G, (size_t)rv[i], gx, gy, gz);
        fx -= gx;
        fy -= gy;
        fz -= gz;
      }
    }
  }
  fx = seg_sum<LA>(fx);
  fy = seg_sum<LA>(fy);
  fz = seg_sum<LA>(fz);
  if (live && q == 0) {
    F[3 * (size_t)a] = fx;
    F[3 * (size_t)a + 1] = fy;
    F[3 * (size_t)a + 2] = fz;
  }
}

// ---------------------------------------------------------------------------
// Stillinger-Weber, multi-species tables (Tomita et al., Jpn. J. Appl. Phys.
// 57, 04FB04 (2018), eq. 4; StillingerWeberPotential._compute_multi_species,
// engine/stillinger_weber.py:1165-1451, kernels :414-560):
//   E_a = 1/2 sum_j V2_{p(a,j)}(r_aj)
//       + sum_{j != k} c_(s_a,s_j,s_k) E_aj E_ak (cos - cos0_(s_a,s_j,s_k))^2,
// with the pair record p(a, j) = (s_lo, s_hi) of the lower-indexed atom, the
// leg factor E_aj = exp(gamma / (r / sigma - a)) of the leg record
// (s_a, s_j), and c = epsilon lambda / 2 of the triple record (each ordered
// partner pair is one term).  The lane of leg j takes the terms (j, k) for
// its energy and both (j, k) and (k, j) for G_aj.
// Tables (d_params): pair (ns, ns, 7) = epsilon, sigma, a, A, B, p, q with
// a = 0 where there is no two-body term; leg (ns, ns, 3) = sigma, a, gamma
// with a = 0 where the record is absent; triple (ns, ns, ns, 2) = c, cos0
// with c = 0 where the term is absent.  The triple table is staged in LDS.
// Same two phases as k_sw_site_c: LA lanes per atom, legs compacted in slot
// order into LDS; dynamic LDS = legs (6 doubles + 3 ints) + triple table.
// ---------------------------------------------------------------------------
static inline size_t sw_m_lds(int blk, int L, int LA, int ns) {
  const size_t legs = (size_t)(blk / LA) * L;
  return legs * 6 * sizeof(double) + (legs * 3 * sizeof(int) + 7) / 8 * 8 +
         sizeof(double) * 2 * (size_t)ns * ns * ns;
}

template <int L, int LA>
__device__ __forceinline__ void k_sw_multi_chunk(
    int c, int n_chunks, int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns,
    const double* __restrict__ tpair, const double* __restrict__ tleg, double rc,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const int* __restrict__ chunk_rng, double* __restrict__ G, double* __restrict__ partial,
    uint64_t* __restrict__ amask, double* __restrict__ lds, const double* __restrict__ s_tri,
    double* __restrict__ esite) {
  static_assert(L >= LA && L % LA == 0 && LA <= 64, "LA lanes cover the row in L / LA rounds");
  const int apb = blockDim.x / LA;
  const int al = (int)threadIdx.x / LA;
  const int q = threadIdx.x & (LA - 1);
  const int seg = (threadIdx.x & 63) & ~(LA - 1);
  double* lx = lds;
  double* ly = lx + apb * L;
  double* lz = ly + apb * L;
  double* lr = lz + apb * L;
  double* lE = lr + apb * L;
  double* lD = lE + apb * L;
  int* lslot = (int*)(lD + apb * L);
  int* lsp = lslot + apb * L;
  int* lpr = lsp + apb * L;
  const int reg = al * L;
  const int a = chunk_rng[2 * c] + al;
  const bool live = a < chunk_rng[2 * c + 1];
  const int ac = live ? a : 0;
  const int nk_row = cnt[ac];  // loaded with the neighbour index
  const int nk = live ? nk_row : 0;
  const int si = spc[ac];
  const Geom& g = geoms[single ? 0 : sid[chunk_rng[2 * c]]];

  // phase 1: legs inside a two-body or a three-body cutoff, in slot order
  int count = 0;
  uint64_t mask = 0;
#pragma unroll
  for (int i = 0; i < L / LA; ++i) {
    const int k = q + LA * i;
    const bool has = live && k < nk;
    double ux = 0.0, uy = 0.0, uz = 0.0, r = 0.0, E = 0.0, dlnE = 0.0;
    int sj = 0, pr = -1;
    bool act = false;
    if (has) {
      const int j = nbr[(size_t)ac * stride + k];
      mic_fast(g, pos, ac, j, ux, uy, uz);
      r = sqrt(ux * ux + uy * uy + uz * uz);
      sj = spc[j];
      if (r < rc && r > 1e-12) {
        const int p = ac < j ? si * ns + sj : sj * ns + si;
        const double* tp = tpair + 7 * (size_t)p;
        if (r / tp[1] < tp[2]) pr = p;
        const double* tl = tleg + 3 * ((size_t)si * ns + sj);
        const double xl = r / tl[0];
        if (xl < tl[1]) {
          const double d = xl - tl[1];
          E = nb_exp(tl[2] / d);
          dlnE = -tl[2] / (tl[0] * d * d);
        }
        act = pr >= 0 || E != 0.0;
      }
    }
    const uint64_t bal = __ballot(act);
    const uint64_t sm = LA == 64 ? bal : (bal >> seg) & ((uint64_t(1) << (LA & 63)) - 1);
    if (act) {
      const int t = reg + count + __popcll(sm & ((uint64_t(1) << q) - 1));
      lx[t] = ux;
      ly[t] = uy;
      lz[t] = uz;
      lr[t] = r;
      lE[t] = E;
      lD[t] = dlnE;
      lslot[t] = k;
      lsp[t] = sj;
      lpr[t] = pr;
    }
    mask |= sm << (LA * i);
    count += __popcll(sm);
  }
  if (live && q == 0) amask[ac] = mask;
  __syncthreads();

  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = q; t < count; t += LA) {
    const double ux = lx[reg + t], uy = ly[reg + t], uz = lz[reg + t], r = lr[reg + t];
    const double E = lE[reg + t], dlnE = lD[reg + t];
    const int sj = lsp[reg + t], pr = lpr[reg + t];
    const double rinv = 1.0 / r;
    double e = 0.0, su = 0.0;
    if (pr >= 0) {
      const double* tp = tpair + 7 * (size_t)pr;
      const double eps = tp[0], sig = tp[1], asw = tp[2], Asw = tp[3], Bsw = tp[4], pp = tp[5],
                   qq = tp[6];
      const double x = r / sig;
      const double tt = 1.0 / (x - asw);
      const double ex = nb_exp(tt);
      const double bx = Bsw * xpow(x, -pp) - xpow(x, -qq);
      const double V2 = eps * Asw * bx * ex;
      const double dV2 = eps * Asw * ex *
                         ((-pp * Bsw * xpow(x, -pp - 1.0) + qq * xpow(x, -qq - 1.0)) +
                          bx * (-tt * tt));
      e = 0.5 * V2;
      su = 0.5 * (dV2 / sig) * rinv;
    }
    double swx = 0.0, swy = 0.0, swz = 0.0;
    if (E != 0.0) {
      const double* row1 = s_tri + 2 * (size_t)((si * ns + sj) * ns);  // (s_a, s_j, s_k)
      const double* col2 = s_tri + 2 * (size_t)(si * ns * ns + sj);    // (s_a, s_k, s_j)
      for (int s2 = 0; s2 < count; ++s2) {
        const double wE = lE[reg + s2];
        if (s2 == t || wE == 0.0) continue;
        const int sk = lsp[reg + s2];
        const double wx = lx[reg + s2], wy = ly[reg + s2], wz = lz[reg + s2];
        const double wri = 1.0 / lr[reg + s2];
        const double c1 = row1[2 * sk], q1 = row1[2 * sk + 1];
        const double c2 = col2[2 * ns * sk], q2 = col2[2 * ns * sk + 1];
        const double cs = (ux * wx + uy * wy + uz * wz) * rinv * wri;
        const double d1 = cs - q1, d2 = cs - q2;
        const double EE = E * wE;
        e += c1 * EE * d1 * d1;
        // dh/du over both orders: E E [(c1 d1^2 + c2 d2^2) dlnE u_hat
        //                              + 2 (c1 d1 + c2 d2) (w_hat - cos u_hat) / r]
        const double rad = EE * (c1 * d1 * d1 + c2 * d2 * d2) * dlnE;
        const double ang = 2.0 * EE * (c1 * d1 + c2 * d2) * rinv;
        su += (rad - ang * cs) * rinv;
        const double aw = ang * wri;
        swx += aw * wx;
        swy += aw * wy;
        swz += aw * wz;
      }
    }
    const double Gx = su * ux + swx, Gy = su * uy + swy, Gz = su * uz + swz;
    const size_t slot = (size_t)ac * stride + lslot[reg + t];
    g_put(G, slot, Gx, Gy, Gz);
    v[0] += e;
    v[1] -= Gx * ux;
    v[2] -= Gx * uy;
    v[3] -= Gx * uz;
    v[4] -= Gy * ux;
    v[5] -= Gy * uy;
    v[6] -= Gy * uz;
    v[7] -= Gz * ux;
    v[8] -= Gz * uy;
    v[9] -= Gz * uz;
  }
  if (esite) {  // per-atom site energy U_a
    const double ea = seg_sum<LA>(v[0]);
    if (live && q == 0) esite[a] = ea;
  }
  wave_partial(v, partial, c, n_chunks);
}

template <int L, int LA>
__global__ __launch_bounds__(256) void k_sw_multi(
    int n_chunks, int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns,
    const double* __restrict__ tables, double rc, const int* __restrict__ cnt,
    const int* __restrict__ nbr, int stride, const int* __restrict__ chunk_rng,
    double* __restrict__ G, double* __restrict__ partial, uint64_t* __restrict__ amask,
    double* __restrict__ esite) {
  extern __shared__ double lds_swm[];
  const int n2 = ns * ns;
  const double* tpair = tables;
  const double* tleg = tpair + 7 * (size_t)n2;
  const double* ttri = tleg + 3 * (size_t)n2;
  const int apb = blockDim.x / LA;
  double* s_tri = lds_swm + (size_t)apb * L * 6 + ((size_t)apb * L * 3 * sizeof(int) + 7) / 8;
  for (int i = threadIdx.x; i < 2 * n2 * ns; i += blockDim.x) s_tri[i] = ttri[i];
  __syncthreads();
  k_sw_multi_chunk<L, LA>(xcd_remap((int)blockIdx.x, n_chunks), n_chunks, n, pos, sid, geoms, single, spc,
                            ns, tpair, tleg, rc, cnt, nbr, stride, chunk_rng, G, partial, amask,
                            lds_swm, s_tri, esite);
}

// ---------------------------------------------------------------------------
// EAM functions.  A function is a row of 44 doubles.  Zhou forms: [code, r_c,
// width, chi, pa[20], pb[20]] with code 3 density, 4 pair, 5 cross pair,
// 6 embedding; element parameters in the order r_e, f_e, rho_e, rho_s,
// alpha, beta, A, B, kappa, lambda, Fn0..Fn3, F0..F3, eta, F_e
// (engine/_eam_kernels.py:129-223).  Cubic splines (scipy PPoly, the
// tables of a setfl file): [code, X offset, n, C offset, 1/h, x_0] with code
// 1 = s(r), 2 = s(r) / r (the r*phi table); X (n knots) and C (4 doubles per
// knot: c0..c3 of the interval starting there) sit after the rows in the same
// buffer, the offsets counted from its start (engine/_eam_kernels.py:128-145,
// 199-205).
// ---------------------------------------------------------------------------
constexpr int ZW = 44;

// Interval i with X[i] <= r < X[i+1] as searchsorted(X, r, 'right') - 1,
// clipped to [0, n - 2] (extrapolation by the end polynomials), then Horner.
// With a uniform grid (1/h > 0) the search starts from floor((r - x_0) / h)
// and steps to the exact interval; otherwise it bisects.
__device__ __forceinline__ void spline_eval(const double* __restrict__ fn,
                                            const double* __restrict__ f, double r, double& v,
                                            double& dv) {
  const double* X = fn + (size_t)f[1];
  const int n = (int)f[2];
  const double* C = fn + (size_t)f[3];
  int i;
  if (f[4] > 0.0) {
    double g = floor((r - f[5]) * f[4]);
    g = fmin(fmax(g, -1.0), (double)(n - 1));  // NaN -> -1
    i = (int)g;
    while (i + 1 < n && X[i + 1] <= r) ++i;
    while (i >= 0 && X[i] > r) --i;
  } else {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (X[mid] <= r)
        lo = mid + 1;
      else
        hi = mid;
    }
    i = lo - 1;
  }
  i = i < 0 ? 0 : i;
  i = i > n - 2 ? n - 2 : i;
  const double dx = r - X[i];
  const double4 c = *reinterpret_cast<const double4*>(C + 4 * (size_t)i);
  v = ((c.x * dx + c.y) * dx + c.z) * dx + c.w;
  dv = (3.0 * c.x * dx + 2.0 * c.y) * dx + c.z;
}

__device__ __forceinline__ void z_decay(double r, double re, double pre, double expo,
                                        double shift, double& val, double& dval) {
  double x = r / re;
  double e = pre * exp(-expo * (x - 1.0));
  double u = x - shift;
  double u2 = u * u, u4 = u2 * u2, u8 = u4 * u4, u16 = u8 * u8;
  double u19 = u16 * u2 * u;
  double den = 1.0 + u19 * u;
  double iden = 1.0 / den;
  val = e * iden;
  dval = (-expo * val - e * 20.0 * u19 * iden * iden) / re;
}

__device__ __forceinline__ void z_embed(const double* __restrict__ p, double rho, double& F,
                                        double& dF) {
  double re = p[2], rs = p[3], eta = p[18];
  double rn = 0.85 * re;
  if (rho < rn) {
    double z = rho / rn - 1.0;
    F = p[10] + p[11] * z + p[12] * z * z + p[13] * z * z * z;
    dF = (p[11] + 2.0 * p[12] * z + 3.0 * p[13] * z * z) / rn;
  } else if (rho >= 1.15 * re) {
    double t = rho / rs, lt = log(t);
    double te = pow(t, eta);
    F = p[19] * (1.0 - eta * lt) * te;
    dF = -p[19] * eta * eta * lt * (te / t) / rs;
  } else {
    double z = rho / re - 1.0;
    F = p[14] + p[15] * z + p[16] * z * z + p[17] * z * z * z;
    dF = (p[15] + 2.0 * p[16] * z + 3.0 * p[17] * z * z) / re;
  }
}

__device__ __forceinline__ void z_feval(const double* __restrict__ fn,
                                        const double* __restrict__ f, double r, double& v,
                                        double& dv) {
  const int code = (int)f[0];
  if (code <= 2) {
    double sv, sd;
    spline_eval(fn, f, r, sv, sd);
    if (code == 1) {
      v = sv;
      dv = sd;
    } else {
      v = sv / r;
      dv = (sd * r - sv) / (r * r);
    }
    return;
  }
  if (code == 6) {
    z_embed(f + 4, r, v, dv);
    return;
  }
  const double* pa = f + 4;
  double val, dval;
  if (code == 3) {
    z_decay(r, pa[0], pa[1], pa[5], pa[9], val, dval);
  } else if (code == 4) {
    double A, dA, B, dB;
    z_decay(r, pa[0], pa[6], pa[4], pa[8], A, dA);
    z_decay(r, pa[0], pa[7], pa[5], pa[9], B, dB);
    val = A - B;
    dval = dA - dB;
  } else {
    const double* pb = f + 24;
    double fa, dfa, fb, dfb, A, dA, B, dB;
    z_decay(r, pa[0], pa[1], pa[5], pa[9], fa, dfa);
    z_decay(r, pb[0], pb[1], pb[5], pb[9], fb, dfb);
    z_decay(r, pa[0], pa[6], pa[4], pa[8], A, dA);
    z_decay(r, pa[0], pa[7], pa[5], pa[9], B, dB);
    double ua = A - B, dua = dA - dB;
    z_decay(r, pb[0], pb[6], pb[4], pb[8], A, dA);
    z_decay(r, pb[0], pb[7], pb[5], pb[9], B, dB);
    double ub = A - B, dub = dA - dB;
    double kc = f[3];
    double q = fb / fa;
    double dq = (dfb * fa - fb * dfa) / (fa * fa);
    val = 0.5 * kc * (q * ua + ub / q);
    dval = 0.5 * kc * (dq * ua + q * dua + dub / q - ub * dq / (q * q));
  }
  const double rc = f[1], w = f[2];
  double s = (r - (rc - w)) / w;
  s = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
  double S = 1.0 - s * s * s * (10.0 - 15.0 * s + 6.0 * s * s);
  double dS = -30.0 * s * s * (1.0 - s) * (1.0 - s) / w;
  v = val * S;
  dv = dval * S + val * dS;
}

// Zhou functions of one pair from a per-species table (the alloy form
// checked by eam_fast_table).  For species s at distance r (x = r / r_e):
//   density  f   = f_e e^{-beta (x-1)} / (1 + (x - lambda)^20)
//   pair     phi = A e^{-alpha (x-1)} / (1 + (x - kappa)^20)
//                - B e^{-beta (x-1)} / (1 + (x - lambda)^20)
// e^{-beta (x-1)} and the lambda denominator are shared: two exponentials
// per species instead of one per function term (a cross pair needs both
// species' densities and pair terms: 4 instead of 8).  Divisions by r_e and
// by the switch widths are products with reciprocals kept in the table; the
// two denominators and the cross pair's density ratio use the hardware
// reciprocal with Newton steps.  The values agree with z_decay / z_feval to
// a few ulp.
// Table: per species (ZF_SP doubles) r_e, f_e, beta, lambda, A, alpha,
// kappa, B, density-row r_c and switch width, 1 / r_e, 1 / width; per
// ordered pair (sa, sj) (ZF_PR doubles) code (4 pair, 5 cross), r_c, width,
// chi, swap (the cross row's first species is sj), 1 / width.
constexpr int ZF_SP = 12, ZF_PR = 6;

struct ZBase {
  double fv, fd, pv, pd;  // unswitched density and pair value, derivative
};

// (x - c)^19 and (x - c)^20
__device__ __forceinline__ void z_pow20(double u, double& u19, double& u20) {
  const double u2 = u * u, u4 = u2 * u2, u8 = u4 * u4, u16 = u8 * u8;
  u19 = u16 * u2 * u;
  u20 = u19 * u;
}

__device__ __forceinline__ ZBase z_base(const double* __restrict__ sp, double r) {
  const double fe = sp[1], beta = sp[2], lam = sp[3], A = sp[4], alpha = sp[5], kap = sp[6],
               B = sp[7], ire = sp[10];
  const double x = r * ire, xm = x - 1.0;
  const double Eb = exp(-beta * xm);
  const double Ea = exp(-alpha * xm);
  double ul19, ul20, uk19, uk20;
  z_pow20(x - lam, ul19, ul20);
  z_pow20(x - kap, uk19, uk20);
  const double idl = fast_rcp(1.0 + ul20), idk = fast_rcp(1.0 + uk20);
  const double tl = 20.0 * ul19 * idl, tk = 20.0 * uk19 * idk;  // -d ln(den) / dx
  ZBase b;
  const double ef = fe * Eb;
  b.fv = ef * idl;
  b.fd = -(beta + tl) * b.fv * ire;
  const double vA = A * Ea * idk, vB = B * Eb * idl;
  b.pv = vA - vB;
  b.pd = (-(alpha + tk) * vA + (beta + tl) * vB) * ire;
  return b;
}

// Switched density of species s at r (value only).
__device__ __forceinline__ double z_dens(const double* __restrict__ sp, double r) {
  const double fe = sp[1], beta = sp[2], lam = sp[3], ire = sp[10];
  const double x = r * ire;
  double ul19, ul20;
  z_pow20(x - lam, ul19, ul20);
  const double v = fe * exp(-beta * (x - 1.0)) * fast_rcp(1.0 + ul20);
  double t = (r - (sp[8] - sp[9])) * sp[11];
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  return v * (1.0 - t * t * t * (10.0 - 15.0 * t + 6.0 * t * t));
}

__device__ __forceinline__ void z_switch(double r, double rc, double w, double iw, double val,
                                         double dval, double& v, double& dv) {
  double s = (r - (rc - w)) * iw;
  s = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
  const double S = 1.0 - s * s * s * (10.0 - 15.0 * s + 6.0 * s * s);
  const double dS = -30.0 * s * s * (1.0 - s) * (1.0 - s) * iw;
  v = val * S;
  dv = dval * S + val * dS;
}

// d f_sj / dr, d f_sa / dr and phi_{sa sj}, phi' at r.  one_sw: every row
// has the switch of species 0's density row (the table's trailing flag),
// formed once and applied to all three functions.
__device__ __forceinline__ void z_pair_fast(const double* __restrict__ fast, int ns, int sa,
                                            int sj, double r, bool one_sw, double& dfj,
                                            double& dfa, double& phi, double& dphi) {
  const double* spj = fast + (size_t)ZF_SP * sj;
  const double* spa = fast + (size_t)ZF_SP * sa;
  const double* pr = fast + (size_t)ZF_SP * ns + (size_t)ZF_PR * (sa * ns + sj);
  const ZBase bj = z_base(spj, r);
  if (one_sw) {
    double t = (r - (fast[8] - fast[9])) * fast[11];
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    const double S = 1.0 - t * t * t * (10.0 - 15.0 * t + 6.0 * t * t);
    const double dS = -30.0 * t * t * (1.0 - t) * (1.0 - t) * fast[11];
    dfj = bj.fd * S + bj.fv * dS;
    double val, dval;
    if (sa == sj) {
      dfa = dfj;
      val = bj.pv;
      dval = bj.pd;
    } else {
      const ZBase ba = z_base(spa, r);
      dfa = ba.fd * S + ba.fv * dS;
      const ZBase& X = pr[4] != 0.0 ? bj : ba;
      const ZBase& Y = pr[4] != 0.0 ? ba : bj;
      const double kc = pr[3];
      const double iX = fast_rcp(X.fv), iY = fast_rcp(Y.fv);
      const double q = Y.fv * iX, iq = X.fv * iY;
      const double dq = (Y.fd * X.fv - Y.fv * X.fd) * iX * iX;
      val = 0.5 * kc * (q * X.pv + Y.pv * iq);
      dval = 0.5 * kc * (dq * X.pv + q * X.pd + Y.pd * iq - Y.pv * dq * iq * iq);
    }
    phi = val * S;
    dphi = dval * S + val * dS;
    return;
  }
  double v;
  z_switch(r, spj[8], spj[9], spj[11], bj.fv, bj.fd, v, dfj);
  double val, dval;
  if (sa == sj) {
    dfa = dfj;
    val = bj.pv;
    dval = bj.pd;
  } else {
    const ZBase ba = z_base(spa, r);
    z_switch(r, spa[8], spa[9], spa[11], ba.fv, ba.fd, v, dfa);
    const ZBase& X = pr[4] != 0.0 ? bj : ba;  // species of the row's first block
    const ZBase& Y = pr[4] != 0.0 ? ba : bj;
    const double kc = pr[3];
    const double iX = fast_rcp(X.fv), iY = fast_rcp(Y.fv);
    const double q = Y.fv * iX, iq = X.fv * iY;
    const double dq = (Y.fd * X.fv - Y.fv * X.fd) * iX * iX;
    val = 0.5 * kc * (q * X.pv + Y.pv * iq);
    dval = 0.5 * kc * (dq * X.pv + q * X.pd + Y.pd * iq - Y.pv * dq * iq * iq);
  }
  z_switch(r, pr[1], pr[2], pr[5], val, dval, phi, dphi);
}

// map: [rho_f[ns], emb_f[ns], phi_f[ns*ns]]
template <int LE>
__global__ __launch_bounds__(256) void k_eam_rho(
    int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns,
    const double* __restrict__ fn, const int* __restrict__ map, double rc,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    double* __restrict__ dF, double* __restrict__ Fe) {
  const int gt = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int a = gt / LE;
  const int l = gt & (LE - 1);
  const bool live = a < n;
  double rho = 0.0;
  if (live) {
    const Geom& g = geoms[single ? 0 : sid[a]];
    const int nk = cnt[a];
    for (int k = l; k < nk; k += LE) {
      const int j = nbr[(size_t)a * stride + k];
      double ux, uy, uz;
      mic_fast(g, pos, a, j, ux, uy, uz);
      double r = sqrt(ux * ux + uy * uy + uz * uz);
      if (r < rc && r > 1e-12) {
        const int sj = ns > 1 ? spc[j] : 0;
        double v, dv;
        z_feval(fn, fn + (size_t)ZW * map[sj], r, v, dv);
        rho += v;
      }
    }
  }
  rho = seg_sum<LE>(rho);
  if (live && l == 0) {
    const int sa = ns > 1 ? spc[a] : 0;
    double F, d;
    z_feval(fn, fn + (size_t)ZW * map[ns + sa], rho > 0.0 ? rho : 0.0, F, d);
    Fe[a] = F;
    dF[a] = d;
  }
}

template <int LE>
__device__ __forceinline__ void k_eam_force_chunk(int c, int n_chunks, 
    const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns,
    const double* __restrict__ fn, const int* __restrict__ map, double rc,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const double* __restrict__ dF, const double* __restrict__ Fe,
    const int* __restrict__ chunk_rng, double* __restrict__ F, double* __restrict__ partial,
    double* __restrict__ esite, int* __restrict__ pair_flag) {
  const int a = chunk_rng[2 * c] + (int)threadIdx.x / LE;
  const int l = threadIdx.x & (LE - 1);
  const bool live = a < chunk_rng[2 * c + 1];
  int bonded = 0;  // a bond inside r_c seen by this lane
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
  double w[6] = {0, 0, 0, 0, 0, 0};  // xx, xy, xz, yy, yz, zz
  if (live) {
    const Geom& g = geoms[single ? 0 : sid[a]];
    const int nk = cnt[a];
    const int sa = ns > 1 ? spc[a] : 0;
    const double dFa = dF[a];
    const double* fa_rho = fn + (size_t)ZW * map[sa];
    if (l == 0) e = Fe[a];
    for (int k = l; k < nk; k += LE) {
      const int j = nbr[(size_t)a * stride + k];
      double ux, uy, uz;
      mic_fast(g, pos, a, j, ux, uy, uz);
      double r = sqrt(ux * ux + uy * uy + uz * uz);
      if (r < rc && r > 1e-12) {
        const int sj = ns > 1 ? spc[j] : 0;
        double v, dfj, dfa, phi, dphi;
        z_feval(fn, fn + (size_t)ZW * map[sj], r, v, dfj);
        if (sj == sa)
          dfa = dfj;
        else
          z_feval(fn, fa_rho, r, v, dfa);
        z_feval(fn, fn + (size_t)ZW * map[2 * ns + sa * ns + sj], r, phi, dphi);
        bonded = 1;
        double c = dphi + dFa * dfj + dF[j] * dfa;
        double s = c / r;
        fx += s * ux;
        fy += s * uy;
        fz += s * uz;
        e += 0.5 * phi;
        double hs = -0.5 * s;
        w[0] += hs * ux * ux;
        w[1] += hs * ux * uy;
        w[2] += hs * ux * uz;
        w[3] += hs * uy * uy;
        w[4] += hs * uy * uz;
        w[5] += hs * uz * uz;
      }
    }
  }
  fx = seg_sum<LE>(fx);
  fy = seg_sum<LE>(fy);
  fz = seg_sum<LE>(fz);
  if (live && l == 0) {
    F[3 * (size_t)a] = fx;
    F[3 * (size_t)a + 1] = fy;
    F[3 * (size_t)a + 2] = fz;
  }
  if (esite) {  // per-atom site energy U_a = F_a + 1/2 sum_j phi_aj
    const double ea = seg_sum<LE>(e);
    if (live && l == 0) esite[a] = ea;
  }
  double v[10] = {e, w[0], w[1], w[2], w[1], w[3], w[4], w[2], w[4], w[5]};
  wave_partial(v, partial, c, n_chunks);
  const int any = __syncthreads_or(bonded);
  if (threadIdx.x == 0) pair_flag[c] = any;
}

template <int LE>
__global__ __launch_bounds__(256) void k_eam_force(int n_chunks, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns,
    const double* __restrict__ fn, const int* __restrict__ map, double rc,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const double* __restrict__ dF, const double* __restrict__ Fe,
    const int* __restrict__ chunk_rng, double* __restrict__ F, double* __restrict__ partial,
    double* __restrict__ esite, int* __restrict__ pair_flag) {
  for (int cc = blockIdx.x; cc < n_chunks; cc += gridDim.x)
    k_eam_force_chunk<LE>(xcd_remap(cc, n_chunks), n_chunks, pos, sid, geoms, single, spc, ns, fn, map, rc, cnt, nbr, stride, dF, Fe, chunk_rng, F, partial, esite, pair_flag);
}

// ---------------------------------------------------------------------------
// Zhou alloy EAM on the per-species table (z_base, z_dens): the density pass
// and the force pass of EAMPotential.compute (engine/eam.py:174-256,
// engine/_eam_kernels.py:231-307).  One block per structure-aligned chunk of
// blockDim / LE atoms (the cell geometry is uniform over the block and
// comes through scalar loads; with a single structure the chunk is
// c * blockDim / LE onwards), LE lanes per atom striding over its row.
// Same terms as k_eam_rho / k_eam_force; the virial partials are the six
// components of the symmetric sum -1/2 sum s u (x) u.
// ---------------------------------------------------------------------------
template <int LE>
__global__ __launch_bounds__(256) void k_eam_rho_z(
    int n_chunks, int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns,
    const double* __restrict__ fn, const int* __restrict__ map,
    const double* __restrict__ fast, double rc, const int* __restrict__ cnt,
    const int* __restrict__ nbr, int stride, const int* __restrict__ chunk_rng,
    double* __restrict__ dF, double* __restrict__ Fe) {
  const int c = xcd_remap((int)blockIdx.x, n_chunks);
  const int apb = (int)blockDim.x / LE;
  const int a0 = single ? c * apb : chunk_rng[2 * c];
  const int a1 = single ? min(a0 + apb, n) : chunk_rng[2 * c + 1];
  const Geom& g = geoms[single ? 0 : sid[a0]];
  const int a = a0 + (int)threadIdx.x / LE;
  const int l = threadIdx.x & (LE - 1);
  const bool live = a < a1;
  const int ac = live ? a : a0;
  const int nk = live ? cnt[ac] : 0;
  const double rc2 = rc * rc;
  const bool ortho = geom_diagonal(g);
  double rho = 0.0;
  for (int k = l; k < nk; k += LE) {
    const int j = nbr[(size_t)ac * stride + k];
    double ux, uy, uz;
    mic_cell(g, ortho, pos, ac, j, ux, uy, uz);
    const double r2 = ux * ux + uy * uy + uz * uz;
    if (r2 < rc2 && r2 > 1e-24) {
      const int sj = ns > 1 ? spc[j] : 0;
      rho += z_dens(fast + (size_t)ZF_SP * sj, r2 * fast_rsq(r2));
    }
  }
  rho = seg_sum<LE>(rho);
  if (live && l == 0) {
    const int sa = ns > 1 ? spc[a] : 0;
    double F, d;
    z_feval(fn, fn + (size_t)ZW * map[ns + sa], rho > 0.0 ? rho : 0.0, F, d);
    Fe[a] = F;
    dF[a] = d;
  }
}

template <int LE>
__global__ __launch_bounds__(256) void k_eam_force_z(
    int n_chunks, int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns,
    const double* __restrict__ fast, double rc, const int* __restrict__ cnt,
    const int* __restrict__ nbr, int stride, const double* __restrict__ dF,
    const double* __restrict__ Fe, const int* __restrict__ chunk_rng, double* __restrict__ F,
    double* __restrict__ partial, double* __restrict__ esite, int* __restrict__ pair_flag) {
  const int c = xcd_remap((int)blockIdx.x, n_chunks);
  const int apb = (int)blockDim.x / LE;
  const int a0 = single ? c * apb : chunk_rng[2 * c];
  const int a1 = single ? min(a0 + apb, n) : chunk_rng[2 * c + 1];
  const Geom& g = geoms[single ? 0 : sid[a0]];
  const int a = a0 + (int)threadIdx.x / LE;
  const int l = threadIdx.x & (LE - 1);
  const bool live = a < a1;
  const int ac = live ? a : a0;
  const int nk = live ? cnt[ac] : 0;
  const int sa = ns > 1 ? spc[ac] : 0;
  const double dFa = dF[ac];
  const double rc2 = rc * rc;
  const bool one_sw = fast[(size_t)ZF_SP * ns + (size_t)ZF_PR * ns * ns] != 0.0;
  const bool ortho = geom_diagonal(g);
  int bonded = 0;  // a bond inside r_c seen by this lane
  double fx = 0.0, fy = 0.0, fz = 0.0;
  double e = live && l == 0 ? Fe[ac] : 0.0;
  double w[6] = {0, 0, 0, 0, 0, 0};  // xx, xy, xz, yy, yz, zz
  for (int k = l; k < nk; k += LE) {
    const int j = nbr[(size_t)ac * stride + k];
    double ux, uy, uz;
    mic_cell(g, ortho, pos, ac, j, ux, uy, uz);
    const double r2 = ux * ux + uy * uy + uz * uz;
    if (r2 < rc2 && r2 > 1e-24) {
      const int sj = ns > 1 ? spc[j] : 0;
      const double dFj = dF[j];
      const double rinv = fast_rsq(r2), r = r2 * rinv;
      double dfj, dfa, phi, dphi;
      z_pair_fast(fast, ns, sa, sj, r, one_sw, dfj, dfa, phi, dphi);
      bonded = 1;
      const double s = (dphi + dFa * dfj + dFj * dfa) * rinv;
      fx += s * ux;
      fy += s * uy;
      fz += s * uz;
      e += 0.5 * phi;
      const double hs = -0.5 * s;
      w[0] += hs * ux * ux;
      w[1] += hs * ux * uy;
      w[2] += hs * ux * uz;
      w[3] += hs * uy * uy;
      w[4] += hs * uy * uz;
      w[5] += hs * uz * uz;
    }
  }
  fx = seg_sum<LE>(fx);
  fy = seg_sum<LE>(fy);
  fz = seg_sum<LE>(fz);
  if (live && l == 0) {
    F[3 * (size_t)a] = fx;
    F[3 * (size_t)a + 1] = fy;
    F[3 * (size_t)a + 2] = fz;
  }
  if (esite) {  // per-atom site energy U_a = F_a + 1/2 sum_j phi_aj
    const double ea = seg_sum<LE>(e);
    if (live && l == 0) esite[a] = ea;
  }
  const double v[8] = {e, w[0], w[1], w[2], w[3], w[4], w[5], 0.0};
  block_partial7(v, partial, c, n_chunks);
  const int any = __syncthreads_or(bonded);
  if (threadIdx.x == 0) pair_flag[c] = any;
}

// Function f of the table at m arguments (tabulation of a potential).
__global__ __launch_bounds__(256) void k_eam_feval(const double* __restrict__ fn, int f, int64_t m,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ v,
                                                   double* __restrict__ dv) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) z_feval(fn, fn + (size_t)ZW * f, x[k], v[k], dv[k]);
}

void eam_feval(const Potential& pot, int f, int64_t m, const double* d_x, double* d_v,
               double* d_dv, hipStream_t st) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_eam_feval, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st,
                     (const double*)pot.d_params.as<double>(), f, m, d_x, d_v, d_dv);
}

// ---------------------------------------------------------------------------
// Host dispatch
// ---------------------------------------------------------------------------

// Launch the dominant kernel; while timing is enabled, through
// hipExtLaunchKernelGGL with start/stop events, which the runtime stamps
// from the kernel's own dispatch packet (no extra barrier packets).
template <typename K, typename... Args>
static void launch_timed(Potential& pot, K kernel, dim3 grid, dim3 block, hipStream_t st,
                         size_t lds, Args... args) {
  Potential::EventPair* ev = pot.prof_next();
  if (ev)
    hipExtLaunchKernelGGL(kernel, grid, block, lds, st, ev->a, ev->b, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
}

// Rows of at least kLongRow slots run the SW site kernel and the gather with
// kRowLanes lanes per atom (k_sw_site_c, k_gather_c).
constexpr int kLongRow = 16;
constexpr int kRowLanes = 8;
static inline size_t sw_c_lds(int blk, int L) {
  return (size_t)(blk / kRowLanes) * L * (5 * sizeof(double) + sizeof(int));
}

// Site-kernel grid: one block per chunk, or at most grid_cap blocks that
// stride over the chunks (persistent form).
static inline dim3 chunk_grid(const Potential& pot, int n_chunks) {
  int g = n_chunks;
  if (pot.grid_cap > 0 && g > pot.grid_cap) g = pot.grid_cap;
  return dim3(g > 0 ? g : 1);
}

// Leading blocks of a gather launch: one per structure, rounded up to a
// multiple of the XCD count (blocks are dispatched to the XCDs round-robin).
static inline int gather_lead(int n_struct) { return (n_struct + kXcds - 1) / kXcds * kXcds; }

// The four-slot Tersoff kernel runs: one species with the flat form's
// record (m = 3, gamma = 1) on rows of four slots, one lane per bond.
static bool q4_path(const Potential& pot, int lanes, bool trip) {
  const double* hp = pot.h_params.data();
  return pot.kind == KIND_TERSOFF && pot.ns == 1 && lanes == 4 && !trip && hp[0] != 0.0 &&
         hp[9] != 0.0 && hp[14] == 3.0 && hp[15] == 1.0;
}

// The single-structure four-slot kernel at block size BLK.
template <int BLK>
static void launch_q4_one(Potential& pot, bool l3z, int n_chunks, hipStream_t st, int n,
                          const double* pos, const Geom* geoms, const double* tb, double rc,
                          const int* cnt, const int* nbr, double* G, double* part,
                          double* esite) {
  launch_timed(pot, l3z ? k_tersoff_q4_one<true, BLK> : k_tersoff_q4_one<false, BLK>,
               dim3(n_chunks > 0 ? n_chunks : 1), dim3(BLK), st, (size_t)0, n, n_chunks, pos,
               geoms, tb, rc, cnt, nbr, G, part, esite);
}

template <int L>
static void launch_bond_kernels(Potential& pot, NList& nl, const ForceWork& w,
                                const double* d_pos, const int* d_spc, double* d_F,
                                double* d_out, bool trip, int blk, double* esite,
                                hipStream_t st) {
  const int n = w.n;
  const int single = w.n_struct == 1 ? 1 : 0;
  double* G = w.G;
  // one block per chunk (the bond kernels do not loop over chunks)
  const dim3 grid(w.n_chunks > 0 ? w.n_chunks : 1), block(blk);
  const int* sid = w.sid;
  const Geom* geoms = w.geoms;
  const int* cnt = nl.cnt.as<int>();
  const int* nbr = nl.nbr.as<int>();
  const int* rng = w.rng;
  double* part = w.partial;
  uint64_t* amask = w.amask;
  const int n_chunks = w.n_chunks, n_struct = w.n_struct;
  if (pot.kind == KIND_TERSOFF) {
    const double* tb = pot.d_params.as<double>();
    const double* tri = tb + (size_t)pot.ns * pot.ns * 9;
    const bool sp1 = pot.ns == 1;
    bool done = false;
    if constexpr (L * L <= 64) {
      if (trip) {
        launch_timed(pot, sp1 ? k_tersoff_trip<L, true> : k_tersoff_trip<L, false>, grid, block,
                     st, (size_t)0, n_chunks, n, d_pos, sid, geoms, single, d_spc, pot.ns, tb, tri, pot.r_cut,
                     cnt, nbr, nl.stride, rng, G, part, amask);
        done = true;
      }
    }
    if constexpr (L == 4) {
      // the flat form (one valid record, m = 3, gamma = 1) on rows of four:
      // one block per chunk (the grid cap does not apply)
      const double* hp = pot.h_params.data();
      if (!done && q4_path(pot, L, false)) {
        const dim3 cg(n_chunks > 0 ? n_chunks : 1);
        if (single) {
          const bool l3z = hp[13] == 0.0;
          switch (blk) {
            case 64: launch_q4_one<64>(pot, l3z, n_chunks, st, n, d_pos, geoms, tb, pot.r_cut, cnt, nbr, G, part, esite); break;
            case 128: launch_q4_one<128>(pot, l3z, n_chunks, st, n, d_pos, geoms, tb, pot.r_cut, cnt, nbr, G, part, esite); break;
            case 256: launch_q4_one<256>(pot, l3z, n_chunks, st, n, d_pos, geoms, tb, pot.r_cut, cnt, nbr, G, part, esite); break;
            default: launch_q4_one<512>(pot, l3z, n_chunks, st, n, d_pos, geoms, tb, pot.r_cut, cnt, nbr, G, part, esite); break;
          }
        } else
          launch_timed(pot, hp[13] == 0.0 ? k_tersoff_q4<true> : k_tersoff_q4<false>, cg, block,
                       st, (size_t)0, n_chunks, n, rng, sid, geoms, d_pos, cnt, nbr, tb,
                       pot.r_cut, G, part, esite);
        done = true;
      }
    }
    if (!done)
      launch_timed(pot, sp1 ? k_tersoff_site<L, true> : k_tersoff_site<L, false>, grid, block, st,
                   (size_t)0, n_chunks, n, d_pos, sid, geoms, single, d_spc, pot.ns, tb, tri, pot.r_cut, cnt,
                   nbr, nl.stride, rng, G, part, amask, esite);
  } else if (pot.sw_multi) {
    constexpr int LA = L < kRowLanes ? L : kRowLanes;
    launch_timed(pot, k_sw_multi<L, LA>, grid, block, st, sw_m_lds(blk, L, LA, pot.ns), n_chunks, n,
                 d_pos, sid, geoms, single, d_spc, pot.ns,
                 (const double*)pot.d_params.as<double>(), pot.r_cut, cnt, nbr, nl.stride, rng, G,
                 part, amask, esite);
  } else if constexpr (L >= kLongRow) {
    const size_t lds = sw_c_lds(blk, L);
    launch_timed(pot, k_sw_site_c<L, kRowLanes>, dim3(n_chunks > 0 ? n_chunks : 1), block, st,
                 lds, n_chunks, n, d_pos, sid,
                 geoms, single, (const double*)pot.d_params.as<double>(), cnt, nbr, nl.stride,
                 rng, G, part, amask, esite);
  } else {
    launch_timed(pot, k_sw_site<L>, grid, block, st, (size_t)0, n_chunks, n, d_pos, sid, geoms,
                 single, (const double*)pot.d_params.as<double>(), cnt, nbr, nl.stride, rng, G,
                 part, amask, esite);
  }
  const int* rev = nl.rev.as<int>();
  const int* first = w.first;
  // Tersoff (r < r_c of the potential) and flat SW (r < a sigma) test a bond
  // and its reverse with the same bitwise distance; multi-species SW legs
  // have per-species cutoffs
  const int sym = pot.sw_multi ? 0 : 1;
  const int lead = gather_lead(n_struct);
  if (pot.sw_multi) {
    constexpr int LA = L < kRowLanes ? L : kRowLanes;
    const int atom_blocks = (int)(((int64_t)n * LA + 255) / 256);
    hipLaunchKernelGGL((k_gather_c<L, LA>), dim3(atom_blocks + lead), dim3(256), 0, st, n, cnt,
                       rev, nl.stride, (const double*)G, d_F, atom_blocks, n_struct, geoms,
                       first, n_chunks, (const double*)part, d_out, (const uint64_t*)amask, sym);
  } else if constexpr (L >= kLongRow) {
    const int atom_blocks = (int)(((int64_t)n * kRowLanes + 255) / 256);
    hipLaunchKernelGGL((k_gather_c<L, kRowLanes>), dim3(atom_blocks + lead), dim3(256), 0, st,
                       n, cnt, rev, nl.stride, (const double*)G, d_F, atom_blocks, n_struct,
                       geoms, first, n_chunks, (const double*)part, d_out,
                       (const uint64_t*)amask, sym);
  } else {
    const int atom_blocks = (int)(((int64_t)n * L + 255) / 256);
    if (L == 4 && n_struct == 1)
      hipLaunchKernelGGL(k_gather_one4, dim3(atom_blocks + lead), dim3(256), 0, st, n,
                         atom_blocks, lead, n_chunks, cnt, rev, (const double*)G, d_F, geoms,
                         first, (const double*)part, d_out);
    else
      hipLaunchKernelGGL(k_gather<L>, dim3(atom_blocks + lead), dim3(256), 0, st, n, cnt, rev,
                         nl.stride, (const double*)G, d_F, atom_blocks, n_struct, geoms, first,
                         n_chunks, (const double*)part, d_out, (const uint64_t*)amask, sym);
  }
}

template <int LE>
static void launch_eam(Potential& pot, NList& nl, const ForceWork& w, const double* d_pos,
                       const int* d_spc, double* d_F, double* esite, hipStream_t st) {
  const int n = w.n;
  const int64_t threads = (int64_t)n * LE;
  const unsigned grid = (unsigned)((threads + 255) / 256);
  double* dF = w.aux;
  double* Fe = dF + n;
  const int single = w.n_struct == 1 ? 1 : 0;
  if (pot.eam_fast) {
    // Zhou alloy table: one block per chunk for both passes (the grid cap
    // does not apply)
    const dim3 cg(w.n_chunks > 0 ? w.n_chunks : 1), blk(pot.block);
    const double* fast = pot.d_eam_fast.as<double>();
    hipLaunchKernelGGL(k_eam_rho_z<LE>, cg, blk, 0, st, w.n_chunks, n, d_pos, w.sid, w.geoms,
                       single, d_spc, pot.ns, (const double*)pot.d_params.as<double>(),
                       (const int*)pot.d_map.as<int>(), fast, pot.r_cut,
                       (const int*)nl.cnt.as<int>(), (const int*)nl.nbr.as<int>(), nl.stride,
                       w.rng, dF, Fe);
    launch_timed(pot, k_eam_force_z<LE>, cg, blk, st, (size_t)0, w.n_chunks, n, d_pos, w.sid,
                 w.geoms, single, d_spc, pot.ns, fast, pot.r_cut, (const int*)nl.cnt.as<int>(),
                 (const int*)nl.nbr.as<int>(), nl.stride, (const double*)dF, (const double*)Fe,
                 w.rng, d_F, w.partial, esite, w.pair_flag);
    return;
  }
  hipLaunchKernelGGL(k_eam_rho<LE>, dim3(grid), dim3(256), 0, st, n, d_pos, w.sid, w.geoms,
                     single, d_spc, pot.ns, pot.d_params.as<double>(), pot.d_map.as<int>(),
                     pot.r_cut, nl.cnt.as<int>(), nl.nbr.as<int>(), nl.stride, dF, Fe);
  launch_timed(pot, k_eam_force<LE>, chunk_grid(pot, w.n_chunks), dim3(pot.block), st,
               (size_t)0, w.n_chunks, d_pos, w.sid, w.geoms, single, d_spc, pot.ns,
               (const double*)pot.d_params.as<double>(), (const int*)pot.d_map.as<int>(),
               pot.r_cut, (const int*)nl.cnt.as<int>(), (const int*)nl.nbr.as<int>(), nl.stride,
               (const double*)dF, (const double*)Fe, w.rng, d_F, w.partial, esite, w.pair_flag);
}

bool eam_fast_table(int ns, const std::vector<double>& rows, const std::vector<int>& map,
                    std::vector<double>& out) {
  auto row = [&](int k) { return rows.data() + (size_t)ZW * k; };
  auto same = [](const double* x, const double* y) {
    return std::memcmp(x, y, 20 * sizeof(double)) == 0;
  };
  out.assign((size_t)ZF_SP * ns + (size_t)ZF_PR * ns * ns + 1, 0.0);
  for (int s = 0; s < ns; ++s) {
    const double* d = row(map[s]);
    if (d[0] != 3.0) return false;
    const double* pa = d + 4;
    double* o = out.data() + (size_t)ZF_SP * s;
    o[0] = pa[0];  // r_e
    o[1] = pa[1];  // f_e
    o[2] = pa[5];  // beta
    o[3] = pa[9];  // lambda
    o[4] = pa[6];  // A
    o[5] = pa[4];  // alpha
    o[6] = pa[8];  // kappa
    o[7] = pa[7];  // B
    o[8] = d[1];
    o[9] = d[2];
    o[10] = 1.0 / pa[0];
    o[11] = 1.0 / d[2];
  }
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b < ns; ++b) {
      const double* f = row(map[2 * ns + a * ns + b]);
      const double* ea = row(map[a]) + 4;
      const double* eb = row(map[b]) + 4;
      double* o = out.data() + (size_t)ZF_SP * ns + (size_t)ZF_PR * (a * ns + b);
      o[0] = f[0];
      o[1] = f[1];
      o[2] = f[2];
      o[3] = f[3];
      o[5] = 1.0 / f[2];
      if (f[0] == 4.0) {
        if (a != b || !same(f + 4, ea)) return false;
      } else if (f[0] == 5.0) {
        if (a == b) return false;
        if (same(f + 4, ea) && same(f + 24, eb))
          o[4] = 0.0;
        else if (same(f + 4, eb) && same(f + 24, ea))
          o[4] = 1.0;
        else
          return false;
      } else {
        return false;
      }
    }
  // one switch: every density and pair row has the same r_c and width
  bool one = true;
  const double rc0 = out[8], w0 = out[9];
  for (int a = 0; a < ns; ++a) {
    const double* o = out.data() + (size_t)ZF_SP * a;
    one = one && o[8] == rc0 && o[9] == w0;
    for (int b = 0; b < ns; ++b) {
      const double* q = out.data() + (size_t)ZF_SP * ns + (size_t)ZF_PR * (a * ns + b);
      one = one && q[1] == rc0 && q[2] == w0;
    }
  }
  out.back() = one ? 1.0 : 0.0;
  return true;
}

ForcePlan force_plan(const Potential& pot, const NList& nl, bool esite) {
  AK_REQUIRE(nl.valid, AK_ERR_STATE, "neighbour list has not been built");
  AK_REQUIRE(pot.block == 64 || pot.block == 128 || pot.block == 256 || pot.block == 512,
             AK_ERR_STATE, "block size must be 64, 128, 256 or 512 threads");
  ForcePlan pl;
  if (pot.kind == KIND_EAM) {
    pl.lanes = pot.eam_lanes > 0 ? pot.eam_lanes : 16;
  } else {
    AK_REQUIRE(nl.stride <= 64, AK_ERR_CAPACITY,
               "more than 64 neighbours within r_cut + skin for a bond-order kernel");
    pl.lanes = nl.stride;
  }
  const int lanes = pl.lanes;
  // lanes per atom of the site kernel: L, or L * L for the Tersoff form with
  // one lane per (bond, partner) pair
  // per-atom site energies come from the lane-per-bond Tersoff form
  pl.trip = pot.kind == KIND_TERSOFF && lanes * lanes <= 64 && pot.bond_tile == 2 && !esite;
  const bool swm = pot.kind == KIND_SW && pot.sw_multi;
  const bool sw_long = pot.kind == KIND_SW && !swm && lanes >= kLongRow;
  const int atom_lanes = pl.trip ? lanes * lanes
                                 : (sw_long || (swm && lanes > kRowLanes) ? kRowLanes : lanes);
  int blk = pot.block;
  // 512 threads per block for the four-slot Tersoff kernel only (every other
  // kernel is built for at most 256)
  if (blk == 512 && !q4_path(pot, lanes, pl.trip)) blk = 256;
  // the long-row and multi-species SW kernels stage every leg of their atoms
  // in LDS: keep it within 48 KiB per block
  while (sw_long && blk > 64 && sw_c_lds(blk, lanes) > 48 * 1024) blk >>= 1;
  while (swm && blk > 64 && sw_m_lds(blk, lanes, atom_lanes, pot.ns) > 48 * 1024) blk >>= 1;
  AK_REQUIRE(atom_lanes <= blk, AK_ERR_STATE, "block smaller than one atom's lanes");
  pl.blk = blk;
  pl.chunk = blk / atom_lanes;
  return pl;
}

void compute_forces_work(Potential& pot, NList& nl, const ForcePlan& pl, const ForceWork& w,
                         const double* d_pos, const int* d_spc, double* d_F, double* d_out,
                         hipStream_t st, double* esite) {
  const int n = w.n;
  AK_REQUIRE(nl.n_atoms == n, AK_ERR_STATE, "neighbour list was built for another size");
  if (n > 0) {
    if (pot.kind == KIND_EAM) {
      switch (pl.lanes) {
        case 4: launch_eam<4>(pot, nl, w, d_pos, d_spc, d_F, esite, st); break;
        case 8: launch_eam<8>(pot, nl, w, d_pos, d_spc, d_F, esite, st); break;
        case 32: launch_eam<32>(pot, nl, w, d_pos, d_spc, d_F, esite, st); break;
        case 64: launch_eam<64>(pot, nl, w, d_pos, d_spc, d_F, esite, st); break;
        default: launch_eam<16>(pot, nl, w, d_pos, d_spc, d_F, esite, st); break;
      }
    } else {
      switch (pl.lanes) {
        case 4: launch_bond_kernels<4>(pot, nl, w, d_pos, d_spc, d_F, d_out, pl.trip, pl.blk, esite, st); break;
        case 8: launch_bond_kernels<8>(pot, nl, w, d_pos, d_spc, d_F, d_out, pl.trip, pl.blk, esite, st); break;
        case 16: launch_bond_kernels<16>(pot, nl, w, d_pos, d_spc, d_F, d_out, pl.trip, pl.blk, esite, st); break;
        case 32: launch_bond_kernels<32>(pot, nl, w, d_pos, d_spc, d_F, d_out, pl.trip, pl.blk, esite, st); break;
        case 64: launch_bond_kernels<64>(pot, nl, w, d_pos, d_spc, d_F, d_out, pl.trip, pl.blk, esite, st); break;
        default: throw Error(AK_ERR_STATE, "unsupported neighbour stride");
      }
    }
  }
  if (pot.kind == KIND_EAM || n == 0)
    hipLaunchKernelGGL(k_struct_final, dim3(w.n_struct), dim3(1024), 0, st, w.geoms, w.first,
                       w.n_chunks, (const double*)w.partial, d_out,
                       pot.kind == KIND_EAM && n > 0 ? (const int*)w.pair_flag
                                                     : (const int*)nullptr);
}

void compute_forces(Potential& pot, NList& nl, SetDev& set, const double* d_pos,
                    const int* d_spc, double* d_F, double* d_out, hipStream_t st,
                    double* esite) {
  const int n = set.n_atoms;
  if (pot.kind != KIND_EAM && nl.valid && nl.stride > kWave) {
    // rows longer than a wavefront (a wide skin, a compressed cell): the bond
    // kernels run on a list of the pairs inside r_cut alone, built from the
    // same positions (the result depends only on those pairs).  Site
    // gradients are defined on the caller's list, so they are refused.
    AK_REQUIRE(!esite, AK_ERR_CAPACITY,
               "site terms and heat fluxes need at most 64 neighbours within r_cut + skin");
    AK_REQUIRE(nl.n_atoms == n, AK_ERR_STATE, "neighbour list was built for another size");
    if (!pot.tight) pot.tight = std::make_unique<NList>();
    NList& t = *pot.tight;
    t.r_cut = t.r_list = pot.r_cut;
    t.skin = 0.0;
    t.valid = false;
    nl_build(t, set, d_pos, st);
    AK_REQUIRE(t.stride <= kWave, AK_ERR_CAPACITY,
               "more than 64 neighbours within r_cut for a bond-order kernel");
    compute_forces(pot, t, set, d_pos, d_spc, d_F, d_out, st, nullptr);
    return;
  }
  const ForcePlan pl = force_plan(pot, nl, esite != nullptr);
  AK_REQUIRE(nl.n_atoms == n, AK_ERR_STATE, "neighbour list was built for another size");
  SetDev::Chunks& ch = set.chunks(pl.chunk, st);
  ForceWork w;
  w.n = n;
  w.n_struct = set.n_struct;
  w.sid = set.d_sid.as<int>();
  w.geoms = set.d_geoms.as<Geom>();
  w.rng = ch.rng.as<int>();
  w.first = ch.first.as<int>();
  w.partial = ch.partial.as<double>();
  w.n_chunks = ch.n;
  // every chunk writes its flag in each call (no clearing needed)
  if (pot.kind == KIND_EAM) {
    pot.pair_flag.ensure(sizeof(int) * (size_t)std::max(1, ch.n));
    w.pair_flag = pot.pair_flag.as<int>();
    if (n > 0) {
      pot.aux.ensure(sizeof(double) * 2 * (size_t)n);
      w.aux = pot.aux.as<double>();
    }
  } else if (n > 0) {
    pot.G.ensure(sizeof(double) * kGRec * (size_t)n * nl.stride);
    pot.amask.ensure(sizeof(uint64_t) * (size_t)n);
    w.G = pot.G.as<double>();
    w.amask = pot.amask.as<uint64_t>();
  }
  compute_forces_work(pot, nl, pl, w, d_pos, d_spc, d_F, d_out, st, esite);
}

// ---------------------------------------------------------------------------
// Site terms and the heat flux (engine/site_terms.py:1-60, 179-475;
// solvers/thermal_conductivity.py:168-303).  With E = sum_a U_a and U_a a
// function of the bond vectors v_ak = r_nbr(a,k) - r_a, the bond kernels'
// G[a, k] = dU_a / dv_ak is the reference's dUdv of the entry (a -> nbr).
// EAM: dU_a / dv_ak = (F'_a rho'_{s_j}(r) + phi'_aj(r) / 2) v / r.
// ---------------------------------------------------------------------------
template <int LE>
__global__ __launch_bounds__(256) void k_eam_site_g(
    int n, const double* __restrict__ pos, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, const int* __restrict__ spc, int ns,
    const double* __restrict__ fn, const int* __restrict__ map, double rc,
    const int* __restrict__ cnt, const int* __restrict__ nbr, int stride,
    const double* __restrict__ dF, double* __restrict__ G) {
  const int gt = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int a = gt / LE;
  const int l = gt & (LE - 1);
  if (a >= n) return;
  const Geom& g = geoms[single ? 0 : sid[a]];
  const int nk = cnt[a];
  const int sa = ns > 1 ? spc[a] : 0;
  const double dFa = dF[a];
  for (int k = l; k < nk; k += LE) {
    const size_t slot = (size_t)a * stride + k;
    const int j = nbr[slot];
    double ux, uy, uz;
    mic_fast(g, pos, a, j, ux, uy, uz);
    const double r = sqrt(ux * ux + uy * uy + uz * uz);
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (r < rc && r > 1e-12) {
      const int sj = ns > 1 ? spc[j] : 0;
      double v, dfj, phi, dphi;
      z_feval(fn, fn + (size_t)ZW * map[sj], r, v, dfj);
      z_feval(fn, fn + (size_t)ZW * map[2 * ns + sa * ns + sj], r, phi, dphi);
      const double s = (dFa * dfj + 0.5 * dphi) / r;
      gx = s * ux;
      gy = s * uy;
      gz = s * uz;
    }
    g_put(G, slot, gx, gy, gz);
  }
}

// Record frame of the heat-flux gate: with gate, the step counters[2] of
// this call records when (step + phase) % every == 0 (phase in [0, every)),
// into frame (step + phase) / every - (phase > 0); -1 = no record.  Without
// gate, frame 0.
__device__ __forceinline__ int flux_frame(const int* __restrict__ gate, int every, int phase) {
  if (!gate) return 0;
  const int t = gate[2] + phase;
  if (t % every != 0) return -1;
  return t / every - (phase > 0 ? 1 : 0);
}

// Per-block sums of V J = sum_a E_a v_a - sum_(a,k) v_ak (G[a, k] . vel_nbr),
// E_a = m_a |v_a|^2 / (2 conv) + U_a (Fan et al., PRB 92, 094301 (2015),
// eq. 21, per neighbour entry).  vscale (nullable) multiplies every
// velocity; masked: only slots whose amask bit is set hold G.  With
// gate != nullptr the kernel runs only when counters[2] % every == 0.
template <int LE>
__global__ __launch_bounds__(256) void k_heat_flux(
    int n, const double* __restrict__ pos, const double* __restrict__ vel,
    const double* __restrict__ mass, const double* __restrict__ esite,
    const double* __restrict__ vscale, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, double rc, const int* __restrict__ cnt,
    const int* __restrict__ nbr, int stride, const double* __restrict__ G,
    const uint64_t* __restrict__ amask, int masked, double conv,
    const int* __restrict__ gate, int every, int phase, double* __restrict__ part) {
  if (flux_frame(gate, every, phase) < 0) return;
  __shared__ double smem[16 * 3];
  const int a = (int)(blockIdx.x * blockDim.x + threadIdx.x) / LE;
  const int l = threadIdx.x & (LE - 1);
  double J[3] = {0.0, 0.0, 0.0};
  if (a < n) {
    const double sc = vscale ? vscale[0] : 1.0;
    const Geom& g = geoms[single ? 0 : sid[a]];
    const int nk = cnt[a];
    const uint64_t ma = masked ? amask[a] : ~uint64_t(0);
    for (int k = l; k < nk; k += LE) {
      if (masked && !((ma >> k) & 1)) continue;
      const size_t slot = (size_t)a * stride + k;
      const int j = nbr[slot];
      double ux, uy, uz;
      mic_fast(g, pos, a, j, ux, uy, uz);
      const double r2 = ux * ux + uy * uy + uz * uz;
      if (!(r2 < rc * rc) || !(r2 > 1e-24)) continue;
      double gx, gy, gz;
      g_get(G, slot, gx, gy, gz);
      const double gv = (gx * vel[3 * (size_t)j] + gy * vel[3 * (size_t)j + 1] +
                         gz * vel[3 * (size_t)j + 2]) *
                        sc;
      J[0] -= ux * gv;
      J[1] -= uy * gv;
      J[2] -= uz * gv;
    }
    if (l == 0) {
      const double vx = vel[3 * (size_t)a] * sc, vy = vel[3 * (size_t)a + 1] * sc,
                   vz = vel[3 * (size_t)a + 2] * sc;
      const double E = 0.5 * mass[a] * (vx * vx + vy * vy + vz * vz) / conv + esite[a];
      J[0] += E * vx;
      J[1] += E * vy;
      J[2] += E * vz;
    }
  }
  block_sum<3>(J, smem);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = J[0];
    part[3 * blockIdx.x + 1] = J[1];
    part[3 * blockIdx.x + 2] = J[2];
  }
}

// J = (sum of the block sums, fixed order) / V -> out[0..2]; with gate,
// into out[3 * (counters[2] / every)] on record steps, and counters[2]
// advances every call.
__global__ __launch_bounds__(256) void k_heat_flux_final(int nb, const double* __restrict__ part,
                                                         double inv_volume, int* __restrict__ gate,
                                                         int every, int phase,
                                                         double* __restrict__ out) {
  __shared__ double smem[16 * 3];
  const int f = flux_frame(gate, every, phase);
  double J[3] = {0.0, 0.0, 0.0};
  const int nt = blockDim.x;
  if (f >= 0)
    for (int b = threadIdx.x; b < nb; b += 4 * nt) {  // four blocks' loads in flight
      double x[4][3];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < 3; ++c) x[u][c] = b + u * nt < nb ? part[3 * (b + u * nt) + c] : 0.0;
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[c] += x[u][c];
    }
  block_sum<3>(J, smem);
  if (threadIdx.x == 0) {
    if (f >= 0) {
      double* o = out + 3 * (size_t)f;
      o[0] = J[0] * inv_volume;
      o[1] = J[1] * inv_volume;
      o[2] = J[2] * inv_volume;
    }
    if (gate) gate[2] += 1;
  }
}

// Per-atom heat currents summed per group (ThermalConductivityCalculator.
// _per_atom_heat_current and _group_labels, solvers/thermal_conductivity.py:
// 305-363).  Atom i carries E_i v_i and, for each neighbour entry (j -> i),
// -vec_ji (dU_j/dvec_ji . v_i) = vec_ij (G[rev(i, k)] . v_i): the term goes to
// the atom whose velocity it holds.  The group of i is labels[i] (or 0) times
// n_reg plus its region: the fractional coordinate along `axis`, wrapped,
// placed between bounds[0..n_reg] (searchsorted right - 1, clipped), or
// with bounds == nullptr in n_reg equal bins, min(floor(f n_reg), n_reg - 1).
// Per block, one sum per group: part[(block * ng + g) * NC + c], NC = 3 (J)
// or 5 (J, the kinetic energy, the atom count).
template <int LE, int NC>
__global__ __launch_bounds__(256) void k_heat_flux_groups(
    int n, const double* __restrict__ pos, const double* __restrict__ vel,
    const double* __restrict__ mass, const double* __restrict__ esite,
    const double* __restrict__ vscale, const int* __restrict__ sid,
    const Geom* __restrict__ geoms, int single, double rc, const int* __restrict__ cnt,
    const int* __restrict__ nbr, const int* __restrict__ rev, int stride,
    const double* __restrict__ G, const uint64_t* __restrict__ amask, int masked, double conv,
    const int* __restrict__ gate, int every, int phase, const int* __restrict__ labels,
    int axis, const double* __restrict__ bounds, int n_reg, int ng, double* __restrict__ part,
    float* __restrict__ atoms, double inv_volume) {
  const int frame = flux_frame(gate, every, phase);
  if (frame < 0) return;
  const int a = (int)(blockIdx.x * blockDim.x + threadIdx.x) / LE;
  const int l = threadIdx.x & (LE - 1);
  double J[3] = {0.0, 0.0, 0.0};
  double ke = 0.0;
  int grp = -1;
  if (a < n) {
    const double sc = vscale ? vscale[0] : 1.0;
    const Geom& g = geoms[single ? 0 : sid[a]];
    const int nk = cnt[a];
    const double vx = vel[3 * (size_t)a] * sc, vy = vel[3 * (size_t)a + 1] * sc,
                 vz = vel[3 * (size_t)a + 2] * sc;
    for (int k = l; k < nk; k += LE) {
      const size_t slot = (size_t)a * stride + k;
      const int r = rev[slot];
      if (r < 0) continue;
      if (masked && !((amask[r / stride] >> (r % stride)) & 1)) continue;
      double ux, uy, uz;
      mic_fast(g, pos, a, nbr[slot], ux, uy, uz);
      const double r2 = ux * ux + uy * uy + uz * uz;
      if (!(r2 < rc * rc) || !(r2 > 1e-24)) continue;
      double gx, gy, gz;
      g_get(G, (size_t)r, gx, gy, gz);
      const double gv = gx * vx + gy * vy + gz * vz;
      J[0] += ux * gv;
      J[1] += uy * gv;
      J[2] += uz * gv;
    }
    if (l == 0) {
      ke = 0.5 * mass[a] * (vx * vx + vy * vy + vz * vz) / conv;
      const double E = ke + esite[a];
      J[0] += E * vx;
      J[1] += E * vy;
      J[2] += E * vz;
    }
    int rg = 0;
    if (axis >= 0) {
      double f = pos[3 * (size_t)a] * g.inv[axis] + pos[3 * (size_t)a + 1] * g.inv[3 + axis] +
                 pos[3 * (size_t)a + 2] * g.inv[6 + axis];
      f -= floor(f);
      if (bounds) {
        for (int k = 1; k < n_reg; ++k)
          if (bounds[k] <= f) rg = k;
      } else {
        rg = min((int)(f * n_reg), n_reg - 1);
      }
    }
    grp = (labels ? labels[a] : 0) * n_reg + rg;
  }
  for (int c = 0; c < 3; ++c) J[c] = seg_sum<LE>(J[c]);
  if (atoms && a < n && l == 0)
    for (int c = 0; c < 3; ++c)
      atoms[((size_t)frame * n + a) * 3 + c] = (float)(J[c] * inv_volume);
  // the block's atoms (blockDim / LE of them, usually in one or two groups)
  // leave their records in LDS; thread q sums the records of group q in atom
  // order and writes the block's partial of every group
  constexpr int kApb = 256 / LE;
  __shared__ double rec[kApb][NC];
  __shared__ int rgrp[kApb];
  const int al = (int)threadIdx.x / LE;
  if (l == 0) {
    rgrp[al] = grp;
    rec[al][0] = J[0];
    rec[al][1] = J[1];
    rec[al][2] = J[2];
    if constexpr (NC == 5) {
      rec[al][3] = ke;
      rec[al][4] = 1.0;
    }
  }
  __syncthreads();
  const int apb = (int)blockDim.x / LE;
  for (int q = threadIdx.x; q < ng; q += blockDim.x) {
    double w[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) w[c] = 0.0;
    for (int t = 0; t < apb; ++t)
      if (rgrp[t] == q)
#pragma unroll
        for (int c = 0; c < NC; ++c) w[c] += rec[t][c];
#pragma unroll
    for (int c = 0; c < NC; ++c) part[((size_t)blockIdx.x * ng + q) * NC + c] = w[c];
  }
}

// Group sums of a record step: (sum of the block sums, fixed order) into
// out[(frame * ng + g) * NC + c], the flux components divided by V; the gate
// is read, not advanced.
// One block per group q: each thread sums its blocks' partials in block
// order, four blocks' loads in flight at a time.
template <int NC>
__global__ __launch_bounds__(256) void k_heat_flux_groups_final(
    int nb, int ng, const double* __restrict__ part, double inv_volume,
    const int* __restrict__ gate, int every, int phase, double* __restrict__ out) {
  __shared__ double smem[16 * NC];
  const int frame = flux_frame(gate, every, phase);
  if (frame < 0) return;
  const size_t f = (size_t)frame;
  const int q = blockIdx.x;
  const int nt = blockDim.x;
  double J[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) J[c] = 0.0;
  for (int b = threadIdx.x; b < nb; b += 4 * nt) {
    double x[4][NC];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < NC; ++c)
        x[u][c] = b + u * nt < nb ? part[((size_t)(b + u * nt) * ng + q) * NC + c] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < NC; ++c) J[c] += x[u][c];
  }
  block_sum<NC>(J, smem);
  if (threadIdx.x == 0)
    for (int c = 0; c < NC; ++c) out[(f * ng + q) * NC + c] = c < 3 ? J[c] * inv_volume : J[c];
}

bool site_masked(const Potential& pot, const NList& nl) {
  return pot.kind != KIND_EAM && (nl.stride >= 8 || pot.sw_multi);
}

void site_grads(Potential& pot, NList& nl, SetDev& set, const double* d_pos, const int* d_spc,
                double* d_F, double* d_out, double* d_esite, hipStream_t st) {
  compute_forces(pot, nl, set, d_pos, d_spc, d_F, d_out, st, d_esite);
  if (pot.kind != KIND_EAM || set.n_atoms == 0) return;
  const int n = set.n_atoms;
  pot.G.ensure(sizeof(double) * kGRec * (size_t)n * nl.stride);
  constexpr int LE = 8;
  const unsigned grid = (unsigned)(((int64_t)n * LE + 255) / 256);
  hipLaunchKernelGGL(k_eam_site_g<LE>, dim3(grid), dim3(256), 0, st, n, d_pos,
                     (const int*)set.d_sid.as<int>(), (const Geom*)set.d_geoms.as<Geom>(),
                     set.n_struct == 1 ? 1 : 0, d_spc, pot.ns,
                     (const double*)pot.d_params.as<double>(), (const int*)pot.d_map.as<int>(),
                     pot.r_cut, (const int*)nl.cnt.as<int>(), (const int*)nl.nbr.as<int>(),
                     nl.stride, (const double*)pot.aux.as<double>(), pot.G.as<double>());
}

void heat_flux(Potential& pot, NList& nl, SetDev& set, const double* d_pos, const double* d_vel,
               const double* d_mass, const double* d_esite, const double* d_vscale,
               double conv, int* d_gate, int every, double* d_part, double* d_out,
               hipStream_t st, const FluxGroups* groups, int phase) {
  const int n = set.n_atoms;
  constexpr int LE = 8;
  const unsigned grid = (unsigned)(((int64_t)n * LE + 255) / 256);
  const double inv_v = 1.0 / set.geoms[0].volume;
  if (groups && groups->ng > 0) {
    auto launch = [&](auto nc) {
      constexpr int NC = decltype(nc)::value;
      hipLaunchKernelGGL((k_heat_flux_groups<LE, NC>), dim3(grid), dim3(256), 0, st, n, d_pos,
                         d_vel, d_mass, d_esite, d_vscale, (const int*)set.d_sid.as<int>(),
                         (const Geom*)set.d_geoms.as<Geom>(), set.n_struct == 1 ? 1 : 0,
                         pot.r_cut, (const int*)nl.cnt.as<int>(), (const int*)nl.nbr.as<int>(),
                         (const int*)nl.rev.as<int>(), nl.stride,
                         (const double*)pot.G.as<double>(),
                         (const uint64_t*)pot.amask.as<uint64_t>(),
                         site_masked(pot, nl) ? 1 : 0, conv, (const int*)d_gate, every, phase,
                         groups->labels, groups->axis, groups->bounds, groups->n_reg,
                         groups->ng, groups->part, groups->atoms, inv_v);
      hipLaunchKernelGGL((k_heat_flux_groups_final<NC>), dim3(groups->ng), dim3(256), 0, st,
                         (int)grid,
                         groups->ng, (const double*)groups->part, inv_v, (const int*)d_gate,
                         every, phase, groups->out);
    };
    if (groups->nc == 5)
      launch(std::integral_constant<int, 5>{});
    else
      launch(std::integral_constant<int, 3>{});
  }
  hipLaunchKernelGGL(k_heat_flux<LE>, dim3(grid), dim3(256), 0, st, n, d_pos, d_vel, d_mass,
                     d_esite, d_vscale, (const int*)set.d_sid.as<int>(),
                     (const Geom*)set.d_geoms.as<Geom>(), set.n_struct == 1 ? 1 : 0, pot.r_cut,
                     (const int*)nl.cnt.as<int>(), (const int*)nl.nbr.as<int>(), nl.stride,
                     (const double*)pot.G.as<double>(), (const uint64_t*)pot.amask.as<uint64_t>(),
                     site_masked(pot, nl) ? 1 : 0, conv, (const int*)d_gate, every, phase, d_part);
  hipLaunchKernelGGL(k_heat_flux_final, dim3(1), dim3(256), 0, st, (int)grid, (const double*)d_part,
                     inv_v, d_gate, every, phase, d_out);
}

}  // namespace ak
