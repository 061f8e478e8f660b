// C ABI of the engine (include/akitsushima_amd.h): handles, argument checks,
// host<->device staging, and the owned-list policy of the reference's force
// functions.
#include "../../include/akitsushima_amd.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

using namespace ak;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_err.clear();
    return AK_OK;
  } catch (const Error& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(AK_ERR_CAPACITY, "host out of memory");
  } catch (const std::exception& e) {
    return fail(AK_ERR_STATE, e.what());
  }
}

// Smallest singular value of a 3x3 matrix (eigenvalues of M^T M by Jacobi).
double sigma_min3(const double* M) {
  double A[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += M[3 * k + i] * M[3 * k + j];
      A[i][j] = s;
    }
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = std::fabs(A[0][1]) + std::fabs(A[0][2]) + std::fabs(A[1][2]);
    if (off < 1e-300) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (A[p][q] == 0.0) continue;
        double th = 0.5 * (A[q][q] - A[p][p]) / A[p][q];
        double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
      }
  }
  double ev = std::min(A[0][0], std::min(A[1][1], A[2][2]));
  return std::sqrt(std::max(ev, 0.0));
}

void matmul3(const double* a, const double* b, double* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
      c[3 * i + j] = s;
    }
}

bool lattice_allclose(const double* a, const double* b) {
  // np.allclose(lattice, built, atol=1e-12) with its default rtol 1e-5
  for (int k = 0; k < 9; ++k)
    if (!(std::fabs(a[k] - b[k]) <= 1e-12 + 1e-5 * std::fabs(b[k]))) return false;
  return true;
}

void check_finite(const double* x, size_t n, const char* what) {
  bool ok = true;  // a branch-free pass (vectorised)
  for (size_t k = 0; k < n; ++k) ok &= std::isfinite(x[k]);
  AK_REQUIRE(ok, AK_ERR_ARG, std::string(what) + " contains non-finite values");
}

}  // namespace

struct ak_nlist {
  NList nl;
  SetDev set;
  hipStream_t stream = nullptr;
  DevBuf d_pos, d_dmax;
  int64_t n_steps = 0;
  ~ak_nlist() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct ak_potential {
  Potential pot;
  hipStream_t stream = nullptr;
  // owned list (engine/neighborlist.py:1103-1163) and its structure
  ak_nlist owned;
  double owned_skin = -1.0;
  // batched layout
  ak_nlist batch;
  std::vector<int64_t> batch_off;
  // staging
  DevBuf d_pos, d_spc, d_F, d_out;
  std::vector<int32_t> last_spc;
  bool spc_valid = false;
  // pinned host staging of the drop-in call: positions in, forces and the
  // eleven result values out (DMA from and to pinned memory instead of the
  // runtime's pageable path)
  double* h_stage = nullptr;
  size_t h_stage_n = 0;
  double* pinned_stage(size_t n_doubles) {
    if (n_doubles > h_stage_n) {
      if (h_stage) AK_HIP(hipHostFree(h_stage));
      h_stage = nullptr;
      h_stage_n = 0;
      AK_HIP(hipHostMalloc((void**)&h_stage, sizeof(double) * n_doubles, hipHostMallocDefault));
      h_stage_n = n_doubles;
    }
    return h_stage;
  }
  ~ak_potential() {
    if (h_stage) (void)hipHostFree(h_stage);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static hipStream_t make_stream() {
  hipStream_t s;
  AK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

static void require_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
    throw Error(AK_ERR_NO_DEVICE, "no HIP device is visible: the akitsushima_amd force "
                                  "engine runs on the GPU only");
}

// Build or keep `L` valid for positions already on the device
// (NeighbourList.check_and_rebuild, engine/neighborlist.py:613-673).
static bool check_and_rebuild(ak_nlist& L, int n, const double* lattice, const double* inv,
                              const double* d_pos, hipStream_t st) {
  L.n_steps++;
  L.nl.n_checks++;
  const int64_t off[2] = {0, (int64_t)n};
  L.set.setup(1, off, lattice, inv, L.nl.r_list, st);
  if (!L.nl.valid || L.nl.n_atoms != n) {
    nl_build(L.nl, L.set, d_pos, st);
    return true;
  }
  const double* Lb = L.nl.built_geoms[0].lat;
  const double* Ib = L.nl.built_geoms[0].inv;
  double sig = 1.0;
  Mat3 M;
  matmul3(Ib, lattice, M.m);
  if (!lattice_allclose(lattice, Lb)) sig = sigma_min3(M.m);
  L.d_dmax.ensure(sizeof(double));
  nl_check_displacement(L.nl, L.set, d_pos, M, sig, nullptr, L.d_dmax.as<double>(), st);
  double dmax = 0.0;
  AK_HIP(hipMemcpyAsync(&dmax, L.d_dmax.p, sizeof(double), hipMemcpyDeviceToHost, st));
  AK_HIP(hipStreamSynchronize(st));
  if (sig * L.nl.r_list - 2.0 * dmax <= L.nl.r_cut) {
    nl_build(L.nl, L.set, d_pos, st);
    return true;
  }
  return false;
}

static void stage_species(ak_potential* P, int n, const int32_t* species) {
  // unchanged species (the common repeated call) need no copy at all
  if (P->spc_valid && (int)P->last_spc.size() == n &&
      (species ? std::equal(species, species + n, P->last_spc.begin())
               : std::all_of(P->last_spc.begin(), P->last_spc.end(),
                             [](int32_t v) { return v == 0; })))
    return;
  std::vector<int32_t> spc(species ? species : nullptr, species ? species + n : nullptr);
  if (!species) spc.assign(n, 0);
  for (int a = 0; a < n; ++a)
    AK_REQUIRE(spc[a] >= 0 && spc[a] < P->pot.ns, AK_ERR_ARG,
               "species index out of range for this potential");
  P->d_spc.ensure(sizeof(int32_t) * std::max(1, n));
  if (n > 0)
    AK_HIP(hipMemcpyAsync(P->d_spc.p, spc.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                          P->stream));
  AK_HIP(hipStreamSynchronize(P->stream));
  P->last_spc.swap(spc);
  P->spc_valid = true;
}

// The list a single-structure call uses, positions already in P->d_pos: the
// caller's (as built or rebuilt by the caller -- the reference's potentials
// use a passed list's candidates without rebuilding it; the current cell is
// used for the minimum image), or the potential's owned list (skin 0.3 A
// when the cell allows it, else none), checked and rebuilt.
static ak_nlist* select_list(ak_potential* P, ak_nlist* L, int n, const double* lattice,
                             const double* inv, hipStream_t st) {
  if (L) {
    AK_REQUIRE(L->nl.r_cut + 1e-12 >= P->pot.r_cut, AK_ERR_ARG,
               "neighbour list r_cut is smaller than the potential's cutoff");
    const int64_t off[2] = {0, (int64_t)n};
    L->set.setup(1, off, lattice, inv, L->nl.r_list, st);
    AK_REQUIRE(L->nl.valid && L->nl.n_atoms == n, AK_ERR_STATE,
               "the neighbour list was not built for this structure");
    return L;
  }
  L = &P->owned;
  CellInfo ci = cell_info(lattice);
  double skin = (P->pot.r_cut + 0.3 < ci.min_width / 2.0) ? 0.3 : 0.0;
  if (skin != P->owned_skin) {
    P->owned.nl.valid = false;
    P->owned.nl.r_cut = P->pot.r_cut;
    P->owned.nl.skin = skin;
    P->owned.nl.r_list = P->pot.r_cut + skin;
    P->owned_skin = skin;
  }
  check_and_rebuild(*L, n, lattice, inv, P->d_pos.as<double>(), st);
  return L;
}

// The owned list's rebuild test without a separate round trip: the force
// call runs on the current list right behind the displacement check, and
// the largest displacement comes back with the results (P->d_out[10],
// copied with the energy and stress).  Only when it shows
// that the list needed a rebuild (then the results are discarded) does the
// caller rebuild and run again; returns false in that case and when the list
// must be built first (not yet built, new size, new skin), true when `run`
// has produced the results (the stream synchronised).
template <class Run>
static bool owned_list_speculative(ak_potential* P, int n, const double* lattice,
                                   const double* inv, hipStream_t st, Run&& run) {
  ak_nlist& L = P->owned;
  const CellInfo ci = cell_info(lattice);
  const double skin = (P->pot.r_cut + 0.3 < ci.min_width / 2.0) ? 0.3 : 0.0;
  if (skin != P->owned_skin || !L.nl.valid || L.nl.n_atoms != n) return false;
  const int64_t off[2] = {0, (int64_t)n};
  L.set.setup(1, off, lattice, inv, L.nl.r_list, st);
  const double* Lb = L.nl.built_geoms[0].lat;
  const double* Ib = L.nl.built_geoms[0].inv;
  double sig = 1.0;
  Mat3 M;
  matmul3(Ib, lattice, M.m);
  if (!lattice_allclose(lattice, Lb)) sig = sigma_min3(M.m);
  nl_check_displacement(L.nl, L.set, P->d_pos.as<double>(), M, sig, nullptr,
                        P->d_out.as<double>() + 10, st);
  const double& dmax_out = run(&L);  // filled by the asynchronous copy
  AK_HIP(hipStreamSynchronize(st));
  const double dmax = dmax_out;
  L.n_steps++;
  L.nl.n_checks++;
  if (sig * L.nl.r_list - 2.0 * dmax <= L.nl.r_cut) {
    nl_build(L.nl, L.set, P->d_pos.as<double>(), st);  // stale list: rebuild, run again
    run(&L);
    AK_HIP(hipStreamSynchronize(st));
  }
  return true;
}

// Minimum-image bond vectors of every slot below cnt (3 planes n x stride).
__global__ __launch_bounds__(256) void k_site_vec(int n, const double* __restrict__ pos,
                                                  Geom g, const int* __restrict__ cnt,
                                                  const int* __restrict__ nbr, int stride,
                                                  double* __restrict__ vec) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * stride) return;
  const int a = (int)(t / stride), k = (int)(t % stride);
  if (k >= cnt[a]) return;
  double ux, uy, uz;
  mic_fast(g, pos, a, nbr[t], ux, uy, uz);
  const size_t plane = (size_t)n * stride;
  vec[t] = ux;
  vec[plane + t] = uy;
  vec[2 * plane + t] = uz;
}

extern "C" {

int ak_abi_version(void) { return AK_ABI_VERSION; }

const char* ak_last_error(void) { return g_err.c_str(); }

int ak_device_count(int32_t* count) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  if (count) *count = n;
  return AK_OK;
}

static int make_potential(ak_potential** out, int kind, int ns, double r_cut,
                          std::vector<double> params, std::vector<int> map) {
  return guarded([&] {
    AK_REQUIRE(out, AK_ERR_ARG, "null output handle");
    require_device();
    auto P = std::make_unique<ak_potential>();
    P->pot.kind = kind;
    P->pot.ns = ns;
    P->pot.r_cut = r_cut;
    P->pot.h_params = std::move(params);
    P->pot.h_map = std::move(map);
    P->stream = make_stream();
    P->pot.d_params.ensure(sizeof(double) * P->pot.h_params.size());
    AK_HIP(hipMemcpy(P->pot.d_params.p, P->pot.h_params.data(),
                     sizeof(double) * P->pot.h_params.size(), hipMemcpyHostToDevice));
    if (!P->pot.h_map.empty()) {
      P->pot.d_map.ensure(sizeof(int) * P->pot.h_map.size());
      AK_HIP(hipMemcpy(P->pot.d_map.p, P->pot.h_map.data(), sizeof(int) * P->pot.h_map.size(),
                       hipMemcpyHostToDevice));
    }
    if (kind == KIND_EAM) {
      std::vector<double> fast;
      if (eam_fast_table(ns, P->pot.h_params, P->pot.h_map, fast)) {
        P->pot.d_eam_fast.ensure(sizeof(double) * fast.size());
        AK_HIP(hipMemcpy(P->pot.d_eam_fast.p, fast.data(), sizeof(double) * fast.size(),
                         hipMemcpyHostToDevice));
        P->pot.eam_fast = true;
      }
    }
    *out = P.release();
  });
}

int ak_tersoff_create(int32_t ns, const double* pair_table, const double* triple_table,
                      double r_cut, ak_potential** out) {
  if (ns < 1 || ns > 16 || !pair_table || !triple_table || !(r_cut > 0.0))
    return fail(AK_ERR_ARG, "ak_tersoff_create: bad arguments");
  std::vector<double> p(pair_table, pair_table + (size_t)ns * ns * 9);
  p.insert(p.end(), triple_table, triple_table + (size_t)ns * ns * ns * 7);
  return make_potential(out, KIND_TERSOFF, ns, r_cut, std::move(p), {});
}

int ak_sw_create(const double* prm, ak_potential** out) {
  if (!prm) return fail(AK_ERR_ARG, "ak_sw_create: null parameters");
  if (!(prm[1] > 0.0) || !(prm[6] > 0.0)) return fail(AK_ERR_ARG, "ak_sw_create: sigma, a > 0");
  return make_potential(out, KIND_SW, 1, prm[6] * prm[1], std::vector<double>(prm, prm + 10),
                        {});
}

int ak_sw_multi_create(int32_t ns, const double* pair, const double* leg, const double* tri,
                       double r_cut, ak_potential** out) {
  if (ns < 1 || ns > 8 || !pair || !leg || !tri || !(r_cut > 0.0))
    return fail(AK_ERR_ARG, "ak_sw_multi_create: bad arguments");
  const size_t n2 = (size_t)ns * ns, n3 = n2 * ns;
  // device layout: pair (ns, ns, 7) with a = 0 where there is no two-body
  // term, leg (ns, ns, 3) with a = 0 where the record is absent, triple
  // (ns, ns, ns, 2) = epsilon lambda / 2 (0 where absent), cos0
  std::vector<double> t(7 * n2 + 3 * n2 + 2 * n3, 0.0);
  for (size_t p = 0; p < n2; ++p) {
    const double* s = pair + 8 * p;
    double* d = t.data() + 7 * p;
    const bool valid = s[0] != 0.0;
    if (valid && !(s[2] > 0.0)) return fail(AK_ERR_ARG, "ak_sw_multi_create: pair sigma > 0");
    d[0] = s[1];
    d[1] = valid ? s[2] : 1.0;
    d[2] = valid ? s[3] : 0.0;
    for (int k = 3; k < 7; ++k) d[k] = s[k + 1];
    const double* l = leg + 3 * p;
    double* dl = t.data() + 7 * n2 + 3 * p;
    if (l[1] > 0.0 && !(l[0] > 0.0)) return fail(AK_ERR_ARG, "ak_sw_multi_create: leg sigma > 0");
    dl[0] = l[1] > 0.0 ? l[0] : 1.0;
    dl[1] = l[1] > 0.0 ? l[1] : 0.0;
    dl[2] = l[2];
  }
  for (size_t q = 0; q < n3; ++q) {
    const double* s = tri + 4 * q;
    double* d = t.data() + 10 * n2 + 2 * q;
    d[0] = s[0] != 0.0 ? s[1] * (0.5 * s[2]) : 0.0;
    d[1] = s[3];
  }
  int rc = make_potential(out, KIND_SW, ns, r_cut, std::move(t), {});
  if (rc == AK_OK) (*out)->pot.sw_multi = true;
  return rc;
}

int ak_eam_create(int32_t ns, int32_t nf, const double* fn, const int32_t* map, double r_cut,
                  ak_potential** out) {
  if (ns < 1 || ns > 16 || nf < 1 || !fn || !map || !(r_cut > 0.0))
    return fail(AK_ERR_ARG, "ak_eam_create: bad arguments");
  std::vector<int> m(map, map + 2 * ns + ns * ns);
  for (int v : m)
    if (v < 0 || v >= nf) return fail(AK_ERR_ARG, "ak_eam_create: index_map out of range");
  int rc = make_potential(out, KIND_EAM, ns, r_cut,
                          std::vector<double>(fn, fn + (size_t)nf * 44), std::move(m));
  if (rc == AK_OK) (*out)->pot.eam_nf = nf;
  return rc;
}

int ak_eam_tables_create(int32_t ns, int32_t nf, const double* fn, const int32_t* map,
                         int64_t n_knots, const double* knots, const double* coeffs,
                         double r_cut, ak_potential** out) {
  if (ns < 1 || ns > 16 || nf < 1 || !fn || !map || !(r_cut > 0.0) || n_knots < 0 ||
      (n_knots > 0 && (!knots || !coeffs)))
    return fail(AK_ERR_ARG, "ak_eam_tables_create: bad arguments");
  std::vector<int> m(map, map + 2 * ns + ns * ns);
  for (int v : m)
    if (v < 0 || v >= nf) return fail(AK_ERR_ARG, "ak_eam_tables_create: index_map out of range");
  // device layout: rows (nf, 44) | knots | coefficients (4 per knot, at a
  // 4-double boundary); spline rows get absolute offsets, 1/h and x_0
  const size_t nr = (size_t)nf * 44, nk = (size_t)n_knots;
  const size_t xo = nr, co = (nr + nk + 3) / 4 * 4;
  std::vector<double> t(co + 4 * nk, 0.0);
  std::copy(fn, fn + nr, t.begin());
  std::copy(knots, knots + nk, t.begin() + xo);
  std::copy(coeffs, coeffs + 4 * nk, t.begin() + co);
  for (int32_t f = 0; f < nf; ++f) {
    double* row = t.data() + 44 * (size_t)f;
    const int code = (int)row[0];
    if (code >= 3 && code <= 6 && row[0] == (double)code) continue;
    if (code != 1 && code != 2)
      return fail(AK_ERR_ARG, "ak_eam_tables_create: function code must be 1-6");
    const double off = row[1], cnt = row[2];
    if (!(off >= 0.0) || !(cnt >= 2.0) || off != std::floor(off) || cnt != std::floor(cnt) ||
        off + cnt > (double)nk)
      return fail(AK_ERR_ARG, "ak_eam_tables_create: spline knots out of range");
    const size_t o = (size_t)off, n = (size_t)cnt;
    const double* x = knots + o;
    for (size_t i = 1; i < n; ++i)
      if (!(x[i] > x[i - 1]))
        return fail(AK_ERR_ARG, "ak_eam_tables_create: spline knots must increase");
    // uniform grid: the search may start at floor((r - x_0) / h)
    const double h = (x[n - 1] - x[0]) / (double)(n - 1);
    bool uniform = true;
    for (size_t i = 0; i < n && uniform; ++i)
      uniform = std::fabs(x[i] - (x[0] + (double)i * h)) <= 1e-9 * h;
    row[1] = (double)(xo + o);
    row[2] = (double)n;
    row[3] = (double)(co + 4 * o);
    row[4] = uniform ? 1.0 / h : 0.0;
    row[5] = x[0];
  }
  int rc = make_potential(out, KIND_EAM, ns, r_cut, std::move(t), std::move(m));
  if (rc == AK_OK) (*out)->pot.eam_nf = nf;
  return rc;
}

int ak_eam_eval(ak_potential* pot, int32_t function, int64_t m, const double* x, double* v,
                double* dv) {
  return guarded([&] {
    AK_REQUIRE(pot && pot->pot.kind == KIND_EAM, AK_ERR_ARG, "ak_eam_eval: not an EAM potential");
    AK_REQUIRE(m >= 0 && (m == 0 || (x && v && dv)), AK_ERR_ARG, "ak_eam_eval: bad arguments");
    AK_REQUIRE(function >= 0 && function < pot->pot.eam_nf, AK_ERR_ARG,
               "ak_eam_eval: no such function");
    if (m == 0) return;
    DevBuf dx, dvv, ddv;
    dx.ensure(sizeof(double) * (size_t)m);
    dvv.ensure(sizeof(double) * (size_t)m);
    ddv.ensure(sizeof(double) * (size_t)m);
    hipStream_t st = pot->stream;
    AK_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * (size_t)m, hipMemcpyHostToDevice, st));
    eam_feval(pot->pot, function, m, dx.as<double>(), dvv.as<double>(), ddv.as<double>(), st);
    AK_HIP(hipGetLastError());
    AK_HIP(hipMemcpyAsync(v, dvv.p, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, st));
    AK_HIP(hipMemcpyAsync(dv, ddv.p, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, st));
    AK_HIP(hipStreamSynchronize(st));
  });
}

int ak_potential_destroy(ak_potential* pot) {
  return guarded([&] { delete pot; });
}

int ak_potential_set_lanes(ak_potential* pot, int32_t lanes) {
  if (!pot) return fail(AK_ERR_ARG, "null potential");
  if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
    return fail(AK_ERR_ARG, "lanes must be 0, 4, 8, 16, 32 or 64");
  pot->pot.eam_lanes = lanes;
  return AK_OK;
}

int ak_potential_set_launch(ak_potential* pot, int32_t bond_tile, int32_t block_threads,
                            int32_t grid_cap) {
  if (!pot) return fail(AK_ERR_ARG, "null potential");
  if (bond_tile < 0 || bond_tile > 2)
    return fail(AK_ERR_ARG, "bond_tile must be 0 (automatic), 1 (lane per bond) or 2 (lane "
                            "per bond and partner leg)");
  if (block_threads != 0 && block_threads != 64 && block_threads != 128 && block_threads != 256 &&
      block_threads != 512)
    return fail(AK_ERR_ARG, "block_threads must be 0 (default 256), 64, 128, 256 or 512");
  if (grid_cap < 0) return fail(AK_ERR_ARG, "grid_cap must be >= 0");
  pot->pot.bond_tile = bond_tile;
  pot->pot.block = block_threads ? block_threads : 256;
  pot->pot.grid_cap = grid_cap;
  return AK_OK;
}

int ak_nlist_create(double r_cut, double skin, ak_nlist** out) {
  return guarded([&] {
    AK_REQUIRE(out, AK_ERR_ARG, "null output handle");
    AK_REQUIRE(r_cut > 0.0, AK_ERR_ARG, "NeighbourList: r_cut must be positive");
    AK_REQUIRE(skin >= 0.0, AK_ERR_ARG, "NeighbourList: skin_width must be >= 0");
    require_device();
    auto L = std::make_unique<ak_nlist>();
    L->nl.r_cut = r_cut;
    L->nl.skin = skin;
    L->nl.r_list = r_cut + skin;
    L->stream = make_stream();
    *out = L.release();
  });
}

int ak_nlist_destroy(ak_nlist* nl) {
  return guarded([&] { delete nl; });
}

int ak_nlist_build(ak_nlist* L, int32_t n, const double* lattice, const double* inv,
                   const double* positions) {
  return guarded([&] {
    AK_REQUIRE(L && lattice && (positions || n == 0) && n >= 0, AK_ERR_ARG, "bad arguments");
    check_finite(lattice, 9, "lattice");
    check_finite(positions, 3 * (size_t)n, "positions");
    L->set.setup(1, std::vector<int64_t>{0, (int64_t)n}.data(), lattice, inv, L->nl.r_list,
                 L->stream);
    L->d_pos.ensure(sizeof(double) * 3 * std::max(1, n));
    if (n > 0)
      AK_HIP(hipMemcpyAsync(L->d_pos.p, positions, sizeof(double) * 3 * n,
                            hipMemcpyHostToDevice, L->stream));
    nl_build(L->nl, L->set, L->d_pos.as<double>(), L->stream);
    AK_HIP(hipStreamSynchronize(L->stream));
  });
}

int ak_nlist_check_and_rebuild(ak_nlist* L, int32_t n, const double* lattice,
                               const double* inv, const double* positions,
                               int32_t* rebuilt) {
  return guarded([&] {
    AK_REQUIRE(L && lattice && (positions || n == 0) && n >= 0, AK_ERR_ARG, "bad arguments");
    check_finite(lattice, 9, "lattice");
    check_finite(positions, 3 * (size_t)n, "positions");
    L->d_pos.ensure(sizeof(double) * 3 * std::max(1, n));
    if (n > 0)
      AK_HIP(hipMemcpyAsync(L->d_pos.p, positions, sizeof(double) * 3 * n,
                            hipMemcpyHostToDevice, L->stream));
    bool r = check_and_rebuild(*L, n, lattice, inv, L->d_pos.as<double>(), L->stream);
    AK_HIP(hipStreamSynchronize(L->stream));
    if (rebuilt) *rebuilt = r ? 1 : 0;
  });
}

int ak_nlist_candidate_pairs(ak_nlist* L, int64_t* n_pairs, int32_t* oi, int32_t* oj,
                             int64_t capacity) {
  return guarded([&] {
    AK_REQUIRE(L && n_pairs, AK_ERR_ARG, "bad arguments");
    int64_t np = nl_half_pairs(L->nl, L->set, L->stream, nullptr, nullptr);
    *n_pairs = np;
    if (oi && oj) {
      AK_REQUIRE(capacity >= np, AK_ERR_CAPACITY, "pair buffers too small");
      nl_half_pairs(L->nl, L->set, L->stream, oi, oj);
    }
  });
}

int ak_nlist_counters(ak_nlist* L, int64_t* n_builds, int64_t* n_steps) {
  if (!L) return fail(AK_ERR_ARG, "null list");
  if (n_builds) *n_builds = L->nl.n_builds;
  if (n_steps) *n_steps = L->n_steps;
  return AK_OK;
}

int ak_compute(ak_potential* P, ak_nlist* nl, int32_t n, const double* lattice,
               const double* inv, const double* positions, const int32_t* species,
               int32_t need_stress, double* energy, double* forces, double* stress) {
  return guarded([&] {
    AK_REQUIRE(P && lattice && energy && (forces || n == 0) && n >= 0, AK_ERR_ARG,
               "bad arguments");
    AK_REQUIRE(positions || n == 0, AK_ERR_ARG, "null positions");
    check_finite(lattice, 9, "lattice");
    hipStream_t st = P->stream;
    // pinned staging [positions 3n | forces 3n | out 11]; the positions are
    // copied in and checked in one pass
    const size_t m = 3 * (size_t)n;
    double* hs = P->pinned_stage(2 * m + 11);
    double* h_pos = hs;
    double* h_F = hs + m;
    double* out = hs + 2 * m;  // energy, stress, the owned list's largest displacement
    {
      bool ok = true;  // a branch-free pass (vectorised)
      for (size_t k = 0; k < m; ++k) {
        const double v = positions[k];
        h_pos[k] = v;
        ok &= std::isfinite(v);
      }
      AK_REQUIRE(ok, AK_ERR_ARG, "positions contains non-finite values");
    }
    stage_species(P, n, species);
    P->d_pos.ensure(sizeof(double) * 3 * std::max(1, n));
    P->d_F.ensure(sizeof(double) * 3 * std::max(1, n));
    P->d_out.ensure(sizeof(double) * 11);
    if (n > 0)
      AK_HIP(hipMemcpyAsync(P->d_pos.p, h_pos, sizeof(double) * m, hipMemcpyHostToDevice, st));
    // the results are valid once the stream has been synchronised; run
    // returns a reference to the largest displacement
    auto run = [&](ak_nlist* Lst) -> const double& {
      compute_forces(P->pot, Lst->nl, Lst->set, P->d_pos.as<double>(), P->d_spc.as<int>(),
                     P->d_F.as<double>(), P->d_out.as<double>(), st);
      AK_HIP(hipMemcpyAsync(out, P->d_out.p, sizeof(double) * 11, hipMemcpyDeviceToHost, st));
      if (n > 0)
        AK_HIP(hipMemcpyAsync(h_F, P->d_F.p, sizeof(double) * m, hipMemcpyDeviceToHost, st));
      return out[10];
    };
    if (nl || !owned_list_speculative(P, n, lattice, inv, st, run)) {
      run(select_list(P, nl, n, lattice, inv, st));
      AK_HIP(hipStreamSynchronize(st));
    }
    if (n > 0) std::memcpy(forces, h_F, sizeof(double) * m);
    *energy = out[0];
    if (stress && need_stress) std::memcpy(stress, out + 1, 9 * sizeof(double));
  });
}

int ak_site_terms(ak_potential* P, ak_nlist* nl, int32_t n, const double* lattice,
                  const double* inv, const double* positions, const int32_t* species,
                  double* e_site, int64_t capacity, int32_t* centre, int32_t* neighbour,
                  double* vec, double* dudv, int64_t* n_entries) {
  return guarded([&] {
    AK_REQUIRE(P && lattice && n > 0 && positions && n_entries, AK_ERR_ARG, "bad arguments");
    AK_REQUIRE(!centre || (neighbour && vec && dudv && capacity >= 0), AK_ERR_ARG,
               "bad entry buffers");
    check_finite(lattice, 9, "lattice");
    check_finite(positions, 3 * (size_t)n, "positions");
    hipStream_t st = P->stream;
    stage_species(P, n, species);
    P->d_pos.ensure(sizeof(double) * 3 * n);
    P->d_F.ensure(sizeof(double) * 3 * n);
    P->d_out.ensure(sizeof(double) * 10);
    AK_HIP(hipMemcpyAsync(P->d_pos.p, positions, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                          st));
    ak_nlist* L = select_list(P, nl, n, lattice, inv, st);
    DevBuf d_es, d_vec;
    d_es.ensure(sizeof(double) * n);
    site_grads(P->pot, L->nl, L->set, P->d_pos.as<double>(), P->d_spc.as<int>(),
               P->d_F.as<double>(), P->d_out.as<double>(), d_es.as<double>(), st);
    const int S = L->nl.stride;
    const size_t ns = (size_t)n * S;
    d_vec.ensure(sizeof(double) * 3 * ns);
    hipLaunchKernelGGL(k_site_vec, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, n,
                       (const double*)P->d_pos.as<double>(), L->set.geoms[0],
                       (const int*)L->nl.cnt.as<int>(), (const int*)L->nl.nbr.as<int>(), S,
                       d_vec.as<double>());
    AK_HIP(hipGetLastError());
    std::vector<int> cnt(n), nbr(ns);
    std::vector<double> G(kGRec * ns), V(3 * ns);
    std::vector<uint64_t> mask(n, ~uint64_t(0));
    const bool masked = site_masked(P->pot, L->nl);
    AK_HIP(hipMemcpyAsync(cnt.data(), L->nl.cnt.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    AK_HIP(hipMemcpyAsync(nbr.data(), L->nl.nbr.p, sizeof(int) * ns, hipMemcpyDeviceToHost, st));
    AK_HIP(hipMemcpyAsync(G.data(), P->pot.G.p, sizeof(double) * kGRec * ns,
                          hipMemcpyDeviceToHost, st));
    AK_HIP(hipMemcpyAsync(V.data(), d_vec.p, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, st));
    if (masked)
      AK_HIP(hipMemcpyAsync(mask.data(), P->pot.amask.p, sizeof(uint64_t) * n,
                            hipMemcpyDeviceToHost, st));
    if (e_site)
      AK_HIP(hipMemcpyAsync(e_site, d_es.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    AK_HIP(hipStreamSynchronize(st));
    // entries: every bond within r_cut (engine/site_terms.py:84-97), atoms in
    // order, each atom's bonds in list order (ascending neighbour index)
    const double rc = P->pot.r_cut;
    int64_t m = 0;
    for (int a = 0; a < n; ++a)
      for (int k = 0; k < cnt[a]; ++k) {
        const size_t t = (size_t)a * S + k;
        const double x = V[t], y = V[ns + t], z = V[2 * ns + t];
        const double r = std::sqrt(x * x + y * y + z * z);
        if (!(r < rc && r > 1e-12)) continue;
        if (centre && m < capacity) {
          const bool has_g = !masked || ((mask[a] >> k) & 1);
          centre[m] = a;
          neighbour[m] = nbr[t];
          vec[3 * m] = x;
          vec[3 * m + 1] = y;
          vec[3 * m + 2] = z;
          dudv[3 * m] = has_g ? G[kGRec * t] : 0.0;
          dudv[3 * m + 1] = has_g ? G[kGRec * t + 1] : 0.0;
          dudv[3 * m + 2] = has_g ? G[kGRec * t + 2] : 0.0;
        }
        ++m;
      }
    *n_entries = m;
    AK_REQUIRE(!centre || m <= capacity, AK_ERR_CAPACITY, "site-term entry buffers too small");
  });
}

int ak_heat_flux(ak_potential* P, ak_nlist* nl, int32_t n, const double* lattice,
                 const double* inv, const double* positions, const double* velocities,
                 const double* masses, const int32_t* species, double* flux, double* e_site) {
  return guarded([&] {
    AK_REQUIRE(P && lattice && n > 0 && positions && velocities && masses && flux, AK_ERR_ARG,
               "bad arguments");
    check_finite(lattice, 9, "lattice");
    check_finite(positions, 3 * (size_t)n, "positions");
    check_finite(velocities, 3 * (size_t)n, "velocities");
    hipStream_t st = P->stream;
    stage_species(P, n, species);
    P->d_pos.ensure(sizeof(double) * 3 * n);
    P->d_F.ensure(sizeof(double) * 3 * n);
    P->d_out.ensure(sizeof(double) * 10);
    AK_HIP(hipMemcpyAsync(P->d_pos.p, positions, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                          st));
    ak_nlist* L = select_list(P, nl, n, lattice, inv, st);
    DevBuf d_es, d_vel, d_m, d_part, d_J;
    d_es.ensure(sizeof(double) * n);
    d_vel.ensure(sizeof(double) * 3 * n);
    d_m.ensure(sizeof(double) * n);
    d_part.ensure(sizeof(double) * 3 * ((size_t)n / 32 + 1));
    d_J.ensure(sizeof(double) * 3);
    AK_HIP(hipMemcpyAsync(d_vel.p, velocities, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                          st));
    AK_HIP(hipMemcpyAsync(d_m.p, masses, sizeof(double) * n, hipMemcpyHostToDevice, st));
    site_grads(P->pot, L->nl, L->set, P->d_pos.as<double>(), P->d_spc.as<int>(),
               P->d_F.as<double>(), P->d_out.as<double>(), d_es.as<double>(), st);
    heat_flux(P->pot, L->nl, L->set, P->d_pos.as<double>(), d_vel.as<double>(),
              d_m.as<double>(), d_es.as<double>(), nullptr, kConvAmu, nullptr, 1,
              d_part.as<double>(), d_J.as<double>(), st);
    AK_HIP(hipGetLastError());
    AK_HIP(hipMemcpyAsync(flux, d_J.p, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
    if (e_site)
      AK_HIP(hipMemcpyAsync(e_site, d_es.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    AK_HIP(hipStreamSynchronize(st));
  });
}

int ak_compute_device(ak_potential* P, ak_nlist* L, int32_t n, const double* lattice,
                      const double* inv, const double* d_pos, const int32_t* d_spc,
                      double* d_F, double* d_es, void* stream) {
  return guarded([&] {
    AK_REQUIRE(P && L && lattice && d_F && d_es && n >= 0, AK_ERR_ARG, "bad arguments");
    AK_REQUIRE(L->nl.r_cut + 1e-12 >= P->pot.r_cut, AK_ERR_ARG,
               "neighbour list r_cut is smaller than the potential's cutoff");
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    const int64_t off[2] = {0, (int64_t)n};
    L->set.setup(1, off, lattice, inv, L->nl.r_list, st);
    AK_REQUIRE(L->nl.valid && L->nl.n_atoms == n, AK_ERR_STATE,
               "the neighbour list was not built for this structure");
    const int* spc = d_spc;
    if (!spc) {
      stage_species(P, n, nullptr);
      spc = P->d_spc.as<int>();
    }
    compute_forces(P->pot, L->nl, L->set, d_pos, spc, d_F, d_es, st);
  });
}

// ---------------------------------------------------------------------------
// Force call without device allocation.  Every buffer of the force kernels
// besides the caller's inputs and outputs (geometry, launch chunks, block
// partials, bond gradients, masks, EAM derivatives) is carved out of one
// file-scope device array, `scratch`, of kScratchBytes.  Host code never
// passes the array's name to a kernel: the address of each device's copy
// comes from hipGetSymbolAddress on the first call on that device (with the
// device current) and is kept per device.  All calls on one device share
// its copy, so calls that overlap on one device race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kScratchDevices] = {};

unsigned char* device_scratch() {
  int dev = 0;
  AK_HIP(hipGetDevice(&dev));
  AK_REQUIRE(dev >= 0 && dev < kScratchDevices, AK_ERR_STATE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    AK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    AK_REQUIRE(p != nullptr, AK_ERR_HIP, "no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[dev];
}

// Geometry, atom -> structure map (all 0) and launch chunks of one structure
// of n atoms, written by a kernel (no memset or host copy).
__global__ void k_work_setup(Geom g, int n, int chunk, int n_chunks, Geom* __restrict__ geoms,
                             int* __restrict__ sid, int* __restrict__ rng,
                             int* __restrict__ first) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    *geoms = g;
    first[0] = 0;
    first[1] = n_chunks;
  }
  if (t < n) sid[t] = 0;
  if (t < n_chunks) {
    rng[2 * t] = t * chunk;
    rng[2 * t + 1] = min((t + 1) * chunk, n);
  }
}
}  // namespace

int ak_compute_device_noalloc(ak_potential* P, ak_nlist* L, int32_t n, const double* lattice,
                              const double* inv, const double* d_pos, const int32_t* d_spc,
                              double* d_F, double* d_es, void* stream) {
  return guarded([&] {
    AK_REQUIRE(P && L && lattice && d_pos && d_spc && d_F && d_es && n >= 0, AK_ERR_ARG,
               "bad arguments (device positions, species, forces and energy/stress are "
               "required)");
    AK_REQUIRE(L->nl.r_cut + 1e-12 >= P->pot.r_cut, AK_ERR_ARG,
               "neighbour list r_cut is smaller than the potential's cutoff");
    AK_REQUIRE(L->nl.valid && L->nl.n_atoms == n, AK_ERR_STATE,
               "the neighbour list was not built for this structure");
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    unsigned char* base = device_scratch();
    // geometry of the cell on the host (no allocation)
    Geom g{};
    std::memcpy(g.lat, lattice, sizeof g.lat);
    if (inv)
      std::memcpy(g.inv, inv, sizeof g.inv);
    else
      invert3(g.lat, g.inv);
    const CellInfo ci = cell_info(g.lat);
    g.volume = ci.volume;
    AK_REQUIRE(L->nl.r_list < ci.min_width / 2.0, AK_ERR_MIN_IMAGE,
               "r_cut + skin_width must be < half the smallest perpendicular cell width for "
               "the minimum-image convention to be valid");
    for (int ax = 0; ax < 3; ++ax) g.nc[ax] = std::max(1, (int)std::floor(ci.perp[ax] / L->nl.r_list));
    g.natoms = n;
    const ForcePlan pl = force_plan(P->pot, L->nl, false);
    const int n_chunks = n > 0 ? (n + pl.chunk - 1) / pl.chunk : 0;
    // scratch layout, 256-byte aligned segments
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t at = off;
      off = (off + bytes + 255) & ~size_t(255);
      return at;
    };
    const bool eam = P->pot.kind == KIND_EAM;
    const size_t o_geom = take(sizeof(Geom));
    const size_t o_sid = take(sizeof(int) * (size_t)std::max(1, n));
    const size_t o_rng = take(sizeof(int) * 2 * (size_t)std::max(1, n_chunks));
    const size_t o_first = take(sizeof(int) * 2);
    const size_t o_part = take(sizeof(double) * 10 * (size_t)std::max(1, n_chunks));
    const size_t o_G = eam ? 0 : take(sizeof(double) * kGRec * (size_t)n * L->nl.stride);
    const size_t o_mask = eam ? 0 : take(sizeof(uint64_t) * (size_t)n);
    const size_t o_aux = eam ? take(sizeof(double) * 2 * (size_t)n) : 0;
    const size_t o_flag = eam ? take(sizeof(int) * (size_t)std::max(1, n_chunks)) : 0;
    if (off > kScratchBytes) {
      char buf[256];
      snprintf(buf, sizeof buf,
               "%d atoms with %d neighbour slots per row need %zu bytes of scratch; the "
               "no-allocation force call holds at most %zu",
               n, L->nl.stride, off, kScratchBytes);
      throw Error(AK_ERR_CAPACITY, buf);
    }
    ForceWork w;
    w.n = n;
    w.n_struct = 1;
    w.n_chunks = n_chunks;
    Geom* geoms = reinterpret_cast<Geom*>(base + o_geom);
    int* sid = reinterpret_cast<int*>(base + o_sid);
    int* rng = reinterpret_cast<int*>(base + o_rng);
    int* first = reinterpret_cast<int*>(base + o_first);
    w.geoms = geoms;
    w.sid = sid;
    w.rng = rng;
    w.first = first;
    w.partial = reinterpret_cast<double*>(base + o_part);
    if (eam) {
      w.aux = reinterpret_cast<double*>(base + o_aux);
      w.pair_flag = reinterpret_cast<int*>(base + o_flag);
    } else {
      w.G = reinterpret_cast<double*>(base + o_G);
      w.amask = reinterpret_cast<uint64_t*>(base + o_mask);
    }
    const int items = std::max(1, std::max(n, n_chunks));
    hipLaunchKernelGGL(k_work_setup, dim3((items + 255) / 256), dim3(256), 0, st, g, n, pl.chunk,
                       n_chunks, geoms, sid, rng, first);
    compute_forces_work(P->pot, L->nl, pl, w, d_pos, d_spc, d_F, d_es, st);
    AK_HIP(hipGetLastError());
  });
}

int ak_compute_batched(ak_potential* P, int32_t S, const int64_t* off, const double* lattices,
                       const double* invs, const double* positions, const int32_t* species,
                       double* energies, double* forces, double* stresses) {
  return guarded([&] {
    AK_REQUIRE(P && S >= 1 && off && lattices && energies, AK_ERR_ARG, "bad arguments");
    const int n = (int)off[S];
    AK_REQUIRE(n >= 0 && (positions || n == 0) && (forces || n == 0), AK_ERR_ARG,
               "bad arguments");
    check_finite(lattices, 9 * (size_t)S, "lattices");
    check_finite(positions, 3 * (size_t)n, "positions");
    hipStream_t st = P->stream;
    stage_species(P, n, species);
    P->d_pos.ensure(sizeof(double) * 3 * std::max(1, n));
    P->d_F.ensure(sizeof(double) * 3 * std::max(1, n));
    P->d_out.ensure(sizeof(double) * 10 * S);
    if (n > 0)
      AK_HIP(hipMemcpyAsync(P->d_pos.p, positions, sizeof(double) * 3 * n,
                            hipMemcpyHostToDevice, st));
    ak_nlist& B = P->batch;
    B.nl.r_cut = P->pot.r_cut;
    B.nl.skin = 0.0;
    B.nl.r_list = P->pot.r_cut;
    B.set.setup(S, off, lattices, invs, B.nl.r_list, st);
    nl_build(B.nl, B.set, P->d_pos.as<double>(), st);
    P->batch_off.assign(off, off + S + 1);
    compute_forces(P->pot, B.nl, B.set, P->d_pos.as<double>(), P->d_spc.as<int>(),
                   P->d_F.as<double>(), P->d_out.as<double>(), st);
    std::vector<double> out(10 * (size_t)S);
    AK_HIP(hipMemcpyAsync(out.data(), P->d_out.p, sizeof(double) * out.size(),
                          hipMemcpyDeviceToHost, st));
    if (n > 0)
      AK_HIP(hipMemcpyAsync(forces, P->d_F.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                            st));
    AK_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < S; ++s) {
      energies[s] = out[10 * s];
      if (stresses) std::memcpy(stresses + 9 * s, out.data() + 10 * s + 1, 9 * sizeof(double));
    }
  });
}

int ak_autotune_batched(ak_potential* P, int32_t S, const int64_t* off, const double* lattices,
                        const double* positions, const int32_t* species, int32_t reps,
                        int32_t* best, double* best_ms, int32_t capacity, int32_t* shapes,
                        double* times_ms, int32_t* n_tried) {
  return guarded([&] {
    AK_REQUIRE(P && S >= 1 && off && lattices && best && best_ms, AK_ERR_ARG, "bad arguments");
    const int n = (int)off[S];
    AK_REQUIRE(n > 0 && positions, AK_ERR_ARG, "the batch holds no atoms");
    check_finite(positions, 3 * (size_t)n, "positions");
    hipStream_t st = P->stream;
    stage_species(P, n, species);
    P->d_pos.ensure(sizeof(double) * 3 * n);
    P->d_F.ensure(sizeof(double) * 3 * n);
    P->d_out.ensure(sizeof(double) * 10 * S);
    AK_HIP(hipMemcpyAsync(P->d_pos.p, positions, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                          st));
    ak_nlist& B = P->batch;
    B.nl.r_cut = P->pot.r_cut;
    B.nl.skin = 0.0;
    B.nl.r_list = P->pot.r_cut;
    B.set.setup(S, off, lattices, nullptr, B.nl.r_list, st);
    nl_build(B.nl, B.set, P->d_pos.as<double>(), st);
    P->batch_off.assign(off, off + S + 1);

    // candidate launch shapes {bond_tile, block, grid_cap, eam lanes}
    std::vector<std::array<int, 4>> cand;
    const int blocks[3] = {64, 128, 256};
    // the bond kernels and the Zhou-table EAM kernels run one block per chunk,
    // so the grid cap is not a candidate
    if (P->pot.kind == KIND_EAM) {
      for (int lanes : {8, 16, 32, 64})
        for (int b : blocks)
          if (lanes <= b) cand.push_back({0, b, 0, lanes});
    } else {
      const int L = B.nl.stride;
      std::vector<int> tiles = {1};
      if (P->pot.kind == KIND_TERSOFF && L * L <= 64) tiles.push_back(2);
      for (int t : tiles)
        for (int b : blocks)
          if ((t == 2 ? L * L : L) <= b) cand.push_back({t, b, 0, 0});
      // the four-slot Tersoff kernel also runs blocks of 512 threads
      if (P->pot.kind == KIND_TERSOFF && L == 4) cand.push_back({1, 512, 0, 0});
    }
    Potential& pot = P->pot;
    const int saved[4] = {pot.bond_tile, pot.block, pot.grid_cap, pot.eam_lanes};
    const bool saved_prof = pot.prof_on;
    pot.prof_on = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    AK_HIP(hipEventCreate(&e0));
    AK_HIP(hipEventCreate(&e1));
    double best_t = 1e300;
    int bi = -1;
    int tried = 0;
    const int R = std::max(1, (int)reps);
    try {
      // three rounds over the candidates, each keeping its best round: the
      // shapes differ by a few per cent, less than one round's noise
      std::vector<double> tk(cand.size(), 1e300);
      for (int round = 0; round < 3; ++round)
        for (size_t k = 0; k < cand.size(); ++k) {
          pot.bond_tile = cand[k][0];
          pot.block = cand[k][1];
          pot.grid_cap = cand[k][2];
          pot.eam_lanes = cand[k][3];
          // first call: chunk tables for this block size, caches warm
          compute_forces(pot, B.nl, B.set, P->d_pos.as<double>(), P->d_spc.as<int>(),
                         P->d_F.as<double>(), P->d_out.as<double>(), st);
          AK_HIP(hipEventRecord(e0, st));
          for (int r = 0; r < R; ++r)
            compute_forces(pot, B.nl, B.set, P->d_pos.as<double>(), P->d_spc.as<int>(),
                           P->d_F.as<double>(), P->d_out.as<double>(), st);
          AK_HIP(hipEventRecord(e1, st));
          AK_HIP(hipEventSynchronize(e1));
          float ms = 0.0f;
          AK_HIP(hipEventElapsedTime(&ms, e0, e1));
          tk[k] = std::min(tk[k], (double)ms / R);
        }
      for (size_t k = 0; k < cand.size(); ++k) {
        if (shapes && times_ms && tried < capacity) {
          for (int q = 0; q < 4; ++q) shapes[4 * tried + q] = cand[k][q];
          times_ms[tried] = tk[k];
        }
        ++tried;
        if (tk[k] < best_t) {
          best_t = tk[k];
          bi = (int)k;
        }
      }
    } catch (...) {
      pot.bond_tile = saved[0];
      pot.block = saved[1];
      pot.grid_cap = saved[2];
      pot.eam_lanes = saved[3];
      pot.prof_on = saved_prof;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    pot.prof_on = saved_prof;
    pot.bond_tile = cand[bi][0];
    pot.block = cand[bi][1];
    pot.grid_cap = cand[bi][2];
    pot.eam_lanes = cand[bi][3];
    // The rounds above queue hundreds of calls between waits. On MI355X
    // (ROCm 7) the next few dozen launches after such a burst cost about
    // twice their usual host time (15 us instead of 8 per force call),
    // whether or not the process pauses in between; calls that each wait
    // for the stream bring the launch cost back, so the tuner ends with
    // twenty of them at the chosen shape.
    for (int r = 0; r < 20; ++r) {
      compute_forces(pot, B.nl, B.set, P->d_pos.as<double>(), P->d_spc.as<int>(),
                     P->d_F.as<double>(), P->d_out.as<double>(), st);
      AK_HIP(hipStreamSynchronize(st));
    }
    for (int q = 0; q < 4; ++q) best[q] = cand[bi][q];
    *best_ms = best_t;
    if (n_tried) *n_tried = tried;
  });
}

int ak_compute_batched_device(ak_potential* P, int32_t S, const int64_t* off,
                              const double* lattices, const double* d_pos,
                              const int32_t* d_spc, double* d_F, double* d_es, int32_t rebuild,
                              void* stream) {
  return guarded([&] {
    AK_REQUIRE(P && S >= 1 && off && lattices && d_F && d_es, AK_ERR_ARG, "bad arguments");
    hipStream_t st = stream ? (hipStream_t)stream : P->stream;
    ak_nlist& B = P->batch;
    bool same = B.nl.valid && P->batch_off.size() == (size_t)S + 1 &&
                std::equal(P->batch_off.begin(), P->batch_off.end(), off);
    B.nl.r_cut = P->pot.r_cut;
    B.nl.skin = 0.0;
    B.nl.r_list = P->pot.r_cut;
    B.set.setup(S, off, lattices, nullptr, B.nl.r_list, st);
    if (!same || rebuild) {
      nl_build(B.nl, B.set, d_pos, st);
      P->batch_off.assign(off, off + S + 1);
    }
    const int* spc = d_spc;
    if (!spc) {
      stage_species(P, (int)off[S], nullptr);
      spc = P->d_spc.as<int>();
    }
    compute_forces(P->pot, B.nl, B.set, d_pos, spc, d_F, d_es, st);
  });
}

}  // extern "C"

namespace ak {
void tersoff_loss_grad_run(Potential& pot, int S, const int64_t* off, const double* lattices,
                           const double* positions, const double* theta,
                           const double* ref_energy, const double* ref_forces,
                           const double* ref_stress, const double* weights, double* loss,
                           double* grad, double* energies, double* forces, double* stresses,
                           hipStream_t st);
void tersoff_fit_set(Potential& pot, int S, const int64_t* off, const double* lattices,
                     const double* positions, const double* ref_energy,
                     const double* ref_forces, const double* ref_stress, hipStream_t st);
void tersoff_fit_eval(Potential& pot, const double* theta, const double* weights, double* loss,
                      double* grad, double* energies, double* forces, double* stresses,
                      hipStream_t st);
void tersoff_fit_terms(Potential& pot, int S, const double* terms, int huber,
                       const double* deltas);
}  // namespace ak

extern "C" {

int ak_md_run(ak_potential* P, int32_t n, const double* lattice, const double* inv,
              double* positions, double* velocities, const double* masses,
              const int32_t* species, double dt, int32_t n_steps, double skin,
              int32_t use_graph, const double* thermostat, double* state,
              const uint8_t* fix_mask, int32_t flux_every, int32_t frame_every, double* frames,
              double* frame_lattices, double* e_pot, double* e_kin, double* e_cons,
              double* flux, double* lattice_out, double* volume, double* pressure,
              int32_t* rebuild_counts, double* cell_history, double* ptensor_history,
              const int32_t* flux_labels, int32_t n_labels, int32_t region_axis,
              const double* region_bounds, int32_t n_regions, double* flux_groups,
              float* atom_flux, int32_t* n_rebuilds, double* seconds) {
  return guarded([&] {
    AK_REQUIRE(P && lattice && positions && velocities && masses && n > 0, AK_ERR_ARG,
               "bad arguments");
    AK_REQUIRE(skin >= 0.0, AK_ERR_ARG, "skin_width must be >= 0");
    AK_REQUIRE(flux_every >= 0, AK_ERR_ARG, "flux_every must be >= 0");
    check_finite(lattice, 9, "lattice");
    check_finite(positions, 3 * (size_t)n, "positions");
    check_finite(velocities, 3 * (size_t)n, "velocities");
    for (int a = 0; a < n; ++a) {
      AK_REQUIRE(masses[a] > 0.0, AK_ERR_ARG, "masses must be positive");
      if (species)
        AK_REQUIRE(species[a] >= 0 && species[a] < P->pot.ns, AK_ERR_ARG,
                   "species index out of range for this potential");
    }
    MDOptions opt;
    if (thermostat) {
      opt.thermostat = (int)thermostat[0];
      AK_REQUIRE(thermostat[0] == (double)opt.thermostat, AK_ERR_ARG, "bad thermostat kind");
      opt.temperature = thermostat[1];
      opt.tau = thermostat[2];
      opt.chain = (int)thermostat[3];
      AK_REQUIRE(thermostat[3] == (double)opt.chain, AK_ERR_ARG, "bad chain length");
      AK_REQUIRE(thermostat[4] >= 0.0 && thermostat[4] < 9007199254740992.0 &&
                     thermostat[4] == std::floor(thermostat[4]),
                 AK_ERR_ARG, "seed must be an integer in [0, 2^53)");
      opt.seed = (uint64_t)thermostat[4];
      if (opt.thermostat == 4) {
        opt.pressure_gpa = thermostat[5];
        opt.tau_baro = thermostat[6];
        AK_REQUIRE(std::isfinite(opt.pressure_gpa), AK_ERR_ARG, "bad NPT pressure");
        opt.cell_dof = (int)thermostat[7];
        opt.baro_thermostat = (int)thermostat[8];
        AK_REQUIRE(thermostat[7] == (double)opt.cell_dof &&
                       thermostat[8] == (double)opt.baro_thermostat,
                   AK_ERR_ARG, "bad NPT cell_dof or baro_thermostat code");
      }
      AK_REQUIRE(thermostat[9] >= 0.0 && thermostat[9] < 2147483647.0 &&
                     thermostat[9] == std::floor(thermostat[9]),
                 AK_ERR_ARG, "velocity_rescale_steps must be an integer >= 0");
      opt.rescale_steps = (int)thermostat[9];
    }
    opt.fixed = fix_mask;
    AK_REQUIRE(frame_every >= 0, AK_ERR_ARG, "frame_every must be >= 0");
    opt.frame_every = frames ? frame_every : 0;
    opt.frames = frames;
    opt.frame_lattices = frame_lattices;
    opt.rebuild_counts = rebuild_counts;
    opt.cell_history = cell_history;
    opt.ptensor_history = ptensor_history;
    opt.flux_labels = flux_labels;
    opt.n_labels = n_labels;
    opt.region_axis = region_axis;
    opt.region_bounds = region_bounds;
    opt.n_regions = n_regions;
    opt.flux_groups = flux_groups;
    opt.atom_flux = atom_flux;
    opt.flux_every = flux_every;
    int nr = 0;
    md_run(P->pot, n, lattice, inv, positions, velocities, masses, species, dt, n_steps, skin,
           use_graph != 0, opt, state, e_pot, e_kin, e_cons, flux, lattice_out, volume,
           pressure, &nr, seconds, P->stream);
    if (n_rebuilds) *n_rebuilds = nr;
  });
}

int ak_nemd_run(ak_potential* P, int32_t n, const double* lattice, const double* inv,
                double* positions, double* velocities, const double* masses,
                const int32_t* species, double dt, int32_t n_steps, double skin,
                int32_t use_graph, const uint8_t* bath, double T_hot, double T_cold,
                double damping, double seed, double* state, int32_t sample_every,
                int32_t sample_phase, int32_t axis, int32_t n_bins, double* e_pot,
                double* e_kin, double* e_cons, double* bath_history, double* bins,
                int32_t* n_rebuilds, double* seconds) {
  return guarded([&] {
    AK_REQUIRE(P && lattice && positions && velocities && masses && bath && n > 0, AK_ERR_ARG,
               "bad arguments");
    AK_REQUIRE(skin >= 0.0, AK_ERR_ARG, "skin_width must be >= 0");
    AK_REQUIRE(sample_every >= 0, AK_ERR_ARG, "sample_every must be >= 0");
    AK_REQUIRE(sample_every == 0 || (axis >= 0 && axis <= 2 && n_bins >= 1 && n_bins <= 1024),
               AK_ERR_ARG, "NEMD bins: axis 0, 1 or 2 and 1 to 1024 bins");
    AK_REQUIRE(seed >= 0.0 && seed < 9007199254740992.0 && seed == std::floor(seed), AK_ERR_ARG,
               "seed must be an integer in [0, 2^53)");
    check_finite(lattice, 9, "lattice");
    check_finite(positions, 3 * (size_t)n, "positions");
    check_finite(velocities, 3 * (size_t)n, "velocities");
    for (int a = 0; a < n; ++a) {
      AK_REQUIRE(masses[a] > 0.0, AK_ERR_ARG, "masses must be positive");
      if (species)
        AK_REQUIRE(species[a] >= 0 && species[a] < P->pot.ns, AK_ERR_ARG,
                   "species index out of range for this potential");
    }
    MDOptions opt;
    opt.thermostat = 5;
    opt.temperature = T_hot;
    opt.T_cold = T_cold;
    opt.tau = damping;
    opt.seed = (uint64_t)seed;
    opt.bath = bath;
    opt.bath_history = bath_history;
    opt.flux_every = sample_every;
    opt.flux_phase = sample_phase;
    if (sample_every > 0) {
      opt.region_axis = axis;
      opt.n_regions = n_bins;
      opt.uniform_bins = true;
      opt.group_kinetic = true;
      opt.flux_groups = bins;
    }
    int nr = 0;
    md_run(P->pot, n, lattice, inv, positions, velocities, masses, species, dt, n_steps, skin,
           use_graph != 0, opt, state, e_pot, e_kin, e_cons, nullptr, nullptr, nullptr, nullptr,
           &nr, seconds, P->stream);
    if (n_rebuilds) *n_rebuilds = nr;
  });
}

int ak_md_nve(ak_potential* P, int32_t n, const double* lattice, const double* inv,
              double* positions, double* velocities, const double* masses,
              const int32_t* species, double dt, int32_t n_steps, double skin,
              int32_t use_graph, double* e_pot, double* e_kin, int32_t* n_rebuilds,
              double* seconds) {
  return guarded([&] {
    AK_REQUIRE(P && lattice && positions && velocities && masses && n > 0, AK_ERR_ARG,
               "bad arguments");
    AK_REQUIRE(skin >= 0.0, AK_ERR_ARG, "skin_width must be >= 0");
    check_finite(lattice, 9, "lattice");
    check_finite(positions, 3 * (size_t)n, "positions");
    check_finite(velocities, 3 * (size_t)n, "velocities");
    for (int a = 0; a < n; ++a) {
      AK_REQUIRE(masses[a] > 0.0, AK_ERR_ARG, "masses must be positive");
      if (species)
        AK_REQUIRE(species[a] >= 0 && species[a] < P->pot.ns, AK_ERR_ARG,
                   "species index out of range for this potential");
    }
    int nr = 0;
    md_run(P->pot, n, lattice, inv, positions, velocities, masses, species, dt, n_steps, skin,
           use_graph != 0, MDOptions{}, nullptr, e_pot, e_kin, nullptr, nullptr, nullptr,
           nullptr, nullptr, &nr, seconds, P->stream);
    if (n_rebuilds) *n_rebuilds = nr;
  });
}

int ak_tersoff_loss_grad(ak_potential* P, int32_t S, const int64_t* off,
                         const double* lattices, const double* positions,
                         const double* theta, const double* ref_energy,
                         const double* ref_forces, const double* ref_stress,
                         const double* weights, double* loss, double* grad,
                         double* energies, double* forces, double* stresses) {
  return guarded([&] {
    AK_REQUIRE(P && S >= 1 && off && lattices && positions && theta && ref_energy &&
                   ref_forces && ref_stress && weights && loss,
               AK_ERR_ARG, "bad arguments");
    AK_REQUIRE(P->pot.kind == KIND_TERSOFF && P->pot.ns == 1, AK_ERR_ARG,
               "the parameter gradient is defined for a single-species Tersoff potential");
    tersoff_loss_grad_run(P->pot, S, off, lattices, positions, theta, ref_energy, ref_forces,
                          ref_stress, weights, loss, grad, energies, forces, stresses,
                          P->stream);
  });
}

int ak_tersoff_fit_load(ak_potential* P, int32_t S, const int64_t* off,
                        const double* lattices, const double* positions,
                        const double* ref_energy, const double* ref_forces,
                        const double* ref_stress) {
  return guarded([&] {
    AK_REQUIRE(P && S >= 1 && off && lattices && positions && ref_energy && ref_forces &&
                   ref_stress,
               AK_ERR_ARG, "bad arguments");
    AK_REQUIRE(P->pot.kind == KIND_TERSOFF && P->pot.ns == 1, AK_ERR_ARG,
               "the parameter gradient is defined for a single-species Tersoff potential");
    check_finite(lattices, 9 * (size_t)S, "lattices");
    check_finite(positions, 3 * (size_t)off[S], "positions");
    tersoff_fit_set(P->pot, S, off, lattices, positions, ref_energy, ref_forces, ref_stress,
                    P->stream);
  });
}

int ak_tersoff_fit_eval(ak_potential* P, const double* theta, const double* weights,
                        double* loss, double* grad, double* energies, double* forces,
                        double* stresses) {
  return guarded([&] {
    AK_REQUIRE(P && theta && weights && loss, AK_ERR_ARG, "bad arguments");
    AK_REQUIRE(P->pot.kind == KIND_TERSOFF && P->pot.ns == 1, AK_ERR_ARG,
               "the parameter gradient is defined for a single-species Tersoff potential");
    tersoff_fit_eval(P->pot, theta, weights, loss, grad, energies, forces, stresses, P->stream);
  });
}

int ak_tersoff_fit_terms(ak_potential* P, int32_t S, const double* terms, int32_t loss_type,
                         const double* huber_deltas) {
  return guarded([&] {
    AK_REQUIRE(P && terms && S >= 1 && (loss_type == 0 || huber_deltas), AK_ERR_ARG,
               "bad arguments");
    tersoff_fit_terms(P->pot, S, terms, loss_type, huber_deltas);
  });
}

}  // extern "C"

extern "C" {

int ak_profile_enable(ak_potential* P, int32_t on) {
  return guarded([&] {
    AK_REQUIRE(P, AK_ERR_ARG, "null potential");
    P->pot.prof_on = on != 0;
    P->pot.prof_used = 0;
  });
}

int ak_profile_read(ak_potential* P, double* total_ms, double* min_ms, int64_t* count) {
  return guarded([&] {
    AK_REQUIRE(P && total_ms && count, AK_ERR_ARG, "bad arguments");
    double tot = 0.0, mn = 0.0;
    for (size_t k = 0; k < P->pot.prof_used; ++k) {
      auto& e = P->pot.prof_pool[k];
      AK_HIP(hipEventSynchronize(e.b));
      float ms = 0.0f;
      AK_HIP(hipEventElapsedTime(&ms, e.a, e.b));
      tot += ms;
      mn = (k == 0 || ms < mn) ? ms : mn;
    }
    *total_ms = tot;
    if (min_ms) *min_ms = mn;
    *count = (int64_t)P->pot.prof_used;
    P->pot.prof_used = 0;
  });
}

}  // extern "C"
