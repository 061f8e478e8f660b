/*
 * akitsushima_amd.h -- C ABI of the MI355X FP64 force engine.
 *
 * Every entry point returns 0 on success or a status code (AK_ERR_*); the
 * message of the last failure on the calling thread is ak_last_error().
 * Arrays are C-contiguous: lattice (3, 3) with lattice vectors as ROWS,
 * positions and forces (N, 3) in Angstrom and eV/Angstrom, stress (3, 3) in
 * eV/Angstrom^3 with the reference's sign convention
 * (sigma = -(1/V) sum f (x) r; docs/adr/0001-stress-sign-and-multiplication-
 * order-conventions.md:7-28).  Species are int32 indices into the species
 * list the potential was created with.
 *
 * Handles are not thread-safe: use one potential handle per thread, as the
 * reference keeps one owned neighbour list per thread
 * (engine/neighborlist.py:1098-1100).
 */
#ifndef AKITSUSHIMA_AMD_H
#define AKITSUSHIMA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AK_ABI_VERSION 1

#define AK_OK 0
#define AK_ERR_ARG 1
#define AK_ERR_MIN_IMAGE 2 /* cell too thin for r_cut (+ skin): the reference's ValueError */
#define AK_ERR_HIP 3
#define AK_ERR_CAPACITY 4
#define AK_ERR_STATE 5
#define AK_ERR_NO_DEVICE 6

typedef struct ak_potential ak_potential;
typedef struct ak_nlist ak_nlist;

int ak_abi_version(void);
const char *ak_last_error(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int ak_device_count(int32_t *count);

/* ---- Potentials ------------------------------------------------------- */

/* Tersoff in the table form of TersoffPotential._compute_multi_species
 * (engine/tersoff.py:1518-1951; the flat single-species dict of
 * TersoffPotential.__init__, engine/tersoff.py:1135-1169, is n_species = 1
 * with m = 3, gamma = 1).
 *   pair_table   (ns, ns, 9): valid, A, B, lambda1, lambda2, R, D, beta, n
 *                for the record (i, j, j);
 *   triple_table (ns, ns, ns, 7): valid, c, d, h, lambda3, m, gamma for the
 *                record (i, j, k);
 *   r_cut: max R + D over the pair records. */
int ak_tersoff_create(int32_t n_species, const double *pair_table,
                      const double *triple_table, double r_cut, ak_potential **out);

/* Stillinger-Weber, flat single-species form of StillingerWeberPotential
 * (engine/stillinger_weber.py:860-946): params = epsilon, sigma, A, B, p, q,
 * a, lambda, gamma, costheta0; r_cut = a * sigma. */
int ak_sw_create(const double *params10, ak_potential **out);

/* Stillinger-Weber, multi-species tables of StillingerWeberPotential
 * (engine/stillinger_weber.py:1255-1315, _compute_multi_species :1165-1451;
 * records as :874-890, Tomita et al. 2018 eq. 4), n_species <= 8:
 *   pair_table (ns, ns, 8): valid, epsilon, sigma, a, A, B, p, q of the
 *                record (i, j, j), valid when it exists with A != 0; a pair
 *                of atoms reads the row of its lower-indexed atom;
 *   leg_table (ns, ns, 3): sigma, a, gamma of the record (i, j, j), a = 0
 *                where that record is absent;
 *   triple_table (ns, ns, ns, 4): valid, epsilon, lambda, costheta0 of the
 *                record (i, j, k), valid when (i, j, k), (i, j, j) and
 *                (i, k, k) exist; both orders of a neighbour pair are terms,
 *                each weighted 1/2;
 *   r_cut: max a * sigma over the records (i, j, j). */
int ak_sw_multi_create(int32_t n_species, const double *pair_table, const double *leg_table,
                       const double *triple_table, double r_cut, ak_potential **out);

/* Alloy EAM with Zhou-Johnson-Wadley functions (EAMPotential built by
 * zhou_eam, engine/eam_zhou.py:128-207, evaluated as the numba kernel
 * engine/_eam_kernels.py:231-307).  functions: (n_functions, 44) rows
 * [code, r_c, width, chi, pa[20], pb[20]] with code 3 density, 4 pair,
 * 5 cross pair, 6 embedding; index_map: rho[ns], embedding[ns], phi[ns*ns]
 * row indices. */
int ak_eam_create(int32_t n_species, int32_t n_functions, const double *functions,
                  const int32_t *index_map, double r_cut, ak_potential **out);

/* Alloy EAM from tabulated functions as well: the EAMPotential that
 * read_setfl builds (engine/eam_zhou.py:257-315), whose functions are scipy
 * CubicSpline PPolys evaluated as the numba kernel does
 * (engine/_eam_kernels.py:128-145, 199-205: interval from
 * searchsorted(x, r, 'right') - 1 clipped to the end intervals, Horner).
 * functions: (n_functions, 44) rows; besides the Zhou rows of ak_eam_create,
 * code 1 = spline s(r) and code 2 = s(r) / r (a setfl r*phi table) as
 * [code, first knot, number of knots, 0...] indexing knots (n_knots) and
 * coeffs (n_knots, 4) = c0..c3 (PPoly order, highest power first) of the
 * interval starting at that knot (the last knot's row is unused). */
int ak_eam_tables_create(int32_t n_species, int32_t n_functions, const double *functions,
                         const int32_t *index_map, int64_t n_knots, const double *knots,
                         const double *coeffs, double r_cut, ak_potential **out);

/* Value and derivative of function row `function` of an EAM potential at m
 * arguments (the tabulation write_setfl does, engine/eam_zhou.py:216-254,
 * evaluated on the device). */
int ak_eam_eval(ak_potential *pot, int32_t function, int64_t m, const double *x, double *v,
                double *dv);

int ak_potential_destroy(ak_potential *pot);

/* ---- Neighbour list (NeighbourList, engine/neighborlist.py:390-714) ---- */

int ak_nlist_create(double r_cut, double skin_width, ak_nlist **out);
int ak_nlist_destroy(ak_nlist *nl);
/* NeighbourList.build (engine/neighborlist.py:574-585); inv_lattice may be
 * NULL (then computed here); passing numpy's inverse reproduces the
 * reference's fractional coordinates bit for bit. */
int ak_nlist_build(ak_nlist *nl, int32_t n_atoms, const double *lattice,
                   const double *inv_lattice, const double *positions);
/* NeighbourList.check_and_rebuild (engine/neighborlist.py:613-673). */
int ak_nlist_check_and_rebuild(ak_nlist *nl, int32_t n_atoms, const double *lattice,
                               const double *inv_lattice, const double *positions,
                               int32_t *rebuilt);
/* NeighbourList.get_candidate_pairs (engine/neighborlist.py:679-714): pairs
 * i < j within r_cut + skin at the last build, canonical (i, j) order.
 * With idx_i == NULL only *n_pairs is set. */
int ak_nlist_candidate_pairs(ak_nlist *nl, int64_t *n_pairs, int32_t *idx_i,
                             int32_t *idx_j, int64_t capacity);
int ak_nlist_counters(ak_nlist *nl, int64_t *n_builds, int64_t *n_steps);

/* ---- Force calls ------------------------------------------------------ */

/* potential.compute(lattice, positions, atom_types, neigh_list)
 * (engine/tersoff.py:1208-1228, engine/stillinger_weber.py:900-922,
 * engine/eam.py:174-256).  nl == NULL: the potential keeps a list of its
 * own with skin 0.3 A (owned_neighbour_list, engine/neighborlist.py:1103-1163).
 * Host buffers; stress may be NULL when need_stress == 0.  A structure with
 * no bond inside r_cut gives zero energy, forces and stress for every kind,
 * EAM included (engine/tersoff.py:1293-1294, engine/eam.py:236-237); n_atoms
 * may be 0. */
int ak_compute(ak_potential *pot, ak_nlist *nl, int32_t n_atoms, const double *lattice,
               const double *inv_lattice, const double *positions, const int32_t *species,
               int32_t need_stress, double *energy, double *forces, double *stress);

/* Device-resident force call: d_positions (N, 3), d_species (N), d_forces
 * (N, 3) and d_energy_stress (10: E, then the 9 stress components) are
 * device pointers; the list must have been built (ak_nlist_build) for this
 * structure.  Runs asynchronously on `stream` (a hipStream_t; NULL = the
 * potential's own stream). */
int ak_compute_device(ak_potential *pot, ak_nlist *nl, int32_t n_atoms,
                      const double *lattice, const double *inv_lattice,
                      const double *d_positions, const int32_t *d_species, double *d_forces,
                      double *d_energy_stress, void *stream);

/* ak_compute_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as ak_compute_device, but
 * d_species is required, and the kernels' scratch (geometry, launch chunks,
 * block partials, bond gradients, EAM embedding derivatives) lives in one
 * 64 MiB device array of the library, per GPU, held until the process
 * ends; inputs that need more (roughly 32 * N * list stride bytes) fail
 * with AK_ERR_CAPACITY.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams or threads) race. */
int ak_compute_device_noalloc(ak_potential *pot, ak_nlist *nl, int32_t n_atoms,
                              const double *lattice, const double *inv_lattice,
                              const double *d_positions, const int32_t *d_species,
                              double *d_forces, double *d_energy_stress, void *stream);

/* Many independent structures in one launch sequence (the phonon / elastic /
 * fitting callers of the force function, solvers/phonon.py:412-482,
 * solvers/elastic.py:245-291).  atom_offsets (n_struct + 1); lattices
 * (n_struct, 3, 3); positions/forces concatenated (sum N, 3); energies
 * (n_struct); stresses (n_struct, 3, 3).  Host buffers. */
int ak_compute_batched(ak_potential *pot, int32_t n_struct, const int64_t *atom_offsets,
                       const double *lattices, const double *inv_lattices,
                       const double *positions, const int32_t *species, double *energies,
                       double *forces, double *stresses);

/* Batched device-resident call for repeated evaluation: the first call with
 * a given layout builds the batch neighbour list; `rebuild` != 0 forces a
 * rebuild.  d_energy_stress is (n_struct, 10). */
int ak_compute_batched_device(ak_potential *pot, int32_t n_struct,
                              const int64_t *atom_offsets, const double *lattices,
                              const double *d_positions, const int32_t *d_species,
                              double *d_forces, double *d_energy_stress, int32_t rebuild,
                              void *stream);

/* Launch-shape tuning knob (the autotuner's output): EAM lanes per atom
 * (4..64, power of two; 0 = default). */
int ak_potential_set_lanes(ak_potential *pot, int32_t lanes);

/* Launch shape of the site / force kernels: the autotuner's output (no
 * counterpart in the reference, whose kernels are CPU loops).
 *   bond_tile     Tersoff site kernel form: 1 = one lane per bond, 2 = one
 *                 lane per (bond, partner leg) pair for list strides up to 8;
 *                 0 = automatic (one lane per bond);
 *   block_threads 64, 128, 256 or 512 threads per block (0 = 256); atoms
 *                 per block = block_threads / lanes per atom; 512 applies to
 *                 the four-slot Tersoff kernel (one species, flat form, rows
 *                 of four), every other kernel then runs 256;
 *   grid_cap      0 = one block per chunk of atoms; > 0 = at most this many
 *                 blocks, each looping over chunks (persistent form).  Only
 *                 the EAM force pass on general function tables (splines,
 *                 rows other than the Zhou alloy form) loops; the bond-order
 *                 kernels and the Zhou-table EAM kernels always run one block
 *                 per chunk and accept the value without using it. */
int ak_potential_set_launch(ak_potential *pot, int32_t bond_tile, int32_t block_threads,
                            int32_t grid_cap);

/* Project-written autotuner (no counterpart in the reference): builds the
 * batch list for the given structures (as ak_compute_batched), times every
 * candidate launch shape -- Tersoff site form and threads per block, or EAM
 * lanes per atom and threads per block -- over `reps` device force calls
 * with HIP events, in three rounds (each candidate keeps its best round), and
 * leaves the fastest set on the potential.  It ends with twenty force calls
 * at that shape on its own stream, each waited for, and returns with that
 * stream idle.  best (4) = bond_tile,
 * block_threads, grid_cap, eam_lanes; *best_ms = its time per call.  When
 * shapes (capacity, 4) and times_ms (capacity) are given they receive every
 * candidate tried; *n_tried counts them. */
int ak_autotune_batched(ak_potential *pot, int32_t n_struct, const int64_t *atom_offsets,
                        const double *lattices, const double *positions,
                        const int32_t *species, int32_t reps, int32_t *best, double *best_ms,
                        int32_t capacity, int32_t *shapes, double *times_ms, int32_t *n_tried);

/* ---- Molecular dynamics ------------------------------------------------ */

/* NVE velocity Verlet, device-resident (NVESimulation.run,
 * solvers/molecular_dynamics.py:1969-1996), with the Verlet-skin rebuild
 * test of engine/neighborlist.py:653-673 evaluated on the device each step.
 * positions/velocities (N, 3) host buffers, updated in place; masses (N) in
 * amu; dt in fs.  e_pot/e_kin (n_steps + 1) receive the energies after each
 * step (index 0 = initial state).  use_graph != 0 replays one captured step
 * with hipGraphLaunch.  *n_rebuilds counts list rebuilds. */
int ak_md_nve(ak_potential *pot, int32_t n_atoms, const double *lattice,
              const double *inv_lattice, double *positions, double *velocities,
              const double *masses, const int32_t *species, double dt, int32_t n_steps,
              double skin_width, int32_t use_graph, double *e_pot, double *e_kin,
              int32_t *n_rebuilds, double *seconds);

/* The same loop with a thermostat and the heat flux, for NVTSimulation.run
 * (solvers/molecular_dynamics.py:1113-1160: velocity Verlet between two
 * half-steps of an M-link Nose-Hoover chain, _nhc_step :635-727) and the
 * production run of ThermalConductivityCalculator.run_md
 * (solvers/thermal_conductivity.py:575-610), LangevinSimulation (BAOAB,
 * :1511-1587), CSVRSimulation (:1589-1687) and NPTSimulation with the
 * isotropic barostat (:2141-2555, cell_dof "iso", baro_thermostat "nhc").
 * thermostat: NULL for NVE, or {kind, temperature K, tau fs, chain links,
 * seed, pressure GPa, tau_baro fs, cell_dof, baro_thermostat,
 * velocity_rescale_steps}: kind 1 =
 * Nose-Hoover chain (tau_nhc, 2-16 links), 2 = Langevin, 3 = CSVR (tau =
 * damping; seed an integer < 2^53 keys the Philox4x32-10 stream), 4 = NPT
 * (the chain of kind 1 plus the barostat; a perpendicular cell width may
 * shrink by up to 5 %); n_dof = 3N - 3.  cell_dof (CELL_DOF_MODES,
 * :2113-2138): 0 iso, 1 aniso, 2 tri, 3 xy, 4 xy_iso, 5 z (the MTTK
 * flexible cell restricted to those strain generators, :2566-2692);
 * baro_thermostat: 0 the 2-link Nose-Hoover chain on the barostat momenta,
 * 1 Langevin (:2596-2617; noise keyed by seed, its heat subtracted from
 * H_NPT).  velocity_rescale_steps (kinds 1 and 4, :1154-1167, 2907-2920):
 * v *= sqrt(T / T_inst) after the end-of-step chain half-step of the first
 * that many steps (counted over restarts through state[45]).
 * fix_mask (3N bytes, 1 = frozen Cartesian component; NULL = none; fix_atoms
 * of every MD class, :248-267): frozen components start at zero and are
 * zeroed after every kick; n_dof = 3N - (number frozen) when any is frozen
 * (_compute_n_dof :341-361); a Langevin run with frozen components keeps
 * its centre-of-mass velocity (:1552-1565).
 * state (AK_MD_STATE_LEN doubles, may be NULL = start from rest) carries the
 * thermostat between calls, as the reference's checkpoint scalars do
 * (_ck_arrays / _CK_SCALARS, :431-445, 823-830, 2228, 2557-2564): read at
 * the start, the final state on return.  Layout: [0, 16) chain positions
 * xi, [16, 32) chain velocities vxi, 32 p_eps (iso), [33, 35) barostat
 * chain xi_b, [35, 37) vxi_b, [37, 43) the strain momenta p_k of a flexible
 * cell (in generator order), 43 the heat Q a stochastic thermostat or the
 * Langevin barostat has added, 44 the random draws used so far (a
 * continued run draws new numbers), 45 the steps run so far, rest reserved
 * (pass zeros).
 * e_cons (n_steps + 1): E_tot (NVE), H_NHC (:947-969), E_tot - Q, Q the heat
 * a stochastic thermostat added (the E-Q column), or H_NPT (:2425-2454).
 * lattice_out (9): the final cell; volume, pressure (n_steps + 1): V [A^3]
 * and P_inst = 2 E_kin / 3V + Tr(sigma) / 3 [eV/A^3] per step (every kind;
 * V is constant without a barostat).
 * rebuild_counts (n_steps + 1, may be NULL): list rebuilds of this call up to
 * each recorded state (the thermo log's NL_builds, less the initial build).
 * cell_history (n_steps + 1, 3, 3) and ptensor_history (n_steps + 1, 6 =
 * xx, yy, zz, yz, xz, xy [eV/A^3]; may be NULL): the cell of each record and,
 * under a flexible-cell NPT, the pressure tensor sum m v v^T / V + sym(sigma)
 * of the recorded velocities (_pressure_tensor, _extra_values :2570-2577,
 * 3027-3037; NaN for other kinds).
 * flux_groups (frames, n_labels * n_regions, 3; NULL = none; needs
 * flux_every > 0): the heat flux of groups of atoms, from the per-atom heat
 * currents (each entry's virial term to the atom whose velocity it carries;
 * ThermalConductivityCalculator._per_atom_heat_current and _group_labels,
 * solvers/thermal_conductivity.py:305-363); atom a belongs to group
 * flux_labels[a] * n_regions + its region (flux_labels may be NULL with
 * n_labels = 1): region_axis 0/1/2 (-1 none, n_regions = 1) and the
 * fractional coordinate, wrapped, between region_bounds[0..n_regions]
 * (searchsorted right - 1, clipped).  The groups sum to flux.
 * atom_flux (frames, N, 3, float32; NULL = none; needs flux_every > 0): the
 * per-atom heat currents J_a / V of every recorded frame, rounded to float32
 * (run_md's save_atom_flux, atom_heat_flux.npy, :551-558, 637-645).
 * frame_every > 0 with frames != NULL: the positions after every
 * frame_every steps, (n_steps / frame_every, N, 3), recorded inside the
 * captured step into device memory (the trajectory frames of traj_every,
 * :128-129, 452-463), and in frame_lattices (n_steps / frame_every, 3, 3;
 * may be NULL) the cell of each frame.
 * flux_every > 0: J [eV/(A^2 fs)] of the state after every step s with
 * s % flux_every == 0 into flux ((n_steps + flux_every - 1) / flux_every, 3),
 * the exact site-term heat flux of ak_heat_flux.  Any output may be NULL. */
int ak_md_run(ak_potential *pot, int32_t n_atoms, const double *lattice,
              const double *inv_lattice, double *positions, double *velocities,
              const double *masses, const int32_t *species, double dt, int32_t n_steps,
              double skin_width, int32_t use_graph, const double *thermostat10,
              double *state, const uint8_t *fix_mask, int32_t flux_every,
              int32_t frame_every, double *frames, double *frame_lattices, double *e_pot,
              double *e_kin, double *e_cons, double *flux, double *lattice_out,
              double *volume, double *pressure, int32_t *rebuild_counts,
              double *cell_history, double *ptensor_history, const int32_t *flux_labels,
              int32_t n_labels, int32_t region_axis, const double *region_bounds,
              int32_t n_regions, double *flux_groups, float *atom_flux, int32_t *n_rebuilds,
              double *seconds);
#define AK_MD_STATE_LEN 48

/* Non-equilibrium MD with two Langevin heat baths (NEMDSimulation.run,
 * solvers/nemd.py:150-232, replacing its velocity Verlet loop, _o_step
 * :113-117 and _sample :239-256): n_steps of velocity Verlet on the device,
 * after each step an exact Ornstein-Uhlenbeck update of the bath atoms,
 * v <- c v + sqrt((1 - c^2) k_B T / m) xi, c = exp(-dt / damping), xi from
 * the Philox4x32-10 stream keyed by seed (blocks (atom, draw, 7, 0/1); the
 * reference draws from numpy's generator instead).  bath (N bytes): 0 none,
 * 1 hot (T_hot), 2 cold (T_cold).  state (AK_MD_STATE_LEN, may be NULL) as
 * for ak_md_run: [43] the heat Q_hot + Q_cold so far, [44] the draws used.
 * e_pot, e_kin, e_cons = E_tot - Q (n_steps + 1, record 0 the initial
 * state); bath_history (n_steps + 1, 4): the heat each bath added in the
 * step (dQ_hot, dQ_cold) and the kinetic energy of each bath after it.
 * sample_every > 0: after the steps s (0-based in this call) with
 * (s + sample_phase) % sample_every == 0, per bin of n_bins equal bins of
 * the wrapped fractional coordinate along axis (min(floor(f n_bins),
 * n_bins - 1)) the sums (sum_i J_i / V (3), kinetic energy, atom count)
 * into bins (samples, n_bins, 5), J_i the per-atom heat currents of
 * ak_md_run's flux_groups.  Any output may be NULL. */
int ak_nemd_run(ak_potential *pot, int32_t n_atoms, const double *lattice,
                const double *inv_lattice, double *positions, double *velocities,
                const double *masses, const int32_t *species, double dt, int32_t n_steps,
                double skin_width, int32_t use_graph, const uint8_t *bath, double T_hot,
                double T_cold, double damping, double seed, double *state,
                int32_t sample_every, int32_t sample_phase, int32_t axis, int32_t n_bins,
                double *e_pot, double *e_kin, double *e_cons, double *bath_history,
                double *bins, int32_t *n_rebuilds, double *seconds);

/* ---- Site terms and heat flux ------------------------------------------ */

/* Site-energy decomposition E = sum_a U_a (potential.site_terms,
 * engine/site_terms.py:1-60, 179-475): e_site (n) = U_a and, for every bond
 * a -> nbr within r_cut (both directions, atoms in order, each atom's bonds
 * by ascending neighbour), centre, neighbour, vec = r_nbr - r_a (minimum
 * image) and dudv = dU_centre / dvec.  Pass centre = NULL to learn the number
 * of entries (*n_entries) first; AK_ERR_CAPACITY if capacity is too small.
 * nl as in ak_compute (NULL: the owned list). */
int ak_site_terms(ak_potential *pot, ak_nlist *nl, int32_t n_atoms, const double *lattice,
                  const double *inv_lattice, const double *positions, const int32_t *species,
                  double *e_site, int64_t capacity, int32_t *centre, int32_t *neighbour,
                  double *vec, double *dudv, int64_t *n_entries);

/* Heat flux J = [sum_a E_a v_a - sum_m vec_m (dudv_m . v_nbr)] / V
 * [eV/(A^2 fs)], E_a = m_a |v_a|^2 / 2 + U_a (Fan et al., PRB 92, 094301
 * (2015), eq. 21), as ThermalConductivityCalculator._compute_heat_flux with
 * the site-term virial flux (solvers/thermal_conductivity.py:168-303);
 * velocities A/fs, masses amu.  e_site (n) may be NULL. */
int ak_heat_flux(ak_potential *pot, ak_nlist *nl, int32_t n_atoms, const double *lattice,
                 const double *inv_lattice, const double *positions,
                 const double *velocities, const double *masses, const int32_t *species,
                 double *flux3, double *e_site);

/* ---- Tersoff parameter gradient ---------------------------------------- */

/* Loss of fitting/loss.py:184-242 (squared) over a batch of structures with
 * the 14 Tersoff parameters theta = (A, B, lambda1, lambda2, lambda3, beta,
 * n, c, d, h, R, D, m, gamma) of one species, and its gradient dL/dtheta by
 * the hand-written backward kernel (no finite differences).
 * ref_energy (n_struct) in eV/atom, ref_forces (sum N, 3), ref_stress
 * (n_struct, 3, 3); weights = w_E, w_F, w_S, r0, eps0.  grad may be NULL.
 * Per-structure E, F, sigma are written when the pointers are non-NULL. */
int ak_tersoff_loss_grad(ak_potential *pot, int32_t n_struct, const int64_t *atom_offsets,
                         const double *lattices, const double *positions,
                         const double *theta14, const double *ref_energy,
                         const double *ref_forces, const double *ref_stress,
                         const double *weights5, double *loss, double *grad14,
                         double *energies, double *forces, double *stresses);

/* The same loss in two steps, for optimisers that call it many times on one
 * batch: ak_tersoff_fit_load uploads the structures and references once
 * (kept on the potential handle), ak_tersoff_fit_eval evaluates loss and
 * gradient at theta14 (grad14 may be NULL) with no host-side staging of the
 * batch.  No counterpart in the reference, which has no gradient. */
int ak_tersoff_fit_load(ak_potential *pot, int32_t n_struct, const int64_t *atom_offsets,
                        const double *lattices, const double *positions,
                        const double *ref_energy, const double *ref_forces,
                        const double *ref_stress);
int ak_tersoff_fit_eval(ak_potential *pot, const double *theta14, const double *weights5,
                        double *loss, double *grad14, double *energies, double *forces,
                        double *stresses);

/* Per-structure terms and loss type of the loaded batch, the options of
 * generic_loss_function (fitting/loss.py:51-248): terms (n_struct, 4) =
 * [trust weight w_s ("weight", linear), unit scale c_s ("scale",
 * squared), stress volume V_s ("volume", default det(lattice)), flags (1:
 * the structure has reference forces, 2: reference stress)];
 * loss_type 0 squared, 1 Huber with huber_deltas = (energy, forces,
 * stress) in eV applied per structure and term (fitting/loss.py:27-48,
 * 166-175).  ak_tersoff_fit_load resets them to weight 1, scale 1, the cell
 * volume, both references present, squared loss. */
int ak_tersoff_fit_terms(ak_potential *pot, int32_t n_struct, const double *terms,
                         int32_t loss_type, const double *huber_deltas);

/* ---- Measurement -------------------------------------------------------- */

/* No counterpart in the reference: timing support for benchmarks.  While
 * enabled, every force call records HIP events around its dominant kernel
 * (the Tersoff/SW site kernel, the EAM force pass) on the launch stream;
 * ak_profile_read waits for them and returns the summed and the smallest
 * duration in milliseconds and the number of launches, then resets. */
int ak_profile_enable(ak_potential *pot, int32_t on);
int ak_profile_read(ak_potential *pot, double *total_ms, double *min_ms, int64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* AKITSUSHIMA_AMD_H */
