/*
 * pdx.h — C ABI of libpdx, the MI355X (gfx950) force hot path of ParticleDynamics.jl.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and returns an
 * int status (PDX_OK = 0). Device arrays are raw device pointers owned by the caller
 * ("It holds no references to particle arrays", src/electrostatics/Solver.jl:67-74);
 * the library owns only its workspaces, allocated at *_create and never on a warm call.
 * Enqueue functions never synchronise; pdx_nl_status / pdx_*_totals are the only
 * device-to-host reads (src/neighbors/DeviceNeighborCUDA.jl:639-651).
 *
 * Layouts follow the reference bit for bit where other consumers read them:
 *   - ELL rows slot-major, 1-based ids: slot t of particle i at neighbors_flat[t*N + i]
 *     (src/neighbors/NeighborTypes.jl:28-31);
 *   - cell_offsets 1-based, last entry N+1; particle_ids_sorted 1-based
 *     (src/neighbors/NeighborAllocation.jl:133-150);
 *   - per-particle virial V is N x 6 column-major (xx, yy, zz, xy, xz, yz)
 *     (src/nonbonded/LJELLEngineKernels.jl:75-77);
 *   - mesh grids column-major, x fastest (src/electrostatics/MeshKernels.jl:4-5),
 *     half spectrum (K1/2+1, K2, K3) (src/electrostatics/Solver.jl:48-49).
 * Coordinates are centred, in [-L/2, L/2) (src/geometry/PeriodicDisplacement.jl:21-25).
 */
#ifndef PDX_H
#define PDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (host-side mirrors of the reference's exceptions) ---- */
#define PDX_OK              0
#define PDX_ERR_ARGUMENT    1  /* ArgumentError (e.g. NeighborAllocation.jl:5-26, Spec.jl:60-101) */
#define PDX_ERR_DIMENSION   2  /* DimensionMismatch */
#define PDX_ERR_CAPACITY    3  /* NeighborCapacityError (NeighborTypes.jl:41-50) */
#define PDX_ERR_COORDINATE  4  /* non-finite / uncentred coordinates (DeviceNeighborCUDA.jl:641) */
#define PDX_ERR_DEVICE      5  /* HIP runtime failure */
#define PDX_ERR_STATE       6  /* no valid list (DeviceNeighborCUDA.jl:645) */

/* ---- device state words of a neighbor list, 0-based (DN_* of DeviceNeighborCUDA.jl:4-33, minus 1) ---- */
#define PDX_DN_REQUEST             0
#define PDX_DN_VALID               1
#define PDX_DN_ERROR               2
#define PDX_DN_STEP                3
#define PDX_DN_LAST_BUILD          4
#define PDX_DN_INTERVAL            5
#define PDX_DN_CHECKS              6
#define PDX_DN_BUILDS              7
#define PDX_DN_REQUIRED            8
#define PDX_DN_REASON              9
#define PDX_DN_ATTEMPTS           10
#define PDX_DN_DISPLACEMENT_BUILDS 11
#define PDX_DN_AGE_BUILDS         12
#define PDX_DN_INVALID_BUILDS     13
#define PDX_DN_SORT_BUILDS        14
#define PDX_DN_OVERLAP_BUILDS     15
#define PDX_DN_LAST_CANDIDATES    16
#define PDX_DN_TOTAL_CANDIDATES   17
#define PDX_DN_LAST_ACCEPTED      18
#define PDX_DN_TOTAL_ACCEPTED     19
#define PDX_DN_MAX_ROW            20
#define PDX_DN_DIAGNOSTICS        21
#define PDX_DN_SECTION            22
#define PDX_DN_NSTATE             23

#define PDX_DN_REASON_INVALID      1
#define PDX_DN_REASON_DISPLACEMENT 2
#define PDX_DN_REASON_AGE          4
#define PDX_DN_CAPACITY_ERROR      1
#define PDX_DN_COORDINATE_ERROR    2

/* ---- neighbor-list flags ---- */
#define PDX_NL_ALLPAIRS     1  /* rows from all pairs, one image (DenseNeighborCUDA.jl:128-151) */
#define PDX_NL_DIAGNOSTICS  2  /* count candidates / accepted pairs (DeviceNeighborCUDA.jl:17-22) */
#define PDX_NL_2D           4  /* D = 2 (DenseNeighborCUDA.jl:77): box[2] is ignored, z may be NULL (all z = 0);
                                  rows, pair forces and the rebuild policy as in 3-D, fz = 0; no mesh */

/* ---- pair models (src/engine/PairModels.jl:13-14, 45-54) ---- */
#define PDX_PAIR_NONE  0
#define PDX_PAIR_LJ    1  /* f/r = 24e(2s^12-s^6)/r^2, u = 4e(s^12-s^6) - u_c */
#define PDX_PAIR_WCA   2  /* same force, u + e, rc^2 = 1.2599211 sigma^2 */
#define PDX_PAIR_SOFT  3  /* harmonic repulsion, rc = sigma */
/* typed models (PairModels.jl:16-19, 61-81); per-particle data set by pdx_pair_set_sigmas /
 * pdx_pair_set_types, ELL rows only (lanes >= 0, rows traversal, FP32 sums) */
#define PDX_PAIR_MIXED_LJ    4  /* MixedSigmaLJ: sigma_ij = (sigma_i + sigma_j)/2, rc = cutoff * sigma_ij
                                   (spec.cutoff is the rcut factor), one epsilon; with shift, u - u(rc)
                                   (the same for every pair: sigma/rc is constant) */
#define PDX_PAIR_MIXED_WCA   5  /* MixedSigmaWCA: the same pairs, u + epsilon */
#define PDX_PAIR_MATRIX_LJ   6  /* PairMatrixLJ{SHIFT}: sigma, epsilon, rc by type pair; shift: u - u(rc_ab) */
#define PDX_PAIR_MATRIX_WCA  7  /* PairMatrixWCA: the same, u + epsilon_ab */

/* ---- evaluation flags ---- */
#define PDX_ENERGY  1  /* also per-particle energy / virial and FP64 totals */
#define PDX_ADD     2  /* add into fx, fy, fz, Epot, V instead of overwriting */

/* A hipStream_t passed as an opaque pointer (NULL = the default stream). */
typedef void *pdx_stream_t;
typedef struct pdx_nl_s *pdx_nl_t;
typedef struct pdx_pair_s *pdx_pair_t;
typedef struct pdx_mesh_s *pdx_mesh_t;

/* Library version (major*10000 + minor*100 + patch) and the last error message of this thread. */
int pdx_version(void);
const char *pdx_last_error(void);

/* ======================================================================= */
/* Neighbor lists: counting-sort cell list + slot-major ELL rows           */
/* ======================================================================= */

/* Device pointers into a list's storage (read-only for callers). */
typedef struct {
  int32_t *neighbors_flat;      /* N*cap, slot-major, 1-based ids */
  int32_t *counts;              /* N (row lengths; kept past cap to report overflow) */
  int32_t *cell_offsets;        /* ncell+1, 1-based, last N+1 */
  int32_t *particle_ids_sorted; /* N, 1-based, cell order (stable: ascending id within a cell) */
  int32_t *cell_of_particle;    /* N, 0-based cell id (cz*ny+cy)*nx+cx */
  float *rref_x, *rref_y, *rref_z; /* reference positions of the last build */
  float *packed;                /* N float4 (x, y, z, q) in storage order, written by pdx_nl_update/_rebuild */
  int64_t *state;               /* PDX_DN_NSTATE words */
} pdx_nl_arrays;

/* Allocate a device list for N particles in a centred box (DeviceNeighborWorkspace,
 * DeviceNeighborCUDA.jl:64-82; _alloc_neighbor_matrix, NeighborAllocation.jl:28-83).
 * cell = cutoff + skin; grid floor(L/cell) >= 3 per dimension unless PDX_NL_ALLPAIRS,
 * which requires cutoff + skin <= L/2. interval: initial adaptive rebuild interval, 1..100. */
int pdx_nl_create(int32_t n, const double box[3], double cutoff, double skin, int32_t cap,
                  int32_t interval, int32_t flags, pdx_nl_t *out);
int pdx_nl_destroy(pdx_nl_t nl);
/* Grid, capacity and geometry: out[0..7] = N, cap, nx, ny, nz, ncell, allpairs, diagnostics. */
int pdx_nl_info(pdx_nl_t nl, int32_t out[8]);
int pdx_nl_arrays_get(pdx_nl_t nl, pdx_nl_arrays *out);

/* Coverage check + predicated rebuild on the device (update_device_neighbors!,
 * DeviceNeighborCUDA.jl:421-427): max |dr|^2 since the last build > skin^2/4, or the
 * build age >= the adaptive interval, requests a rebuild; every rebuild kernel is
 * enqueued and returns at once when none is requested (graph-capturable). Also packs
 * (x, y, z, q) into arrays.packed; q may be NULL (w = 0). */
int pdx_nl_update(pdx_nl_t nl, const float *x, const float *y, const float *z, const float *q,
                  pdx_stream_t stream);
/* The two halves of pdx_nl_update, for callers that overlap work with the rebuild: the
 * displacement pass with the packing of this step's coordinates (and the cell snapshot), then
 * the decision and the predicated rebuild (decide_device_neighbors! + _device_rebuild!,
 * DeviceNeighborCUDA.jl:421-427). Work that needs only the packed coordinates or the snapshot
 * may run on another stream between the two. */
int pdx_nl_check(pdx_nl_t nl, const float *x, const float *y, const float *z, const float *q,
                 pdx_stream_t stream);
int pdx_nl_build(pdx_nl_t nl, pdx_stream_t stream);
/* The list keeps a snapshot of the cells (offsets, ids) of its last completed build: the
 * coverage check (pdx_nl_check / pdx_nl_update) copies the cells of a build completed since
 * the previous check, before this step's rebuild can overwrite them (no counterpart in the
 * reference: it lets a spread read cells that no concurrent rebuild writes; see
 * pdx_mesh_spread_snapshot). pdx_nl_snapshot takes the copy explicitly (predicated on the
 * device build counter, graph-capturable); the check makes it unnecessary. */
int pdx_nl_snapshot(pdx_nl_t nl, pdx_stream_t stream);
/* Velocity-Verlet step around a force evaluation (the reference's fused NVE step: prefix
 * kick-drift + displacement, src/engine/Fusion.jl:1-6, Run.jl:247-294):
 * pdx_nl_check_kick_drift = v += c_i f, x += dt v wrapped into [-L/2, L/2)
 * (_nve_kick_drift3_kernel!, src/simulation/NVEFusion.jl:3-24), then pdx_nl_check on the new x
 * (x, y, z, vx, vy, vz updated in place; fx, fy, fz hold the forces of the old x);
 * pdx_nve_kick = the closing v += c_i f with the new forces (_nve_kick3_kernel!,
 * src/SimulationCore.jl:581-591), advancing the list's step counter. c_i = dt / (2 m): inv_mass
 * (device, per particle, c_i = dt/2 inv_mass[i]) or, when NULL, the scalar mass. Both apply
 * only while the list is valid (a latched error stops the motion). 2-D lists ignore z arrays. */
int pdx_nl_check_kick_drift(pdx_nl_t nl, float *x, float *y, float *z, const float *q, float *vx, float *vy,
                            float *vz, const float *fx, const float *fy, const float *fz, double dt, double mass,
                            const float *inv_mass, pdx_stream_t stream);
int pdx_nve_kick(pdx_nl_t nl, float *vx, float *vy, float *vz, const float *fx, const float *fy, const float *fz,
                 double dt, double mass, const float *inv_mass, pdx_stream_t stream);
/* Unconditional rebuild (build_neighbors_dense!, NeighborAllocation.jl:198-214). */
int pdx_nl_rebuild(pdx_nl_t nl, const float *x, const float *y, const float *z, const float *q,
                   pdx_stream_t stream);
/* Spatial order of the list's last build (the permutation OrderingWorkspace applies,
 * src/engine/Ordering.jl:19-79): perm (device, N) receives, slot by slot, the 0-based id of the
 * particle that belongs there, in cell order (cells z, y, x) and inside a cell by position (x, y,
 * z; the id only for coincident particles), so the order is a function of the configuration, not
 * of the labels. The caller permutes its particle arrays by it (new[s] = old[perm[s]]) and then
 * calls pdx_nl_invalidate. Host boundary (reads the state; PDX_ERR_STATE without a valid list). */
int pdx_nl_spatial_order(pdx_nl_t nl, int32_t *perm, pdx_stream_t stream);
/* The particles were relabelled (reordered): the next check rebuilds, and no cell snapshot is kept
 * until the second check after it. Host boundary. */
int pdx_nl_invalidate(pdx_nl_t nl, pdx_stream_t stream);
/* Advance the device step counter (advance_device_neighbor_step!, DeviceNeighborCUDA.jl:404-407). */
int pdx_nl_advance(pdx_nl_t nl, pdx_stream_t stream);
/* Clear a latched capacity error and request the failed rebuild again (rearm_capacity_rebuild!, :622). */
int pdx_nl_rearm(pdx_nl_t nl, pdx_stream_t stream);
/* Replace the row storage by one of capacity cap (resize_capacity!, :628-635). Host boundary. */
int pdx_nl_resize_capacity(pdx_nl_t nl, int32_t cap);
/* Change the box (set_device_box!, :89-116); the next check rebuilds. Host boundary. */
int pdx_nl_set_box(pdx_nl_t nl, const double box[3], pdx_stream_t stream);
/* Lanes per particle of the row build (the builder lanes of ENGINE_SETTINGS, src/engine/
 * StateEngine.jl:7-27): 1, 2, 4, 8, 16 or 32; the default fills about 2^19 threads. The rows are
 * the same set per particle whatever the lanes. A launch shape, set before capturing a graph. */
int pdx_nl_set_row_lanes(pdx_nl_t nl, int32_t lanes);
/* Boundary readback of the state words (device_neighbor_status, :639-651). Synchronises
 * `stream`. Returns PDX_ERR_CAPACITY / PDX_ERR_COORDINATE / PDX_ERR_STATE when check != 0
 * and the list is unusable; state is filled in every case. */
int pdx_nl_status(pdx_nl_t nl, int64_t state[PDX_DN_NSTATE], int32_t check, pdx_stream_t stream);

/* ======================================================================= */
/* Pair forces over ELL rows (LJ / WCA / soft, optional real-space Ewald)  */
/* ======================================================================= */

typedef struct {
  int32_t model;          /* PDX_PAIR_* */
  int32_t shift;          /* LJ potential-shift modifier (PairModels.jl:20-25) */
  double epsilon, sigma, cutoff;   /* cutoff: LJ only; WCA/SOFT fix their own */
  /* real-space Ewald Coulomb, C q_i q_j (1/r + P_e(r^2)) (PairModels.jl:90-110, PairFunctor.jl:127-156) */
  int32_t coulomb;        /* 0 off, 1 Gaussian (Ewald) split, 2 prolate split (ESP; split_shape) */
  int32_t coulomb_shift;  /* potential-shift modifier on the Coulomb energy (Spec.jl:25-29) */
  double coulomb_C;       /* prefactor / relative permittivity */
  double coulomb_cutoff;  /* real-space cutoff */
  double alpha;           /* Ewald splitting parameter */
  int32_t lanes;          /* traversal: 0 automatic; 1..64 lanes of a wave per particle; -2,-4,-8,-16:
                             blocks of 64 particles whose rows are split over that many waves */
  double split_shape;     /* prolate split: bandlimit c of ProlateSplit(cutoff, c) (Splits.jl:48-60); the
                             real-space pair C q_i q_j E_s(r/rc)/r vanishes at rc, self coefficient chi(0)/rc */
} pdx_pair_spec;

/* Build a pair model (UniformLJ/UniformWCA/ChargedPairs constructors, PairModels.jl:16-18,
 * 84-89; coulomb_polynomial, PairFunctor.jl:65-84: our own Chebyshev fit of the Ewald
 * long-range correction in r^2, FP32 Horner on device). */
int pdx_pair_create(const pdx_pair_spec *spec, pdx_pair_t *out);
int pdx_pair_destroy(pdx_pair_t p);
/* Polynomial sizes and fit error: out[0]=n_pe, out[1]=n_pf, out[2]=max rel err (energy), out[3]=(force). */
int pdx_pair_info(pdx_pair_t p, double out[4]);
/* Traversal shape of later pair-force calls (the force lanes of ENGINE_SETTINGS, src/engine/
 * StateEngine.jl:7-27): the `lanes` of pdx_pair_spec (0 automatic). A launch shape, set before
 * capturing a graph: the first call with more workgroups may grow the energy workspace. */
int pdx_pair_set_lanes(pdx_pair_t p, int32_t lanes);

/* Forces (+ per-particle half energies and half virials with PDX_ENERGY) over the rows of nl
 * (_lj3_ell_kernel_virial!, LJELLEngineKernels.jl:44-79; _pair_cooperative_kernel!,
 * PairTraversal.jl:66-125). Positions are read from arrays.packed (written by pdx_nl_update).
 * Epot, V may be NULL. totals (device, 7 doubles: E, Wxx, Wyy, Wzz, Wxy, Wxz, Wyz), may be
 * NULL, receives FP64 sums of the per-particle values (written, or added with PDX_ADD).
 * With Coulomb, Epot includes the self energy -C alpha/sqrt(pi) q_i^2. No-op while the list
 * is invalid (_device_if_valid!, DeviceNeighborCUDA.jl:157-162). */
int pdx_pair_forces(pdx_pair_t p, pdx_nl_t nl, float *fx, float *fy, float *fz, float *Epot,
                    float *V, double *totals, int32_t flags, pdx_stream_t stream);

/* Per-particle diameters of the mixed-sigma models (data[1] of MixedSigmaLJ/WCA,
 * PairModels.jl:36): n host values in storage order, finite and positive, copied to the device.
 * Host boundary (synchronises); set again after the caller reorders its particles. */
int pdx_pair_set_sigmas(pdx_pair_t p, int32_t n, const float *sigma);
/* Type data of the pair-matrix models (lj_type_table, src/clusterpairs/PairFunctor.jl:166-179;
 * data of PairMatrixLJ/WCA, PairModels.jl:37): ntypes x ntypes column-major sigma, epsilon and
 * rcut (host, 1 <= ntypes <= 64), and the 1-based type id of each of the n particles (host,
 * storage order). Copied to the device; host boundary. */
int pdx_pair_set_types(pdx_pair_t p, int32_t n, int32_t ntypes, const int32_t *typ, const double *sigma,
                       const double *epsilon, const double *rcut);
/* The largest pair cutoff of the model (its list must reach it): uniform models their cutoff,
 * mixed sigma the factor times the largest diameter, pair matrices the largest rcut; with
 * Coulomb at least the real-space cutoff. Needs the sigmas / types of a typed model first. */
int pdx_pair_cutoff(pdx_pair_t p, double *rc);

/* pdx_pair_forces with the position source explicit (the reference's pair-force call takes the
 * coordinate arrays, Run._enqueue_pair_force!, src/engine/Run.jl:215-238): x, y, z, q must be the
 * arrays the last pdx_nl_update / pdx_nl_check of nl packed (compared by address; PDX_ERR_STATE
 * otherwise, q may be NULL when that check had none). The kernels read the packed copy. */
int pdx_pair_compute(pdx_pair_t p, pdx_nl_t nl, const float *x, const float *y, const float *z, const float *q,
                     float *fx, float *fy, float *fz, float *Epot, float *V, double *totals, int32_t flags,
                     pdx_stream_t stream);

/* ======================================================================= */
/* Smooth PME mesh engine (B-spline spread, 3D FFT, influence, gather)     */
/* ======================================================================= */

#define PDX_INFLUENCE_SPME   1  /* smooth PME: G = (C/V) g(k) D1 D2 D3, analytic differentiation */
#define PDX_INFLUENCE_P3M_AD 2  /* optimal influence for analytic differentiation (alias sums) */
#define PDX_INFLUENCE_P3M_IK 3  /* optimal influence for ik differentiation, three field transforms */
#define PDX_INFLUENCE_ESP    4  /* ESP: prolate split and window, optimal AD influence (alias sums) */
#define PDX_INFLUENCE_ESP_DECONVOLUTION 5  /* ESP with the plain window deconvolution */
#define PDX_NET_CHARGE_ERROR 0
#define PDX_NET_CHARGE_BACKGROUND 1

typedef struct {
  double coulomb_prefactor;      /* carries the E*L/Q^2 unit */
  double relative_permittivity;  /* C = prefactor / eps_r */
  double real_cutoff;
  double alpha;                  /* Gaussian split; <= 0: from real_space_tolerance */
  double real_space_tolerance;   /* erfc(alpha rc) = tol (alpha_from_tolerance, Spec.jl:154-163) */
  int32_t order;                 /* B-spline order P, 3..12; ESP: prolate window width, 4..8 */
  int32_t mesh[3];               /* K1, K2, K3; mesh[0] <= 0: from mesh_spacing (resolve_mesh, Spec.jl:141-146) */
  double mesh_spacing;
  int32_t influence;             /* PDX_INFLUENCE_* (INFLUENCE_KINDS, Spec.jl:5-24) */
  int32_t net_charge;            /* PDX_NET_CHARGE_* */
  double net_charge_tolerance;
  int32_t real_space_modifier;   /* 0 none, 1 potential shift */
  int32_t alias_order;           /* P3M / ESP alias sums |n_d| <= alias_order, 0..6 (Spec.jl:80) */
  /* ESP (ProlateSplit / ProlateWindow, Splits.jl:30-60, Windows.jl:51-80): the split's bandlimit c
   * (the real-space part vanishes at real_cutoff) and the window's c1; the window must satisfy
   * order <= 2 c1/pi - 1/2 (Spec.jl:74-75) and be resolved in FP32 (its Nyquist attenuation at
   * least 0.02 eps32^(2/3), Windows.jl:310-330). Ignored by the other influence kinds. */
  double split_shape;
  double window_shape;
} pdx_mesh_spec;

/* Allocate the mesh workspace and compute the FP64 influence function for the box
 * (ElectrostaticsSolver, Solver.jl:129-158; compute_influence!, Influence.jl:190-227). */
int pdx_mesh_create(const pdx_mesh_spec *spec, int32_t n, const double box[3], pdx_mesh_t *out);
int pdx_mesh_destroy(pdx_mesh_t m);
/* ESP defaults of a target relative force error (the ESP descriptor's rules for an FP32 state,
 * src/workflow/Electrostatics.jl:102-215,288-301): out[0] = split shape c with psi(1)/||psi|| = tol,
 * out[1] = window width (order > 0 keeps it; else Table 2 of Liang et al.: 5, 5, 6, 7, 8 for
 * 1e-3 .. 1e-5), out[2] = window shape c(tol/2), or the smallest larger shape an FP32 mesh
 * resolves. The default mesh spacing is pi real_cutoff / c. Host only. */
int pdx_esp_rule(double tolerance, int32_t order, double out[3]);
/* Resolved parameters: out[0..8] = K1, K2, K3, P, alpha, C, self_coefficient, bg (per unit q),
 * lines per plane of the plane transforms (0: line passes). */
int pdx_mesh_info(pdx_mesh_t m, double out[9]);
/* Host charges: total charge, sum q^2, net-charge policy (set_solver_charges!, Solver.jl:166-185).
 * Also fixes the fixed-point scale of the spread from sum |q|. */
int pdx_mesh_set_charges(pdx_mesh_t m, const float *q_host, int32_t n);
/* New box, same mesh: influence function recomputed (refresh_geometry!, Solver.jl:193-209). */
int pdx_mesh_set_box(pdx_mesh_t m, const double box[3]);
/* Spread q into the int64 fixed-point grid (_spread_lines_kernel!, MeshKernels.jl:247-276),
 * deterministic: integer atomics commute. order: 1-based traversal order (e.g.
 * particle_ids_sorted) or NULL for storage order. The grid is cleared first. */
int pdx_mesh_spread(pdx_mesh_t m, const float *x, const float *y, const float *z, const float *q,
                    const int32_t *order, pdx_stream_t stream);
/* Spread from a neighbor list's cells, owner computes: each workgroup accumulates one 16^3
 * block of the grid in LDS (64-bit fixed point, exact) and stores it; no global atomics. On
 * meshes holding at least one particle per 12 nodes (whose tile blocks fit the 1 GiB set aside
 * at pdx_mesh_create) a tile takes only the particles whose base node it holds, into a block
 * with the stencil's halo, and a second pass sums the blocks' overlapping entries; otherwise it
 * takes every particle reaching it. Positions and charges are the list's packed (x, y, z, q) of
 * the last pdx_nl_update on the same step; the list must cover the mesh's particles and box.
 * Same grid as pdx_mesh_spread (bitwise). */
int pdx_mesh_spread_cells(pdx_mesh_t m, pdx_nl_t nl, pdx_stream_t stream);
/* The owner-computes spread from nl's cell snapshot instead of its current cells, with a full
 * skin of margin: it may run on another stream beside pdx_nl_build of the same step (the
 * snapshot's build is at most one step behind a rebuild trigger). The positions are those
 * packed by this step's pdx_nl_check. Returns PDX_ERR_STATE before the second check after the
 * list's creation or a box change (the first builds, the second snapshots the build). Same
 * grid as pdx_mesh_spread. */
int pdx_mesh_spread_snapshot(pdx_mesh_t m, pdx_nl_t nl, pdx_stream_t stream);
/* Workgroup size of the owner-computes spreads that take every particle reaching a tile (sparse
 * meshes; the dense meshes' own-particle form always runs 1024): 512 (the default: 16 particles
 * at a time per tile), 1024 (32 at a time, twice the waves to launch) or 256 (more room for a
 * kernel running beside it). */
int pdx_mesh_set_spread_threads(pdx_mesh_t m, int32_t threads);
/* Workgroup size of the plane passes of pdx_mesh_solve (meshes whose planes fit LDS): 1024 (the
 * default, the faster transform on its own) or 512 (smaller workgroups that share the CUs better
 * with a kernel running beside the transform, such as the pair forces of the same evaluation).
 * The same potential bit for bit; the FP64 energy and virial to FP64 rounding (their partial
 * sums are grouped by workgroup). No counterpart in the reference (a launch shape). */
int pdx_mesh_set_transform_threads(pdx_mesh_t m, int32_t threads);
/* Forward R2C 3D FFT, multiply by G, FP64 mesh energy and virial (with PDX_ENERGY), inverse
 * C2R FFT into the potential grid (mesh_transform!, Solver.jl:305-341; _solve_kernel!,
 * MeshKernels.jl:57-81). */
int pdx_mesh_solve(pdx_mesh_t m, int32_t flags, pdx_stream_t stream);
/* Forces F_i = -q_i grad phi (and E_i += q_i phi/2 + bg q_i with PDX_ENERGY), added into
 * fx.. (or written with set != 0) (_gather_ad_lines_kernel!, MeshKernels.jl:313-333).
 * Vglobal (6 floats, may be NULL) receives the reciprocal virial + background term
 * (_store_global_virial_kernel!, Solver.jl:382-393). */
int pdx_mesh_gather(pdx_mesh_t m, const float *x, const float *y, const float *z, const float *q,
                    const int32_t *order, float *fx, float *fy, float *fz, float *Epot,
                    float *Vglobal, int32_t flags, int32_t set, pdx_stream_t stream);
/* pdx_mesh_gather at the positions of nl's packed array, for a mesh whose last spread was an
 * owner-computes spread (pdx_mesh_spread_cells / _snapshot) from nl on the same positions
 * (PDX_ERR_STATE otherwise): that spread listed each particle under the 16^3 (or 8^3) tile of
 * its base node, and one workgroup per tile gathers its particles from the tile's potential
 * staged in LDS (same forces as pdx_mesh_gather up to FP32 rounding). Smooth PME and P3M-ad
 * meshes (the ik influence has its own gather). Same role as pdx_mesh_gather:
 * _gather_ad_lines_kernel! (MeshKernels.jl:313-333). */
int pdx_mesh_gather_cells(pdx_mesh_t m, pdx_nl_t nl, float *fx, float *fy, float *fz, float *Epot,
                          float *Vglobal, int32_t flags, int32_t set, pdx_stream_t stream);
/* Boundary readback: out[0] mesh energy, out[1..6] reciprocal virial (with the background
 * term on the diagonal), out[7] self energy, out[8] background energy (FP64). */
int pdx_mesh_totals(pdx_mesh_t m, double out[9], pdx_stream_t stream);
/* Debug/validation views: copy the potential grid (K1*K2*K3 floats) or the half spectrum
 * of the last solve to device memory `dst`. which: 0 fixed-point charge grid as double,
 * 1 potential grid (float). */
int pdx_mesh_copy_grid(pdx_mesh_t m, int32_t which, void *dst, pdx_stream_t stream);
/* Stand-alone batched 3D FFT of a real grid (K1,K2,K3) (validation of the transform;
 * the cuFFT wrapper tests, test/test_electrostatics.jl:260-279). dir=-1: R2C real->half
 * spectrum (float2); dir=+1: C2R (unnormalised). */
int pdx_mesh_fft(pdx_mesh_t m, int32_t dir, const void *in, void *out, pdx_stream_t stream);

/* Electrostatic exceptions (normalize_exceptions, src/electrostatics/Spec.jl:200-224): pairs
 * (i, j) whose Coulomb interaction is scaled by s >= 0 (0: excluded). The real-space sums gain
 * C q_i q_j (s - 1) / r at the minimum image, without cutoff (the mesh carries the rest of the
 * pair), and the pair's energy carries no potential shift (_coulomb_real_kernel!,
 * src/electrostatics/RealSpace.jl:83-99). Host arrays, 0-based ids; a pair listed twice must
 * carry the same scale; s == 1 is dropped; nexc = 0 clears. Needs the real-space Coulomb term
 * and the lanes layout (lanes >= 0). */
int pdx_pair_set_exceptions(pdx_pair_t p, int32_t n, int32_t nexc, const int32_t *i, const int32_t *j,
                            const double *scale);

/* ---- traversal of the pair kernel ---- */
/* ROWS: the ELL rows of the list (the reference's path). CELLS: no rows are read; one
 * wave64 per particle walks the 27 cells of the list's grid (width >= cutoff + skin, so the
 * cells cover every pair within the cutoff until the list asks for a rebuild), ballots the
 * candidates within the cutoff into an LDS queue and evaluates them 64 at a time. Forces,
 * energies and virials are those of the rows path up to the order of the FP32 sums; the
 * order is fixed (cells, then ids), so results are deterministic. Needs a cell grid (not
 * PDX_NL_ALLPAIRS) and FP32 accumulation without bonds. */
#define PDX_TRAVERSAL_ROWS  0
#define PDX_TRAVERSAL_CELLS 1
int pdx_pair_set_traversal(pdx_pair_t p, int32_t mode);

/* ---- accumulation of the per-particle sums (force, energy, virial) ---- */
/* The precision study of a bead-spring melt: pair arithmetic stays FP32, the sums over a
 * particle's row (and bonds) run in FP32, in 64-bit fixed point (2^32 scale: exact and
 * independent of the summation order) or in FP64; per-particle outputs are FP32 and the
 * totals FP64 in every mode. Not available with real-space Coulomb. */
#define PDX_ACC_FP32    0
#define PDX_ACC_FIXED64 1
#define PDX_ACC_FP64    2
int pdx_pair_set_accumulation(pdx_pair_t p, int32_t mode);

/* FENE bonds (_fene3_EV!, src/bonded/FENEBondCUDA.jl): u = -k R0^2/2 ln(1 - r^2/R0^2),
 * f/r = -k / (1 - r^2/R0^2), added to the bonded particles' pair sums (bonded pairs also
 * keep their pair interaction, the Kremer-Grest convention, src/SimulationState.jl:152-155).
 * bi, bj: host, nbond pairs of 0-based ids. */
typedef struct pdx_bonds_s *pdx_bonds_t;
int pdx_bonds_create(int32_t n, int32_t nbond, const int32_t *bi, const int32_t *bj, double k, double R0,
                     pdx_bonds_t *out);
int pdx_bonds_destroy(pdx_bonds_t b);
/* Attach (or detach with NULL) bonds to a pair model: pdx_pair_forces then adds them. */
int pdx_pair_set_bonds(pdx_pair_t p, pdx_bonds_t b);
/* Bonded exclusions (BondExclusionFeature, src/engine/Bonded.jl:60-92; the reference's
 * exclude_bonded): with exclude != 0, pairs joined by a bond of the attached FENE bonds get no
 * pair (LJ / WCA / soft) term; the bond term stays. Uncharged models only (charged pairs are
 * excluded through pdx_pair_set_exceptions). */
int pdx_pair_set_bond_exclusion(pdx_pair_t p, int32_t exclude);
/* Domain check of the last evaluations (the reference's DomainError,
 * FENEBondCUDA.jl:14-39): first_invalid receives the lowest particle id with a bond at
 * r >= R0, or -1; with reset the device word is cleared. Synchronises the stream. Returns
 * PDX_ERR_COORDINATE when a bond was out of range. */
int pdx_bonds_status(pdx_bonds_t b, int32_t *first_invalid, int32_t reset, pdx_stream_t stream);

/* ======================================================================= */
/* Cluster-pair lists and the Newton-3 tile kernel                         */
/* ======================================================================= */
/* The cluster-pair nonbonded engine (ClusterPairList / ClusterPairEngine,
 * src/clusterpairs/List.jl:112-179, Engine.jl:5-75): the short-range forces of a pair model
 * created by pdx_pair_create (LJ / WCA / soft, real-space Coulomb polynomial, Coulomb
 * exceptions) evaluated over 8 x 8 cluster tiles, each unique pair once.
 *   slots:   super-cluster S owns slots 64S .. 64S+63 (eight clusters of eight); padding
 *            slots hold perm = -1 (List.jl:4-8);
 *   outer:   per super-cluster, the j-clusters (cluster index >= the i-cluster's) whose
 *            bounding box comes within rlist = rmax + skin of an i-cluster, with an 8-bit
 *            i-cluster mask, in a segment of `entry_capacity` entries (List.jl:10-13);
 *   working: the outer entries pruned to rmax + prune_buffer by the force pass itself
 *            (List.jl:183-188, Kernels.jl:349-394).
 * Rebuild when a particle moved more than skin/2 since the build, prune when one moved more
 * than prune_buffer/2 since the last prune (Engine.jl:5-12); both decided on the device. */
typedef struct pdx_cluster_s *pdx_cluster_t;

/* Device pointers into a cluster list (read-only for callers). */
typedef struct {
  int32_t *perm;      /* nslot: slot -> particle (0-based), -1 padding */
  int32_t *slot_of;   /* N: particle -> slot */
  float *packed;      /* nslot float4 (x, y, z, q) in slot order, NaN coordinates on padding */
  int32_t *n_outer;   /* NSmax: outer entries per super-cluster */
  int32_t *outer;     /* NSmax * entry_capacity: (cj << 8) | i-cluster mask */
  int32_t *control;   /* 16 control words (PDX_CL_*) */
  int32_t *order;     /* N: 1-based ids in slot order (the mesh traversal order, mesh_order) */
  int32_t *working;   /* NSmax * entry_capacity: working entries (cj << 8) | pruned i-cluster mask; chunk k
                         of super-cluster S (of K = waves * splits) holds n_working[S*K + k] of them from
                         the start of its share floor(n_outer[S] k / K) of S's segment */
  int32_t *n_working; /* NSmax * K */
} pdx_cluster_arrays;

#define PDX_CL_NS      0  /* super-clusters in use */
#define PDX_CL_ERROR   1  /* 0, 1 outer-list capacity exceeded, 2 non-finite coordinates */
#define PDX_CL_REQUIRED 2 /* entries the largest super-cluster needed after an overflow */
#define PDX_CL_BUILDS  6  /* builds so far */
#define PDX_CL_PRUNES  7  /* pruning passes so far (each build is followed by one) */
#define PDX_CL_OUTER   8  /* outer entries of the last build */

/* List for n particles in a centred orthorhombic box (ClusterPairEngine(...), Engine.jl:50-95;
 * setup!, List.jl:355-418): rmax is the larger of the pair model's cutoffs; every box side
 * must exceed 2 (rmax + skin); 0 <= prune_buffer <= skin. entry_capacity: outer entries per
 * super-cluster (0: an estimate from the density). waves: wave64s per super-cluster in the
 * force pass, 1, 2, 4 or 8 (0: 4). The pair model must outlive the list. */
int pdx_cluster_create(int32_t n, const double box[3], pdx_pair_t model, double skin, double prune_buffer,
                       int32_t entry_capacity, int32_t waves, pdx_cluster_t *out);
int pdx_cluster_destroy(pdx_cluster_t c);
/* out: N, column grid x, y, z bins, super-cluster bound NSmax, slots, entry capacity, waves,
 * splits, chunks per super-cluster (waves * splits). */
int pdx_cluster_info(pdx_cluster_t c, int64_t out[10]);
/* Launch shape of the force pass: waves (1, 2, 4, 8) per workgroup and splits (1, 2, 4, 8)
 * workgroups per super-cluster; a super-cluster's entries are cut into waves * splits contiguous
 * chunks, one per wave (the reference's split blocks for small systems, Kernels.jl:76-86).
 * Default: 4 waves and splits for about 2048 waves in flight. The next update rebuilds. Host
 * boundary: synchronises the device. */
int pdx_cluster_set_launch(pdx_cluster_t c, int32_t waves, int32_t splits);
int pdx_cluster_arrays_get(pdx_cluster_t c, pdx_cluster_arrays *out);
/* Keep the list valid for the current coordinates and pack them in slot order
 * (update_list!, Engine.jl:41-60): displacement decision, predicated build (clustering,
 * bounding boxes, search), reference positions; graph-capturable. q may be NULL. */
int pdx_cluster_update(pdx_cluster_t c, const float *x, const float *y, const float *z, const float *q,
                       pdx_stream_t stream);
/* Forces (and with PDX_ENERGY per-particle half energies, half virials and FP64 totals, the
 * layout of pdx_pair_forces) of the coordinates of the last update (compute_forces!,
 * Engine.jl:79-92): the pruning pass over the outer list when one is due, else the plain pass
 * over the working list, then the scatter to storage order (_scatter_kernel!,
 * Kernels.jl:207-240) with the Coulomb self energy and exceptions. PDX_ADD adds. */
int pdx_cluster_forces(pdx_cluster_t c, float *fx, float *fy, float *fz, float *Epot, float *V, double *totals,
                       int32_t flags, pdx_stream_t stream);
/* The next update rebuilds (request_build!, List.jl:552-558). */
int pdx_cluster_request_build(pdx_cluster_t c, pdx_stream_t stream);
/* Synchronising read of the control words: out = NS, outer entries, working entries, builds,
 * prunes, error, required capacity, capacity. With check, returns PDX_ERR_CAPACITY /
 * PDX_ERR_COORDINATE for a latched error (forces of that step were not written). */
int pdx_cluster_status(pdx_cluster_t c, int64_t out[8], int32_t check, pdx_stream_t stream);
/* Replace the outer/working storage by segments of entry_capacity entries, clear a latched
 * capacity error and request a build (ensure_capacity!, List.jl:434-450). Host boundary:
 * synchronises the device. */
int pdx_cluster_resize(pdx_cluster_t c, int32_t entry_capacity);

/* ======================================================================= */
/* Learned correction of coarse-mesh PME forces (BF16 MFMA MLP)            */
/* ======================================================================= */
/* No counterpart in the reference: the correction and its training are new. */

/* 80 features per particle from the charge grid of the mesh's last spread (call after
 * pdx_mesh_spread / _spread_cells on the same step): q_i Q(n) at the 64 nodes
 * n = floor(u) - 3 + {0..3}^3 an order-4 B-spline at the particle's grid coordinate u covers
 * (x fastest), q_i (u - floor(u)) per dimension (3), q_i, q_i^2 and 11 zeros. feats: device
 * N x 80 FP32, row-major. */
int pdx_mesh_features(pdx_mesh_t m, const float *x, const float *y, const float *z, const float *q,
                      float *feats, pdx_stream_t stream);

typedef struct pdx_mlp_s *pdx_mlp_t;
/* Number of parameters of the 80-64-64-3 SiLU network, packed as W1 (80x64, row = input), b1,
 * W2 (64x64), b2, W3 (64x3), b3. */
int pdx_mlp_nparams(void);
/* Network with seeded He initialisation (FP32 master weights, BF16 copies for the MFMA tiles). */
int pdx_mlp_create(int32_t seed, pdx_mlp_t *out);
int pdx_mlp_destroy(pdx_mlp_t m);
/* Host parameter vectors (packed as above); setting resets the optimiser state. */
int pdx_mlp_set_params(pdx_mlp_t m, const float *params, int32_t n);
int pdx_mlp_get_params(pdx_mlp_t m, float *params, int32_t n);
/* Feature standardisation x' = (x - mu) isd (host arrays of 80) and output scale: the network
 * predicts residual / out_scale. */
int pdx_mlp_set_norm(pdx_mlp_t m, const float *mu, const float *isd, float out_scale);
/* Inference on N feature rows (device): out (N x 3, may be NULL) = out_scale MLP(x'); when fx,
 * fy, fz are given the correction is added to them. */
int pdx_mlp_forward(pdx_mlp_t m, const float *feats, int32_t n, float *out, float *fx, float *fy, float *fz,
                    pdx_stream_t stream);
/* Inference straight from the mesh (no counterpart in the reference): the features of
 * pdx_mesh_features are formed from mesh's charge grid of its last spread inside the network's
 * kernel, one 16-particle tile at a time, so no N x 80 feature rows are written or read. Call
 * after the spread (the grid holds until the next spread; the transform only reads it). out
 * (N x 3) and/or fx, fy, fz (added to) as in pdx_mlp_forward; bitwise the same values. n <= the
 * mesh's particle count. */
int pdx_mlp_forward_mesh(pdx_mlp_t m, pdx_mesh_t mesh, const float *x, const float *y, const float *z,
                         const float *q, int32_t n, float *out, float *fx, float *fy, float *fz, pdx_stream_t stream);
/* fx, fy, fz += the columns of out (N x 3): a correction computed ahead on another stream, added
 * once the forces are complete. */
int pdx_mlp_add(const float *out, int32_t n, float *fx, float *fy, float *fz, pdx_stream_t stream);
/* Gradient of loss = mean (MLP(x') - t / out_scale)^2 over the batch rows[0..n) (device ids;
 * NULL: rows 0..n-1) of feats / targets (N x 3): grad (device, nparams) and loss (device
 * double), both may be NULL (internal buffers). Per-wave partials are reduced in a fixed order. */
int pdx_mlp_gradient(pdx_mlp_t m, const float *feats, const float *targets, const int32_t *rows, int32_t n,
                     float *grad, double *loss, pdx_stream_t stream);
/* One Adam step (beta 0.9 / 0.999) on that gradient. */
int pdx_mlp_train_step(pdx_mlp_t m, const float *feats, const float *targets, const int32_t *rows, int32_t n,
                       float lr, double *loss, pdx_stream_t stream);

/* ======================================================================= */
/* Batched replica boxes: velocity Verlet, one persistent launch           */
/* ======================================================================= */
/* No single counterpart in the reference: each replica is what one Run with UniformLJ /
 * UniformWCA and the NVE kick-drift / kick kernels advances (src/engine/Run.jl:247-294,
 * src/simulation/NVEFusion.jl:3-24, src/SimulationCore.jl:581-591); the batch runs them all
 * in one launch, one workgroup per replica for every step. */

#define PDX_VV_FORCES_VALID 4  /* fx, fy, fz already hold the forces of the current x */

typedef struct {
  int32_t model;    /* PDX_PAIR_LJ or PDX_PAIR_WCA */
  int32_t shift;    /* LJ: energy shifted to zero at the cutoff */
  int32_t n;        /* atoms, 1..8192 */
  int32_t reserved;
  double epsilon, sigma;
  double cutoff;    /* LJ cutoff; WCA uses 2^(1/6) sigma (rc^2 = 1.2599211 sigma^2) */
  double box[3];    /* orthorhombic box, each side >= 2 rc */
} pdx_replica_spec;

typedef struct pdx_rep_s *pdx_rep_t;
/* Replica set: atoms of replica k occupy [offset_k, offset_k + n_k) of the concatenated SoA
 * arrays (offsets = prefix sums of n). block: workgroup size 256 / 512 / 1024 (0: 512);
 * stencil: cells per cutoff, 1 (27 cells of side >= rc) or 2 (125 cells of side >= rc/2)
 * (0: 1); boxes too small for the stencil use all pairs. */
int pdx_replicas_create(int32_t nrep, const pdx_replica_spec *specs, int32_t block, int32_t stencil,
                        pdx_rep_t *out);
int pdx_replicas_destroy(pdx_rep_t r);
/* Launch shape for later runs (the autotuner's knob). */
int pdx_replicas_set_launch(pdx_rep_t r, int32_t block, int32_t stencil);
/* out: total atoms, replicas, largest replica, LDS bytes per workgroup, block, stencil,
 * replicas on all pairs, largest cell count. */
int pdx_replicas_info(pdx_rep_t r, int64_t out[8]);
/* nrep + 1 offsets (host). */
int pdx_replicas_offsets(pdx_rep_t r, int64_t *offsets);
/* Advance every replica nsteps velocity-Verlet steps (v += c f; x += dt v, wrapped to
 * [-L/2, L/2); forces; v += c f; c = dt / (2 mass)). Without PDX_VV_FORCES_VALID the forces
 * of the current x are computed first (nsteps = 0: forces only). With PDX_ENERGY, out
 * (device, 4 doubles per replica) receives the final potential energy, kinetic energy,
 * virial trace sum r.F over pairs, and n. Positions, velocities and forces are updated in
 * place (device, FP32). Within a replica the result depends only on that replica's data. */
int pdx_replicas_run(pdx_rep_t r, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *fx,
                     float *fy, float *fz, double mass, double dt, int32_t nsteps, int32_t flags, double *out,
                     pdx_stream_t stream);

/* ======================================================================= */
/* FP64 Ewald summation on the device (the converged reference force)      */
/* ======================================================================= */

/* Direct FP64 Ewald sum for an orthorhombic box, tin-foil, optional background, following
 * test/ewald_reference.jl:50-126 (real-space images to rcut = sqrt(digits ln 10)/alpha,
 * reciprocal vectors to kmax = 2 alpha sqrt(digits ln 10)). Forces are evaluated for the
 * `nsample` particles in `sample` (0-based ids, device) into F (3*nsample doubles, xyz
 * interleaved); energy and virial (out: E, Wxx, Wyy, Wzz, Wxy, Wxz, Wyz on the host) are
 * computed over all particles when want_energy != 0. x, y, z, q are device FP64 arrays. */
int pdx_ewald_fp64(int32_t n, const double *x, const double *y, const double *z, const double *q,
                   const double box[3], double C, double alpha, double digits, int32_t background,
                   int32_t nsample, const int32_t *sample, double *F, int32_t want_energy,
                   double out[7], pdx_stream_t stream);

/* pdx_ewald_fp64 without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under graph capture of its caller's other
 * work, tight training loops that re-measure the learned correction's targets, harnesses that
 * forbid hipMalloc). Same arguments and the same kernels and launch shapes as pdx_ewald_fp64,
 * so the same results bit for bit. Its workspace (structure factors, block partials, result
 * words) is a static device array of 4.1 MiB per device, held until the process ends; the
 * half-space wave-vector box (ceil(kmax L_1/2pi) + 2)(2 ceil(kmax L_2/2pi) + 3)(2 ceil(kmax
 * L_3/2pi) + 3), kmax = 2 alpha sqrt(digits ln 10), must hold at most 262 144 indices, else
 * PDX_ERR_ARGUMENT. Every call on one device shares that array: calls that overlap on one
 * device (other host threads, other streams) race. */
int pdx_ewald_fp64_noalloc(int32_t n, const double *x, const double *y, const double *z, const double *q,
                           const double box[3], double C, double alpha, double digits, int32_t background,
                           int32_t nsample, const int32_t *sample, double *F, int32_t want_energy,
                           double out[7], pdx_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PDX_H */
