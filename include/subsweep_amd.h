/*
 * subsweep_amd — C ABI of the MI355X sweep radiative-transfer + hydrogen
 * chemistry library (libsubsweep_amd.so).
 *
 * This is the boundary a Rust `SweepPlugin` replacement would bind through
 * `extern "C"`: the reference's plugin registers `init_sweep_system` (startup
 * stage InitSweep) and `run_sweep_system` (stage Sweep), src/sweep/mod.rs:122-169,
 * and keeps all hot-path state inside one non-Send `Sweep<HydrogenOnly>`
 * resource, src/sweep/mod.rs:139,171-191. Each entry point below replaces one
 * piece of that resource's life cycle; the reference location is cited on each.
 *
 * Conventions
 *  - SI units throughout, like the reference's unit system (src/units/mod.rs:9-110).
 *  - Cell arrays have N entries, dense index = ParticleId.index
 *    (src/sweep/active_list.rs:16-46 requires ids 0..N-1).
 *  - Per-(cell, direction) arrays are direction-major: element [d * N + c].
 *  - Caller-owned host buffers are copied into handle-owned device buffers.
 *  - Every call returns an int status: SSW_OK (0) or a negative SSW_ERR_*;
 *    ssw_last_error() gives the message of the last failure on this thread.
 *  - One handle is used from one thread at a time. Distinct handles run on
 *    distinct HIP streams and may be driven concurrently.
 */
#ifndef SUBSWEEP_AMD_H
#define SUBSWEEP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSW_OK 0
#define SSW_ERR_INVALID (-1)    /* bad argument or call order                  */
#define SSW_ERR_OOM (-2)        /* device allocation failed                    */
#define SSW_ERR_BUDGET (-3)     /* request exceeds the configured HBM budget   */
#define SSW_ERR_CYCLE (-4)      /* dependency cycle in a direction's face graph */
#define SSW_ERR_HIP (-5)        /* HIP runtime error                           */
#define SSW_ERR_GRID (-6)       /* face lists not mirror-symmetric             */
#define SSW_ERR_NO_DEVICE (-7)  /* no HIP device visible                       */

/* Face connection kinds, the single-rank subset of ParticleType
 * (src/sweep/grid/cell.rs:16-23): Local(id), LocalPeriodic(id), Boundary. */
#define SSW_FACE_LOCAL 0
#define SSW_FACE_PERIODIC 1
#define SSW_FACE_BOUNDARY 2

/* Transport precision of the sweep (the chemistry is always FP64). */
#define SSW_TRANSPORT_FP64 64
#define SSW_TRANSPORT_FP32 32

/* SweepParameters (src/sweep/parameters.rs:8-46) plus the cosmology scale
 * factor the reference hands to HydrogenOnly (src/sweep/mod.rs:685-690). */
typedef struct ssw_sweep_parameters {
    /* DirectionsSpecification::Num: 1, 16, 21, 32, 64 or 84 (HEALPix bins,
     * src/sweep/direction/mod.rs:58-75) ...                                  */
    int32_t num_directions;
    /* ... or DirectionsSpecification::Explicit: num_directions x 3 doubles,
     * normalised by the library (src/sweep/direction/mod.rs:97-109). NULL for Num. */
    const double* explicit_directions;
    int32_t num_timestep_levels;           /* >= 1, < 32                         */
    int32_t periodic;                      /* 0: periodic faces act as boundary
                                              (src/voronoi/constructor/mod.rs:125-135) */
    double max_timestep;                   /* s                                   */
    double significant_rate_threshold;     /* s^-1, default 0                     */
    double timestep_safety_factor;         /* default 0.1                         */
    double chemistry_timestep_safety_factor; /* default 0.1                       */
    int32_t prevent_cooling;               /* default 1                           */
    int32_t transport_precision;           /* SSW_TRANSPORT_FP64 / _FP32          */
    double scale_factor;                   /* cosmological a (1 if non-cosmological) */
} ssw_sweep_parameters;

typedef struct ssw_sweep ssw_sweep;

/* The direction table used for DirectionsSpecification::Num (HEALPix bins,
 * used as tabulated, src/sweep/direction/mod.rs:58-75,
 * src/sweep/direction/healpix.rs:1-231): num_directions x 3 doubles into out.
 * A caller that shards by direction passes a block of rows as explicit
 * directions. */
int ssw_direction_table(int32_t num_directions, double* out);

/* set_source_terms_system (src/source_systems.rs:48-80) on the device: each
 * of the nsrc sources (positions src_pos, nsrc x 3, rates src_rate) goes to
 * its nearest cell (squared Euclidean distance to the n cell positions pts,
 * n x 3, non-periodic, ties to the lower cell index), and source_out (n)
 * receives each cell's summed rate, added in source order; cell_out (nsrc,
 * may be NULL) the chosen cells. Host arrays in and out; bitwise the result
 * of the host placement (libssw_grid's ssw_assign_sources). The kiddo
 * k-d tree's order among exactly equidistant cells is not reproduced. */
int ssw_assign_sources_device(int64_t n, const double* pts, int64_t nsrc, const double* src_pos,
                              const double* src_rate, double* source_out, int64_t* cell_out);

/* Fills *p with the reference defaults (src/sweep/parameters.rs:64-78):
 * threshold 0, safety factors 0.1, prevent_cooling true, FP64 transport,
 * scale factor 1. num_directions/levels/max_timestep are left for the caller. */
void ssw_default_parameters(ssw_sweep_parameters* p);

/* Creates the sweep state; replaces Sweep::new (src/sweep/mod.rs:194-232) as
 * called from init_sweep_system (src/sweep/mod.rs:634-692). */
int ssw_create(const ssw_sweep_parameters* params, ssw_sweep** out);

/* Uploads the grid: the `Cell` components (src/sweep/grid/cell.rs:92-97)
 * flattened to CSR. face_offsets has N+1 entries; for face f of cell c,
 * face_neighbour[f] is the dense neighbour index (ignored for boundary faces),
 * face_kind[f] an SSW_FACE_* value, face_area[f] the area (m^2) and
 * face_normal[3f..3f+2] the unit normal pointing from c to the neighbour.
 * Local faces must come in mirror pairs (c->n with normal v, n->c with -v),
 * as the reference's Voronoi grid produces them (src/voronoi/cell.rs:63-71).
 * Builds the per-direction dependency graph once and checks it for cycles. */
int ssw_upload_grid(ssw_sweep* s, int64_t num_cells, const int64_t* face_offsets,
                    const int32_t* face_neighbour, const uint8_t* face_kind,
                    const double* face_area, const double* face_normal,
                    const double* cell_volume, const double* cell_size);

/* Uploads the per-cell state the reference copies from ECS components into
 * each Site (src/sweep/mod.rs:656-671): density (kg m^-3), ionised hydrogen
 * fraction, temperature (K) and source rate (photons s^-1, src/source_systems.rs:48-80).
 * Resets fluxes, timestep levels and the timestep state like Sweep::new. */
int ssw_upload_state(ssw_sweep* s, const double* density, const double* ionized_fraction,
                     const double* temperature, const double* source);

/* One call of Sweep::run_sweeps (src/sweep/mod.rs:258-272): every sweep of the
 * timestep-level cadence, chemistry after each, then the level update. Writes
 * the elapsed simulation time (s). The reference skips the very first
 * run_sweep_system call (src/sweep/mod.rs:708-714); that is the caller's job. */
int ssw_step(ssw_sweep* s, double* time_elapsed);

/* ssw_step of n handles at once (no counterpart in the reference, which
 * steps one simulation; for callers that hold many independent problems, as
 * the request loop does). The handles are cut into packs of about four with
 * similar plan lengths, one pack per stream on up to four streams; within a
 * pack the grouped launches of handles with the same transport precision and
 * group block are packed into shared launches (up to 8 problems per launch),
 * so that small problems fill the device together. Each handle's results are
 * bitwise those of ssw_step on it alone; a handle's stream continues after
 * the packed step. time_elapsed (n entries, may be NULL) receives each
 * handle's elapsed time. All handles on one device, none twice, none used by
 * another thread during the call. */
int ssw_step_batch(ssw_sweep* const* handles, int32_t n, double* time_elapsed);
/* Packed sweeps whose handles, sweep levels and replay plans repeat are
 * captured as one graph the second time they come and replayed from then on
 * (handles with graph replay off are always launched directly): graphs
 * captured and graph launches so far, process-wide. */
int ssw_packed_graph_stats(int64_t* captured, int64_t* replayed);

/* Copies results back, like the tail of run_sweep_system (src/sweep/mod.rs:715-738).
 * Any pointer may be NULL. photon_rate = sum over directions of the incoming
 * rate (no source, no periodic term, src/sweep/mod.rs:727-730). */
int ssw_download(ssw_sweep* s, double* ionized_fraction, double* temperature,
                 double* species_timestep, double* photon_rate, uint8_t* level,
                 double* change_timescale);

/* ssw_download without any allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same outputs,
 * bitwise equal to ssw_download's, staged through a fixed 64 MiB device
 * array that each device the library runs on holds until the process ends,
 * so it handles at most 8388608 cells (SSW_ERR_INVALID above that). It calls
 * no hipMalloc or hipFree. Every call on one device shares that array, so
 * calls that overlap on one device (handles driven from several threads)
 * race: serialise them. */
int ssw_download_noalloc(ssw_sweep* s, double* ionized_fraction, double* temperature,
                         double* species_timestep, double* photon_rate, uint8_t* level,
                         double* change_timescale);

/* The optional chemistry outputs of sweep_optional_output_system
 * (src/sweep/chemistry_output.rs:25-67): per cell, the Solver of
 * Sweep::get_solver (src/sweep/mod.rs:612-632; rate = sum over directions of
 * incoming + source/D + periodic, current x and T, no floor) evaluated with
 * a timestep of 1 yr: PhotoionizationRate (s^-1), HeatingRate (photoheating
 * minus cooling, W m^-3), RecombinationRate (alpha_B n_e x, s^-1) and
 * CollisionalIonizationRate (beta n_e (1 - x), s^-1). SI units, caller's
 * cell order; any pointer may be NULL. The components they fill are
 * src/components.rs:50-68. */
int ssw_download_chemistry_outputs(ssw_sweep* s, double* photoionization_rate, double* heating_rate,
                                   double* recombination_rate, double* collisional_ionization_rate);

/* rotate_directions_system (src/sweep/direction/mod.rs:158-175), which the
 * reference runs after every run_sweep_system call when
 * SweepParameters::rotate_directions is set (src/sweep/mod.rs:143-148,
 * src/sweep/parameters.rs:20-22): every direction d becomes matrix . d
 * (matrix row-major 3 x 3, multiply_by_matrix :150-155), the dependency graph
 * and schedule are rebuilt for the rotated directions, and the incoming,
 * outgoing and periodic rates of every cell are remapped as the reference's
 * remap (:190-205) does: its inner loop overwrites values[i] for every j, so
 * direction i ends up with the old rate of the last direction if the last
 * rotated direction is the one most aligned with old direction i (the last
 * maximum of 1 + old_i . new_j, kernel_f :177-188), and with 0 otherwise.
 * Later sweeps add their corrections to these rates, as the reference's
 * accumulated rates do. A handle with rotated directions does not save or
 * load checkpoints. */
int ssw_rotate_directions(ssw_sweep* s, const double* matrix);

/* get_rotation_matrix (src/sweep/direction/mod.rs:112-135): the row-major
 * rotation by `angle` (radians) about the unit vector `axis`, into out[9]. */
int ssw_rotation_matrix(const double* axis, double angle, double* out);

/* get_random_rotation_matrix (src/sweep/direction/mod.rs:137-148) for its
 * three random draws, phi in [0, 2 pi), u in [0, 1) and psi in [0, 2 pi)
 * (the reference takes them from a StdRng seeded with 1337, :207-210): the
 * rotation by psi about the axis of polar angle acos(2u - 1) and azimuth phi. */
int ssw_random_rotation_matrix(double phi, double u, double psi, double* out);

/* Waits until all work queued on the handle's stream has finished. */
int ssw_synchronize(ssw_sweep* s);

/* Parity hook: per-(direction, cell) upwind dependency counts of
 * Sweep::init_counts (src/sweep/mod.rs:346-386) for a sweep at `sweep_level`
 * under the current timestep levels; inactive cells get 0xFFFFFFFF. */
int ssw_download_counts(ssw_sweep* s, int32_t sweep_level, uint32_t* counts);

/* Parity hook: Kahn depth of every (direction, cell) in the dependency graph
 * restricted to the cells active at `sweep_level` (the level-set order in which
 * the device solves them); inactive cells get 0xFFFFFFFF. */
int ssw_download_levelsets(ssw_sweep* s, int32_t sweep_level, uint32_t* depth);

/* Parity hook: incoming, outgoing and periodic-source rates per
 * (direction, cell) as Site holds them (src/sweep/site.rs:12-23). Any may be NULL. */
int ssw_download_fluxes(ssw_sweep* s, double* incoming, double* outgoing, double* periodic);

typedef struct ssw_info {
    int64_t num_cells;
    int64_t num_faces;
    int32_t num_directions;
    int32_t max_depth;             /* Kahn depth bound of the full graph          */
    int64_t updates_last_step;     /* cell-direction updates of the last step     */
    int64_t sweeps_last_step;
    double downwind_local;         /* mean downwind faces per cell-dir by kind    */
    double downwind_periodic;
    double downwind_boundary;
    int64_t device_bytes;          /* device memory held by this handle           */
    int64_t chemistry_failures;    /* cells whose chemistry hit the depth limit   */
    int32_t transport_mode;        /* SSW_TRANSPORT_MODE_GEOMETRIC / _COMPILED    */
    int32_t list_builds_last_step; /* sweeps of the last step whose active-set
                                      list was rebuilt (compaction)            */
    int64_t overflow_entries;      /* compiled upwind entries beyond the planes   */
    int64_t graph_capture_failures; /* replays launched one by one because their
                                       graph could not be captured            */
    int32_t replay_block;          /* compiled replay block size               */
    int32_t graph_replay;          /* 1 if replays are captured as graphs      */
    int32_t locality_order;        /* 1 if cells are stored in position order  */
    int32_t narrow_runs;           /* 1 if runs of narrow level sets replay in
                                      one workgroup (compiled replay)          */
    int64_t peak_device_bytes;     /* most device memory the handle has held   */
    int32_t group_levels;          /* level sets per grouped launch of full
                                      sweeps (1: one launch per level set)     */
    int32_t group_fallback;        /* 1 if groups were requested but a cone did
                                      not fit (full sweeps per level set)      */
    int64_t group_tasks;           /* tasks a grouped full sweep solves, cones
                                      overlapping (>= the sweep's tasks)       */
    int32_t group_block;           /* grouped replay: root tasks per workgroup */
    int32_t group_blocks_per_cu;   /* grouped replay: workgroups one CU can hold
                                      at once (register and LDS limits)      */
    /* TimescaleCounter of the last step (src/chemistry/timescale.rs:70-111,
     * counted in update_chemistry, src/sweep/mod.rs:572, reset per run_sweeps,
     * :267): chemistry updates whose change timescale is below max_timestep,
     * by limiting process (a tie of the photon-rate and chemistry timescales
     * goes to the chemistry, src/chemistry/timescale.rs:32-38). */
    int64_t limiting_temperature;
    int64_t limiting_ionization_fraction;
    int64_t limiting_photon_rate;
    /* grouped replay of a full sweep: launches and workgroups over them
     * (mean workgroups per launch = group_workgroups / group_launches) */
    int64_t group_launches;
    int64_t group_workgroups;
    int32_t compute_units;         /* CUs of the handle's device               */
    int32_t direction_sweeps;      /* 1 if sweeps run one workgroup per
                                      direction (ssw_set_direction_sweeps)   */
} ssw_info;

int ssw_get_info(ssw_sweep* s, ssw_info* info);

/* Checkpoint of the dynamic state: density, x, T, source, attenuation,
 * species timestep, previous rate, change timescale and level per cell, the
 * transport state X and I per (direction, cell) and the periodic sources, all
 * in the caller's cell order, plus the timestep-level cadence. A handle of the
 * same problem (grid, directions, levels, precision, maximum timestep; any
 * transport mode, launch shape or cell order) continues from it with bitwise
 * the same results. No counterpart in the reference (it restarts from
 * snapshots of the ECS components, src/io/output/mod.rs); for restart files
 * and for moving a problem between handles.
 *   ssw_checkpoint_bytes  size of the checkpoint (0 before ssw_upload_state)
 *   ssw_save_checkpoint   writes it into buf (at least that many bytes)
 *   ssw_load_checkpoint   replaces the state of a handle after its
 *                         ssw_upload_state; SSW_ERR_INVALID for a checkpoint
 *                         of another problem */
int64_t ssw_checkpoint_bytes(ssw_sweep* s);
int ssw_save_checkpoint(ssw_sweep* s, void* buf, int64_t bytes);
int ssw_load_checkpoint(ssw_sweep* s, const void* buf, int64_t bytes);

/* Replay form of the transport (no counterpart in the reference, which has one
 * CPU code path); set before ssw_upload_grid. Both forms solve the same level
 * sets with the same arithmetic and give bitwise the same results.
 *  SSW_TRANSPORT_MODE_GEOMETRIC  task records (16 B per cell-direction);
 *      every replay re-reads the face geometry, 16 lanes per task.
 *  SSW_TRANSPORT_MODE_COMPILED   per task its upwind (neighbour, weight) list
 *      in level-position order (about 16 + 4 + 12 x 8 B per cell-direction in
 *      FP32, twice the entry size in FP64), one lane per task.
 *  SSW_TRANSPORT_MODE_AUTO       compiled when its peak need (schedule
 *      arrays, then active-set lists and transport state) is at most 85% of
 *      the free device memory at ssw_upload_grid, else geometric (default).
 * Values 1 and 2 selected a persistent one-launch-per-sweep kernel that did
 * not terminate on MI355X; it has been removed and they return SSW_ERR_INVALID. */
#define SSW_TRANSPORT_MODE_AUTO 0
#define SSW_TRANSPORT_MODE_GEOMETRIC 3
#define SSW_TRANSPORT_MODE_COMPILED 4
int ssw_set_transport_mode(ssw_sweep* s, int32_t mode);

/* Block size of the compiled replay kernel (64, 128 or 256 threads, one task
 * per thread; 0: the default of 64). A launch-shape knob for the autotuner:
 * results do not depend on it. No counterpart in the reference. */
int ssw_set_replay_block(ssw_sweep* s, int32_t block);

/* Optional cell positions (num_cells x 3 doubles, any units), given before
 * ssw_upload_grid: the device then stores cells in the Morton (Z-curve) order
 * of their positions, so that the upwind neighbours gathered by one wave of
 * tasks share cache lines. Every array the API takes or returns stays in the
 * caller's cell order, and results do not depend on it (bitwise). NULL or
 * n <= 0 clears it. No counterpart in the reference (its cells carry a
 * Position component, which is what a caller passes here). */
int ssw_set_cell_positions(ssw_sweep* s, int64_t num_cells, const double* positions);

/* Replays of an unchanged active set are captured as one HIP graph from their
 * second use on (enabled = 1, default) or launched one level set at a time
 * (0). The library's own allocations, frees and stream set-up wait while a
 * handle captures (they would invalidate the capture), so handles driven
 * from several threads capture too. No counterpart in the reference. */
int ssw_set_graph_replay(ssw_sweep* s, int32_t enabled);

/* Compiled replay: each run of two or more consecutive level sets with at
 * most 1024 active tasks each (the narrow ends of the dependency graph) is
 * solved by one 1024-lane workgroup in one launch, level set after level set
 * with a workgroup barrier in between, instead of one launch per level set
 * (enabled = 1, default; 0 = one launch per level set). Same arithmetic per
 * task, bitwise the same results. Set before ssw_upload_grid. No counterpart
 * in the reference. */
int ssw_set_narrow_runs(ssw_sweep* s, int32_t enabled);

/* Compiled replay: groups of `levels` consecutive level sets (2 to 8, default
 * 4; 8 where 8 consecutive level sets hold at most 2048 tasks each; 1 = one
 * launch per level set) run as one launch each. A workgroup owns a
 * block of root tasks (ssw_set_group_block) of one level set of the group and
 * solves their upwind cone inside the group level set by level set with a
 * workgroup barrier in between; a task in several cones is solved by each of
 * them from the same inputs with the same arithmetic, so results are bitwise
 * the same as without groups. The cones of the full graph are built once per
 * grid at ssw_upload_grid; a restricted sweep filters them to its active set
 * whenever that changes. If a cone exceeds its bound the sweeps stay per
 * level set (ssw_info.group_fallback). Set before ssw_upload_grid. No
 * counterpart in the reference. */
int ssw_set_group_levels(ssw_sweep* s, int32_t levels);

/* Root tasks per workgroup of the grouped replay: 32, 64 or 128 (0 = 64, the
 * default); the workgroup has twice as many lanes. Set before
 * ssw_upload_grid. No counterpart in the reference. */
int ssw_set_group_block(ssw_sweep* s, int32_t roots);

/* Compiled replay: every sweep as one launch of one workgroup per direction,
 * each solving its direction's level sets in order with a workgroup barrier
 * in between (enabled = 1; 0 = the grouped or per-level-set replay,
 * default). A direction's tasks depend only on tasks of the same direction
 * (src/sweep/mod.rs:453-503 hands corrections only to neighbours in the same
 * direction), so no dependency crosses workgroups. X is kept in each
 * direction's Kahn order, and each workgroup holds the values it solved last
 * in LDS, from which most upwind values are read. A restricted sweep passes
 * every task and solves those of active cells. Same arithmetic per task,
 * bitwise the same results. Set before ssw_upload_grid. No counterpart in
 * the reference. */
int ssw_set_direction_sweeps(ssw_sweep* s, int32_t enabled);

/* Raw transport state in its storage precision and scale (FP32 or FP64 per
 * the transport precision, direction-major): X = outgoing rate / downwind
 * effective area and I = incoming rate. For bitwise comparisons. */
int ssw_download_transport(ssw_sweep* s, void* X, void* I);

/* Kernel timing of the last `ssw_step` when timing is enabled, from HIP events
 * recorded on the handle's stream: one event pair around each sweep's whole
 * transport replay, so transport_avg_ms is the replay time per launched
 * level set (kernel time plus the gaps between the launches), together with
 * the transport total and launch count, the total of the chemistry kernels,
 * and the total of the active-set list builds. */
int ssw_set_timing(ssw_sweep* s, int32_t enabled);
int ssw_get_timing(ssw_sweep* s, double* transport_avg_ms, double* transport_total_ms,
                   int64_t* transport_launches, double* chemistry_total_ms, double* build_total_ms);

/* Host-side helpers, the same code the device kernels run.
 * ssw_timestep_level: TimestepLevel::from_max_timestep_and_desired_timestep
 * (src/sweep/timestep_level.rs:27-36), ceil(log2(max/desired)) saturated to
 * [0, num_levels-1], evaluated exactly from the binary exponent.
 * ssw_chemistry_timestep: one Solver::perform_timestep of a single cell
 * (src/chemistry/hydrogen_only/mod.rs:394-441); state = {x_HII, T} in/out,
 * returns the timescale, *process = 0 temperature / 1 ionisation fraction,
 * *failed = 1 when the sub-stepping exceeded its depth (timescale dt/10). */
int32_t ssw_timestep_level(int32_t num_levels, double max_timestep, double desired_timestep);
double ssw_chemistry_timestep(double* state, double density, double volume, double length, double rate,
                              double scale_factor, int32_t prevent_cooling, double timestep, double safety,
                              int32_t* process, int32_t* failed);

/* Device memory a handle would hold for a problem of num_cells cells,
 * num_faces directed faces and num_periodic_cells cells with a periodic face,
 * in the given transport mode (GEOMETRIC or COMPILED): an upper estimate of
 * the handle's peak from ssw_upload_grid through ssw_step and ssw_download,
 * used for admission against an HBM budget, for handles with the default
 * launch shape (grouped replay of 4 level sets with 64 roots in the compiled
 * mode; the diagnostic downloads of
 * counts, level sets, fluxes and transport state allocate transient arrays
 * beyond it). No counterpart in the reference. */
int64_t ssw_estimate_device_bytes(const ssw_sweep_parameters* params, int64_t num_cells, int64_t num_faces,
                                  int64_t num_periodic_cells, int32_t transport_mode);

/* Source rates of star particles (host only), the reference's Arepo input
 * path (src/arepo_postprocess/sources.rs:57-128); the rates then go to the
 * nearest cells like any point source (src/source_systems.rs:48-80).
 * ssw_bpass_source_rate: bpass_lookup (src/arepo_postprocess/bpass.rs:11-56),
 *   ionising photons/s of a stellar population of `mass_solar` solar masses,
 *   age (years) and metallicity (mass fraction) clamped into the table by a
 *   factor 1.00001, bilinear in age and metallicity.
 * ssw_time_between_scale_factors: CosmologyParams::time_difference_between_
 *   scalefactors (src/cosmology.rs:80-107), seconds from scale factor a0 to
 *   a1 in a flat cosmology (NaN if omega_lambda + omega_0 is off 1 by > 1e-2).
 * ssw_star_sources: per particle, is_star = formation scale factor > 0 (the
 *   others are wind particles, sources.rs:117-118) and rate = bpass(age from
 *   the formation scale factor to `scale_factor`, metallicity, mass) x
 *   escape_fraction, 0 for wind particles (new_bpass_source, :57-76).
 *   Returns the number of stars or SSW_ERR_INVALID. */
double ssw_bpass_source_rate(double age_years, double metallicity, double mass_solar);
double ssw_time_between_scale_factors(double a0, double a1, double omega_lambda, double omega_0, double h);
int64_t ssw_star_sources(int64_t n, const double* metallicity, const double* formation_scale_factor,
                         const double* mass_solar, double escape_fraction, double omega_lambda, double omega_0,
                         double h, double scale_factor, double* rate, uint8_t* is_star);

/* Request loop (no counterpart in the reference, which runs one simulation
 * per process): a server with `slots` worker threads runs independent solves
 * concurrently, each in its own handle and HIP stream, so their kernels
 * overlap on the device. Requests are admitted in submission order when a
 * worker is free and their device-memory estimate fits in what is left of
 * the HBM budget; a request above the whole budget fails with
 * SSW_ERR_BUDGET. Each request runs create, upload grid and state, `steps`
 * steps and a download into its output buffers (any may be NULL). Input and
 * output buffers must stay valid until ssw_server_wait returns for the
 * request. Results do not depend on what else runs: every kernel's
 * arithmetic is independent of scheduling.
 *
 * Resident grids: a request with a non-zero grid_key leaves its handle (grid,
 * dependency graph, schedule and replay lists on the device) in the server's
 * cache when it completes; a later request with the same grid_key and the
 * same parameters, transport mode and replay block reuses it, uploading only
 * its state (the caller promises that equal keys name equal grids and cell
 * positions). Cached handles count against the HBM budget and are freed,
 * least recently used first, when a request needs their memory. A request
 * with grid_key 0 frees its handle at the end. */
typedef struct ssw_server ssw_server;

typedef struct ssw_request {
    ssw_sweep_parameters params;
    int32_t transport_mode;        /* SSW_TRANSPORT_MODE_*; AUTO: compiled if it
                                      fits budget / slots, else geometric      */
    int32_t steps;                 /* ssw_step calls                          */
    int32_t replay_block;          /* as ssw_set_replay_block (0: default)     */
    int32_t direction_sweeps;      /* 1: as ssw_set_direction_sweeps (0: off)  */
    int64_t num_cells;             /* grid, as for ssw_upload_grid            */
    const int64_t* face_offsets;
    const int32_t* face_neighbour;
    const uint8_t* face_kind;
    const double* face_area;
    const double* face_normal;
    const double* cell_volume;
    const double* cell_size;
    const double* density;         /* state, as for ssw_upload_state          */
    const double* ionized_fraction;
    const double* temperature;
    const double* source;
    const double* cell_position;   /* optional, as for ssw_set_cell_positions */
    double* out_ionized_fraction;  /* results, as for ssw_download            */
    double* out_temperature;
    double* out_timestep;
    double* out_photon_rate;
    uint8_t* out_level;
    double* out_change_timescale;
    int64_t grid_key;              /* 0: no reuse; else the grid's identity    */
} ssw_request;

typedef struct ssw_request_result {
    int32_t status;                /* SSW_OK or the failing call's SSW_ERR_*  */
    int32_t transport_mode;        /* mode the request ran in                 */
    double submit_s;               /* seconds since server creation           */
    double start_s;                /* admitted                                */
    double end_s;                  /* results downloaded, memory released     */
    int64_t updates;               /* cell-direction updates of all steps     */
    int64_t device_bytes;          /* most device memory its handle held      */
    int64_t estimated_bytes;       /* the estimate it was admitted with       */
    char error[256];
    int32_t reused;                /* 1 if it ran on a cached handle          */
    int32_t reserved1;
} ssw_request_result;

int ssw_server_create(int32_t slots, int64_t hbm_budget_bytes, ssw_server** out);
/* A server whose requests are stepped together: the slot workers set
 * requests up (handle, grid, state) and download their results, and
 * slots / max_batch stepper threads (1 to 4) step the set-up requests, each
 * taking up to max_batch per iteration, with ssw_step_batch (their grouped
 * launches packed into shared launches). max_batch < 2 gives
 * ssw_server_create's server. Results are bitwise those of the unpacked
 * server. */
int ssw_server_create_packed(int32_t slots, int64_t hbm_budget_bytes, int32_t max_batch, ssw_server** out);
/* Packed servers: stepper iterations so far and the steps they took (steps /
 * batches = mean requests per packed step). */
int ssw_server_pack_stats(ssw_server* server, int64_t* batches, int64_t* batched_steps);
/* Queues a copy of *request (not of its arrays); *ticket identifies it. */
int ssw_server_submit(ssw_server* server, const ssw_request* request, int64_t* ticket);
/* Blocks until the request has finished; each ticket can be waited on once. */
int ssw_server_wait(ssw_server* server, int64_t ticket, ssw_request_result* result);
/* Highest reserved budget (running requests' estimates plus cached handles)
 * and highest number of concurrently running requests so far. */
int ssw_server_stats(ssw_server* server, int64_t* peak_reserved_bytes, int32_t* peak_running);
/* Runs every queued request to completion, then stops the workers. */
void ssw_server_destroy(ssw_server* server);

/* Releases the handle and its device memory. */
void ssw_destroy(ssw_sweep* s);

/* Message of the last error on the calling thread ("" if none). */
const char* ssw_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SUBSWEEP_AMD_H */
