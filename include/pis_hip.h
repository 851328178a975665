/*
 * pis_hip.h -- C ABI of the MI355X (gfx950) Lennard-Jones force and velocity-Verlet
 * engine.  Plain pointers and sizes only; no C++ or framework types cross this line.
 *
 * Reference interfaces replaced (vivekadishankara/pis, Rust):
 *   trait PotentialManager                  src/potentials/potential.rs:12-136
 *     fn compute_potential(&self, &mut Atoms) -> f64   potential.rs:13
 *     fn verlet_step_nve(&self, &mut Atoms, dt) -> f64  potential.rs:15-33
 *   LJVOffsetManager (the shipped manager)   src/potentials/lennard_jones.rs:182-244
 *   manager construction from pair_coeff     src/system.rs:121-165, readers/input_file/commands.rs:147-384
 *   Atoms / SimulationBox state              src/atoms/new.rs:9-17, src/simulation_box.rs:5-15
 *   thermo reductions                        src/atoms/properties.rs:17-65
 *   PisError                                 src/errors.rs:5-100
 *
 * Conventions (identical to the reference's in-memory layout, so a Rust shim can pass
 * Matrix3xX::as_ptr() directly):
 *   - positions / velocities / forces: AoS xyz, 3*n_atoms doubles, original atom order;
 *   - box matrices h and h_inv: 3x3 column-major (nalgebra), box vectors as columns;
 *     h_inv is produced by the caller exactly as SimulationBox::new does (the engine
 *     never recomputes it, so minimum image and wrap see the caller's bits);
 *   - atom types are 1-based (properties.rs:10);
 *   - errors: every entry point returns pis_status; pis_last_error() gives the message
 *     of the last failure on the calling thread.  No exception crosses the ABI.
 *
 * Threading: a context is not re-entrant; separate contexts may be used from separate
 * threads.  All device work of a context runs on one HIP stream owned by the context.
 */
#ifndef PIS_HIP_H
#define PIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIS_ABI_VERSION 3

/* Status codes.  1-4 map to PisError variants (errors.rs:71-99). */
typedef enum pis_status {
    PIS_OK = 0,
    PIS_ERR_NO_ATOMS = 1,          /* PisError::NoAtomsDefined (errors.rs:71-72) */
    PIS_ERR_MISSING_POTENTIAL = 2, /* PisError::PotentialNotInitialized / MissingPotential */
    PIS_ERR_NAN_ENERGY = 3,        /* PisError::NaNEnergy (errors.rs:90-91) */
    PIS_ERR_INVALID_ATOM_TYPE = 4, /* PisError::InvalidAtomType (errors.rs:94-95) */
    PIS_ERR_INVALID_ARGUMENT = 5,
    PIS_ERR_UNSUPPORTED = 6, /* triclinic or non-periodic box, thin box with many atoms */
    PIS_ERR_HIP = 7,
    PIS_ERR_OUT_OF_MEMORY = 8,
    PIS_ERR_NO_DEVICE = 9,
    PIS_ERR_CAPACITY = 10, /* neighbour/staging capacity still exceeded after regrowth */
    PIS_ERR_NOT_FOUND = 11,
    PIS_ERR_TIMEOUT = 12,
    PIS_ERR_IO = 13 /* PisError::DumpWriteError (errors.rs:50-55) */
} pis_status;

typedef enum pis_precision {
    PIS_PRECISION_FP64 = 0,  /* every operation in FP64 (the parity path) */
    PIS_PRECISION_MIXED = 1  /* FP32 pair arithmetic; positions, velocities, minimum
                                image, wrap, energy and virial sums in FP64 */
} pis_precision;

/* One pair_coeff line (system.rs:136-163): LennardJones::new(eps, sigma, rcut, shift).
 * (type_i, type_j) is stored under (min, max) (potential.rs:181-192). */
typedef struct pis_pair_coeff {
    uint32_t type_i;
    uint32_t type_j;
    double epsilon;
    double sigma;
    double rcut;
} pis_pair_coeff;

typedef struct pis_system_desc {
    uint64_t n_atoms;
    uint32_t n_types;
    uint32_t n_pairs;
    const uint32_t* type_ids; /* n_atoms, 1-based */
    const double* masses;     /* n_types, per type (Atoms::masses, new.rs:12) */
    const pis_pair_coeff* pairs;
    double h[9];
    double h_inv[9];
    uint8_t pbc[3];       /* all three must be 1 (from_lammps_data always sets them) */
    uint8_t shift;        /* LennardJones::shift; the shipped path always sets 1 */
    uint8_t precision;    /* pis_precision */
    uint8_t tile[3];      /* cells per tile along x, y, z, as an autotuner picked them (each
                             clamped to ncell - 2; ignored if the tile's halo would exceed
                             the staging tables); any zero: no request, the planner's model
                             (or pis_set_tile_override) picks.  It outranks the override. */
    double skin;          /* Verlet skin; list cutoff is rcut_ij + skin, rebuilt when an atom
                             moved more than skin/2 since the last build */
    double kb;            /* Boltzmann constant used for temperature (constants.rs:3 is
                             0.0083144621; reduced units use 1.0) */
    uint32_t neighbor_capacity; /* 0 = automatic; otherwise entries per atom */
    uint32_t flags;       /* reserved, must be 0 */
} pis_system_desc;

/* Per-step thermo record (Simulation::output, simulation.rs:79-86). */
typedef struct pis_thermo {
    double pe;          /* PotentialManager return value: sum of shifted pair energies */
    double ke;          /* kinetic_energy (properties.rs:17-23) */
    double etot;        /* NVE Hamiltonian pe + ke (simulation.rs:98-100) */
    double temperature; /* 2 ke / (3 N kb) (properties.rs:28-30) */
    double pressure;    /* (2 ke + sum_i r_i.F_i) / (3 V) (properties.rs:61-65) */
    double virial_abs;  /* sum_i r_i . F_i, the reference's absolute-position virial */
    double virial_pair; /* sum_{i<j} r_ij . F_ij, the periodic pair virial */
} pis_thermo;

/* Host-side plan of a system (no device needed): what the engine will build. */
typedef struct pis_plan_info {
    int32_t ncell[3];       /* cell grid, cell edge >= max rcut + skin */
    int32_t tile[3];        /* cells per tile (last tile per dim may be smaller) */
    int32_t ntile[3];
    int32_t mode;           /* 0 tile-relative staging, 1 staged raw + per-pair minimum
                               image, 2 whole system staged (thin boxes) */
    int32_t neighbor_capacity;
    int32_t stage_capacity; /* atoms a tile may stage in LDS */
    double r2_list_max;     /* largest r^2 whose correctly rounded sqrt <= max rcut + skin */
    double density;
} pis_plan_info;

typedef struct pis_stats {
    uint64_t steps;
    uint64_t rebuilds;
    uint64_t force_launches;
    double force_kernel_ms; /* summed HIP-event time of force launches while timing is on */
    uint64_t timed_force_launches;
    int32_t neighbor_capacity;
    int32_t max_neighbors; /* largest per-atom list seen at the last build */
} pis_stats;

typedef struct pis_ctx pis_ctx;

int pis_abi_version(void);
const char* pis_last_error(void);
const char* pis_status_string(pis_status s);
pis_status pis_device_count(int32_t* count);

/* Host-only: validate a descriptor and report the grid/tile/capacity plan. */
pis_status pis_plan_system(const pis_system_desc* desc, pis_plan_info* out);

/* Autotuner hook (no reference counterpart): cells per tile used by contexts created
 * afterwards, {tx, ty, tz}; NULL or zeros restore the built-in heuristic.  Process-wide. */
pis_status pis_set_tile_override(const int32_t* tile);

/* Single-system context: the GPU LJVOffsetManager replacement. */
pis_status pis_ctx_create(const pis_system_desc* desc, pis_ctx** out);
void pis_ctx_destroy(pis_ctx* ctx);

/* Upload state (original order, AoS xyz).  vel / force may be NULL (zeros). */
pis_status pis_set_state(pis_ctx* ctx, const double* pos, const double* vel, const double* force);
/* Download state in original order; any pointer may be NULL. */
pis_status pis_get_state(pis_ctx* ctx, double* pos, double* vel, double* force);

/* PotentialManager::compute_potential (lennard_jones.rs:187-243): forces of the current
 * positions are ADDED into forces_inout (3n, may be NULL) as the reference accumulates
 * into Atoms::forces; *pe receives the potential energy.  The device force array is
 * overwritten with the computed forces (what verlet_step_nve uses as F(t)). */
pis_status pis_compute_potential(pis_ctx* ctx, double* forces_inout, double* pe);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * forces and PE of pis_compute_potential for one system whose positions (AoS xyz, 3n
 * doubles) and 1-based atom types (n uint32, desc->type_ids is ignored) are already in
 * device memory.  Forces are ADDED into d_forces_inout (3n doubles, device, may be NULL);
 * *d_pe (device, may be NULL) receives the PE.  Every buffer the kernels need beyond these
 * lives in a static 64 MiB device array per GPU, held until the process ends: a system
 * that needs more is rejected with PIS_ERR_CAPACITY before any device work.  Work goes on
 * `stream` (a hipStream_t, NULL for the null stream) and the call returns once it has
 * finished.  Calls on one GPU share the array and are serialised by a lock; it keeps the
 * last call's contents, and each call clears what it uses.  No hipMalloc, hipFree or
 * allocating library call is made.  The result is bit-identical to pis_ctx_create +
 * pis_set_state + pis_compute_potential on the same input. */
pis_status pis_compute_potential_noalloc(const pis_system_desc* desc, const double* d_pos, const uint32_t* d_types,
                                         double* d_forces_inout, double* d_pe, void* stream);

/* nsteps of PotentialManager::verlet_step_nve (potential.rs:15-33) with state resident
 * on the device.  per_step_out (nsteps records, may be NULL) receives thermo after each
 * step; its pe is what verlet_step_nve returns. */
pis_status pis_verlet_step_nve(pis_ctx* ctx, double dt, uint64_t nsteps, pis_thermo* per_step_out);

/* Rebuild the neighbour list at the current positions and return it in original atom
 * indices as CSR (offsets: n+1 entries; idx: capacity `cap`; *total gets the number of
 * entries even when it exceeds cap).  Used by parity tests of the pair set. */
pis_status pis_get_neighbor_list(pis_ctx* ctx, uint64_t* offsets, uint32_t* idx, uint64_t cap,
                                 uint64_t* total);

/* Rebuild the neighbour list at the current positions and return, per atom in original
 * order, its length (counts[n]) and the order-independent row hash hashes[n] = sum over the
 * row of mix64(j) mod 2^64 (j the partner's original index; mix64 is splitmix64's output
 * function).  Compares the pair set of build_neighbour_list (lennard_jones.rs:344-416) at
 * sizes where the full list is too large to move: no reference counterpart. */
pis_status pis_get_list_digest(pis_ctx* ctx, uint32_t* counts, uint64_t* hashes);

pis_status pis_get_stats(pis_ctx* ctx, pis_stats* out);
/* Replay one captured hipGraph per step in pis_verlet_step_nve / pis_batch_verlet_step_nve
 * (1) instead of launching each kernel (0, the default).  Ignored while kernel timing is on. */
pis_status pis_set_graph_mode(pis_ctx* ctx, int32_t enable);
/* Turn per-launch HIP-event timing of the force kernel on (1) or off (0); resets sums. */
pis_status pis_set_kernel_timing(pis_ctx* ctx, int32_t enable);

/* Measurement utility (no reference counterpart): `reps` launches of a plain streaming copy
 * kernel that reads and writes exactly `bytes` (rounded down to 32); *ms_per_copy gets the
 * average HIP-event time.  Used to scale hardware byte counters and to measure the
 * achievable HBM bandwidth on the running device. */
pis_status pis_copy_calibration(uint64_t bytes, int32_t reps, double* ms_per_copy);

/* ---- Batched ensembles: many independent systems advanced by one launch per kernel
 * per step (no reference counterpart beyond running PotentialManager per system). */
pis_status pis_batch_create(const pis_system_desc* descs, uint32_t n_systems, pis_ctx** out);
uint32_t pis_batch_size(const pis_ctx* ctx);
pis_status pis_sys_set_state(pis_ctx* ctx, uint32_t sys, const double* pos, const double* vel,
                             const double* force);
pis_status pis_sys_get_state(pis_ctx* ctx, uint32_t sys, double* pos, double* vel, double* force);
/* Initial forces for every system; pe_out[n_systems] may be NULL. */
pis_status pis_batch_compute_potential(pis_ctx* ctx, double* pe_out);
/* nsteps for every system, dt per system (dts[n_systems]); thermo_out is
 * [nsteps][n_systems] records or NULL. */
pis_status pis_batch_verlet_step_nve(pis_ctx* ctx, const double* dts, uint64_t nsteps,
                                     pis_thermo* thermo_out);

/* ---- NVT: Nose-Hoover chain thermostat around the NVE step (potential.rs:35-58,
 * ensemble/nvt.rs:4-119, the target ramp of simulation.rs:53-57).  The chain state lives
 * on the device and is updated by the step kernels; velocities are scaled by
 * exp(-dt xi0 / 2) before and after each NVE step.  The chain's kB is the system's kb. */
typedef struct {
    double t_start; /* fix nvt ... temp T_start T_end tau (NHThermostatChainArgs) */
    double t_end;
    double tau;
    int32_t chain; /* chain length (the reference uses 3) */
    int32_t reserved;
} pis_nhc_params;
/* NHThermostatChain::new_from_args for every system of the context (resets the chain). */
pis_status pis_nvt_init(pis_ctx* ctx, const pis_nhc_params* params);
/* nsteps of verlet_step_nvt_nhc with the target ramp over these nsteps; the per-step
 * record's etot is the reference's NVT Hamiltonian (PE + KE + chain KE). */
pis_status pis_verlet_step_nvt(pis_ctx* ctx, double dt, uint64_t nsteps, pis_thermo* per_step_out);
pis_status pis_batch_verlet_step_nvt(pis_ctx* ctx, const double* dts, uint64_t nsteps,
                                     pis_thermo* thermo_out);
/* Steps ramp_step0 .. ramp_step0 + nsteps - 1 of a run of ramp_total steps: the target
 * ramp of Simulation::step (simulation.rs:53-57) continues across calls, so a run can be
 * advanced in pieces (between trajectory dumps) with the reference's temperatures. */
pis_status pis_batch_verlet_step_nvt_ramp(pis_ctx* ctx, const double* dts, uint64_t nsteps,
                                          uint64_t ramp_step0, uint64_t ramp_total,
                                          pis_thermo* thermo_out);
/* The chain's thermostat velocities xi[chain] and current target temperature of system
 * `sys` (NHThermostatChain::xi / target_temperature, nvt.rs:9-13), for callers that keep a
 * host copy of the chain (either pointer may be NULL). */
pis_status pis_nvt_get_chain(pis_ctx* ctx, uint32_t sys, double* xi, double* target);

/* ---- NPT: MTK barostat (ensemble/npt.rs:10-88) around the NVT step ---------------------
 * `fix ... npt temp T0 T1 tau_t iso p0 p1 tau_p`: target pressure p0 * I (commands.rs:429-446;
 * p1 is parsed and unused, like the reference's end_pressure). */
typedef struct pis_mtk_params {
    double p_start;     /* target pressure p0 */
    double tau;         /* barostat time constant; w = 3N * 0.0083144621 * T0 * tau^2 */
    double reserved[2]; /* must be 0 */
} pis_mtk_params;
/* MTKBarostat::new_from_args + NHThermostatChain::new_from_args for every system
 * (replaces Simulation::run's construction, simulation.rs:17-22). */
pis_status pis_npt_init(pis_ctx* ctx, const pis_nhc_params* chain, const pis_mtk_params* barostat);
/* nsteps of verlet_step_npt_mtk (potential.rs:112-135) with the chain's target ramp as in
 * pis_batch_verlet_step_nvt_ramp (ramp_total 0: this call is the whole run).  The box is
 * general after the first step (the symmetrised pressure tensor has off-diagonal terms);
 * the list is rebuilt every step for the new box, as the reference's force call does.
 * The record's etot is the NPT Hamiltonian (simulation.rs:106-114).  PIS_ERR_CAPACITY if
 * the box gets thinner than the cell grid planned for it. */
pis_status pis_batch_verlet_step_npt(pis_ctx* ctx, const double* dts, uint64_t nsteps,
                                     uint64_t ramp_step0, uint64_t ramp_total,
                                     pis_thermo* thermo_out);
/* Current box matrix h (column-major, SimulationBox::h) and barostat momentum of system
 * `sys` (either pointer may be NULL). */
pis_status pis_npt_get_box(pis_ctx* ctx, uint32_t sys, double* h_out, double* momentum_out);

/* ---- Serving: continuous batching of NVE jobs on one GPU (no reference counterpart; the
 * reference runs one Simulation::run per process, simulation.rs:33-105).  A worker thread
 * keeps the admitted jobs' state resident and advances all of them together in rounds
 * of `round_steps` batched steps; between rounds it retires finished jobs and admits
 * queued ones in arrival order while `hbm_budget_bytes` and `max_systems` allow.  A job's
 * result is bit-identical to the same job run alone (pis_ctx_create + compute_potential +
 * verlet_step_nve with the same precision). */
typedef struct pis_server pis_server;
typedef struct {
    uint64_t job_id;
    int32_t status; /* PIS_OK, or the error that ended the job (pis_last_error has text) */
    int32_t reserved;
    uint64_t n_atoms;
    uint64_t steps;
    double submit_s, admit_s, done_s; /* seconds since pis_server_create */
    pis_thermo last;                  /* thermo of the job's final step */
} pis_job_result;
typedef struct {
    uint64_t submitted, completed, failed, rounds, atom_steps;
    uint64_t resident, pending, bytes_in_use, peak_bytes;
    uint32_t peak_resident, reserved;
    uint64_t graph_instantiations; /* round step graphs captured and instantiated (after
                                      every change of the resident set) */
    double graph_build_s;          /* host seconds those captures and instantiations took */
    uint64_t pool_bytes;           /* device memory the server's pool holds */
} pis_server_stats;

pis_status pis_server_create(uint64_t hbm_budget_bytes, uint32_t max_systems, uint32_t round_steps,
                             int32_t precision, pis_server** out);
/* Queues a job (arrays are copied; vel may be NULL for zero velocities).  The job's
 * desc->precision must equal the server's.  PIS_ERR_OUT_OF_MEMORY if the job alone exceeds
 * the budget. */
pis_status pis_server_submit(pis_server* srv, const pis_system_desc* desc, const double* pos,
                             const double* vel, double dt, uint64_t nsteps, uint64_t* job_id);
/* Blocks until the job ends (timeout_s < 0: no limit; PIS_ERR_TIMEOUT otherwise), then
 * collects it: result (may be NULL) and final pos/vel/force [3n] (each may be NULL). */
pis_status pis_server_wait(pis_server* srv, uint64_t job_id, double timeout_s, pis_job_result* result,
                           double* pos, double* vel, double* force);
pis_status pis_server_get_stats(pis_server* srv, pis_server_stats* out);
/* Stops the worker; queued and running jobs end with an error. */
pis_status pis_server_destroy(pis_server* srv);

/* ---- Trajectory frames with asynchronous D2H ----------------------------------------
 *
 * Simulation::run writes a frame every dump_step steps (simulation.rs:76-78, default 1,
 * simulation_context.rs:91).  pis_snapshot_positions enqueues, behind the steps already
 * queued on the context, a copy of system sys's positions (original order, AoS xyz) and
 * their transfer into host_pos on a copy stream of the context's own; it returns at once,
 * so the next steps run while the frame travels.  host_pos (3 n doubles) must stay valid
 * until pis_snapshot_wait returns; it is asynchronous only for pinned memory
 * (pis_host_alloc).  One snapshot is in flight per context: a second call first waits for
 * the previous one.  No reference counterpart for the two allocation calls. */
pis_status pis_snapshot_positions(pis_ctx* ctx, uint32_t sys, double* host_pos);
pis_status pis_snapshot_wait(pis_ctx* ctx);
pis_status pis_host_alloc(size_t bytes, void** out);
pis_status pis_host_free(void* p);

/* ---- Trajectory writer (host only) ------------------------------------------------
 *
 * DumpTraj::write_step (src/writers/dump_traj.rs:33-74): appends one LAMMPS text block to
 * the open file descriptor fd -- "ITEM: TIMESTEP", step, "ITEM: NUMBER OF ATOMS", n_atoms,
 * "ITEM: BOX BOUNDS pp pp pp", "0 h(k,k)" for k = 0..2 (h column-major, 9 doubles),
 * "ITEM: ATOMS id type x y z" and "i+1 type x y z" per atom (pos AoS xyz, original order).
 * Numbers are printed as Rust's `{}` prints them (shortest round-trip digits, positional).
 * threads <= 0: one formatting thread per hardware thread.  PIS_ERR_IO if a write fails. */
pis_status pis_dump_write_block(int fd, uint64_t step, uint64_t n_atoms, const double* pos,
                                const uint32_t* type_ids, const double* h, int32_t threads);
/* One f64 as Rust's `{}` prints it (at most 330 characters plus the terminating NUL);
 * *len receives the length.  PIS_ERR_INVALID_ARGUMENT if cap is too small. */
pis_status pis_format_f64(double x, char* out, size_t cap, size_t* len);

#ifdef __cplusplus
}
#endif
#endif /* PIS_HIP_H */
