/*
 * aurora_hip.h - C ABI of the MI355X-native AuroraPIC particle loop.
 *
 * The reference (AuroraPIC, C++20) has no FFI: its hot path is the call
 * surface of Species / Pusher / NullCollisionModel / FieldSolver classes,
 * dispatched by a RuntimePolicy whose "gpu" backend is reserved and
 * rejected (include/pic/Runtime.hpp:9, src/Runtime.cpp:73-74).  These entry
 * points are what that backend binds: plain pointers and sizes, FP64 at the
 * boundary, device-resident state between calls.  Each declaration names
 * the reference interface it replaces (paths relative to the reference
 * repository root).
 *
 * Conventions
 *  - Every function returns an aurora_status (0 = success).  On failure
 *    aurora_hip_last_error() returns the message; messages reuse the
 *    reference's exception text so callers can keep matching substrings.
 *  - A context is driven by one host thread.  Work is queued on the
 *    context's HIP stream; calls that return host data synchronise it.
 *  - Positions are in the reference's units (metres in SI, normalised
 *    otherwise), origin at node 0.  Velocities are 3 components per
 *    particle (1D3V and 2D3V carry the out-of-plane components).
 *  - Particle state on the device: SoA, integer cell index + FP32 (or FP64)
 *    in-cell offsets + FP32 (or FP64) staggered velocity v^{n-1/2}.
 *  - Charge density is accumulated in int64 fixed point (2^-40 of a
 *    particle per node), so it is bit-exact and independent of particle
 *    order, sorting and batching.
 */
#ifndef AURORA_HIP_H
#define AURORA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aurora_hip_ctx aurora_hip_ctx;

typedef enum aurora_status {
    AURORA_OK = 0,
    AURORA_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    AURORA_ERR_HIP = 2,              /* HIP runtime failure */
    AURORA_ERR_RATE_ABOVE_BOUND = 3, /* Collision.cpp:617-623, 677-684 */
    AURORA_ERR_CANDIDATE_LIMIT = 4,  /* Collision.cpp:864-870 */
    AURORA_ERR_CAPACITY = 5,         /* Simulation.cpp:945-963 */
    AURORA_ERR_B_OUT_OF_RANGE = 6,   /* PrescribedField.cpp:95-100 */
    AURORA_ERR_ENERGY_ABOVE_MAX = 7, /* Swarm.cpp:1021-1044 */
    AURORA_ERR_UNSUPPORTED = 8,      /* configuration outside this build */
    AURORA_ERR_NOT_FOUND = 9,        /* unknown species / request handle */
    AURORA_ERR_POPULATION = 10       /* branching swarm limit / extinction, Swarm.cpp:1069-1100 */
} aurora_status;

enum { AURORA_PERIODIC = 0, AURORA_DIRICHLET = 1 };            /* Types.hpp:31 */
enum { AURORA_PB_ABSORB = 1, AURORA_PB_REFLECT = 2,
       AURORA_PB_PERIODIC = 3 };                               /* Types.hpp:33 */
enum { AURORA_VELOCITY_HALF = 0, AURORA_VELOCITY_CENTRED = 1 };
enum { AURORA_ELASTIC = 0, AURORA_EXCITATION = 1, AURORA_IONIZATION = 2,
       AURORA_ATTACHMENT = 3, AURORA_CHARGE_EXCHANGE = 4 };    /* Config.hpp:42-48 */
enum { AURORA_SCATTER_ISOTROPIC = 0, AURORA_SCATTER_BACKWARD = 1,
       AURORA_SCATTER_HENYEY_GREENSTEIN = 2 };                 /* Config.hpp:33-37 */
enum { AURORA_FRAME_PROJECTILE = 0,
       AURORA_FRAME_CENTER_OF_MASS = 1 };                      /* Config.hpp:38-41 */

/* ---------------------------------------------------------------- context */
/* Runtime.cpp:22-76 (RuntimePolicy with backend = gpu).  device: HIP
 * ordinal.  The context owns a HIP stream at default priority. */
int aurora_hip_create(int device, aurora_hip_ctx** out);
int aurora_hip_destroy(aurora_hip_ctx* ctx);
/* Message of the last failure on this context (empty when none). */
const char* aurora_hip_last_error(const aurora_hip_ctx* ctx);
int aurora_hip_synchronize(aurora_hip_ctx* ctx);
/* Library and device description ("gfx950, 256 CUs ...") */
int aurora_hip_device_info(aurora_hip_ctx* ctx, char* buf, size_t len);

/* ------------------------------------------------------------------- mesh */
/* Grid(nx, L, bc) (Grid.cpp:19-22), Mesh2D (Mesh2D.cpp:49-70),
 * Mesh3D (Mesh3D.cpp:39-50) and FieldSolver(permittivity)
 * (FieldSolver.cpp:375-382).  n: nodes per axis; length: domain extents;
 * depth: out-of-plane depth used by the 2D deposit (Species2D.cpp:160-170),
 * ignored otherwise.  Binding a mesh drops all species. */
int aurora_hip_mesh_bind(aurora_hip_ctx* ctx, int dim, const int64_t* n,
                         const double* length, int boundary,
                         double permittivity, double depth);
/* Particle boundary policy per face (left,right,bottom,top,back,front),
 * ParticleBoundaryConfig3D (Simulation3D.hpp:17-24), already resolved. */
int aurora_hip_particle_boundaries(aurora_hip_ctx* ctx, const int* policy);

/* ---------------------------------------------------------------- species */
/* Species*(cfg) (Species.hpp:11, Species3D.hpp:27).  precision: 32 or 64
 * (offsets and velocities).  capacity: maximum particle slots, the
 * max_particles_per_species of the reference (Config.hpp:191). */
int aurora_hip_species_add(aurora_hip_ctx* ctx, double charge, double mass,
                           double weight, int64_t capacity, int precision,
                           int* species_id);
/* Replaces the species' particles (Species*::particles() assignment /
 * external-state ingestion, Simulation3D.cpp:270-286).  position:
 * count*dim, velocity: count*3, alive: count (NULL = all alive).
 * velocity_kind selects v^{n-1/2} (AURORA_VELOCITY_HALF) or the
 * time-centred v (AURORA_VELOCITY_CENTRED, then v_half = v until
 * aurora_hip_initialize).  Particle uids are assigned 0..count-1. */
int aurora_hip_species_upload(aurora_hip_ctx* ctx, int species_id,
                              int64_t count, const double* position,
                              const double* velocity, const uint8_t* alive,
                              int velocity_kind);
/* Species*::initialize (Species3D.cpp:28-162) for the measured workloads:
 * count particles uniform over the domain, Maxwellian velocities
 * (per-component sigma thermal_velocity, plus drift3), generated on the
 * device from a Philox4x32-10 stream keyed by seed.  Velocities are
 * time-centred until aurora_hip_initialize. */
int aurora_hip_species_init_maxwellian(aurora_hip_ctx* ctx, int species_id,
                                       int64_t count, double thermal_velocity,
                                       const double* drift3, uint64_t seed);
/* Slots in use (live + dead) and live particles (Species::live_count,
 * Species.cpp:353-355). */
int aurora_hip_species_count(aurora_hip_ctx* ctx, int species_id,
                             int64_t* slots, int64_t* live);
/* Copies the state back in storage order: position (slots*dim),
 * velocity_half (slots*3), alive (slots), uid (slots).  Any output may be
 * NULL.  capacity is the number of slots the caller's arrays hold. */
int aurora_hip_species_download(aurora_hip_ctx* ctx, int species_id,
                                int64_t capacity, double* position,
                                double* velocity_half, uint8_t* alive,
                                uint32_t* uid, int64_t* slots);
/* Time-centred velocity (synchronize_boris / synchronize_leapfrog,
 * Pusher.cpp:172-186) from the current field, storage order, slots*3. */
int aurora_hip_species_velocity(aurora_hip_ctx* ctx, int species_id,
                                double dt, int64_t capacity,
                                double* velocity);
/* Raw device state in storage order: cell (slots; linear index
 * (k ny + j) nx + i of the cell's lower corner node, -1 = dead), offsets
 * (slots*dim, the stored values widened exactly to double), velocity_half
 * (slots*3).  The representation the fixed-point deposit is defined on. */
int aurora_hip_species_download_raw(aurora_hip_ctx* ctx, int species_id,
                                    int64_t capacity, int32_t* cell,
                                    double* offsets, double* velocity_half,
                                    int64_t* slots);
/* Fixed-point charge accumulator of one species (node count int64). */
int aurora_hip_species_charge_fixed(aurora_hip_ctx* ctx, int species_id,
                                    int64_t* acc);

/* -------------------------------------------------------------- fields */
/* Uniform B for the Boris pusher (Simulation3DConfig::magnetic_field,
 * Simulation3D.hpp:46).  Zero with no profile keeps the leapfrog kick. */
int aurora_hip_set_uniform_b(aurora_hip_ctx* ctx, const double* b3);
/* TabulatedVectorField1D(axis, coordinates, values)
 * (PrescribedField.cpp:60-87); axis 0/1/2 = x/y/z; values rows*3. */
int aurora_hip_set_b_profile(aurora_hip_ctx* ctx, int axis, int64_t rows,
                             const double* coordinates, const double* values);
/* Sinusoidal electrode drive of the 1D Dirichlet solve,
 * Simulation::electrode_potential (Simulation.cpp:678-685): each side
 * offset + amplitude * sin(2 pi f t + phase). */
int aurora_hip_set_electrodes(aurora_hip_ctx* ctx, const double* left4,
                              const double* right4);
/* Overwrites the node electric field (E: dim arrays of node count),
 * e.g. to drive a kernel with a prescribed field. */
int aurora_hip_set_electric(aurora_hip_ctx* ctx, const double* E);
/* Node charge density, potential and field (any may be NULL).  phi is
 * only valid after a solve with want_phi (or aurora_hip_initialize). */
int aurora_hip_get_fields(aurora_hip_ctx* ctx, double* rho, double* phi,
                          double* E);
/* Overwrites rho (node count) for a standalone field solve. */
int aurora_hip_set_rho(aurora_hip_ctx* ctx, const double* rho);

/* ---------------------------------------------------------- operations */
/* Mesh::clear_charge + Species*::deposit_charge for every species
 * (Species.cpp:161-185, Mesh2D.cpp:90-144, Mesh3D.cpp:62-126). */
int aurora_hip_deposit(aurora_hip_ctx* ctx);
/* FieldSolver::solve (FieldSolver.cpp:384-405): 1D periodic spectral,
 * 1D Dirichlet tridiagonal with (phi_left, phi_right), 2D/3D periodic FFT.
 * want_phi also produces the potential (the field always). */
int aurora_hip_field_solve(aurora_hip_ctx* ctx, double phi_left,
                           double phi_right, int want_phi);
/* Gather + kick + drift + particle boundaries for every species
 * (Simulation3D.cpp:344-352, Simulation.cpp:1025-1079). */
int aurora_hip_push(aurora_hip_ctx* ctx, double dt);
/* Simulation*::initialize after the particles exist
 * (Simulation3D.cpp:288-299, Simulation.cpp:759-767): deposit, solve at
 * t = 0, then v_half from the time-centred velocity. */
int aurora_hip_initialize(aurora_hip_ctx* ctx, double dt);
/* nsteps of Simulation*::step (Simulation3D.cpp:342-365,
 * Simulation2D.cpp:1337-1364, Simulation.cpp:1008-1138): fused push and
 * deposit, field solve at t+dt, MCC where a model is loaded.  The
 * diagnostic synchronisation is computed on demand instead of every step
 * (aurora_hip_species_velocity, aurora_hip_energy).  On 1D meshes with an
 * MCC model that produces ionization products, the slot counts stay on the
 * device for the call: products are reserved and placed there, and an
 * overflow of a species' capacity (all models' products checked at once;
 * that step places none) returns AURORA_ERR_CAPACITY at the end of the
 * call, as the other errors the device detects do. */
int aurora_hip_step(aurora_hip_ctx* ctx, double dt, int64_t nsteps);
/* Simulation3D::sample (Simulation3D.cpp:551-579) and its 1D/2D
 * counterparts (Diagnostics.cpp:37-59, 114-140), FP64: kinetic energy per
 * species (time-centred v), field energy and charge L1 over node volumes
 * (a 2D node's area times the out-of-plane depth), live count. */
int aurora_hip_energy(aurora_hip_ctx* ctx, double dt, double* kinetic,
                      double* field_energy, double* charge_l1,
                      int64_t* live);
/* Boundary-loss ledger of a species since aurora_hip_initialize: absorbed
 * counts per face (6) and, in 1D, the represented kinetic energy at each
 * wall (2): species_boundary_losses() (BoundaryLoss1D, Simulation.hpp:20-25,
 * accumulated as Simulation.cpp:1046-1091). */
int aurora_hip_boundary_losses(aurora_hip_ctx* ctx, int species_id,
                               int64_t* absorbed6, double* energy2);
/* 1D electric work on a species since aurora_hip_initialize:
 * species_power_transfer() (SpeciesPower1D, Simulation.hpp:27-29): the
 * longitudinal kinetic-energy change of the kick up to a wall crossing, and
 * of every survivor from the start of the step to its synchronised velocity
 * after the field solve (Simulation.cpp:1046-1131).  Summed in a fixed
 * order on the device: identical from run to run.  0 outside 1D. */
int aurora_hip_power_transfer(aurora_hip_ctx* ctx, int species_id, double* electric_work);
/* Simulation time and step count kept by the context. */
int aurora_hip_time(aurora_hip_ctx* ctx, double* time, int64_t* step);
/* Tile shape (cells per axis) of the cell sort and of the LDS-tiled 2D and
 * 3D pushes, and its halo in cells; tx = 0 keeps the sort but disables the
 * tiled push.  Built 3D shapes: 8x8x8 (halo 1, 2), 8x8x4 (halo 1, 2),
 * 4x4x4, 16x8x4 (halo 2); 2D (tz ignored; 1 names them): 8x8 (halo 1, 2;
 * 512 threads, or 256 with prefetch), 16x16 and 16x8 (halo 2).
 * No reference counterpart (the reference keeps particles unsorted). */
int aurora_hip_set_tiling(aurora_hip_ctx* ctx, int tx, int ty, int tz,
                          int halo);
/* Launch variant of the LDS-tiled push: workgroup size (256 or 512) and
 * flags (bit 0: the next particle's state is loaded while the current one
 * is processed; bit 1: registers capped at 64 so four 512-thread workgroups
 * share a CU; bit 2: the field box in LDS by component instead of float4
 * nodes; bit 3: non-temporal particle stores).  Built: 512/0 for every
 * shape; 512/1, 256/1, 256/0, 512/3, 512/5, 512/7, 512/11, 512/13, 512/15
 * for 8x8x4 halo 2; 512/3 for 8x8x4 halo 1; 512/1, 256/1, 512/5, 512/13 for 8x8x8
 * halo 2; 512/1, 512/3, 512/5, 512/7, 512/13, 512/15 for 8x8x8 halo 1.
 * Results do not depend on it. */
int aurora_hip_set_push_variant(aurora_hip_ctx* ctx, int block, int prefetch);
/* Cell sort every interval steps (0 = never) and an immediate sort. */
int aurora_hip_set_sort_interval(aurora_hip_ctx* ctx, int64_t interval);
/* Fixed policy: at a sort step, a tile-sorted species is re-sorted only if
 * at least this fraction of its particle pushes since its last sort ended
 * outside their tile's LDS box (default 0 = always sort).  Sorting never
 * changes results (integer deposit), only speed. */
int aurora_hip_set_sort_threshold(aurora_hip_ctx* ctx, double fraction);
/* Sort policy.  AURORA_SORT_FIXED (default): every sort_interval steps.
 * AURORA_SORT_ADAPTIVE: each tile-sorted species is timed on the device
 * (events around its push and its sort) and re-sorted at the first step
 * whose push costs more than the average step cost of its current cycle,
 * the last sort included -- the renewal rule that minimises the long-run
 * cost per step when pushes slow down as particles leave their tiles;
 * max_interval > 0 caps the steps between sorts.  sort_interval is not
 * used by this policy. */
enum { AURORA_SORT_FIXED = 0, AURORA_SORT_ADAPTIVE = 1 };
int aurora_hip_set_sort_policy(aurora_hip_ctx* ctx, int policy, int64_t max_interval);
/* How a sort orders particles.  AURORA_SORT_LOCAL (default): a species
 * still tile-sorted from its last sort is re-sorted by a counting sort over
 * its old tile segments (particles move at most into a neighbouring tile
 * between sorts; those that moved further are placed through a short radix
 * sort), and by the full radix sort otherwise or when over 1/8 of the
 * particles moved further.  AURORA_SORT_RADIX: always the full stable radix
 * sort.  Both are deterministic and give the same set of particles per tile
 * (the order inside a tile can differ); results do not depend on it.  No
 * reference counterpart (the reference keeps particles in creation order). */
enum { AURORA_SORT_LOCAL = 0, AURORA_SORT_RADIX = 1 };
int aurora_hip_set_sort_method(aurora_hip_ctx* ctx, int method);
int aurora_hip_sort(aurora_hip_ctx* ctx);

/* ----------------------------------------------------------- checkpoint */
/* Checkpoint / restart of the context's device state (the reference's
 * restart files, Simulation3D.cpp:367-469, Simulation.cpp:1812-2900, guarded
 * by a physics signature like Collision.cpp:365-577): a versioned binary
 * dump of time, step, every species' particles as stored, boundary and MCC
 * counters.  Loading requires a context configured identically (mesh,
 * species, magnetic field, particle boundaries, electrodes, MCC model and
 * seed), else the physics-signature mismatch error; it then deposits and
 * solves, so continuing reproduces the uninterrupted run to the bit.
 * Batched swarms and ensembles are not included. */
int aurora_hip_checkpoint_save(aurora_hip_ctx* ctx, const char* path);
int aurora_hip_checkpoint_load(aurora_hip_ctx* ctx, const char* path);
int aurora_hip_physics_signature(aurora_hip_ctx* ctx, uint64_t* signature);

/* ------------------------------------------------------------------ MCC */
/* NullCollisionModel(config, particle_mass) (Collision.cpp:315-577) for
 * the target species.  Tables: per channel rows energies (J, strictly
 * increasing) and cross sections (m^2).  product_species: for ionization
 * the (secondary electron, ion) species ids, else -1.  neutral_temperature
 * > 0 samples bounded-Maxwellian neutrals (Collision.cpp:332-363).
 * seed keys the Philox stream (seed, particle uid, step). */
int aurora_hip_mcc_load(aurora_hip_ctx* ctx, int target_species, int nchan,
                        const int* kind, const double* threshold,
                        const int64_t* rows, const double* const* energy,
                        const double* const* sigma, double neutral_density,
                        double max_frequency, int64_t max_candidates,
                        double neutral_mass, double neutral_temperature,
                        const int* product_species, uint64_t seed);
/* Further named null-collision models (Config::collision_models,
 * Simulation.cpp:405-520): each on its own target species, with the
 * arguments of aurora_hip_mcc_load; *model receives its index (model 0 is
 * the one aurora_hip_mcc_load loads, which clears the others).  The step
 * applies the models in index order to the particles present at the start
 * of the step (Simulation.cpp:835-927), each on its own random stream, then
 * places all their ionization products after one all-or-nothing capacity
 * check (Simulation.cpp:929-1005).  Two models may not target one species. */
int aurora_hip_mcc_add(aurora_hip_ctx* ctx, int target_species, int nchan,
                       const int* kind, const double* threshold,
                       const int64_t* rows, const double* const* energy,
                       const double* const* sigma, double neutral_density,
                       double max_frequency, int64_t max_candidates,
                       double neutral_mass, double neutral_temperature,
                       const int* product_species, uint64_t seed, int* model);
/* aurora_hip_mcc_channel_options and aurora_hip_mcc_stats for the
 * context's model of index `model`. */
int aurora_hip_mcc_model_options(aurora_hip_ctx* ctx, int model, int channel,
                                 int angular, int energy_frame, int64_t rows,
                                 const double* energy,
                                 const double* mean_cosine);
int aurora_hip_mcc_model_stats(aurora_hip_ctx* ctx, int model, uint64_t* counts,
                               double* energy_change);
/* Config::velocity_dimensions (Config.hpp:172) of a 1D run: 3 (default)
 * collides 1D3V velocities; 1 applies the scalar 1V collide
 * (NullCollisionModel::collide(double&), Collision.cpp:831-905) to v_x, in
 * the step and in aurora_hip_mcc_collide, and rejects ionization and
 * anisotropic channels as the reference does. */
int aurora_hip_set_velocity_dimensions(aurora_hip_ctx* ctx, int dims);
/* Per-channel options of CollisionChannelConfig (Config.hpp:82-108) for
 * the context's MCC model (model = 0) or the swarm gas (model = 1):
 * angular_scattering (AURORA_SCATTER_*; anisotropic only for elastic
 * channels, Collision.cpp:455-461), with the Henyey-Greenstein mean-cosine
 * table (rows, energy J strictly increasing and covering the cross-section
 * table, mean cosine in (-1, 1); MeanCosineTable, Collision.cpp:239-313),
 * and energy_frame (AURORA_FRAME_*; centre of mass needs a positive
 * neutral mass and an elastic or charge-exchange channel,
 * Collision.cpp:483-500).  Reset to isotropic / projectile on every load. */
int aurora_hip_mcc_channel_options(aurora_hip_ctx* ctx, int model, int channel,
                                   int angular, int energy_frame, int64_t rows,
                                   const double* energy,
                                   const double* mean_cosine);
/* Accumulated CollisionStepStatistics (Collision.hpp:50-63) since load:
 * counts = [candidates, nulls, per-channel collisions (nchan)],
 * energy_change per channel (J, weighted by the species weight). */
int aurora_hip_mcc_stats(aurora_hip_ctx* ctx, uint64_t* counts,
                         double* energy_change);
/* NullCollisionModel::collide on an array of 3V velocities (in place) with
 * the stream keyed (seed, stream_id, uid = index, step): the same kernel
 * the step uses, exposed for per-particle parity checks.  stats as above
 * plus secondaries at counts[2+nchan]; sec_velocity receives up to
 * count*max_secondaries secondary velocities (may be NULL). */
int aurora_hip_mcc_collide(aurora_hip_ctx* ctx, int64_t count,
                           double* velocity, double dt, uint64_t step,
                           uint32_t stream_id, uint64_t* counts,
                           double* energy_change);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  aurora_hip_mcc_collide over device-resident FP32
 * velocities (count*3, interleaved, updated in place) that allocates
 * nothing: the same kernels and random streams, queued on the context's
 * stream without synchronising.  Device outputs, overwritten: counts
 * (nchan + 3: candidates, nulls, per channel, secondaries), energy_change
 * (nchan, J, unweighted) and status (an aurora_status: frequency bound,
 * candidate limit or non-finite velocity).  Its only scratch is a
 * fixed per-device array of 2048 x 4 tally rows of 19 words (1.19 MiB),
 * which bounds the launch to 2048 workgroups (particles are walked
 * grid-stride);
 * count is at most 4294967295.  Calls that overlap on one GPU share that
 * array and race. */
int aurora_hip_mcc_collide_device(aurora_hip_ctx* ctx, int64_t count,
                                  float* velocity, double dt, uint64_t step,
                                  uint32_t stream_id, uint64_t* counts,
                                  double* energy_change, uint32_t* status);

/* ------------------------------------------------------------ swarms */
/* Homogeneous electron swarm requests (run_field, Swarm.cpp:940-1250,
 * fixed population), batched: many requests of uneven particle count
 * advance together in persistent launches; each result is identical
 * whether the request runs alone or mixed with others. */
typedef struct aurora_hip_swarm_request {
    double reduced_field_td;      /* E/N (Td) */
    double timestep;              /* s; max_frequency * timestep <= 0.1 */
    int64_t particles;
    int64_t steps;
    int64_t burn_in;
    int64_t blocks;               /* uncertainty_blocks */
    double initial_mean_energy_ev;
    double max_energy_ev;
    uint64_t seed;
} aurora_hip_swarm_request;

typedef struct aurora_hip_swarm_result {
    double drift;                 /* m/s, -<v_x> */
    double drift_se;
    double mean_energy_ev;
    double mean_energy_se;
    double ionization_rate;       /* 1/s per electron */
    double ionization_rate_se;
    double rate[8];
    double rate_se[8];
    uint64_t collisions[8];
    uint64_t candidates;
    uint64_t nulls;
    double max_energy_ev;
    int status;                   /* aurora_status of the request */
    double longitudinal_diffusion; /* m^2/s, var(dx) / 2T over the sampling window */
    double transverse_diffusion;   /* m^2/s, var(dy, dz) / 4T (Swarm.cpp:1139-1176) */
} aurora_hip_swarm_result;

/* The gas of every swarm request on this context (GasDataset +
 * collision_config, Swarm.cpp:561-593). neutral_temperature (K, >= 0)
 * is SwarmBenchmarkConfig::neutral_temperature (Swarm.cpp:451-460);
 * a positive value samples Maxwellian neutral velocities. */
int aurora_hip_swarm_gas(aurora_hip_ctx* ctx, int nchan, const int* kind,
                         const double* threshold, const int64_t* rows,
                         const double* const* energy,
                         const double* const* sigma, double neutral_density,
                         double max_frequency, int64_t max_candidates,
                         double neutral_mass, double neutral_temperature);
/* Admits a request when its device footprint fits the HBM budget
 * (bytes; 0 = no limit); returns AURORA_ERR_CAPACITY otherwise. */
int aurora_hip_swarm_submit(aurora_hip_ctx* ctx,
                            const aurora_hip_swarm_request* request,
                            uint64_t hbm_budget, int64_t* handle);
/* Advances every admitted, unfinished request by up to max_steps steps in
 * one launch; *active receives the number still running. */
int aurora_hip_swarm_advance(aurora_hip_ctx* ctx, int64_t max_steps,
                             int64_t* active);
/* aurora_hip_swarm_advance without waiting: queues the launch on the
 * context's stream and returns at once, so other device work (another
 * context's ensemble, say) runs beside it and requests can be submitted
 * meanwhile (they join the next launch).  One launch is in flight per
 * context: a further launch first completes the previous one.
 * aurora_hip_swarm_collect completes the launch in flight (wait = 1) or
 * checks it (wait = 0, *busy = 1 while it runs); once it has completed, the
 * requests it finished report done in aurora_hip_swarm_poll. */
int aurora_hip_swarm_launch(aurora_hip_ctx* ctx, int64_t max_steps,
                            int64_t* active);
int aurora_hip_swarm_collect(aurora_hip_ctx* ctx, int wait, int* busy);
/* done = 1 once the request finished; result then holds its outputs and
 * the request's device memory is released. */
int aurora_hip_swarm_poll(aurora_hip_ctx* ctx, int64_t handle, int* done,
                          aurora_hip_swarm_result* result);
/* Device bytes a request of this size holds while admitted. */
uint64_t aurora_hip_swarm_footprint(const aurora_hip_swarm_request* request);

/* Branching / resampled population (population_model = branching_resampled:
 * run_field, Swarm.cpp:940-1250 with 1063-1104, systematic_resample 616-643,
 * block_slope_and_error 136-197).  Ionization appends the secondary at the
 * parent's weight, attachment removes the electron, and every step
 * resamples systematically back to `particles` electrons of equal weight,
 * so the total weight W tracks the physical growth.  Runs one request to
 * completion on the context's stream with the swarm gas (which may hold
 * attachment channels), seeded like swarm_submit.  result carries the drift,
 * energy and per-channel rates (weighted by the electron weight; no
 * diffusion in this model); growth the rest. */
typedef struct aurora_hip_swarm_growth {
    double initial_weight;            /* initial_total_electron_weight */
    double final_weight;              /* final_total_electron_weight */
    int64_t final_particles;
    double growth_rate, growth_rate_se;          /* 1/s, slope of log W */
    double attachment_rate, attachment_rate_se;  /* 1/s per electron */
    double net_creation_rate, net_creation_rate_se;
    int townsend_available;                      /* drift > 0 */
    double townsend_growth, townsend_growth_se;  /* growth / drift, 1/m */
    double townsend_rate_balance, townsend_rate_balance_se;  /* net / drift */
} aurora_hip_swarm_growth;
int aurora_hip_swarm_branching(aurora_hip_ctx* ctx, const aurora_hip_swarm_request* request,
                               int64_t population_limit, aurora_hip_swarm_result* result,
                               aurora_hip_swarm_growth* growth);

/* ---------------------------------------------------------- ensembles */
/* Batched 1D ensembles: many independent instances of one 1D discharge
 * (Simulation::step, Simulation.cpp:1008-1138, with the MCC step and
 * ionization products, Simulation.cpp:771-1006) advanced together.  The
 * reference runs one Simulation per process and has no batched form; the
 * per-instance operations are the context's own step, so an instance's
 * particles are bit-identical to running it alone on a context.
 * The template is the context's 1D mesh, FP32 species, particle
 * boundaries, electrodes and MCC model at creation.  capacity: slots per
 * instance, one per species. */
int aurora_hip_ensemble_create(aurora_hip_ctx* ctx, int64_t instances,
                               const int64_t* capacity, int64_t* handle);
/* Copies the context's current particles and fields (e.g. after
 * aurora_hip_initialize) into one instance; mcc_seed keys its collision
 * RNG streams.  All instances must be loaded at the same time/step. */
int aurora_hip_ensemble_load(aurora_hip_ctx* ctx, int64_t handle,
                             int64_t instance, uint64_t mcc_seed);
/* Steps captured per HIP graph (default 50; 0 = direct launches).  Dead
 * slots are compacted (stably) after every graph. */
int aurora_hip_ensemble_graph(aurora_hip_ctx* ctx, int64_t handle,
                              int steps_per_graph);
/* Advances every instance nsteps steps (field solves at t + dt with the
 * context's electrode potentials). */
int aurora_hip_ensemble_step(aurora_hip_ctx* ctx, int64_t handle, double dt,
                             int64_t nsteps);
int aurora_hip_ensemble_count(aurora_hip_ctx* ctx, int64_t handle,
                              int64_t instance, int species, int64_t* slots,
                              int64_t* live);
/* One instance's particles: position (m), v_half (3 per particle), alive,
 * uid; slots receives the slot count (live + dead). */
int aurora_hip_ensemble_download(aurora_hip_ctx* ctx, int64_t handle,
                                 int64_t instance, int species,
                                 int64_t capacity, double* position,
                                 double* velocity_half, uint8_t* alive,
                                 uint32_t* uid, int64_t* slots);
/* MCC counters [candidates, nulls, per channel..., secondaries], per
 * channel energy change (J, weighted by the species weight), wall counts and
 * energies [species][left,right] of one instance (CollisionStepStatistics,
 * BoundaryLoss1D).  The FP64 sums are formed in a fixed order (per-workgroup
 * rows folded in row order), so they repeat to the bit from
 * run to run; the order differs from a lone context's (whose slots are not
 * compacted), so they equal a lone run's to rounding. */
int aurora_hip_ensemble_stats(aurora_hip_ctx* ctx, int64_t handle,
                              int64_t instance, uint64_t* counts,
                              double* energy_change, uint64_t* absorbed,
                              double* absorbed_energy);
int aurora_hip_ensemble_fields(aurora_hip_ctx* ctx, int64_t handle,
                               int64_t instance, double* rho, double* phi,
                               double* E);
int aurora_hip_ensemble_time(aurora_hip_ctx* ctx, int64_t handle,
                             double* time, int64_t* step);
int aurora_hip_ensemble_destroy(aurora_hip_ctx* ctx, int64_t handle);

/* -------------------------------------------------------------- timing */
/* The attainable HBM bandwidth (GB/s, bytes read + written per second) of
 * hand-written streaming kernels on this device, best of `repeats` timed
 * launches over `bytes` of traffic: patterns 0, 2, 3 copy one array with
 * 16-byte accesses (grid-stride; one element per thread; the same with
 * non-temporal loads and stores); patterns 1, 4, 5 read and rewrite seven
 * 4-byte arrays in place (the 3D push's particle streams) the same three
 * ways.  No reference counterpart: it calibrates the HBM roofline the
 * kernels are measured against. */
int aurora_hip_stream_bandwidth(aurora_hip_ctx* ctx, int pattern, int64_t bytes,
                                int repeats, double* gbs);
/* Per-kernel HIP-event timing on the context stream (off by default). */
int aurora_hip_kernel_timing(aurora_hip_ctx* ctx, int enable);
/* Total milliseconds and launch count of kernel `name` since enabled. */
int aurora_hip_kernel_time(aurora_hip_ctx* ctx, const char* name,
                           double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* AURORA_HIP_H */
