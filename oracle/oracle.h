/*
 * oracle.h - CPU restatement of AuroraPIC's per-step particle loop (FP64).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker for the HIP
 * path: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library (libaurora_hip.so) never links or calls it.
 *
 * Every function restates an algorithm of the reference (cited file:line,
 * paths relative to the reference repository root) or states the
 * definition the GPU path must reproduce bit-for-bit (fixed-point deposit,
 * Philox4x32-10 uniforms).  Pins (see tests/test_oracle_*.py):
 *   or_deposit_cic        reference CIC tests + live reference (oracle/_ref)
 *   or_deposit_fixed      exact-sum invariants, permutation invariance,
 *                         agreement with or_deposit_cic to 2^-40 per particle
 *   or_gather             affine exactness + live reference
 *   or_boris_advance      cyclotron-angle closed form + live reference
 *   or_table_eval         reference interpolation tests + live reference
 *   or_poisson_*          analytic sine modes + live reference
 *   or_philox4x32_10      Random123 known-answer vectors (Salmon et al.)
 *   or_mcc_collide_3v     closed-form rates/conservation + live reference
 *                         statistics (mt19937_64 streams cannot be matched)
 *   or_swarm_run          zero-cross-section closed form + live reference
 */
#ifndef AURORA_ORACLE_H
#define AURORA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_PERIODIC = 0, OR_DIRICHLET = 1 };
enum { OR_PB_ABSORB = 1, OR_PB_REFLECT = 2, OR_PB_PERIODIC = 3 };

/* ---- mesh geometry (Grid.cpp:19-30, Mesh2D.cpp:49-88, Mesh3D.cpp:39-60) */
double or_spacing(int64_t n, double L, int bc);
double or_node_volume(int dim, const int64_t* n, const double* d, int bc,
                      int64_t i, int64_t j, int64_t k);

/* ---- A2-A4: reference FP64 CIC deposit, particle order.
 * pos: np*dim (x[,y[,z]] per particle). rho accumulates (not cleared). */
void or_deposit_cic(int dim, const int64_t* n, const double* L, int bc,
                    int64_t np, const double* pos, const uint8_t* alive,
                    double charge, double weight, double* rho);

/* ---- fixed-point deposit definition the GPU reproduces bit-exactly.
 * cell: linear cell index (-1 = dead); off: np*dim offsets in cell units.
 * Shape weights in FP64 from the offsets, quantised to 2^-40 units with
 * round-to-nearest-even (3D: the product wx * fl(wy * wz) rounded once, by
 * one fma); acc accumulates int64 per node. */
void or_deposit_fixed(int dim, const int64_t* n, int bc, int64_t np,
                      const int32_t* cell, const double* off, int64_t* acc);
/* rho[node] += Sum_s (double)acc_s[node] * (q_s w_s 2^-40) / V(node), in
 * species order, exactly as the device conversion kernel does. */
void or_fixed_to_rho(int dim, const int64_t* n, const double* L, int bc,
                     int nspecies, const int64_t* const* acc,
                     const double* charge_weight, double* rho);

/* ---- A5: reference gather (FieldSolver.cpp:909-1038). E: dim arrays
 * concatenated (Ex | Ey | Ez), each n-node long. out: np*dim. */
void or_gather(int dim, const int64_t* n, const double* L, int bc,
               int64_t np, const double* pos, const double* E, double* out);

/* ---- A6: Pusher.cpp:20-29 boris_advance */
void or_boris_advance(const double* v, const double* E, const double* B,
                      double qm, double h, double* out);

/* ---- A7: PrescribedField.cpp:89-113. Returns 0, or -1 when q lies
 * outside [coords[0], coords[rows-1]] (reference throws out_of_range). */
int or_table_eval(int64_t rows, const double* coords, const double* values,
                  double q, double* out3);

/* ---- A10-A12 Poisson (FP64). All write phi and E (dim arrays). */
void or_poisson_periodic_1d(int64_t n, double L, double eps,
                            const double* rho, double* phi, double* E);
void or_poisson_dirichlet_1d(int64_t n, double L, double eps,
                             const double* rho, double phi_left,
                             double phi_right, double* phi, double* E);
void or_poisson_periodic_2d(int64_t nx, int64_t ny, double Lx, double Ly,
                            double eps, const double* rho, double* phi,
                            double* Ex, double* Ey);
void or_poisson_periodic_3d(int64_t nx, int64_t ny, int64_t nz, double Lx,
                            double Ly, double Lz, double eps,
                            const double* rho, double* phi, double* Ex,
                            double* Ey, double* Ez);

/* ---- structured 2D/3D step (Simulation3D.cpp:342-365,
 * Simulation2D.cpp:1337-1364): kick (Boris when any B), drift, particle
 * boundaries, for one species.  B: uniform (3) when brows == 0, otherwise
 * a 1D table along axis baxis with brows rows.  pb: per-face particle
 * boundary (left,right,bottom,top,back,front).  absorbed: 6 counters.
 * Returns 0, or -1 when a B-table query falls outside the table. */
int or_push_species(int dim, const int64_t* n, const double* L, int bc,
                    int64_t np, double* pos, double* vh, uint8_t* alive,
                    const double* E, double qm, double dt,
                    const double* Buni, int baxis, int64_t brows,
                    const double* bcoords, const double* bvalues,
                    const int* pb, int64_t* absorbed);
/* synchronize (Pusher.cpp:172-186) -> time-centred v (np*3). */
int or_sync_species(int dim, const int64_t* n, const double* L, int bc,
                    int64_t np, const double* pos, const double* vh,
                    const uint8_t* alive, const double* E, double qm,
                    double dt, const double* Buni, int baxis, int64_t brows,
                    const double* bcoords, const double* bvalues, double* v);
/* half-step init: v_half from v (Pusher.cpp:66-68, 158-160). */
int or_step3d(const int64_t* n, const double* L, int nspecies, const int64_t* np,
              double* const* pos, double* const* vh, uint8_t* const* alive,
              double* const* v, const double* qm, const double* charge,
              const double* weight, double eps, double dt, const double* Buni,
              int baxis, int64_t brows, const double* bcoords, const double* bvalues,
              int threads, double* rho, double* phi, double* E);
int or_init_half_species(int dim, const int64_t* n, const double* L, int bc,
                         int64_t np, const double* pos, const double* v,
                         const uint8_t* alive, const double* E, double qm,
                         double dt, const double* Buni, int baxis,
                         int64_t brows, const double* bcoords,
                         const double* bvalues, double* vh);

/* ---- Philox4x32-10 (Salmon et al., SC'11): ctr[4], key[2] -> out[4] */
void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key,
                      uint32_t* out);
/* uniform in (0,1): (u >> 8) * 2^-24 + 2^-25, identical on the device */
double or_u01(uint32_t u);

/* ---- A13 MCC model (Collision.cpp:315-993), cold or thermal neutrals. */
enum { OR_ELASTIC = 0, OR_EXCITATION = 1, OR_IONIZATION = 2,
       OR_ATTACHMENT = 3, OR_CHARGE_EXCHANGE = 4 };
/* angular scattering (Config.hpp:33-37) and collision energy frame
 * (Config.hpp:38-41) per channel */
enum { OR_ISOTROPIC = 0, OR_BACKWARD = 1, OR_HENYEY_GREENSTEIN = 2 };
enum { OR_FRAME_PROJECTILE = 0, OR_FRAME_CENTER_OF_MASS = 1 };
typedef struct {
    int nchan;
    int kind[8];
    double threshold[8];          /* J (same units as 0.5 m v^2) */
    int64_t rows[8];
    const double* energy[8];      /* strictly increasing, J */
    const double* sigma[8];       /* m^2 */
    double neutral_density;
    double max_frequency;
    int64_t max_candidates;
    double particle_mass;
    double neutral_mass;          /* 0 = infinitely heavy, cold */
    double neutral_vstd;          /* thermal neutral sigma, 0 = cold */
    int angular[8];               /* OR_ISOTROPIC / OR_BACKWARD / OR_HENYEY_GREENSTEIN */
    int frame[8];                 /* OR_FRAME_* */
    int64_t mc_rows[8];           /* Henyey-Greenstein mean-cosine table */
    const double* mc_energy[8];   /* J, strictly increasing */
    const double* mc_value[8];    /* mean cosine in (-1, 1) */
} or_mcc_model;

typedef struct {
    uint64_t candidates, nulls;
    uint64_t chan[8];
    double energy_change[8];
    int nsec;                     /* secondaries produced (<= 8 stored) */
    double sec_v[8][3];
    double ion_v[8][3];
    int status;                   /* 0 ok; 1 rate above bound; 2 candidate
                                     limit; 3 non-finite velocity */
    double margin;                /* smallest relative distance of any branch
                                     decision of the call to its boundary
                                     (waiting time vs the time left, channel
                                     selection vs the cumulative rates, energy
                                     vs a threshold, scatter sign, neutral
                                     rejection); 1 when none was taken */
} or_mcc_stats;

/* RNG stream for one particle-step: key=(seed lo, seed hi),
 * ctr=(block, uid, step lo, stream). Draws are consumed in a fixed order
 * (the device kernel consumes the same order). */
typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int used;
} or_stream;
void or_stream_init(or_stream* s, uint64_t seed, uint32_t stream,
                    uint32_t uid, uint64_t step);
double or_stream_u01(or_stream* s);

double or_xs_eval(int64_t rows, const double* e, const double* s, double x);
void or_mcc_collide_3v(const or_mcc_model* m, double* v, double dt,
                       or_stream* rng, or_mcc_stats* st);
void or_mcc_collide_1v(const or_mcc_model* m, double* v, double dt,
                       or_stream* rng, or_mcc_stats* st);

/* ---- A15 swarm run_field (Swarm.cpp:940-1250), fixed population. */
typedef struct {
    double neutral_density, max_frequency, timestep, reduced_field_td;
    int64_t steps, burn_in, particles, blocks;
    double initial_mean_energy_ev, max_energy_ev;
} or_swarm_cfg;
typedef struct {
    double drift, drift_se, mean_energy_ev, mean_energy_se;
    double ion_rate, ion_rate_se;
    double rate[8], rate_se[8];
    uint64_t collisions[8], candidates, nulls;
    double max_energy_ev;
    int status;                   /* 0 ok; 4 energy above max; else MCC */
    double longitudinal_diffusion, transverse_diffusion;  /* m^2/s */
} or_swarm_result;
/* initial velocities vel0 (particles*3) are inputs; RNG stream per particle
 * step keyed (seed, stream=instance, uid=particle index, step). */
void or_swarm_run(const or_mcc_model* m, const or_swarm_cfg* c,
                  const double* vel0, uint64_t seed, uint32_t instance,
                  or_swarm_result* r);
void or_block_mean_error(int64_t n, const double* samples, int64_t blocks,
                         double* mean, double* se);

/* ---- swarm run_field, branching / resampled population (Swarm.cpp:1063-1104,
 * 1214-1245; systematic_resample 616-643; block_slope_and_error 136-197). */
typedef struct {
    double initial_weight, final_weight;
    int64_t final_particles;
    double growth_rate, growth_rate_se;             /* 1/s */
    double attachment_rate, attachment_rate_se;
    double net_creation_rate, net_creation_rate_se;
    int townsend_available;
    double townsend_growth, townsend_growth_se;     /* 1/m */
    double townsend_rate_balance, townsend_rate_balance_se;
} or_swarm_growth;
/* Slope of samples against time (spacing dt) per equal block, mean and
 * standard error over the blocks.  Returns 0, or -1 when a block would hold
 * fewer than two samples or n is not a multiple of blocks. */
int or_block_slope_error(int64_t n, const double* samples, double dt, int64_t blocks,
                         double* slope, double* se);
/* Systematic resampling of M equal-weight electrons to N: output j takes
 * source ceil((u + j) M / N) - 1 (u in [0, 1) the run's draw for the step),
 * the index where the cumulative weight first reaches the j-th threshold. */
int64_t or_systematic_source(int64_t j, int64_t M, int64_t N, double u);
/* status: 0 ok, 4 energy above max, 5 population limit, 6 extinct, else MCC.
 * Streams: collisions of slot p at step s (seed, instance, p, s); the
 * resampling draw of step s (seed, instance + 1, 0xFFFFFFFF, s). */
void or_swarm_branching_run(const or_mcc_model* m, const or_swarm_cfg* c,
                            int64_t population_limit, const double* vel0, uint64_t seed,
                            uint32_t instance, or_swarm_result* r, or_swarm_growth* g);

#ifdef __cplusplus
}
#endif
#endif
